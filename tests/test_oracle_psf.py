"""The PSF-interpolation oracle (oracle/psf.py) pinned by closed forms.

The reference has no test for this path.  Bilinear interpolation reproduces
an affine function of the grid indices exactly, so a table
psf_k(a, b) = c_k + g_k (a - ca) + h_k (b - cb) gives an analytic patch PSF:
at pixel offset (d_ra, d_dec) rotated by -phi_patch to (x, y) it is
sum_k w_k (c_k + g_k x / d0 + h_k y / d1).  This pins the rotation sense,
which image axis is ra / dec, the patch centres and their order, the
radial weights and the roll to the origin.
"""
import numpy as np
import pytest

from oracle import psf as op


def test_polar_round_trip():
    rng = np.random.default_rng(0)
    cc = rng.normal(size=(50, 2)) * 7
    back = op.to_ra_dec(op.to_r_phi(cc))
    assert np.allclose(back, cc, rtol=0, atol=1e-12)
    # phi is measured from the +y axis: (0, 1) has phi 0, (-1, 0) has phi pi/2
    assert np.allclose(op.to_r_phi(np.array([0.0, 2.0])), [2.0, 0.0])
    assert np.allclose(op.to_r_phi(np.array([-1.0, 0.0])), [1.0, np.pi / 2])


def test_interpolation_weights():
    rs = np.array([0.0, 2.0, 5.0, 9.0])
    assert np.allclose(op.get_interpolation_weights(rs, -1.0), [1, 0, 0, 0])
    assert np.allclose(op.get_interpolation_weights(rs, 12.0), [0, 0, 0, 1])
    assert np.allclose(op.get_interpolation_weights(rs, 3.5), [0, 0.5, 0.5, 0])
    assert np.allclose(op.get_interpolation_weights(rs, 2.0), [0, 1, 0, 0])
    for r in np.linspace(0.01, 8.99, 37):
        w = op.get_interpolation_weights(rs, r)
        assert np.isclose(w.sum(), 1.0) and np.isclose(w @ rs, r)


def _affine_infos(E=2):
    rs = np.array([8.0, 0.0, 30.0])          # unsorted on purpose
    ph = pw = 81
    ca, cb = 40.0, 40.0
    a, b = np.meshgrid(np.arange(ph), np.arange(pw), indexing="ij")
    coef = [(5.0, 0.01, -0.02), (3.0, -0.015, 0.005), (1.0, 0.02, 0.03)]
    infos = []
    for e in range(E):
        psfs = np.stack([c + g * (a - ca) + h * (b - cb) for (c, g, h) in coef]) * (1 + e)
        infos.append(dict(psfs=psfs, rs=rs, patch_center_ids=np.tile([ca, cb], (3, 1)),
                          patch_deltas=(0.8, 1.25), pointing_center=(11.0, 9.0)))
    return infos, coef, rs


def test_affine_table_gives_closed_form():
    H = W = 32
    npatch, dist = 4, (0.5, 0.75)
    infos, coef, rs = _affine_infos()
    out = op.psf_interpolator((2, H, W), (1.0,) + dist, npatch, infos)
    assert out.shape == (npatch * npatch, 2, H, W)
    ps = H // npatch
    for p in range(npatch * npatch):
        dec = ((p % npatch) * ps + ps / 2) * dist[0]
        ra = (((npatch * npatch - 1 - p) // npatch) * ps + ps / 2) * dist[1]
        cx, cy = ra - 11.0, dec - 9.0
        rp, php = np.hypot(cx, cy), np.arctan2(cy, cx) - np.pi / 2
        w = op.get_interpolation_weights(np.sort(rs), rp)
        order = np.argsort(rs)
        for i in (0, 5, 17, 31):
            for j in (0, 9, 16, 30):
                d_dec = (i - H / 2) * dist[0]
                d_ra = -(j - W / 2) * dist[1]
                rr, ang = np.hypot(d_ra, d_dec), np.arctan2(d_dec, d_ra) - np.pi / 2 - php
                x, y = rr * np.cos(ang + np.pi / 2), rr * np.sin(ang + np.pi / 2)
                want = sum(wk * (coef[order[k]][0] + coef[order[k]][1] * x / 0.8 +
                                 coef[order[k]][2] * y / 1.25)
                           for k, wk in enumerate(w))
                # rolled to the origin by (-H/2, -W/2)
                got = out[p, 0, (i - H // 2) % H, (j - W // 2) % W]
                assert np.isclose(got, want, rtol=1e-12, atol=1e-12)
                assert np.isclose(out[p, 1, (i - H // 2) % H, (j - W // 2) % W], 2 * want,
                                  rtol=1e-12, atol=1e-12)


def test_nearest_mode_clamps_outside_queries():
    """Far outside the table the value is the nearest edge value."""
    psfs = np.zeros((1, 5, 5))
    psfs[0, 4, :] = 7.0                       # the row at a = 4
    info = dict(psfs=psfs, rs=np.array([0.0]), patch_center_ids=np.array([[2.0, 2.0]]),
                patch_deltas=(1.0, 1.0), pointing_center=(0.0, 0.0))
    f = op.get_psf(info["psfs"], info["rs"], info["patch_center_ids"], info["patch_deltas"],
                   info["pointing_center"])
    # patch centre on the +y axis (phi = 0, no rotation): x offset 100 -> a = 102 -> row 4
    v = f(0.0, 3.0, np.array([[100.0]]), np.array([[0.0]]))
    assert np.isclose(v[0, 0], 7.0)
    v = f(0.0, 3.0, np.array([[-100.0]]), np.array([[0.0]]))
    assert np.isclose(v[0, 0], 0.0)


def test_invalid_inputs_raise():
    infos, _, _ = _affine_infos(1)
    with pytest.raises(ValueError):
        op.psf_interpolator((1, 30, 30), (1, 1, 1), 4, infos)
    with pytest.raises(ValueError):
        op.psf_interpolator((1, 18, 18), (1, 1, 1), 2, [dict(infos[0], rs=np.array([0.0]))])
