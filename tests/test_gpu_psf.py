"""eROSITA PSF interpolation on the GPU against the oracle, its window
layout against the cut of the full frames, and both as linpatch kernels."""
import numpy as np
import pytest
import torch

from oracle import psf as op

pytestmark = pytest.mark.gpu


def _infos(E=2, seed=0):
    rng = np.random.default_rng(seed)
    rs = np.array([0.0, 6.0, 15.0, 3.0, 25.0])
    ph, pw = 61, 57
    a, b = np.meshgrid(np.arange(ph) - 30.0, np.arange(pw) - 28.0, indexing="ij")
    infos = []
    for e in range(E):
        psfs = []
        for k, r in enumerate(rs):
            s1, s2 = 1.5 + 0.2 * r + 0.3 * e, 1.2 + 0.1 * r
            g = np.exp(-0.5 * ((a / s1) ** 2 + (b / s2) ** 2)) + 0.01 * rng.random((ph, pw))
            psfs.append(g / g.sum())
        infos.append(dict(psfs=np.array(psfs), rs=rs,
                          patch_center_ids=np.tile([30.3, 27.8], (rs.size, 1)),
                          patch_deltas=(0.7, 0.9), pointing_center=(20.5, 18.25)))
    return infos


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_full_frames_vs_oracle(ja, dtype, tol):
    H = W = 64
    infos = _infos()
    dom = ja.Domain((2, H, W), (1.0, 0.6, 0.6))
    got = ja.psf_interpolator(dom, 4, infos, dtype=dtype)
    want = op.psf_interpolator((2, H, W), (1.0, 0.6, 0.6), 4, infos)
    assert got.shape == want.shape == (16, 2, H, W)
    assert _rel(got, want) <= tol


def test_window_layout_is_the_linpatch_cut(ja):
    H = W = 64
    npatch = 4
    infos = _infos()
    dom = ja.Domain((2, H, W), (1.0, 0.6, 0.6))
    full = ja.psf_interpolator(dom, npatch, infos)
    win = ja.psf_interpolator(dom, npatch, infos, layout="window")
    w = 2 * (H // npatch)
    c = (H - w) // 2
    cut = np.fft.fftshift(full, axes=(-2, -1))[..., c:c + w, c:c + w]
    assert win.shape == (16, 2, w, w)
    assert np.array_equal(win, cut)


def test_interpolated_psfs_drive_linpatch(ja):
    """Full frames (the reference route) and windows give the same PSF
    operator."""
    H = W = 64
    npatch, margin = 4, 8
    infos = _infos()
    dom = ja.Domain((2, H, W), (1.0, 1.0, 1.0))
    full = ja.psf_interpolator(dom, npatch, infos)
    win = ja.psf_interpolator(dom, npatch, infos, layout="window")
    x = np.random.default_rng(1).random((2, H, W))
    a = ja.LinpatchOperator((2, H, W), npatch, margin, dtype=torch.float64)
    a.set_kernel(full, layout="full")
    b = ja.LinpatchOperator((2, H, W), npatch, margin, dtype=torch.float64)
    b.set_kernel(win, layout="window")
    ya = a.forward(torch.as_tensor(x, device="cuda")).cpu().numpy()
    yb = b.forward(torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert _rel(ya, yb) <= 1e-13
    # and the reference entry point with the full frames
    yc = ja.linpatch_convolve(x, dom, full, npatch, margin)
    assert _rel(np.asarray(yc), ya) <= 1e-13


def test_errors(ja):
    infos = _infos(1)
    with pytest.raises(ValueError):
        ja.psf_interpolator(ja.Domain((1, 30, 30), (1, 1, 1)), 4, infos)
    bad = dict(infos[0], patch_center_ids=np.tile([100.0, 1.0], (5, 1)))
    with pytest.raises(ValueError):
        ja.psf_interpolator(ja.Domain((1, 32, 32), (1, 1, 1)), 4, [bad])
    with pytest.raises(ValueError):
        ja.psf_interpolator(ja.Domain((1, 32, 32), (1, 1, 1)), 4, [dict(infos[0], rs=[0.0])])
