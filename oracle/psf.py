"""eROSITA PSF interpolation onto linpatch patches (float64 restatement).

Test infrastructure only.  Follows jubik/instruments/erosita/erosita_psf_utils.py:
to_r_phi (:17-28), to_ra_dec (:31-38), get_interpolation_weights (:41-72),
to_patch_coordinates (:75-87), get_psf (:90-164) and psf_interpolator
(:185-264).  jax.scipy.ndimage.map_coordinates(order=1, mode='nearest') is
restated with scipy.ndimage.map_coordinates, whose order-1 'nearest' mode is
the same clamped bilinear interpolation.
"""
import numpy as np
from scipy.ndimage import map_coordinates


def to_r_phi(cc):
    """(x, y) -> (r, phi), phi measured from the y axis (erosita_psf_utils.py:17-28)."""
    r = np.sqrt(cc[..., 0] ** 2 + cc[..., 1] ** 2)
    phi = np.angle(cc[..., 0] + 1j * cc[..., 1]) - np.pi / 2.0
    return np.stack((r, phi), axis=-1)


def to_ra_dec(rp):
    """(r, phi) -> (x, y) (erosita_psf_utils.py:31-38)."""
    x = rp[..., 0] * np.cos(rp[..., 1] + np.pi / 2.0)
    y = rp[..., 0] * np.sin(rp[..., 1] + np.pi / 2.0)
    return np.stack((x, y), axis=-1)


def get_interpolation_weights(rs, r):
    """Linear weights over the sorted radii rs for radius r, the end PSFs
    repeated outside [rs[0], rs[-1]] (erosita_psf_utils.py:41-72; the three
    cases are summed as the reference sums its conds)."""
    rs = np.asarray(rs, dtype=float)
    res = np.zeros(rs.shape)
    if r <= rs[0]:
        res[0] += 1.0
    if r >= rs[-1]:
        res[-1] += 1.0
    for i in range(rs.size - 1):
        if rs[i] < r <= rs[i + 1]:
            wgt = (r - rs[i]) / (rs[i + 1] - rs[i])
            res[i] += 1.0 - wgt
            res[i + 1] += wgt
    return res


def to_patch_coordinates(dcoords, patch_center, patch_delta):
    """Sky offsets -> index coordinates of a PSF grid (erosita_psf_utils.py:75-87)."""
    res = (dcoords + np.asarray(patch_center) * np.asarray(patch_delta)) / np.asarray(patch_delta)
    return np.moveaxis(res, -1, 0)


def get_psf(psfs, rs, patch_center_ids, patch_deltas, pointing_center):
    """The PSF evaluator of erosita_psf_utils.py:90-164."""
    rs = np.asarray(rs, dtype=float)
    if psfs.shape[0] != rs.size or patch_center_ids.shape[0] != rs.size:
        raise ValueError
    sort = np.argsort(rs)
    rs, psfs, patch_center_ids = rs[sort], psfs[sort], patch_center_ids[sort]
    patch_deltas = np.array(list(patch_deltas), dtype=float)
    pointing_center = np.array(list(pointing_center), dtype=float)
    for pp, cc in zip(psfs, patch_center_ids):
        if np.any(cc > np.array(list(pp.shape))) or np.any(cc < 0):
            raise ValueError

    def psf(ra, dec, dra, ddec):
        cc = np.array([ra, dec], dtype=float) - pointing_center
        rp = to_r_phi(cc)
        wgts = get_interpolation_weights(rs, rp[0])
        ic = to_r_phi(np.stack((dra, ddec), axis=-1))
        ic[..., 1] = ic[..., 1] - rp[1]
        ic = to_ra_dec(ic)
        res = np.zeros(ic.shape[:-1])
        for ids, pp, ww in zip(patch_center_ids, psfs, wgts):
            q = to_patch_coordinates(ic, ids, patch_deltas)
            res += ww * map_coordinates(pp, q, order=1, mode="nearest")
        return res
    return psf


def psf_interpolator(shape, distances, npatch, psf_infos):
    """(P, E, H, W) full-frame patch PSFs, rolled to the origin
    (erosita_psf_utils.py:185-264); one info dict per energy."""
    H, W = shape[-2], shape[-1]
    dist = (distances[-2], distances[-1])
    for ss in (H, W):
        if ss % npatch != 0 or ss % 2 != 0:
            raise ValueError
    patch_shp = (H // npatch, W // npatch)
    all_patches = []
    for info in psf_infos:
        func = get_psf(info["psfs"], info["rs"], info["patch_center_ids"], info["patch_deltas"],
                       info["pointing_center"])
        c_p = [(np.arange(ss) - ss / 2) * dd for ss, dd in zip((H, W), dist)]
        c_p = np.meshgrid(*c_p, indexing="ij")
        d_dec, d_ra = c_p[0], -1.0 * c_p[1]
        centers = [np.array([(i * ss + ss / 2) * dd for i in range(npatch)])
                   for ss, dd in zip(patch_shp, dist)]
        centers = np.meshgrid(*centers, indexing="xy")
        c_dec = centers[0].flatten()
        c_ra = np.flip(centers[1].flatten())
        pats = [np.roll(np.roll(func(ra, dec, d_ra, d_dec), -H // 2, axis=0), -W // 2, axis=1)
                for ra, dec in zip(c_ra, c_dec)]
        all_patches.append(np.array(pats))
    return np.moveaxis(np.stack(all_patches), 0, -3)
