"""float64 restatement of jubik/response.py (TEST INFRASTRUCTURE ONLY)."""
import numpy as np


def build_exposure_function(exposures, exposure_cut=None):
    """jubik/response.py:12-45.

    Entries below the cut are zeroed *in place* in the caller's array, then
    the returned closure multiplies by the (shared) exposure array.
    """
    if exposure_cut is not None:
        if exposure_cut < 0:
            raise ValueError("exposure_cut should be non-negative or None!")
        exposures[exposures < exposure_cut] = 0

    def exposure(x):
        return exposures * x
    return exposure


def build_readout_function(flags, threshold=None, keys=None):
    """jubik/response.py:48-112.  Returns {key: x[i][flags'[i] != 0]}.

    The flags below the threshold are zeroed in place; read-out order is C
    order over (E, H, W) of each module.
    """
    if threshold is not None and threshold < 0:
        raise ValueError("threshold should be positive!")
    if threshold is not None:
        flags[flags < threshold] = 0
    masked = flags == 0
    if keys is None:
        keys = ['masked input']
    elif len(keys) != flags.shape[0]:
        raise ValueError("length of keys should match the number of flag maps.")

    def readout(x):
        if masked.ndim != 4:
            raise ValueError("flags should have shape (n, m, q, l)!")
        if np.ndim(x) != 4:
            raise ValueError("input should have shape (n, m, q, l)!")
        return {key: np.asarray(x)[i][~masked[i]] for i, key in enumerate(keys)}
    return readout


def readout_transpose(vec, flags_nonzero):
    """Transpose of the read-out: scatter each module's vector back into a
    zero (M, E, H, W) array at the kept positions (C order)."""
    out = np.zeros(flags_nonzero.shape, dtype=np.float64)
    for i, v in enumerate(vec):
        out[i][flags_nonzero[i]] = v
    return out
