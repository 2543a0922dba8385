"""Batched 128-point DFTs: FP32 FFT and split-BF16 matrix-core DFT against
numpy's FFT in float64 (accuracy of each, and the MFMA lane layout through
unit impulses)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_impulses_give_dft_columns(ja):
    d = ja.DFT128()
    x = np.zeros((256, 128), np.complex64)
    for i in range(256):
        x[i, i % 128] = 1.0 if i < 128 else 1j
    want = np.fft.fft(x.astype(np.complex128), axis=1)
    for method, tol in (("fft", 1e-6), ("mfma", 3e-5)):
        got = d(torch.as_tensor(x, device="cuda"), method).cpu().numpy()
        assert _rel(got, want) <= tol, method


def test_random_vectors_accuracy(ja):
    rng = np.random.default_rng(0)
    B = 4096
    x = (rng.normal(size=(B, 128)) + 1j * rng.normal(size=(B, 128))).astype(np.complex64)
    want = np.fft.fft(x.astype(np.complex128), axis=1)
    d = ja.DFT128()
    xt = torch.as_tensor(x, device="cuda")
    e_fft = _rel(d(xt, "fft").cpu().numpy(), want)
    e_mfma = _rel(d(xt, "mfma").cpu().numpy(), want)
    assert e_fft <= 1e-6
    # three BF16 products per FP32 product: about 16 mantissa bits
    assert e_mfma <= 3e-5
    assert e_mfma > e_fft


def test_batch_checks(ja):
    d = ja.DFT128()
    with pytest.raises(ValueError):
        d(torch.zeros((100, 128), dtype=torch.complex64, device="cuda"))
