"""GPU Poisson sampler and mock data.

The counts are pinned in distribution: chi-square goodness of fit against
the exact Poisson pmf (scipy.stats) on both sides of the inversion /
transformed-rejection switch at lam = 12, moments at large lam, edge cases,
seed reproducibility, index purity, and the checks of the reference's
test/test_mock_data.py:94-166 (integer, non-negative, keyed, seeded).
"""
import numpy as np
import pytest
import torch
from scipy import stats

pytestmark = pytest.mark.gpu

N_DRAW = 1 << 20


def _chi2_pvalue(counts, lam):
    kmax = int(counts.max()) + 1
    obs = np.bincount(counts, minlength=kmax + 1).astype(np.float64)
    ks = np.arange(kmax + 1)
    exp = stats.poisson.pmf(ks, lam) * counts.size
    exp[-1] += stats.poisson.sf(kmax, lam) * counts.size
    # merge sparse bins from both ends until every bin expects >= 10
    o, e = [], []
    acc_o = acc_e = 0.0
    for oi, ei in zip(obs, exp):
        acc_o += oi
        acc_e += ei
        if acc_e >= 10:
            o.append(acc_o)
            e.append(acc_e)
            acc_o = acc_e = 0.0
    if acc_e > 0:
        o[-1] += acc_o
        e[-1] += acc_e
    o, e = np.array(o), np.array(e)
    chi2 = float(np.sum((o - e) ** 2 / e))
    return float(stats.chi2.sf(chi2, len(o) - 1))


@pytest.mark.parametrize("lam", [0.05, 0.5, 3.0, 11.9, 12.0, 30.0, 250.0, 4000.0])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_poisson_chi_square(ja, lam, dtype):
    lt = torch.full((N_DRAW,), lam, dtype=dtype, device="cuda")
    c = ja.poisson_sample(lt, seed=12345).cpu().numpy().astype(np.int64)
    assert c.min() >= 0
    assert _chi2_pvalue(c, lam) > 1e-6


def test_poisson_moments_large_rate(ja):
    lam = 1e6
    c = ja.poisson_sample(torch.full((N_DRAW,), lam, dtype=torch.float64, device="cuda"),
                          seed=3).cpu().numpy().astype(np.float64)
    assert abs(c.mean() - lam) < 6 * np.sqrt(lam / N_DRAW)
    assert abs(c.var() / lam - 1) < 6 * np.sqrt(2.0 / N_DRAW)


def test_poisson_edge_cases_and_purity(ja):
    rng = np.random.default_rng(0)
    lam = np.concatenate([np.zeros(100), rng.uniform(0, 40, 100000), [1e-12, 11.999999, 12.0]])
    full = ja.poisson_sample(lam, seed=9)
    assert full.dtype == np.int32
    assert np.all(full[:100] == 0)
    # a draw depends only on (seed, index, lam): a prefix reproduces it
    part = ja.poisson_sample(lam[:5000], seed=9)
    assert np.array_equal(part, full[:5000])
    assert np.array_equal(ja.poisson_sample(lam, seed=9), full)
    assert not np.array_equal(ja.poisson_sample(lam, seed=10), full)
    # neighbouring draws are uncorrelated
    u = full[100:-3].astype(np.float64) - lam[100:-3]
    corr = np.corrcoef(u[:-1], u[1:])[0, 1]
    assert abs(corr) < 5 / np.sqrt(u.size)
    with pytest.raises(ValueError):
        ja.poisson_sample(np.array([1.0, -1.0]), seed=0)


def test_create_mock_counts(ja):
    rng = np.random.default_rng(4)
    lam = rng.uniform(0.1, 50.0, size=(2, 3, 24, 24))
    keep = rng.uniform(size=lam.shape) > 0.1
    d1 = ja.create_mock_counts(lam, keep, seed=42, keys=("tm_1", "tm_2"))
    d2 = ja.create_mock_counts(lam, keep, seed=42, keys=("tm_1", "tm_2"))
    d3 = ja.create_mock_counts(lam, keep, seed=43, keys=("tm_1", "tm_2"))
    assert set(d1.tree) == {"tm_1", "tm_2"}
    for i, k in enumerate(("tm_1", "tm_2")):
        arr = d1[k]
        assert np.issubdtype(arr.dtype, np.integer)
        assert (arr >= 0).all() and arr.size == keep[i].sum()
        assert np.array_equal(arr, d2[k])
    assert any(not np.array_equal(d1[k], d3[k]) for k in d1.tree)
    # the mean over pixels tracks lam
    allc = np.concatenate([d1[k] for k in d1.tree]).astype(np.float64)
    alll = np.concatenate([lam[i][keep[i]] for i in range(2)])
    assert abs(allc.sum() - alll.sum()) < 6 * np.sqrt(alll.sum())
