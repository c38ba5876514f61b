"""Device Philox and samplers against the published generator's known answers
and against the Poisson / binomial laws (scipy.stats pmfs).  The sampler
draws are the simulation kernels' queued large draws (their own stream per
cell and transition), so the rejection samplers' retries run with the block
numbering the kernels use; a retry cut short there shows as excess mass at
floor(mean) in the chi-square tests."""
import numpy as np
import pytest
from scipy import stats

from beep_amd import _lib

pytestmark = pytest.mark.gpu

# Known-answer vectors of Philox-4x32-10 (Random123 kat_vectors).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_philox_kat(gpu_lib):
    for ctr, key, want in KAT:
        got = _lib.debug_philox(np.array([ctr], dtype=np.uint32), key[0], key[1])
        assert tuple(int(v) for v in got[0]) == want


def _chi2_pvalue(x, pmf, support_lo, support_hi):
    """Chi-square goodness of fit with bins merged until expected >= 5."""
    n = len(x)
    ks = np.arange(support_lo, support_hi + 1)
    exp = n * pmf(ks)
    obs = np.bincount(x - support_lo, minlength=len(ks))[:len(ks)].astype(float)
    # tails into the end bins
    exp[0] += n * (1 - exp.sum() / n) * 0  # mass outside is checked separately
    bins_o, bins_e, acc_o, acc_e = [], [], 0.0, 0.0
    for o, e in zip(obs, exp):
        acc_o += o
        acc_e += e
        if acc_e >= 5:
            bins_o.append(acc_o); bins_e.append(acc_e); acc_o = acc_e = 0.0
    if acc_e > 0 and bins_e:
        bins_o[-1] += acc_o; bins_e[-1] += acc_e
    bins_o, bins_e = np.array(bins_o), np.array(bins_e)
    bins_e *= bins_o.sum() / bins_e.sum()
    chi2 = ((bins_o - bins_e) ** 2 / bins_e).sum()
    return stats.chi2.sf(chi2, len(bins_o) - 1)


@pytest.mark.parametrize("lam", [1e-3, 0.3, 2.5, 9.99, 10.0, 37.5, 850.0, 1.2e5])
def test_poisson_law(gpu_lib, lam):
    n = 200_000
    x = _lib.debug_poisson(np.full(n, lam, dtype=np.float32), seed=1234567 + int(lam * 1000), key=7)
    lam32 = float(np.float32(lam))
    assert x.min() >= 0
    m, v = x.mean(), x.var()
    assert abs(m - lam32) < 5 * np.sqrt(lam32 / n) + 1e-12
    assert abs(v - lam32) < 6 * lam32 * np.sqrt(2.0 / n) + 6 * np.sqrt(lam32 / n) + 1e-9
    lo = max(0, int(lam32 - 8 * np.sqrt(lam32) - 10))
    hi = int(lam32 + 8 * np.sqrt(lam32) + 10)
    assert x.max() <= hi
    p = _chi2_pvalue(x, lambda k: stats.poisson.pmf(k, lam32), lo, hi)
    assert p > 1e-4, p


@pytest.mark.parametrize("nn,p", [(5, 0.3), (40, 0.1), (100, 0.5), (1000, 0.02), (5000, 0.7), (100000, 0.37),
                                  (7, 0.99), (300, 0.995), (20, 0.5)])
def test_binomial_law(gpu_lib, nn, p):
    n = 200_000
    x = _lib.debug_binomial(np.full(n, nn), np.full(n, p), seed=99 + nn, key=3)
    assert x.min() >= 0 and x.max() <= nn
    mu, var = nn * p, nn * p * (1 - p)
    assert abs(x.mean() - mu) < 5 * np.sqrt(var / n) + 1e-12
    lo = max(0, int(mu - 8 * np.sqrt(var) - 5))
    hi = min(nn, int(mu + 8 * np.sqrt(var) + 5))
    pv = _chi2_pvalue(x, lambda k: stats.binom.pmf(k, nn, p), lo, hi)
    assert pv > 1e-4, pv


def test_streams_independent_of_batch(gpu_lib):
    """A draw depends on its (seed, key, cell) counter only: the same cell gives
    the same value whether the batch holds it alone or among others."""
    lam = np.linspace(0.01, 200.0, 4096).astype(np.float32)
    full = _lib.debug_poisson(lam, seed=42, key=11)
    again = _lib.debug_poisson(lam, seed=42, key=11)
    assert np.array_equal(full, again)
    other = _lib.debug_poisson(lam, seed=42, key=12)
    assert not np.array_equal(full, other)
