"""PMCMC particle filter on the device: the bootstrap against the oracle's
sequential restatement on the same uniforms, the single-particle case against
the plain simulate + log-likelihood, particle_sample against the oracle's
ancestry trace (with population continuity across the moved section starts),
and the law of the estimate against the oracle's particle filter."""
import numpy as np
import pytest
from scipy import stats

from beep_amd import _lib, configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SECTIONS = [(0, 25), (25, 50), (50, 75), (75, 100)]


def _device_uniforms(n, seed, stream, sec):
    ctr = np.zeros((n, 4), dtype=np.uint32)
    ctr[:, 0] = 0xFFFE | (stream << 16)
    ctr[:, 1] = np.arange(n)
    ctr[:, 2] = sec
    ctr[:, 3] = 0x7FFFFFFE
    x = _lib.debug_philox(ctr, seed & 0xFFFFFFFF, seed >> 32)[:, 0]
    return (x.astype(np.float64) + 0.5) * 2.0 ** -32


@pytest.mark.parametrize("n", [1, 7, 1000, 1024, 1025, 4096, 5000])
@pytest.mark.parametrize("invT", [1.0, 0.3])
def test_bootstrap_matches_oracle(gpu_lib, n, invT):
    rng = np.random.default_rng(n)
    seed, stream, sec = 77 + n, 5, 3
    for spread in (1.0, 30.0):                     # mild and very uneven weights
        L = -50.0 + spread * rng.standard_normal(n)
        bp, ob = _lib.bootstrap(L, invT, seed, stream, sec)
        bpo, obo = O.bootstrap(L, invT, _device_uniforms(n, seed, stream, sec))
        assert np.array_equal(bp, bpo)
        assert abs(ob - obo) <= 1e-12 * max(1.0, abs(obo))
        # every ancestor kept its own place
        assert (bp[bp] == bp).all()


def test_bootstrap_rejects_nan(gpu_lib):
    L = np.zeros(8)
    L[3] = np.nan
    with pytest.raises(_lib.BeepError):
        _lib.bootstrap(L, 1.0, 1, 0, 0)


def _spec():
    spec, truth = configs.seir_single()
    spec = dict(spec)
    spec["sections"] = SECTIONS
    return spec, truth


def test_single_particle_is_the_plain_likelihood(gpu_lib):
    """N = 1: no resampling, the estimate is invT times the log-likelihood of
    one trajectory, and that trajectory is the full-horizon simulation with
    the same key."""
    spec, truth = _spec()
    ctx = _lib.Context(spec, capacity=2, nslot=1)
    ctx.set_param([0], truth[None, :])
    ob, bp = ctx.pmcmc_obs_prob([0], 0, [41], invT=0.7, seed=9, stream=12)
    assert (bp == 0).all()
    ctx.simulate([1], [0], [41], seed=9, stream=12)
    L = ctx.obs_loglik([1])[0]
    assert abs(ob - 0.7 * L) <= 1e-12 * abs(L)
    assert np.array_equal(ctx.get_transnum(0), ctx.get_transnum(1))
    ctx.close()


@pytest.mark.parametrize("stepwise", [False, True])
def test_particle_sample_and_moves(gpu_lib, stepwise):
    spec, truth = _spec()
    N = 64
    keep = _lib.KEEP_ALL | (_lib.STEPWISE if stepwise else 0)
    ctx = _lib.Context(spec, capacity=N + 1, nslot=1, keep=keep)
    ctx.set_param([0], truth[None, :])
    inst = np.arange(N)
    ob, bp = ctx.pmcmc_obs_prob(inst, 0, 1000 + inst, invT=1.0, seed=3, stream=1)
    assert np.isfinite(ob)
    assert bp.shape == (len(SECTIONS) + 1, N) and (bp[0] == inst).all()
    for k in range(1, len(SECTIONS) + 1):
        assert (bp[k][bp[k]] == bp[k]).all()
    tn = [ctx.get_transnum(p) for p in inst]
    pop = [ctx.get_pop(p) for p in inst]
    exp_tn = O.particle_sample(bp, SECTIONS, tn)
    exp_pop = O.particle_sample(bp, SECTIONS, pop)
    ctx.pmcmc_particle_sample(N)
    got_tn, got_pop = ctx.get_transnum(N), ctx.get_pop(N)
    assert np.array_equal(got_tn, exp_tn) and np.array_equal(got_pop, exp_pop)
    # The sampled trajectory is one continuous history: pop[s+1] follows from
    # pop[s] and the transitions of step s, across the section starts where
    # the states were moved from the ancestors.
    fr = [t["from"] for t in spec["trans"]]
    to = [t["to"] for t in spec["trans"]]
    P, Tn = got_pop.astype(np.int64), got_tn.astype(np.int64)
    for s in range(P.shape[0] - 1):
        nxt = P[s].copy()
        for tr in range(len(fr)):
            nxt[:, fr[tr], :] -= Tn[s, :, tr, :]
            nxt[:, to[tr], :] += Tn[s, :, tr, :]
        assert np.array_equal(nxt, P[s + 1]), s
    # particle_sample into particle 0 itself (as the reference does)
    ctx.pmcmc_particle_sample(0)
    assert np.array_equal(ctx.get_transnum(0), exp_tn)
    ctx.close()


def test_estimate_law_matches_oracle(gpu_lib):
    """The particle-filter estimate over independent runs has the oracle's law
    (two-sample KS), and the same mean within sampling error."""
    spec, truth = _spec()
    N, R = 16, 200
    ctx = _lib.Context(spec, capacity=N, nslot=1)
    ctx.set_param([0], truth[None, :])
    inst = np.arange(N)
    dev = [ctx.pmcmc_obs_prob(inst, 0, 5000 + inst, invT=1.0, seed=21, stream=r)[0] for r in range(R)]
    ctx.close()
    orc = O.Oracle(spec)
    tb = orc.tables(truth)
    rng = np.random.default_rng(8)
    ora = []
    for r in range(R):
        seeds = rng.integers(1, 2 ** 31, size=(len(SECTIONS), N))
        u = rng.uniform(size=(len(SECTIONS), N))
        ora.append(O.pmcmc_obs_prob(orc, tb, N, 1.0, seeds, u)[0])
    dev, ora = np.array(dev), np.array(ora)
    assert np.isfinite(dev).all() and np.isfinite(ora).all()
    assert stats.ks_2samp(dev, ora).pvalue > 1e-3
    se = np.sqrt(dev.var() / R + ora.var() / R)
    assert abs(dev.mean() - ora.mean()) < 5 * se
