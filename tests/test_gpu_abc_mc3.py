"""ABC-SMC importance weights and MC3 chain swaps on the device against the
oracle's restatements of ABCSMC::calculate_particle_weight / MVN::probability
and MC3::swap_states."""
import numpy as np
import pytest

from beep_amd import _lib, configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pd(rng, n):
    Q = rng.normal(0, 1, (n, n))
    return Q @ Q.T / n + np.eye(n)


@pytest.mark.parametrize("n_prev,n_new,nvar", [(1, 5, 1), (37, 130, 5), (300, 200, 22), (200, 129, 32)])
def test_abcsmc_weights_match_oracle(gpu_lib, n_prev, n_new, nvar):
    rng = np.random.default_rng(n_prev + nvar)
    prev = rng.normal(0, 1, (n_prev, nvar))
    w = rng.uniform(0.2, 3.0, n_prev)
    theta = prev[rng.integers(0, n_prev, n_new)] + rng.normal(0, 0.3, (n_new, nvar))
    inv_M = np.linalg.inv(_pd(rng, nvar))
    size = 0.8
    logprior = rng.normal(-3, 1, n_new)
    got = _lib.abcsmc_weights(prev, w, theta, inv_M, size, logprior)
    want = np.array([O.abcsmc_weight(theta[i], prev, w, inv_M, size, logprior[i]) for i in range(n_new)])
    assert np.all(np.isfinite(got))
    assert np.allclose(got, want, rtol=1e-11, atol=0), np.abs(got / want - 1).max()


def test_abcsmc_weights_generation_size(gpu_lib):
    """A 4096-particle generation with 22 parameters (16.8 M kernel densities),
    checked on a sample of proposals."""
    rng = np.random.default_rng(11)
    n, nvar = 4096, 22
    prev = rng.normal(0, 1, (n, nvar))
    w = rng.uniform(0.5, 1.5, n)
    theta = prev + rng.normal(0, 0.2, (n, nvar))
    inv_M = np.linalg.inv(_pd(rng, nvar))
    logprior = np.zeros(n)
    got = _lib.abcsmc_weights(prev, w, theta, inv_M, 1.0, logprior)
    for i in (0, 1, 777, 4095):
        want = O.abcsmc_weight(theta[i], prev, w, inv_M, 1.0, 0.0)
        assert abs(got[i] / want - 1) < 1e-11


def _mc3_uniforms(nd, seed, stream):
    ctr = np.zeros((nd, 4), dtype=np.uint32)
    ctr[:, 0] = 0xFFFD | (stream << 16)
    ctr[:, 1] = np.arange(nd)
    ctr[:, 3] = 0x7FFFFFFD
    x = _lib.debug_philox(ctr, seed & 0xFFFFFFFF, seed >> 32)[:, 0]
    return (x.astype(np.float64) + 0.5) * 2.0 ** -32


def test_mc3_swap_matches_oracle(gpu_lib):
    """Chains of three runs with tempered EF values: the device's swap
    sequence, on the same uniforms, matches the oracle's, and relabelling
    by the returned order carries each EF with its state."""
    spec, truth = configs.seir_single()
    nrun, per = 3, 5
    n = nrun * per
    ctx = _lib.Context(spec, capacity=n, nslot=n)
    rng = np.random.default_rng(2)
    params = np.repeat(truth[None, :], n, 0)
    params[:, 0] = rng.uniform(1.6, 2.6, n)
    ctx.set_param(range(n), params)
    inst = rng.permutation(n).astype(np.int32)
    ctx.simulate(range(n), range(n), 40 + np.arange(n), seed=5, stream=2)
    EF_inst = -2 * ctx.obs_loglik(range(n))          # also stores each instance's EF on the device
    run = np.repeat(np.arange(nrun), per).astype(np.int32)
    invT = np.tile(np.linspace(1.0, 0.02, per), nrun)
    EF = EF_inst[inst]
    seed, stream = 77, 9
    order, ntr, nac = ctx.mc3_swap(inst, run, invT, seed=seed, stream=stream)
    nd = 5 * (n - nrun)
    o_order, o_ntr, o_nac, o_EF = O.mc3_swap(list(run), list(invT), list(EF), list(_mc3_uniforms(nd, seed, stream)))
    assert list(order) == o_order and list(ntr) == o_ntr and list(nac) == o_nac
    assert np.allclose(EF[order], o_EF, rtol=0, atol=0)
    assert sorted(order) == list(range(n)) and nac.sum() > 0
    for r in range(nrun):            # swaps stay within a run
        assert set(order[r * per:(r + 1) * per]) == set(range(r * per, (r + 1) * per))
    ctx.close()
