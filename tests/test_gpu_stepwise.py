"""The step-wise path (populations in HBM, per-step contraction + transition
kernels in a hipGraph): MFMA accumulator layout, agreement with the fused
kernel, and parity of the dense (MFMA) contraction with the FP64 oracle."""
import numpy as np
import pytest
from scipy import stats

import toml_models
from beep_amd import _lib, configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,K", [(64, 64), (200, 128), (9, 320), (1000, 320)])
def test_mfma_gemm_matches_numpy(gpu_lib, rows, K):
    rng = np.random.default_rng(rows + K)
    A = rng.standard_normal((rows, K)).astype(np.float32)
    B = rng.standard_normal((K, K)).astype(np.float32)
    C = _lib.debug_gemm(A, B)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    scale = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)
    assert np.all(np.abs(C - ref) <= 1e-5 * scale + 1e-6)


def _spatial_high_R():
    spec, truth = configs.spatial_abcsmc()
    th = truth.copy()
    th[:10] = 2.0
    return spec, th


@pytest.mark.parametrize("make", [_spatial_high_R, toml_models.toml_sird_ages], ids=["spatial_100x4", "sird_ages_tv"])
def test_stepwise_sparse_equals_fused(gpu_lib, make):
    """Same keys and seeds: the step-wise (sparse) path reproduces the fused
    kernel's trajectories (same arithmetic order, same Philox streams), also
    with an age matrix that varies by step."""
    spec, th = make()
    a = _lib.Context(spec, capacity=4, nslot=1)
    b = _lib.Context(spec, capacity=4, nslot=1, keep=_lib.KEEP_ALL | _lib.STEPWISE)
    for ctx in (a, b):
        ctx.set_param([0], th[None, :])
        ctx.simulate([0, 1, 2, 3], [0, 0, 0, 0], [5, 6, 7, 8], seed=31)
    same = 0
    for i in range(4):
        ta, tb = a.get_transnum(i), b.get_transnum(i)
        same += np.array_equal(ta, tb)
        # populations conserved on the step-wise path too
        pb = b.get_pop(i)
        assert (pb.sum(axis=2) == pb[0].sum(axis=1)[None]).all()
    assert same == 4
    # MBP on the step-wise path: unchanged parameters reproduce the trajectory
    b.mbp([0], [1], [0], [5], seed=32)
    assert np.array_equal(b.get_transnum(0), b.get_transnum(1))
    a.close()
    b.close()


def _small_dense(A=64):
    return configs.dense_pmcmc(A=A, period=40, with_obs=True)


@pytest.mark.parametrize("A", [64, 100], ids=["A64", "A100_padded_to_128"])
def test_dense_replay_matches_oracle(gpu_lib, A):
    """A replay on the dense (MFMA) path against the FP64 oracle; at A = 100
    the GEMM's K and N are padded to 128, the padding rows and columns of the
    operands zero."""
    spec, truth = _small_dense(A)
    orc = O.Oracle(spec)
    th = truth.copy()
    th[0] = 2.5
    tb = orc.tables(th)
    st = orc.simulate(tb, seed=17)
    ctx = _lib.Context(spec, capacity=1, nslot=1)
    ctx.set_param([0], th[None, :])
    ctx.put_transnum(0, 0, st.transnum.astype(np.int32))
    assert np.array_equal(ctx.get_pop(0).astype(np.float64), st.pop)
    tm = ctx.get_transmean(0).astype(np.float64)
    big = st.transmean > 1e-6
    rel = np.abs(tm[big] / st.transmean[big] - 1)
    assert np.quantile(rel, 0.999) < 1e-4 and np.median(rel) < 1e-6
    L = ctx.obs_loglik([0])[0]
    Lr, _ = orc.obs_loglik(st)
    assert abs(L / Lr - 1) < 1e-6
    ctx.close()


def test_dense_distribution_matches_oracle(gpu_lib):
    spec, truth = _small_dense()
    th = truth.copy()
    th[0] = 2.5
    n = 1500
    orc = O.Oracle(spec)
    tb = orc.tables(th)
    ref = np.array([orc.simulate(tb, seed=7000 + r).pop[-1].sum(axis=(0, 2)) for r in range(n)])
    ctx = _lib.Context(spec, capacity=n, nslot=1)
    ctx.set_param([0], th[None, :])
    ctx.simulate(np.arange(n), np.zeros(n), np.arange(n) + 1, seed=5)
    gpu = np.array([ctx.get_state(r).sum(axis=(0, 2)) for r in range(n)])
    for c in (0, 4, 8):
        d, p = stats.ks_2samp(gpu[:, c], ref[:, c])
        assert p > 1e-4, (c, d, p)
    ctx.close()


def test_pmcmc_sections_continue_state(gpu_lib):
    """simulate(ti, tf) in sections continues from the state: the same keys
    give the same trajectory as one full-horizon call."""
    spec, truth = configs.seir_single()
    ctx = _lib.Context(spec, capacity=2, nslot=1)
    ctx.set_param([0], truth[None, :])
    ctx.simulate([0], [0], [9], seed=4)
    for t0, t1 in ((0, 30), (30, 55), (55, 100)):
        ctx.simulate([1], [0], [9], t0, t1, seed=4)
    assert np.array_equal(ctx.get_transnum(0), ctx.get_transnum(1))
    ctx.close()
