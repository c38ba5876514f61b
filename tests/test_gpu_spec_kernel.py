"""The fused kernel specialised at context creation for the model's
compartment structure gives trajectories bit-identical to the generic fused
kernel -- simulation, coupled proposals (with and without simulate-afresh
masks), replay and mixed batches -- on models with 1, 4 and 9 age groups."""
import numpy as np
import pytest

import toml_models
from beep_amd import _lib, configs, spec as specmod

pytestmark = pytest.mark.gpu


def _pair(spec, cap, keep=_lib.KEEP_ALL):
    a = _lib.Context(spec, capacity=cap, nslot=cap, keep=keep)
    b = _lib.Context(spec, capacity=cap, nslot=cap, keep=keep | _lib.GENERIC_KERNEL)
    assert a.path == _lib.PATH_FUSED_SPEC and b.path == _lib.PATH_FUSED
    return a, b


def _same(a, b, insts, pop=True):
    for i in insts:
        assert np.array_equal(a.get_transnum(i), b.get_transnum(i)), i
        assert np.array_equal(a.get_transmean(i), b.get_transmean(i)), i
        if pop:
            assert np.array_equal(a.get_pop(i), b.get_pop(i)), i
        assert np.array_equal(a.get_state(i), b.get_state(i)), i


MODELS = [
    ("seir_single", lambda: configs.seir_single()),
    ("spatial_100x4", lambda: configs.spatial_abcsmc()),
    ("seir_areas_30x9", lambda: configs.seir_areas(30, 9, 80, seed=3)),
    ("seir_areas_50x1", lambda: configs.seir_areas(50, 1, 120, seed=4)),
    ("toml_sird_ages_tv_matrix", toml_models.toml_sird_ages),
]


@pytest.mark.parametrize("name,make", MODELS)
def test_simulate_mbp_replay_identical(gpu_lib, name, make):
    spec, truth = make()
    n = 8
    a, b = _pair(spec, 2 * n)
    rng = np.random.default_rng(5)
    params = specmod.sample_prior(spec, rng, n)
    params[0] = truth
    inst = np.arange(n)
    for ctx in (a, b):
        ctx.set_param(inst, params)
        ctx.simulate(inst, inst, 100 + inst, seed=11, stream=3)
    _same(a, b, inst)
    # sections continue from the state identically
    for ctx in (a, b):
        S = ctx.S
        ctx.simulate(inst, inst, 300 + inst, 0, S // 3, seed=12, stream=4)
        ctx.simulate(inst, inst, 300 + inst, S // 3, S, seed=12, stream=4)
    _same(a, b, inst)
    # coupled proposals from instance i to n + i at new parameters, half with masks
    params2 = specmod.sample_prior(spec, rng, n)
    mask = (rng.uniform(size=(n, a.S, a.A)) < 0.2).astype(np.uint8)
    for ctx in (a, b):
        ctx.set_param(n + inst, params2)
        ctx.mbp(inst[: n // 2], n + inst[: n // 2], n + inst[: n // 2], 500 + inst[: n // 2], seed=13, stream=5)
        ctx.mbp(inst[n // 2:], n + inst[n // 2:], n + inst[n // 2:], 500 + inst[n // 2:], simu_mask=mask[n // 2:],
                seed=13, stream=5)
    _same(a, b, n + inst)
    # replay of a stored trajectory
    tn = a.get_transnum(0)
    for ctx in (a, b):
        ctx.put_transnum(1, 0, tn)
    _same(a, b, [1])
    a.close()
    b.close()


def test_batch_identical(gpu_lib):
    """Mixed batches of simulations and proposals over two structures run
    through the specialised kernels and match the generic ones."""
    specs = [configs.seir_areas(20, 4, 60, seed=7), configs.spatial_abcsmc(A=60, period=80, seed=9)]
    ctxs = []
    for spec, truth in specs:
        a, b = _pair(spec, 8)
        for ctx in (a, b):
            ctx.set_param(range(8), np.repeat(truth[None, :], 8, 0))
        ctxs.append((a, b))
    for side in (0, 1):
        items = []
        for k, pair in enumerate(ctxs):
            ctx = pair[side]
            for i in range(4):
                items.append((ctx, _lib.WORK_SIMULATE, i, 0, i, 1000 * k + i, 0, ctx.S))
        _lib.run_batch(items, seed=21, stream=2)
        items = []
        for k, pair in enumerate(ctxs):
            ctx = pair[side]
            for i in range(4):
                items.append((ctx, _lib.WORK_MBP, 4 + i, i, 4 + i, 2000 * k + i, 0, ctx.S))
        _lib.run_batch(items, seed=22, stream=3)
    for a, b in ctxs:
        _same(a, b, range(8))
        a.close()
        b.close()


STEP_MODELS = [
    ("spatial_100x4_sparse", lambda: configs.spatial_abcsmc(A=100, period=60)),
    ("seir_areas_30x9", lambda: configs.seir_areas(30, 9, 60, seed=3)),
    ("dense_40x9", lambda: configs.dense_pmcmc(A=40, period=40, with_obs=False)),
    ("toml_sird_ages_tv_matrix", toml_models.toml_sird_ages),
]


@pytest.mark.parametrize("name,make", STEP_MODELS)
def test_stepwise_spec_identical(gpu_lib, name, make):
    """The step-wise path's specialised transition kernel (inside the captured
    hipGraph) matches the generic one: simulation, proposals and replay."""
    spec, truth = make()
    n = 6
    keep = _lib.KEEP_ALL | _lib.STEPWISE
    a = _lib.Context(spec, capacity=2 * n, nslot=2 * n, keep=keep)
    b = _lib.Context(spec, capacity=2 * n, nslot=2 * n, keep=keep | _lib.GENERIC_KERNEL)
    assert a.path == _lib.PATH_STEPWISE_SPEC and b.path == _lib.PATH_STEPWISE
    rng = np.random.default_rng(9)
    params = specmod.sample_prior(spec, rng, n)
    params[0] = truth
    inst = np.arange(n)
    for ctx in (a, b):
        ctx.set_param(inst, params)
        ctx.simulate(inst, inst, 40 + inst, seed=3, stream=6)
        ctx.simulate(inst, inst, 40 + inst, seed=3, stream=7)    # a second replay of the captured graph
    _same(a, b, inst)
    params2 = specmod.sample_prior(spec, rng, n)
    for ctx in (a, b):
        ctx.set_param(n + inst, params2)
        ctx.mbp(inst, n + inst, n + inst, 90 + inst, seed=4, stream=8)
    _same(a, b, n + inst)
    tn = a.get_transnum(0)
    for ctx in (a, b):
        ctx.put_transnum(1, 0, tn)
    _same(a, b, [1])
    a.close()
    b.close()


def test_large_instance_identical(gpu_lib):
    """An instance whose specialised kernel needs more than 64 KB of LDS
    (400 areas x 4 age groups) runs specialised and matches the generic
    kernel."""
    spec, truth = configs.seir_areas(400, 4, 40, seed=12)
    a = _lib.Context(spec, capacity=3, nslot=3, keep=_lib.KEEP_ALL)
    b = _lib.Context(spec, capacity=3, nslot=3, keep=_lib.KEEP_ALL | _lib.GENERIC_KERNEL)
    assert b.path == _lib.PATH_FUSED
    for ctx in (a, b):
        ctx.set_param(range(3), np.repeat(truth[None, :], 3, 0))
        ctx.simulate(range(3), range(3), [7, 8, 9], seed=2, stream=1)
    assert a.path == _lib.PATH_FUSED_SPEC
    _same(a, b, range(3))
    a.close()
    b.close()
