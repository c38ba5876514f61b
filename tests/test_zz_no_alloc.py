"""beep_simulate_noalloc, the entry point that allocates no device memory (its
job descriptors go to the library's fixed per-device array), against
beep_simulate: the same instances' slots and keys give bit-identical
trajectories on two calls in a row, the second with fewer instances and a
shorter step range; more instances than the array holds, and step-wise
contexts, are refused."""
import numpy as np
import pytest

from beep_amd import _lib, configs, spec as specmod

pytestmark = pytest.mark.gpu


def _same(ctx, a, b):
    for i, j in zip(a, b):
        assert np.array_equal(ctx.get_transnum(int(i)), ctx.get_transnum(int(j))), (i, j)
        assert np.array_equal(ctx.get_pop(int(i)), ctx.get_pop(int(j))), (i, j)
        assert np.array_equal(ctx.get_state(int(i)), ctx.get_state(int(j))), (i, j)


def test_noalloc_equals_simulate(gpu_lib):
    spec, truth = configs.seir_areas(30, 4, 60, seed=5)
    rng = np.random.default_rng(2)
    n1, n2 = 300, 37
    ctx = _lib.Context(spec, capacity=2 * n1, nslot=2)
    ctx.set_param([0, 1], np.vstack([truth[None, :], specmod.sample_prior(spec, rng, 1)]))
    a = np.arange(n1)
    slots = a % 2
    keys = 1000 + a
    ctx.simulate(a, slots, keys, seed=3, stream=4)
    ctx.simulate_noalloc(n1 + a, slots, keys, seed=3, stream=4)
    _same(ctx, a, n1 + a)
    # Second call, smaller, over part of the horizon: the array still holds
    # the first call's descriptors past n2, which this call must not read.
    b = np.arange(n2)
    ctx.simulate(b, b % 2, 5000 + b, t0=0, t1=25, seed=7, stream=1)
    ctx.simulate_noalloc(n1 + b, b % 2, 5000 + b, t0=0, t1=25, seed=7, stream=1)
    _same(ctx, b, n1 + b)
    # Continued from step 25 on the same instances: state carried over.
    ctx.simulate(b, b % 2, 5000 + b, t0=25, t1=ctx.S, seed=7, stream=1)
    ctx.simulate_noalloc(n1 + b, b % 2, 5000 + b, t0=25, t1=ctx.S, seed=7, stream=1)
    _same(ctx, b, n1 + b)
    ctx.close()


def test_noalloc_refusals(gpu_lib):
    spec, truth = configs.seir_areas(12, 1, 20, seed=3)
    ctx = _lib.Context(spec, capacity=4, nslot=1)
    ctx.set_param([0], truth[None, :])
    big = np.arange(20000)
    with pytest.raises(_lib.BeepError, match="at most"):
        ctx.simulate_noalloc(big, np.zeros_like(big), big)
    ctx.close()
    step = _lib.Context(spec, capacity=2, nslot=1, keep=_lib.KEEP_ALL | _lib.STEPWISE)
    step.set_param([0], truth[None, :])
    with pytest.raises(_lib.BeepError, match="step-wise"):
        step.simulate_noalloc([0], [0], [1])
    step.close()
