"""Mixed batches across contexts (uneven area counts, age groups and
horizons; simulations and MBP proposals in one launch) give the results of
the single-context calls with the same keys, whatever the autotuned
workgroup size."""
import os

import numpy as np
import pytest

from beep_amd import _lib, configs

pytestmark = pytest.mark.gpu


def _contexts():
    out = []
    for (A, G, period, seed) in ((12, 1, 60, 1), (50, 4, 80, 2), (23, 9, 45, 3)):
        spec, truth = configs.seir_areas(A, G, period, seed)
        ctx = _lib.Context(spec, capacity=6, nslot=2)
        th2 = truth.copy()
        th2[0] *= 1.3
        ctx.set_param([0, 1], np.stack([truth, th2]))
        out.append(ctx)
    return out


def _work(ctxs):
    items = []
    for ci, ctx in enumerate(ctxs):
        S = ctx.S
        items.append((ctx, _lib.WORK_SIMULATE, 0, 0, 0, 100 + ci, 0, S))          # full horizon
        items.append((ctx, _lib.WORK_SIMULATE, 1, 0, 1, 200 + ci, 0, S // 2))     # shorter horizon
    return items


def test_batch_equals_single_context_calls(gpu_lib, tmp_path):
    a, b = _contexts(), _contexts()
    # reference: single-context calls
    for ci, ctx in enumerate(a):
        ctx.simulate([0], [0], [100 + ci], 0, ctx.S, seed=9, stream=3)
        ctx.simulate([1], [1], [200 + ci], 0, ctx.S // 2, seed=9, stream=3)
    _lib.run_batch(_work(b), seed=9, stream=3)
    for ca, cb in zip(a, b):
        for i in (0, 1):
            assert np.array_equal(ca.get_transnum(i), cb.get_transnum(i))
    # MBP proposals mixed with fresh simulations in one batch
    for ci, ctx in enumerate(a):
        ctx.mbp([0], [2], [1], [300 + ci], seed=10, stream=4)
        ctx.simulate([3], [0], [400 + ci], 0, ctx.S, seed=10, stream=4)
    items = []
    for ci, ctx in enumerate(b):
        items.append((ctx, _lib.WORK_MBP, 2, 0, 1, 300 + ci, 0, ctx.S))
        items.append((ctx, _lib.WORK_SIMULATE, 3, 0, 0, 400 + ci, 0, ctx.S))
    os.environ["BEEP_TUNE_CACHE"] = str(tmp_path / "tune.txt")
    stats = _lib.run_batch(items, seed=10, stream=4, autotune=True)
    assert stats[1] >= 1
    for ca, cb in zip(a, b):
        for i in (2, 3):
            assert np.array_equal(ca.get_transnum(i), cb.get_transnum(i))
    assert (tmp_path / "tune.txt").exists()
    for ctx in a + b:
        ctx.close()
