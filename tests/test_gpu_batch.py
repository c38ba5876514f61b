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


def test_autotune_leaves_continued_sections_untouched(gpu_lib, tmp_path):
    """A class with a job continuing from its state (t0 > 0) is not re-run by
    the autotuner: its result equals the single-context call."""
    os.environ["BEEP_TUNE_CACHE"] = str(tmp_path / "tune2.txt")
    out = []
    for batched in (False, True):
        spec, truth = configs.seir_areas(200, 9, 60, seed=5)     # a size class no other test tunes
        ctx = _lib.Context(spec, capacity=3, nslot=1)
        ctx.set_param([0], truth[None, :])
        ctx.simulate([0, 1], [0, 0], [7, 8], 0, 30, seed=5, stream=1)
        if batched:
            items = [(ctx, _lib.WORK_SIMULATE, 0, 0, 0, 7, 30, 60), (ctx, _lib.WORK_SIMULATE, 1, 0, 0, 8, 30, 60)]
            _lib.run_batch(items, seed=5, stream=2, autotune=True)
        else:
            ctx.simulate([0, 1], [0, 0], [7, 8], 30, 60, seed=5, stream=2)
        out.append([(ctx.get_transnum(i), ctx.get_state(i)) for i in (0, 1)])
        ctx.close()
    for (ta, sa), (tb, sb) in zip(*out):
        assert np.array_equal(ta, tb) and np.array_equal(sa, sb)


def test_written_instances_must_be_distinct(gpu_lib):
    spec, truth = configs.seir_areas(12, 1, 40, seed=6)
    ctx = _lib.Context(spec, capacity=4, nslot=1)
    ctx.set_param([0], truth[None, :])
    with pytest.raises(_lib.BeepError, match="written twice") as e:
        ctx.simulate([1, 1], [0, 0], [1, 2])
    assert e.value.code == -1
    ctx.simulate([0, 1], [0, 0], [1, 2])
    with pytest.raises(_lib.BeepError, match="both read and written"):
        ctx.mbp([0, 2], [2, 3], [0, 0], [3, 4])          # the second proposal's initial state is the first's proposal
    with pytest.raises(_lib.BeepError, match="written twice"):
        _lib.run_batch([(ctx, _lib.WORK_SIMULATE, 2, 0, 0, 5, 0, ctx.S), (ctx, _lib.WORK_MBP, 2, 0, 0, 6, 0, ctx.S)])
    with pytest.raises(_lib.BeepError, match="written by the batch"):
        _lib.run_batch([(ctx, _lib.WORK_SIMULATE, 0, 0, 0, 5, 0, ctx.S), (ctx, _lib.WORK_MBP, 2, 0, 0, 6, 0, ctx.S)])
    ctx.close()


def test_stepwise_contexts_run_outside_the_batch(gpu_lib):
    """beep_run_batch refuses step-wise contexts; batch.run sends their work
    through the context's own calls, with the single-call results."""
    from beep_amd import batch
    spec, truth = configs.seir_areas(40, 4, 50, seed=7)
    outs = []
    for mode in ("single", "run"):
        fused = _lib.Context(spec, capacity=2, nslot=1)
        step = _lib.Context(spec, capacity=3, nslot=1, keep=_lib.KEEP_ALL | _lib.STEPWISE)
        for c in (fused, step):
            c.set_param([0], truth[None, :])
        assert step.path in (_lib.PATH_STEPWISE, _lib.PATH_STEPWISE_SPEC)
        if mode == "single":
            fused.simulate([0], [0], [11], seed=3, stream=9)
            step.simulate([0, 1], [0, 0], [12, 13], seed=3, stream=9)
            step.mbp([0], [2], [0], [14], seed=3, stream=9)
        else:
            with pytest.raises(_lib.BeepError, match="step-wise context"):
                _lib.run_batch([(step, _lib.WORK_SIMULATE, 0, 0, 0, 12, 0, step.S)], seed=3, stream=9)
            batch.run([(fused, _lib.WORK_SIMULATE, 0, 0, 0, 11, 0, fused.S),
                       (step, _lib.WORK_SIMULATE, 0, 0, 0, 12, 0, step.S),
                       (step, _lib.WORK_SIMULATE, 1, 0, 0, 13, 0, step.S)], seed=3, stream=9)
            batch.run([(step, _lib.WORK_MBP, 2, 0, 0, 14, 0, step.S)], seed=3, stream=9)
        outs.append([fused.get_transnum(0)] + [step.get_transnum(i) for i in (0, 1, 2)])
        fused.close()
        step.close()
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
