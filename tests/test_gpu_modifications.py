"""Prediction-mode model modifications on the device: every kind at once
(trans_rate_fac on some areas' dps, beta_fac, efoi_fac, spline_fac on R(t),
spline_set on the external force) replayed against the FP64 oracle, both
kernel paths; and modifications whose factors are all one leave
trajectories bit for bit as without them."""
import numpy as np
import pytest

from beep_amd import _lib
from oracle import oracle as O
from test_modifications import mod_spec

pytestmark = pytest.mark.gpu

MODS = [
    {"type": "trans_rate_fac", "start": 5, "end": 20, "factor": 1.75, "trans": [2], "areas": [1, 3, 5], "dps": [0, 2]},
    {"type": "beta_fac", "start": 10, "end": 30, "factor": 0.4, "areas": [0, 1, 2, 3, 4, 5]},
    {"type": "efoi_fac", "start": 0, "end": 15, "factor": 3.0, "areas": [7, 8]},
    {"type": "spline_fac", "start": 25, "end": 40, "factor": 0.6, "spline": 0},
    {"type": "spline_set", "start": 20, "end": 25, "factor": 0.0, "spline": 1},
]


@pytest.mark.parametrize("keep", [_lib.KEEP_ALL, _lib.KEEP_ALL | _lib.GENERIC_KERNEL, _lib.KEEP_ALL | _lib.STEPWISE],
                         ids=["specialised", "generic", "stepwise"])
def test_modified_replay_matches_oracle(gpu_lib, keep):
    spec, truth = mod_spec()
    spec = dict(spec, modifications=MODS)
    orc = O.Oracle(spec)
    tb = orc.tables(truth)
    st = orc.simulate(tb, seed=5)
    ctx = _lib.Context(spec, capacity=1, nslot=1, keep=keep)
    ctx.set_param([0], truth[None, :])
    ctx.put_transnum(0, 0, st.transnum.astype(np.int32))
    assert np.array_equal(ctx.get_pop(0).astype(np.float64), st.pop)
    tm = ctx.get_transmean(0).astype(np.float64)
    ref = st.transmean
    # FP32 means: relative 2e-5, plus the rounding of the summands' scale
    # (populations up to 5e4 times factors up to 3).
    scale = np.maximum(np.abs(ref), np.abs(ref).max(axis=(1, 3), keepdims=True) * 1e-3)
    assert (np.abs(tm - ref) <= 2e-5 * scale + 1e-9).all(), np.abs(tm - ref).max()
    # The modifications did act: the replay without them gives other means.
    plain = O.Oracle(mod_spec()[0]).replay(O.Oracle(mod_spec()[0]).tables(truth), st.transnum)
    assert not np.allclose(plain.transmean[5:20, 1, 2, 0], ref[5:20, 1, 2, 0])
    ctx.close()


def test_unit_factors_change_nothing(gpu_lib):
    spec, truth = mod_spec()
    ones = [dict(m, factor=1.0) for m in MODS if m["type"] != "spline_set"]
    out = []
    for s in (spec, dict(spec, modifications=ones)):
        ctx = _lib.Context(s, capacity=4, nslot=1)
        ctx.set_param([0], truth[None, :])
        ctx.simulate(range(4), [0] * 4, [11, 12, 13, 14], seed=6, stream=2)
        out.append([ctx.get_transnum(i) for i in range(4)])
        ctx.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
