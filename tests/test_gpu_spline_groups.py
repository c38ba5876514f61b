"""R(t) per area and the external force per area on the device: a trajectory
replayed under two R groups and two external-force groups (each with its own
spline, the second with its own age weights) against the FP64 oracle, on the
specialised fused and step-wise kernels; groups whose values coincide leave
trajectories bit for bit as without groups; the generic kernels refuse
grouped models."""
import numpy as np
import pytest

from beep_amd import _lib, configs
from oracle import oracle as O
from test_modifications import mod_spec
from test_spline_groups import _add_param, _const_spline

pytestmark = pytest.mark.gpu


def _grouped():
    spec, truth = mod_spec()
    spec = dict(spec)
    pR2, truth = _add_param(spec, truth, "R2", 1.5)
    sR2 = _const_spline(spec, pR2)
    pe2, truth = _add_param(spec, truth, "efoi2", 6000.0)
    se2 = _const_spline(spec, pe2)
    A = spec["A"]
    spec["R_splines"] = [{"spline": 0, "areas": list(range(6))}, {"spline": sR2, "areas": list(range(6, A))}]
    spec["efoi_splines"] = [
        {"spline": spec["efoi_spline"], "areas": list(range(0, A, 2)), "agedist": spec["efoi_agedist"]},
        {"spline": se2, "areas": list(range(1, A, 2)), "agedist": [0.4, 0.3, 0.2, 0.1]}]
    return spec, truth


@pytest.mark.parametrize("keep", [_lib.KEEP_ALL, _lib.KEEP_ALL | _lib.STEPWISE], ids=["fused", "stepwise"])
def test_grouped_replay_matches_oracle(gpu_lib, keep):
    spec, truth = _grouped()
    orc = O.Oracle(spec)
    st = orc.simulate(orc.tables(truth), seed=9)
    ctx = _lib.Context(spec, capacity=1, nslot=1, keep=keep)
    ctx.set_param([0], truth[None, :])
    ctx.put_transnum(0, 0, st.transnum.astype(np.int32))
    assert np.array_equal(ctx.get_pop(0).astype(np.float64), st.pop)
    tm = ctx.get_transmean(0).astype(np.float64)
    ref = st.transmean
    scale = np.maximum(np.abs(ref), np.abs(ref).max(axis=(1, 3), keepdims=True) * 1e-3)
    assert (np.abs(tm - ref) <= 2e-5 * scale + 1e-9).all(), np.abs(tm - ref).max()
    # The groups did act: without them the infection means differ.
    plain = O.Oracle(mod_spec()[0])
    base = plain.replay(plain.tables(mod_spec()[1]), st.transnum).transmean
    assert not np.allclose(base[:, 7, 0], ref[:, 7, 0])
    ctx.close()


def test_coinciding_groups_change_nothing(gpu_lib):
    """One age group (the NGM ratio then does not depend on a group's
    population weights): two R groups on the same spline and two external-
    force groups on the same spline and weights give every area the values
    of the ungrouped model, and the same trajectories, through the grouped
    kernel code."""
    spec, truth = configs.seir_areas(12, 1, 40, seed=21)
    spec = dict(spec)
    pe, truth = _add_param(spec, truth, "efoi", 2000.0)
    se = _const_spline(spec, pe)
    spec.update(efoi_spline=se, efoi_agedist=[1.0])
    A = spec["A"]
    g = dict(spec, R_splines=[{"spline": 0, "areas": list(range(0, A, 3))},
                              {"spline": 0, "areas": [c for c in range(A) if c % 3]}],
             efoi_splines=[{"spline": se, "areas": list(range(5)), "agedist": [1.0]},
                           {"spline": se, "areas": list(range(5, A)), "agedist": [1.0]}])
    out = []
    for s in (spec, g):
        ctx = _lib.Context(s, capacity=4, nslot=1)
        ctx.set_param([0], truth[None, :])
        ctx.simulate(range(4), [0] * 4, [21, 22, 23, 24], seed=8, stream=3)
        out.append([ctx.get_transnum(i) for i in range(4)])
        ctx.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_generic_kernel_refuses_groups(gpu_lib):
    spec, truth = _grouped()
    with pytest.raises(_lib.BeepError, match="specialised kernels only"):
        _lib.Context(spec, capacity=1, nslot=1, keep=_lib.GENERIC_KERNEL)
