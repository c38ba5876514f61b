"""R(t) per area and the external force per strain and area (R_spline and
efoi_spline entries naming an area or a strain: Model::update_R_ref /
update_efoi_ref, src/model.cc:384-480; beta per area, src/model.cc:1318-1347;
efoi per (strain, area), src/state.cc:410-424) in the oracle and the host
model, on the CPU.

Each property is pinned by another route to the same numbers: one group is
the model without groups; with one age group the NGM ratio does not depend on
the population weights, so two areas' betas stand in the ratio of their R
values; an R group's weights equal a susceptibility scaled by its areas'
population fractions;
and an external-force group with another spline equals the single spline
with efoi_fac (a prediction-mode factor, tested on its own) on its areas."""
import numpy as np
import pytest

from beep_amd import _lib, configs
from oracle import derive, oracle as O
from test_modifications import mod_spec


def _add_param(spec, truth, name, value):
    spec["params"] = spec["params"] + [{"name": name, "prior": ("fixed", value)}]
    return len(spec["params"]) - 1, np.append(truth, value)


def _const_spline(spec, p):
    spec["splines"] = spec["splines"] + [configs._spline([p, p], [0, spec["period"]])]
    return len(spec["splines"]) - 1


def _trajectory(spec, truth, seed=3):
    orc = O.Oracle(spec)
    return orc.simulate(orc.tables(truth), seed=seed).transnum


def _means(spec, truth, tn):
    orc = O.Oracle(spec)
    return orc.replay(orc.tables(truth), tn).transmean


def test_one_group_is_the_model_without_groups():
    spec, truth = mod_spec()
    A = spec["A"]
    g = dict(spec, R_splines=[{"spline": spec["R_spline"], "areas": list(range(A))}],
             efoi_splines=[{"spline": spec["efoi_spline"], "areas": list(range(A)), "agedist": spec["efoi_agedist"]}])
    a, b = _lib.spec_tables(spec, truth), _lib.spec_tables(g, truth)
    for k in ("beta_af", "efoi", "rate", "N"):
        assert np.array_equal(a[k], b[k]), k
    tn = _trajectory(spec, truth)
    assert np.allclose(_means(g, truth, tn), _means(spec, truth, tn), rtol=1e-14, atol=0)


def test_R_groups_one_age_group_stand_in_the_ratio_of_R():
    spec, truth = configs.seir_areas(10, 1, 30, seed=5)
    spec = dict(spec)
    pR2, truth = _add_param(spec, truth, "R2", 1.4)
    s2 = _const_spline(spec, pR2)
    truth[0] = 2.6
    east = [0, 2, 4, 6, 8]
    west = [c for c in range(10) if c not in east]
    g = dict(spec, R_splines=[{"spline": 0, "areas": east}, {"spline": s2, "areas": west}])
    t = _lib.spec_tables(g, truth)
    d = derive.derive(g, derive.normalise(g), truth)
    S = g["period"] * g["spd"]
    assert np.allclose(t["beta_af"], d["beta"] * d["af"][np.arange(S) // g["spd"]], rtol=1e-12, atol=0)
    beta = t["beta_af"] / d["af"][np.arange(S) // g["spd"]]
    assert np.allclose(beta[:, west] / beta[:, [0]], 1.4 / 2.6, rtol=1e-13)
    assert np.allclose(beta[:, east], beta[:, [0]], rtol=0, atol=0)


def test_R_group_ratio_weights_its_own_areas():
    """Four age groups, areas with different age structures: each group's
    NGM ratio weights the dps by its own areas' populations
    (democatpos_dist).  The weights enter the next-generation matrix only as
    a product with the susceptibility, so a group's beta equals the
    ungrouped model's with susceptibility scaled by (group fraction / overall
    fraction) per dp."""
    spec, truth = configs.seir_areas(12, 4, 30, seed=9)
    spec = dict(spec)
    pR2, truth = _add_param(spec, truth, "R2", 1.7)
    s2 = _const_spline(spec, pR2)
    grp = [[0, 1, 2, 3, 4, 5, 6], [7, 8, 9, 10, 11]]
    # The second group's areas are older.
    p = np.asarray(spec["pop"], dtype=np.int64)
    p[grp[1], 0, :] = (p[grp[1], 0, :] * np.array([0.3, 0.6, 1.5, 3.0])).astype(np.int64)
    spec["pop"] = p.tolist()
    g = dict(spec, R_splines=[{"spline": 0, "areas": grp[0]}, {"spline": s2, "areas": grp[1]}])
    t = _lib.spec_tables(g, truth)
    S = g["period"] * g["spd"]
    d = derive.derive(g, derive.normalise(g), truth)
    beta = t["beta_af"] / d["af"][np.arange(S) // g["spd"]]
    pop = np.asarray(spec["pop"], dtype=np.float64).sum(axis=1)                # [A][G]
    every = pop.sum(axis=0) / pop.sum()
    betas = []
    for k, areas in enumerate(grp):
        frac = pop[areas].sum(axis=0) / pop[areas].sum()
        y, ty = dict(spec, R_spline=[0, s2][k]), truth.copy()
        sus = []
        for q in range(spec["G"]):
            iq, ty = _add_param(y, ty, f"w{q}", frac[q] / every[q])
            sus.append(iq)
        y["sus"] = sus
        dy = derive.derive(y, derive.normalise(y), ty)
        by = _lib.spec_tables(y, ty)["beta_af"] / dy["af"][np.arange(S) // y["spd"]]
        assert np.allclose(beta[:, areas], by[:, [0]], rtol=1e-12, atol=0)
        betas.append(by[0, 0] / [truth[0], 1.7][k])
    # The two groups' populations differ, so do their ratios (the test has teeth).
    assert not np.isclose(betas[0], betas[1], rtol=1e-2)


def test_efoi_group_equals_efoi_fac_on_its_areas():
    """R = 0: infections come from the external force alone.  Areas 7 and 8
    following a spline three times the other's equal the single spline with
    efoi_fac 3 on them."""
    spec, truth = mod_spec(R=0.0)
    spec = dict(spec)
    pe2, truth = _add_param(spec, truth, "efoi2", 6000.0)
    s2 = _const_spline(spec, pe2)
    A = spec["A"]
    rest = [c for c in range(A) if c not in (7, 8)]
    g = dict(spec, efoi_splines=[{"spline": spec["efoi_spline"], "areas": rest, "agedist": spec["efoi_agedist"]},
                                 {"spline": s2, "areas": [7, 8], "agedist": spec["efoi_agedist"]}])
    m = dict(spec, modifications=[{"type": "efoi_fac", "start": 0, "end": spec["period"], "factor": 3.0,
                                   "areas": [7, 8]}])
    tn = _trajectory(g, truth)
    assert np.allclose(_means(g, truth, tn), _means(m, truth, tn), rtol=1e-13, atol=0)
    # The host's table of each cell's external force against the oracle's.
    t = _lib.spec_tables(g, truth)
    d = derive.derive(g, derive.normalise(g), truth)
    assert np.allclose(t["efoi"], d["efoi_cell"], rtol=1e-14, atol=0)
    assert np.allclose(t["efoi"][:, 7] / t["efoi"][:, 0], 3.0, rtol=1e-14)


def test_efoi_group_age_weights():
    """A group's own age weights: cells take their dp's weight."""
    spec, truth = mod_spec(R=0.0)
    A, G = spec["A"], spec["G"]
    ad2 = [0.4, 0.3, 0.2, 0.1]
    g = dict(spec, efoi_splines=[{"spline": spec["efoi_spline"], "areas": list(range(6)), "agedist": spec["efoi_agedist"]},
                                 {"spline": spec["efoi_spline"], "areas": list(range(6, A)), "agedist": ad2}])
    t = _lib.spec_tables(g, truth)
    e = t["efoi"]
    assert np.allclose(e[:, 8, :] / e[:, 1, :], np.array(ad2) / np.array(spec["efoi_agedist"]), rtol=1e-14)


@pytest.mark.parametrize("bad, msg", [
    (lambda s: dict(s, R_splines=[{"spline": 0, "areas": [0, 1]}]), "no area|out of range"),
    (lambda s: dict(s, R_splines=[{"spline": 0, "areas": list(range(s["A"]))}, {"spline": 0, "areas": []}]),
     "no area"),
    (lambda s: dict(s, efoi_splines=[{"spline": 1, "areas": [0], "agedist": [0.25] * 4}]), "out of range"),
    (lambda s: dict(s, efoi_splines=[{"spline": 1, "areas": list(range(s["A"])), "agedist": [-1, 1, 1, 1]}]),
     "non-negative"),
])
def test_group_errors(bad, msg):
    spec, truth = mod_spec()
    with pytest.raises(_lib.BeepError, match=msg):
        _lib.spec_tables(bad(spec), truth)
