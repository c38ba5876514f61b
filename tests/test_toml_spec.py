"""The TOML front end (beep_amd/toml_spec.py) against expectations worked by
hand from the reference's input rules: Erlang stage names and order
(src/inputs_comps.cc:22-139), default names and one parameter per name
(src/inputs.cc:1146-1326), spline knots (src/inputs.cc:1363-1474,872-910),
initial state (src/data.cc:419-457), observation windows and models
(src/inputs.cc:273-475, src/data.cc:1119-1353,1981-2060)."""
import glob
import math
import os

import numpy as np
import pytest

from beep_amd import spec as specmod
from beep_amd import toml_spec as T

HERE = os.path.dirname(__file__)
FIX = os.path.join(HERE, "golden", "toml")
REF_EXAMPLES = "/root/reference/examples"


def _fixture():
    return T.load(os.path.join(FIX, "seir_erlang.toml"), datadir=os.path.join(FIX, "data_seir_erlang"))


def test_fixture_structure():
    spec, values, meta = _fixture()
    # Erlang(2) E becomes stages E, E1; the internal E->E1 transition comes
    # first; E->I leaves the last stage, S->E enters the first.
    assert [c["name"] for c in spec["comps"]] == ["S", "E", "E1", "I", "R"]
    assert [(t["from"], t["to"], t["infection"]) for t in spec["trans"]] == [
        (1, 2, False), (0, 1, True), (2, 3, False), (3, 4, False)]
    assert [c["shape"] for c in spec["comps"]] == [1, 2, 2, 1, 1]
    names = [p["name"] for p in spec["params"]]
    assert names == ["E mean occupancy", "I mean occupancy", "I Infectivity", "R-spline", "efoi-spline"]
    assert spec["comps"][1]["mean"] == spec["comps"][2]["mean"] == [0]
    assert spec["comps"][3]["inf"] == 2 and spec["comps"][0]["inf"] == -1
    assert spec["params"][0]["prior"] == ("uniform", 1.0, 10.0)
    assert spec["params"][2]["prior"] == ("fixed", 1.0)          # value only -> Fixed(value)
    assert np.array_equal(values, [3.0, 4.0, 1.0, 2.2, 2.0])
    # One knot value for the default bp 'start|end': both knots share one parameter.
    assert spec["splines"][0] == {"t": [0.0, 70.0], "param": [3, 3], "multfac": [1.0, 1.0], "factor": -1}
    assert spec["efoi_spline"] == 1 and spec["R_spline"] == 0
    # Everyone starts in the infection transition's source compartment.
    assert spec["pop"] == [[[20000], [0], [0], [0], [0]]]
    assert spec["spd"] == 2 and spec["period"] == 70 and specmod.nsteps(spec) == 140
    assert sorted(meta["ignored"]) == ["description", "outputdir", "state_outputs"]
    assert meta["prediction"]["modification"][0]["type"] == "beta_fac"


def test_fixture_observations():
    spec, _, _ = _fixture()
    obs = spec["obs"]
    ei = [o for o in obs if o["datatable"] == 0]
    r = [o for o in obs if o["datatable"] == 1]
    # Weekly E->I counts: consecutive 7-day windows of 2 steps per day; the
    # transition is E1->I (index 2).
    assert [(o["sett_i"], o["sett_f"]) for o in ei] == [(14 * k, 14 * k + 14) for k in range(10)]
    assert all(o["tr"] == [2] and o["type"] == "normal_percent" and o["percent"] == 10.0 for o in ei)
    assert [o["y"] for o in ei] == [3, 9, 31, 88, 230, 512, 890, 1104, 921, 602]
    # Populations at the start of each listed day; the day-70 row lies past
    # the end of the analysis and is dropped.
    assert [(o["sett_i"], o["sett_f"]) for o in r] == [(14 * k, 14 * k + 1) for k in range(10)]
    assert all(o["co"] == [4] and o["type"] == "normal" and o["sd"] == 20.0 for o in r)
    assert {o["graph"] for o in ei} == {0} and {o["graph"] for o in r} == {1}


def _write(d, name, text):
    p = os.path.join(d, name)
    with open(p, "w") as f:
        f.write(text)
    return p


BASE = """
time_format = "year-month-day"
start = "2020-03-01"
end = "2020-03-29"
steps_per_unit_time = "1"
datadir = "unused"
comps = [{name="S"}, {name="I", dist="Exp", mean_value="5", mean_prior="Uniform(2,8)", inf_value="1"}, {name="H", mean_value="10"}, {name="R"}]
trans = [{from="S", to="I", infection="yes"}, {from="I", to="H", prob_value="0.2", prob_prior="Uniform(0,0.5)"},
         {from="I", to="R", prob_value="0.8", prob_prior="Uniform(0.5,1)"}, {from="H", to="R"}]
R_spline = [{bp="[bp:R_spline.csv]", value="[value:R_spline.csv]", prior="[prior:R_spline.csv]"}]
areas = "areas.csv"
"""


def _model(tmp_path, extra):
    d = str(tmp_path)
    _write(d, "areas.csv", "area,population\nX,1000.5\n")
    _write(d, "R_spline.csv", 'bp,value,prior\nstart,2.5,"Uniform(1,4)"\n2020-03-15,1.0,"Uniform(0.2,2)"\n'
                              '2020-03-29,1.2,"Uniform(0.2,2)"\n')
    _write(d, "I-H.csv", "date,Data,Data SD\n2020-03-01,1,2\n2020-03-08,.,2\n2020-03-15,7,3\n2020-03-22,*,4\n"
                         "2020-03-29,9,5\n")
    _write(d, "H.csv", "date,Data\n2020-02-28,4\n2020-03-02,5\n2020-03-28,6\n")
    p = _write(d, "m.toml", BASE + extra)
    return T.load(p, datadir=d)


def test_dates_tables_and_obsmodels(tmp_path):
    spec, values, _ = _model(tmp_path, """
data_tables = [{type="transition", observation="I->H", obsmodel="normal load", file="I-H.csv", threshold="3"},
               {type="population", observation="H", obsmodel="poisson", file="H.csv"}]
""")
    # Knots from the table column, times in days from start; more than two
    # knots: parameter names carry the knot time.
    sp = spec["splines"][0]
    assert sp["t"] == [0.0, 14.0, 28.0]
    names = [p["name"] for p in spec["params"]]
    assert [names[i] for i in sp["param"]] == ["R-spline t=0", "R-spline t=14", "R-spline t=28"]
    assert [spec["params"][i]["prior"] for i in sp["param"]] == [("uniform", 1.0, 4.0), ("uniform", 0.2, 2.0),
                                                                  ("uniform", 0.2, 2.0)]
    assert [values[i] for i in sp["param"]] == [2.5, 1.0, 1.2]
    assert spec["efoi_spline"] == -1
    # H has a mean but no dist: one stage.  Branch probabilities by default name.
    assert [c["shape"] for c in spec["comps"]] == [1, 1, 1, 1]
    assert "I->H branch prob." in names and "I->R branch prob." in names
    assert spec["pop"][0][0] == [1001]                               # (unsigned)(1000.5 + 0.5)
    ih = [o for o in spec["obs"] if o["datatable"] == 0]
    # Rows at days 0, 7, 14, 21, 28: day 28 is the end of the analysis, so
    # four 7-day windows; '.' is unknown, '*' below the threshold; 'normal
    # load' takes the SD column.
    assert [(o["sett_i"], o["sett_f"]) for o in ih] == [(0, 7), (7, 14), (14, 21), (21, 28)]
    assert [o["y"] for o in ih] == [1.0, T.OBS_UNKNOWN, 7.0, T.OBS_THRESH]
    assert [o["sd"] for o in ih] == [2.0, 2.0, 3.0, 4.0] and ih[3]["thresh"] == 3.0
    h = [o for o in spec["obs"] if o["datatable"] == 1]
    # The row before start is dropped; populations on days 1 and 27.
    assert [(o["sett_i"], o["y"], o["type"]) for o in h] == [(1, 5.0, "poisson"), (27, 6.0, "poisson")]


@pytest.mark.parametrize("extra,msg", [
    ('ages = [{cats="0-19|20+"}]\n', "'ages'"),
    ('data_tables = [{type="transition", observation="I->H", obsmodel="negbin 2", file="I-H.csv"}]\n',
     "not recognised"),
    ('data_tables = [{type="transition", observation="I->X", obsmodel="poisson", file="I-H.csv"}]\n',
     "not a compartment or a transition"),
    ('data_tables = [{type="transition", observation="I->H", obsmodel="poisson", file="I-H.csv", geo_dep="r"}]\n',
     "cannot find the column 'r' in the areas file"),
])
def test_rejections(tmp_path, extra, msg):
    with pytest.raises(T.TomlError, match=msg):
        _model(tmp_path, extra)


def test_observation_spline(tmp_path):
    """obs_spline named by a data table's factor_spline (src/model.cc:611-629,
    1429-1439): a spline whose knot parameters carry the spline's name and
    knot time, attached to every observation of that table; none on the
    others; refused on a population table (src/obsmodel.cc:100) or when the
    name is not found."""
    import toml_models
    spec, values = toml_models.toml_seird_branch()
    names = [p["name"] for p in spec["params"]]
    se = [o for o in spec["obs"] if o["datatable"] == 1]
    k = se[0]["spline"]
    sp = spec["splines"][k]
    assert sp["t"] == [0.0, 42.0, 84.0]
    assert [names[i] for i in sp["param"]] == ["ascertainment t=0", "ascertainment t=42", "ascertainment t=84"]
    assert [values[i] for i in sp["param"]] == [0.6, 0.9, 0.8]
    assert all(o["spline"] == k for o in se)
    assert all(o.get("spline", -1) == -1 for o in spec["obs"] if o["datatable"] == 0)
    text = open(os.path.join(FIX, "seird_branch.toml")).read()
    bad = text.replace('factor_spline="ascertainment"', 'factor_spline="detection"')
    with pytest.raises(T.TomlError, match="cannot find the observation spline 'detection'"):
        T.load(_write(str(tmp_path), "a.toml", bad), datadir=os.path.join(FIX, "data_seird_branch"))
    bad = text.replace('obsmodel="poisson", file="I-D.csv"',
                       'obsmodel="poisson", file="I-D.csv"}, {type="population", observation="D", '
                       'obsmodel="poisson", factor_spline="ascertainment", file="I-D.csv"')
    with pytest.raises(T.TomlError, match="population table"):
        T.load(_write(str(tmp_path), "b.toml", bad), datadir=os.path.join(FIX, "data_seird_branch"))


def test_prob_reach_recorded(tmp_path):
    """prob_reach is an output-only derived quantity (src/model.cc:1932-1957):
    recorded in the metadata, its compartment checked."""
    _, _, meta = _model(tmp_path, 'prob_reach = [{name="ihr", comp="H"}]\n')
    assert meta["derived"]["prob_reach"] == [{"name": "ihr", "comp": "H"}]
    with pytest.raises(T.TomlError, match="cannot find the compartment 'X'"):
        _model(tmp_path, 'prob_reach = [{name="ihr", comp="X"}]\n')


def test_multi_area():
    """Three areas: the raw matrix as diagonal + off-diagonal rows
    (src/data_matrix.cc:206-229), the mixing spline as the geo spline, one
    observation graph per data column, geo_dep by area and by region."""
    import toml_models
    spec, values = toml_models.toml_sir_3areas()
    assert spec["A"] == 3 and spec["G"] == 1
    assert spec["M"] == {"diag": [1.0, 1.5, 0.8], "rowptr": [0, 1, 3, 4], "to": [1, 0, 2, 1],
                         "val": [0.2, 0.2, 0.3, 0.3]}
    assert spec["pop"] == [[[30000], [0], [0]], [[80000], [0], [0]], [[45000], [0], [0]]]
    names = [p["name"] for p in spec["params"]]
    g = spec["splines"][spec["geo_spline"]]
    assert [names[i] for i in g["param"]] == ["geo-mixing-modify"] * 2 and values[g["param"][0]] == 0.6
    si = [o for o in spec["obs"] if o["datatable"] == 0]
    ir = [o for o in spec["obs"] if o["datatable"] == 1]
    assert len(si) == 3 * 8 and len(ir) == 2 * 8
    assert [(o["area"], o["graph"]) for o in si[::8]] == [([0], 0), ([1], 1), ([2], 2)]
    assert [(o["area"], o["graph"]) for o in ir[::8]] == [([0], 3), ([1, 2], 4)]
    # Covariates (src/data.cc:461-556): the areas file's 'Density' column
    # (headings match without case), log-transformed; area time-varying
    # values from rows at days 0, 20, 35, 56, each row holding its days up to
    # and including the next row's, the next row overwriting the shared day.
    dens, mob = spec["covars"]
    assert names[dens["param"]] == "Covariate: density" and values[dens["param"]] == 0.3
    assert names[mob["param"]] == "mobility effect" and values[mob["param"]] == -0.5
    assert np.array_equal(dens["value"], np.log([[0.5] * 56, [4.0] * 56, [1.5] * 56]))
    want = [[v0] * 20 + [v1] * 15 + [v2] * 21 for v0, v1, v2 in ((1.0, 0.6, 0.8), (1.2, 0.5, 0.9), (0.9, 0.7, 0.85))]
    assert np.array_equal(mob["value"], want)


def test_area_and_level_effects():
    """area_effect and level_effect of the three-area fixture, worked by hand:
    weights are the areas' population shares (30/155, 80/155, 45/155) and the
    levels' shares of the 3 x 56 area-days (closed: Centre and South on days
    21-41, South on days 42-55, 56 of 168); '*' is set so that the weighted
    values sum to one; values are returned times their weight
    (src/model.cc:990-1003, 2838-2878); MDir(k) gives gamma(beta * frac, 1)
    with beta = (n - 1) / k^2 - 1 (src/model.cc:3025-3050)."""
    import toml_models
    spec, values = toml_models.toml_sir_3areas()
    names = [p["name"] for p in spec["params"]]
    ae, le = spec["area_effect"], spec["level_effect"]
    assert [names[i] for i in ae] == ["Area effect North", "Area effect Centre", "Area effect South"]
    f = [30 / 155, 80 / 155, 45 / 155]
    star = (1 - 1.2 * f[0] - 0.9 * f[1]) / f[2]
    assert [values[i] for i in ae] == pytest.approx([1.2 * f[0], 0.9 * f[1], star * f[2]], rel=1e-14)
    beta = 2 / 0.09 - 1
    assert [spec["params"][i]["prior"] for i in ae] == [("gamma", pytest.approx(beta * x, rel=1e-14), 1.0) for x in f]
    assert [names[i] for i in le["param"]] == ["open", "closed"]
    want = [[0, 0, 0]] * 21 + [[0, 1, 1]] * 21 + [[0, 0, 1]] * 14
    assert le["map"] == want
    fl = [112 / 168, 56 / 168]
    assert [values[i] for i in le["param"]] == pytest.approx([(1 - 0.6 * fl[1]), 0.6 * fl[1]], rel=1e-14)
    beta = 1 / 0.16 - 1
    assert [spec["params"][i]["prior"] for i in le["param"]] == [("gamma", pytest.approx(beta * x, rel=1e-14), 1.0)
                                                                 for x in fl]
    groups = spec["dirichlet"][-2:]
    assert groups[0]["param"] == ae and groups[0]["frac"] == pytest.approx(f, rel=1e-15)
    assert groups[1]["param"] == le["param"] and groups[1]["frac"] == pytest.approx(fl, rel=1e-15)


@pytest.mark.parametrize("entry,msg", [
    ('area_effect = {value="1.1|0.9", prior="MDir(0.5)"}', "exactly one of the values must be '\\*'"),
    ('area_effect = {value="1.1|*", prior="MDir(0.5)|Fixed(1)"}', "all be MDir"),
    ('area_effect = {value="1.1|*", prior="MDir(3)"}', "too diffuse"),
    ('area_effect = {value="5|*", prior="MDir(0.5)"}', "sum to more than one"),
    ('level_effect = {file="lev.csv", param="a|b", value="*|2", prior="MDir(0.5)"}', "not contained in 'param'"),
])
def test_effect_rejections(tmp_path, entry, msg):
    d = str(tmp_path)
    _write(d, "areas.csv", "area,population\nX,100\nY,300\n")
    _write(d, "g.csv", "1,0.1\n0.1,1\n")
    _write(d, "lev.csv", "date,X,Y\n2020-03-01,a,b\n2020-03-10,a,c\n")
    text = BASE.replace('R_spline = [{bp="[bp:R_spline.csv]", value="[value:R_spline.csv]", '
                        'prior="[prior:R_spline.csv]"}]', 'R_spline = [{value="2", prior="Uniform(1,3)"}]')
    with pytest.raises(T.TomlError, match=msg):
        T.load(_write(d, "m.toml", text + 'geo_mixing_matrix = "g.csv"\n' + entry + "\n"), datadir=d)


def test_time_varying_covariate(tmp_path):
    """tv_covars with a 'Date' heading: the table's dates relative to start,
    one value for every area; a log transform needs positive values; a
    one-row table leaves the values zero, as the reference's loop over row
    pairs does not run."""
    d = str(tmp_path)
    p = _base_model(d, "m.toml", 'prob_value="0.8", prob_prior="Uniform(0.5,1)"',
                    'R_spline = [{value="2", prior="Uniform(1,3)"}]')
    _write(d, "tv.csv", "Date,temp\n2020-02-28,2\n2020-03-05,4\n2020-03-30,8\n")
    text = open(p).read() + 'tv_covars = [{name="temp", file="tv.csv", func="log", value="0.2", prior="Uniform(0,1)"}]\n'
    spec, values, _ = T.load(_write(d, "a.toml", text), datadir=d)
    (cv,) = spec["covars"]
    assert [p["name"] for p in spec["params"]][cv["param"]] == "TV Covariate: temp" and values[cv["param"]] == 0.2
    # Days -2..4 from the first row, 4..29 from the second, clipped to 0..27.
    assert np.array_equal(cv["value"], [np.log([2.0] * 4 + [4.0] * 24)])
    _write(d, "tv.csv", "Date,temp\n2020-02-28,2\n2020-03-05,-1\n2020-03-30,8\n")
    with pytest.raises(T.TomlError, match="positive"):
        T.load(_write(d, "b.toml", text), datadir=d)
    _write(d, "tv.csv", "Date,temp\n2020-03-02,2\n2020-03-30,8\n")
    with pytest.raises(T.TomlError, match="cover the analysis period"):
        T.load(_write(d, "c.toml", text), datadir=d)
    _write(d, "tv.csv", "Date,temp\n2020-03-05,3\n")
    spec, _, _ = T.load(_write(d, "d.toml", text.replace('func="log"', 'func="linear"')), datadir=d)
    assert np.array_equal(spec["covars"][0]["value"], [[0.0] * 28])
    with pytest.raises(T.TomlError, match="func=exp"):
        T.load(_write(d, "e.toml", text.replace('func="log"', 'func="exp"')), datadir=d)


def test_age_groups():
    """Three age groups: 'age:' lists give one parameter per group named
    '<age> <root>'; one Dirichlet group per age; '*' per age; the contact
    matrix raw; efoi weights age_dist / (sum * population share)
    (src/inputs.cc:1421-1442); per-age columns select dp."""
    import toml_models
    spec, values = toml_models.toml_sird_ages()
    names = [p["name"] for p in spec["params"]]
    assert spec["G"] == 3 and spec["N"] == [[2.0, 0.8, 0.3], [0.8, 1.2, 0.4], [0.3, 0.4, 0.9]]
    mean = spec["comps"][1]["mean"]
    assert [names[i] for i in mean] == ["child I mean occupancy", "adult I mean occupancy", "elderly I mean occupancy"]
    assert [values[i] for i in mean] == [4.0, 5.0, 7.0]
    pr, pd = spec["trans"][1]["prob"], spec["trans"][2]["prob"]
    assert [names[i] for i in pd] == ["child I->D branch prob.", "adult I->D branch prob.", "elderly I->D branch prob."]
    assert spec["dirichlet"][:3] == [{"param": [pr[g], pd[g]]} for g in range(3)]
    assert [values[i] for i in pd] == pytest.approx([0.001, 0.01, 0.1], abs=1e-15)
    assert spec["pop"] == [[[24000, 60000, 16000], [0, 0, 0], [0, 0, 0], [0, 0, 0]]]
    share = np.array([24000, 60000, 16000]) / 100000
    assert spec["efoi_agedist"] == pytest.approx(list(np.array([1, 2, 1]) / (4 * share)), rel=1e-15)
    si = [o for o in spec["obs"] if o["datatable"] == 0]
    assert [(o["dp"], o["graph"]) for o in si[::9]] == [([0], 0), ([1], 1), ([2], 2)]
    d = [o for o in spec["obs"] if o["datatable"] == 1]
    assert all(o["dp"] == [0, 1, 2] and o["co"] == [3] for o in d) and len(d) == 21


def test_age_matrix_modifications(tmp_path):
    """age_mixing_modify (row_column on 'elderly') and age_mixing_perturb (one
    entry per age group) of the age fixture, worked by hand: one spline each,
    knot parameters 'MatMod <agecat> factor t=<bp>' by default; at each knot
    the perturbations form a modified Dirichlet group weighted by the age
    shares (0.24, 0.6, 0.16) with gamma(beta * share, 1) priors,
    beta = 2 / 0.3^2 - 1; '*' is filled so the weighted values sum to one and
    values are returned times their weight (src/inputs.cc:944-1011,
    src/model.cc:2786-2825,2838-2878,3025-3050)."""
    import toml_models
    spec, values = toml_models.toml_sird_ages()
    names = [p["name"] for p in spec["params"]]
    mods = spec["age_mod"]
    assert [m["ages"] for m in mods] == [[2], [0], [1], [2]]
    rc = spec["splines"][mods[0]["spline"]]
    assert rc["t"] == [0.0, 28.0, 63.0] and [names[i] for i in rc["param"]] == ["shield 0", "shield 28", "shield 63"]
    assert [values[i] for i in rc["param"]] == [1.0, 0.5, 0.7]
    share = [0.24, 0.6, 0.16]
    beta = 2 / 0.09 - 1
    raw = [[0.24, 0.6, 0.16], [0.6 * 0.24, 1 - 0.6 * 0.24 - 1.2 * 0.16, 1.2 * 0.16], [0.8 * 0.24, 1 - 0.8 * 0.24 - 0.16, 0.16]]
    for k, age in enumerate(("child", "adult", "elderly")):
        sp = spec["splines"][mods[k + 1]["spline"]]
        assert sp["t"] == [0.0, 21.0, 63.0] and sp["smooth_type"] == 2 and sp["smooth"] == [0.3] * 3
        assert [names[i] for i in sp["param"]] == [f"MatMod {age} factor t={t}" for t in (0, 21, 63)]
        assert [values[i] for i in sp["param"]] == pytest.approx([raw[p][k] for p in range(3)], rel=1e-14)
        assert [spec["params"][i]["prior"] for i in sp["param"]] == [("gamma", pytest.approx(beta * share[k]), 1.0)] * 3
    for p in range(3):
        grp = spec["dirichlet"][3 + p]
        assert grp["param"] == [spec["splines"][m["spline"]]["param"][p] for m in mods[1:]]
        assert grp["frac"] == pytest.approx(share, rel=1e-15)
    # Every age group perturbed exactly once (src/model.cc:2796-2813).
    text = open(os.path.join(FIX, "sird_ages.toml")).read().replace('{agecat="adult"', '{agecat="child"')
    with pytest.raises(T.TomlError, match="exactly once"):
        T.load(_write(str(tmp_path), "m.toml", text), datadir=os.path.join(FIX, "data_sird_ages"))


def test_demographic_categories():
    """Two age groups and two sexes (democats): dp enumerates the
    combinations with age fastest (src/data.cc:55-73); populations are the
    age columns times each area's sex fractions, the columns divided by their
    sum (0.48/0.52 and 49/51), rounded (src/data.cc:419-457,821-870);
    'Sex:' and 'age,Sex:' lists name their parameters '<values> <root>'
    (src/inputs.cc:1217-1326); susceptibility by age and by sex are modified
    Dirichlet groups weighted by population shares, a dp's the product
    (src/model.cc:1144-1178,2768-2784); efoi age weights and data columns by
    sex or age select dps (src/data.cc:1491-1520)."""
    import toml_models
    spec, values = toml_models.toml_seir_sex_ages()
    names = [p["name"] for p in spec["params"]]
    assert spec["G"] == 4 and spec["nage"] == 2 and spec["dp_age"] == [0, 1, 0, 1]
    assert spec["N"] == [[1.5, 0.5], [0.5, 1.0]]
    north = [round(30000 * 0.48), round(10000 * 0.48), round(30000 * 0.52), round(10000 * 0.52)]
    south = [round(20000 * 0.49), round(25000 * 0.49), round(20000 * 0.51), round(25000 * 0.51)]
    assert [spec["pop"][0][0], spec["pop"][1][0]] == [north, south]
    assert [names[i] for i in spec["comps"][1]["mean"]] == ["M E mean occupancy"] * 2 + ["F E mean occupancy"] * 2
    assert [names[i] for i in spec["comps"][2]["mean"]] == [
        "young,M I mean occupancy", "old,M I mean occupancy", "young,F I mean occupancy", "old,F I mean occupancy"]
    assert [values[i] for i in spec["comps"][2]["mean"]] == [4.0, 6.0, 5.0, 7.0]
    pop = np.array(north) + np.array(south)
    age_share = [(pop[0] + pop[2]) / pop.sum(), (pop[1] + pop[3]) / pop.sum()]
    sex_share = [(pop[0] + pop[1]) / pop.sum(), (pop[2] + pop[3]) / pop.sum()]
    sy, so = names.index("sus young"), names.index("sus old")
    sm, sf = names.index("sus M"), names.index("sus F")
    assert spec["sus"] == [sy, so, sy, so] and spec["sus_extra"] == [[sm, sm, sf, sf]]
    assert spec["dirichlet"][0] == {"param": [sy, so], "frac": pytest.approx(age_share, rel=1e-15)}
    assert spec["dirichlet"][1] == {"param": [sm, sf], "frac": pytest.approx(sex_share, rel=1e-15)}
    assert values[sy] == pytest.approx(0.8 * age_share[0]) and values[so] == pytest.approx(1 - 0.8 * age_share[0])
    assert spec["params"][sm]["prior"] == ("gamma", pytest.approx((1 / 0.09 - 1) * sex_share[0]), 1.0)
    w = [1 / (4 * age_share[0]), 3 / (4 * age_share[1])]
    assert spec["efoi_agedist"] == pytest.approx([w[0], w[1], w[0], w[1]], rel=1e-14)
    se = [o for o in spec["obs"] if o["datatable"] == 0]
    r = [o for o in spec["obs"] if o["datatable"] == 1]
    assert {tuple(o["dp"]) for o in se[:8]} == {(0, 1)} and {tuple(o["dp"]) for o in se[8:]} == {(2, 3)}
    assert {tuple(o["dp"]) for o in r[:8]} == {(0, 2)} and {tuple(o["dp"]) for o in r[8:]} == {(1, 3)}
    # Marginal totals per sex over days 14-42, one graph (src/data.cc:1355-1455).
    mg = [o for o in spec["obs"] if o["datatable"] == 2]
    assert [(o["y"], o["dp"], o["sett_i"], o["sett_f"], o["tr"]) for o in mg] == [(2900.0, [0, 1], 14, 42, [1]),
                                                                                  (3100.0, [2, 3], 14, 42, [1])]
    assert len({o["graph"] for o in mg}) == 1
    # Infected fractions per area: the factor is one over the area's initial
    # population (src/data.cc:1226-1235).
    fr = [o for o in spec["obs"] if o["datatable"] == 3]
    assert [o["factor"] for o in fr[::8]] == [1 / sum(north), 1 / sum(south)]
    assert all(o["co"] == [2] and o["sett_f"] == o["sett_i"] + 1 for o in fr)


def test_strains():
    """strains (src/inputs.cc:662-690): the last category, dp = strain x 2 +
    age; Rfactor parameters named 'Rfactor_<strain>' by default; strain
    susceptibility a modified Dirichlet group with weights 1
    (src/model.cc:2768-2784); everyone starts in the first strain's dps
    (src/data.cc:440-456); data columns per strain."""
    import toml_models
    spec, values = toml_models.toml_sird_strains()
    names = [p["name"] for p in spec["params"]]
    assert spec["G"] == 4 and spec["dp_age"] == [0, 1, 0, 1]
    assert [names[i] for i in spec["strain_Rfactor"]] == ["Rfactor_s1", "Rfactor_s2"]
    assert [spec["params"][i]["prior"] for i in spec["strain_Rfactor"]] == [("fixed", 1.0), ("uniform", 0.5, 2.0)]
    assert spec["pop"][0][0] == [60000, 40000, 0, 0]
    s1, s2 = names.index("sus s1"), names.index("sus s2")
    assert spec["sus"] == [s1, s1, s2, s2] and {"param": [s1, s2], "frac": [1.0, 1.0]} in spec["dirichlet"]
    assert values[s2] == pytest.approx(0.3) and spec["params"][s1]["prior"] == ("gamma", 3.0, 1.0)
    si = [o for o in spec["obs"] if o["datatable"] == 0]
    assert {tuple(o["dp"]) for o in si[:10]} == {(0, 1)} and {tuple(o["dp"]) for o in si[10:]} == {(2, 3)}


def test_multi_area_needs_matrix(tmp_path):
    d = str(tmp_path)
    _write(d, "areas.csv", "area,population\nX,100\nY,200\n")
    text = BASE.replace('R_spline = [{bp="[bp:R_spline.csv]", value="[value:R_spline.csv]", '
                        'prior="[prior:R_spline.csv]"}]', 'R_spline = [{value="2", prior="Uniform(1,3)"}]')
    with pytest.raises(T.TomlError, match="geo_mixing_matrix"):
        T.load(_write(d, "m.toml", text), datadir=d)
    _write(d, "g.csv", "1,0.5\n0.5\n")
    with pytest.raises(T.TomlError, match="2 x 2"):
        T.load(_write(d, "n.toml", text + 'geo_mixing_matrix = "g.csv"\n'), datadir=d)


def _base_model(d, name, trans_I_R, spline):
    _write(d, "areas.csv", "area,population\nX,100\n")
    text = BASE.replace('prob_value="0.8", prob_prior="Uniform(0.5,1)"', trans_I_R)
    text = text.replace('R_spline = [{bp="[bp:R_spline.csv]", value="[value:R_spline.csv]", '
                        'prior="[prior:R_spline.csv]"}]', spline)
    return _write(d, name, text)


def test_dirichlet_branches(tmp_path):
    d = str(tmp_path)
    spline = 'R_spline = [{value="2", prior="Uniform(1,3)"}]'
    # A Dirichlet prior on one branch only: refused.
    p = _base_model(d, "a.toml", 'prob_value="*", prob_prior="Dir(1)"', spline)
    with pytest.raises(T.TomlError, match="Dir"):
        T.load(p, datadir=d)
    # Dir(*) with Dir(1): a flat and an alpha prior in one group, refused as
    # the reference does (src/model.cc:2924-2931).
    text = open(p).read().replace('prob_value="0.2", prob_prior="Uniform(0,0.5)"', 'prob_value="0.3", prob_prior="Dir(*)"')
    with pytest.raises(T.TomlError, match="all Dir"):
        T.load(_write(d, "m.toml", text), datadir=d)
    # Dir(*) on every branch out of I: one Dirichlet group, flat priors, '*'
    # filled with one minus the others.
    text = text.replace('prob_prior="Dir(1)"', 'prob_prior="Dir(*)"')
    p = _write(d, "b.toml", text)
    spec, values, _ = T.load(p, datadir=d)
    names = [q["name"] for q in spec["params"]]
    ih, ir = names.index("I->H branch prob."), names.index("I->R branch prob.")
    assert spec["dirichlet"] == [{"param": [ih, ir]}]
    assert spec["params"][ih]["prior"] == spec["params"][ir]["prior"] == ("dirflat",)
    assert values[ih] == 0.3 and values[ir] == pytest.approx(0.7, abs=1e-15)
    assert [t.get("prob") for t in spec["trans"]] == [None, [ih], [ir], None]
    # Dir(alpha): a gamma(alpha, 1) prior per branch (src/model.cc:907-913,1624-1637).
    spec, _, _ = T.load(_write(d, "g.toml", text.replace('prob_value="0.3", prob_prior="Dir(*)"',
                                                          'prob_value="0.3", prob_prior="Dir(0.4)"')
                                                  .replace('prob_prior="Dir(*)"', 'prob_prior="Dir(2.5)"')), datadir=d)
    assert spec["params"][ih]["prior"] == ("gamma", 0.4, 1.0) and spec["params"][ir]["prior"] == ("gamma", 2.5, 1.0)
    # Dir(mean,sd): one mean '*', one sd; alpha0 = m(1-m)/sd^2 - 1 from the
    # branch whose sd is set, then gamma(alpha0 * mean_i, 1) (src/model.cc:2936-3020).
    spec, _, _ = T.load(_write(d, "h.toml", text.replace('prob_value="0.3", prob_prior="Dir(*)"',
                                                          'prob_value="0.3", prob_prior="Dir(0.25,0.05)"')
                                                  .replace('prob_prior="Dir(*)"', 'prob_prior="Dir(*,*)"')), datadir=d)
    a0 = 0.25 * 0.75 / 0.05 ** 2 - 1
    assert spec["params"][ih]["prior"] == ("gamma", pytest.approx(a0 * 0.25), 1.0)
    assert spec["params"][ir]["prior"] == ("gamma", pytest.approx(a0 * 0.75), 1.0)
    with pytest.raises(T.TomlError, match="standard deviation"):
        T.load(_write(d, "i.toml", text.replace('prob_value="0.3", prob_prior="Dir(*)"',
                                                 'prob_value="0.3", prob_prior="Dir(0.25,0.05)"')
                                         .replace('prob_prior="Dir(*)"', 'prob_prior="Dir(*,0.1)"')), datadir=d)
    # Values only: fixed probabilities, '*' filled the same way, no group.
    text = text.replace('prob_prior="Dir(*)"', "")
    spec, values, _ = T.load(_write(d, "c.toml", text), datadir=d)
    assert spec["dirichlet"] == [] and spec["params"][ir]["prior"] == ("fixed", pytest.approx(0.7, abs=1e-15))


def test_smoothing_prior(tmp_path):
    """log_smooth between three knots: Pr gains, for each neighbouring pair,
    -(log x2 - log x1)^2 / (2 s^2) - log x2 (src/model.cc:1684-1690)."""
    from oracle import derive
    d = str(tmp_path)
    spline = ('R_spline = [{bp="start|2020-03-15|end", value="2|1|1.5", prior="Uniform(0.5,3)", '
              'smooth_type="log_smooth", smooth="0.5"}]')
    p = _base_model(d, "s.toml", 'prob_value="0.8", prob_prior="Uniform(0.5,1)"', spline)
    spec, values, _ = T.load(p, datadir=d)
    sp = spec["splines"][0]
    assert sp["smooth_type"] == derive.SMOOTH_LOG and sp["smooth"] == [0.5, 0.5, 0.5]
    want = sum(-(np.log(b) - np.log(a)) ** 2 / (2 * 0.25) - np.log(b) for a, b in ((2.0, 1.0), (1.0, 1.5)))
    assert derive.spline_prior(spec, values) == pytest.approx(want, rel=1e-14)
    # 'smooth' (normal on differences) and a fixed pair that contributes nothing.
    spline2 = spline.replace('log_smooth', 'smooth').replace('prior="Uniform(0.5,3)"', 'prior="Fixed(2)|Fixed(1)|Uniform(0.5,3)"')
    spec, values, _ = T.load(_base_model(d, "t.toml", 'prob_value="0.8", prob_prior="Uniform(0.5,1)"', spline2), datadir=d)
    assert derive.spline_prior(spec, values) == pytest.approx(-(1.0 - 1.5) ** 2 / (2 * 0.25), rel=1e-14)
    with pytest.raises(T.TomlError, match="smooth_type"):
        T.load(_base_model(d, "u.toml", 'prob_value="0.8", prob_prior="Uniform(0.5,1)"',
                           spline.replace("log_smooth", "wiggly")), datadir=d)


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="the reference's examples are not present")
def test_reference_examples_host_tables(gpu_lib):
    """Every reference example the reader loads: the C++ host's State::set_param
    tables, prior and bounds equal the FP64 oracle's for the file's values and
    two prior draws (the same check as the fixtures', on the reference's own
    models: Erlang stages, branches, ages, democats, effects, covariates,
    perturbed mixing, observation splines)."""
    import tomli
    from beep_amd import _lib
    from oracle import derive
    n = 0
    for path in sorted(glob.glob(os.path.join(REF_EXAMPLES, "*.toml"))):
        with open(path, "rb") as f:
            dd = tomli.load(f)
        try:
            spec, values, _ = T.load(path, datadir=os.path.join(os.path.dirname(REF_EXAMPLES), dd.get("datadir", "")))
        except T.TomlError:
            continue
        nz = derive.normalise(spec)
        rng = np.random.default_rng(11)
        draws = list(specmod.sample_prior(spec, rng, 2))
        if not np.isnan(values).any():
            draws.insert(0, values)
        S = spec["period"] * spec["spd"]
        for pv in draws:
            t = _lib.spec_tables(spec, pv)
            d = derive.derive(spec, nz, pv)
            ba = d["beta"] * d["af"][np.arange(S) // spec["spd"]]
            assert np.allclose(t["beta_af"], ba, rtol=1e-10, atol=0), path
            assert np.allclose(t["rate"], d["rate"], rtol=1e-14), path
            assert np.allclose(t["sus"], d["sus"], rtol=1e-14), path
            assert t["Pr"] == pytest.approx(derive.prior(spec, pv), rel=1e-12, abs=1e-9), path
            assert t["inbounds"] == derive.inbounds(spec, pv), path
        n += 1
    assert n >= 30


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="the reference's examples are not present")
def test_reference_examples():
    """Every example either loads or is refused with TomlError; EX.toml (SEIR
    with Erlang E, weekly E->I counts and R populations) matches its files."""
    loaded = []
    for path in sorted(glob.glob(os.path.join(REF_EXAMPLES, "*.toml"))):
        import tomli
        with open(path, "rb") as f:
            dd = tomli.load(f)
        datadir = os.path.join(os.path.dirname(REF_EXAMPLES), dd.get("datadir", ""))
        try:
            spec, values, _ = T.load(path, datadir=datadir)
        except T.TomlError:
            continue
        loaded.append(os.path.basename(path))
        assert spec["obs"] and len(values) == len(spec["params"])
    assert "EX.toml" in loaded
    spec, values, _ = T.load(os.path.join(REF_EXAMPLES, "EX.toml"), datadir=os.path.join(REF_EXAMPLES, "Data_EX"))
    assert [c["name"] for c in spec["comps"]] == ["S", "E", "E1", "I", "R"]
    assert spec["period"] == 140 and spec["spd"] == 2
    rows = {}
    for f in ("E-I.csv", "R.csv", "population.csv"):
        with open(os.path.join(REF_EXAMPLES, "Data_EX", f)) as fh:
            rows[f] = [line.strip().split(",") for line in fh.read().strip().splitlines()[1:]]
    ei_rows = [r for r in rows["E-I.csv"] if 0 <= int(r[0]) and int(r[0]) + 7 <= 140]
    r_rows = [r for r in rows["R.csv"] if 0 <= int(r[0]) < 140]
    ei = [o for o in spec["obs"] if o["datatable"] == 0]
    r = [o for o in spec["obs"] if o["datatable"] == 1]
    assert [o["y"] for o in ei] == [float(x[1]) for x in ei_rows]
    assert [o["y"] for o in r] == [float(x[1]) for x in r_rows]
    assert spec["pop"][0][0] == [int(float(rows["population.csv"][0][1]) + 0.5)]
    assert not math.isnan(values[[p["name"] for p in spec["params"]].index("R-spline")])


def _sex_ages(tmp_path, extra, areas=None):
    """The two-area, age-by-sex fixture read in prediction mode with extra
    top-level keys; areas replaces its areas file."""
    import shutil
    data = tmp_path / "data"
    shutil.copytree(os.path.join(FIX, "data_seir_sex_ages"), data)
    if areas is not None:
        (data / "areas.csv").write_text(areas)
    p = tmp_path / "m.toml"
    p.write_text(open(os.path.join(FIX, "seir_sex_ages.toml")).read() + "\n" + extra + "\n")
    return T.load(str(p), datadir=str(data), prediction=True)


MODS_TOML = '''modification = [
{type="trans_rate_fac", trans="S->E", start="7", end="21", factor="1.5", geo_filt="area:South", democats_filt="Sex:F"},
{type="beta_fac", start="28", factor="0.5"},
{type="efoi_fac", end="14", factor="2", geo_filt="region:Coast|Hills,area:North"},
{type="spline_fac", name="R_t", start="35", end="42", factor="0.8"},
{type="spline_set", name="efoi(t)", start="42", value="0"},
{type="trans_rate_fac", trans="I->R", factor="1.2", democats_filt="age:old,Sex:M"}
]'''
AREAS_REGION = "area,region,young,old,M,F\nNorth,Coast,30000,10000,0.48,0.52\nSouth,Hills,20000,25000,49,51\n"


def test_prediction_modifications(tmp_path):
    """Modifications read in prediction mode (src/inputs.cc:1986-2061,
    src/data.cc:1716-1826, src/model.cc:1495-1575), worked by hand: dps
    enumerate (age, Sex) with age fastest, so Sex:F is dps 2 and 3 and
    old-and-M is dp 1; a comma joins terms that must all hold; times are in
    days (one step per day here); the beta and efoi factors act on the one
    strain 'all'; splines are found by their default names."""
    spec, values, meta = _sex_ages(tmp_path, MODS_TOML, AREAS_REGION)
    tr = {(spec["comps"][t["from"]]["name"], spec["comps"][t["to"]]["name"]): k for k, t in enumerate(spec["trans"])}
    m = spec["modifications"]
    assert [x["type"] for x in m] == ["trans_rate_fac", "beta_fac", "efoi_fac", "spline_fac", "spline_set",
                                      "trans_rate_fac"]
    assert (m[0]["start"], m[0]["end"], m[0]["factor"]) == (7, 21, 1.5)
    assert m[0]["trans"] == [tr[("S", "E")]] and m[0]["areas"] == [1] and m[0]["dps"] == [2, 3]
    assert (m[1]["start"], m[1]["end"], m[1]["strain"], m[1]["areas"]) == (28, 56, 0, [0, 1])
    assert (m[2]["start"], m[2]["end"], m[2]["factor"], m[2]["areas"]) == (0, 14, 2.0, [0])
    assert (m[3]["spline"], m[3]["start"], m[3]["end"]) == (spec["R_spline"], 35, 42)
    assert (m[4]["spline"], m[4]["factor"], m[4]["start"], m[4]["end"]) == (spec["efoi_spline"], 0.0, 42, 56)
    assert m[5]["trans"] == [tr[("I", "R")]] and m[5]["dps"] == [1] and m[5]["areas"] == [0, 1]
    assert meta["prediction"]["pred_start"] == 0
    # Without prediction mode the entries are only recorded.
    spec0, _, meta0 = T.load(str(tmp_path / "m.toml"), datadir=str(tmp_path / "data"))
    assert "modifications" not in spec0 and len(meta0["prediction"]["modification"]) == 6
    # The host model takes the list as it is: its tables equal the oracle's.
    from oracle import derive
    from beep_amd import _lib
    t = _lib.spec_tables(spec, values)
    d = derive.derive(spec, derive.normalise(spec), values)
    assert np.allclose(t["beta_af"], d["beta"] * d["af"][np.arange(56)], rtol=1e-12, atol=0)
    bm, em, tm = derive.modification_tables(spec)
    assert tm[7:21, 1, tr[("S", "E")]][:, [2, 3]].min() == tm[7:21, 1, tr[("S", "E")]][:, [2, 3]].max() == 1.5
    assert tm[:, :, tr[("I", "R")], 1].min() == 1.2 and (tm[:, 0, tr[("S", "E")]] == 1).all()
    assert (em[:14, 0] == 2).all() and (em[:14, 1] == 1).all() and (em[14:] == 1).all()


@pytest.mark.parametrize("entry, msg", [
    ('{type="pop_fac", factor="2"}', "not recognised"),
    ('{type="trans_rate_fac", trans="S->I", factor="2"}', "'S->I' is not recognised"),
    ('{type="trans_rate_fac", trans="S->E"}', "numeric 'factor'"),
    ('{type="beta_fac", factor="-1"}', "not positive"),
    ('{type="beta_fac", factor="2", geo_filt="area:West"}', "'West' is not found"),
    ('{type="beta_fac", factor="2", geo_filt="area:North|North"}', "must not be repeated"),
    ('{type="beta_fac", factor="2", democats_filt="Sex:X"}', "cannot find the value 'X'"),
    ('{type="beta_fac", factor="2", democats_filt="Gender:F"}', "cannot find the value 'Gender'"),
    ('{type="beta_fac", factor="2", strain="delta"}', "'strain=\"delta\"' is not recognised"),
    ('{type="spline_fac", name="R(t)", factor="2"}', "cannot find spline with the name 'R\\(t\\)'"),
    ('{type="spline_set", name="R_t", factor="2"}', "numeric 'value'"),
    ('{type="beta_fac", start="30", end="20", factor="2"}', "in that order"),
    ('{type="beta_fac", factor="2", colour="red"}', "colour"),
])
def test_modification_rejections(tmp_path, entry, msg):
    with pytest.raises(T.TomlError, match=msg):
        _sex_ages(tmp_path, f"modification = [{entry}]")


def test_modification_start_before_prediction_start(tmp_path):
    with pytest.raises(T.TomlError, match="cannot be before 'pred_start'"):
        _sex_ages(tmp_path, 'prediction_start = "30"\nmodification = [{type="beta_fac", start="20", factor="2"}]')
    (tmp_path / "b").mkdir()
    _, _, meta = _sex_ages(tmp_path / "b", 'prediction_start = "10"\n'
                                          'modification = [{type="beta_fac", start="20", factor="2"}]')
    assert meta["prediction"]["pred_start"] == 10


def test_prediction_end_extends_the_period(tmp_path):
    """prediction_end beyond 'end' lengthens the period (src/details.cc:92)
    and the parameter splines reach it (src/inputs.cc:1475-1493): the last
    knot moves when the last two knots are specified alike (one value for
    both), else a knot copying the last is added at the new end, as its
    own parameter when the knots are unnamed."""
    spec, values, _ = _sex_ages(tmp_path, 'prediction_end = "70"')
    assert spec["period"] == 70
    assert spec["splines"][0]["t"] == [0.0, 70.0] and spec["splines"][0]["param"][0] == spec["splines"][0]["param"][1]
    assert spec["splines"][1]["t"] == [0.0, 70.0]
    (tmp_path / "b").mkdir()
    spec2, _, _ = _sex_ages(tmp_path / "b", 'prediction_end = "50"')       # before 'end': no change
    assert spec2["period"] == 56 and spec2["splines"][0]["t"] == [0.0, 56.0]
    # A 3-knot R(t) whose last two knots differ gains a fourth.
    p = tmp_path / "c.toml"
    src = open(os.path.join(FIX, "seir_sex_ages.toml")).read()
    src = src.replace('R_spline = [{ value="2.1", prior="Uniform(0.5,4)"}]',
                      'R_spline = [{ bp="start|28|end", value="2.1|1.5|1.2", prior="Uniform(0.5,4)"}]')
    p.write_text(src + '\nprediction_end = "70"\n')
    spec3, values3, _ = T.load(str(p), datadir=os.path.join(FIX, "data_seir_sex_ages"), prediction=True)
    sp = spec3["splines"][0]
    names = [q["name"] for q in spec3["params"]]
    assert sp["t"] == [0.0, 28.0, 56.0, 70.0]
    assert [names[k] for k in sp["param"]] == ["R-spline t=0", "R-spline t=28", "R-spline t=56", "R-spline t=70"]
    assert values3[sp["param"][3]] == 1.2 and spec3["params"][sp["param"][3]]["prior"] == ("uniform", 0.5, 4.0)
    # Without prediction mode prediction_end changes nothing.
    spec4, _, _ = T.load(str(p), datadir=os.path.join(FIX, "data_seir_sex_ages"))
    assert spec4["period"] == 56 and spec4["splines"][0]["t"] == [0.0, 28.0, 56.0]


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="the reference's examples are not present")
def test_reference_prediction_examples():
    """EX_F3 and EX_F4 read in prediction mode: a doubled (EX_F3) or reduced
    (EX_F4) S->E rate from day 80 (40), on a period extended to
    prediction_end (EX_F3: 100 days)."""
    f3, _, meta = T.load(os.path.join(REF_EXAMPLES, "EX_F3.toml"), datadir=os.path.join(REF_EXAMPLES, "Data_EX_F3"),
                         prediction=True)
    spd = f3["spd"]
    assert f3["period"] == 100
    (m,) = f3["modifications"]
    se = next(k for k, t in enumerate(f3["trans"]) if f3["comps"][t["from"]]["name"] == "S")
    assert (m["type"], m["trans"], m["start"], m["end"], m["factor"]) == ("trans_rate_fac", [se], 80 * spd,
                                                                         100 * spd, 2.0)
    assert meta["prediction"]["pred_start"] == 80
    f4, _, _ = T.load(os.path.join(REF_EXAMPLES, "EX_F4.toml"), datadir=os.path.join(REF_EXAMPLES, "Data_EX_F4"),
                      prediction=True)
    assert [(x["start"], x["factor"]) for x in f4["modifications"]] == [(40 * f4["spd"], 0.3)]


def test_data_table_filters(tmp_path):
    """geo_filt and democats_filt on data tables (Data::get_datafilter,
    src/data.cc:1460-1575): a per-sex table over the South only, a per-area
    table of the old only, a plain table of the young women of the Coast
    region; columns are read by their names, the selections worked by hand
    (dps (age, Sex) with age fastest: young-M 0, old-M 1, young-F 2, old-F 3)."""
    import shutil
    data = tmp_path / "data"
    shutil.copytree(os.path.join(FIX, "data_seir_sex_ages"), data)
    (data / "areas.csv").write_text(AREAS_REGION)
    src = open(os.path.join(FIX, "seir_sex_ages.toml")).read()
    start = src.index("data_tables = [")
    src = src[:start] + '''data_tables = [
{type="transition", observation="S->E", timestep="7", obsmodel="normal 10%", democats_dep="Sex", geo_filt="area:South", file="S-E.csv"},
{type="population", observation="R", obsmodel="normal 20", geo_dep="area", democats_filt="age:old", file="R-area.csv"},
{type="population", observation="R", obsmodel="normal 20", geo_filt="region:Coast", democats_filt="age:young,Sex:F", file="R-all.csv"}
]
'''
    (data / "R-area.csv").write_text("time,North,South\n0,1,2\n7,3,4\n")
    (data / "R-all.csv").write_text("time,Data\n0,5\n14,6\n")
    p = tmp_path / "f.toml"
    p.write_text(src)
    spec, _, _ = T.load(str(p), datadir=str(data))
    t0, t1, t2 = ([o for o in spec["obs"] if o["datatable"] == d] for d in range(3))
    # Per-sex columns over the South's cells only.
    assert {(tuple(o["area"]), tuple(o["dp"])) for o in t0} == {((1,), (0, 1)), ((1,), (2, 3))}
    # Per-area columns of the old; the young women of the Coast.
    assert [(o["area"], o["dp"], o["y"]) for o in t1] == [([0], [1, 3], 1.0), ([0], [1, 3], 3.0),
                                                           ([1], [1, 3], 2.0), ([1], [1, 3], 4.0)]
    assert [(o["area"], o["dp"], o["y"]) for o in t2] == [([0], [2], 5.0), ([0], [2], 6.0)]
    # A filter that leaves a column with no cells drops it; one that leaves none is an error.
    bad = src.replace('geo_filt="region:Coast", democats_filt="age:young,Sex:F"', 'geo_filt="region:Coast,area:South"')
    p.write_text(bad)
    with pytest.raises(T.TomlError, match="does not generate any data columns"):
        T.load(str(p), datadir=str(data))


def test_R_and_efoi_splines_per_area_and_strain(tmp_path):
    """R_spline and efoi_spline entries naming an area ('geo_filt', an area
    code) or a strain (Model::update_R_ref / update_efoi_ref,
    src/model.cc:384-480, find_spline's parameter names src/inputs.cc:
    1392-1403): one spline per area, parameters named '<root> <area>', the
    groups in spec["R_splines"] / spec["efoi_splines"], spline names with the
    area for modifications; every area (and strain) covered exactly once."""
    text = open(os.path.join(FIX, "seir_sex_ages.toml")).read()
    text = text.replace('R_spline = [{ value="2.1", prior="Uniform(0.5,4)"}]',
                        'R_spline = [{ geo_filt="North", value="2.1", prior="Uniform(0.5,4)"},\n'
                        '            { geo_filt="South", value="1.6", prior="Uniform(0.5,4)"}]')
    text = text.replace('efoi_spline = [{ value="3", age_dist="1|3"}]',
                        'efoi_spline = [{ geo_filt="North", value="3", age_dist="1|3"}, { geo_filt="South", value="5"}]')
    p = tmp_path / "g.toml"
    p.write_text(text + '\nmodification = [{type="spline_fac", name="R_t South", start="10", factor="0.5"}]\n')
    spec, values, _ = T.load(str(p), datadir=os.path.join(FIX, "data_seir_sex_ages"), prediction=True)
    names = [q["name"] for q in spec["params"]]
    rs, es = spec["R_splines"], spec["efoi_splines"]
    assert [g["areas"] for g in rs] == [[0], [1]] and [g["areas"] for g in es] == [[0], [1]]
    assert [names[spec["splines"][g["spline"]]["param"][0]] for g in rs] == ["R-spline North", "R-spline South"]
    assert [values[spec["splines"][g["spline"]]["param"][0]] for g in rs] == [2.1, 1.6]
    assert [names[spec["splines"][g["spline"]]["param"][0]] for g in es] == ["efoi-spline North", "efoi-spline South"]
    assert es[1]["agedist"] == [1.0] * 4 and es[0]["agedist"] != [1.0] * 4
    assert spec["modifications"][0]["spline"] == rs[1]["spline"]
    # The host model takes the groups: South's beta over North's is 1.6 / 2.1
    # times the ratio of the two groups' NGM ratios.
    from beep_amd import _lib
    t = _lib.spec_tables(spec, values)
    assert t["beta_af"].shape == (56, 2) and not np.allclose(t["beta_af"][:, 0], t["beta_af"][:, 1])
    # Every area exactly once.
    bad = text.replace('{ geo_filt="South", value="1.6", prior="Uniform(0.5,4)"}', '{ value="1.6"}')
    p.write_text(bad)
    with pytest.raises(T.TomlError, match="already has a spline"):
        T.load(str(p), datadir=os.path.join(FIX, "data_seir_sex_ages"))
    bad = text.replace(',\n            { geo_filt="South", value="1.6", prior="Uniform(0.5,4)"}', '')
    p.write_text(bad)
    with pytest.raises(T.TomlError, match="must be set for all areas"):
        T.load(str(p), datadir=os.path.join(FIX, "data_seir_sex_ages"))
    bad = text.replace('geo_filt="North", value="2.1"', 'geo_filt="East", value="2.1"')
    p.write_text(bad)
    with pytest.raises(T.TomlError, match="cannot find 'area=East'"):
        T.load(str(p), datadir=os.path.join(FIX, "data_seir_sex_ages"))


def test_efoi_spline_per_strain(tmp_path):
    text = open(os.path.join(FIX, "sird_strains.toml")).read()
    text = text.replace('efoi_spline = [{ value="2"}]', 'efoi_spline = [{ strain="s1", value="2"}, { strain="s2", value="0.5"}]')
    p = tmp_path / "s.toml"
    p.write_text(text)
    spec, values, _ = T.load(str(p), datadir=os.path.join(FIX, "data_sird_strains"))
    es = spec["efoi_splines"]
    names = [q["name"] for q in spec["params"]]
    assert [g["strains"] for g in es] == [[0], [1]]
    assert [names[spec["splines"][g["spline"]]["param"][0]] for g in es] == ["efoi-spline s1", "efoi-spline s2"]
    from beep_amd import _lib
    t = _lib.spec_tables(spec, values)
    dpm = spec["G"] // 2
    assert np.allclose(t["efoi"][:, :, dpm:] / t["efoi"][:, :, :dpm], 0.25, rtol=1e-14)
    p.write_text(text.replace('strain="s2"', 'strain="s3"'))
    with pytest.raises(T.TomlError, match="cannot find 'strain=s3'"):
        T.load(str(p), datadir=os.path.join(FIX, "data_sird_strains"))
