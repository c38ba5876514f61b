"""The C ABI library (CPU only): exports every symbol the header declares,
and the C++ host model's derived tables equal the oracle's FP64 ones."""
import os
import re

import numpy as np
import pytest

import toml_models
from beep_amd import _lib, configs, spec as specmod
from oracle import derive

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported(gpu_lib):
    hdr = open(os.path.join(ROOT, "include", "beep_gpu.h")).read()
    declared = set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(beep_\w+)\s*\(", hdr, re.M))
    assert declared and declared == set(_lib.SYMBOLS)
    for name in declared:
        assert hasattr(gpu_lib, name), name


@pytest.mark.parametrize("mk", [configs.seir_single, configs.spatial_abcsmc, configs.dense_pmcmc,
                                toml_models.toml_seir_erlang, toml_models.toml_seird_branch,
                                toml_models.toml_sir_3areas, toml_models.toml_sird_ages,
                                toml_models.toml_seir_sex_ages, toml_models.toml_sird_strains])
def test_host_tables_match_oracle(gpu_lib, mk):
    spec, truth = mk()
    nz = derive.normalise(spec)
    rng = np.random.default_rng(3)
    for pv in [truth] + list(specmod.sample_prior(spec, rng, 3)):
        t = _lib.spec_tables(spec, pv)
        d = derive.derive(spec, nz, pv)
        S, spd = spec["period"] * spec["spd"], spec["spd"]
        ba = d["beta"] * d["af"][np.arange(S) // spd]
        assert np.allclose(t["beta_af"], ba, rtol=1e-12, atol=0)
        assert np.allclose(t["rate"], d["rate"], rtol=1e-14)
        assert np.allclose(t["sus"], d["sus"]) and np.allclose(t["inf"], d["inf"])
        assert np.allclose(t["N"], d["Ntime"] if spec.get("age_mod") else d["Ntime"][0], rtol=1e-12)
        assert np.allclose(t["efoi"], d["efoi_cell"], rtol=1e-14, atol=0)
        assert np.allclose(t["efoi"], (d["efoi"][:, None] * d["agedist"][None, :])[:, None, :], rtol=1e-14, atol=0)
        assert np.allclose(t["geo"], d["geo"])
        assert t["Pr"] == pytest.approx(derive.prior(spec, pv), rel=1e-14)
        assert t["inbounds"] == derive.inbounds(spec, pv)


def test_spec_errors_are_reported(gpu_lib):
    spec, truth = configs.seir_single()
    bad = dict(spec)
    bad["trans"] = [dict(t) for t in spec["trans"]]
    bad["trans"][1]["to"] = 9
    with pytest.raises(_lib.BeepError) as e:
        _lib.spec_tables(bad, truth)
    assert e.value.code == -1 and "out of range" in str(e.value)
    bad = dict(spec)
    bad["trans"] = [dict(t, infection=False) for t in spec["trans"]]
    with pytest.raises(_lib.BeepError, match="no infection transition"):
        _lib.spec_tables(bad, truth)
    th = truth.copy()
    th[1] = 0.0       # zero mean residence time
    with pytest.raises(_lib.BeepError):
        _lib.spec_tables(spec, th)


def test_compact_observation_table_matches_expanded_list(gpu_lib):
    """An obs_table fills the same observation structs as its expanded obs
    list (values, windows, element lists), and the host model accepts it."""
    spec, truth = configs.seir_areas(13, 4, 30, seed=3, daily_obs=True)
    assert "obs_table" in spec and len(spec["obs_table"]["y"]) == 13 * 30
    bufa, bufb = _lib._SpecBuffers(spec), _lib._SpecBuffers(specmod.expand_obs(spec))   # own the arrays
    a, b = bufa.spec, bufb.spec
    assert a.nobs == b.nobs == 13 * 30
    for i in range(a.nobs):
        x, y = a.obs[i], b.obs[i]
        for f in ("y", "sd", "shape", "factor", "thresh", "percent", "type", "datatable", "graph", "sett_i", "sett_f",
                  "narea", "ndp", "ntr", "nco", "factor_spline"):
            assert getattr(x, f) == getattr(y, f), (i, f)
        assert [x.area[k] for k in range(x.narea)] == [y.area[k] for k in range(y.narea)]
        assert [x.dp[k] for k in range(x.ndp)] == [y.dp[k] for k in range(y.ndp)]
        assert [x.tr[k] for k in range(x.ntr)] == [y.tr[k] for k in range(y.ntr)]
    t = _lib.spec_tables(spec, truth)
    assert np.isfinite(t["beta_af"]).all()
