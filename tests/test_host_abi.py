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
        assert np.allclose(t["efoi"], d["efoi"][:, None] * d["agedist"][None, :])
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
