"""The C++ host mirror of State / ObservationModel / Mbp (beep_amd/csrc/host)
run through its self-check program on the GPU; the posterior-sample files it
writes (State::save format) are re-scored by the FP64 oracle."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from beep_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "beep_amd", "bin", "gpu_state_check")


def _spec():
    spec, _ = configs.seir_single(obs=False)
    obs = []
    for d in range(100):
        y = math.floor(600.0 * math.exp(-0.5 * ((d - 45.0) / 12.0) ** 2) + 0.5)
        obs.append({"y": float(y), "sd": max(5.0, 0.1 * y), "type": "normal", "datatable": 0, "graph": 0,
                    "sett_i": d, "sett_f": d + 1, "area": [0], "dp": [0], "tr": [2]})
    spec["obs"] = obs
    return spec


def _read_sample(prefix, S):
    with open(prefix + "_parameter.csv") as f:
        lines = f.read().strip().split("\n")
    assert lines[0] == "Parameter,Value" and lines[1].startswith("Error function,")
    EF = float(lines[1].split(",")[1])
    params = [float(l.split(",")[1]) for l in lines[2:]]
    with open(prefix + "_transition.csv") as f:
        rows = f.read().strip().split("\n")
    assert rows[0].startswith("Time,Transitions")
    tn = np.array([[int(v) for v in r.split(",")[1:]] for r in rows[1:]])
    assert tn.shape[0] == S
    return EF, np.array(params), tn


def test_host_mirror_self_check(tmp_path):
    assert os.path.exists(BIN), "build() has not built the host mirror's check program"
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["failures"] == 0
    assert res["al_identity"] == pytest.approx(1.0)
    # The saved sample, re-scored by the oracle.
    spec = _spec()
    EF, params, tn = _read_sample(str(tmp_path / "sample0"), 100)
    assert params.tolist() == [2.1, 5.0, 7.0, 1.0]
    orc = O.Oracle(spec)
    st = orc.replay(orc.tables(params), tn.reshape(100, 1, 3, 1).astype(np.float64))
    L, _ = orc.obs_loglik(st)
    assert EF == pytest.approx(-2 * L, rel=1e-9)
    assert res["EF"] == pytest.approx(EF, rel=1e-15)
    assert res["likelihood"] == pytest.approx(orc.likelihood(st), rel=1e-5)
