"""Posterior sample files in the reference's format (State::save / load,
src/state.cc:709-808; sample_information.txt, src/mpi.cc:715-723 and
src/simulate.cc:77-138): written and read back exactly, with the reference's
checks on names, row lengths and the information file."""
import os

import numpy as np
import pytest

from beep_amd import configs, samples


def test_sample_round_trip(tmp_path):
    spec, truth = configs.seir_areas(7, 4, 30, seed=2)
    rng = np.random.default_rng(1)
    S, A, T, G = 30, 7, 3, 4
    tn = rng.integers(0, 50, (S, A, T, G)).astype(np.int32)
    pv = truth * rng.uniform(0.5, 1.5, len(truth))
    prefix = str(tmp_path / "sample3")
    samples.write_sample(prefix, spec, pv, tn, EF=1234.5678901234567)
    lines = open(prefix + "_parameter.csv").read().splitlines()
    assert lines[0] == "Parameter,Value" and lines[1].startswith("Error function,")
    assert [l.split(",")[0] for l in lines[2:]] == [p["name"] for p in spec["params"]]
    tl = open(prefix + "_transition.csv").read().splitlines()
    assert tl[0] == "Time,Transitions (ordered by area x trans x democatpos)" and len(tl) == 1 + S
    row = tl[5].split(",")
    assert len(row) == 1 + A * T * G and [int(x) for x in row[1:1 + G]] == list(tn[4, 0, 0, :])
    EF, pv2, tn2 = samples.read_sample(prefix, spec, S)
    assert EF == 1234.5678901234567 and np.array_equal(pv2, pv) and np.array_equal(tn2, tn)
    # A posterior that ends at step 12: its file holds 12 rows, and State::load
    # gives zero counts from there on (more rows than that is an error).
    samples.write_sample(prefix, spec, pv, tn[:12])
    _, _, tn3 = samples.read_sample(prefix, spec, 12)
    assert np.array_equal(tn3[:12], tn[:12]) and not tn3[12:].any()
    with pytest.raises(ValueError, match="Problem loading"):
        samples.read_sample(prefix, spec, 11)


def test_sample_errors(tmp_path):
    spec, truth = configs.seir_areas(3, 1, 10, seed=2)
    prefix = str(tmp_path / "s")
    samples.write_sample(prefix, spec, truth, np.zeros((10, 3, 3, 1), np.int32))
    other = dict(spec, params=[dict(p) for p in spec["params"]])
    other["params"][1]["name"] = "renamed"
    with pytest.raises(ValueError, match="Problem loading"):
        samples.read_sample(prefix, other, 10)
    with open(prefix + "_transition.csv", "a") as f:
        f.write("10.5,1,2\n")
    with pytest.raises(ValueError, match="Problem loading"):
        samples.read_sample(prefix, spec, 10)


def test_sample_information(tmp_path):
    samples.write_information(str(tmp_path), 40, "model.toml", 0, 120, 1, 100)
    text = open(tmp_path / "sample_information.txt").read().splitlines()
    assert text == ["# Samples = '40'", "TOML file = 'model.toml'", "start = '0'", "end = '120'",
                    "division_per_time = '1'", "timesteps = '100'"]
    assert samples.read_information(str(tmp_path), toml_file="model.toml", start=0, end=120,
                                    division_per_time=1) == (40, 100)
    with pytest.raises(ValueError, match="must be the same"):
        samples.read_information(str(tmp_path), toml_file="other.toml")
    with pytest.raises(ValueError, match="must be the same"):
        samples.read_information(str(tmp_path), division_per_time=2)
    with open(tmp_path / "sample_information.txt", "w") as f:
        f.write("# Samples = 40\n")
    with pytest.raises(ValueError, match="problem loading the posterior"):
        samples.read_information(str(tmp_path))
