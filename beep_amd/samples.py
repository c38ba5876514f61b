"""Posterior samples on disk, in the reference's format, and their ingest
into device contexts.

Files of one sample K (State::save, src/state.cc:709-739):
  sampleK_parameter.csv    "Parameter,Value", "Error function,<EF>", then one
                           "<name>,<value>" row per parameter in model order;
  sampleK_transition.csv   "Time,Transitions (ordered by area x trans x
                           democatpos)", then per step its time and the
                           transition counts [area][transition][dp].
The directory's sample_information.txt (Mpi::gather_samples,
src/mpi.cc:715-723) lists six lines of `key = 'value'`: # Samples, TOML file,
start, end, division_per_time, timesteps; Simulate::find_posterior_info
(src/simulate.cc:77-138) reads them back and checks them against the model.

Values are written with 17 significant digits so that a sample reloads to
the same parameters bit for bit (the reference writes the stream default; a
file it wrote reads the same way).  Loading follows State::load
(src/state.cc:743-808): counts of the steps from `timesteps` on are zero, and
the trajectory is replayed on the device (initialise_from_particle), which
rebuilds the populations, infectivity and means."""
import os

import numpy as np

INFO_KEYS = ("# Samples", "TOML file", "start", "end", "division_per_time", "timesteps")


def param_names(spec):
    return [p["name"] for p in spec["params"]]


def write_sample(prefix, spec, paramval, transnum, EF=0.0):
    """Writes prefix_parameter.csv and prefix_transition.csv; transnum is
    [S][A][T][G] (steps beyond those given are not written)."""
    names = param_names(spec)
    pv = np.asarray(paramval, dtype=np.float64)
    if len(pv) != len(names):
        raise ValueError("one value per parameter")
    with open(prefix + "_parameter.csv", "w") as f:
        f.write("Parameter,Value\n")
        f.write(f"Error function,{float(EF)!r}\n")
        for n, v in zip(names, pv):
            f.write(f"{n},{float(v)!r}\n")
    tn = np.asarray(transnum)
    spd = spec["spd"]
    with open(prefix + "_transition.csv", "w") as f:
        f.write("Time,Transitions (ordered by area x trans x democatpos)\n")
        for s in range(tn.shape[0]):
            row = tn[s].reshape(-1)
            f.write(repr((s + 0.5) / spd) + "," + ",".join(str(int(x)) for x in row) + "\n")


def read_sample(prefix, spec, ndivision_post):
    """State::load's reading: (EF, parameter vector, transnum [S][A][T][G])
    with the steps from ndivision_post on zero.  Errors as the reference's:
    a parameter name out of order, a row of the wrong length or extra lines
    raise ValueError naming the file."""
    names = param_names(spec)
    A, G, T = spec["A"], spec["G"], len(spec["trans"])
    S = spec["period"] * spec["spd"]
    fn = prefix + "_parameter.csv"
    with open(fn) as f:
        lines = [l.rstrip("\r\n") for l in f]
    if len(lines) < 2:
        raise ValueError(f"Problem loading the file '{fn}'")
    spl = lines[1].split(",")
    if len(spl) != 2:
        raise ValueError(f"Problem loading the file '{fn}'")
    EF = float(spl[1])
    pv = np.zeros(len(names))
    for th, name in enumerate(names):
        if 2 + th >= len(lines):
            raise ValueError(f"Problem loading the file '{fn}'")
        spl = lines[2 + th].split(",")
        if len(spl) != 2 or spl[0] != name:
            raise ValueError(f"Problem loading the file '{fn}'")
        pv[th] = float(spl[1])
    if any(l.strip() for l in lines[2 + len(names):]):
        raise ValueError(f"Problem loading the file '{fn}'")
    fn = prefix + "_transition.csv"
    tn = np.zeros((S, A, T, G), dtype=np.int32)
    with open(fn) as f:
        lines = [l.rstrip("\r\n") for l in f]
    if ndivision_post > S or len(lines) < 1 + ndivision_post:
        raise ValueError(f"Problem loading the file '{fn}'")
    for s in range(ndivision_post):
        spl = lines[1 + s].split(",")
        if len(spl) != 1 + A * T * G:
            raise ValueError(f"Problem loading the file '{fn}'")
        tn[s] = np.array([float(x) for x in spl[1:]], dtype=np.float64).astype(np.int32).reshape(A, T, G)
    if any(l.strip() for l in lines[1 + ndivision_post:]):
        raise ValueError(f"Problem loading the file '{fn}'")
    return EF, pv, tn


def write_information(directory, nsample, toml_file, start, end, division_per_time, timesteps):
    with open(os.path.join(directory, "sample_information.txt"), "w") as f:
        for k, v in zip(INFO_KEYS, (nsample, toml_file, start, end, division_per_time, timesteps)):
            f.write(f"{k} = '{v}'\n")


def read_information(directory, toml_file=None, start=None, end=None, division_per_time=None):
    """Simulate::find_posterior_info: (number of samples, timesteps), checking
    the model's TOML file, start, end and steps per unit time when given."""
    fn = os.path.join(directory, "sample_information.txt")
    with open(fn) as f:
        lines = [l.rstrip("\r\n") for l in f]
    if len(lines) < 6:
        raise ValueError("There was a problem loading the posterior samples.")
    vals = {}
    for k, line in zip(INFO_KEYS, lines[:6]):
        spl = line.split("=")
        if len(spl) != 2:
            raise ValueError("There was a problem loading the posterior samples.")
        q = spl[1].split("'")
        if len(q) != 3 or spl[0].strip() != k:
            raise ValueError("There was a problem loading the posterior samples.")
        vals[k] = q[1]
    for key, want in (("TOML file", toml_file), ("start", start), ("end", end),
                      ("division_per_time", division_per_time)):
        if want is not None and str(want) != vals[key]:
            raise ValueError(f"The {key} '{want}' must be the same as that used for inference '{vals[key]}'")
    return int(vals["# Samples"]), int(vals["timesteps"])


def save_posterior(directory, spec, ctx, inst, slots, paramvals, EF=None, timesteps=None, toml_file="model.toml",
                   start=0):
    """Writes the device instances `inst` (their parameter vectors
    paramvals) as samples 0.. of a posterior directory, with its
    sample_information.txt; counts of steps from `timesteps` on are not
    written (a posterior ends where its data end)."""
    os.makedirs(directory, exist_ok=True)
    S = spec["period"] * spec["spd"]
    timesteps = S if timesteps is None else timesteps
    for k, i in enumerate(inst):
        tn = ctx.get_transnum(int(i))[:timesteps]
        write_sample(os.path.join(directory, f"sample{k}"), spec, paramvals[k], tn,
                     0.0 if EF is None else EF[k])
    write_information(directory, len(inst), toml_file, start, start + spec["period"], spec["spd"], timesteps)


def load_posterior(directory, spec, ctx, inst, slots, **check):
    """State::load of every sample of a posterior directory into the device
    instances `inst` and parameter slots `slots` (one each): parameters set
    (one batch), counts replayed (populations, infectivity and means rebuilt
    on the device).  The counts of the steps from `timesteps` on are zero, so
    the replay leaves each instance's current state at the populations where
    the samples end -- the start of a prediction (Simulate::model_modification,
    src/simulate.cc:142-165, continues from there).  Returns (EF [n],
    parameter vectors [n][nparam], timesteps)."""
    n, ndiv = read_information(directory, **check)
    S = spec["period"] * spec["spd"]
    if len(inst) != n or len(slots) != n:
        raise ValueError(f"the posterior holds {n} samples")
    if not 0 < ndiv < S:
        raise ValueError("the samples must end before the model's last step")
    EF = np.zeros(n)
    pvs = np.zeros((n, len(spec["params"])))
    tns = []
    for k in range(n):
        EF[k], pvs[k], tn = read_sample(os.path.join(directory, f"sample{k}"), spec, ndiv)
        tns.append(tn)
    ctx.set_param(np.asarray(slots, dtype=np.int32), pvs)
    for k in range(n):
        ctx.put_transnum(int(inst[k]), int(slots[k]), tns[k])
    return EF, pvs, ndiv
