"""Model specification: the content of Data / Model once BEEP's TOML is read.

A spec is a plain dict (JSON-serialisable once arrays are lists):

  A, G          areas, demographic groups (dp; the age groups unless democats)
  nage, dp_age  age groups and each dp's age group (absent: dp == age group)
  period, spd   days, steps per day (details.division_per_time)
  comps         [{"name", "inf": param or -1, "mean": [G] params or None, "shape": int}]
                after Erlang expansion (src/inputs_comps.cc:33-46)
  trans         [{"from", "to", "infection": bool, "prob": [G] params or None}]
                internal Erlang transitions first (src/inputs_comps.cc:77-88)
  params        [{"name", "prior": ("fixed", v) | ("uniform", lo, hi) | ("exp", rate)
                  | ("normal", mu, sd) | ("dirflat",) | ("gamma", shape, rate)}]
  splines       [{"t": [...], "param": [...], "multfac": [...], "factor": param or -1}]
  R_spline, Rfactor, efoi_spline, efoi_agedist, geo_spline
  N             [nage][nage] age contact table or None (unit matrix for one age group)
  M             {"diag": [A], "rowptr": [A+1], "to": [nnz], "val": [nnz]} raw, or None for A == 1
  pop           [A][C][G] initial populations
  covars        [{"param", "value": [A][period]}] (centred on load)
  area_effect   [A] params multiplying each area's factor, or None
  level_effect  {"param": [nlevel], "map": [period][A] level index}, or None
  age_mod       [{"spline", "ages": [...]}]: at step s the listed age groups' rows and
                columns of the normalised contact matrix are multiplied by the
                spline's value (age_mixing_modify / age_mixing_perturb)
  sus           [G] params or None
  sus_extra     [[G] params]: further susceptibility factors (one per category)
  strain_Rfactor [nstrain] params (-1 = 1), or absent for one strain; strains are the
                last category (dp = strain * G / nstrain + dp within the strain)
  dirichlet     [{"param": [...], "frac": [...]}]
  obs           [{"y", "sd", "type", "shape", "factor", "thresh", "percent", "datatable",
                  "graph", "sett_i", "sett_f", "area": [...], "dp": [...], "tr": [...], "co": [...],
                  "spline": obs_spline weighting each step's transition counts, or -1}]
  obs_table     a compact alternative to obs for many one-area observations sharing
                their dp and transition lists: {"y", "sd", "sett_i", "sett_f", "area",
                "graph": arrays [nobs]; "type": one obs type; "datatable": int;
                "dp": [...], "tr": [...]}.  `expand_obs` turns it into obs entries.
  sections      [(ti, tf), ...] PMCMC observation sections

`param_index(spec, name)` finds parameters by name.
"""
import numpy as np

PRIOR_CODES = {"fixed": 0, "uniform": 1, "exp": 2, "normal": 3, "dirflat": 4, "gamma": 5}
OBS_CODES = {"normal": 0, "normal_percent": 1, "poisson": 2, "negbin": 3}
OBS_UNKNOWN = 999999997.0
OBS_THRESH = 999999998.0


def param_index(spec, name):
    for i, p in enumerate(spec["params"]):
        if p["name"] == name:
            return i
    raise KeyError(name)


def nsteps(spec):
    return spec["period"] * spec["spd"]


def dims(spec):
    return dict(A=spec["A"], G=spec["G"], C=len(spec["comps"]), T=len(spec["trans"]), S=nsteps(spec))


def sample_prior(spec, rng, n):
    """Draws n parameter vectors from the priors (uniform / fixed / exp / normal;
    flat-Dirichlet weights as Gamma(1,1)), like Model::sample_from_prior."""
    P = len(spec["params"])
    out = np.empty((n, P))
    for i, p in enumerate(spec["params"]):
        pr = p["prior"]
        kind = pr[0]
        if kind == "fixed":
            out[:, i] = pr[1]
        elif kind == "uniform":
            out[:, i] = rng.uniform(pr[1], pr[2], n)
        elif kind == "exp":
            out[:, i] = rng.exponential(1.0 / pr[1], n)
        elif kind == "normal":
            out[:, i] = rng.normal(pr[1], pr[2], n)
        elif kind == "dirflat":
            out[:, i] = rng.gamma(1.0, 1.0, n)
        elif kind == "gamma":
            out[:, i] = rng.gamma(pr[1], 1.0 / pr[2], n)
        else:
            raise ValueError(kind)
    return out


def prior_mid(spec):
    """A parameter vector at the middle of each prior (fixed values kept)."""
    v = []
    for p in spec["params"]:
        pr = p["prior"]
        if pr[0] == "fixed":
            v.append(pr[1])
        elif pr[0] == "uniform":
            v.append(0.5 * (pr[1] + pr[2]))
        elif pr[0] == "exp":
            v.append(1.0 / pr[1])
        elif pr[0] == "normal":
            v.append(pr[1])
        elif pr[0] == "gamma":
            v.append(pr[1] / pr[2])
        else:
            v.append(1.0)
    return np.array(v, dtype=np.float64)


def expand_obs(spec):
    """The spec with an obs_table turned into the equivalent obs entries (a
    shallow copy; specs without a table are returned as they are)."""
    tab = spec.get("obs_table")
    if tab is None:
        return spec
    out = dict(spec)
    dp, tr = list(tab["dp"]), list(tab["tr"])
    out["obs"] = [{"y": float(y), "sd": float(sd), "type": tab["type"], "datatable": int(tab.get("datatable", 0)),
                   "graph": int(gr), "sett_i": int(a), "sett_f": int(b), "area": [int(c)], "dp": dp, "tr": tr}
                  for y, sd, gr, a, b, c in zip(tab["y"], tab["sd"], tab["graph"], tab["sett_i"], tab["sett_f"],
                                                tab["area"])]
    del out["obs_table"]
    return out


def obs_elements(spec):
    """Per observation, the (kind, s, area, x, dp) elements it sums in the
    reference's accumulation order (src/obsmodel.cc:74-109): kind 0 = transnum
    (x = transition), kind 1 = pop (x = compartment)."""
    out = []
    for o in spec.get("obs", []):
        el = []
        for s in range(o["sett_i"], o["sett_f"]):
            for c in o["area"]:
                for tr in o.get("tr", []):
                    for dp in o["dp"]:
                        el.append((0, s, c, tr, dp))
                for co in o.get("co", []):
                    for dp in o["dp"]:
                        el.append((1, s, c, co, dp))
        out.append(el)
    return out
