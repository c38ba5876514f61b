"""BEEP model files (TOML + CSV tables) -> model spec for the device path.

Reads the part of the reference's input format that the device path models
and returns the spec dict of `beep_amd.spec`, the parameter values given in
the file, and what was recorded without effect.  The rules follow the
reference's readers:

  time              src/details.cc:50-106 (time_format, start/end, gettime,
                    steps_per_unit_time defaults to 2)
  comps / trans     src/inputs_comps.cc:10-140,270-365 (Erlang stages named
                    E, E1, ...; internal stage transitions first; a transition
                    leaves the last stage of its source and enters the first
                    stage of its destination)
  parameters        src/inputs.cc:1146-1205,1284-1326, src/model.cc:836-935
                    (value / prior / param strings, '|' lists, a single value
                    repeated, default names, Fixed(value) when only a value
                    is given; one parameter per distinct name)
  splines           src/inputs.cc:1363-1474, 872-910 (bp defaults to
                    start|end, times relative to start, factor -> multfac)
  [col:file] values src/inputs.cc:1499-1551
  initial state     src/data.cc:288-332,419-457 (everyone in the infection
                    transition's source compartment, rounded to the nearest
                    integer)
  data tables       src/inputs.cc:273-475, src/data.cc:1119-1353,1981-2060,
                    1589-1623, src/model.cc:1381-1425, src/utils.cc:575-583
                    (rows in the analysis period, transition counts over
                    consecutive windows of `timestep` days, populations at
                    the start of a day, observations named 'X->Y' or 'X' by
                    base compartment name, '.' unknown, '*' below threshold)

Supported: one or more age groups (`ages`, `age_mixing_matrix`, `age:` parameter lists,
`democats_dep="age"` tables) and areas (`areas` with population columns, a
`geo_mixing_matrix` for more than one, `geo_mixing_modify`, `geo_dep` tables),
`comps` with dist Exp / Erlang, `trans` with branching probabilities,
`R_spline` (one entry, or one per area by `geo_filt`) and `efoi_spline`
(one entry, or one per area and/or strain, each with its own `age_dist`),
`area_covars`, `tv_covars`
and `area_tv_covars` (func log or linear), `area_effect` and `level_effect`
(fixed or MDir(k) priors), `age_mixing_modify` (row_column) and
`age_mixing_perturb` (a contact matrix that varies by step), `obs_spline`
(named by a table's `factor_spline`), `democats`, `init_pop` and `strains`,
`data_tables` of type
'transition' or 'population' with obsmodel 'normal <sd>', 'normal <P>%',
'normal load' or 'poisson'.  Presentation keys are ignored.  `modification`,
`prediction_start` and `prediction_end` act only in the reference's
prediction mode (src/details.cc:80-92, src/model.cc:1044-1049,1349-1356,
1445-1585, src/state.cc:415,446): load(..., prediction=True) reads them
(every modification type, with geo_filt and democats_filt; the period
extended to prediction_end, the parameter splines extended to it),
otherwise they are only recorded in the returned metadata, as is
`prob_reach`, an output-only derived quantity.  Everything else raises
TomlError naming the key.
"""
import csv
import datetime
import math
import os

import numpy as np

try:
    import tomllib as _toml
except ImportError:          # Python < 3.11
    import tomli as _toml

OBS_UNKNOWN = 999999997.0
OBS_THRESH = 999999998.0

IGNORED_KEYS = {"description", "outputdir", "state_outputs", "plot_param_values", "time_labels", "area_plot",
                "output_prop", "mode"}
PREDICTION_KEYS = {"modification", "prediction_start", "prediction_end"}
# Quantities the reference derives from the posterior for its outputs only
# (Model::calculate_probreach, src/model.cc:1932-1957,2140-2160).
DERIVED_KEYS = {"prob_reach"}
SUPPORTED_KEYS = {"time_format", "start", "end", "datadir", "steps_per_unit_time", "comps", "trans", "R_spline",
                  "efoi_spline", "areas", "data_tables", "threshold_str", "nodata_str", "geo_mixing_matrix",
                  "geo_mixing_modify", "ages", "age_mixing_matrix", "area_covars", "tv_covars", "area_tv_covars",
                  "area_effect", "level_effect", "age_mixing_modify", "age_mixing_perturb", "obs_spline",
                  "democats", "init_pop", "strains"}


class TomlError(ValueError):
    """A model file the device path cannot take (unsupported key or bad value)."""


def _s(v):
    """The reference reads every field as a string (InputNode::stringfield)."""
    if v is None:
        return ""
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, float) and v.is_integer():
        return str(int(v))
    return str(v).strip()


def _split(s, d):
    return [] if s == "" else [x.strip() for x in s.split(d)]


def _col(head, name):
    """Index of a table column, headings compared without case
    (Data::find_column_noerror, src/data.cc:2080-2086); None when absent."""
    low = name.lower()
    return next((i for i, h in enumerate(head) if h.lower() == low), None)


class _Reader:
    def __init__(self, doc, datadir, prediction=False):
        self.doc = doc
        self.datadir = datadir
        fmt = _s(doc.get("time_format", ""))
        if fmt == "number":
            self.time_col = "time"
        elif fmt == "year-month-day":
            self.time_col = "date"
        else:
            raise TomlError(f"time_format '{fmt}' is not supported (number or year-month-day)")
        self.fmt = fmt
        self.labels = {_s(t.get("name")): _s(t.get("time")) for t in doc.get("time_labels", [])}
        if "start" not in doc or "end" not in doc:
            raise TomlError("'start' and 'end' must be set")
        self.start = self._abs_time(_s(doc["start"]))
        self.end = self._abs_time(_s(doc["end"]))
        self.period = self.end - self.start
        if self.period <= 0:
            raise TomlError("'end' must be after 'start'")
        # Prediction mode simulates on to 'prediction_end' when that is later
        # (src/details.cc:80-92); 'end' keeps naming the end of the data.
        self.data_period = self.period
        self.prediction = prediction
        if prediction and "prediction_end" in doc:
            self.period = max(self.period, self._abs_time(_s(doc["prediction_end"])) - self.start)
        spd = _s(doc.get("steps_per_unit_time", "2"))      # src/details.cc:97
        self.spd = int(spd)
        if self.spd <= 0 or str(self.spd) != spd:
            raise TomlError(f"steps_per_unit_time '{spd}' must be a positive integer")

    # --- time (Details::gettime, src/details.cc:130-175) ---
    def _abs_time(self, st):
        if st in self.labels:
            st = self.labels[st]
        if self.fmt == "number":
            try:
                return int(st)
            except ValueError:
                raise TomlError(f"time '{st}' is not an integer") from None
        try:
            return datetime.date.fromisoformat(st).toordinal()
        except ValueError:
            raise TomlError(f"time '{st}' is not in year-month-day format") from None

    def time(self, st):
        """Time relative to start, as 'start', 'end', a label or a value."""
        if st == "start":
            return 0
        if st == "end":
            return self.data_period
        return self._abs_time(st) - self.start

    # --- tables ---
    def table(self, name, heading=True):
        path = os.path.join(self.datadir, name)
        if not os.path.exists(path):
            raise TomlError(f"cannot open the file '{path}'")
        delim = "\t" if name.endswith(".txt") else ","
        with open(path, newline="") as f:
            rows = [[c.strip() for c in r] for r in csv.reader(f, delimiter=delim) if r]
        if not rows:
            raise TomlError(f"the file '{name}' contains no information")
        return (rows[0], rows[1:]) if heading else (None, rows)

    def from_table(self, s):
        """'[column:file.csv]' -> the column's values joined by '|' (src/inputs.cc:1499-1551)."""
        s = s.strip()
        if len(s) > 6 and s[0] == "[" and s[-1] == "]":
            spl = s[1:-1].split(":")
            if len(spl) != 2 or not (spl[1].endswith(".csv") or spl[1].endswith(".txt")):
                raise TomlError(f"'{s}' does not use the format '[column:file.csv]'")
            head, rows = self.table(spl[1])
            if spl[0] not in head:
                raise TomlError(f"the heading '{spl[0]}' in file '{spl[1]}' could not be found")
            c = head.index(spl[0])
            return "|".join(r[c] for r in rows)
        return s

    def split_n(self, s, n, what):
        """Inputs::split_string: n '|' items, one item repeated (src/inputs.cc:1554-1578)."""
        spl = _split(self.from_table(s), "|")
        if not spl:
            return [""] * n
        if len(spl) == 1 and n > 1:
            spl = spl * n
        if len(spl) != n:
            raise TomlError(f"{what}: '{s}' has {len(spl)} elements instead of {n}")
        return spl


def _prior(st, value, name):
    """A prior string as the spec's tuple (Model::add_parameter, src/model.cc:839-935)."""
    if st == "":
        if value == "":
            raise TomlError(f"parameter '{name}': a prior or a value must be set")
        st = f"Fixed({value})"
    if "(" not in st or not st.endswith(")"):
        raise TomlError(f"parameter '{name}': prior '{st}' is not recognised")
    kind, args = st[:-1].split("(", 1)
    vals = _split(args, ",")
    kind = kind.strip().lower()
    try:
        if kind == "fixed" and len(vals) == 1:
            return ("fixed", float(vals[0]))
        if kind == "uniform" and len(vals) == 2:
            return ("uniform", float(vals[0]), float(vals[1]))
        if kind == "exp" and len(vals) == 1:
            return ("exp", float(vals[0]))
        if kind == "normal" and len(vals) == 2:
            return ("normal", float(vals[0]), float(vals[1]))
    except ValueError:
        pass
    # Dir / MDir priors tie parameters into Dirichlet groups
    # (src/model.cc:2900-3090), which this reader does not build.
    raise TomlError(f"parameter '{name}': prior '{st}' is not supported")


class _Params:
    """Parameters by name, in order of first use (one per distinct name)."""

    def __init__(self):
        self.specs, self.values, self.index = [], [], {}

    def add(self, name, value, prior):
        if name == "":
            raise TomlError("a parameter must have a name")
        pr = _prior(prior, value, name)
        try:
            v = float(value) if value != "" else math.nan
        except ValueError:
            raise TomlError(f"parameter '{name}': the value '{value}' is not supported") from None
        return self.add_tuple(name, v, pr)

    def add_tuple(self, name, v, pr):
        if name in self.index:
            i = self.index[name]
            if self.specs[i]["prior"] != pr or not (self.values[i] == v or (math.isnan(v) and math.isnan(self.values[i]))):
                raise TomlError(f"parameter '{name}' is given two different definitions")
            return i
        self.index[name] = len(self.specs)
        self.specs.append({"name": name, "prior": pr})
        self.values.append(v)
        return self.index[name]


def _no_dep(entry, keys, where):
    for k in keys:
        v = _s(entry.get(k, ""))
        if ":" in v.split("[")[0]:
            raise TomlError(f"{where}: the dependency in '{k}={v}' is not supported")


class _Demo:
    """Demographic categories (Inputs::find_demographic_category,
    src/inputs.cc:616-660): age first, then the 'democats' entries.  dp
    enumerates their combinations with the first category varying fastest
    (src/data.cc:55-73); dp_age is each dp's age group."""

    def __init__(self, cats):
        self.cats = cats                       # [(name, [values])]
        self.dps = []
        idx = [0] * len(cats)
        while True:
            self.dps.append(tuple(idx))
            k = 0
            while k < len(cats):
                idx[k] += 1
                if idx[k] < len(cats[k][1]):
                    break
                idx[k] = 0
                k += 1
            if k == len(cats):
                break
        self.G = len(self.dps)
        self.ages = cats[0][1]
        self.nage = len(self.ages)
        self.dp_age = [t[0] for t in self.dps]

    def find(self, name):
        """A category by name, without case (Data::get_datafilter, src/data.cc:1491-1493)."""
        return next((k for k, (n, _) in enumerate(self.cats) if n.lower() == name.lower()), None)


def _ct_specs(rd, entry, keys, par_root, demo, where):
    """Per dp, the (name, value, prior) strings of a compartment or transition
    parameter (Inputs::get_dep, find_demo_ref and find_ct_param_spec,
    src/inputs.cc:1107-1143,1217-1326): 'cat: v0 | v1 | ...' (or 'c1,c2: ...',
    the first category varying fastest) gives one parameter per value, named
    '<value> <root>' unless named; without a dependency one parameter serves
    every dp."""
    strs, dep = [], ""
    for k in keys:
        st = _s(entry.get(k, ""))
        spl = st.split("[")[0].split(":")
        if len(spl) > 2:
            raise TomlError(f"{where}: there was a problem with '{st}'")
        if len(spl) == 2:
            if "[" in st:
                raise TomlError(f"{where}: a dependency with a table value ('{st}') is not supported")
            d = spl[0].strip()
            st = spl[1].strip()
            if dep not in ("", d):
                raise TomlError(f"{where}: conflicting dependencies")
            dep = d
        strs.append(st)
    if dep == "":
        return [tuple(strs)] * demo.G
    deps = []
    for d in (x.strip() for x in dep.split(",")):
        k = next((k for k, (n, _) in enumerate(demo.cats) if n == d), None)
        if k is None:
            raise TomlError(f"{where}: the dependency '{d}' cannot be found")
        if len(demo.cats[k][1]) == 1:
            raise TomlError(f"{where}: the dependency '{dep}:' cannot be used because '{d}' only contains one value")
        if k in deps:
            raise TomlError(f"{where}: the value '{dep}' contains multiple dependencies")
        deps.append(k)
    combos = [()]
    for k in deps:                                 # deps[0] varies fastest
        combos = [c + (v,) for v in range(len(demo.cats[k][1])) for c in combos]
    n = len(combos)
    names = rd.split_n(strs[0], n, where)
    vals = rd.split_n(strs[1], n, where)
    pris = rd.split_n(strs[2], n, where)
    out = []
    for t in demo.dps:
        i = combos.index(tuple(t[k] for k in deps))
        dflt = ",".join(demo.cats[k][1][combos[i][j]] for j, k in enumerate(deps)) + " " + par_root
        out.append((names[i] or dflt, vals[i], pris[i]))
    return out


def _check_keys(entry, allowed, where):
    for k in entry:
        if k not in allowed:
            raise TomlError(f"{where}: '{k}' is not supported")


def _compartments(rd, doc, P, demo):
    """comps -> expanded compartments (src/inputs_comps.cc:22-63,303-365)."""
    if "comps" not in doc:
        raise TomlError("the file must define 'comps'")
    comps = []          # {"base", "num", "shape", "mean", "inf", "name"}
    bases = []
    for e in doc["comps"]:
        _check_keys(e, {"name", "dist", "k", "mean_param", "mean_value", "mean_prior", "inf_param", "inf_value",
                        "inf_prior", "plot_name", "line_colour"}, "comps")
        nam = _s(e.get("name"))
        if nam == "":
            raise TomlError("in 'comps' every compartment needs a 'name'")
        if nam in bases:
            raise TomlError(f"the compartment '{nam}' cannot be defined multiple times")
        bases.append(nam)
        dist = _s(e.get("dist", ""))
        if dist == "":
            k = None
        elif dist == "Exp":
            k = 1
        elif dist == "Erlang":
            ks = _s(e.get("k", ""))
            if not ks.isdigit() or int(ks) < 1:
                raise TomlError(f"compartment '{nam}': Erlang needs an integer shape 'k'")
            k = int(ks)
        else:
            raise TomlError(f"compartment '{nam}': dist '{dist}' is not recognised")
        mean = None
        if any(x in e for x in ("mean_param", "mean_value", "mean_prior")):
            specs = _ct_specs(rd, e, ("mean_param", "mean_value", "mean_prior"), f"{nam} mean occupancy", demo,
                              f"compartment '{nam}'")
            mean = [P.add(nm or f"{nam} mean occupancy", v, pr) for nm, v, pr in specs]
        inf = -1
        _no_dep(e, ("inf_param", "inf_value", "inf_prior"), f"compartment '{nam}'")
        if any(x in e for x in ("inf_param", "inf_value", "inf_prior")):
            pname = _s(e.get("inf_param", "")) or f"{nam} Infectivity"
            inf = P.add(pname, _s(e.get("inf_value", "")), _s(e.get("inf_prior", "")))
        stages = 1 if k is None else k
        for num in range(stages):
            comps.append({"base": nam, "num": num, "shape": k, "mean": mean, "inf": inf,
                          "name": nam + (str(num) if num > 0 else "")})
    return comps


def _transitions(rd, doc, comps, P, demo):
    """trans -> internal Erlang transitions first, then the file's (src/inputs_comps.cc:66-139,268-300)."""
    if "trans" not in doc:
        raise TomlError("the file must define 'trans'")
    trans = []
    for c, co in enumerate(comps):
        if co["shape"] is not None and co["num"] < co["shape"] - 1:
            trans.append({"from": c, "to": c + 1, "infection": False, "prob": None})

    def find(name, last):
        for c, co in enumerate(comps):
            if co["base"] == name and (co["shape"] is None or co["num"] == ((co["shape"] - 1) if last else 0)):
                return c
        return None

    ninf = 0
    for e in doc["trans"]:
        _check_keys(e, {"from", "to", "infection", "prob_param", "prob_value", "prob_prior"}, "trans")
        fr, to = _s(e.get("from")), _s(e.get("to"))
        if fr == to:
            raise TomlError("in 'trans' the 'from' and 'to' values cannot be the same")
        cf, ct = find(fr, True), find(to, False)
        if cf is None:
            raise TomlError(f"in 'trans' cannot find the 'from' compartment '{fr}'")
        if ct is None:
            raise TomlError(f"in 'trans' cannot find the 'to' compartment '{to}'")
        inf = _s(e.get("infection", ""))
        if inf not in ("", "yes", "no"):
            raise TomlError("in 'trans' the value of 'infection' can be 'yes', 'no' or omitted")
        branch = None
        if any(x in e for x in ("prob_param", "prob_value", "prob_prior")):
            root = f"{fr}->{to} branch prob."
            branch = [(nm or root, v, pr) for nm, v, pr in
                      _ct_specs(rd, e, ("prob_param", "prob_value", "prob_prior"), root, demo, f"transition {fr}->{to}")]
        ninf += inf == "yes"
        trans.append({"from": cf, "to": ct, "infection": inf == "yes", "prob": None, "base": f"{fr}->{to}",
                      "branch": branch})
    if ninf != 1:
        raise TomlError("exactly one infection transition must exist")
    return trans, _branches(trans, comps, P, demo.G)


def _dir_kind(prior):
    """A branch prior's Dirichlet kind (Model::add_parameter, src/model.cc:
    901-930): 'flat' for Dir(*) (density -x), ('alpha', a) for Dir(a) (a
    gamma(a, 1) density), ('meansd', mean, sd) for Dir(mean,sd) (strings,
    either may be '*'), None otherwise."""
    t = prior.replace(" ", "").lower()
    if not (t.startswith("dir(") and t.endswith(")")):
        return None
    a = t[4:-1]
    if a == "*":
        return "flat"
    if "," in a:
        m, sd = a.split(",", 1)
        return ("meansd", m, sd)
    try:
        return ("alpha", float(a))
    except ValueError:
        raise TomlError(f"the prior '{prior}' is not recognised") from None


def _dir_meansd_alpha(means, sds, where):
    """Dir(mean,sd) priors of a branch group (src/model.cc:2936-3020, each
    weight 1): one mean may be '*' (one minus the others); exactly one sd is
    set, and with that branch's mean m and sd s, alpha0 = m(1-m)/s^2 - 1;
    branch i gets a gamma(alpha0 * mean_i, 1) prior."""
    star = [i for i, m in enumerate(means) if m == "*"]
    if len(star) != 1:
        raise TomlError(f"{where}: exactly one Dirichlet mean must be '*'")
    try:
        mv = [math.nan if m == "*" else float(m) for m in means]
    except ValueError:
        raise TomlError(f"{where}: a Dirichlet mean is not a number") from None
    rest = sum(m for i, m in enumerate(mv) if i != star[0])
    if rest > 1:
        raise TomlError(f"{where}: the Dirichlet mean '*' cannot be set")
    mv[star[0]] = 1.0 - rest
    given = [i for i, x in enumerate(sds) if x != "*"]
    if len(given) != 1:
        raise TomlError(f"{where}: exactly one Dirichlet standard deviation must be set")
    try:
        sd = float(sds[given[0]])
    except ValueError:
        raise TomlError(f"{where}: the Dirichlet standard deviation is not a number") from None
    mu = mv[given[0]]
    alpha0 = mu * (1 - mu) / (sd * sd) - 1
    if not alpha0 > 0:
        raise TomlError(f"{where}: the Dirichlet standard deviation is too large for its mean")
    return [alpha0 * m / sum(mv) for m in mv]


def _branches(trans, comps, P, G):
    """Branch probabilities of the transitions leaving one compartment, per
    age group: one value may be '*', one minus the others (src/model.cc:
    2894-2915); with Dir(*) or Dir(alpha) priors they form a Dirichlet group that
    the model normalises (Model::dirichlet_correct, src/model.cc:2685-2695).
    Sets t["prob"] to the G parameter indices of each branching transition."""
    groups = []
    by_src = {}
    for t in trans:
        t["prob"] = [None] * G if t.get("branch") is not None else None
        if t["prob"] is not None:
            by_src.setdefault(t["from"], []).append(t)
    for src, ts in by_src.items():
        where = f"branch probabilities out of '{comps[src]['base']}'"
        for g in range(G):
            priors = [t["branch"][g][2] for t in ts]
            if any(p.replace(" ", "").lower().startswith("mdir(") for p in priors):
                raise TomlError(f"{where}: MDir priors are not used for branch probabilities")
            kinds = [_dir_kind(p) for p in priors]
            dirs = [k is not None for k in kinds]
            alpha = [k[0] == "alpha" if isinstance(k, tuple) else False for k in kinds]
            meansd = [k[0] == "meansd" if isinstance(k, tuple) else False for k in kinds]
            # All fixed, all Dir(*), all Dir(a) or all Dir(mean,sd) (src/model.cc:2924-2931).
            if any(dirs) and not (all(k == "flat" for k in kinds) or all(alpha) or all(meansd)):
                raise TomlError(f"{where}: the priors must all be fixed, all Dir(*), all Dir(alpha) or all Dir(mean,sd)")
            if all(meansd):
                kinds = [("alpha", a) for a in _dir_meansd_alpha([k[1] for k in kinds], [k[2] for k in kinds], where)]
            vals = [t["branch"][g][1] for t in ts]
            star = [k for k, v in enumerate(vals) if v == "*"]
            if len(star) > 1:
                raise TomlError(f"{where}: at most one value can be '*'")
            num = []
            for k, v in enumerate(vals):
                if k in star or v == "":
                    num.append(math.nan)
                else:
                    try:
                        num.append(float(v))
                    except ValueError:
                        raise TomlError(f"branch probability value '{v}' is not a number") from None
            if star:
                rest = sum(x for k, x in enumerate(num) if k != star[0])
                if math.isnan(rest) or rest > 1:
                    raise TomlError(f"{where} cannot fill the value '*'")
                num[star[0]] = 1.0 - rest
            if all(dirs):
                prs = [("dirflat",) if k == "flat" else ("gamma", k[1], 1.0) for k in kinds]
                idx = [P.add_tuple(t["branch"][g][0], num[k], prs[k]) for k, t in enumerate(ts)]
                if {"param": idx} not in groups:        # parameters shared by every age group: one group
                    groups.append({"param": idx})
            else:
                idx = [P.add(t["branch"][g][0], "" if math.isnan(num[k]) else repr(num[k]), t["branch"][g][2])
                       for k, t in enumerate(ts)]
            for k, t in enumerate(ts):
                t["prob"][g] = idx[k]
    return groups


def _modified_dirichlet(P, names, vals, priors, frac, where):
    """A modified Dirichlet group (Model::add_dirichlet with DIR_MODIFIED,
    src/model.cc:2699-2730,2838-2878,3025-3050): one parameter per distinct
    name, its weight the sum of its items' fractions.  Values are on the
    corrected scale: one must be '*', set so that sum(value * frac) = 1, and
    the returned values are value * frac (Model::simulated_values,
    src/model.cc:990-1003), which Model::dirichlet_correct maps back.  Priors
    are all fixed or all MDir(k), the latter giving each parameter a
    gamma(beta * frac, 1) prior with beta = (n - 1) / k^2 - 1.  Returns the
    parameter index of each item and the group."""
    order, info = [], {}
    for nm, v, pr, f in zip(names, vals, priors, frac):
        if nm in info:
            if info[nm][:2] != [v, pr]:
                raise TomlError(f"{where}: parameter '{nm}' has multiple definitions")
            info[nm][2] += f
        else:
            order.append(nm)
            info[nm] = [v, pr, f]
    n = len(order)
    mdir = [info[nm][1].replace(" ", "").lower().startswith("mdir(") for nm in order]
    fixed = [info[nm][1] == "" for nm in order]
    if not (all(mdir) or all(fixed)):
        raise TomlError(f"{where}: the priors must all be MDir(k) or all fixed (a value and no prior)")
    raw = [math.nan] * n
    if info[order[0]][0] != "":
        star = [i for i, nm in enumerate(order) if info[nm][0] == "*"]
        if len(star) != 1:
            raise TomlError(f"{where}: exactly one of the values must be '*'")
        tot = 0.0
        for i, nm in enumerate(order):
            if i == star[0]:
                continue
            try:
                raw[i] = float(info[nm][0]) * info[nm][2]
            except ValueError:
                raise TomlError(f"{where}: the value '{info[nm][0]}' is not a number") from None
            tot += raw[i]
        if tot > 1:
            raise TomlError(f"{where}: the value '*' cannot be set (the weighted values sum to more than one)")
        raw[star[0]] = 1.0 - tot
    if all(mdir):
        ks = {info[nm][1].replace(" ", "")[5:-1] for nm in order}
        try:
            kappa = float(info[order[0]][1].replace(" ", "")[5:-1])
        except ValueError:
            raise TomlError(f"{where}: the prior '{info[order[0]][1]}' is not recognised") from None
        if len(ks) != 1:
            raise TomlError(f"{where}: the MDir priors must share one argument")
        beta = (n - 1) / (kappa * kappa) - 1
        prs = []
        for nm in order:
            if beta * info[nm][2] < 0.1:
                raise TomlError(f"{where}: the prior '{info[nm][1]}' is too diffuse")
            prs.append(("gamma", beta * info[nm][2], 1.0))
    else:
        if any(math.isnan(r) for r in raw):
            raise TomlError(f"{where}: a value must be set for every parameter")
        prs = [("fixed", r) for r in raw]
    idx = {nm: P.add_tuple(nm, raw[i], prs[i]) for i, nm in enumerate(order)}
    return [idx[nm] for nm in names], {"param": [idx[nm] for nm in order], "frac": [info[nm][2] for nm in order]}


def _effects(rd, doc, P, area_head, area_rows, pop):
    """area_effect (Inputs::find_area_effect, src/inputs.cc:1075-1102: one
    parameter per area, default names 'Area effect <code>', weights the areas'
    population shares) and level_effect (Inputs::find_level_effect,
    src/inputs.cc:746-768; Data::read_level_effect, src/data.cc:571-623: a
    table of level names per day and area, each row holding from its day
    until the next row's, level 0 before the first row; weights the levels'
    shares of area-days).  Both form modified Dirichlet groups.  Returns the
    area effect parameters, the level effect and the groups."""
    A, period = len(area_rows), rd.period
    codes = [r[_col(area_head, "area")] for r in area_rows]
    groups, area_eff, level = [], None, None
    if "area_effect" in doc:
        e = doc["area_effect"]
        _check_keys(e, {"param", "value", "prior"}, "area_effect")
        _no_dep(e, ("param", "value", "prior"), "area_effect")
        names = rd.split_n(_s(e.get("param", "")), A, "area_effect")
        if not _s(e.get("param", "")):
            names = ["Area effect " + c for c in codes]
        vals = rd.split_n(_s(e.get("value", "")), A, "area_effect")
        priors = rd.split_n(_s(e.get("prior", "")), A, "area_effect")
        tot = [sum(sum(g) for g in pa) for pa in pop]
        frac = [t / sum(tot) for t in tot]
        area_eff, grp = _modified_dirichlet(P, names, vals, priors, frac, "area_effect")
        groups.append(grp)
    if "level_effect" in doc:
        e = doc["level_effect"]
        _check_keys(e, {"file", "param", "value", "prior"}, "level_effect")
        names = _split(_s(e.get("param", "")), "|")
        if not names:
            raise TomlError("in 'level_effect' a value for 'param' must be set")
        n = len(names)
        vals = rd.split_n(_s(e.get("value", "")), n, "level_effect")
        priors = rd.split_n(_s(e.get("prior", "")), n, "level_effect")
        head, rows = rd.table(_s(e.get("file", "")))
        tc = _col(head, rd.time_col)
        if tc is None:
            raise TomlError(f"the level effect file needs a column '{rd.time_col}'")
        cols = [_col(head, c) for c in codes]
        if None in cols:
            raise TomlError(f"the level effect file has no column '{codes[cols.index(None)]}'")
        lmap, cur, t = [], [0] * A, 0
        for r in range(len(rows) + 1):
            t_new = period if r == len(rows) else rd.time(rows[r][tc])
            if t_new < t or t_new > period:
                raise TomlError("in the level effect file the dates are out of the analysis range")
            lmap += [list(cur)] * (t_new - t)
            t = t_new
            if r == len(rows):
                break
            for c in range(A):
                v = rows[r][cols[c]]
                if v not in names:
                    raise TomlError(f"in 'level_effect' the value '{v}' is not contained in 'param'")
                cur[c] = names.index(v)
        cnt = np.bincount(np.asarray(lmap).reshape(-1), minlength=n)
        idx, grp = _modified_dirichlet(P, names, vals, priors, [float(x) / (period * A) for x in cnt], "level_effect")
        level = {"param": idx, "map": lmap}
        groups.append(grp)
    return area_eff, level, groups


def _spline(rd, P, key, entry, default_name=None, params=None, suffix=""):
    """One spline entry (src/inputs.cc:1363-1474): knots, parameters, multfac.
    default_name(i, bp) names knot i's parameter when 'param' is not given;
    params, when given, are the knots' parameter indices, already added;
    suffix (' <area>', ' <strain>') follows the default name's root."""
    if default_name is None:
        _check_keys(entry, {"name", "bp", "param", "value", "prior", "factor", "factor_param", "factor_value",
                            "factor_prior", "smooth", "smooth_type"} |
                    ({"age_dist", "geo_filt", "strain"} if key in ("efoi_spline", "R_spline") else set()), key)
    bp = _bp(rd, entry, key, extend=params is None)
    if any(b > a for a, b in zip(bp[1:], bp[:-1])):
        raise TomlError(f"in '{key}' times in 'bp' are not in the right order")
    n = len(bp)
    fac = rd.split_n(_s(entry.get("factor", "")) or "1", n, key)
    sm = rd.split_n(_s(entry.get("smooth", "")), n, key)
    if params is None:
        names = rd.split_n(_s(entry.get("param", "")), n, key)
        vals = rd.split_n(_s(entry.get("value", "")), n, key)
        pris = rd.split_n(_s(entry.get("prior", "")), n, key)
        if bp[-1] < rd.period:
            # Inputs::extend_spline_for_prediction (src/inputs.cc:1475-1493):
            # the last knot moves to the prediction's end when the last two
            # knots are specified alike, else a copy of the last is added.
            lists = (names, vals, pris, fac, sm)
            if n >= 2 and all(x[-2] == x[-1] for x in lists):
                bp[-1] = rd.period
            else:
                bp.append(rd.period)
                for x in lists:
                    x.append(x[-1])
                n += 1
        params = []
        for i in range(n):
            nm = names[i]
            if nm == "":
                if default_name is not None:
                    nm = default_name(i, bp)
                else:
                    nm = key.replace("_", "-") + suffix
                    if n > 2:
                        nm += f" t={bp[i]}"
            params.append(P.add(nm, vals[i], pris[i]))
    factor = -1
    if any(x in entry for x in ("factor_param", "factor_value", "factor_prior")):
        factor = P.add(_s(entry.get("factor_param", "")) or f"{key} factor", _s(entry.get("factor_value", "")),
                       _s(entry.get("factor_prior", "")))
    out = {"t": [float(b) for b in bp], "param": params, "multfac": [float(f) for f in fac], "factor": factor}
    # Smoothing prior between neighbouring knots (Inputs::find_smoothtype,
    # src/inputs.cc:1343-1359; Model::spline_prior, src/model.cc:1662-1704).
    st = _s(entry.get("smooth_type", ""))
    kinds = {"": 0, "no_smooth": 0, "smooth": 1, "log_smooth": 2}
    if st not in kinds:
        raise TomlError(f"in '{key}' the value of 'smooth_type={st}' is not recognised")
    if kinds[st]:
        try:
            smv = [float(x) if x != "" else math.nan for x in sm]
        except ValueError:
            raise TomlError(f"in '{key}' the smoothing values '{entry.get('smooth')}' are not numbers") from None
        for i in range(n - 1):
            pa, pb = P.specs[params[i]]["prior"], P.specs[params[i + 1]]["prior"]
            if params[i] != params[i + 1] and bp[i] != bp[i + 1] and not (pa[0] == "fixed" and pb[0] == "fixed") \
                    and math.isnan(smv[i]):
                raise TomlError(f"for the spline '{key}' a value for 'smooth' must be set")
        out.update(smooth_type=kinds[st], smooth=smv)
    return out


def _bp(rd, entry, key, extend=False):
    """Knot times of a spline entry (Inputs::get_breakpoints, src/inputs.cc:872-905):
    from 'start' to 'end'; in prediction mode the last may end before the
    prediction's end when the spline is extended there (extend)."""
    bp = [rd.time(t) for t in _split(rd.from_table(_s(entry.get("bp", "")) or "start|end"), "|")]
    if not bp or bp[0] != 0:
        raise TomlError(f"in '{key}' 'bp' must start at 'start'")
    if any(b > rd.data_period for b in bp):
        raise TomlError(f"in '{key}' a time in 'bp' is outside of the analysis time period")
    if bp[-1] != rd.period and not (extend and rd.prediction):
        raise TomlError(f"in '{key}' 'bp' must end at 'end'" +
                        (" (the prediction's end: only splines of parameters are extended there)"
                         if rd.period != rd.data_period else ""))
    return bp


def _age_mods(rd, doc, P, ages, pop_age, splines):
    """age_mixing_modify (type 'row_column') and age_mixing_perturb
    (Inputs::find_genQ, src/inputs.cc:944-1011): one spline each, knot
    parameters named 'MatMod <agecat> factor t=<bp>' by default, that
    multiplies the rows and columns of its age groups in the contact matrix
    (Model::create_Ntime, src/model.cc:1254-1301).  The perturbations'
    parameters at each knot form a modified Dirichlet group weighted by the
    groups' population shares, and every age group must be perturbed exactly
    once (Model::set_dirichlet, src/model.cc:2786-2825).  Appends the splines
    and returns the modifications and the Dirichlet groups."""
    mods, groups = [], []
    G = len(ages)

    def agecats(e, key):
        if "bp" not in e:
            raise TomlError(f"in '{key}' a value for 'bp' must be set")
        st = _s(e.get("agecat", ""))
        out = []
        for a in _split(st, "|"):
            if a not in ages:
                raise TomlError(f"in '{key}' the value '{a}' in 'agecat' does not specify a valid age classification")
            out.append(ages.index(a))
        if not out:
            raise TomlError(f"in '{key}' a value for 'agecat' must be set")
        return st, out

    for e in doc.get("age_mixing_modify", []):
        _check_keys(e, {"type", "agecat", "bp", "param", "value", "prior", "smooth", "factor"}, "age_mixing_modify")
        if _s(e.get("type", "")) != "row_column":
            raise TomlError(f"in 'age_mixing_modify' the value of 'type={_s(e.get('type', ''))}' is not recognised")
        st, idx = agecats(e, "age_mixing_modify")
        splines.append(_spline(rd, P, "age_mixing_modify", e, lambda i, bp, st=st: f"MatMod {st} factor t={bp[i]}"))
        mods.append({"spline": len(splines) - 1, "ages": idx})
    pert = doc.get("age_mixing_perturb", [])
    if pert:
        share = pop_age / pop_age.sum()
        info, used = [], [0] * G
        for e in pert:
            _check_keys(e, {"agecat", "bp", "param", "value", "prior", "smooth", "smooth_type", "factor"},
                        "age_mixing_perturb")
            st, idx = agecats(e, "age_mixing_perturb")
            for a in idx:
                used[a] += 1
            bp = _bp(rd, e, "age_mixing_perturb")
            n = len(bp)
            names = [nm or f"MatMod {st} factor t={bp[i]}"
                     for i, nm in enumerate(rd.split_n(_s(e.get("param", "")), n, "age_mixing_perturb"))]
            info.append((e, idx, names, rd.split_n(_s(e.get("value", "")), n, "age_mixing_perturb"),
                         rd.split_n(_s(e.get("prior", "")), n, "age_mixing_perturb"),
                         float(sum(share[a] for a in idx))))
        if any(u != 1 for u in used):
            a = next(a for a in range(G) if used[a] != 1)
            raise TomlError(f"'age_mixing_perturb' must use age category '{ages[a]}' exactly once")
        nk = len(info[0][2])
        if any(len(x[2]) != nk for x in info):
            raise TomlError("every 'age_mixing_perturb' entry must have the same number of knots")
        params = [[None] * nk for _ in info]
        for p in range(nk):
            idx, grp = _modified_dirichlet(P, [x[2][p] for x in info], [x[3][p] for x in info],
                                           [x[4][p] for x in info], [x[5] for x in info],
                                           f"'age_mixing_perturb' at knot {p}")
            if grp not in groups:
                groups.append(grp)
            for j in range(len(info)):
                params[j][p] = idx[j]
        for j, (e, idx, _, _, _, _) in enumerate(info):
            splines.append(_spline(rd, P, "age_mixing_perturb", e, lambda i, bp: "", params[j]))
            mods.append({"spline": len(splines) - 1, "ages": idx})
    return mods, groups


def _regions(area_head, area_rows, geo_dep):
    """Data::create_geomap (src/data.cc:1830-1858): 'all', or the areas grouped
    by their value in column geo_dep ('NA' leaves an area out)."""
    if geo_dep == "all":
        return [("all", list(range(len(area_rows))))]
    col = _col(area_head, geo_dep)
    if col is None:
        raise TomlError(f"cannot find the column '{geo_dep}' in the areas file")
    out = {}
    for r, row in enumerate(area_rows):
        if row[col] != "NA":
            out.setdefault(row[col], []).append(r)
    return list(out.items())


def _covariates(rd, doc, P, area_head, area_rows):
    """area_covars / tv_covars / area_tv_covars (Inputs::find_covariates,
    src/inputs.cc:773-868; Data::read_covars, src/data.cc:461-556): the
    values per area and day, log-transformed when func='log'.  Centring is
    done by the host model.  A time-varying table's row r holds days t_r to
    t_(r+1) inclusive, the next row overwriting the shared day; the last
    row's own value is not used, as in the reference."""
    A, period = len(area_rows), rd.period
    out = []
    kinds = (("area_covars", "Covariate: "), ("tv_covars", "TV Covariate: "),
             ("area_tv_covars", "Area TV Covariate: "))
    for key, prefix in kinds:
        for e in doc.get(key, []):
            allowed = {"name", "param", "value", "prior", "func"} | (set() if key == "area_covars" else {"file"})
            _check_keys(e, allowed, key)
            name = _s(e.get("name", ""))
            if name == "":
                raise TomlError(f"in '{key}' every covariate needs a 'name'")
            func = _s(e.get("func", ""))
            if func not in ("log", "linear"):
                raise TomlError(f"in '{key}' the value 'func={func}' is not recognised")
            pname = _s(e.get("param", "")) or prefix + name
            par = P.add(pname, _s(e.get("value", "")), _s(e.get("prior", "")))
            val = np.zeros((A, period))
            if key == "area_covars":
                col = _col(area_head, name)
                if col is None:
                    raise TomlError(f"the areas file has no column '{name}' for the covariate")
                for c, row in enumerate(area_rows):
                    val[c, :] = float(row[col])
            else:
                head, rows = rd.table(_s(e.get("file", "")))
                tc = _col(head, rd.time_col)
                if tc is None:
                    raise TomlError(f"the covariate file '{e.get('file')}' needs a column '{rd.time_col}'")
                names = [name] if key == "tv_covars" else [r[_col(area_head, "area")] for r in area_rows]
                cols = [_col(head, nm) for nm in names]
                for nm, c in zip(names, cols):
                    if c is None:
                        raise TomlError(f"the covariate file '{e.get('file')}' has no column '{nm}'")
                times = [rd.time(r[tc]) for r in rows]
                if not rows:
                    raise TomlError(f"the covariate file '{e.get('file')}' has no rows")
                # One row: no interval, the values stay zero (the reference's
                # loop over row pairs does not run).
                if len(rows) > 1 and (times[0] > 0 or times[-1] < period - 1):
                    raise TomlError(f"the covariate file '{e.get('file')}' must cover the analysis period")
                for r in range(len(rows) - 1):
                    ti, tf = times[r], times[r + 1]
                    if tf <= ti:
                        raise TomlError(f"in the covariate file '{e.get('file')}' the times must increase")
                    lo, hi = max(ti, 0), min(tf, period - 1)
                    for c in range(A):
                        v = float(rows[r][cols[0 if key == "tv_covars" else c]])
                        val[c, lo:hi + 1] = v
            if func == "log":
                if (val <= 0).any():
                    raise TomlError(f"the values of covariate '{name}' must be positive for log transformation")
                val = np.log(val)
            out.append({"param": par, "value": val.tolist()})
    return out


def _initial_population(rd, doc, demo, comps, trans, head, rows, strain_k=None):
    """Initial populations [A][C][G] (Data::read_initial_population,
    src/data.cc:288-457).  From the areas file: everyone in the infection's
    source compartment, one column per age group times, for every other
    category, the fraction of its value in the area (Data::get_demo_frac,
    src/data.cc:821-870: one column per value, each divided by their sum),
    rounded to the nearest integer.  From an 'init_pop' file: one row per
    area, compartment and category value (columns 'area', 'compartment',
    'population' and one per category with more than one value), the number
    placed in the first stage of the compartment; every combination must be
    given."""
    A, G, C = len(rows), demo.G, len(comps)
    codes = [r[_col(head, "area")] for r in rows]
    if "init_pop" in doc:
        ph, prow = rd.table(_s(doc["init_pop"]))
        cols = {k: _col(ph, k) for k in ("area", "compartment", "population")}
        for k, c in cols.items():
            if c is None:
                raise TomlError(f"the 'init_pop' file needs a column '{k}'")
        dcol = [(_col(ph, n) if len(v) > 1 else None) for n, v in demo.cats]
        for (n, v), c in zip(demo.cats, dcol):
            if len(v) > 1 and c is None:
                raise TomlError(f"the 'init_pop' file needs a column '{n}'")
        pop = [[[None] * G for _ in range(C)] for _ in range(A)]
        for r in prow:
            if r[cols["area"]] not in codes:
                raise TomlError(f"in the 'init_pop' file the area '{r[cols['area']]}' is not recognised")
            a = codes.index(r[cols["area"]])
            t = []
            for (n, v), c in zip(demo.cats, dcol):
                if c is None:
                    t.append(0)
                elif r[c] not in v:
                    raise TomlError(f"in the 'init_pop' file '{r[c]}' is not a value of '{n}'")
                else:
                    t.append(v.index(r[c]))
            dp = demo.dps.index(tuple(t))
            stages = [c for c, co in enumerate(comps) if co["base"] == r[cols["compartment"]]]
            if not stages:
                raise TomlError(f"in the 'init_pop' file the compartment '{r[cols['compartment']]}' is not recognised")
            try:
                v = float(r[cols["population"]])
            except ValueError:
                raise TomlError(f"the population '{r[cols['population']]}' is not a number") from None
            if v < 0:
                raise TomlError("populations must be non-negative")
            for i, c in enumerate(stages):
                pop[a][c][dp] = v if i == 0 else 0
        for c in range(C):
            if any(pop[a][c][dp] is None for a in range(A) for dp in range(G)):
                raise TomlError(f"the initial populations for '{comps[c]['base']}' are not all specified")
        return [[[int(v) for v in row] for row in pa] for pa in pop]
    age_cols = [_col(head, a) for a in demo.ages]
    for a, c in zip(demo.ages, age_cols):
        if c is None:
            raise TomlError(f"the areas file needs a column '{a}'")

    def num(x):
        try:
            v = float(x)
        except ValueError:
            raise TomlError(f"the population '{x}' is not a number") from None
        if v < 0:
            raise TomlError("populations must be non-negative")
        return v

    fracs = {}
    for k in range(1, len(demo.cats)):
        if k == strain_k:
            continue
        vc = [_col(head, v) for v in demo.cats[k][1]]
        if None in vc:
            raise TomlError(f"the areas file needs a column for each value of '{demo.cats[k][0]}'")
        vals = [[num(r[c]) for c in vc] for r in rows]
        fracs[k] = [[x / sum(v) for x in v] for v in vals]
    sus = trans[[t["infection"] for t in trans].index(True)]["from"]
    pop = []
    for a, row in enumerate(rows):
        pa = []
        for t in demo.dps:
            if strain_k is not None and t[strain_k] > 0:      # everyone starts in strain 0's dps
                pa.append(0)
                continue
            v = num(row[age_cols[t[0]]])
            for k in range(1, len(demo.cats)):
                if k != strain_k:
                    v *= fracs[k][a][t[k]]
            pa.append(int(v + 0.5))
        pop.append([pa if c == sus else [0] * G for c in range(C)])
    return pop


def _obs_splines(rd, doc, P, splines):
    """obs_spline entries (Model::add_splines, src/model.cc:611-629): a
    spline per entry that some data table names in 'factor_spline'; its knot
    parameters are named after the spline ('<name> t=<bp>' with more than two
    knots, src/inputs.cc:1388-1401).  Appends the splines; returns name ->
    spline index."""
    used = {_s(e.get("factor_spline", "")) for e in doc.get("data_tables", [])} - {""}
    out = {}
    for e in doc.get("obs_spline", []):
        _check_keys(e, {"name", "bp", "param", "value", "prior", "smooth", "smooth_type", "factor", "factor_param",
                        "factor_value", "factor_prior"}, "obs_spline")
        name = _s(e.get("name", ""))
        if name == "":
            raise TomlError("each spline in 'obs_spline' must have 'name' set")
        if name in used:
            splines.append(_spline(rd, P, "obs_spline", e,
                                   lambda i, bp, name=name: name + (f" t={bp[i]}" if len(bp) > 2 else "")))
            out[name] = len(splines) - 1
    missing = used - set(out)
    if missing:
        raise TomlError(f"in 'data_tables' cannot find the observation spline '{sorted(missing)[0]}'")
    return out


def _area_sel(head, rows, expr, where):
    """Data::create_area_sel (src/data.cc:1783-1826) over the areas file:
    'col:v1|v2,col2:v3' -> the areas that every comma-separated term selects
    (a term takes the areas whose value in the column is listed; column
    'all' has the one value 'all'); '' -> every area."""
    if expr == "":
        return list(range(len(rows)))
    parts = expr.split(",")
    flag = [0] * len(rows)
    for part in parts:
        sp = part.split(":")
        if len(sp) != 2:
            raise TomlError(f"{where} there was a problem with the expression '{expr}'")
        vals = sp[1].split("|")
        if len(set(vals)) != len(vals):
            raise TomlError(f"{where} a value in '{expr}' must not be repeated.")
        regions = dict(_regions(head, rows, sp[0]))
        for v in vals:
            if v not in regions:
                raise TomlError(f"{where} the value '{v}' is not found in '{sp[0]}'.")
            for i in regions[v]:
                flag[i] += 1
    return [i for i in range(len(rows)) if flag[i] == len(parts)]


def _dp_sel(demo, expr, where):
    """Data::create_dp_sel (src/data.cc:1718-1781): 'cat:v1|v2,cat2:v3' ->
    the dps whose value of each named category is listed; '' -> all."""
    if expr == "":
        return list(range(demo.G))
    parts = expr.split(",")
    flag = [[0] * len(v) for _, v in demo.cats]
    for part in parts:
        sp = part.split(":")
        if len(sp) != 2:
            raise TomlError(f"{where} the expression '{expr}' must specify the demographic category and its value")
        k = next((k for k, (n, _) in enumerate(demo.cats) if n == sp[0]), None)
        if k is None:
            raise TomlError(f"{where} cannot find the value '{sp[0]}' in 'ages' or 'democats'")
        vals = sp[1].split("|")
        if len(set(vals)) != len(vals):
            raise TomlError(f"{where} a value in '{expr}' must not be repeated.")
        for v in vals:
            if v not in demo.cats[k][1]:
                raise TomlError(f"{where} cannot find the value '{v}' in 'name=\"{sp[0]}\"'")
            flag[k][demo.cats[k][1].index(v)] += 1
        for kk in range(len(demo.cats)):
            if kk != k:
                flag[kk] = [f + 1 for f in flag[kk]]
    return [dp for dp, t in enumerate(demo.dps) if all(flag[k][t[k]] == len(parts) for k in range(len(demo.cats)))]


MOD_KEYS = {"start", "end", "strain", "geo_filt", "democats_filt", "type", "trans", "factor", "value", "name"}


def _modifications(rd, entries, area_head, area_rows, demo, comps, trans, spline_names, strain_names):
    """'modification' entries (Inputs::find_modification,
    src/inputs.cc:1986-2061; filters resolved as src/data.cc:2142-2150) as
    the host model's modification dicts: type, the steps [start, end) (the
    entry's times relative to 'start', by default the whole period, times
    steps per day), factor (spline_set's 'value'), and the transitions,
    areas, dps, spline and strain it acts on (Model::setup_modification,
    src/model.cc:1445-1577)."""
    out = []
    for e in entries:
        _check_keys(e, MOD_KEYS, "modification")
        where = "In 'modification'"
        start = rd.time(_s(e["start"])) if _s(e.get("start", "")) else 0
        end = rd.time(_s(e["end"])) if _s(e.get("end", "")) else rd.period
        if not 0 <= start <= end <= rd.period:
            raise TomlError(f"{where} 'start' and 'end' must lie in the period, in that order")
        typ = _s(e.get("type", ""))
        md = {"type": typ, "start": start * rd.spd, "end": end * rd.spd,
              "areas": _area_sel(area_head, area_rows, _s(e.get("geo_filt", "")), f"{where} for 'geo_filt'"),
              "dps": _dp_sel(demo, _s(e.get("democats_filt", "")), f"{where} for 'democats_filt'")}
        key = "value" if typ == "spline_set" else "factor"
        if typ not in ("trans_rate_fac", "beta_fac", "efoi_fac", "spline_fac", "spline_set"):
            raise TomlError(f"'{typ}' in 'modification' is not recognised (it should be 'trans_rate_fac' or 'efoi_fac').")
        try:
            md["factor"] = float(_s(e[key]))
        except (KeyError, ValueError):
            raise TomlError(f"{where} a numeric '{key}' must be set") from None
        if md["factor"] < 0:
            raise TomlError(f"{where} for '{key}' the expression '{_s(e[key])}' is not positive.")
        if typ == "trans_rate_fac":
            name = _s(e.get("trans", ""))
            md["trans"] = [k for k, t in enumerate(trans)
                           if comps[t["from"]]["name"] + "->" + comps[t["to"]]["name"] == name]
            if not md["trans"]:
                raise TomlError(f"The transition '{name}' is not recognised")
        if typ in ("beta_fac", "efoi_fac"):
            st = _s(e.get("strain", ""))
            if st == "":
                if strain_names and len(strain_names) > 1:
                    raise TomlError("In 'modification' a value for 'strain' must be set")
                md["strain"] = 0
            else:
                if not strain_names or st not in strain_names:
                    raise TomlError(f"In 'modification' the value 'strain=\"{st}\"' is not recognised")
                md["strain"] = strain_names.index(st)
        if typ in ("spline_fac", "spline_set"):
            name = _s(e.get("name", ""))
            if name not in spline_names:
                raise TomlError(f"In 'modification' cannot find spline with the name '{name}'")
            md["spline"] = spline_names[name]
        out.append(md)
    return out


def _filters(e, demo, area_head, area_rows):
    """A data table's columns as (name, areas, dps) (Data::get_datafilter,
    src/data.cc:1460-1575): one per value of a category with democats_dep
    (its dps), one per region with geo_dep (its areas), else 'Data'; each
    narrowed by the table's geo_filt (areas) and democats_filt (dps) as
    Data::create_area_sel / create_dp_sel read them, and a column that keeps
    no area or no dp dropped."""
    geo_dep, demo_dep = _s(e.get("geo_dep", "")), _s(e.get("democats_dep", ""))
    geo_filt, demo_filt = _s(e.get("geo_filt", "")), _s(e.get("democats_filt", ""))
    file = _s(e.get("file", ""))
    if geo_dep and demo_dep:
        raise TomlError("in 'data_tables' values for 'democats_dep' and 'geo_dep' cannot both be set")

    def join(a, b):
        return a + ("," + b if b else "")

    cols = []
    if demo_dep:
        k = demo.find(demo_dep)
        if k is None:
            raise TomlError(f"'democats_dep={demo_dep}' is not a demographic category")
        name, vals = demo.cats[k]
        if len(vals) == 1:
            raise TomlError(f"'democats_dep=\"{demo_dep}\"' can only be used if '{demo_dep}' has more than one value")
        areas = _area_sel(area_head, area_rows, geo_filt, f"In 'geo_filt' for '{file}'")
        for v in vals:
            cols.append((v, areas, _dp_sel(demo, join(f"{name}:{v}", demo_filt), f"In 'democats_filt' for '{file}'")))
    elif geo_dep:
        dps = _dp_sel(demo, demo_filt, f"In 'democats_filt' for '{file}'")
        for region, _ in _regions(area_head, area_rows, geo_dep):
            cols.append((region, _area_sel(area_head, area_rows, join(f"{geo_dep}:{region}", geo_filt),
                                           f"In 'geo_filt' for '{file}'"), dps))
    else:
        cols.append(("Data", _area_sel(area_head, area_rows, geo_filt, f"In 'geo_filt' for '{file}'"),
                     _dp_sel(demo, demo_filt, f"In 'democats_filt' for '{file}'")))
    cols = [c for c in cols if c[1] and c[2]]
    if not cols:
        raise TomlError(f"In 'data_tables' the filter for '{file}' does not generate any data columns")
    return cols


def _obs_y(raw, thresh, thr_str, nodata_str, file):
    """A data value: a number, '*' below the threshold, '.' unknown (src/utils.cc:575-583)."""
    if raw == thr_str:
        if thresh == "":
            raise TomlError(f"the file '{file}' has thresholded data but no 'threshold' is set")
        return OBS_THRESH
    if raw == nodata_str:
        return OBS_UNKNOWN
    try:
        return float(raw)
    except ValueError:
        raise TomlError(f"in the file '{file}' the value '{raw}' is not a number") from None


def _marginal(rd, e, head, rows, file, demo, area_head, area_rows, tr_list, co_list, d, ngraph, otype, sd, percent,
              factor, thresh, thr_str, nodata_str):
    """A marginal table (Data::load_marginal_datatable, src/data.cc:1355-1455):
    for each value of the category (or region) a row naming it in the first
    column and its total in the second, observed over the steps from 'start'
    (default the analysis start) to 'end' (default its end); one graph for
    the table.  Returns the observations and the next graph index."""
    if not (_s(e.get("geo_dep", "")) or _s(e.get("democats_dep", ""))):
        raise TomlError("for marginal distributions in 'data_tables' either 'democats_dep' or 'geo_dep' must be set")
    t0 = rd.time(_s(e["start"])) if "start" in e else 0
    t1 = rd.time(_s(e["end"])) if "end" in e else rd.period
    ti, tf = t0 * rd.spd, t1 * rd.spd
    if ti < 0 or tf > rd.period * rd.spd or ti >= tf:
        raise TomlError(f"in 'data_tables' the file '{file}' covers times outside the analysis")
    names = [head[0]] + [r[0] for r in rows]         # the heading row names no value
    obs = []
    for name, areas, dps in _filters(e, demo, area_head, area_rows):
        if name not in names[1:]:
            raise TomlError(f"in 'data_tables' cannot find the value '{name}' in the first column of '{file}'")
        row = rows[names[1:].index(name)]
        o = {"y": _obs_y(row[1], thresh, thr_str, nodata_str, file), "type": otype, "factor": factor, "datatable": d,
             "graph": ngraph, "area": list(areas), "dp": list(dps), "thresh": float(thresh) if thresh else 0.0,
             "sett_i": ti, "sett_f": tf, "tr": list(tr_list), "co": list(co_list)}
        if sd is not None:
            o["sd"] = sd
        if percent is not None:
            o["percent"] = percent
        obs.append(o)
    return obs, ngraph + 1


def _observations(rd, doc, comps, trans, with_data, area_head, area_rows, demo, obs_spl, pop):
    """data_tables -> spec observations: time series of transitions,
    populations and population fractions, and marginal totals."""
    thr_str = _s(doc.get("threshold_str", "*"))
    nodata_str = _s(doc.get("nodata_str", "."))
    obs = []
    ngraph = 0                         # one graph per data column (Data::graph)
    tables = doc.get("data_tables", [])
    for d, e in enumerate(tables):
        _check_keys(e, {"type", "observation", "timestep", "obsmodel", "file", "factor", "shift", "start", "end",
                        "threshold", "epsilon_factor", "plot_name", "line_colour", "geo_dep", "democats_dep",
                        "geo_filt", "democats_filt", "factor_spline"}, "data_tables")
        typ = _s(e.get("type"))
        if typ not in ("transition", "population", "population_fraction", "marginal"):
            raise TomlError(f"in 'data_tables' the value of 'type={typ}' is not recognised")
        file = _s(e.get("file"))
        obsname = _s(e.get("observation"))
        tr_list, co_list = [], []
        for st in _split(obsname, ","):
            cs = [c for c, co in enumerate(comps) if co["base"] == st]
            if cs:
                if typ == "transition":
                    raise TomlError(f"observation '{obsname}' names compartments in a transition table")
                co_list += cs
                continue
            ts = [t for t, tr in enumerate(trans) if comps[tr["from"]]["base"] + "->" + comps[tr["to"]]["base"] == st]
            if not ts:
                raise TomlError(f"the observation '{st}' is not a compartment or a transition")
            if typ in ("population", "population_fraction"):
                raise TomlError(f"observation '{obsname}' names transitions in a population table")
            tr_list += ts
        # Observation model grammar (src/inputs.cc:412-474).
        om = _s(e.get("obsmodel", ""))
        spl = om.split(" ")
        sd = percent = None
        load_sd = False
        if om == "":
            raise TomlError("in 'data_tables' a value for 'obsmodel' must be set")
        if len(spl) > 2:
            raise TomlError(f"obsmodel '{om}' is invalid")
        if spl[0].lower() == "normal":
            otype = "normal"
            if len(spl) == 2:
                v = spl[1].lower()
                if v == "load":
                    load_sd = True
                elif v.endswith("%"):
                    otype, percent = "normal_percent", float(v[:-1])
                else:
                    sd = float(v)
            if sd is None and percent is None and not load_sd:
                raise TomlError(f"obsmodel '{om}' needs a standard deviation")
        elif om == "poisson":
            otype = "poisson"
        elif om == "negbin":
            # The reference asks for a shape here but compares the whole
            # string with 'negbin', so no negative-binomial table loads.
            raise TomlError("in 'data_tables' the value 'negbin' must contain a definition for the shape parameter")
        else:
            raise TomlError(f"obsmodel '{om}' is not recognised")
        eps = float(_s(e.get("epsilon_factor", "")) or 0.01)
        if eps != 0.01:
            raise TomlError("'epsilon_factor' other than 0.01 is not supported")
        factor = float(_s(e.get("factor", "")) or 1.0)
        shift = int(_s(e.get("shift", "")) or 0)
        thresh = _s(e.get("threshold", ""))
        if typ == "marginal" and "shift" in e:
            raise TomlError("in 'data_tables' for 'marginal' data the property 'shift' cannot be used")
        if not with_data:
            continue
        head, rows = rd.table(file)
        if typ == "marginal":
            o_list, ngraph = _marginal(rd, e, head, rows, file, demo, area_head, area_rows, tr_list, co_list, d,
                                       ngraph, otype, sd, percent, factor, thresh, thr_str, nodata_str)
            obs += o_list
            continue
        tc = _col(head, rd.time_col)
        if tc is None:
            raise TomlError(f"the file '{file}' needs a column '{rd.time_col}'")
        filters = _filters(e, demo, area_head, area_rows)
        cols = [(name, _col(head, name), areas, dps) for name, areas, dps in filters if _col(head, name) is not None]
        if not cols:
            raise TomlError(f"the data in '{file}' was not used: expected columns such as '{filters[0][0]}'")
        for name, _, _, _ in cols:
            if load_sd and _col(head, name + " SD") is None:
                raise TomlError(f"the file '{file}' needs a column '{name} SD' for 'normal load'")
        # Rows in the analysis period (Data::table_select_dates, src/data.cc:1981-2060).
        times = [rd.time(r[tc]) + shift for r in rows]
        start = rd.time(_s(e["start"])) if "start" in e else None
        end = rd.time(_s(e["end"])) if "end" in e else None
        r0 = 0
        while r0 < len(rows) and (times[r0] < 0 or (start is not None and times[r0] < start)):
            r0 += 1
        r1 = len(rows) - 1
        while r1 >= 0 and (times[r1] >= rd.period or (end is not None and times[r1] >= end)):
            r1 -= 1
        if r0 > r1:
            raise TomlError(f"the file '{file}' has no data in the analysis period")
        timestep = _s(e.get("timestep", ""))
        ts = int(timestep) if timestep else None
        if typ == "transition":
            if r0 == r1 and ts is None:
                raise TomlError(f"the file '{file}' has one row: 'timestep' must be set")
            if r1 > r0:
                steps = {times[r + 1] - times[r] for r in range(r0, r1)}
                if len(steps) != 1 or min(steps) <= 0:
                    raise TomlError(f"in the file '{file}' measurements must be time ordered with an equal timestep")
                tstep = steps.pop()
                if ts is not None and ts != tstep:
                    raise TomlError(f"'timestep' does not agree with the file '{file}'")
                ts = tstep
            if times[r1] + ts > rd.period:
                r1 -= 1
        for name, dc, areas, dps in cols:
            sc = _col(head, name + " SD") if load_sd else None
            graph = ngraph
            ngraph += 1
            fac = factor
            if typ == "population_fraction":
                # The column's initial population over its areas and dps (src/data.cc:1226-1235).
                fac *= 1.0 / sum(pop[a][c][dp] for a in areas for c in range(len(comps)) for dp in dps)
            col = _table_column(rd, rows, times, r0, r1, ts, "transition" if typ == "transition" else "population",
                                dc, sc, areas, dps, graph, d, otype, sd, percent, fac, thresh, thr_str, nodata_str,
                                tr_list, co_list, file)
            fs = _s(e.get("factor_spline", ""))
            if fs:
                # Population observations cannot take a spline (emsgEC("ObsModel",3), src/obsmodel.cc:100).
                if typ != "transition":
                    raise TomlError(f"'factor_spline' in the population table '{file}' is not supported")
                for o in col:
                    o["spline"] = obs_spl[fs]
            obs += col
    return obs


def _table_column(rd, rows, times, r0, r1, ts, typ, dc, sc, areas, dps, graph, d, otype, sd, percent, factor,
                  thresh, thr_str, nodata_str, tr_list, co_list, file):
    """The observations of one data column (Data::load_timeseries_datatable,
    src/data.cc:1236-1310): populations at the start of a day, transition
    counts over consecutive windows of `ts` days.  sc: the column of its
    standard deviations ('normal load'), or None."""
    obs = []
    ti = times[r0] * rd.spd if typ == "transition" else None
    for r in range(r0, r1 + 1):
        raw = rows[r][dc]
        if raw == thr_str:
            if thresh == "":
                raise TomlError(f"the file '{file}' has thresholded data but no 'threshold' is set")
            y = OBS_THRESH
        elif raw == nodata_str:
            y = OBS_UNKNOWN
        else:
            try:
                y = float(raw)
            except ValueError:
                raise TomlError(f"in the file '{file}' the value '{raw}' is not a number") from None
        o = {"y": y, "type": otype, "factor": factor, "datatable": d, "graph": graph, "area": list(areas),
             "dp": list(dps), "thresh": float(thresh) if thresh else 0.0}
        if sd is not None:
            o["sd"] = sd
        if percent is not None:
            o["percent"] = percent
        if sc is not None:
            o["sd"] = float(rows[r][sc])
        if typ == "transition":
            tf = (times[r] + ts) * rd.spd
            o.update(sett_i=ti, sett_f=tf, tr=list(tr_list))
            ti = tf
        else:
            t0 = times[r] * rd.spd
            o.update(sett_i=t0, sett_f=t0 + 1, co=list(co_list))
        obs.append(o)
    return obs


def load(path, datadir=None, with_data=True, prediction=False):
    """Reads a BEEP model file.  Returns (spec, values, meta): the spec dict,
    the parameter values the file gives (NaN where only a prior is given), and
    {"ignored": [...], "prediction": {...}, "derived": {...}} for keys
    recorded without effect on the simulation or the likelihood.
    datadir overrides the file's 'datadir' (which the reference reads relative
    to the working directory); with_data=False skips the data tables' files,
    as the reference's simulation mode does.  prediction=True reads the model
    as the reference's prediction mode does: the 'modification' entries go
    into spec["modifications"] and meta["prediction"]["pred_start"] is the
    day (relative to 'start') the modified simulations resume from
    (src/model.cc:1493-1585)."""
    with open(path, "rb") as f:
        doc = _toml.load(f)
    meta = {"ignored": [], "prediction": {}, "derived": {}}
    for k in doc:
        if k in IGNORED_KEYS:
            meta["ignored"].append(k)
        elif k in PREDICTION_KEYS:
            meta["prediction"][k] = doc[k]
        elif k in DERIVED_KEYS:
            meta["derived"][k] = doc[k]
        elif k not in SUPPORTED_KEYS:
            raise TomlError(f"the key '{k}' is not supported by the device path")
    if datadir is None:
        datadir = _s(doc.get("datadir", ""))
        if datadir == "":
            raise TomlError("'datadir' is not set")
    rd = _Reader(doc, datadir, prediction)
    P = _Params()
    # Demographic categories (Inputs::find_demographic_category,
    # src/inputs.cc:617-660): 'ages' (without it one group whose population
    # column is 'population'), then 'democats'; each may carry susceptibility
    # parameters.
    cats, sus_specs = [("age", ["population"])], {}
    entries = []
    if "ages" in doc:
        ag = doc["ages"]
        if not isinstance(ag, dict):
            raise TomlError("'ages' must be a table with 'cats'")
        _check_keys(ag, {"cats", "sus_param", "sus_value", "sus_prior"}, "ages")
        entries.append(("age", ag, "ages"))
    dcs = doc.get("democats", [])
    if isinstance(dcs, dict):
        dcs = [dcs]
    for e in dcs:
        _check_keys(e, {"name", "cats", "sus_param", "sus_value", "sus_prior"}, "democats")
        if _s(e.get("name", "")) == "":
            raise TomlError("in 'democats' every category needs a 'name'")
        entries.append((_s(e["name"]), e, "democats"))
    # Strains (src/inputs.cc:662-690): the last category.
    strain_rf = None
    if "strains" in doc:
        e = doc["strains"]
        if not isinstance(e, dict):
            raise TomlError("'strains' must be a table with 'cats'")
        _check_keys(e, {"cats", "sus_param", "sus_value", "sus_prior", "Rfactor_param", "Rfactor_value",
                        "Rfactor_prior"}, "strains")
        entries.append(("strain", e, "strains"))
    for name, e, key in entries:
        vals = _split(_s(e.get("cats", "")), "|")
        if not vals or "" in vals:
            raise TomlError(f"in '{key}' the value of 'cats' must list the categories")
        if name == "age":
            cats[0] = ("age", vals)
        else:
            if any(n.lower() == name.lower() for n, _ in cats):
                raise TomlError(f"the demographic category '{name}' is defined twice")
            cats.append((name, vals))
        if any(k in e for k in ("sus_param", "sus_value", "sus_prior")):
            n = len(vals)
            sus_specs[len(cats) - 1 if name != "age" else 0] = (
                rd.split_n(_s(e.get("sus_param", "")), n, key), rd.split_n(_s(e.get("sus_value", "")), n, key),
                rd.split_n(_s(e.get("sus_prior", "")), n, key))
    demo = _Demo(cats)
    ages, G = demo.ages, demo.G
    strain_k = len(cats) - 1 if "strains" in doc else None
    if strain_k is not None:
        e, names = doc["strains"], cats[strain_k][1]
        n = len(names)
        if n > 1:
            pn = rd.split_n(_s(e.get("Rfactor_param", "")), n, "strains")
            pv = rd.split_n(_s(e.get("Rfactor_value", "")), n, "strains")
            pp = rd.split_n(_s(e.get("Rfactor_prior", "")), n, "strains")
            strain_rf = [P.add(pn[i] or f"Rfactor_{names[i]}", pv[i], pp[i]) if (pv[i] or pp[i]) else -1
                         for i in range(n)]
    comps = _compartments(rd, doc, P, demo)
    trans, dirichlet = _transitions(rd, doc, comps, P, demo)

    # The areas file (src/data.cc:288-457), read first: spline entries name
    # areas by code.
    if "areas" not in doc:
        raise TomlError("'areas' must be set")
    head, rows = rd.table(_s(doc["areas"]))
    if _col(head, "area") is None:
        raise TomlError("the areas file needs a column 'area'")
    A = len(rows)
    codes = [r[_col(head, "area")] for r in rows]
    strain_names = cats[strain_k][1] if strain_k is not None else ["all"]
    nst = len(strain_names) if strain_k is not None else 1

    def spline_areas(e, key):
        area = _s(e.get("geo_filt", ""))
        if area == "":
            return "", list(range(A))
        if area not in codes:
            raise TomlError(f"In '{key}' cannot find 'area={area}'")
        return area, [codes.index(area)]

    # Splines by the names modifications use (Model::add_splines,
    # src/model.cc:505-628): 'name', else R_t, efoi(t), m(t), followed by the
    # area (and the strain) an entry is for.  R(t) (Model::update_R_ref,
    # src/model.cc:384-430): one entry per area or one for every area, each
    # area exactly once.
    splines, spline_names = [], {}
    r_entries = doc.get("R_spline", [])
    if isinstance(r_entries, dict):
        r_entries = [r_entries]
    if not r_entries:
        raise TomlError("A value for 'R_spline' must be set.")
    rgroups, rseen = [], [False] * A
    for e in r_entries:
        if _s(e.get("strain", "")):
            raise TomlError("In 'R_spline' a value for 'strain' cannot be set")
        if "age_dist" in e:
            raise TomlError("In 'R_spline' a value for 'age_dist' cannot be set")
        area, areas = spline_areas(e, "R_spline")
        name = (_s(e.get("name", "")) or "R_t") + (f" {area}" if area and A > 1 else "")
        for c in areas:
            if rseen[c]:
                raise TomlError(f"For '{name}' the area '{codes[c]}' already has a spline specified")
            rseen[c] = True
        splines.append(_spline(rd, P, "R_spline", e, suffix=f" {area}" if area else ""))
        spline_names.setdefault(name, len(splines) - 1)
        rgroups.append({"spline": len(splines) - 1, "areas": areas})
    if not all(rseen):
        raise TomlError("'R_spline' must be set for all areas and infection transitions.")
    # The external force (Model::update_efoi_ref, src/model.cc:432-480): one
    # entry per (strain, area), per area, per strain or for everything, each
    # (strain, area) exactly once; its own age_dist.
    e_entries = doc.get("efoi_spline", [])
    if isinstance(e_entries, dict):
        e_entries = [e_entries]
    egroups, eseen = [], [[False] * A for _ in range(nst)]
    for e in e_entries:
        area, areas = spline_areas(e, "efoi_spline")
        st = _s(e.get("strain", ""))
        if st:
            if strain_k is None or st not in strain_names:
                raise TomlError(f"In 'efoi_spline' cannot find 'strain={st}'")
            sts = [strain_names.index(st)]
        else:
            sts = list(range(nst))
        name = (_s(e.get("name", "")) or "efoi(t)") + (f" {area}" if area and A > 1 else "") + \
            (f" {st}" if st and nst > 1 else "")
        for k in sts:
            for c in areas:
                if eseen[k][c]:
                    raise TomlError(f"For '{name}' the area '{codes[c]}' already has a spline specified")
                eseen[k][c] = True
        splines.append(_spline(rd, P, "efoi_spline", e, suffix=(f" {area}" if area else "") + (f" {st}" if st else "")))
        spline_names.setdefault(name, len(splines) - 1)
        egroups.append({"spline": len(splines) - 1, "areas": areas, "strains": sts, "entry": e})
    if e_entries and not all(all(x) for x in eseen):
        raise TomlError("'efoi_spline' must be set for all areas and strains.")
    efoi = egroups[0]["spline"] if egroups else -1

    geo = -1
    if "geo_mixing_modify" in doc:
        if len(doc["geo_mixing_modify"]) != 1:
            raise TomlError("only one spline can represent 'geo_mixing'")
        splines.append(_spline(rd, P, "geo_mixing_modify", doc["geo_mixing_modify"][0]))
        geo = len(splines) - 1
        spline_names.setdefault(_s(doc["geo_mixing_modify"][0].get("name", "")) or "m(t)", geo)

    # The initial population (src/data.cc:288-457).
    pop = _initial_population(rd, doc, demo, comps, trans, head, rows, strain_k)
    pop_dp = np.array(pop, dtype=np.float64).sum(axis=(0, 1))              # [G]
    pop_age = np.array([pop_dp[[dp for dp in range(G) if demo.dp_age[dp] == a]].sum()
                        for a in range(demo.nage)])
    # Susceptibility (Model::set_dirichlet, src/model.cc:2768-2784): per
    # category a modified Dirichlet group weighted by the category's
    # population shares; a dp's susceptibility is the product over categories
    # (Model::create_susceptibility, src/model.cc:1144-1178).
    sus_terms = []
    for k, (nm, vl, pr) in sorted(sus_specs.items()):
        share = [float(pop_dp[[dp for dp in range(G) if demo.dps[dp][k] == v]].sum() / pop_dp.sum())
                 for v in range(len(demo.cats[k][1]))]
        if k == strain_k:                         # the last category's weights are 1 (src/model.cc:2778-2779)
            share = [1.0] * len(share)
        if any(x == "" for x in nm):
            raise TomlError(f"the susceptibility parameters of '{demo.cats[k][0]}' need names ('sus_param')")
        idx, grp = _modified_dirichlet(P, nm, vl, pr, share, f"susceptibility of '{demo.cats[k][0]}'")
        dirichlet.append(grp)
        sus_terms.append([idx[t[k]] for t in demo.dps])
    for e in doc.get("prob_reach", []):          # Inputs::find_probreach, src/inputs.cc:1328-1340
        _check_keys(e, {"name", "comp"}, "prob_reach")
        if _s(e.get("name", "")) == "" or _s(e.get("comp", "")) not in {c["base"] for c in comps}:
            raise TomlError(f"in 'prob_reach' cannot find the compartment '{_s(e.get('comp', ''))}'")
    area_eff, level, eff_groups = _effects(rd, doc, P, head, rows, pop)
    age_mod, mod_groups = _age_mods(rd, doc, P, ages, pop_age, splines)
    for m, e in zip(age_mod, doc.get("age_mixing_modify", []) + doc.get("age_mixing_perturb", [])):
        spline_names.setdefault(f"MatMod {_s(e.get('agecat', ''))} factor", m["spline"])
    obs_spl = _obs_splines(rd, doc, P, splines)
    for name, k in obs_spl.items():
        spline_names.setdefault(name, k)

    # Age mixing (Data::age_mixing_matrix, src/data_matrix.cc:138-175): an
    # age x age table without a heading, normalised later by the host model.
    G, nage = demo.nage, demo.nage
    N = None
    if "age_mixing_matrix" in doc:
        if "|" in _s(doc["age_mixing_matrix"]):
            raise TomlError("more than one 'age_mixing_matrix' (time-varying mixing) is not supported")
        _, nrows = rd.table(_s(doc["age_mixing_matrix"]), heading=False)
        if len(nrows) != G or any(len(r) != G for r in nrows):
            raise TomlError(f"the age mixing matrix must be {G} x {G} (one row and column per age group)")
        try:
            N = [[float(x) for x in r] for r in nrows]
        except ValueError:
            raise TomlError("the age mixing matrix holds a value that is not a number") from None
    elif G != 1:
        raise TomlError("when more than one age group exists a value for 'age_mixing_matrix' must be set")
    G = demo.G

    # efoi by age (src/inputs.cc:1421-1442): 'age_dist' normalised by the
    # population's age distribution, else 1 for every age group; each dp
    # takes its age group's (src/state.cc:415-427).
    def age_weights(e):
        w = [1.0] * nage
        if "age_dist" in e:
            if nage <= 1:
                raise TomlError("in 'efoi_spline' the value of 'age_dist' is not required")
            ad = [float(x) for x in rd.split_n(_s(e["age_dist"]), nage, "efoi_spline")]
            tot, agedist = sum(ad), pop_age / pop_age.sum()
            w = [float(ad[a] / (tot * agedist[a])) for a in range(nage)]
        return [w[a] for a in demo.dp_age]

    for eg in egroups:
        eg["agedist"] = age_weights(eg.pop("entry"))
    efoi_ad = egroups[0]["agedist"] if egroups else [1.0] * G

    # Geographic mixing (Data::load_geo_mixing_matrix, src/data_matrix.cc:206-229):
    # an A x A table without a heading; the diagonal and, per row, the
    # non-zero off-diagonal entries.  Normalised later by the host model.
    M = None
    if "geo_mixing_matrix" in doc:
        _, mrows = rd.table(_s(doc["geo_mixing_matrix"]), heading=False)
        if len(mrows) != A or any(len(r) != A for r in mrows):
            raise TomlError(f"the geographic mixing matrix must be {A} x {A} (one row and column per area)")
        try:
            mat = [[float(x) for x in r] for r in mrows]
        except ValueError:
            raise TomlError("the geographic mixing matrix holds a value that is not a number") from None
        M = {"diag": [mat[j][j] for j in range(A)], "rowptr": [0], "to": [], "val": []}
        for j in range(A):
            for i in range(A):
                if i != j and mat[j][i] != 0:
                    M["to"].append(i)
                    M["val"].append(mat[j][i])
            M["rowptr"].append(len(M["to"]))
    elif A != 1:
        raise TomlError("a value for 'geo_mixing_matrix' must be set when there is more than one area")

    spec = {
        "name": os.path.splitext(os.path.basename(path))[0], "A": A, "G": G, "period": rd.period, "spd": rd.spd,
        "comps": [{"name": c["name"], "inf": c["inf"], "mean": c["mean"], "shape": c["shape"] or 1} for c in comps],
        "trans": [{"from": t["from"], "to": t["to"], "infection": t["infection"],
                   **({"prob": t["prob"]} if t["prob"] is not None else {})} for t in trans],
        "params": P.specs,
        "splines": splines, "R_spline": rgroups[0]["spline"], "Rfactor": -1, "efoi_spline": efoi,
        "efoi_agedist": efoi_ad,
        "geo_spline": geo, "N": N, "M": M, "pop": pop, "dirichlet": dirichlet + eff_groups + mod_groups,
        "sus": sus_terms[0] if sus_terms else None, "sus_extra": sus_terms[1:],
        "nage": nage, "dp_age": demo.dp_age, "strain_Rfactor": strain_rf,
        "covars": _covariates(rd, doc, P, head, rows), "area_effect": area_eff, "level_effect": level,
        "age_mod": age_mod,
        "obs": _observations(rd, doc, comps, trans, with_data, head, rows, demo, obs_spl, pop),
    }
    if len(rgroups) > 1:
        spec["R_splines"] = rgroups
    if len(egroups) > 1:
        spec["efoi_splines"] = egroups
    if prediction:
        mods = doc.get("modification", [])
        if isinstance(mods, dict):
            mods = [mods]
        strain_names = cats[strain_k][1] if strain_k is not None else ["all"]
        spec["modifications"] = _modifications(rd, mods, head, rows, demo, comps, trans, spline_names,
                                               strain_names)
        # The day the simulations resume from (src/model.cc:1493-1585,
        # src/details.cc:80-85): the earliest modification, else the end;
        # 'prediction_start' when given, which no modification may precede.
        ps = min((m["start"] // rd.spd for m in spec["modifications"]), default=rd.period)
        if "prediction_start" in doc:
            given = rd.time(_s(doc["prediction_start"]))
            if given > rd.period:
                raise TomlError("'prediction_start' cannot be after 'end'")
            if spec["modifications"] and ps < given:
                raise TomlError("Dates in 'modification' cannot be before 'pred_start'")
            ps = given
        meta["prediction"]["pred_start"] = ps
    return spec, np.array(P.values), meta
