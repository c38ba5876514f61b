"""Synthetic model specs for the benchmark workloads (fixed seeds).

  seir_single      single-area well-mixed SEIR, N = 1e5, 100 daily steps,
                   beta = 0.3 (R = 2.1 with mean infectious period 7),
                   daily I->R observations with sd = max(5, 0.1 count)
  spatial_abcsmc   100 areas x 4 age groups, 8 compartments, 10-nearest-neighbour
                   mixing, Rt spline with 10 knots, 200 days, uniform priors
  dense_pmcmc      317 areas x 9 age groups, 10 compartments (Erlang E), dense
                   gravity mixing, density and temperature covariates, 300 days
  seir_areas       the SEIR structure of seir_single over A areas and G age
                   groups with 10-nearest-neighbour mixing (uneven batches, serving)

Observed values are synthetic: seir_single's come from a numpy tau-leap run of
the same SEIR model; the spatial configs use smooth per-area epidemic curves.
"""
import numpy as np

from .spec import nsteps


def _spline(name_params, t):
    return {"t": list(map(float, t)), "param": list(name_params), "multfac": [1.0] * len(t), "factor": -1}


def _knn_mixing(xy, k, scale):
    """Raw geographic mixing: diagonal 1, exp(-d/scale) to the k nearest areas."""
    A = len(xy)
    d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
    rowptr, to, val = [0], [], []
    for c in range(A):
        order = np.argsort(d[c], kind="stable")
        nb = [j for j in order if j != c][:k]
        nb.sort()
        for j in nb:
            to.append(int(j))
            val.append(float(np.exp(-d[c, j] / scale)))
        rowptr.append(len(to))
    return {"diag": [1.0] * A, "rowptr": rowptr, "to": to, "val": val}


def _dense_mixing(W):
    A = W.shape[0]
    rowptr, to, val = [0], [], []
    for c in range(A):
        for j in range(A):
            if j != c and W[c, j] != 0:
                to.append(j)
                val.append(float(W[c, j]))
        rowptr.append(len(to))
    return {"diag": [float(W[c, c]) for c in range(A)], "rowptr": rowptr, "to": to, "val": val}


def seir_single(period=100, obs=True):
    """One area, one age group: S -> E -> I -> R with E mean 5, I mean 7, R = 2.1."""
    params = [
        {"name": "R", "prior": ("uniform", 1.0, 4.0)},
        {"name": "mE", "prior": ("fixed", 5.0)},
        {"name": "mI", "prior": ("fixed", 7.0)},
        {"name": "infI", "prior": ("fixed", 1.0)},
    ]
    spec = {
        "name": "seir_single", "A": 1, "G": 1, "period": period, "spd": 1,
        "comps": [{"name": "S", "inf": -1, "mean": None}, {"name": "E", "inf": -1, "mean": [1]},
                  {"name": "I", "inf": 3, "mean": [2]}, {"name": "R", "inf": -1, "mean": None}],
        "trans": [{"from": 0, "to": 1, "infection": True}, {"from": 1, "to": 2}, {"from": 2, "to": 3}],
        "params": params,
        "splines": [_spline([0, 0], [0, period])],
        "R_spline": 0, "Rfactor": -1, "efoi_spline": -1, "geo_spline": -1,
        "N": None, "M": None,
        "pop": [[[99990], [0], [10], [0]]],
        "obs": [],
    }
    truth = np.array([2.1, 5.0, 7.0, 1.0])
    if obs:
        y = _seir_synthetic_counts(period, seed=20250101)
        spec["obs"] = [{"y": float(v), "sd": max(5.0, 0.1 * v), "type": "normal", "datatable": 0, "graph": 0,
                        "sett_i": d, "sett_f": d + 1, "area": [0], "dp": [0], "tr": [2]} for d, v in enumerate(y)]
    return spec, truth


def _seir_synthetic_counts(period, seed):
    """Daily I->R counts of one numpy tau-leap run of the SEIR model (data only)."""
    rng = np.random.default_rng(seed)
    S, E, I, R, N, beta = 99990, 0, 10, 0, 1e5, 0.3
    out = []
    for _ in range(period):
        ninf = rng.poisson(beta * S * I / N) if S > 0 and I > 0 else 0
        nei = rng.poisson(E / 5.0) if E > 0 else 0
        nir = rng.poisson(I / 7.0) if I > 0 else 0
        out.append(nir)
        S, E, I, R = S - ninf, E + ninf - nei, I + nei - nir, R + nir
    return out


AGE_FRAC4 = (0.25, 0.30, 0.30, 0.15)


def _age_pops(rng, A, fracs, median, sigma):
    tot = np.exp(np.log(median) + sigma * rng.standard_normal(A))
    return np.array([[int(np.floor(t * f + 0.5)) for f in fracs] for t in tot])     # src/data.cc:454 rounding


def _epi_curves(rng, A, S, scale):
    t = np.arange(S)
    peak = rng.uniform(0.3, 0.7, A) * S
    width = rng.uniform(0.08, 0.2, A) * S
    amp = scale * rng.uniform(0.2, 1.0, A)
    return np.floor(amp[:, None] * np.exp(-0.5 * ((t[None, :] - peak[:, None]) / width[:, None]) ** 2) + 0.5)


def spatial_abcsmc(A=100, period=200, seed=20250102, with_obs=True):
    """100 areas on random coordinates, 4 age groups, S E A I H D R RH."""
    rng = np.random.default_rng(seed)
    G = 4
    xy = rng.uniform(0, 1, (A, 2))
    ages = _age_pops(rng, A, AGE_FRAC4, 5e4, 0.8)
    nknot = 10
    params = [{"name": f"R{k}", "prior": ("uniform", 0.5, 3.0)} for k in range(nknot)]
    P = {p["name"]: i for i, p in enumerate(params)}

    def add(name, prior):
        P[name] = len(params)
        params.append({"name": name, "prior": prior})
        return P[name]

    mE, mI, mH = add("mE", ("uniform", 2, 10)), add("mI", ("uniform", 2, 10)), add("mH", ("uniform", 2, 10))
    mA = add("mA", ("fixed", 5.0))
    wEA, wEI = add("wEA", ("uniform", 0.05, 0.95)), add("wEI", ("uniform", 0.05, 0.95))
    wIH, wIR = add("wIH", ("uniform", 0.05, 0.95)), add("wIR", ("uniform", 0.05, 0.95))
    wHD, wHR = add("wHD", ("uniform", 0.05, 0.95)), add("wHR", ("uniform", 0.05, 0.95))
    infA, infI = add("infA", ("fixed", 0.5)), add("infI", ("fixed", 1.0))
    S_, E_, A_, I_, H_, D_, R_, RH_ = range(8)
    comps = [
        {"name": "S", "inf": -1, "mean": None}, {"name": "E", "inf": -1, "mean": [mE] * G},
        {"name": "A", "inf": infA, "mean": [mA] * G}, {"name": "I", "inf": infI, "mean": [mI] * G},
        {"name": "H", "inf": -1, "mean": [mH] * G}, {"name": "D", "inf": -1, "mean": None},
        {"name": "R", "inf": -1, "mean": None}, {"name": "RH", "inf": -1, "mean": None},
    ]
    trans = [
        {"from": S_, "to": E_, "infection": True},
        {"from": E_, "to": A_, "prob": [wEA] * G}, {"from": E_, "to": I_, "prob": [wEI] * G},
        {"from": A_, "to": R_},
        {"from": I_, "to": H_, "prob": [wIH] * G}, {"from": I_, "to": R_, "prob": [wIR] * G},
        {"from": H_, "to": D_, "prob": [wHD] * G}, {"from": H_, "to": RH_, "prob": [wHR] * G},
    ]
    pop = np.zeros((A, 8, G), dtype=np.int64)
    pop[:, S_, :] = ages
    seeds = rng.choice(A, 5, replace=False)
    for c in seeds:
        pop[c, I_, 1] += 2
        pop[c, S_, 1] -= 2
    Nraw = 0.2 * np.ones((G, G)) + 0.8 * np.eye(G)
    S = period
    spec = {
        "name": "spatial_abcsmc", "A": A, "G": G, "period": period, "spd": 1,
        "comps": comps, "trans": trans, "params": params,
        "splines": [_spline([P[f"R{k}"] for k in range(nknot)], np.linspace(0, period, nknot))],
        "R_spline": 0, "Rfactor": -1, "efoi_spline": -1, "geo_spline": -1,
        "N": Nraw.tolist(), "M": _knn_mixing(xy, 10, 0.1), "pop": pop.tolist(),
        "dirichlet": [{"param": [wEA, wEI]}, {"param": [wIH, wIR]}, {"param": [wHD, wHR]}],
        "obs": [],
    }
    if with_obs:
        obs = []
        for dt, (tr, scale) in enumerate(((4, 40.0), (6, 8.0))):      # I->H, H->D per area per day
            y = _epi_curves(rng, A, S, scale)
            for c in range(A):
                for d in range(S):
                    obs.append({"y": float(y[c, d]), "type": "normal_percent", "percent": 10.0, "datatable": dt,
                                "graph": dt * A + c, "sett_i": d, "sett_f": d + 1, "area": [c],
                                "dp": list(range(G)), "tr": [tr]})
        spec["obs"] = obs
    truth = np.array([1.8, 1.6, 1.2, 0.9, 0.8, 1.1, 1.3, 1.0, 0.9, 0.8, 5.0, 6.0, 8.0, 5.0,
                      0.4, 0.6, 0.2, 0.8, 0.3, 0.7, 0.5, 1.0])
    return spec, truth


def dense_pmcmc(A=317, period=300, seed=20250103, with_obs=True):
    """317 areas x 9 age groups, dense gravity mixing, covariates, Erlang-2 E."""
    rng = np.random.default_rng(seed)
    G = 9
    xy = rng.uniform(0, 1, (A, 2))
    fr = np.array([0.11, 0.12, 0.13, 0.13, 0.13, 0.13, 0.11, 0.09, 0.05])
    ages = _age_pops(rng, A, fr, 5e4, 0.8)
    tot = ages.sum(1).astype(float)
    d2 = ((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    W = tot[:, None] * tot[None, :] / d2
    nn = d2.min(1)
    W[np.arange(A), np.arange(A)] = tot ** 2 / nn                        # non-zero diagonal, pop_i^2 / d_nn^2
    params = [{"name": "R", "prior": ("uniform", 1.0, 3.0)}]
    P = {"R": 0}

    def add(name, prior):
        P[name] = len(params)
        params.append({"name": name, "prior": prior})
        return P[name]

    mE, mA, mI, mH, mU = (add("mE", ("uniform", 2, 8)), add("mA", ("fixed", 5.0)), add("mI", ("uniform", 2, 8)),
                          add("mH", ("uniform", 2, 10)), add("mU", ("uniform", 2, 10)))
    wEA, wEI = add("wEA", ("uniform", 0.05, 0.95)), add("wEI", ("uniform", 0.05, 0.95))
    wIR, wIH = add("wIR", ("uniform", 0.05, 0.95)), add("wIH", ("uniform", 0.05, 0.95))
    wHU, wHR = add("wHU", ("uniform", 0.05, 0.95)), add("wHR", ("uniform", 0.05, 0.95))
    wUD, wUR = add("wUD", ("uniform", 0.05, 0.95)), add("wUR", ("uniform", 0.05, 0.95))
    infA, infI = add("infA", ("fixed", 0.5)), add("infI", ("fixed", 1.0))
    bden, btemp = add("b_density", ("uniform", -0.5, 0.5)), add("b_temp", ("uniform", -0.1, 0.1))
    S_, E_, E1_, A_, I_, H_, U_, D_, R_, RH_ = range(10)
    comps = [
        {"name": "S", "inf": -1, "mean": None},
        {"name": "E", "inf": -1, "mean": [mE] * G, "shape": 2}, {"name": "E1", "inf": -1, "mean": [mE] * G, "shape": 2},
        {"name": "A", "inf": infA, "mean": [mA] * G}, {"name": "I", "inf": infI, "mean": [mI] * G},
        {"name": "H", "inf": -1, "mean": [mH] * G}, {"name": "U", "inf": -1, "mean": [mU] * G},
        {"name": "D", "inf": -1, "mean": None}, {"name": "R", "inf": -1, "mean": None},
        {"name": "RH", "inf": -1, "mean": None},
    ]
    trans = [
        {"from": E_, "to": E1_},                                        # internal Erlang transition first
        {"from": S_, "to": E_, "infection": True},
        {"from": E1_, "to": A_, "prob": [wEA] * G}, {"from": E1_, "to": I_, "prob": [wEI] * G},
        {"from": A_, "to": R_},
        {"from": I_, "to": R_, "prob": [wIR] * G}, {"from": I_, "to": H_, "prob": [wIH] * G},
        {"from": H_, "to": U_, "prob": [wHU] * G}, {"from": H_, "to": RH_, "prob": [wHR] * G},
        {"from": U_, "to": D_, "prob": [wUD] * G}, {"from": U_, "to": RH_, "prob": [wUR] * G},
    ]
    pop = np.zeros((A, 10, G), dtype=np.int64)
    pop[:, S_, :] = ages
    for c in rng.choice(A, min(A, 10), replace=False):
        pop[c, I_, 3] += 3
        pop[c, S_, 3] -= 3
    Nraw = 0.3 * np.ones((G, G)) + np.diag(np.linspace(1.5, 0.6, G))
    days = np.arange(period)
    dens = rng.standard_normal(A)
    temp = 10 + 8 * np.sin(2 * np.pi * days / 365.0)
    spec = {
        "name": "dense_pmcmc", "A": A, "G": G, "period": period, "spd": 1,
        "comps": comps, "trans": trans, "params": params,
        "splines": [_spline([0, 0], [0, period])],
        "R_spline": 0, "Rfactor": -1, "efoi_spline": -1, "geo_spline": -1,
        "N": Nraw.tolist(), "M": _dense_mixing(W), "pop": pop.tolist(),
        "covars": [{"param": bden, "value": np.repeat(dens[:, None], period, 1).tolist()},
                   {"param": btemp, "value": np.repeat(temp[None, :], A, 0).tolist()}],
        "dirichlet": [{"param": [wEA, wEI]}, {"param": [wIR, wIH]}, {"param": [wHU, wHR]}, {"param": [wUD, wUR]}],
        "obs": [],
        "sections": [(k * 30, min(period, (k + 1) * 30)) for k in range((period + 29) // 30)],
    }
    if with_obs:
        obs = []
        y = _epi_curves(rng, A, period, 20.0)
        for c in range(A):
            for d in range(period):
                obs.append({"y": float(y[c, d]), "type": "normal_percent", "percent": 10.0, "datatable": 0,
                            "graph": c, "sett_i": d, "sett_f": d + 1, "area": [c], "dp": list(range(G)), "tr": [6]})
        spec["obs"] = obs
    truth = np.array([1.6, 4.0, 5.0, 5.0, 7.0, 8.0, 0.4, 0.6, 0.8, 0.2, 0.3, 0.7, 0.4, 0.6, 0.5, 1.0, 0.1, -0.02])
    return spec, truth


def daily_observations(transnum, tr, percent=10.0, datatable=0):
    """Daily observations of transition tr per area (all age groups summed),
    with normal errors of `percent` percent, from a transnum trajectory
    [S][A][T][G] -- e.g. one simulation at the true parameters, so that a
    likelihood of the model at those parameters is not degenerate.  Areas
    whose series is zero throughout are left unobserved (their percentage
    error would be zero)."""
    tn = np.asarray(transnum)
    S, A = tn.shape[0], tn.shape[1]
    y = tn[:, :, tr, :].sum(-1)
    G = tn.shape[3]
    return [{"y": float(y[d, c]), "type": "normal_percent", "percent": percent, "datatable": datatable,
             "graph": c, "sett_i": d, "sett_f": d + 1, "area": [c], "dp": list(range(G)), "tr": [tr]}
            for c in range(A) if y[:, c].max() > 0 for d in range(S)]


def seir_areas(A, G, period, seed, obs_every=0, daily_obs=False):
    """SEIR in every area with 10-NN mixing, G in {1, 4, 9} age groups.
    obs_every: Poisson I->R observations per area every that many days.
    daily_obs: daily I->R counts per area (all age groups) with normal errors
    of sd max(5, 0.1 count), as the single-area SEIR's, stored as a compact
    observation table."""
    rng = np.random.default_rng(seed)
    fr = {1: (1.0,), 4: AGE_FRAC4, 9: (0.11, 0.12, 0.13, 0.13, 0.13, 0.13, 0.11, 0.09, 0.05)}[G]
    ages = _age_pops(rng, A, fr, 5e4, 0.8)
    xy = rng.uniform(0, 1, (A, 2))
    params = [{"name": "R", "prior": ("uniform", 1.0, 4.0)}, {"name": "mE", "prior": ("uniform", 2.0, 8.0)},
              {"name": "mI", "prior": ("uniform", 3.0, 10.0)}, {"name": "infI", "prior": ("fixed", 1.0)}]
    pop = np.zeros((A, 4, G), dtype=np.int64)
    pop[:, 0, :] = ages
    for c in rng.choice(A, min(A, 3), replace=False):
        g = G // 2
        pop[c, 2, g] += 5
        pop[c, 0, g] -= 5
    spec = {
        "name": f"seir_areas_{A}x{G}x{period}", "A": A, "G": G, "period": period, "spd": 1,
        "comps": [{"name": "S", "inf": -1, "mean": None}, {"name": "E", "inf": -1, "mean": [1] * G},
                  {"name": "I", "inf": 3, "mean": [2] * G}, {"name": "R", "inf": -1, "mean": None}],
        "trans": [{"from": 0, "to": 1, "infection": True}, {"from": 1, "to": 2}, {"from": 2, "to": 3}],
        "params": params,
        "splines": [_spline([0, 0], [0, period])],
        "R_spline": 0, "Rfactor": -1, "efoi_spline": -1, "geo_spline": -1,
        "N": (0.2 * np.ones((G, G)) + 0.8 * np.eye(G)).tolist() if G > 1 else None,
        "M": _knn_mixing(xy, min(10, A - 1), 0.1) if A > 1 else None,
        "pop": pop.tolist(), "obs": [],
    }
    if obs_every:
        y = _epi_curves(rng, A, period, 30.0)
        spec["obs"] = [{"y": float(y[c, d]), "type": "poisson", "datatable": 0, "graph": c, "sett_i": d,
                        "sett_f": d + 1, "area": [c], "dp": list(range(G)), "tr": [2]}
                       for c in range(A) for d in range(0, period, obs_every)]
    elif daily_obs:
        y = _epi_curves(rng, A, period, 30.0).reshape(-1)          # [A][period], area-major
        c, d = np.divmod(np.arange(A * period), period)
        spec["obs_table"] = {"y": y, "sd": np.maximum(5.0, 0.1 * y), "type": "normal", "datatable": 0,
                             "graph": c, "area": c, "sett_i": d, "sett_f": d + 1, "dp": list(range(G)), "tr": [2]}
    truth = np.array([2.2, 4.0, 6.0, 1.0])
    return spec, truth


def steps(spec):
    return nsteps(spec)


def uneven_batch(n=16384, seed=4, frac_mc3=0.5, ladder=8):
    """The uneven workload: n instances of the SEIR-in-every-area model with
    area counts uniform in 10..400, horizons uniform in 50..400 days and age
    groups in {1, 4, 9}, half ABC-SMC particles and half MC3 chains.
    Returns {"abc": groups, "mc3": ladders}:
      groups   ABC-SMC particles with the same (areas, age groups) share one
               model: {A, G, S (the longest horizon), horizons}; each particle
               is simulated over its own horizon;
      ladders  MC3 chains in ladders of `ladder` temperatures, each ladder one
               model {A, G, S} (the draw of its first chain)."""
    rng = np.random.default_rng(seed)
    A = rng.integers(10, 401, n)
    S = rng.integers(50, 401, n)
    G = rng.choice([1, 4, 9], n)
    mc3 = rng.random(n) < frac_mc3
    groups = {}
    for a, s, g in zip(A[~mc3], S[~mc3], G[~mc3]):
        d = groups.setdefault((int(a), int(g)), {"A": int(a), "G": int(g), "S": 0, "horizons": []})
        d["S"] = max(d["S"], int(s))
        d["horizons"].append(int(s))
    idx = np.flatnonzero(mc3)
    nl = len(idx) // ladder
    ladders = [{"A": int(A[idx[l * ladder]]), "G": int(G[idx[l * ladder]]), "S": int(S[idx[l * ladder]])}
               for l in range(nl)]
    return {"abc": [groups[k] for k in sorted(groups)], "mc3": ladders, "ladder": ladder}
