"""TEST INFRASTRUCTURE ONLY: FP64 numpy restatement of the reference's
set-up and State::set_param derivations, used to check the product's C++ host
model and to feed the C++ oracle (beep_oracle.cc).

Restates, for one strain with age as the only demographic category:
  Data::geo_normalise               src/data_matrix.cc:55-96
  Data::age_mixing_matrix           src/data_matrix.cc:138-202
  Data::agematrix_normalise         src/data_matrix.cc:100-134
  Data::read_covars (centring)      src/data.cc:557-565
  Data::set_obs_sd                  src/data.cc:1589-1623
  Model::set_infected_uninfected    src/model.cc:110-166
  Model::dirichlet_correct          src/model.cc:2685-2695
  Model::spline_prior               src/model.cc:1662-1704
  Model::create_transrate           src/model.cc:1116-1140
  Model::create_disc_spline         src/model.cc:1020-1053
  Model::create_susceptibility      src/model.cc:1144-1178
  Model::create_areafactor          src/model.cc:1182-1250
  Model::create_Ntime               src/model.cc:1254-1301
  Model::calculate_beta_from_R      src/model.cc:1318-1362 with calculate_Vinv,
                                    calculate_F, calculate_NGM (2485-2576) and
                                    largest_eigenvalue (src/matrix.cc:1139-1219)
  Model::prior / inbounds           src/model.cc:1590-1659, 2625-2669, with
                                    gamma_probability (src/utils.cc:115-120)
                                    for MDir weights

Pins: the gamma prior against the reference's own gamma_probability (golden
vectors) and scipy; the area factor against loops over effects, levels and
uncentred covariates; beta-from-R against the SEIR closed form rho = mean infectious period
and against numpy.linalg.eig of the same next-generation matrix; the
normalisations against their defining identities (row sums of the per-capita
mixing equal one; population-weighted age mixing averages one).
"""
import math

import numpy as np

EPSILON_FACTOR = 0.01
VTINY = 1e-15


def normalise(spec):
    """Data-level quantities: normalised mixing, populations, infected states."""
    A, G = spec["A"], spec["G"]
    comps, trans = spec["comps"], spec["trans"]
    C = len(comps)
    pop = np.asarray(spec["pop"], dtype=np.float64).reshape(A, C, G)
    area_pop = pop.sum(axis=(1, 2))
    popsize = area_pop.sum()
    pop_age = pop.sum(axis=(0, 1))                     # per dp
    # Age mixing over the age groups' populations, then between dps the
    # mixing of their age groups.
    dp_age = np.asarray(spec["dp_age"]) if spec.get("dp_age") is not None else np.arange(G)
    nage = spec.get("nage", G) if spec.get("dp_age") is not None else G
    pop_a = np.array([pop_age[dp_age == a].sum() for a in range(nage)])
    if spec.get("N") is not None:
        raw = np.asarray(spec["N"], dtype=np.float64)
        m = raw / (pop_a[:, None] / popsize)
        N_age = 0.5 * (m + m.T)
    else:
        N_age = np.ones((1, 1))
    s = sum(pop_a[aa] * N_age[aa, a] for a in range(nage) for aa in range(nage)) / (nage * pop_a.sum())
    N_age = N_age / s
    N = N_age[np.ix_(dp_age, dp_age)]
    ns = len(spec.get("strain_Rfactor") or [1])          # strains mix within a strain only
    strain = np.arange(G) // (G // ns)
    N = N * (strain[:, None] == strain[None, :])
    # Geographic mixing.
    M = spec.get("M")
    if M is None:
        diag, rowptr, to, val = np.ones(1), np.array([0, 0]), np.zeros(0, int), np.zeros(0)
    else:
        diag = np.array(M["diag"], dtype=np.float64)
        rowptr, to, val = np.array(M["rowptr"]), np.array(M["to"]), np.array(M["val"], dtype=np.float64)
    diag = diag.copy()
    val = val.copy()
    for c in range(A):
        tot = diag[c] * area_pop[c] + sum(val[j] * area_pop[to[j]] for j in range(rowptr[c], rowptr[c + 1]))
        diag[c] /= tot
        val[rowptr[c]:rowptr[c + 1]] /= tot
    factor = (1.0 / area_pop) / diag
    # Infected states.
    itr = [k for k, t in enumerate(trans) if t.get("infection")][0]
    exits = [[k for k, t in enumerate(trans) if t["from"] == c] for c in range(C)]
    fwd = [False] * C
    lst = [trans[itr]["to"]]
    fwd[lst[0]] = True
    i = 0
    while i < len(lst):
        for k in exits[lst[i]]:
            if trans[k].get("infection"):
                continue
            cc = trans[k]["to"]
            if not fwd[cc] and cc != trans[itr]["from"]:
                lst.append(cc)
                fwd[cc] = True
        i += 1
    back = [False] * C
    for co in range(C):
        if comps[co].get("inf", -1) < 0:
            continue
        l2 = [co]
        back[co] = True
        i = 0
        while i < len(l2):
            for k, t in enumerate(trans):
                if not t.get("infection") and t["to"] == l2[i] and not back[t["from"]]:
                    back[t["from"]] = True
                    l2.append(t["from"])
            i += 1
    inf_state = [c for c in range(C) if fwd[c] and back[c]]
    covars = []
    for cv in spec.get("covars", []):
        v = np.asarray(cv["value"], dtype=np.float64).reshape(A, spec["period"])
        covars.append((cv["param"], v - v.mean()))
    return dict(N=N, N_age=N_age, dp_age=dp_age, strain=strain, nstrain=ns, diag=diag, rowptr=rowptr, to=to, val=val, factor=factor,
                area_pop=area_pop, popsize=popsize, Rdist=pop_age / popsize, inf_state=inf_state, itr=itr, exits=exits,
                covars=covars)


def dirichlet_correct(spec, pv):
    pd = np.array(pv, dtype=np.float64)
    for d in spec.get("dirichlet", []):
        par = d["param"]
        frac = d.get("frac", [1.0] * len(par))
        tot = sum(pv[t] for t in par)
        for t, f in zip(par, frac):
            pd[t] = pv[t] / (f * tot)
    return pd


def disc_spline(spec, sp, pd):
    S = spec["period"] * spec["spd"]
    spl = spec["splines"][sp]
    t_k, par, mf = spl["t"], spl["param"], spl.get("multfac", [1.0] * len(spl["t"]))
    factor = pd[spl["factor"]] if spl.get("factor", -1) >= 0 else 1.0
    out = np.zeros(S)
    p = 0
    for s in range(S):
        t = (s + 0.5) * spec["period"] / S
        while p + 1 < len(t_k) and t > t_k[p + 1]:
            p += 1
        if p + 1 == len(t_k):
            out[s] = out[s - 1] if s > 0 else 0.0
        else:
            f = (t - t_k[p]) / (t_k[p + 1] - t_k[p])
            out[s] = factor * (pd[par[p]] * mf[p] * (1 - f) + pd[par[p + 1]] * mf[p + 1] * f)
    if spec.get("modifications"):
        # Prediction mode (src/model.cc:1044-1049): spline_mult, then spline_set.
        mult, setv = spline_modifications(spec)
        out = out * mult[sp]
        out = np.where(np.isnan(setv[sp]), out, setv[sp])
    return out


def spline_modifications(spec):
    """Model::model_modification's spline tables (src/model.cc:1484-1490,
    1559-1575): per spline and step the product of the spline_fac factors and
    the spline_set value (NaN = not set); a step both set and multiplied is
    the reference's error."""
    S = spec["period"] * spec["spd"]
    ns = len(spec["splines"])
    mult, setv = np.ones((ns, S)), np.full((ns, S), np.nan)
    for md in spec.get("modifications") or []:
        if md["type"] not in ("spline_fac", "spline_set"):
            continue
        sp = md["spline"]
        for s in range(md["start"], md["end"]):
            if md["type"] == "spline_fac":
                if not np.isnan(setv[sp, s]):
                    raise ValueError("In 'modification' a spline cannot both be directly set and have a factor change")
                mult[sp, s] *= md["factor"]
            else:
                if mult[sp, s] != 1:
                    raise ValueError("In 'modification' a spline cannot both be directly set and have a factor change")
                setv[sp, s] = md["factor"]
    return mult, setv


def modification_tables(spec):
    """Model::model_modification's multipliers (src/model.cc:1446-1556):
    beta_mult and efoi_mult [S][A][G] (each dp takes its strain's), and
    transmean_mult [S][A][T][G]; None for a kind with no modification."""
    A, G, T = spec["A"], spec["G"], len(spec["trans"])
    S = spec["period"] * spec["spd"]
    ns = max(1, len(spec.get("strain_Rfactor") or [1]))
    dpm = G // ns
    bm = em = tm = None
    for md in spec.get("modifications") or []:
        areas = md.get("areas") or list(range(A))
        if md["type"] == "trans_rate_fac":
            tm = np.ones((S, A, T, G)) if tm is None else tm
            dps = md.get("dps") or list(range(G))
            for tr in md["trans"]:
                for s in range(md["start"], md["end"]):
                    for c in areas:
                        for dp in dps:
                            tm[s, c, tr, dp] *= md["factor"]
        elif md["type"] in ("beta_fac", "efoi_fac"):
            st = md.get("strain", 0)
            if md["type"] == "beta_fac":
                bm = np.ones((S, A, G)) if bm is None else bm
                tab = bm
            else:
                em = np.ones((S, A, G)) if em is None else em
                tab = em
            for s in range(md["start"], md["end"]):
                for c in areas:
                    tab[s, c, st * dpm:(st + 1) * dpm] *= md["factor"]
    return bm, em, tm


def largest_eigenvalue(M, rng=None):
    """Power iteration with sum normalisation and the reference's convergence
    test; the start vector is random like the reference's ran() warm start."""
    n = M.shape[0]
    rng = rng or np.random.default_rng(0)
    vec = rng.uniform(1e-10, 1.0, n)
    ev = 0.0
    for loop in range(10001):
        v2 = M @ vec
        ev_new = v2.sum()
        v2 = v2 / ev_new
        if loop % 10 == 0:
            lim = 100 * VTINY * (10 if loop > 100 else 1) * (10 if loop > 1000 else 1)
            if abs(ev_new - ev) < lim and np.all(np.abs(v2 - vec) < lim):
                return ev_new
        vec, ev = v2, ev_new
    raise RuntimeError("eigen-vector convergence problem")


def ngm(spec, nz, tr_rate, sus, inf, N=None, dist=None):
    """Next-generation matrix of calculate_R_beta_ratio_using_NGM for the
    age contact matrix N (default the normalised one) and the dp weights dist
    of the R spline's areas (default every area's, nz["Rdist"])."""
    N = nz["N"] if N is None else N
    dist = nz["Rdist"] if dist is None else dist
    G = spec["G"]
    trans = spec["trans"]
    ist = nz["inf_state"]
    ref = {c: i for i, c in enumerate(ist)}
    NV = len(ist) * G
    V = np.zeros((NV, NV))
    for k, t in enumerate(trans):
        fi, ti = ref.get(t["from"]), ref.get(t["to"])
        if fi is None:
            continue
        for g in range(G):
            j = fi * G + g
            V[j, j] += tr_rate[k, g]
            if ti is not None:
                V[ti * G + g, j] -= tr_rate[k, g]
    Vinv = np.linalg.inv(V)
    F = np.zeros((G, NV))
    for si, co in enumerate(ist):
        if inf[co] == 0:
            continue
        for dp in range(G):
            for q in range(G):
                F[q, si * G + dp] += dist[q] * sus[q] * inf[co] * N[q, dp]
    return (F @ Vinv)[:, :G]


def derive(spec, nz, pv):
    """FP64 tables of State::set_param for parameter vector pv."""
    A, G, period, spd = spec["A"], spec["G"], spec["period"], spec["spd"]
    S = period * spd
    comps, trans = spec["comps"], spec["trans"]
    C, T = len(comps), len(trans)
    pd = dirichlet_correct(spec, pv)
    rate = np.zeros((T, G))
    for k, t in enumerate(trans):
        if t.get("infection"):
            continue
        c = comps[t["from"]]
        shape = c.get("shape", 1)
        for g in range(G):
            r = 1.0 / (pd[c["mean"][g]] / shape)
            if len(nz["exits"][t["from"]]) > 1:
                r *= pd[t["prob"][g]]
            rate[k, g] = r
    sus = np.array([pd[p] if p >= 0 else 1.0 for p in spec["sus"]]) if spec.get("sus") is not None else np.ones(G)
    for f in spec.get("sus_extra") or []:          # a product over categories (src/model.cc:1160-1163)
        sus = sus * np.array([pd[p] if p >= 0 else 1.0 for p in f])
    inf = np.array([pd[c["inf"]] if c.get("inf", -1) >= 0 else 0.0 for c in comps])
    af = np.ones((period, A))
    if spec.get("area_effect") is not None:
        af *= pd[np.asarray(spec["area_effect"])][None, :]
    for par, v in nz["covars"]:
        af *= np.exp(pd[par] * v.T)
    le = spec.get("level_effect")
    if le is not None:
        af *= pd[np.asarray(le["param"])][np.asarray(le["map"])]
    area_av = (af.mean(axis=0) * nz["area_pop"]).sum() / nz["popsize"]
    af /= area_av
    # Model::create_Ntime (src/model.cc:1254-1301): modifications multiply
    # their age groups' rows and columns by the spline value at each step.
    Na = np.repeat(nz["N_age"][None, :, :], S, axis=0)
    for am in spec.get("age_mod") or []:
        f = disc_spline(spec, am["spline"], pd)
        for s in range(S):
            for a in am["ages"]:
                Na[s, a, :] *= f[s]
                Na[s, :, a] *= f[s]
    Ntime = Na[:, nz["dp_age"][:, None], nz["dp_age"][None, :]]
    Ntime = Ntime * (nz["strain"][:, None] == nz["strain"][None, :])[None, :, :]
    if nz["nstrain"] > 1:
        return _derive_strains(spec, nz, pd, rate, sus, inf, af, Ntime, S, A, G)
    # Each R spline's ratio, over the dp distribution of its own areas
    # (Model::calculate_beta_from_R, src/model.cc:1318-1347, with
    # SplineInfo::democatpos_dist, src/model.cc:409-424), recomputed where the
    # matrix changes by more than TINY (src/model.cc:1335-1340, Model::equal
    # 1303-1314); beta of an area is its spline's.
    Rfac = pd[spec["Rfactor"]] if spec.get("Rfactor", -1) >= 0 else 1.0
    beta_SA = np.zeros((S, A))
    ratios = []
    for spl, areas, dist in r_groups(spec, nz):
        ratio = np.zeros(S)
        for s in range(S):
            if s == 0 or np.any(np.abs(Ntime[s] - Ntime[s - 1]) > 1e-8):
                ratio[s] = largest_eigenvalue(ngm(spec, nz, rate, sus, inf, Ntime[s], dist))
            else:
                ratio[s] = ratio[s - 1]
        beta_SA[:, areas] = (Rfac * disc_spline(spec, spl, pd) / ratio)[:, None]
        ratios.append(ratio)
    ratio = ratios[0]
    efoi, agedist, efoi_cell = efoi_tables(spec, nz, pd)
    geo = disc_spline(spec, spec["geo_spline"], pd) if spec.get("geo_spline", -1) >= 0 else np.ones(S)
    bm, em, tm = modification_tables(spec)
    return dict(rate=rate, sus=sus, inf=inf, af=af, beta=beta_SA, efoi=efoi, agedist=agedist, efoi_cell=efoi_cell,
                geo=geo, Ntime=Ntime, ratio=ratio[0] if not spec.get("age_mod") else ratio, beta_mult=bm,
                efoi_mult=em, transmean_mult=tm)


def r_groups(spec, nz):
    """The R splines as (spline, areas, dp distribution of their areas'
    initial populations): spec["R_splines"], else R_spline for every area."""
    A, G = spec["A"], spec["G"]
    rs = spec.get("R_splines")
    if not rs:
        return [(spec["R_spline"], list(range(A)), nz["Rdist"])]
    if nz["nstrain"] > 1 and len(rs) > 1:
        raise ValueError("R splines per area are not supported with strains")
    pop = np.asarray(spec["pop"], dtype=np.float64)                      # [A][C][G]
    out = []
    for e in rs:
        areas = list(e["areas"])
        d = pop[areas].sum(axis=(0, 1))
        out.append((e["spline"], areas, d / d.sum()))
    return out


def efoi_tables(spec, nz, pd):
    """The external force (State::set_transmean, src/state.cc:410-424): efoi
    [S] and agedist [G] of the single efoi_spline, and the force of every
    cell, efoi_cell [S][A][G] = efoi of its (strain, area)'s spline / popsize
    times that spline's age weight, from spec["efoi_splines"] when given
    (Model::update_efoi_ref, src/model.cc:432-480)."""
    A, G = spec["A"], spec["G"]
    S = spec["period"] * spec["spd"]
    es = spec.get("efoi_splines")
    if not es:
        if spec.get("efoi_spline", -1) >= 0:
            efoi = disc_spline(spec, spec["efoi_spline"], pd) / nz["popsize"]
        else:
            efoi = np.zeros(S)
        agedist = np.asarray(spec.get("efoi_agedist", [1.0 / G] * G), dtype=np.float64)
        return efoi, agedist, np.repeat((efoi[:, None] * agedist[None, :])[:, None, :], A, axis=1)
    ns = nz["nstrain"]
    strain = np.asarray(nz["strain"])
    cell = np.zeros((S, A, G))
    seen = np.zeros((ns, A), dtype=int)
    for e in es:
        v = disc_spline(spec, e["spline"], pd) / nz["popsize"]
        ad = np.asarray(e["agedist"], dtype=np.float64)
        for st in e.get("strains", range(ns)):
            for c in e["areas"]:
                seen[st, c] += 1
                g = strain == st
                cell[:, c, g] = v[:, None] * ad[None, g]
    if (seen != 1).any():
        raise ValueError("every (strain, area) needs exactly one external-force spline")
    return np.zeros(S), np.zeros(G), cell


def _derive_strains(spec, nz, pd, rate, sus, inf, af, Ntime, S, A, G):
    """Strains (Model::calculate_beta_from_R, src/model.cc:1318-1348; calculate_Vinv
    and calculate_F per strain, src/model.cc:2485-2556): each strain's NGM over
    its own dps with the population fractions pooled over strains and that
    strain's susceptibility; beta_st = Rfactor_st * R(t) / ratio_st.  Returned
    in the device's form: Rfactor_st / ratio_st folded into the strain's block
    of Ntime and beta = R(t) (times the area factor by the caller)."""
    ns = nz["nstrain"]
    dpm = G // ns
    Rd = np.zeros(dpm)
    for g in range(G):
        Rd[g % dpm] += nz["Rdist"][g]
    ist = nz["inf_state"]
    ref = {c: i for i, c in enumerate(ist)}
    NV = len(ist) * dpm
    ratio = np.zeros((S, ns))
    for st in range(ns):
        V = np.zeros((NV, NV))
        for k, t in enumerate(spec["trans"]):
            fi, ti = ref.get(t["from"]), ref.get(t["to"])
            if fi is None:
                continue
            for q in range(dpm):
                V[fi * dpm + q, fi * dpm + q] += rate[k, st * dpm + q]
                if ti is not None:
                    V[ti * dpm + q, fi * dpm + q] -= rate[k, st * dpm + q]
        Vinv = np.linalg.inv(V)
        for s in range(S):
            if s > 0 and not np.any(np.abs(Ntime[s] - Ntime[s - 1]) > 1e-8):
                ratio[s, st] = ratio[s - 1, st]
                continue
            F = np.zeros((dpm, NV))
            for si, co in enumerate(ist):
                if inf[co] == 0:
                    continue
                for d in range(dpm):
                    for q in range(dpm):
                        F[q, si * dpm + d] += Rd[q] * sus[st * dpm + q] * inf[co] * Ntime[s, st * dpm + q, st * dpm + d]
            ratio[s, st] = largest_eigenvalue((F @ Vinv)[:, :dpm])
    Rfac = np.array([pd[p] if p >= 0 else 1.0 for p in spec["strain_Rfactor"]])
    fold = (Rfac[None, :] / ratio)[:, nz["strain"]]                    # [S][G]
    Ntime = Ntime * fold[:, :, None]
    ((spl, _, _),) = r_groups(spec, nz)
    beta = disc_spline(spec, spl, pd)
    beta_SA = np.repeat(beta[:, None], A, axis=1)
    efoi, agedist, efoi_cell = efoi_tables(spec, nz, pd)
    geo = disc_spline(spec, spec["geo_spline"], pd) if spec.get("geo_spline", -1) >= 0 else np.ones(S)
    bm, em, tm = modification_tables(spec)
    return dict(rate=rate, sus=sus, inf=inf, af=af, beta=beta_SA, efoi=efoi, agedist=agedist, efoi_cell=efoi_cell,
                geo=geo, Ntime=Ntime, ratio=ratio, beta_mult=bm, efoi_mult=em, transmean_mult=tm)


def prior(spec, pv):
    Pr = 0.0
    for p, v in zip(spec["params"], pv):
        pr = p["prior"]
        if pr[0] == "uniform":
            Pr += np.log(1.0 / (pr[2] - pr[1]))
        elif pr[0] == "exp":
            Pr += np.log(pr[1]) - pr[1] * v
        elif pr[0] == "normal":
            Pr += -0.5 * np.log(2 * np.pi * pr[2] ** 2) - (v - pr[1]) ** 2 / (2 * pr[2] ** 2)
        elif pr[0] == "dirflat":
            Pr += -v
        elif pr[0] == "gamma":          # gamma_probability, src/utils.cc:115-120
            a, b = pr[1], pr[2]
            Pr += (a - 1) * np.log(v) - b * v + a * np.log(b) - math.lgamma(a)
    return Pr + spline_prior(spec, pv)


SMOOTH_NONE, SMOOTH_NORMAL, SMOOTH_LOG = 0, 1, 2


def spline_prior(spec, pv):
    """Model::spline_prior (src/model.cc:1662-1704): between neighbouring knots
    with different parameters, times and not both fixed, a normal density on
    the difference (smooth) or a log-normal one on the ratio (log_smooth),
    on the Dirichlet-corrected values."""
    pd = dirichlet_correct(spec, pv)
    Pr = 0.0
    for sp in spec["splines"]:
        kind = sp.get("smooth_type", SMOOTH_NONE)
        if kind == SMOOTH_NONE:
            continue
        par, t, mf, sm = sp["param"], sp["t"], sp.get("multfac", [1.0] * len(sp["t"])), sp["smooth"]
        for i in range(len(par) - 1):
            p1, p2 = par[i], par[i + 1]
            if p1 == p2 or t[i] == t[i + 1]:
                continue
            if spec["params"][p1]["prior"][0] == "fixed" and spec["params"][p2]["prior"][0] == "fixed":
                continue
            v = sm[i]
            if kind == SMOOTH_LOG:
                mu = np.log(pd[p1] * mf[i])
                logx = np.log(pd[p2] * mf[i + 1])
                Pr += -(logx - mu) ** 2 / (2 * v * v) - logx
            else:
                d = pd[p1] - pd[p2]
                Pr += -d * d / (2 * v * v)
    return Pr


def inbounds(spec, pv):
    for p, v in zip(spec["params"], pv):
        pr = p["prior"]
        if pr[0] == "fixed" and v != pr[1]:
            return False
        if pr[0] == "uniform" and (v < pr[1] or v > pr[2]):
            return False
        if pr[0] in ("exp", "dirflat", "gamma") and v < 0:
            return False
    return True


def obs_sd(spec):
    obs = spec.get("obs", [])
    gmax = {}
    for o in obs:
        y = o["y"]
        if y not in (999999997.0, 999999998.0):
            gmax[o.get("graph", 0)] = max(gmax.get(o.get("graph", 0), 0.0), y)
    sd = []
    for o in obs:
        if o["type"] == "normal_percent" and not (o.get("sd", 0) > 0):
            sd.append((o["y"] * (1 - EPSILON_FACTOR) + gmax.get(o.get("graph", 0), 0.0) * EPSILON_FACTOR)
                      * 0.01 * o["percent"])
        else:
            sd.append(o.get("sd", 0.0))
    return np.array(sd)
