"""TEST INFRASTRUCTURE ONLY: Python side of the FP64 CPU oracle.

Wraps liboracle.so (beep_oracle.cc, the restatement of State / ObservationModel
/ Mbp) and the numpy set-up restatement (derive.py).  Imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from . import derive

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libbeepref_utils.so")


def build(ref=False):
    """Compiles liboracle.so (and, if asked and the reference tree is present,
    the reference's utils.cc into _ref/)."""
    targets = ["all"] + (["ref"] if ref and os.path.isdir("/root/reference/src") else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.or_seed.argtypes = [C.c_uint]
        for f in ("or_simulate",):
            getattr(L, f).argtypes = [C.c_void_p, C.c_void_p, dp, C.c_void_p, C.c_int, C.c_int]
        L.or_replay.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_void_p]
        L.or_likelihood.argtypes = [C.c_void_p, C.c_void_p]
        L.or_likelihood.restype = C.c_double
        L.or_obs_loglik.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, ip, C.c_int, dp]
        L.or_obs_loglik.restype = C.c_double
        L.or_obs_prob.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double]
        L.or_obs_prob.restype = C.c_double
        L.or_mbp.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_void_p, C.c_void_p, C.POINTER(C.c_ubyte), C.c_int]
        L.or_get_al.argtypes = [C.c_int] + [C.c_double] * 6
        L.or_get_al.restype = C.c_double
        L.or_poisson_stream.argtypes = [C.c_uint, C.c_double, C.c_int, ip]
        L.or_binomial_stream.argtypes = [C.c_uint, C.c_double, C.c_uint, C.c_int, ip]
        for f in ("or_normal_probability",):
            getattr(L, f).argtypes = [C.c_double] * 3
            getattr(L, f).restype = C.c_double
        L.or_poisson_probability.argtypes = [C.c_int, C.c_double]
        L.or_poisson_probability.restype = C.c_double
        L.or_negative_binomial_probability.argtypes = [C.c_uint, C.c_double, C.c_double]
        L.or_negative_binomial_probability.restype = C.c_double
        L.or_update_pop.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        _lib = L
    return _lib


class _OrModel(C.Structure):
    _fields_ = [("A", C.c_int), ("G", C.c_int), ("C", C.c_int), ("T", C.c_int), ("S", C.c_int), ("spd", C.c_int),
                ("inf_tr", C.c_int), ("stochastic", C.c_int), ("dt", C.c_double),
                ("from_", C.POINTER(C.c_int)), ("to", C.POINTER(C.c_int)), ("Mdiag", C.POINTER(C.c_double)),
                ("Mptr", C.POINTER(C.c_int)), ("Mto", C.POINTER(C.c_int)), ("Mval", C.POINTER(C.c_double)),
                ("Mfac", C.POINTER(C.c_double)), ("nstrain", C.c_int)]


class _OrTables(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("inf", "rate", "sus", "beta", "af", "efoi", "agedist", "geo",
                                                     "Ntime", "beta_mult", "efoi_mult", "transmean_mult",
                                                     "efoi_cell")]


class _OrState(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("pop", "transnum", "transmean", "Imap", "Idiag")]


class _OrObs(C.Structure):
    _fields_ = [("nobs", C.c_int)] + [(n, C.POINTER(C.c_double)) for n in ("y", "sd")] + \
               [("type", C.POINTER(C.c_int))] + [(n, C.POINTER(C.c_double)) for n in ("shape", "factor", "thresh")] + \
               [(n, C.POINTER(C.c_int)) for n in ("el_ptr", "el_kind", "el_s", "el_c", "el_x", "el_dp")] + \
               [("el_w", C.POINTER(C.c_double))]


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


OBS_TYPES = {"normal": 0, "normal_percent": 1, "poisson": 2, "negbin": 3}


class State:
    """Trajectory arrays in the reference's order."""

    def __init__(self, A, G, C_, T, S):
        self.pop = np.zeros((S, A, C_, G))
        self.transnum = np.zeros((S, A, T, G))
        self.transmean = np.zeros((S, A, T, G))
        self.Imap = np.zeros((S, A * G))
        self.Idiag = np.zeros((S, A * G))
        self._c = _OrState(_d(self.pop), _d(self.transnum), _d(self.transmean), _d(self.Imap), _d(self.Idiag))
        self.tb = None          # the tables the trajectory was made with


class Oracle:
    def __init__(self, spec, stochastic=True):
        self.spec = spec
        self.nz = derive.normalise(spec)
        A, G = spec["A"], spec["G"]
        C_, T = len(spec["comps"]), len(spec["trans"])
        S = spec["period"] * spec["spd"]
        self.A, self.G, self.C, self.T, self.S = A, G, C_, T, S
        nz = self.nz
        self._from = np.array([t["from"] for t in spec["trans"]], dtype=np.int32)
        self._to = np.array([t["to"] for t in spec["trans"]], dtype=np.int32)
        self._diag = np.ascontiguousarray(nz["diag"], dtype=np.float64)
        self._ptr = np.ascontiguousarray(nz["rowptr"], dtype=np.int32)
        self._mto = np.ascontiguousarray(nz["to"], dtype=np.int32) if len(nz["to"]) else np.zeros(1, np.int32)
        self._val = np.ascontiguousarray(nz["val"], dtype=np.float64) if len(nz["val"]) else np.zeros(1)
        self._fac = np.ascontiguousarray(nz["factor"], dtype=np.float64)
        self.model = _OrModel(A, G, C_, T, S, spec["spd"], nz["itr"], 1 if stochastic else 0,
                              spec["period"] / S, _i(self._from), _i(self._to), _d(self._diag), _i(self._ptr),
                              _i(self._mto), _d(self._val), _d(self._fac), len(spec.get("strain_Rfactor") or [1]))
        self.pop_init = np.ascontiguousarray(np.asarray(spec["pop"], dtype=np.float64).reshape(A, C_, G))
        self._build_obs()

    def _build_obs(self):
        spec = self.spec
        obs = spec.get("obs", [])
        sd = derive.obs_sd(spec)
        ptr, kind, s_, c_, x_, dp_, sp_ = [0], [], [], [], [], [], []
        for o in obs:
            n0 = len(kind)
            for s in range(o["sett_i"], o["sett_f"]):
                for c in o["area"]:
                    for tr in o.get("tr", []):
                        for dp in o["dp"]:
                            kind.append(0); s_.append(s); c_.append(c); x_.append(tr); dp_.append(dp)
                    for co in o.get("co", []):
                        for dp in o["dp"]:
                            kind.append(1); s_.append(s); c_.append(c); x_.append(co); dp_.append(dp)
            sp_ += [o.get("spline", -1)] * (len(kind) - n0)
            ptr.append(len(kind))
        self._el_s, self._el_sp = np.array(s_, dtype=np.int64), np.array(sp_, dtype=np.int64)
        self._obs_arrays = [
            np.array([o["y"] for o in obs] or [0.0]), np.ascontiguousarray(sd if len(sd) else np.zeros(1)),
            np.array([OBS_TYPES[o["type"]] if isinstance(o["type"], str) else o["type"] for o in obs] or [0],
                     dtype=np.int32),
            np.array([o.get("shape", 1.0) for o in obs] or [1.0]), np.array([o.get("factor", 1.0) for o in obs] or [1.0]),
            np.array([o.get("thresh", 0.0) for o in obs] or [0.0]),
            np.array(ptr, dtype=np.int32), np.array(kind or [0], dtype=np.int32), np.array(s_ or [0], dtype=np.int32),
            np.array(c_ or [0], dtype=np.int32), np.array(x_ or [0], dtype=np.int32), np.array(dp_ or [0], dtype=np.int32),
        ]
        a = self._obs_arrays
        self.obs = _OrObs(len(obs), _d(a[0]), _d(a[1]), _i(a[2]), _d(a[3]), _d(a[4]), _d(a[5]),
                          _i(a[6]), _i(a[7]), _i(a[8]), _i(a[9]), _i(a[10]), _i(a[11]), None)
        self.obs_sd = sd

    def tables(self, pv):
        """State::set_param (FP64) as the oracle's table struct (arrays kept alive on it)."""
        d = derive.derive(self.spec, self.nz, np.asarray(pv, dtype=np.float64))
        arrs = [np.ascontiguousarray(d[k], dtype=np.float64) for k in
                ("inf", "rate", "sus", "beta", "af", "efoi", "agedist", "geo", "Ntime")]
        # Prediction-mode modifications: null pointers when the spec has none.
        mods = [None if d.get(k) is None else np.ascontiguousarray(d[k], dtype=np.float64)
                for k in ("beta_mult", "efoi_mult", "transmean_mult")]
        # External-force splines per strain or area: each cell's force (null
        # without them, where efoi * agedist is used).
        mods.append(np.ascontiguousarray(d["efoi_cell"]) if self.spec.get("efoi_splines") else None)
        t = _OrTables(*[_d(a) for a in arrs], *[None if a is None else _d(a) for a in mods])
        t._keep = arrs + [a for a in mods if a is not None]
        t.derived = d
        t.pv = np.asarray(pv, dtype=np.float64)
        return t

    def new_state(self):
        return State(self.A, self.G, self.C, self.T, self.S)

    def simulate(self, tb, seed, ti=0, tf=None, state=None):
        """State::simulate(ti, tf) with the libstdc++ engine seeded by `seed`."""
        L = lib()
        st = state or self.new_state()
        L.or_seed(seed)
        rc = L.or_simulate(C.byref(self.model), C.byref(tb), _d(self.pop_init), C.byref(st._c), ti,
                           self.S if tf is None else tf)
        if rc:
            raise RuntimeError(f"oracle simulate failed: {rc}")
        st.tb = tb
        return st

    def replay(self, tb, transnum):
        st = self.new_state()
        st.transnum[...] = transnum
        rc = lib().or_replay(C.byref(self.model), C.byref(tb), _d(self.pop_init), C.byref(st._c))
        if rc:
            raise RuntimeError(f"oracle replay failed: {rc}")
        st.tb = tb
        return st

    def likelihood(self, st):
        return lib().or_likelihood(C.byref(self.model), C.byref(st._c))

    def obs_loglik(self, st, ti=0, tf=None, sel=None, tb=None):
        """ObservationModel::calculate over steps [ti, tf); observation
        splines take their values from tb (default the tables the state was
        simulated or replayed with)."""
        values = np.zeros(max(1, len(self.spec.get("obs", []))))
        w = None
        if len(self._el_sp) and (self._el_sp >= 0).any():
            tb = tb if tb is not None else st.tb
            spl = {sp: derive.disc_spline(self.spec, sp, derive.dirichlet_correct(self.spec, tb.pv))
                   for sp in set(self._el_sp[self._el_sp >= 0].tolist())}
            w = np.ones(len(self._el_sp))
            for sp, v in spl.items():
                sel_e = self._el_sp == sp
                w[sel_e] = v[self._el_s[sel_e]]
        self._w = w
        self.obs.el_w = _d(w) if w is not None else None
        if sel is not None:
            sel = np.ascontiguousarray(sel, dtype=np.int32)
        L = lib().or_obs_loglik(C.byref(self.model), C.byref(self.obs), C.byref(st._c), ti,
                                self.S if tf is None else tf, _i(sel) if sel is not None else None,
                                len(sel) if sel is not None else 0, _d(values))
        return L, values[:len(self.spec.get("obs", []))]

    def mbp(self, tbp, initial, seed, simu=None, inf_update=1):
        L = lib()
        pro = self.new_state()
        L.or_seed(seed)
        mp = None
        if simu is not None:
            simu = np.ascontiguousarray(simu, dtype=np.uint8)
            mp = simu.ctypes.data_as(C.POINTER(C.c_ubyte))
        rc = L.or_mbp(C.byref(self.model), C.byref(tbp), _d(self.pop_init), C.byref(initial._c), C.byref(pro._c), mp,
                      inf_update)
        if rc:
            raise RuntimeError(f"oracle mbp failed: {rc}")
        pro.tb = tbp
        return pro


def bootstrap(L, invT, u):
    """PMCMC::bootstrap (src/pmcmc.cc:211-268), sequentially in FP64: the
    weights exp(invT (L - Lmax)) and their running sum (224-233), the
    estimate invT Lmax + log(sum / N) (235), draw p at z = u[p] * sum found by
    the linear search 'first pp with z <= wsum[pp]' (245-248), the first draw
    of a particle keeps it in place and later ones become extras (250-251),
    and the particles no draw chose take extras from the back, in increasing
    order (254-259).  Returns (ancestors, estimate term)."""
    import bisect
    import math
    N = len(L)
    Lmax = max(L)
    wsum, tot = [], 0.0
    for p in range(N):
        tot += math.exp(invT * (L[p] - Lmax))
        wsum.append(tot)
    obprob = invT * Lmax + math.log(tot / N)
    bp = [-1] * N
    extra = []
    for p in range(N):
        z = u[p] * tot
        pp = bisect.bisect_left(wsum, z)          # == the linear search on nondecreasing wsum
        if pp == N:
            raise RuntimeError("bootstrap: draw beyond the weights")
        if bp[pp] == -1:
            bp[pp] = pp
        else:
            extra.append(pp)
    for p in range(N):
        if bp[p] == -1:
            bp[p] = extra.pop()
    if extra:
        raise RuntimeError("bootstrap: extras left over")
    return np.array(bp, dtype=np.int32), obprob


def particle_sample(backpart, sections, traj):
    """Mpi::particle_sample (src/mpi.cc:331-385) on host arrays: backpart
    [(nsec+1)][N]; traj[p] an array [S, ...] of particle p's per-step rows.
    Returns the gathered [S, ...] trajectory."""
    nsec = len(sections)
    out = np.array(traj[0], copy=True)
    p = backpart[nsec][0]
    for sec in range(nsec - 1, -1, -1):
        ti, tf = sections[sec]
        out[ti:tf] = traj[p][ti:tf]
        p = backpart[sec][p]
    return out


def pmcmc_obs_prob(orc, tb, N, invT, seeds, uniforms):
    """PMCMC::obs_prob (src/pmcmc.cc:180-207) with the oracle's states: per
    section, each particle continues from its ancestor's populations at the
    section start (Mpi::end_sec_swap, src/mpi.cc:215-327), simulates the
    section with libstdc++ draws seeded by seeds[sec][p], is scored by the
    section's observations, and the bootstrap draws with uniforms[sec].
    Returns (estimate, backpart, states)."""
    sections = orc.spec["sections"]
    obs = orc.spec.get("obs", [])
    states = [orc.new_state() for _ in range(N)]
    backpart = [list(range(N))]
    ob = 0.0
    for sec, (ti, tf) in enumerate(sections):
        if sec > 0:
            bp = backpart[sec]
            start = [st.pop[ti].copy() for st in states]
            for p in range(N):
                states[p].pop[ti] = start[bp[p]]
        sel = [i for i, o in enumerate(obs) if o["sett_i"] >= ti and o["sett_f"] <= tf]
        L = []
        for p in range(N):
            orc.simulate(tb, seed=int(seeds[sec][p]), ti=ti, tf=tf, state=states[p])
            L.append(orc.obs_loglik(states[p], ti, tf, sel=sel)[0])
        bp, term = bootstrap(L, invT, uniforms[sec])
        backpart.append(list(bp))
        ob += term
    return ob, np.array(backpart, dtype=np.int32), states


def mvn_probability(pend, pstart, inv_M, size):
    """MVN::probability (src/mvn.cc:185-195) over the MVN's variables, in the
    reference's summation order."""
    s = 0.0
    nvar = len(pend)
    for v1 in range(nvar):
        for v2 in range(nvar):
            val1 = pend[v1] - pstart[v1]
            val2 = pend[v2] - pstart[v2]
            s += (val1 / size) * inv_M[v1][v2] * (val2 / size)
    return -0.5 * s


def abcsmc_weight(theta, prev, w_prev, inv_M, size, logprior):
    """ABCSMC::calculate_particle_weight (src/abcsmc.cc:260-271)."""
    import math
    s = 0.0
    for j in range(len(prev)):
        s += w_prev[j] * math.exp(mvn_probability(theta, prev[j], inv_M, size))
    return math.exp(logprior) / s


def mc3_swap(run, invT, EF, u):
    """MC3::swap_states (src/mc3.cc:153-215) on one set of chains: five sweeps
    of neighbouring swaps within a run, draw k compared with u[k].  Returns
    (order, ntr, nac, EF after the swaps)."""
    import math
    n = len(run)
    order, ntr, nac, EF = list(range(n)), [0] * n, [0] * n, list(EF)
    k = 0
    for _loop in range(5):
        for i in range(n - 1):
            if run[i] == run[i + 1]:
                x = 0.5 * (invT[i] - invT[i + 1]) * (EF[i] - EF[i + 1])
                al = math.exp(x) if x < 709.0 else math.inf      # C++ exp overflows to +inf: accepted
                ntr[i] += 1
                if u[k] < al:
                    nac[i] += 1
                    order[i], order[i + 1] = order[i + 1], order[i]
                    EF[i], EF[i + 1] = EF[i + 1], EF[i]
                k += 1
    return order, ntr, nac, EF


def get_al(mode, EFp, EFi, Prp, Pri, invT=1.0, EFcut=0.0):
    return lib().or_get_al(mode, EFp, EFi, Prp, Pri, invT, EFcut)


def poisson_stream(seed, lam, n):
    out = np.zeros(n, dtype=np.int32)
    lib().or_poisson_stream(seed, lam, n, _i(out))
    return out


def binomial_stream(seed, p, nn, n):
    out = np.zeros(n, dtype=np.int32)
    lib().or_binomial_stream(seed, p, nn, n, _i(out))
    return out


def normal_probability(x, mean, var):
    return lib().or_normal_probability(x, mean, var)


def poisson_probability(i, lam):
    return lib().or_poisson_probability(i, lam)


def negative_binomial_probability(v, m, shape):
    return lib().or_negative_binomial_probability(v, m, shape)


def obs_prob(value, y, sd, type_, shape=1.0, thresh=0.0):
    return lib().or_obs_prob(value, y, sd, type_, shape, thresh)


_ref = None


def ref_lib():
    """The reference's own src/utils.cc (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        L = C.CDLL(REF_LIB)
        ip = C.POINTER(C.c_int)
        L.ref_poisson_stream.argtypes = [C.c_int, C.c_double, C.c_int, ip]
        L.ref_binomial_stream.argtypes = [C.c_int, C.c_double, C.c_uint, C.c_int, ip]
        L.ref_normal_probability.argtypes = [C.c_double] * 3
        L.ref_normal_probability.restype = C.c_double
        L.ref_poisson_probability.argtypes = [C.c_int, C.c_double]
        L.ref_poisson_probability.restype = C.c_double
        L.ref_negative_binomial_probability.argtypes = [C.c_uint, C.c_double, C.c_double]
        L.ref_negative_binomial_probability.restype = C.c_double
        L.ref_gamma_probability.argtypes = [C.c_double] * 3
        L.ref_gamma_probability.restype = C.c_double
        _ref = L
    return _ref
