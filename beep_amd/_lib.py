"""ctypes binding of the C ABI in include/beep_gpu.h (libbeep_gpu.so).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point."""
import ctypes as C
import os

import numpy as np

from . import spec as specmod

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbeep_gpu.so")

BEEP_OK = 0
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ENAN", -5: "EBOUNDS", -6: "ELARGE"}
KEEP_TRANSNUM, KEEP_TRANSMEAN, KEEP_POP, KEEP_ALL = 1, 2, 4, 7
STEPWISE = 16
GENERIC_KERNEL = 32
SIM_PLAIN = 2
PATH_STEPWISE, PATH_FUSED, PATH_FUSED_SPEC, PATH_STEPWISE_SPEC = 0, 1, 2, 3

# Exported symbols declared by include/beep_gpu.h.
SYMBOLS = [
    "beep_spec_tables", "beep_create", "beep_destroy", "beep_last_error", "beep_get_dims", "beep_device_bytes", "beep_set_param",
    "beep_get_tables", "beep_simulate", "beep_obs_loglik", "beep_obs_values", "beep_pois_loglik", "beep_mbp",
    "beep_accept", "beep_get_transnum", "beep_get_transmean", "beep_get_pop", "beep_get_state", "beep_put_state",
    "beep_put_transnum", "beep_copy_instances", "beep_synchronize", "beep_debug_philox", "beep_debug_poisson",
    "beep_debug_binomial", "beep_timing", "beep_debug_gemm", "beep_run_batch", "beep_copy_states",
    "beep_derive_slots", "beep_sum_transnum", "beep_pmcmc_obs_prob", "beep_pmcmc_particle_sample", "beep_bootstrap", "beep_path",
    "beep_abcsmc_weights", "beep_mc3_swap", "beep_debug_check_launch", "beep_batch_accept", "beep_batch_mc3_swap",
    "beep_batch_set_param", "beep_simulate_noalloc",
]


class BeepError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Prior(C.Structure):
    _fields_ = [("type", C.c_int32), ("a", C.c_double), ("b", C.c_double)]


class Comp(C.Structure):
    _fields_ = [("inf_param", C.c_int32), ("mean_param", C.POINTER(C.c_int32)), ("shape", C.c_int32)]


class Trans(C.Structure):
    _fields_ = [("from_", C.c_int32), ("to", C.c_int32), ("infection", C.c_int32),
                ("prob_param", C.POINTER(C.c_int32))]


class Spline(C.Structure):
    _fields_ = [("nknot", C.c_int32), ("t", C.POINTER(C.c_double)), ("param", C.POINTER(C.c_int32)),
                ("multfac", C.POINTER(C.c_double)), ("factor_param", C.c_int32), ("smooth_type", C.c_int32),
                ("smooth", C.POINTER(C.c_double))]


class Dirichlet(C.Structure):
    _fields_ = [("n", C.c_int32), ("param", C.POINTER(C.c_int32)), ("frac", C.POINTER(C.c_double))]


class Covar(C.Structure):
    _fields_ = [("param", C.c_int32), ("value", C.POINTER(C.c_double))]


class Obs(C.Structure):
    _fields_ = [("y", C.c_double), ("sd", C.c_double), ("shape", C.c_double), ("factor", C.c_double),
                ("thresh", C.c_double), ("percent", C.c_double), ("type", C.c_int32), ("datatable", C.c_int32),
                ("graph", C.c_int32), ("sett_i", C.c_int32), ("sett_f", C.c_int32),
                ("narea", C.c_int32), ("area", C.POINTER(C.c_int32)), ("ndp", C.c_int32), ("dp", C.POINTER(C.c_int32)),
                ("ntr", C.c_int32), ("tr", C.POINTER(C.c_int32)), ("nco", C.c_int32), ("co", C.POINTER(C.c_int32)),
                ("factor_spline", C.c_int32)]


class AgeMod(C.Structure):
    _fields_ = [("spline", C.c_int32), ("nage", C.c_int32), ("ages", C.POINTER(C.c_int32))]


class Mod(C.Structure):
    _fields_ = [("type", C.c_int32), ("start", C.c_int32), ("end", C.c_int32), ("factor", C.c_double),
                ("ntr", C.c_int32), ("tr", C.POINTER(C.c_int32)), ("narea", C.c_int32), ("area", C.POINTER(C.c_int32)),
                ("ndp", C.c_int32), ("dp", C.POINTER(C.c_int32)), ("spline", C.c_int32), ("strain", C.c_int32)]


MOD_TYPES = {"trans_rate_fac": 0, "beta_fac": 1, "efoi_fac": 2, "spline_fac": 3, "spline_set": 4}


class Spec(C.Structure):
    _fields_ = [("A", C.c_int32), ("G", C.c_int32), ("C", C.c_int32), ("T", C.c_int32),
                ("period", C.c_int32), ("steps_per_day", C.c_int32), ("nparam", C.c_int32),
                ("prior", C.POINTER(Prior)), ("comp", C.POINTER(Comp)), ("trans", C.POINTER(Trans)),
                ("nspline", C.c_int32), ("spline", C.POINTER(Spline)), ("R_spline", C.c_int32),
                ("Rfactor_param", C.c_int32), ("efoi_spline", C.c_int32), ("efoi_agedist", C.POINTER(C.c_double)),
                ("geo_spline", C.c_int32), ("N_raw", C.POINTER(C.c_double)), ("M_diag", C.POINTER(C.c_double)),
                ("M_rowptr", C.POINTER(C.c_int32)), ("M_to", C.POINTER(C.c_int32)), ("M_val", C.POINTER(C.c_double)),
                ("pop_init", C.POINTER(C.c_int32)), ("ncovar", C.c_int32), ("covar", C.POINTER(Covar)),
                ("sus_param", C.POINTER(C.c_int32)), ("ndirichlet", C.c_int32), ("dirichlet", C.POINTER(Dirichlet)),
                ("nobs", C.c_int32), ("obs", C.POINTER(Obs)), ("nsection", C.c_int32),
                ("section_ti", C.POINTER(C.c_int32)), ("section_tf", C.POINTER(C.c_int32)),
                ("area_effect_param", C.POINTER(C.c_int32)), ("nlevel", C.c_int32),
                ("level_param", C.POINTER(C.c_int32)), ("level_map", C.POINTER(C.c_int32)),
                ("nagemod", C.c_int32), ("agemod", C.POINTER(AgeMod)),
                ("nage", C.c_int32), ("dp_age", C.POINTER(C.c_int32)),
                ("nsus_extra", C.c_int32), ("sus_extra", C.POINTER(C.c_int32)),
                ("nstrain", C.c_int32), ("strain_Rfactor", C.POINTER(C.c_int32)),
                ("nmod", C.c_int32), ("mod", C.POINTER(Mod)),
                ("nrgroup", C.c_int32), ("rgroup_spline", C.POINTER(C.c_int32)), ("rgroup_area", C.POINTER(C.c_int32)),
                ("negroup", C.c_int32), ("egroup_spline", C.POINTER(C.c_int32)), ("egroup_cell", C.POINTER(C.c_int32)),
                ("egroup_agedist", C.POINTER(C.c_double))]


_lib = None


def lib():
    """The loaded library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i32p, u32p, f64p, f32p, u8p = (C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                                      C.POINTER(C.c_float), C.POINTER(C.c_uint8))
        sig = {
            "beep_spec_tables": (C.c_int, [C.POINTER(Spec), f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p,
                                           i32p, C.c_char_p, C.c_int]),
            "beep_create": (C.c_int, [C.POINTER(Spec), C.c_int, C.c_size_t, C.c_int, C.c_int, C.c_int, C.POINTER(P)]),
            "beep_destroy": (C.c_int, [P]),
            "beep_last_error": (C.c_char_p, [P]),
            "beep_get_dims": (C.c_int, [P, i32p]),
            "beep_device_bytes": (C.c_size_t, [P]),
            "beep_set_param": (C.c_int, [P, C.c_int, i32p, f64p, f64p, i32p]),
            "beep_get_tables": (C.c_int, [P, C.c_int, f64p, f64p, f64p, f64p, f64p, f64p, f64p]),
            "beep_simulate": (C.c_int, [P, C.c_int, i32p, i32p, u32p, C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_int]),
            "beep_simulate_noalloc": (C.c_int, [P, C.c_int, i32p, i32p, u32p, C.c_int, C.c_int, C.c_uint64,
                                                C.c_uint32]),
            "beep_obs_loglik": (C.c_int, [P, C.c_int, i32p, C.c_int, f64p, f64p]),
            "beep_obs_values": (C.c_int, [P, C.c_int, f64p]),
            "beep_pois_loglik": (C.c_int, [P, C.c_int, i32p, f64p]),
            "beep_mbp": (C.c_int, [P, C.c_int, i32p, i32p, i32p, u32p, u8p, C.c_uint64, C.c_uint32, C.c_int]),
            "beep_accept": (C.c_int, [P, C.c_int, i32p, i32p, C.c_int, C.c_double, C.c_double, f64p, u32p, C.c_uint64,
                                      C.c_uint32, f64p, u8p]),
            "beep_get_transnum": (C.c_int, [P, C.c_int, i32p]),
            "beep_get_transmean": (C.c_int, [P, C.c_int, f32p]),
            "beep_get_pop": (C.c_int, [P, C.c_int, i32p]),
            "beep_get_state": (C.c_int, [P, C.c_int, i32p]),
            "beep_put_state": (C.c_int, [P, C.c_int, i32p]),
            "beep_put_transnum": (C.c_int, [P, C.c_int, C.c_int, i32p]),
            "beep_copy_instances": (C.c_int, [P, C.c_int, i32p, i32p]),
            "beep_synchronize": (C.c_int, [P]),
            "beep_timing": (C.c_int, [P, f64p]),
            "beep_debug_philox": (C.c_int, [C.c_int, C.c_int, u32p, C.c_uint32, C.c_uint32, u32p]),
            "beep_debug_poisson": (C.c_int, [C.c_int, C.c_int, f32p, C.c_uint64, C.c_uint32, i32p]),
            "beep_debug_binomial": (C.c_int, [C.c_int, C.c_int, i32p, f64p, C.c_uint64, C.c_uint32, i32p]),
            "beep_debug_gemm": (C.c_int, [C.c_int, C.c_int, C.c_int, f32p, f32p, f32p]),
            "beep_run_batch": (C.c_int, [C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, f64p, f64p]),
            "beep_batch_accept": (C.c_int, [C.c_int, C.c_void_p, C.c_int, f64p, C.c_double, f64p, C.c_uint64,
                                            C.c_uint32, f64p, u8p]),
            "beep_batch_mc3_swap": (C.c_int, [C.c_int, i32p, C.c_void_p, i32p, i32p, f64p, C.c_uint64, C.c_uint32,
                                              i32p, i32p, i32p]),
            "beep_batch_set_param": (C.c_int, [C.c_int, C.c_void_p, i32p, C.c_void_p, f64p, i32p]),
            "beep_copy_states": (C.c_int, [P, C.c_int, i32p, i32p]),
            "beep_derive_slots": (C.c_int, [P, C.c_int, i32p, i32p, f64p, i32p]),
            "beep_sum_transnum": (C.c_int, [P, C.c_int, i32p, C.c_int, C.c_int, C.c_int, f64p]),
            "beep_pmcmc_obs_prob": (C.c_int, [P, C.c_int, i32p, C.c_int, u32p, C.c_double, C.c_uint64, C.c_uint32,
                                              f64p, i32p]),
            "beep_pmcmc_particle_sample": (C.c_int, [P, C.c_int]),
            "beep_path": (C.c_int, [P]),
            "beep_abcsmc_weights": (C.c_int, [C.c_int, C.c_int, C.c_int, f64p, f64p, C.c_int, f64p, f64p,
                                              C.c_double, f64p, f64p]),
            "beep_mc3_swap": (C.c_int, [P, C.c_int, i32p, i32p, f64p, C.c_uint64, C.c_uint32, i32p, i32p, i32p]),
            "beep_bootstrap": (C.c_int, [C.c_int, C.c_int, f64p, C.c_double, C.c_uint64, C.c_uint32, C.c_int, i32p,
                                         f64p]),
            "beep_debug_check_launch": (C.c_int, [C.c_int, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_int64),
                                                  C.c_char_p, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _u32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))


class _SpecBuffers:
    """Holds the numpy arrays a ctypes Spec points into."""

    def __init__(self, spec):
        keep = self.keep = []

        def i32(x):
            a = _i32(x)
            keep.append(a)
            return _p(a, C.c_int32)

        def f64(x):
            a = _f64(x)
            keep.append(a)
            return _p(a, C.c_double)

        null_i32 = C.POINTER(C.c_int32)()
        null_f64 = C.POINTER(C.c_double)()
        A, G = spec["A"], spec["G"]
        comps, trans, params = spec["comps"], spec["trans"], spec["params"]
        s = Spec()
        s.A, s.G, s.C, s.T = A, G, len(comps), len(trans)
        s.period, s.steps_per_day = spec["period"], spec["spd"]
        s.nparam = len(params)
        pri = (Prior * max(1, len(params)))()
        for i, p in enumerate(params):
            pr = p["prior"]
            pri[i].type = specmod.PRIOR_CODES[pr[0]]
            pri[i].a = float(pr[1]) if len(pr) > 1 else 0.0
            pri[i].b = float(pr[2]) if len(pr) > 2 else 0.0
        keep.append(pri)
        s.prior = pri
        cs = (Comp * len(comps))()
        for i, c in enumerate(comps):
            cs[i].inf_param = c.get("inf", -1)
            cs[i].mean_param = i32(c["mean"]) if c.get("mean") is not None else null_i32
            cs[i].shape = c.get("shape", 1)
        keep.append(cs)
        s.comp = cs
        ts = (Trans * len(trans))()
        for i, t in enumerate(trans):
            ts[i].from_, ts[i].to = t["from"], t["to"]
            ts[i].infection = 1 if t.get("infection") else 0
            ts[i].prob_param = i32(t["prob"]) if t.get("prob") is not None else null_i32
        keep.append(ts)
        s.trans = ts
        spl = spec["splines"]
        ss = (Spline * max(1, len(spl)))()
        for i, sp in enumerate(spl):
            ss[i].nknot = len(sp["t"])
            ss[i].t = f64(sp["t"])
            ss[i].param = i32(sp["param"])
            ss[i].multfac = f64(sp.get("multfac", [1.0] * len(sp["t"])))
            ss[i].factor_param = sp.get("factor", -1)
            ss[i].smooth_type = sp.get("smooth_type", 0)
            ss[i].smooth = f64(sp["smooth"]) if sp.get("smooth_type", 0) else null_f64
        keep.append(ss)
        s.nspline, s.spline = len(spl), ss
        s.R_spline = spec["R_spline"]
        s.Rfactor_param = spec.get("Rfactor", -1)
        s.efoi_spline = spec.get("efoi_spline", -1)
        s.efoi_agedist = f64(spec.get("efoi_agedist", [1.0 / G] * G))
        s.geo_spline = spec.get("geo_spline", -1)
        s.N_raw = f64(np.asarray(spec["N"]).reshape(-1)) if spec.get("N") is not None else null_f64
        M = spec.get("M")
        if M is not None:
            s.M_diag, s.M_rowptr, s.M_to, s.M_val = f64(M["diag"]), i32(M["rowptr"]), i32(M["to"]), f64(M["val"])
        else:
            s.M_diag, s.M_rowptr, s.M_to, s.M_val = null_f64, null_i32, null_i32, null_f64
        s.pop_init = i32(np.asarray(spec["pop"]).reshape(-1))
        cov = spec.get("covars", [])
        cv = (Covar * max(1, len(cov)))()
        for i, c in enumerate(cov):
            cv[i].param = c["param"]
            cv[i].value = f64(np.asarray(c["value"]).reshape(-1))
        keep.append(cv)
        s.ncovar, s.covar = len(cov), cv
        s.sus_param = i32(spec["sus"]) if spec.get("sus") is not None else null_i32
        dirs = spec.get("dirichlet", [])
        ds = (Dirichlet * max(1, len(dirs)))()
        for i, d in enumerate(dirs):
            ds[i].n = len(d["param"])
            ds[i].param = i32(d["param"])
            ds[i].frac = f64(d.get("frac", [1.0] * len(d["param"])))
        keep.append(ds)
        s.ndirichlet, s.dirichlet = len(dirs), ds
        obs = spec.get("obs", [])
        if spec.get("obs_table") is not None:
            if obs:
                raise ValueError("a spec holds obs or obs_table, not both")
            os_, nobs = _obs_from_table(spec["obs_table"], keep)
            keep.append(os_)
            s.nobs, s.obs = nobs, os_
            obs = None
        else:
            os_ = (Obs * max(1, len(obs)))()
        for i, o in enumerate(obs or []):
            x = os_[i]
            x.y = o["y"]
            x.sd = o.get("sd", 0.0)
            x.shape = o.get("shape", 1.0)
            x.factor = o.get("factor", 1.0)
            x.thresh = o.get("thresh", 0.0)
            x.percent = o.get("percent", 0.0)
            x.type = specmod.OBS_CODES[o["type"]] if isinstance(o["type"], str) else o["type"]
            x.datatable = o.get("datatable", 0)
            x.graph = o.get("graph", 0)
            x.sett_i, x.sett_f = o["sett_i"], o["sett_f"]
            x.narea, x.area = len(o["area"]), i32(o["area"])
            x.ndp, x.dp = len(o["dp"]), i32(o["dp"])
            tr, co = o.get("tr", []), o.get("co", [])
            x.ntr, x.tr = len(tr), i32(tr) if tr else null_i32
            x.nco, x.co = len(co), i32(co) if co else null_i32
            x.factor_spline = o.get("spline", -1) + 1       # 0 = none
        if obs is not None:
            keep.append(os_)
            s.nobs, s.obs = len(obs), os_
        secs = spec.get("sections", [])
        s.nsection = len(secs)
        s.section_ti = i32([a for a, b in secs]) if secs else null_i32
        s.section_tf = i32([b for a, b in secs]) if secs else null_i32
        am = spec.get("age_mod") or []
        ams = (AgeMod * max(1, len(am)))()
        for i, x in enumerate(am):
            ams[i].spline, ams[i].nage, ams[i].ages = x["spline"], len(x["ages"]), i32(x["ages"])
        keep.append(ams)
        s.nagemod, s.agemod = len(am), ams
        if spec.get("dp_age") is not None:
            s.nage, s.dp_age = spec["nage"], i32(spec["dp_age"])
        else:
            s.nage, s.dp_age = G, null_i32
        sr = spec.get("strain_Rfactor")
        s.nstrain, s.strain_Rfactor = (len(sr), i32(sr)) if sr else (1, null_i32)
        se = spec.get("sus_extra") or []
        s.nsus_extra = len(se)
        s.sus_extra = i32(np.asarray(se, dtype=np.int32).reshape(-1)) if se else null_i32
        ae = spec.get("area_effect")
        s.area_effect_param = i32(ae) if ae is not None else null_i32
        le = spec.get("level_effect")
        if le is not None:
            s.nlevel, s.level_param = len(le["param"]), i32(le["param"])
            s.level_map = i32(np.asarray(le["map"]).reshape(-1))
        else:
            s.nlevel, s.level_param, s.level_map = 0, null_i32, null_i32
        mods = spec.get("modifications") or []
        ma = (Mod * max(1, len(mods)))()
        for i, md in enumerate(mods):
            ma[i].type = MOD_TYPES[md["type"]]
            ma[i].start, ma[i].end, ma[i].factor = md["start"], md["end"], float(md["factor"])
            for n, k in (("ntr", "trans"), ("narea", "areas"), ("ndp", "dps")):
                v = md.get(k) or []
                setattr(ma[i], n, len(v))
                setattr(ma[i], {"ntr": "tr", "narea": "area", "ndp": "dp"}[n], i32(v) if v else null_i32)
            ma[i].spline = md.get("spline", -1)
            ma[i].strain = md.get("strain", 0)
        keep.append(ma)
        s.nmod, s.mod = len(mods), ma
        # R(t) per area group and the external force per (strain, area) group.
        rs = spec.get("R_splines")
        if rs:
            rarea = np.full(spec["A"], -1, dtype=np.int32)
            for k, e in enumerate(rs):
                rarea[np.asarray(e["areas"], dtype=np.int64)] = k
            s.nrgroup, s.rgroup_spline, s.rgroup_area = len(rs), i32([e["spline"] for e in rs]), i32(rarea)
        else:
            s.nrgroup, s.rgroup_spline, s.rgroup_area = 0, null_i32, null_i32
        es = spec.get("efoi_splines")
        if es:
            ns = s.nstrain
            ecell = np.full((ns, spec["A"]), -1, dtype=np.int32)
            for k, e in enumerate(es):
                for st in e.get("strains", range(ns)):
                    ecell[st, np.asarray(e["areas"], dtype=np.int64)] = k
            s.negroup, s.egroup_spline = len(es), i32([e["spline"] for e in es])
            s.egroup_cell = i32(ecell.reshape(-1))
            s.egroup_agedist = f64(np.concatenate([np.asarray(e["agedist"], dtype=np.float64) for e in es]))
        else:
            s.negroup, s.egroup_spline, s.egroup_cell, s.egroup_agedist = 0, null_i32, null_i32, null_f64
        self.spec = s


def _obs_from_table(tab, keep):
    """The ctypes observation array of a compact obs_table, filled column by
    column through numpy views of the struct array (no per-observation Python
    work): one area each, shared dp and transition lists."""
    y = np.asarray(tab["y"], dtype=np.float64)
    n = len(y)
    arr = (Obs * max(1, n))()
    size = C.sizeof(Obs)

    def col(name, dt):
        return np.ndarray((n,), dtype=dt, buffer=arr, offset=getattr(Obs, name).offset, strides=(size,))

    area = _i32(tab["area"])
    dp, tr = _i32(tab["dp"]), _i32(tab["tr"])
    areas = np.arange(int(area.max()) + 1 if n else 1, dtype=np.int32)
    keep.extend([area, dp, tr, areas])
    col("y", np.float64)[:] = y
    col("sd", np.float64)[:] = np.asarray(tab["sd"], dtype=np.float64)
    col("shape", np.float64)[:] = 1.0
    col("factor", np.float64)[:] = 1.0
    t = tab["type"]
    col("type", np.int32)[:] = specmod.OBS_CODES[t] if isinstance(t, str) else t
    col("datatable", np.int32)[:] = int(tab.get("datatable", 0))
    col("graph", np.int32)[:] = _i32(tab["graph"])
    col("sett_i", np.int32)[:] = _i32(tab["sett_i"])
    col("sett_f", np.int32)[:] = _i32(tab["sett_f"])
    col("narea", np.int32)[:] = 1
    col("area", np.uint64)[:] = np.uint64(areas.ctypes.data) + area.astype(np.uint64) * np.uint64(4)
    col("ndp", np.int32)[:] = len(dp)
    col("dp", np.uint64)[:] = dp.ctypes.data
    col("ntr", np.int32)[:] = len(tr)
    col("tr", np.uint64)[:] = tr.ctypes.data
    return arr, n


def spec_tables(spec, paramval):
    """Host-side FP64 tables of State::set_param (no device needed)."""
    L = lib()
    buf = _SpecBuffers(spec)
    A, G = spec["A"], spec["G"]
    C_, T = len(spec["comps"]), len(spec["trans"])
    S = spec["period"] * spec["spd"]
    pv = _f64(paramval)
    Nshape = (S, G, G) if spec.get("age_mod") else (G, G)      # a time-varying age matrix: one per step
    out = dict(rate=np.zeros((T, G)), sus=np.zeros(G), inf=np.zeros(C_), N=np.zeros(Nshape),
               beta_af=np.zeros((S, A)), efoi=np.zeros((S, A, G)), geo=np.zeros(S))
    Pr = np.zeros(1)
    inb = np.zeros(1, dtype=np.int32)
    err = C.create_string_buffer(256)
    rc = L.beep_spec_tables(C.byref(buf.spec), _p(pv, C.c_double),
                            *[_p(out[k], C.c_double) for k in ("rate", "sus", "inf", "N", "beta_af", "efoi", "geo")],
                            _p(Pr, C.c_double), _p(inb, C.c_int32), err, 256)
    if rc != BEEP_OK:
        raise BeepError(rc, err.value.decode())
    out["Pr"] = float(Pr[0])
    out["inbounds"] = bool(inb[0])
    return out


class Work(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("kind", C.c_int32), ("inst", C.c_int32), ("init", C.c_int32),
                ("slot", C.c_int32), ("key", C.c_uint32), ("t0", C.c_int32), ("t1", C.c_int32)]


WORK_SIMULATE, WORK_MBP = 0, 1
BATCH_AUTOTUNE = 1


# beep_work records as a numpy structured type (same offsets and size), for
# callers that keep their work lists as column arrays.
WORK_DTYPE = np.dtype({"names": [f[0] for f in Work._fields_],
                       "formats": [np.uint64, np.int32, np.int32, np.int32, np.int32, np.uint32, np.int32, np.int32],
                       "offsets": [getattr(Work, f[0]).offset for f in Work._fields_],
                       "itemsize": C.sizeof(Work)})


def work_records(ctx, kind, inst, init, slot, key, t0, t1):
    """A beep_work array from columns (arrays or scalars, broadcast to one
    length); ctx holds context handles (handles()).  Slices keep the record
    layout; np.concatenate does not (it packs the records to 36 bytes), so
    a batch made of parts is assembled by assigning into one array."""
    cols = np.broadcast_arrays(*(np.asarray(v) for v in (ctx, kind, inst, init, slot, key, t0, t1)))
    rec = np.zeros(cols[0].shape[0] if cols[0].ndim else 1, dtype=WORK_DTYPE)
    for name, v in zip(WORK_DTYPE.names, cols):
        rec[name] = v
    return rec


def _work_ptr(items):
    """(n, pointer to beep_work records, first context handle, the object
    holding the memory) for a list of items or a work_records array."""
    if isinstance(items, np.ndarray):
        assert items.dtype == WORK_DTYPE and items.flags.c_contiguous
        n = len(items)
        return n, items.ctypes.data, int(items["ctx"][0]) if n else 0, items
    arr = _work_array(items)
    return len(items), arr, (items[0][0]._h.value if items else 0), arr


def _work_array(items):
    """The beep_work array of (ctx, kind, inst, init, slot, key, t0, t1)
    items, filled column by column."""
    n = len(items)
    arr = (Work * max(1, n))()
    size = C.sizeof(Work)

    def col(name, dt):
        return np.ndarray((n,), dtype=dt, buffer=arr, offset=getattr(Work, name).offset, strides=(size,))

    cols = list(zip(*items)) if n else [()] * 8
    col("ctx", np.uint64)[:] = np.array([c if isinstance(c, (int, np.integer)) else c._h.value for c in cols[0]],
                                        dtype=np.uint64)
    for k, (name, dt) in enumerate((("kind", np.int32), ("inst", np.int32), ("init", np.int32), ("slot", np.int32),
                                    ("key", np.uint32), ("t0", np.int32), ("t1", np.int32)), start=1):
        col(name, dt)[:] = np.asarray(cols[k], dtype=np.int64).astype(dt)
    return arr


BATCH_LOGLIK = 2


def run_batch(items, seed=1, stream=0, autotune=False, loglik=False):
    """Runs (ctx, kind, inst, init, slot, key, t0, t1) work items across
    contexts in one batch; returns (wall ms, fused launches, step-wise
    contexts, cell-steps), and with loglik=True also each item's observation
    log-likelihood.  items may also be a work_records array."""
    n, arr, h0, _keep = _work_ptr(items)
    stats = np.zeros(4)
    L = np.zeros(max(1, n))
    flags = (BATCH_AUTOTUNE if autotune else 0) | (BATCH_LOGLIK if loglik else 0)
    rc = lib().beep_run_batch(n, arr, seed, stream, flags, _p(L, C.c_double) if loglik else None,
                              _p(stats, C.c_double))
    if rc != BEEP_OK:
        raise BeepError(rc, lib().beep_last_error(C.c_void_p(h0)).decode())
    return (stats, L[:n]) if loglik else stats


def handles(ctxs):
    """The contexts' handles as a uint64 array (reusable across batch calls)."""
    return np.array([c._h.value for c in ctxs], dtype=np.uint64)


def batch_set_param(ctxs, slots, paramvals):
    """State::set_param for slots of many contexts in one call: row i of the
    2-D array paramvals holds the parameter vector of slot slots[i] of context
    ctxs[i] (its first nparam entries).  ctxs: Context objects or handles().
    Returns (Pr, inbounds)."""
    h = ctxs if isinstance(ctxs, np.ndarray) else handles(ctxs)
    n = len(h)
    pv = _f64(paramvals)
    assert pv.ndim == 2 and pv.shape[0] == n
    ptrs = np.uint64(pv.ctypes.data) + np.arange(n, dtype=np.uint64) * np.uint64(pv.strides[0])
    slots = _i32(slots)
    Pr = np.zeros(n)
    inb = np.zeros(n, dtype=np.int32)
    rc = lib().beep_batch_set_param(n, h.ctypes.data, _p(slots, C.c_int32), ptrs.ctypes.data, _p(Pr, C.c_double),
                                    _p(inb, C.c_int32))
    if rc != BEEP_OK:
        raise BeepError(rc, lib().beep_last_error(C.c_void_p(int(h[0]))).decode())
    return Pr, inb.astype(bool)


def batch_accept(items, invT, dlogq, mode=1, EFcut=0.0, seed=1, stream=0):
    """Metropolis-Hastings acceptance of MBP items across contexts on the
    device: (al, accepted).  items may also be a work_records array."""
    n, arr, h0, _keep = _work_ptr(items)
    invT, dlogq = _f64(np.broadcast_to(invT, (n,))), _f64(np.broadcast_to(dlogq, (n,)))
    al = np.zeros(n)
    acc = np.zeros(n, dtype=np.uint8)
    rc = lib().beep_batch_accept(n, arr, mode, _p(invT, C.c_double), float(EFcut), _p(dlogq, C.c_double), seed, stream,
                                 _p(al, C.c_double), _p(acc, C.c_uint8))
    if rc != BEEP_OK:
        raise BeepError(rc, lib().beep_last_error(C.c_void_p(h0)).decode())
    return al, acc.astype(bool)


def batch_mc3_swap(ladders, seed=1, stream=0):
    """MC3 swaps over many ladders at once: ladders is a list of lists of
    (ctx, inst, run, invT) chains in temperature order.  Returns per ladder
    (order, ntr, nac)."""
    ptr = np.cumsum([0] + [len(l) for l in ladders]).astype(np.int32)
    chains = [ch for l in ladders for ch in l]
    order, ntr, nac = batch_mc3_swap_arrays(ptr, handles([ch[0] for ch in chains]), [ch[1] for ch in chains],
                                            [ch[2] for ch in chains], [ch[3] for ch in chains], seed, stream)
    return [(order[a:b], ntr[a:b], nac[a:b]) for a, b in zip(ptr[:-1], ptr[1:])]


def batch_mc3_swap_arrays(ladder_ptr, h, inst, run, invT, seed=1, stream=0):
    """batch_mc3_swap on column arrays: ladder l holds chains ladder_ptr[l] ..
    ladder_ptr[l+1]-1; chain i is instance inst[i] of the context with handle
    h[i] (handles()).  Returns (order, ntr, nac) over all chains."""
    ptr = _i32(ladder_ptr)
    h = np.ascontiguousarray(h, dtype=np.uint64)
    n = len(h)
    inst, run, invT = _i32(inst), _i32(np.broadcast_to(run, (n,))), _f64(np.broadcast_to(invT, (n,)))
    order, ntr, nac = (np.zeros(n, dtype=np.int32) for _ in range(3))
    rc = lib().beep_batch_mc3_swap(len(ptr) - 1, _p(ptr, C.c_int32), h.ctypes.data,
                                   _p(inst, C.c_int32), _p(run, C.c_int32), _p(invT, C.c_double), seed, stream,
                                   _p(order, C.c_int32), _p(ntr, C.c_int32), _p(nac, C.c_int32))
    if rc != BEEP_OK:
        raise BeepError(rc, lib().beep_last_error(C.c_void_p(int(h[0]) if n else None)).decode())
    return order, ntr, nac


class Context:
    """One device context: a model, `capacity` instances and `nslot` parameter
    slots (the state of many State objects at once)."""

    def __init__(self, spec, capacity, nslot=None, keep=KEEP_ALL, device=0, hbm_budget=0):
        L = lib()
        self.spec = spec
        self.capacity = capacity
        self.nslot = nslot if nslot is not None else capacity
        buf = _SpecBuffers(spec)
        h = C.c_void_p()
        rc = L.beep_create(C.byref(buf.spec), device, hbm_budget, capacity, self.nslot, keep, C.byref(h))
        self._h = h
        if rc != BEEP_OK:
            msg = L.beep_last_error(h).decode() if h.value else "beep_create failed"
            if h.value:
                L.beep_destroy(h)
            self._h = C.c_void_p()
            raise BeepError(rc, msg)
        d = np.zeros(10, dtype=np.int32)
        L.beep_get_dims(h, _p(d, C.c_int32))
        self.A, self.G, self.C, self.T, self.S, self.nparam, self.nobs, self.ndt = [int(v) for v in d[:8]]

    def close(self):
        if self._h and self._h.value:
            lib().beep_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != BEEP_OK:
            raise BeepError(rc, lib().beep_last_error(self._h).decode())

    @property
    def device_bytes(self):
        return lib().beep_device_bytes(self._h)

    def set_param(self, slots, paramvals):
        slots = _i32(np.atleast_1d(slots))
        pv = _f64(np.atleast_2d(paramvals))
        n = len(slots)
        assert pv.shape == (n, self.nparam)
        Pr = np.zeros(n)
        inb = np.zeros(n, dtype=np.int32)
        self._chk(lib().beep_set_param(self._h, n, _p(slots, C.c_int32), _p(pv, C.c_double), _p(Pr, C.c_double),
                                       _p(inb, C.c_int32)))
        return Pr, inb.astype(bool)

    def get_tables(self, slot):
        A, G, C_, T, S = self.A, self.G, self.C, self.T, self.S
        Nshape = (S, G, G) if self.spec.get("age_mod") else (G, G)
        out = dict(rate=np.zeros((T, G)), sus=np.zeros(G), inf=np.zeros(C_), N=np.zeros(Nshape),
                   beta_af=np.zeros((S, A)), efoi=np.zeros((S, A, G)), geo=np.zeros(S))
        self._chk(lib().beep_get_tables(self._h, slot, *[_p(out[k], C.c_double) for k in
                                                         ("rate", "sus", "inf", "N", "beta_af", "efoi", "geo")]))
        return out

    def simulate(self, inst, slot, key, t0=0, t1=None, seed=1, stream=0, plain=False):
        inst, slot, key = _i32(inst), _i32(slot), np.ascontiguousarray(np.asarray(key, dtype=np.uint32))
        t1 = self.S if t1 is None else t1
        self._chk(lib().beep_simulate(self._h, len(inst), _p(inst, C.c_int32), _p(slot, C.c_int32),
                                      _p(key, C.c_uint32), t0, t1, seed, stream, SIM_PLAIN if plain else 0))

    def simulate_noalloc(self, inst, slot, key, t0=0, t1=None, seed=1, stream=0):
        """simulate() through beep_simulate_noalloc: no device memory is
        allocated (job descriptors in the library's fixed per-device array)."""
        inst, slot, key = _i32(inst), _i32(slot), np.ascontiguousarray(np.asarray(key, dtype=np.uint32))
        t1 = self.S if t1 is None else t1
        self._chk(lib().beep_simulate_noalloc(self._h, len(inst), _p(inst, C.c_int32), _p(slot, C.c_int32),
                                              _p(key, C.c_uint32), t0, t1, seed, stream))

    def obs_loglik(self, inst, section=-1, with_dt=False):
        inst = _i32(inst)
        L = np.zeros(len(inst))
        EFdt = np.zeros((len(inst), self.ndt))
        self._chk(lib().beep_obs_loglik(self._h, len(inst), _p(inst, C.c_int32), section, _p(L, C.c_double),
                                        _p(EFdt, C.c_double) if with_dt else None))
        return (L, EFdt) if with_dt else L

    def obs_values(self, inst):
        out = np.zeros(self.nobs)
        self._chk(lib().beep_obs_values(self._h, inst, _p(out, C.c_double)))
        return out

    def pois_loglik(self, inst):
        inst = _i32(inst)
        L = np.zeros(len(inst))
        self._chk(lib().beep_pois_loglik(self._h, len(inst), _p(inst, C.c_int32), _p(L, C.c_double)))
        return L

    def mbp(self, init, prop, slot, key, simu_mask=None, seed=1, stream=0, inf_update=1):
        init, prop, slot = _i32(init), _i32(prop), _i32(slot)
        key = np.ascontiguousarray(np.asarray(key, dtype=np.uint32))
        mp = C.POINTER(C.c_uint8)()
        if simu_mask is not None:
            simu_mask = np.ascontiguousarray(np.asarray(simu_mask, dtype=np.uint8))
            assert simu_mask.shape == (len(init), self.S, self.A)
            mp = _p(simu_mask, C.c_uint8)
        self._chk(lib().beep_mbp(self._h, len(init), _p(init, C.c_int32), _p(prop, C.c_int32), _p(slot, C.c_int32),
                                 _p(key, C.c_uint32), mp, seed, stream, inf_update))

    def accept(self, init, prop, mode, dlogq, key, invT=1.0, EFcut=0.0, seed=1, stream=0):
        init, prop = _i32(init), _i32(prop)
        dlogq = _f64(dlogq)
        key = np.ascontiguousarray(np.asarray(key, dtype=np.uint32))
        al = np.zeros(len(init))
        acc = np.zeros(len(init), dtype=np.uint8)
        self._chk(lib().beep_accept(self._h, len(init), _p(init, C.c_int32), _p(prop, C.c_int32), mode, invT, EFcut,
                                    _p(dlogq, C.c_double), _p(key, C.c_uint32), seed, stream, _p(al, C.c_double),
                                    _p(acc, C.c_uint8)))
        return al, acc.astype(bool)

    def get_transnum(self, inst):
        out = np.zeros((self.S, self.A, self.T, self.G), dtype=np.int32)
        self._chk(lib().beep_get_transnum(self._h, inst, _p(out, C.c_int32)))
        return out

    def get_transmean(self, inst):
        out = np.zeros((self.S, self.A, self.T, self.G), dtype=np.float32)
        self._chk(lib().beep_get_transmean(self._h, inst, _p(out, C.c_float)))
        return out

    def get_pop(self, inst):
        out = np.zeros((self.S, self.A, self.C, self.G), dtype=np.int32)
        self._chk(lib().beep_get_pop(self._h, inst, _p(out, C.c_int32)))
        return out

    def get_state(self, inst):
        out = np.zeros((self.A, self.C, self.G), dtype=np.int32)
        self._chk(lib().beep_get_state(self._h, inst, _p(out, C.c_int32)))
        return out

    def put_state(self, inst, pop):
        pop = _i32(pop)
        assert pop.size == self.A * self.C * self.G
        self._chk(lib().beep_put_state(self._h, inst, _p(pop, C.c_int32)))

    def put_transnum(self, inst, slot, transnum):
        tn = _i32(transnum)
        assert tn.size == self.S * self.A * self.T * self.G
        self._chk(lib().beep_put_transnum(self._h, inst, slot, _p(tn, C.c_int32)))

    def copy_instances(self, dst, src):
        dst, src = _i32(dst), _i32(src)
        self._chk(lib().beep_copy_instances(self._h, len(dst), _p(dst, C.c_int32), _p(src, C.c_int32)))

    def synchronize(self):
        self._chk(lib().beep_synchronize(self._h))

    def copy_states(self, dst, src):
        dst, src = _i32(dst), _i32(src)
        self._chk(lib().beep_copy_states(self._h, len(dst), _p(dst, C.c_int32), _p(src, C.c_int32)))

    def derive_slots(self, dst, src, factor, from_step):
        dst, src, fs = _i32(dst), _i32(src), _i32(from_step)
        f = _f64(factor)
        self._chk(lib().beep_derive_slots(self._h, len(dst), _p(dst, C.c_int32), _p(src, C.c_int32),
                                          _p(f, C.c_double), _p(fs, C.c_int32)))

    def sum_transnum(self, inst, tr, t0, t1):
        inst = _i32(inst)
        out = np.zeros((len(inst), t1 - t0))
        self._chk(lib().beep_sum_transnum(self._h, len(inst), _p(inst, C.c_int32), tr, t0, t1,
                                          _p(out, C.c_double)))
        return out

    def pmcmc_obs_prob(self, inst, slot, keys, invT=1.0, seed=1, stream=0):
        """PMCMC::obs_prob over the particles `inst` (one parameter slot):
        returns (log-likelihood estimate, ancestry [(nsection+1), n])."""
        inst, keys = _i32(inst), _u32(keys)
        nsec = len(self.spec.get("sections", []))
        ob = np.zeros(1)
        bp = np.zeros((nsec + 1, len(inst)), dtype=np.int32)
        self._chk(lib().beep_pmcmc_obs_prob(self._h, len(inst), _p(inst, C.c_int32), int(slot), _p(keys, C.c_uint32),
                                            float(invT), int(seed), int(stream), _p(ob, C.c_double),
                                            _p(bp, C.c_int32)))
        return float(ob[0]), bp

    def pmcmc_particle_sample(self, dst):
        """Mpi::particle_sample of the last particle-filter run into instance dst."""
        self._chk(lib().beep_pmcmc_particle_sample(self._h, int(dst)))

    def mc3_swap(self, inst, run, invT, seed=1, stream=0):
        """MC3::swap_states over chains in instances `inst`: returns (order,
        ntr, nac); chain i takes the state of chain order[i]."""
        inst, run, invT = _i32(inst), _i32(run), _f64(invT)
        n = len(inst)
        order, ntr, nac = (np.zeros(n, dtype=np.int32) for _ in range(3))
        self._chk(lib().beep_mc3_swap(self._h, n, _p(inst, C.c_int32), _p(run, C.c_int32), _p(invT, C.c_double),
                                      int(seed), int(stream), _p(order, C.c_int32), _p(ntr, C.c_int32),
                                      _p(nac, C.c_int32)))
        return order, ntr, nac

    @property
    def path(self):
        """PATH_STEPWISE, PATH_FUSED or PATH_FUSED_SPEC."""
        return lib().beep_path(self._h)

    def timing(self):
        """(simulate/MBP kernel ms, observation kernel ms, Poisson log-lik kernel ms, launches)."""
        out = np.zeros(4)
        self._chk(lib().beep_timing(self._h, _p(out, C.c_double)))
        return out


def abcsmc_weights(prev, w_prev, theta, inv_M, size, logprior, device=0):
    """ABCSMC::calculate_particle_weight on the device for every row of theta."""
    prev, w_prev, theta, inv_M, logprior = _f64(prev), _f64(w_prev), _f64(theta), _f64(inv_M), _f64(logprior)
    n_prev, nvar = prev.shape
    out = np.zeros(len(theta))
    rc = lib().beep_abcsmc_weights(device, nvar, n_prev, _p(prev, C.c_double), _p(w_prev, C.c_double), len(theta),
                                   _p(theta, C.c_double), _p(inv_M, C.c_double), float(size),
                                   _p(logprior, C.c_double), _p(out, C.c_double))
    if rc:
        raise BeepError(rc, "beep_abcsmc_weights")
    return out


def bootstrap(L, invT=1.0, seed=1, stream=0, sec=0, device=0):
    """PMCMC::bootstrap on the device: (ancestors, the section's term)."""
    L = np.ascontiguousarray(np.asarray(L, dtype=np.float64))
    bp = np.zeros(len(L), dtype=np.int32)
    ob = np.zeros(1)
    rc = lib().beep_bootstrap(device, len(L), _p(L, C.c_double), float(invT), int(seed), int(stream), int(sec),
                              _p(bp, C.c_int32), _p(ob, C.c_double))
    if rc:
        raise BeepError(rc, "beep_bootstrap")
    return bp, float(ob[0])


CHECK_FUSED, CHECK_STEPWISE = 0, 1


def check_launch(kind, args):
    """The host-side launch checks (no device needed): the grid a fused launch
    (args njobs, block, lds, lds limit) or a step-wise launch (args A, G, Apad,
    n, capacity) would use; raises BeepError(EINVAL) with the reason."""
    a = np.ascontiguousarray(np.asarray(args, dtype=np.int64))
    g = np.zeros(6, dtype=np.int64)
    err = C.create_string_buffer(256)
    rc = lib().beep_debug_check_launch(kind, _p(a, C.c_int64), len(a), _p(g, C.c_int64), err, 256)
    if rc != BEEP_OK:
        raise BeepError(rc, err.value.decode())
    return g[:1] if kind == CHECK_FUSED else g


def debug_philox(ctr, k0, k1, device=0):
    ctr = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32).reshape(-1, 4))
    out = np.zeros_like(ctr)
    rc = lib().beep_debug_philox(device, len(ctr), _p(ctr, C.c_uint32), k0, k1, _p(out, C.c_uint32))
    if rc:
        raise BeepError(rc, "beep_debug_philox")
    return out


def debug_poisson(lam, seed=1, key=0, device=0):
    lam = np.ascontiguousarray(np.asarray(lam, dtype=np.float32))
    out = np.zeros(len(lam), dtype=np.int32)
    rc = lib().beep_debug_poisson(device, len(lam), _p(lam, C.c_float), seed, key, _p(out, C.c_int32))
    if rc:
        raise BeepError(rc, "beep_debug_poisson")
    return out


def debug_gemm(A, B, device=0):
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    rows, K = A.shape
    out = np.zeros((rows, K), dtype=np.float32)
    rc = lib().beep_debug_gemm(device, rows, K, _p(A, C.c_float), _p(B, C.c_float), _p(out, C.c_float))
    if rc:
        raise BeepError(rc, "beep_debug_gemm")
    return out


def debug_binomial(nn, p, seed=1, key=0, device=0):
    nn = _i32(nn)
    p = _f64(p)
    out = np.zeros(len(nn), dtype=np.int32)
    rc = lib().beep_debug_binomial(device, len(nn), _p(nn, C.c_int32), _p(p, C.c_double), seed, key,
                                   _p(out, C.c_int32))
    if rc:
        raise BeepError(rc, "beep_debug_binomial")
    return out
