"""Throughput of the batched simulation + likelihood + MBP path on one GPU.

Headline workload (one "step" = one sweep): the uneven batch of 16384
instances of the SEIR-in-every-area model -- areas uniform in 10..400,
horizons uniform in 50..400 days, age groups in {1, 4, 9}, 10-nearest-
neighbour mixing, daily I->R observations per area -- half ABC-SMC particles
and half MC3 chains in ladders of 8 temperatures.  A sweep: fresh parameters
for every particle and chain proposal (derived and uploaded), the particles'
simulations and the chains' MBP proposals in one mixed batch, every
instance's observation log-likelihood, the chains' Metropolis-Hastings
acceptance and the ladder swaps.  Inputs are resident in HBM; the parameter
derivation and upload are inside the timed step.

Metric: particle-timesteps/s = instances x daily steps per sweep / sweep wall
time.  Also reported: the plain path (one instance per launch per step) on a
sample, the fused phase against the HBM roofline with the algorithmic bytes
per particle-step of the transition path, the FP64 CPU oracle on all host
cores over a sample of the same work, the 100-area spatial ABC-SMC
workload, and serving p50/p99 latency.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --workload spatial|dense|serving|copy   (single workloads)
"""
import argparse
import gc
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-timesteps/sec batched vs plain GPU path; p50/p99 request latency"


def bytes_per_particle_step(A, G, C, T, mbp=False):
    """Algorithmic HBM bytes of the transition path per instance per step:
    int32 populations read and written, int32 transnum written, the FP32
    infection driver per cell, the area factor per area and the rates per
    (transition, group); an MBP proposal also reads the initial instance's
    transnum (int32) and transmean (FP32)."""
    AG = A * G
    b = 4 * AG * (2 * C + T + 1) + 4 * A + 4 * T * G
    return b + (8 * A * T * G if mbp else 0)


def cpu_info():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if model is None:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return {"nproc": os.cpu_count(), "available": avail, "model": model}


def _cpu_uneven_worker(args):
    """One host core: the oracle's simulate + observation log-likelihood
    (ABC-SMC items) or MBP proposal + log-likelihood (MC3 items), each with
    its parameter derivation, over the sample items k, k + stride, ...;
    models are set up before the timer starts."""
    sample, k, stride, seconds, seed = args
    from beep_amd import configs, spec as specmod
    from oracle import oracle as O
    rng = np.random.default_rng(1000 + k)
    mine = sample[k::stride]
    models = []
    for kind, A, G, S, h, gseed in mine:
        spec, _ = configs.seir_areas(A, G, S, seed=gseed, daily_obs=True)
        orc = O.Oracle(specmod.expand_obs(spec))
        init = None
        if kind == "mc3":
            init = orc.simulate(orc.tables(specmod.sample_prior(spec, rng, 1)[0]), seed=7 + k)
        models.append((kind, orc, spec, h, init))
    n = pts = 0
    t0 = time.perf_counter()
    i = 0
    while models and time.perf_counter() - t0 < seconds:
        kind, orc, spec, h, init = models[i % len(models)]
        pv = specmod.sample_prior(spec, rng, 1)[0]
        tb = orc.tables(pv)
        if kind == "abc":
            st = orc.simulate(tb, seed=100 + i, tf=h)
            pts += h
        else:
            st = orc.mbp(tb, init, seed=100 + i)
            pts += orc.S
        orc.obs_loglik(st)
        n += 1
        i += 1
    return n, pts, time.perf_counter() - t0


def cpu_baseline_uneven(workload, seconds=12.0, cores=None, nsample=None, seed=4):
    """The FP64 CPU oracle (the reference's algorithm and libstdc++
    samplers) on every available host core, one process each, over a bounded
    random sample of the uneven workload's items (ABC-SMC and MC3 in their
    proportions), timing the same work as a GPU sweep: parameter derivation,
    simulation or MBP, observation log-likelihood."""
    import multiprocessing as mp
    info = cpu_info()
    cores = cores or info["available"]
    rng = np.random.default_rng(seed)
    items = [("abc", g["A"], g["G"], g["S"], h, seed * 100003 + gi) for gi, g in enumerate(workload["abc"])
             for h in g["horizons"]]
    items += [("mc3", l["A"], l["G"], l["S"], l["S"], seed * 100003 + 50000 + li) for li, l in enumerate(workload["mc3"])
              for _ in range(workload["ladder"])]
    nsample = nsample or 3 * cores
    sample = [items[i] for i in rng.choice(len(items), size=min(nsample, len(items)), replace=False)]
    jobs = [(sample, k, cores, seconds, seed) for k in range(cores)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_uneven_worker, jobs)
    wall = time.perf_counter() - t0
    per_core = [r[1] / r[2] for r in res if r[2] > 0]
    n = sum(r[0] for r in res)
    return {"value": float(sum(per_core)), "unit": "particle-timesteps/s", "cores": cores, "kind": "port",
            "nproc": info["nproc"], "cpu_model": info["model"],
            "single_core": float(np.mean(per_core)) if per_core else None,
            "sample": f"{len(sample)} items drawn from the uneven workload (ABC-SMC simulations and MC3 MBP proposals "
                      f"in their proportions, each with parameter derivation and observation log-likelihood), "
                      f"{n} evaluations, {cores} processes x {seconds:.0f} s (wall {wall:.1f} s)"}


def _cpu_spatial_worker(args):
    spec, params, idx, stride, seconds = args
    from oracle import oracle as O
    orc = O.Oracle(spec)
    n, t0 = 0, time.perf_counter()
    i = idx
    while time.perf_counter() - t0 < seconds and i < len(params):
        tb = orc.tables(params[i])
        st = orc.simulate(tb, seed=1000 + i)
        orc.obs_loglik(st)
        n += 1
        i += stride
    return n, n * orc.S, time.perf_counter() - t0


def cpu_baseline_spatial(spec, params, seconds=8.0, cores=None):
    import multiprocessing as mp
    cores = cores or cpu_info()["available"]
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_spatial_worker, [(spec, params, k, cores, seconds) for k in range(cores)])
    return {"value": float(sum(r[1] / r[2] for r in res)), "cores": cores,
            "sample": f"{sum(r[0] for r in res)} particles, {cores} processes x {seconds:.0f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="uneven", choices=["uneven", "spatial", "dense", "serving", "copy"])
    ap.add_argument("--instances", type=int, default=16384)
    ap.add_argument("--plain-sample", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="headline only (no spatial / serving keys)")
    ap.add_argument("--generic", action="store_true", help="use the generic fused kernel, not the specialised one")
    ap.add_argument("--rates", default="5,20,80")
    ap.add_argument("--requests", type=int, default=120)
    ap.add_argument("--particles", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path shards by instance: "
              f"one GPU's share is what is measured here)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "spatial":
        print(json.dumps(bench_spatial(args, cpu=not args.no_cpu)), flush=True)
    elif args.workload == "dense":
        bench_dense(args)
    elif args.workload == "serving":
        bench_serving(args)
    elif args.workload == "copy":
        bench_copy(args)
    else:
        bench_uneven(args)


def bench_uneven(args):
    from beep_amd import _lib, configs
    from beep_amd.batch import UnevenBatch
    workload = configs.uneven_batch(args.instances, seed=4)
    # The CPU baseline runs first, in forked processes, before this process
    # initialises the GPU.
    cpu = None if args.no_cpu else cpu_baseline_uneven(workload)
    if cpu is not None:
        print(f"uneven: cpu baseline {cpu['value']:.0f} particle-timesteps/s on {cpu['cores']} cores",
              file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ub = UnevenBatch(workload, seed=4, keep_extra=_lib.GENERIC_KERNEL if args.generic else 0)
    setup = time.perf_counter() - t0
    print(f"uneven: {len(ub.ctx)} contexts set up in {setup:.1f} s", file=sys.stderr, flush=True)
    # The set-up's Python objects (specs, contexts) live for the whole run:
    # moved out of the collector's generations, so that a full collection over
    # them does not land inside a timed sweep.
    gc.collect()
    gc.freeze()
    pts = ub.particle_steps()
    abc, mbp = ub.sweep_items()
    algo_bytes = sum((it[7] - it[6]) * bytes_per_particle_step(it[0].A, it[0].G, it[0].C, it[0].T) for it in abc)
    algo_bytes += sum(it[0].S * bytes_per_particle_step(it[0].A, it[0].G, it[0].C, it[0].T, mbp=True) for it in mbp)
    c0 = abc[0][0]
    # Warm-up sweeps: the first also tunes the workgroup size of each size class.
    for w in range(max(1, args.warmup)):
        ub.sweep(1000 + w, autotune=(w == 0))
    sweeps, fused_ms, ll_ms = [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = ub.sweep(k)
        tm = c0.timing()
        sweeps.append(r)
        fused_ms.append(tm[0])
        ll_ms.append(tm[1])
        print(f"uneven: sweep {k}: {r['total_ms']:.1f} ms (set_param {r['set_param_ms']:.1f}, batch {r['batch_ms']:.1f}, "
              f"accept {r['accept_ms']:.1f}, swap {r['swap_ms']:.1f}); fused phase {tm[0]:.1f} ms, "
              f"log-lik {tm[1]:.1f} ms", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    assert all(r["loglik_finite"] for r in sweeps)
    value = pts * args.steps / wall
    # Plain path: one instance per launch per step on a sample of particles.
    ns = min(args.plain_sample, len(abc))
    sample = [abc[i] for i in np.random.default_rng(5).choice(len(abc), ns, replace=False)]
    t1 = time.perf_counter()
    plain_pts = 0
    for (ctx, _kind, inst, _init, slot, key, a, b) in sample:
        ctx.simulate([inst], [slot], [key], a, b, seed=4, stream=777, plain=True)
        ctx.obs_loglik([inst])
        plain_pts += b - a
    plain_wall = time.perf_counter() - t1
    plain_value = plain_pts / plain_wall
    fused = float(np.mean(fused_ms))
    achieved = algo_bytes / (fused * 1e-3) / 1e9
    kernel = {_lib.PATH_FUSED: "k_fused<MODE_SIM> + k_fused<MODE_MBP> (generic)",
              _lib.PATH_FUSED_SPEC: "beep_fused_spec_sim + beep_fused_spec_mbp"}[c0.path]
    mean = {k: float(np.mean([r[k] for r in sweeps])) for k in ("set_param_ms", "batch_ms", "accept_ms", "swap_ms")}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(1, args.warmup),
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 counts, fp32 rates, fp64 log-likelihood and acceptance",
        "data": "synthetic (prior and random-walk parameters, synthetic daily observations)",
        "config": {"workload": "uneven batch: 16384 SEIR-in-every-area instances, areas 10-400, horizons 50-400 days, "
                               "age groups {1,4,9}, half ABC-SMC particles and half MC3 chains (ladders of 8); one "
                               "sweep = fresh parameters + simulations / MBP proposals + observation log-likelihoods "
                               "+ MH acceptance + ladder swaps",
                   "instances": args.instances, "abc_particles": len(abc), "mc3_chains": len(mbp),
                   "ladders": len(ub.mc3_ctx), "contexts": len(ub.ctx), "particle_steps_per_sweep": pts,
                   "cell_steps_per_sweep": sweeps[-1]["cell_steps"], "device_gb": ub.device_bytes / 1e9,
                   "setup_s": setup},
        "sweep_ms": mean,
        "cell_steps_per_s": sweeps[-1]["cell_steps"] * args.steps / wall,
        "plain": {"value": plain_value, "sample_instances": ns, "batched_over_plain": value / plain_value},
        "kernels_ms": {"fused_phase": fused, "loglik_phase": float(np.mean(ll_ms))},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_sweep": algo_bytes,
                     "note": "fused phase of the batch (size-class launches over 4 streams) timed with HIP events on "
                             "the context stream from the streams' start to their join; traffic counters in profiles/"},
        "mc3": {"accepted_per_sweep": float(np.mean([r["accepted"] for r in sweeps])),
                "swaps_per_sweep": float(np.mean([r["swaps"] for r in sweeps]))},
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    ub.close()
    if not args.no_extras:
        try:
            out["spatial_abcsmc"] = bench_spatial(args, cpu=False)
        except Exception as e:          # an extra key must not hide the headline
            out["spatial_abcsmc"] = {"error": str(e)}
        try:
            out["serving"] = bench_serving(args, rates=[20.0], nreq=60, emit=False)
        except Exception as e:
            out["serving"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


def bench_spatial(args, cpu=True):
    """4096 ABC-SMC particles of the 100-area x 4-age-group spatial model
    (8 compartments, 10-NN mixing, 10-knot R(t), 200 days): simulate +
    observation log-likelihood per step; also one generation with fresh
    parameters including set_param, and the plain path on a sample."""
    from beep_amd import _lib, configs, spec as specmod
    spec, _truth = configs.spatial_abcsmc()
    P = args.particles
    rng = np.random.default_rng(20250102)
    params = specmod.sample_prior(spec, rng, P)
    cpub = cpu_baseline_spatial(spec, params) if cpu else None
    ctx = _lib.Context(spec, capacity=P, nslot=P, keep=_lib.KEEP_ALL | (_lib.GENERIC_KERNEL if args.generic else 0))
    slots = np.arange(P, dtype=np.int32)
    ctx.set_param(slots, params)
    inst, keys = np.arange(P, dtype=np.int32), np.arange(P, dtype=np.uint32)
    S, A, G, C_, T = ctx.S, ctx.A, ctx.G, ctx.C, ctx.T

    def step(k):
        ctx.simulate(inst, slots, keys, 0, S, seed=20250102, stream=k & 0xFFFF)
        L = ctx.obs_loglik(inst)
        return ctx.timing(), L

    step(10000)
    ctx.synchronize()
    sim_ms, obs_ms = [], []
    t0 = time.perf_counter()
    nstep = max(1, args.steps)
    for k in range(nstep):
        tm, L = step(k)
        sim_ms.append(tm[0])
        obs_ms.append(tm[1])
    wall = time.perf_counter() - t0
    assert np.isfinite(L).all()
    value = nstep * P * S / wall
    # One ABC-SMC generation with fresh parameters: set_param + simulate + log-likelihood.
    fresh = specmod.sample_prior(spec, rng, P)
    t1 = time.perf_counter()
    ctx.set_param(slots, fresh)
    t2 = time.perf_counter()
    step(50000)
    t3 = time.perf_counter()
    ns = 6
    t4 = time.perf_counter()
    ctx.simulate(inst[:ns], slots[:ns], keys[:ns], 0, S, seed=20250102, stream=777, plain=True)
    for i in range(ns):
        ctx.obs_loglik(inst[i:i + 1])
    plain_value = ns * S / (time.perf_counter() - t4)
    sim = float(np.mean(sim_ms))
    bpp = bytes_per_particle_step(A, G, C_, T)
    achieved = bpp * P * S / (sim * 1e-3) / 1e9
    out = {"workload": "spatial ABC-SMC: 100 areas x 4 age groups, 8 compartments, 10-NN mixing, 10-knot Rt spline, "
                       "200 days, 4096 particles", "value": value, "unit": "particle-timesteps/s",
           "ms_per_step": wall / nstep * 1e3,
           "fresh_generation_ms": {"set_param": (t2 - t1) * 1e3, "simulate_and_loglik": (t3 - t2) * 1e3,
                                   "total": (t3 - t1) * 1e3},
           "plain": {"value": plain_value, "batched_over_plain": value / plain_value},
           "kernels_ms": {"simulate_fused": sim, "obs_loglik": float(np.mean(obs_ms))},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "bytes_per_particle_step": bpp}}
    if cpub is not None:
        out["cpu_baseline"] = cpub
    ctx.close()
    return out


def bench_copy(args):
    """A plain device-to-device copy of known size (the calibration kernel for
    the FETCH_SIZE / WRITE_SIZE counters): 256 MiB read + 256 MiB written per
    launch."""
    import torch
    n = 64 << 20
    x = torch.arange(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert bool((y[-4:] == x[-4:]).all())
    out = {"metric": "copy GB/s (counter calibration)", "value": 2 * 4 * n * args.steps / wall / 1e9, "unit": "GB/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "device copy", "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n}}
    print(json.dumps(out), flush=True)


def bench_dense(args):
    """317 areas x 9 age groups with dense gravity mixing, 300 days in ten
    30-day observation sections: one step is one PMCMC likelihood estimate
    (PMCMC::obs_prob) over 1024 particles sharing one parameter vector --
    per section the step-wise simulation (contraction as an MFMA GEMM,
    replayed from a hipGraph), the section log-likelihood, the FP64 bootstrap
    and the state moves, all on the device."""
    from beep_amd import _lib, configs
    spec, truth = configs.dense_pmcmc(with_obs=False)
    c0 = _lib.Context(spec, capacity=1, nslot=1, keep=_lib.KEEP_TRANSNUM)
    c0.set_param([0], truth[None, :])
    c0.simulate([0], [0], [123456], seed=17, stream=1)
    spec["obs"] = configs.daily_observations(c0.get_transnum(0), tr=6)
    c0.close()
    P = 1024
    ctx = _lib.Context(spec, capacity=P, nslot=1)
    ctx.set_param([0], truth[None, :])
    inst = np.arange(P, dtype=np.int32)
    keys = inst.astype(np.uint32)
    for w in range(max(1, args.warmup)):
        ctx.pmcmc_obs_prob(inst, 0, keys, invT=1.0, seed=3, stream=900 + w)
    ctx.synchronize()
    ms, est = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ob, _bp = ctx.pmcmc_obs_prob(inst, 0, keys, invT=1.0, seed=3, stream=k)
        ms.append(ctx.timing()[0])
        est.append(ob)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    S = ctx.S
    ctx.simulate(inst, np.zeros(P, dtype=np.int32), keys, seed=3, stream=999)
    sim_ms = ctx.timing()[0]
    flop = 2.0 * ctx.A * ctx.A * ctx.G * P * S
    out = {"metric": METRIC,
           "value": args.steps * P * S / wall, "unit": "particle-timesteps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": max(1, args.warmup), "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32 counts, fp32 rates, fp64 log-likelihood",
           "data": "synthetic",
           "config": {"workload": "dense PMCMC: 317 areas x 9 age groups, 10 compartments, dense gravity mixing, "
                                  "300 days, 10 sections, 1024 particles, particle-filter likelihood per step",
                      "observations": "daily I->H per area from one simulation at the true parameters"},
           "pf_device_ms": float(np.mean(ms)), "log_likelihood_estimates": est,
           "simulate_ms_full_horizon": sim_ms,
           "contraction_flop_per_particle_step": 2.0 * ctx.A * ctx.A * ctx.G,
           "contraction_tflops_algorithmic_over_whole_simulate": flop / (sim_ms * 1e-3) / 1e12}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_serving(args, rates=None, nreq=None, emit=True):
    """Prediction requests with Poisson arrivals at each offered rate: model
    sizes from the single-area SEIR to 100 areas x 4 age groups, 100-1000
    posterior samples, an R(t) factor in [0.5, 1.5] after a random change
    point, 28-180 day horizons; p50/p99 arrival-to-result latency and
    particle-timesteps/s."""
    from beep_amd import serving
    rates = rates or [float(x) for x in args.rates.split(",")]
    nreq = nreq or args.requests
    srv = serving.PredictionServer(serving.CATALOGUE, t_post=100, max_horizon=180, nposterior=1000,
                                   hbm_budget=96 << 30, chunk=14)
    print("serving: models loaded", file=sys.stderr, flush=True)
    warm = serving.make_requests(8, rate=1000.0, catalogue_size=len(serving.CATALOGUE), seed=1)
    srv.run(warm, realtime=False)
    rows = []
    for rate in rates:
        reqs = serving.make_requests(nreq, rate=rate, catalogue_size=len(serving.CATALOGUE), seed=int(rate))
        t0 = time.perf_counter()
        srv.run(reqs, realtime=True)
        wall = time.perf_counter() - t0
        pts = sum(r.nsamp * r.horizon for r in reqs)
        row = {"offered_rate": rate, "requests": len(reqs), "wall_s": wall, "particle_timesteps_per_s": pts / wall}
        row.update(serving.latency_stats(reqs))
        rows.append(row)
        print(f"serving: rate {rate}/s done: {row}", file=sys.stderr, flush=True)
    srv.close()
    best = max(rows, key=lambda r: r["particle_timesteps_per_s"])
    out = {"metric": "serving: particle-timesteps/sec with p50/p99 request latency",
           "value": best["particle_timesteps_per_s"], "unit": "particle-timesteps/s", "n_gpus": 1, "steps": len(rows),
           "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "int32 counts, fp32 rates", "data": "synthetic",
           "config": {"workload": "prediction serving", "catalogue": serving.CATALOGUE,
                      "posterior_samples_per_model": 1000, "chunk_days": 14},
           "rates": rows}
    if emit:
        print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
