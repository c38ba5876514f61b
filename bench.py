"""Throughput of the batched simulation + observation log-likelihood path.

Workload (one "step"): 4096 ABC-SMC particles of the 100-area x 4-age-group
spatial model (8 compartments, 10-nearest-neighbour mixing, 10-knot R(t)
spline), each simulated over 200 daily tau-leap steps and scored by the
observation model; parameters drawn from the uniform priors.
Metric: particle-timesteps/s = particles x daily steps / wall time, inputs
resident in HBM.  Also reported: the plain path (one instance per launch per
step) on a sample, the fused kernel's roofline, and the FP64 CPU oracle on
this host's cores as a baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PARTICLES = 4096
METRIC = "particle-timesteps/sec batched vs plain GPU path; p50/p99 request latency"
# HBM bytes per launch of the simulation kernel on this workload (4096
# particles), from rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, each scaled
# by the ratio of known bytes to counter reading on a plain copy kernel:
# the specialised kernel (profiles/beep_fused_spec_sim_spatial_abcsmc.md) and
# the generic one (profiles/k_fused_spatial_abcsmc_traffic.md).
SPEC_TRAFFIC_BYTES = 35_483_868_282
GENERIC_TRAFFIC_BYTES = 41_895_375_565


def algorithmic_bytes_per_particle_step(A, G, C, T):
    """HBM bytes the fused kernel must move per particle per step with the
    State trajectory kept: transnum (int32) + transmean (fp32) per (transition,
    cell), pop (int32) per (compartment, cell), and the per-step tables it
    reads (beta*areafactor per area, efoi per age group, m(t))."""
    AG = A * G
    return 4 * AG * (2 * T + C) + 4 * A + 4 * G + 4


def _cpu_worker(args):
    """One host core: simulate + score particles idx, idx+stride, ... for
    `seconds`; returns (particles done, particle-timesteps, seconds)."""
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


def cpu_cores():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline(spec, params, seconds=10.0, cores=None):
    """FP64 CPU oracle (the reference's libstdc++ samplers), one process per
    host core with a static split of the particles (as the reference's MPI
    ranks split them), over a bounded sample of the same particles.  Runs in
    forked processes before this process touches the GPU."""
    import multiprocessing as mp
    cores = cores or cpu_cores()
    jobs = [(spec, params, k, cores, seconds) for k in range(cores)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t0
    n = sum(r[0] for r in res)
    pts = sum(r[1] for r in res)
    per_core = [r[1] / r[2] for r in res]
    return {"value": sum(per_core), "unit": "particle-timesteps/s", "cores": cores, "kind": "port",
            "sample": f"{n} of the {len(params)} particles (same priors), 200 steps each, simulate + observation "
                      f"log-likelihood, {cores} processes x {seconds:.0f} s (wall {wall:.1f} s); single core "
                      f"{float(np.mean(per_core)):.0f} particle-timesteps/s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=PARTICLES)
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--generic", action="store_true", help="use the generic fused kernel, not the specialised one")
    ap.add_argument("--workload", default="spatial", choices=["spatial", "uneven", "dense", "serving", "copy"])
    ap.add_argument("--rates", default="5,20,80")
    ap.add_argument("--requests", type=int, default=120)
    ap.add_argument("--instances", type=int, default=16384)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path shards by particle: "
              f"one GPU's share is what is measured here)", file=sys.stderr)
        sys.exit(2)

    from beep_amd import _lib, configs, spec as specmod

    if args.workload == "uneven":
        return bench_uneven(args)
    if args.workload == "dense":
        return bench_dense(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "copy":
        return bench_copy(args)
    spec, _truth = configs.spatial_abcsmc()
    P = args.particles
    rng = np.random.default_rng(20250102)
    params = specmod.sample_prior(spec, rng, P)
    # The CPU baseline runs first, in forked processes, before anything here
    # initialises the GPU.
    cpu = None if args.no_cpu else cpu_baseline(spec, params)
    ctx = _lib.Context(spec, capacity=P, nslot=P, keep=_lib.KEEP_ALL | (_lib.GENERIC_KERNEL if args.generic else 0))
    kernel = {_lib.PATH_FUSED: "k_fused<MODE_SIM> (generic)", _lib.PATH_FUSED_SPEC: "beep_fused_spec_sim (specialised)",
              _lib.PATH_STEPWISE: "step-wise"}[ctx.path]
    slots = np.arange(P, dtype=np.int32)
    ctx.set_param(slots, params)
    inst = np.arange(P, dtype=np.int32)
    keys = np.arange(P, dtype=np.uint32)
    S, A, G, C, T = ctx.S, ctx.A, ctx.G, ctx.C, ctx.T

    def step(k):
        ctx.simulate(inst, slots, keys, 0, S, seed=20250102, stream=k & 0xFFFF)
        L = ctx.obs_loglik(inst)
        return ctx.timing(), L

    for w in range(args.warmup):
        step(10000 + w)
    ctx.synchronize()
    sim_ms, obs_ms = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        tm, L = step(k)
        sim_ms.append(tm[0])
        obs_ms.append(tm[1])
    ctx.synchronize()
    wall = time.perf_counter() - t0
    assert np.isfinite(L).all()
    pts = args.steps * P * S
    value = pts / wall

    # Plain path: one instance per launch per step, on a sample of particles.
    ns = args.plain_sample
    t1 = time.perf_counter()
    ctx.simulate(inst[:ns], slots[:ns], keys[:ns], 0, S, seed=20250102, stream=777, plain=True)
    for i in range(ns):
        ctx.obs_loglik(inst[i:i + 1])
    ctx.synchronize()
    plain_wall = time.perf_counter() - t1
    plain_value = ns * S / plain_wall

    sim_avg = float(np.mean(sim_ms))
    bytes_per = algorithmic_bytes_per_particle_step(A, G, C, T)
    achieved = bytes_per * P * S / (sim_avg * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 counts, fp32 rates, fp64 log-likelihood",
        "data": "synthetic (uniform-prior parameters, synthetic observations)",
        "config": {"workload": "spatial ABC-SMC: 100 areas x 4 age groups, 8 compartments, 10-NN mixing, "
                               "10-knot Rt spline, 200 days, 4096 particles",
                   "particles": P, "days": S, "areas": A, "age_groups": G},
        "plain": {"value": plain_value, "sample_instances": ns, "batched_over_plain": value / plain_value},
        "kernels_ms": {"simulate_fused": sim_avg, "obs_loglik": float(np.mean(obs_ms))},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (None if P != PARTICLES else GENERIC_TRAFFIC_BYTES if args.generic
                                 else SPEC_TRAFFIC_BYTES),
                     "bytes_per_particle_step": bytes_per, "algorithmic_bytes_per_launch": bytes_per * P * S},
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    ctx.close()
    print(json.dumps(out), flush=True)


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


def bench_uneven(args):
    """Mixed batch: ABC-SMC particles (fresh simulations over their own
    horizons) and MC3 chains (MBP proposals coupled to their current
    trajectories) of SEIR-in-every-area models with 10..400 areas, 1/4/9 age
    groups and 50..400-day horizons, one batch call per sweep."""
    import time
    from beep_amd import _lib, configs
    from beep_amd.batch import UnevenBatch
    groups = configs.uneven_batch(args.instances, seed=4)
    t0 = time.perf_counter()
    ub = UnevenBatch(groups, seed=4, keep_extra=_lib.GENERIC_KERNEL if args.generic else 0)
    setup = time.perf_counter() - t0
    print(f"uneven: {len(ub.ctx)} contexts set up in {setup:.1f} s", file=sys.stderr, flush=True)
    _lib.run_batch(ub.init_items, seed=4, stream=1)             # the chains' current trajectories
    items = ub.sim_items + ub.mbp_items
    _lib.run_batch(items, seed=4, stream=2, autotune=True)      # tunes the size classes not yet cached
    pts = ub.particle_steps(items)
    walls, cells = [], 0
    for k in range(args.steps):
        st = _lib.run_batch(items, seed=4, stream=10 + k)
        walls.append(st[0])
        cells = st[3]
        print(f"uneven: sweep {k} {st[0]:.1f} ms, {int(st[1])} fused launches, {int(st[2])} step-wise contexts",
              file=sys.stderr, flush=True)
    wall = float(np.mean(walls)) * 1e-3
    out = {"metric": "particle-timesteps/sec (uneven batch: ABC-SMC simulations + MC3 MBP proposals)",
           "value": pts / wall, "unit": "particle-timesteps/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "int32 counts, fp32 rates", "data": "synthetic",
           "config": {"workload": "uneven batch", "instances": args.instances, "topologies": len(groups),
                      "contexts": len(ub.ctx), "abc_particles": len(ub.sim_items), "mc3_chains": len(ub.mbp_items),
                      "particle_steps_per_sweep": pts, "cell_steps_per_sweep": cells,
                      "device_gb": ub.device_bytes / 1e9, "setup_s": setup},
           "cell_steps_per_s": cells / wall}
    ub.close()
    print(json.dumps(out), flush=True)


def bench_dense(args):
    """317 areas x 9 age groups with dense gravity mixing, 300 days in ten
    30-day observation sections: one step is one PMCMC likelihood estimate
    (PMCMC::obs_prob) over 1024 particles sharing one parameter vector --
    per section the step-wise simulation (contraction as an MFMA GEMM,
    replayed from a hipGraph), the section log-likelihood, the FP64 bootstrap
    and the state moves, all on the device."""
    import time
    from beep_amd import _lib, configs
    # Observations: daily I->H counts per area from one simulation at the true
    # parameters (synthetic data the model can explain).
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
    # Contraction alone: one full-horizon simulation, its GEMM flops over the
    # simulate kernels' device time.
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


def bench_serving(args):
    """Prediction requests with Poisson arrivals at each offered rate: model
    sizes from the single-area SEIR to 100 areas x 4 age groups, 100-1000
    posterior samples, an R(t) factor in [0.5, 1.5] after a random change
    point, 28-180 day horizons; p50/p99 arrival-to-result latency and
    particle-timesteps/s."""
    import time
    from beep_amd import serving
    srv = serving.PredictionServer(serving.CATALOGUE, t_post=100, max_horizon=180, nposterior=1000,
                                   hbm_budget=96 << 30, chunk=14)
    print("serving: models loaded", file=sys.stderr, flush=True)
    warm = serving.make_requests(8, rate=1000.0, catalogue_size=len(serving.CATALOGUE), seed=1)
    srv.run(warm, realtime=False)
    rows = []
    for rate in [float(x) for x in args.rates.split(",")]:
        reqs = serving.make_requests(args.requests, rate=rate, catalogue_size=len(serving.CATALOGUE), seed=int(rate))
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
    out = {"metric": "serving: particle-timesteps/sec with p50/p99 request latency", "value": best["particle_timesteps_per_s"],
           "unit": "particle-timesteps/s", "n_gpus": 1, "steps": len(rows), "warmup": 1, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32 counts, fp32 rates",
           "data": "synthetic", "config": {"workload": "prediction serving", "catalogue": serving.CATALOGUE,
                                           "posterior_samples_per_model": 1000, "chunk_days": 14},
           "rates": rows}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
