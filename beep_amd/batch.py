"""Uneven batches: ABC-SMC particles and MC3 chains of many model sizes in
one mixed launch (the batch-submitting counterpart of the reference drivers,
which advance one particle at a time: src/abcsmc.cc:42-81, src/mc3.cc:60-66)."""
import time

import numpy as np

from . import _lib, configs, spec as specmod


def run(items, seed=1, stream=0, autotune=False):
    """Runs (ctx, kind, inst, init, slot, key, t0, t1) work items: those of
    fused-path contexts in one mixed batch (beep_run_batch), those of step-wise
    contexts (which the batch refuses) through their own context's simulate /
    mbp calls, grouped by (kind, t0, t1).  Returns the batch statistics with
    [0] the wall time of both parts and [2] the step-wise calls made."""
    fused = [it for it in items if it[0].path in (_lib.PATH_FUSED, _lib.PATH_FUSED_SPEC)]
    step = [it for it in items if it[0].path not in (_lib.PATH_FUSED, _lib.PATH_FUSED_SPEC)]
    t0 = time.perf_counter()
    stats = _lib.run_batch(fused, seed=seed, stream=stream, autotune=autotune) if fused else np.zeros(4)
    groups = {}
    for it in step:
        ctx, kind, inst, init, slot, key, a, b = it
        groups.setdefault((id(ctx), kind, a, b), (ctx, []))[1].append(it)
    for (_, kind, a, b), (ctx, its) in groups.items():
        inst = [it[2] for it in its]
        slot = [it[4] for it in its]
        keys = [it[5] for it in its]
        if kind == _lib.WORK_SIMULATE:
            ctx.simulate(inst, slot, keys, a, b, seed=seed, stream=stream)
        else:
            ctx.mbp([it[3] for it in its], inst, slot, keys, seed=seed, stream=stream)
        stats[3] += len(its) * ctx.A * ctx.G * ((b - a) if kind == _lib.WORK_SIMULATE else ctx.S)
    stats = np.array(stats, dtype=float)
    stats[0] = (time.perf_counter() - t0) * 1e3
    stats[2] = len(groups)
    return stats


class UnevenBatch:
    """The uneven workload on one device (configs.uneven_batch): ABC-SMC
    particles of SEIR-in-every-area models grouped by (areas, age groups),
    one context per group keeping transnum, and MC3 chains in ladders of
    tempered chains, one context per ladder keeping transnum and transmean
    (what MBP proposals couple to).  Every model observes daily I->R counts
    per area.

    One sweep (`sweep`) is the work of one ABC-SMC generation and one MC3
    update of every chain, the counterpart of the reference drivers'
    per-particle loops (src/abcsmc.cc:42-81; src/mc3.cc:60-71 with
    Mbp::mc3_mcmc_updates and MC3::swap_states):
      1. fresh parameters: ABC-SMC particles drawn from the prior (the
         generation's proposals), MC3 proposals by a symmetric random walk
         reflected into the prior's bounds; all slots derived and uploaded
         in one batch_set_param call;
      2. one mixed batch: the particles' simulations over their horizons and
         the chains' MBP proposals coupled to their current trajectories,
         then every item's observation log-likelihood;
      3. Metropolis-Hastings acceptance of the proposals at each chain's
         temperature on the device (FP64);
      4. neighbour swaps along every ladder on the device.
    Chain states never move: accepting a proposal or a swap relabels which
    instance (and slot) holds a chain's current state."""

    def __init__(self, workload, seed=4, keep_extra=0, invT_power=3.0):
        self.workload = workload
        self.rng = np.random.default_rng(seed)
        rng = self.rng
        self.ctx = []
        self.abc_items = []
        key = 1
        # ABC-SMC particle groups.
        self.abc_ctx, self.abc_n = [], []
        for gi, g in enumerate(workload["abc"]):
            spec, _truth = configs.seir_areas(g["A"], g["G"], g["S"], seed=seed * 100003 + gi, daily_obs=True)
            n = len(g["horizons"])
            ctx = _lib.Context(spec, capacity=n, nslot=n, keep=_lib.KEEP_TRANSNUM | keep_extra)
            self.ctx.append(ctx)
            self.abc_ctx.append(ctx)
            self.abc_n.append(n)
            for i, h in enumerate(g["horizons"]):
                self.abc_items.append((ctx, _lib.WORK_SIMULATE, i, 0, i, key, 0, h))
                key += 1
        self.spec0 = spec
        self.nparam = len(spec["params"])
        # MC3 ladders: chain k of a ladder holds its state in instance cur[l][k]
        # (slot of the same number); its proposals go to alt[l][k].
        L = workload["ladder"]
        self.L = L
        self.mc3_ctx = []
        for li, ld in enumerate(workload["mc3"]):
            spec, _truth = configs.seir_areas(ld["A"], ld["G"], ld["S"], seed=seed * 100003 + 50000 + li,
                                              daily_obs=True)
            ctx = _lib.Context(spec, capacity=2 * L, nslot=2 * L,
                               keep=_lib.KEEP_TRANSNUM | _lib.KEEP_TRANSMEAN | keep_extra)
            self.ctx.append(ctx)
            self.mc3_ctx.append(ctx)
        nl = len(self.mc3_ctx)
        self.cur = np.tile(np.arange(L, dtype=np.int32), (nl, 1))
        self.alt = self.cur + L
        self.mc3_keys = np.arange(nl * L, dtype=np.uint32).reshape(nl, L) + key
        self.invT = (1.0 - np.arange(L) / L) ** invT_power          # posterior chain first
        pri = spec["params"]
        self.lo = np.array([p["prior"][1] if p["prior"][0] == "uniform" else p["prior"][1] for p in pri])
        self.hi = np.array([p["prior"][2] if p["prior"][0] == "uniform" else p["prior"][1] for p in pri])
        self.step_sd = 0.05 * (self.hi - self.lo)
        self.h_abc = _lib.handles([c for c, n in zip(self.abc_ctx, self.abc_n) for _ in range(n)])
        self.abc_slots = np.concatenate([np.arange(n, dtype=np.int32) for n in self.abc_n]) if self.abc_n else \
            np.zeros(0, np.int32)
        self.h_mc3 = np.repeat(_lib.handles(self.mc3_ctx), L) if nl else np.zeros(0, np.uint64)
        self.mc3_S = np.array([c.S for c in self.mc3_ctx], dtype=np.int64)
        # The sweep's work as one beep_work record array: the particles'
        # records are fixed; the chains' take their instances from cur / alt
        # at each sweep (sweep_records).
        na = len(self.abc_items)
        self.rec = np.zeros(na + nl * L, dtype=_lib.WORK_DTYPE)
        if na:
            self.rec[:na] = _lib.work_records(_lib.handles([it[0] for it in self.abc_items]), _lib.WORK_SIMULATE,
                                              [it[2] for it in self.abc_items], 0, [it[4] for it in self.abc_items],
                                              [it[5] for it in self.abc_items], 0, [it[7] for it in self.abc_items])
        if nl:
            self.rec[na:] = _lib.work_records(self.h_mc3, _lib.WORK_MBP, 0, 0, 0, self.mc3_keys.reshape(-1), 0,
                                              np.repeat(self.mc3_S, L))
        self.ladder_ptr = np.arange(0, nl * L + 1, L, dtype=np.int32)
        self.invT_chains = np.tile(self.invT, nl)
        # Chains' current parameters and prior densities, and their first states.
        self.pv = specmod.sample_prior(self.spec0, rng, nl * L).reshape(nl, L, self.nparam) if nl else \
            np.zeros((0, L, self.nparam))
        if nl:
            Pr, _ = _lib.batch_set_param(self.h_mc3, self.cur.reshape(-1), self.pv.reshape(-1, self.nparam))
            self.Pr = Pr.reshape(nl, L)
            items = [(c, _lib.WORK_SIMULATE, int(self.cur[l, k]), 0, int(self.cur[l, k]), int(self.mc3_keys[l, k]), 0,
                      c.S) for l, c in enumerate(self.mc3_ctx) for k in range(L)]
            _lib.run_batch(items, seed=seed, stream=1, loglik=True)

    def _reflect(self, x):
        lo, hi = self.lo, self.hi
        span = np.where(hi > lo, hi - lo, 1.0)
        y = np.abs((x - lo) % (2 * span))
        y = np.where(y > span, 2 * span - y, y)
        return np.where(hi > lo, lo + y, lo)

    def sweep_items(self):
        """The sweep's batch items (ABC-SMC simulations, then MC3 proposals)."""
        mbp = [(c, _lib.WORK_MBP, int(self.alt[l, k]), int(self.cur[l, k]), int(self.alt[l, k]),
                int(self.mc3_keys[l, k]), 0, c.S) for l, c in enumerate(self.mc3_ctx) for k in range(self.L)]
        return self.abc_items, mbp

    def sweep_records(self):
        """sweep_items as one beep_work record array (the particles, then the
        chains) and the number of particle records."""
        na = len(self.abc_items)
        mbp = self.rec[na:]
        mbp["inst"] = self.alt.reshape(-1)
        mbp["init"] = self.cur.reshape(-1)
        mbp["slot"] = self.alt.reshape(-1)
        return self.rec, na

    def particle_steps(self):
        return sum(it[7] - it[6] for it in self.abc_items) + int(self.mc3_S.sum()) * self.L

    def sweep(self, k, seed=4, autotune=False):
        """One sweep; returns its timings (ms) and counts."""
        import time
        stream = (100 + k) & 0xFFFF
        nl, L = len(self.mc3_ctx), self.L
        out = {}
        t0 = time.perf_counter()
        # 1. parameters
        abc_pv = specmod.sample_prior(self.spec0, self.rng, len(self.abc_slots))
        prop = self._reflect(self.pv + self.rng.normal(0.0, 1.0, self.pv.shape) * self.step_sd)
        h = np.concatenate([self.h_abc, self.h_mc3])
        slots = np.concatenate([self.abc_slots, self.alt.reshape(-1)])
        Pr, _ = _lib.batch_set_param(h, slots, np.concatenate([abc_pv, prop.reshape(-1, self.nparam)]))
        Pr_prop = Pr[len(self.abc_slots):].reshape(nl, L)
        t1 = time.perf_counter()
        # 2. simulations, proposals, log-likelihoods
        rec, na = self.sweep_records()
        stats, Lik = _lib.run_batch(rec, seed=seed, stream=stream, autotune=autotune, loglik=True)
        mbp = rec[na:]
        t2 = time.perf_counter()
        # 3. acceptance at each chain's temperature (symmetric proposal: the
        # log ratio is the prior's)
        nacc = 0
        if nl:
            al, acc = _lib.batch_accept(mbp, self.invT_chains, (Pr_prop - self.Pr).reshape(-1), mode=1,
                                        seed=seed, stream=stream)
            acc = acc.reshape(nl, L)
            nacc = int(acc.sum())
            cur = np.where(acc, self.alt, self.cur)
            self.alt = np.where(acc, self.cur, self.alt)
            self.cur = cur
            self.pv = np.where(acc[..., None], prop, self.pv)
            self.Pr = np.where(acc, Pr_prop, self.Pr)
        t3 = time.perf_counter()
        # 4. swaps along the ladders
        nswap = 0
        if nl:
            order, _ntr, nac = _lib.batch_mc3_swap_arrays(self.ladder_ptr, self.h_mc3, self.cur.reshape(-1), 0,
                                                          self.invT_chains, seed=seed, stream=stream)
            order = order.reshape(nl, L)
            nswap = int(nac.sum())
            rows = np.arange(nl)[:, None]
            self.cur, self.alt = self.cur[rows, order], self.alt[rows, order]
            self.pv, self.Pr = self.pv[rows, order], self.Pr[rows, order]
        t4 = time.perf_counter()
        out.update(set_param_ms=(t1 - t0) * 1e3, batch_ms=(t2 - t1) * 1e3, accept_ms=(t3 - t2) * 1e3,
                   swap_ms=(t4 - t3) * 1e3, total_ms=(t4 - t0) * 1e3, batch_kernel_wall_ms=float(stats[0]),
                   fused_launches=int(stats[1]), cell_steps=float(stats[3]), accepted=nacc, swaps=nswap,
                   loglik_finite=bool(np.isfinite(Lik).all()))
        return out

    @property
    def device_bytes(self):
        return sum(c.device_bytes for c in self.ctx)

    def close(self):
        for c in self.ctx:
            c.close()
