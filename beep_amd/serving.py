"""Continuous-batching prediction server.

Requests are scenario / counterfactual predictions in the sense of the
reference's prediction mode (Simulate::model_modification,
src/simulate.cc:142-165): for each of N posterior samples, continue the
sample's state from the posterior time t_post over a horizon, with the R(t)
spline multiplied by a factor from a change-point day on (the `spline_fac`
modification, src/model.cc:1557-1575).

Everything a request needs stays in HBM: each model of the catalogue has a
context holding its posterior samples (their parameter slots and their
states at t_post) and a pool of work instances and slots sized from the HBM
budget at start-up.  A request is admitted when its model's pool has N free
work instances; the scheduler then advances every active request by up to
`chunk` steps per tick in one mixed batch launch (beep_run_batch), so a
request that arrives while others run joins at the next tick and its work
overlaps theirs on the device.  Draws are keyed by (request id, sample), so a
request's result does not depend on what it was batched with.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, configs, samples, spec as specmod

INFECTION_TR = 0          # S -> E in the SEIR catalogue models


@dataclass
class Request:
    rid: int
    model: int
    nsamp: int
    factor: float
    change_day: int          # days after t_post
    horizon: int             # days after t_post
    arrival: float = 0.0     # seconds since the server clock started
    # filled by the server
    admitted: float = -1.0
    done: float = -1.0
    result: np.ndarray = None          # [horizon] mean daily infections over samples
    _inst: list = field(default_factory=list)
    _slots: list = field(default_factory=list)
    _t: int = 0


class Model:
    """A catalogue entry: context, posterior samples, free work pool.  The
    posterior is loaded from a directory of saved samples (the reference's
    files, beep_amd.samples) when posterior_dir is given; otherwise it is
    made on the device: prior draws simulated up to t_post."""

    def __init__(self, idx, A, G, t_post, S, nposterior, pool, seed, posterior_dir=None):
        self.idx, self.A, self.G = idx, A, G
        self.t_post = t_post
        spec, _truth = configs.seir_areas(A, G, S, seed=1000 + idx)
        self.spec = spec
        cap = nposterior + pool
        self.ctx = _lib.Context(spec, capacity=cap, nslot=cap, keep=_lib.KEEP_TRANSNUM)
        inst = np.arange(nposterior)
        if posterior_dir is not None:
            # State::load of each sample (Simulate::model_modification,
            # src/simulate.cc:142-165): parameters and counts up to where the
            # samples end, replayed on the device; the prediction continues
            # from that step.
            _EF, post, ndiv = samples.load_posterior(posterior_dir, spec, self.ctx, inst, inst)
            if ndiv != t_post:
                raise ValueError(f"the samples in {posterior_dir} end at step {ndiv}, not at t_post = {t_post}")
        else:
            rng = np.random.default_rng(seed)
            post = specmod.sample_prior(spec, rng, nposterior)
            self.ctx.set_param(inst, post)
            # Posterior states at t_post (the state a loaded sample carries).
            self.ctx.simulate(inst, inst, inst + 7_000_000, 0, t_post, seed=seed, stream=1)
        self.post = post
        self.nposterior = nposterior
        self.free = list(range(nposterior, cap))       # work instance == work slot index

    def save_posterior(self, directory):
        """Writes the posterior samples (parameters, counts up to t_post) and
        sample_information.txt in the reference's format."""
        inst = np.arange(self.nposterior)
        samples.save_posterior(directory, self.spec, self.ctx, inst, inst, self.post, timesteps=self.t_post)

    def close(self):
        self.ctx.close()


class PredictionServer:
    def __init__(self, catalogue, t_post=100, max_horizon=180, nposterior=1000, hbm_budget=64 << 30,
                 chunk=14, seed=5, max_pool=4096, posterior_dirs=None):
        """catalogue: list of (areas, age groups).  The work pool of each model
        is sized so the whole server fits hbm_budget bytes.  posterior_dirs:
        per model, a directory of saved posterior samples to load (or None
        for a posterior made on the device)."""
        self.S = t_post + max_horizon
        self.t_post = t_post
        self.chunk = chunk
        self.seed = seed
        self.models = []
        per_inst = []
        for (A, G) in catalogue:
            C, T = 4, 3
            per_inst.append(4 * A * G * (C + T * self.S) + 4 * (self.S * (A + G + 1) + 32))
        fixed = sum(nposterior * p for p in per_inst)
        share = max(0, hbm_budget - fixed) / max(1, len(catalogue))
        for i, (A, G) in enumerate(catalogue):
            pool = int(min(max_pool, max(1000, share // per_inst[i])))
            pdir = posterior_dirs[i] if posterior_dirs is not None else None
            self.models.append(Model(i, A, G, t_post, self.S, nposterior, pool, seed + i, posterior_dir=pdir))
        self.pending, self.active, self.finished = [], [], []

    def save_posteriors(self, root):
        """Each model's posterior samples into root/model<i>; returns the
        directories (for posterior_dirs)."""
        import os
        dirs = []
        for m in self.models:
            d = os.path.join(root, f"model{m.idx}")
            m.save_posterior(d)
            dirs.append(d)
        return dirs

    def close(self):
        for m in self.models:
            m.close()

    # ---- request lifecycle ----
    def _admit(self, r, now):
        m = self.models[r.model]
        if len(m.free) < r.nsamp:
            return False
        r._inst = [m.free.pop() for _ in range(r.nsamp)]
        r._slots = list(r._inst)
        samples = [(r.rid * 7919 + k) % m.nposterior for k in range(r.nsamp)]
        m.ctx.copy_states(r._inst, samples)
        m.ctx.derive_slots(r._slots, samples, [r.factor] * r.nsamp, [self.t_post + r.change_day] * r.nsamp)
        r._t = self.t_post
        r.admitted = now
        self.active.append(r)
        return True

    def _finish(self, r, now):
        m = self.models[r.model]
        t1 = self.t_post + r.horizon
        sums = m.ctx.sum_transnum(r._inst, INFECTION_TR, self.t_post, t1)
        r.result = sums.mean(axis=0)
        m.free.extend(r._inst)
        r._inst, r._slots = [], []
        r.done = now

    def _tick_items(self):
        items = []
        for r in self.active:
            m = self.models[r.model]
            t1 = min(r._t + self.chunk, self.t_post + r.horizon)
            for k, (inst, slot) in enumerate(zip(r._inst, r._slots)):
                items.append((m.ctx, _lib.WORK_SIMULATE, inst, 0, slot, (r.rid << 10) + k, r._t, t1))
        return items

    def run(self, requests, realtime=True):
        """Serves `requests` (arrival times in seconds); returns them with
        their admission / completion times and results."""
        for r in requests:
            m = self.models[r.model]
            if r.nsamp > len(m.free) + sum(len(a._inst) for a in self.active if a.model == r.model):
                raise ValueError(f"request {r.rid} needs {r.nsamp} work instances; model {r.model} has "
                                 f"{m.ctx.capacity - m.nposterior}")
            if not (0 < r.horizon <= self.S - self.t_post) or not (0 <= r.change_day <= r.horizon):
                raise ValueError(f"request {r.rid}: horizon or change day out of range")
        clock0 = time.perf_counter()
        queue = sorted(requests, key=lambda r: r.arrival)
        qi = 0
        waiting = []
        while qi < len(queue) or waiting or self.active:
            now = time.perf_counter() - clock0
            if not realtime and not self.active and not waiting and qi < len(queue):
                now = max(now, queue[qi].arrival)
            while qi < len(queue) and queue[qi].arrival <= now:
                waiting.append(queue[qi])
                qi += 1
            still = []
            for r in waiting:                       # FIFO per model, as slots and memory free up
                if not self._admit(r, now):
                    still.append(r)
            waiting = still
            if not self.active:
                if qi < len(queue) and realtime:
                    time.sleep(max(0.0, min(0.01, queue[qi].arrival - now)))
                continue
            items = self._tick_items()
            _lib.run_batch(items, seed=self.seed, stream=3)
            for r in self.active:
                r._t = min(r._t + self.chunk, self.t_post + r.horizon)
            now = time.perf_counter() - clock0
            keep = []
            for r in self.active:
                if r._t >= self.t_post + r.horizon:
                    self._finish(r, now)
                    self.finished.append(r)
                else:
                    keep.append(r)
            self.active = keep
        return requests


def make_requests(n, rate, catalogue_size, seed=11, max_horizon=180):
    """Poisson arrivals at `rate` requests/s with sizes and counterfactuals
    drawn as in the serving workload."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        horizon = int(rng.integers(28, max_horizon + 1))
        out.append(Request(rid=i + 1, model=int(rng.integers(0, catalogue_size)), nsamp=int(rng.integers(100, 1001)),
                           factor=float(rng.uniform(0.5, 1.5)), change_day=int(rng.integers(0, horizon)),
                           horizon=horizon, arrival=float(t[i])))
    return out


CATALOGUE = [(1, 1), (10, 1), (25, 4), (40, 1), (50, 4), (75, 4), (100, 1), (100, 4)]


def latency_stats(reqs):
    lat = np.array([r.done - r.arrival for r in reqs])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3)}
