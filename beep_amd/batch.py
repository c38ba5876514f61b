"""Uneven batches: ABC-SMC particles and MC3 chains of many model sizes in
one mixed launch (the batch-submitting counterpart of the reference drivers,
which advance one particle at a time: src/abcsmc.cc:42-81, src/mc3.cc:60-66)."""
import numpy as np

from . import _lib, configs, spec as specmod


class UnevenBatch:
    """Contexts for each (areas, age groups) topology of an uneven workload:
    one for its ABC-SMC particles (keeping transnum, what the observation
    model reads) and one for its MC3 chains (keeping transnum and transmean,
    what MBP proposals couple to), parameters drawn from the priors."""

    def __init__(self, groups, seed=4, keep_extra=0):
        self.groups = groups
        self.ctx = []
        rng = np.random.default_rng(seed)
        self.sim_items, self.mbp_items, self.init_items = [], [], []
        key = 1
        for gi, g in enumerate(groups):
            spec, _truth = configs.seir_areas(g["A"], g["G"], g["S"], seed=seed * 100003 + gi)
            nsim, nmc3 = len(g["sim_horizons"]), g["n_mc3"]
            if nsim:
                ctx = _lib.Context(spec, capacity=nsim, nslot=nsim, keep=_lib.KEEP_TRANSNUM | keep_extra)
                ctx.set_param(np.arange(nsim), specmod.sample_prior(spec, rng, nsim))
                self.ctx.append(ctx)
                for i, h in enumerate(g["sim_horizons"]):
                    self.sim_items.append((ctx, _lib.WORK_SIMULATE, i, 0, i, key, 0, h))
                    key += 1
            if nmc3:
                ctx = _lib.Context(spec, capacity=2 * nmc3, nslot=2 * nmc3,
                                   keep=_lib.KEEP_TRANSNUM | _lib.KEEP_TRANSMEAN | keep_extra)
                ctx.set_param(np.arange(2 * nmc3), specmod.sample_prior(spec, rng, 2 * nmc3))
                self.ctx.append(ctx)
                for k in range(nmc3):
                    init, prop = 2 * k, 2 * k + 1
                    self.init_items.append((ctx, _lib.WORK_SIMULATE, init, 0, init, key, 0, g["S"]))
                    self.mbp_items.append((ctx, _lib.WORK_MBP, prop, init, prop, key, 0, g["S"]))
                    key += 1

    @staticmethod
    def particle_steps(items):
        return sum((it[7] - it[6]) if it[1] == _lib.WORK_SIMULATE else it[0].S for it in items)

    @property
    def device_bytes(self):
        return sum(c.device_bytes for c in self.ctx)

    def close(self):
        for c in self.ctx:
            c.close()
