"""Slice-time and fixed-tree MBP proposals on the device (Mbp::mbp_slicetime
/ mbp_fixedtree, src/mbp.cc:656-716): a mask over every cell draws the whole
proposal afresh (it is the simulation with the same key), an empty mask at
unchanged parameters reproduces the chain (accepted with al = 1), and each
chain's sim_frac adapts by the closed form from its own acceptance."""
import math

import numpy as np
import pytest

from beep_amd import _lib, configs, mbp_drivers as md

pytestmark = pytest.mark.gpu


def _chains(n):
    spec, truth = configs.seir_areas(20, 4, 60, seed=11, daily_obs=True)
    ctx = _lib.Context(spec, capacity=3 * n, nslot=1)
    ctx.set_param([0], truth[None, :])
    init = np.arange(n)
    ctx.simulate(init, np.zeros(n), 100 + init, seed=3, stream=2)
    ctx.obs_loglik(init)
    return spec, ctx, init


def test_full_mask_is_the_simulation(gpu_lib):
    n = 16
    spec, ctx, init = _chains(n)
    prop, fresh = init + n, init + 2 * n
    pps = [md.SliceTime(0, ctx.S, sim_frac=1.0) for _ in range(n)]
    keys = 900 + init
    md.propose(ctx, init, prop, np.zeros(n), pps, keys, np.random.default_rng(1), seed=5, stream=7,
               adapt_frac=False)
    ctx.simulate(fresh, np.zeros(n), keys, seed=5, stream=7)
    for i in range(n):
        assert np.array_equal(ctx.get_transnum(int(prop[i])), ctx.get_transnum(int(fresh[i])))
    ctx.close()


def test_empty_mask_reproduces_the_chain(gpu_lib):
    n = 16
    spec, ctx, init = _chains(n)
    prop = init + n
    tree = md.area_split(spec, np.random.default_rng(3))
    pps = [md.SliceTime(5, 40, sim_frac=0.0) if i % 2 else md.FixedTree(1 + i % 5, sim_frac=0.0) for i in range(n)]
    al, acc = md.propose(ctx, init, prop, np.zeros(n), pps, 700 + init, np.random.default_rng(2), seed=5, stream=8,
                         tree=tree)
    assert (al == 1.0).all() and acc.all()
    for i in range(n):
        assert np.array_equal(ctx.get_transnum(int(prop[i])), ctx.get_transnum(int(init[i])))
        assert pps[i].ntr == 1 and pps[i].nac == 1.0 and pps[i].sim_frac == 0.0
    ctx.close()


def test_partial_masks_adapt_per_chain(gpu_lib):
    n = 32
    spec, ctx, init = _chains(n)
    prop = init + n
    tree = md.area_split(spec, np.random.default_rng(4))
    pps = [md.SliceTime(10, 50, sim_frac=0.5) if i % 2 else md.FixedTree(1 + i % 7, sim_frac=0.5) for i in range(n)]
    al, acc = md.propose(ctx, init, prop, np.zeros(n), pps, 300 + init, np.random.default_rng(6), seed=9, stream=3,
                         tree=tree, fast=True)
    EFp = -2 * ctx.obs_loglik(prop)
    EFi = -2 * ctx.obs_loglik(init)
    for i in range(n):
        want = math.exp(min(0.0, -0.5 * (EFp[i] - EFi[i])))
        assert al[i] == pytest.approx(want, rel=1e-12, abs=1e-300)
        assert pps[i].sim_frac == pytest.approx(min(1.0, 0.5 * math.exp(al[i] - 0.33)))
        assert pps[i].ntr == 1 and pps[i].nac == pytest.approx(al[i])
    ctx.close()
