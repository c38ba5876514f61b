"""Parity with the FP64 oracle at the benchmark workloads' real sizes, through
the device operations the benchmarks themselves run:

  dense 317 x 9 x 300 (10 compartments, dense gravity mixing): instances
      simulated on the step-wise path (MFMA contraction over the zero-padded
      320-area operands, rows spanning two 64-row tiles), every trajectory
      kept; each device trajectory replayed through the oracle;
  uneven workload: the bench's own driver (UnevenBatch) over a workload
      holding the largest model of the benchmark (400 areas x 9 age groups x
      400 days) next to small ones: ABC-SMC simulations over their own
      horizons and MC3 coupled proposals in one mixed batch with their
      log-likelihoods; each device trajectory replayed through the oracle at
      its slot's parameters.

The bar: populations bit-exact, transition means within the FP32 bound
(relative to the magnitudes of their summands), observation log-likelihood
within 1e-6 relative."""
import copy

import numpy as np
import pytest

from beep_amd import _lib, configs, spec as specmod
from beep_amd.batch import UnevenBatch
from oracle import oracle as O
from test_gpu_at_size import check_replay, transmean_scale

pytestmark = pytest.mark.gpu


def test_dense_317x9_simulation_against_oracle(gpu_lib):
    """Config 3's model: 8 instances (two at the true parameters, six at
    prior draws) simulated together on the device; the oracle replays the
    device's counts and gets the device's populations, means and
    log-likelihoods."""
    spec, truth = configs.dense_pmcmc()
    orc = O.Oracle(spec)
    rng = np.random.default_rng(3)
    params = np.vstack([truth, truth, specmod.sample_prior(spec, rng, 6)])
    n = len(params)
    ctx = _lib.Context(spec, capacity=n, nslot=n)
    assert ctx.path in (_lib.PATH_STEPWISE, _lib.PATH_STEPWISE_SPEC)
    assert (ctx.A, ctx.G, ctx.S) == (317, 9, 300)
    ctx.set_param(range(n), params)
    ctx.simulate(range(n), range(n), 77 + np.arange(n), seed=5, stream=6)
    for i in (0, 1, 5):
        tb = orc.tables(params[i])
        st = orc.replay(tb, ctx.get_transnum(i).astype(np.float64))
        check_replay(ctx, i, orc, tb, st, f"dense simulation {i}")
    # Two instances at the same parameters with different keys differ.
    assert not np.array_equal(ctx.get_transnum(0), ctx.get_transnum(1))
    ctx.close()


def _workload():
    return {"abc": [{"A": 10, "G": 1, "S": 50, "horizons": [50]},
                    {"A": 173, "G": 4, "S": 120, "horizons": [120, 61]},
                    {"A": 400, "G": 9, "S": 400, "horizons": [400, 231]}],
            "mc3": [{"A": 400, "G": 9, "S": 400}, {"A": 57, "G": 4, "S": 90}],
            "ladder": 8}


def test_uneven_batch_against_oracle(gpu_lib):
    """One sweep's batch of the uneven driver with the benchmark's largest
    model: every particle's and two chains' per ladder trajectories replayed
    through the oracle at their parameters."""
    ub = UnevenBatch(_workload(), seed=7)
    # The sweep's fresh parameters, drawn as UnevenBatch.sweep draws them.
    rng = copy.deepcopy(ub.rng)
    abc_pv = specmod.sample_prior(ub.spec0, rng, len(ub.abc_slots))
    prop = ub._reflect(ub.pv + rng.normal(0.0, 1.0, ub.pv.shape) * ub.step_sd)
    nl, L = len(ub.mc3_ctx), ub.L
    h = np.concatenate([ub.h_abc, ub.h_mc3])
    slots = np.concatenate([ub.abc_slots, ub.alt.reshape(-1)])
    _lib.batch_set_param(h, slots, np.concatenate([abc_pv, prop.reshape(-1, ub.nparam)]))
    rec, na = ub.sweep_records()
    _stats, Lik = _lib.run_batch(rec, seed=4, stream=9, loglik=True)
    assert np.isfinite(Lik).all()
    oracles = {}

    def oracle_of(ctx):
        if id(ctx) not in oracles:
            oracles[id(ctx)] = O.Oracle(specmod.expand_obs(ctx.spec))
        return oracles[id(ctx)]

    checked = 0
    # Particles: counts over their horizons, zero after; the final state; the
    # log-likelihood.
    for i, (ctx, _kind, inst, _init, slot, _key, _a, hz) in enumerate(ub.abc_items):
        orc = oracle_of(ctx)
        tb = orc.tables(abc_pv[i])
        tn = ctx.get_transnum(inst)
        assert not tn[hz:].any(), i
        st = orc.replay(tb, tn.astype(np.float64))
        assert np.array_equal(ctx.get_state(inst).astype(np.float64), st.pop[-1]), f"particle {i}: final state"
        Lr, _ = orc.obs_loglik(st)
        assert abs(Lik[i] / Lr - 1) < 1e-6, (i, Lik[i], Lr)
        checked += 1
    # Chains' proposals: means (kept), final state, log-likelihood.
    for l, ctx in enumerate(ub.mc3_ctx):
        orc = oracle_of(ctx)
        for k in (0, L - 1):
            inst = int(ub.alt[l, k])
            tb = orc.tables(prop[l, k])
            st = orc.replay(tb, ctx.get_transnum(inst).astype(np.float64))
            assert np.array_equal(ctx.get_state(inst).astype(np.float64), st.pop[-1]), f"chain {l},{k}: final state"
            tm = ctx.get_transmean(inst).astype(np.float64)
            mag = transmean_scale(orc, tb, st)
            err = np.abs(tm - st.transmean)
            assert (err <= 2e-5 * st.transmean + 1e-5 * mag + 1e-12).all(), (l, k)
            Lr, _ = orc.obs_loglik(st)
            j = na + l * L + k
            assert abs(Lik[j] / Lr - 1) < 1e-6, (l, k, Lik[j], Lr)
            checked += 1
    assert checked == 5 + 2 * nl
    assert max(c.A * c.G for c in ub.ctx) == 3600
    ub.close()
