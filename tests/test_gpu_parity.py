"""Parity of the HIP path with the FP64 oracle.

Fixed trajectories (the oracle's own stochastic draws loaded into the device):
populations bit-exact, transition means within FP32 rounding, observation
log-likelihood within 1e-6 relative.  Stochastic simulation: per-compartment
means, variances and two-sample KS statistics over 10^4 replicates."""
import numpy as np
import pytest
from scipy import stats

import toml_models
from beep_amd import _lib, configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ctx(spec, cap, nslot=None, keep=_lib.KEEP_ALL):
    return _lib.Context(spec, capacity=cap, nslot=nslot or 1, keep=keep)


def _transmean_magnitude(orc, tb, st):
    """The means recomputed from the absolute values of every summand
    (|inf * pop|, |weights|): the scale of the FP32 rounding error."""
    d, nz = tb.derived, orc.nz
    A, G, T, S = orc.A, orc.G, orc.T, orc.S
    itr = nz["itr"]
    out = np.abs(st.transmean).copy()
    for s in range(S):
        pop = st.pop[s]
        dinf = np.einsum("c,acg->ag", np.abs(d["inf"]), np.abs(pop))
        idg = nz["diag"][:, None] * dinf
        imp = np.zeros((A, G))
        for c in range(A):
            for j in range(nz["rowptr"][c], nz["rowptr"][c + 1]):
                imp[nz["to"][j]] += abs(nz["val"][j]) * dinf[c]
        m = d["geo"][s]
        I = (idg + imp) * abs(m) + np.abs(nz["factor"])[:, None] * idg * abs(1 - m)
        nmi = I @ np.abs(d["Ntime"][s]).T
        S_ = np.maximum(pop[:, orc.spec["trans"][itr]["from"], :], 0)
        lam = (orc.spec["period"] / S) * S_ * d["sus"][None, :] * (
            np.abs(d["beta"][s])[:, None] * d["af"][s // orc.spec["spd"]][:, None] * nmi + abs(d["efoi"][s]))
        out[s][:, itr, :] = np.maximum(out[s][:, itr, :], lam)
    return out


@pytest.mark.parametrize("mk", [configs.seir_single, configs.spatial_abcsmc, toml_models.toml_seir_erlang,
                                toml_models.toml_seird_branch, toml_models.toml_sir_3areas, toml_models.toml_sird_ages,
                                toml_models.toml_seir_sex_ages, toml_models.toml_sird_strains])
def test_fixed_trajectory_replay(gpu_lib, mk):
    spec, truth = mk()
    orc = O.Oracle(spec)
    tb = orc.tables(truth)
    st = orc.simulate(tb, seed=20250101)
    ctx = _ctx(spec, 1)
    ctx.set_param([0], truth[None, :])
    ctx.put_transnum(0, 0, st.transnum.astype(np.int32))
    # Integer state update: bit-exact.
    pop = ctx.get_pop(0)
    assert np.array_equal(pop.astype(np.float64), st.pop)
    # Transition means: FP32 rates against the FP64 oracle.
    # The reference updates Imap incrementally in FP64 and leaves residues of
    # ~1e-22 where the exact infectivity is zero (means ~1e-19 instead of 0);
    # the device recomputes it from the populations.  Means above 1e-9 are
    # compared relatively, the rest absolutely.
    # Negative populations (uncapped Poisson draws, src/state.cc:620-630) make
    # the force of infection a sum of terms of both signs, so the FP32 error is
    # bounded relative to the sum of the terms' magnitudes.
    tm = ctx.get_transmean(0).astype(np.float64)
    ref = st.transmean
    mag = _transmean_magnitude(orc, tb, st)
    err = np.abs(tm - ref)
    assert (err <= 2e-5 * ref + 1e-5 * mag + 1e-12).all(), (err - 2e-5 * ref - 1e-5 * mag).max()
    big = ref > 1e-3 * np.maximum(mag, 1e-300)
    assert np.median(np.abs(tm[big] / ref[big] - 1)) < 1e-6
    # Observation log-likelihood: FP64 on device vs FP64 oracle.
    L_gpu = ctx.obs_loglik([0])[0]
    L_ref, values = orc.obs_loglik(st)
    assert np.array_equal(ctx.obs_values(0), values)
    assert abs(L_gpu / L_ref - 1) < 1e-6, (L_gpu, L_ref)
    # State::likelihood with the device's FP32 means.
    Lp = ctx.pois_loglik([0])[0]
    Lp_ref = orc.likelihood(st)
    assert abs(Lp / Lp_ref - 1) < 1e-5, (Lp, Lp_ref)
    ctx.close()


def test_obs_sections_and_datatables(gpu_lib):
    spec, truth = configs.spatial_abcsmc()
    spec["sections"] = [(0, 50), (50, 120), (120, 200)]
    orc = O.Oracle(spec)
    st = orc.simulate(orc.tables(truth), seed=5)
    ctx = _ctx(spec, 1)
    ctx.set_param([0], truth[None, :])
    ctx.put_transnum(0, 0, st.transnum.astype(np.int32))
    L, EFdt = ctx.obs_loglik([0], with_dt=True)
    Lr, _ = orc.obs_loglik(st)
    assert abs(L[0] / Lr - 1) < 1e-12
    # per data table (get_EF_datatable)
    obs = spec["obs"]
    for d in range(2):
        sel = [i for i, o in enumerate(obs) if o["datatable"] == d]
        Ld, _ = orc.obs_loglik(st, sel=sel)
        assert abs(EFdt[0, d] / (-2 * Ld) - 1) < 1e-12
    for k, (ti, tf) in enumerate(spec["sections"]):
        sel = [i for i, o in enumerate(obs) if o["sett_i"] >= ti and o["sett_f"] <= tf]
        Ls, _ = orc.obs_loglik(st, ti, tf, sel=sel)
        Lg = ctx.obs_loglik([0], section=k)[0]
        assert abs(Lg / Ls - 1) < 1e-12
    ctx.close()


def test_more_than_eight_datatables(gpu_lib):
    """Eleven data tables (the reference loops over data.ndatatable without a
    bound, src/obsmodel.cc:309-322): every table's -2 log-likelihood against
    the oracle's sum over that table's observations, and the total."""
    spec, truth = configs.spatial_abcsmc()
    spec["obs"] = [dict(o, datatable=i % 11) for i, o in enumerate(spec["obs"])]
    orc = O.Oracle(spec)
    st = orc.simulate(orc.tables(truth), seed=8)
    ctx = _ctx(spec, 2)
    assert ctx.ndt == 11
    ctx.set_param([0], truth[None, :])
    for inst in (0, 1):
        ctx.put_transnum(inst, 0, st.transnum.astype(np.int32))
    L, EFdt = ctx.obs_loglik([0, 1], with_dt=True)
    Lr, _ = orc.obs_loglik(st)
    assert abs(L[0] / Lr - 1) < 1e-12 and L[1] == L[0]
    for d in range(11):
        sel = [i for i, o in enumerate(spec["obs"]) if o["datatable"] == d]
        Ld, _ = orc.obs_loglik(st, sel=sel)
        assert abs(EFdt[0, d] / (-2 * Ld) - 1) < 1e-12, d
        assert EFdt[1, d] == EFdt[0, d]
    ctx.close()


def _ks_compare(a, b, label):
    d, p = stats.ks_2samp(a, b)
    assert p > 1e-4, (label, d, p)
    ma, mb = a.mean(), b.mean()
    se = np.sqrt(a.var() / len(a) + b.var() / len(b)) + 1e-12
    assert abs(ma - mb) < 5 * se, (label, ma, mb, se)
    if a.var() > 0 and b.var() > 0:
        assert 0.85 < a.var() / b.var() < 1.18, (label, a.var(), b.var())


def test_simulate_distribution_seir(gpu_lib):
    """10^4 device replicates vs 10^4 oracle replicates of the single-area SEIR."""
    spec, truth = configs.seir_single()
    n = 10_000
    orc = O.Oracle(spec)
    tb = orc.tables(truth)
    ref = np.zeros((n, 4, spec["period"]))
    for r in range(n):
        st = orc.simulate(tb, seed=1000 + r)
        ref[r] = st.pop[:, 0, :, 0].T
    ctx = _ctx(spec, n, keep=_lib.KEEP_ALL)
    ctx.set_param([0], truth[None, :])
    ctx.simulate(np.arange(n), np.zeros(n), np.arange(n) + 77, seed=2024)
    gpu = np.zeros_like(ref)
    for r in range(n):
        gpu[r] = ctx.get_pop(r)[:, 0, :, 0].T
    for t in (10, 30, 60, 99):
        for c, name in enumerate("SEIR"):
            _ks_compare(gpu[:, c, t], ref[:, c, t], f"{name}@{t}")
    ctx.close()


def test_simulate_distribution_strains(gpu_lib):
    """Two strains drawing on pooled susceptibles: 5000 device replicates vs
    5000 oracle replicates; per strain, the cumulative infections and the
    infected at days 30 and 69, and the pooled susceptibles."""
    spec, truth = toml_models.toml_sird_strains()
    n = 5000
    orc = O.Oracle(spec)
    tb = orc.tables(truth)
    itr = [t["infection"] for t in spec["trans"]].index(True)
    S = spec["period"] * spec["spd"]

    def summary(pop, tn):
        out = []
        for t in (30, S - 1):
            out += [tn[:t, 0, itr, 0:2].sum(), tn[:t, 0, itr, 2:4].sum(), pop[t, 0, 1, 0:2].sum(),
                    pop[t, 0, 1, 2:4].sum(), pop[t, 0, 0, 0:2].sum()]
        return out

    ref = np.array([summary((st := orc.simulate(tb, seed=3000 + r)).pop, st.transnum) for r in range(n)])
    ctx = _ctx(spec, n, keep=_lib.KEEP_ALL)
    ctx.set_param([0], truth[None, :])
    ctx.simulate(np.arange(n), np.zeros(n), np.arange(n) + 91, seed=2025)
    gpu = np.array([summary(ctx.get_pop(r), ctx.get_transnum(r)) for r in range(n)])
    assert np.all(gpu[:, 4] > 0) and ref[:, 1].mean() > 10
    for k in range(ref.shape[1]):
        _ks_compare(gpu[:, k], ref[:, k], f"summary {k}")
    ctx.close()


def test_simulate_distribution_spatial(gpu_lib):
    """Per-compartment totals of the spatial model over 2000 replicates."""
    spec, truth = configs.spatial_abcsmc()
    truth = truth.copy()
    truth[:10] = 2.0
    n = 2000
    orc = O.Oracle(spec)
    tb = orc.tables(truth)
    S = spec["period"]
    ts = (40, 80, 120, 199)
    ref = np.zeros((n, len(ts), 8))
    for r in range(n):
        st = orc.simulate(tb, seed=5000 + r)
        ref[r] = st.pop[list(ts)].sum(axis=(1, 3))
    ctx = _ctx(spec, n)
    ctx.set_param([0], truth[None, :])
    ctx.simulate(np.arange(n), np.zeros(n), np.arange(n) + 1, seed=99)
    gpu = np.zeros_like(ref)
    for r in range(n):
        gpu[r] = ctx.get_pop(r)[list(ts)].sum(axis=(1, 3))
    for i, t in enumerate(ts):
        for c in (0, 3, 4, 6):
            _ks_compare(gpu[:, i, c], ref[:, i, c], f"c{c}@{t}")
    ctx.close()


def test_observation_spline_follows_the_instance_slot(gpu_lib):
    """An observation spline takes its values from the parameter slot an
    instance was last simulated or replayed with, and a copied trajectory
    takes its source's: two slots with different spline knots give the
    oracle's likelihood for each, a copy reproduces its source's, and an
    instance never run with a slot is refused."""
    spec, truth = toml_models.toml_seird_branch()
    names = [p["name"] for p in spec["params"]]
    th2 = truth.copy()
    for k in ("ascertainment t=0", "ascertainment t=42", "ascertainment t=84"):
        th2[names.index(k)] = 0.35
    orc = O.Oracle(spec)
    ctx = _lib.Context(spec, capacity=4, nslot=2)
    ctx.set_param([0, 1], np.stack([truth, th2]))
    tn = orc.simulate(orc.tables(truth), seed=77).transnum
    L = []
    for inst, (slot, th) in enumerate(((0, truth), (1, th2))):
        ctx.put_transnum(inst, slot, tn.astype(np.int32))
        L.append(orc.obs_loglik(orc.replay(orc.tables(th), tn))[0])
    Lg = ctx.obs_loglik([0, 1])
    assert Lg[0] == pytest.approx(L[0], rel=1e-9) and Lg[1] == pytest.approx(L[1], rel=1e-9)
    assert abs(L[0] - L[1]) > 1.0
    ctx.copy_instances([2], [1])
    assert ctx.obs_loglik([2])[0] == pytest.approx(L[1], rel=1e-9)
    with pytest.raises(_lib.BeepError, match="no parameter slot"):
        ctx.obs_loglik([3])
    ctx.close()


def test_batch_independence(gpu_lib):
    """An instance's trajectory does not depend on its slot or its batch."""
    spec, truth = configs.spatial_abcsmc()
    ctx = _ctx(spec, 8, nslot=2)
    ctx.set_param([0, 1], np.stack([truth, truth]))
    ctx.simulate([3], [1], [555], seed=7)
    alone = ctx.get_transnum(3)
    ctx.simulate([0, 1, 2, 5, 6], [0, 0, 1, 0, 1], [1, 2, 555, 4, 5], seed=7)
    mixed = ctx.get_transnum(2)
    assert np.array_equal(alone, mixed)
    ctx.simulate([0], [0], [1], seed=7, plain=True)
    plain = ctx.get_transnum(0)
    ctx.simulate([7], [0], [1], seed=7)
    assert np.array_equal(plain, ctx.get_transnum(7))
    ctx.close()


def test_mbp_identity_and_law(gpu_lib):
    """Unchanged parameters give n_p == n_i exactly; changed parameters give
    proposals distributed as fresh simulations (the MBP property)."""
    spec, truth = configs.seir_single()
    n = 10_000
    ctx = _ctx(spec, 3 * n, nslot=2)
    th2 = truth.copy()
    th2[0] = 2.5
    ctx.set_param([0, 1], np.stack([truth, th2]))
    init = np.arange(n)
    ctx.simulate(init, np.zeros(n), init + 1, seed=11)
    prop = init + n
    ctx.mbp(init, prop, np.zeros(n), init + 1, seed=12)
    for r in (0, 17, 999):
        assert np.array_equal(ctx.get_transnum(r), ctx.get_transnum(r + n))
    ctx.mbp(init, prop, np.ones(n), init + 1, seed=13)
    fresh = init + 2 * n
    ctx.simulate(fresh, np.ones(n), fresh + 1, seed=14)
    a = np.array([ctx.get_pop(r)[-1, 0, 3, 0] for r in prop])
    b = np.array([ctx.get_pop(r)[-1, 0, 3, 0] for r in fresh])
    _ks_compare(a, b, "R final after MBP")
    ctx.close()


def test_mbp_identity_and_law_strains(gpu_lib):
    """MBP with two strains drawing on pooled susceptibles: unchanged
    parameters reproduce the trajectory exactly, and proposals at a higher
    second-strain Rfactor follow the oracle's MBP law (final recovered and
    dead of each strain, 4000 device vs 4000 oracle proposals)."""
    spec, truth = toml_models.toml_sird_strains()
    names = [p["name"] for p in spec["params"]]
    n = 4000
    ctx = _ctx(spec, 3 * n, nslot=2)
    th2 = truth.copy()
    th2[names.index("Rfactor_s2")] = 1.9
    ctx.set_param([0, 1], np.stack([truth, th2]))
    init = np.arange(n)
    ctx.simulate(init, np.zeros(n), init + 1, seed=21)
    prop = init + n
    ctx.mbp(init, prop, np.zeros(n), init + 1, seed=22)
    for r in (0, 17, 999):
        assert np.array_equal(ctx.get_transnum(r), ctx.get_transnum(r + n))
        assert np.array_equal(ctx.get_pop(r), ctx.get_pop(r + n))
    ctx.mbp(init, prop, np.ones(n), init + 1, seed=23)
    orc = O.Oracle(spec)
    tb1, tb2 = orc.tables(truth), orc.tables(th2)
    ref = [orc.mbp(tb2, orc.simulate(tb1, seed=10000 + r), seed=20000 + r).pop[-1, 0] for r in range(n)]
    for co in (2, 3):
        for strain in (slice(0, 2), slice(2, 4)):
            a = np.array([ctx.get_pop(r)[-1, 0, co, strain].sum() for r in prop])
            b = np.array([p[co, strain].sum() for p in ref])
            _ks_compare(a, b, f"compartment {co} strain {strain} after MBP")
    ctx.close()


def test_mbp_against_oracle_law(gpu_lib):
    """Device MBP proposals and oracle MBP proposals (libstdc++ samplers) from
    their own initial states have the same law."""
    spec, truth = configs.seir_single()
    n = 4000
    th2 = truth.copy()
    th2[0] = 1.7
    orc = O.Oracle(spec)
    tb1, tb2 = orc.tables(truth), orc.tables(th2)
    ref = []
    for r in range(n):
        st = orc.simulate(tb1, seed=300 + r)
        pr = orc.mbp(tb2, st, seed=900000 + r)
        ref.append(pr.pop[-1, 0, :, 0])
    ref = np.array(ref)
    ctx = _ctx(spec, 2 * n, nslot=2)
    ctx.set_param([0, 1], np.stack([truth, th2]))
    init = np.arange(n)
    ctx.simulate(init, np.zeros(n), init + 1, seed=21)
    ctx.mbp(init, init + n, np.ones(n), init + 1, seed=22)
    gpu = np.array([ctx.get_pop(r)[-1, 0, :, 0] for r in init + n])
    for c in range(4):
        _ks_compare(gpu[:, c], ref[:, c], f"c{c}")
    ctx.close()


def test_accept_matches_closed_form(gpu_lib):
    spec, truth = configs.seir_single()
    n = 64
    ctx = _ctx(spec, 2 * n, nslot=1)
    ctx.set_param([0], truth[None, :])
    ctx.simulate(np.arange(2 * n), np.zeros(2 * n), np.arange(2 * n) + 5, seed=3)
    L = ctx.obs_loglik(np.arange(2 * n))
    EF = -2 * L
    rng = np.random.default_rng(0)
    dlogq = rng.normal(0, 1, n)
    init, prop = np.arange(n), np.arange(n) + n
    for mode, invT, EFcut in ((1, 0.3, 0.0), (0, 1.0, float(np.median(EF)))):
        al, acc = ctx.accept(init, prop, mode, dlogq, np.arange(n), invT=invT, EFcut=EFcut, seed=8)
        for i in range(n):
            want = O.get_al(mode, EF[prop[i]], EF[init[i]], dlogq[i], 0.0, invT, EFcut)
            assert abs(al[i] - want) <= 1e-12 * max(1.0, want)
        assert acc[al >= 1.0].all() and not acc[al <= 0.0].any()
    ctx.close()


def test_all_observation_kinds(gpu_lib):
    """ObservationModel::obs_prob's branches (src/obsmodel.cc:245-305) on one
    fixed trajectory: normal with given sd and with a percentage, Poisson,
    negative binomial, below-threshold for normal and count models, unknown
    values, factors, multi-day windows over several areas, subsets of age
    groups and transitions, population terms, three data tables -- device
    FP64 log-likelihood and per-table EF against the FP64 oracle."""
    spec, truth = configs.seir_areas(6, 4, 60, seed=5)
    spec = dict(spec)
    kinds = []
    for k in range(60):
        d0 = (7 * k) % 53
        base = {"sett_i": d0, "sett_f": d0 + 1 + (k % 7), "area": [k % 6, (k + 2) % 6][: 1 + k % 2],
                "dp": [[0], [1, 2], [0, 1, 2, 3]][k % 3], "tr": [[0], [1], [0, 1], [2]][k % 4], "datatable": k % 3,
                "graph": k % 3, "factor": [1.0, 0.5, 2.0][k % 3]}
        t = k % 8
        if t == 0:
            base.update(type="normal", sd=3.0 + k % 5)
        elif t == 1:
            base.update(type="normal_percent", percent=10.0)
        elif t == 2:
            base.update(type="poisson")
        elif t == 3:
            base.update(type="negbin", shape=2.5)
        elif t == 4:
            base.update(type="poisson", thresh=4.0)
        elif t == 5:
            base.update(type="normal", sd=2.0, thresh=3.0)
        elif t == 6:
            base.update(type="poisson")
        else:
            base.update(type="negbin", shape=0.7, tr=[], co=[2])      # population of compartment 2
        kinds.append(base)
    spec["obs"] = [dict(o, y=0.0) for o in kinds]
    orc0 = O.Oracle(spec)
    st = orc0.simulate(orc0.tables(truth), seed=31)
    _L, values = orc0.obs_loglik(st)
    obs = []
    for k, (o, v) in enumerate(zip(kinds, values)):
        o = dict(o)
        if k % 8 in (4, 5):
            o["y"] = 999999998.0                                      # below threshold
        elif k % 16 == 6:
            o["y"] = 999999997.0                                      # unknown
        else:
            o["y"] = float(np.floor(max(v, 0.0) * (0.8 + 0.05 * (k % 9)) + 0.5))
        if o["type"] == "normal_percent" and o["y"] == 0:
            o["y"] = 1.0
        obs.append(o)
    spec["obs"] = obs
    orc = O.Oracle(spec)
    tb = orc.tables(truth)
    st = orc.simulate(tb, seed=31)
    L_ref, values = orc.obs_loglik(st)
    ctx = _ctx(spec, 1)
    ctx.set_param([0], truth[None, :])
    ctx.put_transnum(0, 0, st.transnum.astype(np.int32))
    assert np.array_equal(ctx.obs_values(0), values)
    L = ctx.obs_loglik([0])[0]
    assert np.isfinite(L_ref) and abs(L / L_ref - 1) < 1e-12, (L, L_ref)
    # per data table: -2 x the sum over its observations
    _Lr, ef = ctx.obs_loglik([0], with_dt=True)
    per = np.zeros(3)
    for k in range(3):
        sel = [i for i, o in enumerate(obs) if o["datatable"] == k]
        per[k] = -2 * orc.obs_loglik(st, sel=sel)[0]
    assert np.allclose(ef[0], per, rtol=1e-12, atol=1e-9), (ef, per)
    ctx.close()


def test_mbp_mask_against_oracle_law(gpu_lib):
    """Mbp::mbp with a simulate-afresh mask [S][A] (src/mbp.cc:252-330 with the
    masks of mbp_slicetime / mbp_fixedtree, 656-716): the joint law of
    (initial, proposal) -- through the change in each area's final recovered
    count -- matches the oracle's, in masked and unmasked areas.  A mask read
    in the wrong layout or ignored changes these laws."""
    spec, truth = configs.seir_areas(5, 1, 50, seed=8)
    S, A = 50, 5
    mask = np.zeros((S, A), dtype=np.uint8)
    mask[10:30, 1] = 1
    mask[:, 3] = 1
    th2 = truth.copy()
    th2[0] *= 1.05
    n = 3000
    orc = O.Oracle(spec)
    tb1, tb2 = orc.tables(truth), orc.tables(th2)
    C_ = len(spec["comps"])
    ref = []
    for r in range(n):
        st = orc.simulate(tb1, seed=500 + r)
        pr = orc.mbp(tb2, st, seed=700000 + r, simu=mask)
        ref.append(pr.pop[-1, :, C_ - 1, 0] - st.pop[-1, :, C_ - 1, 0])
    ref = np.array(ref)
    ctx = _ctx(spec, 2 * n, nslot=2)
    ctx.set_param([0, 1], np.stack([truth, th2]))
    init = np.arange(n)
    ctx.simulate(init, np.zeros(n), init + 1, seed=31)
    ctx.mbp(init, init + n, np.ones(n), init + 1, simu_mask=np.repeat(mask[None], n, 0), seed=32)
    gpu = np.array([ctx.get_pop(r + n)[-1, :, C_ - 1, 0] - ctx.get_pop(r)[-1, :, C_ - 1, 0] for r in init])
    for a in range(A):
        d, pv = stats.ks_2samp(gpu[:, a], ref[:, a])
        assert pv > 1e-4, (a, d, pv)
        se = np.sqrt(gpu[:, a].var() / n + ref[:, a].var() / n) + 1e-12
        assert abs(gpu[:, a].mean() - ref[:, a].mean()) < 5 * se, a
    # masked areas move far more than coupled ones
    assert gpu[:, 3].std() > 2 * gpu[:, 0].std()
    ctx.close()
