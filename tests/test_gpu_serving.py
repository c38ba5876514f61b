"""Continuous-batching prediction server: a request's result is the same
alone and in a mix of concurrent requests; a counterfactual only changes the
prediction from its change-point on."""
import numpy as np
import pytest

from beep_amd import serving

pytestmark = pytest.mark.gpu

CAT = [(10, 1), (30, 4)]


def _server():
    return serving.PredictionServer(CAT, t_post=40, max_horizon=60, nposterior=200, hbm_budget=2 << 30, chunk=7,
                                    max_pool=1200)


def test_result_same_alone_and_in_the_mix(gpu_lib):
    target = serving.Request(rid=77, model=1, nsamp=150, factor=1.3, change_day=20, horizon=45, arrival=0.0)
    srv = _server()
    alone = srv.run([serving.Request(**{k: getattr(target, k) for k in
                                        ("rid", "model", "nsamp", "factor", "change_day", "horizon", "arrival")})],
                    realtime=False)[0].result
    srv.close()
    mix = serving.make_requests(12, rate=1000.0, catalogue_size=len(CAT), seed=3, max_horizon=60)
    for r in mix:
        r.nsamp = min(r.nsamp, 300)
    t = serving.Request(**{k: getattr(target, k) for k in
                           ("rid", "model", "nsamp", "factor", "change_day", "horizon")}, arrival=mix[4].arrival)
    srv = _server()
    out = srv.run(mix + [t], realtime=False)
    srv.close()
    # the target overlapped other requests on the device
    others = [r for r in out if r is not t]
    assert any(r.admitted <= t.done and r.done >= t.admitted for r in others)
    assert np.array_equal(t.result, alone)
    assert all(r.result is not None and len(r.result) == r.horizon for r in out)


def test_counterfactual_changes_only_after_change_point(gpu_lib):
    base = serving.Request(rid=5, model=1, nsamp=200, factor=1.0, change_day=25, horizon=50)
    cf = serving.Request(rid=5, model=1, nsamp=200, factor=1.5, change_day=25, horizon=50)
    srv = _server()
    b = srv.run([base], realtime=False)[0].result
    c = srv.run([cf], realtime=False)[0].result
    srv.close()
    assert np.array_equal(b[:25], c[:25])
    assert c[25:].sum() > b[25:].sum()


def _copy(r, **kw):
    d = {k: getattr(r, k) for k in ("rid", "model", "nsamp", "factor", "change_day", "horizon", "arrival")}
    d.update(kw)
    return serving.Request(**d)


def test_saved_loaded_served_equals_served(gpu_lib, tmp_path):
    """Posterior samples written in the reference's files, loaded into a new
    server (parameters, counts replayed on the device) and served give the
    results of serving them directly."""
    srv = _server()
    dirs = srv.save_posteriors(str(tmp_path))
    assert open(tmp_path / "model1" / "sample_information.txt").readline().strip() == "# Samples = '200'"
    reqs = serving.make_requests(6, rate=1000.0, catalogue_size=len(CAT), seed=8, max_horizon=60)
    for r in reqs:
        r.nsamp = min(r.nsamp, 200)
    direct = srv.run([_copy(r) for r in reqs], realtime=False)
    srv.close()
    srv2 = serving.PredictionServer(CAT, t_post=40, max_horizon=60, nposterior=200, hbm_budget=2 << 30, chunk=7,
                                    max_pool=1200, posterior_dirs=dirs)
    loaded = srv2.run([_copy(r) for r in reqs], realtime=False)
    srv2.close()
    for a, b in zip(direct, loaded):
        assert np.array_equal(a.result, b.result), a.rid


def test_full_catalogue_alone_and_in_the_mix(gpu_lib):
    """The serving workload's full catalogue (1 to 100 areas, 1 or 4 age
    groups, 1000 posterior samples, horizons up to 180 days): the largest
    request -- 1000 samples of the 100 x 4 model over 180 days -- gives the
    same prediction alone and among concurrent requests on every model."""
    def server():
        return serving.PredictionServer(serving.CATALOGUE, t_post=100, max_horizon=180, nposterior=1000,
                                        hbm_budget=24 << 30, chunk=14, max_pool=2500)
    big = serving.Request(rid=4242, model=len(serving.CATALOGUE) - 1, nsamp=1000, factor=0.7, change_day=30,
                          horizon=180, arrival=0.0)
    srv = server()
    alone = srv.run([_copy(big)], realtime=False)[0].result
    srv.close()
    mix = serving.make_requests(24, rate=1000.0, catalogue_size=len(serving.CATALOGUE), seed=13)
    t = _copy(big, arrival=mix[5].arrival)
    srv = server()
    out = srv.run(mix + [t], realtime=False)
    srv.close()
    assert any(r.admitted <= t.done and r.done >= t.admitted for r in out if r is not t)
    assert len({r.model for r in out}) >= 6
    assert np.array_equal(t.result, alone) and len(alone) == 180
