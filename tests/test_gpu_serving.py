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
