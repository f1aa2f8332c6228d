"""Serving loop (continuous batching): requests on windows of resident regional
meteorologies, admitted while others run.  Each served result is bit-identical
to the same request served alone and to the same request stepped by f3d_step
(the batched and the single launch run the same tile code), and within the
FP32 bar of the oracle."""
import copy

import numpy as np
import pytest

from fall3d_amd import lib as F
from fall3d_amd import scenario as S
from fall3d_amd import serve
from oracle.oracle import OracleRun

pytestmark = pytest.mark.gpu


def _regions():
    return [serve.Region(S.config_c2(nx=72, ny=64, nz=10, nc=3, seed=101 + k)) for k in range(2)]


def _requests(regions, n, end_time=900.0, seed=3):
    rng = np.random.default_rng(seed)
    times = np.sort(rng.uniform(0.0, 0.4, size=n))
    return [serve.make_request(k, float(t), regions, rng, nx=41, ny=37, nz=8, nc=3,
                               end_time=end_time) for k, t in enumerate(times)]


def test_served_requests_equal_alone_and_oracle():
    regions = _regions()
    reqs = _requests(regions, 7)
    fresh = [copy.deepcopy(r) for r in reqs]
    srv = serve.Server(regions, reserve_bytes=1 << 30, chunk=3)
    done = srv.run(reqs)
    assert len(done) == len(reqs)
    assert all(r.steps_done == r.nsteps and r.nsteps > 0 for r in done)
    by_id = {r.rid: r for r in done}
    for k in (0, 3, 6):
        # served alone: the same request through a server with nothing else
        q = copy.deepcopy(fresh[k])
        q.arrival = 0.0
        alone = serve.Server(regions, reserve_bytes=1 << 30, chunk=5).run([q])[0]
        assert alone.nsteps == by_id[k].nsteps
        assert np.array_equal(alone.cdry[:, :, 1:], by_id[k].cdry[:, :, 1:]), k
        # stepped alone by f3d_step: the same bits
        q = copy.deepcopy(fresh[k])
        ctx = srv._open(q)
        try:
            ctx.step(q.nsteps)
            cdry = ctx.ground_load()
        finally:
            ctx.close()
        assert np.array_equal(cdry[:, :, 1:], by_id[k].cdry[:, :, 1:]), k
    # against the oracle on the window's host meteorology
    r = by_id[2]
    o = OracleRun(regions[r.region].window_scenario(fresh[2]))
    o.meteo_update()
    for _ in range(r.nsteps):
        o.step()
    load = r.cdry[:, :, 1:].sum(-1)
    ref = o.cdry[:, :, 1:].sum(-1)
    e_load = float(np.abs(load - ref).sum() / np.abs(ref).sum())
    assert e_load <= 1e-4, e_load
    st = serve.latency_stats(done)
    assert st["n"] == len(reqs) and st["p99"] >= st["p50"] > 0


def test_meteo_window_matches_host_window():
    regions = _regions()
    req = _requests(regions, 1, seed=11)[0]
    sc = regions[req.region].window_scenario(req)
    a = F.Run(sc, precision=F.FP32)
    srv = serve.Server(regions)
    ctx = srv._open(copy.deepcopy(req))
    try:
        a.meteo_update()
        assert ctx.dt == a.dt
        ca, cb = a.ctx.coef(), ctx.coef()
        for key in ("vx", "rho", "rkhor", "rkh1", "rkver", "vset"):
            assert np.array_equal(ca[key], cb[key]), key
    finally:
        a.close()
        ctx.close()
