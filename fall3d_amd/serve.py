"""Forecast serving with continuous batching.

A request is a forecast (default 6 h) on a window of one of a few regional
meteorologies that stay resident in HBM (one FP32 context per region holds
its fields); the request's own context takes its meteorology from the
region with f3d_set_meteo_window, so no meteorology crosses PCIe after
start-up.  The serving loop

* admits arrived requests while free HBM (f3d_mem_info) stays above a
  reserve, oldest first;
* steps every admitted request together -- one f3d_batch over the active
  contexts, its steps replayed from HIP graphs -- in chunks of at most
  `chunk` steps, never past the step a request ends on;
* retires requests that reached their end time: their ground load is read
  back (cdry only) and their context freed, and the batch is rebuilt over
  the remaining ones, so requests arriving while others run join between
  chunks.

Each request is stepped exactly as f3d_step would step it alone (same tile
shape, same kernels), so its result does not depend on what else is in the
batch; tests/test_gpu_serve.py checks that bit for bit.

No counterpart in the reference (one forecast per run).
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time

import numpy as np

from . import lib as F
from . import scenario as S


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # s after the start of the serving run
    region: int
    i0: int                   # window origin (0-based) in the region grid
    j0: int
    nx: int
    ny: int
    nz: int
    diam: np.ndarray
    rhop: np.ndarray
    sphe: np.ndarray
    isrc: np.ndarray
    jsrc: np.ndarray
    ksrc: np.ndarray
    tmrat: np.ndarray
    flowrate: float
    end_time: float
    # filled in by the server
    nsteps: int = 0
    steps_done: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    cdry: np.ndarray = None   # final deposit cdry(nx,ny,0:nc)
    ctx: object = None

    @property
    def nc(self):
        return int(self.diam.size)

    @property
    def latency(self):
        return self.finished - self.arrival


class Region:
    """A regional meteorology resident in HBM (an FP32 context on its grid)."""

    def __init__(self, sc: S.Scenario):
        self.sc = sc
        self.ctx = F.Context(sc.nx, sc.ny, sc.nz, sc.nc, sc.dx, sc.dy, sc.zlayer, F.FP32)
        self.ctx.set_classes(sc.diam, sc.rhop, sc.sphe, sc.modv)
        self.ctx.set_models(sc.modkv, sc.modkh, sc.rkv0, sc.rkh0, sc.csh)
        self.ctx.set_meteo(sc.vx, sc.vy, sc.vz, sc.tempe, sc.vptemp, sc.rho, sc.ustar, sc.zi,
                           sc.rl, sc.vdep)

    def window_scenario(self, req: Request) -> S.Scenario:
        """The request as a standalone scenario (host meteo arrays sliced from
        the region's), for reference runs."""
        sc = self.sc
        i, j, nx, ny, nz = req.i0, req.j0, req.nx, req.ny, req.nz
        w3 = lambda a: np.asfortranarray(a[i:i + nx, j:j + ny, :nz])  # noqa: E731
        w2 = lambda a: np.asfortranarray(a[i:i + nx, j:j + ny])  # noqa: E731
        vdep = (w3(sc.vdep) if req.nc == sc.nc else np.zeros((nx, ny, req.nc), order="F"))
        return dataclasses.replace(
            sc, name=f"request{req.rid}", nx=nx, ny=ny, nz=nz, nc=req.nc,
            zlayer=sc.zlayer[:nz].copy(), diam=req.diam, rhop=req.rhop, sphe=req.sphe,
            fc=np.ones(req.nc) / req.nc, vx=w3(sc.vx), vy=w3(sc.vy), vz=w3(sc.vz),
            tempe=w3(sc.tempe), vptemp=w3(sc.vptemp), rho=w3(sc.rho), ustar=w2(sc.ustar),
            zi=w2(sc.zi), rl=w2(sc.rl), topg=w2(sc.topg), vdep=vdep, isrc=req.isrc,
            jsrc=req.jsrc, ksrc=req.ksrc, tmrat=req.tmrat, flowrate=req.flowrate,
            end_time=req.end_time)


def make_request(rid, arrival, regions, rng, nx=200, ny=200, nz=30, nc=16, end_time=6 * 3600.0):
    """A request drawn from the ensemble mixture: random region and window,
    nc classes (2 um - 2 mm), column height ~ U(6,18) km, log10 MER ~
    N(6,0.7), Suzuki A ~ U(2,6), vent at the window centre offset by
    U(+-10 km) per axis."""
    r = int(rng.integers(len(regions)))
    sc = regions[r].sc
    nx, ny, nz = min(nx, sc.nx), min(ny, sc.ny), min(nz, sc.nz)
    i0 = int(rng.integers(0, sc.nx - nx + 1))
    j0 = int(rng.integers(0, sc.ny - ny + 1))
    d, rhop, sphe, fc = S.make_classes(nc, 2e-6, 2e-3)
    H = rng.uniform(6000.0, 18000.0)
    mer = 10.0 ** rng.normal(6.0, 0.7)
    A = rng.uniform(2.0, 6.0)
    ox, oy = rng.uniform(-10000.0, 10000.0, size=2)
    zl = sc.zlayer[:nz]
    x0 = (nx - 1) * sc.dx / 2 + ox
    y0 = (ny - 1) * sc.dy / 2 + oy
    topg = sc.topg[i0:i0 + nx, j0:j0 + ny]
    src = S.suzuki_source(nx, ny, sc.dx, sc.dy, zl, x0, y0, min(H, zl[-1] * 0.95), mer, A, 5.0,
                          fc, topg=topg)
    return Request(rid, arrival, r, i0, j0, nx, ny, nz, d, rhop, sphe, src[0], src[1], src[2],
                   src[3], float(mer), float(end_time))


def poisson_arrivals(rate, duration, rng):
    """Arrival times (s) of a Poisson process of the given rate."""
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append(t)


def estimate_fp32_bytes(nx, ny, nz, nc):
    """Device bytes an FP32 context of this shape allocates (f3d_create):
    the FP64 meteorology and coefficients (10 fields and vset), the 2-D
    fields and deposits, the two FP32 state buffers with their ghost frame,
    and the FP32 step fields (7 per-cell and nc per-class fields in the
    tile-major layout of 64 records per tile column, the separate copies),
    with 5% and 1 MiB of slack.  Used for a request shape not seen before;
    once a context of the shape exists its measured size replaces it."""
    n = nx * ny * nz
    nxy = nx * ny
    g = (nx + 2) * (ny + 2) * (nz + 2)
    # tile columns: 58 own columns each, 60 at either end of the domain
    ftx = 1 if nx <= 62 else 1 + -(-(nx - 62) // 58)
    ftl = ftx * 64 * ny
    b64 = 8 * (10 * n + n * nc + nxy * (5 + nc + 2 * (nc + 1) + nc))
    b32 = 4 * (2 * g * nc + (5 + nc) * ftl * nz + (9 + nc) * n + 2 * nc * nxy + nxy)
    return int(1.05 * (b64 + b32)) + (1 << 20)


class Server:
    """Continuous-batching forecast server over resident regions."""

    def __init__(self, regions, reserve_bytes=8 << 30, max_active=256, chunk=8, tile=None,
                 graph=True):
        self.regions = list(regions)
        self.reserve = int(reserve_bytes)
        self.max_active = int(max_active)
        self.chunk = int(chunk)
        self.tile = tile
        self.graph = graph
        self._bytes = {}  # context bytes by request shape

    # -- one request's context --
    def _open(self, req: Request):
        reg = self.regions[req.region]
        sc = reg.sc
        ctx = F.Context(req.nx, req.ny, req.nz, req.nc, sc.dx, sc.dy, sc.zlayer[:req.nz], F.FP32)
        if self.tile is not None:
            ctx.set_tile(*self.tile)
        ctx.set_classes(req.diam, req.rhop, req.sphe, sc.modv)
        ctx.set_models(sc.modkv, sc.modkh, sc.rkv0, sc.rkh0, sc.csh)
        dt = ctx.set_meteo_window(reg.ctx, req.i0, req.j0)
        ctx.set_source(req.isrc, req.jsrc, req.ksrc, req.tmrat, req.flowrate)
        # Fall3d.f90:120: step while time < end_time
        n, t = 0, 0.0
        while t < req.end_time:
            t += dt
            n += 1
        req.nsteps = n
        return ctx

    def _fits(self, req: Request):
        key = (req.nx, req.ny, req.nz, req.nc)
        need = self._bytes.get(key)
        if need is None:
            need = estimate_fp32_bytes(req.nx, req.ny, req.nz, req.nc)
        free, _ = F.mem_info()
        return free - need > self.reserve

    def run(self, requests, log=None):
        """Serve the requests (arrival times relative to now); returns them
        with admission/finish times, steps and final deposits filled in."""
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue, active, done = collections.deque(), [], []
        batch = None
        self.busy_s = 0.0
        self.updates = 0
        while pending or queue or active:
            while pending and pending[0].arrival <= now():
                queue.append(pending.popleft())
            admitted = False
            while queue and len(active) < self.max_active and self._fits(queue[0]):
                req = queue.popleft()
                req.ctx = self._open(req)
                key = (req.nx, req.ny, req.nz, req.nc)
                self._bytes.setdefault(key, req.ctx.device_bytes)
                req.admitted = now()
                active.append(req)
                admitted = True
            if not active:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival - now()))
                    continue
                if queue:
                    raise RuntimeError("a request does not fit in device memory")
                break
            if admitted or batch is None:
                if batch is not None:
                    batch.close()
                batch = F.Batch([r.ctx for r in active], tile=None, graph=self.graph)
                if self.tile is not None:
                    batch.set_tile(self.tile[0], self.tile[1], self.tile[3] if len(self.tile) > 3 else 2)
            k = min(self.chunk, min(r.nsteps - r.steps_done for r in active))
            t1 = time.perf_counter()
            batch.step(k)
            batch.sync()
            self.busy_s += time.perf_counter() - t1
            for r in active:
                r.steps_done += k
                self.updates += k * r.nx * r.ny * r.nz * r.nc
            fin = [r for r in active if r.steps_done >= r.nsteps]
            if fin:
                batch.close()
                batch = None
                tf = now()
                for r in fin:
                    r.cdry = r.ctx.ground_load()
                    r.finished = tf
                    r.ctx.close()
                    r.ctx = None
                    done.append(r)
                    if log:
                        log(f"request {r.rid} done: {r.nsteps} steps, latency {r.latency:.2f} s, "
                            f"{len(active) - len(fin)} active, {len(queue)} queued")
                active = [r for r in active if r.steps_done < r.nsteps]
        if batch is not None:
            batch.close()
        self.wall_s = now()
        return done


def latency_stats(done):
    lat = np.array([r.latency for r in done])
    return dict(n=int(lat.size), p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                mean=float(lat.mean()), max=float(lat.max()))
