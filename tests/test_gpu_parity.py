"""GPU parity of libfall3d_amd against the reference (golden vectors) and the
oracle, through the C ABI.

Tolerances:
* per-routine FP64 drop-ins: bit-exact (same operation order, no
  contraction, IEEE division);
* FP64 context: coefficients without transcendental calls bit-exact;
  settling velocity (device pow) and everything downstream within 1e-11
  relative L-infinity;
* FP32 fused context: relative L1 of concentration and ground load
  <= 1e-4, mass budget terms <= 1e-4 relative.
"""
import ctypes
import dataclasses
import os

import numpy as np
import pytest

from fall3d_amd import lib as F
from fall3d_amd import scenario as S
from oracle.oracle import OracleRun
from tests.golden_util import GOLDEN, load, run_fixtures, scenario_from

pytestmark = pytest.mark.gpu


def rel_linf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    assert np.array_equal(nan_a, nan_b), "NaN pattern differs"
    m = ~nan_a
    den = np.max(np.abs(b[m])) if m.any() else 0.0
    return 0.0 if den == 0 else float(np.max(np.abs(a[m] - b[m])) / den)


def rel_l1(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.sum(np.abs(b))
    return 0.0 if den == 0 else float(np.sum(np.abs(a - b)) / den)


def interior(c):
    return c[1:-1, 1:-1, 1:-1, :]


# ---------------- per-routine drop-ins: bit-exact ----------------

def test_dropins_bitexact_vs_reference_routines():
    g = load(os.path.join(GOLDEN, "routines_9x7x6.npz"))
    L = F.lib()
    dp = ctypes.POINTER(ctypes.c_double)

    def P(a):
        return a.ctypes.data_as(dp)

    def i(x):
        return ctypes.byref(ctypes.c_int(x))

    def d(x):
        return ctypes.byref(ctypes.c_double(x))

    A = {k: np.asfortranarray(g[k]) for k in g}
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    dx, dy, dt = float(g["dx"]), float(g["dy"]), float(g["dt"])
    args = (i(nx), i(ny), i(nz), i(nz), i(1))
    c = A["c0"].copy(order="F")
    L.setbcc_(P(c), *args, P(A["vx"]), P(A["vy"]), P(A["vz"]), P(A["vset"]), P(A["vdep"]),
              P(A["rkver"]), P(A["dz"]))
    assert np.array_equal(c, g["setbcc"]), F.last_error()
    cb = A["setbcc"]
    work = np.zeros(64)
    calls = {
        "advctzc": lambda c: L.advctzc_(P(c), *args, P(A["vz"]), P(A["vset"]), P(A["dz"]), d(dt), i(1)),
        "advctx": lambda c: L.advctx_(P(c), *args, P(A["vx"]), d(dx), d(dt), P(work), i(1)),
        "advcty": lambda c: L.advcty_(P(c), *args, P(A["vy"]), d(dy), d(dt), P(work), i(1)),
        "diffz": lambda c: L.diffz_(P(c), *args, P(A["dz"]), d(dt), P(A["rkver"]), P(A["rho"]), P(work)),
        "diffx": lambda c: L.diffx_(P(c), *args, d(dx), d(dt), P(A["rkh"]), P(A["rho"]), P(work)),
        "diffy": lambda c: L.diffy_(P(c), *args, d(dy), d(dt), P(A["rkh"]), P(A["rho"]), P(work)),
    }
    for name, call in calls.items():
        c = cb.copy(order="F")
        call(c)
        assert np.array_equal(c, g[name]), (name, F.last_error(), np.abs(c - g[name]).max())
    cdry = A["cdry0"].copy(order="F")
    L.accum_(P(cb), P(cdry), i(2), i(nx), i(ny), i(nz), i(nz), i(1), P(A["vz"]), d(dt),
             P(A["vset"]), P(A["vdep"]))
    assert np.array_equal(cdry, g["accum"])
    om = ctypes.c_double(1.25)
    L.cmass2d_(P(A["c2"]), P(A["vx"]), P(A["vy"]), P(A["vz"]), P(A["vset2"]), i(nx), i(ny), i(nz),
               i(2), i(1), i(2), i(1), i(nz), d(dx), d(dy), P(A["dz"]), d(dt), ctypes.byref(om))
    # tree summation order: within a few ulps of the serial sum
    assert abs(om.value - float(g["cmass2d_outmass"])) <= 1e-13 * abs(float(g["cmass2d_outmass"]))
    outs = [ctypes.c_double() for _ in range(4)]
    L.cmass3d_(P(A["c2"]), P(A["cload"]), P(A["cdr"]), P(A["cwt"]), i(nx), i(ny), i(nz), i(2),
               d(dx), d(dy), P(A["dz"]), *[ctypes.byref(o) for o in outs])
    np.testing.assert_allclose([o.value for o in outs], g["cmass3d"], rtol=1e-13)


# ---------------- FP64 context vs the reference's golden runs ----------------

@pytest.mark.parametrize("path", run_fixtures(), ids=os.path.basename)
def test_fp64_context_vs_golden(path):
    g = load(path)
    sc = scenario_from(g)
    r = F.Run(sc, precision=F.FP64)
    try:
        r.meteo_update()
        co = r.ctx.coef()
        for name in ("vx", "rho", "rkhor", "rkh1", "rkver"):
            assert np.array_equal(co[name], g[name], equal_nan=True), name
        assert rel_linf(co["vset"], g["vset"]) <= 1e-12
        assert abs(r.dt - float(g["dt"])) <= 1e-12 * float(g["dt"])
        if "gravity" in g:  # the radial wind (vx above) and setdivu on the device
            assert np.array_equal(r.ctx.divu(), g["divu"])
        r.step(1)
        c, cdry = r.state()
        assert rel_linf(c, g["c_step1"]) <= 1e-11
        assert rel_linf(cdry, g["cdry_step1"]) <= 1e-11
        r.step(int(g["nsteps"]) - 1)
        c, cdry = r.state()
        assert rel_linf(c, g["c_final"]) <= 1e-11
        assert rel_linf(cdry, g["cdry_final"]) <= 1e-11
        b = r.budget()
        for k in ("massvol", "massdep", "masswet", "outmass", "erumass"):
            ref = float(g["budget_" + k])
            if np.isnan(ref):
                assert np.isnan(b[k])
            else:
                assert abs(b[k] - ref) <= 1e-11 * max(abs(ref), 1e-300), k
    finally:
        r.close()


# ---------------- FP32 fused context ----------------

# the SIMILARITY fixture holds the reference's NaN
FINITE = [p for p in run_fixtures() if "similarity" not in p]
WET = [p for p in run_fixtures() if "wetdep" in p]


@pytest.mark.parametrize("path", WET, ids=os.path.basename)
def test_batch_refuses_wet_context(path):
    # the persistent batched kernel carries no wet deposition: a context
    # with precipitation steps through f3d_step only
    sc = scenario_from(load(path))
    r = F.Run(sc, precision=F.FP32)
    try:
        r.meteo_update()
        with pytest.raises(F.F3DError, match="wet deposition"):
            F.Batch([r.ctx])
    finally:
        r.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_output_products_vs_oracle(prec):
    # load, deposit thickness and z-cumulative concentration
    # (printres_nc.f90:150-241) from the device state against the oracle's
    # own run, with wet deposition on, so cwet enters the load
    path = WET[0]
    sc = scenario_from(load(path))
    o = OracleRun(sc)
    o.meteo_update()
    for _ in range(6):
        o.step()
    r = F.Run(sc, precision=F.FP64 if prec == "fp64" else F.FP32)
    try:
        r.meteo_update()
        r.step(6)
        got = r.ctx.products()
        ref = o.products()
        for name, a, b in zip(("load", "thick", "colmass"), got, ref):
            assert np.all(np.isfinite(a)), name
            if prec == "fp64":
                assert rel_linf(a, b) <= 1e-11, name
            else:
                assert rel_l1(a, b) <= 1e-4, name
        assert np.any(o.cwet[:, :, 1:] > 0)
    finally:
        r.close()


def test_wetdep_dropin_vs_oracle():
    # the wetdep_ drop-in (Sources/wetdep.f90 argument list) against the
    # oracle's restatement: rain over part of the domain, PBL heights that
    # cut the column, a class above 100 um left alone; device pow: 1e-13
    from oracle.oracle import lib as olib
    from oracle.oracle import P as OP
    rng = np.random.default_rng(5)
    nx, ny, nz = 13, 9, 6
    Fa = lambda a: np.asfortranarray(a, dtype=np.float64)  # noqa: E731
    c0 = Fa(rng.uniform(0.0, 1.0, (nx + 2, ny + 2, nz + 2)))
    cwet0 = Fa(rng.uniform(0.0, 0.1, (nx, ny, 3)))
    zl = Fa(np.arange(nz) * 700.0)
    dz = Fa(np.full(nz + 1, 700.0))
    zi = Fa(rng.uniform(300.0, 4000.0, (nx, ny)))
    prate = Fa(np.where(rng.random((nx, ny)) < 0.6, rng.uniform(0.1, 40.0, (nx, ny)), 0.0))
    diam = Fa([5e-6, 2e-4])
    L = F.lib()
    ip = lambda v: ctypes.byref(ctypes.c_int(v))  # noqa: E731
    for ic in (1, 2):
        c, cw = c0.copy(order="F"), cwet0.copy(order="F")
        L.wetdep_(F._P(c), F._P(cw), ip(ic), ip(nx), ip(ny), ip(nz), ip(nz), ip(1),
                  ctypes.byref(ctypes.c_double(90.0)), F._P(diam), F._P(zl), F._P(dz),
                  F._P(zi), F._P(prate))
        assert F.last_error() in ("", None) or "wetdep" not in F.last_error()
        co, cwo = c0.copy(order="F"), cwet0.copy(order="F")
        olib().or_wetdep(OP(co), OP(cwo), ic, nx, ny, nz, 90.0, OP(diam), OP(zl), OP(dz),
                         OP(zi), OP(prate))
        assert np.all(np.abs(c - co) <= 1e-13 * np.abs(co)), ic
        assert np.all(np.abs(cw - cwo) <= 1e-13 * np.abs(cwo)), ic
        if ic == 2:
            assert np.array_equal(c, c0) and np.array_equal(cw, cwo)
        else:
            assert not np.array_equal(c, c0)


@pytest.mark.parametrize("path", FINITE, ids=os.path.basename)
def test_fp32_context_vs_golden(path):
    g = load(path)
    sc = scenario_from(g)
    r = F.Run(sc, precision=F.FP32)
    try:
        r.meteo_update()
        r.step(int(g["nsteps"]))
        c, cdry = r.state()
        ref_c, ref_d = g["c_final"], g["cdry_final"]
        e_c = rel_l1(interior(c), interior(ref_c))
        e_load = rel_l1(cdry[:, :, 1:].sum(-1), ref_d[:, :, 1:].sum(-1))
        assert e_c <= 1e-4, e_c
        assert e_load <= 1e-4, e_load
        b = r.budget()
        for k in ("massvol", "massdep", "masswet", "outmass"):
            ref = float(g["budget_" + k])
            assert abs(b[k] - ref) <= 1e-4 * abs(ref) + 1e-12 * float(g["budget_erumass"]), k
    finally:
        r.close()


TILES = [(58, 10), (58, 18), (58, 26), (58, 10, 0, 2), (58, 18, 0, 2), (58, 26, 0, 2),
         (58, 18, 8, 1), (58, 18, 8, 2)]


def _tile_id(t):
    return (f"{t[0]}x{t[1]}" + (f"w{t[2]}" if len(t) > 3 and t[2] else "")
            + (f"m{t[3]}" if len(t) > 3 else ""))


def _compare_fp32_with_oracle(sc, nsteps, tile=None):
    o = OracleRun(sc)
    o.meteo_update()
    r = F.Run(sc, precision=F.FP32, tile=tile)
    try:
        r.meteo_update()
        assert abs(r.dt - o.dt) <= 1e-12 * o.dt
        for _ in range(nsteps):
            o.step()
        r.step(nsteps)
        c, cdry = r.state()
        bo, br = o.budget(), r.budget()
        load_ref = o.cdry[:, :, 0]
        return dict(c_l1=rel_l1(interior(c), interior(o.c)),
                    load_l1=rel_l1(cdry[:, :, 1:].sum(-1), load_ref),
                    rho_m=abs((br["massvol"] + br["massdep"] + br["outmass"]) -
                              (bo["massvol"] + bo["massdep"] + bo["outmass"])) / bo["erumass"])
    finally:
        r.close()


@pytest.mark.parametrize("parity_steps", [7, 30])
def test_fp32_vs_oracle_moderate_grid(parity_steps):
    sc = S.config_small(nx=45, ny=38, nz=12, nc=5, seed=21, mean_u=7.0, mean_v=-3.0)
    m = _compare_fp32_with_oracle(sc, parity_steps)
    assert m["c_l1"] <= 1e-4, m
    assert m["load_l1"] <= 1e-4, m
    assert m["rho_m"] <= 1e-5, m


def _lonlat_band(sc):
    # LON-LAT rows between colatitudes 35 and 50 degrees: Hm1 = sin(colat)
    # per row (reagrd.f90:247-270), so the scaling of meteo.f90:108-112
    # (vx/Hm1, rho*Hm1, rkh1 = rkhor/Hm1^2) and the setsrc volume differ
    # from the UTM case
    colat = np.radians(np.linspace(35.0, 50.0, sc.ny))
    h = np.repeat(np.sin(colat)[None, :], sc.nx, axis=0)
    return dataclasses.replace(sc, hm1_cols=np.asfortranarray(h))


def test_fp32_vs_oracle_lonlat_hm1():
    sc = _lonlat_band(S.config_small(nx=130, ny=47, nz=9, nc=3, seed=33, mean_u=4.0, mean_v=-3.0))
    m = _compare_fp32_with_oracle(sc, 10)
    assert m["c_l1"] <= 1e-4 and m["load_l1"] <= 1e-4, m
    assert m["rho_m"] <= 1e-5, m


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lonlat_terrain_corvver_vs_oracle(prec):
    # a LON-LAT run over terrain: Hm1 per row, the CMAQ constant of LON-LAT
    # grids and corvver's terrain-slope correction of vz (meteo.f90:116-124)
    sc = _lonlat_band(S.config_small(nx=66, ny=29, nz=8, nc=3, seed=51, mean_u=6.0, mean_v=3.0,
                                     modkh=S.KH_CMAQ))
    x = np.arange(sc.nx)[:, None] / sc.nx
    y = np.arange(sc.ny)[None, :] / sc.ny
    topg = np.asfortranarray(1200.0 * (1.0 + np.sin(2 * np.pi * x) * np.cos(3 * np.pi * y)))
    sc = dataclasses.replace(sc, lonlat=1, topg=topg)
    o = OracleRun(sc)
    o.meteo_update()
    r = F.Run(sc, precision=F.FP64 if prec == "fp64" else F.FP32)
    try:
        r.meteo_update()
        assert abs(r.dt - o.dt) <= 1e-11 * o.dt
        for _ in range(8):
            o.step()
        r.step(8)
        c, cdry = r.state()
        if prec == "fp64":
            assert rel_linf(interior(c), interior(o.c)) <= 1e-11
            assert rel_linf(cdry[:, :, 1:], o.cdry[:, :, 1:]) <= 1e-11
        else:
            assert rel_l1(interior(c), interior(o.c)) <= 1e-4
            assert rel_l1(cdry[:, :, 1:].sum(-1), o.cdry[:, :, 1:].sum(-1)) <= 1e-4
    finally:
        r.close()


def test_fp64_vs_oracle_lonlat_hm1():
    sc = _lonlat_band(S.config_small(nx=21, ny=17, nz=7, nc=3, seed=34, mean_u=-5.0, mean_v=2.0))
    o = OracleRun(sc)
    o.meteo_update()
    r = F.Run(sc, precision=F.FP64)
    try:
        r.meteo_update()
        assert abs(r.dt - o.dt) <= 1e-11 * o.dt
        for _ in range(8):
            o.step()
        r.step(8)
        c, cdry = r.state()
        assert rel_linf(interior(c), interior(o.c)) <= 1e-11
        assert rel_linf(cdry[:, :, 1:], o.cdry[:, :, 1:]) <= 1e-11
    finally:
        r.close()


def test_setvdrydep_vs_reference():
    # setvdrydep on the device (meteo.f90:99-100) against the reference's
    # own vectors: every land-use category, stable / neutral / unstable /
    # calm columns, the four Feng modes, a class above 100 um (untouched),
    # three first-level heights.  Device log/pow/atan/exp: 1e-13 relative.
    d = np.load(os.path.join(GOLDEN, "drydep_12x9x6.npz"))
    nx, ny, nc, nz = int(d["nx"]), int(d["ny"]), int(d["nc"]), 4
    tempe = np.repeat(d["tempe1"][:, :, None], nz, axis=2)
    rho = np.repeat(d["rho1"][:, :, None], nz, axis=2)
    zero3 = np.zeros((nx, ny, nz))
    for n, z1 in enumerate(d["zlayer1"]):
        zl = z1 + np.arange(nz) * 500.0
        ctx = F.Context(nx, ny, nz, nc, 3000.0, 3000.0, zl, precision=F.FP64)
        try:
            ctx.set_classes(d["diam"], np.full(nc, 2000.0), np.full(nc, 0.9), S.VS_GANSER)
            ctx.set_models(S.KV_CONSTANT, S.KH_CONSTANT, 100.0, 1000.0, 0.2275)
            ctx.set_landuse(d["ldu"], d["znt"])
            ctx.set_meteo(zero3, zero3, zero3, tempe, tempe, rho, d["ustar"],
                          np.zeros((nx, ny)), d["rl"], d["vdep0"])
            got, ref = ctx.vdep(), d[f"vdep_{n}"]
            assert np.all(np.abs(got - ref) <= 1e-13 * np.abs(ref)), n
            assert (got[:, :, 4] == 7.0).all()
        finally:
            ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_run_with_device_drydep_vs_oracle(prec):
    # a run whose vdep comes from setvdrydep on the device, against the
    # oracle run given the oracle's own setvdrydep: first level 25 m above
    # the ground (vdep nonzero), mixed land use, unstable and stable columns
    from oracle.oracle import lib as olib
    from oracle.oracle import P as OP
    sc = S.config_small(nx=70, ny=31, nz=8, nc=4, seed=45, mean_u=5.0, mean_v=-2.0)
    sc.zlayer[0] = 25.0
    rng = np.random.default_rng(46)
    sc.rl[:] = rng.uniform(-0.02, 0.02, sc.rl.shape)
    sc.ustar[:] = rng.uniform(0.05, 0.6, sc.ustar.shape)
    ldu = np.asfortranarray(rng.integers(1, 25, (sc.nx, sc.ny)).astype(np.float64))
    znt = np.asfortranarray(rng.uniform(0.01, 1.0, (sc.nx, sc.ny)))
    vdep = np.zeros((sc.nx, sc.ny, sc.nc), order="F")
    Fa = lambda a: np.asfortranarray(a, dtype=np.float64)  # noqa: E731
    olib().or_setvdrydep(sc.nx, sc.ny, sc.nc, float(sc.zlayer[0]), OP(Fa(sc.diam)), OP(ldu),
                         OP(znt), OP(Fa(sc.rl)), OP(Fa(sc.ustar)), OP(Fa(sc.tempe[:, :, 0])),
                         OP(Fa(sc.rho[:, :, 0])), OP(vdep))
    assert np.any(vdep[:, :, :] < 0.0)
    o = OracleRun(dataclasses.replace(sc, vdep=vdep))
    o.meteo_update()
    for _ in range(9):
        o.step()
    r = F.Run(sc, precision=F.FP64 if prec == "fp64" else F.FP32)
    try:
        r.ctx.set_landuse(ldu, znt)
        r.meteo_update()
        assert np.all(np.abs(r.ctx.vdep() - vdep) <= 1e-13 * np.abs(vdep))
        r.step(9)
        c, cdry = r.state()
        if prec == "fp64":
            assert rel_linf(interior(c), interior(o.c)) <= 1e-11
            assert rel_linf(cdry[:, :, 1:], o.cdry[:, :, 1:]) <= 1e-11
        else:
            assert rel_l1(interior(c), interior(o.c)) <= 1e-4
            assert rel_l1(cdry[:, :, 1:].sum(-1), o.cdry[:, :, 1:].sum(-1)) <= 1e-4
    finally:
        r.close()


@pytest.mark.parametrize("tile", TILES, ids=_tile_id)
def test_fp32_every_tile_shape(tile):
    # a grid larger than every tile in both directions (interior and edge
    # tiles), both sweep orders (11 steps), winds of both signs
    sc = S.config_small(nx=150, ny=61, nz=10, nc=3, seed=27, mean_u=-3.0, mean_v=2.0)
    m = _compare_fp32_with_oracle(sc, 11, tile)
    assert m["c_l1"] <= 1e-4 and m["load_l1"] <= 1e-4, m


@pytest.mark.parametrize("tile", TILES, ids=_tile_id)
def test_fp32_tile_edges_partial_tiles(tile):
    # grid sizes that are not multiples of any tile edge, nz at the small end
    sc = S.config_small(nx=37, ny=11, nz=2, nc=2, seed=5, mean_u=-4.0, mean_v=6.0)
    m = _compare_fp32_with_oracle(sc, 12, tile)
    assert m["c_l1"] <= 1e-4 and m["load_l1"] <= 1e-4, m


@pytest.mark.parametrize("shape", [(150, 20), (60, 61), (128, 64)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_fp32_tile_kinds_one_axis_without_interior_tiles(shape):
    # 58x18 tiles: 150x20 has interior columns but no interior rows (y-side
    # and corner tiles only), 60x61 the reverse (x-side and corner tiles),
    # 128x64 one tile column and two tile rows of interior tiles
    nx, ny = shape
    sc = S.config_small(nx=nx, ny=ny, nz=7, nc=3, seed=41, mean_u=6.0, mean_v=-5.0)
    m = _compare_fp32_with_oracle(sc, 9, (58, 18, 0, 2))
    assert m["c_l1"] <= 1e-4 and m["load_l1"] <= 1e-4, m


@pytest.mark.parametrize("shape", [(61, 20), (62, 21), (63, 19), (120, 39), (121, 38), (179, 22)],
                         ids=lambda s: f"{s[0]}x{s[1]}")
def test_fp32_every_tile_shape_gives_the_same_bits(shape):
    # sizes around the tiling's steps: the first and last tiles own their
    # halo (t tiles cover 58t+4 columns and (RY-6)t+4 rows), so 62 columns
    # fit one tile and 63 need two, 120 two and 121 three; every cell takes
    # the same operands whichever tile, lane and row compute it, so every
    # tile shape, slab count and wave count gives the same bits
    nx, ny = shape
    sc = S.config_small(nx=nx, ny=ny, nz=5, nc=3, seed=nx + ny, mean_u=-4.0, mean_v=3.0)
    outs = {}
    for tile in TILES:
        r = F.Run(sc, precision=F.FP32, tile=tile)
        try:
            r.meteo_update()
            r.step(7)
            c, cdry = r.state()
            outs[_tile_id(tile)] = (c, cdry, r.budget())
        finally:
            r.close()
    ref_id = _tile_id(TILES[0])
    c0, d0, b0 = outs[ref_id]
    assert np.all(np.isfinite(c0)) and c0.max() > 0
    for k, (c, d, b) in outs.items():
        assert np.array_equal(c, c0), (k, np.argwhere(c != c0)[:5])
        assert np.array_equal(d[:, :, 1:], d0[:, :, 1:]), k
        for key in ("massvol", "massdep", "outmass"):
            assert b[key] == b0[key], (k, key)


@pytest.mark.parametrize("tile", [(58, 18, 0, 2), (58, 10, 0, 2), (58, 26, 0, 1), (58, 18, 8, 2)],
                         ids=_tile_id)
def test_fp32_deterministic(tile):
    # grids where every tile is an edge tile in x and y, many steps: repeated
    # runs are bit-identical (regression test for a cross-phase LDS hazard
    # on the y-edge rows)
    sc = S.config_small(nx=41, ny=37, nz=9, nc=3, seed=9, mean_u=5.0, mean_v=-4.0)
    outs = []
    for _ in range(3):
        r = F.Run(sc, precision=F.FP32, tile=tile)
        r.meteo_update()
        r.step(40)
        c, cdry = r.state()
        outs.append((c, cdry, r.budget()))
        r.close()
    for o in outs[1:]:
        assert np.array_equal(outs[0][0], o[0])
        assert np.array_equal(outs[0][1], o[1])
        assert outs[0][2] == o[2]


def test_settling_chain_matches_serial_oracle():
    """The warm-started chunked chain reproduces the serial SAVEd-cd chain."""
    sc = S.config_small(nx=40, ny=30, nz=20, nc=6, seed=3)
    o = OracleRun(sc)
    o.meteo_update()
    r = F.Run(sc, precision=F.FP64)
    try:
        r.meteo_update()
        vs = r.ctx.coef()["vset"]
        assert rel_linf(vs, o.vset) <= 1e-12
        assert abs(r.ctx.cd_state - o.cd.value) <= 1e-12 * abs(o.cd.value)
    finally:
        r.close()
