"""Whole-horizon runs of the two single-run configurations.

The plain configuration (60x60x20, 6 classes 2 um - 2 mm, log-profile wind,
Suzuki column of 10 km, 6 h): the oracle (the FP64 C restatement) runs the
whole 6 h beside the FP64 and the FP32 context, and the three are compared
at every simulated hour (Fall3d.f90:79-121 loop: time += dt until time >=
end, the last step may overshoot):

* FP64 context: mass residual rho_M <= 1e-10, concentration and ground load
  <= 1e-11 relative L-infinity, every hour;
* FP32 context: concentration and ground load <= 1e-4 relative L1, every
  hour; rho_M <= 1e-5 (FP32 state drift).

The precision form (400x400x40, 20 classes, 12 h, 4084 steps): the FP64
context against the oracle after 200 steps (rho_M <= 1e-10, L-infinity
<= 1e-11), then the FP32 context against the FP64 context at every
simulated hour to 12 h: ground load relative L1 <= 1e-4 every hour,
concentration relative L1 <= 1e-4 at 6 h and 12 h.  The per-hour figures,
the FP32 rho_M among them, are printed.
"""
import os

import numpy as np
import pytest

from fall3d_amd import lib as F
from fall3d_amd import metrics as M
from fall3d_amd import scenario as S
from oracle.oracle import OracleRun

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1200)]


def _threads(nc):
    return max(1, min(len(os.sched_getaffinity(0)), nc, 16))


def _interior(c):
    return c[1:-1, 1:-1, 1:-1, :]


def _gpu(r):
    b = r.budget()
    c, cdry = r.state()
    return _interior(c), cdry[:, :, 1:].sum(-1), b


def _oracle(o):
    b = o.budget()
    return _interior(o.c), o.cdry[:, :, 1:].sum(-1), b


def _fmt(m):
    # a budget term below the reference's noise floor compares as None
    return ", ".join(f"{k} {'-' if v is None else format(v, '.2e')}" for k, v in m.items())


@pytest.fixture(scope="module")
def plain_run():
    sc = S.config_c1()
    o = OracleRun(sc, nthreads=_threads(sc.nc))
    o.meteo_update()
    r64, r32 = F.Run(sc, precision=F.FP64), F.Run(sc, precision=F.FP32)
    rows = []
    try:
        r64.meteo_update()
        r32.meteo_update()
        dts = (r64.dt, r32.dt, o.dt)
        nhours = int(round(sc.end_time / 3600.0))
        for h in range(1, nhours + 1):
            end = 3600.0 * h
            r64.run_to(end)
            r32.run_to(end)
            while o.time < end:
                o.step()
            oc, ol, ob = _oracle(o)
            c64, l64, b64 = _gpu(r64)
            c32, l32, b32 = _gpu(r32)
            row = dict(hour=h, iiter=(r64.ctx.clock()["iiter"], r32.ctx.clock()["iiter"],
                                      o.iiter),
                       fp64=M.compare(l64, ol, b64, ob, c64, oc),
                       fp32=M.compare(l32, ol, b32, ob, c32, oc),
                       load_max=float(ol.max()))
            print(f"[horizon] plain 60x60x20x6 hour {h} (step {o.iiter}): fp64 "
                  f"{_fmt(row['fp64'])} | fp32 {_fmt(row['fp32'])}", flush=True)
            rows.append(row)
    finally:
        r64.close()
        r32.close()
    return dict(rows=rows, dts=dts, sc=sc)


def test_plain_config_full_horizon_fp64_vs_oracle(plain_run):
    dt64, dt32, dto = plain_run["dts"]
    assert dt64 == dt32
    assert abs(dt64 - dto) <= 1e-12 * dto
    rows = plain_run["rows"]
    assert len(rows) == 6
    # the whole 6 h: about 645 steps of dt = 33.5 s
    assert rows[-1]["iiter"][2] * dto >= plain_run["sc"].end_time
    for row in rows:
        assert row["iiter"][0] == row["iiter"][1] == row["iiter"][2], row
        m = row["fp64"]
        assert row["load_max"] > 0
        assert m["rho_M"] <= 1e-10, (row["hour"], m)
        assert m["c_rel_linf"] <= 1e-11 and m["load_rel_linf"] <= 1e-11, (row["hour"], m)


def test_plain_config_full_horizon_fp32_vs_oracle(plain_run):
    for row in plain_run["rows"]:
        m = row["fp32"]
        assert m["c_rel_l1"] <= 1e-4 and m["load_rel_l1"] <= 1e-4, (row["hour"], m)
        assert m["rho_M"] <= 1e-5, (row["hour"], m)


ORACLE_STEPS = 200


@pytest.fixture(scope="module")
def precision_run():
    sc = S.config_c2()
    r64, r32 = F.Run(sc, precision=F.FP64), F.Run(sc, precision=F.FP32)
    out = dict(hours=[])
    try:
        r64.meteo_update()
        r32.meteo_update()
        assert r64.dt == r32.dt
        r64.step(ORACLE_STEPS)
        r32.step(ORACLE_STEPS)
        g64, g32 = _gpu(r64), _gpu(r32)
        o = OracleRun(sc, nthreads=_threads(sc.nc))
        o.meteo_update()
        print(f"[horizon] precision form: oracle meteo stage done, dt {o.dt:.6f} s", flush=True)
        for k in range(ORACLE_STEPS):
            o.step()
            if (k + 1) % 50 == 0:
                print(f"[horizon] oracle step {k + 1}", flush=True)
        oc, ol, ob = _oracle(o)
        out["oracle"] = dict(fp64=M.compare(g64[1], ol, g64[2], ob, g64[0], oc),
                             fp32=M.compare(g32[1], ol, g32[2], ob, g32[0], oc),
                             dt=(r64.dt, o.dt), load_max=float(ol.max()))
        print(f"[horizon] precision form, {ORACLE_STEPS} steps vs oracle: fp64 "
              f"{_fmt(out['oracle']['fp64'])} | fp32 {_fmt(out['oracle']['fp32'])}", flush=True)
        del o, oc, ol, g64, g32
        nhours = int(round(sc.end_time / 3600.0))
        for h in range(1, nhours + 1):
            end = 3600.0 * h
            r64.run_to(end)
            r32.run_to(end)
            full = h in (6, nhours)
            if full:
                c64, l64, b64 = _gpu(r64)
                c32, l32, b32 = _gpu(r32)
                m = M.compare(l32, l64, b32, b64, c32, c64)
                del c64, c32
            else:
                b64, b32 = r64.budget(), r32.budget()
                l64 = r64.ctx.ground_load()[:, :, 1:].sum(-1)
                l32 = r32.ctx.ground_load()[:, :, 1:].sum(-1)
                m = M.compare(l32, l64, b32, b64)
            it = (r64.ctx.clock()["iiter"], r32.ctx.clock()["iiter"])
            out["hours"].append(dict(hour=h, iiter=it, fp32_vs_fp64=m,
                                     load_max=float(l64.max())))
            print(f"[horizon] precision form hour {h} (step {it[0]}): fp32 vs fp64 {_fmt(m)}",
                  flush=True)
        c64, _ = r64.state()
        out["final_fp64_ok"] = bool(np.isfinite(c64).all() and (c64 >= 0).all())
        del c64
    finally:
        r64.close()
        r32.close()
    return out


def test_precision_form_fp64_vs_oracle_over_hundreds_of_steps(precision_run):
    o = precision_run["oracle"]
    dt64, dto = o["dt"]
    assert abs(dt64 - dto) <= 1e-12 * dto
    assert o["load_max"] > 0
    m = o["fp64"]
    assert m["rho_M"] <= 1e-10, m
    assert m["c_rel_linf"] <= 1e-11 and m["load_rel_linf"] <= 1e-11, m
    m = o["fp32"]
    assert m["c_rel_l1"] <= 1e-4 and m["load_rel_l1"] <= 1e-4, m


def test_precision_form_fp32_over_12_hours(precision_run):
    hours = precision_run["hours"]
    assert len(hours) == 12
    # 12 h at dt = 10.58 s: 4084 steps
    assert hours[-1]["iiter"][0] == hours[-1]["iiter"][1] >= 4000
    for row in hours:
        m = row["fp32_vs_fp64"]
        assert row["load_max"] > 0
        assert m["load_rel_l1"] <= 1e-4, (row["hour"], m)
        if "c_rel_l1" in m:
            assert m["c_rel_l1"] <= 1e-4, (row["hour"], m)
        assert m["rho_M"] <= 1e-5, (row["hour"], m)
    assert precision_run["final_fp64_ok"]
