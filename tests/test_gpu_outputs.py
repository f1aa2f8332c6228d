"""The output side on the device, through the C ABI: the concentration
outputs (C_GRND, CON terrain following and a.s.l., the flight-level cuts)
of a stepped context against the oracle's restatement of printres_nc.f90,
setcut.f90 and ter2asl.f90 (itself bit-exact against the reference's
routines, tests/test_outputs.py) on the context's own state, bit for bit;
and the restart file written from one context and read into another
(wrirst.f90 / rearst.f90)."""
import os

import numpy as np
import pytest

from fall3d_amd import lib as F
from fall3d_amd import scenario as S
from oracle.oracle import output_concentration
from tests.golden_util import GOLDEN, load, scenario_from

pytestmark = pytest.mark.gpu

FL = [0.0, 50.0, 100.0, 150.0, 200.0, 250.0, 300.0, 350.0, 400.0, 700.0]


def _lonlat_scenario():
    # the LON-LAT golden run: Hm1 per row and terrain up to 1.5 km
    return scenario_from(load(os.path.join(GOLDEN, "run_12x9x6x3_lonlat_rams.npz")))


def _pressure(sc):
    # standard atmosphere at each level's height above sea level
    z = sc.zlayer[None, None, :] + sc.topg[:, :, None]
    return np.asfortranarray(S.standard_atm(z)[0])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_concentration_outputs_bitexact_vs_oracle(prec):
    sc = _lonlat_scenario()
    p = _pressure(sc)
    r = F.Run(sc, precision=F.FP64 if prec == "fp64" else F.FP32)
    try:
        r.meteo_update()
        r.step(12)
        r.ctx.set_pressure(p)
        cg, con, cfl = r.ctx.concentration(fl=FL)
        _, asl, _ = r.ctx.concentration(asl=True)
        c, _ = r.state()
    finally:
        r.close()
    hm1 = sc.hm1
    cg_o, con_o, cfl_o = output_concentration(c, hm1, sc.zlayer, sc.topg, p, FL)
    _, asl_o, _ = output_concentration(c, hm1, sc.zlayer, sc.topg, asl=True)
    assert np.array_equal(cg, cg_o)
    assert np.array_equal(con, con_o)
    assert np.array_equal(asl, asl_o)
    assert np.array_equal(cfl, cfl_o)
    assert cg.max() > 0 and np.any(asl == 0) and np.any(cfl[:, :, 1:-1] > 0)


def test_concentration_outputs_refuse_missing_fields():
    sc = S.config_small(seed=3)
    r = F.Run(sc, precision=F.FP32)
    try:
        r.meteo_update()
        with pytest.raises(F.F3DError, match="f3d_set_pressure"):
            r.ctx.concentration(fl=[100.0])
        with pytest.raises(F.F3DError, match="f3d_set_topography"):
            r.ctx.concentration(asl=True)
        with pytest.raises(F.F3DError, match="out of range"):
            r.ctx.concentration(member=1)
    finally:
        r.close()


def test_restart_carries_state_and_clock_between_contexts(tmp_path):
    sc = S.config_small(nx=14, ny=11, nz=7, nc=3, seed=41, mean_u=4.0, mean_v=-2.0)
    path = str(tmp_path / "run.rst.nc")
    a = F.Run(sc, precision=F.FP64)
    b = F.Run(sc, precision=F.FP64)
    try:
        a.meteo_update()
        b.meteo_update()
        a.step(9)
        clock_a = a.ctx.clock()
        ca, cdry_a = a.state()
        a.ctx.write_restart(path, year=2020, month=3, day=7, run_start=0, run_end=3600,
                            extent=(0.0, 0.0, 52000.0, 30000.0))
        cload, info = b.ctx.read_restart(path)
        cb, cdry_b = b.state()
        q = ca.astype(np.float32).astype(np.float64)
        assert np.array_equal(cb, q)
        assert not np.any(cdry_b)
        # LOAD: cdry + cwet with the class sums in slot 0, as 32-bit floats
        want = cdry_a.copy(order="F")
        want[:, :, 0] = cdry_a[:, :, 1:].sum(-1)
        assert np.allclose(cload, want.astype(np.float32), rtol=1e-7, atol=0)
        assert cload[:, :, 1:].max() > 0
        assert info["erumass"] == clock_a["erumass"] and info["outmass"] == clock_a["outmass"]
        assert (info["year"], info["month"], info["day"], info["run_end"]) == (2020, 3, 7, 3600)
        assert info["xmax"] == 52000.0
        clock_b = b.ctx.clock()
        assert clock_b["erumass"] == clock_a["erumass"]
        assert clock_b["outmass"] == clock_a["outmass"]
        # from equal states the two contexts step on identically
        a.ctx.set_state(q, cdry_a)
        b.ctx.set_clock(clock_a["time"], clock_a["iiter"], clock_a["erumass"], clock_a["outmass"])
        a.step(4)
        b.step(4)
        assert np.array_equal(a.state()[0], b.state()[0])
        # a context of another grid refuses the file
        other = F.Run(S.config_small(nx=13, ny=11, nz=7, nc=3, seed=41), precision=F.FP64)
        try:
            with pytest.raises(F.F3DError, match="restart inconsistency"):
                other.ctx.read_restart(path)
        finally:
            other.close()
    finally:
        a.close()
        b.close()


def test_restart_of_an_ensemble_member(tmp_path):
    base = S.config_c2(nx=30, ny=26, nz=8, nc=4, seed=0xF3D0011)
    mem = [S.config_ensemble_member(base, 700 + k) for k in range(3)]
    ens = F.Ensemble(base, mem)
    path = str(tmp_path / "m1.rst.nc")
    try:
        ens.meteo_update()
        ens.step(5)
        c1, _ = ens.state(1)
        ens.ctx.write_restart(path, member=1)
        er = ens.budgets()
    finally:
        ens.close()
    c, cload, info = F.restart_read(path, base.nx, base.ny, base.nz, base.nc)
    assert np.array_equal(c, c1.astype(np.float32).astype(np.float64))
    assert info["erumass"] == er[1]["erumass"] > 0
    assert info["outmass"] == er[1]["outmass"]
