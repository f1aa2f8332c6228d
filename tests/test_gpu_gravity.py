"""The gravity current on the device (addradialwind.f90, setdivu.f90,
divcorr.f90), beyond the composed golden runs that test_gpu_parity.py and
test_gpu_dropin_link.py replay (run_12x10x8x3_gravity*, written by the
reference's own routines):

* the radial wind's scalars of a meteo read (radius, levels, crater node)
  equal the oracle's, and so does the last divcorr factor;
* the FP32 context follows the oracle over a run in which divcorr fires
  at every log time, unclamped and clamped;
* turning the gravity current off restores the plain meteo stage;
* a batch refuses a context with the gravity current (divcorr belongs to
  f3d_step's end of step).
"""
import dataclasses
import os

import numpy as np
import pytest

from fall3d_amd import lib as F
from fall3d_amd.metrics import rel_l1
from oracle.oracle import OracleRun, log_time
from tests.golden_util import GOLDEN, load, scenario_from

pytestmark = pytest.mark.gpu

FIXTURES = ["run_12x10x8x3_gravity", "run_12x10x8x3_gravity_clamped", "run_12x10x8x3_gravity_lonlat"]


def _sc(name):
    return scenario_from(load(os.path.join(GOLDEN, name + ".npz")))


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_radial_wind_scalars_and_divcorr_factor_vs_oracle(name, prec):
    sc = _sc(name)
    o = OracleRun(sc)
    o.meteo_update()
    r = F.Run(sc, precision=F.FP64 if prec == "fp64" else F.FP32)
    try:
        r.meteo_update()
        diag, _, active = r.ctx.gravity_state()
        assert active and o.gactive == 1
        assert np.array_equal(diag, o.gdiag), (diag, o.gdiag)
        assert r.dt == o.dt
        n = 0
        while n < 15:
            o.step()
            r.step(1)
            n += 1
            if log_time(o.iiter, o.time, o.dt):
                _, cd, _ = r.ctx.gravity_state()
                # the budgets behind the factor are tree sums on the device
                tol = 1e-12 if prec == "fp64" else 1e-6
                assert abs(cd - o.corrdiv) <= tol, (n, cd, o.corrdiv)
        c, cdry = r.state()
        if prec == "fp64":
            assert np.abs(c - o.c).max() <= 1e-11 * np.abs(o.c).max()
        else:
            inner = (slice(1, -1),) * 3
            assert rel_l1(c[inner], o.c[inner]) <= 1e-4
            assert rel_l1(cdry[:, :, 1:].sum(-1), o.cdry[:, :, 1:].sum(-1)) <= 1e-4
    finally:
        r.close()


def test_divcorr_fires_unclamped_and_clamped():
    # the fixtures cover both branches of divcorr.f90:33
    seen = {}
    for name in FIXTURES[:2]:
        o = OracleRun(_sc(name))
        o.meteo_update()
        cds = []
        for _ in range(15):
            o.step()
            if log_time(o.iiter, o.time, o.dt):
                cds.append(o.corrdiv)
        seen[name] = cds
    assert any(0.0 < abs(cd) < 0.05 for cd in seen[FIXTURES[0]]), seen
    assert any(abs(cd) == 0.05 for cd in seen[FIXTURES[1]]), seen


def test_gravity_off_restores_plain_meteo_stage():
    sc = _sc(FIXTURES[0])
    plain = dataclasses.replace(sc, gravity=None, zs=None)
    a, b = F.Run(sc, precision=F.FP64), F.Run(plain, precision=F.FP64)
    try:
        a.ctx.set_gravity_current(None)
        a.meteo_update()
        b.meteo_update()
        assert np.array_equal(a.ctx.coef()["vx"], b.ctx.coef()["vx"])
        assert a.dt == b.dt
        # the divergence on request, of the plain wind: the oracle's setdivu
        o = OracleRun(dataclasses.replace(sc, zs=sc.zs, gravity=dict(sc.gravity, c_flow_rate=0.0)))
        o.meteo_update()
        assert np.array_equal(b.ctx.divu(), o.divu)
    finally:
        a.close()
        b.close()


def test_batch_refuses_gravity_context():
    sc = _sc(FIXTURES[0])
    r = F.Run(sc, precision=F.FP32)
    try:
        r.meteo_update()
        with pytest.raises(F.F3DError, match="gravity current"):
            F.Batch([r.ctx])
    finally:
        r.close()
