"""The CPU legs of bench.py compute the workload they claim to time.

The one-core reference leg steps the reference's own routines (oracle/_ref)
on coefficients and dt injected from a meteo stage instead of running its
own; its state after the timed steps must equal the oracle's, bit for bit,
on the same coefficients.
"""
import numpy as np
import pytest

from fall3d_amd import scenario as S
from oracle.oracle import OracleRun, RefRun


@pytest.mark.skipif(not RefRun.available(), reason="reference build (oracle/_ref) not present")
def test_reference_leg_on_injected_coefficients_equals_oracle(monkeypatch):
    import bench

    sc = S.config_c2(nx=48, ny=40, nz=8, nc=3, seed=17)
    o = OracleRun(sc)
    o.meteo_update()
    coef = dict(vx=o.vx, rho=o.rho, rkhor=o.rkhor, rkh1=o.rkh1, rkver=o.rkver, vset=o.vset)
    seen = {}
    step = RefRun.step

    def spy(self):
        step(self)
        seen["run"] = self

    monkeypatch.setattr(RefRun, "step", spy)
    res = bench.ref_cpu_baseline(sc, 0.2, coef, o.dt)
    n = int(res["sample"].split()[0])
    assert res["value"] > 0 and res["cores"] == 1 and res["kind"] == "reference"
    for _ in range(n):
        o.step()
    c, cdry = seen["run"].state()
    assert o.c.max() > 0
    assert np.array_equal(c, o.c)
    assert np.array_equal(cdry[:, :, 1:], o.cdry[:, :, 1:])
