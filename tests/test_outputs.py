"""Output side of the step on the CPU: the concentration outputs of the
oracle (C_GRND, CON terrain following and a.s.l., the flight-level cuts)
against the reference's own setcut_FL, standard_atm and ter2asl
(tests/golden/cuts_10x8x12x2.npz) and against properties the definitions
fix; the restart file of the library's host writer and reader against an
independent netCDF classic implementation (scipy.io.netcdf_file), in the
layout of wrirst.f90 / rearst.f90."""
import os

import numpy as np
import pytest

from oracle.oracle import output_concentration, standard_atm
from tests.golden_util import GOLDEN

CUTS = os.path.join(GOLDEN, "cuts_10x8x12x2.npz")


def _g():
    return dict(np.load(CUTS, allow_pickle=False))


def test_standard_atmosphere_bitexact_vs_reference():
    g = _g()
    for z, p, t in zip(g["zstd"], g["pstd"], g["tstd"]):
        assert standard_atm(z) == (p, t), z


def test_flight_level_cuts_bitexact_vs_reference():
    g = _g()
    _, _, cfl = output_concentration(g["c"], g["hm1"], g["zlayer"], g["topg"], g["p"], g["fl"])
    assert np.array_equal(cfl, g["cfl"])
    # the ground level lies below the first level's pressure almost
    # everywhere, FL700 above the model top everywhere: both cut to zero
    assert np.all(cfl[:, :, -1] == 0) and np.mean(cfl[:, :, 0] == 0) > 0.9
    assert np.all(cfl[:, :, 2:-1] > 0)


def test_con_asl_bitexact_vs_reference():
    g = _g()
    _, con, _ = output_concentration(g["c"], g["hm1"], g["zlayer"], g["topg"], asl=True)
    assert np.array_equal(con, g["con_asl"])
    assert np.any(con == 0) and np.any(con > 0)


def _state(seed, nx=7, ny=6, nz=9, nc=3):
    rng = np.random.default_rng(seed)
    c = np.asfortranarray(rng.uniform(0.0, 1.0, (nx + 2, ny + 2, nz + 2, nc)))
    hm1 = np.asfortranarray(rng.uniform(0.6, 1.0, (nx, ny)))
    zl = np.concatenate([[0.0], np.cumsum(rng.uniform(300.0, 900.0, nz - 1))])
    return c, hm1, zl


def test_concentrations_are_class_sums_over_hm1():
    c, hm1, zl = _state(3)
    cg, con, _ = output_concentration(c, hm1, zl)
    w = c[1:-1, 1:-1, 1:-1, :].sum(-1)
    assert np.allclose(cg, 1e3 * w[:, :, 0] / hm1, rtol=1e-15, atol=0)
    assert np.allclose(con, 1e3 * w / hm1[:, :, None], rtol=1e-15, atol=0)
    # the ghost frame never enters
    c2 = c.copy(order="F")
    c2[0], c2[-1], c2[:, 0], c2[:, -1], c2[:, :, 0], c2[:, :, -1] = 7.0, 7.0, 7.0, 7.0, 7.0, 7.0
    cg2, con2, _ = output_concentration(c2, hm1, zl)
    assert np.array_equal(cg2, cg) and np.array_equal(con2, con)


def test_asl_on_flat_terrain_is_the_identity():
    c, hm1, zl = _state(4)
    _, tf, _ = output_concentration(c, hm1, zl)
    _, asl, _ = output_concentration(c, hm1, zl, topg=np.zeros(hm1.shape), asl=True)
    assert np.array_equal(asl, tf)


def test_asl_shifted_by_one_level():
    # terrain exactly one (uniform) level high: the a.s.l. level k takes the
    # terrain-following level k-1, and the levels under the terrain are zero
    nx, ny, nz, nc = 5, 4, 8, 2
    rng = np.random.default_rng(5)
    c = np.asfortranarray(rng.uniform(0.0, 1.0, (nx + 2, ny + 2, nz + 2, nc)))
    hm1 = np.ones((nx, ny), order="F")
    zl = np.arange(nz) * 500.0
    _, tf, _ = output_concentration(c, hm1, zl)
    _, asl, _ = output_concentration(c, hm1, zl, topg=np.full((nx, ny), 500.0), asl=True)
    assert np.all(asl[:, :, 0] == 0)
    assert np.array_equal(asl[:, :, 1:], tf[:, :, :-1])


def test_flight_level_at_a_level_pressure_takes_that_level():
    # pressure of the standard atmosphere at the levels, flight levels whose
    # nominal pressure is that of level 3 (s = 0 there) and halfway between
    # two levels in pressure (the mean of the two)
    c, hm1, zl = _state(6)
    nx, ny, nz = hm1.shape[0], hm1.shape[1], zl.size
    pz = np.array([standard_atm(z)[0] for z in zl])
    p = np.asfortranarray(np.broadcast_to(pz, (nx, ny, nz)))
    fl3 = zl[2] / 30.48
    _, tf, cfl = output_concentration(c, hm1, zl, p=p, fl=[fl3])
    pn = standard_atm(fl3 * 30.48)[0]
    k = int(np.argmin(np.abs(pz - pn)))
    assert abs(pz[k] - pn) <= 1e-9 * pn
    assert np.allclose(cfl[:, :, 0], tf[:, :, k], rtol=1e-12, atol=0) or \
        np.allclose(cfl[:, :, 0], tf[:, :, k - 1], rtol=1e-12, atol=0)


# ---------------- restart file ----------------

def _lib_or_skip():
    from fall3d_amd import lib as F
    if not os.path.exists(F.LIB_PATH):
        pytest.skip("libfall3d_amd.so not built")
    return F


def _restart_state(seed, nx=6, ny=5, nz=4, nc=3):
    rng = np.random.default_rng(seed)
    c = np.asfortranarray(rng.uniform(0.0, 2.0, (nx + 2, ny + 2, nz + 2, nc)) ** 3)
    cload = np.asfortranarray(rng.uniform(0.0, 5.0, (nx, ny, nc + 1)))
    return c, cload


def test_restart_written_by_the_library_reads_as_netcdf(tmp_path):
    from scipy.io import netcdf_file
    F = _lib_or_skip()
    c, cload = _restart_state(7)
    nx, ny, nz, nc = 6, 5, 4, 3
    path = str(tmp_path / "run.rst.nc")
    F.restart_write(path, c, cload, lonlat=1, year=2011, month=6, day=4, run_start=0,
                    run_end=21600, xmin=-72.5, ymin=-41.0, xmax=-70.0, ymax=-39.5,
                    erumass=1.5e12, outmass=3.25e9)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.dimensions == {"nx": nx, "ny": ny, "nc": nc, "nc1": nc + 1, "nx2": nx + 2,
                                "ny2": ny + 2, "nz2": nz + 2}
        con, load = f.variables["CON"], f.variables["LOAD"]
        assert con.dimensions == ("nc", "nz2", "ny2", "nx2")
        assert load.dimensions == ("nc1", "ny", "nx")
        assert con.data.dtype == np.dtype(">f4") and load.data.dtype == np.dtype(">f4")
        # the file's C order is the transpose of the Fortran arrays
        assert np.array_equal(con.data.T, c.astype(np.float32))
        assert np.array_equal(load.data.T, cload.astype(np.float32))
        assert con.units == b"gr/m3" and load.units == b"kg/m2"
        a = f._attributes
        assert a["COORDINATES"] == b"LON-LAT"
        assert int(a["YEAR"]) == 2011 and int(a["RUN_END"]) == 21600
        assert float(a["ERUPTED_MASS"]) == 1.5e12 and float(a["OUT_MASS"]) == 3.25e9
        assert float(a["LONMIN"]) == -72.5 and float(a["LATMAX"]) == -39.5


def test_restart_written_by_netcdf_reads_in_the_library(tmp_path):
    from scipy.io import netcdf_file
    F = _lib_or_skip()
    c, cload = _restart_state(8)
    nx, ny, nz, nc = 6, 5, 4, 3
    path = str(tmp_path / "ref.rst.nc")
    with netcdf_file(path, "w", version=2) as f:
        for name, n in (("nx", nx), ("ny", ny), ("nc", nc), ("nc1", nc + 1), ("nx2", nx + 2),
                        ("ny2", ny + 2), ("nz2", nz + 2)):
            f.createDimension(name, n)
        f.TITLE = "Fall3d 6.2 results"
        f.COORDINATES = "UTM"
        f.RUN_START = np.int32(3600)
        f.RUN_END = np.int32(7200)
        # real(rp) attributes are NC_DOUBLE; a float one is converted on read
        f.ERUPTED_MASS = np.float64(2.0e11)
        f.OUT_MASS = np.float32(1.0e8)
        f.XMIN, f.YMIN = np.float64(1.0e5), np.float64(2.0e5)
        f.XMAX, f.YMAX = np.float64(3.0e5), np.float64(4.0e5)
        v = f.createVariable("CON", "f4", ("nc", "nz2", "ny2", "nx2"))
        v[:] = c.T.astype(np.float32)
        v = f.createVariable("LOAD", "f4", ("nc1", "ny", "nx"))
        v[:] = cload.T.astype(np.float32)
    c2, cl2, info = F.restart_read(path, nx, ny, nz, nc)
    assert np.array_equal(c2, c.astype(np.float32).astype(np.float64))
    assert np.array_equal(cl2, cload.astype(np.float32).astype(np.float64))
    assert info["run_start"] == 3600 and info["run_end"] == 7200
    assert info["erumass"] == 2.0e11 and info["outmass"] == 1.0e8
    assert (info["xmin"], info["ymin"], info["xmax"], info["ymax"]) == (1.0e5, 2.0e5, 3.0e5, 4.0e5)


def test_restart_round_trip_and_refusals(tmp_path):
    F = _lib_or_skip()
    c, cload = _restart_state(9)
    path = str(tmp_path / "a.rst.nc")
    F.restart_write(path, c, cload, erumass=5.0, outmass=0.25, run_end=60)
    c2, cl2, info = F.restart_read(path, 6, 5, 4, 3)
    assert np.array_equal(c2, c.astype(np.float32).astype(np.float64))
    assert np.array_equal(cl2, cload.astype(np.float32).astype(np.float64))
    assert info["erumass"] == 5.0 and info["outmass"] == 0.25 and info["run_end"] == 60
    # rearst.f90:141-144: a file of another grid is refused
    for shape in ((7, 5, 4, 3), (6, 5, 5, 3), (6, 5, 4, 2)):
        with pytest.raises(F.F3DError, match="restart inconsistency"):
            F.restart_read(path, *shape)
    bad = tmp_path / "not.nc"
    bad.write_bytes(b"HDF\x01" + b"\0" * 64)
    with pytest.raises(F.F3DError, match="not a netCDF classic file"):
        F.restart_read(str(bad), 6, 5, 4, 3)
    trunc = tmp_path / "trunc.nc"
    trunc.write_bytes(open(path, "rb").read()[:-100])
    with pytest.raises(F.F3DError, match="no float CON|truncated"):
        F.restart_read(str(trunc), 6, 5, 4, 3)
