"""ctypes binding of libfall3d_amd.so (the C ABI in include/fall3d_amd.h).

`Context` wraps one device-resident f3d_ctx.  `Run` mirrors the reference
driver (Sources/Fall3d.f90:79-121) over a `Scenario`: one meteo read
(meteo.f90), then steps (setsrc + sizloop), and budgets (endstep/writim).

The library must be built (`python -m fall3d_amd.build`); importing this
module does not fall back to anything when it is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfall3d_amd.so")

FP64, FP32 = 0, 1
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_L = None

# every symbol include/fall3d_amd.h declares
DROPIN_SYMBOLS = ["setbcc_", "advctzc_", "advctx_", "advcty_", "diffz_", "diffx_", "diffy_",
                  "accum_", "wetdep_", "cmass2d_", "cmass3d_"]
CTX_SYMBOLS = ["f3d_create", "f3d_create_ensemble", "f3d_set_member_source",
               "f3d_ensemble_budget", "f3d_get_member_state", "f3d_destroy", "f3d_set_classes", "f3d_set_models", "f3d_set_meteo",
               "f3d_get_dt", "f3d_set_dt", "f3d_set_source", "f3d_step", "f3d_budget",
               "f3d_get_state", "f3d_set_state", "f3d_get_coef", "f3d_set_landuse",
               "f3d_get_vdep", "f3d_set_topography", "f3d_set_precipitation",
               "f3d_output_products", "f3d_set_pressure", "f3d_output_concentration",
               "f3d_write_restart", "f3d_read_restart", "f3d_restart_write", "f3d_restart_read",
               "f3d_get_clock",
               "f3d_set_clock", "f3d_get_cd_state", "f3d_set_cd_state", "f3d_sync",
               "f3d_set_tile", "f3d_set_tile_config", "f3d_timing", "f3d_kernel_time", "f3d_device_bytes",
               "f3d_last_error", "f3d_hbm_copy", "f3d_batch_create", "f3d_batch_destroy",
               "f3d_batch_set_tile", "f3d_batch_get_tile", "f3d_batch_use_graph",
               "f3d_batch_step", "f3d_batch_sync", "f3d_batch_autotune", "f3d_autotune",
               "f3d_set_meteo_window", "f3d_mem_info", "f3d_cmass3d_noalloc",
               "f3d_batch_timing", "f3d_batch_kernel_time", "f3d_set_gravity_current",
               "f3d_get_gravity_current", "f3d_get_divu"]


class F3DError(RuntimeError):
    pass


class _Gravity(ctypes.Structure):
    """f3d_gravity: the GRAVITY_CURRENT block (readat.f90:149-200)."""
    _fields_ = [(n, ctypes.c_double) for n in
                ("c_flow_rate", "lambda_grav", "k_entrain", "brunt_vaisala", "center_x",
                 "center_y", "xorigr", "yorigr", "dx", "dy", "beg_time")]


class _RunInfo(ctypes.Structure):
    _fields_ = [("year", ctypes.c_int), ("month", ctypes.c_int), ("day", ctypes.c_int),
                ("run_start", ctypes.c_int), ("run_end", ctypes.c_int),
                ("xmin", ctypes.c_double), ("ymin", ctypes.c_double),
                ("xmax", ctypes.c_double), ("ymax", ctypes.c_double),
                ("erumass", ctypes.c_double), ("outmass", ctypes.c_double)]


class _Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("nc", ctypes.c_int), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("zlayer", _dp), ("hm1", _dp), ("lonlat", ctypes.c_int),
                ("dlon", ctypes.c_double), ("dlat", ctypes.c_double)]


def lib():
    """Load libfall3d_amd.so; raises if it has not been built."""
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise F3DError(f"{LIB_PATH} is not built: run `python -m fall3d_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        c, d, i, dp, ip, vp = (ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _ip,
                               ctypes.c_void_p)
        L.f3d_last_error.restype = ctypes.c_char_p
        L.f3d_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(_Grid), i]
        L.f3d_create_ensemble.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(_Grid), i]
        L.f3d_set_member_source.argtypes = [vp, i, i, ip, ip, ip, dp, d]
        L.f3d_ensemble_budget.argtypes = [vp, dp]
        L.f3d_get_member_state.argtypes = [vp, i, dp, dp]
        L.f3d_destroy.argtypes = [vp]
        L.f3d_destroy.restype = None
        L.f3d_set_classes.argtypes = [vp, dp, dp, dp, i]
        L.f3d_set_models.argtypes = [vp, i, i, d, d, d]
        L.f3d_set_meteo.argtypes = [vp] + [dp] * 10
        L.f3d_set_meteo_window.argtypes = [vp, vp, i, i]
        L.f3d_mem_info.argtypes = [ctypes.POINTER(ctypes.c_longlong)] * 2
        L.f3d_mem_info.restype = c
        L.f3d_set_meteo_window.restype = c
        L.f3d_get_dt.argtypes = [vp, dp, dp]
        L.f3d_set_dt.argtypes = [vp, d]
        L.f3d_set_source.argtypes = [vp, i, ip, ip, ip, dp, d]
        L.f3d_step.argtypes = [vp, i]
        L.f3d_budget.argtypes = [vp, dp]
        L.f3d_get_state.argtypes = [vp, dp, dp]
        L.f3d_set_state.argtypes = [vp, dp, dp]
        L.f3d_get_coef.argtypes = [vp] + [dp] * 6
        L.f3d_set_landuse.argtypes = [vp, dp, dp]
        L.f3d_get_vdep.argtypes = [vp, dp]
        L.f3d_set_topography.argtypes = [vp, dp]
        L.f3d_set_precipitation.argtypes = [vp, dp]
        L.f3d_set_gravity_current.argtypes = [vp, ctypes.POINTER(_Gravity), i, dp]
        L.f3d_get_gravity_current.argtypes = [vp, dp, dp, ip]
        L.f3d_get_divu.argtypes = [vp, dp]
        L.f3d_output_products.argtypes = [vp, i, dp, dp, dp]
        L.f3d_set_pressure.argtypes = [vp, dp]
        L.f3d_output_concentration.argtypes = [vp, i, dp, dp, i, i, dp, dp]
        L.f3d_write_restart.argtypes = [vp, i, ctypes.c_char_p, ctypes.POINTER(_RunInfo)]
        L.f3d_read_restart.argtypes = [vp, ctypes.c_char_p, dp, ctypes.POINTER(_RunInfo)]
        L.f3d_restart_write.argtypes = [ctypes.c_char_p, i, i, i, i, i, dp, dp,
                                        ctypes.POINTER(_RunInfo)]
        L.f3d_restart_read.argtypes = [ctypes.c_char_p, i, i, i, i, dp, dp,
                                       ctypes.POINTER(_RunInfo)]
        L.f3d_get_clock.argtypes = [vp, dp, ip, dp, dp]
        L.f3d_set_clock.argtypes = [vp, d, i, d, d]
        L.f3d_get_cd_state.argtypes = [vp, dp]
        L.f3d_set_cd_state.argtypes = [vp, d]
        L.f3d_sync.argtypes = [vp]
        L.f3d_set_tile.argtypes = [vp, i, i]
        L.f3d_set_tile_config.argtypes = [vp, i, i, i, i]
        L.f3d_timing.argtypes = [vp, i]
        L.f3d_kernel_time.argtypes = [vp, i, dp, ctypes.POINTER(ctypes.c_long)]
        L.f3d_device_bytes.argtypes = [vp]
        L.f3d_device_bytes.restype = ctypes.c_longlong
        L.f3d_cmass3d_noalloc.argtypes = [vp, vp, vp, vp, i, i, i, i, i, d, d, vp, vp, vp]
        L.f3d_cmass3d_noalloc.restype = c
        L.f3d_hbm_copy.argtypes = [ctypes.c_longlong, i, dp]
        L.f3d_hbm_copy.restype = c
        L.f3d_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), i]
        L.f3d_batch_destroy.argtypes = [vp]
        L.f3d_batch_destroy.restype = None
        L.f3d_batch_set_tile.argtypes = [vp, i, i, i]
        L.f3d_batch_get_tile.argtypes = [vp, ip, ip, ip]
        L.f3d_batch_use_graph.argtypes = [vp, i]
        L.f3d_batch_step.argtypes = [vp, i]
        L.f3d_batch_sync.argtypes = [vp]
        L.f3d_batch_timing.argtypes = [vp, i]
        L.f3d_batch_kernel_time.argtypes = [vp, i, dp, ctypes.POINTER(ctypes.c_long)]
        L.f3d_batch_autotune.argtypes = [vp, i, ip, ip, ip, dp]
        L.f3d_autotune.argtypes = [vp, i, ip, ip, ip, dp]
        for name in ("f3d_batch_create", "f3d_batch_set_tile", "f3d_batch_get_tile",
                     "f3d_batch_use_graph", "f3d_batch_step", "f3d_batch_sync",
                     "f3d_batch_autotune", "f3d_autotune"):
            getattr(L, name).restype = c
        for name in ("f3d_create", "f3d_create_ensemble", "f3d_set_member_source",
                     "f3d_ensemble_budget", "f3d_get_member_state", "f3d_set_classes", "f3d_set_models", "f3d_set_meteo",
                     "f3d_get_dt", "f3d_set_dt", "f3d_set_source", "f3d_step", "f3d_budget",
                     "f3d_get_state", "f3d_set_state", "f3d_get_coef", "f3d_set_landuse",
                     "f3d_get_vdep", "f3d_set_topography", "f3d_set_precipitation",
                     "f3d_set_gravity_current", "f3d_get_gravity_current", "f3d_get_divu",
                     "f3d_output_products", "f3d_set_pressure", "f3d_output_concentration",
                     "f3d_write_restart", "f3d_read_restart", "f3d_restart_write",
                     "f3d_restart_read",
                     "f3d_get_clock",
                     "f3d_set_clock", "f3d_get_cd_state", "f3d_set_cd_state", "f3d_sync",
                     "f3d_set_tile", "f3d_set_tile_config", "f3d_timing", "f3d_kernel_time"):
            getattr(L, name).restype = c
        _L = L
    return _L


def restart_write(path, c, cload, lonlat=0, **info):
    """Restart file (netCDF classic, wrirst's layout) from host arrays c
    (nx+2,ny+2,nz+2,nc) and cload (nx,ny,nc+1); info: year, month, day,
    run_start, run_end, xmin, ymin, xmax, ymax, erumass, outmass."""
    c, cload = _F(c), _F(cload)
    nx, ny, nz, nc = c.shape[0] - 2, c.shape[1] - 2, c.shape[2] - 2, c.shape[3]
    ri = _RunInfo(**{k: info.get(k, 0) for k, _ in _RunInfo._fields_})
    _check(lib().f3d_restart_write(os.fsencode(path), nx, ny, nz, nc, int(lonlat), _P(c),
                                   _P(cload), ctypes.byref(ri)), "f3d_restart_write")


def restart_read(path, nx, ny, nz, nc):
    """(c, cload, attributes) of a restart file of the given grid."""
    c = np.zeros((nx + 2, ny + 2, nz + 2, nc), order="F")
    cload = np.zeros((nx, ny, nc + 1), order="F")
    ri = _RunInfo()
    _check(lib().f3d_restart_read(os.fsencode(path), nx, ny, nz, nc, _P(c), _P(cload),
                                  ctypes.byref(ri)), "f3d_restart_read")
    return c, cload, {k: getattr(ri, k) for k, _ in _RunInfo._fields_}


def hbm_copy(nbytes=1 << 30, reps=5):
    """Average time (ms) of a plain device copy of nbytes in + nbytes out
    (f3d_hbm_copy): the calibration kernel for HBM bandwidth and counters."""
    ms = ctypes.c_double()
    _check(lib().f3d_hbm_copy(int(nbytes), int(reps), ctypes.byref(ms)), "f3d_hbm_copy")
    return ms.value


def mem_info():
    """(free, total) device memory in bytes (f3d_mem_info)."""
    fr, tot = ctypes.c_longlong(), ctypes.c_longlong()
    _check(lib().f3d_mem_info(ctypes.byref(fr), ctypes.byref(tot)), "f3d_mem_info")
    return fr.value, tot.value


def last_error():
    return lib().f3d_last_error().decode(errors="replace")


def _check(rc, what):
    if rc != 0:
        raise F3DError(f"{what} failed ({rc}): {last_error()}")


def _F(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _P(a):
    if a is None:
        return None
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    assert a.dtype == np.float64
    assert a.flags.f_contiguous if a.ndim > 1 else a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Context:
    """One device-resident scenario (f3d_ctx)."""

    def __init__(self, nx, ny, nz, nc, dx, dy, zlayer, precision=FP32, hm1=None, lonlat=0,
                 members=1, dlon=0.0, dlat=0.0):
        L = lib()
        self.nx, self.ny, self.nz, self.nc = nx, ny, nz, nc
        self.precision = precision
        self.members = members
        self._zl = _F(zlayer)
        self._hm1 = None if hm1 is None else _F(hm1)
        g = _Grid(nx, ny, nz, nc, float(dx), float(dy), _P(self._zl), _P(self._hm1), lonlat,
                  float(dlon), float(dlat))
        h = ctypes.c_void_p()
        if members == 1:
            _check(L.f3d_create(ctypes.byref(h), ctypes.byref(g), precision), "f3d_create")
        else:
            _check(L.f3d_create_ensemble(ctypes.byref(h), ctypes.byref(g), members),
                   "f3d_create_ensemble")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().f3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_classes(self, diam, rhop, sphe, modv):
        self._cls = [_F(diam), _F(rhop), _F(sphe)]
        _check(lib().f3d_set_classes(self.h, *[_P(a) for a in self._cls], int(modv)),
               "f3d_set_classes")

    def set_models(self, modkv, modkh, rkv0, rkh0, csh=0.2275):
        _check(lib().f3d_set_models(self.h, int(modkv), int(modkh), float(rkv0), float(rkh0),
                                    float(csh)), "f3d_set_models")

    def set_meteo(self, vx, vy, vz, tempe, vptemp, rho, ustar, zi, rl, vdep):
        arrs = [_F(a) for a in (vx, vy, vz, tempe, vptemp, rho, ustar, zi, rl, vdep)]
        _check(lib().f3d_set_meteo(self.h, *[_P(a) for a in arrs]), "f3d_set_meteo")
        return self.dt

    def set_meteo_window(self, region, i0, j0):
        """Meteorology from the window (i0+1.., j0+1.., 1..nz) of a region
        context's fields in HBM (f3d_set_meteo_window)."""
        _check(lib().f3d_set_meteo_window(self.h, region.h, int(i0), int(j0)),
               "f3d_set_meteo_window")
        return self.dt

    @property
    def dt(self):
        dt, dtc = ctypes.c_double(), ctypes.c_double()
        _check(lib().f3d_get_dt(self.h, ctypes.byref(dt), ctypes.byref(dtc)), "f3d_get_dt")
        return dt.value

    @property
    def dtc(self):
        dt, dtc = ctypes.c_double(), ctypes.c_double()
        _check(lib().f3d_get_dt(self.h, ctypes.byref(dt), ctypes.byref(dtc)), "f3d_get_dt")
        return dtc.value

    def set_dt(self, dt):
        _check(lib().f3d_set_dt(self.h, float(dt)), "f3d_set_dt")

    def set_source(self, isrc, jsrc, ksrc, tmrat, flowrate):
        self._src = [np.ascontiguousarray(a, dtype=np.int32) for a in (isrc, jsrc, ksrc)]
        self._tm = _F(tmrat)
        _check(lib().f3d_set_source(self.h, len(self._src[0]), *[_P(a) for a in self._src],
                                    _P(self._tm), float(flowrate)), "f3d_set_source")

    def set_member_source(self, m, isrc, jsrc, ksrc, tmrat, flowrate):
        src = [np.ascontiguousarray(a, dtype=np.int32) for a in (isrc, jsrc, ksrc)]
        tm = _F(tmrat)
        _check(lib().f3d_set_member_source(self.h, int(m), len(src[0]), *[_P(a) for a in src],
                                           _P(tm), float(flowrate)), "f3d_set_member_source")

    def step(self, n=1):
        _check(lib().f3d_step(self.h, int(n)), "f3d_step")

    def sync(self):
        _check(lib().f3d_sync(self.h), "f3d_sync")

    def budget(self):
        out = np.zeros(6)
        _check(lib().f3d_budget(self.h, _P(out)), "f3d_budget")
        keys = ("massvol", "massdep", "massdry", "masswet", "outmass", "erumass")
        return dict(zip(keys, out.tolist()))

    def state(self, member=0):
        nx, ny, nz, nc = self.nx, self.ny, self.nz, self.nc
        c = np.zeros((nx + 2, ny + 2, nz + 2, nc), order="F")
        cdry = np.zeros((nx, ny, nc + 1), order="F")
        _check(lib().f3d_get_member_state(self.h, int(member), _P(c), _P(cdry)),
               "f3d_get_member_state")
        return c, cdry

    def ground_load(self, member=0):
        """Deposit per class and its total, cdry(nx,ny,0:nc) (no state copy)."""
        cdry = np.zeros((self.nx, self.ny, self.nc + 1), order="F")
        _check(lib().f3d_get_member_state(self.h, int(member), None, _P(cdry)),
               "f3d_get_member_state")
        return cdry

    def ensemble_budget(self):
        out = np.zeros(6 * self.members)
        _check(lib().f3d_ensemble_budget(self.h, _P(out)), "f3d_ensemble_budget")
        keys = ("massvol", "massdep", "massdry", "masswet", "outmass", "erumass")
        return [dict(zip(keys, out[6 * m:6 * m + 6].tolist())) for m in range(self.members)]

    def set_state(self, c=None, cdry=None):
        c = None if c is None else _F(c)
        cdry = None if cdry is None else _F(cdry)
        _check(lib().f3d_set_state(self.h, _P(c), _P(cdry)), "f3d_set_state")

    def coef(self):
        nx, ny, nz, nc = self.nx, self.ny, self.nz, self.nc
        a = [np.zeros((nx, ny, nz), order="F") for _ in range(5)]
        vset = np.zeros((nx, ny, nz, nc), order="F")
        _check(lib().f3d_get_coef(self.h, *[_P(x) for x in a], _P(vset)), "f3d_get_coef")
        return dict(vx=a[0], rho=a[1], rkhor=a[2], rkh1=a[3], rkver=a[4], vset=vset)

    def set_landuse(self, ldu, znt):
        """Land use (USGS category) and roughness length per column: later
        meteo reads compute vdep on the device (setvdrydep)."""
        self._ldu, self._znt = _F(ldu), _F(znt)
        _check(lib().f3d_set_landuse(self.h, _P(self._ldu), _P(self._znt)), "f3d_set_landuse")

    def set_topography(self, topg):
        """Terrain height per column: LON-LAT grids correct vz for its slope
        at every later meteo read (corvver)."""
        self._topg = _F(topg)
        _check(lib().f3d_set_topography(self.h, _P(self._topg)), "f3d_set_topography")

    def set_precipitation(self, prate):
        """Precipitation rate per column (mm/h): wet deposition in every later
        step (through f3d_step; batches refuse such contexts)."""
        self._prate = _F(prate)
        _check(lib().f3d_set_precipitation(self.h, _P(self._prate)), "f3d_set_precipitation")

    def set_gravity_current(self, zs, **prm):
        """Gravity current on (prm: c_flow_rate, lambda_grav, k_entrain,
        brunt_vaisala, center_x, center_y, xorigr, yorigr, dx, dy, beg_time;
        zs: the source heights) or off (zs None)."""
        if zs is None:
            _check(lib().f3d_set_gravity_current(self.h, None, 0, None), "f3d_set_gravity_current")
            return
        gc = _Gravity(*[float(prm[n]) for n, _ in _Gravity._fields_])
        self._zs = np.ascontiguousarray(zs, dtype=np.float64)
        _check(lib().f3d_set_gravity_current(self.h, ctypes.byref(gc), int(self._zs.size),
                                             _P(self._zs)), "f3d_set_gravity_current")

    def gravity_state(self):
        """(diag[9], corrdiv, active) of the last meteo read and divcorr."""
        diag = np.zeros(9)
        cd, act = ctypes.c_double(), ctypes.c_int()
        _check(lib().f3d_get_gravity_current(self.h, _P(diag), ctypes.byref(cd), ctypes.byref(act)),
               "f3d_get_gravity_current")
        return diag, cd.value, bool(act.value)

    def divu(self):
        """The wind divergence (nx,ny,nz) of setdivu."""
        out = np.zeros((self.nx, self.ny, self.nz), order="F")
        _check(lib().f3d_get_divu(self.h, _P(out)), "f3d_get_divu")
        return out

    def products(self, member=0):
        """Output products of a member: load (nx,ny,0:nc) kg/m2, deposit
        thickness (nx,ny) cm, z-cumulative concentration (nx,ny) g/m2."""
        load = np.zeros((self.nx, self.ny, self.nc + 1), order="F")
        thick = np.zeros((self.nx, self.ny), order="F")
        colmass = np.zeros((self.nx, self.ny), order="F")
        _check(lib().f3d_output_products(self.h, int(member), _P(load), _P(thick), _P(colmass)),
               "f3d_output_products")
        return load, thick, colmass

    def set_pressure(self, p):
        """Pressure (nx,ny,nz) in Pa, for the flight-level cuts."""
        self._p = _F(p)
        _check(lib().f3d_set_pressure(self.h, _P(self._p)), "f3d_set_pressure")

    def concentration(self, member=0, asl=False, fl=None):
        """Concentration outputs of a member in g/m3: C_GRND (nx,ny), CON
        (nx,ny,nz) (terrain following, or a.s.l. with asl) and, for the
        flight levels fl (hundreds of feet), the cuts (nx,ny,len(fl))."""
        cg = np.zeros((self.nx, self.ny), order="F")
        con = np.zeros((self.nx, self.ny, self.nz), order="F")
        fl = None if fl is None else np.ascontiguousarray(fl, dtype=np.float64)
        nfl = 0 if fl is None else int(fl.size)
        cfl = np.zeros((self.nx, self.ny, max(nfl, 1)), order="F")
        _check(lib().f3d_output_concentration(self.h, int(member), _P(cg), _P(con), int(bool(asl)),
                                              nfl, None if fl is None else _P(fl),
                                              _P(cfl) if nfl else None),
               "f3d_output_concentration")
        return cg, con, (cfl if nfl else None)

    def write_restart(self, path, member=0, year=2000, month=1, day=1, run_start=0, run_end=0,
                      extent=(0.0, 0.0, 0.0, 0.0)):
        """Restart file of a member (netCDF classic: CON, LOAD and the run's
        attributes, as wrirst writes it)."""
        info = _RunInfo(int(year), int(month), int(day), int(run_start), int(run_end),
                        *[float(v) for v in extent], 0.0, 0.0)
        _check(lib().f3d_write_restart(self.h, int(member), os.fsencode(path), ctypes.byref(info)),
               "f3d_write_restart")

    def read_restart(self, path):
        """Restores the state and the erupted/outflow mass from a restart
        file; returns (cload (nx,ny,0:nc), attributes)."""
        cload = np.zeros((self.nx, self.ny, self.nc + 1), order="F")
        info = _RunInfo()
        _check(lib().f3d_read_restart(self.h, os.fsencode(path), _P(cload), ctypes.byref(info)),
               "f3d_read_restart")
        return cload, {k: getattr(info, k) for k, _ in _RunInfo._fields_}

    def vdep(self):
        out = np.zeros((self.nx, self.ny, self.nc), order="F")
        _check(lib().f3d_get_vdep(self.h, _P(out)), "f3d_get_vdep")
        return out

    def clock(self):
        t, it, em, om = ctypes.c_double(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        _check(lib().f3d_get_clock(self.h, ctypes.byref(t), ctypes.byref(it), ctypes.byref(em),
                                   ctypes.byref(om)), "f3d_get_clock")
        return dict(time=t.value, iiter=it.value, erumass=em.value, outmass=om.value)

    def set_clock(self, time, iiter, erumass, outmass):
        _check(lib().f3d_set_clock(self.h, float(time), int(iiter), float(erumass),
                                   float(outmass)), "f3d_set_clock")

    @property
    def cd_state(self):
        v = ctypes.c_double()
        _check(lib().f3d_get_cd_state(self.h, ctypes.byref(v)), "f3d_get_cd_state")
        return v.value

    @cd_state.setter
    def cd_state(self, v):
        _check(lib().f3d_set_cd_state(self.h, float(v)), "f3d_set_cd_state")

    def set_tile(self, tx, ty, waves=0, slabs=0):
        _check(lib().f3d_set_tile_config(self.h, int(tx), int(ty), int(waves), int(slabs)),
               "f3d_set_tile")

    def timing(self, enable=True):
        _check(lib().f3d_timing(self.h, 1 if enable else 0), "f3d_timing")

    def kernel_time(self, reset=False):
        ms, n = ctypes.c_double(), ctypes.c_long()
        _check(lib().f3d_kernel_time(self.h, 1 if reset else 0, ctypes.byref(ms),
                                     ctypes.byref(n)), "f3d_kernel_time")
        return ms.value, n.value

    @property
    def device_bytes(self):
        return int(lib().f3d_device_bytes(self.h))

    def autotune(self, trials=3):
        """Pick the fastest tile shape for this context (f3d_autotune)."""
        tx, ty, sl, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().f3d_autotune(self.h, int(trials), ctypes.byref(tx), ctypes.byref(ty),
                                  ctypes.byref(sl), ctypes.byref(ms)), "f3d_autotune")
        return (tx.value, ty.value, sl.value), ms.value


class Run:
    """Driver over one device context with the reference's control flow:
    meteo read, then time += dt; iiter += 1; erumass += flowrate*dt;
    setsrc; sizloop (Fall3d.f90:79-121)."""

    def __init__(self, sc, precision=FP32, cd_state=1.0, tile=None):
        self.sc = sc
        lonlat = int(getattr(sc, "lonlat", 0))
        self.ctx = Context(sc.nx, sc.ny, sc.nz, sc.nc, sc.dx, sc.dy, sc.zlayer, precision,
                           hm1=getattr(sc, "hm1_cols", None), lonlat=lonlat,
                           dlon=getattr(sc, "dlon", 0.0), dlat=getattr(sc, "dlat", 0.0))
        if lonlat:
            self.ctx.set_topography(sc.topg)
        if tile is not None:
            self.ctx.set_tile(*tile)
        self.ctx.set_classes(sc.diam, sc.rhop, sc.sphe, sc.modv)
        self.ctx.set_models(sc.modkv, sc.modkh, sc.rkv0, sc.rkh0, sc.csh)
        self.ctx.set_source(sc.isrc, sc.jsrc, sc.ksrc, sc.tmrat, sc.flowrate)
        if getattr(sc, "prate", None) is not None:
            self.ctx.set_precipitation(sc.prate)
        if getattr(sc, "gravity", None) is not None:
            self.ctx.set_gravity_current(sc.zs, **sc.gravity)
        self.ctx.cd_state = cd_state
        self.nsteps = 0

    def meteo_update(self):
        sc = self.sc
        return self.ctx.set_meteo(sc.vx, sc.vy, sc.vz, sc.tempe, sc.vptemp, sc.rho, sc.ustar,
                                  sc.zi, sc.rl, sc.vdep)

    @property
    def dt(self):
        return self.ctx.dt

    def step(self, n=1):
        self.ctx.step(n)
        self.nsteps += n

    def run_to(self, end_time):
        """Steps until time >= end_time (Fall3d.f90:120; the last step may overshoot)."""
        dt = self.ctx.dt
        t = self.ctx.clock()["time"]
        n = 0
        while t < end_time:
            t += dt
            n += 1
        if n:
            self.step(n)
        return n

    def state(self):
        return self.ctx.state()

    def budget(self):
        return self.ctx.budget()

    def close(self):
        self.ctx.close()


class Ensemble:
    """Members on one grid and meteorology (one dt) that differ only in their
    source; one fused launch per step advances all of them."""

    def __init__(self, base, members):
        sc = base
        self.sc = sc
        self.members = list(members)
        self.ctx = Context(sc.nx, sc.ny, sc.nz, sc.nc, sc.dx, sc.dy, sc.zlayer, FP32,
                           members=len(self.members))
        self.ctx.set_classes(sc.diam, sc.rhop, sc.sphe, sc.modv)
        self.ctx.set_models(sc.modkv, sc.modkh, sc.rkv0, sc.rkh0, sc.csh)
        for m, msc in enumerate(self.members):
            self.ctx.set_member_source(m, msc.isrc, msc.jsrc, msc.ksrc, msc.tmrat, msc.flowrate)

    def meteo_update(self):
        sc = self.sc
        return self.ctx.set_meteo(sc.vx, sc.vy, sc.vz, sc.tempe, sc.vptemp, sc.rho, sc.ustar,
                                  sc.zi, sc.rl, sc.vdep)

    def step(self, n=1):
        self.ctx.step(n)

    def state(self, m):
        return self.ctx.state(m)

    def budgets(self):
        return self.ctx.ensemble_budget()

    def close(self):
        self.ctx.close()


class Batch:
    """A set of FP32 contexts stepped together (f3d_batch): one persistent
    launch per step over all their work items, steps replayed from HIP
    graphs.  Each context keeps its own clock, sources and budgets."""

    def __init__(self, contexts, tile=None, graph=True):
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = ctypes.c_void_p()
        _check(lib().f3d_batch_create(ctypes.byref(h), arr, len(self.contexts)),
               "f3d_batch_create")
        self.h = h
        if tile is not None:
            self.set_tile(*tile)
        if not graph:
            _check(lib().f3d_batch_use_graph(self.h, 0), "f3d_batch_use_graph")

    def set_tile(self, tx, ty, slabs=2):
        _check(lib().f3d_batch_set_tile(self.h, int(tx), int(ty), int(slabs)),
               "f3d_batch_set_tile")

    def tile(self):
        tx, ty, sl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().f3d_batch_get_tile(self.h, ctypes.byref(tx), ctypes.byref(ty),
                                        ctypes.byref(sl)), "f3d_batch_get_tile")
        return tx.value, ty.value, sl.value

    def autotune(self, trials=3):
        tx, ty, sl, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().f3d_batch_autotune(self.h, int(trials), ctypes.byref(tx), ctypes.byref(ty),
                                        ctypes.byref(sl), ctypes.byref(ms)), "f3d_batch_autotune")
        return (tx.value, ty.value, sl.value), ms.value

    def step(self, n=1):
        _check(lib().f3d_batch_step(self.h, int(n)), "f3d_batch_step")

    def timing(self, enable=True):
        _check(lib().f3d_batch_timing(self.h, 1 if enable else 0), "f3d_batch_timing")

    def kernel_time(self, reset=False):
        ms, n = ctypes.c_double(), ctypes.c_long()
        _check(lib().f3d_batch_kernel_time(self.h, 1 if reset else 0, ctypes.byref(ms),
                                           ctypes.byref(n)), "f3d_batch_kernel_time")
        return ms.value, n.value

    def sync(self):
        _check(lib().f3d_batch_sync(self.h), "f3d_batch_sync")

    def close(self):
        if getattr(self, "h", None):
            lib().f3d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
