"""ctypes binding of the C restatement (liboracle.so) and of the
reference build (_ref/libf3dref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.

`OracleRun` and `RefRun` expose the same driver-level API (meteo update,
one step, budgets, state) so tests can compare them line by line.  Both
mirror the reference driver loop (Sources/Fall3d.f90:79-121): time += dt,
iiter += 1, erumass += flowrate*dt, setsrc, sizloop.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libf3dref.so")
# the reference's sizloop/setsrc/meteo-stage objects and the harness linked
# against libfall3d_amd.so for setbcc_ ... cmass3d_ (oracle/Makefile dropin)
DROPIN_PATH = os.path.join(HERE, "_ref", "libf3dref_dropin.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def P(a):
    """Pointer to a contiguous float64 (or int32) numpy buffer."""
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    assert a.dtype == np.float64, a.dtype
    # Fortran layout: multi-dimensional arrays must be column-major
    assert a.flags.f_contiguous if a.ndim > 1 else a.flags.c_contiguous, "need Fortran order"
    return a.ctypes.data_as(_dp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        d, i, dp, ip = ctypes.c_double, ctypes.c_int, _dp, _ip
        L.or_finddt.restype = d
        L.or_finddt.argtypes = [i, i, i, i, dp, d, d, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_setvset.restype = ctypes.c_long
        L.or_setvset.argtypes = [i, i, i, i, dp, dp, dp, dp, dp, dp, i, dp, dp]
        L.or_kappa3.argtypes = [i, i, i, i, i, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                d, d, d, d, d, d, d, i, dp, dp]
        L.or_setpsi.argtypes = [dp, dp, dp, i, i]
        L.or_get_gama.argtypes = [d, d, dp]
        L.or_ganser_consts.argtypes = [d, d, d, dp, dp]
        L.or_meteo_scale.argtypes = [i, i, i, dp, dp, dp, dp, dp]
        L.or_setvdrydep.argtypes = [i, i, i, d, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_corvver.argtypes = [i, i, i, dp, dp, d, d, dp, dp, dp]
        L.or_setsrc.argtypes = [dp, i, i, i, i, i, ip, ip, ip, dp, dp, dp, dp, d, d, d]
        L.or_sizloop.argtypes = [dp, dp, i, i, i, i, i, d, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                 dp, d, d, dp, i, dp, dp, dp, dp, dp]
        L.or_wetdep.argtypes = [dp, dp, i, i, i, i, d, dp, dp, dp, dp, dp]
        L.or_output_products.argtypes = [i, i, i, i, dp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_cmass3d.argtypes = [dp, dp, dp, dp, i, i, i, i, d, d, dp, dp, dp, dp, dp]
        L.or_cmass2d.argtypes = [dp, dp, dp, dp, dp, i, i, i, i, d, d, dp, d, dp]
        L.or_getc0.argtypes = [dp, dp, i, i, i]
        L.or_setbcc.argtypes = [dp, i, i, i, dp, dp, dp, dp, dp, dp, dp]
        L.or_advctzc.argtypes = [dp, i, i, i, dp, dp, dp, d]
        L.or_advctx.argtypes = [dp, i, i, i, dp, d, d]
        L.or_advcty.argtypes = [dp, i, i, i, dp, d, d]
        L.or_diffz.argtypes = [dp, i, i, i, dp, d, dp, dp]
        L.or_diffx.argtypes = [dp, i, i, i, d, d, dp, dp]
        L.or_diffy.argtypes = [dp, i, i, i, d, d, dp, dp]
        L.or_accum.argtypes = [dp, dp, i, i, i, i, d, dp, dp]
        L.or_clip.argtypes = [dp, ctypes.c_long]
        L.or_standard_atm.argtypes = [d, dp, dp]
        L.or_output_concentration.argtypes = [i, i, i, i, dp, dp, dp, dp, dp, i, dp, dp, dp, i,
                                              dp]
        L.or_setdivu.argtypes = [i, i, i, dp, dp, dp, dp, d, d, dp, dp]
        L.or_addradialwind.restype = i
        L.or_addradialwind.argtypes = [i, i, i, dp, dp, dp, d, d, dp, dp, d, d, d, dp]
        L.or_divcorr.restype = d
        L.or_divcorr.argtypes = [dp, i, i, i, i, dp, d, d, d, d, d]
        _lib = L
    return _lib


def F(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def mass_flow_rate(tmrat):
    """sum(tmrat(1:nsou,1:nc)) in array order (addradialwind.f90:70): one
    running sum down the columns, as the reference's SUM forms it."""
    s = 0.0
    for v in np.asarray(tmrat, dtype=np.float64).ravel(order="F"):
        s = s + float(v)
    return s


def gravity_prm(g):
    """The ten parameters of or_addradialwind from a scenario's gravity dict."""
    return np.array([g["c_flow_rate"], g["lambda_grav"], g["k_entrain"], g["brunt_vaisala"],
                     g["center_x"], g["center_y"], g["xorigr"], g["yorigr"], g["dx"], g["dy"]])


def log_time(iiter, time, dt):
    """endstep.f90:53: the log (writim, divcorr) runs after the first step
    and whenever mod(time, 600) <= dt."""
    return iiter == 1 or np.fmod(time, 600.0) <= dt


class _RunBase:
    def __init__(self, sc):
        self.sc = sc
        self.nx, self.ny, self.nz, self.nc = sc.nx, sc.ny, sc.nz, sc.nc
        self.time = 0.0
        self.iiter = 0
        self.erumass = 0.0
        self.outmass = 0.0
        self.dt = 0.0
        self.dtc = 0.0

    def run(self, nsteps=None, end_time=None):
        end = self.sc.end_time if end_time is None else end_time
        n = 0
        while True:
            self.step()
            n += 1
            if nsteps is not None and n >= nsteps:
                break
            if nsteps is None and self.time >= end:
                break
        return n


def _khn_spacing(sc):
    """(dx, dy) as kappa3's CMAQ Khn takes them (kappa3.f90:226): m on UTM
    grids; on LON-LAT grids the spacing in degrees of reagrd.f90:155-156,
    given as sc.dlon/sc.dlat or recovered from dX1 = Rearth*(dx*pi/180)
    (reagrd.f90:157-158, Rearth = 6356 km, Numeric.f90:14)."""
    if not int(getattr(sc, "lonlat", 0)):
        return sc.dx, sc.dy
    m_per_deg = 6356.0e3 * (3.14159265358979323846 / 180.0)
    dlon, dlat = getattr(sc, "dlon", 0.0), getattr(sc, "dlat", 0.0)
    return (dlon if dlon > 0 else sc.dx / m_per_deg, dlat if dlat > 0 else sc.dy / m_per_deg)


def output_concentration(c, hm1, zlayer, topg=None, p=None, fl=None, asl=False):
    """C_GRND (nx,ny), CON (nx,ny,nz) and the flight-level cuts (nx,ny,nfl)
    of printres_nc.f90 (setcut.f90, ter2asl.f90) from a state c
    (nx+2,ny+2,nz+2,nc) in the reference's operation order."""
    nx, ny, nz, nc = c.shape[0] - 2, c.shape[1] - 2, c.shape[2] - 2, c.shape[3]
    cg = np.zeros((nx, ny), order="F")
    con = np.zeros((nx, ny, nz), order="F")
    nfl = 0 if fl is None else len(fl)
    cfl = np.zeros((nx, ny, max(nfl, 1)), order="F")
    fla = np.ascontiguousarray(fl if fl is not None else [0.0], dtype=np.float64)
    lib().or_output_concentration(nx, ny, nz, nc, P(F(c)), P(F(hm1)), P(F(zlayer)),
                                  None if topg is None else P(F(topg)),
                                  None if p is None else P(F(p)), nfl, P(fla), P(cg), P(con),
                                  int(bool(asl)), P(cfl))
    return cg, con, (cfl if nfl else None)


def standard_atm(z):
    """(p, T) of the ICAO standard atmosphere at z m (setcut.f90:61-107)."""
    p, T = ctypes.c_double(), ctypes.c_double()
    lib().or_standard_atm(float(z), ctypes.byref(p), ctypes.byref(T))
    return p.value, T.value


class OracleRun(_RunBase):
    """Driver over the C restatement (FP64, serial semantics)."""

    def __init__(self, sc, nthreads=1, cd_state=1.0):
        super().__init__(sc)
        L = lib()
        self.nthreads = nthreads
        nx, ny, nz, nc = self.nx, self.ny, self.nz, self.nc
        self.dzv = F(sc.dz)
        self.zl = F(sc.zlayer)
        self.hm1 = F(sc.hm1)
        self.c = np.zeros((nx + 2, ny + 2, nz + 2, nc), order="F")
        self.cdry = np.zeros((nx, ny, nc + 1), order="F")
        self.cwet = np.zeros((nx, ny, nc + 1), order="F")
        # wet deposition inputs (sizloop.f90:65-67); prate None: off
        self._diam = F(sc.diam)
        self._zi = F(sc.zi)
        self._prate = None if getattr(sc, "prate", None) is None else F(sc.prate)
        self.psi = np.zeros(nc)
        L.or_setpsi(P(self.psi), P(F(sc.sphe)), P(F(sc.diam)), sc.modv, nc)
        self.cd = ctypes.c_double(cd_state)
        self.isrc = np.ascontiguousarray(sc.isrc, dtype=np.int32)
        self.jsrc = np.ascontiguousarray(sc.jsrc, dtype=np.int32)
        self.ksrc = np.ascontiguousarray(sc.ksrc, dtype=np.int32)
        self.tmrat = F(sc.tmrat)  # (nsou,nc); UTM: tmrat*Hm1 = tmrat (source.f90:75-77)

    def meteo_update(self):
        """kappa3 -> setvset -> (setvdrydep: vdep given) -> Hm1 scaling ->
        corvver (LON-LAT) -> finddt (meteo.f90:27,81,100,108-112,116-124,149)."""
        L, sc = lib(), self.sc
        nx, ny, nz, nc = self.nx, self.ny, self.nz, self.nc
        self.vx = F(sc.vx).copy(order="F")
        self.vy = F(sc.vy)
        self.vz = F(sc.vz).copy(order="F")
        self.rho = F(sc.rho).copy(order="F")
        lonlat = int(getattr(sc, "lonlat", 0))
        self.rkver = np.zeros((nx, ny, nz), order="F")
        self.rkhor = np.zeros((nx, ny, nz), order="F")
        self.rkh1 = np.zeros((nx, ny, nz), order="F")
        self.vset = np.zeros((nx, ny, nz, nc), order="F")
        self.vdep = F(sc.vdep)
        # kappa3 takes (dX1, dX2) in m, and (dx, dy) for CMAQ's Khn: degrees
        # on LON-LAT grids (kappa3.f90:151,193-226)
        kdx, kdy = _khn_spacing(sc)
        L.or_kappa3(sc.modkv, sc.modkh, nx, ny, nz, P(self.dzv), P(self.zl), P(self.vx),
                    P(self.vy), P(F(sc.vptemp)), P(F(sc.zi)), P(F(sc.rl)), P(F(sc.ustar)),
                    P(self.hm1), sc.dx, sc.dy, kdx, kdy, sc.rkv0, sc.rkh0, sc.csh, lonlat,
                    P(self.rkver), P(self.rkhor))
        st = L.or_setvset(nx, ny, nz, nc, P(F(sc.tempe)), P(self.rho), P(F(sc.diam)),
                          P(F(sc.rhop)), P(self.psi), P(F(sc.sphe)), sc.modv, P(self.vset),
                          ctypes.byref(self.cd))
        if st != 0:
            raise RuntimeError(f"vsettl did not converge (cell {-st - 1})")
        L.or_meteo_scale(nx, ny, nz, P(self.hm1), P(self.vx), P(self.rho), P(self.rkhor),
                         P(self.rkh1))
        if lonlat:
            L.or_corvver(nx, ny, nz, P(self.hm1), P(F(sc.topg)), sc.dx, sc.dy, P(self.vx),
                         P(self.vy), P(self.vz))
        if getattr(sc, "gravity", None) is not None:
            # meteo.f90:140,145: the radial wind, then the divergence
            self.vy = F(sc.vy).copy(order="F")
            g = sc.gravity
            self.gdiag = np.zeros(9)
            self.gactive = L.or_addradialwind(
                nx, ny, nz, P(self.vx), P(self.vy), P(self.hm1), sc.dx, sc.dy, P(self.zl),
                P(gravity_prm(g)), self.time - g["beg_time"], mass_flow_rate(sc.tmrat),
                float(np.max(sc.zs)), P(self.gdiag))
            self.divu = np.zeros((nx, ny, nz), order="F")
            L.or_setdivu(nx, ny, nz, P(self.vx), P(self.vy), P(self.vz), P(self.hm1), sc.dx,
                         sc.dy, P(self.dzv), P(self.divu))
        dtc = ctypes.c_double()
        self.dt = L.or_finddt(nx, ny, nz, nc, P(self.dzv), sc.dx, sc.dy, P(self.vx),
                              P(self.vy), P(self.vz), P(self.rkh1), P(self.rkhor),
                              P(self.rkver), P(self.vset), ctypes.byref(dtc))
        self.dtc = dtc.value
        return self.dt

    def step(self):
        L, sc = lib(), self.sc
        self.time += self.dt
        self.iiter += 1
        self.erumass += sc.flowrate * self.dt
        L.or_setsrc(P(self.c), self.nx, self.ny, self.nz, self.nc, sc.nsou, P(self.isrc),
                    P(self.jsrc), P(self.ksrc), P(self.tmrat), P(self.dzv), P(self.zl),
                    P(self.hm1), sc.dx, sc.dy, self.dt)
        om = ctypes.c_double(self.outmass)
        L.or_sizloop(P(self.c), P(self.cdry), self.nx, self.ny, self.nz, self.nc, self.iiter,
                     self.dt, P(self.vx), P(self.vy), P(self.vz), P(self.vset), P(self.vdep),
                     P(self.rkver), P(self.rkhor), P(self.rkh1), P(self.rho), P(self.dzv),
                     sc.dx, sc.dy, ctypes.byref(om), self.nthreads, P(self.cwet),
                     P(self._diam), P(self.zl), P(self._zi),
                     None if self._prate is None else P(self._prate))
        self.outmass = om.value
        if getattr(sc, "gravity", None) is not None and log_time(self.iiter, self.time, self.dt):
            # endstep.f90:53-71: writim's cmass3d, then divcorr
            b = self.budget()
            self.corrdiv = L.or_divcorr(P(self.c), self.nx, self.ny, self.nz, self.nc,
                                        P(self.divu), self.dt, self.erumass, b["massvol"],
                                        b["massdep"], self.outmass)

    def products(self):
        """Output products (printres_nc.f90:150-241): load (nx,ny,0:nc),
        deposit thickness (nx,ny), z-cumulative concentration (nx,ny)."""
        nx, ny, nc = self.nx, self.ny, self.nc
        load = np.zeros((nx, ny, nc + 1), order="F")
        thick = np.zeros((nx, ny), order="F")
        colmass = np.zeros((nx, ny), order="F")
        lib().or_output_products(nx, ny, self.nz, nc, P(self.c), P(self.cdry), P(self.cwet),
                                 P(self.hm1), P(F(self.sc.rhop)), P(self.dzv), P(load),
                                 P(thick), P(colmass))
        return load, thick, colmass

    def budget(self):
        """getc0 + cload + cmass3d (endstep.f90:35,80-106; writim.f90:18)."""
        L = lib()
        nx, ny, nc = self.nx, self.ny, self.nc
        L.or_getc0(P(self.cdry), P(self.cwet), nx, ny, nc)
        cload = np.asfortranarray(self.cdry + self.cwet)
        out = [ctypes.c_double() for _ in range(4)]
        L.or_cmass3d(P(self.c), P(cload), P(self.cdry), P(self.cwet), nx, ny, self.nz, nc,
                     self.sc.dx, self.sc.dy, P(self.dzv), *[ctypes.byref(o) for o in out])
        self.cload = cload
        return dict(massvol=out[0].value, massdep=out[1].value, massdry=out[2].value,
                    masswet=out[3].value, outmass=self.outmass, erumass=self.erumass)


class RefRun(_RunBase):
    """Driver over the flang-compiled reference routines (oracle/_ref).
    The reference keeps its state in Fortran module variables, so only one
    RefRun is live per process; the SAVEd vsettl drag coefficient persists
    for the life of the process (setvset.f90:89)."""

    _L = None
    PATH = REF_PATH

    @classmethod
    def available(cls):
        return os.path.exists(cls.PATH)

    @classmethod
    def L(cls):
        if cls._L is None:
            L = ctypes.CDLL(cls.PATH)
            d, i, dp, ip = ctypes.c_double, ctypes.c_int, _dp, _ip
            L.ref_alloc.argtypes = [i, i, i, i, i]
            L.ref_set_grid.argtypes = [d, d, dp]
            L.ref_set_models.argtypes = [i, i, i, d, d]
            L.ref_set_classes.argtypes = [dp, dp, dp]
            L.ref_get_psi.argtypes = [dp]
            L.ref_set_meteo.argtypes = [dp] * 11
            L.ref_scale_and_finddt.argtypes = [dp, dp]
            L.ref_set_dt.argtypes = [d]
            L.ref_set_source.argtypes = [ip, ip, ip, dp, d]
            L.ref_budget.argtypes = [dp] * 6
            L.ref_get_state.argtypes = [dp, dp, dp, ip]
            L.ref_set_state.argtypes = [dp, dp]
            L.ref_get_coef.argtypes = [dp] * 6
            L.ref_set_coef.argtypes = [dp] * 9
            L.ref_setvdrydep.argtypes = [dp, dp, dp]
            L.ref_set_wetdep.argtypes = [dp]
            L.ref_get_cwet.argtypes = [dp]
            L.ref_corvver.argtypes = [dp] * 5
            L.ref_set_clock.argtypes = [d, i, d, d]
            L.ref_set_hm1.argtypes = [i, d, d, dp]
            L.ref_get_vz.argtypes = [dp]
            L.ref_get_scaled.argtypes = [dp, dp]
            L.ref_setcut_fl.argtypes = [dp, dp, dp, d, dp]
            L.ref_standard_atm.argtypes = [d, dp, dp]
            L.ref_ter2asl.argtypes = [dp, dp]
            L.ref_set_gravity.argtypes = [dp, d, dp]
            L.ref_get_wind.argtypes = [dp, dp, dp]
            cls._L = L
        return cls._L

    def __init__(self, sc):
        super().__init__(sc)
        L = self.L()
        nx, ny, nz, nc = self.nx, self.ny, self.nz, self.nc
        L.ref_alloc(nx, ny, nz, nc, sc.nsou)
        L.ref_set_grid(sc.dx, sc.dy, P(F(sc.zlayer)))
        lonlat = int(getattr(sc, "lonlat", 0))
        if lonlat or getattr(sc, "hm1_cols", None) is not None:
            self._hm1 = F(sc.hm1)
            kdx, kdy = _khn_spacing(sc)
            L.ref_set_hm1(lonlat, kdx, kdy, P(self._hm1))
        L.ref_set_models(sc.modkv, sc.modkh, sc.modv, sc.rkv0, sc.rkh0)
        L.ref_set_classes(P(F(sc.diam)), P(F(sc.rhop)), P(F(sc.sphe)))
        self.psi = np.zeros(nc)
        L.ref_get_psi(P(self.psi))
        L.ref_set_meteo(P(F(sc.vx)), P(F(sc.vy)), P(F(sc.vz)), P(F(sc.tempe)),
                        P(F(sc.vptemp)), P(F(sc.rho)), P(F(sc.ustar)), P(F(sc.zi)),
                        P(F(sc.rl)), P(F(sc.vdep)), P(F(sc.topg)))
        self._isrc = np.ascontiguousarray(sc.isrc, dtype=np.int32)
        self._jsrc = np.ascontiguousarray(sc.jsrc, dtype=np.int32)
        self._ksrc = np.ascontiguousarray(sc.ksrc, dtype=np.int32)
        self._tm = F(sc.tmrat)
        L.ref_set_source(P(self._isrc), P(self._jsrc), P(self._ksrc), P(self._tm), sc.flowrate)
        if getattr(sc, "prate", None) is not None:
            self._prate = F(sc.prate)
            L.ref_set_wetdep(P(self._prate))
        self._gravity = getattr(sc, "gravity", None) is not None
        if self._gravity:
            self._gprm = np.ascontiguousarray(gravity_prm(sc.gravity)[:8])
            self._zs = np.ascontiguousarray(sc.zs, dtype=np.float64)
            L.ref_set_gravity(P(self._gprm), sc.gravity["beg_time"], P(self._zs))

    def meteo_update(self):
        L = self.L()
        L.ref_kappa3()
        L.ref_setvset()
        dt, dtc = ctypes.c_double(), ctypes.c_double()
        L.ref_scale_and_finddt(ctypes.byref(dt), ctypes.byref(dtc))
        self.dt, self.dtc = dt.value, dtc.value
        nx, ny, nz, nc = self.nx, self.ny, self.nz, self.nc
        self.vx = np.zeros((nx, ny, nz), order="F")
        self.rho = np.zeros_like(self.vx)
        self.rkhor = np.zeros_like(self.vx)
        self.rkh1 = np.zeros_like(self.vx)
        self.rkver = np.zeros_like(self.vx)
        self.vset = np.zeros((nx, ny, nz, nc), order="F")
        L.ref_get_coef(P(self.vx), P(self.rho), P(self.rkhor), P(self.rkh1), P(self.rkver),
                       P(self.vset))
        self.vz = np.zeros_like(self.vx)
        L.ref_get_vz(P(self.vz))
        self.vy = np.zeros_like(self.vx)
        self.divu = np.zeros_like(self.vx)
        vxw = np.zeros_like(self.vx)
        L.ref_get_wind(P(vxw), P(self.vy), P(self.divu))
        return self.dt

    def step(self):
        L = self.L()
        L.ref_step()
        self.time += self.dt
        self.iiter += 1
        if self._gravity and log_time(self.iiter, self.time, self.dt):
            # endstep.f90:53-71: writim's cmass3d, then divcorr
            self.budget()
            L.ref_divcorr()

    def state(self):
        L = self.L()
        nx, ny, nz, nc = self.nx, self.ny, self.nz, self.nc
        c = np.zeros((nx + 2, ny + 2, nz + 2, nc), order="F")
        cdry = np.zeros((nx, ny, nc + 1), order="F")
        t = ctypes.c_double()
        it = ctypes.c_int()
        L.ref_get_state(P(c), P(cdry), ctypes.byref(t), ctypes.byref(it))
        return c, cdry

    def cwet_state(self):
        cwet = np.zeros((self.nx, self.ny, self.nc + 1), order="F")
        self.L().ref_get_cwet(P(cwet))
        return cwet

    def budget(self):
        L = self.L()
        out = [ctypes.c_double() for _ in range(6)]
        L.ref_budget(*[ctypes.byref(o) for o in out])
        keys = ["massvol", "massdep", "massdry", "masswet", "outmass", "erumass"]
        return {k: o.value for k, o in zip(keys, out)}


class DropinRun(RefRun):
    """The reference's own driver-side code (setsrc.f90, sizloop.f90, the
    meteo-stage routines, through the same harness) with setbcc_, advctzc_,
    advctx_, advcty_, diffz_, diffx_, diffy_, accum_, wetdep_, cmass2d_ and
    cmass3d_ resolved from libfall3d_amd.so: the routine-level drop-in as a
    reference build would link it.  Its own module state, separate from
    RefRun's (ctypes loads each library with local symbol scope)."""

    _L = None
    PATH = DROPIN_PATH
