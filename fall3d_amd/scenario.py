"""Synthetic FALL3D scenarios: grid, meteorology, particle classes and source.

The reference reads these from pre-processed files (DBS meteo netCDF, .grn
granulometry, .src source).  Here they are synthesised in-process with the
same recipes the reference's pre-processors use:

* terrain-following grid and dz extension  -- Sources/reagrd.f90:206-220
* ICAO standard atmosphere                 -- Utilities/SetDbs/read_PROF_data.f90:184-231
* potential temperature TP = T(1e5/p)^0.285 -- Utilities/SetDbs/read_PROF_data.f90:83
* Suzuki column, 100 points                -- Utilities/SetSrc/solvesuzuki.f90:25-45
* bilinear/linear node distribution        -- Utilities/SetSrc/getsrc.f90:23-106
* phi-spaced classes, density linear in phi -- Utilities/SetTgsd/setfrac.f90:18-44
* vz(:,:,1) = 0 after the meteo read       -- Sources/reamet.f90:153
* meteo is float32 in the DBS file         -- Utilities/SetDbs/write_DBS_grid.f90:111-126

All meteo is float32-quantised and promoted to float64 so every consumer
(the FP64 and FP32 GPU paths and any CPU check) sees identical inputs.
Arrays are Fortran-ordered (i fastest), as the reference lays them out.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

__all__ = ["Scenario", "standard_atm", "make_classes", "suzuki_profile", "getshape", "suzuki_source", "config_c1",
           "config_c2", "config_small", "config_ensemble_member", "config_uneven",
           "lonlat_grid", "khn_spacing", "R_EARTH"]

KV_CONSTANT, KV_SIMILARITY, KV_SURFACE = 0, 1, 2
KH_CONSTANT, KH_RAMS, KH_CMAQ = 0, 1, 2
VS_ARASTOOPOUR, VS_GANSER, VS_WILSON, VS_DELLINO = 1, 2, 3, 4
R_EARTH = 6356.0e3  # Numeric.f90:14


@dataclasses.dataclass
class Scenario:
    name: str
    nx: int
    ny: int
    nz: int
    nc: int
    dx: float
    dy: float
    zlayer: np.ndarray          # (nz,)
    # particle classes
    diam: np.ndarray            # (nc,) m
    rhop: np.ndarray            # (nc,) kg/m3
    sphe: np.ndarray            # (nc,)
    fc: np.ndarray              # (nc,) mass fractions
    # meteorology (nx,ny,nz) Fortran order, float64 holding float32 values
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    tempe: np.ndarray
    vptemp: np.ndarray
    rho: np.ndarray
    ustar: np.ndarray           # (nx,ny)
    zi: np.ndarray
    rl: np.ndarray
    topg: np.ndarray
    vdep: np.ndarray            # (nx,ny,nc)
    # source
    isrc: np.ndarray            # (nsou,) 1-based node indices
    jsrc: np.ndarray
    ksrc: np.ndarray
    tmrat: np.ndarray           # (nsou,nc) kg/s, Fortran order
    flowrate: float
    # models
    modkv: int = KV_CONSTANT
    modkh: int = KH_CONSTANT
    modv: int = VS_GANSER
    rkv0: float = 500.0
    rkh0: float = 5000.0
    csh: float = 0.2275
    end_time: float = 6 * 3600.0
    # Hm1 per column (nx,ny) for LON-LAT grids (sin of the colatitude,
    # reagrd.f90:247-270); None: UTM, Hm1 = 1
    hm1_cols: Optional[np.ndarray] = None
    # LON-LAT coordinates: kappa3's CMAQ constant and the terrain-slope
    # correction of vz (corvver) follow them
    lonlat: int = 0
    # LON-LAT grid spacing in degrees (reagrd.f90:155-156); dx, dy hold the
    # scaled spacing Rearth*(dlon*pi/180) in m.  0: not given
    dlon: float = 0.0
    dlat: float = 0.0
    # precipitation rate (nx,ny) in mm/h: wet deposition on (wetdep.f90)
    prate: Optional[np.ndarray] = None
    # gravity current (the GRAVITY_CURRENT block, readat.f90:149-200): the
    # fields of f3d_gravity (c_flow_rate, lambda_grav, k_entrain,
    # brunt_vaisala, center_x, center_y, xorigr, yorigr, dx, dy, beg_time;
    # coordinates in the grid's own units) and the source heights zs
    # (nsou,) in m; None: off
    gravity: Optional[dict] = None
    zs: Optional[np.ndarray] = None

    @property
    def dz(self) -> np.ndarray:
        """dz(0:nz), reagrd.f90:216-220."""
        nz = self.nz
        dz = np.empty(nz + 1)
        dz[1:nz] = self.zlayer[1:] - self.zlayer[:-1]
        dz[0] = dz[1]
        dz[nz] = dz[nz - 1]
        return dz

    @property
    def hm1(self) -> np.ndarray:
        if self.hm1_cols is not None:
            return np.asfortranarray(self.hm1_cols, dtype=np.float64)
        return np.ones((self.nx, self.ny), order="F")

    @property
    def nsou(self) -> int:
        return int(self.isrc.size)

    @property
    def updates_per_step(self) -> int:
        return self.nx * self.ny * self.nz * self.nc

    def write_binary(self, path):
        """Raw little-endian stream read by the Fortran driver
        (fall3d_amd/fortran/f3d_driver.f90)."""
        i4 = lambda *v: np.asarray(v, dtype="<i4").tobytes()  # noqa: E731
        f8 = lambda a: np.asarray(a, dtype="<f8").ravel(order="F").tobytes()  # noqa: E731
        with open(path, "wb") as fh:
            fh.write(i4(self.nx, self.ny, self.nz, self.nc, self.nsou, self.modv, self.modkv,
                        self.modkh))
            fh.write(f8([self.dx, self.dy, self.rkv0, self.rkh0, self.csh, self.flowrate,
                         self.end_time]))
            for a in (self.zlayer, self.diam, self.rhop, self.sphe, self.vx, self.vy, self.vz,
                      self.tempe, self.vptemp, self.rho, self.ustar, self.zi, self.rl, self.vdep):
                fh.write(f8(a))
            if self.nsou:
                for a in (self.isrc, self.jsrc, self.ksrc):
                    fh.write(np.asarray(a, dtype="<i4").tobytes())
                fh.write(f8(self.tmrat))


def _q32(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float32).astype(np.float64))


def standard_atm(z):
    """ICAO standard atmosphere, three layers (read_PROF_data.f90:184-231)."""
    z = np.asarray(z, dtype=np.float64)
    goo, Roo, Moo = 9.80665, 8.31432, 0.0289644
    Too, poo, Loo = 288.15, 101325.0, -0.0065
    zo1, To1, po1 = 11000.0, 216.65, 22632.0
    zo2, To2, po2, Lo2 = 20000.0, 216.65, 5474.0, 0.001
    T = np.where(z < zo1, Too + Loo * z, np.where(z < zo2, To1, To2 + Lo2 * (z - zo2)))
    p1 = poo * (Too / T) ** (goo * Moo / (Roo * Loo))
    p2 = po1 * np.exp(-goo * Moo * (z - zo1) / (Roo * To1))
    p3 = po2 * (To2 / T) ** (goo * Moo / (Roo * Lo2))
    P = np.where(z < zo1, p1, np.where(z < zo2, p2, p3))
    rho = P * Moo / (Roo * T)
    return P, T, rho


def make_classes(nc, dmin_m, dmax_m, rho_fine=2500.0, rho_coarse=1000.0, sphe=0.9,
                 phi_mean=4.0, phi_sigma=2.5):
    """nc classes log-spaced in diameter; density linear in phi = -log2(d/mm)
    between the coarsest (rho_coarse) and finest (rho_fine) class, as SetTgsd
    does (setfrac.f90:26-30); Gaussian mass fractions in phi."""
    d = np.exp(np.linspace(math.log(dmin_m), math.log(dmax_m), nc))
    phi = -np.log2(d * 1e3)
    if nc > 1:
        w = (phi - phi.min()) / (phi.max() - phi.min())
    else:
        w = np.zeros(1)
    rhop = rho_coarse + w * (rho_fine - rho_coarse)
    fc = np.exp(-0.5 * ((phi - phi_mean) / phi_sigma) ** 2)
    fc = fc / fc.sum()
    return d, rhop, np.full(nc, sphe), fc


def suzuki_profile(H, mer, A, L, ns=100):
    """Suzuki column (solvesuzuki.f90:25-45): ns points at z = is*H/ns,
    is = 1..ns, with S(z) = ((1 - z/H) exp(A (z/H - 1)))^L normalised so
    that sum(S) = mer.  Returns (z, S)."""
    dzs = H / ns
    z = dzs * np.arange(1, ns + 1)
    S = ((1.0 - z / H) * np.exp(A * (z / H - 1.0))) ** L
    return z, mer * S / S.sum()


def getshape(s, t):
    """Bilinear shape functions of the 4-node cell at (s,t) in (-1,1)^2,
    nodes ordered (ix,iy), (ix+1,iy), (ix+1,iy+1), (ix,iy+1) (getsrc.f90:
    getshape)."""
    return [0.25 * (1 - s) * (1 - t), 0.25 * (1 + s) * (1 - t),
            0.25 * (1 + s) * (1 + t), 0.25 * (1 - s) * (1 + t)]


def suzuki_source(nx, ny, dx, dy, zlayer, x0, y0, H, mer, A, L, fc, ns=100, topg=None):
    """Suzuki column (solvesuzuki.f90:25-45) of ns points above (x0,y0),
    distributed to nodes by bilinear shape functions in (x,y) and linear
    weights in z (getsrc.f90:23-106).  Returns 1-based (i,j,k) and tmrat.
    A point above the top level is an error, as in getsrc (runend)."""
    nz = len(zlayer)
    z, S = suzuki_profile(H, mer, A, L, ns)
    ix = int(math.floor(x0 / dx)) + 1
    iy = int(math.floor(y0 / dy)) + 1
    ix = min(max(ix, 1), nx - 1)
    iy = min(max(iy, 1), ny - 1)
    s = 2.0 * ((x0 - (ix - 1) * dx) / dx) - 1.0
    t = 2.0 * ((y0 - (iy - 1) * dy) / dy) - 1.0
    shape = getshape(s, t)
    corners = [(ix, iy), (ix + 1, iy), (ix + 1, iy + 1), (ix, iy + 1)]
    src = {}
    for q, (ci, cj) in enumerate(corners):
        zp = 0.0 if topg is None else float(topg[ci - 1, cj - 1])
        for p in range(ns):
            zz = z[p] - zp
            if zz < zlayer[0]:
                k, w = 1, 0.0
            elif zz >= zlayer[nz - 1]:
                raise ValueError(f"source point at {zz:.1f} m above the top level "
                                 f"{zlayer[nz - 1]:.1f} m")
            else:
                k = int(np.searchsorted(zlayer, zz, side="right"))
                w = (zz - zlayer[k - 1]) / (zlayer[k] - zlayer[k - 1])
            for kk, ww in ((k, 1.0 - w), (k + 1, w)):
                if ww > 0.0 and shape[q] > 0.0:
                    src[(ci, cj, kk)] = src.get((ci, cj, kk), 0.0) + ww * shape[q] * S[p]
    keys = sorted(src.keys(), key=lambda t3: (t3[2], t3[1], t3[0]))
    isrc = np.array([k3[0] for k3 in keys], dtype=np.int32)
    jsrc = np.array([k3[1] for k3 in keys], dtype=np.int32)
    ksrc = np.array([k3[2] for k3 in keys], dtype=np.int32)
    tot = np.array([src[k3] for k3 in keys])
    tmrat = np.asfortranarray(np.outer(tot, fc))
    return isrc, jsrc, ksrc, tmrat


def _thermo(nx, ny, zlayer, topg):
    z = topg[:, :, None] + zlayer[None, None, :]
    P, T, rho = standard_atm(z)
    tp = T * (1e5 / P) ** 0.285
    return T, tp, rho


def _finish(name, nx, ny, nz, dx, dy, zlayer, classes, u, v, w, topg, src, flowrate,
            end_time, **models):
    d, rhop, sphe, fc = classes
    nc = len(d)
    T, tp, rho = _thermo(nx, ny, zlayer, topg)
    w = np.array(w, copy=True)
    w[:, :, 0] = 0.0                              # reamet.f90:153
    # neutral surface layer: u* from the 10 m wind by the log law, z0 = 0.1 m
    u10 = np.hypot(u[:, :, 0], v[:, :, 0])
    ustar = 0.4 * np.maximum(u10, 0.1) / math.log(10.0 / 0.1)
    isrc, jsrc, ksrc, tmrat = src
    return Scenario(
        name=name, nx=nx, ny=ny, nz=nz, nc=nc, dx=float(dx), dy=float(dy),
        zlayer=np.asarray(zlayer, dtype=np.float64),
        diam=d, rhop=rhop, sphe=sphe, fc=fc,
        vx=_q32(u), vy=_q32(v), vz=_q32(w), tempe=_q32(T), vptemp=_q32(tp), rho=_q32(rho),
        ustar=_q32(ustar), zi=_q32(np.full((nx, ny), 1000.0)), rl=_q32(np.zeros((nx, ny))),
        topg=_q32(topg), vdep=np.zeros((nx, ny, nc), order="F"),
        isrc=isrc, jsrc=jsrc, ksrc=ksrc, tmrat=tmrat, flowrate=float(flowrate),
        end_time=float(end_time), **models)


def _log_wind(nx, ny, zlayer, u10, v_ratio, z0=0.1):
    zp = np.maximum(zlayer, 10.0)
    prof = u10 * np.log(zp / z0) / math.log(10.0 / z0)
    u = np.broadcast_to(prof[None, None, :], (nx, ny, len(zlayer))).copy(order="F")
    return u, v_ratio * u


def _fourier_wind(rng, nx, ny, zlayer, dx, dy, nmodes, rms_h, rms_v):
    """Seeded random Fourier modes: nmodes separable modes per component with
    integer wavenumbers 1..nmodes on each axis, rescaled to the target rms."""
    Lx, Ly, H = nx * dx, ny * dy, max(zlayer[-1], 1.0)
    x = np.arange(nx) * dx
    y = np.arange(ny) * dy
    z = np.asarray(zlayer)

    def field(rms):
        X = np.empty((nmodes, nx))
        Y = np.empty((nmodes, ny))
        Z = np.empty((nmodes, len(z)))
        for m in range(nmodes):
            kx, ky, kz = rng.integers(1, nmodes + 1, size=3)
            px, py, pz = rng.uniform(0, 2 * np.pi, size=3)
            a = rng.standard_normal()
            X[m] = a * np.cos(2 * np.pi * kx * x / Lx + px)
            Y[m] = np.cos(2 * np.pi * ky * y / Ly + py)
            Z[m] = np.cos(np.pi * kz * z / H + pz)
        # sum over modes of a*fx(x)*fy(y)*fz(z), as one contraction
        XY = X[:, :, None] * Y[:, None, :]
        f = np.asfortranarray(np.tensordot(XY, Z, axes=([0], [0])))
        r = math.sqrt(float(np.mean(f * f)))
        return f * (rms / r if r > 0 else 0.0)

    u = field(rms_h)
    v = field(rms_h)
    w = field(rms_v)
    return u, v, w


def lonlat_grid(sc: Scenario, lat_min: float, dlon: float, dlat: float) -> Scenario:
    """sc on a LON-LAT grid of dlon x dlat degrees whose first row lies at
    latitude lat_min: dX1 = Rearth*(dx*pi/180) (reagrd.f90:155-158) and
    Hm1(:,j) = sin(colat) with lat = lat_min + (j-1)*dy (reagrd.f90:247-262);
    kappa3's CMAQ constant and corvver follow the coordinate system."""
    dx = R_EARTH * (dlon * math.pi / 180.0)
    dy = R_EARTH * (dlat * math.pi / 180.0)
    lat = lat_min + np.arange(sc.ny) * dlat
    colat = (90.0 - lat) * math.pi / 180.0
    hm1 = np.asfortranarray(np.repeat(np.sin(colat)[None, :], sc.nx, axis=0))
    return dataclasses.replace(sc, dx=dx, dy=dy, dlon=dlon, dlat=dlat, hm1_cols=hm1, lonlat=1)


def khn_spacing(sc: Scenario):
    """The grid spacing kappa3's CMAQ Khn takes (kappa3.f90:226): dx, dy in m
    on UTM grids, in degrees on LON-LAT grids (given, or dX1 / (Rearth*pi/180)
    as the library derives it)."""
    if not sc.lonlat:
        return sc.dx, sc.dy
    m_per_deg = R_EARTH * (math.pi / 180.0)
    return (sc.dlon if sc.dlon > 0 else sc.dx / m_per_deg,
            sc.dlat if sc.dlat > 0 else sc.dy / m_per_deg)


def config_c1():
    """60x60x20, 5 km, flat, top 20 km, 6 classes 2 um - 2 mm, log wind
    u=10, v=3 m/s at 10 m, Suzuki A=4 L=5 H=10 km MER 1e6 kg/s, 6 h."""
    nx = ny = 60
    nz = 20
    dx = dy = 5000.0
    zl = np.array([(k * 20000.0) / (nz - 1) for k in range(nz)])
    classes = make_classes(6, 2e-6, 2e-3)
    u, v = _log_wind(nx, ny, zl, 10.0, 0.3)
    topg = np.zeros((nx, ny), order="F")
    src = suzuki_source(nx, ny, dx, dy, zl, (nx - 1) * dx / 2, (ny - 1) * dy / 2,
                        10000.0, 1e6, 4.0, 5.0, classes[3])
    return _finish("c1_60x60x20x6", nx, ny, nz, dx, dy, zl, classes, u, v,
                   np.zeros_like(u), topg, src, 1e6, 6 * 3600.0,
                   modkv=KV_CONSTANT, modkh=KH_CONSTANT, rkv0=500.0, rkh0=5000.0)


def config_c2(nx=400, ny=400, nz=40, nc=20, seed=0xF3D0002):
    """Precision form: 400x400x40, 20 classes 2 um - 4 mm, Fourier-mode wind
    (8 modes, 15 m/s rms horizontal, 0.1 m/s vertical), sinusoidal terrain up
    to 2 km, 12 h."""
    rng = np.random.default_rng(seed)
    dx = dy = 5000.0
    zl = np.array([(k * 20000.0) / (nz - 1) for k in range(nz)])
    classes = make_classes(nc, 2e-6, 4e-3)
    u, v, w = _fourier_wind(rng, nx, ny, zl, dx, dy, 8, 15.0, 0.1)
    x = np.arange(nx) * dx
    y = np.arange(ny) * dy
    topg = np.asfortranarray(1000.0 * (1.0 + np.sin(2 * np.pi * x / (nx * dx))[:, None]
                                       * np.sin(2 * np.pi * y / (ny * dy))[None, :]))
    src = suzuki_source(nx, ny, dx, dy, zl, (nx - 1) * dx / 2, (ny - 1) * dy / 2,
                        10000.0, 1e6, 4.0, 5.0, classes[3])
    return _finish(f"c2_{nx}x{ny}x{nz}x{nc}", nx, ny, nz, dx, dy, zl, classes, u, v, w, topg,
                   src, 1e6, 12 * 3600.0, modkv=KV_CONSTANT, modkh=KH_RAMS, rkv0=500.0)


def config_small(nx=9, ny=7, nz=6, nc=3, seed=1, modkv=KV_CONSTANT, modkh=KH_RAMS,
                 modv=VS_GANSER, rms_h=12.0, rms_v=0.3, mean_u=0.0, mean_v=0.0):
    """Tiny seeded scenario for unit and composed-step parity tests."""
    rng = np.random.default_rng(seed)
    dx, dy = 4000.0, 3000.0
    zl = np.array([(k * 8000.0) / (nz - 1) for k in range(nz)])
    classes = make_classes(nc, 5e-6, 1e-3)
    u, v, w = _fourier_wind(rng, nx, ny, zl, dx, dy, 4, rms_h, rms_v)
    u += mean_u
    v += mean_v
    topg = np.zeros((nx, ny), order="F")
    src = suzuki_source(nx, ny, dx, dy, zl, (nx - 1) * dx / 2, (ny - 1) * dy / 2,
                        zl[-1] * 0.6, 1e5, 4.0, 5.0, classes[3])
    return _finish(f"small_{nx}x{ny}x{nz}x{nc}_s{seed}", nx, ny, nz, dx, dy, zl, classes,
                   u, v, w, topg, src, 1e5, 3600.0, modkv=modkv, modkh=modkh, modv=modv,
                   rkv0=50.0, rkh0=2000.0)


def with_gravity_current(sc: Scenario, mer=1e7, H=None, beg_time=-1800.0, c_flow_rate=430.0,
                         lambda_grav=0.2, k_entrain=0.1, brunt_vaisala=0.02):
    """The scenario with a Suzuki column of mass rate mer at the domain
    centre and the gravity current of its umbrella on (UTM grid, origin 0:
    the reference places node i at xorigr + i*dx, addradialwind.f90:159).
    The source heights zs are the levels of the source nodes."""
    H = sc.zlayer[-1] * 0.6 if H is None else H
    x0, y0 = (sc.nx - 1) * sc.dx / 2, (sc.ny - 1) * sc.dy / 2
    src = suzuki_source(sc.nx, sc.ny, sc.dx, sc.dy, sc.zlayer, x0, y0, H, mer, 4.0, 5.0, sc.fc)
    zs = np.asarray(sc.zlayer)[src[2] - 1].astype(np.float64)
    grav = dict(c_flow_rate=c_flow_rate, lambda_grav=lambda_grav, k_entrain=k_entrain,
                brunt_vaisala=brunt_vaisala, center_x=x0, center_y=y0, xorigr=0.0, yorigr=0.0,
                dx=sc.dx, dy=sc.dy, beg_time=beg_time)
    return dataclasses.replace(sc, name=f"{sc.name}_gc", isrc=src[0], jsrc=src[1], ksrc=src[2],
                               tmrat=src[3], flowrate=float(mer), gravity=grav, zs=zs)


def config_ensemble_member(base: Scenario, seed: int):
    """Ensemble member on a shared grid and meteo: column height ~ U(6,18) km,
    log10 MER ~ N(6,0.7), Suzuki A ~ U(2,6), vent offset ~ U(+-10 km) per axis."""
    rng = np.random.default_rng(seed)
    H = rng.uniform(6000.0, 18000.0)
    mer = 10.0 ** rng.normal(6.0, 0.7)
    A = rng.uniform(2.0, 6.0)
    ox, oy = rng.uniform(-10000.0, 10000.0, size=2)
    x0 = (base.nx - 1) * base.dx / 2 + ox
    y0 = (base.ny - 1) * base.dy / 2 + oy
    src = suzuki_source(base.nx, base.ny, base.dx, base.dy, base.zlayer, x0, y0,
                        min(H, base.zlayer[-1] * 0.95), mer, A, 5.0, base.fc)
    return dataclasses.replace(base, name=f"{base.name}_m{seed}", isrc=src[0], jsrc=src[1],
                               ksrc=src[2], tmrat=src[3], flowrate=float(mer))


def config_uneven(seed: int, nmin=80, nmax=320, zmin=16, zmax=48, cmin=4, cmax=24):
    """Uneven-batch scenario: nx=ny ~ U{nmin..nmax}, nz ~ U{zmin..zmax},
    nc ~ U{cmin..cmax}, own Fourier meteo and source."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(nmin, nmax + 1))
    nz = int(rng.integers(zmin, zmax + 1))
    nc = int(rng.integers(cmin, cmax + 1))
    s = config_c2(nx=n, ny=n, nz=nz, nc=nc, seed=int(rng.integers(1, 2**31)))
    return dataclasses.replace(s, name=f"uneven_{seed}_{n}x{n}x{nz}x{nc}", end_time=6 * 3600.0)
