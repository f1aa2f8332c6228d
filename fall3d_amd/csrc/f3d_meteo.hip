// Coefficient stage of a meteo read, FP64 on the device (meteo.f90:27-149):
// kappa3 (eddy diffusivities), setvset (settling velocities), the Hm1
// scaling and finddt (the CFL step).  Built with -ffp-contract=off: the
// arithmetic is the reference's operation for operation; transcendental
// calls (pow) are the device library's and may differ from the host libm
// in the last bit.
#include "f3d_internal.h"

namespace f3d {
namespace {

__device__ __forceinline__ double fmax_f(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fmin_f(double a, double b) { return a < b ? a : b; }

constexpr double kVkarm = 0.4, kBeta1 = 9.2, kBeta2 = 13.0, kGi = 9.81, kAlfa = 0.28;
constexpr double kRlamb = 30.0, kRkhmin = 1e2, kRkvmin = 1.0, kHinf = 5.0, kHsup = 5e3;
constexpr double kKhf = 8000.0;

// SIMILARITY gradient terms at (i,j,k), kappa3.f90:65-70
__device__ __forceinline__ void similarity_terms(int nx, int ny, int i, int j, int k,
                                                 const double *dz, const double *vx,
                                                 const double *vy, const double *vptemp,
                                                 double &dvdz2, double &rich)
{
    const double den = dz[k] + dz[k - 1];
    const double a = (vx[midx(nx, ny, i, j, k + 1)] - vx[midx(nx, ny, i, j, k - 1)]) / den;
    const double b = (vy[midx(nx, ny, i, j, k + 1)] - vy[midx(nx, ny, i, j, k - 1)]) / den;
    dvdz2 = a * a + b * b;
    dvdz2 = fmax_f(dvdz2, 1e-10);
    rich = kGi * ((vptemp[midx(nx, ny, i, j, k + 1)] - vptemp[midx(nx, ny, i, j, k - 1)]) / den) /
           (vptemp[midx(nx, ny, i, j, k)] * dvdz2);
}

// kappa3.f90:50-236, one thread per cell; khn_dx, khn_dy: the grid spacing
// of CMAQ's Khn (kappa3.f90:226), degrees on LON-LAT grids
__global__ void k_kappa3(Grid g, int modkv, int modkh, const double *dz, const double *zlayer,
                         const double *vx, const double *vy, const double *vptemp,
                         const double *zi, const double *rl, const double *ustar,
                         const double *hm1, const double *kmh, double khn_dx, double khn_dy,
                         double rkv0, double rkh0, double csh, int lonlat, double *rkver,
                         double *rkhor)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.N) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int i = (int)(t % nx) + 1;
    const int j = (int)((t / nx) % ny) + 1;
    const int k = (int)(t / g.NXY) + 1;
    const long p2 = (i - 1) + (long)nx * (j - 1);

    if (modkv == 0) {
        rkver[t] = rkv0;
    } else if (modkv == 1) {
        if (k >= 2) {
            double dvdz2 = 0.0, rich = 0.0;
            if (k < nz)
                similarity_terms(nx, ny, i, j, k, dz, vx, vy, vptemp, dvdz2, rich);
            else if (nz >= 3)  // stale values of the last k < nz iterate, (nx,ny,nz-1)
                similarity_terms(nx, ny, nx, ny, nz - 1, dz, vx, vy, vptemp, dvdz2, rich);
            const double h = fmin_f(fmax_f(zi[p2], kHinf), kHsup);
            const double zl = zlayer[k - 1];
            const double zeta = zl / h;
            double r;
            if (zeta < 1.0) {
                if (rl[p2] >= 0.0)
                    r = kVkarm * ustar[p2] * zl * (1.0 - zeta) / (1.0 + kBeta1 * zl * rl[p2]);
                else
                    r = kVkarm * ustar[p2] * zl * (1.0 - zeta) * sqrt(1.0 - kBeta2 * zl * rl[p2]);
            } else {
                const double lc = 1.0 / ((1.0 / (kVkarm * zl)) + (1.0 / kRlamb));
                if (rich >= 0.0)
                    r = lc * lc * sqrt(dvdz2) / (1.0 + 10.0 * rich * (1.0 + 8.0 * rich));
                else
                    r = lc * lc * sqrt(dvdz2) * sqrt(1.0 - 18.0 * rich);
                r = fmax_f(r, kRkvmin);
            }
            rkver[t] = r;
        }
    } else if (modkv == 2) {
        const double gh = 11.6, betah = 7.8, prt = 0.95;
        const double zl = zlayer[k - 1];
        double zeta = zl * rl[p2];
        double r;
        if (zeta < 0.0)
            r = kVkarm * ustar[p2] * zl * sqrt(1.0 - gh * zeta) / prt;
        else if (zeta > 0.0)
            r = kVkarm * ustar[p2] * zl / (prt + betah * zeta);
        else
            r = kVkarm * ustar[p2] * zl / prt;
        const double h = fmin_f(fmax_f(zi[p2], kHinf), kHsup);
        zeta = zl / h;
        if (zeta > 1.0) zeta = 1.0;
        rkver[t] = r * (1.0 - zeta);
    }

    if (modkh == 0) {
        rkhor[t] = rkh0;
    } else {
        const double dXX = hm1[p2] * g.dx;
        const double dX2 = g.dy;
        double dvxdx, dvydx, dvxdy, dvydy;
        if (i == 1) {
            dvxdx = (vx[midx(nx, ny, 2, j, k)] - vx[midx(nx, ny, 1, j, k)]) / dXX;
            dvydx = (vy[midx(nx, ny, 2, j, k)] - vy[midx(nx, ny, 1, j, k)]) / dXX;
        } else if (i == nx) {
            dvxdx = (vx[midx(nx, ny, nx, j, k)] - vx[midx(nx, ny, nx - 1, j, k)]) / dXX;
            dvydx = (vy[midx(nx, ny, nx, j, k)] - vy[midx(nx, ny, nx - 1, j, k)]) / dXX;
        } else {
            dvxdx = (vx[midx(nx, ny, i + 1, j, k)] - vx[midx(nx, ny, i - 1, j, k)]) / (2.0 * dXX);
            dvydx = (vy[midx(nx, ny, i + 1, j, k)] - vy[midx(nx, ny, i - 1, j, k)]) / (2.0 * dXX);
        }
        if (j == 1) {
            dvxdy = (vx[midx(nx, ny, i, 2, k)] - vx[midx(nx, ny, i, 1, k)]) / dX2;
            dvydy = (vy[midx(nx, ny, i, 2, k)] - vy[midx(nx, ny, i, 1, k)]) / dX2;
        } else if (j == ny) {
            dvxdy = (vx[midx(nx, ny, i, ny, k)] - vx[midx(nx, ny, i, ny - 1, k)]) / dX2;
            dvydy = (vy[midx(nx, ny, i, ny, k)] - vy[midx(nx, ny, i, ny - 1, k)]) / dX2;
        } else {
            dvxdy = (vx[midx(nx, ny, i, j + 1, k)] - vx[midx(nx, ny, i, j - 1, k)]) / (2.0 * dX2);
            dvydy = (vy[midx(nx, ny, i, j + 1, k)] - vy[midx(nx, ny, i, j - 1, k)]) / (2.0 * dX2);
        }
        double r;
        if (modkh == 1) {
            // Kmh = 0.075*((dXX*dX2)**(2/3)) precomputed per column on the host
            const double s = dvxdy + dvydx;
            r = csh * csh * dXX * dX2 * sqrt(s * s + 2.0 * (dvxdx * dvxdx + dvydy * dvydy));
            r = fmax_f(r, kmh[p2]);
            r = fmax_f(r, kRkhmin);
        } else {
            const double Khn0 = lonlat ? 3.6e-2 : 4000.0;
            const double a = dvxdx - dvydy, b = dvydx + dvxdy;
            const double Kht = kAlfa * kAlfa * dXX * dX2 * sqrt(a * a + b * b);
            const double Khn = kKhf * (Khn0 / hm1[p2] / khn_dx) * (Khn0 / khn_dy);
            r = (1.0 / Kht) + (1.0 / Khn);
            r = fmax_f(1.0 / r, kRkhmin);
        }
        rkhor[t] = r;
    }
}

// Sutherland viscosity per cell (setvset.f90:35-37); the reference evaluates
// it inside the class loop, with the same value for every class.  Tair0 is
// the default-REAL(4) literal 291.15 (setvset.f90:23).
__global__ void k_visc(long n, const double *tempe, double *visc)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double visc0 = 1.827e-5;
    const double Tair0 = (double)291.15f;
    const double T = tempe[t];
    visc[t] = visc0 * ((Tair0 + 120.0) / (T + 120.0)) * (pow(T / Tair0, 1.5));
}

// vsettl (setvset.f90:65-181) continuing from drag coefficient cd.
// Returns 1 on convergence, 0 otherwise.
__device__ int vsettl(double diam, double rhop, double rhoa, double visc, int model, double psi,
                      double rk1, double rk2, double &cd, double &vout)
{
    const double gi = 9.81, eps = 1e-3;
    double vold = 1e5, v = 0.0;
    if (model == 4) {
        v = pow((diam * diam * diam * gi * (rhop - rhoa) * rhoa * (pow(psi, 1.6))) / (visc * visc),
                0.5206);
        vout = (1.2065 * visc * v) / (diam * rhoa);
        return 1;
    }
    for (int it = 1; it <= 100; ++it) {
        v = sqrt(4.0 * gi * diam * rhop / (3.0 * cd * rhoa));
        const double rey = rhoa * v * diam / visc;
        if (model == 2) {
            cd = 24.0 / (rey * rk1) * (1.0 + 0.1118 * pow(rey * rk1 * rk2, 0.6567)) +
                 0.4305 * rk2 / (1.0 + 3305.0 / (rey * rk1 * rk2));
        } else if (model == 1) {
            if (rey <= 988.947)
                cd = 24.0 / rey * (1.0 + 0.15 * pow(rey, 0.687));
            else
                cd = 0.44;
        } else {
            if (rey <= 100.0) {
                cd = 24.0 / rey * pow(psi, -0.828) + 2.0 * sqrt(1.0 - psi);
            } else if (rey > 100.0 && rey < 1000.0) {
                const double cd100 = 0.24 * pow(psi, -0.828) + 2.0 * sqrt(1.0 - psi);
                const double a = (1.0 - cd100) / 900.0;
                const double b = 1.0 - 1000.0 * a;
                cd = a * rey + b;
            } else {
                cd = 1.0;
            }
        }
        if (it > 1 && fabs(v - vold) <= eps) {
            vout = sqrt(4.0 * gi * diam * rhop / (3.0 * cd * rhoa));
            return 1;
        }
        vold = v;
    }
    vout = v;
    return 0;
}

// setvset.f90:25-60.  The reference threads one SAVEd drag coefficient
// through every cell in the order (class, i, j, k) (setvset.f90:25-33,89).
// Chain position p = ((q*nx + i-1)*ny + j-1)*nz + k-1.  Thread t owns
// positions [t*L, (t+1)*L) and first replays the `warm` positions before
// them from cd = 1: the map cd_in -> cd_out of one cell contracts by at
// least 4x (two or more Picard updates, each halving the error of log cd),
// so after the warm-up the carried cd equals the serial chain's to the last
// bit.  The chunk that starts the chain uses the carried state instead.
__global__ void k_setvset(Grid g, int np, int modv, const met::ClassConst *cls,
                          const double *visc, const double *rho, double *vset, double *cd_io,
                          int *status, long P, int L, int warm)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long p0 = t * (long)L;
    if (p0 >= P) return;
    const long p1 = min(p0 + (long)L, P);
    long p = p0 - warm;
    double cd;
    if (p <= 0) {
        p = 0;
        cd = cd_io[0];
    } else {
        cd = 1.0;
    }
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    for (; p < p1; ++p) {
        const int k = (int)(p % nz) + 1;
        long r = p / nz;
        const int j = (int)(r % ny) + 1;
        r /= ny;
        const int i = (int)(r % nx) + 1;
        const int q = (int)(r / nx);
        const long m = midx(nx, ny, i, j, k);
        const met::ClassConst c = cls[q];
        double v;
        const int ok = vsettl(c.diam, c.rhop, rho[m], visc[m], modv, c.psi, c.rk1, c.rk2, cd, v);
        if (p >= p0) {
            vset[(long)q * g.N + m] = -v;  // setvset.f90:60
            if (!ok) atomicOr(status, 1);
        }
    }
    if (p1 == P) cd_io[1] = cd;
}

// meteo.f90:108-112
__global__ void k_scale(Grid g, const double *hm1, double *vx, double *rho, const double *rkhor,
                        double *rkh1)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.N) return;
    const double h = hm1[t % g.NXY];
    vx[t] = vx[t] / h;
    rho[t] = rho[t] * h;
    rkh1[t] = rkhor[t] / (h * h);
}

// finddt.f90:26-48: max over cells and classes; max is exact, so the
// reduction order cannot change the result.
__global__ void k_finddt(Grid g, const double *dz, double dX1, double dX2, const double *vx,
                         const double *vy, const double *vz, const double *rkh1,
                         const double *rkhor, const double *rkver, const double *vset,
                         unsigned long long *bits)
{
    __shared__ double sh[256];
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    double m = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < g.N;
         t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t / g.NXY) + 1;
        const double ddz = fmin_f(dz[k - 1], dz[k]);
        const double base = (2.0 * rkh1[t] / (dX1 * dX1)) + (fabs(vx[t]) / dX1) +
                            (2.0 * rkhor[t] / (dX2 * dX2)) + (fabs(vy[t]) / dX2) +
                            (2.0 * rkver[t] / (ddz * ddz));
        for (int q = 0; q < g.nc; ++q) {
            const double *vs = vset + (long)q * g.N;
            double dvsetdz;
            if (k == 1)
                dvsetdz = (vs[t + g.NXY] - vs[t]) / (dz[1]);
            else if (k == nz)
                dvsetdz = (vs[t] - vs[t - g.NXY]) / (dz[nz - 1]);
            else
                dvsetdz = (vs[t + g.NXY] - vs[t - g.NXY]) / (dz[k] + dz[k - 1]);
            const double r = base + (fabs((vz[t] + vs[t])) / ddz) + (fabs(dvsetdz));
            m = fmax_f(m, r);
        }
        (void)nx;
        (void)ny;
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax_f(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(bits, (unsigned long long)__double_as_longlong(sh[0]));
}

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// setvdrydep.f90:54-264 (Feng 2008 dry deposition), one thread per column:
// the aerodynamic resistance at the first level with the stability
// corrections, Sutherland viscosity, the Re* term, and per class up to
// 100 um the resistances in series.  Classes coarser than 100 um keep
// their vdep.  Same expression order as the reference; the file is built
// without contraction.
__constant__ double kFengA[6][4] = {
    {0.0048, 0.0315, 1.2891, 1.0338}, {0.0037, 0.0120, 1.3977, 1.0707},
    {0.0042, 0.2928, 1.3970, 0.9155}, {0.0032, 0.1201, 1.1838, 1.0096},
    {0.0043, 0.1337, 1.2834, 1.1595}, {0.0045, 0.0925, 1.2654, 1.0891}};
__constant__ double kFengB[6][4] = {
    {1.0, 2.7925, 2.6878, 1.2644}, {1.0, 2.2413, 2.5838, 1.3247},
    {1.0, 3.8581, 2.5580, 1.0364}, {1.0, 3.4407, 2.8033, 1.2069},
    {1.0, 3.5456, 2.7157, 1.4863}, {1.0, 3.2920, 2.7227, 1.4240}};

// Feng mode of a diameter (setvdrydep.f90:134-152): 0 above 100 um
__device__ __forceinline__ int feng_mode(double d)
{
    return d > 100e-6 ? 0 : (d > 10e-6 ? 4 : (d > 2.5e-6 ? 3 : (d > 0.1e-6 ? 2 : 1)));
}

// USGS land-use category, truncated as Fortran INT does, onto Feng's six
// classes (setvdrydep.f90:167-216)
__device__ __forceinline__ int feng_land(double ldu)
{
    const int u = (int)ldu;
    if (u == 1) return 1;
    if (u == 10 || u == 19) return 3;
    if (u == 16) return 5;
    if (u >= 20 && u <= 23) return 2;
    if (u == 24) return 4;
    return 6;
}

__global__ void k_setvdrydep(Grid g, int nc, const met::ClassConst *cls, double zlayer1,
                             const double *ldu, const double *znt, const double *rl,
                             const double *ustar, const double *tempe, const double *rho,
                             double *vdep)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.NXY) return;
    const double vkarm = 0.4, visc0 = 1.827e-5, Tair0 = 291.15;
    const int land = feng_land(ldu[t]);
    const double us = ustar[t], z0 = znt[t];
    const double zeta = rl[t] * zlayer1, zeta0 = rl[t] * z0;
    double ra;
    if (us > 0.0) {
        ra = log(zlayer1 / z0);
        if (zeta > 0) {
            ra = ra + 4.7 * (zeta - zeta0);
        } else if (zeta < 0) {
            const double e = pow(1.0 - 15.0 * zeta, 0.25);
            const double e0 = pow(1.0 - 15.0 * zeta0, 0.25);
            ra = ra + log((e0 * e0 + 1) * ((e0 + 1) * (e0 + 1)) / ((e * e + 1) * ((e + 1) * (e + 1))));
            ra = ra + 2 * (atan(e) - atan(e0));
        }
        ra = ra / (us * vkarm);
    } else {
        ra = 1e20;
    }
    const double T = tempe[t];  // level 1
    const double visc = visc0 * ((Tair0 + 120.0) / (T + 120.0)) * pow(T / Tair0, 1.5);
    double Re = us * z0 * rho[t] / visc;
    if (Re > 40300.0) Re = 40300.0;
    const double x = (Re - 40300.0) / 15330.0;
    const double vd1 = 0.0226 * us * exp(-(0.5 * x * x));
    for (int ic = 0; ic < nc; ++ic) {
        const int m = feng_mode(cls[ic].diam);
        if (m == 0) continue;
        double vd = vd1 + kFengA[land - 1][m - 1] * pow(us, kFengB[land - 1][m - 1]);
        vd = ra + 1.0 / vd;
        vdep[t + (long)g.NXY * ic] = -1.0 / vd;
    }
}

// corvver.f90:16-36 (LON-LAT grids, meteo.f90:116-124), one thread per
// cell: vz minus the terrain-slope projection of the scaled horizontal wind,
// forward differences of topg clamped at the last column / row
__global__ void k_corvver(Grid g, const double *hm1, const double *topg, double dX1, double dX2,
                          const double *vx, const double *vy, double *vz)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.N) return;
    const long c = t % g.NXY;
    const int i = (int)(c % g.nx) + 1, j = (int)(c / g.nx) + 1;
    const double dyinv = 1.0 / dX2;
    const double dxinv = 1.0 / (hm1[c] * dX1);
    const long ir = (i < g.nx ? i : g.nx - 1) + (long)g.nx * (j - 1);  // (min(i+1,nx), j)
    const long ju = (i - 1) + (long)g.nx * (j < g.ny ? j : g.ny - 1);  // (i, min(j+1,ny))
    const double zx = dxinv * (topg[ir] - topg[c]);
    const double zy = dyinv * (topg[ju] - topg[c]);
    const double vs = zx * vx[t] + zy * vy[t];
    vz[t] = vz[t] - vs;
}

}  // namespace

namespace met {

int kappa3(hipStream_t s, const Grid &g, int modkv, int modkh, const double *dz,
           const double *zlayer, const double *vx, const double *vy, const double *vptemp,
           const double *zi, const double *rl, const double *ustar, const double *hm1,
           const double *kmh, double khn_dx, double khn_dy, double rkv0, double rkh0,
           double csh, int lonlat, double *rkver, double *rkhor)
{
    hipLaunchKernelGGL(k_kappa3, dim3(blocks_for(g.N, 256)), dim3(256), 0, s, g, modkv, modkh, dz,
                       zlayer, vx, vy, vptemp, zi, rl, ustar, hm1, kmh, khn_dx, khn_dy, rkv0,
                       rkh0, csh, lonlat, rkver, rkhor);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int setvset(hipStream_t s, const Grid &g, int np, int modv, const ClassConst *cls,
            const double *tempe, const double *rho, double *visc, double *vset, double *cd_io,
            int *status, int chunk, int warm)
{
    hipLaunchKernelGGL(k_visc, dim3(blocks_for(g.N, 256)), dim3(256), 0, s, g.N, tempe, visc);
    F3D_CHECK_LAUNCH();
    const long P = (long)np * g.N;
    const long nthreads = (P + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_setvset, dim3(blocks_for(nthreads, 128)), dim3(128), 0, s, g, np, modv,
                       cls, visc, rho, vset, cd_io, status, P, chunk, warm);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int setvdrydep(hipStream_t s, const Grid &g, const ClassConst *cls, double zlayer1,
               const double *ldu, const double *znt, const double *rl, const double *ustar,
               const double *tempe, const double *rho, double *vdep)
{
    hipLaunchKernelGGL(k_setvdrydep, dim3((unsigned)((g.NXY + 255) / 256)), dim3(256), 0, s, g,
                       g.nc, cls, zlayer1, ldu, znt, rl, ustar, tempe, rho, vdep);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int corvver(hipStream_t s, const Grid &g, const double *hm1, const double *topg, double dX1,
            double dX2, const double *vx, const double *vy, double *vz)
{
    hipLaunchKernelGGL(k_corvver, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, g, hm1, topg,
                       dX1, dX2, vx, vy, vz);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int scale(hipStream_t s, const Grid &g, const double *hm1, double *vx, double *rho,
          const double *rkhor, double *rkh1)
{
    hipLaunchKernelGGL(k_scale, dim3(blocks_for(g.N, 256)), dim3(256), 0, s, g, hm1, vx, rho,
                       rkhor, rkh1);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int finddt(hipStream_t s, const Grid &g, const double *dz, double dX1, double dX2,
           const double *vx, const double *vy, const double *vz, const double *rkh1,
           const double *rkhor, const double *rkver, const double *vset,
           unsigned long long *dtcn_bits)
{
    const unsigned nb = (unsigned)std::min<long>(blocks_for(g.N, 256), 2048);
    hipLaunchKernelGGL(k_finddt, dim3(nb), dim3(256), 0, s, g, dz, dX1, dX2, vx, vy, vz, rkh1,
                       rkhor, rkver, vset, dtcn_bits);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

}  // namespace met
}  // namespace f3d
