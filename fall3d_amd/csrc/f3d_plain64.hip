// Plain FP64 transport kernels: one kernel per reference routine, the
// reference's arithmetic operation for operation.  Built with
// -ffp-contract=off so that no multiply-add is fused and every rounding is
// the reference's; division and square root are the IEEE-correct
// sequences.  Each launch covers all classes of the step (grid.y = class):
// classes are independent inside sizloop's class loop, so running routine
// r for every class before routine r+1 gives the same bits as the
// reference's class-by-class order (Sources/sizloop.f90:31-73).
//
// Launch shapes:
//   setbcc       one thread per face cell
//   advctzc      one thread per (i,j) column, register window along k
//   diffz        one thread per (i,j) column, register window along k
//   advctx/diffx one workgroup per (j,k) line staged in LDS (the line is
//                owned whole, so the in-place Jacobi update is race free)
//   advcty/diffy one workgroup per (TI-wide i-tile, k) holding the whole
//                j extent of its TI columns in LDS
//   accum, clip  one thread per element
#include "f3d_internal.h"

namespace f3d {
namespace {

__device__ __forceinline__ double fmax_f(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fmin_f(double a, double b) { return a < b ? a : b; }

// minmod limiter, advctx.f90:40 (sign(1,x) is copysign)
__device__ __forceinline__ double minmod(double dr, double dl)
{
    return 0.5 * (copysign(1.0, dr) + copysign(1.0, dl)) * fmin_f(fabs(dr), fabs(dl));
}

// ---------------- setbcc.f90:19-109 ----------------
__global__ void k_setbcc(Grid g, double *c, const double *vx, const double *vy, const double *vz,
                         const double *vset, const double *vdep, const double *rkver,
                         const double *dz)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int q = blockIdx.y;
    double *cc = c + q * g.G;
    const double *vs = vset + q * g.N;
    const double *vd = vdep + q * g.NXY;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nb = g.NXY, nxf = (long)ny * nz, nyf = (long)nx * nz;
    if (t < nb) {
        const int i = (int)(t % nx) + 1, j = (int)(t / nx) + 1;
        if ((vz[midx(nx, ny, i, j, 1)] + vs[midx(nx, ny, i, j, 1)]) >= 0.0)
            cc[cidx(nx, ny, i, j, 0)] = 0.0;
        else
            cc[cidx(nx, ny, i, j, 0)] = cc[cidx(nx, ny, i, j, 1)] *
                (1.0 + dz[0] * vd[t] / rkver[midx(nx, ny, i, j, 1)]);
        if ((vz[midx(nx, ny, i, j, nz)] + vs[midx(nx, ny, i, j, nz)]) <= 0.0)
            cc[cidx(nx, ny, i, j, nz + 1)] = 0.0;
        else
            cc[cidx(nx, ny, i, j, nz + 1)] = cc[cidx(nx, ny, i, j, nz)];
        return;
    }
    t -= nb;
    if (t < nxf) {
        const int j = (int)(t % ny) + 1, k = (int)(t / ny) + 1;
        cc[cidx(nx, ny, 0, j, k)] =
            (vx[midx(nx, ny, 1, j, k)] >= 0.0) ? 0.0 : cc[cidx(nx, ny, 1, j, k)];
        cc[cidx(nx, ny, nx + 1, j, k)] =
            (vx[midx(nx, ny, nx, j, k)] <= 0.0) ? 0.0 : cc[cidx(nx, ny, nx, j, k)];
        return;
    }
    t -= nxf;
    if (t < nyf) {
        const int i = (int)(t % nx) + 1, k = (int)(t / nx) + 1;
        cc[cidx(nx, ny, i, 0, k)] =
            (vy[midx(nx, ny, i, 1, k)] >= 0.0) ? 0.0 : cc[cidx(nx, ny, i, 1, k)];
        cc[cidx(nx, ny, i, ny + 1, k)] =
            (vy[midx(nx, ny, i, ny, k)] <= 0.0) ? 0.0 : cc[cidx(nx, ny, i, ny, k)];
        return;
    }
    t -= nyf;
    if (t == 0) {
        cc[cidx(nx, ny, 0, 0, 0)] = cc[cidx(nx, ny, 1, 1, 1)];
        cc[cidx(nx, ny, nx + 1, 0, 0)] = cc[cidx(nx, ny, nx, 1, 1)];
        cc[cidx(nx, ny, 0, ny + 1, 0)] = cc[cidx(nx, ny, 1, ny, 1)];
        cc[cidx(nx, ny, nx + 1, ny + 1, 0)] = cc[cidx(nx, ny, nx, ny, 1)];
        cc[cidx(nx, ny, 0, 0, nz + 1)] = cc[cidx(nx, ny, 1, 1, nz)];
        cc[cidx(nx, ny, nx + 1, 0, nz + 1)] = cc[cidx(nx, ny, nx, 1, nz)];
        cc[cidx(nx, ny, 0, ny + 1, nz + 1)] = cc[cidx(nx, ny, 1, ny, nz)];
        cc[cidx(nx, ny, nx + 1, ny + 1, nz + 1)] = cc[cidx(nx, ny, nx, ny, nz)];
    }
}

// ---------------- advctzc.f90:28-81 ----------------
// W2(k) of advctzc.f90:164-169 from c(k-1), c(k), c(k+1), vk(k), vk(max(1,k-1))
__device__ __forceinline__ double w2z(double cm, double c0, double cp, double vk, double vk_1,
                                      const ZCoef &z, double dt)
{
    const double vk_12 = 0.5 * (vk + vk_1);
    const double dr = (cp - c0) * z.dzk;
    const double dl = (c0 - cm) * z.dzk_1;
    const double sigma = minmod(dr, dl);
    const double Crk = fabs(vk_12) * dt * z.dzk_1;
    return dt * 0.5 * z.dzm1 * (1.0 - Crk) * sigma * fabs(vk_12) * z.dzz;
}

__global__ void k_advctzc(Grid g, double *c, const double *vz, const double *vset,
                          const ZCoef *zc, double dznz, double dt)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.NXY) return;
    const int q = blockIdx.y;
    const int i = (int)(t % nx) + 1, j = (int)(t / nx) + 1;
    const long SZ = (long)(nx + 2) * (ny + 2);
    double *cc = c + q * g.G + cidx(nx, ny, i, j, 0);
    const double *vzc = vz + t;
    const double *vsc = vset + q * g.N + t;

    double cm1 = cc[0], c0 = cc[SZ], cp1 = cc[2 * SZ];
    double vk = vzc[0] + vsc[0];
    const ZCoef z1 = zc[1];
    double W1k = dt * (vk * c0 - vk * cm1) * z1.dzz;  // vk_1 = vk(max(1,0)) = vk(1)
    double W2k = w2z(cm1, c0, cp1, vk, vk, z1, dt);
    double W2km1 = W2k;                               // W2(max(1,k-1)) at k = 1
    for (int k = 1; k <= nz; ++k) {
        double W1kp1, W2kp1, vkp1 = 0.0, cp2 = 0.0;
        if (k < nz) {
            cp2 = cc[(long)(k + 2) * SZ];
            vkp1 = vzc[(long)k * g.NXY] + vsc[(long)k * g.NXY];
            const ZCoef z = zc[k + 1];
            W1kp1 = dt * (vkp1 * cp1 - vk * c0) * z.dzz;
            W2kp1 = w2z(c0, cp1, cp2, vkp1, vk, z, dt);
        } else {
            W1kp1 = dt * vk * (cp1 - c0) * dznz;  // advctzc.f90:175-181
            W2kp1 = W2k;                         // W2(min(nz,k+1)) at k = nz
        }
        double nv;
        if (vk > 0.0)
            nv = c0 - W1k - (W2k - W2km1);
        else
            nv = c0 - W1kp1 - (W2kp1 - W2k);
        cc[(long)k * SZ] = nv;
        cm1 = c0;
        c0 = cp1;
        cp1 = cp2;
        vk = vkp1;
        W1k = W1kp1;
        W2km1 = W2k;
        W2k = W2kp1;
    }
}

// ---------------- advctx.f90 / advcty.f90 line update ----------------
// New value at position i (1..n) of a line held in LDS with element stride
// st: sc[0..n+1] is the pre-sweep line, sv[1..n] the velocity.  The five
// fluxes of the update are evaluated with the reference's expressions
// (advctx.f90:32-53), so they carry the bits its work array would hold,
// including W1(n+1) built from the last velocity vx(n) (:46).
__device__ __forceinline__ double adv_line_update(const double *sc, const double *sv, int st,
                                                  int i, int n, double h, double dt)
{
    const double auxv = dt / h;
#define SC(p) sc[(p) * st]
#define SV(p) sv[(p) * st]
    auto W1 = [&](int f) -> double {
        if (f == n + 1) return auxv * SV(n) * (SC(n + 1) - SC(n));
        const double vf = SV(f), vf_1 = SV(f > 1 ? f - 1 : 1);
        return auxv * (vf * SC(f) - vf_1 * SC(f - 1));
    };
    auto W2 = [&](int p) -> double {
        const double vp = SV(p), vp_1 = SV(p > 1 ? p - 1 : 1);
        const double v12 = 0.5 * (vp + vp_1);
        const double dr = (SC(p + 1) - SC(p)) / h;
        const double dl = (SC(p) - SC(p - 1)) / h;
        const double sigma = minmod(dr, dl);
        const double Cr = auxv * fabs(v12);
        return dt * 0.5 * (1.0 - Cr) * sigma * fabs(v12);
    };
    const double ci = SC(i);
    if (SV(i) > 0.0) return ci - W1(i) - (W2(i) - W2(i > 1 ? i - 1 : 1));
    return ci - W1(i + 1) - (W2(i < n ? i + 1 : n) - W2(i));
#undef SC
#undef SV
}

// diffx.f90:20-53 / diffy.f90 line update: sc[0..n+1] line, lr[1..n] =
// 1/rho (l_rho), pk[1..n] = rho*rkh (the product the reference forms in
// place); the end faces use the un-averaged products (:27-28,44-45).
__device__ __forceinline__ double diff_line_update(const double *sc, const double *lr,
                                                   const double *pk, int st, int i, int n,
                                                   double dtdx2)
{
#define SC(p) sc[(p) * st]
#define LR(p) lr[(p) * st]
#define PK(p) pk[(p) * st]
    double gradr, gradl;
    if (i == 1) {
        gradr = PK(2) * (SC(2) * LR(2) - SC(1) * LR(1));
        gradl = PK(1) * (SC(1) * LR(1) - SC(0) * LR(1));
        gradl = fmax_f(0.0, gradl);
    } else if (i == n) {
        gradr = PK(n) * (SC(n + 1) * LR(n) - SC(n) * LR(n));
        gradl = PK(n - 1) * (SC(n) * LR(n) - SC(n - 1) * LR(n - 1));
        gradr = fmin_f(0.0, gradr);
    } else {
        const double rksr = 0.5 * (PK(i + 1) + PK(i));
        const double rksl = 0.5 * (PK(i) + PK(i - 1));
        gradr = rksr * (SC(i + 1) * LR(i + 1) - SC(i) * LR(i));
        gradl = rksl * (SC(i) * LR(i) - SC(i - 1) * LR(i - 1));
    }
    return SC(i) + dtdx2 * (gradr - gradl);
#undef SC
#undef LR
#undef PK
}

// one workgroup per (j,k) line, grid.y = class
__global__ void k_advctx(Grid g, double *c, const double *vx, double dx, double dt)
{
    extern __shared__ double sm[];
    const int nx = g.nx, ny = g.ny;
    const int line = blockIdx.x, q = blockIdx.y;
    const int j = line % ny + 1, k = line / ny + 1;
    double *sc = sm;             // [0..nx+1]
    double *sv = sm + (nx + 2);  // [1..nx]
    double *cl = c + q * g.G + cidx(nx, ny, 0, j, k);
    const double *vl = vx + midx(nx, ny, 1, j, k) - 1;
    for (int i = threadIdx.x; i < nx + 2; i += blockDim.x) sc[i] = cl[i];
    for (int i = threadIdx.x + 1; i <= nx; i += blockDim.x) sv[i] = vl[i];
    __syncthreads();
    for (int i = threadIdx.x + 1; i <= nx; i += blockDim.x)
        cl[i] = adv_line_update(sc, sv, 1, i, nx, dx, dt);
}

__global__ void k_diffx(Grid g, double *c, const double *rkh, const double *rho, double dtdx2)
{
    extern __shared__ double sm[];
    const int nx = g.nx, ny = g.ny;
    const int line = blockIdx.x, q = blockIdx.y;
    const int j = line % ny + 1, k = line / ny + 1;
    double *sc = sm, *lr = sm + (nx + 2), *pk = sm + 2 * (nx + 2);
    double *cl = c + q * g.G + cidx(nx, ny, 0, j, k);
    const long m0 = midx(nx, ny, 1, j, k) - 1;
    for (int i = threadIdx.x; i < nx + 2; i += blockDim.x) sc[i] = cl[i];
    for (int i = threadIdx.x + 1; i <= nx; i += blockDim.x) {
        const double r = rho[m0 + i];
        lr[i] = 1.0 / r;
        pk[i] = r * rkh[m0 + i];
    }
    __syncthreads();
    for (int i = threadIdx.x + 1; i <= nx; i += blockDim.x)
        cl[i] = diff_line_update(sc, lr, pk, 1, i, nx, dtdx2);
}

// one workgroup per (TI-wide i-tile, k) holding the whole j extent;
// LDS element (j, ti) sits at j*TI + ti.  grid.y = class
__global__ void k_advcty(Grid g, double *c, const double *vy, double dy, double dt, int TI)
{
    extern __shared__ double sm[];
    const int nx = g.nx, ny = g.ny;
    const int ntile = (nx + TI - 1) / TI;
    const int tile = blockIdx.x % ntile, k = blockIdx.x / ntile + 1, q = blockIdx.y;
    const int i0 = tile * TI + 1;
    const int tin = min(TI, nx - i0 + 1);
    double *sc = sm;
    double *sv = sm + (long)(ny + 2) * TI;
    double *cb = c + q * g.G;
    for (int e = threadIdx.x; e < (ny + 2) * TI; e += blockDim.x) {
        const int ti = e % TI, jj = e / TI;
        if (ti < tin) {
            sc[e] = cb[cidx(nx, ny, i0 + ti, jj, k)];
            if (jj >= 1 && jj <= ny) sv[e] = vy[midx(nx, ny, i0 + ti, jj, k)];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ny * TI; e += blockDim.x) {
        const int ti = e % TI, jj = e / TI + 1;
        if (ti < tin)
            cb[cidx(nx, ny, i0 + ti, jj, k)] = adv_line_update(sc + ti, sv + ti, TI, jj, ny, dy, dt);
    }
}

__global__ void k_diffy(Grid g, double *c, const double *rkh, const double *rho, double dtdy2,
                        int TI)
{
    extern __shared__ double sm[];
    const int nx = g.nx, ny = g.ny;
    const int ntile = (nx + TI - 1) / TI;
    const int tile = blockIdx.x % ntile, k = blockIdx.x / ntile + 1, q = blockIdx.y;
    const int i0 = tile * TI + 1;
    const int tin = min(TI, nx - i0 + 1);
    double *sc = sm;
    double *lr = sm + (long)(ny + 2) * TI;
    double *pk = sm + 2L * (ny + 2) * TI;
    double *cb = c + q * g.G;
    for (int e = threadIdx.x; e < (ny + 2) * TI; e += blockDim.x) {
        const int ti = e % TI, jj = e / TI;
        if (ti < tin) {
            sc[e] = cb[cidx(nx, ny, i0 + ti, jj, k)];
            if (jj >= 1 && jj <= ny) {
                const double r = rho[midx(nx, ny, i0 + ti, jj, k)];
                lr[e] = 1.0 / r;
                pk[e] = r * rkh[midx(nx, ny, i0 + ti, jj, k)];
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ny * TI; e += blockDim.x) {
        const int ti = e % TI, jj = e / TI + 1;
        if (ti < tin)
            cb[cidx(nx, ny, i0 + ti, jj, k)] =
                diff_line_update(sc + ti, lr + ti, pk + ti, TI, jj, ny, dtdy2);
    }
}

// ---------------- diffz.f90:19-62 ----------------
__global__ void k_diffz(Grid g, double *c, const ZCoef *zc, const double *ldz, double dt,
                        const double *rkver, const double *rho)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.NXY) return;
    const int q = blockIdx.y;
    const int i = (int)(t % nx) + 1, j = (int)(t / nx) + 1;
    const long SZ = (long)(nx + 2) * (ny + 2);
    double *cc = c + q * g.G + cidx(nx, ny, i, j, 0);
    const double *K = rkver + t;
    const double *R = rho + t;
    const long S = g.NXY;
#define KK(k) K[(long)((k) - 1) * S]
#define RR(k) R[(long)((k) - 1) * S]
    // old c(k-1), c(k), c(k+1) kept in registers: work(k) needs pre-sweep values
    double cm1 = cc[0], c0 = cc[SZ], cp1 = cc[2 * SZ];
    double Rm1 = 0.0, R0 = RR(1), Rp1 = RR(2);
    double Km1 = 0.0, K0 = KK(1), Kp1 = KK(2);
    {
        const double rkst = Kp1 * Rp1;
        const double rksb = K0 * R0;
        const double gradt = rkst * ((cp1 / Rp1) - (c0 / R0)) * ldz[1];
        double gradb = rksb * ((c0 / R0) - (cm1 / R0)) * ldz[0];
        gradb = fmax_f(0.0, gradb);
        cc[SZ] = c0 + dt * (gradt - gradb) * zc[1].dzz;
    }
    for (int k = 2; k <= nz; ++k) {
        cm1 = c0;
        c0 = cp1;
        cp1 = cc[(long)(k + 1) * SZ];
        Rm1 = R0;
        R0 = Rp1;
        Km1 = K0;
        K0 = Kp1;
        double w;
        if (k < nz) {
            Rp1 = RR(k + 1);
            Kp1 = KK(k + 1);
            const double rkst = 0.5 * (Rp1 * Kp1 + R0 * K0);
            const double rksb = 0.5 * (R0 * K0 + Rm1 * Km1);
            const double gradt = rkst * ((cp1 / Rp1) - (c0 / R0)) * ldz[k];
            const double gradb = rksb * ((c0 / R0) - (cm1 / Rm1)) * ldz[k - 1];
            w = dt * (gradt - gradb) * zc[k].dzz;
        } else {
            const double rkst = K0 * R0;
            const double rksb = Km1 * Rm1;
            double gradt = rkst * ((cp1 / R0) - (c0 / R0)) * ldz[nz];
            const double gradb = rksb * ((c0 / R0) - (cm1 / Rm1)) * ldz[nz - 1];
            gradt = fmin_f(0.0, gradt);
            w = dt * (gradt - gradb) * ldz[nz];
        }
        cc[(long)k * SZ] = c0 + w;
    }
#undef KK
#undef RR
}

// ---------------- accum.f90:17-25 ----------------
// wetdep.f90:28-50 (Jung and Shao 2006): one thread per column and class
// (grid.y); below the PBL height zi a fraction lambda = min(dt*a*P**b, 1)
// of c goes into cwet, level by level in the reference's order.  c points
// at the first class's slab, cwet at its deposit slot; classes whose wet
// flag is 0 (coarser than 100 um) are left alone.
__global__ void k_wetdep(Grid g, double *c, double *cwet, const int *wet, double dt,
                         const double *zlayer, const double *dz, const double *zi,
                         const double *prate)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (t >= g.NXY || !wet[q]) return;
    const int i = (int)(t % g.nx) + 1, j = (int)(t / g.nx) + 1;
    const double a = 8.4e-5, b = 0.79;
    double lambda = dt * a * pow(prate[t], b);
    lambda = lambda < 1.0 ? lambda : 1.0;
    double *cc = c + (long)g.G * q;
    double *w = cwet + (long)g.NXY * q + t;
    double acc = *w;
    for (int k = 1; k <= g.nz; ++k)
        if (zlayer[k - 1] <= zi[t]) {
            const long o = cidx(g.nx, g.ny, i, j, k);
            acc = acc + dz[k] * lambda * cc[o];
            cc[o] = cc[o] * (1.0 - lambda);
        }
    *w = acc;
}

__global__ void k_accum(Grid g, const double *c, double *cdry, int ic0, double dt,
                        const double *vset, const double *vdep)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.NXY) return;
    const int q = blockIdx.y;
    const int nx = g.nx, ny = g.ny;
    const int i = (int)(t % nx) + 1, j = (int)(t / nx) + 1;
    double *cd = cdry + (long)(ic0 + q) * g.NXY + t;
    const double s = vset[q * g.N + t] + vdep[q * g.NXY + t];
    *cd = *cd - dt * c[q * g.G + cidx(nx, ny, i, j, 1)] * fmin_f(0.0, s);
}

// ---------------- sizloop.f90:77-78 ----------------
__global__ void k_clip(double *c, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x)
        c[t] = fmax_f(0.0, c[t]);
}

// ---------------- setsrc.f90:20-35 ----------------
// one thread per class, sources in file order (duplicates accumulate in order)
__global__ void k_setsrc(Grid g, double *c, int nsou, const int *isrc, const int *jsrc,
                         const int *ksrc, const double *tmrat, const double *zlayer,
                         const double *dz, const double *hm1, double dt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nc) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    for (int is = 0; is < nsou; ++is) {
        const int i = isrc[is], j = jsrc[is], k = ksrc[is];
        if (k >= 1 && k <= nz) {
            const double dzeta = (k == 1) ? 0.5 * dz[k] + zlayer[k - 1] : 0.5 * (dz[k - 1] + dz[k]);
            const double vol0 = hm1[(i - 1) + (long)nx * (j - 1)] * g.dx * g.dy * dzeta;
            double *cc = c + q * g.G + cidx(nx, ny, i, j, k);
            *cc = *cc + dt * tmrat[is + (long)nsou * q] / vol0;
        }
    }
}

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// LDS tile width for the y sweeps: as wide as fits in `budget` bytes
int y_tile(int ny, int arrays, long budget)
{
    int ti = 16;
    while (ti > 1 && (long)(ny + 2) * ti * arrays * 8 > budget) ti /= 2;
    return ti;
}

constexpr long kLdsBudget = 64 * 1024;

}  // namespace

namespace p64 {

int setbcc(hipStream_t s, const Grid &g, int ncls, double *c, const double *vx, const double *vy,
           const double *vz, const double *vset, const double *vdep, const double *rkver,
           const double *dz)
{
    const long n = g.NXY + (long)g.ny * g.nz + (long)g.nx * g.nz + 1;
    hipLaunchKernelGGL(k_setbcc, dim3(blocks_for(n, 256), ncls), dim3(256), 0, s, g, c, vx, vy,
                       vz, vset, vdep, rkver, dz);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int advctzc(hipStream_t s, const Grid &g, int ncls, double *c, const double *vz,
            const double *vset, const ZCoef *zc, double dznz, double dt)
{
    hipLaunchKernelGGL(k_advctzc, dim3(blocks_for(g.NXY, 128), ncls), dim3(128), 0, s, g, c, vz,
                       vset, zc, dznz, dt);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

static int line_block(int n) { return n >= 192 ? 256 : (n >= 96 ? 128 : 64); }

int advctx(hipStream_t s, const Grid &g, int ncls, double *c, const double *vx, double dx,
           double dt)
{
    const size_t lds = sizeof(double) * (2 * (size_t)g.nx + 3);
    if ((long)lds > kLdsBudget) {
        set_error("advctx: nx=%d exceeds the LDS line budget", g.nx);
        return F3D_EINVAL;
    }
    hipLaunchKernelGGL(k_advctx, dim3((unsigned)(g.ny * g.nz), ncls), dim3(line_block(g.nx)), lds,
                       s, g, c, vx, dx, dt);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int diffx(hipStream_t s, const Grid &g, int ncls, double *c, double dx, double dt,
          const double *rkh, const double *rho)
{
    const size_t lds = sizeof(double) * 3 * ((size_t)g.nx + 2);
    if ((long)lds > kLdsBudget) {
        set_error("diffx: nx=%d exceeds the LDS line budget", g.nx);
        return F3D_EINVAL;
    }
    const double dtdx2 = dt / (dx * dx);  // diffx.f90:18
    hipLaunchKernelGGL(k_diffx, dim3((unsigned)(g.ny * g.nz), ncls), dim3(line_block(g.nx)), lds,
                       s, g, c, rkh, rho, dtdx2);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int advcty(hipStream_t s, const Grid &g, int ncls, double *c, const double *vy, double dy,
           double dt)
{
    const int TI = y_tile(g.ny, 2, kLdsBudget);
    const size_t lds = sizeof(double) * 2 * (size_t)(g.ny + 2) * TI;
    if ((long)lds > kLdsBudget) {
        set_error("advcty: ny=%d exceeds the LDS column budget", g.ny);
        return F3D_EINVAL;
    }
    const int ntile = (g.nx + TI - 1) / TI;
    hipLaunchKernelGGL(k_advcty, dim3((unsigned)(ntile * g.nz), ncls), dim3(256), lds, s, g, c,
                       vy, dy, dt, TI);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int diffy(hipStream_t s, const Grid &g, int ncls, double *c, double dy, double dt,
          const double *rkh, const double *rho)
{
    const int TI = y_tile(g.ny, 3, kLdsBudget);
    const size_t lds = sizeof(double) * 3 * (size_t)(g.ny + 2) * TI;
    if ((long)lds > kLdsBudget) {
        set_error("diffy: ny=%d exceeds the LDS column budget", g.ny);
        return F3D_EINVAL;
    }
    const int ntile = (g.nx + TI - 1) / TI;
    const double dtdy2 = dt / (dy * dy);  // diffy.f90:18
    hipLaunchKernelGGL(k_diffy, dim3((unsigned)(ntile * g.nz), ncls), dim3(256), lds, s, g, c,
                       rkh, rho, dtdy2, TI);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int diffz(hipStream_t s, const Grid &g, int ncls, double *c, const ZCoef *zc, const double *ldz,
          double dt, const double *rkver, const double *rho)
{
    hipLaunchKernelGGL(k_diffz, dim3(blocks_for(g.NXY, 128), ncls), dim3(128), 0, s, g, c, zc,
                       ldz, dt, rkver, rho);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int accum(hipStream_t s, const Grid &g, int ncls, const double *c, double *cdry, int ic0,
          double dt, const double *vset, const double *vdep)
{
    hipLaunchKernelGGL(k_accum, dim3(blocks_for(g.NXY, 256), ncls), dim3(256), 0, s, g, c, cdry,
                       ic0, dt, vset, vdep);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int wetdep(hipStream_t s, const Grid &g, int ncls, double *c, double *cwet, const int *wet,
           double dt, const double *zlayer, const double *dz, const double *zi,
           const double *prate)
{
    hipLaunchKernelGGL(k_wetdep, dim3(blocks_for(g.NXY, 256), ncls), dim3(256), 0, s, g, c, cwet,
                       wet, dt, zlayer, dz, zi, prate);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int clip(hipStream_t s, double *c, long n)
{
    const unsigned nb = (unsigned)std::min<long>(blocks_for(n, 256), 8192);
    hipLaunchKernelGGL(k_clip, dim3(nb), dim3(256), 0, s, c, n);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int setsrc(hipStream_t s, const Grid &g, double *c, int nsou, const int *isrc, const int *jsrc,
           const int *ksrc, const double *tmrat, const double *zlayer, const double *dz,
           const double *hm1, double dt)
{
    if (nsou <= 0) return F3D_OK;
    hipLaunchKernelGGL(k_setsrc, dim3(blocks_for(g.nc, 64)), dim3(64), 0, s, g, c, nsou, isrc,
                       jsrc, ksrc, tmrat, zlayer, dz, hm1, dt);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

}  // namespace p64
}  // namespace f3d
