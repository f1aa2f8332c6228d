// Deterministic FP64 reductions for the mass budgets.
//
// Every reduction runs on a fixed grid of red::kPartials blocks: each thread
// sums its grid-stride share in index order, each block folds its threads by
// a fixed LDS tree into one partial, and a single-block kernel folds the
// partials by a fixed tree.  The order depends only on the problem size, so
// a scenario's budget is bit-identical whatever else runs on the device.
// The summation order differs from the reference's serial loops (cmass2d.f90,
// cmass3d.f90), which changes the sums by O(1e-16) relative.
#include "f3d_internal.h"

namespace f3d {
namespace {

constexpr int kBS = 256;

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// cmass2d.f90:38-105: one "site" per x-face (j,k) pair, y-face (i,k) pair
// and top cell (i,j), for every class.  Terms are dt*area*velo*conc in the
// reference's association; influx terms enter with a minus sign.
__global__ void __launch_bounds__(kBS) k_cmass2d64(Grid g, int ncls, const double *c,
                                                    const double *vx, const double *vy,
                                                    const double *vz, const double *vset,
                                                    const double *dz, double dt, double *part)
{
    __shared__ double sh[kBS];
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long nxf = (long)ny * nz, nyf = (long)nx * nz, ntop = g.NXY;
    const long per = nxf + nyf + ntop;
    const long total = per * ncls;
    double acc = 0.0;
    for (long t = (long)blockIdx.x * kBS + threadIdx.x; t < total; t += (long)gridDim.x * kBS) {
        const int q = (int)(t / per);
        long r = t - (long)q * per;
        const double *cc = c + q * g.G;
        if (r < nxf) {
            const int j = (int)(r % ny) + 1, k = (int)(r / ny) + 1;
            const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
            const double area = faty * g.dy * 0.5 * (dz[k] + dz[k - 1]);
            acc += -(dt * area * vx[midx(nx, ny, 1, j, k)] * cc[cidx(nx, ny, 0, j, k)]);
            acc += dt * area * vx[midx(nx, ny, nx, j, k)] * cc[cidx(nx, ny, nx + 1, j, k)];
            continue;
        }
        r -= nxf;
        if (r < nyf) {
            const int i = (int)(r % nx) + 1, k = (int)(r / nx) + 1;
            const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
            const double area = fatx * g.dx * 0.5 * (dz[k] + dz[k - 1]);
            acc += -(dt * area * vy[midx(nx, ny, i, 1, k)] * cc[cidx(nx, ny, i, 0, k)]);
            acc += dt * area * vy[midx(nx, ny, i, ny, k)] * cc[cidx(nx, ny, i, ny + 1, k)];
            continue;
        }
        r -= nyf;
        {
            const int i = (int)(r % nx) + 1, j = (int)(r / nx) + 1;
            const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
            const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
            const double area = fatx * faty * g.dx * g.dy;
            const double velo = vz[midx(nx, ny, i, j, nz)] + vset[q * g.N + midx(nx, ny, i, j, nz)];
            acc += dt * area * velo * cc[cidx(nx, ny, i, j, nz + 1)];
        }
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// getc0, endstep.f90:80-106: slot 0 = sum of the particle slots, in order;
// grid.y = member
__global__ void k_getc0(Grid g, double *cdry, double *cwet)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.NXY) return;
    const long mo = (long)blockIdx.y * (g.nc + 1) * g.NXY;
    double d = 0.0, w = 0.0;
    for (int ic = 1; ic <= g.nc; ++ic) {
        d = d + cdry[mo + t + ic * g.NXY];
        w = w + cwet[mo + t + ic * g.NXY];
    }
    cdry[mo + t] = d;
    cwet[mo + t] = w;
}

// cmass3d.f90:24-58 with cload(:,:,0) = cdry + cwet (endstep.f90:35) when
// cload0 is null; grid = (kMemberPartials, members); four partials per
// block: vol, dep, dry, wet, stored at part[(4*m + quantity)*nb + block].
template <typename T>
__global__ void __launch_bounds__(kBS) k_cmass3d(Grid g, const T *c, const double *cload0,
                                                 const double *cdry0, const double *cwet0,
                                                 long dstride, const double *dz, double *part)
{
    __shared__ double sh[kBS];
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int m = blockIdx.y;
    const long SZ = (long)(nx + 2) * (ny + 2);
    const T *cm = c + (long)m * g.nc * g.G;
    const double *d0 = cdry0 + m * dstride, *w0 = cwet0 + m * dstride;
    const double *l0 = cload0 ? cload0 + m * dstride : nullptr;
    double av = 0.0, ad = 0.0, adr = 0.0, aw = 0.0;
    for (long t = (long)blockIdx.x * kBS + threadIdx.x; t < g.NXY; t += (long)gridDim.x * kBS) {
        const int i = (int)(t % nx) + 1, j = (int)(t / nx) + 1;
        const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
        const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
        const double ff = fatx * faty;
        for (int q = 0; q < g.nc; ++q) {
            const T *cc = cm + q * g.G + cidx(nx, ny, i, j, 0);
            double zmass = 0.0;
            for (int k = 1; k <= nz - 1; ++k)
                zmass = zmass + 0.5 * dz[k] * ((double)cc[k * SZ] + (double)cc[(k + 1) * SZ]);
            av += ff * zmass;
        }
        const double d = d0[t], w = w0[t];
        const double cl = l0 ? l0[t] : d + w;
        ad += ff * cl;
        adr += ff * d;
        aw += ff * w;
    }
    const double sv = block_sum(av, sh);
    const double sd = block_sum(ad, sh);
    const double sdr = block_sum(adr, sh);
    const double sw = block_sum(aw, sh);
    if (threadIdx.x == 0) {
        const long nb = gridDim.x;
        part[(4L * m + 0) * nb + blockIdx.x] = sv;
        part[(4L * m + 1) * nb + blockIdx.x] = sd;
        part[(4L * m + 2) * nb + blockIdx.x] = sdr;
        part[(4L * m + 3) * nb + blockIdx.x] = sw;
    }
}

// dst[r] (+)= sum of part[r*n .. r*n+n): one block per row, fixed tree
__global__ void __launch_bounds__(1024) k_finish(const double *part, int n, double *dst,
                                                 int accumulate)
{
    const int r = blockIdx.x;
    red::block_sum_row(part + (long)r * n, n, dst + r, accumulate);
}

}  // namespace

namespace red {

int finish_rows(hipStream_t s, const double *partials, int n, int nrows, double *dst,
                int accumulate)
{
    hipLaunchKernelGGL(k_finish, dim3(nrows), dim3(1024), 0, s, partials, n, dst, accumulate);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int cmass2d64(hipStream_t s, const Grid &g, int ncls, const double *c, const double *vx,
              const double *vy, const double *vz, const double *vset, const double *dz,
              double dt, double *scratch, double *outmass)
{
    hipLaunchKernelGGL(k_cmass2d64, dim3(kPartials), dim3(kBS), 0, s, g, ncls, c, vx, vy, vz,
                       vset, dz, dt, scratch);
    F3D_CHECK_LAUNCH();
    return finish_rows(s, scratch, kPartials, 1, outmass, 1);
}

int getc0(hipStream_t s, const Grid &g, double *cdry, double *cwet)
{
    hipLaunchKernelGGL(k_getc0, dim3((unsigned)((g.NXY + 255) / 256), g.nm), dim3(256), 0, s, g,
                       cdry, cwet);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

// Output products of printres_nc.f90:150-241 for one member, one thread
// per column, in the reference's summation order: load (nx,ny,0:nc) =
// cload/Hm1 with cload = cdry + cwet and slot 0 from the class sums
// (endstep.f90:35, getc0); the deposit thickness 1e2*sum(cload/rhop)/Hm1
// (cm); the z-cumulative particle concentration 1e3*trapezoid of
// sum_ic(c)/Hm1 over levels 1..nz (g/m2).  Null outputs are skipped.
template <typename T>
__global__ void k_products(Grid g, const T *c, const double *cdry, const double *cwet,
                           const double *hm1, const met::ClassConst *cls, const double *dz,
                           double *load, double *thick, double *colmass)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.NXY) return;
    const int nc = g.nc;
    const int i = (int)(t % g.nx) + 1, j = (int)(t / g.nx) + 1;
    const double h = hm1[t];
    if (load) {
        double d0 = 0.0, w0 = 0.0;
        for (int q = 1; q <= nc; ++q) {
            d0 = d0 + cdry[q * g.NXY + t];
            w0 = w0 + cwet[q * g.NXY + t];
        }
        load[t] = (d0 + w0) / h;
        for (int q = 1; q <= nc; ++q)
            load[q * g.NXY + t] = (cdry[q * g.NXY + t] + cwet[q * g.NXY + t]) / h;
    }
    if (thick) {
        double s = 0.0;
        for (int q = 1; q <= nc; ++q)
            s = s + (cdry[q * g.NXY + t] + cwet[q * g.NXY + t]) / cls[q - 1].rhop;
        thick[t] = 1e2 * s / h;
    }
    if (colmass) {
        auto level = [&](int k) {
            double w = 0.0;
            for (int q = 0; q < nc; ++q) w = w + (double)c[(long)q * g.G + cidx(g.nx, g.ny, i, j, k)];
            return w / h;
        };
        double acc = 0.0, wk = level(1);
        for (int k = 1; k <= g.nz - 1; ++k) {
            const double wn = level(k + 1);
            acc = acc + dz[k] * 0.5 * (wk + wn);
            wk = wn;
        }
        colmass[t] = 1e3 * acc;
    }
}

template <typename T>
static int products_impl(hipStream_t s, const Grid &g, const T *c, const double *cdry,
                         const double *cwet, const double *hm1, const met::ClassConst *cls,
                         const double *dz, double *load, double *thick, double *colmass)
{
    hipLaunchKernelGGL((k_products<T>), dim3((unsigned)((g.NXY + 255) / 256)), dim3(256), 0, s, g, c,
                       cdry, cwet, hm1, cls, dz, load, thick, colmass);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int products(hipStream_t s, const Grid &g, const double *c, const double *cdry,
             const double *cwet, const double *hm1, const met::ClassConst *cls, const double *dz,
             double *load, double *thick, double *colmass)
{
    return products_impl<double>(s, g, c, cdry, cwet, hm1, cls, dz, load, thick, colmass);
}
int products(hipStream_t s, const Grid &g, const float *c, const double *cdry,
             const double *cwet, const double *hm1, const met::ClassConst *cls, const double *dz,
             double *load, double *thick, double *colmass)
{
    return products_impl<float>(s, g, c, cdry, cwet, hm1, cls, dz, load, thick, colmass);
}

template <typename T>
static int cmass3d_impl(hipStream_t s, const Grid &g, const T *c, const double *cload0,
                        const double *cdry0, const double *cwet0, long dstride, const double *dz,
                        double *scratch, double *out)
{
    hipLaunchKernelGGL(k_cmass3d<T>, dim3(kMemberPartials, g.nm), dim3(kBS), 0, s, g, c, cload0,
                       cdry0, cwet0, dstride, dz, scratch);
    F3D_CHECK_LAUNCH();
    return finish_rows(s, scratch, kMemberPartials, 4 * g.nm, out, 0);
}

int cmass3d(hipStream_t s, const Grid &g, const double *c, const double *cload0,
            const double *cdry0, const double *cwet0, long dstride, const double *dz,
            double *scratch, double *out)
{
    return cmass3d_impl<double>(s, g, c, cload0, cdry0, cwet0, dstride, dz, scratch, out);
}

int cmass3d(hipStream_t s, const Grid &g, const float *c, const double *cload0,
            const double *cdry0, const double *cwet0, long dstride, const double *dz,
            double *scratch, double *out)
{
    return cmass3d_impl<float>(s, g, c, cload0, cdry0, cwet0, dstride, dz, scratch, out);
}

}  // namespace red
}  // namespace f3d
