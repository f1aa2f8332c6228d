// Pieces of the fused FP32 step around the flux kernel (f3d_flux32.hip):
// the FP32 step fields built from the FP64 coefficients after a meteo read
// (per-cell velocities as Courant numbers, the 12-byte diffusion records
// and vz+vset per class in the flux kernel's tile-major layout, the z-level
// table), the cmass2d outflux and the completion of the ghost frame, setsrc on the
// FP32 state, and the FP32 <-> FP64 state conversions.  The deposit and
// the budgets accumulate in FP64.
#include <algorithm>

#include "f3d_ctx.h"

namespace f3d {
namespace f32 {
namespace {

// ---------------- cmass2d outflux and the ghost frame ----------------
// cmass2d terms (cmass2d.f90:38-105) of the step's clipped ghosts, FP64:
// the x-face ghosts the flux kernel left in the ghost buffer; the y-face and
// top ghosts as setbcc forms them (setbcc.f90:114-169) from c_in's first
// and last rows and its top level (whole rows: coalesced).  Block bx of nbx
// for member mem: a grid-stride share of the face cells of every class; the
// block's outflux sum goes to part[mem*nbx + bx].
__device__ __forceinline__ void outflux_block(const FrameArgs &a, int bx, int nbx, int mem)
{
    __shared__ double sh[256];
    const Grid &g = a.g;
    const double *vx = a.vx, *vy = a.vy, *vz = a.vz, *vset = a.vset, *dz = a.dz;
    const double dt = a.dt;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    // face cells of a class: x faces (k, j), y faces (k, i), top (j, i); the
    // flux kernel's parameters hold per * nc under 2^31, so the cell index
    // and its quotients are 32-bit
    const unsigned nxf = (unsigned)ny * nz, nyf = (unsigned)nx * nz, nxy = (unsigned)g.NXY;
    const unsigned per = a.per.d, total = per * (unsigned)g.nc;
    const long gf = (long)nz * ny;
    double acc = 0.0;
    for (unsigned t = (unsigned)bx * 256 + threadIdx.x; t < total; t += (unsigned)nbx * 256) {
        const unsigned q = fdiv(t, a.per);  // class
        unsigned r = t - q * per;
        if (r < nxf) {
            const unsigned k0 = fdiv(r, a.ny);
            const int j = (int)(r - k0 * ny) + 1, k = (int)k0 + 1;
            // the ghost buffer is [side][k][j]: cell r of the face
            const float *gb = a.gbuf + ((long)mem * g.nc + q) * ghost_floats(g);
            const float gl = gb[r], gr = gb[gf + r];
            // a zero ghost (inflow, or no ash) adds nothing: its wind is
            // not read
            if (gl == 0.f && gr == 0.f) continue;
            const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
            const double area = faty * g.dy * 0.5 * (dz[k] + dz[k - 1]);
            if (gl != 0.f) acc += -(dt * area * vx[midx(nx, ny, 1, j, k)] * (double)gl);
            if (gr != 0.f) acc += dt * area * vx[midx(nx, ny, nx, j, k)] * (double)gr;
            continue;
        }
        r -= nxf;
        if (r < nyf) {
            const unsigned k0 = fdiv(r, a.nx);
            const int i = (int)(r - k0 * nx) + 1, k = (int)k0 + 1;
            const float *ci = a.cin + ((long)mem * g.nc + q) * g.G;
            const float gd = (vy[midx(nx, ny, i, 1, k)] >= 0.0) ? 0.f : fmaxf(0.f, ci[cidx(nx, ny, i, 1, k)]);
            const float gu = (vy[midx(nx, ny, i, ny, k)] <= 0.0) ? 0.f : fmaxf(0.f, ci[cidx(nx, ny, i, ny, k)]);
            const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
            const double area = fatx * g.dx * 0.5 * (dz[k] + dz[k - 1]);
            acc += -(dt * area * vy[midx(nx, ny, i, 1, k)] * (double)gd);
            acc += dt * area * vy[midx(nx, ny, i, ny, k)] * (double)gu;
            continue;
        }
        r -= nyf;
        // the top ghost is zero wherever vz+vset <= 0 at the top (tfac = 0;
        // settling columns): nothing more is read there
        const float tf = a.tfac[(long)q * nxy + r];
        if (tf == 0.f) continue;
        const unsigned j0 = fdiv(r, a.nx);
        const int i = (int)(r - j0 * nx) + 1, j = (int)j0 + 1;
        const float *ci = a.cin + ((long)mem * g.nc + q) * g.G;
        const float gt = fmaxf(0.f, tf * ci[cidx(nx, ny, i, j, nz)]);
        if (gt == 0.f) continue;
        const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
        const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
        const double area = fatx * faty * g.dx * g.dy;
        const double velo = vz[midx(nx, ny, i, j, nz)] + vset[(long)q * g.N + midx(nx, ny, i, j, nz)];
        acc += dt * area * velo * (double)gt;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[(long)mem * nbx + bx] = sh[0];
}

__global__ void __launch_bounds__(256) k_outflux(FrameArgs a)
{
    outflux_block(a, blockIdx.x, gridDim.x, blockIdx.y);  // y: member
}

// the outflux of many contexts: workgroup b belongs to the (context,
// member) row r with rofs[r] <= b < rofs[r+1], which holds that row's
// outflux_blocks workgroups
__global__ void __launch_bounds__(256) k_outflux_batch(const FrameArgs *fa, const int2 *rows,
                                                       const int *rofs, int nrows)
{
    const int b = blockIdx.x;
    int lo = 0, hi = nrows;  // the last row whose first workgroup is <= b
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (rofs[mid] <= b) lo = mid;
        else hi = mid;
    }
    const int2 r = rows[lo];
    outflux_block(fa[r.x], b - rofs[lo], rofs[lo + 1] - rofs[lo], r.y);
}

// Completes the ghost frame of a state c the flux kernel wrote from cin
// (setbcc.f90:19-195 on cin, then the clip of sizloop.f90:77-78): the
// x-face ghosts from the ghost buffer, the y-face ghosts and the bottom and
// top planes from cin's faces and the wind and deposition factors, the
// corners from cin's interior corners, and the edges (which setbcc never
// sets) clipped.  Slab blockIdx.y (member*nc + class).
__global__ void __launch_bounds__(256) k_frame_fill(FrameFill a)
{
    const Grid &g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long q = blockIdx.y, cls = q % g.nc, nxy = g.NXY;
    float *cs = a.c + q * g.G;
    const float *ci = a.cin + q * g.G;
    const float *gb = a.gbuf + q * ghost_floats(g);
    const long gf = (long)nz * ny;
    const long nxf = (long)ny * nz, nyf = (long)nx * nz;
    const long total = 2 * nxf + 2 * nyf + 2 * nxy + 4L * nz + 4L * ny + 4L * nx + 8;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        long r = t;
        if (r < 2 * nxf) {
            const int f = (int)(r / nxf);
            r -= f * nxf;
            const int j = (int)(r % ny) + 1, k = (int)(r / ny) + 1;
            cs[cidx(nx, ny, f ? nx + 1 : 0, j, k)] = gb[f * gf + (long)(k - 1) * ny + (j - 1)];
            continue;
        }
        r -= 2 * nxf;
        if (r < 2 * nyf) {
            const int f = (int)(r / nyf);
            r -= f * nyf;
            const int i = (int)(r % nx) + 1, k = (int)(r / nx) + 1;
            const int jf = f ? ny : 1;
            const double v = a.vy[midx(nx, ny, i, jf, k)];
            const bool in = f ? v <= 0.0 : v >= 0.0;
            cs[cidx(nx, ny, i, f ? ny + 1 : 0, k)] = in ? 0.f : fmaxf(0.f, ci[cidx(nx, ny, i, jf, k)]);
            continue;
        }
        r -= 2 * nyf;
        if (r < 2 * nxy) {
            const int f = (int)(r / nxy);
            r -= f * nxy;
            const int i = (int)(r % nx) + 1, j = (int)(r / nx) + 1;
            cs[cidx(nx, ny, i, j, f ? nz + 1 : 0)] =
                f ? fmaxf(0.f, a.tfac[cls * nxy + r] * ci[cidx(nx, ny, i, j, nz)])
                  : fmaxf(0.f, ci[cidx(nx, ny, i, j, 1)] * a.bfac[cls * nxy + r]);
            continue;
        }
        r -= 2 * nxy;
        long o;
        float v;
        if (r < 4L * nz) {
            const int e = (int)(r / nz), k = (int)(r % nz) + 1;
            o = cidx(nx, ny, (e & 1) ? nx + 1 : 0, (e & 2) ? ny + 1 : 0, k);
            v = cs[o];
        } else if ((r -= 4L * nz) < 4L * ny) {
            const int e = (int)(r / ny), j = (int)(r % ny) + 1;
            o = cidx(nx, ny, (e & 1) ? nx + 1 : 0, j, (e & 2) ? nz + 1 : 0);
            v = cs[o];
        } else if ((r -= 4L * ny) < 4L * nx) {
            const int e = (int)(r / nx), i = (int)(r % nx) + 1;
            o = cidx(nx, ny, i, (e & 1) ? ny + 1 : 0, (e & 2) ? nz + 1 : 0);
            v = cs[o];
        } else {
            // corners copy the adjacent interior corner (setbcc.f90:173-195)
            const int e = (int)(r - 4L * nx);
            const int i = (e & 1) ? nx + 1 : 0, j = (e & 2) ? ny + 1 : 0, k = (e & 4) ? nz + 1 : 0;
            o = cidx(nx, ny, i, j, k);
            v = ci[cidx(nx, ny, i ? nx : 1, j ? ny : 1, k ? nz : 1)];
        }
        cs[o] = fmaxf(0.f, v);
    }
}

// outmass[member] += the sum of its partials, for many contexts (the tree
// of red::finish_rows, row by row)
__global__ void __launch_bounds__(1024) k_finish_batch(const FrameArgs *fa, const int2 *rows)
{
    const int2 r = rows[blockIdx.x];
    const FrameArgs &a = fa[r.x];
    const int nb = outflux_blocks(a.g);
    red::block_sum_row(a.part + (long)r.y * nb, nb, a.outmass + r.y, 1);
}

// FP32 step fields from the FP64 coefficients
__global__ void k_fields3d(long N, const double *vx, const double *vy, const double *vz,
                           const double *rho, const double *rkh1, const double *rkhor,
                           const double *rkver, float *o)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const double r = rho[t];
    o[t] = (float)vx[t];
    o[N + t] = (float)vy[t];
    o[2 * N + t] = (float)vz[t];
    o[3 * N + t] = (float)(1.0 / r);
    o[4 * N + t] = (float)(r * rkh1[t]);
    o[5 * N + t] = (float)(r * rkhor[t]);
    o[6 * N + t] = (float)(r * rkver[t]);
}

// Courant numbers of the horizontal sweeps, formed once per meteo read
// (they depend on the wind and dt only): ax = dt/dx * vx, ay = dt/dy * vy
// (advctx.f90:24-36, advcty.f90).  The flux kernel forms the limiter factor
// of each face from the two cells beside it, and the face-averaged
// diffusivities from the products rho*rkh1 and rho*rkhor.
__global__ void k_fields_face(Grid g, const double *vx, const double *vy, double dt, float *ax,
                              float *ay)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.N) return;
    ax[t] = (float)(dt / g.dx * vx[t]);
    ay[t] = (float)(dt / g.dy * vy[t]);
}

// the flux kernel's records in its tile-major layout (Fields::cA): record r
// of a level is lane r%64 of tile column (r/64)%ftx in row r/(64*ftx); the
// lane's column is clamped to the domain, as the kernel clamps ghost columns
__global__ void k_fields_pack(Grid g, Fields f, float2 *cA, float *cD, float *wzt)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = f.ftl * g.nz;
    if (r >= n) return;
    const int li = (int)(r & 63);
    const long t = r >> 6;
    const int bx = (int)(t % f.ftx);
    const long jk = t / f.ftx;  // (j-1) + ny*(k-1)
    const int ci = min(max(kFluxTileCols * bx + li, 1), g.nx);
    const long src = (ci - 1) + (long)g.nx * jk;
    cA[r] = make_float2(f.ax[src], f.ay[src]);
    cD[3 * r] = f.lr[src];
    cD[3 * r + 1] = f.pkz[src];
    cD[3 * r + 2] = f.pky[src];
    for (int ic = 0; ic < g.nc; ++ic) wzt[ic * n + r] = f.wz[ic * g.N + src];
}

// 1/Hm1^2 per column: the flux kernel forms rho*rkh1 = rho*rkhor/Hm1^2
// (meteo.f90:108-112) from rho*rkhor instead of reading a field for it
__global__ void k_fields_h2(long nxy, const double *hm1, float *h2)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nxy) return;
    const double h = hm1[t];
    h2[t] = (float)(1.0 / (h * h));
}

// per-level constants of advctzc (advctzc.f90:153-181) and diffz
// (diffz.f90:19-62), products formed in FP64 and rounded once
__global__ void k_zlev(int nz, const ZCoef *zc, double dt, ZLev *zl)
{
    for (int k = threadIdx.x; k <= nz + 1; k += blockDim.x) {
        ZLev z{};
        if (k >= 1 && k <= nz) {
            const ZCoef c = zc[k];
            const double a2 = dt * 0.5 * c.dzm1 * c.dzz;
            z.dtdzz = (float)(dt * c.dzz);
            z.a2 = (float)a2;
            z.a2bc = (float)(a2 * dt * c.dzk_1);
            z.dzk = (float)c.dzk;
            z.hldz = (float)(0.5 * c.dzk);
            z.ldz = (float)c.dzk;
        }
        if (k == 0) z.ldz = (float)zc[1].dzk_1;
        zl[k] = z;
    }
}

__global__ void k_fields_class(Grid g, const double *vz, const double *vset, const double *vdep,
                               const double *rkver, const double *dz, float *wz, float *bfac,
                               float *tfac, double *vsum1)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (t < g.N) wz[(long)q * g.N + t] = (float)(vz[t] + vset[(long)q * g.N + t]);
    if (t < g.NXY) {
        const long top = (long)(g.nz - 1) * g.NXY + t;
        const double vs1 = vset[(long)q * g.N + t];
        const double vd = vdep[(long)q * g.NXY + t];
        // setbcc.f90:108-113 and 126-130
        bfac[(long)q * g.NXY + t] =
            ((vz[t] + vs1) >= 0.0) ? 0.f : (float)(1.0 + dz[0] * vd / rkver[t]);
        tfac[(long)q * g.NXY + t] = ((vz[top] + vset[(long)q * g.N + top]) <= 0.0) ? 0.f : 1.f;
        const double s = vs1 + vd;  // accum.f90:22
        vsum1[(long)q * g.NXY + t] = s < 0.0 ? s : 0.0;
    }
}

// setsrc.f90:20-35.  Member m owns source points [off[m], off[m+1]) and
// its tmrat block (nsou_m, nc) starts at nc*off[m].  Points are grouped by
// target cell (grp[g] = {member, first, count} into perm, in point order):
// one thread per (group, class) adds its points in the reference order, so
// every cell sees the same sequence of roundings as the serial loop.
// item t = (cell group, class) of one context's sources
__device__ __forceinline__ void setsrc_item(const SrcArgs &a, int t)
{
    const Grid &g = a.g;
    float *c = a.c;
    const int *off = a.off, *perm = a.perm, *isrc = a.isrc, *jsrc = a.jsrc, *ksrc = a.ksrc;
    const int4 *grp = a.grp;
    const double *tmrat = a.tmrat, *zlayer = a.zlayer, *dz = a.dz, *hm1 = a.hm1;
    const double dt = a.dt;
    const int q = t % g.nc;
    const int4 gr = grp[t / g.nc];
    const int mem = gr.x;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int o = off[mem], nsou = off[mem + 1] - o;
    const double *tm = tmrat + (long)g.nc * o;
    const int P0 = perm[gr.y];
    const int i = isrc[P0], j = jsrc[P0], k = ksrc[P0];
    if (k < 1 || k > nz) return;
    const double dzeta = (k == 1) ? 0.5 * dz[k] + zlayer[k - 1] : 0.5 * (dz[k - 1] + dz[k]);
    const double vol0 = hm1[(i - 1) + (long)nx * (j - 1)] * g.dx * g.dy * dzeta;
    float *cc = c + ((long)mem * g.nc + q) * g.G + cidx(nx, ny, i, j, k);
    float v = *cc;
    for (int n = 0; n < gr.z; ++n) {
        const int is = perm[gr.y + n] - o;
        v = (float)((double)v + dt * tm[is + (long)nsou * q] / vol0);
    }
    *cc = v;
}

__global__ void k_setsrc32(SrcArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ngrp * a.g.nc) return;
    setsrc_item(a, t);
}

// the sources of many contexts: items [first[k], first[k+1]) are context k's
__global__ void k_setsrc_batch(const SrcArgs *sa, const int *first, int nctx)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= first[nctx]) return;
    int lo = 0, hi = nctx - 1;  // the context whose range holds t
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (first[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    setsrc_item(sa[lo], t - first[lo]);
}

__global__ void k_to64(const float *a, double *b, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        b[t] = (double)a[t];
}

__global__ void k_from64(const double *a, float *b, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        b[t] = (float)a[t];
}

}  // namespace

// records per level of the flux kernel's tile-major layout
static size_t tiled_level(const Grid &g)
{
    return (size_t)g.ny * (size_t)flux_tiles(g.nx, kFluxTileCols + 6) * 64;
}

// the flux kernel's tile-major groups lead the buffer ((ax, ay), the 12-byte
// records, vz+vset per class: each starts on a multiple of 64 records, so
// every region row is 256-byte aligned), followed by the z-level table and
// the separate fields
static size_t separate_offset(const Grid &g)
{
    const size_t zl = (size_t)(g.nz + 2) * (sizeof(ZLev) / sizeof(float));
    const size_t t = tiled_level(g) * (size_t)g.nz;
    return ((size_t)(5 + g.nc) * t + zl + 63) & ~(size_t)63;
}

size_t fields_floats(const Grid &g)
{
    return separate_offset(g) + (size_t)9 * g.N + (size_t)g.nc * g.N +
           (size_t)2 * g.nc * g.NXY + (size_t)g.NXY;
}

Fields fields_view(const Grid &g, const float *b0, const double *vsum1)
{
    Fields f;
    f.ftx = flux_tiles(g.nx, kFluxTileCols + 6);
    f.ftl = (long)tiled_level(g);
    const size_t t = (size_t)f.ftl * (size_t)g.nz;
    f.cA = reinterpret_cast<const float2 *>(b0);
    f.cD = b0 + 2 * t;
    f.wzt = b0 + 5 * t;
    f.zl = reinterpret_cast<const ZLev *>(b0 + (5 + g.nc) * t);
    f.base = b0;
    f.bytes = sizeof(float) * fields_floats(g);
    const float *b = b0 + separate_offset(g);
    f.vx = b;
    f.vy = b + g.N;
    f.vz = b + 2 * g.N;
    f.lr = b + 3 * g.N;
    f.pkx = b + 4 * g.N;
    f.pky = b + 5 * g.N;
    f.pkz = b + 6 * g.N;
    f.wz = b + 7 * g.N;
    f.bfac = b + 7 * g.N + (long)g.nc * g.N;
    f.tfac = f.bfac + (long)g.nc * g.NXY;
    const float *face = f.tfac + (long)g.nc * g.NXY;
    f.ax = face;
    f.ay = face + g.N;
    f.h2 = face + 2 * g.N;
    f.vsum1 = vsum1;
    return f;
}

int build_fields(hipStream_t s, const Grid &g, const double *vx, const double *vy,
                 const double *vz, const double *rho, const double *rkh1, const double *rkhor,
                 const double *rkver, const double *vset, const double *vdep, const double *dz,
                 const double *hm1, const ZCoef *zc, double dt, float *f32buf, double *vsum1)
{
    Fields f = fields_view(g, f32buf, vsum1);
    // vx .. pkz are contiguous from f.vx
    hipLaunchKernelGGL(k_fields3d, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, g.N, vx,
                       vy, vz, rho, rkh1, rkhor, rkver, const_cast<float *>(f.vx));
    F3D_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fields_class, dim3((unsigned)((g.N + 255) / 256), g.nc), dim3(256), 0, s, g,
                       vz, vset, vdep, rkver, dz, const_cast<float *>(f.wz),
                       const_cast<float *>(f.bfac), const_cast<float *>(f.tfac), vsum1);
    F3D_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fields_face, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, g, vx,
                       vy, dt, const_cast<float *>(f.ax), const_cast<float *>(f.ay));
    F3D_CHECK_LAUNCH();
    const long nt = f.ftl * g.nz;
    hipLaunchKernelGGL(k_fields_pack, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, g, f,
                       const_cast<float2 *>(f.cA), const_cast<float *>(f.cD),
                       const_cast<float *>(f.wzt));
    F3D_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fields_h2, dim3((unsigned)((g.NXY + 255) / 256)), dim3(256), 0, s,
                       (long)g.NXY, hm1, const_cast<float *>(f.h2));
    F3D_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_zlev, dim3(1), dim3(64), 0, s, g.nz, zc, dt, const_cast<ZLev *>(f.zl));
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

bool tile_supported(TileShape ts) { return flux_supported(ts); }

int fused_step(hipStream_t s, const StepArgs &a, const Fields &f, TileShape ts)
{
    if (flux_supported(ts)) return flux_step(s, a, f, ts);
    set_error("fused step: unsupported tile %dx%d", ts.tx, ts.ty);
    return F3D_EINVAL;
}

FrameArgs frame_args(const StepArgs &a, const Fields &f, const double *vx64, const double *vy64,
                     const double *vz64, const double *vset64, const double *dz, double *scratch,
                     double *outmass)
{
    FrameArgs fa;
    fa.g = a.g;
    fa.per = fast_div((unsigned)((long)a.g.ny * a.g.nz + (long)a.g.nx * a.g.nz + a.g.NXY));
    fa.ny = fast_div((unsigned)a.g.ny);
    fa.nx = fast_div((unsigned)a.g.nx);
    fa.gbuf = a.gbuf;
    fa.cin = a.cin;
    fa.tfac = f.tfac;
    fa.vx = vx64;
    fa.vy = vy64;
    fa.vz = vz64;
    fa.vset = vset64;
    fa.dz = dz;
    fa.dt = a.dt;
    fa.part = scratch;
    fa.outmass = outmass;
    return fa;
}

int outflux(hipStream_t s, const StepArgs &a, const Fields &f, const double *vx64,
            const double *vy64, const double *vz64, const double *vset64, const double *dz,
            double *scratch, double *outmass)
{
    const FrameArgs fa = frame_args(a, f, vx64, vy64, vz64, vset64, dz, scratch, outmass);
    const int nb = outflux_blocks(a.g);
    hipLaunchKernelGGL(k_outflux, dim3(nb, a.g.nm), dim3(256), 0, s, fa);
    F3D_CHECK_LAUNCH();
    return red::finish_rows(s, scratch, nb, a.g.nm, outmass, 1);
}

int frame_fill(hipStream_t s, const FrameFill &a)
{
    const Grid &g = a.g;
    const long total = 2L * g.nz * (g.nx + g.ny) + 2L * g.NXY + 4L * (g.nx + g.ny + g.nz) + 8;
    const unsigned nb = (unsigned)std::min<long>((total + 255) / 256, 1024);
    hipLaunchKernelGGL(k_frame_fill, dim3(nb, (unsigned)(g.nc * g.nm)), dim3(256), 0, s, a);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int outflux_batch(hipStream_t s, const FrameArgs *fa, const int2 *rows, const int *rofs, int nrows,
                  int nblocks)
{
    if (nrows <= 0) return F3D_OK;
    hipLaunchKernelGGL(k_outflux_batch, dim3(nblocks), dim3(256), 0, s, fa, rows, rofs, nrows);
    F3D_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_finish_batch, dim3(nrows), dim3(1024), 0, s, fa, rows);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

SrcArgs src_args(const Grid &g, float *c, const int *off, const int4 *grp, int ngrp,
                 const int *perm, const int *isrc, const int *jsrc, const int *ksrc,
                 const double *tmrat, const double *zlayer, const double *dz, const double *hm1,
                 double dt)
{
    SrcArgs a;
    a.g = g;
    a.c = c;
    a.off = off;
    a.grp = grp;
    a.ngrp = ngrp;
    a.perm = perm;
    a.isrc = isrc;
    a.jsrc = jsrc;
    a.ksrc = ksrc;
    a.tmrat = tmrat;
    a.zlayer = zlayer;
    a.dz = dz;
    a.hm1 = hm1;
    a.dt = dt;
    return a;
}

int setsrc(hipStream_t s, const Grid &g, float *c, const int *off, const int4 *grp, int ngrp,
           const int *perm, const int *isrc, const int *jsrc, const int *ksrc,
           const double *tmrat, const double *zlayer, const double *dz, const double *hm1,
           double dt)
{
    const int n = ngrp * g.nc;
    if (n == 0) return F3D_OK;
    const SrcArgs a = src_args(g, c, off, grp, ngrp, perm, isrc, jsrc, ksrc, tmrat, zlayer, dz, hm1, dt);
    hipLaunchKernelGGL(k_setsrc32, dim3((n + 255) / 256), dim3(256), 0, s, a);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int setsrc_batch(hipStream_t s, const SrcArgs *sa, const int *first, int nctx, int nitems)
{
    if (nitems <= 0) return F3D_OK;
    hipLaunchKernelGGL(k_setsrc_batch, dim3((nitems + 255) / 256), dim3(256), 0, s, sa, first, nctx);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int to_fp64(hipStream_t s, const float *c32, double *c64, long n)
{
    hipLaunchKernelGGL(k_to64, dim3(1024), dim3(256), 0, s, c32, c64, n);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

int from_fp64(hipStream_t s, const double *c64, float *c32, long n)
{
    hipLaunchKernelGGL(k_from64, dim3(1024), dim3(256), 0, s, c64, c32, n);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

}  // namespace f32
}  // namespace f3d
