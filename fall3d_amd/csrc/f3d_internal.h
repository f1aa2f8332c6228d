// Internal declarations shared by the HIP translation units of libfall3d_amd.
//
// Layout conventions (reference: Sources/Master.f90:92-136):
//   c    (0:nx+1,0:ny+1,0:nz+1,nc)  column-major with a one-cell ghost frame
//   3-D  (nx,ny,nz)                  meteo and coefficients, no ghosts
//   vset (nx,ny,nz,nc)   vdep (nx,ny,nc)   cdry (nx,ny,0:nc)   dz (0:nz)
// Device buffers keep exactly this layout so that host arrays of the
// reference can be copied in and out without reshaping.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <algorithm>
#include <cstdint>

#include "../../include/fall3d_amd.h"

namespace f3d {

void set_error(const char *fmt, ...);

// propagate a nonzero F3D_* status
#define TRY(x)                   \
    do {                         \
        int rc_ = (x);           \
        if (rc_) return rc_;     \
    } while (0)

#define F3D_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::f3d::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                             __FILE__, __LINE__);                                         \
            return F3D_EHIP;                                                              \
        }                                                                                 \
    } while (0)

#define F3D_CHECK_LAUNCH()                                                                \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess) {                                                           \
            ::f3d::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),   \
                             __FILE__, __LINE__);                                         \
            return F3D_EHIP;                                                              \
        }                                                                                 \
    } while (0)

// Grid of one scenario as the kernels see it.
struct Grid {
    int nx, ny, nz, nc;
    long G;       // ghosted slab (nx+2)(ny+2)(nz+2)
    long N;       // nx*ny*nz
    long NXY;     // nx*ny
    double dx, dy;
    int nm;       // ensemble members sharing grid, meteo and dt (1 for a single run);
                  // member m, class ic owns concentration slab m*nc + ic and
                  // deposit slots cdry + m*(nc+1)*NXY
};

__host__ __device__ inline long cidx(int nx, int ny, int i, int j, int k)
{
    return (long)i + (long)(nx + 2) * ((long)j + (long)(ny + 2) * (long)k);
}
__host__ __device__ inline long midx(int nx, int ny, int i, int j, int k)
{
    return (long)(i - 1) + (long)nx * ((long)(j - 1) + (long)ny * (long)(k - 1));
}

// Per-level vertical coefficients of advctzc/diffz (advctzc.f90:151-153,
// diffz.f90:82-85), computed once on the host with IEEE division.
struct ZCoef {
    double dzz;    // 2/(dz(k)+dz(k-1))
    double dzk;    // 1/dz(k)
    double dzk_1;  // 1/dz(k-1)
    double dzm1;   // dz(k-1)
};

// ---------------- plain FP64 kernels (f3d_plain64.hip) ----------------
// Every launcher processes `ncls` classes in one launch; class q uses
// c + q*G, vset + q*N, vdep + q*NXY.  Within a class the arithmetic is the
// reference routine's, operation for operation.
namespace p64 {
int setbcc(hipStream_t s, const Grid &g, int ncls, double *c, const double *vx, const double *vy,
           const double *vz, const double *vset, const double *vdep, const double *rkver,
           const double *dz);
int advctzc(hipStream_t s, const Grid &g, int ncls, double *c, const double *vz,
            const double *vset, const ZCoef *zc, double dznz, double dt);
int advctx(hipStream_t s, const Grid &g, int ncls, double *c, const double *vx, double dx,
           double dt);
int advcty(hipStream_t s, const Grid &g, int ncls, double *c, const double *vy, double dy,
           double dt);
int diffz(hipStream_t s, const Grid &g, int ncls, double *c, const ZCoef *zc, const double *ldz,
          double dt, const double *rkver, const double *rho);
int diffx(hipStream_t s, const Grid &g, int ncls, double *c, double dx, double dt,
          const double *rkh, const double *rho);
int diffy(hipStream_t s, const Grid &g, int ncls, double *c, double dy, double dt,
          const double *rkh, const double *rho);
// cdry slot of class q is cdry + (ic0 + q)*NXY
int accum(hipStream_t s, const Grid &g, int ncls, const double *c, double *cdry, int ic0,
          double dt, const double *vset, const double *vdep);
// wetdep.f90:28-50 for ncls classes from the slab c and deposit slot cwet
// of the first (grid.y = class); wet[q] = 0 skips class q
int wetdep(hipStream_t s, const Grid &g, int ncls, double *c, double *cwet, const int *wet,
           double dt, const double *zlayer, const double *dz, const double *zi,
           const double *prate);
int clip(hipStream_t s, double *c, long n);
int setsrc(hipStream_t s, const Grid &g, double *c, int nsou, const int *isrc, const int *jsrc,
           const int *ksrc, const double *tmrat, const double *zlayer, const double *dz,
           const double *hm1, double dt);
}  // namespace p64

// ---------------- deterministic FP64 reductions (f3d_reduce.hip) ----------------
namespace red {
constexpr int kPartials = 1024;  // fixed number of per-block partials
// outmass += cmass2d over ncls classes (cmass2d.f90:38-105); scratch holds kPartials doubles
int cmass2d64(hipStream_t s, const Grid &g, int ncls, const double *c, const double *vx,
              const double *vy, const double *vz, const double *vset, const double *dz,
              double dt, double *scratch, double *outmass);
// per-member partial counts: a member's reduction tree depends only on its
// own sizes, never on how many members share the launch
constexpr int kMemberPartials = 256;
constexpr int kFramePartials = 1024;
// getc0 (endstep.f90:80-106) for every member: slot 0 of cdry/cwet = sum of
// the class slots; member m's deposit is cdry + m*(nc+1)*NXY
int getc0(hipStream_t s, const Grid &g, double *cdry, double *cwet);
// cmass3d (cmass3d.f90:24-58) for every member, without the final dx*dy
// scaling: out[4*m + 0..3] = sums for massvol, massdep, massdry, masswet.
// Member m's slot-0 deposit arrays start at cdry0/cwet0 + m*dstride; cload0
// may be null, then cload(:,:,0) = cdry0 + cwet0 (endstep.f90:35).
// scratch holds 4*nm*kMemberPartials doubles.
int cmass3d(hipStream_t s, const Grid &g, const double *c, const double *cload0,
            const double *cdry0, const double *cwet0, long dstride, const double *dz,
            double *scratch, double *out);
int cmass3d(hipStream_t s, const Grid &g, const float *c, const double *cload0,
            const double *cdry0, const double *cwet0, long dstride, const double *dz,
            double *scratch, double *out);
// final stage: dst[r] (+)= sum(partials[r*n .. r*n+n)) for r < nrows, fixed tree
int finish_rows(hipStream_t s, const double *partials, int n, int nrows, double *dst,
                int accumulate);
// the tree of finish_rows for one row, by a block of 1024 threads: *dst (+)=
// sum of row[0 .. n)
__device__ __forceinline__ void block_sum_row(const double *row, int n, double *dst, int accumulate)
{
    __shared__ double sh[1024];
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) v += row[t];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *dst = accumulate ? *dst + sh[0] : sh[0];
}
inline int finish_sum(hipStream_t s, const double *partials, int n, double *dst, int accumulate)
{
    return finish_rows(s, partials, n, 1, dst, accumulate);
}
}  // namespace red

// ---------------- meteo stage (f3d_meteo.hip), FP64 ----------------
namespace met {
int kappa3(hipStream_t s, const Grid &g, int modkv, int modkh, const double *dz,
           const double *zlayer, const double *vx, const double *vy, const double *vptemp,
           const double *zi, const double *rl, const double *ustar, const double *hm1,
           const double *kmh, double khn_dx, double khn_dy, double rkv0, double rkh0,
           double csh, int lonlat, double *rkver, double *rkhor);
// Per-class constants prepared on the host: rk1, rk2 (Ganser) and psi.
struct ClassConst {
    double diam, rhop, psi, rk1, rk2;
};
// setvset chain (setvset.f90:25-60) with warm-started chunks.
// visc is an N-double scratch for the per-cell Sutherland viscosity;
// cd_io[0] is the carried drag coefficient, cd_io[1] receives the last one.
int setvset(hipStream_t s, const Grid &g, int np, int modv, const ClassConst *cls,
            const double *tempe, const double *rho, double *visc, double *vset, double *cd_io,
            int *status, int chunk, int warm);
// setvdrydep.f90:54-264 for every column and class (meteo.f90:99-100);
// tempe and rho at level 1, before the Hm1 scaling; classes above 100 um
// keep their vdep
int setvdrydep(hipStream_t s, const Grid &g, const ClassConst *cls, double zlayer1,
               const double *ldu, const double *znt, const double *rl, const double *ustar,
               const double *tempe, const double *rho, double *vdep);
int scale(hipStream_t s, const Grid &g, const double *hm1, double *vx, double *rho,
          const double *rkhor, double *rkh1);
// corvver.f90:16-36 on the scaled vx (LON-LAT grids, meteo.f90:116-124)
int corvver(hipStream_t s, const Grid &g, const double *hm1, const double *topg, double dX1,
            double dX2, const double *vx, const double *vy, double *vz);
// dtcn = max over cells and classes of the CFL rate (finddt.f90:26-48);
// dtcn_bits is a uint64 holding the bit pattern of a non-negative double.
int finddt(hipStream_t s, const Grid &g, const double *dz, double dX1, double dX2,
           const double *vx, const double *vy, const double *vz, const double *rkh1,
           const double *rkhor, const double *rkver, const double *vset,
           unsigned long long *dtcn_bits);
}  // namespace met

namespace red {
// Output products of printres_nc.f90:150-241 for one member (c, cdry, cwet
// at that member's slabs and slots): load (nx,ny,0:nc) in kg/m2, deposit
// thickness (nx,ny) in cm, z-cumulative concentration (nx,ny) in g/m2;
// null outputs are skipped
int products(hipStream_t s, const Grid &g, const double *c, const double *cdry,
             const double *cwet, const double *hm1, const met::ClassConst *cls, const double *dz,
             double *load, double *thick, double *colmass);
int products(hipStream_t s, const Grid &g, const float *c, const double *cdry,
             const double *cwet, const double *hm1, const met::ClassConst *cls, const double *dz,
             double *load, double *thick, double *colmass);
}  // namespace red

}  // namespace f3d
