// Device-resident context behind the f3d_* C ABI.
#pragma once

#include <array>
#include <vector>

#include "f3d_internal.h"

namespace f3d {

struct DeviceBuffer {
    void *p = nullptr;
    size_t bytes = 0;
    long long *cnt = nullptr;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { release(); }
    int alloc(size_t n, long long *counter);
    void release();
    template <typename T>
    T *as() const
    {
        return static_cast<T *>(p);
    }
};

int get_gama(double diam, double sphe, double *gama_out);
int class_constants(int modv, double diam, double rhop, double sphe, met::ClassConst *out);
Grid make_grid(int nx, int ny, int nz, int nc, double dx, double dy);
void vertical_coefs(int nz, const double *zlayer, std::vector<double> &dz,
                    std::vector<ZCoef> &zc, std::vector<double> &ldz);
// dst(nx,ny,nz) = the window (i0+1.., j0+1.., 1..) of src(snx,sny,*)
int window_copy(hipStream_t s, double *dst, const double *src, int nx, int ny, int nz, int snx,
                int sny, int i0, int j0);

// ---------------- FP32 fused path (f3d_fused32.hip) ----------------
namespace f32 {

// Per-level FP32 constants of the vertical operators (dt included)
struct ZLev {
    float dtdzz;  // dt*2/(dz(k)+dz(k-1))
    float a2;     // dt*0.5*dz(k-1)*dzz(k)        (advctzc W2)
    float a2bc;   // a2*dt/dz(k-1)                 (advctzc W2, Courant part)
    float dzk;    // 1/dz(k)
    float hldz;   // 0.5/dz(k)                      (averaged diffz face)
    float ldz;    // 1/dz(k); level 0: 1/dz(0)
    float pad0, pad1;
};

// FP32 step fields, built from the FP64 coefficients after a meteo read.
// 3-D fields share the (nx,ny,nz) layout; per-class fields are stacked.
struct Fields {
    const float *vx, *vy, *vz;  // scaled velocities (vx/Hm1)
    const float *lr;            // 1/rho (rho scaled by Hm1)
    const float *pkx;           // rho*rkh1  (diffx face products)
    const float *pky;           // rho*rkhor (diffy face products)
    const float *pkz;           // rho*rkver (diffz face products)
    const float *wz;            // vz+vset per class (nx,ny,nz,nc), summed in FP64
    const float *bfac;          // bottom ghost factor (nx,ny,nc): 0, or 1+dz0*vdep/rkver(1)
    const float *tfac;          // top ghost factor (nx,ny,nc): 0 or 1 (setbcc.f90:126-130)
    const double *vsum1;        // min(0, vset(:,:,1)+vdep) (nx,ny,nc), FP64 deposit factor
    // Courant numbers dt/dx*vx, dt/dy*vy (dt-dependent), see k_fields_face
    const float *ax, *ay;
    // Read by the flux kernel, in its tile-major layout: for each level,
    // row j and tile column bx (flux_tiles columns), the 64 columns of the
    // tile's region 58*bx .. 58*bx+63 (clamped to 1..nx) are 64 consecutive
    // records, so one region row is one aligned, whole-line load.  Record
    // index ((k-1)*ny + j-1)*ftx*64 + bx*64 + lane; ftl = ny*ftx*64 records
    // per level.
    const float2 *cA;  // (ax, ay)
    const float *cD;   // (1/rho, rho*rkver, rho*rkhor), 12-byte records
    const float *wzt;  // vz+vset per class, nz*ftl records per class
    int ftx;           // tile columns of the layout
    long ftl;          // records per level
    const float *h2;   // 1/Hm1^2 per column (nx,ny): rho*rkh1 = rho*rkhor*h2
    const ZLev *zl;    // per level 0..nz+1
    const float *base;  // the buffer all FP32 fields live in, and its size
    size_t bytes;
};

// output columns of one tile of the flux kernel (a 64-lane region row
// minus the halo of 3 on either side)
constexpr int kFluxTileCols = 58;

// Tiles of the flux kernel across n cells, for regions of `region` cells
// (64 columns, or RY rows) with `region - 6` output cells.  Tile b's region
// starts at cell (region - 6)*b.  The first tile owns cells 1 .. region-4:
// its halo on that side is the ghost at cell 0, in lane (row) 0.  The last
// tile owns up to cell n, the ghost n+1 at most in its last lane (row).  An
// inner tile owns lanes 3 .. region-4.  So t tiles cover (region-6)*t + 4
// cells.
__host__ __device__ inline int flux_tiles(int n, int region)
{
    const int own = region - 6;
    return n <= region - 2 ? 1 : 1 + (n - (region - 2) + own - 1) / own;
}

// Build Fields from the FP64 coefficient arrays (meteo.f90:108-112 applied).
int build_fields(hipStream_t s, const Grid &g, const double *vx, const double *vy,
                 const double *vz, const double *rho, const double *rkh1, const double *rkhor,
                 const double *rkver, const double *vset, const double *vdep, const double *dz,
                 const double *hm1, const ZCoef *zc, double dt, float *f32buf, double *vsum1);
size_t fields_floats(const Grid &g);
Fields fields_view(const Grid &g, const float *f32buf, const double *vsum1);

struct StepArgs {
    Grid g;
    const float *cin;
    float *cout;
    double *cdry;       // (nx,ny,0:nc); class q accumulates into slot q+1
    double dt;
    int iiter;
    const ZCoef *zc;    // per level (FP64, converted in kernel)
    // optional second stream and fork/join events: edge tiles of the flux
    // kernel run beside the interior tiles
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // wet deposition (sizloop.f90:65-67), when wet: per column the
    // scavenged fraction lambda (FP64) and the top level below the PBL
    // height (0: dry column), per class 1 for particles up to 100 um;
    // cwet (nx,ny,0:nc) per member like cdry; dz64 = dz(0:nz)
    int wet = 0;
    const double *lam = nullptr;
    const int *kwet = nullptr, *wetcls = nullptr;
    double *cwet = nullptr;
    const double *dz64 = nullptr;
    // the step's x-face ghosts of c_out, clipped, per slab ghost_floats(g):
    // [x low | x high] x nz levels x ny cells (cell j-1); the flux kernel
    // writes them here instead of into c_out's frame (frame_fill puts them
    // there)
    float *gbuf = nullptr;
};

__host__ __device__ inline long ghost_floats(const Grid &g) { return 2L * g.nz * g.ny; }

// wet deposition columns for the fused step: lambda = min(dt*a*P**b, 1)
// (wetdep.f90:38-40) and the top level with zlayer(k) <= zi (0 when
// lambda is 0)
int wet_columns(hipStream_t s, const Grid &g, double dt, const double *prate, const double *zi,
                const double *zlayer, double *lam, int *kwet);

// Tile geometry of the fused kernel (chosen by the autotuner).  tx = 58
// selects the flux kernel (64-lane region rows, ty interior rows).
struct TileShape {
    int tx, ty;
    int nw = 0;  // waves per workgroup of the flux kernel (0: its default)
    int mc = 0;  // member*class slabs per workgroup of the flux kernel (0: 1)
};
bool tile_supported(TileShape ts);

// One fused launch: setbcc + advctzc + advctx/advcty (parity order) + diffz
// + diffx/diffy (parity order) + accum + clip for every class, c_in -> c_out.
int fused_step(hipStream_t s, const StepArgs &a, const Fields &f, TileShape ts);
// Flux form of the same step (f3d_flux32.hip): tiles 58 x ty with ty in
// {10, 18, 26}.
bool flux_supported(TileShape ts);
int flux_step(hipStream_t s, const StepArgs &a, const Fields &f, TileShape ts);
// Batched form of the flux step: every (scenario, tile, slab group) of a set
// of FP32 contexts as one work list, drawn by persistent workgroups.
// Scenarios fall in two parity groups (group[k]): in a step one group takes
// an odd and the other an even step, so contexts on different iiter
// parities share the batch.
struct FluxBatchPlan {
    TileShape ts{58, 18, 0, 2};
    DeviceBuffer par[2];   // kernel parameters per scenario, even / odd iiter
    DeviceBuffer items;    // per group: interior items, then edge items (int4 each)
    DeviceBuffer xoff;     // per list: the offsets of its per-XCD sublists
    DeviceBuffer counter;  // work counters of the four lists' per-XCD sublists
    int nint[2] = {0, 0}, nedge[2] = {0, 0};
    int off_int[2] = {0, 0}, off_edge[2] = {0, 0};
    int grid_int = 0, grid_edge = 0;  // persistent grid sizes (resident capacity)
    bool mixed = false;  // one list of interior and edge items per group
};
int flux_batch_plan(FluxBatchPlan &b, const std::vector<StepArgs> &even,
                    const std::vector<StepArgs> &odd, const std::vector<Fields> &f,
                    const std::vector<int> &group, TileShape ts);
// zero the counters; group 0 steps with parity odd0, group 1 with the other;
// interior lists on s, edge lists on side (fork, then join unless the caller
// joins the side stream itself)
int flux_batch_step(hipStream_t s, hipStream_t side, hipEvent_t fork, hipEvent_t join,
                    const FluxBatchPlan &b, bool odd0, bool join_side = true);
// cmass2d outflux (cmass2d.f90:38-105) of a step, after its flux launch:
// the clipped x-face ghosts from the ghost buffer, the y-face and top ghosts
// from c_in's faces; outmass[m] += member m's outflux (deterministic per
// member).
int outflux(hipStream_t s, const StepArgs &a, const Fields &f, const double *vx64,
            const double *vy64, const double *vz64, const double *vset64, const double *dz,
            double *scratch, double *outmass);
// the same for many contexts in two launches: fa[k] describes context k's
// step, rows[y] = (context, member) for y < nrows; each member's partials
// and tree are the ones outflux uses
// n / d for n < 2^31 by a multiply-high: division by an invariant integer
// (Granlund and Montgomery, PLDI 1994), with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1, n / d = (umulhi(n, m) + n) >> l.
struct FastDiv {
    unsigned d, m, l;
};
inline FastDiv fast_div(unsigned d)
{
    unsigned l = 0;
    while ((1ull << l) < d) ++l;
    const unsigned long long m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
    return FastDiv{d, (unsigned)m, l};
}
__device__ __forceinline__ unsigned fdiv(unsigned n, const FastDiv &f)
{
    return (__umulhi(n, f.m) + n) >> f.l;
}
struct FrameArgs {
    Grid g;
    FastDiv per, ny, nx;  // face cells per class (ny*nz + nx*nz + nx*ny), ny, nx
    const float *gbuf;  // the step's x-face ghosts (StepArgs::gbuf)
    const float *cin;   // its input, whose faces give the y-face and top ghosts
    const float *tfac;  // top ghost factor per class (Fields::tfac)
    const double *vx, *vy, *vz, *vset, *dz;
    double dt;
    double *part;     // outflux_blocks(g) per member, at most kFramePartials
    double *outmass;  // per member
};
// Workgroups (and partial sums) of one member's outflux, 256 lanes each:
// one per 2048 face cells, at most red::kFramePartials, so that a launch of
// one or a few members still spreads over the device.  A function of the
// grid alone, so a member's budget is the same bits alone, in an ensemble or
// in a batch.
__host__ __device__ inline int outflux_blocks(const Grid &g)
{
    const long cells = ((long)g.ny * g.nz + (long)g.nx * g.nz + (long)g.nx * g.ny) * g.nc;
    const long nb = (cells + 2047) / 2048;
    return (int)(nb < 1 ? 1 : (nb > red::kFramePartials ? red::kFramePartials : nb));
}
FrameArgs frame_args(const StepArgs &a, const Fields &f, const double *vx64, const double *vy64,
                     const double *vz64, const double *vset64, const double *dz, double *scratch,
                     double *outmass);
// rows: (context, member); rofs: nrows+1 prefix sums of the rows'
// outflux_blocks, the last one the launch's workgroup count
int outflux_batch(hipStream_t s, const FrameArgs *fa, const int2 *rows, const int *rofs, int nrows,
                  int nblocks);
// Completes the ghost frame of a state c the flux kernel wrote from cin
// (which must still hold that step's input): the x-face ghosts from gbuf,
// the rest of setbcc (setbcc.f90:19-195) formed from cin's faces, clipped
// (sizloop.f90:77-78), and the frame's edges (which setbcc never sets)
// clipped, for every member*class slab.
struct FrameFill {
    Grid g;
    float *c;
    const float *cin, *gbuf;
    const double *vy;
    const float *bfac, *tfac;  // Fields::bfac, Fields::tfac
};
int frame_fill(hipStream_t s, const FrameFill &a);
// setsrc for every (member, class); off[m]..off[m+1] are member m's points,
// grp/perm group them by target cell in point order
int setsrc(hipStream_t s, const Grid &g, float *c, const int *off, const int4 *grp, int ngrp,
           const int *perm, const int *isrc, const int *jsrc, const int *ksrc,
           const double *tmrat, const double *zlayer, const double *dz, const double *hm1,
           double dt);
// the same for many contexts in one launch: sa[k] describes context k's
// sources, its items are [first[k], first[k+1]) of nitems = first[nctx]
struct SrcArgs {
    Grid g;
    float *c;
    const int *off;
    const int4 *grp;
    int ngrp;
    const int *perm, *isrc, *jsrc, *ksrc;
    const double *tmrat, *zlayer, *dz, *hm1;
    double dt;
};
SrcArgs src_args(const Grid &g, float *c, const int *off, const int4 *grp, int ngrp,
                 const int *perm, const int *isrc, const int *jsrc, const int *ksrc,
                 const double *tmrat, const double *zlayer, const double *dz, const double *hm1,
                 double dt);
int setsrc_batch(hipStream_t s, const SrcArgs *sa, const int *first, int nctx, int nitems);
int to_fp64(hipStream_t s, const float *c32, double *c64, long n);
int from_fp64(hipStream_t s, const double *c64, float *c32, long n);

}  // namespace f32

}  // namespace f3d

// The context
struct f3d_ctx {
    int precision = F3D_FP64;
    f3d::Grid g{};
    int lonlat = 0;
    // the spacing kappa3's CMAQ Khn takes: degrees on LON-LAT grids, m on UTM
    double khn_dx = 0.0, khn_dy = 0.0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;                     // edge tiles of the flux kernel
    unsigned gen = 0;  // bumped when fields or sources change (batched steps re-plan)
    // bumped by every API entry on the context (check_ctx); a batch orders
    // its stream after the context's streams again only when this moved, or
    // another batch stepped the context, since it last did
    mutable unsigned long api_seq = 0;
    const void *last_batch = nullptr;  // the batch whose work the streams last waited for
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    long long bytes = 0;

    // host copies
    std::vector<double> zlayer, dz, ldz, hm1, kmh;
    std::vector<f3d::ZCoef> zc;
    bool uniform_hm1 = true;

    // classes and models
    int modv = 2;
    bool classes_set = false;
    std::vector<f3d::met::ClassConst> cls;
    int modkv = 0, modkh = 0;
    double rkv0 = 0.0, rkh0 = 0.0, csh = 0.2275;

    // device: grid
    f3d::DeviceBuffer d_dz, d_ldz, d_zlayer, d_hm1, d_kmh, d_zc;
    // device: meteo and coefficients (FP64)
    f3d::DeviceBuffer d_vx, d_vy, d_vz, d_tempe, d_vptemp, d_rho, d_ustar, d_zi, d_rl, d_vdep;
    // land use (USGS category) and roughness length per column: when set,
    // every meteo read computes vdep on the device (meteo.f90:99-100)
    f3d::DeviceBuffer d_ldu, d_znt;
    bool landuse_set = false;
    // terrain height per column (nx,ny): on LON-LAT grids every meteo read
    // corrects vz for the terrain slope (corvver, meteo.f90:116-124)
    f3d::DeviceBuffer d_topg;
    bool topg_set = false;
    // pressure (nx,ny,nz) in Pa (Master.f90:103), for the flight-level cuts
    f3d::DeviceBuffer d_p;
    bool p_set = false;
    // precipitation rate (nx,ny, mm/h) and per-class wet flags (diameter
    // up to 100 um): wet deposition in the FP64 step (sizloop.f90:65-67)
    f3d::DeviceBuffer d_prate, d_wet;
    f3d::DeviceBuffer d_lam, d_kwet;  // FP32 step: per-column lambda and top wet level
    f3d::DeviceBuffer d_products;  // output products of one member (lazily allocated)
    bool wet_on = false;
    std::vector<double> diam_h;
    f3d::DeviceBuffer d_rkhor, d_rkh1, d_rkver, d_vset, d_visc;
    f3d::DeviceBuffer d_cls, d_cd, d_status, d_dtcn;
    // device: state
    f3d::DeviceBuffer d_c64;          // FP64 path
    f3d::DeviceBuffer d_c32[2];       // FP32 path ping-pong
    int cur = 0;                      // FP32: index of the current state buffer
    // FP32: the last step's x- and y-face ghosts (f32::StepArgs::gbuf);
    // frame_dirty while they are not yet in d_c32[cur]'s frame
    f3d::DeviceBuffer d_gbuf;
    bool frame_dirty = false;
    f3d::DeviceBuffer d_cdry, d_cwet;
    f3d::DeviceBuffer d_scratch, d_outmass, d_budget;
    // FP32 step fields
    f3d::DeviceBuffer d_f32, d_vsum1;
    f3d::f32::TileShape tile{58, 18, 0, 2};

    // source points per member (host copies) and their device concatenation:
    // member m owns points [src_off[m], src_off[m+1])
    std::vector<std::vector<int>> src_i, src_j, src_k;
    std::vector<std::vector<double>> src_tm;  // (nsou_m, nc) per member
    std::vector<double> flowrate;             // per member
    int nsou = 0;                             // total points
    f3d::DeviceBuffer d_isrc, d_jsrc, d_ksrc, d_tmrat, d_srcoff;
    f3d::DeviceBuffer d_srcgrp, d_srcperm;  // source points grouped by target cell
    int nsrcgrp = 0;

    // clock (shared by the members of an ensemble: one dt)
    double time = 0.0, dt = 0.0, dtc = 0.0;
    std::vector<double> erumass;              // per member
    int iiter = 0;
    bool meteo_ready = false;
    double cd_state = 1.0;

    // gravity current (f3d_set_gravity_current): parameters, source
    // heights, the wind divergence of the last meteo read (setdivu), the
    // radial wind's scalars of that read and the last divcorr factor
    bool gravity_on = false, grav_active = false;
    f3d_gravity grav{};
    std::vector<double> grav_zs;
    f3d::DeviceBuffer d_divu;
    std::array<double, 9> grav_diag{};
    double grav_corrdiv = 0.0;

    // timing of the dominant step kernel
    bool timing = false;
    std::vector<hipEvent_t> ev_start, ev_stop;
    int ev_next = 0;
    double t_ms = 0.0;
    long t_count = 0;
};

// Pieces of the FP32 step, shared by f3d_step and the batched step.
namespace f3d {
f32::Fields fp32_fields(const f3d_ctx *c);
// step arguments reading state buffer cur_in and writing the other
f32::StepArgs fp32_args(f3d_ctx *c, int cur_in, int iiter);
int fp32_setsrc(f3d_ctx *c, hipStream_t s, const f32::StepArgs &a);
int fp32_outflux(f3d_ctx *c, hipStream_t s, const f32::StepArgs &a);
// gravity current: the radial wind and the divergence of a meteo read
// (after the Hm1 scaling and corvver), and divcorr after a step at the log
// times
int gravity_meteo(f3d_ctx *c);
int gravity_endstep(f3d_ctx *c);
int fp32_complete_frame(f3d_ctx *c);
f32::FrameArgs fp32_frame_args(f3d_ctx *c, const f32::StepArgs &a);
f32::SrcArgs fp32_src_args(f3d_ctx *c, const f32::StepArgs &a);
}  // namespace f3d

// non-null and alive; sets the error message otherwise
bool check_ctx(const f3d_ctx *c);
