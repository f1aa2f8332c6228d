// Fused FP32 transport step in flux form (gfx950).
//
// The step of the tile and row kernels -- setbcc, advctzc, the parity-ordered
// horizontal advection, diffz, the parity-ordered horizontal diffusion, accum
// and the clip, for every class, c_in -> c_out -- written so that every sweep
// evaluates each limiter and each flux once:
//
// * advctx/advcty.  With A = dt/h*v, F the limiter face factor
//   0.5*dt/h*(1-dt/h*|v12|)*|v12| and Phi(i) = A(i)*c(i) + F(i)*minmod(c(i+1)
//   -c(i), c(i)-c(i-1)), the update of advctx.f90:40-56 is c + Phi(i-1) -
//   Phi(i) where v(i) > 0 and c + Phi(i) - Phi(i+1) elsewhere.  The minmod
//   of the reference, 0.5*(sign(a)+sign(b))*min(|a|,|b|), is med3(a, b, 0).
// * diffx/diffy.  With G(i) = D(i-1/2)*(c(i)/rho(i) - c(i-1)/rho(i-1)), D the
//   face-averaged rho*K times dt/h^2, the update of diffx.f90:30-44 is
//   c + G(i+1) - G(i).
// * advctzc and diffz run down each column in registers; advctzc keeps the
//   combined flux U(k) = W1(k) + W2(k) - W2(k-1), whose upwind choice gives
//   the update c - U(k) (w > 0) or c - U(k+1), and diffz reuses the face flux
//   of the level below.
//
// Layout: a tile is 58 columns by RY-6 rows with a halo of 3.  A region row
// is 64 columns, one wavefront, lane = column, so x-neighbours come from DPP
// lane shifts (no LDS, no barrier).  Wave w owns NS = RY/NW contiguous rows;
// y-neighbours are its own registers except across band edges, which go
// through three small LDS planes.  Diffusion lags advection by one level and
// the last sweep by two, so a level costs two workgroup barriers.  Tiles whose
// region meets the domain frame take the EX / EY paths (ghost columns / rows
// from setbcc, line-end clamps of the limiter and of the diffusion faces).
#include <algorithm>
#include <vector>

#include "f3d_ctx.h"

namespace f3d {
namespace f32 {
namespace {

constexpr int kXCD = 8;
constexpr int RX = 64;       // region row = one wavefront
constexpr int TXR = RX - 6;  // output columns per tile
static_assert(TXR == kFluxTileCols, "the field layout's tile columns are the kernel's");
constexpr int kMaxLevels = 64;
// s_waitcnt immediate of gfx9: vmcnt(0), expcnt and lgkmcnt left open
constexpr int kWaitVm0 = 0x0F70;
// cache-policy operand of the buffer builtins on gfx950: the nt bit
constexpr int kNonTemporal = 2;

struct FluxParams {
    Grid g;
    const float *cin;
    float *cout;
    double *cdry;
    // all FP32 fields live in one buffer; byte offsets of the groups in it:
    // cA (ax, ay: 8-byte records), cD (1/rho, rho*rkver, rho*rkhor:
    // 12-byte records) and wz (vz+vset, class 0), all three in the
    // tile-major layout (one region row = 64 consecutive, aligned records;
    // ftl records per level); bfac and tfac (class 0) and h2
    // (1/Hm1^2 per column: rho*rkh1 = rho*rkhor/Hm1^2, meteo.f90:108-112,
    // so no rho*rkh1 field is read; 1 on UTM grids).  The face
    // coefficients (limiter factors, face-averaged diffusivities) are formed
    // from the two cells on either side of the face, which the kernel holds
    // anyway, so no per-face field is read.  The kernel reads each record
    // whole, once per level: a partial read of a cache line would fetch the
    // line again for the rest.
    const float *fbase;
    unsigned fbytes;
    unsigned oA, oD, oW, oBF, oTF, oH;
    unsigned ftx, ftl;  // tile columns and records per level of the tile-major layout
    const double *vsum1;
    // wet deposition (sizloop.f90:65-67), read by the WET tile code only
    const double *lam;
    const int *kwet, *wetcls;
    double *cwet;
    const double *dz64;
    // the x-face ghosts of c_out (one column per level, a cache line per
    // cell in c's layout) go to a compact buffer, per slab two faces (x low,
    // x high) of nz levels of ny cells, gslab floats; the rest of the frame
    // follows from c_in's faces, where the outflux and the frame
    // completion read it
    float *gbuf;
    unsigned gslab;
    const ZLev *zl;
    double dt64;
    float dt, dtdx2, dtdy2;
    int ntx, nty;
    int bx0, bx1, by0, by1;  // the rectangle of interior tiles
    int nlogical;
    int per_xcd;
    // K_ALL launches: the interior part's logical blocks and their count per
    // XCD (nlogical and per_xcd then count the edge part)
    int nlogical_i;
    int per_xcd_i;
};

// value held by lane i-1 (lane 0 reads 0)
__device__ __forceinline__ float from_left(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}
// value held by lane i+1 (lane 63 reads 0)
__device__ __forceinline__ float from_right(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, true));
}

__device__ __forceinline__ float minmod(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.f); }

// Limiter factor of the face between two cells with Courant numbers A and
// Am = dt/h*v (advctx.f90:33-43): with a = 0.5*|A + Am| the Courant number
// of the face velocity, the second-order flux is 0.5*a*(1-a) times the
// minmod of the two differences.  At the line start the reference takes
// Am = A (v(max(1,i-1))), which the clamped column of the ghost supplies.
__device__ __forceinline__ float face_factor(float A, float Am)
{
    const float a = 0.5f * fabsf(A + Am);
    return 0.5f * a * (1.f - a);
}

// advctx on one region row.  EX: lo/hi mark the first/last domain column,
// where the reference clamps W2 to the line (advctx.f90:50-52) and A to
// v(1) / v(nx); cells outside the domain keep their value.
template <bool EX>
__device__ __forceinline__ float adv_x(float a, float A, float Fc, bool inner, bool lo, bool hi)
{
    const float dr = from_right(a) - a;
    const float dl = from_left(dr);
    const float w2 = Fc * minmod(dr, dl);
    const float phi = fmaf(A, a, w2);
    float pl = from_left(phi), pr = from_right(phi);
    if (EX) {
        const float al = from_left(a), ar = from_right(a);
        pl = lo ? fmaf(A, al, w2) : pl;
        pr = hi ? fmaf(A, ar, w2) : pr;
    }
    const float out = (A > 0.f) ? a + (pl - phi) : a + (phi - pr);
    return EX ? (inner ? out : a) : out;
}

// diffx on one region row.  EX: at the line ends the reference uses the
// un-averaged products rho*K and one-sided fluxes clamped by sign
// (diffx.f90:20-29, 45-53).  The per-row factors of a level are formed once
// for the row's slabs (XDiff): the left face's factor Dl and clamp ml and the
// right face's Dr and mr.  In the interior, Dl = D (the face-averaged
// product) and Dr = D of the right neighbour, with open clamps (-inf, +inf),
// so Dl*(q - ql) and Dr*(qr - q) are the same products the interior
// formula forms (D*(q - ql) here and on the right neighbour); at the first
// column Dl = dtdx2*pk with the flux clamped to inflow (max 0), Dr =
// dtdx2*pk of the right neighbour; at the last, Dl = dtdx2*pk of the left
// neighbour, Dr = dtdx2*pk with the flux clamped to outflow (min 0).  Cells
// outside the domain keep their value (a y-ghost row's q feeds diffy of the
// first domain row).  Every lane shift is taken before the per-lane
// choices: a shift evaluated under a lane-dependent branch would read 0 from
// the lanes the branch leaves out.
struct XDiff {
    float Dl, Dr, ml, mr;
};
template <bool EX>
__device__ __forceinline__ XDiff diff_x_factors(float D, float pk, float dtdx2, bool lo, bool hi)
{
    XDiff k;
    k.Dl = D;
    if (!EX) return k;
    const float pkl = from_left(pk), pkr = from_right(pk), Dr = from_right(D);
    k.Dl = lo ? dtdx2 * pk : (hi ? dtdx2 * pkl : D);
    k.Dr = lo ? dtdx2 * pkr : (hi ? dtdx2 * pk : Dr);
    k.ml = lo ? 0.f : -__builtin_inff();
    k.mr = (hi && !lo) ? 0.f : __builtin_inff();
    return k;
}
template <bool EX>
__device__ __forceinline__ float diff_x(float d, float lr, const XDiff &k, bool inner)
{
    const float q = d * lr;
    const float ql = from_left(q);
    if (!EX) {
        const float G = k.Dl * (q - ql);
        return d + (from_right(G) - G);
    }
    const float qr = from_right(q);
    const float gl = fmaxf(k.Dl * (q - ql), k.ml);
    const float gr = fminf(k.Dr * (qr - q), k.mr);
    return inner ? d + (gr - gl) : d;
}

// Buffer resources (gfx9 raw buffers: stride 0, range in bytes): loads and
// stores take the lane's 32-bit byte offset in a VGPR and the level offset
// in an SGPR, so no per-lane 64-bit address is ever formed.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef __amdgpu_buffer_rsrc_t rsrc_t;

__device__ __forceinline__ rsrc_t mkbuf(const void *base, size_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0,
                                             (int)(unsigned)bytes, 0x00020000);
}
__device__ __forceinline__ float bld(rsrc_t r, unsigned voff, unsigned soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
// c_in, read once per step apart from the halo rows of neighbouring tiles:
// non-temporal, so it does not displace the coefficient records and vz+vset
// (read by every slab group of a tile, and by every ensemble member) in L2
__device__ __forceinline__ float bldnt(rsrc_t r, unsigned voff, unsigned soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, kNonTemporal));
}
__device__ __forceinline__ float2 bld2(rsrc_t r, unsigned voff, unsigned soff)
{
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
    return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
typedef unsigned u32x3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ float3 bld3(rsrc_t r, unsigned voff, unsigned soff)
{
    const u32x3 v = __builtin_amdgcn_raw_buffer_load_b96(r, (int)voff, (int)soff, 0);
    return make_float3(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z));
}
__device__ __forceinline__ float4 bld4(rsrc_t r, unsigned voff, unsigned soff)
{
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                       __uint_as_float(v.w));
}
// c_out is written once per step and read only by the next one: non-temporal.
// The whole offset is in the VGPR: a lane that must not write carries
// kNoStore, which stays past the buffer's range whatever level offset is
// added (slabs are under 2 GiB), so the range check drops the store and no
// branch is needed around it.
constexpr unsigned kNoStore = 0x80000000u;
__device__ __forceinline__ void bst(float x, rsrc_t r, unsigned voff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), r, (int)voff, 0, kNonTemporal);
}

// per-row flags of edge tiles
enum : unsigned {
    F_INNER = 1u,  // domain cell
    F_XLO = 2u, F_XHI = 4u, F_YLO = 8u, F_YHI = 16u,  // first/last domain column/row
    F_GHOST = 32u,  // setbcc face ghost
    F_GX = 64u,     // ... on an x face (else a y face)
    F_GLO = 128u,   // ... on the low face (i = 0 or j = 0)
    F_OWN = 256u,   // a domain cell of the tile's own columns and rows (stored)
};

// LDS planes of one tile: SP, SF, SQ (one per slab), SA, SY and, on odd
// steps, the two SC planes
constexpr int flux_planes(int mc, bool odd) { return 3 * mc + 2 + (odd ? 2 : 0); }
// edge tiles: the x-face ghosts of every level, staged in LDS and written
// out once per tile (a ghost column is one lane of a region row, so a store
// per row and level would cost a memory instruction for one value):
// [face][region row][level][slab], then [face][region row] the row's
// byte offset in the ghost buffer (kNoStore where the tile does not own it)
constexpr int flux_xstage(int ry, int mc) { return 2 * ry * kMaxLevels * mc + 2 * ry; }

template <int RY, int NW, int MC, bool ODD, bool EX, bool EY, bool WET, bool SCL = false>
struct FluxTile {
    static constexpr int NS = RY / NW;
    static constexpr bool EDGE = EX || EY;
    // SCL: the group's two slabs are two members of one class (ensembles
    // with an even member count take slabs class-major): they share the
    // class's vz+vset loads, its ghost factors and advctzc's face factors
    static constexpr bool SC2 = SCL && MC == 2;
    static constexpr int cv(int c) { return SC2 ? 0 : c; }
    // groups of at most 768 lanes (a budget of 168 registers) run the
    // steady levels in pairs, prefetch c a level early and store each
    // level's outputs after its loads, interior and edge tiles alike; the
    // single and the batched launch run the same tile code, so a scenario's
    // result is the same bits in both
    static constexpr bool PIPE = NW * 64 <= 768;
    static constexpr int PL = RX * RY;  // one LDS plane
    static_assert(RY % NW == 0, "region rows must divide evenly over the waves");

    const FluxParams &p;
    int nz, li, w;
    unsigned NXY, SZ;
    float *SP, *SF, *SQ;  // MC planes each
    // ay of each band's top row (the face below the next band's first row),
    // and rho*rkhor of the level awaiting its last sweep (every row: held in
    // a register, the lane would outlive the rest of its record and be
    // copied out of the next load into the same registers, which waits for
    // that load)
    float *SA, *SY;
    // odd steps: (rho*rkh1, 1/rho) of the level awaiting its last sweep, own
    // rows only (kept in LDS rather than registers, which would keep the
    // loaded record live into a third level)
    float *SC;
    float *SG;            // x-face ghost staging (edge tiles)
    rsrc_t rf;            // the field buffer
    const ZLev *zt;       // the z-level table, staged in LDS
    float dtop, ldz0;
    // per class: slabs of c_in / c_out, offsets of its per-class fields,
    // class and member index; a pair's second class may be a duplicate of
    // the first (odd class count), computed but not stored
    rsrc_t rin[MC], rout[MC];
    unsigned oW[MC], oBF[MC], oTF[MC];
    int icv[MC], memv[MC];
    bool live[MC];

    // per-row geometry: byte offsets of the (clamped) column in c, in the
    // meteo arrays and of the rows above/below; edge flags
    // m4: the column in 2-D fields; mt4/mt8/mt12: the lane's record in the
    // tile-major layout, times the record size
    unsigned co4[NS], so4[NS], m4[NS], mt4[NS], mt8[NS], mt12[NS], fl[NS];
    unsigned FTL;  // records per level of the tile-major layout
    int lj[NS];
    int gjb;  // the region's first row in the grid (wave-uniform)
    // the row is one of the tile's own rows (rows 3 .. RY-4; from row 1 in
    // the first tile row, whose row 0 is the ghost row 0; up to row ny in
    // the last tile row): wave-uniform
    bool orow[NS];
    // Class-independent coefficients: (ax, ay) of the current and the next
    // level, (lr, pz, pky) of the current level and the one below, and the
    // column's 1/Hm1^2, which turns pky = rho*rkhor into pkx = rho*rkh1.
    // Every load fills a register after the last use of the value it
    // replaces, every loaded lane is used (a dead lane is written back late
    // and makes its register's next use wait), and each value is loaded a
    // level before it is needed.  gv: wind at the face cell of a ghost (edge
    // tiles), one level ahead.
    float2 CA[NS], CAn[NS];
    float3 Dn[NS], D0[NS];
    float h2[NS];
    float gv[NS];
    int kwl[NS];  // WET: the column's top wet level (0: dry column)
    // per class: advctzc window c(s), c(s+1), w(s), U(s), W2(s), delta(s) (after
    // level nz, c0 holds the top ghost c(nz+1)), prefetched c(s+2), w(s+1);
    // diffz window b(k), q(k) = b/rho and the face flux it leaves for the
    // level above (before level 1: b(0)); carried state and q of the level
    // awaiting its last sweep; the advected value and its y flux
    float c0[MC][NS], cp1[MC][NS], vk[MC][NS], Uc[MC][NS], W2k[MC][NS], dprev[MC][NS];
    float nc2[MC][NS], nwz[MC][NS];
    float b0[MC][NS], q0[MC][NS], gprev[MC][NS];
    float dc[MC][NS], qc[MC][NS], P[MC][NS], phi[MC][NS];
    // clipped outputs of the last sweep, awaiting flush()
    float sto[MC][NS];
    // edge rows (EY): the y fluxes with W2 clamped to the line, formed in
    // phase 2 (phase 3 must not read the SP plane: the next level's phase 1
    // may already be rewriting it)
    float apl[MC][NS], apr[MC][NS];

    __device__ __forceinline__ FluxTile(const FluxParams &pp, const int *qv, int gi0, int gj0,
                                        float *sm, const ZLev *zt_)
        : p(pp)
    {
        zt = zt_;
        const int nx = p.g.nx, ny = p.g.ny, nc = p.g.nc;
        nz = p.g.nz;
        NXY = (unsigned)p.g.NXY;
        SZ = (unsigned)((nx + 2) * (ny + 2));
        li = threadIdx.x & (RX - 1);
        w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        gjb = gj0;
        SP = sm;
        SF = sm + MC * PL;
        SQ = sm + 2 * MC * PL;
        SA = sm + 3 * MC * PL;
        SY = SA + PL;
        SC = SY + PL;
        SG = sm + flux_planes(MC, ODD) * PL;
        const size_t G = (size_t)p.g.G, N = (size_t)p.g.N;
        rf = mkbuf(p.fbase, p.fbytes);
        const int nm = p.g.nm, nq = nc * nm;
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            live[c] = qv[c] < nq;
            // slabs are taken class-major (slab s = class*nm + member; with
            // one member s is the class), so the slabs of a group and the
            // groups of a tile that run at once are members of few classes;
            // the state and deposits stay member-major (q)
            const int s = live[c] ? qv[c] : qv[0];
            icv[c] = s / nm;
            memv[c] = s - icv[c] * nm;
            const int q = memv[c] * nc + icv[c];
            rin[c] = mkbuf(p.cin + (size_t)q * G, 4 * G);
            // the duplicate slab of an odd class count writes nothing
            rout[c] = mkbuf(p.cout + (size_t)q * G, live[c] ? 4 * G : 0);
            oW[c] = p.oW + 4u * (unsigned)icv[c] * (unsigned)nz * p.ftl;
            oBF[c] = p.oBF + 4u * (unsigned)icv[c] * NXY;
            oTF[c] = p.oTF + 4u * (unsigned)icv[c] * NXY;
        }
        const int gi = gi0 + li;
        FTL = p.ftl;
        const unsigned btx = (unsigned)(gi0 / TXR);  // the tile's column
        // own lanes and rows: 3 .. R-4 of a region of R, widened to the ghost
        // at the domain's ends (the first tile starts at the ghost column or
        // row 0, the last reaches the ghost n+1 in its last lane or row)
        const int oxl = gi0 <= 0 ? 1 : 3, oxh = (nx - gi0 <= RX - 2) ? nx - gi0 : RX - 4;
        const int oyl = gj0 <= 0 ? 1 : 3, oyh = (ny - gj0 <= RY - 2) ? ny - gj0 : RY - 4;
        const bool olane = li >= oxl && li <= oxh;
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
            lj[sl] = w * NS + sl;
            const int gj = gj0 + lj[sl];
            const int ci = EX ? min(max(gi, 1), nx) : gi;
            const int cj = EY ? min(max(gj, 1), ny) : gj;
            const unsigned mo = (unsigned)((ci - 1) + nx * (cj - 1));
            co4[sl] = 4u * (unsigned)cidx(nx, ny, ci, cj, 0);
            m4[sl] = 4u * mo;
            const unsigned mt = ((unsigned)(cj - 1) * p.ftx + btx) * 64u + (unsigned)li;
            mt4[sl] = 4u * mt;
            mt8[sl] = 8u * mt;
            orow[sl] = lj[sl] >= oyl && lj[sl] <= oyh;
            // rows more than one away from the own rows reach them only
            // through the advection (c, vz+vset, ax, ay): their diffusion
            // feeds rows whose y diffusion is never stored.  Their diffusion
            // records all read the level's first record, so such a row
            // fetches one line per level instead of six.
            mt12[sl] = (lj[sl] < oyl - 1 || lj[sl] > oyh + 1) ? 0u : 12u * mt;
            unsigned f = 0;
            const bool inx = !EX || (gi >= 1 && gi <= nx);
            const bool iny = !EY || (gj >= 1 && gj <= ny);
            if (inx && iny) f |= F_INNER;
            if (EX && gi == 1) f |= F_XLO;
            if (EX && gi == nx) f |= F_XHI;
            if (EY && gj == 1) f |= F_YLO;
            if (EY && gj == ny) f |= F_YHI;
            // setbcc face ghosts (setbcc.f90:139-169): x faces inside the y
            // range, y faces inside the x range
            if (EX && iny && (gi == 0 || gi == nx + 1)) f |= F_GHOST | F_GX | (gi == 0 ? F_GLO : 0u);
            if (EY && inx && (gj == 0 || gj == ny + 1)) f |= F_GHOST | (gj == 0 ? F_GLO : 0u);
            if (olane && orow[sl] && (f & F_INNER)) f |= F_OWN;
            fl[sl] = f;
            // output cells: the tile's own columns, inside the domain
            so4[sl] = (olane && (f & F_INNER)) ? co4[sl] : kNoStore;
            // the row's place in the ghost buffer for each x face, where the
            // tile owns the face's ghost in this row: the column beside the
            // ghost is an output column and the row an output row
            if (EX && li == 0) {
                unsigned *so = reinterpret_cast<unsigned *>(SG) + 2 * RY * kMaxLevels * MC;
#pragma unroll
                for (int f = 0; f < 2; ++f) {
                    const int lg = (f == 0 ? 0 : nx + 1) - gi0;  // the ghost column's lane
                    const int la = lg + (f == 0 ? 1 : -1);
                    const bool own = lg >= 0 && lg < RX && la >= oxl && la <= oxh && orow[sl] &&
                                     gj >= 1 && gj <= ny;
                    so[f * RY + lj[sl]] = own ? 4u * ((unsigned)f * (unsigned)nz * (unsigned)ny + (unsigned)(gj - 1))
                                              : kNoStore;
                }
            }
        }
    }

    __device__ __forceinline__ bool has(int sl, unsigned b) const { return (fl[sl] & b) != 0u; }
    // the first / last domain row (y-side tiles): the same for every lane of
    // a row, so the y line-end clamps branch on it instead of selecting per
    // lane
    __device__ __forceinline__ bool row_ylo(int sl) const { return EY && gjb + lj[sl] == 1; }
    __device__ __forceinline__ bool row_yhi(int sl) const { return EY && gjb + lj[sl] == p.g.ny; }
    // offset of the wind at a ghost's face cell (ax for x faces, ay for y
    // faces) in the (ax, ay) records; only ghost lanes use it, so the others
    // read past the buffer's range, which returns 0 without a memory request
    // (an x-side row then fetches one line instead of four)
    __device__ __forceinline__ unsigned ghost_wind(int sl) const
    {
        return has(sl, F_GHOST) ? mt8[sl] + (has(sl, F_GX) ? 0u : 4u) : kNoStore;
    }
    __device__ __forceinline__ bool inner(int sl) const { return !EDGE || has(sl, F_INNER); }
    __device__ __forceinline__ float *row(float *plane, int c, int r) const
    {
        return plane + c * PL + r * RX + li;
    }

    __device__ __forceinline__ void init()
    {
        const ZLev z1 = zt[1];
        ldz0 = zt[0].ldz;
        dtop = p.dt * zt[nz].ldz;
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
            CA[sl] = bld2(rf, mt8[sl], p.oA);  // level 1
            CAn[sl] = make_float2(0.f, 0.f);
            Dn[sl] = bld3(rf, mt12[sl], p.oD);
            D0[sl] = make_float3(0.f, 0.f, 0.f);
            h2[sl] = bld(rf, m4[sl], p.oH);
            if (WET) kwl[sl] = p.kwet[m4[sl] / 4u];
            gv[sl] = EDGE ? bld(rf, ghost_wind(sl), p.oA) : 0.f;
            float bfv[MC], v1v[MC];
#pragma unroll
            for (int c = 0; c < MC; ++c) {
                bfv[c] = c == cv(c) ? bld(rf, m4[sl], oBF[c]) : bfv[cv(c)];
                v1v[c] = c == cv(c) ? bld(rf, mt4[sl], oW[c]) : v1v[cv(c)];
            }
#pragma unroll
            for (int c = 0; c < MC; ++c) {
                const float c1 = bld(rin[c], co4[sl], 4 * SZ);
                const float c2 = bld(rin[c], co4[sl], 8 * SZ);
                const float gb = c1 * bfv[c];  // c(0), setbcc.f90:108-113
                const float v1 = v1v[c];
                c0[c][sl] = c1;
                cp1[c][sl] = c2;
                if (c == cv(c)) vk[c][sl] = v1;
                const float d1 = (c2 - c1) * z1.dzk;
                const float d0 = (c1 - gb) * ldz0;
                const float av = fabsf(v1);
                W2k[c][sl] = minmod(d1, d0) * av * (z1.a2 - z1.a2bc * av);
                Uc[c][sl] = z1.dtdzz * (v1 * c1 - v1 * gb);  // U(1) = W1(1)
                dprev[c][sl] = d1;
                gprev[c][sl] = gb;
                nc2[c][sl] = (3 <= nz) ? bld(rin[c], co4[sl], 12 * SZ) : 0.f;
                if (c == cv(c)) nwz[c][sl] = bld(rf, mt4[sl], oW[c] + 4u * FTL);
                b0[c][sl] = q0[c][sl] = 0.f;
                dc[c][sl] = qc[c][sl] = 0.f;
                P[c][sl] = phi[c][sl] = 0.f;
                sto[c][sl] = 0.f;
                apl[c][sl] = apr[c][sl] = 0.f;
            }
        }
    }

    // One level of the pipeline.  STEADY promises 4 <= s <= nz-3 (no level
    // is at a boundary of any stage).
    template <bool STEADY>
    __device__ __forceinline__ void level(const int s)
    {
        const int nzl = nz;

        // ======== phase 1: advctzc at level s (and advctx on even steps) ========
        if (STEADY || s <= nzl) {
            const ZLev zn = zt[s + 1];
            const unsigned lz = 4u * FTL * (unsigned)s;  // byte offset of level s+1 in a tiled field
            // c(s+3) first: carrying it into the window copies the load, so
            // it is issued a whole level before that copy
            float nc3[MC][NS];
            if (PIPE) {
#pragma unroll
                for (int sl = 0; sl < NS; ++sl)
#pragma unroll
                    for (int c = 0; c < MC; ++c)
                        nc3[c][sl] = (STEADY || s + 3 <= nzl)
                                         ? bldnt(rin[c], co4[sl], 4u * SZ * (unsigned)(s + 3)) : 0.f;
                if (STEADY) __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                const float Ax = CA[sl].x;
                const float Fx = ODD ? 0.f : face_factor(Ax, from_left(Ax));
                bool zero = false;
                if (EDGE) {
                    // ghost faces: zero when the wind at the face cell blows
                    // out of the domain through it, else the adjacent cell
                    const float v = gv[sl];
                    zero = has(sl, F_GLO) ? v >= 0.f : v <= 0.f;
                }
#pragma unroll
                for (int c = 0; c < MC; ++c) {
                    float cp2 = 0.f, vkn = 0.f, W1n, W2n, dn = 0.f;
                    const float vkc = vk[cv(c)][sl];
                    if (STEADY || s < nzl) {
                        // top ghost c(nz+1) = tfac*c(nz) (setbcc.f90:126-130)
                        cp2 = (STEADY || s + 2 <= nzl) ? nc2[c][sl]
                                                       : bld(rf, m4[sl], oTF[cv(c)]) * cp1[c][sl];
                        vkn = nwz[cv(c)][sl];
                        W1n = zn.dtdzz * (vkn * cp1[c][sl] - vkc * c0[c][sl]);
                        dn = (cp2 - cp1[c][sl]) * zn.dzk;
                        const float av = 0.5f * fabsf(vkn + vkc);
                        W2n = minmod(dn, dprev[c][sl]) * av * (zn.a2 - zn.a2bc * av);
                    } else {
                        W1n = dtop * vkc * (cp1[c][sl] - c0[c][sl]);  // advctzc.f90:175-181
                        W2n = W2k[c][sl];
                    }
                    const float Un = W1n + (W2n - W2k[c][sl]);
                    float a = c0[c][sl] - ((vkc > 0.f) ? Uc[c][sl] : Un);
                    if (EDGE) a = inner(sl) ? a : (has(sl, F_GHOST) && !zero ? c0[c][sl] : 0.f);
                    c0[c][sl] = cp1[c][sl];
                    cp1[c][sl] = cp2;
                    if (!SC2) vk[c][sl] = vkn;
                    Uc[c][sl] = Un;
                    W2k[c][sl] = W2n;
                    dprev[c][sl] = dn;
                    if (!ODD) a = adv_x<EX>(a, Ax, Fx, inner(sl), has(sl, F_XLO), has(sl, F_XHI));
                    P[c][sl] = a;
                }
                // a shared class window moves once both slabs have used it
                if (SC2) vk[0][sl] = (STEADY || s < nzl) ? nwz[0][sl] : 0.f;
            }
#pragma unroll
            for (int sl = 0; sl < NS; ++sl)
                if (sl == 0 || sl == NS - 1)
#pragma unroll
                    for (int c = 0; c < MC; ++c) *row(SP, c, lj[sl]) = P[c][sl];
            *row(SA, 0, lj[NS - 1]) = CA[NS - 1].y;
            // Reloads go after the last use of the registers they refill: a
            // load issued while the old value is still live lands in another
            // register and is copied back, and the copy waits for the load.
            if (STEADY) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                if (EDGE && (STEADY || s + 1 <= nzl))
                    gv[sl] = bld(rf, ghost_wind(sl), p.oA + 2u * lz);
#pragma unroll
                for (int c = 0; c < MC; ++c) {
                    if (PIPE)
                        nc2[c][sl] = nc3[c][sl];
                    else if (STEADY || s + 3 <= nzl)
                        nc2[c][sl] = bldnt(rin[c], co4[sl], 4u * SZ * (unsigned)(s + 3));
                    if ((STEADY || s + 2 <= nzl) && c == cv(c))
                        nwz[c][sl] = bld(rf, mt4[sl], oW[c] + lz + 4u * FTL);
                }
                if (STEADY || s + 1 <= nzl) CAn[sl] = bld2(rf, mt8[sl], p.oA + 2u * lz);
            }
            // x-face ghosts (the advection left them unchanged) to the LDS
            // stage
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                if (EX && has(sl, F_GHOST) && has(sl, F_GX)) {
                    float *st = SG + ((has(sl, F_GLO) ? 0 : RY) + lj[sl]) * (kMaxLevels * MC) + (s - 1) * MC;
#pragma unroll
                    for (int c = 0; c < MC; ++c) st[c] = P[c][sl];
                }
            }
        }
        __syncthreads();

        // ======== phase 2: advcty fluxes at level s; last sweep of level s-2 ========
        // The band-edge neighbours are read first so the LDS reads overlap.
        // Every row is computed (no wave-uniform skips, so no loop-carried
        // value is assigned under a branch); the region's first and last
        // rows read a clamped neighbour and their results are never used.
        float ePm[MC], ePp[MC], eqm[MC], eqp[MC], elr[NS], epk[NS];
        if (ODD) {
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                epk[sl] = *row(SC, 0, lj[sl]);
                elr[sl] = *row(SC, 1, lj[sl]);
            }
        }
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            ePm[c] = *row(SP, c, max(lj[0] - 1, 0));
            ePp[c] = *row(SP, c, min(lj[NS - 1] + 1, RY - 1));
            eqm[c] = *row(SQ, c, max(lj[0] - 1, 0));
            eqp[c] = *row(SQ, c, min(lj[NS - 1] + 1, RY - 1));
        }
        // ay of the row below the band (the region's first row has none and
        // pairs with itself: its fluxes are never used), rho*rkhor of the
        // rows below and above the band
        const float eay = *row(SA, 0, max(lj[0] - 1, 0));
        const float eym = *row(SY, 0, max(lj[0] - 1, 0));
        const float eyp = *row(SY, 0, min(lj[NS - 1] + 1, RY - 1));
        float pyk[NS];
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) pyk[sl] = *row(SY, 0, lj[sl]);
        if (STEADY || s <= nzl) {
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                const float ay = CA[sl].y;
                const float aym = (sl > 0) ? CA[sl - 1].y : (lj[0] > 0 ? eay : ay);
                const float Fy = face_factor(ay, aym);
#pragma unroll
                for (int c = 0; c < MC; ++c) {
                    const float Pm = (sl > 0) ? P[c][sl - 1] : ePm[c];
                    const float Pp = (sl < NS - 1) ? P[c][sl + 1] : ePp[c];
                    const float w2 = Fy * minmod(Pp - P[c][sl], P[c][sl] - Pm);
                    phi[c][sl] = fmaf(ay, P[c][sl], w2);
                    if (EY) {
                        // advcty.f90:49-51: W2 clamps to the line at j = 1 / ny,
                        // A at the ghost row is v(1) / v(ny)
                        apl[c][sl] = fmaf(ay, Pm, w2);
                        apr[c][sl] = fmaf(ay, Pp, w2);
                    }
                    if (sl == 0 || sl == NS - 1) *row(SF, c, lj[sl]) = phi[c][sl];
                }
            }
        }
        if (STEADY || s >= 3) {
            const int k = s - 2;
            const float d2 = p.dtdy2, hd2 = 0.5f * p.dtdy2;
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                if (!orow[sl]) continue;  // wave-uniform
                // face-averaged rho*rkhor times dt/dy^2 below and above the row
                const float pm = (sl > 0) ? pyk[sl - 1] : eym;
                const float pp = (sl < NS - 1) ? pyk[sl + 1] : eyp;
                const float Dl = hd2 * (pyk[sl] + pm);
                const float Du = hd2 * (pp + pyk[sl]);
                XDiff xk{};
                if (ODD)
                    xk = diff_x_factors<EX>(0.5f * p.dtdx2 * (epk[sl] + from_left(epk[sl])), epk[sl],
                                            p.dtdx2, has(sl, F_XLO), has(sl, F_XHI));
#pragma unroll
                for (int c = 0; c < MC; ++c) {
                    // diffy (diffy.f90:30-44; line ends 20-29, 45-53)
                    const float qm = (sl > 0) ? qc[c][sl - 1] : eqm[c];
                    const float qp = (sl < NS - 1) ? qc[c][sl + 1] : eqp[c];
                    const float q = qc[c][sl];
                    float glo = Dl * (q - qm);
                    float ghi = Du * (qp - q);
                    if (EY && row_ylo(sl)) {
                        glo = fmaxf(0.f, d2 * pyk[sl] * (q - qm));
                        ghi = d2 * pp * (qp - q);
                    } else if (EY && row_yhi(sl)) {
                        glo = d2 * pm * (q - qm);
                        ghi = fminf(0.f, d2 * pyk[sl] * (qp - q));
                    }
                    float out = dc[c][sl] + (ghi - glo);
                    // the last sweep's result is only stored, except on odd
                    // steps where diffx follows: a ghost column must then
                    // still hold its value for the first and last column
                    if (EX && ODD) out = inner(sl) ? out : dc[c][sl];
                    if (ODD) out = diff_x<EX>(out, elr[sl], xk, inner(sl));
                    // sizloop.f90:65-67, wetdep.f90:28-50 (WET launches only):
                    // below the PBL height a fraction lambda of c goes to cwet,
                    // level by level; the output lane owns its column's slot
                    if (WET && k <= kwl[sl] && p.wetcls[icv[c]] && stored(sl, c)) {
                        const size_t t2 = m4[sl] / 4u;
                        const double lam = p.lam[t2];
                        double *cw = p.cwet + ((size_t)memv[c] * (p.g.nc + 1) + icv[c] + 1) * NXY + t2;
                        *cw = *cw + p.dz64[k] * lam * (double)out;
                        out = (float)((double)out * (1.0 - lam));
                    }
                    // sizloop.f90:77-78; interior tiles store in flush()
                    if (PIPE)
                        sto[c][sl] = fmaxf(0.f, out);
                    else
                        bst(fmaxf(0.f, out), rout[c], so4[sl] + 4u * SZ * (unsigned)k);
                    if (!STEADY && k == 1 && stored(sl, c)) {
                        // accum.f90:22: deposit from c(:,:,1) before the clip, FP64
                        const size_t t2 = m4[sl] / 4u;
                        double *cd = p.cdry + ((size_t)memv[c] * (p.g.nc + 1) + icv[c] + 1) * NXY + t2;
                        *cd = *cd - p.dt64 * (double)out * p.vsum1[(size_t)icv[c] * NXY + t2];
                    }
                }
            }
        }
        __syncthreads();

        // ======== phase 3: advcty (+ advctx on odd steps) at level s; diffz and
        // the first horizontal diffusion at level s-1 ========
        if (!STEADY && s >= nzl + 2) {
            flush(s);
            return;
        }
        const int k = s - 1;
        const ZLev zk = zt[k > 0 ? k : 0];
        // band-edge fluxes first; every row is computed, as in phase 2
        float epl[MC], epr[MC];
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            epl[c] = *row(SF, c, max(lj[0] - 1, 0));
            epr[c] = *row(SF, c, min(lj[NS - 1] + 1, RY - 1));
        }
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
            const float lrn = Dn[sl].x, pzn = Dn[sl].y;  // level s (unused at s = nz+1)
            const float3 Dk = D0[sl];                    // level k (unused at s = 1)
            const float lr0 = Dk.x, pz0 = Dk.y;
            const float ay = CA[sl].y;
            const float Ax = CA[sl].x;
            const float Fx = ODD ? face_factor(Ax, from_left(Ax)) : 0.f;
#pragma unroll
            for (int c = 0; c < MC; ++c) {
                float bnew;
                if (STEADY || s <= nzl) {
                    const float Pc = P[c][sl], ph = phi[c][sl];
                    float pl = (sl > 0) ? phi[c][sl - 1] : epl[c];
                    float pr = (sl < NS - 1) ? phi[c][sl + 1] : epr[c];
                    if (EY) {
                        if (row_ylo(sl)) pl = apl[c][sl];
                        if (row_yhi(sl)) pr = apr[c][sl];
                    }
                    bnew = (ay > 0.f) ? Pc + (pl - ph) : Pc + (ph - pr);
                    if (EDGE) bnew = inner(sl) ? bnew : Pc;
                    if (ODD)
                        bnew = adv_x<EX>(bnew, Ax, Fx, inner(sl), has(sl, F_XLO), has(sl, F_XHI));
                } else {
                    bnew = c0[c][sl];  // b(nz+1): the top ghost of c_in
                }
                // diffz (diffz.f90:19-62) at level k = s-1, input b(k+1) = bnew
                if (!STEADY && s == 1) {
                    // bottom face: c(0)/rho(1), un-averaged rho*K(1), inflow only
                    const float q1 = bnew * lrn;
                    gprev[c][sl] = fmaxf(0.f, pzn * (q1 - gprev[c][sl] * lrn) * ldz0);
                    b0[c][sl] = bnew;
                    q0[c][sl] = q1;
                    continue;
                }
                float d, qn = 0.f;
                if (STEADY || k < nzl) {
                    qn = bnew * lrn;
                    const float dq = qn - q0[c][sl];
                    const float gav = (pzn + pz0) * zk.hldz * dq;
                    // level 1 uses the un-averaged rho*K(2) on its top face
                    const float gt = (!STEADY && k == 1) ? pzn * dq * zk.ldz : gav;
                    d = b0[c][sl] + zk.dtdzz * (gt - gprev[c][sl]);
                    // level nz takes the un-averaged rho*K(nz-1) on its bottom face
                    gprev[c][sl] = (!STEADY && k == nzl - 1) ? pz0 * dq * zk.ldz : gav;
                } else {
                    const float gt = fminf(0.f, pz0 * (bnew * lr0 - q0[c][sl]) * zk.ldz);
                    d = b0[c][sl] + dtop * (gt - gprev[c][sl]);
                }
                if (EDGE) d = inner(sl) ? d : b0[c][sl];
                q0[c][sl] = qn;
                b0[c][sl] = bnew;
                dc[c][sl] = d;  // first horizontal diffusion below (even steps)
            }
            if (!STEADY && s == 1) {
                D0[sl] = Dn[sl];
                Dn[sl] = bld3(rf, mt12[sl], p.oD + 12u * FTL);  // level 2 (nz >= 2)
                continue;
            }
            D0[sl] = Dn[sl];
            // first horizontal diffusion at level k: the face-averaged
            // rho*rkh1 times dt/dx^2 (diffx.f90:35-36); x-side tiles also
            // take the raw product at the line ends
            const float pkx = Dk.z * h2[sl];
            if (!ODD) {
                const XDiff xk = diff_x_factors<EX>(0.5f * p.dtdx2 * (pkx + from_left(pkx)), pkx, p.dtdx2,
                                                    has(sl, F_XLO), has(sl, F_XHI));
#pragma unroll
                for (int c = 0; c < MC; ++c) dc[c][sl] = diff_x<EX>(dc[c][sl], lr0, xk, inner(sl));
            } else {
                *row(SC, 0, lj[sl]) = pkx;
                *row(SC, 1, lj[sl]) = lr0;
            }
#pragma unroll
            for (int c = 0; c < MC; ++c) qc[c][sl] = dc[c][sl] * lr0;
            *row(SY, 0, lj[sl]) = Dk.z;
        }
        // after their last use: (lr, pz, pky) of level s+1; its (ax, ay)
        // arrived during phase 1
        if (STEADY) __builtin_amdgcn_sched_barrier(0);
        if (STEADY || s + 1 <= nzl) {
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                if (STEADY || s >= 2) Dn[sl] = bld3(rf, mt12[sl], p.oD + 12u * FTL * (unsigned)s);
                CA[sl] = CAn[sl];
            }
        }
#pragma unroll
        for (int sl = 0; sl < NS; ++sl)
            if ((sl == 0 || sl == NS - 1) && (STEADY || s >= 2))
#pragma unroll
                for (int c = 0; c < MC; ++c) *row(SQ, c, lj[sl]) = qc[c][sl];
        flush(s);
    }

    // PIPE tiles store the outputs of the last sweep (level s-2) after
    // every load of the level is issued: vector-memory counters complete in
    // order, so a store issued before a load makes every wait for that load
    // also wait for the store's write.  (1024-lane groups store at once:
    // holding the outputs costs registers they do not have.)
    __device__ __forceinline__ bool stored(int sl, int c) const
    {
        return has(sl, F_OWN) && (c == 0 || live[c]);
    }
    __device__ __forceinline__ void flush(const int s)
    {
        if (!PIPE || !(s >= 3)) return;
        const unsigned ko = 4u * SZ * (unsigned)(s - 2);
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
            if (!orow[sl]) continue;  // wave-uniform
#pragma unroll
            for (int c = 0; c < MC; ++c)
                bst(sto[c][sl], rout[c], so4[sl] + ko);
        }
    }

    // x-side tiles: the staged x-face ghosts of the tile's own rows,
    // clipped, to the ghost buffer ([face][level][row] per slab: each
    // level's rows are one run of consecutive cells; rows the tile does not
    // own carry kNoStore)
    __device__ __forceinline__ void flush_xstage()
    {
        __syncthreads();
        const int ny = p.g.ny;
        const unsigned *so = reinterpret_cast<const unsigned *>(SG) + 2 * RY * kMaxLevels * MC;
        // the ghost slabs of the group (a duplicate slab writes nothing)
        rsrc_t rgb[MC];
#pragma unroll
        for (int c = 0; c < MC; ++c)
            rgb[c] = mkbuf(p.gbuf + (size_t)(memv[c] * p.g.nc + icv[c]) * p.gslab,
                           live[c] ? 4 * (size_t)p.gslab : 0);
#pragma unroll
        for (int f = 0; f < 2; ++f) {
#pragma unroll
            for (int c = 0; c < MC; ++c) {
                for (int t = threadIdx.x; t < nz * RY; t += NW * 64) {
                    const int k = t / RY, r = t - k * RY;
                    // kNoStore plus a level offset stays past the range
                    const unsigned o = so[f * RY + r] + 4u * (unsigned)k * (unsigned)ny;
                    const float v = SG[(f * RY + r) * (kMaxLevels * MC) + k * MC + c];
                    bst(fmaxf(0.f, v), rgb[c], o);
                }
            }
        }
    }

    // Prologue, steady and epilogue levels run as three loops so each keeps
    // its own register allocation.
    __device__ __forceinline__ void run()
    {
        init();
        const int nzl = nz;
        int s = 1;
        {
            for (; s <= 3; ++s) level<false>(s);
            // drain the loads of the prologue: the steady loop then enters
            // with the same pending loads as its own back edge, and its
            // waits count only the loads issued one level back
            __builtin_amdgcn_s_waitcnt(kWaitVm0);
            // two levels per iteration: a register reloaded in one level
            // and its partner in the next swap roles, so no loaded value is
            // copied (a copy would wait for the load)
            if (PIPE)
                for (; s + 1 <= nzl - 3; s += 2) {
                    level<true>(s);
                    level<true>(s + 1);
                }
            for (; s <= nzl - 3; ++s) level<true>(s);
        }
        for (; s <= nzl + 2; ++s) level<false>(s);
        if (EX) flush_xstage();
    }
};

template <int RY, int NW, int MC, bool ODD, bool EX, bool EY, bool WET = false, bool SCL = false>
__device__ __forceinline__ void flux_tile(const FluxParams &p, const int *qv, const int gi0,
                                          const int gj0, float *sm, const ZLev *zt)
{
    FluxTile<RY, NW, MC, ODD, EX, EY, WET, SCL> t(p, qv, gi0, gj0, sm, zt);
    t.run();
}

// Tile kinds.  The rectangle [bx0, bx1] x [by0, by1] holds the interior
// tiles (region inside the domain); a column outside [bx0, bx1] meets an x
// face of the frame, a row outside [by0, by1] a y face.  K_INTERIOR launches
// the interior tiles; K_EDGE the x-side, y-side and corner tiles in one
// launch (so the small kinds do not each wait for a launch's tail), picking
// the tile's code path per workgroup; K_ALL both parts in one launch, each
// XCD taking its share of the edge tiles first and then its share of the
// interior tiles.
enum : int { K_INTERIOR = 0, K_EDGE = 1, K_ALL = 2 };

// the j-th edge column (outside [bx0, bx1]) / edge row (outside [by0, by1])
__device__ __forceinline__ int edge_col(const FluxParams &p, int c)
{
    return (c < p.bx0) ? c : p.bx1 + 1 + (c - p.bx0);
}
__device__ __forceinline__ int edge_row(const FluxParams &p, int r)
{
    return (r < p.by0) ? r : p.by1 + 1 + (r - p.by0);
}

// One workgroup per (tile of kind KIND, group of MC consecutive
// member*class slabs).  Consecutive tiles run down y, so the 6-row halos of
// neighbours meet in L2.
template <int RY, int NW, int MC, bool ODD, int KIND, bool WET, bool SCL = false>
__global__ void __launch_bounds__(NW * 64) k_flux(FluxParams p)
{
    const int b = blockIdx.x;
    int L;  // XCD-aware logical block of its part
    bool edge = KIND == K_EDGE;
    if (KIND == K_ALL) {
        const int x = b % kXCD, i = b / kXCD;
        edge = i < p.per_xcd;
        if (edge) {
            L = x * p.per_xcd + i;
            if (L >= p.nlogical) return;
        } else {
            L = x * p.per_xcd_i + (i - p.per_xcd);
            if (L >= p.nlogical_i) return;
        }
    } else {
        L = (b % kXCD) * p.per_xcd + b / kXCD;
        if (L >= p.nlogical) return;
    }
    const int nq = p.g.nc * p.g.nm;
    const int ngrp = (nq + MC - 1) / MC;
    const int grp = L % ngrp;
    int qv[MC];
#pragma unroll
    for (int c = 0; c < MC; ++c) qv[c] = grp * MC + c;
    const int t = L / ngrp;
    const int nmy = p.by1 - p.by0 + 1;          // interior rows (>= 0)
    const int nry = p.by0 + (p.nty - 1 - p.by1);  // edge rows
    const int nmx = p.bx1 - p.bx0 + 1;
    const int nex = p.ntx - nmx;
    int bx, by, kind = 0;  // 0 interior, 1 x side, 2 y side, 3 corner
    if (!edge) {
        by = p.by0 + t % nmy;
        bx = p.bx0 + t / nmy;
    } else {
        // x-side tiles first, then y-side, then corners, each down y
        int e = t;
        if (e < nex * nmy) {
            kind = 1;
            by = p.by0 + e % nmy;
            bx = edge_col(p, e / nmy);
        } else if ((e -= nex * nmy) < nmx * nry) {
            kind = 2;
            by = edge_row(p, e % nry);
            bx = p.bx0 + e / nry;
        } else {
            e -= nmx * nry;
            kind = 3;
            by = edge_row(p, e % nry);
            bx = edge_col(p, e / nry);
        }
    }
    const int gi0 = bx * TXR, gj0 = by * (RY - 6);
    __shared__ float sm[flux_planes(MC, ODD) * RX * RY + (KIND != K_INTERIOR ? flux_xstage(RY, MC) : 0)];
    __shared__ ZLev zt[kMaxLevels + 2];
    for (int k = threadIdx.x; k <= p.g.nz + 1; k += NW * 64) zt[k] = p.zl[k];
    __syncthreads();
    if (KIND == K_INTERIOR || (KIND == K_ALL && !edge))
        flux_tile<RY, NW, MC, ODD, false, false, WET, SCL>(p, qv, gi0, gj0, sm, zt);
    else if (kind == 1)
        flux_tile<RY, NW, MC, ODD, true, false, WET, SCL>(p, qv, gi0, gj0, sm, zt);
    else if (kind == 2)
        flux_tile<RY, NW, MC, ODD, false, true, WET, SCL>(p, qv, gi0, gj0, sm, zt);
    else
        flux_tile<RY, NW, MC, ODD, true, true, WET, SCL>(p, qv, gi0, gj0, sm, zt);
}

template <int RY, int NW, int MC, int KIND, bool WET, bool SCL = false>
void launch_part_t(hipStream_t s, FluxParams p, int ntiles, bool odd)
{
    if (ntiles <= 0) return;
    const int ngrp = (p.g.nc * p.g.nm + MC - 1) / MC;
    p.nlogical = ntiles * ngrp;
    p.per_xcd = (p.nlogical + kXCD - 1) / kXCD;
    const unsigned nb = (unsigned)(p.per_xcd * kXCD);
    if (odd)
        hipLaunchKernelGGL((k_flux<RY, NW, MC, true, KIND, WET, SCL>), dim3(nb), dim3(NW * 64), 0, s, p);
    else
        hipLaunchKernelGGL((k_flux<RY, NW, MC, false, KIND, WET, SCL>), dim3(nb), dim3(NW * 64), 0, s, p);
}

// edge and interior tiles in one launch: the edge part's logical blocks
// first on every XCD, then the interior part's
template <int RY, int NW, int MC, bool WET, bool SCL = false>
void launch_all_t(hipStream_t s, FluxParams p, int nedge, int nint, bool odd)
{
    if (nedge + nint <= 0) return;
    const int ngrp = (p.g.nc * p.g.nm + MC - 1) / MC;
    p.nlogical = nedge * ngrp;
    p.per_xcd = (p.nlogical + kXCD - 1) / kXCD;
    p.nlogical_i = nint * ngrp;
    p.per_xcd_i = (p.nlogical_i + kXCD - 1) / kXCD;
    const unsigned nb = (unsigned)((p.per_xcd + p.per_xcd_i) * kXCD);
    if (odd)
        hipLaunchKernelGGL((k_flux<RY, NW, MC, true, K_ALL, WET, SCL>), dim3(nb), dim3(NW * 64), 0, s, p);
    else
        hipLaunchKernelGGL((k_flux<RY, NW, MC, false, K_ALL, WET, SCL>), dim3(nb), dim3(NW * 64), 0, s, p);
}

// the wet-deposition instantiation only when the step has precipitation;
// dry ensembles of an even member count in the default 58x18, two-slab
// configuration take the shared-class instantiation (same arithmetic, the
// class's loads and factors shared by the pair)
template <int RY, int NW, int MC, int KIND>
void launch_part(hipStream_t s, const FluxParams &p, int ntiles, bool odd, bool wet)
{
    constexpr bool kShared = RY == 24 && NW == 12 && MC == 2;
    if (wet)
        launch_part_t<RY, NW, MC, KIND, true>(s, p, ntiles, odd);
    else if (kShared && p.g.nm > 1 && p.g.nm % 2 == 0)
        launch_part_t<RY, NW, MC, KIND, false, kShared>(s, p, ntiles, odd);
    else
        launch_part_t<RY, NW, MC, KIND, false>(s, p, ntiles, odd);
}

template <int RY, int NW, int MC>
void launch_all(hipStream_t s, const FluxParams &p, int nedge, int nint, bool odd, bool wet)
{
    constexpr bool kShared = RY == 24 && NW == 12 && MC == 2;
    if (wet)
        launch_all_t<RY, NW, MC, true>(s, p, nedge, nint, odd);
    else if (kShared && p.g.nm > 1 && p.g.nm % 2 == 0)
        launch_all_t<RY, NW, MC, false, kShared>(s, p, nedge, nint, odd);
    else
        launch_all_t<RY, NW, MC, false>(s, p, nedge, nint, odd);
}

template <int RY, int NW, int MC>
int launch_flux(hipStream_t s, const StepArgs &a, const FluxParams &p0, bool odd)
{
    FluxParams p = p0;
    const int nx = p.g.nx, ny = p.g.ny, ty = RY - 6;
    p.ntx = flux_tiles(nx, RX);
    p.nty = flux_tiles(ny, RY);
    // tile bx is interior when 1 <= gi0 and gi0 + 63 <= nx (gi0 = bx*58),
    // likewise in y with RY rows
    p.bx0 = 1;
    p.bx1 = min(p.ntx - 1, (nx - (RX - 1)) >= 0 ? (nx - (RX - 1)) / TXR : -1);
    p.by0 = 1;
    p.by1 = min(p.nty - 1, (ny - (RY - 1)) >= 0 ? (ny - (RY - 1)) / ty : -1);
    // an axis without interior tiles: every column (row) is an edge one
    if (p.bx1 < p.bx0) {
        p.bx0 = p.ntx;
        p.bx1 = p.ntx - 1;
    }
    if (p.by1 < p.by0) {
        p.by0 = p.nty;
        p.by1 = p.nty - 1;
    }
    const int nmx = p.bx1 - p.bx0 + 1, nmy = p.by1 - p.by0 + 1;
    const int nex = p.ntx - nmx, ney = p.nty - nmy;
    const int nint = nmx * nmy;
    const int nedge = p.ntx * p.nty - nint;
    // groups of up to 768 lanes: edge and interior tiles in one launch.  Two
    // launches on two streams ran 6-14% apart depending on which hardware
    // queues the streams were given.  1024-lane groups take one slab per
    // edge group (two would exceed the register budget), so their edge
    // tiles go to the side stream, beside the interior launch.
    constexpr int MCE = (NW * 64 > 768) ? 1 : MC;
    if constexpr (MCE == MC) {
        launch_all<RY, NW, MC>(s, p, nedge, nint, odd, a.wet != 0);
        F3D_CHECK_LAUNCH();
    } else {
        const bool fork = a.side && a.fork && a.join && nint > 0 && nedge > 0;
        hipStream_t se = fork ? a.side : s;
        if (fork) {
            F3D_HIP(hipEventRecord(a.fork, s));
            F3D_HIP(hipStreamWaitEvent(se, a.fork, 0));
        }
        launch_part<RY, NW, MCE, K_EDGE>(se, p, nedge, odd, a.wet != 0);
        F3D_CHECK_LAUNCH();
        launch_part<RY, NW, MC, K_INTERIOR>(s, p, nint, odd, a.wet != 0);
        F3D_CHECK_LAUNCH();
        if (fork) {
            F3D_HIP(hipEventRecord(a.join, se));
            F3D_HIP(hipStreamWaitEvent(s, a.join, 0));
        }
    }
    return F3D_OK;
}

}  // namespace

// Configurations of the flux kernel: ty output rows (region rows RY = ty + 6,
// two per wavefront: 8, 12 or 16 waves; or, for ty = 18, three per
// wavefront: 8 waves of up to 256 registers) and mc = 1 or 2 member*class
// slabs per workgroup.
bool flux_supported(TileShape ts)
{
    if (ts.tx != TXR) return false;
    const bool three = ts.ty == 18 && ts.nw == 8;
    if (ts.nw != 0 && ts.nw != (ts.ty + 6) / 2 && !three) return false;
    return (ts.ty == 10 || ts.ty == 18 || ts.ty == 26) && (ts.mc == 0 || ts.mc == 1 || ts.mc == 2);
}

// Kernel parameters of one (member-)scenario step: grid, state buffers,
// field offsets, dt.
static int make_params(const StepArgs &a, const Fields &f, FluxParams &p)
{
    p = FluxParams{};
    p.g = a.g;
    p.cin = a.cin;
    p.cout = a.cout;
    p.cdry = a.cdry;
    p.lam = a.lam;
    p.kwet = a.kwet;
    p.wetcls = a.wetcls;
    p.cwet = a.cwet;
    p.dz64 = a.dz64;
    p.gbuf = a.gbuf;
    p.gslab = (unsigned)ghost_floats(a.g);
    const char *fb = reinterpret_cast<const char *>(f.base);
    auto off = [&](const void *q) {
        return (unsigned)(reinterpret_cast<const char *>(q) - fb);
    };
    // every byte offset the kernel forms (field buffer, class slab of c) is
    // 32-bit
    if (f.bytes >= (size_t)1 << 32 || (size_t)a.g.G * 4 >= (size_t)1 << 31) {
        set_error("flux kernel: fields (%zu B) exceed 4 GiB or a class slab 2 GiB", f.bytes);
        return F3D_EINVAL;
    }
    // the outflux indexes the face cells of all classes in 32 bits
    if (((long)a.g.ny * a.g.nz + (long)a.g.nx * a.g.nz + a.g.NXY) * a.g.nc >= 1L << 31) {
        set_error("outflux: %d classes of %dx%dx%d face cells exceed 2^31", a.g.nc, a.g.nx, a.g.ny,
                  a.g.nz);
        return F3D_EINVAL;
    }
    if (a.g.nz > kMaxLevels) {
        set_error("flux kernel: nz=%d exceeds the supported %d levels", a.g.nz, kMaxLevels);
        return F3D_EINVAL;
    }
    p.fbase = f.base;
    p.fbytes = (unsigned)f.bytes;
    p.oA = off(f.cA);
    p.oD = off(f.cD);
    p.oH = off(f.h2);
    p.oW = off(f.wzt);
    p.ftx = (unsigned)f.ftx;
    p.ftl = (unsigned)f.ftl;
    if (f.ftx != flux_tiles(a.g.nx, RX)) {
        set_error("flux kernel: field layout of %d tile columns, grid needs %d", f.ftx,
                  flux_tiles(a.g.nx, RX));
        return F3D_EINVAL;
    }
    p.oBF = off(f.bfac);
    p.oTF = off(f.tfac);
    p.vsum1 = f.vsum1;
    p.zl = f.zl;
    p.dt64 = a.dt;
    p.dt = (float)a.dt;
    p.dtdx2 = (float)(a.dt / (a.g.dx * a.g.dx));
    p.dtdy2 = (float)(a.dt / (a.g.dy * a.g.dy));
    return F3D_OK;
}

int flux_step(hipStream_t s, const StepArgs &a, const Fields &f, TileShape ts)
{
    FluxParams p;
    TRY(make_params(a, f, p));
    const bool odd = (a.iiter % 2) != 0;
    const int mc = ts.mc ? ts.mc : 1;
    if (ts.ty == 10 && mc == 1) return launch_flux<16, 8, 1>(s, a, p, odd);
    if (ts.ty == 10 && mc == 2) return launch_flux<16, 8, 2>(s, a, p, odd);
    if (ts.ty == 18 && ts.nw == 8) {
        if (mc == 1) return launch_flux<24, 8, 1>(s, a, p, odd);
        return launch_flux<24, 8, 2>(s, a, p, odd);
    }
    if (ts.ty == 18 && mc == 1) return launch_flux<24, 12, 1>(s, a, p, odd);
    if (ts.ty == 18 && mc == 2) return launch_flux<24, 12, 2>(s, a, p, odd);
    if (ts.ty == 26 && mc == 1) return launch_flux<32, 16, 1>(s, a, p, odd);
    if (ts.ty == 26 && mc == 2) return launch_flux<32, 16, 2>(s, a, p, odd);
    set_error("flux kernel: unsupported tile %dx%d (%d slabs per group)", ts.tx, ts.ty, mc);
    return F3D_EINVAL;
}

// ---------------- batched form: many scenarios, one persistent launch ----------------
namespace {

__global__ void k_zero_counters(unsigned *c) { c[threadIdx.x] = 0u; }

// a list of the batch plan: kXCD sublists, kXCD + 1 offsets, kXCD counters;
// four lists (two parity groups, interior and edge)
constexpr int kListCtrs = kXCD;
constexpr int kBatchCounters = 4 * kListCtrs;

// Persistent form of k_flux over a work list: item = (scenario, bx, by,
// slab group).  The list is dealt out in chunks to kXCD sublists, one per
// XCD (workgroup b of a launch runs on XCD b % kXCD): a
// workgroup draws items from its own XCD's sublist through that sublist's
// atomic counter, so the slab groups of one tile, and neighbouring tiles,
// share that XCD's L2; once its sublist is exhausted it draws from the
// others in turn.  Every item writes cells no other item writes, so the
// result does not depend on which workgroup takes which item.  LIST: 0
// interior items only, 1 edge items only, 2 both (the tile's path is picked
// per item).
enum : int { L_INTERIOR = 0, L_EDGE = 1, L_MIXED = 2 };
// one batch item's tile code, by tile kind
template <int RY, int NW, int MC, bool ODD, int LIST>
__device__ __forceinline__ void batch_tile(const FluxParams &p, const int *qv, int gi0, int gj0, bool ex,
                                           bool ey, float *sm, const ZLev *zt)
{
    if (LIST == L_INTERIOR || (LIST == L_MIXED && !ex && !ey)) {
        flux_tile<RY, NW, MC, ODD, false, false>(p, qv, gi0, gj0, sm, zt);
    } else {
        if (ex && ey)
            flux_tile<RY, NW, MC, ODD, true, true>(p, qv, gi0, gj0, sm, zt);
        else if (ex)
            flux_tile<RY, NW, MC, ODD, true, false>(p, qv, gi0, gj0, sm, zt);
        else
            flux_tile<RY, NW, MC, ODD, false, true>(p, qv, gi0, gj0, sm, zt);
    }
}

template <int RY, int NW, int MC, bool ODD, int LIST>
__global__ void __launch_bounds__(NW * 64)
    k_flux_batch(const FluxParams *scen, const int4 *items, const int *xoff, unsigned *counter)
{
    constexpr int PW = sizeof(FluxParams) / sizeof(int);
    static_assert(sizeof(FluxParams) % sizeof(int) == 0, "parameter block of whole dwords");
    static_assert(PW <= NW * 64, "one thread per parameter dword");
    __shared__ float sm[flux_planes(MC, ODD) * RX * RY + (LIST != L_INTERIOR ? flux_xstage(RY, MC) : 0)];
    __shared__ ZLev zt[kMaxLevels + 2];
    __shared__ int sp[PW];
    __shared__ int sitem[4];
    // thread 0's place: the sublist it draws from and how many it has left
    int x = (int)(blockIdx.x % kXCD), left = kXCD;
    for (;;) {
        if (threadIdx.x == 0) {
            int4 w = make_int4(-1, 0, 0, 0);
            while (left > 0) {
                const int n = xoff[x + 1] - xoff[x];
                const int it = (int)atomicAdd(counter + x, 1u);
                if (it < n) {
                    w = items[xoff[x] + it];
                    break;
                }
                x = (x + 1) % kXCD;
                --left;
            }
            sitem[0] = w.x;
            sitem[1] = w.y;
            sitem[2] = w.z;
            sitem[3] = w.w;
        }
        __syncthreads();
        const int sc = __builtin_amdgcn_readfirstlane(sitem[0]);
        if (sc < 0) return;  // every sublist exhausted (the same value for every wave)
        const int bx = __builtin_amdgcn_readfirstlane(sitem[1]);
        const int by = __builtin_amdgcn_readfirstlane(sitem[2]);
        const int grp = __builtin_amdgcn_readfirstlane(sitem[3]);
        if (threadIdx.x < PW) sp[threadIdx.x] = reinterpret_cast<const int *>(scen + sc)[threadIdx.x];
        __syncthreads();
        FluxParams p;
#pragma unroll
        for (int k = 0; k < PW; ++k)
            reinterpret_cast<int *>(&p)[k] = __builtin_amdgcn_readfirstlane(sp[k]);
        for (int k = threadIdx.x; k <= p.g.nz + 1; k += NW * 64) zt[k] = p.zl[k];
        __syncthreads();
        int qv[MC];
#pragma unroll
        for (int c = 0; c < MC; ++c) qv[c] = grp * MC + c;
        const int gi0 = bx * TXR, gj0 = by * (RY - 6);
        const bool ex = gi0 < 1 || gi0 + RX - 1 > p.g.nx;
        const bool ey = gj0 < 1 || gj0 + RY - 1 > p.g.ny;
        // the last group of an odd slab count holds one slab: it runs the
        // one-slab tile code (the same arithmetic per slab, about 0.63 of a
        // two-slab group's time, where a dead second slab saves only 5%)
        if (MC == 2 && grp * MC + 1 >= p.g.nc * p.g.nm)
            batch_tile<RY, NW, 1, ODD, LIST>(p, qv, gi0, gj0, ex, ey, sm, zt);
        else
            batch_tile<RY, NW, MC, ODD, LIST>(p, qv, gi0, gj0, ex, ey, sm, zt);
        __syncthreads();  // LDS is reused by the next item
    }
}

template <int RY, int NW, int MC, int LIST>
int batch_grid(int &grid)
{
    int occ = 0;
    F3D_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_flux_batch<RY, NW, MC, false, LIST>,
                                                         NW * 64, 0));
    int dev = 0, ncu = 0;
    F3D_HIP(hipGetDevice(&dev));
    F3D_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    grid = std::max(1, occ) * ncu;
    return F3D_OK;
}

// parity group gr steps with parity odd: one mixed list on s, or its edge
// list on se and its interior list on s.  List l = 2*gr (interior or mixed)
// or 2*gr + 1 (edge) has its kXCD + 1 sublist offsets at xoff[l*(kXCD+1)]
// and its counters at counter[l*kXCD].
template <int RY, int NW, int MC>
int batch_launch(hipStream_t s, hipStream_t se, const FluxBatchPlan &b, int gr, bool odd)
{
    constexpr int MCE = (NW * 64 > 768) ? 1 : MC;
    const FluxParams *par = b.par[odd ? 1 : 0].as<FluxParams>();
    const int4 *items = b.items.as<int4>();
    const int *xo = b.xoff.as<int>();
    unsigned *ctr = b.counter.as<unsigned>();
    const int li = 2 * gr, le = 2 * gr + 1;
    const int ne = b.nedge[gr], ni = b.nint[gr];
    if (b.mixed) {
        if (ni > 0) {
            const unsigned g = (unsigned)std::min(b.grid_int, ni);
            const int *x = xo + li * (kXCD + 1);
            unsigned *c = ctr + li * kListCtrs;
            if (odd)
                hipLaunchKernelGGL((k_flux_batch<RY, NW, MC, true, L_MIXED>), dim3(g), dim3(NW * 64), 0, s,
                                   par, items, x, c);
            else
                hipLaunchKernelGGL((k_flux_batch<RY, NW, MC, false, L_MIXED>), dim3(g), dim3(NW * 64), 0, s,
                                   par, items, x, c);
            F3D_CHECK_LAUNCH();
        }
        return F3D_OK;
    }
    if (ne > 0) {
        const unsigned g = (unsigned)std::min(b.grid_edge, ne);
        const int *x = xo + le * (kXCD + 1);
        unsigned *c = ctr + le * kListCtrs;
        if (odd)
            hipLaunchKernelGGL((k_flux_batch<RY, NW, MCE, true, L_EDGE>), dim3(g), dim3(NW * 64), 0, se,
                               par, items, x, c);
        else
            hipLaunchKernelGGL((k_flux_batch<RY, NW, MCE, false, L_EDGE>), dim3(g), dim3(NW * 64), 0, se,
                               par, items, x, c);
        F3D_CHECK_LAUNCH();
    }
    if (ni > 0) {
        const unsigned g = (unsigned)std::min(b.grid_int, ni);
        const int *x = xo + li * (kXCD + 1);
        unsigned *c = ctr + li * kListCtrs;
        if (odd)
            hipLaunchKernelGGL((k_flux_batch<RY, NW, MC, true, L_INTERIOR>), dim3(g), dim3(NW * 64), 0, s,
                               par, items, x, c);
        else
            hipLaunchKernelGGL((k_flux_batch<RY, NW, MC, false, L_INTERIOR>), dim3(g), dim3(NW * 64), 0, s,
                               par, items, x, c);
        F3D_CHECK_LAUNCH();
    }
    return F3D_OK;
}

// item lists of one scenario for region rows RY: interior tiles in groups of
// mc slabs, edge tiles in groups of mce slabs; the slab groups of a tile,
// and the tiles down a tile column, are consecutive
void scenario_items(int sc, const Grid &g, int RY, int mc, int mce, std::vector<int4> &iint,
                    std::vector<int4> &iedge)
{
    const int ty = RY - 6;
    const int ntx = flux_tiles(g.nx, RX), nty = flux_tiles(g.ny, RY);
    const int nq = g.nc * g.nm;
    for (int bx = 0; bx < ntx; ++bx)
        for (int by = 0; by < nty; ++by) {
            const int gi0 = bx * TXR, gj0 = by * ty;
            const bool edge = gi0 < 1 || gi0 + RX - 1 > g.nx || gj0 < 1 || gj0 + RY - 1 > g.ny;
            const int m = edge ? mce : mc;
            for (int gr = 0; gr < (nq + m - 1) / m; ++gr)
                (edge ? iedge : iint).push_back(make_int4(sc, bx, by, gr));
        }
}

// Deal list [o, o + n) of all out to kXCD sublists, in place: the list is
// cut into chunks of about kChunk items (a chunk keeps the slab groups of a
// tile, and neighbouring tiles, together) and chunk t goes to sublist
// t % kXCD, so every sublist holds the same mix of deep and shallow items
// (the list is sorted deepest first) and ends on shallow ones.  x receives
// the kXCD + 1 absolute offsets of the sublists.
static void xcd_deal(std::vector<int4> &all, int o, int n, int *x)
{
    constexpr int kChunk = 32;
    std::vector<int4> part[kXCD];
    for (int k = 0; k < n; ++k) part[(k / kChunk) % kXCD].push_back(all[o + k]);
    int at = o;
    for (int d = 0; d < kXCD; ++d) {
        x[d] = at;
        std::copy(part[d].begin(), part[d].end(), all.begin() + at);
        at += (int)part[d].size();
    }
    x[kXCD] = at;
}

template <int RY, int NW, int MC>
int batch_plan_t(FluxBatchPlan &b, const std::vector<StepArgs> &ae, const std::vector<StepArgs> &ao,
                 const std::vector<Fields> &f, const std::vector<int> &group)
{
    constexpr int MCE = (NW * 64 > 768) ? 1 : MC;
    const size_t n = ae.size();
    std::vector<FluxParams> pe(n), po(n);
    std::vector<int4> iint[2], iedge[2];
    for (size_t k = 0; k < n; ++k) {
        TRY(make_params(ae[k], f[k], pe[k]));
        TRY(make_params(ao[k], f[k], po[k]));
        const int gr = group[k] ? 1 : 0;
        scenario_items((int)k, ae[k].g, RY, MC, MCE, iint[gr], iedge[gr]);
    }
    // deeper columns first (cost per item grows with nz); lists laid out as
    // [group 0 interior | group 0 edge | group 1 interior | group 1 edge],
    // each cut into per-XCD sublists.  When interior and edge tiles take the
    // same slab count, each group is one mixed list drawn by one persistent
    // launch: two persistent launches side by side would split the device by
    // arrival, not by work.
    auto deeper = [&](const int4 &x, const int4 &y) { return ae[x.x].g.nz > ae[y.x].g.nz; };
    b.mixed = MCE == MC;
    std::vector<int4> all;
    int xo[4 * (kXCD + 1)];
    for (int gr = 0; gr < 2; ++gr) {
        if (b.mixed) {
            iint[gr].insert(iint[gr].end(), iedge[gr].begin(), iedge[gr].end());
            iedge[gr].clear();
        }
        std::stable_sort(iint[gr].begin(), iint[gr].end(), deeper);
        std::stable_sort(iedge[gr].begin(), iedge[gr].end(), deeper);
        b.nint[gr] = (int)iint[gr].size();
        b.nedge[gr] = (int)iedge[gr].size();
        b.off_int[gr] = (int)all.size();
        all.insert(all.end(), iint[gr].begin(), iint[gr].end());
        b.off_edge[gr] = (int)all.size();
        all.insert(all.end(), iedge[gr].begin(), iedge[gr].end());
        xcd_deal(all, b.off_int[gr], b.nint[gr], xo + (2 * gr) * (kXCD + 1));
        xcd_deal(all, b.off_edge[gr], b.nedge[gr], xo + (2 * gr + 1) * (kXCD + 1));
    }
    if (b.par[0].alloc(sizeof(FluxParams) * n, nullptr) || b.par[1].alloc(sizeof(FluxParams) * n, nullptr) ||
        b.items.alloc(sizeof(int4) * std::max<size_t>(1, all.size()), nullptr) ||
        b.xoff.alloc(sizeof(xo), nullptr) ||
        b.counter.alloc(sizeof(unsigned) * kBatchCounters, nullptr))
        return F3D_ENOMEM;
    F3D_HIP(hipMemcpy(b.par[0].p, pe.data(), sizeof(FluxParams) * n, hipMemcpyHostToDevice));
    F3D_HIP(hipMemcpy(b.par[1].p, po.data(), sizeof(FluxParams) * n, hipMemcpyHostToDevice));
    if (!all.empty())
        F3D_HIP(hipMemcpy(b.items.p, all.data(), sizeof(int4) * all.size(), hipMemcpyHostToDevice));
    F3D_HIP(hipMemcpy(b.xoff.p, xo, sizeof(xo), hipMemcpyHostToDevice));
    if (b.mixed) {
        TRY((batch_grid<RY, NW, MC, L_MIXED>(b.grid_int)));
        b.grid_edge = 0;
    } else {
        TRY((batch_grid<RY, NW, MC, L_INTERIOR>(b.grid_int)));
        TRY((batch_grid<RY, NW, MCE, L_EDGE>(b.grid_edge)));
    }
    return F3D_OK;
}

}  // namespace

int flux_batch_plan(FluxBatchPlan &b, const std::vector<StepArgs> &even,
                    const std::vector<StepArgs> &odd, const std::vector<Fields> &f,
                    const std::vector<int> &group, TileShape ts)
{
    if (even.empty() || even.size() != odd.size() || even.size() != f.size() ||
        group.size() != f.size()) {
        set_error("batch plan: inconsistent scenario lists");
        return F3D_EINVAL;
    }
    if (!flux_supported(ts)) {
        set_error("batch plan: the batched step runs the 58-column kernel (got %dx%d)", ts.tx, ts.ty);
        return F3D_EINVAL;
    }
    b.ts = ts;
    const int mc = ts.mc ? ts.mc : 1;
    if (ts.ty == 10) return mc == 1 ? batch_plan_t<16, 8, 1>(b, even, odd, f, group) : batch_plan_t<16, 8, 2>(b, even, odd, f, group);
    if (ts.ty == 18) return mc == 1 ? batch_plan_t<24, 12, 1>(b, even, odd, f, group) : batch_plan_t<24, 12, 2>(b, even, odd, f, group);
    return mc == 1 ? batch_plan_t<32, 16, 1>(b, even, odd, f, group) : batch_plan_t<32, 16, 2>(b, even, odd, f, group);
}

int flux_batch_step(hipStream_t s, hipStream_t side, hipEvent_t fork, hipEvent_t join,
                    const FluxBatchPlan &b, bool odd0, bool join_side)
{
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(kBatchCounters), 0, s, b.counter.as<unsigned>());
    F3D_CHECK_LAUNCH();
    const int ne = b.nedge[0] + b.nedge[1], ni = b.nint[0] + b.nint[1];
    const bool split = side && fork && join && ni > 0 && ne > 0;
    hipStream_t se = split ? side : s;
    if (split) {
        F3D_HIP(hipEventRecord(fork, s));
        F3D_HIP(hipStreamWaitEvent(se, fork, 0));
    }
    const TileShape ts = b.ts;
    const int mc = ts.mc ? ts.mc : 1;
    for (int gr = 0; gr < 2; ++gr) {
        const bool odd = gr == 0 ? odd0 : !odd0;
        int rc;
        if (ts.ty == 10)
            rc = mc == 1 ? batch_launch<16, 8, 1>(s, se, b, gr, odd) : batch_launch<16, 8, 2>(s, se, b, gr, odd);
        else if (ts.ty == 18)
            rc = mc == 1 ? batch_launch<24, 12, 1>(s, se, b, gr, odd) : batch_launch<24, 12, 2>(s, se, b, gr, odd);
        else
            rc = mc == 1 ? batch_launch<32, 16, 1>(s, se, b, gr, odd) : batch_launch<32, 16, 2>(s, se, b, gr, odd);
        if (rc) return rc;
    }
    if (split && join_side) {
        F3D_HIP(hipEventRecord(join, se));
        F3D_HIP(hipStreamWaitEvent(s, join, 0));
    }
    return F3D_OK;
}

}  // namespace f32
}  // namespace f3d
