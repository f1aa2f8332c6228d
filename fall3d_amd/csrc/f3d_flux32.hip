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
#include "f3d_ctx.h"

namespace f3d {
namespace f32 {
namespace {

constexpr int kXCD = 8;
constexpr int RX = 64;       // region row = one wavefront
constexpr int TXR = RX - 6;  // output columns per tile

struct FluxParams {
    Grid g;
    const float *cin;
    float *cout;
    double *cdry;
    const float4 *cA;  // (ax, fx, ay, fy)
    const float4 *cD;  // (dfx, dfy, 1/rho, rho*rkver)
    const float2 *cE;  // (rho*rkh1, rho*rkhor), raw products for line ends
    const float *wz;   // vz+vset per class
    const float *bfac, *tfac;
    const double *vsum1;
    const ZLev *zl;
    double dt64;
    float dt, dtdx2, dtdy2;
    int ntx, nty;
    // interior tiles form the rectangle [bx0, bx1] x [by0, by1]; the others
    // (regions meeting the domain frame) go to the edge kernel
    int bx0, bx1, by0, by1;
    int nlogical;
    int per_xcd;
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
// (diffx.f90:20-29, 45-53).
template <bool EX>
__device__ __forceinline__ float diff_x(float d, float lr, float D, float pk, float dtdx2,
                                        bool inner, bool lo, bool hi)
{
    const float q = d * lr;
    const float ql = from_left(q);
    const float G = D * (q - ql);
    const float Gr = from_right(G);
    if (!EX) return d + (Gr - G);
    const float qr = from_right(q);
    const float pkl = from_left(pk), pkr = from_right(pk);
    const float gl = lo ? fmaxf(0.f, dtdx2 * pk * (q - ql)) : (hi ? dtdx2 * pkl * (q - ql) : G);
    const float gr = lo ? dtdx2 * pkr * (qr - q) : (hi ? fminf(0.f, dtdx2 * pk * (qr - q)) : Gr);
    return inner ? d + (gr - gl) : d;
}

template <int RY, int NW, bool ODD, bool EX, bool EY>
__device__ __forceinline__ void flux_tile(const FluxParams &p, const int q, const int ic,
                                          const int mem, const int gi0, const int gj0, float *sm)
{
    constexpr int NS = RY / NW;
    constexpr bool EDGE = EX || EY;
    static_assert(RY % NW == 0, "region rows must divide evenly over the waves");
    const int nx = p.g.nx, ny = p.g.ny, nz = p.g.nz;
    const unsigned NXY = (unsigned)p.g.NXY;
    const unsigned SZ = (unsigned)((nx + 2) * (ny + 2));
    const int li = threadIdx.x & (RX - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gi = gi0 + li;
    float *SP = sm, *SF = sm + RX * RY, *SQ = sm + 2 * RX * RY;

    const float *cin = p.cin + (size_t)q * p.g.G;
    float *cout = p.cout + (size_t)q * p.g.G;
    const float *wz = p.wz + (size_t)ic * p.g.N;
    const float *cDf = reinterpret_cast<const float *>(p.cD);
    const float dt = p.dt;

    // per-row geometry and flags
    unsigned coff[NS], moff[NS], mup[NS], mdn[NS];
    bool inner[NS], xlo[NS], xhi[NS], ylo[NS], yhi[NS], gst[NS], gzx[NS];
    int lj[NS];
#pragma unroll
    for (int sl = 0; sl < NS; ++sl) {
        lj[sl] = w * NS + sl;
        const int gj = gj0 + lj[sl];
        const int ci = EX ? min(max(gi, 1), nx) : gi;
        const int cj = EY ? min(max(gj, 1), ny) : gj;
        coff[sl] = (unsigned)cidx(nx, ny, ci, cj, 0);
        moff[sl] = (unsigned)((ci - 1) + nx * (cj - 1));
        // row above / below (clamped), for the face j+1/2 and line-end products
        mup[sl] = (unsigned)((ci - 1) + nx * (min(cj + 1, ny) - 1));
        mdn[sl] = (unsigned)((ci - 1) + nx * (max(cj - 1, 1) - 1));
        const bool inx = !EX || (gi >= 1 && gi <= nx);
        const bool iny = !EY || (gj >= 1 && gj <= ny);
        inner[sl] = inx && iny;
        xlo[sl] = EX && gi == 1;
        xhi[sl] = EX && gi == nx;
        ylo[sl] = EY && gj == 1;
        yhi[sl] = EY && gj == ny;
        // setbcc face ghosts (setbcc.f90:139-169): x faces inside the y range,
        // y faces inside the x range; gzx selects which wind decides
        const bool xg = EX && iny && (gi == 0 || gi == nx + 1);
        const bool yg = EY && inx && (gj == 0 || gj == ny + 1);
        gst[sl] = xg || yg;
        gzx[sl] = xg;
    }

    // advctzc column window (per row): c(s), c(s+1), w(s), U(s), W2(s), delta(s);
    // after level nz, c0 holds the top ghost c(nz+1)
    float c0[NS], cp1[NS], vk[NS], Uc[NS], W2k[NS], dprev[NS];
    float nc2[NS], nwz[NS];
    float4 Ac[NS], An[NS];  // (ax, fx, ay, fy) at level s, prefetched level s+1
    // diffz window: b(k), q(k) = b/rho and the face flux gb(k+1) it leaves for
    // the level above (before level 1: b(0))
    float b0[NS], q0[NS], gprev[NS];
    float4 D0[NS], Dn[NS];  // (dfx, dfy, lr, pz) at levels k and s
    // carried between phase 3 and the next phase 2: level k state and q,
    // face products of diffy at j-1/2 and j+1/2, raw products (EY)
    float dc[NS], qc[NS], dyl[NS], dyu[NS], pyo[NS], pym[NS], pyp[NS], lrc[NS], dxc[NS], pxc[NS];
    float P[NS], phi[NS], apl[NS], apr[NS];

    {
        const ZLev z1 = p.zl[1];
        const float ldz0 = p.zl[0].ldz;
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
            const unsigned o = coff[sl], m = moff[sl];
            const float c1 = cin[o + SZ];
            const float c2 = cin[o + 2 * SZ];
            const float bf = p.bfac[(size_t)ic * NXY + m];
            const float gb = c1 * bf;  // c(0), setbcc.f90:108-113
            const float v1 = wz[m];
            c0[sl] = c1;
            cp1[sl] = c2;
            vk[sl] = v1;
            const float d1 = (c2 - c1) * z1.dzk;
            const float d0 = (c1 - gb) * ldz0;
            const float av = fabsf(v1);
            W2k[sl] = minmod(d1, d0) * av * (z1.a2 - z1.a2bc * av);
            Uc[sl] = z1.dtdzz * (v1 * c1 - v1 * gb);  // U(1) = W1(1)
            dprev[sl] = d1;
            gprev[sl] = gb;
            nc2[sl] = (3 <= nz) ? cin[o + 3 * SZ] : 0.f;
            nwz[sl] = wz[NXY + m];
            An[sl] = p.cA[m];
            b0[sl] = q0[sl] = 0.f;
            D0[sl] = Dn[sl] = make_float4(0.f, 0.f, 0.f, 0.f);
            dc[sl] = qc[sl] = dyl[sl] = dyu[sl] = pyo[sl] = pym[sl] = pyp[sl] = 0.f;
            lrc[sl] = dxc[sl] = pxc[sl] = 0.f;
            P[sl] = phi[sl] = apl[sl] = apr[sl] = 0.f;
        }
    }
    const float dtop = dt * p.zl[nz].ldz;

    for (int s = 1; s <= nz + 2; ++s) {
        // ======== phase 1: advctzc at level s (and advctx on even steps) ========
        if (s <= nz) {
            const ZLev zn = p.zl[s + 1];
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                Ac[sl] = An[sl];
                Dn[sl] = p.cD[(unsigned)(s - 1) * NXY + moff[sl]];  // used after two barriers
                float cp2 = 0.f, vkn = 0.f, W1n, W2n, dn = 0.f;
                if (s < nz) {
                    // top ghost c(nz+1) (setbcc.f90:126-130)
                    cp2 = (s + 2 <= nz) ? nc2[sl] : p.tfac[(size_t)ic * NXY + moff[sl]] * cp1[sl];
                    vkn = nwz[sl];
                    W1n = zn.dtdzz * (vkn * cp1[sl] - vk[sl] * c0[sl]);
                    dn = (cp2 - cp1[sl]) * zn.dzk;
                    const float av = 0.5f * fabsf(vkn + vk[sl]);
                    W2n = minmod(dn, dprev[sl]) * av * (zn.a2 - zn.a2bc * av);
                } else {
                    W1n = dtop * vk[sl] * (cp1[sl] - c0[sl]);  // advctzc.f90:175-181
                    W2n = W2k[sl];
                }
                const float Un = W1n + (W2n - W2k[sl]);
                float a = c0[sl] - ((vk[sl] > 0.f) ? Uc[sl] : Un);
                if (EDGE) {
                    // ghost faces: zero when the wind at the face cell blows inward
                    const bool gi0s = (gi == 0), gj0s = (gj0 + lj[sl] == 0);
                    const float v = gzx[sl] ? Ac[sl].x : Ac[sl].z;
                    const bool zero = gzx[sl] ? (gi0s ? v >= 0.f : v <= 0.f) : (gj0s ? v >= 0.f : v <= 0.f);
                    a = inner[sl] ? a : (gst[sl] && !zero ? c0[sl] : 0.f);
                }
                c0[sl] = cp1[sl];
                cp1[sl] = cp2;
                vk[sl] = vkn;
                Uc[sl] = Un;
                W2k[sl] = W2n;
                dprev[sl] = dn;
                // prefetch level s+1 coefficients, c(s+3), w(s+2)
                if (s + 3 <= nz) nc2[sl] = cin[coff[sl] + (unsigned)(s + 3) * SZ];
                if (s + 2 <= nz) nwz[sl] = wz[(unsigned)(s + 1) * NXY + moff[sl]];
                if (s + 1 <= nz) An[sl] = p.cA[(unsigned)s * NXY + moff[sl]];
                if (!ODD) a = adv_x<EX>(a, Ac[sl].x, Ac[sl].y, inner[sl], xlo[sl], xhi[sl]);
                P[sl] = a;
            }
#pragma unroll
            for (int sl = 0; sl < NS; ++sl)
                if (sl == 0 || sl == NS - 1) SP[lj[sl] * RX + li] = P[sl];
        }
        __syncthreads();

        // ======== phase 2: advcty fluxes at level s; last sweep of level s-2 ========
        if (s <= nz) {
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                if (lj[sl] < 1 || lj[sl] >= RY - 1) continue;  // wave-uniform
                const float Pm = (sl > 0) ? P[sl - 1] : SP[(lj[sl] - 1) * RX + li];
                const float Pp = (sl < NS - 1) ? P[sl + 1] : SP[(lj[sl] + 1) * RX + li];
                const float w2 = Ac[sl].w * minmod(Pp - P[sl], P[sl] - Pm);
                phi[sl] = fmaf(Ac[sl].z, P[sl], w2);
                if (EY) {
                    apl[sl] = fmaf(Ac[sl].z, Pm, w2);
                    apr[sl] = fmaf(Ac[sl].z, Pp, w2);
                }
                if (sl == 0 || sl == NS - 1) SF[lj[sl] * RX + li] = phi[sl];
            }
        }
        if (s >= 3) {
            const int k = s - 2;
            float *co = cout + (unsigned)k * SZ;
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                if (lj[sl] < 3 || lj[sl] >= RY - 3) continue;  // wave-uniform
                // diffy (diffy.f90:30-44; line ends 20-29, 45-53)
                const float qm = (sl > 0) ? qc[sl - 1] : SQ[(lj[sl] - 1) * RX + li];
                const float qp = (sl < NS - 1) ? qc[sl + 1] : SQ[(lj[sl] + 1) * RX + li];
                const float Du = (sl < NS - 1) ? dyl[sl + 1] : dyu[sl];
                float glo = dyl[sl] * (qc[sl] - qm);
                float ghi = Du * (qp - qc[sl]);
                if (EY) {
                    const float d2 = p.dtdy2;
                    glo = ylo[sl] ? fmaxf(0.f, d2 * pyo[sl] * (qc[sl] - qm))
                                  : (yhi[sl] ? d2 * pym[sl] * (qc[sl] - qm) : glo);
                    ghi = ylo[sl] ? d2 * pyp[sl] * (qp - qc[sl])
                                  : (yhi[sl] ? fminf(0.f, d2 * pyo[sl] * (qp - qc[sl])) : ghi);
                }
                float out = dc[sl] + (ghi - glo);
                if (EDGE) out = inner[sl] ? out : dc[sl];
                if (ODD)
                    out = diff_x<EX>(out, lrc[sl], dxc[sl], pxc[sl], p.dtdx2, inner[sl], xlo[sl],
                                     xhi[sl]);
                const bool store = li >= 3 && li < RX - 3 && (!EDGE || inner[sl]);
                if (store) {
                    co[coff[sl]] = fmaxf(0.f, out);  // sizloop.f90:77-78
                    if (k == 1) {
                        // accum.f90:22: deposit from c(:,:,1) before the clip, FP64
                        const size_t t2 = moff[sl];
                        double *cd = p.cdry + ((size_t)mem * (p.g.nc + 1) + ic + 1) * NXY + t2;
                        *cd = *cd - p.dt64 * (double)out * p.vsum1[(size_t)ic * NXY + t2];
                    }
                }
            }
        }
        __syncthreads();

        // ======== phase 3: advcty (+ advctx on odd steps) at level s; diffz and
        // the first horizontal diffusion at level s-1 ========
        if (s >= nz + 2) continue;
        const int k = s - 1;
        const ZLev zk = p.zl[k > 0 ? k : 0];
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
            if (lj[sl] < 2 || lj[sl] >= RY - 2) continue;  // wave-uniform
            float bnew;
            if (s <= nz) {
                float pl = (sl > 0) ? phi[sl - 1] : SF[(lj[sl] - 1) * RX + li];
                float pr = (sl < NS - 1) ? phi[sl + 1] : SF[(lj[sl] + 1) * RX + li];
                if (EY) {
                    pl = ylo[sl] ? apl[sl] : pl;
                    pr = yhi[sl] ? apr[sl] : pr;
                }
                bnew = (Ac[sl].z > 0.f) ? P[sl] + (pl - phi[sl]) : P[sl] + (phi[sl] - pr);
                if (EDGE) bnew = inner[sl] ? bnew : P[sl];
                if (ODD) bnew = adv_x<EX>(bnew, Ac[sl].x, Ac[sl].y, inner[sl], xlo[sl], xhi[sl]);
            } else {
                bnew = c0[sl];  // b(nz+1): the top ghost of c_in
            }
            // diffz (diffz.f90:19-62) at level k = s-1, input b(k+1) = bnew
            const float lrn = Dn[sl].z, pzn = Dn[sl].w;  // level s (unused at s = nz+1)
            if (s == 1) {
                // bottom face: c(0)/rho(1), un-averaged rho*K(1), inflow only
                const float q1 = bnew * lrn;
                gprev[sl] = fmaxf(0.f, pzn * (q1 - gprev[sl] * lrn) * p.zl[0].ldz);
                b0[sl] = bnew;
                q0[sl] = q1;
                D0[sl] = Dn[sl];
                continue;
            }
            const float lr0 = D0[sl].z, pz0 = D0[sl].w;
            float d, qn = 0.f;
            if (k < nz) {
                qn = bnew * lrn;
                const float dq = qn - q0[sl];
                const float gav = (pzn + pz0) * zk.hldz * dq;
                // level 1 uses the un-averaged rho*K(2) on its top face
                const float gt = (k == 1) ? pzn * dq * zk.ldz : gav;
                d = b0[sl] + zk.dtdzz * (gt - gprev[sl]);
                // level nz takes the un-averaged rho*K(nz-1) on its bottom face
                gprev[sl] = (k == nz - 1) ? pz0 * dq * zk.ldz : gav;
            } else {
                const float gt = fminf(0.f, pz0 * (bnew * lr0 - q0[sl]) * zk.ldz);
                d = b0[sl] + dtop * (gt - gprev[sl]);
            }
            if (EDGE) d = inner[sl] ? d : b0[sl];
            const float4 Dk = D0[sl];
            // shift the diffz window
            q0[sl] = qn;
            b0[sl] = bnew;
            D0[sl] = Dn[sl];
            // first horizontal diffusion at level k
            if (!ODD) {
                const float pk = EX ? p.cE[(unsigned)(k - 1) * NXY + moff[sl]].x : 0.f;
                d = diff_x<EX>(d, Dk.z, Dk.x, pk, p.dtdx2, inner[sl], xlo[sl], xhi[sl]);
            } else {
                lrc[sl] = Dk.z;
                dxc[sl] = Dk.x;
                if (EX) pxc[sl] = p.cE[(unsigned)(k - 1) * NXY + moff[sl]].x;
            }
            dc[sl] = d;
            qc[sl] = d * Dk.z;
            dyl[sl] = Dk.y;
            if (sl == NS - 1) dyu[sl] = cDf[4 * ((unsigned)(k - 1) * NXY + mup[sl]) + 1];
            if (EY) {
                pyo[sl] = p.cE[(unsigned)(k - 1) * NXY + moff[sl]].y;
                pym[sl] = p.cE[(unsigned)(k - 1) * NXY + mdn[sl]].y;
                pyp[sl] = p.cE[(unsigned)(k - 1) * NXY + mup[sl]].y;
            }
        }
#pragma unroll
        for (int sl = 0; sl < NS; ++sl)
            if ((sl == 0 || sl == NS - 1) && lj[sl] >= 2 && lj[sl] < RY - 2 && s >= 2)
                SQ[lj[sl] * RX + li] = qc[sl];
    }
}

// Interior tiles (no frame in the region) and edge tiles run as two kernels
// so the interior one keeps the register budget of the plain path.
template <int RY, int NW, bool ODD, bool EDGEK>
__global__ void __launch_bounds__(NW * 64) k_flux(FluxParams p)
{
    const int b = blockIdx.x;
    const int L = (b % kXCD) * p.per_xcd + b / kXCD;  // XCD-aware logical block
    if (L >= p.nlogical) return;
    const int nc = p.g.nc;
    const int nq = nc * p.g.nm;
    const int q = L % nq;
    const int ic = q % nc, mem = q / nc;
    const int t = L / nq;
    int bx, by;
    if (!EDGEK) {
        // consecutive tiles run down y so the 6-row halos are shared in L2
        const int nyi = p.by1 - p.by0 + 1;
        by = p.by0 + t % nyi;
        bx = p.bx0 + t / nyi;
    } else {
        // edge tiles: full rows below by0, the side columns of rows
        // by0..by1, full rows above by1
        const int nside = p.bx0 + (p.ntx - 1 - p.bx1);
        const int nmid = (p.by1 >= p.by0) ? (p.by1 - p.by0 + 1) : 0;
        int e = t;
        if (e < p.by0 * p.ntx) {
            by = e / p.ntx;
            bx = e % p.ntx;
        } else if ((e -= p.by0 * p.ntx) < nmid * nside) {
            by = p.by0 + e / nside;
            const int c = e % nside;
            bx = (c < p.bx0) ? c : p.bx1 + 1 + (c - p.bx0);
        } else {
            e -= nmid * nside;
            by = p.by0 + nmid + e / p.ntx;
            bx = e % p.ntx;
        }
    }
    const int gi0 = 1 + bx * TXR - 3, gj0 = 1 + by * (RY - 6) - 3;
    __shared__ float sm[3 * RX * RY];
    if (!EDGEK) {
        flux_tile<RY, NW, ODD, false, false>(p, q, ic, mem, gi0, gj0, sm);
    } else {
        const bool ex = gi0 < 1 || gi0 + RX - 1 > p.g.nx;
        const bool ey = gj0 < 1 || gj0 + RY - 1 > p.g.ny;
        if (ex && ey)
            flux_tile<RY, NW, ODD, true, true>(p, q, ic, mem, gi0, gj0, sm);
        else if (ex)
            flux_tile<RY, NW, ODD, true, false>(p, q, ic, mem, gi0, gj0, sm);
        else
            flux_tile<RY, NW, ODD, false, true>(p, q, ic, mem, gi0, gj0, sm);
    }
}

template <int RY, int NW, bool EDGEK>
void launch_one(hipStream_t s, FluxParams p, int ntiles, bool odd)
{
    if (ntiles <= 0) return;
    p.nlogical = ntiles * p.g.nc * p.g.nm;
    p.per_xcd = (p.nlogical + kXCD - 1) / kXCD;
    const unsigned nb = (unsigned)(p.per_xcd * kXCD);
    if (odd)
        hipLaunchKernelGGL((k_flux<RY, NW, true, EDGEK>), dim3(nb), dim3(NW * 64), 0, s, p);
    else
        hipLaunchKernelGGL((k_flux<RY, NW, false, EDGEK>), dim3(nb), dim3(NW * 64), 0, s, p);
}

template <int RY, int NW>
int launch_flux(hipStream_t s, const FluxParams &p0, bool odd)
{
    FluxParams p = p0;
    const int nx = p.g.nx, ny = p.g.ny, ty = RY - 6;
    p.ntx = (nx + TXR - 1) / TXR;
    p.nty = (ny + ty - 1) / ty;
    // tile bx is interior when 1 <= gi0 and gi0 + 63 <= nx (gi0 = bx*58 - 2),
    // likewise in y with RY rows
    p.bx0 = 1;
    p.bx1 = min(p.ntx - 1, (nx - (RX - 3)) >= 0 ? (nx - (RX - 3)) / TXR : -1);
    p.by0 = 1;
    p.by1 = min(p.nty - 1, (ny - (RY - 3)) >= 0 ? (ny - (RY - 3)) / ty : -1);
    if (p.bx1 < p.bx0 || p.by1 < p.by0) {
        // no interior tile: everything is edge (rows 0 .. nty-1 below "by0")
        p.bx0 = 0;
        p.bx1 = p.ntx - 1;
        p.by0 = p.nty;
        p.by1 = p.nty - 1;
    }
    const int nint = (p.bx1 - p.bx0 + 1) * (p.by1 >= p.by0 ? p.by1 - p.by0 + 1 : 0);
    const int nedge = p.ntx * p.nty - nint;
    launch_one<RY, NW, false>(s, p, nint, odd);
    F3D_CHECK_LAUNCH();
    launch_one<RY, NW, true>(s, p, nedge, odd);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

}  // namespace

bool flux_supported(TileShape ts)
{
    return ts.tx == TXR && (ts.ty == 10 || ts.ty == 18 || ts.ty == 26);
}

int flux_step(hipStream_t s, const StepArgs &a, const Fields &f, TileShape ts)
{
    FluxParams p{};
    p.g = a.g;
    p.cin = a.cin;
    p.cout = a.cout;
    p.cdry = a.cdry;
    p.cA = f.cA;
    p.cD = f.cD;
    p.cE = f.cE;
    p.wz = f.wz;
    p.bfac = f.bfac;
    p.tfac = f.tfac;
    p.vsum1 = f.vsum1;
    p.zl = f.zl;
    p.dt64 = a.dt;
    p.dt = (float)a.dt;
    p.dtdx2 = (float)(a.dt / (a.g.dx * a.g.dx));
    p.dtdy2 = (float)(a.dt / (a.g.dy * a.g.dy));
    const bool odd = (a.iiter % 2) != 0;
    if (ts.ty == 10) return launch_flux<16, 8>(s, p, odd);
    if (ts.ty == 18) return launch_flux<24, 8>(s, p, odd);
    if (ts.ty == 26) return launch_flux<32, 8>(s, p, odd);
    set_error("flux kernel: unsupported tile %dx%d", ts.tx, ts.ty);
    return F3D_EINVAL;
}

}  // namespace f32
}  // namespace f3d
