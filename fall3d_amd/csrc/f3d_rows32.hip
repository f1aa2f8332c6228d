// Fused FP32 transport step, row-per-wavefront form (gfx950).
//
// Same step as the tile kernel in f3d_fused32.hip -- setbcc, advctzc, the
// parity-ordered horizontal advection, diffz, the parity-ordered horizontal
// diffusion, accum and the clip, for every class, c_in -> c_out -- laid out
// for CDNA's 64-lane wavefront:
//
// * A tile is 58 interior columns by TY rows; its region (halo 3) is 64
//   columns wide, so one region row is exactly one wavefront (lane = column)
//   and every pipeline stage covers whole rows.  Rows outside a stage's
//   region are skipped by a wave-uniform branch instead of being carried as
//   idle lanes.
// * Row lj of the region belongs to wave lj % NW for the whole kernel, and
//   the column's k-window (advctzc, diffz) lives in that lane's registers.
// * The class-independent factors of the horizontal sweeps (dt/dx*v, the
//   limiter face factor 0.5*dt/dx*(1-Cr)*|v12|, the diffusion face products)
//   are precomputed per meteo read (k_fields_face) and shared by all
//   classes; each sweep then costs two limiter evaluations, one flux
//   difference and a select.
// * Next-level global data is prefetched into registers one level ahead; a
//   workgroup barrier waits for LDS only, so the loads stay in flight across
//   the level's four barriers.
// Tiles whose region touches the domain frame take the EDGE path (ghost
// columns, end-of-line limiter clamps, un-averaged end faces of diffx/diffy).
#include "f3d_ctx.h"

namespace f3d {
namespace f32 {
namespace {

constexpr int kXCD = 8;
constexpr int RX = 64;       // region row = one wavefront
constexpr int TXR = RX - 6;  // interior columns per tile

struct RowParams {
    Grid g;
    const float *cin;
    float *cout;
    double *cdry;
    Fields f;
    const ZCoef *zc;
    double dt64;
    float dt;
    int ntx, nty;
    int nlogical;
    int per_xcd;
};

struct LevelCoef {
    float dtdzz;  // dt*2/(dz(k)+dz(k-1))
    float a2;     // dt*0.5*dz(k-1)*dzz(k)
    float bc;     // dt/dz(k-1)
    float dzk;    // 1/dz(k)
    float dzk_1;  // 1/dz(k-1)
    float ldz;    // 1/dz(k)
};

// minmod of advctx.f90:40: same-signed differences give the one of smaller
// magnitude, anything else (opposite signs or a zero) gives 0
__device__ __forceinline__ float minmod2(float a, float b)
{
    const bool same = (__float_as_uint(a) ^ __float_as_uint(b)) < 0x80000000u;
    const float m = fminf(fabsf(a), fabsf(b));
    return same ? copysignf(m, a) : 0.f;
}

// advctzc W2(k) (advctzc.f90:164-169)
__device__ __forceinline__ float w2z(float cm, float c0, float cp, float vk, float vk_1,
                                     const LevelCoef &z)
{
    const float av = 0.5f * fabsf(vk + vk_1);
    const float sigma = minmod2((cp - c0) * z.dzk, (c0 - cm) * z.dzk_1);
    return z.a2 * (1.f - av * z.bc) * sigma * av;
}

// advctx/advcty update at an LDS position (stride st along the sweep):
// P state, A = dt/h*v (clamped to the domain at the ends), F the limiter
// face factor at face p-1/2; lo/hi flag the first/last domain cell, where
// the reference clamps W2 (advctx.f90:50-52).
__device__ __forceinline__ float adv_face(const float *P, const float *A, const float *Fc, int st,
                                          bool lo, bool hi)
{
    auto W2 = [&](int o) -> float { return Fc[o] * minmod2(P[o + st] - P[o], P[o] - P[o - st]); };
    const float c0 = P[0];
    const float w2i = W2(0);
    if (A[0] > 0.f) {
        const float w1 = A[0] * P[0] - A[-st] * P[-st];
        const float w2m = lo ? w2i : W2(-st);
        return c0 - w1 - (w2i - w2m);
    }
    const float w1 = A[st] * P[st] - A[0] * P[0];
    const float w2p = hi ? w2i : W2(st);
    return c0 - w1 - (w2p - w2i);
}

// interior diffx/diffy update with face products D at face p-1/2 (dt/h^2 included)
__device__ __forceinline__ float diff_face(const float *P, const float *LR, const float *D, int st)
{
    const float q0 = P[0] * LR[0];
    const float gr = D[st] * (P[st] * LR[st] - q0);
    const float gl = D[0] * (q0 - P[-st] * LR[-st]);
    return P[0] + (gr - gl);
}

// diffx/diffy update with raw products PK (diffx.f90:20-53), used where a
// line ends inside the region (the end faces are not averaged)
__device__ __forceinline__ float diff_raw(const float *P, const float *LR, const float *PK, int st,
                                          float dtdh2, bool lo, bool hi)
{
    float gr, gl;
    if (lo) {
        gr = PK[st] * (P[st] * LR[st] - P[0] * LR[0]);
        gl = fmaxf(0.f, PK[0] * (P[0] * LR[0] - P[-st] * LR[0]));
    } else if (hi) {
        gr = fminf(0.f, PK[0] * (P[st] * LR[0] - P[0] * LR[0]));
        gl = PK[-st] * (P[0] * LR[0] - P[-st] * LR[-st]);
    } else {
        gr = 0.5f * (PK[st] + PK[0]) * (P[st] * LR[st] - P[0] * LR[0]);
        gl = 0.5f * (PK[0] + PK[-st]) * (P[0] * LR[0] - P[-st] * LR[-st]);
    }
    return P[0] + dtdh2 * (gr - gl);
}

template <int RY, int NW, bool ODD, bool EDGE>
__device__ __forceinline__ void rows_tile(const RowParams &p, const int q, const int ic,
                                          const int mem, const int gi0, const int gj0, float *sm,
                                          const LevelCoef *zt)
{
    constexpr int RN = RX * RY;
    constexpr int NSLOT = RY / NW;
    static_assert(RY % NW == 0, "region rows must divide evenly over the waves");
    const int nx = p.g.nx, ny = p.g.ny, nz = p.g.nz;
    const unsigned NXY = (unsigned)p.g.NXY, N = (unsigned)p.g.N;
    const unsigned SZ = (unsigned)((nx + 2) * (ny + 2));
    const int li = threadIdx.x & (RX - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gi = gi0 + li;

    float *sA = sm, *sB = sm + RN, *sC = sm + 2 * RN, *sD = sm + 3 * RN;
    float *sAX = sm + 4 * RN, *sFX = sm + 5 * RN, *sAY = sm + 6 * RN, *sFY = sm + 7 * RN;
    float *sLR = sm + 8 * RN, *sDX = sm + 9 * RN, *sDY = sm + 10 * RN;

    const float *cin = p.cin + (size_t)q * p.g.G;
    float *cout = p.cout + (size_t)q * p.g.G;
    const float *fax = p.f.ax, *ffx = p.f.fx, *fay = p.f.ay, *ffy = p.f.fy;
    const float *fvz = p.f.vz, *flr = p.f.lr, *fpz = p.f.pkz;
    // diffusion planes: face products inside, raw products where lines end
    const float *fdx = EDGE ? p.f.pkx : p.f.dfx;
    const float *fdy = EDGE ? p.f.pky : p.f.dfy;
    const float *wz = p.f.wz + (size_t)ic * N;

    const bool lane_c = li >= 2 && li < RX - 2;  // diffz region along the row

    int kind[NSLOT];  // EDGE only: 0 other, 1 interior, 2 x-ghost, 3 y-ghost
    unsigned m2[NSLOT], coff[NSLOT];
#pragma unroll
    for (int sl = 0; sl < NSLOT; ++sl) {
        const int lj = w + NW * sl;
        const int gj = gj0 + lj;
        kind[sl] = 1;
        if (EDGE) {
            const bool inx = gi >= 1 && gi <= nx, iny = gj >= 1 && gj <= ny;
            kind[sl] = (inx && iny) ? 1
                     : (iny && (gi == 0 || gi == nx + 1)) ? 2
                     : (inx && (gj == 0 || gj == ny + 1)) ? 3 : 0;
        }
        const int ci = EDGE ? min(max(gi, 1), nx) : gi;
        const int cj = EDGE ? min(max(gj, 1), ny) : gj;
        m2[sl] = (unsigned)((ci - 1) + nx * (cj - 1));
        coff[sl] = (unsigned)cidx(nx, ny, ci, cj, 0);
    }

    float cm1[NSLOT], c0[NSLOT], cp1[NSLOT], vk[NSLOT], W1k[NSLOT], W2km1[NSLOT], W2k[NSLOT];
    float gtopm[NSLOT];
    float bm2[NSLOT], bm1[NSLOT], lrm2[NSLOT], lrm1[NSLOT], pzm2[NSLOT], pzm1[NSLOT];
    float nc2[NSLOT], nvz[NSLOT], nax[NSLOT], nfx[NSLOT], nay[NSLOT], nfy[NSLOT];
    float nlr[NSLOT], npz[NSLOT], ndx[NSLOT], ndy[NSLOT];

    const LevelCoef z1 = zt[1];
#pragma unroll
    for (int sl = 0; sl < NSLOT; ++sl) {
        const int lj = w + NW * sl;
        const bool crow = lj >= 2 && lj < RY - 2;
        cm1[sl] = c0[sl] = cp1[sl] = vk[sl] = W1k[sl] = W2km1[sl] = W2k[sl] = 0.f;
        gtopm[sl] = 0.f;
        bm2[sl] = bm1[sl] = lrm2[sl] = lrm1[sl] = pzm2[sl] = pzm1[sl] = 0.f;
        nc2[sl] = nvz[sl] = 0.f;
        nlr[sl] = npz[sl] = ndx[sl] = ndy[sl] = 0.f;
        const unsigned o = coff[sl], m = m2[sl];
        if (kind[sl] == 1) {
            const float c1 = cin[o + SZ];
            const float c2 = cin[o + 2 * SZ];
            const float bf = p.f.bfac[(size_t)ic * NXY + m];
            gtopm[sl] = p.f.tfac[(size_t)ic * NXY + m];
            const float gb = c1 * bf;  // setbcc.f90:108-113
            cm1[sl] = gb;
            c0[sl] = c1;
            cp1[sl] = c2;
            const float v1 = wz[m];
            vk[sl] = v1;
            W1k[sl] = z1.dtdzz * (v1 * c1 - v1 * gb);
            W2k[sl] = w2z(gb, c1, c2, v1, v1, z1);
            W2km1[sl] = W2k[sl];
            bm1[sl] = gb;  // b(0): becomes bm2 after the level-1 shift
            nc2[sl] = (3 <= nz) ? cin[o + 3 * SZ] : 0.f;
            nvz[sl] = wz[NXY + m];
                        if (crow && lane_c) {
                nlr[sl] = flr[m];
                npz[sl] = fpz[m];
            }
        } else if (EDGE && kind[sl] >= 2) {
            nc2[sl] = cin[o + SZ];  // face neighbour at level 1
        }
        nax[sl] = fax[m];
        nfx[sl] = ffx[m];
        nay[sl] = fay[m];
        nfy[sl] = ffy[m];
    }

    const float dt = p.dt;
    const float dtdx2 = (float)(p.dt64 / (p.g.dx * p.g.dx));
    const float dtdy2 = (float)(p.dt64 / (p.g.dy * p.g.dy));

    for (int s = 1; s <= nz + 1; ++s) {
        const unsigned lv = (unsigned)(s - 1) * NXY;
        // ---------------- stage A: advctzc at level s, planes for level s ----------------
        if (s <= nz) {
            const LevelCoef zn = zt[s + 1];
            const float dtop = dt * zt[nz].ldz;
#pragma unroll
            for (int sl = 0; sl < NSLOT; ++sl) {
                const int lj = w + NW * sl;
                const int r = lj * RX + li;
                float a = 0.f;
                if (kind[sl] == 1) {
                    float W1n, W2n, vkn = 0.f, cp2 = 0.f;
                    if (s < nz) {
                        cp2 = (s + 2 <= nz) ? nc2[sl] : gtopm[sl] * cp1[sl];  // top ghost (setbcc.f90:126-130)
                        vkn = nvz[sl];
                        W1n = zn.dtdzz * (vkn * cp1[sl] - vk[sl] * c0[sl]);
                        W2n = w2z(c0[sl], cp1[sl], cp2, vkn, vk[sl], zn);
                    } else {
                        W1n = dtop * vk[sl] * (cp1[sl] - c0[sl]);  // advctzc.f90:175-181
                        W2n = W2k[sl];
                    }
                    if (vk[sl] > 0.f)
                        a = c0[sl] - W1k[sl] - (W2k[sl] - W2km1[sl]);
                    else
                        a = c0[sl] - W1n - (W2n - W2k[sl]);
                    cm1[sl] = c0[sl];
                    c0[sl] = cp1[sl];
                    cp1[sl] = cp2;
                    vk[sl] = vkn;
                    W1k[sl] = W1n;
                    W2km1[sl] = W2k[sl];
                    W2k[sl] = W2n;
                    if (s + 3 <= nz) nc2[sl] = cin[coff[sl] + (unsigned)(s + 3) * SZ];
                    if (s + 2 <= nz) {
                        nvz[sl] = wz[lv + 2 * NXY + m2[sl]];
                                            }
                } else if (EDGE && kind[sl] >= 2) {
                    // face ghost (setbcc.f90:139-169) from the adjacent cell;
                    // dt/h*v carries the sign of the wind
                    const bool zero = (kind[sl] == 2) ? ((gi == 0) ? (nax[sl] >= 0.f) : (nax[sl] <= 0.f))
                                                      : ((gj0 + lj == 0) ? (nay[sl] >= 0.f) : (nay[sl] <= 0.f));
                    a = zero ? 0.f : nc2[sl];
                    if (s + 1 <= nz) nc2[sl] = cin[coff[sl] + (unsigned)(s + 1) * SZ];
                }
                sA[r] = a;
                sAX[r] = nax[sl];
                sFX[r] = nfx[sl];
                sAY[r] = nay[sl];
                sFY[r] = nfy[sl];
                if (s + 1 <= nz) {
                    const unsigned o1 = lv + NXY + m2[sl];
                    nax[sl] = fax[o1];
                    nfx[sl] = ffx[o1];
                    nay[sl] = fay[o1];
                    nfy[sl] = ffy[o1];
                }
            }
        }
        __syncthreads();

        // ---------------- stage B1: first horizontal advection at level s ----------------
        if (s <= nz) {
#pragma unroll
            for (int sl = 0; sl < NSLOT; ++sl) {
                const int lj = w + NW * sl;
                const int r = lj * RX + li;
                const int gj = gj0 + lj;
                if (ODD) {
                    // advcty on rows [2, RY-2): read rows lj-2 .. lj+2
                    if (lj >= 2 && lj < RY - 2) {
                        float out = sA[r];
                        if (kind[sl] == 1)
                            out = adv_face(&sA[r], &sAY[r], &sFY[r], RX, EDGE && gj == 1,
                                           EDGE && gj == ny);
                        sB[r] = out;
                    }
                } else {
                    // advctx on every row, lanes [2, RX-2)
                    if (lane_c) {
                        float out = sA[r];
                        if (kind[sl] == 1)
                            out = adv_face(&sA[r], &sAX[r], &sFX[r], 1, EDGE && gi == 1,
                                           EDGE && gi == nx);
                        sB[r] = out;
                    }
                }
            }
        }
        __syncthreads();

        // ---------------- B2 then stage C: diffz at level s-1 ----------------
#pragma unroll
        for (int sl = 0; sl < NSLOT; ++sl) {
            const int lj = w + NW * sl;
            if (lj < 2 || lj >= RY - 2) continue;  // wave-uniform
            if (!lane_c) continue;
            const int r = lj * RX + li;
            const int gj = gj0 + lj;
            float bnew;
            if (s <= nz) {
                bnew = sB[r];
                if (kind[sl] == 1) {
                    if (ODD)
                        bnew = adv_face(&sB[r], &sAX[r], &sFX[r], 1, EDGE && gi == 1,
                                        EDGE && gi == nx);
                    else
                        bnew = adv_face(&sB[r], &sAY[r], &sFY[r], RX, EDGE && gj == 1,
                                        EDGE && gj == ny);
                }
            } else {
                bnew = (kind[sl] == 1) ? gtopm[sl] * cm1[sl] : bm1[sl];  // b(nz+1) = top ghost
            }
            if (kind[sl] == 1) {
                const float lrn = nlr[sl], pzn = npz[sl];  // level s (unused at s = nz+1)
                if (s >= 2) {
                    const int k = s - 1;
                    float dnew;
                    if (k == 1) {
                        const float gt = pzn * (bnew * lrn - bm1[sl] * lrm1[sl]) * zt[1].ldz;
                        const float gb =
                            fmaxf(0.f, pzm1[sl] * (bm1[sl] * lrm1[sl] - bm2[sl] * lrm1[sl]) * zt[0].ldz);
                        dnew = bm1[sl] + zt[1].dtdzz * (gt - gb);
                    } else if (k < nz) {
                        const float gt = 0.5f * (pzn + pzm1[sl]) * (bnew * lrn - bm1[sl] * lrm1[sl]) * zt[k].ldz;
                        const float gb = 0.5f * (pzm1[sl] + pzm2[sl]) *
                                         (bm1[sl] * lrm1[sl] - bm2[sl] * lrm2[sl]) * zt[k - 1].ldz;
                        dnew = bm1[sl] + zt[k].dtdzz * (gt - gb);
                    } else {
                        const float gt =
                            fminf(0.f, pzm1[sl] * (bnew * lrm1[sl] - bm1[sl] * lrm1[sl]) * zt[nz].ldz);
                        const float gb = pzm2[sl] * (bm1[sl] * lrm1[sl] - bm2[sl] * lrm2[sl]) * zt[nz - 1].ldz;
                        dnew = bm1[sl] + dt * zt[nz].ldz * (gt - gb);
                    }
                    sC[r] = dnew;
                    sLR[r] = lrm1[sl];
                    sDX[r] = ndx[sl];
                    sDY[r] = ndy[sl];
                }
                bm2[sl] = bm1[sl];
                bm1[sl] = bnew;
                lrm2[sl] = lrm1[sl];
                lrm1[sl] = lrn;
                pzm2[sl] = pzm1[sl];
                pzm1[sl] = pzn;
                // prefetch for the next level: lr/pz(s+1), diffusion planes(s)
                if (s + 1 <= nz) {
                    nlr[sl] = flr[lv + NXY + m2[sl]];
                    npz[sl] = fpz[lv + NXY + m2[sl]];
                }
                if (s <= nz) {
                    ndx[sl] = fdx[lv + m2[sl]];
                    ndy[sl] = fdy[lv + m2[sl]];
                }
            } else {
                if (s >= 2) {
                    sC[r] = bm1[sl];
                    sLR[r] = 0.f;
                    sDX[r] = 0.f;
                    sDY[r] = 0.f;
                }
                bm1[sl] = bnew;
            }
        }
        __syncthreads();
        if (s < 2) continue;

        // ---------------- stage D1: first horizontal diffusion at level s-1 ----------------
#pragma unroll
        for (int sl = 0; sl < NSLOT; ++sl) {
            const int lj = w + NW * sl;
            const int r = lj * RX + li;
            const int gj = gj0 + lj;
            if (ODD) {
                // diffy on rows [3, RY-3), lanes [2, RX-2)
                if (lj < 3 || lj >= RY - 3 || !lane_c) continue;
                float out = sC[r];
                if (!EDGE)
                    out = diff_face(&sC[r], &sLR[r], &sDY[r], RX);
                else if (kind[sl] == 1)
                    out = diff_raw(&sC[r], &sLR[r], &sDY[r], RX, dtdy2, gj == 1, gj == ny);
                sD[r] = out;
            } else {
                // diffx on rows [2, RY-2), lanes [3, RX-3)
                if (lj < 2 || lj >= RY - 2 || li < 3 || li >= RX - 3) continue;
                float out = sC[r];
                if (!EDGE)
                    out = diff_face(&sC[r], &sLR[r], &sDX[r], 1);
                else if (kind[sl] == 1)
                    out = diff_raw(&sC[r], &sLR[r], &sDX[r], 1, dtdx2, gi == 1, gi == nx);
                sD[r] = out;
            }
        }
        __syncthreads();

        // ---------------- stage D2: last sweep, output, deposit ----------------
        {
            const int k = s - 1;
            float *co = cout + (unsigned)k * SZ;
#pragma unroll
            for (int sl = 0; sl < NSLOT; ++sl) {
                const int lj = w + NW * sl;
                if (lj < 3 || lj >= RY - 3) continue;  // wave-uniform
                if (li < 3 || li >= RX - 3) continue;
                const int gj = gj0 + lj;
                if (EDGE && (gi > nx || gj > ny)) continue;
                const int r = lj * RX + li;
                float out;
                if (ODD) {
                    out = EDGE ? diff_raw(&sD[r], &sLR[r], &sDX[r], 1, dtdx2, gi == 1, gi == nx)
                               : diff_face(&sD[r], &sLR[r], &sDX[r], 1);
                } else {
                    out = EDGE ? diff_raw(&sD[r], &sLR[r], &sDY[r], RX, dtdy2, gj == 1, gj == ny)
                               : diff_face(&sD[r], &sLR[r], &sDY[r], RX);
                }
                co[coff[sl]] = fmaxf(0.f, out);  // sizloop.f90:77-78
                if (k == 1) {
                    // accum.f90:22, deposit from c(:,:,1) before the clip, FP64
                    const size_t t2 = m2[sl];
                    double *cd = p.cdry + ((size_t)mem * (p.g.nc + 1) + ic + 1) * NXY + t2;
                    *cd = *cd - p.dt64 * (double)out * p.f.vsum1[(size_t)ic * NXY + t2];
                }
            }
        }
    }
}

template <int RY, int NW, bool ODD>
__global__ void __launch_bounds__(NW * 64) k_rows(RowParams p)
{
    constexpr int RN = RX * RY;
    const int b = blockIdx.x;
    const int L = (b % kXCD) * p.per_xcd + b / kXCD;  // XCD-aware logical block
    if (L >= p.nlogical) return;
    const int nc = p.g.nc;
    const int nq = nc * p.g.nm;
    const int q = L % nq;
    const int ic = q % nc, mem = q / nc;
    const int tile = L / nq;
    const int bx = tile % p.ntx, by = tile / p.ntx;
    const int gi0 = 1 + bx * TXR - 3, gj0 = 1 + by * (RY - 6) - 3;

    __shared__ float sm[11 * RN];
    __shared__ LevelCoef zt[66];
    const int nz = p.g.nz;
    for (int k = threadIdx.x; k <= nz + 1 && k < 66; k += NW * 64) {
        LevelCoef z{};
        if (k >= 1 && k <= nz) {
            const ZCoef c = p.zc[k];
            z.dtdzz = (float)(p.dt64 * c.dzz);
            z.a2 = (float)(p.dt64 * 0.5 * c.dzm1 * c.dzz);
            z.bc = (float)(p.dt64 * c.dzk_1);
            z.dzk = (float)c.dzk;
            z.dzk_1 = (float)c.dzk_1;
            z.ldz = (float)c.dzk;
        }
        if (k == 0) z.ldz = (float)p.zc[1].dzk_1;
        zt[k] = z;
    }
    __syncthreads();
    const bool edge = gi0 < 1 || gj0 < 1 || gi0 + RX - 1 > p.g.nx || gj0 + RY - 1 > p.g.ny;
    if (edge)
        rows_tile<RY, NW, ODD, true>(p, q, ic, mem, gi0, gj0, sm, zt);
    else
        rows_tile<RY, NW, ODD, false>(p, q, ic, mem, gi0, gj0, sm, zt);
}

template <int RY, int NW>
int launch_rows(hipStream_t s, const RowParams &p0, bool odd)
{
    RowParams p = p0;
    p.ntx = (p.g.nx + TXR - 1) / TXR;
    p.nty = (p.g.ny + (RY - 6) - 1) / (RY - 6);
    p.nlogical = p.ntx * p.nty * p.g.nc * p.g.nm;
    p.per_xcd = (p.nlogical + kXCD - 1) / kXCD;
    const unsigned nb = (unsigned)(p.per_xcd * kXCD);
    if (odd)
        hipLaunchKernelGGL((k_rows<RY, NW, true>), dim3(nb), dim3(NW * 64), 0, s, p);
    else
        hipLaunchKernelGGL((k_rows<RY, NW, false>), dim3(nb), dim3(NW * 64), 0, s, p);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

}  // namespace

bool rows_supported(TileShape ts)
{
    return ts.tx == TXR && (ts.ty == 10 || ts.ty == 18 || ts.ty == 26);
}

int rows_step(hipStream_t s, const StepArgs &a, const Fields &f, TileShape ts)
{
    RowParams p{};
    p.g = a.g;
    p.cin = a.cin;
    p.cout = a.cout;
    p.cdry = a.cdry;
    p.f = f;
    p.zc = a.zc;
    p.dt64 = a.dt;
    p.dt = (float)a.dt;
    const bool odd = (a.iiter % 2) != 0;
    if (ts.ty == 10) return launch_rows<16, 8>(s, p, odd);
    if (ts.ty == 18) return launch_rows<24, 8>(s, p, odd);
    if (ts.ty == 26) return launch_rows<32, 8>(s, p, odd);
    set_error("row kernel: unsupported tile %dx%d", ts.tx, ts.ty);
    return F3D_EINVAL;
}

}  // namespace f32
}  // namespace f3d
