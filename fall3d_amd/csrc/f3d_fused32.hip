// Fused FP32 transport step for MI355X (gfx950).
//
// One launch performs, for every class, the whole class step of
// Sources/sizloop.f90:31-78 -- setbcc, advctzc, advctx/advcty in iiter
// parity order, diffz, diffx/diffy in parity order, accum and the clip --
// reading c_in and writing c_out (ping-pong buffers, so neighbouring tiles
// always read pre-step values).  A second small kernel writes the ghost frame
// of c_out and the cmass2d outflux.  Deposit and budgets accumulate in FP64.
//
// Work decomposition: one workgroup (256 lanes, 4 waves) per (tile, class).
// A tile is TX x TY columns; the workgroup holds a (TX+6) x (TY+6) region
// (halo 3 = advection radius 2 + diffusion radius 1) and marches up the
// column one level at a time as a software pipeline:
//   A  advctzc at level s        per column, register window along k
//   B1 first horizontal advection at level s        LDS plane
//   B2 second horizontal advection at level s       LDS plane -> registers
//   C  diffz at level s-1        per column, register window along k
//   D1 first horizontal diffusion at level s-1      LDS plane
//   D2 second horizontal diffusion at level s-1     -> c_out, deposit
// Each region column is owned by one lane for the whole kernel (lane =
// column % 256), so the z windows live in registers and only one plane per
// stage is exchanged through LDS (four barriers per level).  Ghost cells
// never change during a step (the reference refreshes them once per class
// step, before advection, sizloop.f90:35), so ghost columns pass through
// every stage with their setbcc value.
//
// Block order is XCD-aware: consecutive logical blocks (the classes of one
// tile) are placed on the same XCD so the tile's meteorology is fetched
// from HBM once and served to its classes from that XCD's L2.
#include "f3d_ctx.h"

namespace f3d {
namespace f32 {
namespace {

constexpr int kNT = 256;
constexpr int kXCD = 8;

__device__ __forceinline__ float minmodf(float dr, float dl)
{
    const float m = fminf(fabsf(dr), fabsf(dl));
    return (dr > 0.f) == (dl > 0.f) && dr != 0.f && dl != 0.f ? copysignf(m, dr) : 0.f;
}

struct LevelCoef {
    float dtdzz;  // dt*2/(dz(k)+dz(k-1))
    float a2;     // dt*0.5*dz(k-1)*dzz(k)
    float bc;     // dt/dz(k-1)
    float dzk;    // 1/dz(k)
    float dzk_1;  // 1/dz(k-1)
    float ldz;    // 1/dz(k)
};

struct Params {
    Grid g;
    const float *cin;
    float *cout;
    double *cdry;
    Fields f;
    const ZCoef *zc;
    double dt64;
    float dt;
    int odd;  // iiter odd: y sweeps first (sizloop.f90:42-61)
    int ntx, nty;
    int nlogical;
    int per_xcd;
};

// advctzc W2(k) (advctzc.f90:164-169), FP32
__device__ __forceinline__ float w2z(float cm, float c0, float cp, float vk, float vk_1,
                                     const LevelCoef &z)
{
    const float v12 = 0.5f * (vk + vk_1);
    const float av = fabsf(v12);
    const float sigma = minmodf((cp - c0) * z.dzk, (c0 - cm) * z.dzk_1);
    return z.a2 * (1.f - av * z.bc) * sigma * av;
}

// advctx/advcty update at a line position held in an LDS plane with stride
// st: P plane values, V velocity (clamped to the domain at the ends),
// auxv = dt/h, kh = 0.5*dt/h; lo/hi flag the first/last domain cell.
__device__ __forceinline__ float adv_update(const float *P, const float *V, int st, float auxv,
                                            float kh, bool lo, bool hi)
{
    auto W2 = [&](int o) -> float {
        const float v12 = 0.5f * (V[o] + V[o - st]);
        const float av = fabsf(v12);
        const float sigma = minmodf(P[o + st] - P[o], P[o] - P[o - st]);
        return kh * (1.f - auxv * av) * sigma * av;
    };
    const float c0 = P[0];
    const float w2i = W2(0);
    if (V[0] > 0.f) {
        const float w1 = auxv * (V[0] * P[0] - V[-st] * P[-st]);
        const float w2m = lo ? w2i : W2(-st);
        return c0 - w1 - (w2i - w2m);
    }
    const float w1 = auxv * (V[st] * P[st] - V[0] * P[0]);
    const float w2p = hi ? w2i : W2(st);
    return c0 - w1 - (w2p - w2i);
}

// diffx/diffy update (diffx.f90:20-53), FP32 with reciprocal density
__device__ __forceinline__ float diff_update(const float *P, const float *LR, const float *PK,
                                             int st, float dtdh2, bool lo, bool hi)
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

// One (tile, class) of the fused step.  ODD selects the sweep order of
// the step (iiter odd: y sweeps first, sizloop.f90:42-61); EDGE is false for
// tiles whose whole region lies inside the domain, which then need none of
// the ghost-column and end-of-line cases.
template <int TX, int TY, bool ODD, bool EDGE>
__device__ __forceinline__ void fused_tile(const Params &p, const int q, const int ic,
                                           const int mem, const int gi0, const int gj0,
                                           float *sA, float *sB, float *sC, float *sD,
                                           float *sVX, float *sVY, float *sLR, float *sPX,
                                           float *sPY, const LevelCoef *zt)
{
    constexpr int RX = TX + 6, RY = TY + 6, RN = RX * RY;
    constexpr int NSLOT = (RN + kNT - 1) / kNT;
    const int nx = p.g.nx, ny = p.g.ny, nz = p.g.nz;
    const unsigned NXY = (unsigned)p.g.NXY, N = (unsigned)p.g.N;
    const unsigned SZ = (unsigned)((nx + 2) * (ny + 2));
    const int tid = threadIdx.x;

    const float *cin = p.cin + (size_t)q * p.g.G;
    float *cout = p.cout + (size_t)q * p.g.G;
    const float *fvx = p.f.vx, *fvy = p.f.vy, *fvz = p.f.vz, *flr = p.f.lr;
    const float *fpx = p.f.pkx, *fpy = p.f.pky, *fpz = p.f.pkz;
    const float *wz = p.f.wz + (size_t)ic * N;

    // per-slot column data
    int kind[NSLOT];       // EDGE only: 0 other, 1 interior, 2 x-ghost, 3 y-ghost
    bool ctr[NSLOT];       // inside the diffz region (halo 1)
    unsigned m2[NSLOT];    // 2-D offset (i-1)+nx*(j-1) (clamped into the domain)
    unsigned coff[NSLOT];  // c offset at k=0 of the column (or of its face neighbour)
#pragma unroll
    for (int sl = 0; sl < NSLOT; ++sl) {
        const int r = tid + kNT * sl;
        const int li = r % RX, lj = r / RX;
        const int gi = gi0 + li, gj = gj0 + lj;
        kind[sl] = (r < RN) ? 1 : 0;
        if (EDGE && r < RN) {
            const bool inx = gi >= 1 && gi <= nx, iny = gj >= 1 && gj <= ny;
            kind[sl] = (inx && iny) ? 1
                     : (iny && (gi == 0 || gi == nx + 1)) ? 2
                     : (inx && (gj == 0 || gj == ny + 1)) ? 3 : 0;
        }
        const int ci = EDGE ? min(max(gi, 1), nx) : gi;
        const int cj = EDGE ? min(max(gj, 1), ny) : gj;
        m2[sl] = (r < RN) ? (unsigned)((ci - 1) + nx * (cj - 1)) : 0u;
        coff[sl] = (r < RN) ? (unsigned)cidx(nx, ny, ci, cj, 0) : 0u;
        ctr[sl] = (r < RN) && li >= 2 && li < RX - 2 && lj >= 2 && lj < RY - 2;
    }

    float cm1[NSLOT], c0[NSLOT], cp1[NSLOT], vk[NSLOT], W1k[NSLOT], W2km1[NSLOT], W2k[NSLOT];
    float gtopm[NSLOT];
    float bm2[NSLOT], bm1[NSLOT], lrm2[NSLOT], lrm1[NSLOT], pzm2[NSLOT], pzm1[NSLOT];
    // next level's global data, loaded one level ahead so that memory latency
    // overlaps the current level's compute
    float nc2[NSLOT], nvz[NSLOT], nvx[NSLOT], nvy[NSLOT];
    float nlr[NSLOT], npz[NSLOT], npx[NSLOT], npy[NSLOT];

    const LevelCoef z1 = zt[1];
#pragma unroll
    for (int sl = 0; sl < NSLOT; ++sl) {
        cm1[sl] = c0[sl] = cp1[sl] = vk[sl] = W1k[sl] = W2km1[sl] = W2k[sl] = 0.f;
        gtopm[sl] = 0.f;
        bm2[sl] = bm1[sl] = lrm2[sl] = lrm1[sl] = pzm2[sl] = pzm1[sl] = 0.f;
        nc2[sl] = nvz[sl] = nvx[sl] = nvy[sl] = 0.f;
        nlr[sl] = npz[sl] = npx[sl] = npy[sl] = 0.f;
        if (kind[sl] == 1) {
            const unsigned o = coff[sl], m = m2[sl];
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
                        if (ctr[sl]) {
                nlr[sl] = flr[m];
                npz[sl] = fpz[m];
            }
        } else if (EDGE && kind[sl] >= 2) {
            nc2[sl] = cin[coff[sl] + SZ];  // face neighbour at level 1
        }
        if (EDGE || kind[sl] == 1) {
            nvx[sl] = fvx[m2[sl]];
            nvy[sl] = fvy[m2[sl]];
        }
    }

    const float dt = p.dt;
    const float auxx = (float)(p.dt64 / p.g.dx), khx = 0.5f * auxx;
    const float auxy = (float)(p.dt64 / p.g.dy), khy = 0.5f * auxy;
    const float dtdx2 = (float)(p.dt64 / (p.g.dx * p.g.dx));
    const float dtdy2 = (float)(p.dt64 / (p.g.dy * p.g.dy));

    for (int s = 1; s <= nz + 1; ++s) {
        const unsigned lv = (unsigned)(s - 1) * NXY;  // level s offset in (nx,ny,nz) fields
        // ---------------- stage A: advctzc at level s ----------------
        if (s <= nz) {
            const LevelCoef zn = zt[s + 1];
            const float dtop = dt * zt[nz].ldz;
#pragma unroll
            for (int sl = 0; sl < NSLOT; ++sl) {
                const int r = tid + kNT * sl;
                if (r >= RN) continue;
                float a = 0.f;
                const float vxs = nvx[sl], vys = nvy[sl];
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
                    // face ghost (setbcc.f90:139-169) from the adjacent cell
                    const int li = r % RX, lj = r / RX;
                    const bool zero = (kind[sl] == 2)
                                          ? ((gi0 + li == 0) ? (vxs >= 0.f) : (vxs <= 0.f))
                                          : ((gj0 + lj == 0) ? (vys >= 0.f) : (vys <= 0.f));
                    a = zero ? 0.f : nc2[sl];
                    if (s + 1 <= nz) nc2[sl] = cin[coff[sl] + (unsigned)(s + 1) * SZ];
                }
                sA[r] = a;
                sVX[r] = vxs;
                sVY[r] = vys;
                if (s + 1 <= nz && (EDGE || kind[sl] == 1)) {
                    nvx[sl] = fvx[lv + NXY + m2[sl]];
                    nvy[sl] = fvy[lv + NXY + m2[sl]];
                }
            }
        }
        __syncthreads();

        // ---------------- stage B1: first horizontal advection at level s ----------------
        if (s <= nz) {
#pragma unroll
            for (int sl = 0; sl < NSLOT; ++sl) {
                const int r = tid + kNT * sl;
                if (r >= RN) continue;
                const int li = r % RX, lj = r / RX;
                float out = sA[r];
                if (kind[sl] == 1) {
                    const int gi = gi0 + li, gj = gj0 + lj;
                    if (ODD) {
                        if (lj >= 2 && lj < RY - 2)
                            out = adv_update(&sA[r], &sVY[r], RX, auxy, khy, EDGE && gj == 1,
                                             EDGE && gj == ny);
                    } else {
                        if (li >= 2 && li < RX - 2)
                            out = adv_update(&sA[r], &sVX[r], 1, auxx, khx, EDGE && gi == 1,
                                             EDGE && gi == nx);
                    }
                }
                sB[r] = out;
            }
        }
        __syncthreads();

        // B2 (owner lanes of the diffz region) then stage C: diffz at level s-1
#pragma unroll
        for (int sl = 0; sl < NSLOT; ++sl) {
            const int r = tid + kNT * sl;
            if (!ctr[sl]) continue;
            float bnew;
            if (s <= nz) {
                bnew = sB[r];
                if (kind[sl] == 1) {
                    const int li = r % RX, lj = r / RX;
                    const int gi = gi0 + li, gj = gj0 + lj;
                    if (ODD)
                        bnew = adv_update(&sB[r], &sVX[r], 1, auxx, khx, EDGE && gi == 1,
                                          EDGE && gi == nx);
                    else
                        bnew = adv_update(&sB[r], &sVY[r], RX, auxy, khy, EDGE && gj == 1,
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
                    sPX[r] = npx[sl];
                    sPY[r] = npy[sl];
                }
                bm2[sl] = bm1[sl];
                bm1[sl] = bnew;
                lrm2[sl] = lrm1[sl];
                lrm1[sl] = lrn;
                pzm2[sl] = pzm1[sl];
                pzm1[sl] = pzn;
                // prefetch for the next level: lr/pz(s+1), pkx/pky(s)
                if (s + 1 <= nz) {
                    nlr[sl] = flr[lv + NXY + m2[sl]];
                    npz[sl] = fpz[lv + NXY + m2[sl]];
                }
                if (s <= nz) {
                    npx[sl] = fpx[lv + m2[sl]];
                    npy[sl] = fpy[lv + m2[sl]];
                }
            } else {
                if (s >= 2) {
                    sC[r] = bm1[sl];
                    sLR[r] = 0.f;
                    sPX[r] = 0.f;
                    sPY[r] = 0.f;
                }
                bm1[sl] = bnew;
            }
        }
        __syncthreads();
        if (s < 2) continue;

        // ---------------- stage D1: first horizontal diffusion at level s-1 ----------------
        {
            // ODD: diffy first on li in [2,RX-2), lj in [3,RY-3); else diffx first on
            // li in [3,RX-3), lj in [2,RY-2)
            constexpr int nli = ODD ? RX - 4 : TX, nlj = ODD ? TY : RY - 4;
            constexpr int li0 = ODD ? 2 : 3, lj0 = ODD ? 3 : 2;
#pragma unroll
            for (int e0 = 0; e0 < nli * nlj; e0 += kNT) {
                const int e = e0 + tid;
                if (e >= nli * nlj) continue;
                const int li = li0 + e % nli, lj = lj0 + e / nli;
                const int r = lj * RX + li;
                const int gi = gi0 + li, gj = gj0 + lj;
                const bool interior = !EDGE || (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny);
                float out = sC[r];
                if (interior) {
                    if (ODD)
                        out = diff_update(&sC[r], &sLR[r], &sPY[r], RX, dtdy2, EDGE && gj == 1,
                                          EDGE && gj == ny);
                    else
                        out = diff_update(&sC[r], &sLR[r], &sPX[r], 1, dtdx2, EDGE && gi == 1,
                                          EDGE && gi == nx);
                }
                sD[r] = out;
            }
        }
        __syncthreads();

        // ---------------- stage D2: last sweep, output, deposit ----------------
        {
            const int k = s - 1;
            float *co = cout + (unsigned)k * SZ;
#pragma unroll
            for (int e0 = 0; e0 < TX * TY; e0 += kNT) {
                const int e = e0 + tid;
                if (e >= TX * TY) continue;
                const int li = 3 + e % TX, lj = 3 + e / TX;
                const int gi = gi0 + li, gj = gj0 + lj;
                if (EDGE && (gi > nx || gj > ny)) continue;
                const int r = lj * RX + li;
                float out;
                if (ODD)
                    out = diff_update(&sD[r], &sLR[r], &sPX[r], 1, dtdx2, EDGE && gi == 1,
                                      EDGE && gi == nx);
                else
                    out = diff_update(&sD[r], &sLR[r], &sPY[r], RX, dtdy2, EDGE && gj == 1,
                                      EDGE && gj == ny);
                co[(unsigned)cidx(nx, ny, gi, gj, 0)] = fmaxf(0.f, out);  // sizloop.f90:77-78
                if (k == 1) {
                    // accum.f90:22, deposit from c(:,:,1) before the clip, FP64
                    const size_t t2 = (gi - 1) + (size_t)nx * (gj - 1);
                    double *cd = p.cdry + ((size_t)mem * (p.g.nc + 1) + ic + 1) * NXY + t2;
                    *cd = *cd - p.dt64 * (double)out * p.f.vsum1[(size_t)ic * NXY + t2];
                }
            }
        }
        // the next level's stage A writes sA/sVX/sVY, last read in B2 before the
        // barrier after stage C; stage D planes are rewritten only after two
        // more barriers
    }
}

template <int TX, int TY, bool ODD>
__global__ void __launch_bounds__(kNT) k_fused(Params p)
{
    constexpr int RX = TX + 6, RY = TY + 6, RN = RX * RY;
    // XCD-aware logical block id: consecutive logical ids (the classes and
    // members of one tile) land on one XCD
    const int b = blockIdx.x;
    const int L = (b % kXCD) * p.per_xcd + b / kXCD;
    if (L >= p.nlogical) return;
    const int nc = p.g.nc;
    const int nq = nc * p.g.nm;  // member-major slabs: q = member*nc + class
    const int q = L % nq;
    const int ic = q % nc, mem = q / nc;
    const int tile = L / nq;
    const int bx = tile % p.ntx, by = tile / p.ntx;
    const int gi0 = 1 + bx * TX - 3, gj0 = 1 + by * TY - 3;

    __shared__ float sA[RN], sB[RN], sC[RN], sD[RN];
    __shared__ float sVX[RN], sVY[RN], sLR[RN], sPX[RN], sPY[RN];
    __shared__ LevelCoef zt[66];

    const int nz = p.g.nz;
    for (int k = threadIdx.x; k <= nz + 1 && k < 66; k += kNT) {
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
        fused_tile<TX, TY, ODD, true>(p, q, ic, mem, gi0, gj0, sA, sB, sC, sD, sVX, sVY, sLR, sPX,
                                      sPY, zt);
    else
        fused_tile<TX, TY, ODD, false>(p, q, ic, mem, gi0, gj0, sA, sB, sC, sD, sVX, sVY, sLR,
                                       sPX, sPY, zt);
}

// ---------------- ghost frame + cmass2d outflux ----------------
// setbcc (setbcc.f90:19-109) values of c_in written to c_out's frame, then
// clipped; cmass2d terms (cmass2d.f90:68-127) of the clipped ghosts, FP64.
// Block bx of nbx for member mem: a grid-stride share of the frame cells of
// every class; the block's outflux sum goes to part[mem*nbx + bx].
__device__ __forceinline__ void frame_block(const FrameArgs &a, int bx, int nbx, int mem)
{
    __shared__ double sh[256];
    const Grid &g = a.g;
    const float *cin = a.cin;
    float *cout = a.cout;
    const double *vx = a.vx, *vy = a.vy, *vz = a.vz, *vset = a.vset, *dz = a.dz;
    const double dt = a.dt;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long nxf = (long)ny * nz, nyf = (long)nx * nz, nxy = g.NXY;
    const long nedge = 4L * nz + 4L * ny + 4L * nx + 8;
    const long per = nxf + nyf + nxy + nedge;
    const long total = per * g.nc;
    double acc = 0.0;
    for (long t = (long)bx * 256 + threadIdx.x; t < total; t += (long)nbx * 256) {
        const int q = (int)(t / per);  // class
        long r = t - (long)q * per;
        const float *ci = cin + ((long)mem * g.nc + q) * g.G;
        float *co = cout + ((long)mem * g.nc + q) * g.G;
        if (r < nxf) {
            const int j = (int)(r % ny) + 1, k = (int)(r / ny) + 1;
            const float gl = (vx[midx(nx, ny, 1, j, k)] >= 0.0) ? 0.f : fmaxf(0.f, ci[cidx(nx, ny, 1, j, k)]);
            const float gr = (vx[midx(nx, ny, nx, j, k)] <= 0.0) ? 0.f : fmaxf(0.f, ci[cidx(nx, ny, nx, j, k)]);
            co[cidx(nx, ny, 0, j, k)] = gl;
            co[cidx(nx, ny, nx + 1, j, k)] = gr;
            const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
            const double area = faty * g.dy * 0.5 * (dz[k] + dz[k - 1]);
            acc += -(dt * area * vx[midx(nx, ny, 1, j, k)] * (double)gl);
            acc += dt * area * vx[midx(nx, ny, nx, j, k)] * (double)gr;
            continue;
        }
        r -= nxf;
        if (r < nyf) {
            const int i = (int)(r % nx) + 1, k = (int)(r / nx) + 1;
            const float gd = (vy[midx(nx, ny, i, 1, k)] >= 0.0) ? 0.f : fmaxf(0.f, ci[cidx(nx, ny, i, 1, k)]);
            const float gu = (vy[midx(nx, ny, i, ny, k)] <= 0.0) ? 0.f : fmaxf(0.f, ci[cidx(nx, ny, i, ny, k)]);
            co[cidx(nx, ny, i, 0, k)] = gd;
            co[cidx(nx, ny, i, ny + 1, k)] = gu;
            const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
            const double area = fatx * g.dx * 0.5 * (dz[k] + dz[k - 1]);
            acc += -(dt * area * vy[midx(nx, ny, i, 1, k)] * (double)gd);
            acc += dt * area * vy[midx(nx, ny, i, ny, k)] * (double)gu;
            continue;
        }
        r -= nyf;
        if (r < nxy) {
            const int i = (int)(r % nx) + 1, j = (int)(r / nx) + 1;
            const float gb = fmaxf(0.f, ci[cidx(nx, ny, i, j, 1)] * a.bfac[(long)q * nxy + r]);
            const float gt = fmaxf(0.f, a.tfac[(long)q * nxy + r] * ci[cidx(nx, ny, i, j, nz)]);
            co[cidx(nx, ny, i, j, 0)] = gb;
            co[cidx(nx, ny, i, j, nz + 1)] = gt;
            const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
            const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
            const double area = fatx * faty * g.dx * g.dy;
            const double velo = vz[midx(nx, ny, i, j, nz)] + vset[(long)q * g.N + midx(nx, ny, i, j, nz)];
            acc += dt * area * velo * (double)gt;
            continue;
        }
        r -= nxy;
        // frame edges are never set by setbcc: carried over (clipped);
        // corners copy the adjacent interior corner (setbcc.f90:173-195)
        long o;
        float v;
        if (r < 4L * nz) {
            const int e = (int)(r / nz), k = (int)(r % nz) + 1;
            const int i = (e & 1) ? nx + 1 : 0, j = (e & 2) ? ny + 1 : 0;
            o = cidx(nx, ny, i, j, k);
            v = fmaxf(0.f, ci[o]);
        } else if ((r -= 4L * nz) < 4L * ny) {
            const int e = (int)(r / ny), j = (int)(r % ny) + 1;
            const int i = (e & 1) ? nx + 1 : 0, k = (e & 2) ? nz + 1 : 0;
            o = cidx(nx, ny, i, j, k);
            v = fmaxf(0.f, ci[o]);
        } else if ((r -= 4L * ny) < 4L * nx) {
            const int e = (int)(r / nx), i = (int)(r % nx) + 1;
            const int j = (e & 1) ? ny + 1 : 0, k = (e & 2) ? nz + 1 : 0;
            o = cidx(nx, ny, i, j, k);
            v = fmaxf(0.f, ci[o]);
        } else {
            r -= 4L * nx;
            const int e = (int)r;
            const int i = (e & 1) ? nx + 1 : 0, j = (e & 2) ? ny + 1 : 0, k = (e & 4) ? nz + 1 : 0;
            o = cidx(nx, ny, i, j, k);
            v = fmaxf(0.f, ci[cidx(nx, ny, i ? nx : 1, j ? ny : 1, k ? nz : 1)]);
        }
        co[o] = v;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[(long)mem * nbx + bx] = sh[0];
}

__global__ void __launch_bounds__(256) k_frame(FrameArgs a)
{
    frame_block(a, blockIdx.x, gridDim.x, blockIdx.y);  // y: member
}

// the frames of many contexts: row y of the grid is (context, member)
__global__ void __launch_bounds__(256) k_frame_batch(const FrameArgs *fa, const int2 *rows)
{
    const int2 r = rows[blockIdx.y];
    frame_block(fa[r.x], blockIdx.x, gridDim.x, r.y);
}

// outmass[member] += the sum of its partials, for many contexts (the tree
// of red::finish_rows, row by row)
__global__ void __launch_bounds__(1024) k_finish_batch(const FrameArgs *fa, const int2 *rows)
{
    const int2 r = rows[blockIdx.x];
    const FrameArgs &a = fa[r.x];
    red::block_sum_row(a.part + (long)r.y * red::kFramePartials, red::kFramePartials,
                       a.outmass + r.y, 1);
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
    const int ci = min(max(kFluxTileCols * bx - 2 + li, 1), g.nx);
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

template <int TX, int TY>
int launch_fused(hipStream_t s, const Params &p0)
{
    Params p = p0;
    p.ntx = (p.g.nx + TX - 1) / TX;
    p.nty = (p.g.ny + TY - 1) / TY;
    p.nlogical = p.ntx * p.nty * p.g.nc * p.g.nm;
    p.per_xcd = (p.nlogical + kXCD - 1) / kXCD;
    const unsigned nb = (unsigned)(p.per_xcd * kXCD);
    if (p.odd)
        hipLaunchKernelGGL((k_fused<TX, TY, true>), dim3(nb), dim3(kNT), 0, s, p);
    else
        hipLaunchKernelGGL((k_fused<TX, TY, false>), dim3(nb), dim3(kNT), 0, s, p);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

}  // namespace

// records per level of the flux kernel's tile-major layout
static size_t tiled_level(const Grid &g)
{
    return (size_t)g.ny * (size_t)((g.nx + kFluxTileCols - 1) / kFluxTileCols) * 64;
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
    f.ftx = (g.nx + kFluxTileCols - 1) / kFluxTileCols;
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

bool tile_supported(TileShape ts)
{
    if (flux_supported(ts)) return true;
    if (ts.nw != 0 || ts.mc != 0) return false;
    const int ok[][2] = {{32, 8}, {32, 16}, {64, 8}, {64, 4}, {16, 16}};
    for (auto &t : ok)
        if (t[0] == ts.tx && t[1] == ts.ty) return true;
    return false;
}

int fused_step(hipStream_t s, const StepArgs &a, const Fields &f, TileShape ts)
{
    if (a.g.nz + 1 >= 66) {
        set_error("fused step: nz=%d exceeds the supported 64 levels", a.g.nz);
        return F3D_EINVAL;
    }
    Params p{};
    p.g = a.g;
    p.cin = a.cin;
    p.cout = a.cout;
    p.cdry = a.cdry;
    p.f = f;
    p.zc = a.zc;
    p.dt64 = a.dt;
    p.dt = (float)a.dt;
    p.odd = (a.iiter % 2) != 0;
    if (flux_supported(ts)) return flux_step(s, a, f, ts);
    if (ts.tx == 32 && ts.ty == 8) return launch_fused<32, 8>(s, p);
    if (ts.tx == 32 && ts.ty == 16) return launch_fused<32, 16>(s, p);
    if (ts.tx == 64 && ts.ty == 8) return launch_fused<64, 8>(s, p);
    if (ts.tx == 16 && ts.ty == 16) return launch_fused<16, 16>(s, p);
    if (ts.tx == 64 && ts.ty == 4) return launch_fused<64, 4>(s, p);
    set_error("fused step: unsupported tile %dx%d", ts.tx, ts.ty);
    return F3D_EINVAL;
}

FrameArgs frame_args(const StepArgs &a, const Fields &f, const double *vx64, const double *vy64,
                     const double *vz64, const double *vset64, const double *dz, double *scratch,
                     double *outmass)
{
    FrameArgs fa;
    fa.g = a.g;
    fa.cin = a.cin;
    fa.cout = a.cout;
    fa.bfac = f.bfac;
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

int frame_and_outflux(hipStream_t s, const StepArgs &a, const Fields &f, const double *vx64,
                      const double *vy64, const double *vz64, const double *vset64,
                      const double *dz, double *scratch, double *outmass)
{
    const FrameArgs fa = frame_args(a, f, vx64, vy64, vz64, vset64, dz, scratch, outmass);
    hipLaunchKernelGGL(k_frame, dim3(red::kFramePartials, a.g.nm), dim3(256), 0, s, fa);
    F3D_CHECK_LAUNCH();
    return red::finish_rows(s, scratch, red::kFramePartials, a.g.nm, outmass, 1);
}

int frame_batch(hipStream_t s, const FrameArgs *fa, const int2 *rows, int nrows)
{
    if (nrows <= 0) return F3D_OK;
    hipLaunchKernelGGL(k_frame_batch, dim3(red::kFramePartials, nrows), dim3(256), 0, s, fa, rows);
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
