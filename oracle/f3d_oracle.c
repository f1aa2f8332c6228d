/*
 * FALL3D-7.0 transport core: FP64 CPU restatement (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker / CPU baseline, never by the
 * product.  Each function restates one reference routine with the same
 * operation order (left-to-right Fortran evaluation, division versus
 * reciprocal-multiply as written there), so that, compiled without
 * contraction, it reproduces the reference's FP64 results bit for bit.
 * Sign(1,x) is copysign(1,x) (flang's lowering, which honours -0).
 */
#include "f3d_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* 1-based Fortran indexing helpers */
#define CI(i, j, k) ((long)(i) + (long)(nx + 2) * ((long)(j) + (long)(ny + 2) * (long)(k)))
#define MI(i, j, k) ((long)((i) - 1) + (long)nx * ((long)((j) - 1) + (long)ny * (long)((k) - 1)))
#define PI2(i, j) ((long)((i) - 1) + (long)nx * (long)((j) - 1))

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline double fmax_f(double a, double b) { return a > b ? a : b; } /* Fortran MAX */
static inline double fmin_f(double a, double b) { return a < b ? a : b; } /* Fortran MIN */

/* minmod limiter: advctx.f90:40 */
static inline double minmod(double dr, double dl)
{
    return 0.5 * (copysign(1.0, dr) + copysign(1.0, dl)) * fmin_f(fabs(dr), fabs(dl));
}

/* setbcc.f90:19-109 (serial: iz_loc = 1, nz_loc = nz) */
void or_setbcc(double *c, int nx, int ny, int nz, const double *vx, const double *vy,
               const double *vz, const double *vset, const double *vdep,
               const double *rkver, const double *dz)
{
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            if ((vz[MI(i, j, 1)] + vset[MI(i, j, 1)]) >= 0.0)
                c[CI(i, j, 0)] = 0.0;
            else
                c[CI(i, j, 0)] = c[CI(i, j, 1)] * (1.0 + dz[0] * vdep[PI2(i, j)] / rkver[MI(i, j, 1)]);
        }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            if ((vz[MI(i, j, nz)] + vset[MI(i, j, nz)]) <= 0.0)
                c[CI(i, j, nz + 1)] = 0.0;
            else
                c[CI(i, j, nz + 1)] = c[CI(i, j, nz)];
        }
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) {
            c[CI(0, j, k)] = (vx[MI(1, j, k)] >= 0.0) ? 0.0 : c[CI(1, j, k)];
            c[CI(nx + 1, j, k)] = (vx[MI(nx, j, k)] <= 0.0) ? 0.0 : c[CI(nx, j, k)];
        }
    for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) {
            c[CI(i, 0, k)] = (vy[MI(i, 1, k)] >= 0.0) ? 0.0 : c[CI(i, 1, k)];
            c[CI(i, ny + 1, k)] = (vy[MI(i, ny, k)] <= 0.0) ? 0.0 : c[CI(i, ny, k)];
        }
    c[CI(0, 0, 0)] = c[CI(1, 1, 1)];
    c[CI(nx + 1, 0, 0)] = c[CI(nx, 1, 1)];
    c[CI(0, ny + 1, 0)] = c[CI(1, ny, 1)];
    c[CI(nx + 1, ny + 1, 0)] = c[CI(nx, ny, 1)];
    c[CI(0, 0, nz + 1)] = c[CI(1, 1, nz)];
    c[CI(nx + 1, 0, nz + 1)] = c[CI(nx, 1, nz)];
    c[CI(0, ny + 1, nz + 1)] = c[CI(1, ny, nz)];
    c[CI(nx + 1, ny + 1, nz + 1)] = c[CI(nx, ny, nz)];
}

/* advctzc.f90:28-81: both passes read the pre-sweep slab (work holds the fluxes) */
void or_advctzc(double *c, int nx, int ny, int nz, const double *vz, const double *vset,
                const double *dz, double dt)
{
    const long nxy = (long)nx * ny;
    double *w1 = (double *)malloc(sizeof(double) * nxy * (nz + 1));
    double *w2 = (double *)malloc(sizeof(double) * nxy * (nz + 1));
#define W(a, i, j, k) a[PI2(i, j) + nxy * ((k) - 1)]
    for (int k = 1; k <= nz; ++k) {
        const double dzz = 2.0 / (dz[k] + dz[k - 1]);
        const double dzk = 1.0 / dz[k];
        const double dzk_1 = 1.0 / dz[k - 1];
        const int kmax = imax(1, k - 1);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double vk = vz[MI(i, j, k)] + vset[MI(i, j, k)];
                const double vk_1 = vz[MI(i, j, kmax)] + vset[MI(i, j, kmax)];
                const double vk_12 = 0.5 * (vk + vk_1);
                W(w1, i, j, k) = dt * (vk * c[CI(i, j, k)] - vk_1 * c[CI(i, j, k - 1)]) * dzz;
                const double dr = (c[CI(i, j, k + 1)] - c[CI(i, j, k)]) * dzk;
                const double dl = (c[CI(i, j, k)] - c[CI(i, j, k - 1)]) * dzk_1;
                const double sigma = minmod(dr, dl);
                const double Crk = fabs(vk_12) * dt * dzk_1;
                W(w2, i, j, k) = dt * 0.5 * dz[k - 1] * (1.0 - Crk) * sigma * fabs(vk_12) * dzz;
            }
    }
    const double dznz = 1.0 / dz[nz];
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const double vk = vz[MI(i, j, nz)] + vset[MI(i, j, nz)];
            W(w1, i, j, nz + 1) = dt * vk * (c[CI(i, j, nz + 1)] - c[CI(i, j, nz)]) * dznz;
        }
    for (int k = 1; k <= nz; ++k) {
        const int kmax = imax(1, k - 1), kmin = imin(nz, k + 1);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double vk = vz[MI(i, j, k)] + vset[MI(i, j, k)];
                double *cc = &c[CI(i, j, k)];
                if (vk > 0.0)
                    *cc = *cc - W(w1, i, j, k) - (W(w2, i, j, k) - W(w2, i, j, kmax));
                else
                    *cc = *cc - W(w1, i, j, k + 1) - (W(w2, i, j, kmin) - W(w2, i, j, k));
            }
    }
#undef W
    free(w1);
    free(w2);
}

/* advctx.f90:24-56: per (j,k) line, flux pass then update pass.  The
 * closing flux W1(nx+1) uses the last loop value vi = vx(nx) (:46). */
void or_advctx(double *c, int nx, int ny, int nz, const double *vx, double dx, double dt)
{
    double *w1 = (double *)malloc(sizeof(double) * (nx + 2));
    double *w2 = (double *)malloc(sizeof(double) * (nx + 2));
    const double auxv = dt / dx;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) {
            double vi = 0.0;
            for (int i = 1; i <= nx; ++i) {
                vi = vx[MI(i, j, k)];
                const double vi_1 = vx[MI(imax(1, i - 1), j, k)];
                const double vi_12 = 0.5 * (vi + vi_1);
                w1[i] = auxv * (vi * c[CI(i, j, k)] - vi_1 * c[CI(i - 1, j, k)]);
                const double dr = (c[CI(i + 1, j, k)] - c[CI(i, j, k)]) / dx;
                const double dl = (c[CI(i, j, k)] - c[CI(i - 1, j, k)]) / dx;
                const double sigma = minmod(dr, dl);
                const double Cri = auxv * fabs(vi_12);
                w2[i] = dt * 0.5 * (1.0 - Cri) * sigma * fabs(vi_12);
            }
            w1[nx + 1] = auxv * vi * (c[CI(nx + 1, j, k)] - c[CI(nx, j, k)]);
            for (int i = 1; i <= nx; ++i) {
                double *cc = &c[CI(i, j, k)];
                if (vx[MI(i, j, k)] > 0.0)
                    *cc = *cc - w1[i] - (w2[i] - w2[imax(1, i - 1)]);
                else
                    *cc = *cc - w1[i + 1] - (w2[imin(nx, i + 1)] - w2[i]);
            }
        }
    free(w1);
    free(w2);
}

/* advcty.f90:24-55: as advctx along j (loop k -> i -> j) */
void or_advcty(double *c, int nx, int ny, int nz, const double *vy, double dy, double dt)
{
    double *w1 = (double *)malloc(sizeof(double) * (ny + 2));
    double *w2 = (double *)malloc(sizeof(double) * (ny + 2));
    const double auxv = dt / dy;
    for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) {
            double vj = 0.0;
            for (int j = 1; j <= ny; ++j) {
                vj = vy[MI(i, j, k)];
                const double vj_1 = vy[MI(i, imax(1, j - 1), k)];
                const double vj_12 = 0.5 * (vj + vj_1);
                w1[j] = auxv * (vj * c[CI(i, j, k)] - vj_1 * c[CI(i, j - 1, k)]);
                const double dr = (c[CI(i, j + 1, k)] - c[CI(i, j, k)]) / dy;
                const double dl = (c[CI(i, j, k)] - c[CI(i, j - 1, k)]) / dy;
                const double sigma = minmod(dr, dl);
                const double Crj = auxv * fabs(vj_12);
                w2[j] = dt * 0.5 * (1.0 - Crj) * sigma * fabs(vj_12);
            }
            w1[ny + 1] = auxv * vj * (c[CI(i, ny + 1, k)] - c[CI(i, ny, k)]);
            for (int j = 1; j <= ny; ++j) {
                double *cc = &c[CI(i, j, k)];
                if (vy[MI(i, j, k)] > 0.0)
                    *cc = *cc - w1[j] - (w2[j] - w2[imax(1, j - 1)]);
                else
                    *cc = *cc - w1[j + 1] - (w2[imin(ny, j + 1)] - w2[j]);
            }
        }
    free(w1);
    free(w2);
}

/* diffz.f90:19-62: c/rho by division; top flux uses l_dz(nz) (:57) */
void or_diffz(double *c, int nx, int ny, int nz, const double *dz, double dt,
              const double *rkver, const double *rho)
{
    double *l_dz = (double *)malloc(sizeof(double) * (nz + 1));
    double *l_dzdz = (double *)malloc(sizeof(double) * (nz + 1));
    double *work = (double *)malloc(sizeof(double) * (nz + 1));
    for (int k = 0; k <= nz; ++k) l_dz[k] = 1.0 / dz[k];
    for (int k = 1; k <= nz; ++k) l_dzdz[k] = 2.0 / (dz[k] + dz[k - 1]);
#define R(k) rho[MI(i, j, k)]
#define K(k) rkver[MI(i, j, k)]
#define C(k) c[CI(i, j, k)]
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            double rkst = K(2) * R(2);
            double rksb = K(1) * R(1);
            double gradt = rkst * ((C(2) / R(2)) - (C(1) / R(1))) * l_dz[1];
            double gradb = rksb * ((C(1) / R(1)) - (C(0) / R(1))) * l_dz[0];
            gradb = fmax_f(0.0, gradb);
            work[1] = dt * (gradt - gradb) * l_dzdz[1];
            for (int k = 2; k <= nz - 1; ++k) {
                rkst = 0.5 * (R(k + 1) * K(k + 1) + R(k) * K(k));
                rksb = 0.5 * (R(k) * K(k) + R(k - 1) * K(k - 1));
                gradt = rkst * ((C(k + 1) / R(k + 1)) - (C(k) / R(k))) * l_dz[k];
                gradb = rksb * ((C(k) / R(k)) - (C(k - 1) / R(k - 1))) * l_dz[k - 1];
                work[k] = dt * (gradt - gradb) * l_dzdz[k];
            }
            rkst = K(nz) * R(nz);
            rksb = K(nz - 1) * R(nz - 1);
            gradt = rkst * ((C(nz + 1) / R(nz)) - (C(nz) / R(nz))) * l_dz[nz];
            gradb = rksb * ((C(nz) / R(nz)) - (C(nz - 1) / R(nz - 1))) * l_dz[nz - 1];
            gradt = fmin_f(0.0, gradt);
            work[nz] = dt * (gradt - gradb) * l_dz[nz];
            for (int k = 1; k <= nz; ++k) C(k) = C(k) + work[k];
        }
#undef R
#undef K
#undef C
    free(l_dz);
    free(l_dzdz);
    free(work);
}

/* diffx.f90:17-53: reciprocal density l_rho, un-averaged K*rho at the ends */
void or_diffx(double *c, int nx, int ny, int nz, double dx, double dt, const double *rkh,
              const double *rho)
{
    double *l_rho = (double *)malloc(sizeof(double) * (nx + 1));
    double *work = (double *)malloc(sizeof(double) * (nx + 1));
    const double dtdx2 = dt / (dx * dx);
#define R(i) rho[MI(i, j, k)]
#define K(i) rkh[MI(i, j, k)]
#define C(i) c[CI(i, j, k)]
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) {
            for (int i = 1; i <= nx; ++i) l_rho[i] = 1.0 / R(i);
            double rksr = K(2) * R(2);
            double rksl = K(1) * R(1);
            double gradr = rksr * (C(2) * l_rho[2] - C(1) * l_rho[1]);
            double gradl = rksl * (C(1) * l_rho[1] - C(0) * l_rho[1]);
            gradl = fmax_f(0.0, gradl);
            work[1] = dtdx2 * (gradr - gradl);
            for (int i = 2; i <= nx - 1; ++i) {
                rksr = 0.5 * (R(i + 1) * K(i + 1) + R(i) * K(i));
                rksl = 0.5 * (R(i) * K(i) + R(i - 1) * K(i - 1));
                gradr = rksr * (C(i + 1) * l_rho[i + 1] - C(i) * l_rho[i]);
                gradl = rksl * (C(i) * l_rho[i] - C(i - 1) * l_rho[i - 1]);
                work[i] = dtdx2 * (gradr - gradl);
            }
            rksr = K(nx) * R(nx);
            rksl = K(nx - 1) * R(nx - 1);
            gradr = rksr * (C(nx + 1) * l_rho[nx] - C(nx) * l_rho[nx]);
            gradl = rksl * (C(nx) * l_rho[nx] - C(nx - 1) * l_rho[nx - 1]);
            gradr = fmin_f(0.0, gradr);
            work[nx] = dtdx2 * (gradr - gradl);
            for (int i = 1; i <= nx; ++i) C(i) = C(i) + work[i];
        }
#undef R
#undef K
#undef C
    free(l_rho);
    free(work);
}

/* diffy.f90:17-53 */
void or_diffy(double *c, int nx, int ny, int nz, double dy, double dt, const double *rkh,
              const double *rho)
{
    double *l_rho = (double *)malloc(sizeof(double) * (ny + 1));
    double *work = (double *)malloc(sizeof(double) * (ny + 1));
    const double dtdy2 = dt / (dy * dy);
#define R(j) rho[MI(i, j, k)]
#define K(j) rkh[MI(i, j, k)]
#define C(j) c[CI(i, j, k)]
    for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) {
            for (int j = 1; j <= ny; ++j) l_rho[j] = 1.0 / R(j);
            double rksu = K(2) * R(2);
            double rksd = K(1) * R(1);
            double gradu = rksu * (C(2) * l_rho[2] - C(1) * l_rho[1]);
            double gradd = rksd * (C(1) * l_rho[1] - C(0) * l_rho[1]);
            gradd = fmax_f(0.0, gradd);
            work[1] = dtdy2 * (gradu - gradd);
            for (int j = 2; j <= ny - 1; ++j) {
                rksu = 0.5 * (R(j + 1) * K(j + 1) + R(j) * K(j));
                rksd = 0.5 * (R(j) * K(j) + R(j - 1) * K(j - 1));
                gradu = rksu * (C(j + 1) * l_rho[j + 1] - C(j) * l_rho[j]);
                gradd = rksd * (C(j) * l_rho[j] - C(j - 1) * l_rho[j - 1]);
                work[j] = dtdy2 * (gradu - gradd);
            }
            rksu = K(ny) * R(ny);
            rksd = K(ny - 1) * R(ny - 1);
            gradu = rksu * (C(ny + 1) * l_rho[ny] - C(ny) * l_rho[ny]);
            gradd = rksd * (C(ny) * l_rho[ny] - C(ny - 1) * l_rho[ny - 1]);
            gradu = fmin_f(0.0, gradu);
            work[ny] = dtdy2 * (gradu - gradd);
            for (int j = 1; j <= ny; ++j) C(j) = C(j) + work[j];
        }
#undef R
#undef K
#undef C
    free(l_rho);
    free(work);
}

/* accum.f90:17-25 (ic is 1-based) */
void or_accum(const double *c, double *cdry, int ic, int nx, int ny, int nz, double dt,
              const double *vset, const double *vdep)
{
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            double *cd = &cdry[PI2(i, j) + (long)nx * ny * ic];
            *cd = *cd - dt * c[CI(i, j, 1)] * fmin_f(0.0, vset[MI(i, j, 1)] + vdep[PI2(i, j)]);
        }
}

/* sizloop.f90:77-78 */
void or_clip(double *c, long n)
{
    for (long q = 0; q < n; ++q) c[q] = fmax_f(0.0, c[q]);
}

/* cmass2d.f90:38-105, serial summation order */
void or_cmass2d(const double *c, const double *vx, const double *vy, const double *vz,
                const double *vset, int nx, int ny, int nz, int nc, double dx, double dy,
                const double *dz, double dt, double *outmass)
{
    const long G = (long)(nx + 2) * (ny + 2) * (nz + 2);
    const long N = (long)nx * ny * nz;
    double m = 0.0;
    for (int ic = 0; ic < nc; ++ic)
        for (int j = 1; j <= ny; ++j) {
            const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
            for (int k = 1; k <= nz; ++k) {
                const double area = faty * dy * 0.5 * (dz[k] + dz[k - 1]);
                m = m - dt * area * vx[MI(1, j, k)] * c[G * ic + CI(0, j, k)];
                m = m + dt * area * vx[MI(nx, j, k)] * c[G * ic + CI(nx + 1, j, k)];
            }
        }
    for (int ic = 0; ic < nc; ++ic)
        for (int i = 1; i <= nx; ++i) {
            const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
            for (int k = 1; k <= nz; ++k) {
                const double area = fatx * dx * 0.5 * (dz[k] + dz[k - 1]);
                m = m - dt * area * vy[MI(i, 1, k)] * c[G * ic + CI(i, 0, k)];
                m = m + dt * area * vy[MI(i, ny, k)] * c[G * ic + CI(i, ny + 1, k)];
            }
        }
    for (int ic = 0; ic < nc; ++ic)
        for (int j = 1; j <= ny; ++j) {
            const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
            for (int i = 1; i <= nx; ++i) {
                const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
                const double area = fatx * faty * dx * dy;
                const double velo = vz[MI(i, j, nz)] + vset[N * ic + MI(i, j, nz)];
                m = m + dt * area * velo * c[G * ic + CI(i, j, nz + 1)];
            }
        }
    *outmass = *outmass + m;
}

/* cmass3d.f90:24-58, summation order i -> j -> ic -> k */
void or_cmass3d(const double *c, const double *cload, const double *cdry, const double *cwet,
                int nx, int ny, int nz, int nc, double dx, double dy, const double *dz,
                double *massvol, double *massdep, double *massdry, double *masswet)
{
    const long G = (long)(nx + 2) * (ny + 2) * (nz + 2);
    double mv = 0.0, md = 0.0, mdr = 0.0, mw = 0.0;
    for (int i = 1; i <= nx; ++i) {
        const double fatx = (i == 1 || i == nx) ? 0.5 : 1.0;
        for (int j = 1; j <= ny; ++j) {
            const double faty = (j == 1 || j == ny) ? 0.5 : 1.0;
            for (int ic = 0; ic < nc; ++ic) {
                double zmass = 0.0;
                for (int k = 1; k <= nz - 1; ++k)
                    zmass = zmass + 0.5 * dz[k] * (c[G * ic + CI(i, j, k)] + c[G * ic + CI(i, j, k + 1)]);
                mv = mv + fatx * faty * zmass;
            }
            md = md + fatx * faty * cload[PI2(i, j)];
            mdr = mdr + fatx * faty * cdry[PI2(i, j)];
            mw = mw + fatx * faty * cwet[PI2(i, j)];
        }
    }
    *massvol = mv * dx * dy;
    *massdep = md * dx * dy;
    *massdry = mdr * dx * dy;
    *masswet = mw * dx * dy;
}

/* getc0, endstep.f90:80-106 */
void or_getc0(double *cdry, double *cwet, int nx, int ny, int np)
{
    const long nxy = (long)nx * ny;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            double d = 0.0, w = 0.0;
            for (int ic = 1; ic <= np; ++ic) {
                d = d + cdry[PI2(i, j) + nxy * ic];
                w = w + cwet[PI2(i, j) + nxy * ic];
            }
            cdry[PI2(i, j)] = d;
            cwet[PI2(i, j)] = w;
        }
}

/* finddt.f90:26-61 (x**2.0 is x*x: pow(x,2) is exact) */
double or_finddt(int nx, int ny, int nz, int nc, const double *dz, double dX1, double dX2,
                 const double *vx, const double *vy, const double *vz, const double *rkh1,
                 const double *rkhor, const double *rkver, const double *vset, double *dtc)
{
    const long N = (long)nx * ny * nz;
    double dtcn = 0.0;
    for (int ic = 0; ic < nc; ++ic) {
        const double *vs = vset + N * ic;
        for (int k = 1; k <= nz; ++k) {
            const double ddz = fmin_f(dz[k - 1], dz[k]);
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double dvsetdz;
                    if (k == 1)
                        dvsetdz = (vs[MI(i, j, 2)] - vs[MI(i, j, 1)]) / (dz[1]);
                    else if (k == nz)
                        dvsetdz = (vs[MI(i, j, nz)] - vs[MI(i, j, nz - 1)]) / (dz[nz - 1]);
                    else
                        dvsetdz = (vs[MI(i, j, k + 1)] - vs[MI(i, j, k - 1)]) / (dz[k] + dz[k - 1]);
                    const double r = (2.0 * rkh1[MI(i, j, k)] / (dX1 * dX1)) + (fabs(vx[MI(i, j, k)]) / dX1)
                                   + (2.0 * rkhor[MI(i, j, k)] / (dX2 * dX2)) + (fabs(vy[MI(i, j, k)]) / dX2)
                                   + (2.0 * rkver[MI(i, j, k)] / (ddz * ddz))
                                   + (fabs((vz[MI(i, j, k)] + vs[MI(i, j, k)])) / ddz)
                                   + (fabs(dvsetdz));
                    dtcn = fmax_f(dtcn, r);
                }
        }
    }
    dtcn = fmax_f(dtcn, 1e-13);
    *dtc = 1.0 / dtcn;
    return 0.9 * (*dtc);
}

/* kappa3.f90:50-236.  SIMILARITY leaves level 1 untouched and reuses the
 * last rich/dvdz2 (from (nx,ny,nz-1)) at level nz, as the loop there does. */
void or_kappa3(int modkv, int modkh, int nx, int ny, int nz, const double *dz,
               const double *zlayer, const double *vx, const double *vy, const double *vptemp,
               const double *zi, const double *rl, const double *ustar, const double *Hm1,
               double dX1, double dX2, double dx, double dy, double rkv0, double rkh0,
               double Csh, int lonlat, double *rkver, double *rkhor)
{
    const double vkarm = 0.4, beta1 = 9.2, beta2 = 13.0, gi = 9.81, alfa = 0.28, rlamb = 30.0;
    const double rkhmin = 1e2, rkvmin = 1.0, hinf = 5.0, hsup = 5e3, Khf = 8000.0;
    if (modkv == 0) {
        for (long q = 0; q < (long)nx * ny * nz; ++q) rkver[q] = rkv0;
    } else if (modkv == 1) {
        double rich = 0.0, dvdz2 = 0.0;
        for (int k = 2; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    if (k < nz) {
                        const double a = (vx[MI(i, j, k + 1)] - vx[MI(i, j, k - 1)]) / (dz[k] + dz[k - 1]);
                        const double b = (vy[MI(i, j, k + 1)] - vy[MI(i, j, k - 1)]) / (dz[k] + dz[k - 1]);
                        dvdz2 = a * a + b * b;
                        dvdz2 = fmax_f(dvdz2, 1e-10);
                        rich = gi * ((vptemp[MI(i, j, k + 1)] - vptemp[MI(i, j, k - 1)]) / (dz[k] + dz[k - 1]))
                             / (vptemp[MI(i, j, k)] * dvdz2);
                    }
                    const double h = fmin_f(fmax_f(zi[PI2(i, j)], hinf), hsup);
                    const double zl = zlayer[k - 1];
                    const double zeta = zl / h;
                    double r;
                    if (zeta < 1.0) {
                        if (rl[PI2(i, j)] >= 0.0)
                            r = vkarm * ustar[PI2(i, j)] * zl * (1.0 - zeta) / (1.0 + beta1 * zl * rl[PI2(i, j)]);
                        else
                            r = vkarm * ustar[PI2(i, j)] * zl * (1.0 - zeta) * sqrt(1.0 - beta2 * zl * rl[PI2(i, j)]);
                    } else {
                        const double lc = 1.0 / ((1.0 / (vkarm * zl)) + (1.0 / rlamb));
                        if (rich >= 0.0) {
                            r = lc * lc * sqrt(dvdz2) / (1.0 + 10.0 * rich * (1.0 + 8.0 * rich));
                            r = fmax_f(r, rkvmin);
                        } else {
                            r = lc * lc * sqrt(dvdz2) * sqrt(1.0 - 18.0 * rich);
                            r = fmax_f(r, rkvmin);
                        }
                    }
                    rkver[MI(i, j, k)] = r;
                }
    } else if (modkv == 2) {
        const double gh = 11.6, betah = 7.8, prt = 0.95;
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    const double zl = zlayer[k - 1];
                    double zeta = zl * rl[PI2(i, j)];
                    double r;
                    if (zeta < 0.0)
                        r = vkarm * ustar[PI2(i, j)] * zl * sqrt(1.0 - gh * zeta) / prt;
                    else if (zeta > 0.0)
                        r = vkarm * ustar[PI2(i, j)] * zl / (prt + betah * zeta);
                    else
                        r = vkarm * ustar[PI2(i, j)] * zl / prt;
                    const double h = fmin_f(fmax_f(zi[PI2(i, j)], hinf), hsup);
                    zeta = zl / h;
                    if (zeta > 1.0) zeta = 1.0;
                    rkver[MI(i, j, k)] = r * (1.0 - zeta);
                }
    }

    if (modkh == 0) {
        for (long q = 0; q < (long)nx * ny * nz; ++q) rkhor[q] = rkh0;
    } else if (modkh == 1 || modkh == 2) {
        const double Khn0 = lonlat ? 3.6e-2 : 4000.0;
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    const double dXX = Hm1[PI2(i, j)] * dX1;
                    double dvxdx, dvydx, dvxdy, dvydy;
                    if (i == 1) {
                        dvxdx = (vx[MI(2, j, k)] - vx[MI(1, j, k)]) / dXX;
                        dvydx = (vy[MI(2, j, k)] - vy[MI(1, j, k)]) / dXX;
                    } else if (i == nx) {
                        dvxdx = (vx[MI(nx, j, k)] - vx[MI(nx - 1, j, k)]) / dXX;
                        dvydx = (vy[MI(nx, j, k)] - vy[MI(nx - 1, j, k)]) / dXX;
                    } else {
                        dvxdx = (vx[MI(i + 1, j, k)] - vx[MI(i - 1, j, k)]) / (2.0 * dXX);
                        dvydx = (vy[MI(i + 1, j, k)] - vy[MI(i - 1, j, k)]) / (2.0 * dXX);
                    }
                    if (j == 1) {
                        dvxdy = (vx[MI(i, 2, k)] - vx[MI(i, 1, k)]) / dX2;
                        dvydy = (vy[MI(i, 2, k)] - vy[MI(i, 1, k)]) / dX2;
                    } else if (j == ny) {
                        dvxdy = (vx[MI(i, ny, k)] - vx[MI(i, ny - 1, k)]) / dX2;
                        dvydy = (vy[MI(i, ny, k)] - vy[MI(i, ny - 1, k)]) / dX2;
                    } else {
                        dvxdy = (vx[MI(i, j + 1, k)] - vx[MI(i, j - 1, k)]) / (2.0 * dX2);
                        dvydy = (vy[MI(i, j + 1, k)] - vy[MI(i, j - 1, k)]) / (2.0 * dX2);
                    }
                    double r;
                    if (modkh == 1) {
                        const double Kmh = 0.075 * pow(dXX * dX2, 2.0 / 3.0);
                        const double s = dvxdy + dvydx;
                        r = Csh * Csh * dXX * dX2 * sqrt(s * s + 2.0 * (dvxdx * dvxdx + dvydy * dvydy));
                        r = fmax_f(r, Kmh);
                        r = fmax_f(r, rkhmin);
                    } else {
                        const double a = dvxdx - dvydy, b = dvydx + dvxdy;
                        const double Kht = alfa * alfa * dXX * dX2 * sqrt(a * a + b * b);
                        const double Khn = Khf * (Khn0 / Hm1[PI2(i, j)] / dx) * (Khn0 / dy);
                        r = (1.0 / Kht) + (1.0 / Khn);
                        r = fmax_f(1.0 / r, rkhmin);
                    }
                    rkhor[MI(i, j, k)] = r;
                }
    }
}

/* get_gama, setpsi.f90:65-127: bisection on the prolate-ellipsoid area */
int or_get_gama(double diam, double sphe, double *gama_out)
{
    const double d = diam * 1e3;
    const double toler = 1e-8;
    double gmin = 1e-3, gmax = 1.0, gama = 0.0;
    const double pi = 4.0 * atan(1.0);
    const double hd = 0.5 * d;
    const double Vp = 4.0 * pi * (hd * hd * hd) / 3.0;
    const double Ap = (pow(pi, 1.0 / 3.0)) * (pow(6.0 * Vp, 2.0 / 3.0)) / sphe;
    for (int it = 1; it <= 1000; ++it) {
        gama = 0.5 * (gmin + gmax);
        const double e = acos(gama);
        const double Ao = 0.5 * pi * d * d * (pow(gama, -4.0 / 3.0)) * (gama * gama + (e / tan(e)));
        if (Ao < Ap)
            gmax = gama;
        else
            gmin = gama;
        if ((it > 1) && (fabs(Ao - Ap) < toler)) {
            *gama_out = gama;
            return 0;
        }
    }
    *gama_out = gama;
    return -1;
}

/* setpsi, setpsi.f90:1-60 */
void or_setpsi(double *psi, const double *sphe, const double *diam, int modv, int nc)
{
    for (int ic = 0; ic < nc; ++ic) {
        double gama;
        if (modv == 1) {
            psi[ic] = 1.0;
        } else if (modv == 2) {
            psi[ic] = sphe[ic];
        } else if (modv == 3) {
            or_get_gama(diam[ic], sphe[ic], &gama);
            psi[ic] = (gama >= 1.0) ? 0.5 * (1.0 + 1.0 / gama) : gama;
        } else if (modv == 4) {
            or_get_gama(diam[ic], sphe[ic], &gama);
            psi[ic] = (gama >= 1.0) ? sphe[ic] / 1.0 : sphe[ic] / sqrt(gama);
        }
    }
}

/* Ganser constants (setvset.f90:119-128) hoisted out of the cell loop:
 * they depend only on the class, so the bits are those of the per-cell form.
 * The exponent -2.0/3.0 is a default-REAL(4) literal in the reference. */
int or_ganser_consts(double diam, double sphe, double psi, double *rk1, double *rk2)
{
    double gama;
    const int st = or_get_gama(diam, sphe, &gama);
    const double dnd = 0.5 * (1.0 + gama) * pow(gama, (double)(-2.0f / 3.0f));
    *rk1 = 1.0 / (dnd / 3.0 + 2.0 / (3.0 * sqrt(psi)));
    *rk2 = pow(10.0, 1.8148 * pow(-log10(psi), 0.5743));
    return st;
}

/* vsettl, setvset.f90:65-181.  cd_state is the SAVEd drag coefficient
 * (:89) carried from the previous call; convergence is absolute (:93). */
static int vsettl_core(double diam, double rhop, double rhoa, double visc, double *vset,
                       int model, double psi, double rk1, double rk2, double *cd_state)
{
    const double gi = 9.81, eps = 1e-3;
    double cd = *cd_state, vold = 1e5, v = 0.0;
    int conv = 0;
    if (model == 1) {
        for (int it = 1; it <= 100; ++it) {
            v = sqrt(4.0 * gi * diam * rhop / (3.0 * cd * rhoa));
            const double rey = rhoa * v * diam / visc;
            if (rey <= 988.947)
                cd = 24.0 / rey * (1.0 + 0.15 * pow(rey, 0.687));
            else
                cd = 0.44;
            if (it > 1 && fabs(v - vold) <= eps) { conv = 1; break; }
            vold = v;
        }
    } else if (model == 2) {
        for (int it = 1; it <= 100; ++it) {
            v = sqrt(4.0 * gi * diam * rhop / (3.0 * cd * rhoa));
            const double rey = rhoa * v * diam / visc;
            cd = 24.0 / (rey * rk1) * (1.0 + 0.1118 * pow(rey * rk1 * rk2, 0.6567))
               + 0.4305 * rk2 / (1.0 + 3305.0 / (rey * rk1 * rk2));
            if (it > 1 && fabs(v - vold) <= eps) { conv = 1; break; }
            vold = v;
        }
    } else if (model == 3) {
        for (int it = 1; it <= 100; ++it) {
            v = sqrt(4.0 * gi * diam * rhop / (3.0 * cd * rhoa));
            const double rey = rhoa * v * diam / visc;
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
            if (it > 1 && fabs(v - vold) <= eps) { conv = 1; break; }
            vold = v;
        }
    } else if (model == 4) {
        v = pow((diam * diam * diam * gi * (rhop - rhoa) * rhoa * (pow(psi, 1.6))) / (visc * visc), 0.5206);
        v = (1.2065 * visc * v) / (diam * rhoa);
        *vset = v;
        return 1;
    }
    *cd_state = cd;
    if (!conv) {
        *vset = v;
        return 0;
    }
    *vset = sqrt(4.0 * gi * diam * rhop / (3.0 * cd * rhoa));
    return 1;
}

int or_vsettl(double diam, double rhop, double rhoa, double visc, double *vset, int model,
              double psi, double sphe, double *cd_state)
{
    double rk1 = 0.0, rk2 = 0.0;
    if (model == 2) or_ganser_consts(diam, sphe, psi, &rk1, &rk2);
    return vsettl_core(diam, rhop, rhoa, visc, vset, model, psi, rk1, rk2, cd_state);
}

/* setvset.f90:25-60: chain order ic -> i -> j -> k, then vset = -vset.
 * Tair0 is the default-REAL(4) literal 291.15 of the reference (:23). */
long or_setvset(int nx, int ny, int nz, int np, const double *tempe, const double *rho,
                const double *diam, const double *rhop, const double *psi, const double *sphe,
                int modv, double *vset, double *cd_state)
{
    const double visc0 = 1.827e-5;
    const double Tair0 = (double)291.15f;
    const long N = (long)nx * ny * nz;
    long bad = 0;
    for (int ic = 0; ic < np && !bad; ++ic) {
        double rk1 = 0.0, rk2 = 0.0;
        if (modv == 2) or_ganser_consts(diam[ic], sphe[ic], psi[ic], &rk1, &rk2);
        for (int i = 1; i <= nx && !bad; ++i)
            for (int j = 1; j <= ny && !bad; ++j)
                for (int k = 1; k <= nz; ++k) {
                    const double T = tempe[MI(i, j, k)];
                    const double visc = visc0 * ((Tair0 + 120.0) / (T + 120.0)) * (pow(T / Tair0, 1.5));
                    double v;
                    if (!vsettl_core(diam[ic], rhop[ic], rho[MI(i, j, k)], visc, &v, modv, psi[ic],
                                     rk1, rk2, cd_state)) {
                        bad = -(1 + N * ic + MI(i, j, k));
                        break;
                    }
                    vset[N * ic + MI(i, j, k)] = v;
                }
    }
    if (bad) return bad;
    for (long q = 0; q < N * np; ++q) vset[q] = -vset[q];
    return 0;
}

/* meteo.f90:108-112 */
void or_meteo_scale(int nx, int ny, int nz, const double *Hm1, double *vx, double *rho,
                    const double *rkhor, double *rkh1)
{
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double h = Hm1[PI2(i, j)];
                vx[MI(i, j, k)] = vx[MI(i, j, k)] / h;
                rho[MI(i, j, k)] = rho[MI(i, j, k)] * h;
                rkh1[MI(i, j, k)] = rkhor[MI(i, j, k)] / (h * h);
            }
}

/* setvdrydep.f90:54-264: dry deposition velocity after Feng (JGR 113,
 * 2008) for classes up to 100 um.  Mode by diameter (:134-152), the USGS
 * land-use class folded onto Feng's six categories (:167-216), the
 * aerodynamic resistance from the log profile with the stable / unstable
 * corrections at the first level (:218-237), Sutherland viscosity with the
 * FP64 reference temperature (:240-241), Re* clamped at 40300 (:245-246),
 * the first term (:250-251), and per class the resistances in series
 * (:256-260).  Classes coarser than 100 um keep their vdep.  tempe1 and
 * rho1 are the first level of tempe and rho (before the Hm1 scaling).
 * The coefficients are Feng's table as the reference sets them (:76-123). */
static const double kFengA[6][4] = {
    {0.0048, 0.0315, 1.2891, 1.0338}, {0.0037, 0.0120, 1.3977, 1.0707},
    {0.0042, 0.2928, 1.3970, 0.9155}, {0.0032, 0.1201, 1.1838, 1.0096},
    {0.0043, 0.1337, 1.2834, 1.1595}, {0.0045, 0.0925, 1.2654, 1.0891}};
static const double kFengB[6][4] = {
    {1.0, 2.7925, 2.6878, 1.2644}, {1.0, 2.2413, 2.5838, 1.3247},
    {1.0, 3.8581, 2.5580, 1.0364}, {1.0, 3.4407, 2.8033, 1.2069},
    {1.0, 3.5456, 2.7157, 1.4863}, {1.0, 3.2920, 2.7227, 1.4240}};

static int feng_mode(double d)
{
    if (d > 100e-6) return 0;
    if (d > 10e-6) return 4;
    if (d > 2.5e-6) return 3;
    if (d > 0.1e-6) return 2;
    return 1;
}

/* USGS category (truncated like Fortran INT) -> Feng land class 1..6 */
static int feng_land(double ldu)
{
    const int u = (int)ldu;
    if (u == 1) return 1;
    if (u == 10 || u == 19) return 3;
    if (u == 16) return 5;
    if (u >= 20 && u <= 23) return 2;
    if (u == 24) return 4;
    return 6;
}

void or_setvdrydep(int nx, int ny, int nc, double zlayer1, const double *diam,
                   const double *ldu, const double *znt, const double *rl, const double *ustar,
                   const double *tempe1, const double *rho1, double *vdep)
{
    const double vkarm = 0.4, visc0 = 1.827e-5, Tair0 = 291.15;
    int any = 0;
    for (int ic = 0; ic < nc; ++ic) any |= feng_mode(diam[ic]) > 0;
    if (!any) return;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const long t = PI2(i, j);
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
                    ra = ra + log((e0 * e0 + 1) * ((e0 + 1) * (e0 + 1)) /
                                  ((e * e + 1) * ((e + 1) * (e + 1))));
                    ra = ra + 2 * (atan(e) - atan(e0));
                }
                ra = ra / (us * vkarm);
            } else {
                ra = 1e20;
            }
            const double T = tempe1[t];
            const double visc = visc0 * ((Tair0 + 120.0) / (T + 120.0)) * pow(T / Tair0, 1.5);
            double Re = us * z0 * rho1[t] / visc;
            if (Re > 40300.0) Re = 40300.0;
            const double x = (Re - 40300.0) / 15330.0;
            const double vd1 = 0.0226 * us * exp(-(0.5 * x * x));
            for (int ic = 0; ic < nc; ++ic) {
                const int m = feng_mode(diam[ic]);
                if (m == 0) continue;
                double vd = vd1 + kFengA[land - 1][m - 1] * pow(us, kFengB[land - 1][m - 1]);
                vd = ra + 1.0 / vd;
                vdep[t + (long)nx * ny * ic] = -1.0 / vd;
            }
        }
}

/* corvver.f90:16-36 (meteo.f90:116-124, LON-LAT grids only): the vertical
 * wind minus the terrain-slope projection of the horizontal wind, on the
 * scaled vx; forward differences of topg, clamped at the last column/row */
void or_corvver(int nx, int ny, int nz, const double *Hm1, const double *topg, double dX1,
                double dX2, const double *vx, const double *vy, double *vz)
{
    const double dyinv = 1.0 / dX2;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const double dxinv = 1.0 / (Hm1[PI2(i, j)] * dX1);
            const double zx = dxinv * (topg[PI2(i < nx ? i + 1 : nx, j)] - topg[PI2(i, j)]);
            const double zy = dyinv * (topg[PI2(i, j < ny ? j + 1 : ny)] - topg[PI2(i, j)]);
            for (int k = 1; k <= nz; ++k) {
                const double vs = zx * vx[MI(i, j, k)] + zy * vy[MI(i, j, k)];
                vz[MI(i, j, k)] = vz[MI(i, j, k)] - vs;
            }
        }
}

/* setsrc.f90:20-35 (1-based source indices) */
void or_setsrc(double *c, int nx, int ny, int nz, int nc, int nsou, const int *isrc,
               const int *jsrc, const int *ksrc, const double *tmrat, const double *dz,
               const double *zlayer, const double *Hm1, double dX1, double dX2, double dt)
{
    const long G = (long)(nx + 2) * (ny + 2) * (nz + 2);
    for (int is = 0; is < nsou; ++is) {
        const int i = isrc[is], j = jsrc[is], k = ksrc[is];
        if (k >= 1 && k <= nz) {
            const double dzeta = (k == 1) ? 0.5 * dz[k] + zlayer[k - 1] : 0.5 * (dz[k - 1] + dz[k]);
            const double vol0 = Hm1[PI2(i, j)] * dX1 * dX2 * dzeta;
            for (int ic = 0; ic < nc; ++ic)
                c[G * ic + CI(i, j, k)] = c[G * ic + CI(i, j, k)] + dt * tmrat[is + (long)nsou * ic] / vol0;
        }
    }
}

/* wetdep.f90:28-50 (Jung and Shao 2006; ic 1-based, classes up to 100 um):
 * below the PBL height zi, a fraction lambda = min(dt*a*P**b, 1) of c goes
 * to cwet(:,:,ic) each step, weighted by dz(iz) */
void or_wetdep(double *c, double *cwet, int ic, int nx, int ny, int nz, double dt,
               const double *diam, const double *zlayer, const double *dz, const double *zi,
               const double *prate)
{
    const double diam_max = 100.0 * 1e-6;
    if (diam[ic - 1] > diam_max) return;
    const double a = 8.4e-5, b = 0.79;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            double lambda = dt * a * pow(prate[PI2(i, j)], b);
            lambda = lambda < 1.0 ? lambda : 1.0;
            for (int k = 1; k <= nz; ++k)
                if (zlayer[k - 1] <= zi[PI2(i, j)]) {
                    double *w = cwet + (long)nx * ny * ic + PI2(i, j);
                    *w = *w + dz[k] * lambda * c[CI(i, j, k)];
                    c[CI(i, j, k)] = c[CI(i, j, k)] * (1.0 - lambda);
                }
        }
}

/* Output products of printres_nc.f90:150-241 (before the netCDF writes):
 * load(nx,ny,0:nc) = cload/Hm1 with cload = cdry + cwet (endstep.f90:35)
 * and slot 0 from the class sums (getc0, endstep.f90:80-106); the deposit
 * thickness 1d2*sum_ic(cload/rhop)/Hm1 (cm, :194-199); the z-cumulative
 * concentration 1d3*sum_{iz<nz} dz(iz)*0.5*(w(iz)+w(iz+1)), w =
 * sum_ic(c)/Hm1 (g/m2, :218-230).  Null outputs are skipped. */
void or_output_products(int nx, int ny, int nz, int nc, const double *c, const double *cdry,
                        const double *cwet, const double *Hm1, const double *rhop,
                        const double *dz, double *load, double *thick, double *colmass)
{
    const long nxy = (long)nx * ny, G = (long)(nx + 2) * (ny + 2) * (nz + 2);
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const long t = PI2(i, j);
            const double h = Hm1[t];
            if (load) {
                double d0 = 0.0, w0 = 0.0;
                for (int q = 1; q <= nc; ++q) {
                    d0 = d0 + cdry[q * nxy + t];
                    w0 = w0 + cwet[q * nxy + t];
                }
                load[t] = (d0 + w0) / h;
                for (int q = 1; q <= nc; ++q) load[q * nxy + t] = (cdry[q * nxy + t] + cwet[q * nxy + t]) / h;
            }
            if (thick) {
                double s = 0.0;
                for (int q = 1; q <= nc; ++q) s = s + (cdry[q * nxy + t] + cwet[q * nxy + t]) / rhop[q - 1];
                thick[t] = 1e2 * s / h;
            }
            if (colmass) {
                double acc = 0.0;
                for (int k = 1; k <= nz - 1; ++k) {
                    double w = 0.0, wn = 0.0;
                    for (int q = 0; q < nc; ++q) {
                        w = w + c[q * G + CI(i, j, k)];
                        wn = wn + c[q * G + CI(i, j, k + 1)];
                    }
                    acc = acc + dz[k] * 0.5 * (w / h + wn / h);
                }
                colmass[t] = 1e3 * acc;
            }
        }
}

/* ICAO standard atmosphere, three layers (setcut.f90:61-107): p (Pa) and
 * T (K) at height z (m a.s.l.). */
void or_standard_atm(double z, double *p, double *T)
{
    const double goo = 9.80665, Roo = 8.31432, Moo = 0.0289644;
    const double zoo = 0.0, Too = 288.15, poo = 101325.0, Loo = -0.0065;
    const double zo1 = 11000.0, To1 = 216.65, po1 = 22632.0;
    const double zo2 = 20000.0, To2 = 216.65, po2 = 5474.0, Lo2 = 0.001;
    if (z < zo1) {
        *T = Too + Loo * (z - zoo);
        *p = poo * pow(Too / *T, goo * Moo / (Roo * Loo));
    } else if (z < zo2) {
        *T = To1;
        *p = po1 * exp(-goo * Moo * (z - zo1) / (Roo * To1));
    } else {
        *T = To2 + Lo2 * (z - zo2);
        *p = po2 * pow(To2 / *T, goo * Moo / (Roo * Lo2));
    }
}

/* Concentration outputs of printres_nc.f90 (g/m3; all classes are
 * particles, np = nc), each from w(k) = sum_ic c(:,:,k,ic) summed in class
 * order:
 *   cgrnd (nx,ny)    C_GRND = 1d3*w(1)/Hm1 (:391-396);
 *   con (nx,ny,nz)   CON = 1d3*(w/Hm1) (:574-584), and with asl the
 *                    ter2asl interpolation to a.s.l. heights (:586,
 *                    ter2asl.f90:25-50) on topg;
 *   cfl (nx,ny,nfl)  the cut of w/Hm1 at flight level fl[l] (hundreds of
 *                    feet, fl*30.48 m) on the pressure p, times 1d3 (:545-560,
 *                    setcut.f90:24-57).
 * Null outputs are skipped. */
void or_output_concentration(int nx, int ny, int nz, int nc, const double *c, const double *Hm1,
                             const double *zlayer, const double *topg, const double *p, int nfl,
                             const double *fl, double *cgrnd, double *con, int asl, double *cfl)
{
    const long nxy = (long)nx * ny, G = (long)(nx + 2) * (ny + 2) * (nz + 2);
    double *w = (double *)malloc(sizeof(double) * (size_t)(nz + 1));
    double *col = (double *)malloc(sizeof(double) * (size_t)(nz + 1));
    double *pnom = (double *)malloc(sizeof(double) * (size_t)(nfl > 0 ? nfl : 1));
    for (int l = 0; l < nfl; ++l) {
        double T;
        or_standard_atm(fl[l] * 30.48, &pnom[l], &T);
    }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const long t = PI2(i, j);
            for (int k = 1; k <= nz; ++k) {
                double s = 0.0;
                for (int q = 0; q < nc; ++q) s = s + c[q * G + CI(i, j, k)];
                w[k] = s;
            }
            if (cgrnd) cgrnd[t] = 1e3 * w[1] / Hm1[t];
            if (con) {
                for (int k = 1; k <= nz; ++k) col[k] = 1e3 * (w[k] / Hm1[t]);
                for (int k = 1; k <= nz; ++k) {
                    double v = col[k];
                    if (asl) {
                        const double zo = fmax_f(0.0, topg[t]), zasl = zlayer[k - 1];
                        v = 0.0;
                        if (!(zasl < zo + zlayer[0]))
                            for (int izz = 1; izz + 1 <= nz; ++izz)
                                if (zasl >= zo + zlayer[izz - 1] && zasl <= zo + zlayer[izz]) {
                                    const double s = (zo + zlayer[izz] - zasl) / (zlayer[izz] - zlayer[izz - 1]);
                                    v = s * col[izz] + (1.0 - s) * col[izz + 1];
                                    break;
                                }
                    }
                    con[(long)(k - 1) * nxy + t] = v;
                }
            }
            for (int l = 0; l < nfl && cfl; ++l) {
                const double pn = pnom[l];
                double v = 0.0;
                if (!(pn > p[MI(i, j, 1)]))
                    for (int iz = 1; iz + 1 <= nz; ++iz)
                        if (pn <= p[MI(i, j, iz)] && pn > p[MI(i, j, iz + 1)]) {
                            const double s = (pn - p[MI(i, j, iz)]) / (p[MI(i, j, iz + 1)] - p[MI(i, j, iz)]);
                            v = (w[iz] / Hm1[t]) * (1.0 - s) + (w[iz + 1] / Hm1[t]) * s;
                            break;
                        }
                cfl[(long)l * nxy + t] = 1e3 * v;
            }
        }
    free(w);
    free(col);
    free(pnom);
}

/* ---- sizloop ---- */
typedef struct {
    double *c, *cdry;
    int nx, ny, nz, ic0, ic1, iiter;
    double dt, dX1, dX2;
    const double *vx, *vy, *vz, *vset, *vdep, *rkver, *rkhor, *rkh1, *rho, *dz;
    /* wet deposition (sizloop.f90:65-67): on when prate is given */
    double *cwet;
    const double *diam, *zlayer, *zi, *prate;
} class_job;

static void run_classes(const class_job *J)
{
    const int nx = J->nx, ny = J->ny, nz = J->nz;
    const long G = (long)(nx + 2) * (ny + 2) * (nz + 2);
    const long N = (long)nx * ny * nz;
    for (int ic = J->ic0; ic < J->ic1; ++ic) {
        double *cc = J->c + G * ic;
        const double *vs = J->vset + N * ic;
        const double *vd = J->vdep + (long)nx * ny * ic;
        or_setbcc(cc, nx, ny, nz, J->vx, J->vy, J->vz, vs, vd, J->rkver, J->dz);
        or_advctzc(cc, nx, ny, nz, J->vz, vs, J->dz, J->dt);
        if (J->iiter % 2 == 0) {
            or_advctx(cc, nx, ny, nz, J->vx, J->dX1, J->dt);
            or_advcty(cc, nx, ny, nz, J->vy, J->dX2, J->dt);
        } else {
            or_advcty(cc, nx, ny, nz, J->vy, J->dX2, J->dt);
            or_advctx(cc, nx, ny, nz, J->vx, J->dX1, J->dt);
        }
        or_diffz(cc, nx, ny, nz, J->dz, J->dt, J->rkver, J->rho);
        if (J->iiter % 2 == 0) {
            or_diffx(cc, nx, ny, nz, J->dX1, J->dt, J->rkh1, J->rho);
            or_diffy(cc, nx, ny, nz, J->dX2, J->dt, J->rkhor, J->rho);
        } else {
            or_diffy(cc, nx, ny, nz, J->dX2, J->dt, J->rkhor, J->rho);
            or_diffx(cc, nx, ny, nz, J->dX1, J->dt, J->rkh1, J->rho);
        }
        if (J->prate)
            or_wetdep(cc, J->cwet, ic + 1, nx, ny, nz, J->dt, J->diam, J->zlayer, J->dz, J->zi,
                      J->prate);
        or_accum(cc, J->cdry, ic + 1, nx, ny, nz, J->dt, vs, vd);
        or_clip(cc, G); /* elementwise: per-class clip == whole-array clip */
    }
}

static void *class_thread(void *arg)
{
    run_classes((const class_job *)arg);
    return NULL;
}

void or_sizloop(double *c, double *cdry, int nx, int ny, int nz, int nc, int iiter, double dt,
                const double *vx, const double *vy, const double *vz, const double *vset,
                const double *vdep, const double *rkver, const double *rkhor,
                const double *rkh1, const double *rho, const double *dz, double dX1,
                double dX2, double *outmass, int nthreads, double *cwet, const double *diam,
                const double *zlayer, const double *zi, const double *prate)
{
    class_job base = {c, cdry, nx, ny, nz, 0, nc, iiter, dt, dX1, dX2,
                      vx, vy, vz, vset, vdep, rkver, rkhor, rkh1, rho, dz,
                      cwet, diam, zlayer, zi, prate};
    if (nthreads <= 1 || nc <= 1) {
        run_classes(&base);
    } else {
        if (nthreads > nc) nthreads = nc;
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
        class_job *jobs = (class_job *)malloc(sizeof(class_job) * nthreads);
        for (int t = 0; t < nthreads; ++t) {
            jobs[t] = base;
            jobs[t].ic0 = (int)((long)nc * t / nthreads);
            jobs[t].ic1 = (int)((long)nc * (t + 1) / nthreads);
            pthread_create(&th[t], NULL, class_thread, &jobs[t]);
        }
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
        free(th);
        free(jobs);
    }
    or_cmass2d(c, vx, vy, vz, vset, nx, ny, nz, nc, dX1, dX2, dz, dt, outmass);
}

/* ---------------- gravity current (addradialwind, setdivu, divcorr) ---------------- */

/* setdivu.f90:16-30: the divergence of the (scaled) wind field; one-sided at
 * the first column, row and level, where the reference takes v(max(1,i-1)) */
void or_setdivu(int nx, int ny, int nz, const double *vx, const double *vy, const double *vz,
                const double *Hm1, double dX1, double dX2, const double *dz, double *divu)
{
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz; ++k) {
                const double vi = vx[MI(i, j, k)], vi_1 = vx[MI(imax(1, i - 1), j, k)];
                const double vj = vy[MI(i, j, k)], vj_1 = vy[MI(i, imax(1, j - 1), k)];
                const double vk = vz[MI(i, j, k)], vk_1 = vz[MI(i, j, imax(1, k - 1))];
                divu[MI(i, j, k)] = (vi - vi_1) / (Hm1[PI2(i, j)] * dX1) + (vj - vj_1) / dX2 +
                                    (vk - vk_1) / dz[k];
            }
}

/* addradialwind.f90:47-173: the zero-divergence radial wind of the umbrella
 * cloud added to the scaled vx, vy between the levels nearest hmin and hmax.
 * prm = {c_flow_rate, lambda_grav, k_entrain, brunt_vaisala, centre x,
 * centre y, xorigr, yorigr, dx, dy} in the grid's coordinates (dx, dy:
 * degrees on LON-LAT grids, m on UTM grids); time_eruption = time -
 * beg_time; mass_flow_rate = sum(tmrat); h_tot = maxval(zs).  Returns 0
 * when the source is off (mass flow rate <= 0.1, the single-precision
 * literal of the reference), 1 after adding the wind; out (optional) =
 * {min_radius, max_radius, radius_grav, c_grav_factor, kmin, kmax, i_wind,
 * j_wind, k_wind}. */
int or_addradialwind(int nx, int ny, int nz, double *vx, double *vy, const double *Hm1,
                     double dX1, double dX2, const double *zlayer, const double *prm,
                     double time_eruption, double mass_flow_rate, double h_tot, double *out)
{
    const double pi = 3.14159265358979323846;
    const double c_flow_rate = prm[0], lambda_grav = prm[1], k_entrain = prm[2],
                 brunt_vaisala = prm[3], cx = prm[4], cy = prm[5], xorigr = prm[6],
                 yorigr = prm[7], dx = prm[8], dy = prm[9];
    const double h11 = Hm1[PI2(1, 1)];
    const double min_radius = 1.5 * sqrt(h11 * dX1 * h11 * dX1 + dX2 * dX2);
    if (mass_flow_rate <= (double)0.1f) return 0;
    const double vol_flow_rate =
        c_flow_rate * sqrt(k_entrain) * pow(mass_flow_rate, 0.75) / pow(brunt_vaisala, 0.625);
    const double c_gravr = 3.0 * lambda_grav * brunt_vaisala * vol_flow_rate / (2.0 * pi);
    double radius_grav = pow(c_gravr * (time_eruption * time_eruption), 0.33333);
    radius_grav = fmax_f(radius_grav, min_radius);
    const double c_gravu = sqrt(2.0 * lambda_grav * brunt_vaisala * vol_flow_rate / (3.0 * pi));
    const double c_grav_factor = 0.75 * c_gravu * sqrt(radius_grav);
    const double th_grav = c_gravu / (sqrt(radius_grav) * lambda_grav * brunt_vaisala);
    const double h_umbr = h_tot / 1.32;
    const double hmin = fmax_f(h_umbr - 0.5 * th_grav, 0.5 * h_tot);
    const double hmax = fmin_f(h_umbr + 0.5 * th_grav, h_tot);
    /* integer conversion truncates toward zero; the reference clamps the
     * low side only */
    int i_wind = (int)((cx - xorigr) / dx), j_wind = (int)((cy - yorigr) / dy);
    i_wind = imax(1, i_wind);
    j_wind = imax(1, j_wind);
    int k_wind = 1;
    double store = zlayer[nz - 1];
    for (int k = 1; k <= nz; ++k) {
        const double dist = fabs(zlayer[k - 1] - h_umbr);
        if (dist <= store) {
            k_wind = k;
            store = dist;
        }
    }
    const double ux = vx[MI(i_wind, j_wind, k_wind)], uy = vy[MI(i_wind, j_wind, k_wind)];
    const double v_wind = sqrt(ux * ux + uy * uy);
    const double max_radius = 1.7777 * c_gravr / fmax_f(v_wind * v_wind, 0.1);
    int kmin = 1, kmax = nz;
    store = zlayer[nz - 1];
    for (int k = 1; k <= nz; ++k) {
        const double dist = fabs(zlayer[k - 1] - hmin);
        if (dist <= store) {
            kmin = k;
            store = dist;
        }
    }
    store = zlayer[nz - 1];
    for (int k = 1; k <= nz; ++k) {
        const double dist = fabs(zlayer[k - 1] - hmax);
        if (dist <= store) {
            kmax = k;
            store = dist;
        }
    }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const double h = Hm1[PI2(i, j)];
            const double deltax = dX1 * h * ((xorigr + i * dx) - cx) / dx;
            const double deltay = dX2 * ((yorigr + j * dy) - cy) / dy;
            const double radius = sqrt(deltax * deltax + deltay * deltay);
            if (fabs(radius) <= min_radius) continue;
            if (fabs(radius) > max_radius) continue;
            const double rvel =
                c_grav_factor / radius * (1.0 + radius * radius / (3.0 * (radius_grav * radius_grav)));
            for (int k = kmin; k <= kmax; ++k) {
                vx[MI(i, j, k)] = vx[MI(i, j, k)] + rvel * deltax / radius / h;
                vy[MI(i, j, k)] = vy[MI(i, j, k)] + rvel * deltay / radius;
            }
        }
    if (out) {
        const double o[9] = {min_radius, max_radius, radius_grav, c_grav_factor, kmin, kmax,
                             i_wind, j_wind, k_wind};
        for (int n = 0; n < 9; ++n) out[n] = o[n];
    }
    return 1;
}

/* divcorr.f90:22-38 (endstep.f90:71, at the log times after writim): the
 * relative mass imbalance (erumass - massvol - massdep - outmass)/erumass,
 * clamped to +-0.05, scales every class in the cells whose |divu*dt| >=
 * 0.005.  Returns the factor corrdiv. */
double or_divcorr(double *c, int nx, int ny, int nz, int nc, const double *divu, double dt,
                  double erumass, double massvol, double massdep, double outmass)
{
    const long G = (long)(nx + 2) * (ny + 2) * (nz + 2);
    const double compmass = massvol + massdep + outmass;
    const double eps = 0.005, eps1 = 10.0 * eps;
    const double corrdiv = fmax_f(fmin_f((erumass - compmass) / erumass, eps1), -eps1);
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double dive = divu[MI(i, j, k)];
                if (fabs(dive * dt) >= eps)
                    for (int ic = 0; ic < nc; ++ic)
                        c[G * ic + CI(i, j, k)] = c[G * ic + CI(i, j, k)] * (1.0 + corrdiv);
            }
    return corrdiv;
}
