// C ABI of libfall3d_amd: the reference-ABI per-routine drop-ins and the
// device-resident context (include/fall3d_amd.h).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "f3d_ctx.h"

namespace f3d {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

// ---------------- host-side class constants ----------------
// get_gama, setpsi.f90:65-127 (bisection on the prolate-ellipsoid area)
int get_gama(double diam, double sphe, double *gama_out)
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

// setpsi.f90:1-60 and the Ganser constants of setvset.f90:119-128; the
// latter depend only on the class, so they are formed once per class with
// the same expressions (the exponent -2.0/3.0 is a default-REAL(4) literal).
int class_constants(int modv, double diam, double rhop, double sphe, met::ClassConst *out)
{
    double gama = 1.0, psi = 1.0;
    int st = 0;
    if (modv == 1) {
        psi = 1.0;
    } else if (modv == 2) {
        psi = sphe;
    } else if (modv == 3) {
        st |= get_gama(diam, sphe, &gama);
        psi = (gama >= 1.0) ? 0.5 * (1.0 + 1.0 / gama) : gama;
    } else if (modv == 4) {
        st |= get_gama(diam, sphe, &gama);
        psi = (gama >= 1.0) ? sphe / 1.0 : sphe / sqrt(gama);
    } else {
        set_error("unknown settling velocity model %d", modv);
        return F3D_EINVAL;
    }
    out->diam = diam;
    out->rhop = rhop;
    out->psi = psi;
    out->rk1 = 0.0;
    out->rk2 = 0.0;
    if (modv == 2) {
        st |= get_gama(diam, sphe, &gama);
        const double dnd = 0.5 * (1.0 + gama) * pow(gama, (double)(-2.0f / 3.0f));
        out->rk1 = 1.0 / (dnd / 3.0 + 2.0 / (3.0 * sqrt(psi)));
        out->rk2 = pow(10.0, 1.8148 * pow(-log10(psi), 0.5743));
    }
    if (st) {
        set_error("get_gama: convergence not achieved (diam=%g, sphe=%g)", diam, sphe);
        return F3D_ECONV;
    }
    return F3D_OK;
}

Grid make_grid(int nx, int ny, int nz, int nc, double dx, double dy)
{
    Grid g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.nc = nc;
    g.G = (long)(nx + 2) * (ny + 2) * (nz + 2);
    g.N = (long)nx * ny * nz;
    g.NXY = (long)nx * ny;
    g.dx = dx;
    g.dy = dy;
    g.nm = 1;
    return g;
}

// dz(0:nz) from zlayer (reagrd.f90:216-220) and the per-level coefficients
void vertical_coefs(int nz, const double *zlayer, std::vector<double> &dz,
                    std::vector<ZCoef> &zc, std::vector<double> &ldz)
{
    dz.assign(nz + 1, 0.0);
    for (int k = 1; k <= nz - 1; ++k) dz[k] = zlayer[k] - zlayer[k - 1];
    dz[0] = dz[1];
    dz[nz] = dz[nz - 1];
    zc.assign(nz + 2, ZCoef{0, 0, 0, 0});
    ldz.assign(nz + 1, 0.0);
    for (int k = 0; k <= nz; ++k) ldz[k] = 1.0 / dz[k];
    for (int k = 1; k <= nz; ++k) {
        zc[k].dzz = 2.0 / (dz[k] + dz[k - 1]);
        zc[k].dzk = 1.0 / dz[k];
        zc[k].dzk_1 = 1.0 / dz[k - 1];
        zc[k].dzm1 = dz[k - 1];
    }
}

int DeviceBuffer::alloc(size_t n, long long *counter)
{
    release();
    if (n == 0) return F3D_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
        return F3D_ENOMEM;
    }
    bytes = n;
    cnt = counter;
    if (cnt) *cnt += (long long)n;
    return F3D_OK;
}

void DeviceBuffer::release()
{
    if (p) {
        (void)hipFree(p);
        if (cnt) *cnt -= (long long)bytes;
    }
    p = nullptr;
    bytes = 0;
    cnt = nullptr;
}

}  // namespace f3d

using namespace f3d;

// ===================== layer 1: reference-ABI drop-ins =====================

namespace {

// Scratch device copies for one drop-in call.  These entry points serve
// tests and the routine-by-routine parity mode; the performance path is the
// device-resident context.
struct Stage {
    std::vector<void *> ptrs;
    ~Stage()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T *up(const T *h, size_t n)
    {
        void *d = nullptr;
        if (hipMalloc(&d, n * sizeof(T) + 16) != hipSuccess) return nullptr;
        ptrs.push_back(d);
        if (h && hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return (T *)d;
    }
};

bool serial_decomp(const char *who, int nz, int nz_loc, int iz_loc)
{
    if (nz_loc != nz || iz_loc != 1) {
        set_error("%s: only the serial decomposition (nz_loc == nz, iz_loc == 1) is supported", who);
        return false;
    }
    return true;
}

bool dims_ok(const char *who, int nx, int ny, int nz)
{
    if (nx < 2 || ny < 2 || nz < 2) {
        set_error("%s: nx, ny, nz must be >= 2 (got %d %d %d)", who, nx, ny, nz);
        return false;
    }
    return true;
}

int copy_back(double *h, const double *d, size_t n)
{
    F3D_HIP(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost));
    return F3D_OK;
}

#define DROPIN_GUARD(name)                                                                   \
    if (!dims_ok(name, *nx, *ny, *nz) || !serial_decomp(name, *nz, *nz_loc, *iz_loc)) return

}  // namespace

extern "C" {

void setbcc_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *vx,
             const double *vy, const double *vz, const double *vset, const double *vdep,
             const double *rkver, const double *dz)
{
    DROPIN_GUARD("setbcc");
    const Grid g = make_grid(*nx, *ny, *nz, 1, 0.0, 0.0);
    Stage st;
    double *dc = st.up(c, g.G);
    const double *dvx = st.up(vx, g.N), *dvy = st.up(vy, g.N), *dvz = st.up(vz, g.N);
    const double *dvs = st.up(vset, g.N), *dvd = st.up(vdep, g.NXY), *drk = st.up(rkver, g.N);
    const double *ddz = st.up(dz, (size_t)*nz + 1);
    if (!dc || !dvx || !dvy || !dvz || !dvs || !dvd || !drk || !ddz) {
        set_error("setbcc: device staging failed");
        return;
    }
    if (p64::setbcc(0, g, 1, dc, dvx, dvy, dvz, dvs, dvd, drk, ddz)) return;
    copy_back(c, dc, g.G);
}

void advctzc_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *vz,
              const double *vset, const double *dz, double *dt, int *lmeth)
{
    (void)lmeth;
    DROPIN_GUARD("advctzc");
    const Grid g = make_grid(*nx, *ny, *nz, 1, 0.0, 0.0);
    std::vector<double> dzv, ldz;
    std::vector<ZCoef> zc;
    std::vector<double> zl(*nz);
    // zlayer is not an argument of advctzc: rebuild the coefficients from dz
    dzv.assign(dz, dz + *nz + 1);
    zc.assign(*nz + 2, ZCoef{0, 0, 0, 0});
    for (int k = 1; k <= *nz; ++k) {
        zc[k].dzz = 2.0 / (dzv[k] + dzv[k - 1]);
        zc[k].dzk = 1.0 / dzv[k];
        zc[k].dzk_1 = 1.0 / dzv[k - 1];
        zc[k].dzm1 = dzv[k - 1];
    }
    Stage st;
    double *dc = st.up(c, g.G);
    const double *dvz = st.up(vz, g.N), *dvs = st.up(vset, g.N);
    const ZCoef *dzc = st.up(zc.data(), zc.size());
    if (!dc || !dvz || !dvs || !dzc) {
        set_error("advctzc: device staging failed");
        return;
    }
    if (p64::advctzc(0, g, 1, dc, dvz, dvs, dzc, 1.0 / dzv[*nz], *dt)) return;
    copy_back(c, dc, g.G);
}

void advctx_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *vx,
             double *dx, double *dt, double *work, int *lmeth)
{
    (void)work;
    (void)lmeth;
    DROPIN_GUARD("advctx");
    const Grid g = make_grid(*nx, *ny, *nz, 1, *dx, 0.0);
    Stage st;
    double *dc = st.up(c, g.G);
    const double *dv = st.up(vx, g.N);
    if (!dc || !dv) {
        set_error("advctx: device staging failed");
        return;
    }
    if (p64::advctx(0, g, 1, dc, dv, *dx, *dt)) return;
    copy_back(c, dc, g.G);
}

void advcty_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *vy,
             double *dy, double *dt, double *work, int *lmeth)
{
    (void)work;
    (void)lmeth;
    DROPIN_GUARD("advcty");
    const Grid g = make_grid(*nx, *ny, *nz, 1, 0.0, *dy);
    Stage st;
    double *dc = st.up(c, g.G);
    const double *dv = st.up(vy, g.N);
    if (!dc || !dv) {
        set_error("advcty: device staging failed");
        return;
    }
    if (p64::advcty(0, g, 1, dc, dv, *dy, *dt)) return;
    copy_back(c, dc, g.G);
}

void diffz_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *dz,
            double *dt, const double *rkver, const double *rho, double *work)
{
    (void)work;
    DROPIN_GUARD("diffz");
    const Grid g = make_grid(*nx, *ny, *nz, 1, 0.0, 0.0);
    std::vector<ZCoef> zc(*nz + 2, ZCoef{0, 0, 0, 0});
    std::vector<double> ldz(*nz + 1);
    for (int k = 0; k <= *nz; ++k) ldz[k] = 1.0 / dz[k];
    for (int k = 1; k <= *nz; ++k) zc[k].dzz = 2.0 / (dz[k] + dz[k - 1]);
    Stage st;
    double *dc = st.up(c, g.G);
    const double *dk = st.up(rkver, g.N), *dr = st.up(rho, g.N);
    const ZCoef *dzc = st.up(zc.data(), zc.size());
    const double *dl = st.up(ldz.data(), ldz.size());
    if (!dc || !dk || !dr || !dzc || !dl) {
        set_error("diffz: device staging failed");
        return;
    }
    if (p64::diffz(0, g, 1, dc, dzc, dl, *dt, dk, dr)) return;
    copy_back(c, dc, g.G);
}

void diffx_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, double *dx,
            double *dt, const double *rkhor, const double *rho, double *work)
{
    (void)work;
    DROPIN_GUARD("diffx");
    const Grid g = make_grid(*nx, *ny, *nz, 1, *dx, 0.0);
    Stage st;
    double *dc = st.up(c, g.G);
    const double *dk = st.up(rkhor, g.N), *dr = st.up(rho, g.N);
    if (!dc || !dk || !dr) {
        set_error("diffx: device staging failed");
        return;
    }
    if (p64::diffx(0, g, 1, dc, *dx, *dt, dk, dr)) return;
    copy_back(c, dc, g.G);
}

void diffy_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, double *dy,
            double *dt, const double *rkhor, const double *rho, double *work)
{
    (void)work;
    DROPIN_GUARD("diffy");
    const Grid g = make_grid(*nx, *ny, *nz, 1, 0.0, *dy);
    Stage st;
    double *dc = st.up(c, g.G);
    const double *dk = st.up(rkhor, g.N), *dr = st.up(rho, g.N);
    if (!dc || !dk || !dr) {
        set_error("diffy: device staging failed");
        return;
    }
    if (p64::diffy(0, g, 1, dc, *dy, *dt, dk, dr)) return;
    copy_back(c, dc, g.G);
}

void accum_(const double *c, double *cdry, int *ic, int *nx, int *ny, int *nz, int *nz_loc,
            int *iz_loc, const double *vz, double *dt, const double *vset, const double *vdep)
{
    (void)vz;
    DROPIN_GUARD("accum");
    const Grid g = make_grid(*nx, *ny, *nz, 1, 0.0, 0.0);
    const size_t slot = (size_t)*ic * g.NXY;
    Stage st;
    const double *dc = st.up(c, g.G);
    double *dd = st.up(cdry + slot, g.NXY);
    const double *dvs = st.up(vset, g.N), *dvd = st.up(vdep, g.NXY);
    if (!dc || !dd || !dvs || !dvd) {
        set_error("accum: device staging failed");
        return;
    }
    if (p64::accum(0, g, 1, dc, dd, 0, *dt, dvs, dvd)) return;
    copy_back(cdry + slot, dd, g.NXY);
}

void wetdep_(double *c, double *cwet, int *ic, int *nx, int *ny, int *nz, int *nz_loc,
             int *iz_loc, double *dt, const double *diam, const double *zlayer, const double *dz,
             const double *zi, const double *prate)
{
    DROPIN_GUARD("wetdep");
    if (diam[*ic - 1] > 100.0 * 1e-6) return;  // wetdep.f90:33-35
    const Grid g = make_grid(*nx, *ny, *nz, 1, 0.0, 0.0);
    const size_t slot = (size_t)*ic * g.NXY;
    const int one = 1;
    Stage st;
    double *dc = st.up(c, g.G);
    double *dw = st.up(cwet + slot, g.NXY);
    const int *dwet = st.up(&one, 1);
    const double *dzl = st.up(zlayer, g.nz), *ddz = st.up(dz, g.nz + 1);
    const double *dzi = st.up(zi, g.NXY), *dp = st.up(prate, g.NXY);
    if (!dc || !dw || !dwet || !dzl || !ddz || !dzi || !dp) {
        set_error("wetdep: device staging failed");
        return;
    }
    if (p64::wetdep(0, g, 1, dc, dw, dwet, *dt, dzl, ddz, dzi, dp)) return;
    copy_back(c, dc, g.G);
    copy_back(cwet + slot, dw, g.NXY);
}

void cmass2d_(const double *c, const double *vx, const double *vy, const double *vz,
              const double *vset, int *nx, int *ny, int *nz, int *nc, int *ic_loc, int *nc_loc,
              int *iz_loc, int *nz_loc, double *dx, double *dy, const double *dz, double *dt,
              double *outmass)
{
    DROPIN_GUARD("cmass2d");
    if (*ic_loc != 1 || *nc_loc != *nc) {
        set_error("cmass2d: only ic_loc == 1, nc_loc == nc is supported");
        return;
    }
    const Grid g = make_grid(*nx, *ny, *nz, *nc, *dx, *dy);
    Stage st;
    const double *dc = st.up(c, g.G * *nc);
    const double *dvx = st.up(vx, g.N), *dvy = st.up(vy, g.N), *dvz = st.up(vz, g.N);
    const double *dvs = st.up(vset, g.N * *nc), *ddz = st.up(dz, (size_t)*nz + 1);
    double *scr = st.up<double>(nullptr, red::kPartials);
    double *om = st.up(outmass, 1);
    if (!dc || !dvx || !dvy || !dvz || !dvs || !ddz || !scr || !om) {
        set_error("cmass2d: device staging failed");
        return;
    }
    if (red::cmass2d64(0, g, *nc, dc, dvx, dvy, dvz, dvs, ddz, *dt, scr, om)) return;
    copy_back(outmass, om, 1);
}

void cmass3d_(const double *c, const double *cload, const double *cdry, const double *cwet,
              int *nx, int *ny, int *nz, int *nc, double *dx, double *dy, const double *dz,
              double *massvol, double *massdep, double *massdry, double *masswet)
{
    if (!dims_ok("cmass3d", *nx, *ny, *nz)) return;
    const Grid g = make_grid(*nx, *ny, *nz, *nc, *dx, *dy);
    Stage st;
    const double *dc = st.up(c, g.G * *nc);
    const double *dl = st.up(cload, g.NXY), *dd = st.up(cdry, g.NXY), *dw = st.up(cwet, g.NXY);
    const double *ddz = st.up(dz, (size_t)*nz + 1);
    double *scr = st.up<double>(nullptr, 4 * red::kPartials);
    double *out = st.up<double>(nullptr, 4);
    if (!dc || !dl || !dd || !dw || !ddz || !scr || !out) {
        set_error("cmass3d: device staging failed");
        return;
    }
    if (red::cmass3d(0, g, dc, dl, dd, dw, 0, ddz, scr, out)) return;
    double o[4];
    if (copy_back(o, out, 4)) return;
    *massvol = o[0] * *dx * *dy;  // cmass3d.f90:191-194
    *massdep = o[1] * *dx * *dy;
    *massdry = o[2] * *dx * *dy;
    *masswet = o[3] * *dx * *dy;
}

const char *f3d_last_error(void) { return f3d::g_err; }

}  // extern "C"

namespace {
// dst(nx,ny,nz) = src(i0+1.., j0+1.., 1..) of a (snx,sny,*) array
__global__ void k_window(double *dst, const double *src, int nx, int ny, int nz, int snx, int sny,
                         int i0, int j0)
{
    const long n = (long)nx * ny * nz;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((long)nx * ny));
        dst[t] = src[(i0 + i) + (long)snx * ((j0 + j) + (long)sny * k)];
    }
}

// plain copy, one float4 per thread: the known-size kernel for HBM calibration
__global__ void k_hbm_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}
}  // namespace

namespace f3d {
int window_copy(hipStream_t s, double *dst, const double *src, int nx, int ny, int nz, int snx,
                int sny, int i0, int j0)
{
    const long n = (long)nx * ny * nz;
    const unsigned nb = (unsigned)std::min<long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_window, dim3(nb), dim3(256), 0, s, dst, src, nx, ny, nz, snx, sny, i0, j0);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}
}  // namespace f3d

extern "C" int f3d_mem_info(long long *free_bytes, long long *total_bytes)
{
    size_t fr = 0, tot = 0;
    F3D_HIP(hipMemGetInfo(&fr, &tot));
    if (free_bytes) *free_bytes = (long long)fr;
    if (total_bytes) *total_bytes = (long long)tot;
    return F3D_OK;
}

extern "C" int f3d_hbm_copy(long long nbytes, int reps, double *avg_ms)
{
    if (nbytes <= 0 || nbytes % 4096 != 0 || reps < 1 || !avg_ms) {
        set_error("f3d_hbm_copy: nbytes must be a positive multiple of 4096, reps >= 1");
        return F3D_EINVAL;
    }
    DeviceBuffer a, b;
    if (a.alloc((size_t)nbytes, nullptr) || b.alloc((size_t)nbytes, nullptr)) return F3D_ENOMEM;
    const long n = (long)(nbytes / 16);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error("f3d_hbm_copy: hipEventCreate failed");
        return F3D_EHIP;
    }
    int rc = F3D_OK;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_hbm_copy, dim3(nb), dim3(256), 0, 0, (const float4 *)a.p, (float4 *)b.p, n);
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_hbm_copy, dim3(nb), dim3(256), 0, 0, (const float4 *)a.p,
                           (float4 *)b.p, n);
    hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) {
        set_error("f3d_hbm_copy: copy kernel failed");
        rc = F3D_EHIP;
    } else {
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        *avg_ms = (double)ms / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}
