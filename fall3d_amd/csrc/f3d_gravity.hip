// Gravity current of the umbrella cloud (readat.f90:149-200 GRAVITY_CURRENT
// block): the zero-divergence radial wind added at each meteo read
// (addradialwind.f90, meteo.f90:140), the wind divergence (setdivu.f90,
// meteo.f90:145) and the mass correction of the non-zero divergence at the
// log times (divcorr.f90, endstep.f90:71).  All FP64.  The scalars of the
// radial wind (current radius, front velocity, the levels of the umbrella)
// are formed once per meteo read on the host in the reference's operation
// order; the per-column wind and the per-cell divergence and correction run
// on the device.
#include <cmath>
#include <vector>

#include "f3d_ctx.h"

using namespace f3d;

namespace {

constexpr double kPi = 3.14159265358979323846;  // Numeric.f90:13

// Scalars of one radial-wind update (addradialwind.f90:52-155).
struct RadialWind {
    double dX1, dX2, xorigr, yorigr, dx, dy, cx, cy;
    double min_radius, max_radius, radius_grav, c_grav_factor;
    int kmin, kmax;
};

// addradialwind.f90:157-173, one thread per column: the radial velocity of
// the gravity current between min_radius and max_radius added to the
// (Hm1-scaled) vx and to vy on the levels kmin..kmax
__global__ void k_radialwind(Grid g, double *vx, double *vy, const double *hm1, RadialWind w)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.NXY) return;
    const int i = (int)(t % g.nx) + 1, j = (int)(t / g.nx) + 1;
    const double h = hm1[t];
    const double deltax = w.dX1 * h * ((w.xorigr + i * w.dx) - w.cx) / w.dx;
    const double deltay = w.dX2 * ((w.yorigr + j * w.dy) - w.cy) / w.dy;
    const double radius = sqrt(deltax * deltax + deltay * deltay);
    if (fabs(radius) <= w.min_radius || fabs(radius) > w.max_radius) return;
    const double rvel =
        w.c_grav_factor / radius * (1.0 + radius * radius / (3.0 * (w.radius_grav * w.radius_grav)));
    for (int k = w.kmin; k <= w.kmax; ++k) {
        const long o = t + g.NXY * (long)(k - 1);
        vx[o] = vx[o] + rvel * deltax / radius / h;
        vy[o] = vy[o] + rvel * deltay / radius;
    }
}

// setdivu.f90:16-30, one thread per cell: the divergence of the scaled wind,
// one-sided at the first column, row and level (v(max(1,i-1)))
__global__ void k_setdivu(Grid g, const double *vx, const double *vy, const double *vz,
                          const double *hm1, double dX1, double dX2, const double *dz, double *divu)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.N) return;
    const long c = t % g.NXY;
    const int i = (int)(c % g.nx) + 1, j = (int)(c / g.nx) + 1, k = (int)(t / g.NXY) + 1;
    const double vi = vx[t], vi_1 = vx[i > 1 ? t - 1 : t];
    const double vj = vy[t], vj_1 = vy[j > 1 ? t - g.nx : t];
    const double vk = vz[t], vk_1 = vz[k > 1 ? t - g.NXY : t];
    divu[t] = (vi - vi_1) / (hm1[c] * dX1) + (vj - vj_1) / dX2 + (vk - vk_1) / dz[k];
}

// divcorr.f90:26-38, one thread per cell of levels 1..nz: every class
// scaled by f = 1 + corrdiv where |divu*dt| >= eps
template <typename T>
__global__ void k_divcorr(Grid g, T *c, const double *divu, double dt, double eps, double f)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.N) return;
    if (!(fabs(divu[t] * dt) >= eps)) return;
    const long cc = t % g.NXY;
    const int i = (int)(cc % g.nx) + 1, j = (int)(cc / g.nx) + 1, k = (int)(t / g.NXY) + 1;
    const long o = cidx(g.nx, g.ny, i, j, k);
    for (int ic = 0; ic < g.nc; ++ic) c[(long)ic * g.G + o] = (T)((double)c[(long)ic * g.G + o] * f);
}

// the level nearest height h, the last one on a tie, starting from the
// reference's "large value" zlayer(nz) (addradialwind.f90:110-118)
int nearest_level(const std::vector<double> &zlayer, double h, int k0)
{
    const int nz = (int)zlayer.size();
    int kk = k0;
    double store = zlayer[nz - 1];
    for (int k = 1; k <= nz; ++k) {
        const double dist = fabs(zlayer[k - 1] - h);
        if (dist <= store) {
            kk = k;
            store = dist;
        }
    }
    return kk;
}

}  // namespace

namespace f3d {

// meteo.f90:138-145 on the device fields of a meteo read, after the Hm1
// scaling and corvver: the radial wind (gravity current on and source on)
// and the divergence.
int gravity_meteo(f3d_ctx *c)
{
    const Grid &g = c->g;
    hipStream_t s = c->stream;
    if (c->gravity_on) {
        const f3d_gravity &p = c->grav;
        // addradialwind.f90:52-96
        const double h11 = c->hm1[0];
        const double min_radius = 1.5 * sqrt(h11 * g.dx * h11 * g.dx + g.dy * g.dy);
        const double time_eruption = c->time - p.beg_time;
        double mass_flow_rate = 0.0;  // sum(tmrat(1:nsou,1:nc)), in array order
        for (double v : c->src_tm[0]) mass_flow_rate = mass_flow_rate + v;
        c->grav_active = mass_flow_rate > (double)0.1f;
        if (c->grav_active) {
            const double vol_flow_rate = p.c_flow_rate * sqrt(p.k_entrain) * pow(mass_flow_rate, 0.75) /
                                         pow(p.brunt_vaisala, 0.625);
            const double c_gravr = 3.0 * p.lambda_grav * p.brunt_vaisala * vol_flow_rate / (2.0 * kPi);
            double radius_grav = pow(c_gravr * (time_eruption * time_eruption), 0.33333);
            radius_grav = radius_grav > min_radius ? radius_grav : min_radius;
            const double c_gravu =
                sqrt(2.0 * p.lambda_grav * p.brunt_vaisala * vol_flow_rate / (3.0 * kPi));
            const double c_grav_factor = 0.75 * c_gravu * sqrt(radius_grav);
            const double th_grav = c_gravu / (sqrt(radius_grav) * p.lambda_grav * p.brunt_vaisala);
            double h_tot = c->grav_zs.empty() ? 0.0 : c->grav_zs[0];  // maxval(zs)
            for (double z : c->grav_zs) h_tot = z > h_tot ? z : h_tot;
            const double h_umbr = h_tot / 1.32;
            const double a = h_umbr - 0.5 * th_grav, b = 0.5 * h_tot;
            const double hmin = a > b ? a : b;
            const double a2 = h_umbr + 0.5 * th_grav;
            const double hmax = a2 < h_tot ? a2 : h_tot;
            // addradialwind.f90:102-121: the crater node and the NBL level;
            // the reference clamps the low side only, this clamps both
            int i_wind = (int)((p.center_x - p.xorigr) / p.dx);
            int j_wind = (int)((p.center_y - p.yorigr) / p.dy);
            i_wind = std::min(std::max(1, i_wind), g.nx);
            j_wind = std::min(std::max(1, j_wind), g.ny);
            const int k_wind = nearest_level(c->zlayer, h_umbr, 1);
            const long o = midx(g.nx, g.ny, i_wind, j_wind, k_wind);
            double uv[2];
            F3D_HIP(hipMemcpyAsync(&uv[0], c->d_vx.as<double>() + o, sizeof(double), hipMemcpyDeviceToHost, s));
            F3D_HIP(hipMemcpyAsync(&uv[1], c->d_vy.as<double>() + o, sizeof(double), hipMemcpyDeviceToHost, s));
            F3D_HIP(hipStreamSynchronize(s));
            const double v_wind = sqrt(uv[0] * uv[0] + uv[1] * uv[1]);
            const double v2 = v_wind * v_wind;
            const double max_radius = 1.7777 * c_gravr / (v2 > 0.1 ? v2 : 0.1);
            RadialWind w{g.dx, g.dy, p.xorigr, p.yorigr, p.dx, p.dy, p.center_x, p.center_y,
                         min_radius, max_radius, radius_grav, c_grav_factor,
                         nearest_level(c->zlayer, hmin, 1), nearest_level(c->zlayer, hmax, g.nz)};
            c->grav_diag = {min_radius, max_radius, radius_grav, c_grav_factor, (double)w.kmin,
                            (double)w.kmax, (double)i_wind, (double)j_wind, (double)k_wind};
            hipLaunchKernelGGL(k_radialwind, dim3((unsigned)((g.NXY + 255) / 256)), dim3(256), 0, s, g,
                               c->d_vx.as<double>(), c->d_vy.as<double>(), c->d_hm1.as<double>(), w);
            F3D_CHECK_LAUNCH();
        }
    }
    if (c->d_divu.p) {
        hipLaunchKernelGGL(k_setdivu, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, g,
                           c->d_vx.as<double>(), c->d_vy.as<double>(), c->d_vz.as<double>(),
                           c->d_hm1.as<double>(), g.dx, g.dy, c->d_dz.as<double>(), c->d_divu.as<double>());
        F3D_CHECK_LAUNCH();
    }
    return F3D_OK;
}

// endstep.f90:53,71 after a step: at the log times (the first step, and
// whenever mod(time, 600) <= dt) with the gravity current on, divcorr on the
// budgets of that step (writim's cmass3d, outmass and erumass).
int gravity_endstep(f3d_ctx *c)
{
    if (!c->gravity_on) return F3D_OK;
    if (!(c->iiter == 1 || std::fmod(c->time, 600.0) <= c->dt)) return F3D_OK;
    double b[6];
    TRY(f3d_budget(c, b));  // massvol, massdep, massdry, masswet, outmass, erumass
    const double compmass = b[0] + b[1] + b[4];
    const double eps = 0.005, eps1 = 10.0 * eps;
    const double r = (c->erumass[0] - compmass) / (c->erumass[0]);
    const double lo = r < eps1 ? r : eps1;
    const double corrdiv = lo > -eps1 ? lo : -eps1;
    c->grav_corrdiv = corrdiv;
    const Grid &g = c->g;
    const unsigned nb = (unsigned)((g.N + 255) / 256);
    if (c->precision == F3D_FP64) {
        hipLaunchKernelGGL((k_divcorr<double>), dim3(nb), dim3(256), 0, c->stream, g, c->d_c64.as<double>(),
                           c->d_divu.as<double>(), c->dt, eps, 1.0 + corrdiv);
    } else {
        hipLaunchKernelGGL((k_divcorr<float>), dim3(nb), dim3(256), 0, c->stream, g,
                           c->d_c32[c->cur].as<float>(), c->d_divu.as<double>(), c->dt, eps, 1.0 + corrdiv);
    }
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}

}  // namespace f3d

extern "C" {

int f3d_set_gravity_current(f3d_ctx *c, const f3d_gravity *gc, int nsou, const double *zs)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!gc) {
        c->gravity_on = false;
        return F3D_OK;
    }
    if (c->g.nm != 1) {
        set_error("f3d_set_gravity_current: single-member contexts only (the radial wind depends "
                  "on the source)");
        return F3D_EINVAL;
    }
    if (nsou < 1 || !zs || !(gc->dx > 0.0) || !(gc->dy > 0.0) || !(gc->brunt_vaisala > 0.0) ||
        !(gc->lambda_grav > 0.0)) {
        set_error("f3d_set_gravity_current: need nsou >= 1 heights, dx, dy, BRUNT_VAISALA and "
                  "LAMBDA_GRAV > 0");
        return F3D_EINVAL;
    }
    if (!c->d_divu.p && c->d_divu.alloc(sizeof(double) * c->g.N, &c->bytes)) {
        set_error("f3d_set_gravity_current: out of device memory");
        return F3D_ENOMEM;
    }
    c->grav = *gc;
    c->grav_zs.assign(zs, zs + nsou);
    c->gravity_on = true;
    return F3D_OK;
}

int f3d_get_gravity_current(f3d_ctx *c, double *diag, double *corrdiv, int *active)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (diag)
        for (int k = 0; k < 9; ++k) diag[k] = c->grav_diag[k];
    if (corrdiv) *corrdiv = c->grav_corrdiv;
    if (active) *active = c->grav_active ? 1 : 0;
    return F3D_OK;
}

int f3d_get_divu(f3d_ctx *c, double *divu)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!c->meteo_ready) {
        set_error("f3d_get_divu: no meteo read yet");
        return F3D_ESTATE;
    }
    const Grid &g = c->g;
    if (!c->d_divu.p) {
        // the divergence of the current (scaled) wind, as the meteo read
        // would have formed it
        if (c->d_divu.alloc(sizeof(double) * g.N, &c->bytes)) {
            set_error("f3d_get_divu: out of device memory");
            return F3D_ENOMEM;
        }
        hipLaunchKernelGGL(k_setdivu, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, c->stream, g,
                           c->d_vx.as<double>(), c->d_vy.as<double>(), c->d_vz.as<double>(),
                           c->d_hm1.as<double>(), g.dx, g.dy, c->d_dz.as<double>(), c->d_divu.as<double>());
        F3D_CHECK_LAUNCH();
    }
    if (divu)
        F3D_HIP(hipMemcpyAsync(divu, c->d_divu.p, sizeof(double) * g.N, hipMemcpyDeviceToHost, c->stream));
    F3D_HIP(hipStreamSynchronize(c->stream));
    return F3D_OK;
}

}  // extern "C"
