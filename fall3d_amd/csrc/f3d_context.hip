// Device-resident context: the f3d_* half of the C ABI.
//
// The context mirrors the reference driver's state (Master.f90 arrays,
// Numeric grid scalars, the Fall3d.f90 clock) in HBM.  f3d_step runs the
// loop body of Fall3d.f90:112-117 on the device; nothing is copied to the
// host during a step.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "f3d_ctx.h"

using namespace f3d;

namespace f3d {

f32::Fields fp32_fields(const f3d_ctx *c)
{
    return f32::fields_view(c->g, c->d_f32.as<float>(), c->d_vsum1.as<double>());
}

f32::StepArgs fp32_args(f3d_ctx *c, int cur_in, int iiter)
{
    f32::StepArgs a{c->g,      c->d_c32[cur_in].as<float>(), c->d_c32[1 - cur_in].as<float>(),
                    c->d_cdry.as<double>(), c->dt, iiter, c->d_zc.as<ZCoef>(), c->side,
                    c->ev_fork, c->ev_join};
    a.gbuf = c->d_gbuf.as<float>();
    if (c->wet_on) {
        a.wet = 1;
        a.lam = c->d_lam.as<double>();
        a.kwet = c->d_kwet.as<int>();
        a.wetcls = c->d_wet.as<int>();
        a.cwet = c->d_cwet.as<double>();
        a.dz64 = c->d_dz.as<double>();
    }
    return a;
}

namespace f32 {
// wetdep.f90:38-46 per column: lambda = min(dt*a*P**b, 1) and the top level
// below the PBL height (the levels rise, so the wet ones are 1..k)
__global__ void k_wet_columns(Grid g, double dt, const double *prate, const double *zi,
                              const double *zlayer, double *lam, int *kwet)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.NXY) return;
    const double a = 8.4e-5, b = 0.79;
    double l = dt * a * pow(prate[t], b);
    l = l < 1.0 ? l : 1.0;
    int kw = 0;
    for (int k = 1; k <= g.nz; ++k)
        if (zlayer[k - 1] <= zi[t]) kw = k;
    lam[t] = l;
    kwet[t] = l > 0.0 ? kw : 0;
}

int wet_columns(hipStream_t s, const Grid &g, double dt, const double *prate, const double *zi,
                const double *zlayer, double *lam, int *kwet)
{
    hipLaunchKernelGGL(k_wet_columns, dim3((unsigned)((g.NXY + 255) / 256)), dim3(256), 0, s, g, dt,
                       prate, zi, zlayer, lam, kwet);
    F3D_CHECK_LAUNCH();
    return F3D_OK;
}
}  // namespace f32

f32::SrcArgs fp32_src_args(f3d_ctx *c, const f32::StepArgs &a)
{
    return f32::src_args(c->g, const_cast<float *>(a.cin), c->d_srcoff.as<int>(),
                         c->d_srcgrp.as<int4>(), c->nsrcgrp, c->d_srcperm.as<int>(),
                         c->d_isrc.as<int>(), c->d_jsrc.as<int>(), c->d_ksrc.as<int>(),
                         c->d_tmrat.as<double>(), c->d_zlayer.as<double>(), c->d_dz.as<double>(),
                         c->d_hm1.as<double>(), c->dt);
}

int fp32_setsrc(f3d_ctx *c, hipStream_t s, const f32::StepArgs &a)
{
    return f32::setsrc(s, c->g, const_cast<float *>(a.cin), c->d_srcoff.as<int>(),
                       c->d_srcgrp.as<int4>(), c->nsrcgrp, c->d_srcperm.as<int>(),
                       c->d_isrc.as<int>(), c->d_jsrc.as<int>(), c->d_ksrc.as<int>(),
                       c->d_tmrat.as<double>(), c->d_zlayer.as<double>(), c->d_dz.as<double>(),
                       c->d_hm1.as<double>(), c->dt);
}

f32::FrameArgs fp32_frame_args(f3d_ctx *c, const f32::StepArgs &a)
{
    return f32::frame_args(a, fp32_fields(c), c->d_vx.as<double>(), c->d_vy.as<double>(),
                           c->d_vz.as<double>(), c->d_vset.as<double>(), c->d_dz.as<double>(),
                           c->d_scratch.as<double>(), c->d_outmass.as<double>());
}

int fp32_outflux(f3d_ctx *c, hipStream_t s, const f32::StepArgs &a)
{
    return f32::outflux(s, a, fp32_fields(c), c->d_vx.as<double>(), c->d_vy.as<double>(),
                        c->d_vz.as<double>(), c->d_vset.as<double>(), c->d_dz.as<double>(),
                        c->d_scratch.as<double>(), c->d_outmass.as<double>());
}

// An FP32 step leaves the x-face ghosts of the new state in the ghost
// buffer and the rest of its ghost frame unwritten (it follows from the
// step's input, the other state buffer, until the next step overwrites
// it); readers of the whole state (ghost frame included) call this first.
// Queued on the context's stream.
int fp32_complete_frame(f3d_ctx *c)
{
    if (c->precision != F3D_FP32 || !c->frame_dirty) return F3D_OK;
    const f32::Fields f = fp32_fields(c);
    f32::FrameFill a{c->g, c->d_c32[c->cur].as<float>(), c->d_c32[1 - c->cur].as<float>(),
                     c->d_gbuf.as<float>(), c->d_vy.as<double>(), f.bfac, f.tfac};
    TRY(f32::frame_fill(c->stream, a));
    c->frame_dirty = false;
    return F3D_OK;
}

}  // namespace f3d

bool check_ctx(const f3d_ctx *c)
{
    if (!c) {
        set_error("null context");
        return false;
    }
    ++c->api_seq;
    return true;
}

namespace {

int upload(f3d::DeviceBuffer &b, const void *h, size_t n, hipStream_t s)
{
    F3D_HIP(hipMemcpyAsync(b.p, h, n, hipMemcpyHostToDevice, s));
    return F3D_OK;
}

void record_start(f3d_ctx *c)
{
    if (!c->timing) return;
    (void)hipEventRecord(c->ev_start[c->ev_next], c->stream);
}

void flush_timing(f3d_ctx *c)
{
    if (c->ev_next == 0) return;
    (void)hipEventSynchronize(c->ev_stop[c->ev_next - 1]);
    for (int i = 0; i < c->ev_next; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_start[i], c->ev_stop[i]) == hipSuccess) {
            c->t_ms += ms;
            c->t_count += 1;
        }
    }
    c->ev_next = 0;
}

void record_stop(f3d_ctx *c)
{
    if (!c->timing) return;
    (void)hipEventRecord(c->ev_stop[c->ev_next], c->stream);
    c->ev_next += 1;
    if (c->ev_next == (int)c->ev_stop.size()) flush_timing(c);
}

int step_fp64(f3d_ctx *c)
{
    const Grid &g = c->g;
    const int nc = g.nc;
    hipStream_t s = c->stream;
    double *cc = c->d_c64.as<double>();
    const double *vx = c->d_vx.as<double>(), *vy = c->d_vy.as<double>(), *vz = c->d_vz.as<double>();
    const double *vset = c->d_vset.as<double>(), *vdep = c->d_vdep.as<double>();
    const double *rkver = c->d_rkver.as<double>(), *rkhor = c->d_rkhor.as<double>();
    const double *rkh1 = c->d_rkh1.as<double>(), *rho = c->d_rho.as<double>();
    const double *dz = c->d_dz.as<double>();
    const double dt = c->dt;
    record_start(c);
    // setsrc.f90:20-35
    TRY(p64::setsrc(s, g, cc, c->nsou, c->d_isrc.as<int>(), c->d_jsrc.as<int>(),
                    c->d_ksrc.as<int>(), c->d_tmrat.as<double>(), c->d_zlayer.as<double>(), dz,
                    c->d_hm1.as<double>(), dt));
    // sizloop.f90:31-73, routine by routine over all classes
    TRY(p64::setbcc(s, g, nc, cc, vx, vy, vz, vset, vdep, rkver, dz));
    TRY(p64::advctzc(s, g, nc, cc, vz, vset, c->d_zc.as<ZCoef>(), c->ldz[g.nz], dt));
    if (c->iiter % 2 == 0) {
        TRY(p64::advctx(s, g, nc, cc, vx, g.dx, dt));
        TRY(p64::advcty(s, g, nc, cc, vy, g.dy, dt));
    } else {
        TRY(p64::advcty(s, g, nc, cc, vy, g.dy, dt));
        TRY(p64::advctx(s, g, nc, cc, vx, g.dx, dt));
    }
    TRY(p64::diffz(s, g, nc, cc, c->d_zc.as<ZCoef>(), c->d_ldz.as<double>(), dt, rkver, rho));
    if (c->iiter % 2 == 0) {
        TRY(p64::diffx(s, g, nc, cc, g.dx, dt, rkh1, rho));
        TRY(p64::diffy(s, g, nc, cc, g.dy, dt, rkhor, rho));
    } else {
        TRY(p64::diffy(s, g, nc, cc, g.dy, dt, rkhor, rho));
        TRY(p64::diffx(s, g, nc, cc, g.dx, dt, rkh1, rho));
    }
    // sizloop.f90:65-67 wet deposition (all classes are particles: np = nc)
    if (c->wet_on)
        TRY(p64::wetdep(s, g, nc, cc, c->d_cwet.as<double>() + g.NXY, c->d_wet.as<int>(), dt,
                        c->d_zlayer.as<double>(), dz, c->d_zi.as<double>(),
                        c->d_prate.as<double>()));
    TRY(p64::accum(s, g, nc, cc, c->d_cdry.as<double>(), 1, dt, vset, vdep));
    // sizloop.f90:77-83
    TRY(p64::clip(s, cc, g.G * nc));
    TRY(red::cmass2d64(s, g, nc, cc, vx, vy, vz, vset, dz, dt, c->d_scratch.as<double>(),
                       c->d_outmass.as<double>()));
    record_stop(c);
    return F3D_OK;
}

int step_fp32(f3d_ctx *c)
{
    hipStream_t s = c->stream;
    const f32::Fields f = fp32_fields(c);
    const f32::StepArgs a = fp32_args(c, c->cur, c->iiter);
    TRY(fp32_setsrc(c, s, a));
    record_start(c);
    TRY(f32::fused_step(s, a, f, c->tile));
    record_stop(c);
    // the flux launch left the x-face ghosts in the ghost buffer; the
    // outflux reads those and c_in's other faces
    TRY(fp32_outflux(c, s, a));
    c->cur = 1 - c->cur;
    c->frame_dirty = true;
    return F3D_OK;
}

// FP32 step fields from the FP64 coefficients; the face coefficients
// depend on dt, so they are rebuilt whenever dt changes
int rebuild_fields(f3d_ctx *c)
{
    if (c->precision != F3D_FP32) return F3D_OK;
    ++c->gen;
    if (c->wet_on)  // lambda depends on dt, the wet levels on zi
        TRY(f32::wet_columns(c->stream, c->g, c->dt, c->d_prate.as<double>(), c->d_zi.as<double>(),
                             c->d_zlayer.as<double>(), c->d_lam.as<double>(),
                             c->d_kwet.as<int>()));
    const Grid &g = c->g;
    TRY(f32::build_fields(c->stream, g, c->d_vx.as<double>(), c->d_vy.as<double>(),
                          c->d_vz.as<double>(), c->d_rho.as<double>(), c->d_rkh1.as<double>(),
                          c->d_rkhor.as<double>(), c->d_rkver.as<double>(), c->d_vset.as<double>(),
                          c->d_vdep.as<double>(), c->d_dz.as<double>(), c->d_hm1.as<double>(),
                          c->d_zc.as<ZCoef>(),
                          c->dt, c->d_f32.as<float>(),
                          c->d_vsum1.as<double>()));
    F3D_HIP(hipStreamSynchronize(c->stream));
    return F3D_OK;
}

}  // namespace

extern "C" {

static int create_impl(f3d_ctx **out, const f3d_grid *gr, int precision, int nm)
{
    if (!out || !gr) {
        set_error("f3d_create: null argument");
        return F3D_EINVAL;
    }
    *out = nullptr;
    if (gr->nx < 2 || gr->ny < 2 || gr->nz < 2 || gr->nc < 1 || !gr->zlayer) {
        set_error("f3d_create: need nx, ny, nz >= 2, nc >= 1 and zlayer (got %d %d %d %d)", gr->nx,
                  gr->ny, gr->nz, gr->nc);
        return F3D_EINVAL;
    }
    if (precision != F3D_FP64 && precision != F3D_FP32) {
        set_error("f3d_create: unknown precision %d", precision);
        return F3D_EINVAL;
    }
    if (precision == F3D_FP32 && gr->nz > 64) {
        set_error("f3d_create: the FP32 fused path supports nz <= 64 (got %d)", gr->nz);
        return F3D_EINVAL;
    }
    if (!(gr->dx > 0.0) || !(gr->dy > 0.0)) {
        set_error("f3d_create: dx and dy must be positive");
        return F3D_EINVAL;
    }
    if (nm < 1 || (nm > 1 && precision != F3D_FP32)) {
        set_error("f3d_create: ensembles (%d members) run on the FP32 fused path only", nm);
        return F3D_EINVAL;
    }
    f3d_ctx *c = new f3d_ctx();
    c->precision = precision;
    c->g = make_grid(gr->nx, gr->ny, gr->nz, gr->nc, gr->dx, gr->dy);
    c->g.nm = nm;
    c->src_i.assign(nm, {});
    c->src_j.assign(nm, {});
    c->src_k.assign(nm, {});
    c->src_tm.assign(nm, {});
    c->flowrate.assign(nm, 0.0);
    c->erumass.assign(nm, 0.0);
    c->lonlat = gr->lonlat;
    // reagrd.f90:155-158: dX1 = Rearth*(dx*pi/180) with dx in degrees
    const double m_per_deg = 6356.0e3 * (3.14159265358979323846 / 180.0);
    c->khn_dx = !gr->lonlat ? gr->dx : (gr->dlon > 0.0 ? gr->dlon : gr->dx / m_per_deg);
    c->khn_dy = !gr->lonlat ? gr->dy : (gr->dlat > 0.0 ? gr->dlat : gr->dy / m_per_deg);
    const Grid &g = c->g;
    c->zlayer.assign(gr->zlayer, gr->zlayer + g.nz);
    vertical_coefs(g.nz, gr->zlayer, c->dz, c->zc, c->ldz);
    c->hm1.assign(g.NXY, 1.0);
    if (gr->hm1) {
        c->hm1.assign(gr->hm1, gr->hm1 + g.NXY);
        for (double h : c->hm1)
            if (h != 1.0) c->uniform_hm1 = false;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        set_error("f3d_create: hipStreamCreate / hipEventCreate failed");
        f3d_destroy(c);
        return F3D_EHIP;
    }
    long long *B = &c->bytes;
    int rc = 0;
    const size_t d8 = sizeof(double);
    rc |= c->d_dz.alloc(d8 * (g.nz + 1), B);
    rc |= c->d_ldz.alloc(d8 * (g.nz + 1), B);
    rc |= c->d_zlayer.alloc(d8 * g.nz, B);
    rc |= c->d_hm1.alloc(d8 * g.NXY, B);
    rc |= c->d_kmh.alloc(d8 * g.NXY, B);
    rc |= c->d_zc.alloc(sizeof(ZCoef) * (g.nz + 2), B);
    for (DeviceBuffer *b : {&c->d_vx, &c->d_vy, &c->d_vz, &c->d_tempe, &c->d_vptemp, &c->d_rho,
                            &c->d_rkhor, &c->d_rkh1, &c->d_rkver, &c->d_visc})
        rc |= b->alloc(d8 * g.N, B);
    for (DeviceBuffer *b : {&c->d_ustar, &c->d_zi, &c->d_rl}) rc |= b->alloc(d8 * g.NXY, B);
    rc |= c->d_vdep.alloc(d8 * g.NXY * g.nc, B);
    rc |= c->d_vset.alloc(d8 * g.N * g.nc, B);
    rc |= c->d_cls.alloc(sizeof(met::ClassConst) * g.nc, B);
    rc |= c->d_cd.alloc(d8 * 2, B);
    rc |= c->d_status.alloc(sizeof(int) * 2, B);
    rc |= c->d_dtcn.alloc(sizeof(unsigned long long), B);
    rc |= c->d_cdry.alloc(d8 * g.NXY * (g.nc + 1) * nm, B);
    rc |= c->d_cwet.alloc(d8 * g.NXY * (g.nc + 1) * nm, B);
    const size_t nscr = std::max<size_t>({4 * (size_t)red::kPartials,
                                          4 * (size_t)nm * red::kMemberPartials,
                                          (size_t)nm * red::kFramePartials});
    rc |= c->d_scratch.alloc(d8 * nscr, B);
    rc |= c->d_outmass.alloc(d8 * nm, B);
    rc |= c->d_budget.alloc(d8 * 4 * nm, B);
    rc |= c->d_srcoff.alloc(sizeof(int) * (nm + 1), B);
    if (precision == F3D_FP64) {
        rc |= c->d_c64.alloc(d8 * g.G * g.nc, B);
    } else {
        rc |= c->d_c32[0].alloc(sizeof(float) * g.G * g.nc * nm, B);
        rc |= c->d_c32[1].alloc(sizeof(float) * g.G * g.nc * nm, B);
        rc |= c->d_f32.alloc(sizeof(float) * f32::fields_floats(g), B);
        rc |= c->d_vsum1.alloc(d8 * g.NXY * g.nc, B);
        rc |= c->d_gbuf.alloc(sizeof(float) * f32::ghost_floats(g) * g.nc * nm, B);
    }
    if (rc) {
        f3d_destroy(c);
        return F3D_ENOMEM;
    }
    // host-side per-column RAMS floor Kmh = 0.075*((Hm1*dX1*dX2)**(2/3)) (kappa3.f90:175)
    c->kmh.resize(g.NXY);
    for (long t = 0; t < g.NXY; ++t) {
        const double dXX = c->hm1[t] * g.dx;
        c->kmh[t] = 0.075 * (pow(dXX * g.dy, 2.0 / 3.0));
    }
    hipStream_t s = c->stream;
    rc |= upload(c->d_dz, c->dz.data(), d8 * (g.nz + 1), s);
    rc |= upload(c->d_ldz, c->ldz.data(), d8 * (g.nz + 1), s);
    rc |= upload(c->d_zlayer, c->zlayer.data(), d8 * g.nz, s);
    rc |= upload(c->d_hm1, c->hm1.data(), d8 * g.NXY, s);
    rc |= upload(c->d_kmh, c->kmh.data(), d8 * g.NXY, s);
    rc |= upload(c->d_zc, c->zc.data(), sizeof(ZCoef) * (g.nz + 2), s);
    // zero state, coefficients and budgets (Master.f90:getmem1 initialises to 0)
    for (DeviceBuffer *b : {&c->d_rkhor, &c->d_rkh1, &c->d_rkver, &c->d_vset, &c->d_vdep,
                            &c->d_cdry, &c->d_cwet, &c->d_outmass, &c->d_c64, &c->d_c32[0],
                            &c->d_c32[1], &c->d_gbuf, &c->d_srcoff})
        if (b->p) rc |= (hipMemsetAsync(b->p, 0, b->bytes, s) != hipSuccess);
    if (hipStreamSynchronize(s) != hipSuccess || rc) {
        set_error("f3d_create: initial upload failed");
        f3d_destroy(c);
        return F3D_EHIP;
    }
    *out = c;
    return F3D_OK;
}

int f3d_create(f3d_ctx **out, const f3d_grid *gr, int precision)
{
    return create_impl(out, gr, precision, 1);
}

int f3d_create_ensemble(f3d_ctx **out, const f3d_grid *gr, int nmembers)
{
    return create_impl(out, gr, F3D_FP32, nmembers);
}

void f3d_destroy(f3d_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    c->side = nullptr;
    for (auto e : c->ev_start) (void)hipEventDestroy(e);
    for (auto e : c->ev_stop) (void)hipEventDestroy(e);
    c->ev_start.clear();
    c->ev_stop.clear();
    hipStream_t s = c->stream;
    c->stream = nullptr;
    delete c;  // DeviceBuffer destructors free device memory
    if (s) (void)hipStreamDestroy(s);
}

int f3d_set_classes(f3d_ctx *c, const double *diam, const double *rhop, const double *sphe,
                    int modv)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!diam || !rhop || !sphe) {
        set_error("f3d_set_classes: null argument");
        return F3D_EINVAL;
    }
    c->diam_h.assign(diam, diam + c->g.nc);
    std::vector<met::ClassConst> cls(c->g.nc);
    for (int q = 0; q < c->g.nc; ++q) TRY(class_constants(modv, diam[q], rhop[q], sphe[q], &cls[q]));
    c->cls = cls;
    c->modv = modv;
    TRY(upload(c->d_cls, cls.data(), sizeof(met::ClassConst) * c->g.nc, c->stream));
    F3D_HIP(hipStreamSynchronize(c->stream));
    c->classes_set = true;
    return F3D_OK;
}

int f3d_set_models(f3d_ctx *c, int modkv, int modkh, double rkv0, double rkh0, double csh)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (modkv < 0 || modkv > 2 || modkh < 0 || modkh > 2) {
        set_error("f3d_set_models: modkv/modkh out of range (%d, %d)", modkv, modkh);
        return F3D_EINVAL;
    }
    c->modkv = modkv;
    c->modkh = modkh;
    c->rkv0 = rkv0;
    c->rkh0 = rkh0;
    c->csh = csh;
    return F3D_OK;
}

static int meteo_stage(f3d_ctx *c);

int f3d_set_meteo(f3d_ctx *c, const double *vx, const double *vy, const double *vz,
                  const double *tempe, const double *vptemp, const double *rho,
                  const double *ustar, const double *zi, const double *rl, const double *vdep)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!c->classes_set) {
        set_error("f3d_set_meteo: f3d_set_classes must come first");
        return F3D_ESTATE;
    }
    if (!vx || !vy || !vz || !tempe || !rho) {
        set_error("f3d_set_meteo: null field");
        return F3D_EINVAL;
    }
    const Grid &g = c->g;
    hipStream_t s = c->stream;
    const size_t d8 = sizeof(double);
    std::vector<double> zeros;
    auto up_or_zero = [&](DeviceBuffer &b, const double *h, size_t n) -> int {
        if (h) return upload(b, h, n * d8, s);
        F3D_HIP(hipMemsetAsync(b.p, 0, n * d8, s));
        return F3D_OK;
    };
    TRY(up_or_zero(c->d_vx, vx, g.N));
    TRY(up_or_zero(c->d_vy, vy, g.N));
    TRY(up_or_zero(c->d_vz, vz, g.N));
    TRY(up_or_zero(c->d_tempe, tempe, g.N));
    TRY(up_or_zero(c->d_vptemp, vptemp, g.N));
    TRY(up_or_zero(c->d_rho, rho, g.N));
    TRY(up_or_zero(c->d_ustar, ustar, g.NXY));
    TRY(up_or_zero(c->d_zi, zi, g.NXY));
    TRY(up_or_zero(c->d_rl, rl, g.NXY));
    TRY(up_or_zero(c->d_vdep, vdep, g.NXY * g.nc));
    return meteo_stage(c);
}

// kappa3, setvset, Hm1 scaling and finddt on the meteo fields in HBM, then
// the FP32 step fields (meteo.f90:27-149)
static int meteo_stage(f3d_ctx *c)
{
    const Grid &g = c->g;
    hipStream_t s = c->stream;
    // meteo.f90:27 kappa3 (on the unscaled wind)
    TRY(met::kappa3(s, g, c->modkv, c->modkh, c->d_dz.as<double>(), c->d_zlayer.as<double>(),
                    c->d_vx.as<double>(), c->d_vy.as<double>(), c->d_vptemp.as<double>(),
                    c->d_zi.as<double>(), c->d_rl.as<double>(), c->d_ustar.as<double>(),
                    c->d_hm1.as<double>(), c->d_kmh.as<double>(), c->khn_dx, c->khn_dy, c->rkv0, c->rkh0,
                    c->csh, c->lonlat, c->d_rkver.as<double>(), c->d_rkhor.as<double>()));
    // meteo.f90:81 setvset, continuing the SAVEd drag-coefficient chain
    double cd2[2] = {c->cd_state, c->cd_state};
    int st0[2] = {0, 0};
    TRY(upload(c->d_cd, cd2, sizeof cd2, s));
    TRY(upload(c->d_status, st0, sizeof st0, s));
    TRY(met::setvset(s, g, g.nc, c->modv, c->d_cls.as<met::ClassConst>(), c->d_tempe.as<double>(),
                     c->d_rho.as<double>(), c->d_visc.as<double>(), c->d_vset.as<double>(),
                     c->d_cd.as<double>(), c->d_status.as<int>(), 256, 64));
    // meteo.f90:99-100 setvdrydep, on the unscaled density, when the land
    // use is known (otherwise vdep is the caller's)
    if (c->landuse_set)
        TRY(met::setvdrydep(s, g, c->d_cls.as<met::ClassConst>(), c->zlayer[0],
                            c->d_ldu.as<double>(), c->d_znt.as<double>(), c->d_rl.as<double>(),
                            c->d_ustar.as<double>(), c->d_tempe.as<double>(),
                            c->d_rho.as<double>(), c->d_vdep.as<double>()));
    // meteo.f90:108-112 Hm1 scaling
    TRY(met::scale(s, g, c->d_hm1.as<double>(), c->d_vx.as<double>(), c->d_rho.as<double>(),
                   c->d_rkhor.as<double>(), c->d_rkh1.as<double>()));
    // meteo.f90:116-124 corvver (LON-LAT grids; flat terrain leaves vz as it is)
    if (c->lonlat && c->topg_set)
        TRY(met::corvver(s, g, c->d_hm1.as<double>(), c->d_topg.as<double>(), g.dx, g.dy,
                         c->d_vx.as<double>(), c->d_vy.as<double>(), c->d_vz.as<double>()));
    // meteo.f90:138-145 the gravity current's radial wind and setdivu
    TRY(gravity_meteo(c));
    // meteo.f90:149 finddt
    F3D_HIP(hipMemsetAsync(c->d_dtcn.p, 0, sizeof(unsigned long long), s));
    TRY(met::finddt(s, g, c->d_dz.as<double>(), g.dx, g.dy, c->d_vx.as<double>(),
                    c->d_vy.as<double>(), c->d_vz.as<double>(), c->d_rkh1.as<double>(),
                    c->d_rkhor.as<double>(), c->d_rkver.as<double>(), c->d_vset.as<double>(),
                    c->d_dtcn.as<unsigned long long>()));
    unsigned long long bits = 0;
    int status[2] = {0, 0};
    F3D_HIP(hipMemcpyAsync(&bits, c->d_dtcn.p, sizeof bits, hipMemcpyDeviceToHost, s));
    F3D_HIP(hipMemcpyAsync(cd2, c->d_cd.p, sizeof cd2, hipMemcpyDeviceToHost, s));
    F3D_HIP(hipMemcpyAsync(status, c->d_status.p, sizeof status, hipMemcpyDeviceToHost, s));
    F3D_HIP(hipStreamSynchronize(s));
    if (status[0]) {
        set_error("No convergence in vsettl routine");  // setvset.f90:40
        return F3D_ECONV;
    }
    c->cd_state = cd2[1];
    double dtcn;
    std::memcpy(&dtcn, &bits, sizeof dtcn);
    // finddt.f90:250-258
    dtcn = dtcn > 1e-13 ? dtcn : 1e-13;
    c->dtc = 1.0 / dtcn;
    c->dt = 0.9 * c->dtc;
    TRY(rebuild_fields(c));
    c->meteo_ready = true;
    return F3D_OK;
}

int f3d_set_meteo_window(f3d_ctx *c, const f3d_ctx *region, int i0, int j0)
{
    if (!check_ctx(c) || !check_ctx(region)) return F3D_EINVAL;
    if (!c->classes_set) {
        set_error("f3d_set_meteo_window: f3d_set_classes must come first");
        return F3D_ESTATE;
    }
    if (!region->meteo_ready) {
        set_error("f3d_set_meteo_window: the region has no meteo");
        return F3D_ESTATE;
    }
    const Grid &g = c->g, &r = region->g;
    // the region's wind and density are stored after the Hm1 scaling of
    // meteo.f90:108-112, which is the identity only for a uniform Hm1 = 1
    if (!region->uniform_hm1 || !c->uniform_hm1 || region->hm1[0] != 1.0) {
        set_error("f3d_set_meteo_window: needs Hm1 = 1 (UTM grids)");
        return F3D_EINVAL;
    }
    if (i0 < 0 || j0 < 0 || i0 + g.nx > r.nx || j0 + g.ny > r.ny || g.nz > r.nz ||
        g.dx != r.dx || g.dy != r.dy) {
        set_error("f3d_set_meteo_window: window %d+%d x %d+%d x %d does not fit the region %dx%dx%d",
                  i0, g.nx, j0, g.ny, g.nz, r.nx, r.ny, r.nz);
        return F3D_EINVAL;
    }
    for (int k = 0; k < g.nz; ++k)
        if (c->zlayer[k] != region->zlayer[k]) {
            set_error("f3d_set_meteo_window: levels differ from the region's");
            return F3D_EINVAL;
        }
    hipStream_t s = c->stream;
    // the region's own stream may still be writing its fields
    F3D_HIP(hipStreamSynchronize(region->stream));
    auto win3 = [&](DeviceBuffer &d, const DeviceBuffer &src, int nz) {
        return window_copy(s, d.as<double>(), src.as<double>(), g.nx, g.ny, nz, r.nx, r.ny, i0, j0);
    };
    TRY(win3(c->d_vx, region->d_vx, g.nz));
    TRY(win3(c->d_vy, region->d_vy, g.nz));
    TRY(win3(c->d_vz, region->d_vz, g.nz));
    TRY(win3(c->d_tempe, region->d_tempe, g.nz));
    TRY(win3(c->d_vptemp, region->d_vptemp, g.nz));
    TRY(win3(c->d_rho, region->d_rho, g.nz));
    TRY(win3(c->d_ustar, region->d_ustar, 1));
    TRY(win3(c->d_zi, region->d_zi, 1));
    TRY(win3(c->d_rl, region->d_rl, 1));
    // dry deposition velocities are per class: taken when the class sets
    // agree in size, else zero
    if (g.nc == r.nc)
        TRY(win3(c->d_vdep, region->d_vdep, g.nc));
    else
        F3D_HIP(hipMemsetAsync(c->d_vdep.p, 0, sizeof(double) * g.NXY * g.nc, s));
    return meteo_stage(c);
}

int f3d_get_dt(f3d_ctx *c, double *dt, double *dtc)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (dt) *dt = c->dt;
    if (dtc) *dtc = c->dtc;
    return F3D_OK;
}

int f3d_set_dt(f3d_ctx *c, double dt)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!(dt > 0.0)) {
        set_error("f3d_set_dt: dt must be positive");
        return F3D_EINVAL;
    }
    c->dt = dt;
    if (c->meteo_ready) TRY(rebuild_fields(c));
    return F3D_OK;
}

}  // extern "C"

// Concatenate the members' source points and upload them
static int upload_sources(f3d_ctx *c)
{
    const Grid &g = c->g;
    std::vector<int> off(g.nm + 1, 0), I, J, K;
    std::vector<double> T;
    for (int m = 0; m < g.nm; ++m) {
        off[m + 1] = off[m] + (int)c->src_i[m].size();
        I.insert(I.end(), c->src_i[m].begin(), c->src_i[m].end());
        J.insert(J.end(), c->src_j[m].begin(), c->src_j[m].end());
        K.insert(K.end(), c->src_k[m].begin(), c->src_k[m].end());
        T.insert(T.end(), c->src_tm[m].begin(), c->src_tm[m].end());
    }
    const int n = off[g.nm];
    // group each member's points by target cell, keeping the point order
    // inside a group (the order setsrc.f90 adds them in)
    std::vector<int> perm;
    std::vector<int4> grp;
    for (int m = 0; m < g.nm; ++m) {
        std::vector<int> idx(off[m + 1] - off[m]);
        for (size_t a = 0; a < idx.size(); ++a) idx[a] = off[m] + (int)a;
        auto key = [&](int P) {
            return ((long long)K[P] * (g.ny + 2) + J[P]) * (g.nx + 2) + I[P];
        };
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key(a) < key(b); });
        for (size_t a = 0; a < idx.size();) {
            size_t b = a;
            while (b < idx.size() && key(idx[b]) == key(idx[a])) ++b;
            grp.push_back(make_int4(m, (int)perm.size(), (int)(b - a), 0));
            perm.insert(perm.end(), idx.begin() + a, idx.begin() + b);
            a = b;
        }
    }
    long long *B = &c->bytes;
    const size_t nn = n > 0 ? n : 1;
    int rc = 0;
    rc |= c->d_isrc.alloc(sizeof(int) * nn, B);
    rc |= c->d_jsrc.alloc(sizeof(int) * nn, B);
    rc |= c->d_ksrc.alloc(sizeof(int) * nn, B);
    rc |= c->d_tmrat.alloc(sizeof(double) * nn * g.nc, B);
    rc |= c->d_srcperm.alloc(sizeof(int) * nn, B);
    rc |= c->d_srcgrp.alloc(sizeof(int4) * (grp.empty() ? 1 : grp.size()), B);
    if (rc) return F3D_ENOMEM;
    TRY(upload(c->d_srcoff, off.data(), sizeof(int) * (g.nm + 1), c->stream));
    if (n > 0) {
        TRY(upload(c->d_isrc, I.data(), sizeof(int) * n, c->stream));
        TRY(upload(c->d_jsrc, J.data(), sizeof(int) * n, c->stream));
        TRY(upload(c->d_ksrc, K.data(), sizeof(int) * n, c->stream));
        TRY(upload(c->d_tmrat, T.data(), sizeof(double) * n * g.nc, c->stream));
        TRY(upload(c->d_srcperm, perm.data(), sizeof(int) * n, c->stream));
        TRY(upload(c->d_srcgrp, grp.data(), sizeof(int4) * grp.size(), c->stream));
    }
    c->nsrcgrp = (int)grp.size();
    ++c->gen;
    F3D_HIP(hipStreamSynchronize(c->stream));
    c->nsou = n;
    return F3D_OK;
}

static int set_member_source(f3d_ctx *c, int m, int nsou, const int *isrc, const int *jsrc,
                             const int *ksrc, const double *tmrat, double flowrate)
{
    const Grid &g = c->g;
    if (m < 0 || m >= g.nm) {
        set_error("member %d out of range (ensemble of %d)", m, g.nm);
        return F3D_EINVAL;
    }
    if (nsou < 0 || (nsou > 0 && (!isrc || !jsrc || !ksrc || !tmrat))) {
        set_error("f3d_set_source: bad arguments");
        return F3D_EINVAL;
    }
    for (int is = 0; is < nsou; ++is) {
        if (isrc[is] < 1 || isrc[is] > g.nx || jsrc[is] < 1 || jsrc[is] > g.ny) {
            // reasrc.f90:76-80
            set_error("f3d_set_source: source point %d outside the grid", is + 1);
            return F3D_EINVAL;
        }
    }
    c->src_i[m].assign(isrc, isrc + nsou);
    c->src_j[m].assign(jsrc, jsrc + nsou);
    c->src_k[m].assign(ksrc, ksrc + nsou);
    c->src_tm[m].assign(tmrat, tmrat + (size_t)nsou * g.nc);
    c->flowrate[m] = flowrate;
    return F3D_OK;
}

extern "C" {

int f3d_set_source(f3d_ctx *c, int nsou, const int *isrc, const int *jsrc, const int *ksrc,
                   const double *tmrat, double flowrate)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    for (int m = 0; m < c->g.nm; ++m)
        TRY(set_member_source(c, m, nsou, isrc, jsrc, ksrc, tmrat, flowrate));
    return upload_sources(c);
}

int f3d_set_member_source(f3d_ctx *c, int member, int nsou, const int *isrc, const int *jsrc,
                          const int *ksrc, const double *tmrat, double flowrate)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    TRY(set_member_source(c, member, nsou, isrc, jsrc, ksrc, tmrat, flowrate));
    return upload_sources(c);
}

int f3d_step(f3d_ctx *c, int nsteps)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!c->meteo_ready) {
        set_error("f3d_step: no meteo read yet (f3d_set_meteo)");
        return F3D_ESTATE;
    }
    for (int n = 0; n < nsteps; ++n) {
        // Fall3d.f90:112-114
        c->time = c->time + c->dt;
        c->iiter = c->iiter + 1;
        for (int m = 0; m < c->g.nm; ++m) c->erumass[m] = c->erumass[m] + c->flowrate[m] * c->dt;
        TRY(c->precision == F3D_FP64 ? step_fp64(c) : step_fp32(c));
        // endstep.f90:71
        TRY(gravity_endstep(c));
    }
    return F3D_OK;
}

}  // extern "C"

// getc0 + cload + cmass3d for every member; out[6*m + ...]
static int budgets(f3d_ctx *c, double *out)
{
    const Grid &g = c->g;
    hipStream_t s = c->stream;
    double *cdry = c->d_cdry.as<double>(), *cwet = c->d_cwet.as<double>();
    const long dstride = (long)(g.nc + 1) * g.NXY;
    TRY(red::getc0(s, g, cdry, cwet));
    if (c->precision == F3D_FP64)
        TRY(red::cmass3d(s, g, c->d_c64.as<double>(), nullptr, cdry, cwet, dstride,
                         c->d_dz.as<double>(), c->d_scratch.as<double>(), c->d_budget.as<double>()));
    else
        TRY(red::cmass3d(s, g, c->d_c32[c->cur].as<float>(), nullptr, cdry, cwet, dstride,
                         c->d_dz.as<double>(), c->d_scratch.as<double>(), c->d_budget.as<double>()));
    std::vector<double> sums(4 * g.nm), om(g.nm);
    F3D_HIP(hipMemcpyAsync(sums.data(), c->d_budget.p, sizeof(double) * 4 * g.nm,
                           hipMemcpyDeviceToHost, s));
    F3D_HIP(hipMemcpyAsync(om.data(), c->d_outmass.p, sizeof(double) * g.nm, hipMemcpyDeviceToHost,
                           s));
    F3D_HIP(hipStreamSynchronize(s));
    if (out) {
        for (int m = 0; m < g.nm; ++m) {
            for (int k = 0; k < 4; ++k) out[6 * m + k] = sums[4 * m + k] * g.dx * g.dy;  // cmass3d.f90:191-194
            out[6 * m + 4] = om[m];
            out[6 * m + 5] = c->erumass[m];
        }
    }
    return F3D_OK;
}

extern "C" {

int f3d_budget(f3d_ctx *c, double *out)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    std::vector<double> all(6 * c->g.nm);
    TRY(budgets(c, all.data()));
    if (out)
        for (int k = 0; k < 6; ++k) out[k] = all[k];
    return F3D_OK;
}

int f3d_ensemble_budget(f3d_ctx *c, double *out)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    return budgets(c, out);
}

int f3d_get_member_state(f3d_ctx *c, int m, double *ch, double *cdry)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    const Grid &g = c->g;
    if (m < 0 || m >= g.nm) {
        set_error("member %d out of range (ensemble of %d)", m, g.nm);
        return F3D_EINVAL;
    }
    hipStream_t s = c->stream;
    const size_t n = (size_t)g.G * g.nc;
    if (ch) {
        if (c->precision == F3D_FP64) {
            F3D_HIP(hipMemcpyAsync(ch, c->d_c64.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
        } else {
            TRY(fp32_complete_frame(c));
            std::vector<float> tmp(n);
            F3D_HIP(hipMemcpyAsync(tmp.data(), c->d_c32[c->cur].as<float>() + n * m,
                                   n * sizeof(float), hipMemcpyDeviceToHost, s));
            F3D_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < n; ++i) ch[i] = (double)tmp[i];
        }
    }
    const size_t nd = (size_t)g.NXY * (g.nc + 1);
    if (cdry)
        F3D_HIP(hipMemcpyAsync(cdry, c->d_cdry.as<double>() + nd * m, sizeof(double) * nd,
                               hipMemcpyDeviceToHost, s));
    F3D_HIP(hipStreamSynchronize(s));
    return F3D_OK;
}

int f3d_get_state(f3d_ctx *c, double *ch, double *cdry)
{
    return f3d_get_member_state(c, 0, ch, cdry);
}

int f3d_set_state(f3d_ctx *c, const double *ch, const double *cdry)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    const Grid &g = c->g;
    if (g.nm != 1) {
        set_error("f3d_set_state: single-member contexts only");
        return F3D_EINVAL;
    }
    hipStream_t s = c->stream;
    const size_t n = (size_t)g.G * g.nc;
    if (ch) {
        if (c->precision == F3D_FP64) {
            F3D_HIP(hipMemcpyAsync(c->d_c64.p, ch, n * sizeof(double), hipMemcpyHostToDevice, s));
        } else {
            std::vector<float> tmp(n);
            for (size_t i = 0; i < n; ++i) tmp[i] = (float)ch[i];
            // into both buffers: a step writes every frame cell of its output
            // but the edges, which stay as set (the clip of the next state
            // read completes them)
            for (int b = 0; b < 2; ++b)
                F3D_HIP(hipMemcpyAsync(c->d_c32[b].p, tmp.data(), n * sizeof(float),
                                       hipMemcpyHostToDevice, s));
            F3D_HIP(hipStreamSynchronize(s));
            c->frame_dirty = false;
        }
    }
    if (cdry)
        F3D_HIP(hipMemcpyAsync(c->d_cdry.p, cdry, sizeof(double) * g.NXY * (g.nc + 1),
                               hipMemcpyHostToDevice, s));
    F3D_HIP(hipStreamSynchronize(s));
    return F3D_OK;
}

int f3d_set_landuse(f3d_ctx *c, const double *ldu, const double *znt)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!ldu || !znt) {
        set_error("f3d_set_landuse: null field");
        return F3D_EINVAL;
    }
    const Grid &g = c->g;
    const size_t b = sizeof(double) * g.NXY;
    if (!c->d_ldu.p && (c->d_ldu.alloc(b, &c->bytes) || c->d_znt.alloc(b, &c->bytes))) {
        set_error("f3d_set_landuse: out of device memory");
        return F3D_ENOMEM;
    }
    TRY(upload(c->d_ldu, ldu, b, c->stream));
    TRY(upload(c->d_znt, znt, b, c->stream));
    F3D_HIP(hipStreamSynchronize(c->stream));
    c->landuse_set = true;
    return F3D_OK;
}

int f3d_set_topography(f3d_ctx *c, const double *topg)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!topg) {
        set_error("f3d_set_topography: null field");
        return F3D_EINVAL;
    }
    const size_t b = sizeof(double) * c->g.NXY;
    if (!c->d_topg.p && c->d_topg.alloc(b, &c->bytes)) {
        set_error("f3d_set_topography: out of device memory");
        return F3D_ENOMEM;
    }
    TRY(upload(c->d_topg, topg, b, c->stream));
    F3D_HIP(hipStreamSynchronize(c->stream));
    c->topg_set = true;
    return F3D_OK;
}

int f3d_set_precipitation(f3d_ctx *c, const double *prate)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!c->classes_set) {
        set_error("f3d_set_precipitation: f3d_set_classes must come first");
        return F3D_ESTATE;
    }
    if (!prate) {
        set_error("f3d_set_precipitation: null field");
        return F3D_EINVAL;
    }
    const Grid &g = c->g;
    if (!c->d_prate.p && (c->d_prate.alloc(sizeof(double) * g.NXY, &c->bytes) ||
                          c->d_wet.alloc(sizeof(int) * g.nc, &c->bytes) ||
                          c->d_lam.alloc(sizeof(double) * g.NXY, &c->bytes) ||
                          c->d_kwet.alloc(sizeof(int) * g.NXY, &c->bytes))) {
        set_error("f3d_set_precipitation: out of device memory");
        return F3D_ENOMEM;
    }
    // wetdep.f90:33-35: particles above 100 um are not scavenged
    std::vector<int> wet(g.nc);
    for (int q = 0; q < g.nc; ++q) wet[q] = c->diam_h[q] > 100.0 * 1e-6 ? 0 : 1;
    TRY(upload(c->d_prate, prate, sizeof(double) * g.NXY, c->stream));
    TRY(upload(c->d_wet, wet.data(), sizeof(int) * g.nc, c->stream));
    F3D_HIP(hipStreamSynchronize(c->stream));
    c->wet_on = true;
    // the FP32 step's per-column form; rebuilt with the fields on every meteo
    // read and dt change
    if (c->meteo_ready) TRY(rebuild_fields(c));
    return F3D_OK;
}

int f3d_output_products(f3d_ctx *c, int member, double *load, double *thick, double *colmass)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    const Grid &g = c->g;
    if (member < 0 || member >= g.nm) {
        set_error("f3d_output_products: member %d out of range (ensemble of %d)", member, g.nm);
        return F3D_EINVAL;
    }
    if (!c->classes_set) {
        set_error("f3d_output_products: f3d_set_classes must come first");
        return F3D_ESTATE;
    }
    const size_t nl = (size_t)g.NXY * (g.nc + 1);
    if (!c->d_products.p &&
        c->d_products.alloc(sizeof(double) * (nl + 2 * (size_t)g.NXY), &c->bytes)) {
        set_error("f3d_output_products: out of device memory");
        return F3D_ENOMEM;
    }
    hipStream_t s = c->stream;
    double *dl = c->d_products.as<double>(), *dt = dl + nl, *dm = dt + g.NXY;
    const long slots = (long)member * (g.nc + 1) * g.NXY;
    const double *cdry = c->d_cdry.as<double>() + slots, *cwet = c->d_cwet.as<double>() + slots;
    const long slab = (long)member * g.nc * g.G;
    const met::ClassConst *cls = c->d_cls.as<met::ClassConst>();
    if (c->precision == F3D_FP64)
        TRY(red::products(s, g, c->d_c64.as<double>() + slab, cdry, cwet, c->d_hm1.as<double>(), cls,
                          c->d_dz.as<double>(), load ? dl : nullptr, thick ? dt : nullptr,
                          colmass ? dm : nullptr));
    else
        TRY(red::products(s, g, c->d_c32[c->cur].as<float>() + slab, cdry, cwet,
                          c->d_hm1.as<double>(), cls, c->d_dz.as<double>(), load ? dl : nullptr,
                          thick ? dt : nullptr, colmass ? dm : nullptr));
    if (load) F3D_HIP(hipMemcpyAsync(load, dl, sizeof(double) * nl, hipMemcpyDeviceToHost, s));
    if (thick) F3D_HIP(hipMemcpyAsync(thick, dt, sizeof(double) * g.NXY, hipMemcpyDeviceToHost, s));
    if (colmass)
        F3D_HIP(hipMemcpyAsync(colmass, dm, sizeof(double) * g.NXY, hipMemcpyDeviceToHost, s));
    F3D_HIP(hipStreamSynchronize(s));
    return F3D_OK;
}

int f3d_get_vdep(f3d_ctx *c, double *vdep)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (!vdep) {
        set_error("f3d_get_vdep: null output");
        return F3D_EINVAL;
    }
    const Grid &g = c->g;
    F3D_HIP(hipMemcpyAsync(vdep, c->d_vdep.p, sizeof(double) * g.NXY * g.nc,
                           hipMemcpyDeviceToHost, c->stream));
    F3D_HIP(hipStreamSynchronize(c->stream));
    return F3D_OK;
}

int f3d_get_coef(f3d_ctx *c, double *vx, double *rho, double *rkhor, double *rkh1,
                 double *rkver, double *vset)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    const Grid &g = c->g;
    hipStream_t s = c->stream;
    const size_t b = sizeof(double) * g.N;
    if (vx) F3D_HIP(hipMemcpyAsync(vx, c->d_vx.p, b, hipMemcpyDeviceToHost, s));
    if (rho) F3D_HIP(hipMemcpyAsync(rho, c->d_rho.p, b, hipMemcpyDeviceToHost, s));
    if (rkhor) F3D_HIP(hipMemcpyAsync(rkhor, c->d_rkhor.p, b, hipMemcpyDeviceToHost, s));
    if (rkh1) F3D_HIP(hipMemcpyAsync(rkh1, c->d_rkh1.p, b, hipMemcpyDeviceToHost, s));
    if (rkver) F3D_HIP(hipMemcpyAsync(rkver, c->d_rkver.p, b, hipMemcpyDeviceToHost, s));
    if (vset) F3D_HIP(hipMemcpyAsync(vset, c->d_vset.p, b * g.nc, hipMemcpyDeviceToHost, s));
    F3D_HIP(hipStreamSynchronize(s));
    return F3D_OK;
}

int f3d_get_clock(f3d_ctx *c, double *time, int *iiter, double *erumass, double *outmass)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (time) *time = c->time;
    if (iiter) *iiter = c->iiter;
    if (erumass) *erumass = c->erumass[0];
    if (outmass) {
        F3D_HIP(hipMemcpyAsync(outmass, c->d_outmass.p, sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        F3D_HIP(hipStreamSynchronize(c->stream));
    }
    return F3D_OK;
}

int f3d_set_clock(f3d_ctx *c, double time, int iiter, double erumass, double outmass)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    c->time = time;
    c->iiter = iiter;
    c->erumass.assign(c->g.nm, erumass);
    F3D_HIP(hipMemcpyAsync(c->d_outmass.p, &outmass, sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    F3D_HIP(hipStreamSynchronize(c->stream));
    return F3D_OK;
}

int f3d_get_cd_state(f3d_ctx *c, double *cd)
{
    if (!check_ctx(c) || !cd) return F3D_EINVAL;
    *cd = c->cd_state;
    return F3D_OK;
}

int f3d_set_cd_state(f3d_ctx *c, double cd)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    c->cd_state = cd;
    return F3D_OK;
}

int f3d_sync(f3d_ctx *c)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    F3D_HIP(hipStreamSynchronize(c->stream));
    return F3D_OK;
}

int f3d_set_tile(f3d_ctx *c, int tx, int ty) { return f3d_set_tile_config(c, tx, ty, 0, 0); }

int f3d_set_tile_config(f3d_ctx *c, int tx, int ty, int waves, int slabs)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    const f32::TileShape ts{tx, ty, waves, slabs};
    if (!f32::tile_supported(ts)) {
        set_error("f3d_set_tile: unsupported tile %dx%d (waves %d, slabs %d)", tx, ty, waves, slabs);
        return F3D_EINVAL;
    }
    c->tile = ts;
    return F3D_OK;
}

int f3d_timing(f3d_ctx *c, int enable)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (enable && c->ev_start.empty()) {
        c->ev_start.resize(512);
        c->ev_stop.resize(512);
        for (size_t i = 0; i < c->ev_start.size(); ++i) {
            F3D_HIP(hipEventCreate(&c->ev_start[i]));
            F3D_HIP(hipEventCreate(&c->ev_stop[i]));
        }
    }
    if (!enable) flush_timing(c);
    c->timing = enable != 0;
    return F3D_OK;
}

int f3d_kernel_time(f3d_ctx *c, int reset, double *avg_ms, long *count)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    flush_timing(c);
    if (avg_ms) *avg_ms = c->t_count ? c->t_ms / (double)c->t_count : 0.0;
    if (count) *count = c->t_count;
    if (reset) {
        c->t_ms = 0.0;
        c->t_count = 0;
    }
    return F3D_OK;
}

long long f3d_device_bytes(const f3d_ctx *c) { return c ? c->bytes : 0; }

}  // extern "C"
