// Batched and autotuned FP32 steps (the batching layer of the C ABI).
//
// A batch is a set of FP32 contexts -- scenarios of any grid size, class
// count and member count, each with its own meteorology, sources and dt --
// stepped together.  One step of the batch is, on the device:
//   setsrc of every context,
//   one persistent launch over the (scenario, tile, slab group) work list of
//     the interior tiles (and one of the edge tiles on a second stream),
//   the outflux of every context,
// and the host side advances each context's clock as f3d_step does
// (Fall3d.f90:112-114).  The device work of a step is captured once per
// iiter parity in a HIP graph and replayed.  Every work item reads and
// writes only its own scenario's buffers, so a scenario's result in the
// batch is the one it gets stepped alone with the same tile shape.
//
// The autotuner times the flux step of each candidate tile shape on the
// actual scenarios (deposit into scratch, output into the buffer the next
// step overwrites, so no state changes) and keeps the fastest.
#include <algorithm>
#include <cfloat>
#include <vector>

#include "f3d_ctx.h"

using namespace f3d;

struct f3d_batch {
    std::vector<f3d_ctx *> ctx;
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // orders the batch stream after each context's stream and back: a
    // context's own calls (f3d_step, f3d_budget, f3d_get_state, a meteo
    // read) and the batch's steps act on one state in call order
    hipEvent_t order = nullptr;
    f32::TileShape ts{58, 18, 0, 2};
    f32::FluxBatchPlan plan;
    bool planned = false;
    std::vector<unsigned> gen;     // context generations the plan was built for
    // each context's api_seq when the batch stream last waited for its streams
    std::vector<unsigned long> joined;
    std::vector<int> cur_odd;      // state buffer read by an odd step, per context
    std::vector<int> group;        // 1: on the other iiter parity than context 0
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // device work of an even / odd step
    bool use_graph = true;
    // outflux arguments of every context for a step in which context 0 takes
    // an even / odd step, and the (context, member) rows of the outflux launch
    DeviceBuffer fdesc[2], frows, frofs;
    int nfrows = 0, nfblocks = 0;  // outflux rows and their workgroups in all
    // source arguments of every context (even / odd as above) and the first
    // item of each context in the one setsrc launch
    DeviceBuffer sdesc[2], sfirst;
    int nsrc = 0;
    // HIP-event timing of the persistent flux launches (steps launch directly
    // while it is on)
    bool timing = false;
    std::vector<hipEvent_t> ev_start, ev_stop;
    int ev_next = 0;
    double t_ms = 0.0;
    long t_count = 0;
};

namespace {

const f32::TileShape kCandidates[] = {{58, 10, 0, 1}, {58, 10, 0, 2}, {58, 18, 0, 1},
                                      {58, 18, 0, 2}, {58, 26, 0, 1}, {58, 26, 0, 2}};

bool check_batch(const f3d_batch *b)
{
    if (!b) {
        set_error("null batch");
        return false;
    }
    return true;
}

void drop_graphs(f3d_batch *b)
{
    for (auto &g : b->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// every context has its meteorology, no wet deposition (the batched kernel
// has no wet instantiation: a context that set precipitation after joining
// the batch is refused here rather than stepped without it) and no gravity
// current (its divcorr belongs to f3d_step's end of step)
int check_members(const f3d_batch *b)
{
    for (size_t k = 0; k < b->ctx.size(); ++k) {
        const f3d_ctx *c = b->ctx[k];
        if (!c->meteo_ready) {
            set_error("batch: a context has no meteo read yet");
            return F3D_ESTATE;
        }
        if (c->wet_on) {
            set_error("batch: context %zu has wet deposition, which runs through f3d_step only", k);
            return F3D_EINVAL;
        }
        if (c->gravity_on) {
            set_error("batch: context %zu has the gravity current (divcorr at the log times), which "
                      "runs through f3d_step only", k);
            return F3D_EINVAL;
        }
    }
    return F3D_OK;
}

// the batch stream waits for the work queued so far on every context's
// stream (and side stream).  A context whose streams hold nothing but the
// wait for this batch's own earlier work (no API call on it since, no other
// batch stepped it) is skipped unless all is set: a batch of 128 contexts
// would otherwise spend about 3.5 ms of host time per f3d_batch_step call
// on 512 event calls.
int after_members(f3d_batch *b, bool all)
{
    for (size_t k = 0; k < b->ctx.size(); ++k) {
        f3d_ctx *c = b->ctx[k];
        if (!all && c->last_batch == b && b->joined[k] == c->api_seq) continue;
        F3D_HIP(hipEventRecord(b->order, c->stream));
        F3D_HIP(hipStreamWaitEvent(b->stream, b->order, 0));
        if (c->side) {
            F3D_HIP(hipEventRecord(b->order, c->side));
            F3D_HIP(hipStreamWaitEvent(b->stream, b->order, 0));
        }
        b->joined[k] = c->api_seq;
    }
    return F3D_OK;
}

// every context's stream waits for the batch's work queued so far
int members_after(f3d_batch *b)
{
    F3D_HIP(hipEventRecord(b->order, b->stream));
    for (f3d_ctx *c : b->ctx) {
        F3D_HIP(hipStreamWaitEvent(c->stream, b->order, 0));
        c->last_batch = b;
    }
    return F3D_OK;
}

// step arguments of every context for an even and an odd step, reading the
// buffer each will read at that parity; cdry_alt, when given, replaces the
// deposit target (autotuning)
void step_lists(f3d_batch *b, std::vector<f32::StepArgs> &even, std::vector<f32::StepArgs> &odd,
                std::vector<f32::Fields> &f, const std::vector<double *> *cdry_alt)
{
    const size_t n = b->ctx.size();
    even.resize(n);
    odd.resize(n);
    f.resize(n);
    b->cur_odd.resize(n);
    b->group.resize(n);
    const int next0 = (b->ctx[0]->iiter + 1) & 1;
    for (size_t k = 0; k < n; ++k) {
        f3d_ctx *c = b->ctx[k];
        // the next step has parity (iiter+1)&1 and reads buffer cur
        const int next_odd = (c->iiter + 1) & 1;
        const int cur_odd = next_odd ? c->cur : 1 - c->cur;
        b->cur_odd[k] = cur_odd;
        b->group[k] = next_odd != next0;
        odd[k] = fp32_args(c, cur_odd, 1);
        even[k] = fp32_args(c, 1 - cur_odd, 0);
        if (cdry_alt) {  // tuning runs: deposits to scratch, no wet deposition
            odd[k].cdry = even[k].cdry = (*cdry_alt)[k];
            odd[k].wet = even[k].wet = 0;
        }
        f[k] = fp32_fields(c);
    }
}

int replan(f3d_batch *b)
{
    drop_graphs(b);
    std::vector<f32::StepArgs> even, odd;
    std::vector<f32::Fields> f;
    step_lists(b, even, odd, f, nullptr);
    b->planned = false;
    TRY(f32::flux_batch_plan(b->plan, even, odd, f, b->group, b->ts));
    const size_t n = b->ctx.size();
    std::vector<int2> rows;
    for (size_t k = 0; k < n; ++k)
        for (int m = 0; m < b->ctx[k]->g.nm; ++m) rows.push_back(make_int2((int)k, m));
    b->nfrows = (int)rows.size();
    std::vector<int> rofs(rows.size() + 1, 0);
    for (size_t r = 0; r < rows.size(); ++r)
        rofs[r + 1] = rofs[r] + f32::outflux_blocks(b->ctx[rows[r].x]->g);
    b->nfblocks = rofs.back();
    if (b->frows.alloc(sizeof(int2) * rows.size(), nullptr) ||
        b->frofs.alloc(sizeof(int) * rofs.size(), nullptr))
        return F3D_ENOMEM;
    F3D_HIP(hipMemcpy(b->frows.p, rows.data(), sizeof(int2) * rows.size(), hipMemcpyHostToDevice));
    F3D_HIP(hipMemcpy(b->frofs.p, rofs.data(), sizeof(int) * rofs.size(), hipMemcpyHostToDevice));
    std::vector<int> first(n + 1, 0);
    for (size_t k = 0; k < n; ++k) first[k + 1] = first[k] + b->ctx[k]->nsrcgrp * b->ctx[k]->g.nc;
    b->nsrc = first[n];
    if (b->sfirst.alloc(sizeof(int) * (n + 1), nullptr)) return F3D_ENOMEM;
    F3D_HIP(hipMemcpy(b->sfirst.p, first.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    for (int o = 0; o < 2; ++o) {
        std::vector<f32::FrameArgs> fa(n);
        std::vector<f32::SrcArgs> sa(n);
        for (size_t k = 0; k < n; ++k) {
            const bool odd_k = b->group[k] ? !o : (o != 0);
            fa[k] = fp32_frame_args(b->ctx[k], odd_k ? odd[k] : even[k]);
            sa[k] = fp32_src_args(b->ctx[k], odd_k ? odd[k] : even[k]);
        }
        if (b->fdesc[o].alloc(sizeof(f32::FrameArgs) * n, nullptr) ||
            b->sdesc[o].alloc(sizeof(f32::SrcArgs) * n, nullptr))
            return F3D_ENOMEM;
        F3D_HIP(hipMemcpy(b->fdesc[o].p, fa.data(), sizeof(f32::FrameArgs) * n, hipMemcpyHostToDevice));
        F3D_HIP(hipMemcpy(b->sdesc[o].p, sa.data(), sizeof(f32::SrcArgs) * n, hipMemcpyHostToDevice));
    }
    b->gen.resize(b->ctx.size());
    for (size_t k = 0; k < b->ctx.size(); ++k) b->gen[k] = b->ctx[k]->gen;
    b->planned = true;
    return F3D_OK;
}

bool plan_current(const f3d_batch *b)
{
    if (!b->planned) return false;
    const int next0 = (b->ctx[0]->iiter + 1) & 1;
    for (size_t k = 0; k < b->ctx.size(); ++k) {
        const f3d_ctx *c = b->ctx[k];
        if (c->gen != b->gen[k]) return false;
        const int next_odd = (c->iiter + 1) & 1;
        if ((next_odd ? c->cur : 1 - c->cur) != b->cur_odd[k]) return false;
        if ((next_odd != next0) != (b->group[k] != 0)) return false;
    }
    return true;
}

void flush_timing(f3d_batch *b)
{
    if (b->ev_next == 0) return;
    (void)hipEventSynchronize(b->ev_stop[b->ev_next - 1]);
    for (int i = 0; i < b->ev_next; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, b->ev_start[i], b->ev_stop[i]) == hipSuccess) {
            b->t_ms += ms;
            b->t_count += 1;
        }
    }
    b->ev_next = 0;
}

// device work of one step in which context 0 (and its parity group) takes
// parity odd0
int record_step(f3d_batch *b, bool odd0)
{
    hipStream_t s = b->stream;
    const int o = odd0 ? 1 : 0;
    // the sources of every context in one launch, then the flux lists (a
    // persistent flux launch holds every CU until its lists are drawn, so
    // nothing runs well beside it), then the outflux of every context from
    // the ghosts the flux launch left
    TRY(f32::setsrc_batch(s, b->sdesc[o].as<f32::SrcArgs>(), b->sfirst.as<int>(), (int)b->ctx.size(),
                          b->nsrc));
    // the events bracket the persistent flux launches (the side stream's
    // list is joined back before the stop event); only when timing, which
    // launches directly (never inside a capture)
    if (b->timing) F3D_HIP(hipEventRecord(b->ev_start[b->ev_next], s));
    TRY(f32::flux_batch_step(s, b->side, b->fork, b->join, b->plan, odd0));
    if (b->timing) {
        F3D_HIP(hipEventRecord(b->ev_stop[b->ev_next], s));
        b->ev_next += 1;
        if (b->ev_next == (int)b->ev_stop.size()) flush_timing(b);
    }
    TRY(f32::outflux_batch(s, b->fdesc[o].as<f32::FrameArgs>(), b->frows.as<int2>(), b->frofs.as<int>(),
                           b->nfrows, b->nfblocks));
    return F3D_OK;
}

int capture(f3d_batch *b, bool odd)
{
    hipGraph_t g = nullptr;
    F3D_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = record_step(b, odd);
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        set_error("batch: graph capture failed: %s", hipGetErrorString(e));
        return F3D_EHIP;
    }
    const hipError_t ei = hipGraphInstantiate(&b->graph[odd ? 1 : 0], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        b->graph[odd ? 1 : 0] = nullptr;
        set_error("batch: graph instantiation failed: %s", hipGetErrorString(ei));
        return F3D_EHIP;
    }
    return F3D_OK;
}

// average time (ms) of the flux step of the next parity with tile shape ts,
// deposits into scratch; nothing of the contexts' state changes
int time_candidate(f3d_batch *b, f32::TileShape ts, int trials, const std::vector<double *> &scr,
                   double *ms)
{
    std::vector<f32::StepArgs> even, odd;
    std::vector<f32::Fields> f;
    step_lists(b, even, odd, f, &scr);
    f32::FluxBatchPlan plan;
    TRY(f32::flux_batch_plan(plan, even, odd, f, b->group, ts));
    const bool next_odd = ((b->ctx[0]->iiter + 1) & 1) != 0;
    hipEvent_t e0, e1;
    F3D_HIP(hipEventCreate(&e0));
    F3D_HIP(hipEventCreate(&e1));
    int rc = f32::flux_batch_step(b->stream, b->side, b->fork, b->join, plan, next_odd);  // warm-up
    if (!rc) rc = hipEventRecord(e0, b->stream) == hipSuccess ? F3D_OK : F3D_EHIP;
    for (int t = 0; t < trials && !rc; ++t)
        rc = f32::flux_batch_step(b->stream, b->side, b->fork, b->join, plan, next_odd);
    if (!rc) rc = hipEventRecord(e1, b->stream) == hipSuccess ? F3D_OK : F3D_EHIP;
    if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? F3D_OK : F3D_EHIP;
    float t_ms = 0.f;
    if (!rc) (void)hipEventElapsedTime(&t_ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        if (rc == F3D_EHIP) set_error("batch autotune: timing failed");
        return rc;
    }
    *ms = t_ms / std::max(1, trials);
    return F3D_OK;
}

}  // namespace

extern "C" {

int f3d_batch_create(f3d_batch **out, f3d_ctx *const *ctxs, int n)
{
    if (!out || !ctxs || n < 1) {
        set_error("f3d_batch_create: need at least one context");
        return F3D_EINVAL;
    }
    *out = nullptr;
    for (int k = 0; k < n; ++k) {
        if (ctxs[k] && ctxs[k]->wet_on) {
            set_error("f3d_batch_create: context %d has wet deposition, which runs through "
                      "f3d_step only", k);
            return F3D_EINVAL;
        }
        if (ctxs[k] && ctxs[k]->gravity_on) {
            set_error("f3d_batch_create: context %d has the gravity current (divcorr at the log "
                      "times), which runs through f3d_step only", k);
            return F3D_EINVAL;
        }
        if (!ctxs[k] || ctxs[k]->precision != F3D_FP32) {
            set_error("f3d_batch_create: context %d is not an FP32 context", k);
            return F3D_EINVAL;
        }
        for (int j = 0; j < k; ++j)
            if (ctxs[j] == ctxs[k]) {
                set_error("f3d_batch_create: context %d appears twice", k);
                return F3D_EINVAL;
            }
    }
    f3d_batch *b = new f3d_batch;
    b->ctx.assign(ctxs, ctxs + n);
    b->joined.assign(n, ~0UL);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->order, hipEventDisableTiming) != hipSuccess) {
        set_error("f3d_batch_create: hipStreamCreate / hipEventCreate failed");
        f3d_batch_destroy(b);
        return F3D_EHIP;
    }
    // the contexts' own streams may hold pending work (meteo, sources)
    for (f3d_ctx *c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            set_error("f3d_batch_create: a context's stream failed");
            f3d_batch_destroy(b);
            return F3D_EHIP;
        }
    *out = b;
    return F3D_OK;
}

void f3d_batch_destroy(f3d_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->side) (void)hipStreamSynchronize(b->side);
    drop_graphs(b);
    for (hipEvent_t e : b->ev_start) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->ev_stop) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->order) (void)hipEventDestroy(b->order);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int f3d_batch_set_tile(f3d_batch *b, int tx, int ty, int slabs)
{
    if (!check_batch(b)) return F3D_EINVAL;
    const f32::TileShape ts{tx, ty, 0, slabs};
    if (!f32::flux_supported(ts)) {
        set_error("f3d_batch_set_tile: unsupported tile %dx%d (slabs %d)", tx, ty, slabs);
        return F3D_EINVAL;
    }
    b->ts = ts;
    b->planned = false;
    drop_graphs(b);
    return F3D_OK;
}

int f3d_batch_get_tile(const f3d_batch *b, int *tx, int *ty, int *slabs)
{
    if (!check_batch(b)) return F3D_EINVAL;
    if (tx) *tx = b->ts.tx;
    if (ty) *ty = b->ts.ty;
    if (slabs) *slabs = b->ts.mc ? b->ts.mc : 1;
    return F3D_OK;
}

int f3d_batch_use_graph(f3d_batch *b, int on)
{
    if (!check_batch(b)) return F3D_EINVAL;
    b->use_graph = on != 0;
    if (!b->use_graph) drop_graphs(b);
    return F3D_OK;
}

int f3d_batch_step(f3d_batch *b, int nsteps)
{
    if (!check_batch(b)) return F3D_EINVAL;
    if (nsteps < 1) return F3D_OK;
    TRY(check_members(b));
    TRY(after_members(b, false));
    for (int n = 0; n < nsteps; ++n) {
        TRY(check_members(b));
        if (!plan_current(b)) TRY(replan(b));
        // Fall3d.f90:112-114 for every context
        for (f3d_ctx *c : b->ctx) {
            c->time = c->time + c->dt;
            c->iiter = c->iiter + 1;
            for (int m = 0; m < c->g.nm; ++m) c->erumass[m] = c->erumass[m] + c->flowrate[m] * c->dt;
        }
        const bool odd = (b->ctx[0]->iiter & 1) != 0;
        if (b->use_graph && !b->timing) {
            if (!b->graph[odd ? 1 : 0]) TRY(capture(b, odd));
            F3D_HIP(hipGraphLaunch(b->graph[odd ? 1 : 0], b->stream));
        } else {
            TRY(record_step(b, odd));
        }
        for (f3d_ctx *c : b->ctx) {
            c->cur = 1 - c->cur;
            c->frame_dirty = true;
        }
    }
    TRY(members_after(b));
    return F3D_OK;
}

int f3d_batch_timing(f3d_batch *b, int enable)
{
    if (!check_batch(b)) return F3D_EINVAL;
    if (enable && b->ev_start.empty()) {
        b->ev_start.resize(256);
        b->ev_stop.resize(256);
        for (size_t i = 0; i < b->ev_start.size(); ++i) {
            F3D_HIP(hipEventCreate(&b->ev_start[i]));
            F3D_HIP(hipEventCreate(&b->ev_stop[i]));
        }
    }
    if (!enable) flush_timing(b);
    b->timing = enable != 0;
    return F3D_OK;
}

int f3d_batch_kernel_time(f3d_batch *b, int reset, double *avg_ms, long *count)
{
    if (!check_batch(b)) return F3D_EINVAL;
    flush_timing(b);
    if (avg_ms) *avg_ms = b->t_count ? b->t_ms / (double)b->t_count : 0.0;
    if (count) *count = b->t_count;
    if (reset) {
        b->t_ms = 0.0;
        b->t_count = 0;
    }
    return F3D_OK;
}

int f3d_batch_sync(f3d_batch *b)
{
    if (!check_batch(b)) return F3D_EINVAL;
    F3D_HIP(hipStreamSynchronize(b->stream));
    return F3D_OK;
}

int f3d_batch_autotune(f3d_batch *b, int trials, int *tx, int *ty, int *slabs, double *best_ms)
{
    if (!check_batch(b)) return F3D_EINVAL;
    TRY(check_members(b));
    // the tuning runs overwrite the ghost buffers: complete the frames first
    for (f3d_ctx *c : b->ctx) TRY(fp32_complete_frame(c));
    TRY(after_members(b, true));
    std::vector<double *> scr(b->ctx.size(), nullptr);
    int rc = F3D_OK;
    for (size_t k = 0; k < b->ctx.size() && !rc; ++k) {
        const Grid &g = b->ctx[k]->g;
        const size_t nb = sizeof(double) * (size_t)g.NXY * (g.nc + 1) * g.nm;
        if (hipMalloc(&scr[k], nb) != hipSuccess) {
            set_error("f3d_batch_autotune: scratch allocation failed");
            rc = F3D_ENOMEM;
        }
    }
    double best = DBL_MAX;
    f32::TileShape pick = b->ts;
    for (const f32::TileShape &ts : kCandidates) {
        if (rc) break;
        double ms = 0.0;
        rc = time_candidate(b, ts, trials, scr, &ms);
        if (!rc && ms < best) {
            best = ms;
            pick = ts;
        }
    }
    (void)hipStreamSynchronize(b->stream);
    for (double *p : scr)
        if (p) (void)hipFree(p);
    if (rc) return rc;
    b->ts = pick;
    b->planned = false;
    drop_graphs(b);
    if (tx) *tx = pick.tx;
    if (ty) *ty = pick.ty;
    if (slabs) *slabs = pick.mc ? pick.mc : 1;
    if (best_ms) *best_ms = best;
    return F3D_OK;
}

int f3d_autotune(f3d_ctx *c, int trials, int *tx, int *ty, int *slabs, double *best_ms)
{
    if (!check_ctx(c)) return F3D_EINVAL;
    if (c->precision != F3D_FP32 || !c->meteo_ready) {
        set_error("f3d_autotune: needs an FP32 context with meteo read");
        return F3D_ESTATE;
    }
    const Grid &g = c->g;
    // the tuning runs overwrite the ghost buffer: complete the frame first
    TRY(fp32_complete_frame(c));
    double *scr = nullptr;
    if (hipMalloc(&scr, sizeof(double) * (size_t)g.NXY * (g.nc + 1) * g.nm) != hipSuccess) {
        set_error("f3d_autotune: scratch allocation failed");
        return F3D_ENOMEM;
    }
    // the next step's arguments, deposit into scratch: cout is the buffer the
    // next step overwrites
    f32::StepArgs a = fp32_args(c, c->cur, c->iiter + 1);
    a.cdry = scr;
    a.wet = 0;  // tuning leaves cwet alone
    const f32::Fields f = fp32_fields(c);
    std::vector<f32::TileShape> cand(std::begin(kCandidates), std::end(kCandidates));
    double best = DBL_MAX;
    f32::TileShape pick = c->tile;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) ? F3D_OK : F3D_EHIP;
    for (const f32::TileShape &ts : cand) {
        if (rc) break;
        rc = f32::fused_step(c->stream, a, f, ts);  // warm-up
        if (!rc) rc = hipEventRecord(e0, c->stream) == hipSuccess ? F3D_OK : F3D_EHIP;
        for (int t = 0; t < trials && !rc; ++t) rc = f32::fused_step(c->stream, a, f, ts);
        if (!rc) rc = hipEventRecord(e1, c->stream) == hipSuccess ? F3D_OK : F3D_EHIP;
        if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? F3D_OK : F3D_EHIP;
        float ms = 0.f;
        if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
        const double avg = ms / std::max(1, trials);
        if (!rc && avg < best) {
            best = avg;
            pick = ts;
        }
    }
    (void)hipStreamSynchronize(c->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(scr);
    if (rc) {
        if (rc == F3D_EHIP) set_error("f3d_autotune: timing failed");
        return rc;
    }
    c->tile = pick;
    if (tx) *tx = pick.tx;
    if (ty) *ty = pick.ty;
    if (slabs) *slabs = pick.mc ? pick.mc : 1;
    if (best_ms) *best_ms = best;
    return F3D_OK;
}

}  // extern "C"
