"""Benchmark of the FALL3D transport step on one MI355X.

The headline is the batched form the target is stated for: the uneven
batch of 128 scenarios (grids 80..320 squared x 16..48 levels, 4..24
classes, each with its own seeded Fourier-mode meteorology, source and dt),
stepped by one persistent launch per step replayed from HIP graphs with the
autotuned tile.  A step advances every scenario by one time step: setsrc +
sizloop (advection, diffusion, boundaries, deposit, clip, outflux).

value     = cell*class updates/s of the batched FP32 step (inputs resident in
            HBM, meteo stage excluded)
roofline  = the batched flux launches' algorithmic HBM bytes per step over
            their HIP-event average time, against 8 TB/s
fp64_plain= the routine-by-routine FP64 drop-in over the same scenarios
cpu_baseline = the oracle's FP64 C restatement (class-parallel threads) on
            this host over a bounded sample of the same scenarios; beside it
            the reference's own Sources/*.f90 routines (amdflang -O2,
            oracle/_ref) on one core
errors    = relative L1/Linf of concentration and ground load and the mass
            residual rho_M against the FP64 CPU oracle on sampled scenarios

The other configurations are reported beside it, each with its rate, its
FP64-plain ratio, its roofline and its errors against the oracle:
plain     = 60x60x20 x 6 classes, log-profile wind, 6 h (the CPU-runnable
            case): every path over the whole 6 h, with the reference's own
            routines on one core beside the C restatement (calibration)
precision = 400x400x40 x 20 classes, Fourier-mode meteo, single run
ensemble  = 64 members x 200x200x30 x 16 classes on one meteorology
serving   = Poisson arrivals of 6 h forecasts on windows of three resident
            400x400x40 regions, continuous batching: p50/p99 latency and
            throughput over an arrival window (30 s by default)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def alg_bytes_per_update(nc, n_share=1, s_c=4, s_v=4, s_m=4):
    """Compulsory HBM bytes per cell*class update of one fused step: read and
    write c, read vz+vset of the class, read the 7 shared meteo/coefficient
    fields (vx, vy, vz, 1/rho, rho*rkh1, rho*rkhor, rho*rkver) once per tile
    for all nc classes (and all n_share members on one meteorology)."""
    return 2 * s_c + (s_v + 7 * s_m / nc) / n_share


def _threads(nc):
    return max(1, min(len(os.sched_getaffinity(0)), nc, 16))


def _gpu_result(r):
    """(interior c, ground load, budget) of a Run."""
    b = r.budget()
    c, cdry = r.state()
    return c[1:-1, 1:-1, 1:-1], cdry[:, :, 1:].sum(-1), b


def _oracle_result(o):
    b = o.budget()
    return o.c[1:-1, 1:-1, 1:-1], o.cdry[:, :, 1:].sum(-1), b


def _errors(got, ref):
    from fall3d_amd import metrics as M
    return M.compare(got[1], ref[1], got[2], ref[2], got[0], ref[0])


def _fp32_run(sc, nsteps, tile=None):
    from fall3d_amd import lib as F
    r = F.Run(sc, precision=F.FP32, tile=tile)
    try:
        r.meteo_update()
        r.step(nsteps)
        return _gpu_result(r)
    finally:
        r.close()


def _oracle_run(sc, nsteps):
    from oracle.oracle import OracleRun
    o = OracleRun(sc, nthreads=_threads(sc.nc))
    o.meteo_update()
    for _ in range(nsteps):
        o.step()
    return _oracle_result(o)


def _fp64_rate(sc, nsteps=2):
    """Updates/s of the FP64 drop-in path on one scenario (after one step)."""
    from fall3d_amd import lib as F
    r = F.Run(sc, precision=F.FP64)
    try:
        r.meteo_update()
        r.step(1)
        r.ctx.sync()
        t0 = time.perf_counter()
        r.step(nsteps)
        r.ctx.sync()
        return sc.updates_per_step * nsteps, time.perf_counter() - t0
    finally:
        r.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed batched steps of the headline (uneven batch)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--uneven-n", type=int, default=128)
    ap.add_argument("--precision-steps", type=int, default=200)
    ap.add_argument("--fp64-steps", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--ref-seconds", type=float, default=10.0,
                    help="step time of the one-core reference legs (0: skip them)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-precision", action="store_true")
    ap.add_argument("--skip-ensemble", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=64)
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="arrival window of the serving run (2 req/s)")
    ap.add_argument("--serve-rate", type=float, default=2.0)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; --gpus above 1 is not supported (the path "
              "would shard (member, class) slabs across GPUs, with one FP64 all-reduce of "
              "each member's outflux)", file=sys.stderr)
        sys.exit(2)

    copy_gbs = copy_calibration()
    head = bench_uneven(args)
    out = {
        "metric": "cell·class updates/s + % HBM peak; ensemble p50/p99 forecast latency",
        "value": head["batched_updates_per_s"],
        "unit": "cell*class updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["batched_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 deposit/budgets/settling/CFL",
        "data": "synthetic (seeded Fourier-mode meteorology, Suzuki sources)",
        "config": {"workload": head["workload"], "scenarios": head["scenarios"],
                   "updates_per_step": head["updates_per_step"], "tile": head["autotuned_tile"],
                   "graphs": True},
        "roofline": head["roofline"],
        "fp64_plain_updates_per_s": head["fp64_plain_updates_per_s"],
        "speedup_vs_fp64_plain": head["speedup_vs_fp64_plain"],
        "cpu_baseline": head.get("cpu_baseline"),
        "errors": head["errors"],
        "per_scenario_launch_updates_per_s": head["per_scenario_updates_per_s"],
        "hbm_copy_gbs": copy_gbs,
    }
    if not args.skip_precision:
        out["precision"] = bench_precision(args)
    if not args.skip_plain:
        out["plain"] = bench_plain(args)
    if not args.skip_ensemble:
        out["ensemble"] = bench_ensemble(args)
    if not args.skip_serve:
        out["serving"] = bench_serving(args)
    print(json.dumps(out), flush=True)


# ---------------- headline: the uneven batch ----------------

def _uneven_scenario(seed):
    from fall3d_amd import scenario as S
    return S.config_uneven(seed)


def bench_uneven(args):
    """128 scenarios of their own size, class count, meteorology, source and
    dt: the persistent batched step (autotuned tile, HIP graphs), beside
    per-scenario f3d_step launches of the same kernels and the FP64 drop-in
    over the same scenarios."""
    import concurrent.futures as cf

    from fall3d_amd import lib as F
    t0 = time.time()
    nproc = max(1, min(16, len(os.sched_getaffinity(0))))
    with cf.ProcessPoolExecutor(nproc) as ex:
        scs = list(ex.map(_uneven_scenario, range(1000, 1000 + args.uneven_n)))
    upd = sum(sc.updates_per_step for sc in scs)
    log(f"uneven: {len(scs)} scenarios built in {time.time() - t0:.1f}s, {upd:.3e} updates/step")
    nsteps = args.steps
    runs = []
    try:
        for sc in scs:
            r = F.Run(sc, precision=F.FP32)
            r.meteo_update()
            runs.append(r)
        # per-scenario launches (f3d_step on each context in turn)
        for r in runs:
            r.step(1)
        for r in runs:
            r.ctx.sync()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            for r in runs:
                r.step(1)
        for r in runs:
            r.ctx.sync()
        el_plain = time.perf_counter() - t0
        # batched: autotune, then graph-replayed persistent steps
        b = F.Batch([r.ctx for r in runs])
        tile, tune_ms = b.autotune(trials=2)
        b.step(args.warmup)
        b.sync()
        t0 = time.perf_counter()
        b.step(nsteps)
        b.sync()
        el_batch = time.perf_counter() - t0
        # the persistent flux launches of the same steps, HIP events around
        # them (direct launches while timing)
        b.timing(True)
        b.kernel_time(reset=True)
        b.step(nsteps)
        b.sync()
        kms, kcount = b.kernel_time()
        b.timing(False)
        b.close()
        steps_done = 1 + nsteps + args.warmup + 2 * nsteps
        # sampled scenarios against the oracle: the smallest and the median
        order = sorted(range(len(scs)), key=lambda k: scs[k].updates_per_step)
        sample = [order[0], order[len(order) // 2]]
        got = {k: _gpu_result(runs[k]) for k in sample}
    finally:
        for r in runs:
            r.close()
    v_batch, v_plain = upd * nsteps / el_batch, upd * nsteps / el_plain
    b_alg = sum(sc.updates_per_step * alg_bytes_per_update(sc.nc) for sc in scs) / upd
    ach = b_alg * upd / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    log(f"uneven: batched {el_batch * 1e3 / nsteps:.2f} ms/step ({v_batch:.3e} updates/s), "
        f"flux launches {kms:.2f} ms, per-scenario {el_plain * 1e3 / nsteps:.2f} ms/step "
        f"({v_plain:.3e}), tile {tile}")
    errors = []
    for k in sample:
        ref = _oracle_run(scs[k], steps_done)
        e = _errors(got[k], ref)
        e.update(scenario=scs[k].name, steps=steps_done)
        errors.append(e)
        log(f"uneven: {scs[k].name} after {steps_done} steps vs oracle: load L1 "
            f"{e['load_rel_l1']:.2e}, c L1 {e['c_rel_l1']:.2e}, rho_M {e['rho_M']:.2e}")
    # the FP64 drop-in over the same scenarios, one context at a time
    n64, t64 = 0, 0.0
    for sc in scs:
        n, t = _fp64_rate(sc, 1)
        n64 += n
        t64 += t
    fp64 = n64 / t64
    log(f"uneven: fp64 plain {fp64:.3e} updates/s over the same scenarios")
    cpu = None
    if not args.skip_cpu:
        cpu = uneven_cpu_baseline(scs, args)
    return {"workload": f"uneven batch: {len(scs)} scenarios, nx=ny 80..320, nz 16..48, "
                        f"nc 4..24, each with its own meteorology, source and dt; one "
                        f"persistent launch per step, HIP-graph replay, autotuned tile",
            "scenarios": len(scs), "updates_per_step": upd,
            "batched_updates_per_s": v_batch, "batched_ms_per_step": el_batch * 1e3 / nsteps,
            "per_scenario_updates_per_s": v_plain,
            "autotuned_tile": "%dx%d slabs %d" % tile, "autotune_best_flux_ms": tune_ms,
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS,
                         "traffic": measured_traffic("uneven128"),
                         "bytes_per_update": b_alg, "kernel_avg_ms": kms, "launches": kcount,
                         "kernel": "k_flux_batch (persistent flux launches of one step)"},
            "fp64_plain_updates_per_s": fp64, "speedup_vs_fp64_plain": v_batch / fp64,
            "errors": errors, "cpu_baseline": cpu}


def _oracle_on_device_coef(sc):
    """An OracleRun of sc holding the device meteo stage's coefficients (equal
    to the oracle's own within 1e-11) and its dt, so that only the step
    phase is run on the host."""
    import numpy as np

    from fall3d_amd import lib as F
    from oracle.oracle import OracleRun
    r64 = F.Run(sc, precision=F.FP64)
    try:
        r64.meteo_update()
        coef = r64.ctx.coef()
    finally:
        r64.close()
    o = OracleRun(sc, nthreads=_threads(sc.nc))
    o.vx, o.rho, o.rkhor, o.rkh1, o.rkver, o.vset = (coef[k] for k in
                                                      ("vx", "rho", "rkhor", "rkh1", "rkver", "vset"))
    o.vy, o.vz, o.vdep = (np.asfortranarray(a) for a in (sc.vy, sc.vz, sc.vdep))
    o.dt = r_dt(sc, coef)
    return o, coef


def uneven_cpu_baseline(scs, args):
    """The oracle over a sample of the uneven batch's own scenarios (every
    16th of the seeded draw) on the device meteo stage's coefficients, each
    stepped until its equal share of about cpu_seconds of step time (at
    least one step, at most 20); the reference's routines on one core over
    the first of them."""
    sample = scs[::16]
    share = args.cpu_seconds / len(sample)
    nupd, tcpu, nsteps, first = 0, 0.0, 0, None
    for sc in sample:
        o, coef = _oracle_on_device_coef(sc)
        if first is None:
            first = (sc, coef, o.dt)
        t_sc, n_sc = 0.0, 0
        while n_sc == 0 or (t_sc < share and n_sc < 20):
            t0 = time.perf_counter()
            o.step()
            t_sc += time.perf_counter() - t0
            n_sc += 1
        tcpu += t_sc
        nsteps += n_sc
        nupd += sc.updates_per_step * n_sc
        del o, coef
    nthreads = max(_threads(sc.nc) for sc in sample)
    log(f"uneven: cpu oracle {nupd / tcpu:.3e} updates/s, {nsteps} steps of {len(sample)} "
        f"scenarios ({tcpu:.1f}s)")
    cpu = dict(value=nupd / tcpu, unit="cell*class updates/s", cores=nthreads, kind="port",
               sample=f"{nsteps} steps over {len(sample)} scenarios of the uneven batch (every "
                      f"16th of the seeded draw, {tcpu:.1f} s of step time; FP64 C "
                      f"restatement, up to {nthreads} class-parallel threads, coefficients "
                      f"from the device meteo stage)")
    if args.ref_seconds > 0:
        cpu["reference_1core"] = ref_cpu_baseline(first[0], args.ref_seconds, first[1], first[2])
    return cpu


# ---------------- the precision-form single run ----------------

def bench_precision(args):
    """400x400x40, 20 classes: the fused FP32 step with its HIP-event
    roofline, the FP64 drop-in, the oracle on this host and the reference's
    routines on one core, and the FP32 and FP64 states against the oracle's
    after the oracle's steps."""
    import numpy as np

    from fall3d_amd import lib as F
    from fall3d_amd import scenario as S
    sc = S.config_c2()
    upd = sc.updates_per_step
    r = F.Run(sc, precision=F.FP32)
    try:
        t0 = time.time()
        r.meteo_update()
        r.ctx.sync()
        meteo_s = time.time() - t0
        r.step(10)
        r.ctx.sync()
        # the library's tile autotuner on this context (no state change)
        (tx, ty, slabs), tune_ms = r.ctx.autotune(trials=3)
        tile_key = f"{tx}x{ty}m{slabs}"
        r.ctx.timing(True)
        r.ctx.kernel_time(reset=True)
        t0 = time.perf_counter()
        r.step(args.precision_steps)
        r.ctx.sync()
        el = time.perf_counter() - t0
        kms, kcount = r.ctx.kernel_time()
        r.ctx.timing(False)
    finally:
        r.close()
    value = upd * args.precision_steps / el
    b_alg = alg_bytes_per_update(sc.nc)
    ach = b_alg * upd / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    log(f"precision: fp32 {el * 1e3 / args.precision_steps:.3f} ms/step, {value:.3e} updates/s, "
        f"flux launch {kms:.3f} ms, tile {tx}x{ty}/{slabs}")
    n64, t64 = _fp64_rate(sc, args.fp64_steps)
    fp64 = n64 / t64
    log(f"precision: fp64 plain {t64 * 1e3 / args.fp64_steps:.3f} ms/step, {fp64:.3e} updates/s")
    res = {"workload": f"{sc.nx}x{sc.ny}x{sc.nz} grid, {sc.nc} classes, single run, fused FP32 "
                       f"step", "updates_per_s": value,
           "ms_per_step": el * 1e3 / args.precision_steps, "meteo_stage_s": meteo_s,
           "autotuned_tile": f"{tx}x{ty} slabs {slabs}", "autotune_best_flux_ms": tune_ms,
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(tile_key),
                        "bytes_per_update": b_alg, "kernel_avg_ms": kms, "launches": kcount,
                        "kernel": "k_flux (interior and edge launches of one step)"},
           "fp64_plain_updates_per_s": fp64, "speedup_vs_fp64_plain": value / fp64}
    if args.skip_cpu:
        return res
    o, coef = _oracle_on_device_coef(sc)
    o_dt = o.dt
    nsteps, tcpu = 0, 0.0
    while tcpu < args.cpu_seconds and nsteps < 50:
        t0 = time.perf_counter()
        o.step()
        tcpu += time.perf_counter() - t0
        nsteps += 1
    nthreads = _threads(sc.nc)
    res["cpu_baseline"] = dict(
        value=upd * nsteps / tcpu, unit="cell*class updates/s", cores=nthreads, kind="port",
        sample=f"{nsteps} steps of the {sc.name} workload (FP64 C restatement, {nthreads} "
               f"class-parallel threads, coefficients from the device meteo stage)")
    log(f"precision: cpu oracle {upd * nsteps / tcpu:.3e} updates/s ({nsteps} steps)")
    ref = _oracle_result(o)
    del o
    # the GPU paths over the same steps against the oracle's state
    g32 = _fp32_run(sc, nsteps)
    r64 = F.Run(sc, precision=F.FP64)
    try:
        r64.meteo_update()
        r64.step(nsteps)
        g64 = _gpu_result(r64)
    finally:
        r64.close()
    res["errors"] = {"steps": nsteps, "fp32": _errors(g32, ref), "fp64": _errors(g64, ref)}
    log(f"precision: after {nsteps} steps vs oracle: fp32 load L1 "
        f"{res['errors']['fp32']['load_rel_l1']:.2e}, rho_M {res['errors']['fp32']['rho_M']:.2e}; "
        f"fp64 rho_M {res['errors']['fp64']['rho_M']:.2e}")
    del g32, g64, ref
    if args.ref_seconds > 0:
        res["cpu_baseline"]["reference_1core"] = ref_cpu_baseline(sc, args.ref_seconds, coef,
                                                                  o_dt)
    return res


# ---------------- the plain configuration, whole 6 h ----------------

def bench_plain(args):
    """60x60x20, 6 classes, log-profile wind, 6 h: the FP32 and FP64 GPU
    paths, the oracle (one thread and class-parallel) and the reference's own
    routines on one core, each over the whole run; errors at 6 h against the
    oracle, and the reference against the oracle (the calibration of the
    restatement used as the CPU baseline elsewhere)."""
    from fall3d_amd import lib as F
    from fall3d_amd import scenario as S
    from oracle.oracle import OracleRun, RefRun
    sc = S.config_c1()
    res = {"workload": f"{sc.nx}x{sc.ny}x{sc.nz} grid, {sc.nc} classes, log-profile wind, "
                       f"Suzuki column of 10 km, MER 1e6 kg/s, 6 h"}
    got = {}
    for name, prec in (("fp32", F.FP32), ("fp64", F.FP64)):
        r = F.Run(sc, precision=prec)
        try:
            r.meteo_update()
            r.ctx.sync()
            t0 = time.perf_counter()
            n = r.run_to(sc.end_time)
            r.ctx.sync()
            el = time.perf_counter() - t0
            got[name] = _gpu_result(r)
        finally:
            r.close()
        res[f"{name}_updates_per_s"] = sc.updates_per_step * n / el
        res["steps"] = n
    res["fp32_vs_fp64_plain"] = res["fp32_updates_per_s"] / res["fp64_updates_per_s"]
    oracle = {}
    for nth in (1, _threads(sc.nc)):
        o = OracleRun(sc, nthreads=nth)
        o.meteo_update()
        t0 = time.perf_counter()
        n = o.run(end_time=sc.end_time)
        oracle[nth] = (sc.updates_per_step * n / (time.perf_counter() - t0), o)
    ref = _oracle_result(oracle[1][1])
    res["oracle_1thread_updates_per_s"] = oracle[1][0]
    res["oracle_threads"] = {"threads": _threads(sc.nc),
                             "updates_per_s": oracle[_threads(sc.nc)][0]}
    res["errors"] = {"fp32": _errors(got["fp32"], ref), "fp64": _errors(got["fp64"], ref)}
    log(f"plain: fp32 {res['fp32_updates_per_s']:.3e}, fp64 {res['fp64_updates_per_s']:.3e}, "
        f"oracle 1 thread {oracle[1][0]:.3e} updates/s; fp32 load L1 "
        f"{res['errors']['fp32']['load_rel_l1']:.2e}, fp64 rho_M {res['errors']['fp64']['rho_M']:.2e}")
    if args.ref_seconds > 0 and RefRun.available():
        rr = RefRun(sc)
        rr.meteo_update()
        t0 = time.perf_counter()
        n = rr.run(end_time=sc.end_time)
        tref = time.perf_counter() - t0
        c, cdry = rr.state()
        b = rr.budget()
        res["reference_1core_updates_per_s"] = sc.updates_per_step * n / tref
        # the calibration of the restatement used as the CPU baseline
        res["oracle_1thread_over_reference_1core"] = oracle[1][0] / (sc.updates_per_step * n / tref)
        res["reference_vs_oracle"] = _errors((c[1:-1, 1:-1, 1:-1], cdry[:, :, 1:].sum(-1), b), ref)
        log(f"plain: reference one core {res['reference_1core_updates_per_s']:.3e} updates/s; "
            f"reference vs oracle c Linf {res['reference_vs_oracle']['c_rel_linf']:.2e}")
    return res


# ---------------- the ensemble ----------------

def bench_ensemble(args, nsteps=10):
    """64 members on one 200x200x30 grid, 16 classes: one fused launch per
    step for all members (FP32 state, FP64 deposits and budgets); member 0
    against the oracle; the FP64 drop-in on the member grid."""
    from fall3d_amd import lib as F
    from fall3d_amd import scenario as S
    base = S.config_c2(nx=200, ny=200, nz=30, nc=16, seed=0xF3D0003)
    mem = [S.config_ensemble_member(base, 5000 + k) for k in range(args.ensemble_members)]
    ens = F.Ensemble(base, mem)
    try:
        ens.meteo_update()
        # the library's tile autotuner on this context (no state change)
        (tx, ty, slabs), tune_ms = ens.ctx.autotune(trials=2)
        ens.step(2)
        ens.ctx.sync()
        ens.ctx.timing(True)
        ens.ctx.kernel_time(reset=True)
        t0 = time.perf_counter()
        ens.step(nsteps)
        ens.ctx.sync()
        el = time.perf_counter() - t0
        kms, _ = ens.ctx.kernel_time()
        ens.ctx.timing(False)
        c, cdry = ens.state(0)
        g0 = (c[1:-1, 1:-1, 1:-1], cdry[:, :, 1:].sum(-1), ens.budgets()[0])
    finally:
        ens.close()
    upd = base.nx * base.ny * base.nz * base.nc * len(mem)
    value = upd * nsteps / el
    b_alg = alg_bytes_per_update(base.nc, n_share=len(mem))
    ach = b_alg * upd / (kms * 1e-3) / 1e9
    e0 = _errors(g0, _oracle_run(mem[0], nsteps + 2))
    n64, t64 = _fp64_rate(mem[0], 2)
    fp64 = n64 / t64
    # the same members as separate single-member contexts, each with its own
    # copy of the meteorology, stepped together by one batch with the same
    # tile: what sharing the meteorology across the members is worth
    sep = _separate_members_rate(mem, (tx, ty, slabs), nsteps)
    log(f"ensemble: {len(mem)} members, {el * 1e3 / nsteps:.2f} ms/step, {value:.3e} updates/s, "
        f"tile {tx}x{ty}/{slabs}; as separate contexts {sep['ms_per_step']:.2f} ms/step; "
        f"member 0 load L1 {e0['load_rel_l1']:.2e}")
    return {"workload": f"{len(mem)} members x {base.nx}x{base.ny}x{base.nz} x {base.nc} classes",
            "autotuned_tile": f"{tx}x{ty} slabs {slabs}", "updates_per_s": value,
            "ms_per_step": el * 1e3 / nsteps,
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "bytes_per_update": b_alg,
                         "kernel_avg_ms": kms, "kernel": "k_flux over all member*class slabs"},
            "fp64_plain_updates_per_s": fp64, "speedup_vs_fp64_plain": value / fp64,
            "separate_members": sep, "shared_meteo_speedup": value / sep["updates_per_s"],
            "errors": {"member": 0, "steps": nsteps + 2, **e0}}


def _separate_members_rate(mem, tile, nsteps):
    """The ensemble's members as len(mem) single-member FP32 contexts, each
    with its own meteo read, batched with the given tile (graph replay)."""
    from fall3d_amd import lib as F
    runs = [F.Run(m, precision=F.FP32) for m in mem]
    b = None
    try:
        for r in runs:
            r.meteo_update()
        b = F.Batch([r.ctx for r in runs])
        b.set_tile(tile[0], tile[1], tile[2])
        b.step(2)
        b.sync()
        t0 = time.perf_counter()
        b.step(nsteps)
        b.sync()
        el = time.perf_counter() - t0
    finally:
        if b is not None:
            b.close()
        for r in runs:
            r.close()
    upd = sum(m.nx * m.ny * m.nz * m.nc for m in mem)
    return {"contexts": len(mem), "updates_per_s": upd * nsteps / el,
            "ms_per_step": el * 1e3 / nsteps,
            "form": "one batch of single-member contexts (per-member meteorology), same tile"}


# ---------------- serving ----------------

def bench_serving(args):
    """Continuous-batching serving of 6 h forecasts on windows of three
    resident 400x400x40 regions over an arrival window; latency from arrival
    to final ground load.  A sampled request is re-run by the FP64 drop-in
    (equal to the oracle within 1e-11, tests/test_gpu_horizon.py)."""
    import copy

    import numpy as np

    from fall3d_amd import lib as F
    from fall3d_amd import scenario as S
    from fall3d_amd import serve
    t0 = time.time()
    regions = [serve.Region(S.config_c2(nx=400, ny=400, nz=40, nc=16, seed=0xF3D0500 + k))
               for k in range(3)]
    rng = np.random.default_rng(0xF3D0005)
    times = serve.poisson_arrivals(args.serve_rate, args.serve_seconds, rng)
    reqs = [serve.make_request(k, t, regions, rng) for k, t in enumerate(times)]
    sample = copy.deepcopy(reqs[len(reqs) // 2]) if reqs else None
    log(f"serving: {len(reqs)} requests over {args.serve_seconds:.0f}s, setup {time.time() - t0:.1f}s")
    srv = serve.Server(regions)
    nfin = [0]

    def progress(msg):
        # every 50th finished request: a long arrival window stays visibly alive
        nfin[0] += 1
        if nfin[0] % 50 == 0:
            log(f"serving: {msg}")

    done = srv.run(reqs, log=progress)
    st = serve.latency_stats(done)
    thr = srv.updates / srv.wall_s
    log(f"serving: p50 {st['p50']:.2f}s p99 {st['p99']:.2f}s, {thr:.3e} updates/s over "
        f"{srv.wall_s:.1f}s (GPU busy {srv.busy_s:.1f}s)")
    err = None
    if sample is not None:
        got = next(r for r in done if r.rid == sample.rid)
        reg = regions[sample.region]
        c64 = F.Context(sample.nx, sample.ny, sample.nz, sample.nc, reg.sc.dx, reg.sc.dy,
                        reg.sc.zlayer[:sample.nz], F.FP64)
        try:
            c64.set_classes(sample.diam, sample.rhop, sample.sphe, reg.sc.modv)
            c64.set_models(reg.sc.modkv, reg.sc.modkh, reg.sc.rkv0, reg.sc.rkh0, reg.sc.csh)
            c64.set_meteo_window(reg.ctx, sample.i0, sample.j0)
            c64.set_source(sample.isrc, sample.jsrc, sample.ksrc, sample.tmrat, sample.flowrate)
            c64.step(got.nsteps)
            ref = c64.ground_load()
        finally:
            c64.close()
        a, b = got.cdry[:, :, 1:].sum(-1), ref[:, :, 1:].sum(-1)
        err = {"request": sample.rid, "steps": got.nsteps, "against": "FP64 drop-in",
               "load_rel_l1": float(np.abs(a - b).sum() / np.abs(b).sum()),
               "load_rel_linf": float(np.abs(a - b).max() / np.abs(b).max())}
    window = ("the configuration's own window" if args.serve_seconds >= 600 else
              "a window of the configuration's 600 s, whose full run is in profiles/")
    return {"workload": f"Poisson {args.serve_rate} req/s over a {args.serve_seconds:.0f} s "
                        f"arrival window ({window}), 6 h forecasts "
                        f"of 200x200x30 x 16 classes on windows of 3 resident 400x400x40 regions",
            "requests": st["n"], "latency_p50_s": st["p50"], "latency_p99_s": st["p99"],
            "latency_max_s": st["max"], "throughput_updates_per_s": thr,
            "busy_updates_per_s": srv.updates / srv.busy_s if srv.busy_s else None,
            "gpu_busy_frac": srv.busy_s / srv.wall_s, "wall_s": srv.wall_s,
            "fp64_check": err}


# ---------------- helpers ----------------

def ref_cpu_baseline(sc, seconds, coef, dt):
    """The reference's own routines (Sources/*.f90 compiled by amdflang -O2
    into oracle/_ref by oracle/Makefile), serial on one core as one process of
    the reference runs, on the coefficients and dt of the device meteo stage
    (the reference's meteo stage is not run: its SAVEd drag coefficient,
    setvset.f90:89, stays untouched for the whole-run leg): steps until
    `seconds` of step time (at most 5).  None where oracle/_ref was not
    built."""
    from oracle.oracle import RefRun, P
    from oracle.oracle import F as FF

    if not RefRun.available():
        return None
    r = RefRun(sc)
    arrs = [coef["vx"], FF(sc.vy), FF(sc.vz), coef["rho"], coef["rkhor"], coef["rkh1"],
            coef["rkver"], coef["vset"], FF(sc.vdep)]
    r.L().ref_set_coef(*[P(a) for a in arrs])
    r.L().ref_set_dt(dt)
    r.dt = dt
    nsteps, tcpu = 0, 0.0
    while tcpu < seconds and nsteps < 5:
        t0 = time.perf_counter()
        r.step()
        tcpu += time.perf_counter() - t0
        nsteps += 1
        log(f"cpu reference step {nsteps}: {tcpu:.1f}s")
    return dict(value=sc.updates_per_step * nsteps / tcpu, unit="cell*class updates/s", cores=1,
                kind="reference",
                sample=f"{nsteps} steps of the {sc.name} workload (the reference's "
                       f"Sources/*.f90 routines, amdflang -O2, serial; coefficients from the "
                       f"device meteo stage)")


def measured_traffic(key):
    """HBM bytes per step of the keyed launch from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summaries under profiles/ (scaled by the copy-kernel
    calibration recorded there); None when not measured."""
    import glob
    val = None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_hbm_traffic.md"))):
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == key:
                val = float(f[2])
    return val


def copy_calibration(nbytes=1 << 30, reps=3):
    """The library's plain copy kernel (f3d_hbm_copy) of nbytes in and nbytes
    out; under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per dispatch give
    the scale from counter KiB to HBM bytes.  Returns the copy's GB/s."""
    from fall3d_amd import lib as F
    ms = F.hbm_copy(nbytes, reps)
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    log(f"copy calibration: {reps} launches, {nbytes} B read + {nbytes} B written each, "
        f"{ms:.3f} ms, {gbs:.0f} GB/s")
    return gbs


def r_dt(sc, coef):
    """finddt.f90:26-61 on the coefficient arrays (host, for the CPU leg)."""
    import ctypes
    from oracle.oracle import lib as olib, P, F as FF
    dtc = ctypes.c_double()
    return olib().or_finddt(sc.nx, sc.ny, sc.nz, sc.nc, P(FF(sc.dz)), sc.dx, sc.dy, P(coef["vx"]),
                            P(FF(sc.vy)), P(FF(sc.vz)), P(coef["rkh1"]), P(coef["rkhor"]),
                            P(coef["rkver"]), P(coef["vset"]), ctypes.byref(dtc))


if __name__ == "__main__":
    main()
