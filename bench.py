"""Benchmark of the FALL3D transport step on one MI355X.

Workload: the precision-form scenario, 400x400x40 grid, 20 particle classes
(2 um - 4 mm), Fourier-mode meteorology with sinusoidal terrain (synthetic,
seeded).  A step is one pass of the transport core over all 20 classes:
setsrc + sizloop (advection, diffusion, boundaries, deposit, clip, outflux).

value     = cell*class updates per second of the FP32 fused path (inputs
            resident in HBM, meteo stage excluded and reported separately)
roofline  = the fused step kernels' algorithmic HBM bytes per step over their
            HIP-event average time, against 8 TB/s
fp64_plain= the routine-by-routine FP64 drop-in on the same workload
cpu_baseline = the oracle's FP64 C restatement (class-parallel threads) on
            this host, on a bounded number of steps of the same workload;
            beside it (reference_1core) the reference's own Sources/*.f90
            routines, compiled by amdflang -O2 into oracle/_ref, on one core
ensemble  = 64 members x 200x200x30 x 16 classes on one meteorology, one
            fused launch per step
uneven_batch = scenarios of 80..320 squared x 16..48 levels x 4..24 classes,
            each with its own meteorology, source and dt: the persistent
            batched step with HIP-graph replay and the autotuned tile, beside
            per-scenario f3d_step launches
serving   = Poisson arrivals of 6 h forecasts on windows of three resident
            400x400x40 regions, continuous batching: p50/p99 latency,
            throughput, error against the FP64 drop-in on a sampled request
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
    write c, read vset of the class, read the 7 shared meteo/coefficient
    fields (vx, vy, vz, 1/rho, rho*rkh1, rho*rkhor, rho*rkver) once per tile
    for all nc classes."""
    return 2 * s_c + (s_v + 7 * s_m / nc) / n_share


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=400)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--nc", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--ref-seconds", type=float, default=10.0,
                    help="step time of the one-core reference leg (0: skip it)")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--tile", default="58x18")
    ap.add_argument("--waves", type=int, default=0,
                    help="waves per workgroup of the flux kernel (0: two region rows per wave)")
    ap.add_argument("--skip-ensemble", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=64)
    ap.add_argument("--skip-uneven", action="store_true")
    ap.add_argument("--uneven-n", type=int, default=128)
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="arrival window of the serving run (2 req/s)")
    ap.add_argument("--serve-rate", type=float, default=2.0)
    ap.add_argument("--slabs", type=int, default=2,
                    help="member*class slabs per workgroup of the 58-column kernel")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; --gpus above 1 is not supported (the path "
              "would shard (member, class) slabs across GPUs, with one FP64 all-reduce of "
              "each member's outflux)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from fall3d_amd import lib as F
    from fall3d_amd import scenario as S

    t0 = time.time()
    sc = S.config_c2(nx=args.nx, ny=args.nx, nz=args.nz, nc=args.nc)
    log(f"scenario {sc.name} built in {time.time() - t0:.1f}s")
    upd = sc.updates_per_step

    # ---------------- FP32 fused path ----------------
    r = F.Run(sc, precision=F.FP32)
    tx, ty = (int(v) for v in args.tile.split("x"))
    r.ctx.set_tile(tx, ty, args.waves, args.slabs)
    t0 = time.time()
    r.meteo_update()
    r.ctx.sync()
    meteo_s = time.time() - t0
    log(f"meteo stage (kappa3+setvset+finddt, FP64 on device) {meteo_s:.3f}s dt={r.dt:.3f}s")
    r.step(args.warmup)
    r.ctx.sync()
    r.ctx.timing(True)
    r.ctx.kernel_time(reset=True)
    r.ctx.sync()
    t0 = time.perf_counter()
    r.step(args.steps)
    r.ctx.sync()
    el = time.perf_counter() - t0
    kms, kcount = r.ctx.kernel_time()
    r.ctx.timing(False)
    value = upd * args.steps / el
    ms_per_step = el * 1e3 / args.steps
    log(f"fp32 fused: {ms_per_step:.3f} ms/step, {value:.3e} updates/s, "
        f"kernel avg {kms:.3f} ms over {kcount} launches")
    b_alg = alg_bytes_per_update(sc.nc)
    achieved = b_alg * upd / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    budget = r.budget()
    fp32_state = None
    r.close()

    # plain 1 GiB copy: achievable HBM bandwidth, and the known-size kernel
    # that calibrates the rocprofv3 FETCH_SIZE / WRITE_SIZE counters
    copy_gbs = copy_calibration()

    # ---------------- FP64 plain drop-in ----------------
    fp64 = None
    coef = None
    if not args.skip_fp64:
        r64 = F.Run(sc, precision=F.FP64)
        r64.meteo_update()
        r64.step(2)
        r64.ctx.sync()
        t0 = time.perf_counter()
        r64.step(args.fp64_steps)
        r64.ctx.sync()
        el64 = time.perf_counter() - t0
        fp64 = upd * args.fp64_steps / el64
        log(f"fp64 plain: {el64 * 1e3 / args.fp64_steps:.3f} ms/step, {fp64:.3e} updates/s")
        coef = r64.ctx.coef() if not args.skip_cpu else None
        r64.close()

    # ---------------- CPU baseline: the oracle on this host ----------------
    cpu = None
    if not args.skip_cpu:
        from oracle.oracle import OracleRun

        ncores = len(os.sched_getaffinity(0))
        nthreads = max(1, min(ncores, sc.nc, 16))
        o = OracleRun(sc, nthreads=nthreads)
        if coef is None:
            r64 = F.Run(sc, precision=F.FP64)
            r64.meteo_update()
            coef = r64.ctx.coef()
            r64.close()
        # coefficients from the device meteo stage (the step phase is timed)
        o.vx, o.rho, o.rkhor, o.rkh1, o.rkver, o.vset = (coef[k] for k in
                                                          ("vx", "rho", "rkhor", "rkh1", "rkver", "vset"))
        o.vy = np.asfortranarray(sc.vy)
        o.vz = np.asfortranarray(sc.vz)
        o.vdep = np.asfortranarray(sc.vdep)
        o.dt = o_dt = r_dt(sc, coef)
        nsteps, tcpu = 0, 0.0
        while tcpu < args.cpu_seconds and nsteps < 50:
            t0 = time.perf_counter()
            o.step()
            tcpu += time.perf_counter() - t0
            nsteps += 1
            log(f"cpu oracle step {nsteps}: {tcpu:.1f}s")
        cpu = dict(value=upd * nsteps / tcpu, unit="cell*class updates/s", cores=nthreads,
                   kind="port",
                   sample=f"{nsteps} steps of the {sc.name} workload (FP64 C restatement, "
                          f"{nthreads} class-parallel threads, coefficients from the device "
                          f"meteo stage)")
        del o
        if args.ref_seconds > 0:
            cpu["reference_1core"] = ref_cpu_baseline(sc, upd, args.ref_seconds, coef,
                                                      o_dt)

    out = {
        "metric": "cell·class updates/s + % HBM peak; ensemble p50/p99 forecast latency",
        "value": value,
        "unit": "cell*class updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 deposit/budgets/settling/CFL",
        "data": "synthetic (seeded Fourier-mode meteorology, Suzuki source)",
        "config": {"workload": f"precision form {sc.nx}x{sc.ny}x{sc.nz} grid, {sc.nc} classes, "
                               f"fused FP32 step", "nx": sc.nx, "ny": sc.ny, "nz": sc.nz,
                   "nc": sc.nc, "tile": args.tile, "slabs": args.slabs, "dt_s": None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"{args.tile}m{args.slabs}"),
                     "bytes_per_update": b_alg, "kernel_avg_ms": kms, "launches": kcount},
        "fp64_plain_updates_per_s": fp64,
        "speedup_vs_fp64_plain": (value / fp64) if fp64 else None,
        "meteo_stage_s": meteo_s,
        "cpu_baseline": cpu,
        "hbm_peak_frac_wall": value * b_alg / 1e9 / HBM_PEAK_GBS,
        "hbm_copy_gbs": copy_gbs,
    }
    if not args.skip_ensemble:
        out["ensemble"] = bench_ensemble(args, fp64)
    if not args.skip_uneven:
        out["uneven_batch"] = bench_uneven(args, fp64)
    if not args.skip_serve:
        out["serving"] = bench_serving(args, fp64)
    print(json.dumps(out), flush=True)


def ref_cpu_baseline(sc, upd, seconds, coef, dt):
    """The reference's own routines (Sources/*.f90 compiled by amdflang -O2
    into oracle/_ref by oracle/Makefile), serial on one core as one process of
    the reference runs, on the coefficients and dt of the device meteo stage
    (as the port leg): steps until `seconds` of step time (at most 5).  None
    where oracle/_ref was not built."""
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
    return dict(value=upd * nsteps / tcpu, unit="cell*class updates/s", cores=1,
                kind="reference",
                sample=f"{nsteps} steps of the {sc.name} workload (the reference's "
                       f"Sources/*.f90 routines, amdflang -O2, serial; coefficients from "
                       f"the device meteo stage)")


def _speedup(v, fp64):
    return (v / fp64) if fp64 else None


def bench_ensemble(args, fp64, nsteps=10):
    """64 members on one 200x200x30 grid, 16 classes: one fused launch per
    step for all members (FP32 state, FP64 deposits and budgets)."""
    from fall3d_amd import lib as F
    from fall3d_amd import scenario as S
    base = S.config_c2(nx=200, ny=200, nz=30, nc=16, seed=0xF3D0003)
    mem = [S.config_ensemble_member(base, 5000 + k) for k in range(args.ensemble_members)]
    ens = F.Ensemble(base, mem)
    try:
        ens.meteo_update()
        # the library's tile autotuner on this context (no state change)
        (tx, ty, slabs), tune_ms = ens.ctx.autotune(trials=2)
        log(f"ensemble: autotuned tile {tx}x{ty}, {slabs} slabs ({tune_ms:.2f} ms per flux step)")
        ens.step(2)
        ens.ctx.sync()
        ens.ctx.timing(True)
        ens.ctx.kernel_time(reset=True)
        t0 = time.perf_counter()
        ens.step(nsteps)
        ens.ctx.sync()
        el = time.perf_counter() - t0
        kms, _ = ens.ctx.kernel_time()
    finally:
        ens.close()
    upd = base.nx * base.ny * base.nz * base.nc * len(mem)
    value = upd * nsteps / el
    b_alg = alg_bytes_per_update(base.nc, n_share=len(mem))
    ach = b_alg * upd / (kms * 1e-3) / 1e9
    log(f"ensemble: {len(mem)} members, {el * 1e3 / nsteps:.2f} ms/step, {value:.3e} updates/s")
    return {"workload": f"{len(mem)} members x {base.nx}x{base.ny}x{base.nz} x {base.nc} classes",
            "autotuned_tile": f"{tx}x{ty} slabs {slabs}", "updates_per_s": value, "ms_per_step": el * 1e3 / nsteps, "kernel_avg_ms": kms,
            "bytes_per_update": b_alg, "achieved_gbs": ach, "hbm_frac": ach / HBM_PEAK_GBS,
            "speedup_vs_fp64_plain": _speedup(value, fp64)}


def _uneven_scenario(seed):
    from fall3d_amd import scenario as S
    return S.config_uneven(seed)


def bench_uneven(args, fp64, nsteps=10):
    """Uneven batch: scenarios of their own size, class count, meteorology,
    source and dt; persistent batched step (autotuned tile, HIP graphs)
    against per-scenario f3d_step launches of the same kernels."""
    import concurrent.futures as cf
    from fall3d_amd import lib as F
    t0 = time.time()
    nproc = max(1, min(16, len(os.sched_getaffinity(0))))
    with cf.ProcessPoolExecutor(nproc) as ex:
        scs = list(ex.map(_uneven_scenario, range(1000, 1000 + args.uneven_n)))
    log(f"uneven: {len(scs)} scenarios built in {time.time() - t0:.1f}s")
    runs = []
    try:
        for sc in scs:
            r = F.Run(sc, precision=F.FP32)
            r.meteo_update()
            runs.append(r)
        upd = sum(sc.updates_per_step for sc in scs)
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
        b.step(2)
        b.sync()
        t0 = time.perf_counter()
        b.step(nsteps)
        b.sync()
        el_batch = time.perf_counter() - t0
        b.close()
    finally:
        for r in runs:
            r.close()
    v_batch, v_plain = upd * nsteps / el_batch, upd * nsteps / el_plain
    log(f"uneven: batched {el_batch * 1e3 / nsteps:.2f} ms/step ({v_batch:.3e} updates/s), "
        f"per-scenario {el_plain * 1e3 / nsteps:.2f} ms/step ({v_plain:.3e}), tile {tile}")
    return {"workload": f"{len(scs)} scenarios, nx=ny 80..320, nz 16..48, nc 4..24",
            "updates_per_step": upd, "batched_updates_per_s": v_batch,
            "per_scenario_updates_per_s": v_plain, "batched_ms_per_step": el_batch * 1e3 / nsteps,
            "per_scenario_ms_per_step": el_plain * 1e3 / nsteps,
            "autotuned_tile": "%dx%d slabs %d" % tile, "autotune_best_flux_ms": tune_ms,
            "speedup_vs_fp64_plain": _speedup(v_batch, fp64)}


def bench_serving(args, fp64):
    """Continuous-batching serving of 6 h forecasts on windows of three
    resident 400x400x40 regions; latency from arrival to final ground load."""
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
    done = srv.run(reqs, log=None)
    st = serve.latency_stats(done)
    thr = srv.updates / srv.wall_s
    log(f"serving: p50 {st['p50']:.2f}s p99 {st['p99']:.2f}s, {thr:.3e} updates/s over "
        f"{srv.wall_s:.1f}s (GPU busy {srv.busy_s:.1f}s)")
    err = None
    if sample is not None:
        # the sampled request again, stepped by the FP64 drop-in
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
        err = {"request": sample.rid, "steps": got.nsteps,
               "load_rel_l1": float(np.abs(a - b).sum() / np.abs(b).sum()),
               "load_rel_linf": float(np.abs(a - b).max() / np.abs(b).max())}
    return {"workload": f"Poisson {args.serve_rate} req/s over {args.serve_seconds:.0f}s, "
                        f"6 h forecasts of 200x200x30 x 16 classes on windows of 3 resident "
                        f"400x400x40 regions",
            "requests": st["n"], "latency_p50_s": st["p50"], "latency_p99_s": st["p99"],
            "latency_max_s": st["max"], "throughput_updates_per_s": thr,
            "busy_updates_per_s": srv.updates / srv.busy_s if srv.busy_s else None,
            "gpu_busy_frac": srv.busy_s / srv.wall_s, "wall_s": srv.wall_s,
            "speedup_vs_fp64_plain": _speedup(srv.updates / srv.busy_s if srv.busy_s else 0, fp64),
            "fp64_check": err}


def measured_traffic(tile):
    """HBM bytes per launch of the step kernel at this tile shape, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE summaries under profiles/ (scaled by the
    copy-kernel calibration recorded there); None when not measured."""
    import glob
    val = None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_hbm_traffic.md"))):
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == tile:
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
