#!/usr/bin/env python3
"""Benchmark of the temporal-blocking stencil sweep on one MI355X.

Default workload: star3d1r FP32, 512^3 interior + 1-cell fixed halo, 1000
timesteps, centre 0.4 / six neighbours 0.1, input uniform [0, 1) from rand_r
seeded with index + N.  A "step" is one full sweep of the 1000 timesteps over
the device-resident grid (the launch schedule of kernel_stencil's GPU branch).

Prints one JSON line: GStencil/s = T * N^3 / t_sweep / 1e9, plus
  roofline     the dominant kernel (the bt-step sweep): algorithmic HBM bytes
               per launch (one read + one write of the grid: 2*4*N^3 B) over
               its mean launch time from HIP events on its stream, against
               8 TB/s; `traffic` = HBM bytes per launch measured by rocprofv3
               (FETCH_SIZE + WRITE_SIZE) when a profile summary provides it.
  cpu_baseline the reference's own OpenMP loop (oracle/_ref, built from
               star3d1r.c) on the host cores, on a bounded sample.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "an5d-artifact_amd"))

import numpy as np  # noqa: E402

METRIC = "GStencil/s & % of bt-scaled HBM roofline; fwd+adj and batched vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
VALU_LANE_OPS = 256 * 4 * 16 * 2.4e9  # unfused FP32/FP64 vector lane-ops per second

WORKLOADS = {
    "star3d1r-f32": dict(stencil="star3d1r", dtype=np.float32, n=[512, 512, 512], T=1000,
                         desc="star3d1r FP32 512^3 T=1000, centre 0.4 / neighbours 0.1"),
    "star3d1r-f64": dict(stencil="star3d1r", dtype=np.float64, n=[512, 512, 512], T=1000,
                         desc="star3d1r FP64 512^3 T=1000, centre 0.4 / neighbours 0.1"),
    "j2d5pt-f32": dict(stencil="j2d5pt", dtype=np.float32, n=[16384, 16384], T=1000,
                       desc="j2d5pt FP32 16384^2 T=1000, (5.1N+12.1W+15C+12.2E+12.3S)/118"),
    "j2d5pt-f32-4096": dict(stencil="j2d5pt", dtype=np.float32, n=[4096, 4096], T=100,
                            desc="j2d5pt FP32 4096^2 T=100, (5.1N+12.1W+15C+12.2E+12.3S)/118"),
}
# the high-order set: seeded coefficients uniform (0, 1] normalised to sum 1
for _name, _n in (("box2d4r", [16384, 16384]), ("star3d4r", [512, 512, 512]), ("j3d27pt", [512, 512, 512])):
    for _dt, _tag in ((np.float32, "f32"), (np.float64, "f64")):
        WORKLOADS[f"{_name}-{_tag}"] = dict(
            stencil=_name, dtype=_dt, n=_n, T=1000,
            desc=f"{_name} {_tag.upper()} {'x'.join(map(str, _n))} T=1000, seeded normalised coefficients")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """CPU model, the hardware threads this process may use and the physical
    cores they belong to (from /proc/cpuinfo and the affinity mask)."""
    allowed = sorted(os.sched_getaffinity(0))
    model, cores, cur = "unknown", set(), {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if cur.get("processor") in allowed:
                    cores.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor"))))
                cur = {}
                continue
            k, v = (x.strip() for x in line.split(":", 1))
            cur[k] = int(v) if k == "processor" else v
            if k == "model name":
                model = v
        if cur.get("processor") in allowed:
            cores.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor"))))
    except OSError:
        pass
    return model, len(allowed), len(cores) or len(allowed)


def cpu_baseline(wl, budget_s=12.0):
    """The reference's own CPU paths (oracle/_ref, built from <stencil>.c with
    -O3 -fopenmp -ffp-contract=off): the OpenMP loop on every core this
    process may use (one thread per core, OMP_PROC_BIND=close,
    OMP_PLACES=cores) and the sequential scop loop on one, each timed on a
    bounded sample of the same grid (GStencil/s is a per-step rate).  It runs
    in a child process, so that its OpenMP placement is its own and the GPU
    process keeps its threads unbound."""
    model, hw_threads, phys = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", phys))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    code = ("import json, sys; sys.path.insert(0, sys.argv[1]); import bench; "
            "print(json.dumps(bench.cpu_baseline_child(sys.argv[2], float(sys.argv[3]))))")
    key = next(k for k, v in WORKLOADS.items() if v is wl)
    r = subprocess.run([sys.executable, "-c", code, ROOT, key, str(budget_s)], env=env, capture_output=True,
                       text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-2000:]}")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out.update({"cores": threads, "cpu_model": model, "physical_cores": phys, "hw_threads": hw_threads,
                "omp": {k: env[k] for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}})
    return out


def cpu_baseline_child(key, budget_s):
    """Body of the CPU baseline (runs in its own process, see cpu_baseline).
    Falls back to the oracle's sequential loop on a smaller grid when the
    reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    wl = WORKLOADS[key]
    name, dt = wl["stencil"], wl["dtype"]
    cells = float(np.prod(wl["n"]))
    if O.ref_available(name, dt):
        n = wl["n"][0]
        A = O.ref_init_grid(name, n, dt)
        t1 = O.ref_run(name, A, n, 1, scop=False) / 1e3
        steps = int(max(1, min(wl["T"], budget_s / max(t1, 1e-6))))
        ms = O.ref_run(name, A, n, steps, scop=False)
        # the sequential scop loop: a fifth of the time budget
        s1 = O.ref_run(name, A, n, 1, scop=True) / 1e3
        steps1 = int(max(1, min(wl["T"], 0.2 * budget_s / max(s1, 1e-6))))
        ms1 = O.ref_run(name, A, n, steps1, scop=True)
        return {"value": steps * cells / (ms / 1e3) / 1e9, "unit": "GStencil/s", "kind": "reference",
                "sample": f"{name}.c OpenMP loop (kernel_stencil scop=false), {'x'.join(map(str, wl['n']))}, "
                          f"{steps} of {wl['T']} steps, reference coefficients and init_grid",
                "build": "gcc -O3 -fopenmp -ffp-contract=off",
                "value_1thread": steps1 * cells / (ms1 / 1e3) / 1e9,
                "sample_1thread": f"{name}.c sequential scop loop, {steps1} steps"}
    n = [min(x, 128) for x in wl["n"]]
    A = O.init_grid(name, n[0], dt) if len(set(n)) == 1 else None
    t0 = time.perf_counter()
    O.run(name, A, n, 1)
    t1 = time.perf_counter() - t0
    steps = int(max(1, min(200, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    O.run(name, A, n, steps)
    dt_s = time.perf_counter() - t0
    return {"value": steps * float(np.prod(n)) / dt_s / 1e9, "unit": "GStencil/s", "kind": "port",
            "sample": f"oracle sequential loop, {'x'.join(map(str, n))}, {steps} steps"}


def error_statement(torch, pyan5d, wl, coefs, divisor, variant, sl, mode):
    """Per-run correctness statement: the first 2*bt steps of the benchmarked
    configuration (same kernel, tiling, coefficients and input) against the
    oracle's C loop in the same precision (max abs, max rel with the smallest
    normal as floor, PASS at 1e-5 FP32 / 1e-12 FP64) and in FP64."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    name, dt, n = wl["stencil"], wl["dtype"], wl["n"]
    rc = pyan5d.resolve_config(name, n, dt, variant=variant, sl=sl, mode=mode)
    T = 2 * rc["bt"]
    A = pyan5d.init_grid(name, n, dt, mode=1)
    ref = A.copy()
    ref64 = A.astype(np.float64)
    g = torch.from_numpy(A).cuda()
    pyan5d.forward_device(name, g, n, T, coefs=coefs, divisor=divisor, bt=rc["bt"], variant=variant, sl=sl,
                          mode=mode)
    torch.cuda.synchronize()
    got = g[T % 2].cpu().numpy().astype(np.float64)
    del g
    O.run(name, ref, n, T, coefs=coefs, divisor=divisor)
    O.run(name, ref64, n, T, coefs=coefs, divisor=divisor)
    c = ref[T % 2].astype(np.float64)
    c64 = ref64[T % 2]
    floor = float(np.finfo(dt).tiny)
    max_abs = float(np.max(np.abs(got - c)))
    max_rel = float(np.max(np.abs(got - c) / np.maximum(np.abs(c), floor)))
    rms = float(np.sqrt(np.mean((got - c) ** 2)))
    err64 = float(np.max(np.abs(got - c64)))
    rel64 = float(np.max(np.abs(got - c64) / np.maximum(np.abs(c64), floor)))
    tol = 1e-5 if dt == np.float32 else 1e-12
    finite = bool(np.all(np.isfinite(got)) == np.all(np.isfinite(c)))
    return {"steps_checked": T, "max_abs": max_abs, "rms": rms, "max_rel": max_rel, "rel_floor": floor,
            "max_err_vs_fp64": err64, "max_rel_vs_fp64": rel64, "tol_rel": tol,
            "bitwise": bool(max_abs == 0.0 and finite), "pass": bool(max_rel <= tol and finite)}


def bench_fwd_adj(torch, pyan5d, presets, name, n=(8192, 8192), T=256, reps=3, abts=(2, 3, 4, 5), sls=(0, 64)):
    """(checkpointed forward + adjoint) / plain forward, 2D, FP32."""
    n = list(n)
    d = pyan5d.info(name)
    coefs, div = presets.coefficients(name, "baseline")
    g = torch.Generator(device="cuda").manual_seed(7)
    shape = [x + 2 * d["rad"] for x in n]
    u0 = torch.rand(shape, device="cuda", generator=g)
    target = torch.rand(shape, device="cuda", generator=g)
    grid = torch.stack([u0, u0])
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    # plain forward: best of the exact-mode bt values
    best = None
    for bt in (0, 2, 3, 4, 6):  # 0: the autotuner's choice
        if bt > pyan5d.variants(name)[0]["btmax"]:
            continue
        pyan5d.forward_device(name, grid, n, T, coefs=coefs, divisor=div, bt=bt)
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record()
        for _ in range(reps):
            pyan5d.forward_device(name, grid, n, T, coefs=coefs, divisor=div, bt=bt)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        if best is None or ms < best[0]:
            best = (ms, bt if bt else pyan5d.resolve_config(name, n)["bt"])
    grad = torch.zeros_like(u0)
    gc = torch.zeros(d["ntaps"] if d["linear"] else 2, dtype=torch.float64, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    best_fa = None
    for abt in abts:
        if abt > 1 and not d["linear"] and abt > 2:
            continue
        for sl in sls:
            try:
                ws = pyan5d.AdjointWorkspace(name, n, T, np.float32, bt=abt, sl=sl)
            except pyan5d.An5dError:
                continue
            pyan5d.adjoint_device(name, u0, target, grad, gc, loss, n, T, ws, coefs=coefs, divisor=div,
                                  bt=abt, sl=sl)
            torch.cuda.synchronize()
            a, b = ev(), ev()
            a.record()
            for _ in range(reps):
                pyan5d.adjoint_device(name, u0, target, grad, gc, loss, n, T, ws, coefs=coefs, divisor=div,
                                      bt=abt, sl=sl)
            b.record()
            torch.cuda.synchronize()
            ms = a.elapsed_time(b) / reps
            log(f"[bench]   {name} fwd+adj bt={abt} sl={sl}: {ms:.3f} ms")
            if best_fa is None or ms < best_fa[0]:
                best_fa = (ms, abt, sl)
            del ws
            torch.cuda.empty_cache()
    cells = float(np.prod(n))
    return {"stencil": name, "n": n, "timesteps": T, "forward_ms": round(best[0], 3), "forward_bt": best[1],
            "forward_gstencil": round(T * cells / best[0] / 1e6, 1), "fwd_adj_ms": round(best_fa[0], 3),
            "fwd_adj_bt": best_fa[1], "fwd_adj_sl": best_fa[2], "ratio": round(best_fa[0] / best[0], 3)}


def bench_batched(torch, pyan5d, presets, reps=2):
    """96 instances: plain per-grid launches (each grid at the configuration
    the autotuner measured for it) vs the batched graph (segment length
    tuned by measurement), forward."""
    work = presets.batch_workload()
    g = torch.Generator(device="cuda").manual_seed(11)
    items = []
    total = 0.0
    for name, n, T in work:
        d = pyan5d.info(name)
        u = torch.rand([x + 2 * d["rad"] for x in n], device="cuda", generator=g)
        coefs, div = presets.coefficients(name, "baseline")
        items.append((name, n, T, torch.stack([u, u]), coefs, div))
        total += T * float(np.prod(n))
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    t0 = time.perf_counter()
    for name, n, T, t, coefs, div in items:
        pyan5d.autotune(name, n, max_measured=6, timesteps=T, coefs=coefs, divisor=div)  # in-memory table
    log(f"[bench]   per-grid autotuning of {len(items)} grids: {time.perf_counter() - t0:.1f} s")

    def timed(fn, r):
        fn()
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record()
        for _ in range(r):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / r

    def plain():
        for name, n, T, t, coefs, div in items:
            pyan5d.forward_device(name, t, n, T, coefs=coefs, divisor=div)

    ms_plain = timed(plain, reps)
    batch = pyan5d.Batch(items)
    sl, _ = batch.tune(reps=1)
    ms_batch = timed(batch.run, reps)
    levels = batch.levels
    batch.close()
    fwd = {"instances": len(items), "updates": total, "levels": levels, "sl": sl,
           "plain_ms": round(ms_plain, 3), "plain_gstencil": round(total / ms_plain / 1e6, 1),
           "batched_ms": round(ms_batch, 3), "batched_gstencil": round(total / ms_batch / 1e6, 1),
           "speedup": round(ms_plain / ms_batch, 3)}
    log(f"[bench]   batched forward {fwd}")
    del items
    torch.cuda.empty_cache()

    # the batched forward + adjoint path is switched off (an5d_batch_adjoint_create
    # refuses every batch), so there is no forward + adjoint leg to time
    fa = {"status": "switched off: the batched forward + adjoint stalled on its first GPU run"}
    return {"forward": fwd, "fwd_adj": fa}


def parse_profile_traffic(kernel_tag):
    """HBM bytes per launch of the dominant kernel from a committed rocprofv3
    counter summary (profiles/*traffic*.json written from the --pmc runs)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel_tag}.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="star3d1r-f32", choices=sorted(WORKLOADS))
    ap.add_argument("--bt", type=int, default=0)
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--sl", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the fwd+adj and batched legs")
    ap.add_argument("--mode", type=int, default=0, help="0 exact (bitwise), 1 fast (FMA, tolerance)")
    ap.add_argument("--timesteps", type=int, default=0,
                    help="override the workload's timesteps (short profiled runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one plain 1 GiB device-to-device copy (scale reference for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--fwd-adj-only", default=None, metavar="STENCIL",
                    help="run only the forward+adjoint leg of one 2D stencil and print its record "
                         "(profiling runs); --adj-bt limits it to one adjoint bt with the default sl")
    ap.add_argument("--adj-bt", type=int, default=0)
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched leg (96 instances, forward and forward+adjoint)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the stencil sweep runs on a single MI355X; "
              "--gpus N > 1 is not supported (multi-GPU sharding is a design note only).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import pyan5d
    from pyan5d import presets

    if args.fwd_adj_only:
        abts = (args.adj_bt,) if args.adj_bt > 0 else (2, 3, 4, 5)
        rec = bench_fwd_adj(torch, pyan5d, presets, args.fwd_adj_only, abts=abts,
                            sls=(0,) if args.adj_bt > 0 else (0, 64))
        print(json.dumps({"fwd_adj": rec}), flush=True)
        return

    if args.batched_only:
        print(json.dumps({"batched": bench_batched(torch, pyan5d, presets)}), flush=True)
        return

    wl = WORKLOADS[args.workload]
    name, dt, n, T = wl["stencil"], wl["dtype"], wl["n"], wl["T"]
    if args.timesteps > 0:
        T = args.timesteps
    d = pyan5d.info(name)
    coefs, divisor = presets.coefficients(name, "baseline")
    dcode = 0 if dt == np.float32 else 1

    # pick the kernel: explicit flags, else the autotuner's table, else the defaults
    variants = pyan5d.variants(name, dcode)
    rc = pyan5d.resolve_config(name, n, dt, bt=args.bt, variant=args.variant, sl=args.sl, mode=args.mode)
    variant, bt, sl = rc["variant"], rc["bt"], rc["sl"]

    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)  # 1 GiB read + 1 GiB written
        torch.cuda.synchronize()
        del src, dst
        log("[bench] calibration copy of 1 GiB done")

    log(f"[bench] workload {wl['desc']}; variant {variants[variant]} bt={bt}")
    host = pyan5d.init_grid(name, n, dt, mode=1)
    grid = torch.from_numpy(host).cuda()
    del host
    stream = torch.cuda.current_stream()
    sched = pyan5d.schedule(T, bt)

    def sweep(events=None):
        c0 = 0
        for k in sched:
            if events is not None and k == bt:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                pyan5d.launch(name, grid, n, k, c0=c0, coefs=coefs, divisor=divisor,
                              variant=variant, sl=sl, stream=stream, mode=args.mode)
                e1.record(stream)
                events.append((e0, e1))
            else:
                pyan5d.launch(name, grid, n, k, c0=c0, coefs=coefs, divisor=divisor,
                              variant=variant, sl=sl, stream=stream, mode=args.mode)
            c0 += 1

    for w in range(args.warmup):
        sweep()
        torch.cuda.synchronize()
        log(f"[bench] warmup {w + 1}/{args.warmup} done")

    events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        sweep(events)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in events]))
    log(f"[bench] {args.steps} sweeps timed: {ms_per_step:.3f} ms/sweep, "
        f"mean {bt}-step launch {kernel_ms * 1e3:.1f} us")

    cells = float(np.prod(n))
    esz = np.dtype(dt).itemsize
    gstencil = T * cells / (ms_per_step / 1e3) / 1e9
    alg_bytes = 2.0 * esz * cells  # per launch of bt steps: one read + one write of the grid
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    # the measured bytes belong to one kernel: variant, depth and segment length
    tag = f"{name}_{'f32' if dt == np.float32 else 'f64'}_v{variant}_bt{bt}_sl{sl}"
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": parse_profile_traffic(tag)}

    del grid
    torch.cuda.empty_cache()
    extra = {}
    if not args.no_extra:
        log("[bench] forward+adjoint legs ...")
        extra["fwd_adj"] = [bench_fwd_adj(torch, pyan5d, presets, nm) for nm in ("gradient2d", "j2d9pt", "j2d9pt-gol")]
        log(f"[bench] {extra['fwd_adj']}")
        torch.cuda.empty_cache()
        log("[bench] batched leg ...")
        extra["batched"] = bench_batched(torch, pyan5d, presets)
        log(f"[bench] {extra['batched']}")
        torch.cuda.empty_cache()

    log("[bench] error statement (first launches vs the oracle) ...")
    errs = error_statement(torch, pyan5d, wl, coefs, divisor, variant, sl, args.mode)
    log(f"[bench] error statement {errs}")

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(wl)
        log(f"[bench] cpu baseline {cpu}")

    out = {
        "metric": METRIC, "value": round(gstencil, 3), "unit": "GStencil/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if dt == np.float32 else "f64",
        "data": "synthetic: uniform [0,1) from rand_r seeded with cell index + N; halo fixed",
        "config": {"workload": wl["desc"], "stencil": name, "n": n, "timesteps": T, "bt": bt,
                   "mode": "fast (FMA, tolerance-checked)" if args.mode else "exact (bitwise vs the C loop)",
                   "variant": variants[variant], "sl": sl,
                   "gflops": round(gstencil * d["fpp"], 2),
                   "bt_scaled_hbm_roof_gstencil": HBM_PEAK_GBS * bt / (2 * esz),
                   "frac_of_bt_scaled_roof": round(gstencil / (HBM_PEAK_GBS * bt / (2 * esz)), 4),
                   # the fused sweep is issue-bound: the reference's FLOPs per update
                   # as unfused vector ops (FP32 and FP64: 256 CUs x 4 SIMDs x 16
                   # lanes x 2.4 GHz = 39.3 T lane-ops/s) bound it from above
                   "valu_roof_gstencil": round(VALU_LANE_OPS / d["fpp"] / 1e9, 1),
                   "frac_of_valu_roof": round(gstencil / (VALU_LANE_OPS / d["fpp"] / 1e9), 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "error": errs,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
