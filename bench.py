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
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "an5d-artifact_amd"))

import numpy as np  # noqa: E402

METRIC = "GStencil/s & % of bt-scaled HBM roofline; fwd+adj and batched vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0

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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(wl, budget_s=12.0):
    """The reference's OpenMP CPU loop (oracle/_ref) timed on a bounded sample:
    the same grid, fewer timesteps.  Falls back to the oracle's sequential loop
    on a smaller grid when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    name, dt = wl["stencil"], wl["dtype"]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if O.ref_available(name, dt):
        n = wl["n"][0]
        A = O.ref_init_grid(name, n, dt)
        t1 = O.ref_run(name, A, n, 1, scop=False) / 1e3
        steps = int(max(1, min(200, budget_s / max(t1, 1e-6))))
        ms = O.ref_run(name, A, n, steps, scop=False)
        cells = float(np.prod(wl["n"]))
        return {"value": steps * cells / (ms / 1e3) / 1e9, "unit": "GStencil/s", "cores": threads,
                "kind": "reference",
                "sample": f"{name}.c OpenMP loop (scop=false), {'x'.join(map(str, wl['n']))}, {steps} steps, "
                          f"OMP_NUM_THREADS={threads}"}
    d = O.info(name)
    n = [min(x, 128) for x in wl["n"]]
    A = O.init_grid(name, n[0], dt) if len(set(n)) == 1 else None
    t0 = time.perf_counter()
    O.run(name, A, n, 1)
    t1 = time.perf_counter() - t0
    steps = int(max(1, min(200, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    O.run(name, A, n, steps)
    dt_s = time.perf_counter() - t0
    return {"value": steps * float(np.prod(n)) / dt_s / 1e9, "unit": "GStencil/s", "cores": 1,
            "kind": "port", "sample": f"oracle sequential loop, {'x'.join(map(str, n))}, {steps} steps"}


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
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the stencil sweep runs on a single MI355X; "
              "--gpus N > 1 is not supported (multi-GPU sharding is a design note only).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import pyan5d
    from pyan5d import presets

    wl = WORKLOADS[args.workload]
    name, dt, n, T = wl["stencil"], wl["dtype"], wl["n"], wl["T"]
    d = pyan5d.info(name)
    coefs, divisor = presets.coefficients(name, "baseline")
    dcode = 0 if dt == np.float32 else 1

    # pick the kernel: explicit flags, else the tuned/default choice
    variants = pyan5d.variants(name, dcode)
    variant = args.variant if args.variant >= 0 else 0
    bt = args.bt if args.bt > 0 else (2 if d["dim"] == 3 else 4)
    bt = min(bt, variants[variant]["btmax"])

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
                              variant=variant, sl=args.sl, stream=stream)
                e1.record(stream)
                events.append((e0, e1))
            else:
                pyan5d.launch(name, grid, n, k, c0=c0, coefs=coefs, divisor=divisor,
                              variant=variant, sl=args.sl, stream=stream)
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
    tag = f"{name}_{'f32' if dt == np.float32 else 'f64'}_bt{bt}"
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": parse_profile_traffic(tag)}

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
                   "variant": variants[variant], "sl": args.sl,
                   "gflops": round(gstencil * d["fpp"], 2),
                   "bt_scaled_hbm_roof_gstencil": HBM_PEAK_GBS * bt / (2 * esz),
                   "frac_of_bt_scaled_roof": round(gstencil / (HBM_PEAK_GBS * bt / (2 * esz)), 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
