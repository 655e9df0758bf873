"""Build the autotuner table for a set of workloads on the current GPU.

    python -m pyan5d.tune [--set bench|all] [--out FILE] [--max-measured K]

For every (stencil, dtype, size, mode, run length) of the set, an5d_autotune
ranks all (kernel shape, register cap, bt, sl) candidates with the analytic
model, times the best K on the GPU (on the field halfway through the run,
with the benchmark's coefficients: pyan5d.presets "baseline") and keeps the
fastest; the table is then written to FILE (default:
the file the library loads, an5d-artifact_amd/tuned.tsv).  This replaces the
reference's Excel model + top-5 compile/run scripts
(compiled/nvcc_compile_top5.sh, compiled/run.sh), which pick one
configuration per stencil by compiling and timing separate binaries.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import EXACT, FAST, autotune, info, presets, tune_clear, tune_load, tune_save

STENCILS_2D = ["star2d1r", "star2d2r", "star2d3r", "star2d4r", "box2d1r", "box2d2r", "box2d3r",
               "box2d4r", "j2d5pt", "j2d9pt", "j2d9pt-gol", "gradient2d"]
STENCILS_3D = ["star3d1r", "star3d2r", "star3d3r", "star3d4r", "box3d1r", "box3d2r", "box3d3r",
               "box3d4r", "j3d27pt"]

# (stencil, dtype, interior size, modes, timesteps of the run)
BENCH = [
    ("star3d1r", np.float32, [512, 512, 512], (EXACT, FAST), 1000),
    ("star3d1r", np.float64, [512, 512, 512], (EXACT, FAST), 1000),
    ("j2d5pt", np.float32, [16384, 16384], (EXACT, FAST), 1000),
    ("j2d5pt", np.float32, [4096, 4096], (EXACT, FAST), 100),
    ("box2d4r", np.float32, [16384, 16384], (EXACT,), 1000),
    ("box2d4r", np.float64, [16384, 16384], (EXACT,), 1000),
    ("star3d4r", np.float32, [512, 512, 512], (EXACT,), 1000),
    ("star3d4r", np.float64, [512, 512, 512], (EXACT,), 1000),
    ("j3d27pt", np.float32, [512, 512, 512], (EXACT,), 1000),
    ("j3d27pt", np.float64, [512, 512, 512], (EXACT,), 1000),
    ("gradient2d", np.float32, [8192, 8192], (EXACT,), 256),
    ("j2d9pt", np.float32, [8192, 8192], (EXACT,), 256),
    ("j2d9pt-gol", np.float32, [8192, 8192], (EXACT,), 256),
]


def workloads(which):
    out = list(BENCH)
    if which == "all":
        seen = {(s, np.dtype(d).name, tuple(n), m) for s, d, n, ms, _ in BENCH for m in ms}
        for s in STENCILS_2D + STENCILS_3D:
            n = [16384, 16384] if s in STENCILS_2D else [512, 512, 512]
            for dt in (np.float32, np.float64):
                if (s, np.dtype(dt).name, tuple(n), EXACT) not in seen:
                    out.append((s, dt, n, (EXACT,), 1000))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--set", choices=("bench", "all"), default="all")
    ap.add_argument("--out", default=None, help="table file (default: the library's tuned.tsv)")
    ap.add_argument("--max-measured", type=int, default=1000)
    ap.add_argument("--fresh", action="store_true", help="start from an empty table")
    args = ap.parse_args(argv)
    if args.fresh:
        tune_clear()
    elif args.out and os.path.exists(args.out):
        tune_load(args.out)
    for name, dt, n, modes, steps in workloads(args.set):
        info(name)  # validates the name
        for mode in modes:
            t0 = time.perf_counter()
            coefs, div = presets.coefficients(name, "baseline")
            cfg, gs = autotune(name, n, dt, mode=mode, max_measured=args.max_measured, timesteps=steps,
                               coefs=coefs, divisor=div)
            print(f"[tune] {name} {np.dtype(dt).name} {'x'.join(map(str, n))} "
                  f"{'fast' if mode == FAST else 'exact'}: variant {cfg['variant']} bt {cfg['bt']} "
                  f"sl {cfg['sl']} -> {gs:.1f} GStencil/s ({time.perf_counter() - t0:.1f} s)", flush=True)
    tune_save(args.out)
    print(f"[tune] table written to {args.out or 'the default tuned.tsv'}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
