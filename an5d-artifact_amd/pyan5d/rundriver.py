"""Run driver and result log in the reference's CSV format.

The reference compiles one binary per configuration and runs the top five
configurations the model predicted for each stencil (compiled/run.sh:13-35):
`<prec>/bin/<stencil>-<bS1>[x<bS2>]-<bT>-<sl> -s 16384|512 -t 1000 -n 5`, takes
field 2 of the `Average` line (GFLOPS) and appends
`Stencil,bS1,bS2,bT,sl,Model,Measured` to `<gpu>/<prec>_run.csv`, with
`Failed` when no Average line appears; summarize.sh:44-66 then reads columns
1 and 7.  Here one binary per stencil serves every configuration: the
configuration goes in through AN5D_VARIANT / AN5D_BT / AN5D_SL, and the list
comes either from a reference-format CSV (`star2d1r,256,,10,256,"7,330"`) or
from this library's model (the autotuner's prior, an5d_model_estimate).

    python -m pyan5d.rundriver --prec float --stencils j2d5pt,star3d1r --top 5 \\
        --out float_run.csv
    python -m pyan5d.rundriver --summarize float_run.csv
"""
from __future__ import annotations

import argparse
import csv
import os
import subprocess
import sys

import numpy as np

from . import FLOAT32, FLOAT64, info, model_estimate, variants

HEADER = ["Stencil", "bS1", "bS2", "bT", "sl", "Model", "Measured"]
_HERE = os.path.dirname(os.path.abspath(__file__))
BIN_DIR = os.path.join(os.path.dirname(_HERE), "bin")


def default_size(name):
    """compiled/run.sh: 16384 for the 2D stencils, 512 for the 3D ones."""
    return 16384 if info(name)["dim"] == 2 else 512


def block_sizes(name, v, bt):
    """The reference's bS1 / bS2 for a kernel shape: the block extent with its
    overlap halo along the fastest dim (64 lanes x CK cells) and, in 3D, along
    the middle dim (NW waves x CJ rows); 2D leaves bS2 empty."""
    b1 = 64 * v["ck"]
    b2 = v["nw"] * v["cj"] if info(name)["dim"] == 3 else None
    return b1, b2


def top_configs(name, prec="float", n=None, k=5, sls=(0, 64, 128, 256)):
    """The k configurations (variant, bt, sl) the model ranks highest for the
    run's size, with their predicted GFLOPS (the reference's Model column)."""
    dt = np.float32 if prec == "float" else np.float64
    d = info(name)
    ext = [n or default_size(name)] * d["dim"]
    cands = []
    for v in variants(name, FLOAT32 if prec == "float" else FLOAT64):
        for bt in range(1, v["btmax"] + 1):
            for sl in (sls if v["kind"] != 1 else (0,)):
                try:
                    gs = model_estimate(name, ext, dt, bt=bt, variant=v["index"], sl=sl)
                except Exception:
                    continue
                if gs > 0:
                    b1, b2 = block_sizes(name, v, bt)
                    cands.append(dict(stencil=name, variant=v["index"], bt=bt, sl=sl, bS1=b1, bS2=b2,
                                      model=gs * d["fpp"]))
    cands.sort(key=lambda c: -c["model"])
    return cands[:k]


def read_reference_list(path):
    """Rows of a reference top-5 list (compiled/v100_float.csv format:
    Stencil,bS1,bS2,bT,sl,"Model", the model in GFLOPS with thousands
    separators) as configurations: bT and sl carry over (temporal depth and
    streamed planes per segment mean the same here), bS1 / bS2 describe a
    CUDA block and are kept only for the log; the kernel shape is the
    library's tuned or default one for that depth."""
    rows = []
    with open(path, newline="") as f:
        for rec in csv.reader(f):
            if not rec or rec[0] == "Stencil":
                continue
            rows.append(dict(stencil=rec[0], variant=-1, bt=int(rec[3]), sl=int(rec[4]), bS1=rec[1],
                             bS2=rec[2] or None, model=float(rec[5].replace(",", ""))))
    return rows


def measured_gflops(log_text):
    """Field 2 of the first Average line once runs of blanks are squeezed
    (compiled/run.sh:29: grep Average | tr -s " " | cut -d " " -f 2), or None
    when there is no such line."""
    for line in log_text.splitlines():
        if "Average" in line:
            fields = [x for x in line.split(" ") if x != ""]
            return fields[1] if len(fields) > 1 else None
    return None


def run_config(cfg, prec, size, steps, reps, bin_dir=BIN_DIR, log_dir=None, timeout=1200):
    """Run one configuration through the stencil's binary; returns the
    measured GFLOPS text or "Failed"."""
    exe = os.path.join(bin_dir, f"{cfg['stencil']}.{prec}")
    env = dict(os.environ)
    if cfg["variant"] >= 0:
        env["AN5D_VARIANT"] = str(cfg["variant"])
    env["AN5D_BT"] = str(cfg["bt"])
    env["AN5D_SL"] = str(cfg["sl"])
    try:
        r = subprocess.run([exe, "-s", str(size), "-t", str(steps), "-n", str(reps)], capture_output=True,
                           text=True, env=env, timeout=timeout)
        text = r.stdout + r.stderr
    except (OSError, subprocess.TimeoutExpired) as e:
        text = f"run failed: {e}\n"
    if log_dir:
        os.makedirs(log_dir, exist_ok=True)
        tag = f"{cfg['stencil']}-v{cfg['variant']}-{cfg['bt']}-{cfg['sl']}"
        with open(os.path.join(log_dir, tag + ".log"), "w") as f:
            f.write(text)
    return measured_gflops(text) or "Failed"


def csv_row(cfg, measured):
    b2 = "" if cfg.get("bS2") is None else str(cfg["bS2"])
    return [cfg["stencil"], str(cfg["bS1"]), b2, str(cfg["bt"]), str(cfg["sl"]), f"{cfg['model']:.0f}", measured]


def write_csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f, lineterminator="\n")
        w.writerow(HEADER)
        w.writerows(rows)


def summarize(path):
    """summarize.sh:44-49: the sorted rows' stencil and measured columns."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f if ln.strip() and not ln.startswith("Stencil,")]
    return [",".join(ln.split(",")[i] for i in (0, 6)) for ln in sorted(lines)]


def main(argv=None):
    ap = argparse.ArgumentParser(description="run the benchmark binaries over model-ranked configurations")
    ap.add_argument("--prec", choices=["float", "double"], default="float")
    ap.add_argument("--stencils", default="j2d5pt,star3d1r")
    ap.add_argument("--top", type=int, default=5)
    ap.add_argument("--list", default=None, metavar="CSV",
                    help="run the configurations of a reference-format top-5 list instead of the model's")
    ap.add_argument("--size", type=int, default=0, help="compute size (default: 16384 in 2D, 512 in 3D)")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--bin-dir", default=BIN_DIR)
    ap.add_argument("--log-dir", default=None)
    ap.add_argument("--out", default=None, help="CSV path (default: <prec>_run.csv)")
    ap.add_argument("--summarize", default=None, metavar="CSV", help="print summarize.sh's view of a CSV")
    a = ap.parse_args(argv)
    if a.summarize:
        print("\n".join(summarize(a.summarize)))
        return 0
    if a.list:
        cfgs = read_reference_list(a.list)
    else:
        cfgs = [c for name in a.stencils.split(",") for c in top_configs(name, a.prec, a.size or None, a.top)]
    rows = []
    for cfg in cfgs:
        name = cfg["stencil"]
        size = a.size or default_size(name)
        print(f"Running {name} variant {cfg['variant']} bt {cfg['bt']} sl {cfg['sl']} for {a.prec}...", flush=True)
        rows.append(csv_row(cfg, run_config(cfg, a.prec, size, a.steps, a.reps, a.bin_dir, a.log_dir)))
    write_csv(a.out or f"{a.prec}_run.csv", rows)
    return 0


if __name__ == "__main__":
    sys.exit(main())
