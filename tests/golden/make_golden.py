#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference itself.

Run in a container that has the reference tree and `make -C oracle ref` built:

    python tests/golden/make_golden.py --ref /root/reference

Outputs (data only, no reference source):

* stencil_<name>_<float|double>.npz  for every stencil x dtype:
    - ``n``, ``T``: interior size per dim and timesteps of the full-array case
    - ``out``: both buffers after the reference's sequential scop loop
      (kernel_stencil(A, n, T, scop=1) from <name>.c + common.h), starting
      from the reference's init_grid (common.h:44-56)
    - ``hash_cases`` / ``hashes``: (n, T) pairs and the SHA-256 of the two
      output buffers for a few more cases (bigger grids, T up to 17)
* schedule.json: for every temporal block depth bt used by the generated host
  launchers in compiled/float/*_host.cu, the step counts of the launches that
  follow the main loop, keyed by parity case and T mod bt.  Parsed from the
  remainder tables at e.g. compiled/float/j2d5pt-256-10-128_host.cu:79-554.
"""
from __future__ import annotations

import argparse
import glob
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

FULL = {2: (16, 5), 3: (6, 4)}
HASHED = {2: [(9, 1), (24, 2), (33, 17), (64, 3)], 3: [(5, 1), (9, 3), (12, 2)]}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make_stencils():
    for name in O.NAMES:
        d = O.info(name)
        for dt, tname in ((np.float32, "float"), (np.float64, "double")):
            n, T = FULL[d["dim"]]
            A = O.ref_init_grid(name, n, dt)
            O.ref_run(name, A, n, T, True)
            cases, hashes = [], []
            for hn, hT in HASHED[d["dim"]]:
                B = O.ref_init_grid(name, hn, dt)
                O.ref_run(name, B, hn, hT, True)
                cases.append((hn, hT))
                hashes.append(sha(B))
            np.savez(os.path.join(HERE, f"stencil_{name}_{tname}.npz"),
                     n=np.int64(n), T=np.int64(T), out=A,
                     hash_cases=np.array(cases, dtype=np.int64),
                     hashes=np.array(hashes))
            print(name, tname, "ok")


def make_schedule(ref):
    table = {}
    for path in sorted(glob.glob(os.path.join(ref, "compiled", "float", "*_host.cu"))):
        text = open(path).read()
        bt = int(re.search(r"kernel0_(\d+)<<<", text).group(1))
        # split at the parity test that follows the main loop
        head = text.index("if ((__c0Len % 2) != (((__c0Len + __side0LenMax - 1) / __side0LenMax) % 2))")
        tail = text[head:]
        match_at = tail.index("else if (__c0Len % __side0LenMax)")
        blocks = {"mismatch": tail[:match_at], "match": tail[match_at:]}
        entry = {}
        for case, body in blocks.items():
            parts = re.split(r"if \(__c0Len % __side0LenMax == (\d+)\)", body)
            per_r = {}
            for k in range(1, len(parts), 2):
                r = int(parts[k])
                per_r[str(r)] = [int(x) for x in re.findall(r"kernel0_(\d+)<<<", parts[k + 1])]
            entry[case] = per_r
        prev = table.get(str(bt))
        if prev is not None and prev != entry:
            raise SystemExit(f"inconsistent remainder tables for bt={bt}: {path}")
        table[str(bt)] = entry
    with open(os.path.join(HERE, "schedule.json"), "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print("schedule bts:", sorted(int(b) for b in table))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    make_schedule(args.ref)
    make_stencils()


if __name__ == "__main__":
    main()
