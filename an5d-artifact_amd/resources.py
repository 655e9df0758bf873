"""Compiler resource table of every compiled kernel instance.

The reference reads register counts and spills off its nvcc builds
(regnum.sh:12-13, show_spill.sh:50: `-Xptxas -v` output per binary).  Here
the same facts come from the code-object metadata the AMDGPU backend writes
into each gfx950 object of the build (VGPRs, AGPRs, SGPRs, private segment =
register spills and stack, spill counts, static LDS, workgroup bound), read
with the ROCm LLVM tools from `build/*.o`:

    python an5d-artifact_amd/resources.py [--build DIR] [--tsv OUT.tsv] [--md OUT.md]

Waves per SIMD from registers: gfx950 gives each SIMD lane 512 registers
(VGPRs and AGPRs from one file, allocated in blocks of 8), at most 8 waves;
a workgroup of W waves needs ceil(W/4) of them per SIMD.  Dynamic LDS (the
sweep kernels' exchange buffers) is set per launch and not in the metadata.
"""
from __future__ import annotations

import argparse
import os
import re
import subprocess
import sys
import tempfile

import yaml

LLVM = "/opt/rocm/llvm/bin"
HERE = os.path.dirname(os.path.abspath(__file__))
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def stencil_names():
    """Stencil ids in build order (the Makefile's STENCILS list)."""
    text = open(os.path.join(HERE, "Makefile")).read()
    m = re.search(r"^STENCILS\s*:=\s*((?:.*\\\n)*.*)$", text, re.M)
    return m.group(1).replace("\\\n", " ").split()


def kernels_of(obj, tmp):
    """The AMDGPU metadata kernel records of one host object's gfx950 bundle."""
    base = os.path.join(tmp, os.path.basename(obj))
    fat, co = base + ".fatbin", base + ".co"
    subprocess.run([f"{LLVM}/llvm-objcopy", f"--dump-section=.hip_fatbin={fat}", obj, base + ".dummy"],
                   check=True, capture_output=True)
    subprocess.run([f"{LLVM}/clang-offload-bundler", "--type=o", f"--targets={TARGET}", f"--input={fat}",
                    f"--output={co}", "--unbundle"], check=True, capture_output=True)
    notes = subprocess.run([f"{LLVM}/llvm-readelf", "--notes", co], check=True, capture_output=True,
                           text=True).stdout
    doc = notes[notes.index("---"):]
    doc = doc[:doc.index("...") + 3] if "..." in doc else doc
    meta = yaml.safe_load(doc)
    return meta.get("amdhsa.kernels", [])


def demangle(names):
    tool = f"{LLVM}/llvm-cxxfilt" if os.path.exists(f"{LLVM}/llvm-cxxfilt") else "c++filt"
    out = subprocess.run([tool], input="\n".join(names), capture_output=True,
                         text=True, check=True).stdout.split("\n")
    return out[:len(names)]


FIELDS = {
    "sweep_kernel": ["bt", "ck", "cj", "nw", "exact", "mids", "unr", "share"],
    "sweep_kernel_cap": ["bt", "ck", "cj", "nw", "exact", "waves_per_eu"],
    "plane_kernel": ["bt", "ck", "cj", "nw", "exact"],
    "direct_kernel": ["ck"],
    "adjoint_kernel": ["bt", "ck", "nw"],
    "rows_kernel": ["ck", "nw", "exact"],
    "rows_kernel_cap": ["ck", "nw", "exact", "waves_per_eu"],
    "rows3_kernel": ["ck", "cj", "nw", "exact"],
}


def describe(pretty, names):
    """(family, stencil, dtype, {param: value}) from a demangled kernel name."""
    m = re.match(r"(?:void )?an5d::(\w+)<(.*)>\(", pretty)
    if not m:
        plain = re.match(r"(?:void )?an5d::(\w+)\(", pretty)
        return (plain.group(1) if plain else pretty), "", "", {}
    fam, args = m.group(1), [x.strip() for x in m.group(2).split(",")]
    stencil, dtype, rest = "", "", args
    if fam in FIELDS and len(args) >= 2 and args[1] in ("float", "double"):
        sid = int(args[0])
        stencil = names[sid] if sid < len(names) else str(sid)
        dtype = "f32" if args[1] == "float" else "f64"
        rest = args[2:]
    params = {}
    keys = FIELDS.get(fam, [f"p{i}" for i in range(len(rest))])
    for k, v in zip(keys, rest):
        params[k] = {"true": "1", "false": "0"}.get(v, v)
    return fam, stencil, dtype, params


def waves_per_simd(vgpr, agpr, wg):
    regs = vgpr + agpr
    alloc = max(8, (regs + 7) // 8 * 8)
    w = min(8, 512 // alloc)
    per_simd = -(-max(1, wg // 64) // 4)
    return (w // per_simd) * per_simd if w >= per_simd else 0


def collect(build):
    names = stencil_names()
    rows = []
    with tempfile.TemporaryDirectory() as tmp:
        for f in sorted(os.listdir(build)):
            if not f.endswith(".o"):
                continue
            try:
                ks = kernels_of(os.path.join(build, f), tmp)
            except (subprocess.CalledProcessError, ValueError):
                continue
            pretty = demangle([k[".name"] for k in ks])
            for k, p in zip(ks, pretty):
                fam, st, dt, params = describe(p, names)
                wg = int(k.get(".max_flat_workgroup_size", 0))
                rows.append(dict(object=f, family=fam, stencil=st, dtype=dt, params=params,
                                 vgpr=int(k.get(".vgpr_count", 0)), agpr=int(k.get(".agpr_count", 0)),
                                 sgpr=int(k.get(".sgpr_count", 0)),
                                 scratch=int(k.get(".private_segment_fixed_size", 0)),
                                 vspill=int(k.get(".vgpr_spill_count", 0)),
                                 sspill=int(k.get(".sgpr_spill_count", 0)),
                                 lds_static=int(k.get(".group_segment_fixed_size", 0)), wg=wg,
                                 waves=waves_per_simd(int(k.get(".vgpr_count", 0)), int(k.get(".agpr_count", 0)),
                                                      wg), name=k[".name"]))
    return rows


def param_text(r):
    return " ".join(f"{k}={v}" for k, v in r["params"].items())


def write_tsv(rows, path):
    with open(path, "w") as f:
        f.write("family\tstencil\tdtype\tparams\tvgpr\tagpr\tsgpr\tscratch_bytes\tvgpr_spills\tsgpr_spills"
                "\tlds_static\twg_size\twaves_per_simd\n")
        for r in rows:
            f.write(f"{r['family']}\t{r['stencil']}\t{r['dtype']}\t{param_text(r)}\t{r['vgpr']}\t{r['agpr']}"
                    f"\t{r['sgpr']}\t{r['scratch']}\t{r['vspill']}\t{r['sspill']}\t{r['lds_static']}\t{r['wg']}"
                    f"\t{r['waves']}\n")


def write_md(rows, path, cmd):
    spill = [r for r in rows if r["scratch"] > 0 or r["vspill"] > 0]
    fams = {}
    for r in rows:
        fams.setdefault(r["family"], []).append(r)
    with open(path, "w") as f:
        f.write("# Compiler resources of every compiled kernel instance (gfx950)\n\n")
        f.write(f"Command: `{cmd}` after `make -C an5d-artifact_amd`.  Full table: `resources.tsv` beside "
                "this file (one row per instance).\n\n")
        f.write(f"{len(rows)} instances; {len(spill)} with register spills or private stack.\n\n")
        f.write("| family | instances | VGPR max | with scratch |\n|---|---|---|---|\n")
        for fam, rs in sorted(fams.items()):
            f.write(f"| {fam} | {len(rs)} | {max(r['vgpr'] for r in rs)} | "
                    f"{sum(1 for r in rs if r['scratch'] > 0)} |\n")
        f.write("\nInstances with scratch (spills or stack), per stencil and dtype:\n\n")
        f.write("| family | stencil | dtype | params | VGPR | AGPR | scratch B | VGPR spills | waves/SIMD |\n")
        f.write("|---|---|---|---|---|---|---|---|---|\n")
        for r in sorted(spill, key=lambda r: (r["stencil"], r["dtype"], r["family"], param_text(r))):
            f.write(f"| {r['family']} | {r['stencil']} | {r['dtype']} | {param_text(r)} | {r['vgpr']} | "
                    f"{r['agpr']} | {r['scratch']} | {r['vspill']} | {r['waves']} |\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--build", default=os.path.join(HERE, "build"))
    ap.add_argument("--tsv", default=None)
    ap.add_argument("--md", default=None)
    a = ap.parse_args(argv)
    rows = collect(a.build)
    if not rows:
        print("no gfx950 kernels found under", a.build, file=sys.stderr)
        return 1
    if a.tsv:
        write_tsv(rows, a.tsv)
    if a.md:
        write_md(rows, a.md, "python an5d-artifact_amd/resources.py --tsv profiles/resources.tsv "
                             "--md profiles/resources.md")
    if not a.tsv and not a.md:
        for r in rows:
            print(r["family"], r["stencil"], r["dtype"], param_text(r), r["vgpr"], r["scratch"], r["waves"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
