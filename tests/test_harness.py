"""The benchmark binaries bin/<stencil>.<type> (this repository's harness)
and the reference's own harness (common.h) linked against liban5d.so
through include/an5d_kernel_stencil.h (oracle/_ref/dropin_j2d5pt.*, built
from the reference tree when it is present).

CPU: usage, argument errors and exit codes, byte for byte against the
reference harness (common.h:21-30,143-177); the library's init_grid against
the reference's (common.h:44-56).  GPU: -c runs of both binaries."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import pyan5d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "an5d-artifact_amd", "bin")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin_j2d5pt.{}")
AVERAGE = re.compile(r"^Average: +([0-9]+\.[0-9]{4}) GFLOPS, +([0-9]+\.[0-9]{4}) ms$", re.M)
RUN_LINE = re.compile(r"^ {6}\d: +[0-9]+\.[0-9]{4} GFLOPS, +[0-9]+\.[0-9]{4} ms$", re.M)

USAGE = ("{cmd} [OPTION] ...\n\n-s\tSpecify compute size\n-t\tSpecify the number of timesteps\n"
         "-n\tSpecify the number of execution\n-c\tEnable comparison to CPU execution\n-h\tShow this usage\n")
ERRORS = [(["-s", "0"], "Larger compute size required: 0\n"),
          (["-s", "-7"], "Larger compute size required: -7\n"),
          (["-t", "0"], "More than one time-step required: 0\n"),
          (["-n", "0"], "More than one execution required: 0\n"),
          (["-s", "64", "-n", "-2"], "More than one execution required: -2\n")]


def run(path, args, timeout=120):
    return subprocess.run([path] + args, capture_output=True, text=True, timeout=timeout)


def binaries():
    out = [os.path.join(BIN, f) for f in ("j2d5pt.float", "j2d5pt.double", "star3d1r.float", "box2d4r.double")]
    return [p for p in out if os.path.exists(p)]


def test_binaries_are_built():
    assert len(binaries()) == 4


@pytest.mark.parametrize("path", binaries())
def test_usage_and_argument_errors(path):
    r = run(path, ["-h"])
    assert r.returncode == 0 and r.stdout == USAGE.format(cmd=path)
    for args, msg in ERRORS:
        r = run(path, args)
        assert r.returncode == 1, (args, r.returncode)
        assert r.stderr.endswith(msg), (args, r.stderr)
        assert r.stdout == ""


@pytest.mark.parametrize("prec", ["float", "double"])
def test_cli_matches_reference_harness(prec):
    """The same arguments give the same bytes and exit codes from this
    harness and from the reference's common.h over the library."""
    ref = DROPIN.format(prec)
    if not os.path.exists(ref):
        pytest.skip("the reference tree was absent when oracle/_ref was built")
    ours = os.path.join(BIN, f"j2d5pt.{prec}")
    for args in [["-h"]] + [a for a, _ in ERRORS]:
        a, b = run(ours, args), run(ref, args)
        assert a.returncode == b.returncode, args
        assert a.stdout.replace(ours, "CMD") == b.stdout.replace(ref, "CMD"), args
        assert a.stderr.replace(ours, "CMD") == b.stderr.replace(ref, "CMD"), args


@pytest.mark.parametrize("name,n,dt", [("j2d5pt", 37, np.float32), ("j2d5pt", 37, np.float64),
                                       ("star3d1r", 11, np.float32), ("box2d4r", 20, np.float64)])
def test_library_init_grid_equals_reference(name, n, dt):
    if not O.ref_available(name, dt):
        pytest.skip("reference build absent")
    got = pyan5d.init_grid(name, [n] * O.info(name)["dim"], dt, mode=0)
    ref = O.ref_init_grid(name, n, dt)
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
    assert np.array_equal(got, O.init_grid(name, n, dt))


@pytest.mark.gpu
@pytest.mark.parametrize("binary,args", [("j2d5pt.float", ["-s", "512", "-t", "50", "-n", "2", "-c"]),
                                         ("j2d5pt.double", ["-s", "300", "-t", "33", "-n", "1", "-c"]),
                                         ("star3d1r.float", ["-s", "64", "-t", "10", "-n", "2", "-c"]),
                                         ("box2d4r.double", ["-s", "200", "-t", "5", "-n", "1", "-c"])])
def test_binary_check_run(binary, args):
    print(f"case binary {binary} {' '.join(args)}", flush=True)
    r = run(os.path.join(BIN, binary), args)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    assert out.startswith("[Device warm-up]\n")
    assert "\n[Device execution]\n" in out and "\n[CPU execution]\n" in out
    reps = int(args[args.index("-n") + 1])
    assert len(RUN_LINE.findall(out)) == reps
    m = AVERAGE.search(out)
    assert m and float(m.group(1)) > 0 and float(m.group(2)) > 0
    assert "Max Error: 0.000000e+00\n" in out and "RMS Error: 0.000000e+00\n" in out
    assert "Verdict: PASS" in out


@pytest.mark.gpu
@pytest.mark.parametrize("prec,size,steps", [("float", 4096, 100), ("double", 1024, 40)])
def test_reference_harness_over_library(prec, size, steps):
    """The reference's own common.h, unmodified, linked to liban5d.so: its
    -c comparison (GPU path vs the library's OpenMP loop) reports no error."""
    ref = DROPIN.format(prec)
    if not os.path.exists(ref):
        pytest.skip("the reference tree was absent when oracle/_ref was built")
    print(f"case dropin {prec} -s {size} -t {steps}", flush=True)
    r = run(ref, ["-s", str(size), "-t", str(steps), "-n", "1", "-c"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert AVERAGE.search(r.stdout)
    assert len(RUN_LINE.findall(r.stdout)) == 1
    assert "Max Error: 0.000000e+00\n" in r.stdout and "RMS Error: 0.000000e+00\n" in r.stdout
