"""The C ABI library loads, exports every symbol include/*.h declares, and
its registry and argument checking work without touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import pyan5d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("an5d.h",):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(an5d_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = pyan5d.lib()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n


def test_kernel_stencil_macro_header_compiles(tmp_path):
    src = tmp_path / "ks.c"
    src.write_text('#define SB_TYPE float\n#include "an5d_kernel_stencil.h"\n'
                   'AN5D_DEFINE_KERNEL_STENCIL("j2d5pt")\n')
    rc = os.system(f"gcc -c -I{ROOT}/include {src} -o {tmp_path}/ks.o")
    assert rc == 0


def test_registry_matches_oracle():
    names = pyan5d.stencil_names()
    assert names == O.NAMES
    for name in names:
        a, b = pyan5d.info(name), O.info(name)
        for k in ("dim", "rad", "fpp", "ntaps", "divisor", "linear"):
            assert a[k] == b[k], (name, k)
        if a["linear"]:
            po, pc = pyan5d.taps(name)
            oo, oc = O.taps(name)
            assert np.array_equal(po, oo)
            assert np.array_equal(pc.astype(np.float32), oc)


@pytest.mark.parametrize("name", O.NAMES)
def test_every_stencil_has_kernels_for_both_dtypes(name):
    for dt in (0, 1):
        vs = pyan5d.variants(name, dt)
        assert vs, name
        kinds = [v["kind"] for v in vs]
        assert kinds.count(1) == 1  # the one-step-per-launch kernel is always there
        # one-step kernels on whole tiles: the LDS-staged plane kernel exactly
        # for the radius 2-4 3D stars, two plane-streamed partial-sum shapes
        # exactly for the radius-1 3D boxes
        d = O.info(name)
        is_star3 = name.startswith("star3d") and d["rad"] >= 2
        is_box3_r1 = name in ("j3d27pt", "box3d1r")
        assert kinds.count(2) == (1 if is_star3 else 2 if is_box3_r1 else 0), (name, kinds)
        for v in vs:
            assert 1 <= v["btmax"] <= 16
            if v["kind"] == 2:
                assert v["btmax"] == 1 and v["lds_bytes"] <= 160 * 1024


def test_argument_errors_do_not_touch_the_gpu():
    with pytest.raises(pyan5d.An5dError):
        pyan5d.stencil_id("nope")
    A = np.zeros((2, 6, 6), np.float32)
    p, _ = pyan5d.make_problem("j2d5pt", [4, 4], 3, np.float32)
    p.n[0] = 0
    ms = ctypes.c_double()
    assert pyan5d.lib().an5d_forward_host(ctypes.byref(p), A.ctypes.data, None, ctypes.byref(ms)) == -1
    assert "extents" in pyan5d.lib().an5d_last_error().decode()
    with pytest.raises(pyan5d.An5dError):
        pyan5d.forward_host("j2d5pt", A, [4, 4], -1)
    p, _ = pyan5d.make_problem("j2d5pt", [4, 4], 3, np.float32)
    p.dtype = 7
    ms = ctypes.c_double()
    assert pyan5d.lib().an5d_forward_host(ctypes.byref(p), A.ctypes.data, None, ctypes.byref(ms)) == -1
    p.dtype = 0
    assert pyan5d.lib().an5d_forward_host(ctypes.byref(p), None, None, ctypes.byref(ms)) == -1
    cfg = pyan5d.make_config(bt=0, variant=99)
    # variant out of range is rejected before any device work
    assert pyan5d.lib().an5d_forward_device(ctypes.byref(p), ctypes.c_void_p(16), None,
                                            ctypes.byref(cfg), None) == -1


def test_zero_steps_is_a_no_op():
    A = np.arange(2 * 36, dtype=np.float32).reshape(2, 6, 6)
    B = A.copy()
    pyan5d.forward_host("j2d5pt", A, [4, 4], 0)
    assert np.array_equal(A, B)
    assert pyan5d.kernel_stencil("j2d5pt", A, 4, 0, scop=True) == 0.0
    assert np.array_equal(A, B)


def test_wrappers_reject_bad_tensors_before_any_launch():
    """CPU tensors, single-buffer tensors, wrong extents, wrong dtypes and
    wrongly sized gradient outputs raise in Python, before the library sees a
    pointer."""
    import torch
    A = np.zeros((2, 6, 6), np.float32)
    with pytest.raises(ValueError):
        pyan5d.forward_host("j2d5pt", A, [0, 4], 3)
    with pytest.raises(ValueError):
        pyan5d.forward_host("j2d5pt", A[:1].copy(), [4, 4], 3)       # one buffer
    with pytest.raises(ValueError):
        pyan5d.forward_host("j2d5pt", A, [4, 5], 3)                  # wrong extents
    with pytest.raises(ValueError):
        pyan5d.forward_host("j2d5pt", A, [4, 4], 3, coefs=[1.0, 2.0])  # wrong tap count
    with pytest.raises(ValueError):
        pyan5d.forward_host("j2d5pt", A.astype(np.float16), [4, 4], 3)  # unsupported dtype
    t = torch.zeros(2, 6, 6)
    with pytest.raises(ValueError):
        pyan5d.forward_device("j2d5pt", t, [4, 4], 3)                # CPU tensor
    with pytest.raises(ValueError):
        pyan5d.launch("j2d5pt", t, [4, 4], 1)
    with pytest.raises((TypeError, ValueError)):
        pyan5d.forward_device("j2d5pt", A, [4, 4], 3)                # not a tensor
    u = torch.zeros(6, 6)
    gc = torch.zeros(5, dtype=torch.float64)
    loss = torch.zeros(1, dtype=torch.float64)
    with pytest.raises(ValueError):
        pyan5d.adjoint_device("j2d5pt", u, u, u, gc, loss, [4, 4], 3, None)
    with pytest.raises(ValueError):
        pyan5d.Batch([("j2d5pt", [4, 4], 3, t, None, -1.0)])
