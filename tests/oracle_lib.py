"""ctypes access to the CPU oracle (oracle/liboracle.so) and, when built, to the
reference's own CPU programs (oracle/_ref/libref_<stencil>_<type>.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

NAMES = [
    "star2d1r", "star2d2r", "star2d3r", "star2d4r",
    "box2d1r", "box2d2r", "box2d3r", "box2d4r",
    "j2d5pt", "j2d9pt", "j2d9pt-gol", "gradient2d",
    "star3d1r", "star3d2r", "star3d3r", "star3d4r",
    "box3d1r", "box3d2r", "box3d3r", "box3d4r",
    "j3d27pt",
]
ID = {n: i for i, n in enumerate(NAMES)}

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        i64p = ctypes.POINTER(ctypes.c_int64)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_stencil_info.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.oracle_stencil_taps.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_float)]
        for suf, ct in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            f = getattr(L, "oracle_run_" + suf)
            f.argtypes = [ctypes.c_int, ctypes.POINTER(ct), i64p, ctypes.c_int, dp, ctypes.c_double]
            f = getattr(L, "oracle_init_grid_" + suf)
            f.argtypes = [ctypes.POINTER(ct), ctypes.c_int, ctypes.c_int, ctypes.c_int]
            f = getattr(L, "oracle_check_" + suf)
            f.argtypes = [ctypes.POINTER(ct), ctypes.POINTER(ct), ctypes.c_size_t, ctypes.c_int, dp]
        L.oracle_adjoint_f64.argtypes = [ctypes.c_int, dp, dp, i64p, ctypes.c_int, dp,
                                         ctypes.c_double, dp, dp, dp]
        _lib = L
    return _lib


def info(name):
    arr = (ctypes.c_int * 6)()
    lib().oracle_stencil_info(ID[name], arr)
    dim, rad, fpp, ntaps, div, linear = list(arr)
    return dict(dim=dim, rad=rad, fpp=fpp, ntaps=ntaps, divisor=div, linear=bool(linear))


def taps(name):
    n = info(name)["ntaps"]
    offs = (ctypes.c_int * (4 * n))()
    coefs = (ctypes.c_float * n)()
    lib().oracle_stencil_taps(ID[name], offs, coefs)
    o = np.array(list(offs), dtype=np.int64).reshape(n, 4)
    return o, np.array(list(coefs), dtype=np.float32)


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _ct(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def grid_shape(name, n):
    """Full (halo-included) shape of one buffer for interior extents n."""
    d = info(name)
    r = d["rad"]
    return tuple(int(x) + 2 * r for x in n)


def run(name, A, n, T, coefs=None, divisor=-1.0):
    """Run T steps of the sequential scop loop in place on A[2, ...]."""
    A = np.ascontiguousarray(A)
    suf = "f32" if A.dtype == np.float32 else "f64"
    n64 = np.array(list(n) + [0] * (3 - len(n)), dtype=np.int64)
    cp = None
    if coefs is not None:
        cbuf = np.ascontiguousarray(np.asarray(coefs, dtype=np.float64))
        cp = _ptr(cbuf, ctypes.c_double)
    rc = getattr(lib(), "oracle_run_" + suf)(ID[name], _ptr(A, _ct(A.dtype)), _ptr(n64, ctypes.c_int64),
                                             int(T), cp, float(divisor))
    if rc != 0:
        raise ValueError("oracle_run failed")
    return A


def init_grid(name, compsize, dtype=np.float32):
    d = info(name)
    shape = (2,) + (compsize + 2 * d["rad"],) * d["dim"]
    A = np.empty(shape, dtype=dtype)
    suf = "f32" if A.dtype == np.float32 else "f64"
    getattr(lib(), "oracle_init_grid_" + suf)(_ptr(A, _ct(dtype)), d["dim"], d["rad"], compsize)
    return A


def check(A, B, T):
    out = (ctypes.c_double * 2)()
    area = A[0].size
    suf = "f32" if A.dtype == np.float32 else "f64"
    getattr(lib(), "oracle_check_" + suf)(_ptr(A, _ct(A.dtype)), _ptr(B, _ct(B.dtype)), area, int(T), out)
    return out[0], out[1]


def adjoint(name, u0, target, n, T, coefs=None, divisor=-1.0):
    """FP64 reverse sweep: returns (loss, dL/du0 over the whole grid, dL/dcoefs)."""
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    target = np.ascontiguousarray(target, dtype=np.float64)
    d = info(name)
    nc = d["ntaps"] if d["linear"] else 2
    gu = np.zeros_like(u0)
    gc = np.zeros(nc, dtype=np.float64)
    loss = ctypes.c_double()
    n64 = np.array(list(n) + [0] * (3 - len(n)), dtype=np.int64)
    cp = None
    if coefs is not None:
        cbuf = np.ascontiguousarray(np.asarray(coefs, dtype=np.float64))
        cp = _ptr(cbuf, ctypes.c_double)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib().oracle_adjoint_f64(ID[name], _ptr(u0, ctypes.c_double), _ptr(target, ctypes.c_double),
                                  _ptr(n64, ctypes.c_int64), int(T), cp, float(divisor),
                                  _ptr(gu, ctypes.c_double), _ptr(gc, ctypes.c_double), ctypes.byref(loss))
    if rc != 0:
        raise ValueError("oracle_adjoint failed")
    return loss.value, gu, gc


# --- the reference's own CPU programs (built by `make -C oracle ref`) -------

def ref_available(name, dtype=np.float32):
    return os.path.exists(_ref_path(name, dtype))


def _ref_path(name, dtype):
    t = "float" if np.dtype(dtype) == np.float32 else "double"
    return os.path.join(REF_DIR, f"libref_{name}_{t}.so")


_ref_cache = {}


def ref_lib(name, dtype=np.float32):
    key = (name, np.dtype(dtype).name)
    if key not in _ref_cache:
        L = ctypes.CDLL(_ref_path(name, dtype), mode=ctypes.RTLD_LOCAL)
        ct = _ct(dtype)
        L.kernel_stencil.argtypes = [ctypes.POINTER(ct), ctypes.c_int, ctypes.c_int, ctypes.c_bool]
        L.kernel_stencil.restype = ctypes.c_double
        L.init_grid.argtypes = [ctypes.POINTER(ct), ctypes.c_int]
        _ref_cache[key] = L
    return _ref_cache[key]


def ref_init_grid(name, compsize, dtype=np.float32):
    d = info(name)
    shape = (2,) + (compsize + 2 * d["rad"],) * d["dim"]
    A = np.empty(shape, dtype=dtype)
    ref_lib(name, dtype).init_grid(_ptr(A, _ct(dtype)), compsize)
    return A


def ref_run(name, A, compsize, T, scop=True):
    """Reference kernel_stencil (scop loop if scop else OpenMP loop), in place; returns ms."""
    assert A.flags.c_contiguous
    return ref_lib(name, A.dtype).kernel_stencil(_ptr(A, _ct(A.dtype)), compsize, T, scop)
