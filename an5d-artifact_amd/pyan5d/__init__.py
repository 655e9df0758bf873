"""Python binding of liban5d (the C ABI in include/an5d.h) via ctypes.

Device-resident calls take torch tensors (ROCm build) for memory and the
current stream; all compute runs in the HIP kernels of liban5d.so.  There is
no CPU fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "liban5d.so")

OK = 0
FLOAT32, FLOAT64 = 0, 1


class An5dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"an5d error {code}: {msg}")
        self.code = code


class Problem(ctypes.Structure):
    _fields_ = [("stencil", ctypes.c_int), ("dtype", ctypes.c_int),
                ("n", ctypes.c_longlong * 3), ("timesteps", ctypes.c_int),
                ("coefs", ctypes.POINTER(ctypes.c_double)), ("divisor", ctypes.c_double)]


class Config(ctypes.Structure):
    _fields_ = [("bt", ctypes.c_int), ("variant", ctypes.c_int), ("sl", ctypes.c_int),
                ("mode", ctypes.c_int)]


EXACT, FAST, EXACT_GENERAL = 0, 1, 2


class Layout(ctypes.Structure):
    _fields_ = [("s0", ctypes.c_longlong), ("s1", ctypes.c_longlong), ("buf_stride", ctypes.c_longlong)]


class StencilInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("dim", ctypes.c_int), ("rad", ctypes.c_int),
                ("fpp", ctypes.c_int), ("ntaps", ctypes.c_int), ("divisor", ctypes.c_int),
                ("linear", ctypes.c_int)]


class VariantInfo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("ck", ctypes.c_int), ("cj", ctypes.c_int),
                ("nw", ctypes.c_int), ("btmax", ctypes.c_int), ("lds_bytes", ctypes.c_longlong),
                ("regcap", ctypes.c_int)]


class BatchGrad(ctypes.Structure):
    _fields_ = [("u0", ctypes.c_void_p), ("target", ctypes.c_void_p), ("grad", ctypes.c_void_p),
                ("gcoef", ctypes.c_void_p), ("loss", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C an5d-artifact_amd` "
                               "(there is no fallback path)")
        # torch's ROCm wheel carries its own libamdhip64.so.  Loaded first, it
        # also serves liban5d (same soname), so torch's tensors, streams and
        # the library share one HIP runtime.  Loaded after liban5d, it would
        # be a second runtime in the process, with no device visible to torch.
        if "torch" not in sys.modules:
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.an5d_last_error.restype = ctypes.c_char_p
        L.an5d_stencil_id.argtypes = [ctypes.c_char_p]
        L.an5d_stencil_get_info.argtypes = [ctypes.c_int, P(StencilInfo)]
        L.an5d_stencil_taps.argtypes = [ctypes.c_int, P(ctypes.c_int), P(ctypes.c_double), ctypes.c_int]
        L.an5d_variant_count.argtypes = [ctypes.c_int, ctypes.c_int]
        L.an5d_variant_get_info.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(VariantInfo)]
        L.an5d_schedule.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_int), ctypes.c_int]
        L.an5d_kernel_stencil.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
        L.an5d_forward_host.argtypes = [P(Problem), ctypes.c_void_p, P(Config), P(ctypes.c_double)]
        L.an5d_forward_device.argtypes = [P(Problem), ctypes.c_void_p, P(Layout), P(Config), ctypes.c_void_p]
        L.an5d_launch.argtypes = [P(Problem), ctypes.c_void_p, P(Layout), P(Config), ctypes.c_int,
                                  ctypes.c_int, ctypes.c_void_p]
        L.an5d_init_grid.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                     ctypes.c_int]
        L.an5d_adjoint_host.argtypes = [P(Problem), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        P(ctypes.c_double), P(ctypes.c_double), P(Config),
                                        P(ctypes.c_double), P(ctypes.c_double)]
        L.an5d_adjoint_workspace.argtypes = [P(Problem), P(Config), P(ctypes.c_size_t)]
        L.an5d_adjoint_device.argtypes = [P(Problem), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                          P(Config), ctypes.c_void_p, P(ctypes.c_double)]
        L.an5d_adjoint_device_noalloc.argtypes = [P(Problem), ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                  P(Config), ctypes.c_void_p]
        L.an5d_batch_create.argtypes = [P(ctypes.c_void_p), ctypes.c_int, P(Problem), P(ctypes.c_void_p),
                                        P(Config)]
        L.an5d_batch_adjoint_create.argtypes = [P(ctypes.c_void_p), ctypes.c_int, P(Problem), P(BatchGrad),
                                                P(Config)]
        L.an5d_batch_tune.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_double)]
        L.an5d_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.an5d_batch_info.argtypes = [ctypes.c_void_p, P(ctypes.c_int), P(ctypes.c_longlong)]
        L.an5d_batch_destroy.argtypes = [ctypes.c_void_p]
        L.an5d_autotune.argtypes = [P(Problem), ctypes.c_void_p, P(Layout), ctypes.c_int, ctypes.c_int,
                                    P(Config), P(ctypes.c_double)]
        L.an5d_model_estimate.argtypes = [P(Problem), P(Config), P(ctypes.c_double)]
        L.an5d_resolve_config.argtypes = [P(Problem), P(Config), P(Config)]
        L.an5d_tune_load.argtypes = [ctypes.c_char_p]
        L.an5d_tune_save.argtypes = [ctypes.c_char_p]
        L.an5d_cpu_forward.argtypes = [P(Problem), ctypes.c_void_p, ctypes.c_int, P(ctypes.c_double)]
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise An5dError(rc, lib().an5d_last_error().decode())
    return rc


def stencil_id(name: str) -> int:
    return _check(lib().an5d_stencil_id(name.encode()))


def info(name):
    st = StencilInfo()
    _check(lib().an5d_stencil_get_info(stencil_id(name), ctypes.byref(st)))
    return dict(name=st.name.decode(), dim=st.dim, rad=st.rad, fpp=st.fpp, ntaps=st.ntaps,
                divisor=st.divisor, linear=bool(st.linear))


def stencil_names():
    out = []
    for i in range(lib().an5d_stencil_count()):
        st = StencilInfo()
        _check(lib().an5d_stencil_get_info(i, ctypes.byref(st)))
        out.append(st.name.decode())
    return out


def taps(name):
    n = info(name)["ntaps"]
    offs = (ctypes.c_int * (4 * n))()
    coefs = (ctypes.c_double * n)()
    _check(lib().an5d_stencil_taps(stencil_id(name), offs, coefs, n))
    return np.array(list(offs), dtype=np.int64).reshape(n, 4), np.array(list(coefs))


def variants(name, dtype=FLOAT32):
    sid = stencil_id(name)
    out = []
    for v in range(_check(lib().an5d_variant_count(sid, dtype))):
        vi = VariantInfo()
        _check(lib().an5d_variant_get_info(sid, dtype, v, ctypes.byref(vi)))
        out.append(dict(index=v, kind=vi.kind, ck=vi.ck, cj=vi.cj, nw=vi.nw, btmax=vi.btmax,
                        lds_bytes=vi.lds_bytes, regcap=vi.regcap))
    return out


def schedule(T, bt):
    n = _check(lib().an5d_schedule(int(T), int(bt), None, 0))
    buf = (ctypes.c_int * max(n, 1))()
    _check(lib().an5d_schedule(int(T), int(bt), buf, n))
    return list(buf)[:n]


def _dtype_code(dt):
    dt = np.dtype(dt)
    if dt == np.float32:
        return FLOAT32
    if dt == np.float64:
        return FLOAT64
    raise ValueError(f"unsupported dtype {dt}")


def grid_extents(name, n):
    """Full extents (interior + halo) of one buffer for interior extents n."""
    d = info(name)
    if len(n) != d["dim"]:
        raise ValueError(f"{name} is {d['dim']}D: n must have {d['dim']} extents, got {list(n)}")
    if any(int(x) < 1 for x in n):
        raise ValueError(f"interior extents must be >= 1, got {list(n)}")
    return [int(x) + 2 * d["rad"] for x in n]


def _check_host(A, name, n, buffers):
    """A host numpy array holding `buffers` grids of the reference layout."""
    if not isinstance(A, np.ndarray):
        raise TypeError("expected a numpy array")
    _dtype_code(A.dtype)
    if not A.flags.c_contiguous:
        raise ValueError("host array must be C-contiguous")
    want = tuple(([buffers] if buffers else []) + grid_extents(name, n))
    if A.shape != want:
        raise ValueError(f"host array shape {A.shape}, expected {want}")


def _check_device(t, what, shape=None, dtype=None, contiguous=False):
    """A torch tensor on the current GPU, of the given shape and dtype, with
    positive strides that do not overlap (unit last stride)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{what} must be a GPU tensor (got device {t.device})")
    if t.device.index != torch.cuda.current_device():
        raise ValueError(f"{what} is on {t.device}, the current device is cuda:{torch.cuda.current_device()}")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{what} must be {dtype}, got {t.dtype}")
    if t.dtype not in (torch.float32, torch.float64):
        raise ValueError(f"{what} must be float32 or float64, got {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what} shape {tuple(t.shape)}, expected {tuple(shape)}")
    if contiguous and not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    st, sh = t.stride(), t.shape
    if st[-1] != 1:
        raise ValueError(f"{what}: last dimension must be contiguous")
    for i in range(len(sh) - 1):
        if st[i] < st[i + 1] * sh[i + 1]:
            raise ValueError(f"{what}: strides {st} overlap or are not outermost-first")


def _check_pair(t, name, n, contiguous=False):
    """Both buffers of a grid: a [2, D0, (D1,) D2] device tensor."""
    _check_device(t, "grid tensor", shape=[2] + grid_extents(name, n), contiguous=contiguous)


def make_problem(name, n, T, dtype, coefs=None, divisor=-1.0):
    p = Problem()
    p.stencil = stencil_id(name)
    p.dtype = _dtype_code(dtype)
    for d, v in enumerate(n):
        p.n[d] = int(v)
    p.timesteps = int(T)
    keep = None
    if coefs is not None:
        keep = np.ascontiguousarray(np.asarray(coefs, dtype=np.float64)).ravel()
        d = info(name)
        want = d["ntaps"] if d["linear"] else 2
        if keep.size != want:
            raise ValueError(f"{name} takes {want} coefficients, got {keep.size}")
        p.coefs = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    p.divisor = float(divisor)
    return p, keep


def make_config(bt=0, variant=-1, sl=0, mode=EXACT):
    c = Config()
    c.bt, c.variant, c.sl, c.mode = int(bt), int(variant), int(sl), int(mode)
    return c


def forward_host(name, A, n, T, coefs=None, divisor=-1.0, bt=0, variant=-1, sl=0, mode=EXACT):
    """GPU forward on a host numpy array A[2, ...] (reference layout), in place; returns ms."""
    _check_host(A, name, n, 2)
    p, keep = make_problem(name, n, T, A.dtype, coefs, divisor)
    cfg = make_config(bt, variant, sl, mode)
    ms = ctypes.c_double()
    _check(lib().an5d_forward_host(ctypes.byref(p), A.ctypes.data, ctypes.byref(cfg), ctypes.byref(ms)))
    del keep
    return ms.value


def cpu_forward(name, A, n, T, coefs=None, divisor=-1.0, threads=0):
    """The library's CPU path (OpenMP loop) on a host array, in place; returns ms."""
    _check_host(A, name, n, 2)
    p, keep = make_problem(name, n, T, A.dtype, coefs, divisor)
    ms = ctypes.c_double()
    _check(lib().an5d_cpu_forward(ctypes.byref(p), A.ctypes.data, int(threads), ctypes.byref(ms)))
    del keep
    return ms.value


def kernel_stencil(name, A, compsize, T, scop=True):
    """The drop-in entry (common.h:42 signature) on a host array; returns ms."""
    _check_host(A, name, [compsize] * info(name)["dim"], 2)
    ms = ctypes.c_double()
    _check(lib().an5d_kernel_stencil(stencil_id(name), _dtype_code(A.dtype), A.ctypes.data,
                                     int(compsize), int(T), 1 if scop else 0, ctypes.byref(ms)))
    return ms.value


def init_grid(name, n, dtype=np.float32, mode=0, seed_base=None):
    """Both buffers of a host grid for interior extents n: mode 0 is the
    reference's init_grid (seed i + compsize), mode 1 uniform [0, 1)."""
    d = info(name)
    shape = (2,) + tuple(int(x) + 2 * d["rad"] for x in n)
    A = np.empty(shape, dtype=dtype)
    if seed_base is None:
        seed_base = int(n[0])
    _check(lib().an5d_init_grid(_dtype_code(dtype), A.ctypes.data, A[0].size, int(seed_base), int(mode)))
    return A


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def layout_of(t, dim):
    """Strides (elements) of a torch tensor shaped [2, D0, (D1,) D2] with unit last stride."""
    st = t.stride()
    if st[-1] != 1:
        raise ValueError("last dimension must be contiguous")
    L = Layout()
    L.buf_stride = st[0]
    L.s0 = st[1]
    L.s1 = st[2] if dim == 3 else st[1]
    return L


def forward_device(name, t, n, T, coefs=None, divisor=-1.0, bt=0, variant=-1, sl=0, stream=None,
                   mode=EXACT):
    """Enqueue T steps on a device tensor t[2, ...]; returns immediately."""
    d = info(name)
    _check_pair(t, name, n)
    p, keep = make_problem(name, n, T, _np_dtype(t), coefs, divisor)
    cfg = make_config(bt, variant, sl, mode)
    L = layout_of(t, d["dim"])
    _check(lib().an5d_forward_device(ctypes.byref(p), ctypes.c_void_p(t.data_ptr()), ctypes.byref(L),
                                     ctypes.byref(cfg), _stream_ptr(stream)))
    del keep


def launch(name, t, n, steps, c0=0, coefs=None, divisor=-1.0, bt=0, variant=-1, sl=0, stream=None,
           mode=EXACT):
    """One launch of `steps` fused steps from buffer c0%2 to (c0+1)%2."""
    d = info(name)
    _check_pair(t, name, n)
    p, keep = make_problem(name, n, steps, _np_dtype(t), coefs, divisor)
    cfg = make_config(bt if bt else steps, variant, sl, mode)
    L = layout_of(t, d["dim"])
    _check(lib().an5d_launch(ctypes.byref(p), ctypes.c_void_p(t.data_ptr()), ctypes.byref(L),
                             ctypes.byref(cfg), int(steps), int(c0), _stream_ptr(stream)))
    del keep


def _np_dtype(t):
    import torch
    if t.dtype == torch.float32:
        return np.float32
    if t.dtype == torch.float64:
        return np.float64
    raise ValueError(f"unsupported tensor dtype {t.dtype}")


def adjoint_host(name, u0, target, n, T, coefs=None, divisor=-1.0, bt=0, variant=-1, sl=0):
    """Loss 1/2||u_T - target||^2 (interior), dL/du0 (interior; halo 0) and
    dL/dcoefs on the GPU.  u0, target: single grids (halo included).
    Returns (loss, grad, gcoef, ms_fwd, ms_total)."""
    u0 = np.ascontiguousarray(u0)
    target = np.ascontiguousarray(target, dtype=u0.dtype)
    _check_host(u0, name, n, 0)
    _check_host(target, name, n, 0)
    d = info(name)
    p, keep = make_problem(name, n, T, u0.dtype, coefs, divisor)
    cfg = make_config(bt, variant, sl)
    grad = np.zeros_like(u0)
    nc = d["ntaps"] if d["linear"] else 2
    gc = np.zeros(nc, dtype=np.float64)
    loss = ctypes.c_double()
    mf, mt = ctypes.c_double(), ctypes.c_double()
    _check(lib().an5d_adjoint_host(ctypes.byref(p), u0.ctypes.data, target.ctypes.data, grad.ctypes.data,
                                   gc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(loss),
                                   ctypes.byref(cfg), ctypes.byref(mf), ctypes.byref(mt)))
    del keep
    return loss.value, grad, gc, mf.value, mt.value


class AdjointWorkspace:
    """Device workspace for adjoint_device (checkpoints + two adjoint buffers)."""

    def __init__(self, name, n, T, dtype, bt=0, variant=-1, sl=0):
        import torch
        p, _ = make_problem(name, n, T, dtype)
        cfg = make_config(bt, variant, sl)
        nbytes = ctypes.c_size_t()
        _check(lib().an5d_adjoint_workspace(ctypes.byref(p), ctypes.byref(cfg), ctypes.byref(nbytes)))
        self.nbytes = nbytes.value
        self.buf = torch.zeros(self.nbytes, dtype=torch.uint8, device="cuda")  # zeroed once


def _check_adjoint_args(name, n, u0, target, grad, gcoef, loss):
    import torch
    ext = grid_extents(name, n)
    _check_device(u0, "u0", shape=ext, contiguous=True)
    _check_device(target, "target", shape=ext, dtype=u0.dtype, contiguous=True)
    _check_device(grad, "grad", shape=ext, dtype=u0.dtype, contiguous=True)
    d = info(name)
    ng = d["ntaps"] if d["linear"] else 2
    _check_device(gcoef, "gcoef", shape=[ng], dtype=torch.float64, contiguous=True)
    _check_device(loss, "loss", dtype=torch.float64, contiguous=True)
    if loss.numel() < 1:
        raise ValueError("loss needs one float64 element")


def adjoint_device(name, u0, target, grad, gcoef, loss, n, T, ws, coefs=None, divisor=-1.0,
                   bt=0, variant=-1, sl=0, stream=None, time_forward=False):
    # ws must have been sized for the same bt
    """Device tensors: u0/target/grad single grids, gcoef/loss float64 tensors."""
    _check_adjoint_args(name, n, u0, target, grad, gcoef, loss)
    if not isinstance(ws, AdjointWorkspace):
        raise TypeError("ws must be an AdjointWorkspace")
    p, keep = make_problem(name, n, T, _np_dtype(u0), coefs, divisor)
    cfg = make_config(bt, variant, sl)
    mf = ctypes.c_double()
    _check(lib().an5d_adjoint_device(ctypes.byref(p), ctypes.c_void_p(u0.data_ptr()),
                                     ctypes.c_void_p(target.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
                                     ctypes.c_void_p(gcoef.data_ptr()), ctypes.c_void_p(loss.data_ptr()),
                                     ctypes.c_void_p(ws.buf.data_ptr()), ws.nbytes, ctypes.byref(cfg),
                                     _stream_ptr(stream), ctypes.byref(mf) if time_forward else None))
    del keep
    return mf.value if time_forward else None


def adjoint_device_noalloc(name, u0, target, grad, gcoef, loss, n, T, coefs=None, divisor=-1.0,
                           bt=0, variant=-1, sl=0, stream=None):
    """adjoint_device without a workspace: the library's fixed device scratch
    array holds it (no allocation; problems beyond its size are rejected)."""
    _check_adjoint_args(name, n, u0, target, grad, gcoef, loss)
    p, keep = make_problem(name, n, T, _np_dtype(u0), coefs, divisor)
    cfg = make_config(bt, variant, sl)
    _check(lib().an5d_adjoint_device_noalloc(ctypes.byref(p), ctypes.c_void_p(u0.data_ptr()),
                                             ctypes.c_void_p(target.data_ptr()),
                                             ctypes.c_void_p(grad.data_ptr()),
                                             ctypes.c_void_p(gcoef.data_ptr()),
                                             ctypes.c_void_p(loss.data_ptr()), ctypes.byref(cfg),
                                             _stream_ptr(stream)))
    del keep


class Batch:
    """Batched forward over many device grids (an5d_batch_create).  `items` is
    a list of (name, n, T, tensor[2, ...], coefs, divisor); bt / sl apply to
    every item (0: the model prior / the batch-wide segment length)."""

    def __init__(self, items, bt=0, sl=0):
        n = len(items)
        probs = (Problem * n)()
        bufs = (ctypes.c_void_p * n)()
        cfgs = (Config * n)()
        self._keep = []
        for i, (name, ext, T, t, coefs, divisor) in enumerate(items):
            _check_pair(t, name, ext, contiguous=True)
            p, keep = make_problem(name, ext, T, _np_dtype(t), coefs, divisor)
            probs[i] = p
            # the captured graph holds the raw pointers: the tensors must outlive it
            self._keep.append((keep, t))
            bufs[i] = t.data_ptr()
            cfgs[i] = make_config(bt, -1, sl)
        h = ctypes.c_void_p()
        _check(lib().an5d_batch_create(ctypes.byref(h), n, probs, bufs, cfgs))
        self.handle = h
        self._info()

    def _info(self):
        lv, un = ctypes.c_int(), ctypes.c_longlong()
        _check(lib().an5d_batch_info(self.handle, ctypes.byref(lv), ctypes.byref(un)))
        self.levels, self.units = lv.value, un.value

    def run(self, stream=None):
        _check(lib().an5d_batch_run(self.handle, _stream_ptr(stream)))

    def tune(self, reps=2):
        """Measured choice of the batch-wide segment length; returns (sl, ms)."""
        sl, ms = ctypes.c_int(), ctypes.c_double()
        _check(lib().an5d_batch_tune(self.handle, int(reps), ctypes.byref(sl), ctypes.byref(ms)))
        self._info()
        return sl.value, ms.value

    def close(self):
        if self.handle:
            lib().an5d_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchAdjoint(Batch):
    """Batched forward + adjoint (an5d_batch_adjoint_create).  `items` is a
    list of (name, n, T, u0, target, grad, gcoef, loss, coefs, divisor): u0,
    target, grad single device grids (FP32), gcoef / loss float64 device
    tensors of ntaps / 1 elements.  bt / sl apply to every item (0: the
    deepest compiled block / the batch-wide segment length)."""

    def __init__(self, items, bt=0, sl=0):
        n = len(items)
        probs = (Problem * n)()
        io = (BatchGrad * n)()
        cfgs = (Config * n)()
        self._keep = []
        for i, (name, ext, T, u0, target, grad, gcoef, loss, coefs, divisor) in enumerate(items):
            _check_adjoint_args(name, ext, u0, target, grad, gcoef, loss)
            p, keep = make_problem(name, ext, T, _np_dtype(u0), coefs, divisor)
            probs[i] = p
            self._keep.append((keep, u0, target, grad, gcoef, loss))
            io[i].u0, io[i].target, io[i].grad = u0.data_ptr(), target.data_ptr(), grad.data_ptr()
            io[i].gcoef, io[i].loss = gcoef.data_ptr(), loss.data_ptr()
            cfgs[i] = make_config(bt, -1, sl)
        h = ctypes.c_void_p()
        _check(lib().an5d_batch_adjoint_create(ctypes.byref(h), n, probs, io, cfgs))
        self.handle = h
        self._info()


def autotune(name, n, dtype=np.float32, mode=EXACT, max_measured=24, tensor=None, timesteps=1, coefs=None,
             divisor=-1.0):
    """Rank configurations with the model, time the best on the GPU, record
    the winner.  `timesteps`: the length of the run being tuned for (the
    candidates are timed on the field halfway through it); `coefs` /
    `divisor`: the run's coefficients (they decide, e.g., whether the
    shared-product form applies).  Returns (config dict, GStencil/s)."""
    p, keep = make_problem(name, n, timesteps, dtype, coefs, divisor)
    best = Config()
    gs = ctypes.c_double()
    ptr, lay = None, None
    if tensor is not None:
        _check_pair(tensor, name, n)
        if _np_dtype(tensor) != np.dtype(dtype):
            raise ValueError("tensor dtype differs from dtype")
        ptr = ctypes.c_void_p(tensor.data_ptr())
        lay = ctypes.byref(layout_of(tensor, info(name)["dim"]))
    _check(lib().an5d_autotune(ctypes.byref(p), ptr, lay, int(mode), int(max_measured),
                               ctypes.byref(best), ctypes.byref(gs)))
    return dict(bt=best.bt, variant=best.variant, sl=best.sl, mode=best.mode), gs.value


def model_estimate(name, n, dtype=np.float32, bt=1, variant=0, sl=0, mode=EXACT):
    p, keep = make_problem(name, n, 1, dtype)
    cfg = make_config(bt, variant, sl, mode)
    gs = ctypes.c_double()
    _check(lib().an5d_model_estimate(ctypes.byref(p), ctypes.byref(cfg), ctypes.byref(gs)))
    return gs.value


def resolve_config(name, n, dtype=np.float32, bt=0, variant=-1, sl=0, mode=EXACT):
    """The (variant, bt, sl) the forward entry points run for this problem:
    explicit values win, then the tuned table, then the defaults."""
    p, keep = make_problem(name, n, 1, dtype)
    cfg = make_config(bt, variant, sl, mode)
    out = Config()
    _check(lib().an5d_resolve_config(ctypes.byref(p), ctypes.byref(cfg), ctypes.byref(out)))
    return dict(bt=out.bt, variant=out.variant, sl=out.sl, mode=out.mode)


def tune_save(path=None):
    _check(lib().an5d_tune_save(path.encode() if path else None))


def tune_load(path=None):
    _check(lib().an5d_tune_load(path.encode() if path else None))


def tune_clear():
    _check(lib().an5d_tune_clear())
