"""an5d_adjoint_device_noalloc (the adjoint with its workspace in the library's
fixed device scratch array) against an5d_adjoint_device with a caller
workspace: bit for bit on consecutive calls, the second one smaller, and
against the FP64 oracle adjoint at the FP32 tolerance."""
import numpy as np
import pytest

import oracle_lib as O
import pyan5d


def _problem(name, n, dt, seed):
    d = O.info(name)
    rng = np.random.default_rng(seed)
    shape = tuple(x + 2 * d["rad"] for x in n)
    u0 = rng.random(shape).astype(dt)
    target = rng.random(shape).astype(dt)
    if d["linear"]:
        c = 1.0 - rng.random(d["ntaps"])
        return u0, target, c / c.sum()
    return u0, target, np.array([1e-4, 0.01])


def _run(torch, fn, name, n, T, u0, target, c, div):
    d = O.info(name)
    tu, tt = torch.from_numpy(u0).cuda(), torch.from_numpy(target).cuda()
    tg = torch.full_like(tu, float("nan"))
    ng = d["ntaps"] if d["linear"] else 2
    gc = torch.full((ng,), float("nan"), dtype=torch.float64, device="cuda")
    loss = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    fn(tu, tt, tg, gc, loss)
    torch.cuda.synchronize()
    return tg.cpu().numpy(), gc.cpu().numpy(), float(loss.item())


@pytest.mark.gpu
def test_noalloc_adjoint_matches_workspace_entry(torch_cuda):
    torch = torch_cuda
    # (stencil, extents, T, dtype, divisor): the second call is smaller than the
    # first, the third switches dtype and stencil
    calls = [("j2d9pt-gol", [200, 300], 17, np.float32, -1.0),
             ("j2d9pt-gol", [64, 80], 9, np.float32, 0.0),
             ("gradient2d", [96, 40], 7, np.float64, -1.0)]
    for i, (name, n, T, dt, div) in enumerate(calls):
        u0, target, c = _problem(name, n, dt, seed=40 + i)
        print(f"case noalloc call {i}: {name} {n} T={T} {np.dtype(dt).name} div={div}", flush=True)
        ws = pyan5d.AdjointWorkspace(name, n, T, dt)

        def with_ws(tu, tt, tg, gc, loss):
            pyan5d.adjoint_device(name, tu, tt, tg, gc, loss, n, T, ws, coefs=c, divisor=div)

        def no_alloc(tu, tt, tg, gc, loss):
            pyan5d.adjoint_device_noalloc(name, tu, tt, tg, gc, loss, n, T, coefs=c, divisor=div)

        g0, gc0, l0 = _run(torch, with_ws, name, n, T, u0, target, c, div)
        g1, gc1, l1 = _run(torch, no_alloc, name, n, T, u0, target, c, div)
        assert np.array_equal(g0.view(np.uint8), g1.view(np.uint8)), (name, n, T)
        assert np.array_equal(gc0.view(np.uint64), gc1.view(np.uint64)), (name, gc0, gc1)
        assert l0 == l1
        ol, og, ogc = O.adjoint(name, u0.astype(np.float64), target.astype(np.float64), n, T,
                                coefs=c, divisor=div)
        tol = 2e-4 if dt == np.float32 else 1e-10
        assert abs(l1 - ol) <= tol * max(1.0, abs(ol))
        assert np.max(np.abs(g1 - og)) <= tol * np.max(np.abs(og))
        assert np.max(np.abs(gc1 - ogc)) <= tol * np.max(np.abs(ogc))
        del ws
        torch.cuda.empty_cache()


def test_noalloc_adjoint_rejects_oversize_problem():
    """Problems whose workspace exceeds the fixed scratch array are refused
    before any device call (runs without a GPU)."""
    import ctypes
    p, keep = pyan5d.make_problem("j2d5pt", [8192, 8192], 256, np.float32)
    cfg = pyan5d.make_config()
    rc = pyan5d.lib().an5d_adjoint_device_noalloc(ctypes.byref(p), ctypes.c_void_p(1), ctypes.c_void_p(1),
                                                  ctypes.c_void_p(1), ctypes.c_void_p(1), ctypes.c_void_p(1),
                                                  ctypes.byref(cfg), None)
    assert rc == -1
    msg = pyan5d.lib().an5d_last_error().decode()
    assert "at most 67108864" in msg, msg


def test_noalloc_wrapper_checks_tensors():
    import torch
    u = torch.zeros(6, 6)
    gc = torch.zeros(5, dtype=torch.float64)
    loss = torch.zeros(1, dtype=torch.float64)
    with pytest.raises(ValueError):
        pyan5d.adjoint_device_noalloc("j2d5pt", u, u, u, gc, loss, [4, 4], 3)
