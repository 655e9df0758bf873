"""GPU adjoint sweep against the FP64 oracle adjoint and central finite
differences (FP64, 64x64), for every 2D stencil and block depth."""
import numpy as np
import pytest

import oracle_lib as O
import pyan5d

pytestmark = pytest.mark.gpu

NAMES_2D = [n for n in O.NAMES if O.info(n)["dim"] == 2]


def problem(name, n, dt, seed):
    d = O.info(name)
    rng = np.random.default_rng(seed)
    shape = tuple(x + 2 * d["rad"] for x in n)
    u0 = rng.random(shape).astype(dt)
    target = rng.random(shape).astype(dt)
    if d["linear"]:
        c = 1.0 - rng.random(d["ntaps"])
        c = c / c.sum()
        return u0, target, c, 0.0
    return u0, target, np.array([1e-4, 0.01]), -1.0


def adj_btmax(name, dt):
    """Deepest compiled adjoint block (mirrors the kernel registry)."""
    d = O.info(name)
    if not d["linear"]:
        return 2
    if d["rad"] == 1:
        return 4 if d["ntaps"] > 5 else 5
    if d["rad"] == 2:
        return 2 if dt == np.float64 else 3
    return 1


def cases():
    out = []
    for name in NAMES_2D:
        for dt in (np.float32, np.float64):
            for bt in range(1, adj_btmax(name, dt) + 1):
                out.append((name, dt, bt))
    return out


@pytest.mark.parametrize("name,dt,bt", cases())
def test_adjoint_vs_oracle(name, dt, bt):
    n = [37, 300]
    T = 2 * bt + 1
    u0, target, c, div = problem(name, n, dt, seed=bt)
    print(f"case adjoint {name} {np.dtype(dt).name} bt={bt} T={T}", flush=True)
    loss, g, gc, _, _ = pyan5d.adjoint_host(name, u0, target, n, T, coefs=c, divisor=div, bt=bt)
    oloss, og, ogc = O.adjoint(name, u0.astype(np.float64), target.astype(np.float64), n, T,
                               coefs=c, divisor=div)
    tol = 1e-10 if dt == np.float64 else 2e-4
    assert abs(loss - oloss) <= tol * max(1.0, abs(oloss))
    assert np.max(np.abs(g - og)) <= tol * max(1e-30, np.max(np.abs(og)))
    assert np.max(np.abs(gc - ogc)) <= tol * max(1e-30, np.max(np.abs(ogc)))


def division_cases():
    out = []
    for name in ("j2d5pt", "j2d9pt", "j2d9pt-gol"):
        for dt in (np.float32, np.float64):
            for bt in sorted({1, adj_btmax(name, dt)}):
                out.append((name, dt, bt, "reference"))
        out.append((name, np.float32, 2, "baseline") if name == "j2d5pt" else None)
    return [c for c in out if c]


@pytest.mark.parametrize("name,dt,bt,preset", division_cases())
def test_adjoint_with_division_vs_oracle(name, dt, bt, preset):
    """The division path: the reference's own coefficients over 118
    (divisor -1: the folded taps s_q c_q / div, the coefficient gradient
    divided by div, the checkpointed forward's corrected division) and, for
    j2d5pt, the benchmark preset (S = 12.3, / 118)."""
    from pyan5d import presets
    n = [37, 300]
    T = 2 * bt + 1
    rng = np.random.default_rng(7 + bt)
    shape = tuple(x + 2 * O.info(name)["rad"] for x in n)
    u0 = rng.random(shape).astype(dt)
    target = rng.random(shape).astype(dt)
    c, div = presets.coefficients(name, preset)
    assert div != 0.0
    print(f"case adjoint-division {name} {np.dtype(dt).name} bt={bt} T={T} preset={preset}", flush=True)
    loss, g, gc, _, _ = pyan5d.adjoint_host(name, u0, target, n, T, coefs=c, divisor=div, bt=bt)
    oloss, og, ogc = O.adjoint(name, u0.astype(np.float64), target.astype(np.float64), n, T,
                               coefs=c, divisor=div)
    tol = 1e-10 if dt == np.float64 else 2e-4
    assert abs(loss - oloss) <= tol * max(1.0, abs(oloss))
    assert np.max(np.abs(g - og)) <= tol * max(1e-30, np.max(np.abs(og)))
    assert np.max(np.abs(gc - ogc)) <= tol * max(1e-30, np.max(np.abs(ogc)))


@pytest.mark.parametrize("name", ["j2d9pt-gol", "j2d9pt", "gradient2d", "j2d5pt"])
def test_adjoint_central_differences_fp64(name):
    """FD on a 64x64 FP64 instance through the GPU forward."""
    n, T = [64, 64], 12
    u0, target, c, div = problem(name, n, np.float64, seed=7)
    d = O.info(name)
    r = d["rad"]

    def loss_of(u, coefs):
        A = np.stack([u, u]).copy()
        pyan5d.forward_host(name, A, n, T, coefs=coefs, divisor=div)
        return 0.5 * float(np.sum((A[T % 2][r:-r, r:-r] - target[r:-r, r:-r]) ** 2))

    print(f"case fd {name}", flush=True)
    loss, g, gc, _, _ = pyan5d.adjoint_host(name, u0, target, n, T, coefs=c, divisor=div)
    assert abs(loss - loss_of(u0, c)) <= 1e-12 * loss
    rng = np.random.default_rng(1)
    eps = 1e-6
    for _ in range(8):
        x = (int(rng.integers(r, 64 + r)), int(rng.integers(r, 64 + r)))
        up, um = u0.copy(), u0.copy()
        up[x] += eps
        um[x] -= eps
        fd = (loss_of(up, c) - loss_of(um, c)) / (2 * eps)
        assert abs(fd - g[x]) <= 1e-6 * max(1.0, abs(fd)), (x, fd, g[x])
    for q in range(len(c)):
        h = eps * max(1.0, abs(c[q]))
        cp, cm = c.copy(), c.copy()
        cp[q] += h
        cm[q] -= h
        fd = (loss_of(u0, cp) - loss_of(u0, cm)) / (2 * h)
        assert abs(fd - gc[q]) <= 1e-5 * max(1.0, abs(fd)), (q, fd, gc[q])
    assert np.all(g[:r, :] == 0) and np.all(g[:, -r:] == 0)


def test_adjoint_device_entry(torch_cuda):
    torch = torch_cuda
    name, n, T = "j2d9pt-gol", [100, 700], 11
    u0, target, c, div = problem(name, n, np.float32, seed=3)
    ws = pyan5d.AdjointWorkspace(name, n, T, np.float32)
    tu, tt = torch.from_numpy(u0).cuda(), torch.from_numpy(target).cuda()
    tg = torch.zeros_like(tu)
    gc = torch.zeros(9, dtype=torch.float64, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    print("case adjoint_device", flush=True)
    pyan5d.adjoint_device(name, tu, tt, tg, gc, loss, n, T, ws, coefs=c, divisor=div)
    torch.cuda.synchronize()
    hl, hg, hgc, _, _ = pyan5d.adjoint_host(name, u0, target, n, T, coefs=c, divisor=div)
    assert np.array_equal(tg.cpu().numpy(), hg)
    assert np.allclose(gc.cpu().numpy(), hgc, rtol=1e-12, atol=0)


def test_autograd_function_gradcheck(torch_cuda):
    """torch.autograd.gradcheck (central differences, FP64) on a small grid,
    for both the initial state and the coefficients."""
    torch = torch_cuda
    from pyan5d.autograd import sweep
    for name in ("j2d9pt-gol", "gradient2d"):
        d = O.info(name)
        n, T = [12, 14], 5
        shape = [x + 2 * d["rad"] for x in n]
        gen = torch.Generator(device="cuda").manual_seed(0)
        u0 = torch.rand(shape, dtype=torch.float64, device="cuda", generator=gen).requires_grad_()
        if d["linear"]:
            c = torch.rand(d["ntaps"], dtype=torch.float64, device="cuda", generator=gen) / d["ntaps"]
            div = 0.0
        else:
            c = torch.tensor([1e-2, 0.5], dtype=torch.float64, device="cuda")
            div = -1.0
        c.requires_grad_()
        mask = torch.zeros(shape, dtype=torch.float64, device="cuda")
        r = d["rad"]
        mask[r:-r, r:-r] = 1.0
        fn = lambda u, cc: sweep(u, cc, name, n, T, div) * mask  # noqa: E731
        print(f"case gradcheck {name}", flush=True)
        # the halo is a boundary condition: only interior cells reach the
        # sweep from the perturbed input (gradcheck perturbs u0 in place, so
        # the halo comes from a separate copy)
        halo = u0.detach().clone() * (1 - mask)
        assert torch.autograd.gradcheck(lambda ui, cc: fn(halo + ui * mask, cc),
                                        (u0, c), eps=1e-6, atol=1e-7, rtol=1e-5, nondet_tol=1e-10)


def test_autograd_matches_oracle_fp32(torch_cuda):
    torch = torch_cuda
    from pyan5d.autograd import sweep
    name, n, T = "j2d9pt", [40, 300], 9
    u0, target, c, div = problem(name, n, np.float32, seed=4)
    tu = torch.from_numpy(u0).cuda().requires_grad_()
    tc = torch.from_numpy(c).cuda().requires_grad_()
    uT = sweep(tu, tc, name, n, T, div)
    r = O.info(name)["rad"]
    loss = 0.5 * ((uT - torch.from_numpy(target).cuda())[r:-r, r:-r] ** 2).sum()
    print("case autograd vs oracle", flush=True)
    loss.backward()
    ol, og, ogc = O.adjoint(name, u0.astype(np.float64), target.astype(np.float64), n, T, coefs=c, divisor=div)
    assert abs(loss.item() - ol) <= 1e-4 * ol
    assert np.max(np.abs(tu.grad.cpu().numpy() - og)) <= 2e-4 * np.max(np.abs(og))
    assert np.max(np.abs(tc.grad.cpu().numpy() - ogc)) <= 2e-4 * np.max(np.abs(ogc))


def test_adjoint_workspace_beyond_device_memory_is_rejected(torch_cuda):
    """One padded grid per bt-step block: 16384^2 FP32 over 20000 steps at
    bt=1 needs about 22 TB, which the library refuses with a message before
    anything is allocated."""
    print("case adjoint workspace too large", flush=True)
    with pytest.raises(pyan5d.An5dError, match="exceeds the device"):
        pyan5d.AdjointWorkspace("j2d5pt", [16384, 16384], 20000, np.float32, bt=1)
