"""GPU parity of the forward sweep: the HIP kernels against the CPU oracle,
bit for bit (every stencil x dtype x kernel shape x several bt), tiling
invariance, the kernel_stencil drop-in, device-resident layouts and the
golden fixtures made from the reference's own programs."""
import os

import numpy as np
import pytest

import oracle_lib as O
import pyan5d

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def small_n(name):
    d = O.info(name)
    if d["dim"] == 2:
        return [37, 300]  # two column tiles, ragged edges
    if d["ntaps"] > 100:
        return [6, 7, 9]
    return [11, 45, 150]


def rand_grid(name, n, dt, seed):
    d = O.info(name)
    rng = np.random.default_rng(seed)
    shape = tuple(x + 2 * d["rad"] for x in n)
    u = rng.random(shape).astype(dt)
    return np.stack([u, u])


def same(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def cases():
    out = []
    for name in O.NAMES:
        for dt in (np.float32, np.float64):
            for v in pyan5d.variants(name, 0 if dt == np.float32 else 1):
                bts = sorted({1, v["btmax"], max(1, v["btmax"] // 2)})
                for bt in bts:
                    out.append((name, dt, v["index"], bt))
    return out


@pytest.mark.parametrize("name,dt,variant,bt", cases())
def test_forward_bitwise_vs_oracle(name, dt, variant, bt):
    n = small_n(name)
    T = 2 * bt + 1 if O.info(name)["ntaps"] <= 100 else 2
    print(f"case {name} {np.dtype(dt).name} variant={variant} bt={bt} n={n} T={T}", flush=True)
    A = rand_grid(name, n, dt, seed=bt * 7 + variant)
    B = A.copy()
    pyan5d.forward_host(name, A, n, T, bt=bt, variant=variant)
    O.run(name, B, n, T)
    assert same(A[T % 2], B[T % 2])


@pytest.mark.parametrize("name,dt", [("j2d5pt", np.float32), ("star3d1r", np.float32),
                                     ("box3d1r", np.float64), ("gradient2d", np.float32)])
def test_tiling_invariance(name, dt):
    d = O.info(name)
    n = [300, 1030] if d["dim"] == 2 else [40, 130, 200]
    T = 13
    base = rand_grid(name, n, dt, seed=3)
    results = []
    for v in pyan5d.variants(name, 0 if dt == np.float32 else 1):
        for bt in range(1, v["btmax"] + 1):
            for sl in (0, 7, 64):
                A = base.copy()
                print(f"case {name} variant={v['index']} bt={bt} sl={sl}", flush=True)
                pyan5d.forward_host(name, A, n, T, bt=bt, variant=v["index"], sl=sl)
                results.append(A[T % 2].copy())
    B = base.copy()
    pyan5d.cpu_forward(name, B, n, T)
    for r in results:
        assert same(r, B[T % 2])


def test_kernel_stencil_dropin_config0_size():
    """j2d5pt FP32 4096^2, 100 steps, reference init; GPU path vs
    the CPU path (OpenMP loop), whole grid, both buffers' halo untouched."""
    print("case j2d5pt 4096^2 T=100 drop-in", flush=True)
    A = O.init_grid("j2d5pt", 4096, np.float32)
    B = A.copy()
    halo = A[:, 0, :].copy()
    pyan5d.kernel_stencil("j2d5pt", A, 4096, 100, scop=True)
    pyan5d.kernel_stencil("j2d5pt", B, 4096, 100, scop=False)
    assert same(A[0], B[0])
    assert np.array_equal(A[:, 0, :], halo)


@pytest.mark.parametrize("name", O.NAMES)
@pytest.mark.parametrize("dt,tname", [(np.float32, "float"), (np.float64, "double")])
def test_golden_fixture_on_gpu(name, dt, tname):
    g = np.load(os.path.join(GOLD, f"stencil_{name}_{tname}.npz"))
    n, T = int(g["n"]), int(g["T"])
    print(f"case golden {name} {tname} n={n} T={T}", flush=True)
    d = O.info(name)
    A = O.init_grid(name, n, dt)
    pyan5d.kernel_stencil(name, A, n, T, scop=True)
    assert same(A[T % 2], g["out"][T % 2])


def test_forward_device_layouts(torch_cuda):
    torch = torch_cuda
    name, n, T = "star3d1r", [20, 33, 70], 9
    d = O.info(name)
    base = rand_grid(name, n, np.float32, seed=11)
    B = base.copy()
    O.run(name, B, n, T)
    # reference layout
    t = torch.from_numpy(base.copy()).cuda()
    print("case forward_device reference layout", flush=True)
    pyan5d.forward_device(name, t, n, T, bt=3)
    torch.cuda.synchronize()
    assert same(t.cpu().numpy()[T % 2], B[T % 2])
    # padded rows and planes (strided view)
    D = [x + 2 * d["rad"] for x in n]
    big = torch.zeros((2, D[0], D[1] + 3, D[2] + 29), dtype=torch.float32, device="cuda")
    view = big[:, :, :D[1], :D[2]]
    view.copy_(torch.from_numpy(base))
    print("case forward_device strided layout", flush=True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        pyan5d.forward_device(name, view, n, T, bt=2, stream=s)
    s.synchronize()
    assert same(view.cpu().numpy()[T % 2], B[T % 2])


def test_single_launch_entry(torch_cuda):
    torch = torch_cuda
    name, n = "j2d9pt", [50, 400]
    base = rand_grid(name, n, np.float32, seed=2)
    B = base.copy()
    O.run(name, B, n, 5)
    t = torch.from_numpy(base.copy()).cuda()
    print("case single launches 3+2", flush=True)
    pyan5d.launch(name, t, n, 3, c0=0)
    pyan5d.launch(name, t, n, 2, c0=1)
    torch.cuda.synchronize()
    assert same(t.cpu().numpy()[0], B[1])


@pytest.mark.parametrize("name", ["j2d5pt", "star3d1r", "box2d4r", "j3d27pt", "gradient2d"])
def test_other_coefficients_and_divisor(name):
    d = O.info(name)
    n = [30, 270] if d["dim"] == 2 else [10, 40, 140]
    rng = np.random.default_rng(1)
    if d["linear"]:
        c = rng.random(d["ntaps"]) + 0.01
        c = c / c.sum()
        div = 0.0
    else:
        c = [1e-4, 0.01]
        div = -1.0
    for dt in (np.float32, np.float64):
        A = rand_grid(name, n, dt, seed=4)
        B = A.copy()
        print(f"case coefs {name} {np.dtype(dt).name}", flush=True)
        pyan5d.forward_host(name, A, n, 6, coefs=c, divisor=div)
        O.run(name, B, n, 6, coefs=c, divisor=div)
        assert same(A[0], B[0])


def test_rectangular_and_tiny_grids():
    for name, n in [("star2d1r", [1, 1]), ("star2d2r", [3, 700]), ("j2d5pt", [900, 5]),
                    ("star3d1r", [1, 2, 3]), ("star3d1r", [70, 3, 9]), ("box3d1r", [5, 60, 2])]:
        A = rand_grid(name, n, np.float32, seed=9)
        B = A.copy()
        print(f"case rect {name} {n}", flush=True)
        pyan5d.forward_host(name, A, n, 7)
        O.run(name, B, n, 7)
        assert same(A[1], B[1])


@pytest.mark.parametrize("name", ["star3d2r", "star3d3r", "star3d4r"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_plane_kernel_ragged_grids(name, dt):
    """The LDS-staged one-step kernel (kind 2) on grids smaller than one tile,
    ragged tile edges in both in-plane dims and segments shorter than 2R+1."""
    vs = [v for v in pyan5d.variants(name, 0 if dt == np.float32 else 1) if v["kind"] == 2]
    assert len(vs) == 1
    v = vs[0]
    for n, sl in [([1, 1, 1], 0), ([3, 17, 130], 1), ([9, 5, 300], 2), ([20, 33, 257], 7), ([13, 16, 128], 0)]:
        A = rand_grid(name, n, dt, seed=sum(n))
        B = A.copy()
        T = 3
        print(f"case plane {name} {np.dtype(dt).name} n={n} sl={sl}", flush=True)
        pyan5d.forward_host(name, A, n, T, variant=v["index"], sl=sl)
        O.run(name, B, n, T)
        assert same(A[T % 2], B[T % 2])
        C = rand_grid(name, n, dt, seed=sum(n))
        D = C.copy()
        pyan5d.forward_host(name, C, n, T, variant=v["index"], sl=sl, mode=pyan5d.FAST)
        O.run(name, D, n, T)
        got, ref = C[T % 2].astype(np.float64), D[T % 2].astype(np.float64)
        rel = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), np.finfo(dt).tiny))
        assert rel <= (1e-5 if dt == np.float32 else 1e-12), rel


@pytest.mark.parametrize("name", ["j3d27pt", "box3d1r"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_box3d_plane_partial_sums_ragged_grids(name, dt):
    """The plane-streamed partial-sum kernels of the radius-1 3D boxes (kind 2
    for a box): bit for bit the oracle on grids smaller than one tile, ragged
    tiles in both in-plane dims, segments of 1-7 planes and the default, in
    exact mode; within 1e-5 / 1e-12 relative in FMA mode; j3d27pt also with
    its /159 divisor and other coefficients."""
    vs = [v for v in pyan5d.variants(name, 0 if dt == np.float32 else 1) if v["kind"] == 2]
    assert len(vs) == 2
    for v in vs:
        for n, sl in [([1, 1, 1], 0), ([3, 17, 130], 1), ([9, 5, 300], 2), ([20, 33, 257], 7),
                      ([13, 16, 128], 0), ([40, 70, 260], 0)]:
            T = 4
            print(f"case box3d partial sums {name} {np.dtype(dt).name} variant {v['index']} n={n} sl={sl}",
                  flush=True)
            A = rand_grid(name, n, dt, seed=sum(n) + v["index"])
            B = A.copy()
            pyan5d.forward_host(name, A, n, T, variant=v["index"], sl=sl)
            O.run(name, B, n, T)
            assert same(A[T % 2], B[T % 2])
            C = rand_grid(name, n, dt, seed=sum(n) + 1)
            D = C.copy()
            pyan5d.forward_host(name, C, n, T, variant=v["index"], sl=sl, mode=pyan5d.FAST)
            O.run(name, D, n, T)
            got, ref = C[T % 2].astype(np.float64), D[T % 2].astype(np.float64)
            rel = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), np.finfo(dt).tiny))
            assert rel <= (1e-5 if dt == np.float32 else 1e-12), rel
        if name == "j3d27pt":
            n, T = [30, 40, 200], 5
            rng = np.random.default_rng(3)
            coefs = rng.random(27).astype(np.float64)
            for div in (159.0, 0.0):
                print(f"case box3d partial sums {name} {np.dtype(dt).name} variant {v['index']} "
                      f"divisor {div}", flush=True)
                A = rand_grid(name, n, dt, seed=9)
                B = A.copy()
                pyan5d.forward_host(name, A, n, T, coefs=coefs, divisor=div, variant=v["index"])
                O.run(name, B, n, T, coefs=coefs, divisor=div)
                assert same(A[T % 2], B[T % 2])


@pytest.mark.parametrize("name", ["j2d5pt", "j2d9pt", "star2d3r", "box2d2r", "gradient2d",
                                  "star3d1r", "box3d1r", "j3d27pt", "star3d4r"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_fast_mode_within_tolerance(name, dt):
    """FMA mode (divisor folded into the coefficients): within 1e-5 (FP32) /
    1e-12 (FP64) relative of the oracle, relative to max(|ref|, smallest normal)."""
    d = O.info(name)
    n = [200, 700] if d["dim"] == 2 else [30, 60, 140]
    T = 40 if d["dim"] == 2 else 12
    A = rand_grid(name, n, dt, seed=21)
    B = A.copy()
    vs = pyan5d.variants(name, 0 if dt == np.float32 else 1)
    bt = min(3, vs[0]["btmax"])
    print(f"case fast {name} {np.dtype(dt).name} bt={bt}", flush=True)
    pyan5d.forward_host(name, A, n, T, bt=bt, mode=pyan5d.FAST)
    O.run(name, B, n, T)
    got, ref = A[T % 2].astype(np.float64), B[T % 2].astype(np.float64)
    floor = np.finfo(dt).tiny
    rel = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor))
    assert rel <= (1e-5 if dt == np.float32 else 1e-12), rel


SHARE_CASES = [("star3d1r", np.float32), ("star3d1r", np.float64), ("star2d1r", np.float32),
               ("star2d1r", np.float64), ("j2d5pt", np.float32), ("j2d5pt", np.float64)]


@pytest.mark.parametrize("name,dt", SHARE_CASES)
def test_shared_product_form_bitwise(name, dt):
    """Radius-1 stars whose face taps carry one coefficient run the
    shared-product kernels (each c * u formed once per cell): bit for bit
    the oracle and the general kernels, for every shape and depth."""
    d = O.info(name)
    n = [37, 70, 300] if d["dim"] == 3 else [57, 700]
    T = 11
    ctr = [i for i, t in enumerate(O.taps(name)[0]) if not t[1:].any()][0]
    coefs = np.full(d["ntaps"], 0.1)
    coefs[ctr] = 0.4
    div = 118.0 if d["divisor"] else 0.0
    rng = np.random.default_rng(3)
    u = rng.random([x + 2 * d["rad"] for x in n]).astype(dt)
    A0 = np.stack([u, u])  # both buffers alike: the halo is one fixed boundary
    ref = A0.copy()
    O.run(name, ref, n, T, coefs=coefs, divisor=div)
    dcode = 0 if dt == np.float32 else 1
    for v in pyan5d.variants(name, dcode):
        if v["kind"] != 0:
            continue
        for bt in sorted({1, 2, min(3, v["btmax"]), v["btmax"]}):
            print(f"case shared-product {name} {np.dtype(dt).name} variant {v['index']} bt {bt}", flush=True)
            A = A0.copy()
            pyan5d.forward_host(name, A, n, T, coefs=coefs, divisor=div, bt=bt, variant=v["index"])
            B = A0.copy()
            pyan5d.forward_host(name, B, n, T, coefs=coefs, divisor=div, bt=bt, variant=v["index"],
                                mode=pyan5d.EXACT_GENERAL)
            assert np.array_equal(A[T % 2].view(np.uint8), ref[T % 2].view(np.uint8)), (v, bt)
            assert np.array_equal(B[T % 2].view(np.uint8), ref[T % 2].view(np.uint8)), (v, bt)


def test_forward_host_from_several_threads():
    """Host calls from several threads on one device, with grids of
    different sizes (the staging grid grows while others are running): every
    result equals the oracle's."""
    import threading
    cases = [([300, 700], 9), ([900, 1200], 7), ([150, 2000], 5), ([1100, 900], 8)]
    print("case threads forward_host j2d5pt x4", flush=True)
    # both buffers hold the same field (the halo is read from either one)
    inputs = []
    for i, (n, _) in enumerate(cases):
        u = np.random.default_rng(i).random((n[0] + 2, n[1] + 2)).astype(np.float32)
        inputs.append(np.stack([u, u]))
    outs = [a.copy() for a in inputs]
    errors = []

    def work(i):
        try:
            n, T = cases[i]
            pyan5d.forward_host("j2d5pt", outs[i], n, T)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for (n, T), a, o in zip(cases, inputs, outs):
        ref = a.copy()
        O.run("j2d5pt", ref, n, T)
        assert np.array_equal(o[T % 2], ref[T % 2])
