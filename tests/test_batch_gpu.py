"""Batched forward, and batched forward + adjoint, against the oracle and the
single-grid entry points: small mixed sets, and the 96-instance benchmark set
(stencils j2d5pt / j2d9pt / star2d2r, sides 256-4096, squares and 1:4
rectangles, T 64-1024)."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as O
import pyan5d
from pyan5d import presets

pytestmark = pytest.mark.gpu

# the adjoint blocks the batched kernels hold (j2d5pt and j2d9pt-gol 4, the
# radius-2 stars 3)
BATCH_ADJ_BT = {"j2d5pt": 4, "j2d9pt-gol": 4, "j2d9pt": 3, "star2d2r": 3}


def _same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


def test_batch_matches_single_grid(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    specs = [("j2d5pt", [300, 500], 37), ("j2d9pt", [64, 1100], 23), ("star2d2r", [700, 90], 50),
             ("j2d9pt-gol", [256, 256], 9), ("j2d5pt", [5, 7], 13), ("j2d9pt", [129, 40], 64),
             ("star2d2r", [1000, 1000], 11), ("j2d5pt", [2048, 512], 30), ("j2d5pt", [40, 40], 0)]
    items, hosts = [], []
    for name, n, T in specs:
        d = O.info(name)
        u = rng.random([x + 2 * d["rad"] for x in n]).astype(np.float32)
        A = np.stack([u, u])
        hosts.append(A)
        coefs, div = presets.coefficients(name, "baseline")
        items.append((name, n, T, torch.from_numpy(A.copy()).cuda(), coefs, div))
    print("case batch create", flush=True)
    b = pyan5d.Batch(items)
    assert b.levels >= 1 and b.units > 0
    print("case batch tune", flush=True)
    sl, ms = b.tune(reps=1)
    assert sl in (32, 64, 128, 256) and ms > 0
    torch.cuda.synchronize()
    for (_, _, _, t, _, _), A in zip(items, hosts):  # the tuner restores the grids
        assert _same_bits(t.cpu().numpy(), A)
    for rep in range(2):  # second run replays the captured graph
        for (_, _, _, t, _, _), A in zip(items, hosts):
            t.copy_(torch.from_numpy(A))
        torch.cuda.synchronize()
        print(f"case batch run {rep}", flush=True)
        b.run()
        torch.cuda.synchronize()
        for (name, n, T, t, coefs, div), A in zip(items, hosts):
            ref = A.copy()
            if T:
                pyan5d.forward_host(name, ref, n, T, coefs=coefs, divisor=div, bt=4)
            got = t.cpu().numpy()
            assert _same_bits(got[T % 2], ref[T % 2]), (name, n, T)
    b.close()


def test_batch_rejects_unsupported(torch_cuda):
    t = torch_cuda.zeros(2, 10, 10, device="cuda")
    with pytest.raises(pyan5d.An5dError):
        pyan5d.Batch([("box2d4r", [2, 2], 3, t, None, -1.0)])


def test_batch_workload_shape():
    w = presets.batch_workload()
    assert len(w) == 96
    assert {x[0] for x in w} == {"j2d5pt", "j2d9pt", "star2d2r"}
    assert all(64 <= x[2] <= 1024 for x in w)
    assert all(min(x[1]) >= 64 and max(x[1]) <= 4096 for x in w)


def _adjoint_items(torch, specs, seed):
    rng = np.random.default_rng(seed)
    items, hosts = [], []
    for name, n, T in specs:
        d = O.info(name)
        shape = [x + 2 * d["rad"] for x in n]
        u0 = rng.random(shape).astype(np.float32)
        tgt = rng.random(shape).astype(np.float32)
        coefs, div = presets.coefficients(name, "baseline")
        tu, tt = torch.from_numpy(u0).cuda(), torch.from_numpy(tgt).cuda()
        tg = torch.full_like(tu, float("nan"))
        gc = torch.full((d["ntaps"],), float("nan"), dtype=torch.float64, device="cuda")
        loss = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        items.append((name, n, T, tu, tt, tg, gc, loss, coefs, div))
        hosts.append((u0, tgt))
    return items, hosts


def _check_vs_oracle(item, host, tol=2e-4):
    name, n, T, _, _, tg, gc, loss, coefs, div = item
    u0, tgt = host
    ol, og, ogc = O.adjoint(name, u0.astype(np.float64), tgt.astype(np.float64), n, T, coefs=coefs, divisor=div)
    g, c, lv = tg.cpu().numpy(), gc.cpu().numpy(), float(loss.item())
    assert abs(lv - ol) <= tol * max(1.0, abs(ol)), (name, n, T, lv, ol)
    assert np.max(np.abs(g - og)) <= tol * np.max(np.abs(og)), (name, n, T)
    assert np.max(np.abs(c - ogc)) <= tol * np.max(np.abs(ogc)), (name, n, T, c, ogc)


def _check_vs_single_grid(torch, item):
    """Same block depth on the single-grid adjoint: the gradient grid bit for
    bit (every cell is the same expression of the same operands); the sums to
    rounding, since the FP32 per-lane partials group other cells (other
    segment boundaries) and the slots are reduced in another order."""
    name, n, T, tu, tt, tg, gc, loss, coefs, div = item
    d = O.info(name)
    bt = min(BATCH_ADJ_BT[name], T)
    ws = pyan5d.AdjointWorkspace(name, n, T, np.float32, bt=bt)
    g2 = torch.zeros_like(tu)
    gc2 = torch.zeros(d["ntaps"], dtype=torch.float64, device="cuda")
    l2 = torch.zeros(1, dtype=torch.float64, device="cuda")
    pyan5d.adjoint_device(name, tu, tt, g2, gc2, l2, n, T, ws, coefs=coefs, divisor=div, bt=bt)
    torch.cuda.synchronize()
    assert _same_bits(tg.cpu().numpy(), g2.cpu().numpy()), (name, n, T)
    c, c2 = gc.cpu().numpy(), gc2.cpu().numpy()
    assert np.max(np.abs(c - c2)) <= 1e-5 * np.max(np.abs(c2)), (name, n, T, c, c2)
    assert abs(loss.item() - l2.item()) <= 1e-12 * abs(l2.item())
    del ws


@pytest.mark.skip(reason="batched forward + adjoint is switched off: its first GPU run stalled twice "
                         "(the second time with single-stream capture) and the cause is not found; "
                         "an5d_batch_adjoint_create returns AN5D_E_UNSUPPORTED")
@pytest.mark.parametrize("two_level", [False, True])
def test_batch_adjoint_small_vs_oracle(torch_cuda, two_level, monkeypatch):
    torch = torch_cuda
    if two_level:  # a budget no workspace meets: every instance with >= 2 blocks recomputes
        monkeypatch.setenv("AN5D_BATCH_MEM", "1")
    specs = [("j2d5pt", [300, 500], 37), ("j2d9pt", [64, 1100], 23), ("star2d2r", [70, 90], 50),
             ("j2d9pt-gol", [256, 256], 9), ("j2d5pt", [5, 7], 13), ("j2d9pt", [129, 40], 2),
             ("j2d5pt", [40, 300], 1), ("star2d2r", [33, 257], 31)]
    items, hosts = _adjoint_items(torch, specs, seed=8 + two_level)
    print(f"case batch adjoint create two_level={two_level}", flush=True)
    b = pyan5d.BatchAdjoint(items)
    for rep in range(2):
        print(f"case batch adjoint run {rep}", flush=True)
        b.run()
        torch.cuda.synchronize()
        if rep == 0:
            first = [(it[5].cpu().numpy(), it[6].cpu().numpy(), float(it[7].item())) for it in items]
            for it in items:  # the replay must rewrite every output
                it[5].fill_(float("nan"))
                it[6].fill_(float("nan"))
                it[7].fill_(float("nan"))
        else:
            for it, (g, c, lv) in zip(items, first):
                assert _same_bits(it[5].cpu().numpy(), g)
                assert np.array_equal(it[6].cpu().numpy(), c) and float(it[7].item()) == lv
    for it, h in zip(items, hosts):
        print(f"case batch adjoint check {it[0]} {it[1]} T={it[2]}", flush=True)
        _check_vs_oracle(it, h)
        _check_vs_single_grid(torch, it)
        r = O.info(it[0])["rad"]
        g = it[5].cpu().numpy()
        assert np.all(g[:r, :] == 0) and np.all(g[-r:, :] == 0) and np.all(g[:, :r] == 0) and np.all(g[:, -r:] == 0)
    b.close()


def test_batch_workload_forward_vs_oracle_and_single_grid(torch_cuda):
    """All 96 benchmark instances: an 8-step prefix (two launches at the
    deepest batched block) bit for bit against the oracle, then the full
    horizons bit for bit against single-grid sweeps."""
    torch = torch_cuda
    work = presets.batch_workload()
    g = np.random.default_rng(11)
    hosts = []
    for name, n, T in work:
        d = O.info(name)
        u = g.random([x + 2 * d["rad"] for x in n]).astype(np.float32)
        hosts.append(np.stack([u, u]))
    pre = 8
    items = [(name, n, pre, torch.from_numpy(A.copy()).cuda(), *presets.coefficients(name, "baseline"))
             for (name, n, T), A in zip(work, hosts)]
    print("case batch96 prefix", flush=True)
    b = pyan5d.Batch(items, bt=4)
    b.run()
    torch.cuda.synchronize()

    def oracle_prefix(i):
        name, n, _ = work[i]
        ref = hosts[i].copy()
        coefs, div = presets.coefficients(name, "baseline")
        O.run(name, ref, n, pre, coefs=coefs, divisor=div)
        return ref[pre % 2]

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        refs = list(ex.map(oracle_prefix, range(len(work))))
    for i, (it, ref) in enumerate(zip(items, refs)):
        assert _same_bits(it[3][pre % 2].cpu().numpy(), ref), (i, work[i])
    b.close()
    del items
    print("case batch96 full", flush=True)
    items = [(name, n, T, torch.from_numpy(A.copy()).cuda(), *presets.coefficients(name, "baseline"))
             for (name, n, T), A in zip(work, hosts)]
    b = pyan5d.Batch(items)
    b.tune(reps=1)
    b.run()
    torch.cuda.synchronize()
    for i, ((name, n, T, t, coefs, div), A) in enumerate(zip(items, hosts)):
        ref = torch.from_numpy(A).cuda()
        pyan5d.forward_device(name, ref, n, T, coefs=coefs, divisor=div)
        torch.cuda.synchronize()
        assert torch.equal(ref[T % 2].view(torch.int32), t[T % 2].view(torch.int32)), (i, work[i])
    b.close()


@pytest.mark.skip(reason="batched forward + adjoint is switched off: its first GPU run stalled twice "
                         "(the second time with single-stream capture) and the cause is not found; "
                         "an5d_batch_adjoint_create returns AN5D_E_UNSUPPORTED")
def test_batch_workload_adjoint(torch_cuda):
    """All 96 benchmark instances differentiated in one batch: the instances
    with sides <= 512 against the FP64 oracle adjoint, every instance's
    gradient grid bit for bit against the single-grid adjoint."""
    torch = torch_cuda
    work = presets.batch_workload()
    items, hosts = _adjoint_items(torch, work, seed=12)
    print("case batch96 adjoint create", flush=True)
    b = pyan5d.BatchAdjoint(items)
    b.tune(reps=1)
    print(f"case batch96 adjoint run ({b.levels} levels)", flush=True)
    b.run()
    torch.cuda.synchronize()
    small = [i for i, (_, n, _) in enumerate(work) if max(n) <= 512]
    assert len(small) >= 20

    def oracle(i):
        name, n, T = work[i]
        coefs, div = presets.coefficients(name, "baseline")
        u0, tgt = hosts[i]
        return O.adjoint(name, u0.astype(np.float64), tgt.astype(np.float64), n, T, coefs=coefs, divisor=div)

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        refs = dict(zip(small, ex.map(oracle, small)))
    for i in small:
        ol, og, ogc = refs[i]
        _, _, _, _, _, tg, gc, loss, _, _ = items[i]
        g, c, lv = tg.cpu().numpy(), gc.cpu().numpy(), float(loss.item())
        assert abs(lv - ol) <= 2e-4 * max(1.0, abs(ol)), (i, work[i])
        assert np.max(np.abs(g - og)) <= 2e-4 * np.max(np.abs(og)), (i, work[i])
        assert np.max(np.abs(c - ogc)) <= 2e-4 * np.max(np.abs(ogc)), (i, work[i])
    b.close()
    print("case batch96 adjoint vs single grid", flush=True)
    for it in items:
        _check_vs_single_grid(torch, it)
        torch.cuda.empty_cache()


def test_batch_adjoint_entry_is_switched_off(torch_cuda):
    torch = torch_cuda
    items, _ = _adjoint_items(torch, [("j2d5pt", [16, 16], 5)], seed=1)
    with pytest.raises(pyan5d.An5dError) as e:
        pyan5d.BatchAdjoint(items)
    assert e.value.code == -2 and "switched off" in str(e.value)
