"""Batched forward: every instance bitwise equal to its own single-grid sweep."""
import numpy as np
import pytest

import oracle_lib as O
import pyan5d
from pyan5d import presets

pytestmark = pytest.mark.gpu


def test_batch_matches_single_grid(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    specs = [("j2d5pt", [300, 500], 37), ("j2d9pt", [64, 1100], 23), ("star2d2r", [700, 90], 50),
             ("j2d9pt-gol", [256, 256], 9), ("j2d5pt", [5, 7], 13), ("j2d9pt", [129, 40], 64),
             ("star2d2r", [1000, 1000], 11), ("j2d5pt", [2048, 512], 30)]
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
    assert b.levels == max(len(pyan5d.schedule(T, 4)) for _, _, T in specs)
    for rep in range(2):  # second run replays the captured graph
        for (_, _, _, t, _, _), A in zip(items, hosts):
            t.copy_(torch.from_numpy(A))
        torch.cuda.synchronize()
        print(f"case batch run {rep}", flush=True)
        b.run()
        torch.cuda.synchronize()
        for (name, n, T, t, coefs, div), A in zip(items, hosts):
            ref = A.copy()
            pyan5d.forward_host(name, ref, n, T, coefs=coefs, divisor=div, bt=4)
            got = t.cpu().numpy()
            assert np.array_equal(got[T % 2].view(np.uint32), ref[T % 2].view(np.uint32)), (name, n, T)
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
