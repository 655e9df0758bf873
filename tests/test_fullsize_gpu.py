"""Parity at the benchmark sizes: star3d1r 512^3 (FP32, FP64), j2d5pt FP32
16384^2, star3d4r 512^3 (FP32, FP64; the LDS-staged one-step kernel),
box2d4r 16384^2 (FP32, FP64; the row partial sums) and j3d27pt 512^3 (FP32,
FP64; the plane partial sums), each with the configuration the benchmark
runs (explicit flags, else the autotuner's table, else the defaults).

- The first launches are compared with the oracle bit for bit (the oracle
  finishes a few steps of these grids in seconds).
- The full 1000-step horizon is checked through properties that need no
  oracle run: the tuned sweep equals a sweep with another kernel shape and
  depth bit for bit (every cell's arithmetic is fixed by the source order, so
  the tiling must not show), and 1000 steps in one call equal 400 + 600.
- The benchmark's own configuration (its coefficient presets and uniform
  [0, 1) input) is compared with the oracle bit for bit over its first
  launches.
- FMA mode stays within the stated tolerance (1e-5 relative in FP32, 1e-12 in
  FP64, floor: the smallest normal) over the first launches against the
  oracle and over the full 1000 steps against the exact-mode run, which
  equals the C loop bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as O
import pyan5d
from pyan5d import presets

pytestmark = pytest.mark.gpu

FULL = [("star3d1r", np.float32, [512, 512, 512]),
        ("star3d1r", np.float64, [512, 512, 512]),
        ("j2d5pt", np.float32, [16384, 16384]),
        ("star3d4r", np.float32, [512, 512, 512]),
        ("star3d4r", np.float64, [512, 512, 512]),
        ("box2d4r", np.float32, [16384, 16384]),
        ("box2d4r", np.float64, [16384, 16384]),
        ("j3d27pt", np.float32, [512, 512, 512]),
        ("j3d27pt", np.float64, [512, 512, 512])]


def same(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name,dt,n", FULL)
def test_first_launches_bitwise_vs_oracle(name, dt, n):
    rc = pyan5d.resolve_config(name, n, dt)
    bt = rc["bt"]
    # two launches where that is cheap for the oracle, one for the 2D grid
    T = bt if O.info(name)["dim"] == 2 else 2 * bt
    print(f"case full-size oracle {name} {np.dtype(dt).name} n={n} T={T} cfg={rc}", flush=True)
    A = O.init_grid(name, n[0], dt)
    B = A.copy()
    pyan5d.forward_host(name, A, n, T, bt=bt, variant=rc["variant"], sl=rc["sl"])
    O.run(name, B, n, T)
    assert same(A[T % 2], B[T % 2])


@pytest.mark.parametrize("name,dt,n", FULL)
def test_full_horizon_shape_and_split_invariance(name, dt, n, torch_cuda):
    torch = torch_cuda
    T = 1000
    rc = pyan5d.resolve_config(name, n, dt)
    dcode = 0 if dt == np.float32 else 1
    vs = pyan5d.variants(name, dcode)
    # another kernel: the first variant (the second if the tuned run uses
    # the first) at bt 2 where it goes that deep
    alt_v = 0 if rc["variant"] != 0 else 1
    alt_bt = min(2, vs[alt_v]["btmax"])
    assert (alt_v, alt_bt) != (rc["variant"], rc["bt"])
    host = pyan5d.init_grid(name, n, dt, mode=1)
    base = torch.from_numpy(host).cuda()
    del host
    out = {}
    print(f"case full-horizon {name} {np.dtype(dt).name} tuned {rc}", flush=True)
    t = base.clone()
    pyan5d.forward_device(name, t, n, T, bt=rc["bt"], variant=rc["variant"], sl=rc["sl"])
    torch.cuda.synchronize()
    out["tuned"] = t[T % 2].clone()
    print(f"case full-horizon {name} {np.dtype(dt).name} variant {alt_v} bt {alt_bt}", flush=True)
    t.copy_(base)
    pyan5d.forward_device(name, t, n, T, bt=alt_bt, variant=alt_v)
    torch.cuda.synchronize()
    assert torch.equal(t[T % 2].view(torch.uint8), out["tuned"].view(torch.uint8))
    print(f"case full-horizon {name} {np.dtype(dt).name} 400 + 600", flush=True)
    t.copy_(base)
    pyan5d.forward_device(name, t, n, 400, bt=rc["bt"], variant=rc["variant"], sl=rc["sl"])
    pyan5d.forward_device(name, t, n, 600, bt=rc["bt"], variant=rc["variant"], sl=rc["sl"])
    torch.cuda.synchronize()
    assert torch.equal(t[0].view(torch.uint8), out["tuned"].view(torch.uint8))
    # the run did move the field (no stage or launch was skipped)
    assert not torch.equal(out["tuned"], base[0])
    del t, base, out
    torch.cuda.empty_cache()


@pytest.mark.parametrize("name,dt,n", [FULL[0], FULL[2]])
def test_fast_mode_full_size_within_tolerance(name, dt, n):
    rc = pyan5d.resolve_config(name, n, dt, mode=pyan5d.FAST)
    T = rc["bt"] + 1  # one full launch and a remainder launch
    print(f"case full-size fast {name} n={n} T={T} cfg={rc}", flush=True)
    A = O.init_grid(name, n[0], dt)
    B = A.copy()
    pyan5d.forward_host(name, A, n, T, bt=rc["bt"], variant=rc["variant"], sl=rc["sl"], mode=pyan5d.FAST)
    O.run(name, B, n, T)
    got, ref = A[T % 2].astype(np.float64), B[T % 2].astype(np.float64)
    rel = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), np.finfo(dt).tiny))
    assert rel <= 1e-5, rel


# the benchmark's configurations: coefficient presets, uniform [0, 1) input
BENCH = [("star3d1r", np.float32, [512, 512, 512]),
         ("star3d1r", np.float64, [512, 512, 512]),
         ("j2d5pt", np.float32, [16384, 16384])]
# FMA mode on star3d1r FP32 drifts to 1.85e-5 relative over 1000 steps (MI355X
# measurement): beyond the 1e-5 tolerance, so that configuration is held to it
# as an expected failure and its FMA-mode numbers carry no parity claim
FAST_FULL = [pytest.param(*BENCH[0], marks=pytest.mark.xfail(
                 strict=True, reason="FMA drift 1.85e-5 > 1e-5 relative after 1000 steps")),
             BENCH[1], BENCH[2]]
TOL = {np.float32: 1e-5, np.float64: 1e-12}


def _rel_err(got, ref, dt):
    got, ref = got.astype(np.float64), ref.astype(np.float64)
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), np.finfo(dt).tiny)))


@pytest.mark.parametrize("name,dt,n", BENCH)
def test_benchmark_preset_first_launches_vs_oracle(name, dt, n):
    rc = pyan5d.resolve_config(name, n, dt)
    coefs, div = presets.coefficients(name, "baseline")
    T = 2 * rc["bt"]
    print(f"case preset oracle {name} {np.dtype(dt).name} n={n} T={T} cfg={rc}", flush=True)
    A = pyan5d.init_grid(name, n, dt, mode=1)
    B = A.copy()
    pyan5d.forward_host(name, A, n, T, coefs=coefs, divisor=div, bt=rc["bt"], variant=rc["variant"],
                        sl=rc["sl"])
    O.run(name, B, n, T, coefs=coefs, divisor=div)
    assert same(A[T % 2], B[T % 2])
    assert not same(A[T % 2], A[(T + 1) % 2])


@pytest.mark.parametrize("name,dt,n", FAST_FULL)
def test_fast_mode_full_horizon_within_tolerance(name, dt, n, torch_cuda):
    """1000 steps in FMA mode against the same 1000 steps in exact mode."""
    torch = torch_cuda
    T = 1000
    coefs, div = presets.coefficients(name, "baseline")
    rf = pyan5d.resolve_config(name, n, dt, mode=pyan5d.FAST)
    rx = pyan5d.resolve_config(name, n, dt)
    host = pyan5d.init_grid(name, n, dt, mode=1)
    base = torch.from_numpy(host).cuda()
    del host
    print(f"case fast full-horizon {name} {np.dtype(dt).name} exact {rx}", flush=True)
    t = base.clone()
    pyan5d.forward_device(name, t, n, T, coefs=coefs, divisor=div, bt=rx["bt"], variant=rx["variant"],
                          sl=rx["sl"])
    torch.cuda.synchronize()
    exact = t[T % 2].cpu().numpy()
    print(f"case fast full-horizon {name} {np.dtype(dt).name} fast {rf}", flush=True)
    t.copy_(base)
    pyan5d.forward_device(name, t, n, T, coefs=coefs, divisor=div, bt=rf["bt"], variant=rf["variant"],
                          sl=rf["sl"], mode=pyan5d.FAST)
    torch.cuda.synchronize()
    fast = t[T % 2].cpu().numpy()
    del t, base
    torch.cuda.empty_cache()
    assert np.all(np.isfinite(fast)) and np.all(np.isfinite(exact))
    rel = _rel_err(fast, exact, dt)
    print(f"case fast full-horizon {name} {np.dtype(dt).name} max rel {rel:.3e}", flush=True)
    assert rel <= TOL[dt], rel
