"""Coefficient presets of the benchmark workloads.

"reference": the literals of each <stencil>.c (coefs=None, source divisor).
"baseline":  the benchmark configurations' definitions:
  j2d5pt   (5.1 N + 12.1 W + 15.0 C + 12.2 E + 12.3 S) / 118
  star3d1r centre 0.4, six neighbours 0.1
  box2d4r / star3d4r / j3d27pt / j2d9pt / j2d9pt-gol / star2d2r: seeded
           uniform (0, 1] coefficients normalised to sum 1, no divisor
           (seed = 1000 + stencil id, numpy default_rng)
  gradient2d eps 1e-4, scale 0.01
Each returns (coefs or None, divisor) for an5d_problem.
"""
import numpy as np

from . import info, stencil_id, taps


def seeded_normalised(name):
    d = info(name)
    rng = np.random.default_rng(1000 + stencil_id(name))
    c = 1.0 - rng.random(d["ntaps"])  # (0, 1]
    return c / c.sum()


def coefficients(name, preset="baseline"):
    if preset == "reference":
        return None, -1.0
    if preset != "baseline":
        raise ValueError(preset)
    if name == "j2d5pt":
        offs, c = taps(name)
        c = c.copy()
        c[4] = 12.3  # S term (source: 5.2)
        return c, 118.0
    if name == "star3d1r":
        offs, _ = taps(name)
        c = np.where((offs[:, 1:] == 0).all(axis=1), 0.4, 0.1)
        return c, 0.0
    if name == "gradient2d":
        return np.array([1e-4, 0.01]), -1.0
    if info(name)["linear"]:
        return seeded_normalised(name), 0.0
    return None, -1.0


def batch_workload(count=96, seed=2024):
    """The batched configuration: `count` instances, stencil uniform over
    {j2d5pt, j2d9pt, star2d2r}, side log-uniform in [256, 4096], half square
    and half 1:4 rectangles, T uniform in [64, 1024].  Returns a list of
    (stencil, [n0, n1], T)."""
    rng = np.random.default_rng(seed)
    kinds = ["j2d5pt", "j2d9pt", "star2d2r"]
    out = []
    for i in range(count):
        name = kinds[int(rng.integers(0, 3))]
        side = int(round(float(np.exp(rng.uniform(np.log(256), np.log(4096))))))
        if rng.random() < 0.5:
            n = [side, side]
        else:
            n = [max(1, side // 4), side] if rng.random() < 0.5 else [side, max(1, side // 4)]
        T = int(rng.integers(64, 1025))
        out.append((name, n, T))
    return out
