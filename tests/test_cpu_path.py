"""The library's CPU path (the drop-in's scop=false branch) against the oracle."""
import numpy as np
import pytest

import oracle_lib as O
import pyan5d


@pytest.mark.parametrize("name", O.NAMES)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_cpu_path_bitwise(name, dt):
    d = O.info(name)
    n = [11, 14] if d["dim"] == 2 else [5, 6, 7]
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    shape = (2,) + tuple(x + 2 * d["rad"] for x in n)
    u = rng.random(shape[1:]).astype(dt)
    A = np.stack([u, u])
    B = A.copy()
    pyan5d.cpu_forward(name, A, n, 5, threads=3)
    O.run(name, B, n, 5)
    assert np.array_equal(A.view(np.uint8), B.view(np.uint8))


def test_cpu_path_other_coefficients():
    n = [9, 10]
    rng = np.random.default_rng(0)
    c = rng.random(5)
    u = rng.random((11, 12)).astype(np.float32)
    A = np.stack([u, u]); B = A.copy()
    pyan5d.cpu_forward("j2d5pt", A, n, 4, coefs=c, divisor=0)
    O.run("j2d5pt", B, n, 4, coefs=c, divisor=0)
    assert np.array_equal(A, B)
    A = np.stack([u, u]); B = A.copy()
    pyan5d.cpu_forward("gradient2d", A, n, 3, coefs=[1e-4, 0.01])
    O.run("gradient2d", B, n, 3, coefs=[1e-4, 0.01])
    assert np.array_equal(A, B)


def test_kernel_stencil_cpu_branch_matches_reference_program():
    if not O.ref_available("j2d5pt"):
        pytest.skip("oracle/_ref not built")
    A = O.ref_init_grid("j2d5pt", 20)
    B = A.copy()
    pyan5d.kernel_stencil("j2d5pt", A, 20, 7, scop=False)
    O.ref_run("j2d5pt", B, 20, 7, scop=False)
    assert np.array_equal(A[1], B[1])
