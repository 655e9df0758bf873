"""Pins the CPU oracle: golden fixtures made from the reference's own
programs, the reference programs themselves (when oracle/_ref is built), and
closed forms that a wrong tap, sign, offset or divisor would violate."""
import hashlib
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DTYPES = [(np.float32, "float"), (np.float64, "double")]


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", O.NAMES)
@pytest.mark.parametrize("dt,tname", DTYPES)
def test_oracle_matches_golden(name, dt, tname):
    g = np.load(os.path.join(GOLD, f"stencil_{name}_{tname}.npz"))
    d = O.info(name)
    n, T = int(g["n"]), int(g["T"])
    A = O.init_grid(name, n, dt)
    O.run(name, A, [n] * d["dim"], T)
    assert np.array_equal(A.view(np.uint8), g["out"].view(np.uint8))
    for (hn, hT), h in zip(g["hash_cases"], g["hashes"]):
        B = O.init_grid(name, int(hn), dt)
        O.run(name, B, [int(hn)] * d["dim"], int(hT))
        assert _sha(B) == str(h), (hn, hT)


@pytest.mark.parametrize("name", O.NAMES)
def test_oracle_matches_reference_build(name):
    if not O.ref_available(name):
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    d = O.info(name)
    for dt in (np.float32, np.float64):
        n = 13 if d["dim"] == 2 else 5
        A = O.ref_init_grid(name, n, dt)
        B = O.init_grid(name, n, dt)
        assert np.array_equal(A.view(np.uint8), B.view(np.uint8))
        O.ref_run(name, A, n, 6, True)
        O.run(name, B, [n] * d["dim"], 6)
        assert np.array_equal(A.view(np.uint8), B.view(np.uint8))


def _impulse_grid(name, n, dt):
    d = O.info(name)
    shape = (2,) + tuple([n + 2 * d["rad"]] * d["dim"])
    A = np.zeros(shape, dtype=dt)
    c = tuple([(n + 2 * d["rad"]) // 2] * d["dim"])
    A[(0,) + c] = 1.0
    A[(1,) + c] = 1.0
    return A, c


@pytest.mark.parametrize("name", [n for n in O.NAMES if O.info(n)["linear"]])
def test_impulse_response_is_transposed_tap_table(name):
    """One step on a unit impulse: out(x) = sum_q c_q [x + d_q == centre] / div.
    Computed here by scattering the source-order taps (float arithmetic of
    one nonzero term is exact up to the divisor's rounding)."""
    d = O.info(name)
    n = 4 * d["rad"] + 3
    A, c = _impulse_grid(name, n, np.float64)
    O.run(name, A, [n] * d["dim"], 1)
    offs, coefs = O.taps(name)
    div = float(d["divisor"]) if d["divisor"] else 1.0
    want = np.zeros_like(A[1])
    for (sign, a, b, k), cf in zip(offs, coefs):
        off = (a, k) if d["dim"] == 2 else (a, b, k)
        x = tuple(ci - oi for ci, oi in zip(c, off))
        want[x] += sign * float(cf)
    want /= div
    # the impulse cell itself stays in the interior; compare the interior only
    r = d["rad"]
    sl = tuple([slice(r, -r)] * d["dim"])
    np.testing.assert_allclose(A[1][sl], want[sl], rtol=1e-15, atol=0)


@pytest.mark.parametrize("name", ["star2d1r", "box2d2r", "box2d3r", "star3d2r", "box3d2r"])
def test_constant_field_fixed_point(name):
    """Sum of coefficients is 1 (up to literal rounding): a constant grid stays constant."""
    d = O.info(name)
    n = 6
    shape = (2,) + tuple([n + 2 * d["rad"]] * d["dim"])
    A = np.full(shape, 3.0, dtype=np.float64)
    O.run(name, A, [n] * d["dim"], 3)
    offs, coefs = O.taps(name)
    s = float(np.sum(offs[:, 0] * coefs.astype(np.float64)))
    assert abs(s - 1.0) < 1e-4
    np.testing.assert_allclose(A, 3.0, rtol=1e-3)


@pytest.mark.parametrize("name", ["j2d5pt", "j2d9pt", "j3d27pt"])
def test_decay_factor_of_divided_stencils(name):
    """Constant field u inside a matching constant halo: one step gives
    u * (sum c)/div exactly as the divided sum (checked in FP64)."""
    d = O.info(name)
    n = 5
    shape = (2,) + tuple([n + 2 * d["rad"]] * d["dim"])
    A = np.full(shape, 2.0, dtype=np.float64)
    O.run(name, A, [n] * d["dim"], 1)
    offs, coefs = O.taps(name)
    want = 2.0 * float(np.sum(offs[:, 0] * coefs.astype(np.float64))) / d["divisor"]
    r = d["rad"]
    sl = tuple([slice(r, -r)] * d["dim"])
    np.testing.assert_allclose(A[1][sl], want, rtol=1e-14)


def test_gradient2d_closed_form():
    """Flat field: out = C + 1/sqrt(1e-4) = C + 100 (C semantics, double sqrt)."""
    n = 5
    A = np.full((2, n + 2, n + 2), 1.5, dtype=np.float32)
    O.run("gradient2d", A, [n, n], 1)
    want = np.float32(1.5 + 1.0 / np.sqrt(np.float64(np.float32(1e-4))))
    assert np.all(A[1, 1:-1, 1:-1] == want)
    assert np.all(A[1, 0, :] == 1.5)


def test_time_loop_and_halo_semantics():
    """Brute force: T steps with ping-pong buffers leave u_T in buffer T%2,
    u_{T-1} in the other, and never touch the halo."""
    rng = np.random.default_rng(5)
    n, r = 7, 1
    u = rng.random((n + 2, n + 2))
    A = np.stack([u, u]).copy()
    O.run("star2d1r", A, [n, n], 4)
    offs, coefs = O.taps("star2d1r")
    cur = u.copy()
    hist = [cur.copy()]
    for _ in range(4):
        nxt = cur.copy()
        for i in range(r, n + r):
            for j in range(r, n + r):
                acc = None
                for (sign, a, b, k), cf in zip(offs, coefs):
                    v = float(np.float32(cf)) * cur[i + a, j + k]
                    acc = v if acc is None else (acc + v if sign > 0 else acc - v)
                nxt[i, j] = acc
        cur = nxt
        hist.append(cur.copy())
    assert np.array_equal(A[0], hist[4])
    assert np.array_equal(A[1], hist[3])


def test_init_grid_restatement():
    """common.h:44-56 restated: seed i+N per cell, 1000*r/RAND_MAX."""
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
    A = O.init_grid("j2d5pt", 5, np.float32)
    for i in (0, 3, 48):
        s = ctypes.c_uint(i + 5)
        r = libc.rand_r(ctypes.byref(s))
        assert A.reshape(2, -1)[0, i] == np.float32(np.float32(1000) * np.float32(r)) / np.float32(2147483647)
    assert np.array_equal(A[0], A[1])


def test_check_result_restatement():
    A = np.zeros((2, 4, 4), np.float32)
    B = np.zeros((2, 4, 4), np.float32)
    B[1, 2, 2] = 3.0
    B[1, 0, 0] = -4.0
    mx, rms = O.check(A, B, 1)
    assert mx == 4.0
    assert abs(rms - np.sqrt(25.0 / 16)) < 1e-12
    mx0, _ = O.check(A, B, 2)
    assert mx0 == 0.0


def _loss(name, u0, target, n, T, coefs=None):
    A = np.stack([u0, u0]).astype(np.float64)
    O.run(name, A, n, T, coefs=coefs)
    d = O.info(name)
    r = d["rad"]
    sl = tuple([slice(r, -r)] * d["dim"])
    return 0.5 * float(np.sum((A[T % 2][sl] - target[sl]) ** 2))


@pytest.mark.parametrize("name", ["j2d5pt", "j2d9pt", "j2d9pt-gol", "box2d2r", "gradient2d", "star3d1r"])
def test_oracle_adjoint_matches_finite_differences(name):
    """Central differences of the oracle forward pin the oracle adjoint,
    for the initial state and for every coefficient."""
    d = O.info(name)
    n = [9, 11] if d["dim"] == 2 else [5, 6, 7]
    T = 4
    rng = np.random.default_rng(3)
    shape = tuple(x + 2 * d["rad"] for x in n)
    u0 = rng.random(shape)
    target = rng.random(shape)
    coefs = (np.array([1e-2, 0.5]) if not d["linear"] else
             np.asarray(O.taps(name)[1], dtype=np.float64))
    loss, gu, gc = O.adjoint(name, u0, target, n, T, coefs=coefs)
    assert abs(loss - _loss(name, u0, target, n, T, coefs)) < 1e-12 * max(1.0, loss)
    eps = 1e-6
    r = d["rad"]
    for _ in range(6):
        x = tuple(int(rng.integers(r, s - r)) for s in shape)
        up, um = u0.copy(), u0.copy()
        up[x] += eps
        um[x] -= eps
        fd = (_loss(name, up, target, n, T, coefs) - _loss(name, um, target, n, T, coefs)) / (2 * eps)
        assert abs(fd - gu[x]) <= 1e-6 * max(1.0, abs(fd)), (x, fd, gu[x])
    for q in range(len(coefs)):
        cp, cm = coefs.copy(), coefs.copy()
        h = eps * max(1.0, abs(coefs[q]))
        cp[q] += h
        cm[q] -= h
        fd = (_loss(name, u0, target, n, T, cp) - _loss(name, u0, target, n, T, cm)) / (2 * h)
        assert abs(fd - gc[q]) <= 1e-5 * max(1.0, abs(fd)), (q, fd, gc[q])
