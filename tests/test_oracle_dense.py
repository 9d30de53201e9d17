"""The CPU oracle's dense kernels (oracle/dense.cpp) against scipy LAPACK/BLAS golden
vectors (tests/golden/dense.npz, made by tests/golden/make_golden.py) and against the
reference's own known-answer tests for choose_rank and swap2perm."""
import os

import numpy as np
import pytest

import oracle

L = oracle.lib()


@pytest.fixture(scope="module")
def G(golden_dir):
    return np.load(os.path.join(golden_dir, "dense.npz"))


def test_choose_rank_kat():
    # tests/tests.cpp:254-262
    errs = np.array([1.0, -0.1, 0.01, -0.001, 1e-4])
    assert L.or_choose_rank(errs, 5, 1e-1) == 2
    assert L.or_choose_rank(errs, 5, 1e-2) == 3
    assert L.or_choose_rank(errs, 5, 1.0) == 0
    assert L.or_choose_rank(errs, 5, 0.0) == 5
    assert L.or_choose_rank(errs, 5, 1e-16) == 5
    # degenerate sizes (src/util.cpp:440-442)
    assert L.or_choose_rank(np.array([3.0]), 1, 0.5) == 1
    assert L.or_choose_rank(np.zeros(1), 0, 0.5) == 0


def test_swap2perm_kat():
    # tests/tests.cpp:349-357
    swap = np.array([3, 3, 2, 5, 4, 5], dtype=np.int32)
    perm = np.zeros(6, dtype=np.int32)
    L.or_swap2perm(6, swap, perm)
    assert perm.tolist() == [3, 0, 2, 5, 4, 1]


@pytest.mark.parametrize("n", [1, 5, 33, 64])
def test_potrf(G, n):
    A = np.asfortranarray(G[f"potrf_{n}_A"].copy())
    assert L.or_potrf(n, A, n) == 0
    Lo = np.tril(A)
    assert np.allclose(Lo, G[f"potrf_{n}_L"], rtol=0, atol=1e-12 * np.abs(G[f"potrf_{n}_L"]).max())
    # closed form: L L^T = A
    assert np.linalg.norm(Lo @ Lo.T - G[f"potrf_{n}_A"]) <= 1e-13 * np.linalg.norm(G[f"potrf_{n}_A"])


def test_potrf_not_spd():
    A = np.asfortranarray(np.array([[1.0, 2.0], [2.0, 1.0]]))
    assert L.or_potrf(2, A, 2) == 2


@pytest.mark.parametrize("m,n", [(7, 5), (40, 64), (1, 9)])
def test_trsm(G, m, n):
    for kind, key, fn in (("llnn", "L", L.or_trsm_llnn), ("rltn", "L", L.or_trsm_rltn), ("runn", "U", L.or_trsm_runn)):
        T = np.asfortranarray(G[f"trsm_{kind}_{m}_{n}_{key}"])
        B = np.asfortranarray(G[f"trsm_{kind}_{m}_{n}_B"].copy())
        fn(m, n, T, T.shape[0], B, m)
        X = G[f"trsm_{kind}_{m}_{n}_X"]
        assert np.abs(B - X).max() <= 1e-13 * max(1.0, np.abs(X).max()), kind


@pytest.mark.parametrize("m,n", [(5, 3), (3, 5), (18, 150), (40, 40), (48, 160), (1, 7), (7, 1)])
def test_geqp3_and_ormqr(G, m, n):
    A0 = G[f"geqp3_{m}_{n}_A"]
    A = np.asfortranarray(A0.copy())
    k = min(m, n)
    jpvt = np.zeros(n, dtype=np.int32)
    tau = np.zeros(max(k, 1))
    L.or_geqp3(m, n, A, m, jpvt, tau)
    # random matrices: no near-ties, so the pivot order must be LAPACK's
    assert np.array_equal(jpvt, G[f"geqp3_{m}_{n}_jpvt"])
    R, Rg = np.triu(A)[:k], np.triu(G[f"geqp3_{m}_{n}_QR"])[:k]
    assert np.abs(R - Rg).max() <= 1e-12 * np.abs(Rg).max()
    assert np.abs(tau[:k] - G[f"geqp3_{m}_{n}_tau"][:k]).max() <= 1e-12
    # closed form: A P = Q R, and Q^T x / Q x against dormqr
    x = G[f"ormqr_{m}_{n}_x"].copy()
    qtx = x.copy()
    L.or_ormqr_vec(1, m, k, A, m, tau, qtx)
    assert np.abs(qtx - G[f"ormqr_{m}_{n}_qtx"]).max() <= 1e-12 * max(1.0, np.abs(x).max())
    qx = x.copy()
    L.or_ormqr_vec(0, m, k, A, m, tau, qx)
    assert np.abs(qx - G[f"ormqr_{m}_{n}_qx"]).max() <= 1e-12 * max(1.0, np.abs(x).max())
    Q = np.eye(m)
    for j in range(m):
        col = Q[:, j].copy()
        L.or_ormqr_vec(0, m, k, A, m, tau, col)
        Q[:, j] = col
    full_R = np.zeros((m, n))
    full_R[:k] = np.triu(A)[:k]
    assert np.linalg.norm(Q @ full_R - A0[:, jpvt]) <= 1e-12 * np.linalg.norm(A0)


@pytest.mark.parametrize("n", [1, 6, 50])
def test_getrf(G, n):
    A = np.asfortranarray(G[f"getrf_{n}_A"].copy())
    piv = np.zeros(n, dtype=np.int32)
    assert L.or_getrf(n, A, n, piv) == 0
    assert np.array_equal(piv, G[f"getrf_{n}_piv"])
    assert np.abs(A - G[f"getrf_{n}_LU"]).max() <= 1e-12 * max(1.0, np.abs(G[f"getrf_{n}_LU"]).max())


def test_trsv_and_gemv_closed_forms():
    rng = np.random.default_rng(3)
    n = 17
    Lm = np.asfortranarray(np.tril(rng.standard_normal((n, n))) + 5 * np.eye(n))
    b = rng.standard_normal(n)
    for kind, M in ((0, Lm), (1, Lm.T)):
        x = b.copy()
        L.or_trsv(kind, n, Lm, n, x)
        assert np.allclose(M @ x, b, rtol=0, atol=1e-12)


def test_geqp3_exact_ties_kat():
    # Exactly tied column norms (orthogonal columns, all arithmetic exact): LAPACK's dgeqp3
    # takes the first maximum by current position, after its column swaps. Answers from
    # scipy.linalg.lapack.dgeqp3 (OpenBLAS), 0-based.
    A = np.asfortranarray(np.diag([1.0, 2.0, 2.0, 1.0]))
    jpvt, tau = np.zeros(4, dtype=np.int32), np.zeros(4)
    L.or_geqp3(4, 4, A, 4, jpvt, tau)
    assert jpvt.tolist() == [1, 2, 0, 3]
    B = np.vstack([np.diag([1.0, 2.0, 2.0, 1.0]), np.zeros((2, 4))])
    B[4, 0], B[5, 3] = 2.0, 3.0
    B = np.asfortranarray(B)
    L.or_geqp3(6, 4, B, 6, jpvt, tau)
    assert jpvt.tolist() == [3, 0, 2, 1]
    assert np.allclose(np.abs(np.diag(B)), [np.sqrt(10), np.sqrt(5), 2.0, 2.0], rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", [0, 1, 15, 16, 17, 33, 1000])
def test_reduction16_sums(n):
    # The RRQR's fixed 16-chain summation order: on integer data every partial sum is
    # exact, so any dropped, repeated or misplaced term changes the result.
    rng = np.random.default_rng(n)
    a = rng.integers(-1000, 1000, n).astype(np.float64)
    b = rng.integers(-1000, 1000, n).astype(np.float64)
    assert L.or_sumsq16(n, a) == float(np.sum(a.astype(np.int64) ** 2))
    assert L.or_dot16(n, a, b) == float(np.sum(a.astype(np.int64) * b.astype(np.int64)))
    # and on random data it is a sum: relative error within n * eps
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    s, d = L.or_sumsq16(n, x), L.or_dot16(n, x, y)
    exact_s = float(np.sum(np.float128(x) ** 2)) if hasattr(np, "float128") else float(np.sum(x * x))
    assert abs(s - exact_s) <= (n + 1) * 2.3e-16 * exact_s + 1e-300
    assert abs(d - float(np.dot(x, y))) <= (n + 1) * 2.3e-16 * float(np.dot(np.abs(x), np.abs(y))) + 1e-300


def test_gemm_syrk_schur_closed_forms():
    rng = np.random.default_rng(7)
    m, n, k = 13, 9, 21
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((k, n)))
    C0 = rng.standard_normal((m, n))
    C = np.asfortranarray(C0.copy())
    L.or_gemm(0, 0, m, n, k, -1.0, A, m, B, k, 1.0, C, m)
    assert np.allclose(C, C0 - A @ B, rtol=0, atol=1e-12)
    At, Bt = np.asfortranarray(A.T), np.asfortranarray(B.T)
    C = np.asfortranarray(C0.copy())
    L.or_gemm(1, 1, m, n, k, 2.0, At, k, Bt, n, 0.5, C, m)
    assert np.allclose(C, 0.5 * C0 + 2.0 * A @ B, rtol=0, atol=1e-12)
    # Schur update C -= op(A) op(B), lower triangle only when asked
    for tA, tB in ((0, 1), (1, 0), (0, 0)):
        Aop = np.asfortranarray(A.T) if tA else A
        Bop = np.asfortranarray(B.T) if tB else B
        C = np.asfortranarray(C0.copy())
        L.or_schur(tA, tB, 0, m, n, k, Aop, Aop.shape[0], Bop, Bop.shape[0], C, m)
        assert np.allclose(C, C0 - A @ B, rtol=0, atol=1e-12)
    S0 = rng.standard_normal((m, m))
    S = np.asfortranarray(S0.copy())
    L.or_schur(0, 1, 1, m, m, k, A, m, A, m, S, m)
    ref = S0 - A @ A.T
    low = np.tril(np.ones((m, m), dtype=bool))
    assert np.allclose(S[low], ref[low], rtol=0, atol=1e-12)
    assert np.array_equal(S[~low], S0[~low])
    # syrk lower (src/util.cpp:148-156): C := alpha A A^T + beta C on the lower triangle
    S = np.asfortranarray(S0.copy())
    L.or_syrk(0, m, k, -1.0, A, m, 1.0, S, m)
    assert np.allclose(S[low], ref[low], rtol=0, atol=1e-12)
    assert np.array_equal(S[~low], S0[~low])
