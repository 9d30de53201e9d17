"""End-to-end behaviour of the CPU oracle (the checker of the GPU path), pinned by the
reference's own tolerance tests and by brute force:
  ApproxTest.Exact  (tests/tests.cpp:562-609): tol 0 / 1e-14 -> ||Ax-b||/||b|| <= 1e-10
  ApproxTest.Approx (tests/tests.cpp:799-856): tol > 0 -> error <= tol * 2e2
  ApproxTest.Repro  (tests/tests.cpp:917-994): bitwise-identical reruns
  dense numpy solves on tiny grids; PCG to 1e-12 (src/is.cpp:39-121)."""
import numpy as np
import pytest

import oracle
from spand_public_amd.problems import linspace_nd, neglapl


def run(n, d, nl, tol, skip=0, seed=2019):
    A, X = neglapl(n, d), linspace_nd(n, d)
    t, _ = oracle.setup_tree(A, X, nl, tol, skip)
    t.factorize()
    b = np.random.default_rng(seed).uniform(-1, 1, A.n)
    x = t.solve(b)
    return A, t, b, x


@pytest.mark.parametrize("n,d", [(5, 2), (10, 2), (20, 2), (5, 3), (15, 3)])
@pytest.mark.parametrize("tol,skip", [(1e-14, 0), (1e-14, 4), (0.0, 1000)])
def test_exact(n, d, tol, skip):
    N = n ** d
    lmin = 1 if N < 1000 else 8
    for nl in range(lmin, lmin + 5):
        A, t, b, x = run(n, d, nl, tol, skip)
        assert np.linalg.norm(A.matvec(x) - b) / np.linalg.norm(b) <= 1e-10


@pytest.mark.parametrize("n,d", [(5, 2), (10, 2), (20, 2), (128, 2), (5, 3), (15, 3)])
def test_approx(n, d):
    N = n ** d
    lmin = 1 if N < 1000 else 8
    for nl in range(lmin, lmin + 5, 2):
        for tol in (0.0, 1e-10, 1e-6, 1e-2, 10.0):
            for skip in (0, 1, 2):
                A, t, b, x = run(n, d, nl, tol, skip)
                err = np.linalg.norm(A.matvec(x) - b) / np.linalg.norm(b)
                assert err <= (5e-12 if tol == 0 else tol * 2e2), (nl, tol, skip, err)


def test_brute_force_tiny():
    for n, d, nl in ((3, 2, 2), (5, 2, 3), (5, 3, 3)):
        A, t, b, x = run(n, d, nl, 0.0)
        xd = np.linalg.solve(A.to_scipy().toarray(), b)
        assert np.allclose(x, xd, rtol=0, atol=1e-12 * np.abs(xd).max())


def test_repro_bitwise():
    ref = None
    for _ in range(3):
        A, t, b, x = run(20, 2, 5, 1e-5, 1)
        if ref is None:
            ref = x
        assert np.array_equal(ref, x)


def test_pcg_to_1e12():
    A, X = neglapl(20, 3), linspace_nd(20, 3)
    t, _ = oracle.setup_tree(A, X, 7, 1e-2)
    t.factorize()
    b = np.random.default_rng(5).standard_normal(A.n)
    x, it, _ = t.cg(A, b, iters=100, tol=1e-12)
    assert np.linalg.norm(A.matvec(x) - b) / np.linalg.norm(b) <= 1e-11
    assert it < 30


def test_not_spd_raises():
    A, X = neglapl(10, 2), linspace_nd(10, 2)
    A.nzval = -A.nzval
    t, _ = oracle.setup_tree(A, X, 3, 1e-2)
    with pytest.raises(RuntimeError, match="Non-SPD"):
        t.factorize()


def test_gmres_properties():
    """gmres (src/is.cpp:123-300): with the exact factorization (tol 0) the preconditioned
    operator is the identity and GMRES stops after one iteration with x = A^{-1} b; with an
    approximate factorization it converges, restarted or not, to the same solution; b = 0
    returns x = 0 without iterating."""
    A, X = neglapl(8, 2), linspace_nd(8, 2)
    b = np.random.default_rng(11).standard_normal(A.n)
    xs = np.linalg.solve(A.to_scipy().toarray(), b)
    t, _ = oracle.setup_tree(A, X, 3, 0.0)
    t.factorize()
    x, it, rel = t.gmres(A, b, iters=50, restart=30, tol=1e-12)
    assert it == 1 and np.linalg.norm(x - xs) <= 1e-12 * np.linalg.norm(xs)
    A, X = neglapl(20, 3), linspace_nd(20, 3)
    t, _ = oracle.setup_tree(A, X, 7, 1e-1)
    t.factorize()
    b = np.random.default_rng(12).standard_normal(A.n)
    sols = []
    for restart in (30, 5, 2):
        x, it, rel = t.gmres(A, b, iters=300, restart=restart, tol=1e-10)
        assert rel < 1e-10, (restart, it, rel)
        assert np.linalg.norm(A.matvec(x) - b) / np.linalg.norm(b) <= 1e-8
        sols.append((restart, it, x))
    # more restarts never help: GMRES(2) needs at least as many iterations as GMRES(30)
    assert sols[2][1] >= sols[0][1]
    x, it, _ = t.gmres(A, np.zeros(A.n), iters=50)
    assert it == 0 and not np.any(x)
