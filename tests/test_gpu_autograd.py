"""Adjoint apply and differentiable solve. These have no counterpart in the reference
(parity unpinned); they are pinned by the symmetry of the SPD apply, by the adjoint
identity <M^{-T} x, y> = <x, M^{-1} y> of the GEN apply (M^{-1} itself being pinned to the
oracle by the GEN parity tests), by the definition of the sampled outer product, and by
central finite differences (h = 1e-6)."""
import numpy as np
import pytest

import spand_public_amd as S
from spand_public_amd.problems import DiffusionFV, gaussian_field, linspace_nd, neglapl, neglapl_unsym, symmetric_graph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def tree_for(A, X, nl, tol, prec=S.FP64, symm="SPD"):
    t = S.Tree(nl)
    t.set_use_geo(True)
    t.set_Xcoo(X)
    t.set_tol(tol)
    t.set_precision(prec)
    if symm != "SPD":
        t.set_symm_kind(symm)
    t.partition(symmetric_graph(A))
    t.assemble(A)
    t.factorize()
    return t


def test_reassembly_matches_a_fresh_tree():
    """assemble() with new values after a factorization, then factorize(): the same
    factor (op sequence, solves bitwise) as a tree built from scratch on those values."""
    say("case re-assembly")
    n = 14
    fv, X = DiffusionFV(n, 2), linspace_nd(n, 2)
    A1 = fv.matrix(np.exp(gaussian_field(n, 2, 1.0, 3.0, 1)))
    A2 = fv.matrix(np.exp(gaussian_field(n, 2, 1.0, 3.0, 2)))
    t = tree_for(A1, X, 5, 1e-3)
    t.assemble(A2)
    t.factorize()
    f = tree_for(A2, X, 5, 1e-3)
    assert np.array_equal(t.ops(), f.ops())
    b = np.random.default_rng(3).standard_normal((A2.n, 2))
    assert np.array_equal(t.solve(b), f.solve(b))


@pytest.mark.parametrize("prec", [S.FP64, S.FP32], ids=["fp64", "fp32"])
def test_spd_apply_is_symmetric(prec):
    """<M^{-1} x, y> = <x, M^{-1} y>, and the transposed apply is the apply."""
    say(f"case spd apply symmetric prec={prec}")
    A, X = neglapl(12, 3), linspace_nd(12, 3)
    t = tree_for(A, X, 6, 1e-2, prec)
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal((A.n, 3)), rng.standard_normal((A.n, 3))
    Mx, My, MTx = t.solve(x), t.solve(y), t.solve_transpose(x)
    assert np.array_equal(MTx, Mx)
    for r in range(3):
        lhs, rhs = Mx[:, r] @ y[:, r], x[:, r] @ My[:, r]
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Mx[:, r]) * np.linalg.norm(y[:, r])


@pytest.mark.parametrize("prec", [S.FP64, S.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("kind,n,d,nl,tol", [("unsym", 12, 2, 4, 1e-2), ("unsym", 10, 3, 5, 1e-2),
                                             ("cd", 14, 3, 6, 1e-2), ("unsym", 20, 2, 5, 0.0)])
def test_gen_transpose_is_the_adjoint(kind, n, d, nl, tol, prec):
    """<M^{-T} x, y> = <x, M^{-1} y> for the PLU preconditioner (the transposed apply runs
    its own phase lists, with P as a permuted destination), and M^{-T} differs from M^{-1}."""
    from spand_public_amd.problems import convdiff
    say(f"case gen transpose {kind} {d}D n={n} levels={nl} tol={tol} prec={prec}")
    A = neglapl_unsym(n, d, 3) if kind == "unsym" else convdiff(n, d)
    t = tree_for(A, linspace_nd(n, d), nl, tol, prec, symm="GEN")
    rng = np.random.default_rng(8)
    x, y = rng.standard_normal((A.n, 3)), rng.standard_normal((A.n, 3))
    MTx, My = t.solve_transpose(x), t.solve(y)
    for r in range(3):
        lhs, rhs = MTx[:, r] @ y[:, r], x[:, r] @ My[:, r]
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(MTx[:, r]) * np.linalg.norm(y[:, r]), (lhs, rhs)
    assert np.linalg.norm(MTx - t.solve(x)) > 1e-6 * np.linalg.norm(MTx)
    # the transposed phases are set up once per factorization: again, and after a refactorization
    assert np.array_equal(t.solve_transpose(x), MTx)
    t.factorize()
    assert np.array_equal(t.solve_transpose(x), MTx)
    if tol == 0.0:  # exact factorization: M^{-T} = A^{-T} (to the factor's precision)
        At = A.to_scipy().toarray().T
        bound = 1e-10 if prec == S.FP64 else 1e-4
        assert np.linalg.norm(At @ MTx - x) <= bound * np.linalg.norm(x)


def test_sampled_outer_matches_definition():
    torch = pytest.importorskip("torch")
    say("case sampled outer")
    A = neglapl_unsym(11, 2, 1)
    M = S.Matrix(A)
    rng = np.random.default_rng(2)
    u, v = rng.standard_normal((A.n, 3)), rng.standard_normal((A.n, 3))
    out = M.sampled_outer(u, v, -2.0)
    cols = np.repeat(np.arange(A.n), np.diff(A.colptr))
    ref = -2.0 * (u[A.rowval] * v[cols]).sum(axis=1)
    assert np.allclose(out, ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())
    du = torch.tensor(u.T.copy(), device="cuda")
    dv = torch.tensor(v.T.copy(), device="cuda")
    dout = torch.zeros(A.nnz, dtype=torch.float64, device="cuda")
    M.sampled_outer_device(du.data_ptr(), dv.data_ptr(), 3, A.n, -2.0, dout.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(dout.cpu().numpy(), out)


def test_fv_solve_gradients_match_finite_differences():
    """L = 1/2 ||x - x*||^2 with x = A(kappa)^{-1} b: dL/dkappa (through the harmonic-mean
    face coefficients) and dL/db against central differences in random directions."""
    torch = pytest.importorskip("torch")
    import spand_public_amd.autograd as ad

    say("case fv autograd vs finite differences")
    n, nrhs = 16, 4
    fv = DiffusionFV(n, 2)
    kappa = np.exp(gaussian_field(n, 2, 1.0, 4.0, 11))
    solver = ad.fv_solver(fv, 5, tol=0.0, pcg_tol=1e-14)
    rng = np.random.default_rng(12)
    dev = "cuda"
    b0 = torch.tensor(rng.standard_normal((fv.N, nrhs)), device=dev)
    xs = torch.tensor(rng.standard_normal((fv.N, nrhs)), device=dev)

    def loss(k, b):
        x = ad.spd_solve(ad.fv_values(fv, k), b, solver)
        return 0.5 * ((x - xs) ** 2).sum()

    k0 = torch.tensor(kappa, device=dev)
    k = k0.clone().requires_grad_()
    b = b0.clone().requires_grad_()
    loss(k, b).backward()
    gk, gb = k.grad.detach(), b.grad.detach()
    h = 1e-6
    with torch.no_grad():
        for q in range(6):
            d = torch.tensor(rng.standard_normal(fv.N), device=dev) * k0
            fd = (loss(k0 + h * d, b0) - loss(k0 - h * d, b0)) / (2 * h)
            an = (gk * d).sum()
            assert abs(float(fd - an)) <= 1e-5 * abs(float(an)), ("kappa", q, float(fd), float(an))
        for q in range(3):
            d = torch.tensor(rng.standard_normal((fv.N, nrhs)), device=dev)
            fd = (loss(k0, b0 + h * d) - loss(k0, b0 - h * d)) / (2 * h)
            an = (gb * d).sum()
            assert abs(float(fd - an)) <= 1e-5 * abs(float(an)), ("b", q, float(fd), float(an))
    assert max(max(its) for its in solver.iters) <= 3  # the exact factorization: PCG in one or two steps
