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


def test_device_reassembly_matches_host_assembly():
    """spand_assemble_values_device with the values on the GPU gives the same factor as a host
    assemble() of the same values (op sequence, solves bitwise), SPD and GEN."""
    torch = pytest.importorskip("torch")
    from spand_public_amd.problems import convdiff

    for symm in ("SPD", "GEN"):
        say(f"case device re-assembly {symm}")
        if symm == "SPD":
            n = 14
            fv, X = DiffusionFV(n, 2), linspace_nd(n, 2)
            A1 = fv.matrix(np.exp(gaussian_field(n, 2, 1.0, 3.0, 1)))
            A2 = fv.matrix(np.exp(gaussian_field(n, 2, 1.0, 3.0, 2)))
        else:
            n = 9
            X = linspace_nd(n, 3)
            A1 = convdiff(n, 3)
            A2 = convdiff(n, 3, pe=3.0)
        t = tree_for(A1, X, 5, 1e-3, symm=symm)
        d_vals = torch.tensor(A2.nzval, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t.assemble_values_device(d_vals.data_ptr(), A2.nnz)
        t.factorize()
        f = tree_for(A2, X, 5, 1e-3, symm=symm)
        assert np.array_equal(t.ops(), f.ops())
        b = np.random.default_rng(3).standard_normal((A2.n, 2))
        assert np.array_equal(t.solve(b), f.solve(b))
        # the trailing matrix of the re-assembled tree reads the device values
        t2 = S.Tree(5)
        t2.set_use_geo(True)
        t2.set_Xcoo(X)
        if symm != "SPD":
            t2.set_symm_kind(symm)
        t2.partition(symmetric_graph(A1))
        t2.assemble(A1)
        t2.assemble_values_device(d_vals.data_ptr(), A2.nnz)
        f2 = S.Tree(5)
        f2.set_use_geo(True)
        f2.set_Xcoo(X)
        if symm != "SPD":
            f2.set_symm_kind(symm)
        f2.partition(symmetric_graph(A2))
        f2.assemble(A2)
        assert (t2.get_trailing_mat() != f2.get_trailing_mat()).nnz == 0


def test_gmres_transpose_solves_the_adjoint_system():
    torch = pytest.importorskip("torch")
    from spand_public_amd.problems import convdiff

    say("case gmres transpose")
    n = 12
    A, X = convdiff(n, 3), linspace_nd(n, 3)
    for prec in (S.FP64, S.FP32):
        t = tree_for(A, X, 6, 1e-2, prec, symm="GEN")
        M = S.Matrix(A)
        B = np.random.default_rng(4).standard_normal((A.n, 3))
        d_b = torch.tensor(B.T.copy(), device="cuda")
        d_x = torch.zeros_like(d_b)
        torch.cuda.synchronize()
        its, info = S.gmres_device(M, t, d_b.data_ptr(), d_x.data_ptr(), 3, A.n, iters=200, restart=30, tol=1e-11,
                                   transpose=True)
        assert info["status"] == 0, info
        x = d_x.cpu().numpy().T
        At = A.to_scipy().T.tocsr()
        res = np.linalg.norm(At @ x - B, axis=0) / np.linalg.norm(B, axis=0)
        assert res.max() <= 1e-9, res
        # and it is not the forward system's solution
        xf, _ = S.gmres(A, B, t, iters=200, restart=30, tol=1e-11)
        assert np.linalg.norm(xf - x) > 1e-3 * np.linalg.norm(x)


def test_transpose_solve_on_a_nonblocking_stream_right_after_factorize():
    """The transposed phases are uploaded at the first transposed solve; a caller's
    non-blocking stream must see them complete."""
    torch = pytest.importorskip("torch")
    from spand_public_amd.problems import convdiff

    say("case transpose solve on a non-blocking stream")
    A, X = convdiff(10, 3), linspace_nd(10, 3)
    t = tree_for(A, X, 5, 1e-2, symm="GEN")
    B = np.random.default_rng(7).standard_normal((A.n, 2))
    ref = t.solve_transpose(B)
    t.factorize()
    s = torch.cuda.Stream()
    d = torch.tensor(np.asfortranarray(B).T.copy(), device="cuda")
    torch.cuda.synchronize()
    t.solve_transpose_device(d.data_ptr(), 2, A.n, stream=s.cuda_stream)
    s.synchronize()
    assert np.array_equal(d.cpu().numpy().T, ref)


def test_precision_change_after_factorize_keeps_the_factor_type():
    say("case set_precision after factorize")
    A, X = neglapl(12, 3), linspace_nd(12, 3)
    t = tree_for(A, X, 6, 1e-2, S.FP32)
    b = np.random.default_rng(9).standard_normal(A.n)
    ref = t.solve(b)
    t.set_precision(S.FP64)  # applies to the next factorization only
    assert np.array_equal(t.solve(b), ref)
    x, info = S.cg(A, b, t, iters=100, tol=1e-12)
    assert info["status"] == 0
    t.factorize()  # now FP64
    f = tree_for(A, X, 6, 1e-2, S.FP64)
    assert np.array_equal(t.solve(b), f.solve(b))


def test_failed_refactorization_leaves_no_factor():
    say("case failed refactorization")
    A, X = neglapl(10, 2), linspace_nd(10, 2)
    t = tree_for(A, X, 3, 1e-2)
    b = np.ones(A.n)
    t.solve(b)
    bad = S.CSC(A.n, A.colptr, A.rowval, -A.nzval)
    t.assemble(bad)
    with pytest.raises(S.SpandError, match="NOT_SPD"):
        t.factorize()
    with pytest.raises(S.SpandError, match="STATE"):
        t.solve(b)
    t.assemble(A)
    t.factorize()
    assert np.allclose(A.matvec(t.solve(b)), b, atol=1e-1)


def fd_check(loss, params, grads, directions, h=1e-6, rtol=1e-5):
    import torch

    worst = 0.0
    with torch.no_grad():
        for which, d in directions:
            plus = [p + h * d if i == which else p for i, p in enumerate(params)]
            minus = [p - h * d if i == which else p for i, p in enumerate(params)]
            fd = (loss(*plus) - loss(*minus)) / (2 * h)
            an = (grads[which] * d).sum()
            rel = abs(float(fd - an)) / abs(float(an))
            worst = max(worst, rel)
            assert rel <= rtol, (which, float(fd), float(an))
    return worst


def test_fv_gradients_with_the_approximate_factorization():
    """The same finite-difference check with tol 1e-2 (the preconditioner is approximate; PCG
    still solves to 1e-13), FP64 and FP32 factors."""
    torch = pytest.importorskip("torch")
    import spand_public_amd.autograd as ad

    for prec in (S.FP64, S.FP32):
        say(f"case fv autograd tol 1e-2 prec={prec}")
        n, nrhs = 24, 3
        fv = DiffusionFV(n, 2)
        kappa = np.exp(gaussian_field(n, 2, 1.0, 4.0, 21))
        solver = ad.fv_solver(fv, 6, tol=1e-2, precision=prec, pcg_tol=1e-13)
        rng = np.random.default_rng(22)
        b0 = torch.tensor(rng.standard_normal((fv.N, nrhs)), device="cuda")
        xs = torch.tensor(rng.standard_normal((fv.N, nrhs)), device="cuda")

        def loss(k, b):
            x = ad.sparse_solve(ad.fv_values(fv, k), b, solver)
            return 0.5 * ((x - xs) ** 2).sum()

        k0 = torch.tensor(kappa, device="cuda")
        k = k0.clone().requires_grad_()
        b = b0.clone().requires_grad_()
        loss(k, b).backward()
        dirs = [(0, torch.tensor(rng.standard_normal(fv.N), device="cuda") * k0) for _ in range(4)]
        dirs += [(1, torch.tensor(rng.standard_normal((fv.N, nrhs)), device="cuda")) for _ in range(2)]
        fd_check(loss, [k0, b0], [k.grad, b.grad], dirs)
        assert max(max(i) for i in solver.iters) > 3  # approximate: PCG needs several iterations


def test_gen_gradients_match_finite_differences():
    """Nonsymmetric A (convection-diffusion values scaled per entry by exp(theta)): the adjoint
    solve runs GMRES on A^T with M^{-T}; dL/dtheta and dL/db against central differences."""
    torch = pytest.importorskip("torch")
    import spand_public_amd.autograd as ad
    from spand_public_amd.problems import convdiff

    for tol in (0.0, 1e-2):
        say(f"case gen autograd tol={tol}")
        n, nrhs = 9, 2
        A, X = convdiff(n, 3), linspace_nd(n, 3)
        solver = ad.SparseSolver(A.colptr, A.rowval, X, 5, tol=tol, symm="GEN", krylov_tol=1e-13)
        rng = np.random.default_rng(31)
        base = torch.tensor(A.nzval, device="cuda")
        th0 = torch.tensor(0.1 * rng.standard_normal(A.nnz), device="cuda")
        b0 = torch.tensor(rng.standard_normal((A.n, nrhs)), device="cuda")
        xs = torch.tensor(rng.standard_normal((A.n, nrhs)), device="cuda")

        def loss(th, b):
            x = ad.sparse_solve(base * torch.exp(th), b, solver)
            return 0.5 * ((x - xs) ** 2).sum()

        th = th0.clone().requires_grad_()
        b = b0.clone().requires_grad_()
        loss(th, b).backward()
        dirs = [(0, torch.tensor(rng.standard_normal(A.nnz), device="cuda")) for _ in range(4)]
        dirs += [(1, torch.tensor(rng.standard_normal((A.n, nrhs)), device="cuda")) for _ in range(2)]
        fd_check(loss, [th0, b0], [th.grad, b.grad], dirs)


def test_two_solves_before_one_backward():
    """L = L(kappa1) + L(kappa2) on one solver: the second forward replaces the factorization,
    and the first solve's backward restores its own before the adjoint solve."""
    torch = pytest.importorskip("torch")
    import spand_public_amd.autograd as ad

    say("case two solves, one backward")
    n, nrhs = 16, 2
    fv = DiffusionFV(n, 2)
    solver = ad.fv_solver(fv, 5, tol=1e-2, pcg_tol=1e-13)
    rng = np.random.default_rng(41)
    k1 = torch.tensor(np.exp(gaussian_field(n, 2, 1.0, 4.0, 1)), device="cuda")
    k2 = torch.tensor(np.exp(gaussian_field(n, 2, 1.0, 4.0, 2)), device="cuda")
    b0 = torch.tensor(rng.standard_normal((fv.N, nrhs)), device="cuda")
    xs = torch.tensor(rng.standard_normal((fv.N, nrhs)), device="cuda")

    def loss(ka, kb):
        x1 = ad.sparse_solve(ad.fv_values(fv, ka), b0, solver)
        x2 = ad.sparse_solve(ad.fv_values(fv, kb), b0, solver)
        return 0.5 * ((x1 - xs) ** 2).sum() + 0.25 * ((x2 + xs) ** 2).sum()

    ka, kb = k1.clone().requires_grad_(), k2.clone().requires_grad_()
    loss(ka, kb).backward()
    dirs = [(0, torch.tensor(rng.standard_normal(fv.N), device="cuda") * k1) for _ in range(3)]
    dirs += [(1, torch.tensor(rng.standard_normal(fv.N), device="cuda") * k2) for _ in range(3)]
    fd_check(loss, [k1, k2], [ka.grad, kb.grad], dirs)
