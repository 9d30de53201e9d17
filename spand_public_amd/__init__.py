"""spaND (sparsified nested dissection) factorization and preconditioner on one MI355X.

Python mirror of the reference's ``Tree`` API (include/tree.h:130-197) and of its Krylov
driver ``cg`` (include/is.h:13), over the C ABI of ``libspand.so`` (include/spand.h). The
numerical work runs in hand-written HIP kernels for gfx950; this module only marshals
arrays. Usage follows the reference driver (tests/spaND.cpp:236-367)::

    t = Tree(nlevels); t.set_tol(1e-2); t.set_use_geo(True); t.set_Xcoo(X)
    t.partition(symmetric_graph(A)); t.assemble(A); t.factorize()
    x = t.solve(b)                       # one preconditioner apply (Tree::solve)
    x, info = cg(A, b, t, iters=100, tol=1e-12)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import problems
from ._lib import KrylovStats, SpandError, lib
from .problems import CSC

__all__ = ["Tree", "Matrix", "cg", "cg_device", "gmres", "gmres_device", "problems", "SpandError", "CSC", "FP64", "FP32"]

FP64 = 0
FP32 = 1
_SYMM = {"SPD": 0, "SYM": 1, "GEN": 2}


def _check(handle, status):
    if status != 0:
        msg = lib().spand_last_error(handle)
        raise SpandError(status, msg.decode() if msg else "")


class Tree:
    """The reference's spaND ``Tree`` (include/tree.h): partition -> assemble -> factorize -> solve."""

    def __init__(self, nlevels: int):
        self._L = lib()
        h = C.c_void_p()
        _check(None, self._L.spand_create(int(nlevels), C.byref(h)))
        self._h = h
        self.nlevels = int(nlevels)
        self._N = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                self._L.spand_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- options (include/tree.h:133-147) --
    def set_tol(self, tol: float):
        _check(self._h, self._L.spand_set_tol(self._h, float(tol)))

    def set_skip(self, skip: int):
        _check(self._h, self._L.spand_set_skip(self._h, int(skip)))

    def set_verb(self, verb: bool):
        _check(self._h, self._L.spand_set_verb(self._h, int(bool(verb))))

    def set_symm_kind(self, kind: str):
        _check(self._h, self._L.spand_set_symm_kind(self._h, _SYMM[kind]))

    def set_use_geo(self, geo: bool):
        _check(self._h, self._L.spand_set_use_geo(self._h, int(bool(geo))))

    def set_Xcoo(self, X):
        X = np.asarray(X, dtype=np.float64)
        dim, N = X.shape
        _check(self._h, self._L.spand_set_Xcoo(self._h, dim, N, np.ascontiguousarray(X.T).reshape(-1)))

    def set_use_sparsify(self, use: bool):
        _check(self._h, self._L.spand_set_use_sparsify(self._h, int(bool(use))))

    def set_monitor_flops(self, on: bool = True):
        """Tree::set_monitor_flops: record the shape of every dense call of the next
        factorizations (flops_log)."""
        _check(self._h, self._L.spand_set_monitor_flops(self._h, int(bool(on))))

    FLOPS_KINDS = ("potf", "panel", "gemm", "syrk", "geqp3", "getf")

    def flops_log(self) -> np.ndarray:
        """The shape log of the last factorization (spand_get_flops_log): an (n, 5) int32 array
        of (level, kind, rows, cols, inner) in the reference's call order; kind indexes
        FLOPS_KINDS."""
        n = C.c_longlong()
        _check(self._h, self._L.spand_get_flops_log(self._h, C.byref(n), None))
        out = np.zeros(5 * max(n.value, 1), dtype=np.int32)
        _check(self._h, self._L.spand_get_flops_log(self._h, C.byref(n), out.ctypes.data_as(C.c_void_p)))
        return out[: 5 * n.value].reshape(n.value, 5)

    def set_precision(self, precision: int):
        """FP64 (default) or FP32 factor storage and arithmetic (no reference counterpart)."""
        _check(self._h, self._L.spand_set_precision(self._h, int(precision)))

    # -- pipeline --
    def partition(self, A: CSC) -> np.ndarray:
        """Returns the ClusterIDs (N x 6: self.lvl, self.sep, l.lvl, l.sep, r.lvl, r.sep)."""
        out = np.zeros(6 * A.n, dtype=np.int32)
        _check(self._h, self._L.spand_partition(self._h, A.n, A.colptr, A.rowval, out.ctypes.data))
        self._N = A.n
        return out.reshape(A.n, 6)

    def assemble(self, A: CSC):
        _check(self._h, self._L.spand_assemble(self._h, A.n, A.colptr, A.rowval, A.nzval))

    def assemble_values_device(self, d_vals: int, nnz: int, stream: int = 0):
        """New values of the assembled pattern from a device array (nnz float64 entries in the
        CSC order given to assemble(), e.g. a torch CUDA tensor's data_ptr())."""
        _check(self._h, self._L.spand_assemble_values_device(self._h, C.c_void_p(d_vals), int(nnz),
                                                              C.c_void_p(stream)))

    def factorize(self):
        _check(self._h, self._L.spand_factorize(self._h))

    def solve(self, b) -> np.ndarray:
        """Returns M^{-1} b for a vector (N,) or a block (N, nrhs) on the host."""
        b = np.asarray(b, dtype=np.float64)
        x = np.asfortranarray(b.reshape(b.shape[0], -1)).copy(order="F")
        _check(self._h, self._L.spand_solve(self._h, x.ctypes.data, x.shape[1], x.shape[0]))
        return x.reshape(b.shape)

    def solve_device(self, d_ptr: int, nrhs: int, ld: int, stream: int = 0):
        """In-place M^{-1} on a device buffer (e.g. a torch CUDA tensor's data_ptr())."""
        _check(self._h, self._L.spand_solve_device(self._h, C.c_void_p(d_ptr), int(nrhs), int(ld),
                                                    C.c_void_p(stream)))

    def solve_transpose(self, b) -> np.ndarray:
        """Returns M^{-T} b (the adjoint apply; equal to solve() for an SPD tree)."""
        b = np.asarray(b, dtype=np.float64)
        x = np.asfortranarray(b.reshape(b.shape[0], -1)).copy(order="F")
        _check(self._h, self._L.spand_solve_transpose(self._h, x.ctypes.data, x.shape[1], x.shape[0]))
        return x.reshape(b.shape)

    def solve_transpose_device(self, d_ptr: int, nrhs: int, ld: int, stream: int = 0):
        """In-place M^{-T} on a device buffer."""
        _check(self._h, self._L.spand_solve_transpose_device(self._h, C.c_void_p(d_ptr), int(nrhs), int(ld),
                                                              C.c_void_p(stream)))

    # -- info --
    def get_N(self) -> int:
        v = C.c_int()
        _check(self._h, self._L.spand_get_N(self._h, C.byref(v)))
        return v.value

    def get_stop(self) -> int:
        v = C.c_int()
        _check(self._h, self._L.spand_get_stop(self._h, C.byref(v)))
        return v.value

    def nnz(self) -> int:
        v = C.c_longlong()
        _check(self._h, self._L.spand_nnz(self._h, C.byref(v)))
        return v.value

    def get_assembly_perm(self) -> np.ndarray:
        p = np.zeros(self.get_N(), dtype=np.int32)
        _check(self._h, self._L.spand_get_assembly_perm(self._h, p))
        return p

    def get_nlevels(self) -> int:
        v = C.c_int()
        _check(self._h, self._L.spand_get_nlevels(self._h, C.byref(v)))
        return v.value

    def get_clusters(self) -> np.ndarray:
        """Tree::get_clusters (src/tree.cpp:265-273) as rows of (order, level, merge level,
        parent order or -1, original size, size, start), in the reference's list order."""
        k = C.c_int()
        _check(self._h, self._L.spand_get_clusters(self._h, None, C.byref(k)))
        out = np.zeros((k.value, 7), dtype=np.int32)
        _check(self._h, self._L.spand_get_clusters(self._h, out.ctypes.data, C.byref(k)))
        return out

    def get_dof2ID(self) -> np.ndarray:
        """Tree::get_dof2ID (src/tree.cpp:275-284): order of each dof's level-0 cluster."""
        ids = np.zeros(self.get_N(), dtype=np.int32)
        _check(self._h, self._L.spand_get_dof2ID(self._h, ids))
        return ids

    def get_trailing_mat(self):
        """Current trailing matrix (scipy CSC, both triangles), Tree::get_trailing_mat."""
        import scipy.sparse as sp

        n = C.c_longlong()
        _check(self._h, self._L.spand_get_trailing_coo(self._h, C.byref(n), None, None, None))
        I = np.zeros(n.value, dtype=np.int32)
        J = np.zeros(n.value, dtype=np.int32)
        V = np.zeros(n.value)
        _check(self._h, self._L.spand_get_trailing_coo(self._h, C.byref(n), I.ctypes.data, J.ctypes.data,
                                                        V.ctypes.data))
        N = self.get_N()
        M = sp.coo_matrix((V, (I, J)), shape=(N, N)).tocsc()
        off = I != J
        Mt = sp.coo_matrix((V[off], (J[off], I[off])), shape=(N, N)).tocsc()
        return (M + Mt).tocsc()

    def level_stats(self) -> np.ndarray:
        out = np.zeros(20 * self.nlevels)
        _check(self._h, self._L.spand_get_level_stats(self._h, out))
        return out.reshape(self.nlevels, 20)

    TUNE_FIELDS = ("n_elim", "elim_med", "elim_max", "n_spars", "rows_med", "rows_max", "cols_med", "cols_max",
                   "rrqr_reg", "small_elems", "step_waves", "trsm_cw", "source")

    def level_tune(self) -> list:
        """Per level of the last factorization: its shape and the launch configuration chosen
        for it (spand_get_level_tune), as dicts."""
        nf = len(self.TUNE_FIELDS)
        out = np.zeros(nf * self.nlevels, dtype=np.int32)
        _check(self._h, self._L.spand_get_level_tune(self._h, out))
        return [dict(zip(self.TUNE_FIELDS, (int(v) for v in row))) for row in out.reshape(self.nlevels, nf)]

    def status_flags(self) -> dict:
        """Run-time fallbacks taken so far (spand_get_status_flags)."""
        f = C.c_uint()
        _check(self._h, self._L.spand_get_status_flags(self._h, C.byref(f)))
        return {"step_throttle_gave_up": bool(f.value & 1)}

    def factor_info(self):
        s, fb, ab = C.c_double(), C.c_longlong(), C.c_longlong()
        _check(self._h, self._L.spand_get_factor_info(self._h, C.byref(s), C.byref(fb), C.byref(ab)))
        return {"seconds": s.value, "factor_bytes": fb.value, "arena_bytes": ab.value}

    # -- kernel timing (HIP events per launch) --
    def kernel_families(self) -> list:
        return [self._L.spand_kernel_family_name(f).decode() for f in range(self._L.spand_kernel_family_count())]

    def set_profiling(self, on=True):
        """True: time every kernel family; False: off; a family name or a list of names:
        time only those families."""
        if on is True:
            mask = (1 << self._L.spand_kernel_family_count()) - 1
        elif not on:
            mask = 0
        else:
            names = [on] if isinstance(on, str) else list(on)
            fams = self.kernel_families()
            mask = 0
            for n in names:
                mask |= 1 << fams.index(n)
        _check(self._h, self._L.spand_set_profiling(self._h, mask))

    def profile_reset(self):
        _check(self._h, self._L.spand_profile_reset(self._h))

    def kernel_profile(self) -> dict:
        """{family: {"launches", "ms", "bytes", "flops"}} accumulated since the last reset."""
        out = {}
        for f in range(self._L.spand_kernel_family_count()):
            n, ms, by, fl = C.c_longlong(), C.c_double(), C.c_double(), C.c_double()
            _check(self._h, self._L.spand_kernel_profile(self._h, f, C.byref(n), C.byref(ms), C.byref(by), C.byref(fl)))
            if n.value:
                out[self._L.spand_kernel_family_name(f).decode()] = {"launches": n.value, "ms": ms.value,
                                                                      "bytes": by.value, "flops": fl.value}
        return out

    def ops(self) -> np.ndarray:
        n = C.c_int()
        _check(self._h, self._L.spand_num_ops(self._h, C.byref(n)))
        info = np.zeros(8 * n.value, dtype=np.int32)
        _check(self._h, self._L.spand_ops_info(self._h, info))
        return info.reshape(n.value, 8)

    def op_data(self, i: int, info_row) -> dict:
        t, _, _, sa, sb, r, c, _ = (int(v) for v in info_row)
        M = np.zeros(max(r * c, 1))
        tau = np.zeros(max(c, 1)) if t == 6 else None
        _check(self._h, self._L.spand_op_data(self._h, int(i), M.ctypes.data if r * c else None,
                                               tau.ctypes.data if tau is not None else None))
        out = {"M": M[: r * c].reshape(c, r).T.copy()}
        if tau is not None:
            out["tau"] = tau[:c]
        if t == 1:  # ScalingPLUQ: U and the row permutation
            U = np.zeros(max(sa * sa, 1))
            p = np.zeros(max(sa, 1), dtype=np.int32)
            _check(self._h, self._L.spand_op_plu(self._h, int(i), U.ctypes.data, p.ctypes.data))
            out["U"] = U[: sa * sa].reshape(sa, sa).T.copy()
            out["p"] = p[:sa]
        return out


class Matrix:
    """Device copy of a sparse operator for the Krylov solvers (the SpMat of include/is.h)."""

    def __init__(self, A: CSC):
        self._L = lib()
        h = C.c_void_p()
        _check(None, self._L.spand_matrix_create(A.n, A.colptr, A.rowval, A.nzval, C.byref(h)))
        self._h = h
        self.n = A.n
        self.nnz = A.nnz

    def sampled_outer(self, u, v, alpha: float = 1.0) -> np.ndarray:
        """alpha * sum_r u[i, r] v[j, r] for every stored entry (i, j), in CSC order (host
        arrays (N,) or (N, nrhs))."""
        u = np.asfortranarray(np.asarray(u, dtype=np.float64).reshape(self.n, -1))
        v = np.asfortranarray(np.asarray(v, dtype=np.float64).reshape(self.n, -1))
        if u.shape != v.shape:
            raise ValueError("sampled_outer: u and v differ in shape")
        out = np.zeros(self.nnz)
        _check(None, self._L.spand_matrix_sampled_outer(self._h, u.ctypes.data, v.ctypes.data, u.shape[1], self.n,
                                                        float(alpha), out.ctypes.data, 0))
        return out

    def sampled_outer_device(self, d_u: int, d_v: int, nrhs: int, ld: int, alpha: float, d_out: int):
        """Device-pointer form of sampled_outer (column-major blocks with leading dimension ld)."""
        _check(None, self._L.spand_matrix_sampled_outer(self._h, C.c_void_p(d_u), C.c_void_p(d_v), int(nrhs),
                                                        int(ld), float(alpha), C.c_void_p(d_out), 1))

    def set_values_device(self, d_vals: int, nnz: int, stream: int = 0):
        """New values of the same pattern from a device array (CSC order)."""
        _check(None, self._L.spand_matrix_set_values_device(self._h, C.c_void_p(d_vals), int(nnz),
                                                            C.c_void_p(stream)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                self._L.spand_matrix_destroy(h)
            except Exception:
                pass
            self._h = None


def cg_device(A: "Matrix", precond: Tree, d_b: int, d_x: int, nrhs: int, ld: int, iters: int = 100,
              tol: float = 1e-12):
    """PCG on device-resident column-major blocks (pointers d_b, d_x with leading dimension
    ld; d_x holds x0 on entry and the solution on return). Returns (iters per column, stats)."""
    its = np.zeros(nrhs, dtype=np.int32)
    st = KrylovStats()
    status = lib().spand_pcg(precond._h, A._h, C.c_void_p(d_b), C.c_void_p(d_x), int(nrhs), int(ld), int(iters),
                             float(tol), 1, its.ctypes.data, C.byref(st))
    if status not in (0, 7):
        _check(precond._h, status)
    return its, {"status": status, **{f: getattr(st, f) for f, _ in KrylovStats._fields_}}


def tune_choose(elem_bytes: int, shape: dict) -> dict:
    """The measured table's launch configuration for a level shape (host only)."""
    keys = Tree.TUNE_FIELDS[:8]
    sh = np.array([int(shape.get(k, 0)) for k in keys], dtype=np.int32)
    out = np.zeros(5, dtype=np.int32)
    _check(None, lib().spand_tune_choose(int(elem_bytes), sh, out))
    return dict(zip(Tree.TUNE_FIELDS[8:], (int(v) for v in out)))


def gmres_device(A: "Matrix", precond: Tree, d_b: int, d_x: int, nrhs: int, ld: int, iters: int = 100,
                 restart: int = 30, tol: float = 1e-12, transpose: bool = False):
    """GMRES(restart) on device-resident column-major blocks, like cg_device; transpose=True
    solves A^T x = b with M^{-T} (spand_gmres_transpose)."""
    its = np.zeros(nrhs, dtype=np.int32)
    st = KrylovStats()
    fn = lib().spand_gmres_transpose if transpose else lib().spand_gmres
    status = fn(precond._h, A._h, C.c_void_p(d_b), C.c_void_p(d_x), int(nrhs), int(ld), int(iters), int(restart),
                float(tol), 1, its.ctypes.data, C.byref(st))
    if status not in (0, 7):
        _check(precond._h, status)
    return its, {"status": status, **{f: getattr(st, f) for f, _ in KrylovStats._fields_}}


def cg(A, b, precond: Tree, x=None, iters: int = 100, tol: float = 1e-12, raise_on_fail: bool = False):
    """Preconditioned CG (src/is.cpp:39-121) on the GPU for one or several right-hand sides.

    Returns (x, info) with info = {"iters": per-column counts, "stats": KrylovStats fields}.
    """
    Am = A if isinstance(A, Matrix) else Matrix(A)
    b = np.asarray(b, dtype=np.float64)
    B = np.asfortranarray(b.reshape(b.shape[0], -1))
    X = np.zeros_like(B, order="F") if x is None else np.asfortranarray(np.asarray(x, np.float64).reshape(B.shape)).copy(order="F")
    nrhs = B.shape[1]
    its = np.zeros(nrhs, dtype=np.int32)
    st = KrylovStats()
    L = lib()
    status = L.spand_pcg(precond._h, Am._h, B.ctypes.data, X.ctypes.data, nrhs, B.shape[0], int(iters), float(tol),
                         0, its.ctypes.data, C.byref(st))
    if status not in (0, 7) or (status == 7 and raise_on_fail):
        _check(precond._h, status)
    info = {"iters": its, "status": status, **{f: getattr(st, f) for f, _ in KrylovStats._fields_}}
    return X.reshape(b.shape), info


def gmres(A, b, precond: Tree, x=None, iters: int = 100, restart: int = 30, tol: float = 1e-12,
          raise_on_fail: bool = False):
    """Left-preconditioned GMRES(restart) (src/is.cpp:123-300) on the GPU, one or several
    right-hand sides. Returns (x, info) like cg; info["relres_max"] is the true residual."""
    Am = A if isinstance(A, Matrix) else Matrix(A)
    b = np.asarray(b, dtype=np.float64)
    B = np.asfortranarray(b.reshape(b.shape[0], -1))
    X = np.zeros_like(B, order="F") if x is None else np.asfortranarray(np.asarray(x, np.float64).reshape(B.shape)).copy(order="F")
    nrhs = B.shape[1]
    its = np.zeros(nrhs, dtype=np.int32)
    st = KrylovStats()
    status = lib().spand_gmres(precond._h, Am._h, B.ctypes.data, X.ctypes.data, nrhs, B.shape[0], int(iters),
                               int(restart), float(tol), 0, its.ctypes.data, C.byref(st))
    if status not in (0, 7) or (status == 7 and raise_on_fail):
        _check(precond._h, status)
    info = {"iters": its, "status": status, **{f: getattr(st, f) for f, _ in KrylovStats._fields_}}
    return X.reshape(b.shape), info
