"""Text exports of a factorized tree and MatrixMarket writers, in the reference's file formats.

* ``write_clustering`` / ``write_merging`` / ``write_clusters`` / ``write_stats``: the
  reference's exports of a tree (src/tree.cpp:12-55), read from ``Tree.get_dof2ID`` and
  ``Tree.get_clusters``. This build records no R diagonals (the reference keeps them only with
  ``set_monitor_Rdiag(true)``, src/tree.cpp:1315), so the Rdiag column of ``write_stats`` is
  empty, as in the reference's default.
* ``write_log_flops``: the shape log of the dense calls (src/tree.cpp:60-77) from
  ``Tree.flops_log`` (recorded under ``Tree.set_monitor_flops``). The time column is 0: the
  calls run batched per level on the GPU, so a per-call wall time does not exist.
* ``sp_mmwrite`` / ``dense_mmwrite`` / ``dense_mmread``: include/mmio.hpp:221-327. Sparse
  entries are written column by column with 20 significant digits; dense entries with the
  stream's default 6 significant digits, as the reference writes them. The sparse reader is
  ``problems.mmread``.

Numbers are printed the way C++ streams print doubles (``%g`` with the stream's precision).
"""
from __future__ import annotations

import numpy as np

__all__ = ["write_clustering", "write_merging", "write_clusters", "write_stats", "write_log_flops", "sp_mmwrite",
           "dense_mmwrite", "dense_mmread"]

_PROPS = ("general", "symmetric", "hermitian", "skew_symmetric")


def _g(v: float, prec: int = 6) -> str:
    """A double as `std::ostream << v` prints it with the given precision (defaultfloat)."""
    return "%.*g" % (prec, v)


def write_clustering(t, X, fn: str) -> None:
    """src/tree.cpp:12-26: "N;dim;nlevels", "X;id", then per dof its coordinates and the
    order of its level-0 cluster. X is dim x N (one column per dof)."""
    X = np.asarray(X, dtype=np.float64)
    ids = t.get_dof2ID()
    N = t.get_N()
    assert X.shape[1] == N, "X must be dim x N"
    with open(fn, "w") as f:
        f.write(f"{N};{X.shape[0]};{t.get_nlevels()}\n")
        f.write("X;id\n")
        for i in range(N):
            f.write("".join(" " + _g(X[d, i]) for d in range(X.shape[0])) + f";{ids[i]}\n")


def write_merging(t, fn: str) -> None:
    """src/tree.cpp:28-37: "child;parent" orders of every listed cluster with a parent."""
    cl = t.get_clusters()
    with open(fn, "w") as f:
        f.write("child;parent\n")
        for r in cl:
            if r[3] >= 0:
                f.write(f"{r[0]};{r[3]}\n")


def write_clusters(t, fn: str) -> None:
    """src/tree.cpp:39-46: "id;lvl;mergeLvl;name" (the name is the order)."""
    cl = t.get_clusters()
    with open(fn, "w") as f:
        f.write("id;lvl;mergeLvl;name\n")
        for r in cl:
            f.write(f"{r[0]};{r[1]};{r[2]};{r[0]}\n")


def write_stats(t, fn: str) -> None:
    """src/tree.cpp:48-56: "id;size;rank;Rdiag" with the original size and the kept size."""
    cl = t.get_clusters()
    with open(fn, "w") as f:
        f.write("id;size;rank;Rdiag\n")
        for r in cl:
            f.write(f"{r[0]};{r[4]};{r[5]};\n")


# shape-log kinds (Tree.FLOPS_KINDS order) -> the reference's four lists
_FLOPS_LIST = {0: "pivot", 5: "pivot", 1: "panel", 2: "gemm", 3: "gemm", 4: "rrqr"}


def write_log_flops(log, fn: str, nlevels: int | None = None) -> None:
    """src/tree.cpp:60-77: per level, the pivot, panel, gemm and rrqr records (each list in
    call order) as "level;kind;rows;cols;inner;time". ``log`` is ``Tree.flops_log()`` (or an
    (n, 5) array of (level, kind, rows, cols, inner)); ``nlevels`` defaults to one past the
    last level in the log."""
    log = np.asarray(log, dtype=np.int64).reshape(-1, 5)
    nl = int(log[:, 0].max()) + 1 if nlevels is None and len(log) else int(nlevels or 0)
    with open(fn, "w") as f:
        for lvl in range(nl):
            rows = log[log[:, 0] == lvl]
            for name in ("pivot", "panel", "gemm", "rrqr"):
                for _, kind, m, n, k in rows:
                    if _FLOPS_LIST[int(kind)] != name:
                        continue
                    if name == "pivot":
                        f.write(f"{lvl};pivot;{m};0;0;0\n")
                    elif name == "gemm":
                        f.write(f"{lvl};gemm;{m};{n};{k};0\n")
                    else:
                        f.write(f"{lvl};{name};{m};{n};0;0\n")


def sp_mmwrite(fn: str, A, prop: str = "general") -> None:
    """include/mmio.hpp:265-295. A is a problems.CSC or a scipy sparse matrix; with
    "symmetric" / "hermitian" only entries with row >= col are written, with
    "skew_symmetric" only row > col."""
    if prop not in _PROPS:
        raise ValueError(f"prop must be one of {_PROPS}")
    if hasattr(A, "colptr"):
        n_rows = n_cols = A.n
        colptr, rowval, vals = A.colptr, A.rowval, A.nzval
    else:
        M = A.tocsc()
        M.sort_indices()
        n_rows, n_cols = M.shape
        colptr, rowval, vals = M.indptr, M.indices, M.data
    keep = []
    for j in range(n_cols):
        for k in range(colptr[j], colptr[j + 1]):
            i = int(rowval[k])
            if prop in ("symmetric", "hermitian") and i < j:
                continue
            if prop == "skew_symmetric" and i <= j:
                continue
            keep.append((i, j, float(vals[k])))
    with open(fn, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real {prop}\n")
        f.write(f"{n_rows} {n_cols} {len(keep)}\n")
        for i, j, v in keep:
            f.write(f"{i + 1} {j + 1} {_g(v, 20)}\n")


def dense_mmwrite(fn: str, M, prop: str = "general") -> None:
    """include/mmio.hpp:300-323: column-major values, 6 significant digits (the stream's
    default precision); the property filters entries as in sp_mmwrite."""
    if prop not in _PROPS:
        raise ValueError(f"prop must be one of {_PROPS}")
    M = np.asarray(M, dtype=np.float64)
    if M.ndim == 1:
        M = M[:, None]
    with open(fn, "w") as f:
        f.write(f"%%MatrixMarket matrix array real {prop}\n")
        f.write(f"{M.shape[0]} {M.shape[1]}\n")
        for j in range(M.shape[1]):
            for i in range(M.shape[0]):
                if prop in ("symmetric", "hermitian") and i < j:
                    continue
                if prop == "skew_symmetric" and i <= j:
                    continue
                f.write(_g(M[i, j]) + "\n")


def dense_mmread(fn: str) -> np.ndarray:
    """include/mmio.hpp:221-263: a general column-major array file."""
    with open(fn) as f:
        header = f.readline().lower().split()
        if header[0] != "%%matrixmarket" or header[2] != "array":
            raise ValueError(f"{fn}: not a MatrixMarket array file")
        if len(header) > 4 and header[4] != "general":
            raise ValueError(f"{fn}: only general array files are read")
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
        m, n = (int(t) for t in line.split())
        vals = np.array([float(t) for t in f.read().split()], dtype=np.float64)
    if vals.size != m * n:
        raise ValueError(f"{fn}: expected {m * n} values, found {vals.size}")
    return vals.reshape(n, m).T.copy()
