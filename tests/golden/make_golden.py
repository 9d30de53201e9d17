"""Writes the golden vectors under tests/golden/ (run once in a container that has scipy and
the reference checkout; the outputs are committed, so tests never need either).

* dense.npz: inputs and LAPACK/BLAS outputs from scipy.linalg.lapack / scipy.linalg.blas
  (OpenBLAS inside scipy) for dpotrf('L'), dtrsm (the three variants spaND calls),
  dgeqp3, dormqr('L','T'/'N') on a vector and dgetrf. They pin the oracle's dense kernels
  (oracle/dense.cpp), which restate the calls of src/util.cpp:122-452.
* mats.json: N, nnz and a SHA-256 of the sorted (row, col, value) triplets of each
  Dirichlet Laplacian fixture mats/neglapl_{d}_{n}.mm of the reference, after the
  reference reader's symmetric expansion (include/mmio.hpp:160-219); the test regenerates
  the matrix with spand_public_amd.problems.neglapl and compares.
* neglapl_2_5.npz / neglapl_3_5.npz: the two smallest fixtures as COO arrays.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np
import scipy
import scipy.linalg.blas as blas
import scipy.linalg.lapack as lapack

HERE = os.path.dirname(os.path.abspath(__file__))
REF_MATS = "/root/reference/mats"


def coo_hash(A):
    """SHA-256 of the column-major sorted (row, col, value) triplets of a CSC matrix."""
    cols = np.repeat(np.arange(A.n, dtype=np.int64), np.diff(A.colptr))
    h = hashlib.sha256()
    h.update(np.asarray(A.rowval, dtype=np.int64).tobytes())
    h.update(cols.tobytes())
    h.update(np.asarray(A.nzval, dtype=np.float64).tobytes())
    return h.hexdigest()


def dense_golden(rng):
    out = {}
    # potrf on SPD matrices
    for n in (1, 5, 33, 64):
        B = rng.standard_normal((n, n))
        A = B @ B.T + n * np.eye(n)
        L, info = lapack.dpotrf(A, lower=1, clean=1)
        assert info == 0
        out[f"potrf_{n}_A"] = A
        out[f"potrf_{n}_L"] = np.tril(L)
    # trsm variants: left/lower/notrans, right/lower/trans, right/upper/notrans
    for (m, n) in ((7, 5), (40, 64), (1, 9)):
        Ls = np.tril(rng.standard_normal((m, m))) + 4 * np.eye(m)
        Bl = rng.standard_normal((m, n))
        out[f"trsm_llnn_{m}_{n}_L"] = Ls
        out[f"trsm_llnn_{m}_{n}_B"] = Bl
        out[f"trsm_llnn_{m}_{n}_X"] = blas.dtrsm(1.0, Ls, Bl, side=0, lower=1, trans_a=0, diag=0)
        Lr = np.tril(rng.standard_normal((n, n))) + 4 * np.eye(n)
        Br = rng.standard_normal((m, n))
        out[f"trsm_rltn_{m}_{n}_L"] = Lr
        out[f"trsm_rltn_{m}_{n}_B"] = Br
        out[f"trsm_rltn_{m}_{n}_X"] = blas.dtrsm(1.0, Lr, Br, side=1, lower=1, trans_a=1, diag=0)
        Ur = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n)
        out[f"trsm_runn_{m}_{n}_U"] = Ur
        out[f"trsm_runn_{m}_{n}_B"] = Br
        out[f"trsm_runn_{m}_{n}_X"] = blas.dtrsm(1.0, Ur, Br, side=1, lower=0, trans_a=0, diag=0)
    # geqp3 (all columns free), random matrices of the shapes spaND sees (wide s x c)
    for (m, n) in ((5, 3), (3, 5), (18, 150), (40, 40), (48, 160), (1, 7), (7, 1)):
        A = rng.standard_normal((m, n)) * np.exp(rng.standard_normal(n))[None, :]
        qr, jpvt, tau, work, info = lapack.dgeqp3(A)
        assert info == 0
        out[f"geqp3_{m}_{n}_A"] = A
        out[f"geqp3_{m}_{n}_QR"] = qr
        out[f"geqp3_{m}_{n}_jpvt"] = jpvt.astype(np.int32) - 1
        out[f"geqp3_{m}_{n}_tau"] = tau
        # Q^T x and Q x through dormqr on one vector
        k = min(m, n)
        x = rng.standard_normal((m, 1))
        lw = max(1, m)
        qtx, _, info1 = lapack.dormqr("L", "T", qr[:, :k], tau[:k], x.copy(), lw)
        qx, _, info2 = lapack.dormqr("L", "N", qr[:, :k], tau[:k], x.copy(), lw)
        assert info1 == 0 and info2 == 0
        out[f"ormqr_{m}_{n}_x"] = x[:, 0]
        out[f"ormqr_{m}_{n}_qtx"] = qtx[:, 0]
        out[f"ormqr_{m}_{n}_qx"] = qx[:, 0]
    # getrf with partial pivoting
    for n in (1, 6, 50):
        A = rng.standard_normal((n, n))
        lu, piv, info = lapack.dgetrf(A)
        assert info == 0
        out[f"getrf_{n}_A"] = A
        out[f"getrf_{n}_LU"] = lu
        out[f"getrf_{n}_piv"] = piv.astype(np.int32)
    return out


def main():
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from spand_public_amd.problems import mmread

    rng = np.random.default_rng(20190101)
    dense = dense_golden(rng)
    np.savez_compressed(os.path.join(HERE, "dense.npz"), **dense)
    mats = {"_generator": f"scipy {scipy.__version__}, numpy {np.__version__}"}
    for fn in sorted(os.listdir(REF_MATS)):
        if not fn.startswith("neglapl_"):
            continue
        A = mmread(os.path.join(REF_MATS, fn))
        _, d, n = fn[:-3].split("_")
        mats[fn] = {"d": int(d), "n": int(n), "N": A.n, "nnz": A.nnz, "sha256": coo_hash(A)}
        if fn in ("neglapl_2_5.mm", "neglapl_3_5.mm"):
            cols = np.repeat(np.arange(A.n), np.diff(A.colptr))
            np.savez_compressed(os.path.join(HERE, fn.replace(".mm", ".npz")), rows=A.rowval, cols=cols,
                                vals=A.nzval, N=A.n)
    with open(os.path.join(HERE, "mats.json"), "w") as f:
        json.dump(mats, f, indent=1, sort_keys=True)
    print("wrote", len(dense), "dense arrays and", len(mats) - 1, "matrix checksums")


if __name__ == "__main__":
    main()
