"""Host-side problem generators and sparse-matrix helpers.

These produce the matrices spaND is run on, as CSC arrays (``colptr``, ``rowval``,
``nzval``) with sorted row indices and int32 indices, the layout of
``Eigen::SparseMatrix<double, ColMajor, int>`` that the reference's ``Tree::partition`` /
``Tree::assemble`` consume (include/spaND.h:59).

* ``neglapl(n, d)``: Dirichlet negative Laplacian, 2D 5-point (4/-1) or 3D 7-point
  (6/-1); identical entries to the reference fixtures ``mats/neglapl_{d}_{n}.mm``.
* ``linspace_nd(n, d)``: grid coordinates with dimension 0 slowest
  (src/util.cpp:488-517). Dof ``id`` of ``neglapl`` sits at ``linspace_nd[:, id]``; the
  Laplacian is axis-symmetric, so the reference's pairing of ``mats/*.mm`` (coordinate 0
  fastest) with ``linspace_nd`` describes the same matrix.
* ``neglapl_unsym(n, d, seed)``: ``neglapl`` with every stored entry perturbed by
  U(-0.1, 0.1) draws, the unsymmetric test matrix of the reference's tests
  (tests/tests.cpp:98-109).
* ``convdiff(n, d, pe, velocity)``: Dirichlet convection-diffusion, 7-point (5-point in
  2D) diffusion plus first-order upwind convection, scaled by h^2 like ``neglapl``. No
  reference counterpart (a nonsymmetric workload for the GEN / PLU path).
* ``DiffusionFV(n, d)``: -div(kappa grad u) with cell-wise kappa, 5-point (2D) / 7-point
  (3D) finite volumes with harmonic-mean face coefficients, the variable-coefficient
  workload; ``gaussian_field`` draws log-kappa. No reference counterpart.
* ``symmetric_graph(A)``: |A| + |A^T| + I pattern (src/util.cpp:47-63).
* ``mmread(path)``: MatrixMarket coordinate reader with the reference's symmetric /
  hermitian expansion (include/mmio.hpp:160-219).
"""
from __future__ import annotations

import numpy as np


class CSC:
    """Square sparse matrix in CSC form (sorted rows, int32 indices, float64 values)."""

    __slots__ = ("n", "colptr", "rowval", "nzval")

    def __init__(self, n, colptr, rowval, nzval):
        self.n = int(n)
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int32)
        self.rowval = np.ascontiguousarray(rowval, dtype=np.int32)
        self.nzval = np.ascontiguousarray(nzval, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.colptr[-1])

    @staticmethod
    def from_coo(n, rows, cols, vals):
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        vals = np.asarray(vals, dtype=np.float64)
        key = cols * n + rows
        order = np.argsort(key, kind="stable")
        key = key[order]
        vals = vals[order]
        # sum duplicates (Eigen setFromTriplets semantics)
        uniq, start = np.unique(key, return_index=True)
        sums = np.add.reduceat(vals, start) if len(vals) else vals
        r = (uniq % n).astype(np.int32)
        c = (uniq // n).astype(np.int64)
        colptr = np.zeros(n + 1, dtype=np.int64)
        np.add.at(colptr, c + 1, 1)
        colptr = np.cumsum(colptr)
        return CSC(n, colptr.astype(np.int32), r, sums)

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csc_matrix((self.nzval, self.rowval, self.colptr), shape=(self.n, self.n))

    def to_csr_arrays(self):
        """CSR of the same matrix (rowptr, colval, nzval), used by the device SpMV."""
        t = self.transpose()
        return t.colptr, t.rowval, t.nzval

    def transpose(self):
        cols = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.colptr))
        return CSC.from_coo(self.n, cols, self.rowval, self.nzval)

    def matvec(self, x):
        x = np.asarray(x, dtype=np.float64)
        cols = np.repeat(np.arange(self.n), np.diff(self.colptr))
        y = np.zeros((self.n,) + x.shape[1:], dtype=np.float64)
        np.add.at(y, self.rowval, self.nzval[:, None] * x[cols] if x.ndim > 1 else self.nzval * x[cols])
        return y


def linspace_nd(n: int, d: int) -> np.ndarray:
    """Coordinates (d x n^d), dimension 0 slowest (src/util.cpp:488-517)."""
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64)] * d, indexing="ij")
    return np.ascontiguousarray(np.stack([g.reshape(-1) for g in grids], axis=0))


def neglapl(n: int, d: int) -> CSC:
    """Dirichlet negative Laplacian on an n^d grid: diagonal 2d, -1 per grid neighbour."""
    N = n ** d
    idx = np.arange(N, dtype=np.int64)
    coords = np.stack(np.unravel_index(idx, (n,) * d), axis=0)
    rows = [idx]
    cols = [idx]
    vals = [np.full(N, 2.0 * d)]
    stride = 1
    for axis in range(d - 1, -1, -1):
        c = coords[axis]
        for sgn in (-1, 1):
            ok = (c + sgn >= 0) & (c + sgn < n)
            rows.append(idx[ok])
            cols.append(idx[ok] + sgn * stride)
            vals.append(np.full(int(ok.sum()), -1.0))
        stride *= n
    return CSC.from_coo(N, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


class MinStd0:
    """std::minstd_rand0 (Lehmer, a = 16807, m = 2^31 - 1), the libstdc++
    ``std::default_random_engine``, with libstdc++'s ``generate_canonical<double, 53>``
    (two draws: (g1 - 1) + (g2 - 1) * r over r^2, r = 2^31 - 2, in double arithmetic)."""

    M = 2147483647

    def __init__(self, seed: int):
        s = seed % self.M
        self.x = s if s != 0 else 1

    def __call__(self) -> int:
        self.x = (16807 * self.x) % self.M
        return self.x

    def canonical(self) -> float:
        r = 2147483646.0
        s, t = 0.0, 1.0
        for _ in range(2):
            s += float(self() - 1) * t
            t *= r
        v = s / t
        return v if v < 1.0 else float(np.nextafter(1.0, 0.0))

    def uniform(self, a: float, b: float) -> float:
        return self.canonical() * (b - a) + a


def neglapl_unsym(n: int, d: int, seed: int) -> CSC:
    """neglapl(n, d) + U(-0.1, 0.1) on every stored entry, drawn in column-major storage
    order (tests/tests.cpp:98-109: default_random_engine seeded with ``seed``)."""
    A = neglapl(n, d)
    g = MinStd0(seed)
    v = A.nzval.copy()
    for k in range(len(v)):
        v[k] += g.uniform(-0.1, 0.1)
    return CSC(A.n, A.colptr, A.rowval, v)


def convdiff(n: int, d: int, pe: float = 10.0, velocity=(1.0, 0.5, 0.25)) -> CSC:
    """-Laplace(u) + w . grad(u) on an n^d grid, Dirichlet, times h^2 (h the grid step):
    diagonal 2d + pe * sum(v), -1 per grid neighbour, and -pe * v_k more on the upwind
    neighbour along axis k (w = pe * v / h; axis 0 is the slowest index, as in
    ``linspace_nd``). Nonsymmetric for pe != 0, an M-matrix for velocities >= 0."""
    v = np.asarray(velocity[:d], dtype=np.float64)
    N = n ** d
    idx = np.arange(N, dtype=np.int64)
    coords = np.stack(np.unravel_index(idx, (n,) * d), axis=0)
    rows = [idx]
    cols = [idx]
    vals = [np.full(N, 2.0 * d + pe * float(np.abs(v).sum()))]
    stride = 1
    for axis in range(d - 1, -1, -1):
        c = coords[axis]
        for sgn in (-1, 1):
            ok = (c + sgn >= 0) & (c + sgn < n)
            # upwind: the neighbour the flow comes from (behind for v > 0)
            up = (v[axis] > 0 and sgn == -1) or (v[axis] < 0 and sgn == 1)
            rows.append(idx[ok])
            cols.append(idx[ok] + sgn * stride)
            vals.append(np.full(int(ok.sum()), -1.0 - (pe * abs(v[axis]) if up else 0.0)))
        stride *= n
    return CSC.from_coo(N, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def gaussian_field(n: int, d: int, sigma: float, corr: float, seed: int) -> np.ndarray:
    """Stationary Gaussian random field on an n^d grid (dimension 0 slowest): white noise
    N(0,1) from numpy's PCG64 with `seed`, smoothed by a periodic Gaussian kernel of
    standard deviation `corr` cells (FFT), then shifted and scaled to mean 0 and standard
    deviation `sigma`. Flattened to n^d values."""
    w = np.random.default_rng(seed).standard_normal((n,) * d)
    k = np.meshgrid(*[np.fft.fftfreq(n)] * d, indexing="ij")
    filt = np.exp(-2.0 * (np.pi * corr) ** 2 * sum(g * g for g in k))
    g = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    g = (g - g.mean()) / max(g.std(), 1e-300) * sigma
    return g.reshape(-1)


class DiffusionFV:
    """-div(kappa grad u) on an n^d grid of unit cells, kappa constant per cell, Dirichlet
    u = 0 one cell outside the grid. A = sum over faces f of t_f (e_a - e_b)(e_a - e_b)^T
    for the interior face between cells a and b, t_f = 2 kappa_a kappa_b / (kappa_a +
    kappa_b) (harmonic mean), plus t_f e_a e_a^T for each boundary face of cell a with
    t_f = kappa_a. kappa = 1 gives ``neglapl(n, d)`` exactly. Stored entries follow the
    ``neglapl`` pattern in CSC order; entry k of ``values(t)`` is sum_f E[k, f] t_f with
    E[k, f] in {-1, 0, 1} (``entry_index``, ``entry_face``, ``entry_sign``)."""

    def __init__(self, n: int, d: int):
        self.n, self.d = int(n), int(d)
        N = n ** d
        self.N = N
        idx = np.arange(N, dtype=np.int64)
        coords = np.stack(np.unravel_index(idx, (n,) * d), axis=0)
        fa, fb, bnd = [], [], []
        stride = 1
        for axis in range(d - 1, -1, -1):
            c = coords[axis]
            ok = c + 1 < n
            fa.append(idx[ok])
            fb.append(idx[ok] + stride)
            bnd.append(idx[c == 0])
            bnd.append(idx[c == n - 1])
            stride *= n
        self.face_a = np.concatenate(fa)
        self.face_b = np.concatenate(fb)
        self.bnd_cell = np.concatenate(bnd)
        self.nfaces = len(self.face_a) + len(self.bnd_cell)
        # the pattern; an entry's CSC position is the rank of its key col * N + row
        P = neglapl(n, d)
        self.colptr, self.rowval = P.colptr, P.rowval
        keys = np.repeat(np.arange(N, dtype=np.int64), np.diff(P.colptr)) * N + P.rowval
        ni, nb = len(self.face_a), len(self.bnd_cell)
        fa, fb = self.face_a, self.face_b
        f = np.arange(ni, dtype=np.int64)
        # interior face f: +t at (a, a) and (b, b), -t at (a, b) and (b, a); boundary: +t at (a, a)
        r = np.concatenate([fa, fb, fa, fb, self.bnd_cell])
        c = np.concatenate([fa, fb, fb, fa, self.bnd_cell])
        self.entry_face = np.concatenate([f, f, f, f, ni + np.arange(nb, dtype=np.int64)])
        self.entry_sign = np.concatenate([np.ones(2 * ni), -np.ones(2 * ni), np.ones(nb)])
        self.entry_index = np.searchsorted(keys, c * N + r)
        self.nnz = int(P.colptr[-1])

    def transmissibility(self, kappa) -> np.ndarray:
        """Face coefficients t (interior faces, then boundary faces) for cell-wise kappa."""
        kappa = np.asarray(kappa, dtype=np.float64).reshape(-1)
        ka, kb = kappa[self.face_a], kappa[self.face_b]
        return np.concatenate([2.0 * ka * kb / (ka + kb), kappa[self.bnd_cell]])

    def values(self, t) -> np.ndarray:
        """Stored entries (CSC order) of A for face coefficients t."""
        return np.bincount(self.entry_index, weights=self.entry_sign * np.asarray(t)[self.entry_face],
                           minlength=self.nnz)

    def matrix(self, kappa) -> CSC:
        return CSC(self.N, self.colptr, self.rowval, self.values(self.transmissibility(kappa)))


def symmetric_graph(A: CSC) -> CSC:
    """|A| + |A^T| + I (src/util.cpp:47-63)."""
    n = A.n
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(A.colptr))
    r = np.concatenate([np.arange(n), cols, A.rowval])
    c = np.concatenate([np.arange(n), A.rowval, cols])
    v = np.concatenate([np.ones(n), np.abs(A.nzval), np.abs(A.nzval)])
    return CSC.from_coo(n, r, c, v)


def mmread(path: str) -> CSC:
    """MatrixMarket coordinate reader (general / symmetric / hermitian / skew)."""
    with open(path) as f:
        header = f.readline().lower().split()
        assert header[0] == "%%matrixmarket" and header[2] == "coordinate", header
        prop = header[4] if len(header) > 4 else "general"
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
        M, N, K = (int(t) for t in line.split())
        data = np.loadtxt(f, ndmin=2)
    assert data.shape[0] == K and M == N
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int64) - 1
    v = data[:, 2].astype(np.float64)
    prop = prop.replace("_", "-")  # the reference's writer spells it skew_symmetric
    if prop in ("symmetric", "hermitian", "skew-symmetric"):
        off = r != c
        sgn = -1.0 if prop == "skew-symmetric" else 1.0
        r, c, v = (np.concatenate([r, c[off]]), np.concatenate([c, r[off]]),
                   np.concatenate([v, sgn * v[off]]))
    return CSC.from_coo(N, r, c, v)


def auto_nlevels(N: int, l: int = 0) -> int:
    """The CLI's level rule round(log2(N/64)) + l for l <= 0 (tests/spaND.cpp:137-140)."""
    return int(round(np.log(N / 64.0) / np.log(2.0))) + l


def seeded_rhs(N: int, nrhs: int, seed: int) -> np.ndarray:
    """b ~ N(0,1), N x nrhs, from numpy's PCG64 with the given seed."""
    return np.random.default_rng(seed).standard_normal((N, nrhs))
