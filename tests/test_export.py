"""MatrixMarket files and tree text exports in the reference's formats (include/mmio.hpp:160-323,
src/tree.cpp:12-55). The number formats are pinned by what libstdc++ prints for
`std::setprecision(20) << v` and for the default `<< v` (the expected strings below were
printed by a g++ 11 build); the readers by round trips and by the reference's matrix files'
layout; the tree exports by the oracle's cluster tree."""
import numpy as np
import pytest

import oracle
from spand_public_amd.export import (dense_mmread, dense_mmwrite, sp_mmwrite, write_clustering, write_clusters,
                                     write_merging, write_stats)
from spand_public_amd.problems import CSC, linspace_nd, mmread, neglapl, neglapl_unsym
from tests.export_oracle import OracleStub, oracle_cluster_rows, oracle_dof2id


def small():
    r = [0, 1, 0, 1, 2, 2]
    c = [0, 0, 1, 1, 1, 2]
    v = [4.0, -1.0, -1.0, 0.1, 2.5, 1e-30]
    return CSC.from_coo(3, np.array(r), np.array(c), np.array(v))


def test_sp_mmwrite_known_text(tmp_path):
    A = small()
    sp_mmwrite(str(tmp_path / "g.mm"), A)
    assert (tmp_path / "g.mm").read_text().splitlines() == [
        "%%MatrixMarket matrix coordinate real general", "3 3 6", "1 1 4", "2 1 -1", "1 2 -1",
        "2 2 0.10000000000000000555", "3 2 2.5", "3 3 1.0000000000000000833e-30"]
    sp_mmwrite(str(tmp_path / "s.mm"), A, "symmetric")
    assert (tmp_path / "s.mm").read_text().splitlines() == [
        "%%MatrixMarket matrix coordinate real symmetric", "3 3 5", "1 1 4", "2 1 -1",
        "2 2 0.10000000000000000555", "3 2 2.5", "3 3 1.0000000000000000833e-30"]
    with pytest.raises(ValueError):
        sp_mmwrite(str(tmp_path / "x.mm"), A, "skew-symmetric")  # the reference spells skew_symmetric


@pytest.mark.parametrize("prop", ["general", "symmetric", "skew_symmetric"])
def test_sp_mm_round_trip(tmp_path, prop):
    if prop == "general":
        A = neglapl_unsym(7, 2, 4)
    elif prop == "symmetric":
        A = neglapl(5, 3)
    else:  # B - B^T: zero diagonal, exactly antisymmetric
        M = neglapl_unsym(6, 2, 1).to_scipy()
        S = (M - M.T).tocoo()
        keep = S.row != S.col
        A = CSC.from_coo(M.shape[0], S.row[keep], S.col[keep], S.data[keep])
    fn = str(tmp_path / "a.mm")
    sp_mmwrite(fn, A, prop)
    B = mmread(fn)
    assert np.array_equal(A.colptr, B.colptr) and np.array_equal(A.rowval, B.rowval)
    assert np.array_equal(A.nzval, B.nzval)  # 20 significant digits round-trip exactly


def test_mmread_reference_style_file(tmp_path):
    """The layout of the reference's mats/*.mm: "hermitian" lower triangle, comments, values
    with a decimal point."""
    A = neglapl(3, 2)
    S = A.to_scipy().tocoo()
    low = [(r, c, v) for r, c, v in zip(S.row, S.col, S.data) if r >= c]
    lines = ["%%MatrixMarket matrix coordinate real hermitian", "%a comment", f"9 9 {len(low)}"]
    lines += [f"{r + 1} {c + 1} {v:.1f}" for r, c, v in sorted(low, key=lambda t: (t[1], t[0]))]
    (tmp_path / "h.mm").write_text("\n".join(lines) + "\n")
    B = mmread(str(tmp_path / "h.mm"))
    assert np.array_equal(B.colptr, A.colptr) and np.array_equal(B.rowval, A.rowval)
    assert np.array_equal(B.nzval, A.nzval)


def test_dense_mm_known_text_and_round_trip(tmp_path):
    M = np.array([[1.0, 2.5], [1e-7, 123456789.0]])
    fn = str(tmp_path / "d.mm")
    dense_mmwrite(fn, M)
    assert (tmp_path / "d.mm").read_text().splitlines() == [
        "%%MatrixMarket matrix array real general", "2 2", "1", "1e-07", "2.5", "1.23457e+08"]
    assert np.array_equal(dense_mmread(fn), np.array([[1.0, 2.5], [1e-7, 1.23457e8]]))
    R = np.random.default_rng(3).integers(-50, 50, (4, 3)).astype(float) / 8  # exact in 6 digits
    dense_mmwrite(fn, R)
    assert np.array_equal(dense_mmread(fn), R)
    dense_mmwrite(fn, np.arange(9.0).reshape(3, 3), "symmetric")
    assert (tmp_path / "d.mm").read_text().splitlines()[2:] == ["0", "3", "6", "4", "7", "8"]


def test_tree_exports_from_oracle(tmp_path):
    n, d, nl = 10, 2, 4
    A, X = neglapl(n, d), linspace_nd(n, d)
    o, _ = oracle.setup_tree(A, X, nl, 1e-2)
    o.factorize()
    t = OracleStub(o, A.n, nl)
    rows, ids = oracle_cluster_rows(o), oracle_dof2id(o, A.n)
    # every dof in exactly one level-0 cluster, whose original size counts it
    assert (ids >= 0).all()
    sizes = dict(zip(o.bottom_clusters(0)[:, 0], o.bottom_clusters(0)[:, 3]))
    assert all(np.count_nonzero(ids == k) == s for k, s in sizes.items())
    # merge levels: 0 for the partition's clusters, and a parent sits one level up
    assert (rows[: len(sizes), 2] == 0).all()
    ml = dict(zip(rows[:, 0], rows[:, 2]))
    assert all(ml[p] == m + 1 for _, _, m, p, *_ in rows if p >= 0)
    write_clustering(t, X, str(tmp_path / "c.txt"))
    L = (tmp_path / "c.txt").read_text().splitlines()
    assert L[0] == f"{A.n};{d};{nl}" and L[1] == "X;id" and len(L) == A.n + 2
    for i in (0, 37, A.n - 1):
        xs, k = L[2 + i].split(";")
        assert [float(v) for v in xs.split()] == list(X[:, i]) and int(k) == ids[i]
    write_merging(t, str(tmp_path / "m.txt"))
    M = (tmp_path / "m.txt").read_text().splitlines()
    assert M[0] == "child;parent"
    assert M[1:] == [f"{r[0]};{r[3]}" for r in rows if r[3] >= 0]
    write_clusters(t, str(tmp_path / "k.txt"))
    K = (tmp_path / "k.txt").read_text().splitlines()
    assert K[0] == "id;lvl;mergeLvl;name" and K[1:] == [f"{r[0]};{r[1]};{r[2]};{r[0]}" for r in rows]
    write_stats(t, str(tmp_path / "s.txt"))
    S = (tmp_path / "s.txt").read_text().splitlines()
    assert S[0] == "id;size;rank;Rdiag" and S[1:] == [f"{r[0]};{r[4]};{r[5]};" for r in rows]
