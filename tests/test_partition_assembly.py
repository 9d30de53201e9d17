"""Partition / ordering / assembly, host side, without a GPU.

The product partitioner (spand_public_amd, src: partition.cpp) must be bit-exact with the
reference: pinned by PartitionTest.Square's exact ClusterIDs (tests/tests.cpp:378-412),
by the ND consistency invariants (tests/tests.cpp:417-481), by equality with the oracle's
independent restatement, and by Assembly.Consistency (trailing matrix == P^T A P exactly,
tests/tests.cpp:488-547)."""
import json
import os

import numpy as np
import pytest

import oracle
import spand_public_amd as S
from spand_public_amd.problems import linspace_nd, neglapl, symmetric_graph

S_ = lambda l, s: (l, s)
SQUARE_SELF = [S_(0, 0), S_(0, 0), S_(1, 0), S_(0, 1), S_(0, 1), S_(0, 0), S_(0, 0), S_(1, 0), S_(0, 1), S_(0, 1),
               S_(2, 0), S_(2, 0), S_(2, 0), S_(2, 0), S_(2, 0), S_(0, 2), S_(0, 2), S_(1, 1), S_(0, 3), S_(0, 3),
               S_(0, 2), S_(0, 2), S_(1, 1), S_(0, 3), S_(0, 3)]
SQUARE_LEFT = [S_(0, 0), S_(0, 0), S_(0, 0), S_(0, 1), S_(0, 1), S_(0, 0), S_(0, 0), S_(0, 0), S_(0, 1), S_(0, 1),
               S_(0, 0), S_(0, 0), S_(1, 0), S_(0, 1), S_(0, 1), S_(0, 2), S_(0, 2), S_(0, 2), S_(0, 3), S_(0, 3),
               S_(0, 2), S_(0, 2), S_(0, 2), S_(0, 3), S_(0, 3)]
SQUARE_RIGHT = [S_(0, 0), S_(0, 0), S_(0, 1), S_(0, 1), S_(0, 1), S_(0, 0), S_(0, 0), S_(0, 1), S_(0, 1), S_(0, 1),
                S_(0, 2), S_(0, 2), S_(1, 1), S_(0, 3), S_(0, 3), S_(0, 2), S_(0, 2), S_(0, 3), S_(0, 3), S_(0, 3),
                S_(0, 2), S_(0, 2), S_(0, 3), S_(0, 3), S_(0, 3)]


def product_tree(A, X, nl):
    t = S.Tree(nl)
    t.set_use_geo(True)
    t.set_Xcoo(X)
    part = t.partition(symmetric_graph(A))
    return t, part


def test_linspace_nd_kat():
    # tests/tests.cpp:301-310
    X2 = linspace_nd(3, 2)
    assert X2.tolist() == [[0, 0, 0, 1, 1, 1, 2, 2, 2], [0, 1, 2, 0, 1, 2, 0, 1, 2]]
    X3 = linspace_nd(2, 3)
    assert X3.tolist() == [[0, 0, 0, 0, 1, 1, 1, 1], [0, 0, 1, 1, 0, 0, 1, 1], [0, 1, 0, 1, 0, 1, 0, 1]]


def test_neglapl_matches_reference_fixtures(golden_dir):
    from tests.golden.make_golden import coo_hash
    mats = json.load(open(os.path.join(golden_dir, "mats.json")))
    for fn, m in mats.items():
        if fn.startswith("_"):
            continue
        A = neglapl(m["n"], m["d"])
        assert (A.n, A.nnz) == (m["N"], m["nnz"]), fn
        assert coo_hash(A) == m["sha256"], fn
    for fn in ("neglapl_2_5", "neglapl_3_5"):
        z = np.load(os.path.join(golden_dir, fn + ".npz"))
        d = int(fn.split("_")[1])
        A = neglapl(5, d).to_scipy().tocoo()
        ref = {(int(r), int(c)): v for r, c, v in zip(z["rows"], z["cols"], z["vals"])}
        got = {(int(r), int(c)): v for r, c, v in zip(A.row, A.col, A.data)}
        assert ref == got


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_partition_square_kat(impl):
    # tests/tests.cpp:378-412
    A, X = neglapl(5, 2), linspace_nd(5, 2)
    if impl == "product":
        _, part = product_tree(A, X, 3)
    else:
        t = oracle.OracleTree(3)
        t.set_Xcoo(X)
        part = t.partition(A)
    for i in range(25):
        assert tuple(part[i, 0:2]) == SQUARE_SELF[i]
        assert tuple(part[i, 2:4]) == SQUARE_LEFT[i]
        assert tuple(part[i, 4:6]) == SQUARE_RIGHT[i]


def should_be_disconnected(l1, l2, s1, s2):
    # src/util.cpp:28-42
    while l2 > l1:
        l1 += 1
        s1 //= 2
    while l1 > l2:
        l2 += 1
        s2 //= 2
    return s1 != s2


@pytest.mark.parametrize("n,d", [(5, 2), (20, 2), (100, 2), (5, 3), (15, 3)])
def test_partition_consistency_and_oracle_equality(n, d):
    # tests/tests.cpp:417-481 (geometric MND), plus bit-exact equality with the oracle
    A, X = neglapl(n, d), linspace_nd(n, d)
    As = A.to_scipy().tocoo()
    for nl in range(1, 8):
        t, part = product_tree(A, X, nl)
        o = oracle.OracleTree(nl)
        o.set_Xcoo(X)
        opart = o.partition(symmetric_graph(A))
        assert np.array_equal(part, opart)
        assert np.array_equal(t.get_assembly_perm(), o.get_assembly_perm())
        sl, ss = part[:, 0], part[:, 1]
        for i, j in zip(As.row, As.col):
            assert not should_be_disconnected(sl[i], sl[j], ss[i], ss[j])
        for i in range(A.n):
            pl, ps, ll, ls_, rl, rs = part[i]
            if pl == 0:
                assert (ll, ls_) == (pl, ps) and (rl, rs) == (pl, ps)
            else:
                assert pl > ll and pl > rl
                while ll < pl - 1:
                    ll, ls_ = ll + 1, ls_ // 2
                while rl < pl - 1:
                    rl, rs = rl + 1, rs // 2
                assert (ll, ls_) == (pl - 1, 2 * ps) and (rl, rs) == (pl - 1, 2 * ps + 1)


@pytest.mark.parametrize("n,d,nl", [(5, 2, 2), (10, 2, 4), (20, 2, 3), (5, 3, 3), (10, 3, 4), (15, 3, 4)])
def test_assembly_trailing_matrix_exact(n, d, nl):
    # tests/tests.cpp:488-547: get_trailing_mat() == P^T A P exactly
    A, X = neglapl(n, d), linspace_nd(n, d)
    t, _ = product_tree(A, X, nl)
    t.assemble(A)
    P = t.get_assembly_perm()
    Ad = A.to_scipy().toarray()
    M = t.get_trailing_mat().toarray()
    assert np.array_equal(Ad[np.ix_(P, P)], M)


def test_cluster_hierarchy_matches_oracle():
    A, X = neglapl(16, 3), linspace_nd(16, 3)
    o, _ = oracle.setup_tree(A, X, 7, 1e-2)
    # the oracle's cluster table before factorization: leaves then parents, in order
    cl = o.clusters()
    t, _ = product_tree(A, X, 7)
    t.assemble(A)
    assert cl.shape[0] > 0
    # every leaf's [start, start+size) is a run of the assembly permutation
    starts = sorted(int(r[1]) for r in cl if r[4] >= 0)
    assert starts[0] == 0
