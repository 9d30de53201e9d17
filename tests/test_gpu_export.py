"""Tree::get_clusters / get_dof2ID (src/tree.cpp:265-284) and the text exports built on them
(src/tree.cpp:12-55): the GPU tree lists the same clusters as the oracle (orders, levels,
merge levels, parents, original and kept sizes, starts), so every export file is
byte-identical to the one written from the oracle's tree."""
import numpy as np
import pytest

import oracle
import spand_public_amd as S
from spand_public_amd.export import write_clustering, write_clusters, write_merging, write_stats
from spand_public_amd.problems import linspace_nd, neglapl, neglapl_unsym, symmetric_graph
from tests.export_oracle import OracleStub, oracle_cluster_rows, oracle_dof2id

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("symm,n,d,nl,tol", [("SPD", 20, 2, 5, 1e-2), ("SPD", 10, 3, 5, 1e-1),
                                             ("GEN", 15, 2, 5, 1e-2)])
def test_clusters_and_exports_match_oracle(tmp_path, symm, n, d, nl, tol):
    print(f"case export {symm} {d}D n={n} levels={nl} tol={tol}", flush=True)
    A = neglapl(n, d) if symm == "SPD" else neglapl_unsym(n, d, 2)
    X = linspace_nd(n, d)
    o, _ = oracle.setup_tree(A, X, nl, tol, symm=symm)
    t = S.Tree(nl)
    t.set_use_geo(True)
    t.set_Xcoo(X)
    t.set_tol(tol)
    if symm != "SPD":
        t.set_symm_kind(symm)
    t.partition(symmetric_graph(A))
    t.assemble(A)
    # every level's clusters are listed from the partition on, with their partition sizes
    assert np.array_equal(t.get_clusters(), oracle_cluster_rows(o))
    assert np.array_equal(t.get_dof2ID(), oracle_dof2id(o, A.n))
    o.factorize()
    t.factorize()
    want = oracle_cluster_rows(o)
    assert len(want) > len(o.bottom_clusters(0)) and (want[:, 5] < want[:, 4]).any()  # compressed
    assert np.array_equal(t.get_clusters(), want)
    assert np.array_equal(t.get_dof2ID(), oracle_dof2id(o, A.n))
    stub = OracleStub(o, A.n, nl)
    for name, w in (("clustering", lambda tr, f: write_clustering(tr, X, f)), ("merging", write_merging),
                    ("clusters", write_clusters), ("stats", write_stats)):
        w(t, str(tmp_path / f"gpu_{name}.txt"))
        w(stub, str(tmp_path / f"ref_{name}.txt"))
        assert (tmp_path / f"gpu_{name}.txt").read_bytes() == (tmp_path / f"ref_{name}.txt").read_bytes(), name
