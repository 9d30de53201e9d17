"""Test helper: the oracle's cluster tree in the shape that spand_public_amd.export reads
(Tree::get_clusters rows and Tree::get_dof2ID), derived from the oracle's own cluster records."""
import numpy as np


def oracle_cluster_rows(o):
    """(order, level, merge level, parent order, original size, size, start) per listed
    cluster; the merge level is level - depth (src/cluster.cpp:178-193), the depth counted
    along the oracle's parent links."""
    parent = {int(r[0]): int(r[5]) for r in o.clusters()}
    rows = []
    for r in o.bottom_clusters():  # order, start, size, original size, level, parent, ...
        depth, p = 0, int(r[5])
        while p >= 0:
            depth += 1
            p = parent[p]
        rows.append([r[0], r[4], r[4] - depth, r[5], r[3], r[2], r[1]])
    return np.array(rows, dtype=np.int32).reshape(-1, 7)


def oracle_dof2id(o, N):
    """src/tree.cpp:275-284 on the oracle: ids[perm[start + i]] = order, level-0 bottom."""
    perm = o.get_assembly_perm()
    ids = np.full(N, -1, dtype=np.int32)
    for r in o.bottom_clusters(0):
        ids[perm[r[1]:r[1] + r[3]]] = r[0]
    return ids


class OracleStub:
    """get_N / get_nlevels / get_clusters / get_dof2ID of an oracle tree."""

    def __init__(self, o, N, nlevels):
        self.o, self.N, self.nl = o, N, nlevels

    def get_N(self):
        return self.N

    def get_nlevels(self):
        return self.nl

    def get_clusters(self):
        return oracle_cluster_rows(self.o)

    def get_dof2ID(self):
        return oracle_dof2id(self.o, self.N)
