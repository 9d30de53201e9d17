// ORACLE — test infrastructure only. CPU FP64 restatement of spaND's partition, assembly,
// factorization, solve and Krylov drivers, following the reference sources cited on each
// function. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it, and only as the checker; the product library (spand_public_amd) never links it.
//
// Parity pins (see tests/test_oracle_*.py): PartitionTest.Square KAT and consistency
// invariants (tests/tests.cpp:378-481), Assembly.Consistency (tests/tests.cpp:488-547),
// ApproxTest.Exact / Approx / Repro tolerances (tests/tests.cpp:562-609, 799-856, 917-994),
// ChooseRank / swap2perm KATs, dense kernels against scipy LAPACK golden vectors, and
// brute-force dense solves on tiny grids.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <list>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "dense.hpp"

namespace oracle {

// ---------------------------------------------------------------------------------------
// Dense matrix (column-major, ld = rows), the stand-in for Eigen::MatrixXd.
struct Mat {
    int rows = 0, cols = 0;
    std::vector<double> d;
    Mat() {}
    Mat(int r, int c) : rows(r), cols(c), d((size_t)r * c, 0.0) {}
    double& operator()(int i, int j) { return d[(size_t)i + (size_t)j * rows]; }
    double operator()(int i, int j) const { return d[(size_t)i + (size_t)j * rows]; }
    double* data() { return d.data(); }
    const double* data() const { return d.data(); }
    static Mat identity(int n) {
        Mat m(n, n);
        for (int i = 0; i < n; i++) m(i, i) = 1.0;
        return m;
    }
};
using pMat = std::unique_ptr<Mat>;

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// ---------------------------------------------------------------------------------------
// Partition IDs: include/partition.h:22-78 (SepID, ClusterID, lexicographic order).
struct SepID {
    int lvl = -1, sep = 0;
    SepID() {}
    SepID(int l, int s) : lvl(l), sep(s) {}
    bool operator==(const SepID& o) const { return lvl == o.lvl && sep == o.sep; }
    bool operator<(const SepID& o) const {
        return lvl < o.lvl || (lvl == o.lvl && sep < o.sep);
    }
};
struct ClusterID {
    SepID self, l, r;
    ClusterID() {}
    ClusterID(SepID s, SepID a, SepID b) : self(s), l(a), r(b) {}
    bool operator==(const ClusterID& o) const { return self == o.self && l == o.l && r == o.r; }
    bool operator<(const ClusterID& o) const {
        return self < o.self || (self == o.self && l < o.l) ||
               (self == o.self && l == o.l && r < o.r);
    }
};
// src/partition.cpp:8-16
static SepID merge(const SepID& s) { return SepID(s.lvl + 1, s.sep / 2); }
static ClusterID merge_if(const ClusterID& c, int lvl) {
    SepID left = c.l.lvl < lvl ? merge(c.l) : c.l;
    SepID right = c.r.lvl < lvl ? merge(c.r) : c.r;
    return ClusterID(c.self, left, right);
}

// src/partition.cpp:139-210 (partition_nd_geo). X is dim x N column-major.
static void partition_nd_geo(const std::vector<int>& colptr, const std::vector<int>& rowval,
                             const std::vector<int>& dofs, std::vector<int>& parts,
                             const double* X, int dim, std::vector<int>& invp) {
    int N = (int)dofs.size();
    if (N == 0) return;
    int bestdim = -1;
    double maxrange = -1.0;
    for (int d = 0; d < dim; d++) {
        double maxi = std::numeric_limits<double>::lowest();
        double mini = std::numeric_limits<double>::max();
        for (int i = 0; i < N; i++) {
            double v = X[(size_t)d + (size_t)dim * dofs[i]];
            maxi = std::max(maxi, v);
            mini = std::min(mini, v);
        }
        double range = maxi - mini;
        if (range > maxrange) {
            bestdim = d;
            maxrange = range;
        }
    }
    auto xv = [&](int i) { return X[(size_t)bestdim + (size_t)dim * i]; };
    for (int i = 0; i < N; i++) parts[i] = dofs[i];
    // Only the N/2-th order statistic of the sorted sequence is used.
    std::sort(parts.begin(), parts.begin() + N, [&](int i, int j) { return xv(i) < xv(j); });
    double midv = xv(parts[N / 2]);
    for (int i = 0; i < N; i++) {
        int dof = dofs[i];
        invp[dof] = -1;
        for (int k = colptr[dof]; k < colptr[dof + 1]; k++) invp[rowval[k]] = -1;
    }
    for (int i = 0; i < N; i++) {
        int dof = dofs[i];
        invp[dof] = i;
        parts[i] = xv(dof) < midv ? 0 : 1;
    }
    for (int i = 0; i < N; i++) {
        int dof = dofs[i];
        if (parts[i] == 0) continue;
        for (int k = colptr[dof]; k < colptr[dof + 1]; k++) {
            int in = invp[rowval[k]];
            if (in == -1) continue;
            if (parts[in] == 0) {
                parts[i] = 2;
                break;
            }
        }
    }
}

// src/partition.cpp:384-478 (partition_modifiedND, geometric branch). colptr/rowval is the
// adjacency graph without self loops (SpMat2CSC, src/partition.cpp:21-39).
static std::vector<ClusterID> partition_modifiedND(int N, const std::vector<int>& colptr,
                                                   const std::vector<int>& rowval,
                                                   int nlevels, const double* X, int dim) {
    std::vector<int> invp(N), parttmp(N);
    std::vector<std::vector<int>> doms(1, std::vector<int>(N));
    std::iota(doms[0].begin(), doms[0].end(), 0);
    SepID top(nlevels - 1, 0);
    std::vector<ClusterID> part(N, ClusterID(top, top, top));
    for (int depth = 0; depth < nlevels - 1; depth++) {
        int level = nlevels - depth - 1;
        std::vector<std::vector<int>> newdoms((size_t)1 << (depth + 1));
        for (int sep = 0; sep < (1 << depth); sep++) {
            SepID idself(level, sep);
            std::vector<int> dofssep = doms[sep];
            std::sort(dofssep.begin(), dofssep.end());
            partition_nd_geo(colptr, rowval, dofssep, parttmp, X, dim, invp);
            SepID idleft(level - 1, 2 * sep), idright(level - 1, 2 * sep + 1);
            for (size_t i = 0; i < dofssep.size(); i++) {
                int j = dofssep[i];
                if (part[j].self == idself) {
                    if (parttmp[i] == 0)
                        part[j].self = idleft;
                    else if (parttmp[i] == 1)
                        part[j].self = idright;
                }
                if (parttmp[i] == 0) {
                    if (part[j].l == idself) part[j].l = idleft;
                    if (part[j].r == idself) part[j].r = idleft;
                } else if (parttmp[i] == 1) {
                    if (part[j].l == idself) part[j].l = idright;
                    if (part[j].r == idself) part[j].r = idright;
                } else if (parttmp[i] == 2) {
                    if (part[j].self == idself) {
                        part[j].l = idleft;
                        part[j].r = idright;
                    }
                }
            }
            std::vector<int> nl, nr;
            for (int i : dofssep) {
                const ClusterID& c = part[i];
                if (c.self == idleft || c.l == idleft || c.r == idleft) nl.push_back(i);
                if (c.self == idright || c.l == idright || c.r == idright) nr.push_back(i);
            }
            newdoms[2 * sep] = std::move(nl);
            newdoms[2 * sep + 1] = std::move(nr);
        }
        doms = std::move(newdoms);
    }
    return part;
}

// ---------------------------------------------------------------------------------------
// Cluster / Edge graph: include/cluster.h:16-117, include/edge.h:32-44, src/cluster.cpp,
// src/edge.cpp.
struct Cluster;
struct Edge {
    Cluster* n1;
    Cluster* n2;
    pMat A21;  // A[n2, n1]: n2->size rows, n1->size cols
    bool original;
    Edge(Cluster* a, Cluster* b, pMat A, bool orig) : n1(a), n2(b), A21(std::move(A)), original(orig) {}
    Mat* A() { return A21.get(); }
};
using pEdge = std::unique_ptr<Edge>;

struct Cluster {
    int start, size, level, order;
    bool eliminated = false, sparsify;
    Cluster* parent = nullptr;
    std::vector<Cluster*> children;
    std::list<pEdge> edgesOut;  // pivot first
    std::list<Edge*> edgesIn;
    std::vector<double> x;  // original_size = x.size()
    Cluster(int st, int sz, int lvl, int ord, bool sp)
        : start(st), size(sz), level(lvl), order(ord), sparsify(sp) {
        reset_size(sz);
    }
    void reset_size(int s) {
        size = s;
        x.assign(s, 0.0);
    }
    int original_size() const { return (int)x.size(); }
    Edge* pivot() { return edgesOut.begin()->get(); }
    // src/cluster.cpp:113-123
    void add_edge(pEdge e) {
        if (e->n1 == e->n2) {
            edgesOut.insert(edgesOut.begin(), std::move(e));
        } else {
            e->n2->edgesIn.push_back(e.get());
            edgesOut.push_back(std::move(e));
        }
    }
    // src/cluster.cpp:162-176
    void sort_edges() {
        edgesOut.sort([](const pEdge& a, const pEdge& b) {
            if (a->n1 == a->n2) return true;
            if (b->n1 == b->n2) return false;
            return a->n2->order < b->n2->order;
        });
        edgesIn.sort([](const Edge* a, const Edge* b) { return a->n1->order < b->n1->order; });
    }
    // src/cluster.cpp:32-45
    void set_eliminated() {
        for (auto& e : edgesOut) {
            Cluster* n = e->n2;
            n->edgesIn.remove_if([this](Edge* e2) { return e2->n1 == this; });
        }
        for (auto* e : edgesIn) {
            Cluster* n = e->n1;
            n->edgesOut.remove_if([this](const pEdge& e2) { return e2->n2 == this; });
        }
        edgesIn.clear();
        edgesOut.clear();
        eliminated = true;
    }
    template <class F>
    void for_out_nbr(F f) {
        auto it = edgesOut.begin();
        ++it;
        for (; it != edgesOut.end(); ++it) f(it->get());
    }
};
using pCluster = std::unique_ptr<Cluster>;

// ---------------------------------------------------------------------------------------
// Recorded operations: include/operations.h, src/operations.cpp. Segments are captured as
// (cluster, size at record time) and resolved to the cluster's x at solve time, which is
// what Eigen's head_x() segment views amount to (src/cluster.cpp:125-127).
enum OpType {
    OP_SCALING_LLT = 0,
    OP_SCALING_PLUQ = 1,
    OP_GEMM_SYMM_OUT = 2,
    OP_GEMM_SYMM_IN = 3,
    OP_GEMM_OUT = 4,
    OP_GEMM_IN = 5,
    OP_ORTHOGONAL = 6,
    OP_MERGE = 7,
    OP_SPLIT = 8,
};

struct Op {
    OpType type;
    Cluster* a = nullptr;  // self (scaling/gemm/orthogonal), parent (merge), original (split)
    Cluster* b = nullptr;  // nbr (gemm), sibling (split)
    int sa = 0, sb = 0;    // segment sizes at record time
    pMat M;                // L (LLT), L (PLUQ), A (gemm), v (orthogonal)
    pMat U;                // U (PLUQ)
    std::vector<double> tau;
    std::vector<int> p, q;
    int level = 0;         // factorization level that recorded the op

    long long nnz() const {
        switch (type) {
            case OP_SCALING_LLT: return (long long)sa * (sa + 1) / 2;
            case OP_SCALING_PLUQ: return (long long)sa * sa + 2LL * sa;
            case OP_GEMM_SYMM_OUT: case OP_GEMM_SYMM_IN: case OP_GEMM_OUT: case OP_GEMM_IN:
                return (long long)sa * sb;
            case OP_ORTHOGONAL: return (long long)sa * sa;
            default: return 0;
        }
    }

    void fwd() {
        double* xa = a ? a->x.data() : nullptr;
        double* xb = b ? b->x.data() : nullptr;
        switch (type) {
            case OP_SCALING_LLT:  // src/operations.cpp:97-99
                if (sa > 0) trsv_lower_notrans(sa, M->data(), sa, xa);
                break;
            case OP_SCALING_PLUQ: {  // src/operations.cpp:114-117
                std::vector<double> t(sa);
                for (int i = 0; i < sa; i++) t[i] = xa[p[i]];
                std::copy(t.begin(), t.end(), xa);
                if (sa > 0) trsv_lower_notrans(sa, M->data(), sa, xa);
                break;
            }
            case OP_GEMM_OUT:       // src/operations.cpp:31-39
            case OP_GEMM_SYMM_OUT:  // src/operations.cpp:55-63
                if (sb > 0 && sa > 0) gemv_sub_notrans(sb, sa, M->data(), sb, xa, xb);
                break;
            case OP_GEMM_SYMM_IN:  // src/operations.cpp:75-83
                if (sb > 0 && sa > 0) gemv_sub_trans(sa, sb, M->data(), sa, xa, xb);
                break;
            case OP_GEMM_IN: break;
            case OP_ORTHOGONAL:  // src/operations.cpp:153-155
                ormqr_vec(true, sa, M->cols, M->data(), sa, tau.data(), xa);
                break;
            case OP_MERGE: {  // src/operations.cpp:167-177
                int k = 0;
                for (Cluster* c : a->children)
                    for (int i = 0; i < c->size; i++) a->x[k++] = c->x[i];
                break;
            }
            case OP_SPLIT:  // src/operations.cpp:194-197
                for (int i = 0; i < b->size; i++) b->x[i] = a->x[a->size + i];
                break;
        }
    }

    void bwd() {
        double* xa = a ? a->x.data() : nullptr;
        double* xb = b ? b->x.data() : nullptr;
        switch (type) {
            case OP_SCALING_LLT:  // src/operations.cpp:101-103
                if (sa > 0) trsv_lower_trans(sa, M->data(), sa, xa);
                break;
            case OP_SCALING_PLUQ: {  // src/operations.cpp:118-121
                if (sa > 0) trsv_upper_notrans(sa, U->data(), sa, xa);
                std::vector<double> t(sa);
                for (int i = 0; i < sa; i++) t[i] = xa[q[i]];
                std::copy(t.begin(), t.end(), xa);
                break;
            }
            case OP_GEMM_OUT: break;
            case OP_GEMM_IN:  // src/operations.cpp:45-53
                if (sb > 0 && sa > 0) gemv_sub_notrans(sa, sb, M->data(), sa, xb, xa);
                break;
            case OP_GEMM_SYMM_OUT:  // src/operations.cpp:65-73
                if (sb > 0 && sa > 0) gemv_sub_trans(sb, sa, M->data(), sb, xb, xa);
                break;
            case OP_GEMM_SYMM_IN:  // src/operations.cpp:85-93
                if (sb > 0 && sa > 0) gemv_sub_notrans(sa, sb, M->data(), sa, xb, xa);
                break;
            case OP_ORTHOGONAL:  // src/operations.cpp:156-158
                ormqr_vec(false, sa, M->cols, M->data(), sa, tau.data(), xa);
                break;
            case OP_MERGE: {  // src/operations.cpp:178-187
                int k = 0;
                for (Cluster* c : a->children)
                    for (int i = 0; i < c->size; i++) c->x[i] = a->x[k++];
                break;
            }
            case OP_SPLIT:  // src/operations.cpp:198-201
                for (int i = 0; i < b->size; i++) a->x[a->size + i] = b->x[i];
                break;
        }
    }
};
using pOp = std::unique_ptr<Op>;

// Shape log (include/util.h:311-342 ProfileFlops, plus kind codes).
struct Shape {
    int level, kind;  // 0 potf, 1 trsm, 2 gemm, 3 syrk, 4 geqp3, 5 getf
    int m, n, k;
};

struct LevelLog {
    int dofs_nd = 0, dofs_left_nd = 0, dofs_left_elim = 0, dofs_left_spars = 0;
    long long fact_nnz = 0;
    int rank_before_sum = 0, rank_after_sum = 0, sparsified = 0, ignored = 0;
    double t_elim = 0, t_scale = 0, t_spars = 0, t_merge = 0;
};

// ---------------------------------------------------------------------------------------
// Tree: include/tree.h, src/tree.cpp.
struct Tree {
    int nlevels;
    double tol = 10.0;  // src/tree.cpp:94
    int skip = 0;
    int symm = 0;       // 0 = SPD/LLT, 2 = GEN/PLU
    bool verb = false;
    bool geo = false;
    int ilvl = 0;
    int max_order = 0;
    int current_bottom = 0;
    std::vector<double> X;
    int dim = 0;
    std::vector<int> perm;
    std::vector<std::list<pCluster>> bottoms;
    std::vector<pCluster> others;
    std::list<pOp> ops;
    std::vector<LevelLog> log;
    std::vector<Shape> shapes;
    bool record_shapes = false;
    std::string last_error;
    std::vector<int> part_out;  // 6 ints per dof: self(lvl,sep), l(lvl,sep), r(lvl,sep)

    explicit Tree(int nl) : nlevels(nl), bottoms(nl), log(nl) {}
    bool symmetry() const { return symm == 0; }
    int get_new_order() { return max_order++; }
    std::list<pCluster>& bottom_current() { return bottoms[current_bottom]; }
    std::list<pCluster>& bottom_original() { return bottoms[0]; }
    int get_N() const { return (int)perm.size(); }
    void shape(int kind, int m, int n, int k) {
        if (record_shapes) shapes.push_back({ilvl, kind, m, n, k});
    }

    int ndofs_left() {
        int n = 0;
        for (auto& c : bottom_current())
            if (!c->eliminated) n += c->size;
        return n;
    }

    // src/tree.cpp:306-419
    void partition(int N, const std::vector<int>& colptr, const std::vector<int>& rowval) {
        if (!geo || (int)X.size() != dim * N) throw std::runtime_error("geometric partition needs coordinates");
        // adjacency without self loops (src/partition.cpp:21-39)
        std::vector<int> cp(N + 1, 0), rv;
        rv.reserve(colptr[N]);
        for (int i = 0; i < N; i++) {
            cp[i + 1] = cp[i];
            for (int k = colptr[i]; k < colptr[i + 1]; k++) {
                if (rowval[k] != i) {
                    rv.push_back(rowval[k]);
                    cp[i + 1]++;
                }
            }
        }
        std::vector<ClusterID> part = partition_modifiedND(N, cp, rv, nlevels, X.data(), dim);
        part_out.resize((size_t)6 * N);
        for (int i = 0; i < N; i++) {
            const ClusterID& c = part[i];
            int* o = &part_out[(size_t)6 * i];
            o[0] = c.self.lvl; o[1] = c.self.sep; o[2] = c.l.lvl; o[3] = c.l.sep; o[4] = c.r.lvl; o[5] = c.r.sep;
            log[c.self.lvl].dofs_nd++;
        }
        log[nlevels - 1].dofs_left_nd = 0;
        for (int l = nlevels - 2; l >= 0; l--) log[l].dofs_left_nd = log[l + 1].dofs_nd + log[l + 1].dofs_left_nd;
        perm.resize(N);
        std::iota(perm.begin(), perm.end(), 0);
        std::vector<ClusterID> pm = part;
        auto cmp = [&pm](int i, int j) { return pm[i] < pm[j]; };
        std::stable_sort(perm.begin(), perm.end(), cmp);
        for (int lvl = 1; lvl < nlevels; lvl++) {
            for (auto& c : pm) c = merge_if(c, lvl);
            std::stable_sort(perm.begin(), perm.end(), cmp);
        }
        std::vector<ClusterID> pp(N);
        for (int k = 0; k < N; k++) pp[k] = part[perm[k]];
        std::map<Cluster*, ClusterID> ids;
        for (int k = 0; k < N;) {
            int kn = k + 1;
            ClusterID id = pp[k];
            while (kn < N && pp[kn] == id) kn++;
            auto c = std::make_unique<Cluster>(k, kn - k, id.self.lvl, get_new_order(),
                                               id.l.lvl == 0 && id.r.lvl == 0);
            ids[c.get()] = id;
            bottoms[0].push_back(std::move(c));
            k = kn;
        }
        for (int lvl = 1; lvl < nlevels; lvl++) {
            auto begin = std::find_if(bottoms[lvl - 1].begin(), bottoms[lvl - 1].end(),
                                      [lvl](const pCluster& s) { return s->level >= lvl; });
            auto end = bottoms[lvl - 1].end();
            std::map<Cluster*, ClusterID> pids;
            for (auto s = begin; s != end; ++s) pids[s->get()] = merge_if(ids.at(s->get()), lvl);
            for (auto k = begin; k != end;) {
                ClusterID idp = pids.at(k->get());
                std::vector<Cluster*> ch{k->get()};
                int cstart = (*k)->start, csize = (*k)->size;
                ++k;
                while (k != end && idp == pids.at(k->get())) {
                    ch.push_back(k->get());
                    csize += (*k)->size;
                    ++k;
                }
                auto p = std::make_unique<Cluster>(cstart, csize, idp.self.lvl, get_new_order(),
                                                   idp.l.lvl == lvl && idp.r.lvl == lvl);
                for (Cluster* c : ch) {
                    c->parent = p.get();
                    p->children.push_back(c);
                }
                ids[p.get()] = idp;
                bottoms[lvl].push_back(std::move(p));
            }
        }
    }

    // src/tree.cpp:505-575 (with symm_perm src/util.cpp:520-546, block2dense :454-486)
    void assemble(int N, const int* colptr, const int* rowval, const double* vals) {
        std::vector<int> pinv(N);
        for (int i = 0; i < N; i++) pinv[perm[i]] = i;
        // CSC of P^T A P with sorted rows
        std::vector<int> cnt(N + 1, 0);
        int nnz = colptr[N];
        for (int j = 0; j < N; j++)
            for (int k = colptr[j]; k < colptr[j + 1]; k++) cnt[pinv[j] + 1]++;
        for (int j = 0; j < N; j++) cnt[j + 1] += cnt[j];
        std::vector<int> prow(nnz);
        std::vector<double> pval(nnz);
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int j = 0; j < N; j++)
            for (int k = colptr[j]; k < colptr[j + 1]; k++) {
                int pj = pinv[j];
                prow[fill[pj]] = pinv[rowval[k]];
                pval[fill[pj]] = vals[k];
                fill[pj]++;
            }
        for (int j = 0; j < N; j++) {
            std::vector<std::pair<int, double>> col;
            for (int k = cnt[j]; k < cnt[j + 1]; k++) col.push_back({prow[k], pval[k]});
            std::sort(col.begin(), col.end(), [](auto& a, auto& b) { return a.first < b.first; });
            for (int k = cnt[j]; k < cnt[j + 1]; k++) {
                prow[k] = col[k - cnt[j]].first;
                pval[k] = col[k - cnt[j]].second;
            }
        }
        std::vector<Cluster*> cmap(N);
        for (auto& s : bottom_original())
            for (int k = s->start; k < s->start + s->size; k++) cmap[k] = s.get();
        for (auto& s : bottom_original()) {
            std::set<Cluster*, bool (*)(Cluster*, Cluster*)> nbrs(
                [](Cluster* a, Cluster* b) { return a->order < b->order; });
            for (int j = s->start; j < s->start + s->size; j++)
                for (int k = cnt[j]; k < cnt[j + 1]; k++) {
                    if (symmetry() && prow[k] < j) continue;
                    nbrs.insert(cmap[prow[k]]);
                }
            nbrs.insert(s.get());
            for (Cluster* n : nbrs) {
                auto A = std::make_unique<Mat>(n->size, s->size);
                for (int c = 0; c < s->size; c++) {
                    int j = s->start + c;
                    for (int k = cnt[j]; k < cnt[j + 1]; k++) {
                        int r = prow[k];
                        if (r >= n->start && r < n->start + n->size) (*A)(r - n->start, c) = pval[k];
                    }
                }
                s->add_edge(std::make_unique<Edge>(s.get(), n, std::move(A), true));
            }
            s->sort_edges();
        }
    }

    // ------------------------------ factorization pieces ------------------------------
    // src/tree.cpp:582-593
    void potf_cluster(Cluster* s) {
        Mat* A = s->pivot()->A();
        shape(0, A->rows, 0, 0);
        int info = potrf_lower(A->rows, A->data(), std::max(1, A->rows));
        if (info != 0) throw std::runtime_error("Error: Non-SPD Pivot (cluster order " + std::to_string(s->order) + ")");
    }
    // src/tree.cpp:614-635 (PLU branch) + split_LU src/util.cpp:213-227
    void getf_cluster(Cluster* s, pMat* L, pMat* U, std::vector<int>* p, std::vector<int>* q) {
        Mat* A = s->pivot()->A();
        int n = A->rows;
        shape(5, n, 0, 0);
        std::vector<int> swap(n);
        int info = getrf(n, A->data(), std::max(1, n), swap.data());
        if (info != 0) throw std::runtime_error("Error: Singular Pivot (cluster order " + std::to_string(s->order) + ")");
        p->resize(n);
        swap2perm(n, swap.data(), p->data());
        q->resize(n);
        std::iota(q->begin(), q->end(), 0);
        *L = std::make_unique<Mat>(n, n);
        *U = std::make_unique<Mat>(n, n);
        for (int j = 0; j < n; j++) {
            double dj = (*A)(j, j);
            double sq = std::sqrt(std::fabs(dj));
            (**L)(j, j) = sq;  // 1 * |d_j|^(1/2)
            for (int i = j + 1; i < n; i++) (**L)(i, j) = (*A)(i, j) * sq;
        }
        for (int i = 0; i < n; i++) {
            double di = (*A)(i, i);
            double sgn = (di > 0) - (di < 0);
            double f = sgn * std::sqrt(std::fabs(di));
            double inv = 1.0 / di;
            (**U)(i, i) = f * 1.0;
            for (int j = i + 1; j < n; j++) (**U)(i, j) = f * (inv * (*A)(i, j));
        }
    }
    // src/tree.cpp:640-663
    void trsm_potf_edgeIn(Edge* e) {
        Mat* Asn = e->A();
        Mat* L = e->n2->pivot()->A();
        shape(1, Asn->cols, Asn->rows, 0);
        trsm_left_lower_notrans(Asn->rows, Asn->cols, L->data(), std::max(1, L->rows), Asn->data(), std::max(1, Asn->rows));
    }
    void trsm_potf_edgeOut(Edge* e) {
        Mat* Ans = e->A();
        Mat* L = e->n1->pivot()->A();
        shape(1, Ans->rows, Ans->cols, 0);
        trsm_right_lower_trans(Ans->rows, Ans->cols, L->data(), std::max(1, L->rows), Ans->data(), std::max(1, Ans->rows));
    }
    // src/tree.cpp:668-689
    void trsm_getf_edgeIn(Edge* e, Mat* L, const std::vector<int>& p) {
        Mat* Asn = e->A();
        Mat t(Asn->rows, Asn->cols);
        for (int j = 0; j < Asn->cols; j++)
            for (int i = 0; i < Asn->rows; i++) t(i, j) = (*Asn)(p[i], j);
        *Asn = t;
        shape(1, Asn->cols, Asn->rows, 0);
        trsm_left_lower_notrans(Asn->rows, Asn->cols, L->data(), std::max(1, L->rows), Asn->data(), std::max(1, Asn->rows));
    }
    void trsm_getf_edgeOut(Edge* e, Mat* U, const std::vector<int>& q) {
        Mat* Ans = e->A();
        Mat t(Ans->rows, Ans->cols);
        for (int j = 0; j < Ans->cols; j++)
            for (int i = 0; i < Ans->rows; i++) t(i, j) = (*Ans)(i, q[j]);
        *Ans = t;
        shape(1, Ans->rows, Ans->cols, 0);
        trsm_right_upper_notrans(Ans->rows, Ans->cols, U->data(), std::max(1, U->rows), Ans->data(), std::max(1, Ans->rows));
    }
    // src/tree.cpp:717-724, 735-742
    void panel_potf(Cluster* s) {
        for (Edge* e : s->edgesIn) trsm_potf_edgeIn(e);
        s->for_out_nbr([&](Edge* e) { trsm_potf_edgeOut(e); });
    }
    void panel_getf(Cluster* s, Mat* L, Mat* U, const std::vector<int>& p, const std::vector<int>& q) {
        for (Edge* e : s->edgesIn) trsm_getf_edgeIn(e, L, p);
        s->for_out_nbr([&](Edge* e) { trsm_getf_edgeOut(e, U, q); });
    }
    // src/tree.cpp:748-793
    void gemm_edges(Edge* e1, Edge* e2, bool t1, bool t2) {
        Cluster* s = t1 ? e1->n2 : e1->n1;
        Cluster* n1 = t1 ? e1->n1 : e1->n2;
        Mat* An1s = e1->A();
        Cluster* n2 = t2 ? e2->n2 : e2->n1;
        Mat* Asn2 = e2->A();
        Edge* found = nullptr;
        for (auto& e : n2->edgesOut)
            if (e->n2 == n1) { found = e.get(); break; }
        if (!found) {
            auto A = std::make_unique<Mat>(n1->size, n2->size);
            n2->add_edge(std::make_unique<Edge>(n2, n1, std::move(A), false));
            for (auto& e : n2->edgesOut)
                if (e->n2 == n1) { found = e.get(); break; }
        }
        Mat* C = found->A();
        int k = t1 ? An1s->rows : An1s->cols;
        const bool syrk = n1 == n2 && symmetry();
        shape(syrk ? 3 : 2, C->rows, C->cols, k);
        if (syrk)
            schur_update(t1, !t1, true, C->rows, C->cols, k, An1s->data(), std::max(1, An1s->rows), An1s->data(),
                         std::max(1, An1s->rows), C->data(), std::max(1, C->rows));
        else
            schur_update(t1, t2, false, C->rows, C->cols, k, An1s->data(), std::max(1, An1s->rows), Asn2->data(),
                         std::max(1, Asn2->rows), C->data(), std::max(1, C->rows));
        (void)s;
    }
    // src/tree.cpp:862-883
    void schur_symmetric(Cluster* self) {
        std::vector<Edge*> out, in(self->edgesIn.begin(), self->edgesIn.end());
        self->for_out_nbr([&](Edge* e) { out.push_back(e); });
        for (Edge* e1 : out)
            for (Edge* e2 : out)
                if (e1->n2->order >= e2->n2->order) gemm_edges(e1, e2, false, true);
        for (Edge* e1 : in)
            for (Edge* e2 : in)
                if (e1->n1->order >= e2->n1->order) gemm_edges(e1, e2, true, false);
        for (Edge* e1 : out)
            for (Edge* e2 : in) gemm_edges(e1, e2, false, false);
    }
    void push_op(OpType t, Cluster* a, Cluster* b, pMat M) {
        auto op = std::make_unique<Op>();
        op->type = t;
        op->a = a;
        op->b = b;
        op->sa = a ? a->size : 0;
        op->sb = b ? b->size : 0;
        op->M = std::move(M);
        op->level = ilvl;
        ops.push_back(std::move(op));
    }
    // src/tree.cpp:895-967
    void eliminate_cluster(Cluster* self) {
        if (symm == 0) {
            potf_cluster(self);
            panel_potf(self);
            schur_symmetric(self);
            push_op(OP_SCALING_LLT, self, nullptr, std::move(self->pivot()->A21));
            self->for_out_nbr([&](Edge* e) { push_op(OP_GEMM_SYMM_OUT, self, e->n2, std::move(e->A21)); });
            for (Edge* e : self->edgesIn) push_op(OP_GEMM_SYMM_IN, self, e->n1, std::move(e->A21));
        } else {
            pMat L, U;
            std::vector<int> p, q;
            getf_cluster(self, &L, &U, &p, &q);
            panel_getf(self, L.get(), U.get(), p, q);
            std::vector<Edge*> out, in(self->edgesIn.begin(), self->edgesIn.end());
            self->for_out_nbr([&](Edge* e) { out.push_back(e); });
            for (Edge* e1 : out)
                for (Edge* e2 : in) gemm_edges(e1, e2, false, false);
            push_op(OP_SCALING_PLUQ, self, nullptr, std::move(L));
            ops.back()->U = std::move(U);
            ops.back()->p = p;
            ops.back()->q = q;
            for (Edge* e : out) push_op(OP_GEMM_OUT, self, e->n2, std::move(e->A21));
            for (Edge* e : in) push_op(OP_GEMM_IN, self, e->n1, std::move(e->A21));
        }
        self->set_eliminated();
    }
    // src/tree.cpp:796-856
    void scale_cluster(Cluster* self) {
        if (symm == 0) {
            potf_cluster(self);
            panel_potf(self);
            push_op(OP_SCALING_LLT, self, nullptr, std::move(self->pivot()->A21));
        } else {
            pMat L, U;
            std::vector<int> p, q;
            getf_cluster(self, &L, &U, &p, &q);
            panel_getf(self, L.get(), U.get(), p, q);
            push_op(OP_SCALING_PLUQ, self, nullptr, std::move(L));
            ops.back()->U = std::move(U);
            ops.back()->p = p;
            ops.back()->q = q;
        }
        self->pivot()->A21 = std::make_unique<Mat>(Mat::identity(self->size));
    }
    // src/tree.cpp:1189-1224 (pred = all-true)
    pMat assemble_Asn(Cluster* self) {
        int rows = self->size, cols = 0;
        for (Edge* e : self->edgesIn) cols += e->n1->size;
        self->for_out_nbr([&](Edge* e) { cols += e->n2->size; });
        auto Asn = std::make_unique<Mat>(rows, cols);
        int c = 0;
        for (Edge* e : self->edgesIn) {
            Mat* A = e->A();
            for (int j = 0; j < A->cols; j++)
                for (int i = 0; i < rows; i++) (*Asn)(i, c + j) = (*A)(i, j);
            c += A->cols;
        }
        self->for_out_nbr([&](Edge* e) {
            Mat* A = e->A();  // n2 x s
            for (int j = 0; j < A->rows; j++)
                for (int i = 0; i < rows; i++) (*Asn)(i, c + j) = (*A)(j, i);
            c += A->rows;
        });
        return Asn;
    }
    // src/tree.cpp:980-1104 (pred all-true, is_pivot_I, keep_sibling)
    Cluster* shrink_split_scatter(Cluster* orig, int rank, const Mat& AsnP) {
        auto psib = std::make_unique<Cluster>(orig->start + rank, orig->size - rank, orig->level,
                                              get_new_order(), orig->sparsify);
        Cluster* sib = psib.get();
        orig->size = rank;
        int col = 0;
        for (Edge* e : orig->edgesIn) {
            int cols = e->n1->size;
            auto A = std::make_unique<Mat>(rank, cols);
            for (int j = 0; j < cols; j++)
                for (int i = 0; i < rank; i++) (*A)(i, j) = AsnP(i, col + j);
            e->A21 = std::move(A);
            col += cols;
        }
        orig->for_out_nbr([&](Edge* e) {
            int cols = e->n2->size;
            auto A = std::make_unique<Mat>(cols, rank);
            for (int j = 0; j < cols; j++)
                for (int i = 0; i < rank; i++) (*A)(j, i) = AsnP(i, col + j);
            e->A21 = std::move(A);
            col += cols;
        });
        sib->add_edge(std::make_unique<Edge>(sib, sib, std::make_unique<Mat>(Mat::identity(sib->size)), true));
        orig->pivot()->A21 = std::make_unique<Mat>(Mat::identity(rank));
        others.push_back(std::move(psib));
        push_op(OP_SPLIT, orig, sib, nullptr);
        return sib;
    }
    // src/tree.cpp:1292-1347
    void sparsify_adaptive_only(Cluster* self) {
        pMat Asn = assemble_Asn(self);
        int rows = Asn->rows, cols = Asn->cols;
        int mn = std::min(rows, cols);
        std::vector<int> jpvt(cols);
        std::vector<double> ht(mn);
        shape(4, rows, cols, 0);  // recorded for empty Asn too, as src/tree.cpp:1312
        if (rows > 0 && cols > 0) geqp3(rows, cols, Asn->data(), rows, jpvt.data(), ht.data());
        std::vector<double> diag(mn);
        for (int i = 0; i < mn; i++) diag[i] = (*Asn)(i, i);
        int rank = choose_rank(diag.data(), mn, tol);
        if (rank >= rows) return;
        auto v = std::make_unique<Mat>(rows, rank);
        for (int j = 0; j < rank; j++)
            for (int i = 0; i < rows; i++) (*v)(i, j) = (*Asn)(i, j);
        push_op(OP_ORTHOGONAL, self, nullptr, std::move(v));
        ops.back()->tau.assign(ht.begin(), ht.begin() + rank);
        Mat AsnP(rank, cols);
        for (int j = 0; j < cols; j++)
            for (int i = 0; i <= std::min(j, rank - 1); i++) AsnP(i, jpvt[j]) = (*Asn)(i, j);
        Cluster* sib = shrink_split_scatter(self, rank, AsnP);
        eliminate_cluster(sib);
    }
    // src/tree.cpp:1417-1433
    void sparsify_cluster(Cluster* self) {
        if (self->sparsify) {
            log[ilvl].rank_before_sum += self->size;
            sparsify_adaptive_only(self);
            log[ilvl].rank_after_sum += self->size;
            log[ilvl].sparsified++;
        } else {
            log[ilvl].ignored++;
        }
    }
    // src/tree.cpp:1106-1184, 1435-1445
    void merge_all() {
        current_bottom++;
        std::map<Cluster*, int> posparent;
        for (auto& snew : bottom_current()) {
            int size = 0;
            for (Cluster* c : snew->children) {
                posparent[c] = size;
                size += c->size;
            }
            snew->reset_size(size);
            push_op(OP_MERGE, snew.get(), nullptr, nullptr);
        }
        for (auto& snewp : bottom_current()) {
            Cluster* snew = snewp.get();
            std::map<Cluster*, bool> merged;
            for (Cluster* sold : snew->children)
                for (auto& eold : sold->edgesOut) {
                    Cluster* nnew = eold->n2->parent;
                    if (!merged.count(nnew)) merged[nnew] = false;
                    if (!merged[nnew]) merged[nnew] = eold->original;
                }
            for (auto& kv : merged)
                snew->add_edge(std::make_unique<Edge>(snew, kv.first, std::make_unique<Mat>(kv.first->size, snew->size), kv.second));
            snew->sort_edges();
            for (Cluster* sold : snew->children) {
                for (auto& eold : sold->edgesOut) {
                    Cluster* nold = eold->n2;
                    Cluster* nnew = nold->parent;
                    Edge* found = nullptr;
                    for (auto& e : snew->edgesOut)
                        if (e->n2 == nnew) { found = e.get(); break; }
                    Mat* D = found->A();
                    Mat* S = eold->A();
                    int r0 = posparent[nold], c0 = posparent[sold];
                    for (int j = 0; j < S->cols; j++)
                        for (int i = 0; i < S->rows; i++) (*D)(r0 + i, c0 + j) = (*S)(i, j);
                }
                sold->edgesOut.clear();
                sold->edgesIn.clear();
            }
        }
    }
    // src/tree.cpp:1447-1551
    void factorize() {
        for (ilvl = 0; ilvl < nlevels; ilvl++) {
            double t0 = now();
            for (auto& s : bottom_current())
                if (s->level == ilvl) eliminate_cluster(s.get());
            log[ilvl].dofs_left_elim = ndofs_left();
            double t1 = now();
            log[ilvl].t_elim = t1 - t0;
            if (ilvl >= skip) {
                for (auto& s : bottom_current())
                    if (s->level > ilvl) scale_cluster(s.get());
                double t2 = now();
                log[ilvl].t_scale = t2 - t1;
                for (auto& s : bottom_current())
                    if (s->level > ilvl) sparsify_cluster(s.get());
                log[ilvl].t_spars = now() - t2;
            }
            if (ilvl < nlevels - 1) {
                double t3 = now();
                merge_all();
                log[ilvl].t_merge = now() - t3;
            }
            log[ilvl].dofs_left_spars = ndofs_left();
            log[ilvl].fact_nnz = nnz();
        }
    }
    long long nnz() const {
        long long n = 0;
        for (auto& op : ops) n += op->nnz();
        return n;
    }
    // src/tree.cpp:158-169
    int get_stop() const {
        int stop = (int)perm.size();
        for (auto& l : log) {
            if (l.dofs_left_elim > 0) stop = std::min(stop, l.dofs_left_elim);
            if (l.dofs_left_spars > 0) stop = std::min(stop, l.dofs_left_spars);
        }
        return stop;
    }
    // src/tree.cpp:1610-1635
    void solve(double* x) {
        int N = get_N();
        std::vector<double> b(N);
        for (int i = 0; i < N; i++) b[i] = x[perm[i]];
        for (auto& c : bottom_original()) {
            if (c->children.empty())
                for (int i = 0; i < c->original_size(); i++) c->x[i] = b[c->start + i];
            else
                std::fill(c->x.begin(), c->x.end(), 0.0);
        }
        for (auto& op : ops) op->fwd();
        for (auto it = ops.rbegin(); it != ops.rend(); ++it) (*it)->bwd();
        for (auto& c : bottom_original())
            if (c->children.empty())
                for (int i = 0; i < c->original_size(); i++) b[c->start + i] = c->x[i];
        for (int i = 0; i < N; i++) x[perm[i]] = b[i];
    }
};

// ---------------------------------------------------------------------------------------
// CSC matrix-vector product y = A x.
static void spmv(int N, const int* cp, const int* ri, const double* v, const double* x, double* y) {
    std::fill(y, y + N, 0.0);
    for (int j = 0; j < N; j++) {
        double xj = x[j];
        for (int k = cp[j]; k < cp[j + 1]; k++) y[ri[k]] += v[k] * xj;
    }
}
static double dot(int n, const double* a, const double* b) {
    double s = 0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

// src/is.cpp:39-121 (PCG; stops on the recursive residual, returns i+1).
static int cg(Tree& t, int N, const int* cp, const int* ri, const double* v, const double* rhs, double* x,
              int iters, double tol, double* t_prec) {
    std::vector<double> r(N), p(N), z(N), tmp(N);
    spmv(N, cp, ri, v, x, tmp.data());
    for (int i = 0; i < N; i++) r[i] = rhs[i] - tmp[i];
    double rhsNorm2 = dot(N, rhs, rhs);
    if (rhsNorm2 == 0) {
        std::fill(x, x + N, 0.0);
        return 0;
    }
    double threshold = tol * tol * rhsNorm2;
    double rn2 = dot(N, r.data(), r.data());
    if (rn2 < threshold) return 0;
    p = r;
    double tp = now();
    t.solve(p.data());
    *t_prec += now() - tp;
    double absNew = dot(N, r.data(), p.data());
    int i = 0;
    while (i < iters) {
        spmv(N, cp, ri, v, p.data(), tmp.data());
        double alpha = absNew / dot(N, p.data(), tmp.data());
        for (int k = 0; k < N; k++) x[k] += alpha * p[k];
        for (int k = 0; k < N; k++) r[k] -= alpha * tmp[k];
        rn2 = dot(N, r.data(), r.data());
        if (t.verb) printf("%d: |Ax-b|/|b| = %3.2e <? %3.2e\n", i, std::sqrt(rn2 / rhsNorm2), tol);
        if (rn2 < threshold) break;
        z = r;
        tp = now();
        t.solve(z.data());
        *t_prec += now() - tp;
        double absOld = absNew;
        absNew = dot(N, r.data(), z.data());
        double beta = absNew / absOld;
        for (int k = 0; k < N; k++) p[k] = z[k] + beta * p[k];
        i++;
    }
    return i + 1;
}

// Eigen's HouseholderSequence primitives for real vectors, as gmres uses them:
// makeHouseholder (Eigen/src/Householder/Householder.h) on x[0..n), and
// applyHouseholderOnTheLeft of (essential, tau) on y[0..n).
static void make_householder(int n, const double* x, double* essential, double& tau, double& beta) {
    const double tail_sq = n == 1 ? 0.0 : dot(n - 1, x + 1, x + 1);
    const double c0 = x[0];
    if (tail_sq <= DBL_MIN) {
        tau = 0.0;
        beta = c0;
        for (int i = 0; i + 1 < n; i++) essential[i] = 0.0;
    } else {
        beta = std::sqrt(c0 * c0 + tail_sq);
        if (c0 >= 0.0) beta = -beta;
        for (int i = 0; i + 1 < n; i++) essential[i] = x[i + 1] / (c0 - beta);
        tau = (beta - c0) / beta;
    }
}
static void apply_householder(int n, double* y, const double* essential, double tau) {
    if (n == 1) {
        y[0] *= 1.0 - tau;
    } else if (tau != 0.0) {
        const double tmp = dot(n - 1, essential, y + 1) + y[0];
        y[0] -= tau * tmp;
        for (int i = 0; i + 1 < n; i++) y[i + 1] -= tau * essential[i] * tmp;
    }
}
// JacobiRotation::makeGivens (real) and applyOnTheLeft(i, i + 1, G.adjoint())
struct Givens {
    double c = 1.0, s = 0.0;
};
static Givens make_givens(double p, double q) {
    Givens g;
    if (q == 0.0) {
        g.c = p < 0.0 ? -1.0 : 1.0;
        g.s = 0.0;
    } else if (p == 0.0) {
        g.c = 0.0;
        g.s = q < 0.0 ? 1.0 : -1.0;
    } else if (std::fabs(p) > std::fabs(q)) {
        const double t = q / p;
        double u = std::sqrt(1.0 + t * t);
        if (p < 0.0) u = -u;
        g.c = 1.0 / u;
        g.s = -t * g.c;
    } else {
        const double t = p / q;
        double u = std::sqrt(1.0 + t * t);
        if (q < 0.0) u = -u;
        g.s = -1.0 / u;
        g.c = -t * g.s;
    }
    return g;
}
static void apply_givens_adjoint(const Givens& g, double& x, double& y) {
    const double x0 = x, y0 = y;
    x = g.c * x0 - g.s * y0;
    y = g.s * x0 + g.c * y0;
}

// gmres (src/is.cpp:123-300): left-preconditioned GMRES(restart) with Householder Arnoldi
// and Givens rotations; stops when |w_k| / ||M^{-1}(b - A x0)|| < tol (the initial
// preconditioned residual norm, kept across restarts). Returns the iteration count;
// *rel_out receives the last |w_k| / ||r0||.
static int gmres(Tree& t, int N, const int* cp, const int* ri, const double* v_, const double* rhs, double* x,
                 int maxIters, int restart, double tol, double* rel_out, double* t_prec) {
    const int m = N;
    *rel_out = 0.0;
    if (std::sqrt(dot(N, rhs, rhs)) <= DBL_MIN) {
        std::fill(x, x + N, 0.0);
        return 0;
    }
    int iters = 0;
    std::vector<double> p0(N), r0(N), tmp(N);
    spmv(N, cp, ri, v_, x, tmp.data());
    for (int i = 0; i < N; i++) p0[i] = rhs[i] - tmp[i];
    r0 = p0;
    double tp = now();
    t.solve(r0.data());
    *t_prec += now() - tp;
    const double r0Norm = std::sqrt(dot(N, r0.data(), r0.data()));
    if (r0Norm == 0.0) return 0;
    std::vector<double> H((size_t)m * (restart + 1), 0.0), w(restart + 1, 0.0), tau(restart + 1, 0.0);
    std::vector<Givens> G(restart);
    std::vector<double> v(m), x_new(m);
    auto Hc = [&](int c) { return H.data() + (size_t)c * m; };
    double beta;
    make_householder(m, r0.data(), Hc(0) + 1, tau[0], beta);
    w[0] = beta;
    for (int k = 1; k <= restart; ++k) {
        ++iters;
        std::fill(v.begin(), v.end(), 0.0);
        v[k - 1] = 1.0;
        for (int i = k - 1; i >= 0; --i) apply_householder(m - i, v.data() + i, Hc(i) + i + 1, tau[i]);
        spmv(N, cp, ri, v_, v.data(), tmp.data());
        v = tmp;
        tp = now();
        t.solve(v.data());
        *t_prec += now() - tp;
        for (int i = 0; i < k; ++i) apply_householder(m - i, v.data() + i, Hc(i) + i + 1, tau[i]);
        if (k < m && std::sqrt(dot(m - k, v.data() + k, v.data() + k)) != 0.0) {
            make_householder(m - k, v.data() + k, Hc(k) + k + 1, tau[k], beta);
            apply_householder(m - k, v.data() + k, Hc(k) + k + 1, tau[k]);
        }
        for (int i = 0; i + 1 < k; ++i) apply_givens_adjoint(G[i], v[i], v[i + 1]);
        if (k < m && v[k] != 0.0) {
            G[k - 1] = make_givens(v[k - 1], v[k]);
            apply_givens_adjoint(G[k - 1], v[k - 1], v[k]);
            apply_givens_adjoint(G[k - 1], w[k - 1], w[k]);
        }
        for (int i = 0; i < k; i++) Hc(k - 1)[i] = v[i];
        const double tol_error = std::fabs(w[k]) / r0Norm;
        *rel_out = tol_error;
        const bool stop = (k == m || tol_error < tol || iters == maxIters);
        if (t.verb) printf("%d: |Ax-b|/|b| = %3.2e <? %3.2e\n", iters, tol_error, tol);
        if (stop || k == restart) {
            std::vector<double> y(w.begin(), w.begin() + k);
            for (int i = k - 1; i >= 0; --i) {  // upper-triangular solve
                double s = y[i];
                for (int j = i + 1; j < k; j++) s -= Hc(j)[i] * y[j];
                y[i] = s / Hc(i)[i];
            }
            std::fill(x_new.begin(), x_new.end(), 0.0);
            for (int i = k - 1; i >= 0; --i) {
                x_new[i] += y[i];
                apply_householder(m - i, x_new.data() + i, Hc(i) + i + 1, tau[i]);
            }
            for (int i = 0; i < N; i++) x[i] += x_new[i];
            if (stop) return iters;
            k = 0;
            spmv(N, cp, ri, v_, x, tmp.data());
            for (int i = 0; i < N; i++) p0[i] = rhs[i] - tmp[i];
            r0 = p0;
            tp = now();
            t.solve(r0.data());
            *t_prec += now() - tp;
            std::fill(H.begin(), H.end(), 0.0);
            std::fill(w.begin(), w.end(), 0.0);
            std::fill(tau.begin(), tau.end(), 0.0);
            make_householder(m, r0.data(), Hc(0) + 1, tau[0], beta);
            w[0] = beta;
        }
    }
    return iters;
}

}  // namespace oracle

// =======================================================================================
// C API for the test harness (ctypes).
using namespace oracle;

extern "C" {

void* or_create(int nlevels) { return new Tree(nlevels); }
void or_destroy(void* h) { delete (Tree*)h; }
void or_set_tol(void* h, double tol) { ((Tree*)h)->tol = tol; }
void or_set_skip(void* h, int skip) { ((Tree*)h)->skip = skip; }
void or_set_symm(void* h, int symm) { ((Tree*)h)->symm = symm; }
void or_set_verb(void* h, int v) { ((Tree*)h)->verb = v != 0; }
void or_set_record_shapes(void* h, int v) { ((Tree*)h)->record_shapes = v != 0; }
void or_set_Xcoo(void* h, int dim, int N, const double* X) {
    Tree* t = (Tree*)h;
    t->dim = dim;
    t->geo = true;
    t->X.assign(X, X + (size_t)dim * N);
}
const char* or_last_error(void* h) { return ((Tree*)h)->last_error.c_str(); }

int or_partition(void* h, int N, const int* colptr, const int* rowval, int* part_out) {
    Tree* t = (Tree*)h;
    try {
        std::vector<int> cp(colptr, colptr + N + 1), rv(rowval, rowval + colptr[N]);
        t->partition(N, cp, rv);
        if (part_out) std::copy(t->part_out.begin(), t->part_out.end(), part_out);
        return 0;
    } catch (std::exception& e) {
        t->last_error = e.what();
        return 1;
    }
}
void or_get_perm(void* h, int* perm) {
    Tree* t = (Tree*)h;
    std::copy(t->perm.begin(), t->perm.end(), perm);
}
int or_assemble(void* h, int N, const int* colptr, const int* rowval, const double* vals) {
    Tree* t = (Tree*)h;
    try {
        t->assemble(N, colptr, rowval, vals);
        return 0;
    } catch (std::exception& e) {
        t->last_error = e.what();
        return 1;
    }
}
int or_factorize(void* h) {
    Tree* t = (Tree*)h;
    try {
        t->factorize();
        return 0;
    } catch (std::exception& e) {
        t->last_error = e.what();
        return 1;
    }
}
void or_solve(void* h, double* x) { ((Tree*)h)->solve(x); }
long long or_nnz(void* h) { return ((Tree*)h)->nnz(); }
int or_get_stop(void* h) { return ((Tree*)h)->get_stop(); }
int or_get_N(void* h) { return ((Tree*)h)->get_N(); }

// Per-level log: 12 numbers per level.
void or_get_log(void* h, double* out) {
    Tree* t = (Tree*)h;
    for (int l = 0; l < t->nlevels; l++) {
        LevelLog& g = t->log[l];
        double* o = out + 12 * l;
        o[0] = g.dofs_nd; o[1] = g.dofs_left_nd; o[2] = g.dofs_left_elim; o[3] = g.dofs_left_spars;
        o[4] = (double)g.fact_nnz; o[5] = g.rank_before_sum; o[6] = g.rank_after_sum; o[7] = g.sparsified;
        o[8] = g.t_elim; o[9] = g.t_scale; o[10] = g.t_spars; o[11] = g.t_merge;
    }
}

// Current (bottom_current) trailing matrix as COO of the lower triangle (+diag) for SPD,
// full for GEN (src/tree.cpp:1730-1763). Returns the number of entries written (or needed
// when out arrays are null).
long long or_trailing_coo(void* h, int* I, int* J, double* V) {
    Tree* t = (Tree*)h;
    long long n = 0;
    for (auto& s : t->bottom_current()) {
        for (auto& e : s->edgesOut) {
            Mat* A = e->A();
            for (int i_ = 0; i_ < A->rows; i_++)
                for (int j_ = 0; j_ < A->cols; j_++) {
                    int i = e->n2->start + i_, j = s->start + j_;
                    if (t->symmetry() && i < j) continue;
                    if (I) { I[n] = i; J[n] = j; V[n] = (*A)(i_, j_); }
                    n++;
                }
        }
    }
    return n;
}

// Op list introspection. info: type, order(a), order(b), sa, sb, rows(M), cols(M), level.
int or_num_ops(void* h) { return (int)((Tree*)h)->ops.size(); }
void or_ops_info(void* h, int* info) {
    Tree* t = (Tree*)h;
    int k = 0;
    for (auto& op : t->ops) {
        int* o = info + 8 * k++;
        o[0] = op->type;
        o[1] = op->a ? op->a->order : -1;
        o[2] = op->b ? op->b->order : -1;
        o[3] = op->sa;
        o[4] = op->sb;
        o[5] = op->M ? op->M->rows : 0;
        o[6] = op->M ? op->M->cols : 0;
        o[7] = op->level;
    }
}
// Copies op i's matrix M (col-major) into out, and tau (if any) into tau_out.
int or_op_data(void* h, int i, double* out, double* tau_out, double* U_out, int* p_out) {
    Tree* t = (Tree*)h;
    auto it = t->ops.begin();
    std::advance(it, i);
    Op* op = it->get();
    if (op->M && out) std::copy(op->M->d.begin(), op->M->d.end(), out);
    if (tau_out) std::copy(op->tau.begin(), op->tau.end(), tau_out);
    if (U_out && op->U) std::copy(op->U->d.begin(), op->U->d.end(), U_out);
    if (p_out) std::copy(op->p.begin(), op->p.end(), p_out);
    return 0;
}

// Shape log: 5 ints per record (level, kind, m, n, k).
long long or_num_shapes(void* h) { return (long long)((Tree*)h)->shapes.size(); }
void or_get_shapes(void* h, int* out) {
    Tree* t = (Tree*)h;
    for (size_t i = 0; i < t->shapes.size(); i++) {
        const Shape& s = t->shapes[i];
        int* o = out + 5 * i;
        o[0] = s.level; o[1] = s.kind; o[2] = s.m; o[3] = s.n; o[4] = s.k;
    }
}

// Cluster table of every cluster ever created (bottoms then others): order, start, size,
// original_size, level, parent order (-1), eliminated, sparsify.
int or_num_clusters(void* h) {
    Tree* t = (Tree*)h;
    int n = (int)t->others.size();
    for (auto& b : t->bottoms) n += (int)b.size();
    return n;
}
// clusters listed by Tree::get_clusters (src/tree.cpp:265-273): the bottoms only (or_clusters
// lists them first, then the sparsification siblings); level >= 0: that bottom's count
int or_num_bottom_clusters(void* h, int level) {
    Tree* t = (Tree*)h;
    if (level >= 0) return level < (int)t->bottoms.size() ? (int)t->bottoms[level].size() : 0;
    int n = 0;
    for (auto& b : t->bottoms) n += (int)b.size();
    return n;
}
void or_clusters(void* h, int* out) {
    Tree* t = (Tree*)h;
    int k = 0;
    auto put = [&](Cluster* c) {
        int* o = out + 8 * k++;
        o[0] = c->order; o[1] = c->start; o[2] = c->size; o[3] = c->original_size();
        o[4] = c->level; o[5] = c->parent ? c->parent->order : -1; o[6] = c->eliminated; o[7] = c->sparsify;
    };
    for (auto& b : t->bottoms)
        for (auto& c : b) put(c.get());
    for (auto& c : t->others) put(c.get());
}

int or_cg(void* h, int N, const int* colptr, const int* rowval, const double* vals, const double* b, double* x,
          int iters, double tol, double* t_prec) {
    double tp = 0;
    int it = cg(*(Tree*)h, N, colptr, rowval, vals, b, x, iters, tol, &tp);
    if (t_prec) *t_prec = tp;
    return it;
}

int or_gmres(void* h, int N, const int* colptr, const int* rowval, const double* vals, const double* b, double* x,
             int iters, int restart, double tol, double* rel, double* t_prec) {
    double tp = 0, r = 0;
    int it = gmres(*(Tree*)h, N, colptr, rowval, vals, b, x, iters, restart, tol, &r, &tp);
    if (rel) *rel = r;
    if (t_prec) *t_prec = tp;
    return it;
}

// Dense kernels, exported for golden-vector tests.
int or_potrf(int n, double* A, int lda) { return potrf_lower(n, A, lda); }
void or_trsm_llnn(int m, int n, const double* L, int ldl, double* B, int ldb) { trsm_left_lower_notrans(m, n, L, ldl, B, ldb); }
void or_trsm_rltn(int m, int n, const double* L, int ldl, double* B, int ldb) { trsm_right_lower_trans(m, n, L, ldl, B, ldb); }
void or_trsm_runn(int m, int n, const double* U, int ldu, double* B, int ldb) { trsm_right_upper_notrans(m, n, U, ldu, B, ldb); }
void or_syrk(int trans, int n, int k, double alpha, const double* A, int lda, double beta, double* C, int ldc) {
    syrk_lower(trans != 0, n, k, alpha, A, lda, beta, C, ldc);
}
void or_gemm(int tA, int tB, int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
             double beta, double* C, int ldc) {
    gemm(tA != 0, tB != 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
void or_schur(int tA, int tB, int lower, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
              double* C, int ldc) {
    schur_update(tA != 0, tB != 0, lower != 0, m, n, k, A, lda, B, ldb, C, ldc);
}
double or_sumsq16(int n, const double* x) { return sumsq16(n, x, 1); }
double or_dot16(int n, const double* a, const double* b) { return dot16(n, a, b); }
void or_geqp3(int m, int n, double* A, int lda, int* jpvt, double* tau) { geqp3(m, n, A, lda, jpvt, tau); }
void or_ormqr_vec(int trans, int m, int k, const double* V, int ldv, const double* tau, double* x) {
    ormqr_vec(trans != 0, m, k, V, ldv, tau, x);
}
int or_getrf(int n, double* A, int lda, int* ipiv) { return getrf(n, A, lda, ipiv); }
void or_swap2perm(int n, const int* swap, int* perm) { swap2perm(n, swap, perm); }
int or_choose_rank(const double* s, int n, double tol) { return choose_rank(s, n, tol); }
void or_trsv(int kind, int n, const double* A, int lda, double* x) {
    if (kind == 0) trsv_lower_notrans(n, A, lda, x);
    else if (kind == 1) trsv_lower_trans(n, A, lda, x);
    else trsv_upper_notrans(n, A, lda, x);
}

}  // extern "C"
