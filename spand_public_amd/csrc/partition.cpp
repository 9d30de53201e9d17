// Geometric modified nested dissection and the cluster ordering of spaND, host side.
//
//   partition_nd_geo     src/partition.cpp:139-210
//   partition_modifiedND src/partition.cpp:384-478 (geometric branch)
//   merge_if / ordering  src/partition.cpp:8-16, src/tree.cpp:336-415
//
// The output (ND ids, assembly permutation, leaf clusters, parent hierarchy) is bit-exact
// with the reference: only comparisons and integer arithmetic are involved, and the stable
// sorts of src/tree.cpp:348,351 are reproduced with std::stable_sort.
#include <algorithm>
#include <limits>
#include <numeric>
#include <stdexcept>

#include "tree.hpp"

namespace spand {

namespace {

struct Sep {
    int lvl, sep;
    bool operator==(const Sep& o) const { return lvl == o.lvl && sep == o.sep; }
    bool operator<(const Sep& o) const { return lvl < o.lvl || (lvl == o.lvl && sep < o.sep); }
};
struct Cid {
    Sep self, l, r;
    bool operator==(const Cid& o) const { return self == o.self && l == o.l && r == o.r; }
    bool operator<(const Cid& o) const {
        if (!(self == o.self)) return self < o.self;
        if (!(l == o.l)) return l < o.l;
        return r < o.r;
    }
};
inline Sep up(Sep s) { return Sep{s.lvl + 1, s.sep / 2}; }
inline Cid merged(const Cid& c, int lvl) {
    return Cid{c.self, c.l.lvl < lvl ? up(c.l) : c.l, c.r.lvl < lvl ? up(c.r) : c.r};
}

// One geometric bisection with vertex separator of the sorted dof list `dofs`:
// result[i] in {0 (left), 1 (right), 2 (separator)}.
void bisect_geo(const std::vector<int>& adjp, const std::vector<int>& adji, const std::vector<int>& dofs,
                const double* X, int dim, std::vector<int>& mark, std::vector<int>& result,
                std::vector<int>& scratch) {
    const int n = (int)dofs.size();
    result.assign(n, 0);
    if (n == 0) return;
    int axis = -1;
    double widest = -1.0;
    for (int d = 0; d < dim; d++) {
        double lo = std::numeric_limits<double>::max(), hi = std::numeric_limits<double>::lowest();
        for (int v : dofs) {
            double c = X[(size_t)d + (size_t)dim * v];
            lo = std::min(lo, c);
            hi = std::max(hi, c);
        }
        if (hi - lo > widest) {
            widest = hi - lo;
            axis = d;
        }
    }
    auto coord = [&](int v) { return X[(size_t)axis + (size_t)dim * v]; };
    // the split value is the (n/2)-th smallest coordinate along the axis
    scratch.assign(dofs.begin(), dofs.end());
    std::nth_element(scratch.begin(), scratch.begin() + n / 2, scratch.end(),
                     [&](int a, int b) { return coord(a) < coord(b); });
    const double split = coord(scratch[n / 2]);
    for (int v : dofs) {
        mark[v] = -1;
        for (int k = adjp[v]; k < adjp[v + 1]; k++) mark[adji[k]] = -1;
    }
    for (int i = 0; i < n; i++) {
        mark[dofs[i]] = i;
        result[i] = coord(dofs[i]) < split ? 0 : 1;
    }
    for (int i = 0; i < n; i++) {
        if (result[i] == 0) continue;
        const int v = dofs[i];
        for (int k = adjp[v]; k < adjp[v + 1]; k++) {
            int j = mark[adji[k]];
            if (j >= 0 && result[j] == 0) {
                result[i] = 2;
                break;
            }
        }
    }
}

}  // namespace

void Tree::partition(int N, const int* colptr, const int* rowval) {
    if (state_ != State::Created) throw Error(ERR_STATE, "partition() must be called once, before assemble()");
    if (!geo_ || (int)X_.size() != dim_ * N)
        throw Error(ERR_UNSUPPORTED, "only geometric partitioning is available: call set_use_geo(1) and set_Xcoo");
    // adjacency without self loops (SpMat2CSC, src/partition.cpp:21-39)
    std::vector<int> adjp(N + 1, 0), adji;
    adji.reserve(colptr[N]);
    for (int j = 0; j < N; j++) {
        for (int k = colptr[j]; k < colptr[j + 1]; k++)
            if (rowval[k] != j) adji.push_back(rowval[k]);
        adjp[j + 1] = (int)adji.size();
    }
    const Sep top{nlevels_ - 1, 0};
    std::vector<Cid> id(N, Cid{top, top, top});
    std::vector<std::vector<int>> domains(1, std::vector<int>(N));
    std::iota(domains[0].begin(), domains[0].end(), 0);
    std::vector<int> mark(N), res, scratch;
    for (int depth = 0; depth + 1 < nlevels_; depth++) {
        const int level = nlevels_ - depth - 1;
        std::vector<std::vector<int>> next((size_t)2 << depth);
        for (int s = 0; s < (1 << depth); s++) {
            std::vector<int>& dofs = domains[s];
            std::sort(dofs.begin(), dofs.end());
            bisect_geo(adjp, adji, dofs, X_.data(), dim_, mark, res, scratch);
            const Sep me{level, s}, L{level - 1, 2 * s}, R{level - 1, 2 * s + 1};
            for (size_t i = 0; i < dofs.size(); i++) {
                Cid& c = id[dofs[i]];
                const int side = res[i];
                if (side == 2) {
                    if (c.self == me) {
                        c.l = L;
                        c.r = R;
                    }
                    continue;
                }
                const Sep to = side == 0 ? L : R;
                if (c.self == me) c.self = to;
                if (c.l == me) c.l = to;
                if (c.r == me) c.r = to;
            }
            for (int v : dofs) {
                const Cid& c = id[v];
                if (c.self == L || c.l == L || c.r == L) next[2 * s].push_back(v);
                if (c.self == R || c.l == R || c.r == R) next[2 * s + 1].push_back(v);
            }
        }
        domains.swap(next);
    }
    part_.resize((size_t)6 * N);
    for (int i = 0; i < N; i++) {
        const Cid& c = id[i];
        int* o = &part_[(size_t)6 * i];
        o[0] = c.self.lvl; o[1] = c.self.sep; o[2] = c.l.lvl; o[3] = c.l.sep; o[4] = c.r.lvl; o[5] = c.r.sep;
    }
    // ordering: stable sort by ND id, then repeatedly by the merged ids (src/tree.cpp:344-352)
    perm_.resize(N);
    std::iota(perm_.begin(), perm_.end(), 0);
    std::vector<Cid> key = id;
    auto by_key = [&key](int a, int b) { return key[a] < key[b]; };
    std::stable_sort(perm_.begin(), perm_.end(), by_key);
    for (int lvl = 1; lvl < nlevels_; lvl++) {
        for (auto& k : key) k = merged(k, lvl);
        std::stable_sort(perm_.begin(), perm_.end(), by_key);
    }
    // leaf clusters = runs of equal ND id in the permuted order (src/tree.cpp:360-370)
    N_ = N;
    cls_.clear();
    bottoms_.assign(nlevels_, {});
    std::vector<Cid> cid_of;  // ClusterID of every cluster, by cluster id
    for (int k = 0; k < N;) {
        const Cid c = id[perm_[k]];
        int e = k + 1;
        while (e < N && id[perm_[e]] == c) e++;
        HCluster h;
        h.start = k;
        h.size = h.osize = e - k;
        h.level = c.self.lvl;
        h.spars = c.l.lvl == 0 && c.r.lvl == 0;
        bottoms_[0].push_back((int)cls_.size());
        cls_.push_back(std::move(h));
        cid_of.push_back(c);
        k = e;
    }
    // parents (src/tree.cpp:381-415)
    for (int lvl = 1; lvl < nlevels_; lvl++) {
        const std::vector<int>& below = bottoms_[lvl - 1];
        size_t b = 0;
        while (b < below.size() && cls_[below[b]].level < lvl) b++;
        for (size_t k = b; k < below.size();) {
            const Cid p = merged(cid_of[below[k]], lvl);
            HCluster h;
            h.start = cls_[below[k]].start;
            h.size = 0;
            h.level = p.self.lvl;
            h.spars = p.l.lvl == lvl && p.r.lvl == lvl;
            const int pid = (int)cls_.size();
            while (k < below.size() && merged(cid_of[below[k]], lvl) == p) {
                h.children.push_back(below[k]);
                h.size += cls_[below[k]].size;
                cls_[below[k]].parent = pid;
                k++;
            }
            h.osize = h.size;
            bottoms_[lvl].push_back(pid);
            cls_.push_back(std::move(h));
            cid_of.push_back(p);
        }
    }
    for (size_t i = 0; i < cls_.size(); i++) cls_[i].order = (int)i;
    state_ = State::Partitioned;
}

}  // namespace spand
