// spaND factorization and preconditioner apply on one GPU.
//
// Host side of the reference's Tree::assemble / factorize / solve (src/tree.cpp:505-575,
// 1447-1551, 1610-1635). The host mirrors the reference's structural bookkeeping exactly
// (edge list order, fill-in creation order, cluster orders, sibling creation), which fixes
// the column order of every RRQR and the op order of the solve. Numerical work is planned
// per ND level and launched as level-batched kernels:
//   Elim:   potrf(all pivots) -> trsm(all panels) -> Schur(all targets, ordered accumulation)
//   Scale:  potrf(all pivots) -> trsm by the lower-order end of every edge -> trsm by the
//           higher-order end (the reference's in-order rounding, src/tree.cpp:1512-1516)
//   Sparsify: dependency waves (a cluster waits for its earlier-ordered sparsified
//           neighbours, src/tree.cpp:1013-1017,1041-1045), one RRQR launch per wave
//   Merge:  zero-fill + block copies into the parents' blocks
// with one host<->device synchronization per level (the ranks decide the next level's sizes).
#include "tree.hpp"

#include "scratch.hpp"
#include "tune.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <numeric>
#include <unordered_map>

#include "kernels.hpp"

namespace spand {

static double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw Error(ERR_NOMEM, std::string(what) + ": out of device memory");
        throw Error(ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// ---------------------------------------------------------------------------------------
size_t arena_chunk_bytes(size_t dflt) {
    const char* e = std::getenv("SPAND_ARENA_CHUNK_BYTES");
    if (!e || !*e) return dflt;
    return (size_t)std::strtoull(e, nullptr, 10);
}

DeviceArena::~DeviceArena() {
    for (auto& c : chunks_) (void)hipFree(c.p);
}
void* DeviceArena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    if (chunk_bytes_ == 0) {
        // one allocation per block (bounds-checking builds); kept until the arena dies
        Chunk c;
        hip_check(hipMalloc((void**)&c.p, bytes), "hipMalloc(arena block)");
        c.size = c.used = bytes;
        chunks_.push_back(c);
        in_use_ += bytes;
        return c.p;
    }
    while (cur_ < chunks_.size()) {
        Chunk& c = chunks_[cur_];
        if (c.used + bytes <= c.size) {
            void* p = c.p + c.used;
            c.used += bytes;
            in_use_ += bytes;
            return p;
        }
        cur_++;
    }
    Chunk c;
    c.size = std::max(chunk_bytes_, bytes);
    hip_check(hipMalloc((void**)&c.p, c.size), "hipMalloc(arena chunk)");
    c.used = bytes;
    chunks_.push_back(c);
    cur_ = chunks_.size() - 1;
    in_use_ += bytes;
    return c.p;
}
void DeviceArena::reset() {
    if (chunk_bytes_ == 0) {
        for (auto& c : chunks_) (void)hipFree(c.p);
        chunks_.clear();
    }
    for (auto& c : chunks_) c.used = 0;
    cur_ = 0;
    in_use_ = 0;
}
size_t DeviceArena::bytes_reserved() const {
    size_t s = 0;
    for (auto& c : chunks_) s += c.size;
    return s;
}

// ---------------------------------------------------------------------------------------
Tree::Tree(int nlevels) : nlevels_(nlevels), stats_(nlevels) {
    if (nlevels <= 0) throw Error(ERR_ARG, "nlevels must be positive");
}

Tree::~Tree() {
    (void)hipFree(d_scatter_off_);
    (void)hipFree(d_scatter_src_);
    (void)hipFree(d_scatter_val_);
    (void)hipFree(d_size_);
    (void)hipFree(d_info_);
    (void)hipFree(d_perm_);
    (void)hipFree(d_ws_);
    (void)hipHostFree(h_stage_);
    (void)hipHostFree(h_sig_);
    drop_apply_graphs();
    if (cap_st_) (void)hipStreamDestroy(cap_st_);
    if (aux_st_) {
        (void)hipStreamSynchronize(aux_st_);
        (void)hipStreamDestroy(aux_st_);
        (void)hipEventDestroy(ev_fork_);
        (void)hipEventDestroy(ev_join_);
    }
}

// SymmKind (include/spaND.h): SPD (0) with ScalingKind::LLT, GEN (2) with ScalingKind::PLU.
// SYM (1, LDLT) is not built.
// The kind decides which blocks assemble() builds (lower only for SPD), so it cannot change
// once the matrix is assembled.
void Tree::set_symm_kind(int k) {
    if (k != 0 && k != 2) throw Error(ERR_UNSUPPORTED, "only SPD (LLT) and GEN (PLU) are available");
    if (k != symm_ && (state_ == State::Assembled || state_ == State::Factorized))
        throw Error(ERR_STATE, "set_symm_kind() must come before assemble()");
    symm_ = k;
}
void Tree::set_Xcoo(int dim, int N, const double* X) {
    if (dim <= 0 || N <= 0 || !X) throw Error(ERR_ARG, "set_Xcoo: bad arguments");
    dim_ = dim;
    X_.assign(X, X + (size_t)dim * N);
}
void Tree::set_precision(int p) {
    if (p != 0 && p != 1) throw Error(ERR_ARG, "precision must be 0 (FP64) or 1 (FP32)");
    prec_ = p;
}

int Tree::find_out_edge(int from, int to) const {
    for (int e : cls_[from].out)
        if (edges_[e].n2 == to) return e;
    return -1;
}

// Cluster::add_edge (src/cluster.cpp:113-123): the pivot goes first, others are appended.
int Tree::new_edge(int n1, int n2, void* A, bool original) {
    int id = (int)edges_.size();
    HEdge e;
    e.n1 = n1;
    e.n2 = n2;
    e.A = A;
    e.original = original;
    edges_.push_back(e);
    if (n1 == n2) {
        cls_[n1].out.insert(cls_[n1].out.begin(), id);
    } else {
        cls_[n1].out.push_back(id);
        cls_[n2].in.push_back(id);
    }
    return id;
}

// Cluster::set_eliminated (src/cluster.cpp:32-45)
void Tree::set_eliminated(int c) {
    HCluster& s = cls_[c];
    for (int e : s.out) {
        int n = edges_[e].n2;
        if (n == c) continue;
        auto& v = cls_[n].in;
        v.erase(std::remove_if(v.begin(), v.end(), [&](int x) { return edges_[x].n1 == c; }), v.end());
    }
    for (int e : s.in) {
        int n = edges_[e].n1;
        auto& v = cls_[n].out;
        v.erase(std::remove_if(v.begin(), v.end(), [&](int x) { return edges_[x].n2 == c; }), v.end());
    }
    s.out.clear();
    s.in.clear();
    s.elim = true;
}

// Host -> device uploads of the factorization go through a pinned staging buffer and are
// asynchronous on the factorization stream, so the host plans the next batch while the
// device works. The staging area is reused from the start at every factorization (the
// previous one ended with a device synchronization); when it runs out, the device is
// synchronized and the buffer grown to hold everything staged since the factorization
// began. Such a factorization ends by growing the buffer to its whole upload volume, so
// that repeated factorizations of one pattern never stall on a growth.
char* Tree::stage_reserve(size_t bytes) {
    const size_t off = (h_stage_used_ + 255) & ~(size_t)255;
    if (off + bytes > h_stage_cap_) {
        // earlier uploads may still read the old buffer: drain before replacing it
        hip_check(hipDeviceSynchronize(), "staging");
        stage_grow(std::max<size_t>(2 * h_stage_cap_, std::max<size_t>(h_stage_total_ + bytes, (size_t)8 << 20)));
        return stage_reserve(bytes);
    }
    h_stage_total_ += off + bytes - h_stage_used_;
    h_stage_used_ = off + bytes;
    return h_stage_ + off;
}

// Replaces the staging buffer by an empty one of cap bytes (no upload may be in flight).
void Tree::stage_grow(size_t cap) {
    (void)hipHostFree(h_stage_);
    h_stage_ = nullptr;
    h_stage_cap_ = 0;
    hip_check(hipHostMalloc((void**)&h_stage_, cap, hipHostMallocDefault), "hipHostMalloc(staging)");
    h_stage_cap_ = cap;
    h_stage_used_ = 0;
}

// Grows the signal buffer (between levels: no launch that writes it is in flight).
void Tree::signal_reserve(size_t n) {
    if (n <= h_sig_cap_) return;
    hip_check(hipDeviceSynchronize(), "signals");
    (void)hipHostFree(h_sig_);
    h_sig_ = nullptr;
    d_sig_ = nullptr;
    h_sig_cap_ = 0;
    const size_t cap = std::max<size_t>(n, 4096);
    hip_check(hipHostMalloc((void**)&h_sig_, cap * sizeof(long long), hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc(signals)");
    hip_check(hipHostGetDevicePointer((void**)&d_sig_, h_sig_, 0), "signals device address");
    std::memset(h_sig_, 0, cap * sizeof(long long));
    h_sig_cap_ = cap;
    sig_seq_ = 0;
}

static inline void cpu_relax() {
#if defined(__x86_64__) || defined(__i386__)
    __builtin_ia32_pause();
#endif
}

// Waits until the step launch numbered target has started on the device. A wait longer
// than two seconds turns the throttle off for the life of the tree (every step is then
// launched up front, as without signals) and is reported by status_flags().
void Tree::wait_signal(long long target) {
    if (!sig_throttle_) return;
    const volatile long long* p = h_sig_;
    if (*p >= target) return;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned n = 1; *p < target; n++) {
        cpu_relax();
        if ((n & 4095) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
            sig_throttle_ = false;
            throttle_gave_up_ = true;
            return;
        }
    }
}

// Algorithmic bytes / flops of the profiled stepped-RRQR waves (init, steps, finish), from
// the rows, columns and steps run that the device decided.
void Tree::price_waves(size_t esz) {
    std::vector<std::vector<int>> dims(level_dims_.size());
    for (const WaveWork& ww : wave_work_) {
        std::vector<int>& d = dims[ww.level];
        if (d.empty()) {
            d.resize(5 * level_dims_[ww.level].second);
            hip_check(hipMemcpy(d.data(), level_dims_[ww.level].first, d.size() * sizeof(int), hipMemcpyDeviceToHost),
                      "read dims");
        }
        double ib = 0, fb = 0, by = 0, fl = 0;
        for (size_t k = ww.first; k < ww.first + ww.n; k++) {
            const double rows = d[5 * k], cols = d[5 * k + 1];
            const int steps = d[5 * k + 3];  // the QR stops at the rank
            ib += rows * cols * 2 * esz + 16.0 * cols;
            fb += (rows * cols * 2 + rows * steps) * esz;
            for (int i = 0; i < std::min(steps, ww.steps); i++) {
                const double m = rows - i, n = cols - i;
                by += (2.0 * m * n + m) * esz + 16.0 * n;
                fl += 4.0 * m * n;
            }
        }
        prof_.set_work(ww.slot_init, ib, 0);
        prof_.set_work(ww.slot_finish, fb, 0);
        prof_.set_work(ww.slot_step, by, fl);
    }
    wave_work_.clear();
    level_dims_.clear();
}

void Tree::stage_upload(void* d, const void* h, size_t bytes) {
    if (bytes == 0) return;
    char* p = stage_reserve(bytes);
    std::memcpy(p, h, bytes);
    hip_check(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, 0), "upload");
}

// Packs several task arrays into one upload. add() only records the array (it must stay
// unchanged until upload_tasks); upload_tasks copies each array once, straight into the
// pinned staging buffer, and issues one asynchronous copy.
struct Packer {
    struct Item {
        const void* p;
        size_t off, bytes;
    };
    std::vector<Item> items;
    size_t size = 0;
    template <class X>
    size_t add(const std::vector<X>& v) {
        const size_t off = (size + 255) & ~(size_t)255;
        items.push_back({v.data(), off, v.size() * sizeof(X)});
        size = off + v.size() * sizeof(X);
        return off;
    }
    template <class X>
    static X* at(void* base, size_t off) {
        return reinterpret_cast<X*>(reinterpret_cast<char*>(base) + off);
    }
};

void* Tree::upload_tasks(const Packer& pk) {
    if (pk.size == 0) return nullptr;
    void* d = task_arena_.alloc(pk.size);
    char* h = stage_reserve(pk.size);
    for (const auto& it : pk.items)
        if (it.bytes) std::memcpy(h + it.off, it.p, it.bytes);
    hip_check(hipMemcpyAsync(d, h, pk.size, hipMemcpyHostToDevice, 0), "upload");
    return d;
}

// ---------------------------------------------------------------------------------------
// Tree::assemble (src/tree.cpp:505-575): P^T A P cut into dense blocks, one edge per
// (leaf cluster, neighbour cluster) pair with order(nbr) >= order(self) (symmetric). The
// dense blocks are never materialized on the host: the device zero-fills the level-0
// region and scatters the nonzeros to precomputed offsets.
void Tree::assemble(int N, const int* colptr, const int* rowval, const double* vals) {
    // Called again after an assembly or a factorization, it re-assembles new values (of the
    // partitioned pattern) from the partition's clusters.
    if (state_ == State::Assembled || state_ == State::Factorized) cls_ = cls0_;
    else if (state_ != State::Partitioned) throw Error(ERR_STATE, "assemble() must follow partition()");
    if (N != N_) throw Error(ERR_ARG, "assemble: matrix size differs from the partitioned one");
    std::vector<int> pinv(N);
    for (int i = 0; i < N; i++) pinv[perm_[i]] = i;
    const long long nnz = colptr[N];
    // permuted CSC with sorted rows
    std::vector<long long> cp(N + 1, 0);
    for (int j = 0; j < N; j++) cp[pinv[j] + 1] += colptr[j + 1] - colptr[j];
    for (int j = 0; j < N; j++) cp[j + 1] += cp[j];
    // (permuted row, index of the entry in the caller's CSC arrays)
    std::vector<std::pair<int, long long>> ent(nnz);
    {
        std::vector<long long> fill(cp.begin(), cp.end() - 1);
        for (int j = 0; j < N; j++) {
            int pj = pinv[j];
            for (int k = colptr[j]; k < colptr[j + 1]; k++) ent[fill[pj]++] = {pinv[rowval[k]], (long long)k};
        }
        for (int j = 0; j < N; j++)
            std::sort(ent.begin() + cp[j], ent.begin() + cp[j + 1],
                      [](const auto& a, const auto& b) { return a.first < b.first; });
    }
    std::vector<int> cmap(N);
    for (int c : bottoms_[0])
        for (int k = cls_[c].start; k < cls_[c].start + cls_[c].size; k++) cmap[k] = c;
    const bool symmetric = symm_ == 0;
    edges_.clear();
    for (auto& c : cls_) {
        c.out.clear();
        c.in.clear();
    }
    edge0_off_.clear();
    level0_elems_ = 0;
    std::vector<long long> soff, ssrc;
    std::vector<double> sval;
    soff.reserve(nnz);
    ssrc.reserve(nnz);
    sval.reserve(nnz);
    std::vector<int> edge_to(cls_.size(), -1);
    for (int self : bottoms_[0]) {
        HCluster& s = cls_[self];
        std::vector<int> nbrs{self};
        for (int j = s.start; j < s.start + s.size; j++)
            for (long long k = cp[j]; k < cp[j + 1]; k++) {
                int i = ent[k].first;
                if (symmetric && i < j) continue;
                nbrs.push_back(cmap[i]);
            }
        std::sort(nbrs.begin(), nbrs.end());
        nbrs.erase(std::unique(nbrs.begin(), nbrs.end()), nbrs.end());
        for (int n : nbrs) {
            int e = new_edge(self, n, nullptr, true);
            edge_to[n] = e;
            edge0_off_.push_back(level0_elems_);
            level0_elems_ += (long long)cls_[n].size * s.size;
        }
        // Cluster::sort_edges (src/cluster.cpp:162-176)
        auto& out = s.out;
        std::stable_sort(out.begin() + 1, out.end(), [&](int a, int b) { return edges_[a].n2 < edges_[b].n2; });
        std::stable_sort(s.in.begin(), s.in.end(), [&](int a, int b) { return edges_[a].n1 < edges_[b].n1; });
        for (int j = s.start; j < s.start + s.size; j++)
            for (long long k = cp[j]; k < cp[j + 1]; k++) {
                int i = ent[k].first;
                int n = cmap[i];
                if (symmetric && i < j && n != self) continue;
                int e = edge_to[n];
                if (e < 0 || edges_[e].n1 != self) continue;
                soff.push_back(edge0_off_[e] + (i - cls_[n].start) + (long long)(j - s.start) * cls_[n].size);
                ssrc.push_back(ent[k].second);
                sval.push_back(vals[ent[k].second]);
            }
        for (int n : nbrs) edge_to[n] = -1;
    }
    nscatter_ = (long long)soff.size();
    h_scatter_off_.swap(soff);
    h_scatter_src_.swap(ssrc);
    h_scatter_val_.swap(sval);
    nnz_in_ = nnz;
    device_inputs_ = false;
    host_vals_stale_ = false;
    // ND statistics (src/tree.cpp:336-342)
    for (auto& st : stats_) st = LevelStats();
    for (int i = 0; i < N; i++) stats_[part_[(size_t)6 * i]].dofs_nd++;
    for (int l = nlevels_ - 2; l >= 0; l--) stats_[l].dofs_left_nd = stats_[l + 1].dofs_nd + stats_[l + 1].dofs_left_nd;
    cls0_ = cls_;
    edges0_ = edges_;
    state_ = State::Assembled;
}

// ---------------------------------------------------------------------------------------
// Device copies of the assembly scatter map and of the permutation (first factorize()).
void Tree::upload_inputs() {
    if (device_inputs_) return;
    (void)hipFree(d_scatter_off_);
    (void)hipFree(d_scatter_src_);
    (void)hipFree(d_scatter_val_);
    (void)hipFree(d_perm_);
    d_scatter_off_ = nullptr;
    d_scatter_src_ = nullptr;
    d_scatter_val_ = nullptr;
    d_perm_ = nullptr;
    if (nscatter_ > 0) {
        hip_check(hipMalloc((void**)&d_scatter_off_, nscatter_ * sizeof(long long)), "hipMalloc(scatter)");
        hip_check(hipMalloc((void**)&d_scatter_src_, nscatter_ * sizeof(long long)), "hipMalloc(scatter)");
        hip_check(hipMalloc((void**)&d_scatter_val_, nscatter_ * sizeof(double)), "hipMalloc(scatter)");
        hip_check(hipMemcpy(d_scatter_off_, h_scatter_off_.data(), nscatter_ * sizeof(long long),
                            hipMemcpyHostToDevice),
                  "upload scatter");
        hip_check(hipMemcpy(d_scatter_src_, h_scatter_src_.data(), nscatter_ * sizeof(long long),
                            hipMemcpyHostToDevice),
                  "upload scatter");
        hip_check(hipMemcpy(d_scatter_val_, h_scatter_val_.data(), nscatter_ * sizeof(double), hipMemcpyHostToDevice),
                  "upload scatter");
    }
    hip_check(hipMalloc((void**)&d_perm_, (size_t)N_ * sizeof(int)), "hipMalloc(perm)");
    hip_check(hipMemcpy(d_perm_, perm_.data(), (size_t)N_ * sizeof(int), hipMemcpyHostToDevice), "upload perm");
    device_inputs_ = true;
}

// Tree::assemble with new values of the assembled pattern, read on the device: vals holds
// the nnz entries in the CSC order given to assemble(). The next factorize() uses them.
void Tree::assemble_values_device(const double* vals, long long nnz, hipStream_t st) {
    if (state_ != State::Assembled && state_ != State::Factorized)
        throw Error(ERR_STATE, "assemble_values_device() must follow assemble()");
    if (nnz != nnz_in_) throw Error(ERR_ARG, "assemble_values_device: nnz differs from the assembled pattern");
    if (!vals && nscatter_ > 0) throw Error(ERR_ARG, "assemble_values_device: null values");
    upload_inputs();
    if (nscatter_ > 0) {
        hip_check(launch_gather_f64(d_scatter_src_, vals, nscatter_, d_scatter_val_, st), "gather values");
        hip_check(hipStreamSynchronize(st), "gather values");
    }
    host_vals_stale_ = true;
    cls_ = cls0_;
    edges_ = edges0_;
    state_ = State::Assembled;
}

// Second stream of the factorization (non-blocking, default priority) and its fork / join
// events, created on first use.
hipStream_t Tree::aux_stream() {
    if (!aux_st_) {
        hip_check(hipStreamCreateWithFlags(&aux_st_, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming), "hipEventCreate");
    }
    return aux_st_;
}

void Tree::factorize() {
    if (state_ != State::Assembled && state_ != State::Factorized)
        throw Error(ERR_STATE, "factorize() must follow assemble()");
    upload_inputs();
    // a factorization that fails part-way leaves no usable factor: solve() refuses until the
    // next factorize() succeeds
    state_ = State::Assembled;
    fprec_ = prec_;
    if (prec_ == 0)
        factorize_impl<double>();
    else
        factorize_impl<float>();
    state_ = State::Factorized;
}

// Host-side records of the solve ops of one level, turned into mv phases at its end.
struct TrsvTask {  // ScalingLLT / ScalingPLUQ of one cluster
    const void* L;   // forward factor (explicit inverse, lower)
    const void* Lb;  // backward factor (lower, applied transposed)
    const int* perm;
    int n, xoff;
};
struct GemvGroup { int xdst, n, k0, k1; };
struct GemvTerm {
    const void* A;
    int rows, cols, xsrc, trans;
};

// One mv phase on the host: groups and terms, then split into chunks and pieces.
struct MvPlan {
    std::vector<MvGroup> groups;
    std::vector<MvTerm> terms;
    std::vector<MvChunk> chunks;
    std::vector<MvPiece> pieces;
    std::vector<int> rchunks;
    int family = KF_GEMV;
    bool inplace = false;  // destinations are also sources: every chunk goes through partials
    SweepWork work;

    void split(size_t esz) {
        const int W = MV_PIECE_WORK;
        for (int gi = 0; gi < (int)groups.size(); gi++) {
            const MvGroup& g = groups[gi];
            for (int o0 = 0; o0 < g.n; o0 += SOLVE_CHUNK) {
                const int no = std::min(SOLVE_CHUNK, g.n - o0);
                const int ci = (int)chunks.size(), p0 = (int)pieces.size();
                int pk0 = -1, pwork = 0;
                auto flush = [&](int kend) {
                    if (pk0 >= 0) {
                        const MvTerm& t = terms[pk0];
                        pieces.push_back({ci, pk0, kend, 0, t.trans ? t.rows : t.cols, 0, 0, 0, 0, 0});
                    }
                    pk0 = -1;
                    pwork = 0;
                };
                for (int k = g.k0; k < g.k1; k++) {
                    const MvTerm& t = terms[k];
                    int lo = 0, hi = t.trans ? t.rows : t.cols;
                    if (t.tri == 1) hi = std::min(hi, o0 + SOLVE_CHUNK);
                    if (t.tri == 2) lo = std::max(lo, o0);
                    const int len = std::max(0, hi - lo);
                    work.mat_bytes += (double)no * len * esz;
                    work.flops += 2.0 * no * len;
                    work.vec_elems += len;
                    if (len > W) {  // a long term alone: pieces of W inner indices
                        flush(k);
                        for (int j = lo; j < hi; j += W)
                            pieces.push_back({ci, k, k + 1, j, std::min(hi, j + W), 0, 0, 0, 0, 0});
                    } else {
                        // short terms share a piece: at most W inner indices and
                        // MV_PIECE_TERMS terms (one lane of a wavefront reads each record)
                        if (pk0 >= 0 && (pwork + len > W || k - pk0 >= MV_PIECE_TERMS)) flush(k);
                        if (pk0 < 0) pk0 = k;
                        pwork += len;
                    }
                }
                flush(g.k1);
                const int p1 = (int)pieces.size();
                if (p1 == p0) continue;
                chunks.push_back({gi, o0, p0, p1});
                for (int q = p0; q < p1; q++) {
                    pieces[q].o0 = o0;
                    pieces[q].no = no;
                    pieces[q].xdst = g.xdst;
                    pieces[q].mode = g.mode;
                }
                work.vec_elems += 2.0 * no;
                if (p1 - p0 == 1 && !inplace && !g.dperm) pieces[p0].direct = 1;
                else rchunks.push_back(ci);
            }
        }
    }
};

// Stable grouping of n items by key (keys in [0, nkeys)): groups in order of first
// appearance, items in their original order inside a group. Returns the item order and
// fills start (group g is order[start[g] .. start[g + 1])) and the group keys.
template <typename KeyOf>
static std::vector<int> group_by_first(int n, int nkeys, KeyOf key, std::vector<int>& start, std::vector<int>& keys) {
    std::vector<int> slot(nkeys, -1), cnt;
    keys.clear();
    for (int k = 0; k < n; k++) {
        int& sl = slot[key(k)];
        if (sl < 0) {
            sl = (int)keys.size();
            keys.push_back(key(k));
            cnt.push_back(0);
        }
        cnt[sl]++;
    }
    start.assign(keys.size() + 1, 0);
    for (size_t g = 0; g < keys.size(); g++) start[g + 1] = start[g] + cnt[g];
    std::vector<int> pos(start.begin(), start.end() - 1), order(n);
    for (int k = 0; k < n; k++) order[pos[slot[key(k)]]++] = k;
    return order;
}

// Triangles of at most TRSM_MFMA_MIN_N rows first (the vector kernel: they have no trailing
// update to speak of), longer ones after (FP32: the matrix-core kernel); returns the split.
static int split_trsm(std::vector<TrsmTask>& v) {
    auto it = std::stable_partition(v.begin(), v.end(), [](const TrsmTask& t) { return t.n <= TRSM_MFMA_MIN_N; });
    return (int)(it - v.begin());
}

static inline double potrf_flops(double n) { return n * n * n / 3.0; }
static inline double getrf_flops(double n) { return 2.0 * n * n * n / 3.0; }

// Splits a triangular solve into independent chunks of at most 64 columns (left solve,
// block n x m, ld n) or 64 rows (right solve, block m x n, ld m); 32 / 16 for n >= 128 / 256.
// A left solve may read its right-hand side rows permuted from another block (src, perm).
static void push_trsm(std::vector<TrsmTask>& v, const void* L, void* B, int n, int m, int side, size_t esz,
                      const void* src = nullptr, const int* perm = nullptr, int cw_policy = 0) {
    if (n == 0 || m == 0) return;
    // long triangles walk many row blocks per chunk: narrower chunks give more workgroups
    const int cw = trsm_chunk_width(cw_policy, n);
    for (int c0 = 0; c0 < m; c0 += cw) {
        const int w = std::min(cw, m - c0);
        const size_t off = (side == 0 ? (size_t)c0 * n : (size_t)c0) * esz;
        char* b = reinterpret_cast<char*>(B) + off;
        const char* sb = src ? reinterpret_cast<const char*>(src) + off : nullptr;
        v.push_back({L, b, n, w, side, side == 0 ? n : m, sb, perm});
    }
}

// Algorithmic work of one launch over a task array: every block touched is read (and
// written, if updated) once, in the factor's element type; flops by the LAPACK counts.
struct Work {
    double bytes = 0, flops = 0;
};
static Work work_potrf(const std::vector<PotrfTask>& v, size_t esz) {
    Work w;
    for (auto& t : v) {
        w.bytes += (double)t.n * (t.n + 1) * esz;
        w.flops += potrf_flops(t.n);
    }
    return w;
}
static Work work_trsm(const std::vector<TrsmTask>& v, size_t esz) {
    Work w;
    for (auto& t : v) {
        w.bytes += ((double)t.n * (t.n + 1) / 2 + 2.0 * t.n * t.m) * esz;
        w.flops += (double)t.m * t.n * t.n;
    }
    return w;
}
static Work work_schur(const std::vector<SchurTile>& tiles, const std::vector<SchurContrib>& cs, size_t esz) {
    Work w;
    for (auto& t : tiles) {
        w.bytes += 2.0 * t.tm * t.tn * esz;
        for (int k = t.k0; k < t.k1; k++) {
            w.bytes += (double)(t.tm + t.tn) * cs[k].k * esz;
            w.flops += 2.0 * t.tm * t.tn * cs[k].k;
        }
    }
    return w;
}
static Work work_copy(const std::vector<CopyTask>& v, size_t esz) {
    Work w;
    for (auto& t : v) w.bytes += (t.src ? 2.0 : 1.0) * t.rows * t.cols * esz;
    return w;
}
static double geqp3_flops(double m, double n) {
    const double k = std::min(m, n);
    return k > 0 ? 2.0 * m * n * k - (2.0 / 3.0) * k * k * k : 0.0;
}

// Stepped RRQR launches queued ahead of the device: enough to keep it busy while the host
// launches the next one, few enough that launches past every cluster's rank stay rare.
constexpr int RRQR_STEPS_AHEAD = 12;

// Clusters with at least 16 rows and more than 10000 Asn elements take the stepped
// multi-workgroup RRQR; smaller ones run in one workgroup. Both paths compute the same
// numbers; SPAND_RRQR_STEPPED_MIN_ROWS / SPAND_RRQR_STEPPED_MIN_ELEMS override the split.
static long long env_or(const char* name, long long dflt) {
    const char* e = std::getenv(name);
    return (e && *e) ? std::strtoll(e, nullptr, 10) : dflt;
}

// One Orthogonal op of the solve: y = -V^T x, z = -T^T y (forward) or -T y (backward),
// x -= V z, with X = T^T stored (compact WY over its kept reflectors).
struct QOp {
    const void* Ve;
    const void* X;
    int m, k, xoff, yoff;
};

// One level's solve-plan inputs, kept on the host for the transposed phases (built and
// uploaded at the first transposed solve).
struct HostPhases {
    std::vector<TrsvTask> sv_elim, sv_scale;
    std::vector<GemvGroup> fwd_groups, bwd_groups;
    std::vector<GemvTerm> fwd_t, bwd_terms;
    std::vector<QOp> qops;
};

// Host side of one level's preconditioner-apply plan, before upload.
struct SolvePlanHost {
    std::vector<MvPlan> fwd, bwd;
    HostPhases inputs;
    std::vector<MergeTask> mt;
    std::vector<int2> mchildren;
    long long max_pieces = 0;
};

// Phase builders. A triangular sweep (ScalingLLT / ScalingPLUQ of each cluster): forward
// x := L x with the gathered source P^T x, backward x := Lb^T x.
static MvPlan trsv_phase(const std::vector<TrsvTask>& v, bool backward) {
    MvPlan m;
    m.family = KF_TRMV;
    m.inplace = true;
    for (auto& t : v) {
        const int k0 = (int)m.terms.size();
        if (!backward) m.terms.push_back({t.L, t.perm, t.n, t.n, t.xoff, 0, 1, 0});
        else m.terms.push_back({t.Lb, nullptr, t.n, t.n, t.xoff, 1, 2, 0});
        m.groups.push_back({t.xoff, t.n, k0, k0 + 1, MV_SET, 0, nullptr});
    }
    return m;
}

// x_dst -= sum_k op(A_k) x_src(k) for every group
static MvPlan gemv_phase(const std::vector<GemvGroup>& gs, const std::vector<GemvTerm>& ts) {
    MvPlan m;
    m.family = KF_GEMV;
    for (auto& g : gs) {
        const int k0 = (int)m.terms.size();
        for (int k = g.k0; k < g.k1; k++)
            m.terms.push_back({ts[k].A, nullptr, ts[k].rows, ts[k].cols, ts[k].xsrc, ts[k].trans, 0, 0});
        m.groups.push_back({g.xdst, g.n, k0, (int)m.terms.size(), MV_SUB, 0, nullptr});
    }
    return m;
}

// Orthogonal ops, step 0..2: y = -V^T x, z = -X y (forward) or -X^T y (backward), x -= V z
static MvPlan q_phase(const std::vector<QOp>& qops, int step, bool backward) {
    MvPlan m;
    m.family = KF_QAPPLY;
    for (auto& q : qops) {
        const int k0 = (int)m.terms.size();
        if (step == 0) {
            m.terms.push_back({q.Ve, nullptr, q.m, q.k, q.xoff, 1, 2, 0});
            m.groups.push_back({q.yoff, q.k, k0, k0 + 1, MV_SETNEG, 0, nullptr});
        } else if (step == 1) {  // X = T^T lower: z = -X y (fwd), -X^T y (bwd)
            m.terms.push_back({q.X, nullptr, q.k, q.k, q.yoff, backward ? 1 : 0, backward ? 2 : 1, 0});
            m.groups.push_back({q.yoff + q.k, q.k, k0, k0 + 1, MV_SETNEG, 0, nullptr});
        } else {
            m.terms.push_back({q.Ve, nullptr, q.m, q.k, q.yoff + q.k, 0, 1, 0});
            m.groups.push_back({q.xoff, q.m, k0, k0 + 1, MV_SUB, 0, nullptr});
        }
    }
    return m;
}

// The transpose of trsv_phase: (L P^T)^T = P L^T (P as a permuted destination), (Lb^T)^T = Lb.
static MvPlan trsv_phase_t(const std::vector<TrsvTask>& v, bool backward) {
    MvPlan m;
    m.family = KF_TRMV;
    m.inplace = true;
    for (auto& t : v) {
        const int k0 = (int)m.terms.size();
        if (!backward) {
            m.terms.push_back({t.L, nullptr, t.n, t.n, t.xoff, 1, 2, 0});
            m.groups.push_back({t.xoff, t.n, k0, k0 + 1, MV_SET, 0, t.perm});
        } else {
            m.terms.push_back({t.Lb, nullptr, t.n, t.n, t.xoff, 0, 1, 0});
            m.groups.push_back({t.xoff, t.n, k0, k0 + 1, MV_SET, 0, nullptr});
        }
    }
    return m;
}

// The transpose of gemv_phase: x_src(k) -= op(A_k)^T x_dst, grouped by source segment.
static MvPlan gemv_phase_t(const std::vector<GemvGroup>& gs, const std::vector<GemvTerm>& ts) {
    struct TT {
        int dst, n;
        GemvTerm t;
    };
    std::vector<TT> items;
    for (auto& g : gs)
        for (int k = g.k0; k < g.k1; k++) {
            const GemvTerm& t = ts[k];
            const int n_in = t.trans ? t.rows : t.cols;  // op(A)'s input length
            items.push_back({t.xsrc, n_in, GemvTerm{t.A, t.rows, t.cols, g.xdst, t.trans ? 0 : 1}});
        }
    std::stable_sort(items.begin(), items.end(), [](const TT& a, const TT& b) { return a.dst < b.dst; });
    MvPlan m;
    m.family = KF_GEMV;
    for (size_t q = 0; q < items.size();) {
        const int k0 = (int)m.terms.size();
        const int dst = items[q].dst, n = items[q].n;
        for (; q < items.size() && items[q].dst == dst; q++) {
            const GemvTerm& t = items[q].t;
            m.terms.push_back({t.A, nullptr, t.rows, t.cols, t.xsrc, t.trans, 0, 0});
        }
        m.groups.push_back({dst, n, k0, (int)m.terms.size(), MV_SUB, 0, nullptr});
    }
    return m;
}

// The level's mv phases: forward {trsv Elim, gemv, trsv Scale, Q 0..2}, backward
// {Q 0..2, trsv Scale', gemv, trsv Elim'}, each split into chunks and pieces. The inputs
// are kept for the transposed phases.
static SolvePlanHost build_solve_plan(size_t esz, HostPhases in, std::vector<MergeTask> mt,
                                      std::vector<int2> mchildren) {
    SolvePlanHost h;
    h.fwd.push_back(trsv_phase(in.sv_elim, false));
    h.fwd.push_back(gemv_phase(in.fwd_groups, in.fwd_t));
    h.fwd.push_back(trsv_phase(in.sv_scale, false));
    for (int q = 0; q < 3; q++) h.fwd.push_back(q_phase(in.qops, q, false));
    for (int q = 0; q < 3; q++) h.bwd.push_back(q_phase(in.qops, q, true));
    h.bwd.push_back(trsv_phase(in.sv_scale, true));
    h.bwd.push_back(gemv_phase(in.bwd_groups, in.bwd_terms));
    h.bwd.push_back(trsv_phase(in.sv_elim, true));
    for (auto* v : {&h.fwd, &h.bwd})
        for (auto& m : *v) {
            m.split(esz);
            h.max_pieces = std::max(h.max_pieces, (long long)m.pieces.size());
        }
    h.mt = std::move(mt);
    h.mchildren = std::move(mchildren);
    h.inputs = std::move(in);
    return h;
}

// The adjoint apply's phases of one level. Forward level l applies F = Q S G E (E first),
// backward B = E' G' S' Q' (Q' first); M^{-T} runs B^T (E'^T, G'^T, S'^T, Q'^T = the forward
// Q) going up and F^T (Q^T = the backward Q, S^T, G^T, E^T) coming down.
static void build_transposed(size_t esz, const HostPhases& in, std::vector<MvPlan>& up, std::vector<MvPlan>& down,
                             long long& max_pieces) {
    up.push_back(trsv_phase_t(in.sv_elim, true));
    up.push_back(gemv_phase_t(in.bwd_groups, in.bwd_terms));
    up.push_back(trsv_phase_t(in.sv_scale, true));
    for (int q = 0; q < 3; q++) up.push_back(q_phase(in.qops, q, false));
    for (int q = 0; q < 3; q++) down.push_back(q_phase(in.qops, q, true));
    down.push_back(trsv_phase_t(in.sv_scale, false));
    down.push_back(gemv_phase_t(in.fwd_groups, in.fwd_t));
    down.push_back(trsv_phase_t(in.sv_elim, false));
    for (auto* v : {&up, &down})
        for (auto& m : *v) {
            m.split(esz);
            max_pieces = std::max(max_pieces, (long long)m.pieces.size());
        }
}

// Uploads phase lists (one Packer) and returns their device descriptors, in order.
static std::vector<std::vector<MvPhase>> upload_phases(Tree& tree, std::vector<std::vector<MvPlan>*> lists,
                                                       void* (Tree::*upload)(const Packer&), Packer& ps,
                                                       void** base) {
    std::vector<std::array<size_t, 5>> offs;
    for (auto* v : lists)
        for (auto& m : *v)
            offs.push_back({ps.add(m.groups), ps.add(m.terms), ps.add(m.chunks), ps.add(m.pieces), ps.add(m.rchunks)});
    void* dsv = (tree.*upload)(ps);
    *base = dsv;
    std::vector<std::vector<MvPhase>> out;
    size_t oi = 0;
    for (auto* v : lists) {
        out.emplace_back();
        for (auto& m : *v) {
            const auto& o = offs[oi++];
            if (m.pieces.empty()) continue;
            MvPhase ph;
            ph.groups = Packer::at<MvGroup>(dsv, o[0]);
            ph.terms = Packer::at<MvTerm>(dsv, o[1]);
            ph.chunks = Packer::at<MvChunk>(dsv, o[2]);
            ph.pieces = Packer::at<MvPiece>(dsv, o[3]);
            ph.rchunks = Packer::at<int>(dsv, o[4]);
            ph.npieces = (int)m.pieces.size();
            ph.nreduce = (int)m.rchunks.size();
            ph.family = m.family;
            ph.w = m.work;
            out.back().push_back(ph);
        }
    }
    return out;
}

// Uploads one level's solve plan and points the level's phases at the device copy. A GEN
// tree keeps the phase inputs on the host until a transposed solve asks for the transposed
// phases; an SPD tree's transposed apply is its forward apply, and it keeps nothing.
void Tree::upload_solve_plan(int level, void* host_plan) {
    SolvePlanHost& h = *static_cast<SolvePlanHost*>(host_plan);
    LevelSolve& lsv = solve_plan_[level];
    Packer ps;
    const size_t s_m = ps.add(h.mt), s_mc = ps.add(h.mchildren);
    void* dsv = nullptr;
    auto ph = upload_phases(*this, {&h.fwd, &h.bwd}, &Tree::upload_tasks, ps, &dsv);
    lsv.fwd = std::move(ph[0]);
    lsv.bwd = std::move(ph[1]);
    for (auto& c : h.mchildren) lsv.w_merge.vec_elems += 2.0 * c.y;
    lsv.merge = Packer::at<MergeTask>(dsv, s_m);
    lsv.n_merge = (int)h.mt.size();
    lsv.children = Packer::at<int2>(dsv, s_mc);
    max_pieces_ = std::max(max_pieces_, h.max_pieces);
    if (fsymm_ != 0) tplans_[level].reset(new HostPhases(std::move(h.inputs)));
}

// Builds and uploads the transposed phases of every level (first transposed solve after a
// factorization). The partial-sum area grows with them if needed.
void Tree::upload_transposed() {
    const size_t esz = fprec_ == 0 ? sizeof(double) : sizeof(float);
    bool any = false;
    for (int l = 0; l < nlevels_; l++) {
        if (!tplans_[l]) continue;
        std::vector<MvPlan> up, down;
        build_transposed(esz, *tplans_[l], up, down, max_pieces_);
        Packer ps;
        void* dsv = nullptr;
        auto ph = upload_phases(*this, {&up, &down}, &Tree::upload_tasks, ps, &dsv);
        solve_plan_[l].tfwd = std::move(ph[0]);
        solve_plan_[l].tbwd = std::move(ph[1]);
        tplans_[l].reset();
        any = true;
    }
    // the uploads run on the null stream; the caller's stream may not wait for it
    if (any) hip_check(hipStreamSynchronize(0), "upload transposed phases");
}

// ---------------------------------------------------------------------------------------
// The factorization (src/tree.cpp:1447-1551), level by level. Each level runs the phases
// below in order; they share the level's task lists through LevelPlan.

// One factorization's state shared by its levels.
struct FactorRun {
    hipStream_t st = 0;  // the library's calls are ordered on the null stream
    bool gen = false;    // GEN / PLU tree
    long long lds_elems = 0, big_min_rows = 0;  // RRQR path limits
    long long spin_max = RRQR_SPIN_MAX;         // polls of a persistent RRQR wait before it gives up
    std::vector<int> hsize;                     // cluster sizes read back from the device
    std::vector<std::future<SolvePlanHost>> pending_solve;  // per level; joined on every exit
};

// A Schur contribution and its target edge, in recorded order.
struct SchurPending {
    int target;
    SchurContrib c;
};

// One level's plan: the task lists each phase builds and the later phases launch or record.
template <typename T>
struct LevelPlan {
    FactorRun* run = nullptr;
    LevelStats* lst = nullptr;
    LevelSolve* lsv = nullptr;
    const std::vector<int>* bot = nullptr;  // the current bottom list
    double t_lvl0 = 0, t_plan1 = 0, t_lvl1 = 0, t_lvl2 = 0;
    // elimination and scaling
    std::vector<PotrfTask> potrfE, potrfS;
    std::vector<GetrfTask> getrfE, getrfS;
    std::vector<TrsmTask> trsmE, trsmA, trsmB;
    std::vector<SchurPending> pcs;
    std::vector<int> fill_edges;
    long long fill_elems = 0;
    std::vector<long long> fill_off;  // by position in fill_edges
    std::vector<TrsvTask> sv_elim, sv_scale;
    std::vector<CopyTask> inv_init;  // identity blocks that become explicit inverses L^{-1}
    std::vector<std::pair<int, GemvTerm>> fwd_list;  // (destination cluster, term), in op order
    std::vector<GemvGroup> bwd_groups;
    std::vector<GemvTerm> bwd_terms;
    LevelShape shape;  // the level's size histogram features and launch configuration
    LevelTune tune;
    int cwp = 0;       // trsm chunk-width policy
    T* fill_base = nullptr;
    std::vector<SchurTile> tiles;
    std::vector<SchurContrib> contribs;
    size_t n_potrf_elim = 0;
    // sparsification
    bool do_sparsify_level = false;
    std::vector<int> live;  // clusters above the level (scaled and sparsified)
    std::vector<int> plist, wave, widx;
    int nwaves = 0;
    std::vector<std::vector<RrqrTask>> wtasks;   // one-workgroup LDS path, per wave
    std::vector<long long> wlds_small;           // dynamic LDS of each wave's small-path launch
    std::vector<std::vector<RrqrBig>> wbig;      // stepped path, per wave
    std::vector<std::vector<RrqrSlab>> wslab;
    std::vector<int> wsteps, wmaxrows;
    std::vector<std::vector<RrqrTask>> wreg;     // register path, per wave (sorted by variant)
    std::vector<std::vector<int>> wregvar;       // the variant of each wreg task
    std::vector<std::pair<int, int>> where;      // (0 small | 1 big | 2 register, index) by plist position
    std::vector<RrqrSeg> segs;
    std::vector<size_t> wlds_step, wlds_wave;
    bool persistent = false;
    int step_waves = 0;
    std::vector<size_t> wflag0;
    // kept ranks and merge
    std::vector<QOp> qops;
    std::vector<WyTask> wy;
    std::vector<int4> wy_tiles;
    std::vector<TrsmTask> wy_trsm;
    std::vector<MergeTask> mt;
    std::vector<int2> mchildren;
    int left = 0;  // dofs left after the level
};

// Cholesky batch order: <= 64, <= POTRF_SMALL_MAX, larger
static void small_first(std::vector<PotrfTask>& v) {
    std::stable_partition(v.begin(), v.end(), [](const PotrfTask& t) { return t.n <= POTRF_SMALL_MAX; });
    std::stable_partition(v.begin(), v.end(), [](const PotrfTask& t) { return t.n <= 64; });
}
// LU pivots that fit LDS first (one launch), the larger ones after (another)
template <typename T>
static void lds_first(std::vector<GetrfTask>& v) {
    std::stable_partition(v.begin(), v.end(), [](const GetrfTask& t) {
        return (size_t)t.n * (t.n + 1) * sizeof(T) <= (size_t)GETRF_LDS_BYTES;
    });
}

// Cholesky pivots of one batch: up to POTRF_SMALL_MAX in one launch (one workgroup each;
// 64 and below in their own), larger ones (sorted last) one 32-column panel per launch pair.
template <typename T>
void Tree::run_potrf(hipStream_t st, void* dt, const std::vector<PotrfTask>& v, size_t off, const char* what) {
    const size_t esz = sizeof(T);
    if (v.empty()) return;
    int ntiny = 0, nsmall = 0, nmax = 0;
    for (auto& t : v) {
        if (t.n <= 64) ntiny++;
        if (t.n <= POTRF_SMALL_MAX) nsmall++;
        nmax = std::max(nmax, t.n);
    }
    const PotrfTask* d = Packer::at<PotrfTask>(dt, off);
    if (ntiny > 0) {
        std::vector<PotrfTask> vs(v.begin(), v.begin() + ntiny);
        Work w = work_potrf(vs, esz);
        hip_check(timed(prof_, KF_POTRF, st, w.bytes, w.flops,
                        [&] { return launch_potrf64<T>(d, ntiny, d_info_, st); }),
                  what);
    }
    if (nsmall > ntiny) {
        std::vector<PotrfTask> vs(v.begin() + ntiny, v.begin() + nsmall);
        Work w = work_potrf(vs, esz);
        hip_check(timed(prof_, KF_POTRF, st, w.bytes, w.flops,
                        [&] { return launch_potrf<T>(d + ntiny, nsmall - ntiny, d_info_, st); }),
                  what);
    }
    const int nbig = (int)v.size() - nsmall;
    if (nbig == 0) return;
    const PotrfTask* db = d + nsmall;
    for (int kb = 0; kb < nmax; kb += 32) {
        double pb = 0, pf = 0, tb = 0, tf = 0;
        for (int k = nsmall; k < (int)v.size(); k++) {
            const double n = v[k].n, nb = std::min(32.0, n - kb), rest = n - kb - nb;
            if (nb <= 0) continue;
            pb += (nb * (nb + 1) + 2.0 * rest * nb) * esz;
            pf += nb * nb * nb / 3.0 + rest * nb * nb;
            tb += (rest * (rest + 1) + 2.0 * rest * nb) * esz;
            tf += rest * (rest + 1) * nb;
        }
        hip_check(timed(prof_, KF_POTRF_PANEL, st, pb, pf,
                        [&] { return launch_potrf_panel<T>(db, nbig, kb, nmax, d_info_, st); }),
                  what);
        hip_check(timed(prof_, KF_POTRF_TRAIL, st, tb, tf,
                        [&] { return launch_potrf_trail<T>(db, nbig, kb, nmax, st); }),
                  what);
    }
}

// One trsm batch (split: the first task that may take the matrix-core kernel).
template <typename T>
void Tree::run_trsm(hipStream_t st, void* dt, const std::vector<TrsmTask>& v, size_t off, int split, const char* what) {
    const size_t esz = sizeof(T);
    if (v.empty()) return;
    Work w = work_trsm(v, esz);
    hip_check(timed(prof_, KF_TRSM, st, w.bytes, w.flops,
                    [&] { return launch_trsm<T>(Packer::at<TrsmTask>(dt, off), (int)v.size(), split, st); }),
              what);
}

// LU pivots of one batch: those that fit LDS in one launch, the larger ones in another.
template <typename T>
void Tree::run_getrf(hipStream_t st, void* dt, const std::vector<GetrfTask>& v, size_t off, const char* what) {
    const size_t esz = sizeof(T);
    if (v.empty()) return;
    int nl = 0, nmax_l = 0;
    double lb = 0, lf = 0, gb = 0, gf = 0;
    for (auto& t : v) {
        const double n = t.n;
        const double by = (3.0 * n * n) * esz + 4.0 * n, fl = getrf_flops(n);
        if ((size_t)t.n * (t.n + 1) * esz <= (size_t)GETRF_LDS_BYTES) {
            nl++;
            nmax_l = std::max(nmax_l, t.n);
            lb += by;
            lf += fl;
        } else {
            gb += by;
            gf += fl;
        }
    }
    const GetrfTask* d = Packer::at<GetrfTask>(dt, off);
    if (nl > 0)
        hip_check(timed(prof_, KF_GETRF, st, lb, lf,
                        [&] { return launch_getrf<T>(d, nl, 1, nmax_l, d_info_, st); }),
                  what);
    if (nl < (int)v.size())
        hip_check(timed(prof_, KF_GETRF, st, gb, gf,
                        [&] { return launch_getrf<T>(d + nl, (int)v.size() - nl, 0, 0, d_info_, st); }),
                  what);
}

// Elimination plan of level ilvl_ (eliminate_cluster of every cluster of the level,
// src/tree.cpp:1494-1506, 895-967): pivots, panels, Schur contributions grouped into target
// tiles, fill-in blocks, the ops and the solve terms.
template <typename T>
void Tree::plan_elimination(LevelPlan<T>& L) {
    FactorRun& R = *L.run;
    const bool gen = R.gen;
    LevelStats& lst = *L.lst;
    const std::vector<int>& bot = *L.bot;
    auto& potrfE = L.potrfE;
    auto& getrfE = L.getrfE;
    auto& trsmE = L.trsmE;
    auto& pcs = L.pcs;
    auto& fill_edges = L.fill_edges;
    auto& fill_elems = L.fill_elems;
    auto& fill_off = L.fill_off;
    auto& sv_elim = L.sv_elim;
    auto& inv_init = L.inv_init;
    auto& fwd_list = L.fwd_list;
    auto& bwd_groups = L.bwd_groups;
    auto& bwd_terms = L.bwd_terms;
    auto& shape = L.shape;
    auto& tune = L.tune;
    auto& fill_base = L.fill_base;
    auto& tiles = L.tiles;
    auto& contribs = L.contribs;
    auto push_trsm = [&L](std::vector<TrsmTask>& v, const void* A, void* B, int n, int m, int side, size_t es,
                          const void* src = nullptr, const int* perm = nullptr) {
        spand::push_trsm(v, A, B, n, m, side, es, src, perm, L.cwp);
    };
    auto target_edge = [&](int n1, int n2) {
        int e = find_out_edge(n2, n1);
        if (e < 0) {
            e = new_edge(n2, n1, nullptr, false);
            fill_edges.push_back(e);
            fill_off.push_back(fill_elems);
            fill_elems += (long long)size(n1) * size(n2);
        }
        return e;
    };
    // the level's launch configuration: trsm chunks from its pivot sizes now, the
    // sparsification knobs once the sparsified clusters' Asn shapes are known (below)
    {
        std::vector<int> piv;
        for (int c : bot)
            if (cls_[c].level == ilvl_) piv.push_back(cls_[c].size);
        shape.n_elim = (int)piv.size();
        if (!piv.empty()) {
            std::sort(piv.begin(), piv.end());
            shape.elim_med = piv[piv.size() / 2];
            shape.elim_max = piv.back();
        }
    }
    tune = choose_tune(sizeof(T), shape);
    apply_tune_env(tune);
    L.cwp = tune.trsm_cw;
    for (int c : bot) {
        if (cls_[c].level != ilvl_) continue;
        HCluster& s = cls_[c];
        const int n = s.size;
        // edges_ may grow (fill-in) below: keep the pivot block pointer, not a reference
        const HEdge piv = edges_[s.out[0]];
        if (piv.ident) throw Error(ERR_STATE, "internal: identity pivot on an eliminated cluster");
        if (gen) {
            // getf_cluster + panel_getf + gemm_edges, PLU branch (src/tree.cpp:929-956)
            void* Ut = arena_.alloc((size_t)std::max(1, n * n) * sizeof(T));
            int* pp = reinterpret_cast<int*>(arena_.alloc((size_t)std::max(1, n) * sizeof(int)));
            getrfE.push_back({piv.A, Ut, pp, n, c});
            lst.flops_potrf += getrf_flops(n);
            log_shape(SK_GETF, n, 0, 0);
            std::vector<int> out(s.out.begin() + 1, s.out.end()), in(s.in.begin(), s.in.end());
            for (int e : in) {  // A_sn := L^{-1} P^T A_sn, into a fresh block
                const int m = size(edges_[e].n1);
                log_shape(SK_PANEL, m, n, 0);
                void* nb = arena_.alloc((size_t)std::max(1, n * m) * sizeof(T));
                push_trsm(trsmE, piv.A, nb, n, m, 0, sizeof(T), edges_[e].A, pp);
                edges_[e].A = nb;
                lst.flops_trsm += (double)n * n * m;
            }
            for (int e : out) {  // A_ns := A_ns U^{-1} = A_ns (U^T)^{-T}
                log_shape(SK_PANEL, size(edges_[e].n2), n, 0);
                push_trsm(trsmE, Ut, edges_[e].A, n, size(edges_[e].n2), 1, sizeof(T));
                lst.flops_trsm += (double)n * n * size(edges_[e].n2);
            }
            for (int e1 : out)
                for (int e2 : in) {
                    int n1 = edges_[e1].n2, n2 = edges_[e2].n1;
                    int tgt = target_edge(n1, n2);
                    SchurContrib sc{edges_[e1].A, edges_[e2].A, n, size(n1), n, 0, 0, 0};
                    pcs.push_back({tgt, sc});
                    lst.flops_schur += 2.0 * size(n1) * size(n2) * n;
                    log_shape(SK_GEMM, size(n1), size(n2), n);
                }
            // record ScalingPLUQ + GemmOut / GemmIn (src/tree.cpp:945-951)
            ops_.push_back({OP_SCALING_PLUQ, c, -1, n, 0, n, n, ilvl_, piv.A, nullptr, Ut, pp});
            if (n > 0) {
                void* Li = arena_.alloc((size_t)n * n * sizeof(T));
                void* Wi = arena_.alloc((size_t)n * n * sizeof(T));
                inv_init.push_back({nullptr, Li, n, n, n, 0});
                inv_init.push_back({nullptr, Wi, n, n, n, 0});
                push_trsm(trsmE, piv.A, Li, n, n, 0, sizeof(T));
                push_trsm(trsmE, Ut, Wi, n, n, 0, sizeof(T));
                lst.flops_trsm += 2.0 * n * n * n;
                sv_elim.push_back({Li, Wi, pp, n, (int)s.xoff});
            }
            factor_bytes_ += (size_t)n * (n + 1) * sizeof(T) + (size_t)n * sizeof(int);
            std::vector<GemvTerm> mine;
            for (int e : out) {
                int nb = edges_[e].n2;
                ops_.push_back({OP_GEMM_OUT, c, nb, n, size(nb), size(nb), n, ilvl_, edges_[e].A, nullptr});
                factor_bytes_ += (size_t)n * size(nb) * sizeof(T);
                if (n > 0 && size(nb) > 0) {
                    fwd_list.push_back({nb, GemvTerm{edges_[e].A, size(nb), n, (int)s.xoff, 0}});
                }
            }
            for (int e : in) {
                int nb = edges_[e].n1;
                ops_.push_back({OP_GEMM_IN, c, nb, n, size(nb), n, size(nb), ilvl_, edges_[e].A, nullptr});
                factor_bytes_ += (size_t)n * size(nb) * sizeof(T);
                if (n > 0 && size(nb) > 0) mine.push_back({edges_[e].A, n, size(nb), (int)cls_[nb].xoff, 0});
            }
            if (!mine.empty()) {
                int k0 = (int)bwd_terms.size();
                for (auto it = mine.rbegin(); it != mine.rend(); ++it) bwd_terms.push_back(*it);
                bwd_groups.push_back({(int)s.xoff, n, k0, (int)bwd_terms.size()});
            }
            set_eliminated(c);
            continue;
        }
        potrfE.push_back({piv.A, n, c, nullptr});
        lst.flops_potrf += potrf_flops(n);
        log_shape(SK_POTF, n, 0, 0);
        std::vector<int> out(s.out.begin() + 1, s.out.end()), in(s.in.begin(), s.in.end());
        for (int e : in) {
            push_trsm(trsmE, piv.A, edges_[e].A, n, size(edges_[e].n1), 0, sizeof(T));
            lst.flops_trsm += (double)n * n * size(edges_[e].n1);
            log_shape(SK_PANEL, size(edges_[e].n1), n, 0);
        }
        for (int e : out) {
            push_trsm(trsmE, piv.A, edges_[e].A, n, size(edges_[e].n2), 1, sizeof(T));
            lst.flops_trsm += (double)n * n * size(edges_[e].n2);
            log_shape(SK_PANEL, size(edges_[e].n2), n, 0);
        }
        // schur_symmetric (src/tree.cpp:862-883) with gemm_edges (:748-793)
        for (int e1 : out)
            for (int e2 : out) {
                int n1 = edges_[e1].n2, n2 = edges_[e2].n2;
                if (cls_[n1].order < cls_[n2].order) continue;
                int tgt = target_edge(n1, n2);
                SchurContrib sc{edges_[e1].A, edges_[e2].A, n, size(n1), size(n2), 0, 1, 0};
                pcs.push_back({tgt, sc});
                lst.flops_schur += (n1 == n2 ? 1.0 : 2.0) * size(n1) * size(n2) * n;
                log_shape(n1 == n2 ? SK_SYRK : SK_GEMM, size(n1), size(n2), n);
            }
        for (int e1 : in)
            for (int e2 : in) {
                int n1 = edges_[e1].n1, n2 = edges_[e2].n1;
                if (cls_[n1].order < cls_[n2].order) continue;
                int tgt = target_edge(n1, n2);
                SchurContrib sc{edges_[e1].A, edges_[e2].A, n, n, n, 1, 0, 0};
                pcs.push_back({tgt, sc});
                lst.flops_schur += (n1 == n2 ? 1.0 : 2.0) * size(n1) * size(n2) * n;
                log_shape(n1 == n2 ? SK_SYRK : SK_GEMM, size(n1), size(n2), n);
            }
        for (int e1 : out)
            for (int e2 : in) {
                int n1 = edges_[e1].n2, n2 = edges_[e2].n1;
                int tgt = target_edge(n1, n2);
                SchurContrib sc{edges_[e1].A, edges_[e2].A, n, size(n1), n, 0, 0, 0};
                pcs.push_back({tgt, sc});
                lst.flops_schur += 2.0 * size(n1) * size(n2) * n;
                log_shape(n1 == n2 ? SK_SYRK : SK_GEMM, size(n1), size(n2), n);
            }
        // record ScalingLLT + GemmSymmOut/In (src/tree.cpp:909-910, 885-892)
        ops_.push_back({OP_SCALING_LLT, c, -1, n, 0, n, n, ilvl_, piv.A, nullptr});
        if (n > 0) {
            // explicit inverse for the solve: L^{-1} = L^{-1} I by the chunked trsm
            void* Li = arena_.alloc((size_t)n * n * sizeof(T));
            inv_init.push_back({nullptr, Li, n, n, n, 0});
            push_trsm(trsmE, piv.A, Li, n, n, 0, sizeof(T));
            lst.flops_trsm += (double)n * n * n;
            sv_elim.push_back({Li, Li, nullptr, n, (int)s.xoff});
        }
        factor_bytes_ += (size_t)n * (n + 1) / 2 * sizeof(T);
        std::vector<GemvTerm> mine;
        for (int e : out) {
            int nb = edges_[e].n2;
            ops_.push_back({OP_GEMM_SYMM_OUT, c, nb, n, size(nb), size(nb), n, ilvl_, edges_[e].A, nullptr});
            factor_bytes_ += (size_t)n * size(nb) * sizeof(T);
            if (n > 0 && size(nb) > 0) {
                fwd_list.push_back({nb, GemvTerm{edges_[e].A, size(nb), n, (int)s.xoff, 0}});
                mine.push_back({edges_[e].A, size(nb), n, (int)cls_[nb].xoff, 1});
            }
        }
        for (int e : in) {
            int nb = edges_[e].n1;
            ops_.push_back({OP_GEMM_SYMM_IN, c, nb, n, size(nb), n, size(nb), ilvl_, edges_[e].A, nullptr});
            factor_bytes_ += (size_t)n * size(nb) * sizeof(T);
            if (n > 0 && size(nb) > 0) {
                fwd_list.push_back({nb, GemvTerm{edges_[e].A, n, size(nb), (int)s.xoff, 1}});
                mine.push_back({edges_[e].A, n, size(nb), (int)cls_[nb].xoff, 0});
            }
        }
        if (!mine.empty()) {
            int k0 = (int)bwd_terms.size();
            for (auto it = mine.rbegin(); it != mine.rend(); ++it) bwd_terms.push_back(*it);
            bwd_groups.push_back({(int)s.xoff, n, k0, (int)bwd_terms.size()});
        }
        set_eliminated(c);
    }
    // fill-in blocks: the Schur targets created at this level, zero-initialized
    if (fill_elems > 0) {
        fill_base = reinterpret_cast<T*>(arena_.alloc((size_t)fill_elems * sizeof(T)));
        for (size_t k = 0; k < fill_edges.size(); k++) edges_[fill_edges[k]].A = fill_base + fill_off[k];
    }
    // Schur tiles grouped by target, contributions in recorded order (built before the
    // scale phase below turns the neighbours' pivots into identities)
    {
        std::vector<int> tstart, torder;
        const std::vector<int> korder = group_by_first(
            (int)pcs.size(), (int)edges_.size(), [&](int k) { return pcs[k].target; }, tstart, torder);
        for (size_t gi = 0; gi < torder.size(); gi++) {
            const int tg = torder[gi];
            const HEdge& e = edges_[tg];
            const int m = size(e.n2), nn = size(e.n1);
            const int k0 = (int)contribs.size();
            for (int q = tstart[gi]; q < tstart[gi + 1]; q++) contribs.push_back(pcs[korder[q]].c);
            const int k1 = (int)contribs.size();
            const int lower = !gen && e.n1 == e.n2 ? 1 : 0;  // syrk on symmetric diagonal targets
            for (int c0 = 0; c0 < nn; c0 += 32)
                for (int r0 = 0; r0 < m; r0 += 32) {
                    int tm = std::min(32, m - r0), tn = std::min(32, nn - c0);
                    if (lower && r0 + tm - 1 < c0) continue;
                    tiles.push_back({e.A, m, r0, c0, tm, tn, lower, k0, k1});
                }
        }
    }
}

// Uploads and launches the elimination group, before the scaling and sparsification are
// planned, so that the device eliminates while the host plans.
template <typename T>
void Tree::launch_elimination(LevelPlan<T>& L) {
    FactorRun& R = *L.run;
    const hipStream_t st = R.st;
    const size_t esz = sizeof(T);
    LevelStats& lst = *L.lst;
    const std::vector<int>& bot = *L.bot;
    auto& potrfE = L.potrfE;
    auto& getrfE = L.getrfE;
    auto& trsmE = L.trsmE;
    auto& fill_elems = L.fill_elems;
    auto& inv_init = L.inv_init;
    auto& fill_base = L.fill_base;
    auto& tiles = L.tiles;
    auto& contribs = L.contribs;
    auto& n_potrf_elim = L.n_potrf_elim;
    // Nothing planned below feeds the elimination group (fill-in zeroing, identity inits
    // of the explicit inverses, pivots, panels, Schur updates): it is launched now, so the
    // device eliminates while the host plans the level's scaling and sparsification.
    // LU pivots that fit LDS first (one launch), the larger ones after (another)
    // pivots up to POTRF_SMALL_MAX in one launch (one workgroup each); larger ones (sorted
    // last) one 32-column panel per launch pair
    L.n_potrf_elim = potrfE.size();
    {
        small_first(potrfE);
        for (PotrfTask& t : potrfE)
            if (t.n > POTRF_SMALL_MAX) t.dbuf = arena_.alloc(32 * 32 * sizeof(T));
        lds_first<T>(getrfE);
        Packer pe;
        const int sp_te = split_trsm(trsmE);
        const size_t o_pe = pe.add(potrfE), o_te = pe.add(trsmE), o_ti = pe.add(tiles), o_co = pe.add(contribs);
        const size_t o_ie = pe.add(inv_init), o_ge = pe.add(getrfE);
        void* de = upload_tasks(pe);
        if (fill_elems > 0)
            hip_check(timed(prof_, KF_FILL, st, (double)fill_elems * esz, 0,
                            [&] { return launch_fill<T>(fill_base, (size_t)fill_elems, T(0), st); }),
                      "fill fill-in");
        if (!inv_init.empty()) {
            Work w = work_copy(inv_init, esz);
            hip_check(timed(prof_, KF_COPY, st, w.bytes, 0,
                            [&] { return launch_copy<T>(Packer::at<CopyTask>(de, o_ie), (int)inv_init.size(), st); }),
                      "identity blocks");
        }
        run_potrf<T>(st, de, potrfE, o_pe, "potrf elim");
        run_getrf<T>(st, de, getrfE, o_ge, "getrf elim");
        run_trsm<T>(st, de, trsmE, o_te, sp_te, "trsm elim");
        if (!tiles.empty()) {
            Work w = work_schur(tiles, contribs, esz);
            hip_check(timed(prof_, KF_SCHUR, st, w.bytes, w.flops,
                            [&] {
                                return launch_schur<T>(Packer::at<SchurTile>(de, o_ti), (int)tiles.size(),
                                                       Packer::at<SchurContrib>(de, o_co), st);
                            }),
                      "schur");
        }
        inv_init.clear();  // the scaling's identity inits follow
    }
    int dofs_left = 0;
    for (int c : bot)
        if (!cls_[c].elim) dofs_left += cls_[c].size;
    lst.dofs_left_elim = dofs_left;
}

// Scaling plan (scale_cluster of every remaining cluster, src/tree.cpp:1510-1519, 796-856).
template <typename T>
void Tree::plan_scaling(LevelPlan<T>& L) {
    FactorRun& R = *L.run;
    const bool gen = R.gen;
    LevelStats& lst = *L.lst;
    const std::vector<int>& bot = *L.bot;
    auto& potrfS = L.potrfS;
    auto& getrfS = L.getrfS;
    auto& trsmA = L.trsmA;
    auto& trsmB = L.trsmB;
    auto& sv_scale = L.sv_scale;
    auto& inv_init = L.inv_init;
    auto& do_sparsify_level = L.do_sparsify_level;
    auto& live = L.live;
    auto& left = L.left;
    auto push_trsm = [&L](std::vector<TrsmTask>& v, const void* A, void* B, int n, int m, int side, size_t es,
                          const void* src = nullptr, const int* perm = nullptr) {
        spand::push_trsm(v, A, B, n, m, side, es, src, perm, L.cwp);
    };
    do_sparsify_level = ilvl_ >= skip_;
    for (int c : bot)
        if (cls_[c].level > ilvl_) live.push_back(c);
    if (do_sparsify_level && monitor_flops_)
        for (int c : live) {  // scale_cluster: the pivot, then the panels of its in- and out-edges
            const HCluster& s = cls_[c];
            log_shape(gen ? SK_GETF : SK_POTF, s.size, 0, 0);
            for (int e : s.in) log_shape(SK_PANEL, size(edges_[e].n1), s.size, 0);
            for (size_t k = 1; k < s.out.size(); k++) log_shape(SK_PANEL, size(edges_[s.out[k]].n2), s.size, 0);
        }
    if (do_sparsify_level && gen) {
        // scale_cluster, PLU branch (src/tree.cpp:831-849): an edge a -> b (block
        // size(b) x size(a)) receives A U_a^{-1} from a and L_b^{-1} P_b^T A from b, in
        // the order of a and b in the bottom list, into a fresh block at b's step
        std::unordered_map<int, std::pair<void*, int*>> lu;  // cluster -> (U^T, perm)
        for (int c : live) {
            HCluster& s = cls_[c];
            const int n = s.size;
            const HEdge& piv = edges_[s.out[0]];
            void* Ut = arena_.alloc((size_t)std::max(1, n * n) * sizeof(T));
            int* pp = reinterpret_cast<int*>(arena_.alloc((size_t)std::max(1, n) * sizeof(int)));
            lu[c] = {Ut, pp};
            getrfS.push_back({piv.A, Ut, pp, n, c});
            lst.flops_potrf += getrf_flops(n);
            ops_.push_back({OP_SCALING_PLUQ, c, -1, n, 0, n, n, ilvl_, piv.A, nullptr, Ut, pp});
            factor_bytes_ += (size_t)n * (n + 1) * sizeof(T) + (size_t)n * sizeof(int);
            if (n > 0) {
                void* Li = arena_.alloc((size_t)n * n * sizeof(T));
                void* Wi = arena_.alloc((size_t)n * n * sizeof(T));
                inv_init.push_back({nullptr, Li, n, n, n, 0});
                inv_init.push_back({nullptr, Wi, n, n, n, 0});
                push_trsm(trsmA, piv.A, Li, n, n, 0, sizeof(T));
                push_trsm(trsmA, Ut, Wi, n, n, 0, sizeof(T));
                lst.flops_trsm += 2.0 * n * n * n;
                sv_scale.push_back({Li, Wi, pp, n, (int)s.xoff});
            }
        }
        for (int c : live) {
            for (size_t k = 1; k < cls_[c].out.size(); k++) {
                HEdge& e = edges_[cls_[c].out[k]];
                const int a = e.n1, b = e.n2;  // block is size(b) x size(a)
                const void* Lb = edges_[cls_[b].out[0]].A;
                lst.flops_trsm += (double)size(a) * size(b) * (size(a) + size(b));
                const bool a_first = cls_[a].order < cls_[b].order;
                void* nb = arena_.alloc((size_t)std::max(1, size(a) * size(b)) * sizeof(T));
                if (a_first) {
                    push_trsm(trsmA, lu[a].first, e.A, size(a), size(b), 1, sizeof(T));
                    push_trsm(trsmB, Lb, nb, size(b), size(a), 0, sizeof(T), e.A, lu[b].second);
                } else {
                    push_trsm(trsmA, Lb, nb, size(b), size(a), 0, sizeof(T), e.A, lu[b].second);
                    push_trsm(trsmB, lu[a].first, nb, size(a), size(b), 1, sizeof(T));
                }
                // a zero-sized block keeps its (unread) pointer
                if (size(a) > 0 && size(b) > 0) e.A = nb;
            }
        }
        for (int c : live) {
            HEdge& piv = edges_[cls_[c].out[0]];
            piv.A = nullptr;
            piv.ident = true;
        }
    } else if (do_sparsify_level) {
        for (int c : live) {
            HCluster& s = cls_[c];
            const HEdge& piv = edges_[s.out[0]];
            potrfS.push_back({piv.A, s.size, c, nullptr});
            lst.flops_potrf += potrf_flops(s.size);
            ops_.push_back({OP_SCALING_LLT, c, -1, s.size, 0, s.size, s.size, ilvl_, piv.A, nullptr});
            factor_bytes_ += (size_t)s.size * (s.size + 1) / 2 * sizeof(T);
            if (s.size > 0) {
                void* Li = arena_.alloc((size_t)s.size * s.size * sizeof(T));
                inv_init.push_back({nullptr, Li, s.size, s.size, s.size, 0});
                push_trsm(trsmA, piv.A, Li, s.size, s.size, 0, sizeof(T));
                lst.flops_trsm += (double)s.size * s.size * s.size;
                sv_scale.push_back({Li, Li, nullptr, s.size, (int)s.xoff});
            }
        }
        for (int c : live) {
            for (size_t k = 1; k < cls_[c].out.size(); k++) {
                const HEdge& e = edges_[cls_[c].out[k]];
                const int a = e.n1, b = e.n2;  // block is size(b) x size(a)
                const void* La = edges_[cls_[a].out[0]].A;
                const void* Lb = edges_[cls_[b].out[0]].A;
                lst.flops_trsm += (double)size(a) * size(b) * (size(a) + size(b));
                const bool a_first = cls_[a].order < cls_[b].order;
                // right solve by L_a (B is size(b) x size(a)), left solve by L_b
                push_trsm(a_first ? trsmA : trsmB, La, e.A, size(a), size(b), 1, sizeof(T));
                push_trsm(a_first ? trsmB : trsmA, Lb, e.A, size(b), size(a), 0, sizeof(T));
            }
        }
        for (int c : live) {
            HEdge& piv = edges_[cls_[c].out[0]];
            piv.A = nullptr;
            piv.ident = true;
        }
    }
}

// Sparsification plan (src/tree.cpp:1521-1533, 1292-1347): dependency waves, the RRQR
// path of each cluster (register, one-workgroup LDS or stepped slabs) and its buffers.
template <typename T>
void Tree::plan_sparsification(LevelPlan<T>& L) {
    FactorRun& R = *L.run;
    const long long lds_elems = R.lds_elems;
    const long long big_min_rows = R.big_min_rows;
    LevelStats& lst = *L.lst;
    auto& shape = L.shape;
    auto& tune = L.tune;
    auto& do_sparsify_level = L.do_sparsify_level;
    auto& live = L.live;
    auto& plist = L.plist;
    auto& wave = L.wave;
    auto& widx = L.widx;
    auto& nwaves = L.nwaves;
    auto& wtasks = L.wtasks;
    auto& wlds_small = L.wlds_small;
    auto& wbig = L.wbig;
    auto& wslab = L.wslab;
    auto& wsteps = L.wsteps;
    auto& wmaxrows = L.wmaxrows;
    auto& wreg = L.wreg;
    auto& wregvar = L.wregvar;
    auto& where = L.where;
    auto& segs = L.segs;
    auto& wlds_step = L.wlds_step;
    auto& wlds_wave = L.wlds_wave;
    auto& persistent = L.persistent;
    auto& step_waves = L.step_waves;
    auto& wflag0 = L.wflag0;
    if (do_sparsify_level) {
        std::vector<int> inP(cls_.size(), -1);
        for (int c : live) {
            if (use_want_sparsify_ && !cls_[c].spars) {
                lst.nignored++;
                continue;
            }
            inP[c] = (int)plist.size();
            plist.push_back(c);
        }
        wave.assign(plist.size(), 0);
        for (size_t p = 0; p < plist.size(); p++) {
            const int c = plist[p];
            int w = 0;
            auto visit = [&](int nb) {
                if (inP[nb] >= 0 && cls_[nb].order < cls_[c].order) w = std::max(w, wave[inP[nb]] + 1);
            };
            for (int e : cls_[c].in) visit(edges_[e].n1);
            for (size_t k = 1; k < cls_[c].out.size(); k++) visit(edges_[cls_[c].out[k]].n2);
            wave[p] = w;
            nwaves = std::max(nwaves, w + 1);
        }
        wtasks.assign(nwaves, {});
        wlds_small.assign(nwaves, 0);
        wbig.assign(nwaves, {});
        wslab.assign(nwaves, {});
        wsteps.assign(nwaves, 0);
        wmaxrows.assign(nwaves, 0);
        wreg.assign(nwaves, {});
        wregvar.assign(nwaves, {});
        where.assign(plist.size(), {0, 0});
        std::vector<std::vector<std::pair<int, int>>> regp(nwaves);  // (variant, plist position)
        std::vector<RrqrTask> ptask(plist.size());
        {  // the sparsification knobs from the level's Asn shapes
            std::vector<int> rows, cols;
            for (int c : plist) {
                const HCluster& s = cls_[c];
                int nc = 0;
                for (int e : s.in) nc += size(edges_[e].n1);
                for (size_t k = 1; k < s.out.size(); k++) nc += size(edges_[s.out[k]].n2);
                rows.push_back(s.size);
                cols.push_back(nc);
            }
            shape.n_spars = (int)plist.size();
            if (!rows.empty()) {
                std::sort(rows.begin(), rows.end());
                std::sort(cols.begin(), cols.end());
                shape.rows_med = rows[rows.size() / 2];
                shape.rows_max = rows.back();
                shape.cols_med = cols[cols.size() / 2];
                shape.cols_max = cols.back();
            }
            const LevelTune t = choose_tune(sizeof(T), shape);
            tune.rrqr_reg = t.rrqr_reg;
            tune.small_elems = t.small_elems;
            tune.step_waves = t.step_waves;
            apply_tune_env(tune);
        }
        const bool reg_path = tune.rrqr_reg != 0;
        const long long big_min_elems = tune.small_elems;
        widx.assign(plist.size(), 0);
        for (size_t p = 0; p < plist.size(); p++) {
            const int c = plist[p];
            const HCluster& s = cls_[c];
            RrqrTask t{};
            t.cluster = c;
            t.seg0 = (int)segs.size();
            int cols = 0;
            for (int e : s.in) {
                segs.push_back({edges_[e].A, edges_[e].n1, 0});
                cols += size(edges_[e].n1);
            }
            for (size_t k = 1; k < s.out.size(); k++) {
                const HEdge& e = edges_[s.out[k]];
                segs.push_back({e.A, e.n2, 1});
                cols += size(e.n2);
            }
            t.nseg = (int)segs.size() - t.seg0;
            if (t.nseg > 1024) throw Error(ERR_UNSUPPORTED, "cluster with more than 1024 neighbours");
            t.max_rows = s.size;
            t.max_cols = cols;
            const int mn = std::min(s.size, cols);
            t.Q = arena_.alloc((size_t)std::max(1, s.size * mn) * sizeof(T));
            t.tau = arena_.alloc((size_t)std::max(1, mn) * sizeof(T));
            lst.flops_rrqr += mn > 0 ? 2.0 * s.size * cols * mn - (2.0 / 3.0) * (double)mn * mn * mn : 0.0;
            const int w = wave[p];
            const int reg_var = reg_path ? rrqr_reg_variant(sizeof(T), s.size, cols) : -1;
            if (reg_var >= 0) {
                // register-resident, one workgroup per cluster
                ptask[p] = t;
                regp[w].push_back({reg_var, (int)p});
            } else if (s.size >= big_min_rows && (long long)s.size * cols > big_min_elems) {
                // stepped multi-workgroup path
                RrqrBig b{};
                b.nslab = std::max(1, (cols + 63) / 64);
                t.scratch = arena_.alloc((size_t)s.size * b.nslab * 64 * sizeof(T));
                t.norms = reinterpret_cast<double*>(arena_.alloc((size_t)2 * std::max(cols, 1) * sizeof(double)));
                b.map = reinterpret_cast<int*>(arena_.alloc((size_t)4 * std::max(cols, 1) * sizeof(int)));
                b.cand_v = reinterpret_cast<double*>(arena_.alloc((size_t)(2 * b.nslab + 1) * sizeof(double)));
                b.cand_p = reinterpret_cast<int*>(arena_.alloc((size_t)2 * b.nslab * sizeof(int)));
                b.t = t;
                where[p] = {1, (int)wbig[w].size()};
                for (int k = 0; k < b.nslab; k++) wslab[w].push_back({(int)wbig[w].size(), k});
                wbig[w].push_back(b);
                wsteps[w] = std::max(wsteps[w], mn);
                wmaxrows[w] = std::max(wmaxrows[w], s.size);
            } else {
                // LDS-resident with an odd column stride (rows | 1), else a global scratch
                if ((long long)(s.size | 1) * cols > lds_elems)
                    t.scratch = arena_.alloc((size_t)s.size * cols * sizeof(T));
                else
                    wlds_small[w] =
                        std::max(wlds_small[w], (long long)(s.size | 1) * cols * (long long)sizeof(T));
                if (cols > 1024) {
                    t.norms = reinterpret_cast<double*>(arena_.alloc((size_t)2 * cols * sizeof(double)));
                    t.jpvt = reinterpret_cast<int*>(arena_.alloc((size_t)cols * sizeof(int)));
                }
                where[p] = {0, (int)wtasks[w].size()};
                wtasks[w].push_back(t);
            }
        }
        for (int w = 0; w < nwaves; w++) {
            std::stable_sort(regp[w].begin(), regp[w].end(),
                             [](const std::pair<int, int>& x, const std::pair<int, int>& y) { return x.first < y.first; });
            for (auto& vp : regp[w]) {
                where[vp.second] = {2, (int)wreg[w].size()};
                wreg[w].push_back(ptask[vp.second]);
                wregvar[w].push_back(vp.first);
            }
        }
    }
    // pivot-column / reflector buffers of the stepped path: LDS unless a wave's largest
    // cluster needs more than STEP_LDS_MAX, then global buffers for the whole wave
    // A wave whose clusters all have at most WAVE_ROWS_MAX rows runs as one persistent
    // launch (candidate-column buffers in vbuf); the per-step launches remain the path
    // for larger clusters and for waves the device cannot hold at once.
    wlds_step.assign(nwaves, 0);
    wlds_wave.assign(nwaves, 0);
    // The persistent launch is the default (SPAND_RRQR_PERSISTENT=0 returns to one launch
    // per step): on the 64^3 Laplacian its steps take 6-12 us against about 10.5 us per
    // step launch, and FP64 and FP32 both gain.
    persistent = env_or("SPAND_RRQR_PERSISTENT", 1) != 0;
    step_waves = tune.step_waves;  // 4, 8, 16; 0: by slab count
    tunes_[ilvl_] = tune;
    shapes_[ilvl_] = shape;
    for (int w = 0; w < nwaves; w++) {
        const size_t vb = (size_t)std::max(wmaxrows[w], 1) * sizeof(T);
        int wmaxcols = 0;
        for (auto& b : wbig[w]) wmaxcols = std::max(wmaxcols, b.t.max_cols);
        const size_t lw = ((size_t)wmaxcols * sizeof(int) + 15) / 16 * 16;  // the persistent launch's p2c
        if (persistent && !wbig[w].empty() && wmaxrows[w] <= WAVE_ROWS_MAX && lw <= (size_t)WAVE_LDS_MAX) {
            wlds_wave[w] = lw;
            wlds_step[w] = 2 * vb;  // the per-step fallback (a wave the device cannot hold at once)
            for (auto& b : wbig[w]) {
                b.vbuf = arena_.alloc(rrqr_relay_bytes(sizeof(T), b.nslab, std::max(b.t.max_rows, 1)));
                b.relay = 1;
            }
        } else if (2 * vb <= (size_t)STEP_LDS_MAX) {
            wlds_step[w] = 2 * vb;
        } else {
            for (auto& b : wbig[w])
                b.vbuf = arena_.alloc((size_t)b.nslab * 2 * std::max(b.t.max_rows, 1) * sizeof(T));
        }
    }
    // dims (one block per level) and done flags of the stepped clusters; the flags are
    // reset before any of this level's launches
    wflag0.assign(nwaves, 0);
    {
        size_t nbig = 0;
        for (int w = 0; w < nwaves; w++) nbig += wbig[w].size();
        int* dblock = nullptr;
        if (nbig > 0) {
            signal_reserve(1 + nbig);
            dblock = reinterpret_cast<int*>(arena_.alloc(nbig * 5 * sizeof(int)));
            size_t f = 1;
            for (int w = 0; w < nwaves; w++) {
                wflag0[w] = f;
                for (auto& b : wbig[w]) {
                    h_sig_[f] = 0;
                    b.done = d_sig_ + f;
                    b.dims = dblock + 5 * (f - 1);
                    f++;
                }
            }
        }
        level_dims_.push_back({dblock, nbig});
    }
    lst.nwaves = nwaves;
}

// Uploads and launches the scaling and the sparsification waves, then reads the level's
// pivot status and the kept ranks back.
template <typename T>
void Tree::launch_level(LevelPlan<T>& L) {
    FactorRun& R = *L.run;
    const hipStream_t st = R.st;
    const bool gen = R.gen;
    const size_t esz = sizeof(T);
    std::vector<int>& hsize = R.hsize;
    LevelStats& lst = *L.lst;
    auto& potrfS = L.potrfS;
    auto& getrfS = L.getrfS;
    auto& trsmA = L.trsmA;
    auto& trsmB = L.trsmB;
    auto& inv_init = L.inv_init;
    auto& live = L.live;
    auto& plist = L.plist;
    auto& wave = L.wave;
    auto& nwaves = L.nwaves;
    auto& wtasks = L.wtasks;
    auto& wlds_small = L.wlds_small;
    auto& wbig = L.wbig;
    auto& wslab = L.wslab;
    auto& wsteps = L.wsteps;
    auto& wmaxrows = L.wmaxrows;
    auto& wreg = L.wreg;
    auto& wregvar = L.wregvar;
    auto& segs = L.segs;
    auto& wlds_step = L.wlds_step;
    auto& wlds_wave = L.wlds_wave;
    auto& persistent = L.persistent;
    auto& step_waves = L.step_waves;
    auto& wflag0 = L.wflag0;
    auto& t_lvl0 = L.t_lvl0;
    auto& t_plan1 = L.t_plan1;
    auto& t_lvl1 = L.t_lvl1;
    for (size_t c = 0; c < cls_.size(); c++) hsize[c] = cls_[c].size;
    small_first(potrfS);
    for (PotrfTask& t : potrfS)
        if (t.n > POTRF_SMALL_MAX) t.dbuf = arena_.alloc(32 * 32 * sizeof(T));
    Packer pk;
    const int sp_ta = split_trsm(trsmA), sp_tb = split_trsm(trsmB);
    size_t o_ps = pk.add(potrfS), o_ta = pk.add(trsmA), o_tb = pk.add(trsmB), o_sg = pk.add(segs);
    size_t o_ii = pk.add(inv_init);
    lds_first<T>(getrfS);
    size_t o_gs = pk.add(getrfS);
    std::vector<size_t> o_w(nwaves), o_b(nwaves), o_s(nwaves), o_r(nwaves);
    for (int w = 0; w < nwaves; w++) {
        o_r[w] = pk.add(wreg[w]);
        o_w[w] = pk.add(wtasks[w]);
        o_b[w] = pk.add(wbig[w]);
        o_s[w] = pk.add(wslab[w]);
    }
    L.t_plan1 = wall();
    void* dt = upload_tasks(pk);
    stage_upload(d_size_, hsize.data(), cls_.size() * sizeof(int));
    if (!inv_init.empty()) {
        Work w = work_copy(inv_init, esz);
        hip_check(timed(prof_, KF_COPY, st, w.bytes, 0,
                        [&] { return launch_copy<T>(Packer::at<CopyTask>(dt, o_ii), (int)inv_init.size(), st); }),
                  "identity blocks");
    }
    run_potrf<T>(st, dt, potrfS, o_ps, "potrf scale");
    run_getrf<T>(st, dt, getrfS, o_gs, "getrf scale");
    run_trsm<T>(st, dt, trsmA, o_ta, sp_ta, "trsm scale A");
    run_trsm<T>(st, dt, trsmB, o_tb, sp_tb, "trsm scale B");
    // stepped-path launch slots, for the post-hoc work attribution below
    std::vector<int> slot_init(nwaves, -1), slot_finish(nwaves, -1);
    std::vector<int> slot_step(nwaves, -1);
    for (int w = 0; w < nwaves; w++) {
        // the register-path clusters of a wave run on the side stream, beside the
        // one-workgroup and stepped launches on the main stream; both join before the
        // next wave (whose clusters read the edges this wave rewrites)
        const bool side = !wreg[w].empty() && (!wtasks[w].empty() || !wbig[w].empty());
        const hipStream_t rst = side ? aux_stream() : st;
        if (side) {
            hip_check(hipEventRecord(ev_fork_, st), "fork");
            hip_check(hipStreamWaitEvent(rst, ev_fork_, 0), "fork");
        }
        if (!wreg[w].empty()) {
            // one launch per variant run (tasks are sorted by variant); a single launch
            // when the variant of the largest rows and columns holds every task
            int mr = 0, mc = 0;
            for (auto& t : wreg[w]) {
                mr = std::max(mr, t.max_rows);
                mc = std::max(mc, t.max_cols);
            }
            const int cover = rrqr_reg_variant(sizeof(T), mr, mc);
            size_t a0 = 0;
            while (a0 < wreg[w].size()) {
                size_t a1 = cover >= 0 ? wreg[w].size() : a0 + 1;
                if (cover < 0)
                    while (a1 < wreg[w].size() && wregvar[w][a1] == wregvar[w][a0]) a1++;
                const int var = cover >= 0 ? cover : wregvar[w][a0];
                Work wk;
                for (size_t x = a0; x < a1; x++) {
                    const RrqrTask& t = wreg[w][x];
                    wk.bytes += (2.0 * t.max_rows * t.max_cols + (double)t.max_rows * std::min(t.max_rows, t.max_cols)) * esz;
                    wk.flops += geqp3_flops(t.max_rows, t.max_cols);
                }
                hip_check(timed(prof_, KF_RRQR_REG, rst, wk.bytes, wk.flops,
                                [&] {
                                    return launch_rrqr_reg<T>(Packer::at<RrqrTask>(dt, o_r[w]) + a0, (int)(a1 - a0),
                                                              Packer::at<RrqrSeg>(dt, o_sg), d_size_, tol_, var, rst);
                                }),
                          "rrqr reg");
                a0 = a1;
            }
        }
        // join at the end of the wave (below), also on early exits of the loop body
        struct Join {
            bool on;
            hipStream_t main, side;
            hipEvent_t ev;
            ~Join() {
                if (on && hipEventRecord(ev, side) == hipSuccess) (void)hipStreamWaitEvent(main, ev, 0);
            }
        } join{side, st, rst, ev_join_};
        if (!wtasks[w].empty()) {
            Work wk;
            for (auto& t : wtasks[w]) {
                wk.bytes += (2.0 * t.max_rows * t.max_cols + (double)t.max_rows * std::min(t.max_rows, t.max_cols)) * esz;
                wk.flops += geqp3_flops(t.max_rows, t.max_cols);
            }
            hip_check(timed(prof_, KF_RRQR, st, wk.bytes, wk.flops,
                            [&] {
                                return launch_rrqr<T>(Packer::at<RrqrTask>(dt, o_w[w]), (int)wtasks[w].size(),
                                                      Packer::at<RrqrSeg>(dt, o_sg), d_size_, tol_,
                                                      (int)std::max(wlds_small[w], (long long)LDS_BYTES), st);
                            }),
                      "rrqr");
        }
        if (wbig[w].empty()) continue;
        const RrqrBig* db = Packer::at<RrqrBig>(dt, o_b[w]);
        const RrqrSlab* ds = Packer::at<RrqrSlab>(dt, o_s[w]);
        const int nsl = (int)wslab[w].size();
        hip_check(timed(prof_, KF_RRQR_INIT, st, 0, 0,
                        [&] { return launch_rrqr_init<T>(db, ds, nsl, Packer::at<RrqrSeg>(dt, o_sg), d_size_, tol_, st); },
                        &slot_init[w]),
                  "rrqr init");
        // the steps of a wave are timed as one batch (two events per wave)
        bool done = false;
        if (wlds_wave[w] > 0) {
            // whole clusters in chunks the device holds at once, one cooperative launch each
            int cap = 0;
            hip_check(rrqr_wave_capacity<T>(wmaxrows[w], wlds_wave[w], &cap), "rrqr wave capacity");
            int maxslab = 0;
            for (auto& b : wbig[w]) maxslab = std::max(maxslab, b.nslab);
            if (cap >= maxslab && cap > 0) {
                const int sl = prof_.start(KF_RRQR_WAVE, st);
                int launches = 0;
                size_t x0 = 0, s0 = 0;
                while (x0 < wbig[w].size()) {
                    size_t x1 = x0, s1 = s0;
                    int steps = 0;
                    while (x1 < wbig[w].size() && (int)(s1 - s0) + wbig[w][x1].nslab <= cap) {
                        steps = std::max(steps, std::min(wbig[w][x1].t.max_rows, wbig[w][x1].t.max_cols));
                        s1 += wbig[w][x1].nslab;
                        x1++;
                    }
                    const hipError_t e = launch_rrqr_wave<T>(db, ds + s0, (int)(s1 - s0), steps, wmaxrows[w], tol_,
                                                             wlds_wave[w], d_info_ + 1, R.spin_max, st);
                    hip_check(e, "rrqr wave");
                    launches++;
                    x0 = x1;
                    s0 = s1;
                }
                prof_.stop(sl, st, 0, 0, launches);
                done = true;
                slot_step[w] = sl;
            }
        }
        if (!done) {
            // Steps are launched at most RRQR_STEPS_AHEAD ahead of the device, and none
            // once every cluster of the wave has flagged its last working step: the
            // ranks (decided on the device) end most clusters well before min(rows, cols).
            slot_step[w] = prof_.start(KF_RRQR_STEP, st);
            std::vector<int> live(wbig[w].size());
            std::iota(live.begin(), live.end(), 0);
            std::vector<long long> seq(wsteps[w], 0);
            int launched = 0;
            for (int i = 0; i < wsteps[w]; i++) {
                if (i >= RRQR_STEPS_AHEAD) wait_signal(seq[i - RRQR_STEPS_AHEAD]);
                size_t k = 0;
                for (int x : live) {
                    const RrqrBig& b = wbig[w][x];
                    const volatile long long* f = h_sig_ + wflag0[w] + x;
                    if (*f == 0 && i < std::min(b.t.max_rows, b.t.max_cols)) live[k++] = x;
                }
                live.resize(k);
                if (live.empty()) break;
                seq[i] = ++sig_seq_;
                hip_check(launch_rrqr_step<T>(db, ds, nsl, i, wmaxrows[w], tol_, wlds_step[w], step_waves, d_sig_,
                                              seq[i], st),
                          "rrqr step");
                launched++;
            }
            prof_.stop(slot_step[w], st, 0, 0, launched);
        }
        hip_check(timed(prof_, KF_RRQR_FINISH, st, 0, 0,
                        [&] {
                            return launch_rrqr_finish<T>(db, ds, nsl, Packer::at<RrqrSeg>(dt, o_sg), d_size_, tol_,
                                                         0, st);
                        },
                        &slot_finish[w]),
                  "rrqr finish");
    }
    int info2[2] = {0, 0};
    hip_check(hipMemcpy(info2, d_info_, 2 * sizeof(int), hipMemcpyDeviceToHost), "read info");
    const int info = info2[0];
    if (info2[1] != 0) throw Error(ERR_HIP, "persistent RRQR: a wait for another slab's candidate gave up");
    if (info != 0) {
        if (gen) throw Error(ERR_SINGULAR, "Error: Singular Pivot (cluster order " + std::to_string(info - 1) + ")");
        throw Error(ERR_NOT_SPD, "Error: Non-SPD Pivot (cluster order " + std::to_string(info - 1) + ")");
    }
    if (!plist.empty())
        hip_check(hipMemcpy(hsize.data(), d_size_, cls_.size() * sizeof(int), hipMemcpyDeviceToHost), "read ranks");
    if (prof_.on())  // the stepped path's sizes are decided on the device: priced at the end
        for (int w = 0; w < nwaves; w++)
            if (!wbig[w].empty())
                wave_work_.push_back({slot_init[w], slot_finish[w], slot_step[w], wsteps[w], wflag0[w] - 1,
                                      wbig[w].size(), ilvl_});
    L.t_lvl1 = wall();
    lst.t_plan = t_plan1 - t_lvl0;
    lst.t_elim_scale = t_lvl1 - t_plan1;
}

// Records the Orthogonal, Split and sibling ops of the kept ranks (src/tree.cpp:1317-1342)
// and plans the compact-WY factors of the solve.
template <typename T>
void Tree::apply_ranks(LevelPlan<T>& L) {
    FactorRun& R = *L.run;
    const bool gen = R.gen;
    std::vector<int>& hsize = R.hsize;
    LevelStats& lst = *L.lst;
    auto& plist = L.plist;
    auto& wave = L.wave;
    auto& wtasks = L.wtasks;
    auto& wbig = L.wbig;
    auto& wreg = L.wreg;
    auto& where = L.where;
    auto& qops = L.qops;
    auto& wy = L.wy;
    auto& wy_tiles = L.wy_tiles;
    auto& wy_trsm = L.wy_trsm;
    auto& t_lvl1 = L.t_lvl1;
    auto& t_lvl2 = L.t_lvl2;
    auto push_trsm = [&L](std::vector<TrsmTask>& v, const void* A, void* B, int n, int m, int side, size_t es,
                          const void* src = nullptr, const int* perm = nullptr) {
        spand::push_trsm(v, A, B, n, m, side, es, src, perm, L.cwp);
    };
    for (size_t p = 0; p < plist.size(); p++) {
        const int c = plist[p];
        const int r = hsize[c];
        const int s0 = cls_[c].size;
        lst.rank_before += s0;
        lst.nsparsified++;
        if (monitor_flops_) {
            // Asn as the sequential loop sees it: neighbours earlier in the list already
            // hold their kept rank
            int ncols = 0;
            for (int e : cls_[c].in) ncols += size(edges_[e].n1);
            for (size_t k = 1; k < cls_[c].out.size(); k++) ncols += size(edges_[cls_[c].out[k]].n2);
            log_shape(SK_RRQR, s0, ncols, 0);
        }
        if (r < s0) {
            const RrqrTask* t = where[p].first == 1   ? &wbig[wave[p]][where[p].second].t
                                : where[p].first == 2 ? &wreg[wave[p]][where[p].second]
                                                      : &wtasks[wave[p]][where[p].second];
            ops_.push_back({OP_ORTHOGONAL, c, -1, s0, 0, s0, r, ilvl_, t->Q, t->tau});
            factor_bytes_ += ((size_t)s0 * r + (size_t)r * (r + 1) / 2) * sizeof(T);
            if (r > 0) {
                // compact WY over all r kept reflectors, for the solve: Q = I - Ve T Ve^T
                void* Ve = arena_.alloc((size_t)s0 * r * sizeof(T));
                void* Sm = arena_.alloc((size_t)r * r * sizeof(T));
                void* X = arena_.alloc((size_t)r * r * sizeof(T));
                const int tw = (int)wy.size();
                wy.push_back({t->Q, t->tau, Ve, Sm, X, s0, r});
                for (int c0 = 0; c0 < r; c0 += 32)
                    for (int a0 = 0; a0 <= c0; a0 += 32) wy_tiles.push_back(make_int4(tw, a0, c0, 0));
                push_trsm(wy_trsm, Sm, X, r, r, 0, sizeof(T));
                qops.push_back({Ve, X, s0, r, (int)cls_[c].xoff, (int)ws_len_});
                ws_len_ += 2 * r;  // the y and z segments of the apply
            }
            HCluster sib;
            sib.start = cls_[c].start + r;
            sib.size = sib.osize = s0 - r;
            sib.level = cls_[c].level;
            sib.spars = cls_[c].spars;
            sib.elim = true;
            sib.order = (int)cls_.size();
            sib.xoff = cls_[c].xoff + r;
            const int sid = (int)cls_.size();
            cls_.push_back(sib);
            others_.push_back(sid);
            cls_[c].size = r;
            ops_.push_back({OP_SPLIT, c, sid, r, s0 - r, 0, 0, ilvl_, nullptr, nullptr});
            ops_.push_back({gen ? OP_SCALING_PLUQ : OP_SCALING_LLT, sid, -1, s0 - r, 0, s0 - r, s0 - r, ilvl_,
                            nullptr, nullptr});
            log_shape(gen ? SK_GETF : SK_POTF, s0 - r, 0, 0);  // eliminate_cluster(sibling)
            // the sibling's pivot is the identity and is never stored (src/tree.cpp:1076-1099)
        }
        lst.rank_after += cls_[c].size;
    }
    L.t_lvl2 = wall();
    lst.t_spars = t_lvl2 - t_lvl1;
}

// Merge (src/tree.cpp:1435-1445, 1106-1184): parents, their merged edges, and the
// compact-WY launches of the Orthogonal ops.
template <typename T>
void Tree::merge_level(LevelPlan<T>& L) {
    FactorRun& R = *L.run;
    const hipStream_t st = R.st;
    LevelStats& lst = *L.lst;
    auto& wy = L.wy;
    auto& wy_tiles = L.wy_tiles;
    auto& wy_trsm = L.wy_trsm;
    auto& mt = L.mt;
    auto& mchildren = L.mchildren;
    auto& left = L.left;
    auto& t_lvl2 = L.t_lvl2;
    if (ilvl_ < nlevels_ - 1) {
        current_bottom_++;
        const std::vector<int>& nb = bottoms_[current_bottom_];
        std::vector<int> pos(cls_.size(), 0);  // a child's row offset inside its parent
        for (int p : nb) {
            HCluster& P = cls_[p];
            int off = 0;
            MergeTask m{(int)ws_len_, (int)mchildren.size(), 0, 0};
            for (int ch : P.children) {
                pos[ch] = off;
                mchildren.push_back(make_int2((int)cls_[ch].xoff, cls_[ch].size));
                off += cls_[ch].size;
            }
            m.c1 = (int)mchildren.size();
            P.size = P.osize = off;
            P.xoff = ws_len_;
            ws_len_ += off;
            ops_.push_back({OP_MERGE, p, -1, off, 0, 0, 0, ilvl_, nullptr, nullptr});
            if (off > 0) mt.push_back(m);
        }
        long long melems = 0;
        std::vector<long long> moff;   // by new edge id - first_new
        const int first_new = (int)edges_.size();
        struct CT { int dst_edge; long long dst_add; const void* src; int rows, cols; };
        std::vector<CT> cts;
        // per parent: its new neighbours in order of first appearance over the children's
        // out-edges (stamped slots instead of searches), each new edge original if any
        // merged child edge is
        std::vector<int> stamp(cls_.size(), -1), eid(cls_.size(), -1);
        std::vector<char> origs;
        std::vector<int> nnews;
        for (int p : nb) {
            nnews.clear();
            origs.clear();
            for (int ch : cls_[p].children)
                for (int e : cls_[ch].out) {
                    const int nn = cls_[edges_[e].n2].parent;
                    if (stamp[nn] != p) {
                        stamp[nn] = p;
                        eid[nn] = (int)nnews.size();
                        nnews.push_back(nn);
                        origs.push_back(0);
                    }
                    if (edges_[e].original) origs[eid[nn]] = 1;
                }
            for (size_t q = 0; q < nnews.size(); q++) {
                const int nn = nnews[q];
                eid[nn] = new_edge(p, nn, nullptr, origs[q] != 0);
                moff.push_back(melems);
                melems += (long long)size(nn) * size(p);
            }
            auto& out = cls_[p].out;
            std::stable_sort(out.begin() + 1, out.end(),
                             [&](int a, int b) { return cls_[edges_[a].n2].order < cls_[edges_[b].n2].order; });
            auto& in = cls_[p].in;
            std::stable_sort(in.begin(), in.end(),
                             [&](int a, int b) { return cls_[edges_[a].n1].order < cls_[edges_[b].n1].order; });
            for (int ch : cls_[p].children) {
                for (int e : cls_[ch].out) {
                    int nold = edges_[e].n2, nn = cls_[nold].parent;
                    int de = eid[nn];
                    long long add = pos[nold] + (long long)pos[ch] * size(nn);
                    if (size(nold) > 0 && size(ch) > 0)
                        cts.push_back({de, add, edges_[e].ident ? nullptr : edges_[e].A, size(nold), size(ch)});
                }
                cls_[ch].out.clear();
                cls_[ch].in.clear();
            }
        }
        T* mbase = reinterpret_cast<T*>(arena_.alloc((size_t)std::max<long long>(melems, 1) * sizeof(T)));
        for (size_t k = first_new; k < edges_.size(); k++) edges_[k].A = mbase + moff[k - first_new];
        std::vector<CopyTask> copies;
        copies.reserve(cts.size());
        for (auto& c : cts) {
            const HEdge& d = edges_[c.dst_edge];
            copies.push_back({c.src, reinterpret_cast<T*>(d.A) + c.dst_add, c.rows, c.cols, size(d.n2), 0});
        }
        Packer pm;
        size_t o_c = pm.add(copies);
        void* dm = upload_tasks(pm);
        if (melems > 0)
            hip_check(timed(prof_, KF_FILL, st, (double)melems * sizeof(T), 0,
                            [&] { return launch_fill<T>(mbase, (size_t)melems, T(0), st); }),
                      "fill merge");
        Work wc = work_copy(copies, sizeof(T));
        hip_check(timed(prof_, KF_COPY, st, wc.bytes, 0,
                        [&] { return launch_copy<T>(Packer::at<CopyTask>(dm, o_c), (int)copies.size(), st); }),
                  "merge copy");
    }
    if (!wy.empty()) {
        Packer pl;
        const int sp_wy = split_trsm(wy_trsm);
        size_t o_w = pl.add(wy), o_t = pl.add(wy_tiles), o_s = pl.add(wy_trsm);
        void* dl = upload_tasks(pl);
        double lb = 0, lf = 0;
        for (auto& w : wy) {
            lb += (3.0 * w.m * w.k + 3.0 * w.k * w.k) * sizeof(T);
            lf += (double)w.m * w.k * w.k;
        }
        hip_check(timed(prof_, KF_LARFT, st, lb, lf,
                        [&] {
                            return launch_wy<T>(Packer::at<WyTask>(dl, o_w), (int)wy.size(),
                                                Packer::at<int4>(dl, o_t), (int)wy_tiles.size(), st);
                        }),
                  "compact WY");
        Work w = work_trsm(wy_trsm, sizeof(T));
        hip_check(timed(prof_, KF_TRSM, st, w.bytes, w.flops,
                        [&] {
                            return launch_trsm<T>(Packer::at<TrsmTask>(dl, o_s), (int)wy_trsm.size(), sp_wy, st);
                        }),
                  "compact WY inverse");
    }
    lst.t_merge = wall() - t_lvl2;
    left = 0;
    for (int c : bottoms_[current_bottom_])
        if (!cls_[c].elim) left += cls_[c].size;
    lst.dofs_left_spars = left;
    lst.fact_nnz = nnz();
}

// The level's preconditioner-apply plan: forward terms grouped now, the rest built on a
// worker thread.
template <typename T>
void Tree::plan_level_solve(LevelPlan<T>& L) {
    FactorRun& R = *L.run;
    const size_t esz = sizeof(T);
    auto& pending_solve = R.pending_solve;
    LevelStats& lst = *L.lst;
    auto& potrfS = L.potrfS;
    auto& sv_elim = L.sv_elim;
    auto& sv_scale = L.sv_scale;
    auto& fwd_list = L.fwd_list;
    auto& bwd_groups = L.bwd_groups;
    auto& bwd_terms = L.bwd_terms;
    auto& n_potrf_elim = L.n_potrf_elim;
    auto& plist = L.plist;
    auto& nwaves = L.nwaves;
    auto& qops = L.qops;
    auto& mt = L.mt;
    auto& mchildren = L.mchildren;
    auto& left = L.left;
    // The forward terms are grouped here (the grouping reads cluster offsets, and the
    // cluster array grows at the next merge); the rest is built on a worker thread from
    // moved copies of this level's lists and uploaded once the factorization is done.
    const double t_sp0 = wall();
    {
        std::vector<GemvGroup> fwd_groups;
        std::vector<GemvTerm> fwd_t;
        {
            std::vector<int> gstart, gkeys;
            const std::vector<int> forder = group_by_first(
                (int)fwd_list.size(), (int)cls_.size(), [&](int k) { return fwd_list[k].first; }, gstart, gkeys);
            for (size_t gi = 0; gi < gkeys.size(); gi++) {
                const int k0 = (int)fwd_t.size();
                for (int q = gstart[gi]; q < gstart[gi + 1]; q++) fwd_t.push_back(fwd_list[forder[q]].second);
                const GemvTerm& t0 = fwd_t[k0];
                const int n = t0.trans ? t0.cols : t0.rows;
                fwd_groups.push_back({(int)cls_[gkeys[gi]].xoff, n, k0, (int)fwd_t.size()});
            }
        }
        HostPhases in{std::move(sv_elim),    std::move(sv_scale),  std::move(fwd_groups), std::move(bwd_groups),
                      std::move(fwd_t),      std::move(bwd_terms), std::move(qops)};
        pending_solve[ilvl_] =
            std::async(std::launch::async, [esz = sizeof(T), in = std::move(in), mt = std::move(mt),
                                            mchildren = std::move(mchildren)]() mutable {
                return build_solve_plan(esz, std::move(in), std::move(mt), std::move(mchildren));
            });
    }
    lst.t_solve_plan = wall() - t_sp0;
    if (verb_)
        printf("Level %d: elim %zu, scale %zu, sparsify %zu in %d waves, %d dofs left\n", ilvl_, n_potrf_elim,
               potrfS.size(), plist.size(), nwaves, left);
}

template <typename T>
void Tree::factorize_impl() {
    const double t_start = wall();
    FactorRun R;
    const hipStream_t st = R.st;  // the null stream
    drop_apply_graphs();
    graphs_on_ = env_or("SPAND_APPLY_GRAPH", 1) != 0;
    arena_.reset();
    task_arena_.reset();
    wave_work_.clear();
    level_dims_.clear();
    R.pending_solve.resize(nlevels_);  // joined on every exit (FactorRun's destructor)
    tplans_.clear();
    tplans_.resize(nlevels_);
    hip_check(hipDeviceSynchronize(), "factorize");
    h_stage_used_ = 0;
    h_stage_total_ = 0;
    ops_.clear();
    cls_ = cls0_;
    edges_ = edges0_;
    others_.clear();
    current_bottom_ = 0;
    solve_plan_.assign(nlevels_, LevelSolve());
    for (auto& s : stats_) {
        int nd = s.dofs_nd, nl = s.dofs_left_nd;
        s = LevelStats();
        s.dofs_nd = nd;
        s.dofs_left_nd = nl;
    }
    factor_bytes_ = 0;
    max_pieces_ = 0;
    // level-0 blocks
    T* base0 = reinterpret_cast<T*>(arena_.alloc((size_t)std::max<long long>(level0_elems_, 1) * sizeof(T)));
    hip_check(timed(prof_, KF_FILL, st, (double)level0_elems_ * sizeof(T), 0,
                    [&] { return launch_fill<T>(base0, (size_t)level0_elems_, T(0), st); }),
              "fill level0");
    if (nscatter_ > 0)
        hip_check(timed(prof_, KF_SCATTER, st, (double)nscatter_ * (8.0 + 8.0 + sizeof(T)), 0,
                        [&] { return launch_scatter<T>(d_scatter_off_, d_scatter_val_, nscatter_, base0, st); }),
                  "scatter");
    for (size_t e = 0; e < edges_.size(); e++) edges_[e].A = base0 + edge0_off_[e];
    // device size table (room for one sibling per cluster)
    const int cap = (int)(2 * cls_.size() + 16);
    if (cap > d_size_cap_) {
        (void)hipFree(d_size_);
        hip_check(hipMalloc((void**)&d_size_, cap * sizeof(int)), "hipMalloc(size)");
        d_size_cap_ = cap;
    }
    // [0]: failed pivot (cluster order + 1), [1]: a persistent RRQR wait that gave up
    if (!d_info_) hip_check(hipMalloc((void**)&d_info_, 2 * sizeof(int)), "hipMalloc(info)");
    hip_check(hipMemset(d_info_, 0, 2 * sizeof(int)), "memset info");
    R.hsize.assign(cap, 0);
    for (int c : bottoms_[0]) cls_[c].xoff = cls_[c].start;
    ws_len_ = N_;
    // small-path RRQR: Asn in LDS up to RRQR_LDS_MAX; each wave's launch gets the LDS of
    // its largest resident task
    R.lds_elems = RRQR_LDS_MAX / (int)sizeof(T);
    R.big_min_rows = env_or("SPAND_RRQR_STEPPED_MIN_ROWS", 16);
    // SPAND_RRQR_SPIN_MAX lowers the bound on the persistent RRQR's waits (the give-up path's
    // test runs it at 1 poll: the waits fail, the launch ends, factorize reports the error)
    R.spin_max = std::max(1LL, env_or("SPAND_RRQR_SPIN_MAX", RRQR_SPIN_MAX));
    tunes_.assign(nlevels_, LevelTune());
    flops_log_.clear();
    shapes_.assign(nlevels_, LevelShape());
    R.gen = symm_ == 2;
    fsymm_ = symm_;

    for (ilvl_ = 0; ilvl_ < nlevels_; ilvl_++) {
        LevelPlan<T> L;
        L.run = &R;
        L.lst = &stats_[ilvl_];
        L.lsv = &solve_plan_[ilvl_];
        L.t_lvl0 = wall();
        L.bot = &bottoms_[current_bottom_];
        plan_elimination(L);
        launch_elimination(L);
        plan_scaling(L);
        plan_sparsification(L);
        launch_level(L);
        apply_ranks(L);
        merge_level(L);
        plan_level_solve(L);
    }
    for (int l = 0; l < nlevels_; l++)
        if (R.pending_solve[l].valid()) {
            SolvePlanHost h = R.pending_solve[l].get();
            upload_solve_plan(l, &h);
        }
    hip_check(hipDeviceSynchronize(), "factorize");
    // a factorization that outgrew the staging buffer leaves one that holds all its uploads
    if (h_stage_total_ > h_stage_cap_) stage_grow(h_stage_total_ + h_stage_total_ / 8);
    t_factor_ = wall() - t_start;
    price_waves(sizeof(T));
    prof_.collect();
}

// ---------------------------------------------------------------------------------------
int Tree::get_stop() const {
    int stop = N_;
    for (auto& l : stats_) {
        if (l.dofs_left_elim > 0) stop = std::min(stop, l.dofs_left_elim);
        if (l.dofs_left_spars > 0) stop = std::min(stop, l.dofs_left_spars);
    }
    return stop;
}

// Tree::nnz (src/tree.cpp:1637-1643) with the reference's per-op counts.
long long Tree::nnz() const {
    long long n = 0;
    for (auto& o : ops_) {
        switch (o.type) {
            case OP_SCALING_LLT: n += (long long)o.sa * (o.sa + 1) / 2; break;
            case OP_SCALING_PLUQ: n += (long long)o.sa * o.sa + 2LL * o.sa; break;
            case OP_GEMM_SYMM_OUT: case OP_GEMM_SYMM_IN: case OP_GEMM_OUT: case OP_GEMM_IN:
                n += (long long)o.sa * o.sb;
                break;
            case OP_ORTHOGONAL: n += (long long)o.sa * o.sa; break;
            default: break;
        }
    }
    return n;
}

template <typename T>
static void d2h_as_double(const void* src, size_t n, double* dst) {
    std::vector<T> tmp(n);
    if (n) hip_check(hipMemcpy(tmp.data(), src, n * sizeof(T), hipMemcpyDeviceToHost), "read block");
    for (size_t i = 0; i < n; i++) dst[i] = (double)tmp[i];
}

void Tree::op_data(int i, double* M, double* tau) const {
    if (state_ != State::Factorized) throw Error(ERR_STATE, "op data needs a completed factorize()");
    if (i < 0 || i >= (int)ops_.size()) throw Error(ERR_ARG, "op index out of range");
    const OpRec& o = ops_[i];
    const size_t n = (size_t)o.rows * o.cols;
    if (M && n) {
        if (o.M == nullptr) {
            for (size_t k = 0; k < n; k++) M[k] = 0.0;
            for (int k = 0; k < std::min(o.rows, o.cols); k++) M[k + (size_t)k * o.rows] = 1.0;
        } else if (fprec_ == 0) {
            d2h_as_double<double>(o.M, n, M);
        } else {
            d2h_as_double<float>(o.M, n, M);
        }
    }
    if (tau && o.tau && o.cols > 0) {
        if (fprec_ == 0) d2h_as_double<double>(o.tau, o.cols, tau);
        else d2h_as_double<float>(o.tau, o.cols, tau);
    }
}

void Tree::op_plu(int i, double* U, int* perm) const {
    if (state_ != State::Factorized) throw Error(ERR_STATE, "op data needs a completed factorize()");
    if (i < 0 || i >= (int)ops_.size()) throw Error(ERR_ARG, "op index out of range");
    const OpRec& o = ops_[i];
    if (o.type != OP_SCALING_PLUQ) throw Error(ERR_ARG, "op is not a ScalingPLUQ");
    const int n = o.sa;
    std::vector<double> ut((size_t)n * n, 0.0);
    if (o.Ut) {
        if (fprec_ == 0) d2h_as_double<double>(o.Ut, ut.size(), ut.data());
        else d2h_as_double<float>(o.Ut, ut.size(), ut.data());
    } else {
        for (int k = 0; k < n; k++) ut[k + (size_t)k * n] = 1.0;
    }
    if (U)
        for (int c = 0; c < n; c++)
            for (int r = 0; r < n; r++) U[r + (size_t)c * n] = ut[c + (size_t)r * n];
    if (perm) {
        if (o.perm && n > 0) hip_check(hipMemcpy(perm, o.perm, (size_t)n * sizeof(int), hipMemcpyDeviceToHost), "read perm");
        else for (int k = 0; k < n; k++) perm[k] = k;
    }
}

// Columns: order, level, merge level (level - depth, src/cluster.cpp:178-193), parent's
// order (-1: none), original size (its x segment), size (after sparsification), start.
// Every level's bottom exists from the partition on, as in the reference
// (src/tree.cpp:368-412); the sizes are the factorization's once it has run.
int Tree::clusters_info(int* out) const {
    if (state_ == State::Created) throw Error(ERR_STATE, "get_clusters() must follow partition()");
    int k = 0;
    for (size_t l = 0; l < bottoms_.size(); l++)
        for (int c : bottoms_[l]) {
            if (out) {
                const HCluster& h = cls_[c];
                int depth = 0;
                for (int p = h.parent; p >= 0; p = cls_[p].parent) depth++;
                int* o = out + (size_t)CLUSTER_COLS * k;
                o[0] = h.order;
                o[1] = h.level;
                o[2] = h.level - depth;
                o[3] = h.parent >= 0 ? cls_[h.parent].order : -1;
                o[4] = h.osize;
                o[5] = h.size;
                o[6] = h.start;
            }
            k++;
        }
    return k;
}

void Tree::dof2id(int* ids) const {
    if (state_ == State::Created) throw Error(ERR_STATE, "get_dof2ID() must follow partition()");
    for (int c : bottoms_[0]) {
        const HCluster& h = cls_[c];  // level-0 starts and sizes: fixed by the partition
        for (int i = 0; i < h.osize; i++) ids[perm_[h.start + i]] = h.order;
    }
}

long long Tree::trailing_coo(int* I, int* J, double* V) const {
    // Before factorize(): the assembled matrix, from the scatter map. After: the current
    // bottom's blocks (empty once every level is eliminated).
    if (state_ == State::Assembled || state_ == State::Factorized) {
        if (state_ == State::Factorized) {
            long long cnt = 0;
            for (int c : bottoms_[current_bottom_]) {
                for (int e : cls_[c].out) {
                    const HEdge& E = edges_[e];
                    const int m = size(E.n2), n = size(E.n1);
                    std::vector<double> blk((size_t)m * n);
                    if (E.ident) {
                        for (int k = 0; k < std::min(m, n); k++) blk[k + (size_t)k * m] = 1.0;
                    } else if (fprec_ == 0) {
                        d2h_as_double<double>(E.A, blk.size(), blk.data());
                    } else {
                        d2h_as_double<float>(E.A, blk.size(), blk.data());
                    }
                    for (int jj = 0; jj < n; jj++)
                        for (int ii = 0; ii < m; ii++) {
                            int gi = cls_[E.n2].start + ii, gj = cls_[c].start + jj;
                            if (symm_ == 0 && gi < gj) continue;
                            if (I) { I[cnt] = gi; J[cnt] = gj; V[cnt] = blk[ii + (size_t)jj * m]; }
                            cnt++;
                        }
                }
            }
            return cnt;
        }
        const std::vector<long long>& off = h_scatter_off_;
        std::vector<double> dev_val;
        if (host_vals_stale_ && nscatter_ > 0) {  // values re-assembled on the device
            dev_val.resize(nscatter_);
            hip_check(hipMemcpy(dev_val.data(), d_scatter_val_, nscatter_ * sizeof(double), hipMemcpyDeviceToHost),
                      "read values");
        }
        const std::vector<double>& val = host_vals_stale_ ? dev_val : h_scatter_val_;
        // map offsets back to (edge, i, j)
        std::vector<std::pair<long long, int>> starts;
        for (size_t e = 0; e < edges0_.size(); e++) starts.push_back({edge0_off_[e], (int)e});
        std::sort(starts.begin(), starts.end());
        long long cnt = 0;
        for (long long k = 0; k < nscatter_; k++) {
            auto it = std::upper_bound(starts.begin(), starts.end(), std::make_pair(off[k], INT32_MAX));
            --it;
            const HEdge& E = edges0_[it->second];
            const int m = cls0_[E.n2].size;
            long long rel = off[k] - it->first;
            int ii = (int)(rel % m), jj = (int)(rel / m);
            int gi = cls0_[E.n2].start + ii, gj = cls0_[E.n1].start + jj;
            if (symm_ == 0 && gi < gj) continue;
            if (I) { I[cnt] = gi; J[cnt] = gj; V[cnt] = val[k]; }
            cnt++;
        }
        return cnt;
    }
    throw Error(ERR_STATE, "trailing matrix needs assemble()");
}

// ---------------------------------------------------------------------------------------
// The SPD apply is its own transpose: every level's backward sweep applies the transposes
// of its forward sweep's ops in reverse order (L^{-T} after L^{-1}, A_ns^T after A_ns, Q after
// Q^T), so M^{-1} = W^T W for the forward sweep W. A GEN tree runs the transposed phase lists.
void Tree::solve_transpose(double* x, int nrhs, int ldx, bool on_device, hipStream_t st) {
    solve_any(x, nrhs, ldx, on_device, st, fsymm_ != 0);
}

void Tree::solve(double* x, int nrhs, int ldx, bool on_device, hipStream_t st) {
    solve_any(x, nrhs, ldx, on_device, st, false);
}

void Tree::apply_any(double* d_x, int nrhs, size_t sx, size_t sr, hipStream_t st, bool transpose, double* ws_ext) {
    if (fprec_ == 0)
        apply_impl<double>(d_x, nrhs, sx, sr, st, transpose, ws_ext);
    else
        apply_impl<float>(d_x, nrhs, sx, sr, st, transpose, ws_ext);
}

size_t Tree::apply_workspace_bytes(int nrhs) const {
    const size_t seg = (size_t)ws_len_ * nrhs, part = (size_t)max_pieces_ * SOLVE_CHUNK * nrhs;
    return (seg + part) * sizeof(double);
}

void Tree::solve_device_noalloc(double* d_x, int nrhs, int ldx, hipStream_t st) {
    if (state_ != State::Factorized) throw Error(ERR_STATE, "solve() must follow factorize()");
    if (!d_x || nrhs <= 0 || nrhs > 64 || ldx < N_) throw Error(ERR_ARG, "solve: need 1 <= nrhs <= 64 and ld >= N");
    const size_t need = apply_workspace_bytes(nrhs);
    if (need > SCRATCH_BYTES)
        throw Error(ERR_ARG, "solve_device_noalloc: the apply needs " + std::to_string(need >> 20) +
                                 " MiB of workspace, more than the static " + std::to_string(SCRATCH_BYTES >> 20) +
                                 " MiB (use spand_solve_device, or fewer right-hand sides per call)");
    apply_any(d_x, nrhs, 1, (size_t)ldx, st, false, scratch_device_ptr());
}

void Tree::solve_any(double* x, int nrhs, int ldx, bool on_device, hipStream_t st, bool transpose) {
    if (transpose && state_ == State::Factorized) upload_transposed();
    if (state_ != State::Factorized) throw Error(ERR_STATE, "solve() must follow factorize()");
    if (nrhs <= 0 || nrhs > 64 || ldx < N_) throw Error(ERR_ARG, "solve: need 1 <= nrhs <= 64 and ld >= N");
    double* dx = x;
    double* tmp = nullptr;
    int ld = ldx;
    if (!on_device) {
        hip_check(hipMalloc((void**)&tmp, (size_t)N_ * nrhs * sizeof(double)), "hipMalloc(solve x)");
        hip_check(hipMemcpy2D(tmp, (size_t)N_ * sizeof(double), x, (size_t)ldx * sizeof(double),
                              (size_t)N_ * sizeof(double), nrhs, hipMemcpyHostToDevice),
                  "upload x");
        dx = tmp;
        ld = N_;
    }
    apply_any(dx, nrhs, 1, (size_t)ld, st, transpose);
    if (!on_device) {
        hip_check(hipStreamSynchronize(st), "solve");
        hip_check(hipMemcpy2D(x, (size_t)ldx * sizeof(double), tmp, (size_t)N_ * sizeof(double),
                              (size_t)N_ * sizeof(double), nrhs, hipMemcpyDeviceToHost),
                  "download x");
        (void)hipFree(tmp);
    }
}

void Tree::solve_interleaved(double* d_x, int nrhs, hipStream_t st, bool transpose) {
    if (state_ != State::Factorized) throw Error(ERR_STATE, "solve() must follow factorize()");
    if (nrhs <= 0 || nrhs > 64) throw Error(ERR_ARG, "solve: need 1 <= nrhs <= 64");
    transpose = transpose && fsymm_ != 0;  // the SPD apply is its own transpose
    if (transpose) upload_transposed();
    apply_any(d_x, nrhs, (size_t)nrhs, 1, st, transpose);
}

// Tree::solve (src/tree.cpp:1610-1635): permute in, forward sweep over the levels, backward
// sweep in reverse, permute out. Split ops need no work: a sibling's segment aliases the
// tail of its original cluster's segment.
// Instantiated graphs of the apply, by launch configuration; dropped at every factorization
// (the plans they launch are rebuilt there).
void Tree::drop_apply_graphs() {
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
}

template <typename T>
void Tree::apply_impl(double* dx, int nrhs, size_t sx, size_t sr, hipStream_t st_caller, bool transpose,
                      double* ws_ext) {
    // ws: the segments (FP64 whatever the factor's element type T), then the partial sums
    // of the largest mv phase
    const size_t seg = (size_t)ws_len_ * nrhs;
    const size_t need = apply_workspace_bytes(nrhs);
    if (!ws_ext && need > ws_cap_) {
        drop_apply_graphs();  // they address the old workspace
        (void)hipFree(d_ws_);
        d_ws_ = nullptr;
        hip_check(hipMalloc(&d_ws_, need), "hipMalloc(solve workspace)");
        ws_cap_ = need;
    }
    double* ws = ws_ext ? ws_ext : reinterpret_cast<double*>(d_ws_);
    double* partial = ws + seg;
    const double vsz = sizeof(double);
    // The whole apply (a fixed list of about 100-300 launches once the factorization is done)
    // is captured once per configuration in a hipGraph and replayed: not while launches are
    // timed, not for the allocation-free apply (instantiation allocates), and not with
    // SPAND_APPLY_GRAPH=0.
    const bool graph = !ws_ext && !prof_.on() && graphs_on_;
    hipStream_t st = st_caller;
    if (graph) {
        for (auto& g : graphs_)
            if (g.dx == dx && g.nrhs == nrhs && g.sx == sx && g.sr == sr && g.transpose == transpose && g.ws == ws) {
                hip_check(hipGraphLaunch(g.exec, st_caller), "apply graph launch");
                return;
            }
        st = capture_stream();
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "apply graph capture");
    }
    auto run = [&](int fam, const SweepWork& w, const char* what, auto&& launch) {
        if (ws_ext || graph) {  // the allocation-free apply creates no events; a capture times nothing
            hip_check(launch(), what);
            return;
        }
        hip_check(timed(prof_, fam, st, w.mat_bytes + w.vec_elems * nrhs * vsz, w.flops * nrhs, launch), what);
    };
    auto run_phase = [&](const MvPhase& ph) {
        run(ph.family, ph.w, "mv", [&] {
            return launch_mv<T>(ph.groups, ph.chunks, ph.pieces, ph.npieces, ph.terms, ph.rchunks, ph.nreduce, ws,
                                partial, nrhs, st);
        });
    };
    SweepWork wp;
    wp.vec_elems = (double)N_ * 2.0 + (double)N_ * 4.0 / vsz / nrhs;
    run(KF_PERM, wp, "perm in", [&] { return launch_perm_in<T>(dx, sx, sr, nrhs, d_perm_, N_, ws, st); });
    for (int l = 0; l < nlevels_; l++) {
        const LevelSolve& p = solve_plan_[l];
        for (const MvPhase& ph : transpose ? p.tfwd : p.fwd) run_phase(ph);
        if (p.n_merge)
            run(KF_MERGE, p.w_merge, "merge",
                [&] { return launch_merge<T>(p.merge, p.n_merge, p.children, ws, nrhs, 1, st); });
    }
    for (int l = nlevels_ - 1; l >= 0; l--) {
        const LevelSolve& p = solve_plan_[l];
        if (p.n_merge)
            run(KF_MERGE, p.w_merge, "merge",
                [&] { return launch_merge<T>(p.merge, p.n_merge, p.children, ws, nrhs, 0, st); });
        for (const MvPhase& ph : transpose ? p.tbwd : p.bwd) run_phase(ph);
    }
    run(KF_PERM, wp, "perm out", [&] { return launch_perm_out<T>(ws, d_perm_, N_, nrhs, dx, sx, sr, st); });
    if (graph) {
        hipGraph_t g = nullptr;
        hip_check(hipStreamEndCapture(st, &g), "apply graph capture");
        ApplyGraph ag{dx, ws, nrhs, sx, sr, transpose, nullptr};
        const hipError_t e = hipGraphInstantiate(&ag.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "apply graph instantiate");
        if (graphs_.size() >= 8) {  // a few configurations per factorization: the oldest goes
            (void)hipGraphExecDestroy(graphs_.front().exec);
            graphs_.erase(graphs_.begin());
        }
        graphs_.push_back(ag);
        hip_check(hipGraphLaunch(ag.exec, st_caller), "apply graph launch");
    }
}

// Stream that apply graphs are captured on (the null stream cannot be captured).
hipStream_t Tree::capture_stream() {
    if (!cap_st_) hip_check(hipStreamCreateWithFlags(&cap_st_, hipStreamNonBlocking), "hipStreamCreate");
    return cap_st_;
}

}  // namespace spand
