// Host-side spaND tree for one GPU: the reference's Tree / Cluster / Edge API
// (include/tree.h:130-197, include/cluster.h:16-117, include/edge.h:32-44) with the
// numerical blocks moved into device arenas.
//
// The host keeps only the structure (cluster sizes, orders, edge lists in the reference's
// list order) and plans every ND level as a handful of level-batched kernel launches; the
// dense blocks (pivots, edges, factors, Householder vectors) live in device memory, in the
// element type chosen by set_precision (FP64 or FP32).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "profile.hpp"
#include "tasks.hpp"
#include "tune.hpp"

namespace spand {

enum Status {
    OK = 0,
    ERR_ARG = 1,
    ERR_NOT_SPD = 2,
    ERR_SINGULAR = 3,
    ERR_HIP = 4,
    ERR_STATE = 5,
    ERR_NOMEM = 6,
    ERR_NOT_CONVERGED = 7,
    ERR_UNSUPPORTED = 8,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);

// Chunked bump allocator of device memory. reset() keeps the chunks for the next
// factorization, so refactorizations of the same structure do not call hipMalloc.
// A chunk size of 0 (environment SPAND_ARENA_CHUNK_BYTES=0) gives one exact-size
// allocation per block, for bounds-checking builds.
size_t arena_chunk_bytes(size_t dflt);
class DeviceArena {
  public:
    explicit DeviceArena(size_t chunk_bytes = (size_t)256 << 20) : chunk_bytes_(arena_chunk_bytes(chunk_bytes)) {}
    ~DeviceArena();
    void* alloc(size_t bytes);
    void reset();
    size_t bytes_in_use() const { return in_use_; }
    size_t bytes_reserved() const;

  private:
    struct Chunk {
        char* p;
        size_t size, used;
    };
    std::vector<Chunk> chunks_;
    size_t cur_ = 0;
    size_t chunk_bytes_;
    size_t in_use_ = 0;
};

struct HCluster {
    int start = 0, size = 0, osize = 0, level = 0, order = 0;
    bool elim = false, spars = false;
    int parent = -1;
    std::vector<int> children;
    std::vector<int> out;  // edge ids; out[0] is the pivot (self edge)
    std::vector<int> in;   // edge ids of edges (n -> self), n != self
    long long xoff = -1;   // segment of the cluster's x in the solve workspace
};

struct HEdge {
    int n1 = -1, n2 = -1;  // block is A[n2, n1]: size(n2) x size(n1), ld = size(n2)
    void* A = nullptr;     // device block; nullptr together with ident => identity pivot
    bool ident = false;
    bool original = false;
};

// One recorded op, kept for introspection (op-by-op comparison with the CPU oracle).
// Types follow the reference's Operation classes (include/operations.h).
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
struct OpRec {
    int type, a, b, sa, sb, rows, cols, level;
    const void* M;
    const void* tau;
    const void* Ut = nullptr;    // ScalingPLUQ: U^T (rows x rows, lower)
    const int* perm = nullptr;   // ScalingPLUQ: row permutation p
};

// Algorithmic work of one solve launch: matrix bytes read once per apply, vector elements
// moved per right-hand side, flops per right-hand side.
struct SweepWork {
    double mat_bytes = 0, vec_elems = 0, flops = 0;
};

// One mv phase of the preconditioner apply (device task arrays, see MvPiece).
struct MvPhase {
    const MvGroup* groups = nullptr;
    const MvTerm* terms = nullptr;
    const MvChunk* chunks = nullptr;
    const MvPiece* pieces = nullptr;
    const int* rchunks = nullptr;
    int npieces = 0, nreduce = 0;
    int family = 0;  // kernel family for the timer
    SweepWork w;
};

// Per-level solve plan: the forward phases (before the merge), the backward phases (after
// the split back to the children), and the merge itself.
struct HostPhases;  // host-side mv phases awaiting upload (tree.cpp)
template <typename T>
struct LevelPlan;   // one level's task lists during factorize (tree.cpp)

struct LevelSolve {
    std::vector<MvPhase> fwd, bwd;
    std::vector<MvPhase> tfwd, tbwd;  // the adjoint apply: transposed bwd ops going up, fwd ops down
    SweepWork w_merge;
    MergeTask* merge = nullptr; int n_merge = 0;
    int2* children = nullptr;
};

struct LevelStats {
    int dofs_nd = 0, dofs_left_nd = 0, dofs_left_elim = 0, dofs_left_spars = 0;
    long long fact_nnz = 0;
    int rank_before = 0, rank_after = 0, nsparsified = 0, nignored = 0, nwaves = 0;
    double t_plan = 0, t_elim_scale = 0, t_spars = 0, t_merge = 0, t_solve_plan = 0;
    double flops_potrf = 0, flops_trsm = 0, flops_schur = 0, flops_rrqr = 0;
};

class Tree {
  public:
    explicit Tree(int nlevels);
    ~Tree();

    // options (include/tree.h:133-147)
    void set_tol(double t) { tol_ = t; }
    void set_skip(int s) { skip_ = s; }
    void set_verb(bool v) { verb_ = v; }
    void set_symm_kind(int k);
    void set_use_geo(bool g) { geo_ = g; }
    void set_Xcoo(int dim, int N, const double* X);
    void set_precision(int p);
    void set_use_sparsify(bool u) { use_want_sparsify_ = u; }

    void partition(int N, const int* colptr, const int* rowval);
    void assemble(int N, const int* colptr, const int* rowval, const double* vals);
    // New values (device array, CSC order of the assembled pattern) for the next factorize().
    void assemble_values_device(const double* vals, long long nnz, hipStream_t st);
    long long assembled_nnz() const { return nnz_in_; }
    void factorize();
    // M^{-1} applied in place to nrhs columns (ld >= N) of x, on device or host memory.
    void solve(double* x, int nrhs, int ldx, bool on_device, hipStream_t st);
    // M^{-T} applied in place (the adjoint apply). The SPD (LLT) preconditioner is
    // symmetric, so this is solve(); a GEN (PLU) tree runs its transposed phase lists.
    void solve_transpose(double* x, int nrhs, int ldx, bool on_device, hipStream_t st);
    void solve_any(double* x, int nrhs, int ldx, bool on_device, hipStream_t st, bool transpose);
    // M^{-1} applied in place to nrhs device columns (ld >= N) with the segments and partial
    // sums in the library's static device workspace (scratch.hpp): no device allocation, no
    // event; the launches are ordered on st. Throws ERR_ARG when the apply needs more than
    // SCRATCH_BYTES.
    void solve_device_noalloc(double* d_x, int nrhs, int ldx, hipStream_t st);
    // Bytes of workspace an apply on nrhs columns needs (segments + partial sums).
    size_t apply_workspace_bytes(int nrhs) const;
    // M^{-1} (or M^{-T}) applied in place to an interleaved device block (element (i, r) at
    // x[i*nrhs+r]).
    void solve_interleaved(double* d_x, int nrhs, hipStream_t st, bool transpose = false);

    int N() const { return N_; }
    int nlevels() const { return nlevels_; }
    int precision() const { return prec_; }
    int factor_precision() const { return fprec_; }
    int get_stop() const;
    long long nnz() const;
    const std::vector<int>& perm() const { return perm_; }
    const std::vector<int>& part() const { return part_; }
    const std::vector<HCluster>& clusters() const { return cls_; }
    const std::vector<int>& bottom(int l) const { return bottoms_[l]; }
    const std::vector<LevelStats>& stats() const { return stats_; }
    // the launch configuration of each level of the last factorization and the level shape
    // it was chosen from
    const std::vector<LevelTune>& tunes() const { return tunes_; }
    const std::vector<LevelShape>& shapes() const { return shapes_; }
    const std::vector<OpRec>& ops() const { return ops_; }
    // Shape log of the dense calls of the last factorization (Tree::tprof_flops,
    // include/tree.h:165, recorded under set_monitor_flops, src/tree.cpp:592-1312): one
    // record per reference call, in the reference's call order.
    enum ShapeKind { SK_POTF = 0, SK_PANEL = 1, SK_GEMM = 2, SK_SYRK = 3, SK_RRQR = 4, SK_GETF = 5 };
    struct ShapeRec {
        int level, kind, m, n, k;
    };
    void set_monitor_flops(bool on) { monitor_flops_ = on; }
    // Run-time fallbacks taken so far (bit 0: the stepped-RRQR launch throttle stopped after
    // a wait for the device longer than two seconds; later steps launched without it).
    unsigned status_flags() const { return throttle_gave_up_ ? 1u : 0u; }
    const std::vector<ShapeRec>& flops_log() const { return flops_log_; }
    double factor_seconds() const { return t_factor_; }
    size_t factor_bytes() const { return factor_bytes_; }
    size_t arena_bytes() const { return arena_.bytes_in_use(); }
    // Copies op i's block (rows x cols, column-major) and tau (if any) to host, as FP64.
    void op_data(int i, double* M, double* tau) const;
    // ScalingPLUQ op i: U (rows x rows, column-major, FP64) and the row permutation p.
    void op_plu(int i, double* U, int* perm) const;
    // Trailing (current) matrix as COO triplets; returns the count (arrays may be null).
    long long trailing_coo(int* I, int* J, double* V) const;
    // Tree::get_clusters (src/tree.cpp:265-273): every level's bottom, in cluster order, as
    // rows of CLUSTER_COLS ints; returns the row count (out may be null).
    static constexpr int CLUSTER_COLS = 7;
    int clusters_info(int* out) const;
    // Tree::get_dof2ID (src/tree.cpp:275-284): dof -> order of its level-0 cluster
    void dof2id(int* ids) const;
    // Per-launch kernel timing (HIP events on the launch stream); off by default.
    KernelProfiler& profiler() { return prof_; }

  private:
    enum class State { Created, Partitioned, Assembled, Factorized };
    template <typename T> void factorize_impl();
    // the phases of one level of factorize_impl (tree.cpp), in order
    template <typename T> void plan_elimination(LevelPlan<T>& L);
    template <typename T> void launch_elimination(LevelPlan<T>& L);
    template <typename T> void plan_scaling(LevelPlan<T>& L);
    template <typename T> void plan_sparsification(LevelPlan<T>& L);
    template <typename T> void launch_level(LevelPlan<T>& L);
    template <typename T> void apply_ranks(LevelPlan<T>& L);
    template <typename T> void merge_level(LevelPlan<T>& L);
    template <typename T> void plan_level_solve(LevelPlan<T>& L);
    template <typename T>
    void run_potrf(hipStream_t st, void* dt, const std::vector<PotrfTask>& v, size_t off, const char* what);
    template <typename T>
    void run_trsm(hipStream_t st, void* dt, const std::vector<TrsmTask>& v, size_t off, int split, const char* what);
    template <typename T>
    void run_getrf(hipStream_t st, void* dt, const std::vector<GetrfTask>& v, size_t off, const char* what);
    // ws_ext: a caller-provided workspace of apply_workspace_bytes(nrhs) (then nothing is
    // allocated and no launch is timed), or nullptr for the tree's own.
    template <typename T>
    void apply_impl(double* d_x, int nrhs, size_t sx, size_t sr, hipStream_t st, bool transpose = false,
                    double* ws_ext = nullptr);
    void apply_any(double* d_x, int nrhs, size_t sx, size_t sr, hipStream_t st, bool transpose,
                   double* ws_ext = nullptr);
    int size(int c) const { return cls_[c].size; }
    int find_out_edge(int from, int to) const;
    int new_edge(int n1, int n2, void* A, bool original);
    void set_eliminated(int c);
    void* upload_tasks(const struct Packer& pk);
    char* stage_reserve(size_t bytes);
    void stage_grow(size_t cap);
    void stage_upload(void* d, const void* h, size_t bytes);
    char* h_stage_ = nullptr;  // pinned staging of host -> device uploads
    size_t h_stage_cap_ = 0, h_stage_used_ = 0;
    size_t h_stage_total_ = 0;  // bytes staged since the last factorization began (across growths)
    // Fine-grained pinned signals of the stepped RRQR: [0] the sequence number of the latest
    // step launch to start, [1..] one done flag per stepped cluster of the current level.
    void signal_reserve(size_t n);
    void price_waves(size_t esz);
    void upload_solve_plan(int level, void* host_plan);  // host_plan: SolvePlanHost (tree.cpp)
    void upload_transposed();
    // per level, the adjoint apply's phases, uploaded at the first transposed solve
    std::vector<std::unique_ptr<HostPhases>> tplans_;
    void wait_signal(long long target);
    // Profiled stepped-RRQR waves, priced after the factorization from the clusters' final
    // dims (5 ints each, one block per level), so that no copy waits inside the level loop.
    struct WaveWork {
        int slot_init, slot_finish, slot_step, steps;
        size_t first, n;  // the wave's clusters in its level's dims block
        int level;
    };
    std::vector<WaveWork> wave_work_;
    std::vector<std::pair<const int*, size_t>> level_dims_;  // per level: dims block, clusters
    long long* h_sig_ = nullptr;
    long long* d_sig_ = nullptr;  // device address of h_sig_
    size_t h_sig_cap_ = 0;
    long long sig_seq_ = 0;
    bool sig_throttle_ = true;  // off after a wait that timed out
    bool throttle_gave_up_ = false;  // a wait timed out (reported by status_flags)

    // options
    int nlevels_;
    double tol_ = 10.0;  // src/tree.cpp:94
    int skip_ = 0;
    bool verb_ = false;
    int symm_ = 0;       // 0 SPD (LLT), 2 GEN (PLU)
    int fsymm_ = 0;      // symm_ of the last factorization
    bool geo_ = false;
    int dim_ = 0;
    std::vector<double> X_;
    int prec_ = 0;       // 0 FP64, 1 FP32: element type of the next factorization
    int fprec_ = 0;      // element type of the stored factor (prec_ at the last factorize())
    bool use_want_sparsify_ = true;

    State state_ = State::Created;
    int N_ = 0;
    std::vector<int> part_;
    std::vector<int> perm_;
    // structure after assembly (restored at every factorize())
    std::vector<HCluster> cls0_;
    std::vector<HEdge> edges0_;
    // live structure
    std::vector<HCluster> cls_;
    std::vector<HEdge> edges_;
    std::vector<std::vector<int>> bottoms_;
    std::vector<int> others_;
    int current_bottom_ = 0;
    int ilvl_ = 0;

    // assembly: level-0 blocks live in one region of `level0_elems` elements; nonzero k of
    // the permuted matrix lands at offset scatter_off[k] of that region
    std::vector<long long> edge0_off_;  // element offset of each assembled edge block
    long long level0_elems_ = 0;
    std::vector<long long> h_scatter_off_;
    std::vector<long long> h_scatter_src_;  // index of each scattered value in the caller's CSC arrays
    std::vector<double> h_scatter_val_;
    long long* d_scatter_off_ = nullptr;
    long long* d_scatter_src_ = nullptr;
    double* d_scatter_val_ = nullptr;
    long long nnz_in_ = 0;           // stored entries of the assembled CSC
    bool host_vals_stale_ = false;   // values last set on the device (assemble_values_device)
    long long nscatter_ = 0;
    bool device_inputs_ = false;  // scatter map and permutation uploaded
    void upload_inputs();

    // device state
    DeviceArena arena_;
    DeviceArena task_arena_{(size_t)64 << 20};
    int* d_size_ = nullptr;
    int* d_info_ = nullptr;
    int d_size_cap_ = 0;
    int* d_perm_ = nullptr;
    void* d_ws_ = nullptr;
    size_t ws_cap_ = 0;
    long long ws_len_ = 0;   // segments in the workspace (per right-hand side)
    long long max_pieces_ = 0;  // largest mv phase, in pieces (sizes the partial-sum buffer)

    hipStream_t aux_stream();
    hipStream_t aux_st_ = nullptr;
    // the apply as instantiated graphs, one per launch configuration
    struct ApplyGraph {
        double* dx;
        double* ws;
        int nrhs;
        size_t sx, sr;
        bool transpose;
        hipGraphExec_t exec;
    };
    std::vector<ApplyGraph> graphs_;
    bool graphs_on_ = true;
    hipStream_t cap_st_ = nullptr;
    hipStream_t capture_stream();
    void drop_apply_graphs();
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;

    KernelProfiler prof_;
    std::vector<LevelSolve> solve_plan_;
    std::vector<OpRec> ops_;
    std::vector<LevelStats> stats_;
    std::vector<LevelTune> tunes_;
    std::vector<LevelShape> shapes_;
    bool monitor_flops_ = false;
    std::vector<ShapeRec> flops_log_;
    void log_shape(int kind, int m, int n, int k) {
        if (monitor_flops_) flops_log_.push_back({ilvl_, kind, m, n, k});
    }
    double t_factor_ = 0;
    size_t factor_bytes_ = 0;
};

}  // namespace spand
