/*
 * spand.h — C ABI of the MI355X (gfx950) spaND factorization and preconditioner.
 *
 * This is the drop-in boundary for the reference's Tree API and its Krylov callers
 * (leopoldcambier/spaND_public). Each entry point names the reference interface it
 * replaces. Plain pointers and sizes only; no C++ exceptions cross this boundary: every
 * call returns a spand_status and spand_last_error() holds the message of the last
 * failure on that handle (the reference throws std::runtime_error or asserts).
 *
 * Threading: a tree handle is not reentrant (like the reference, whose const solve()
 * mutates per-cluster buffers, src/cluster.cpp:82-96). Use one handle per host thread.
 * Matrices are CSC with 0-based int32 indices and sorted rows, the layout of
 * Eigen::SparseMatrix<double, ColMajor, int> (include/spaND.h:59). Right-hand sides are
 * column-major with a leading dimension.
 */
#ifndef SPAND_H
#define SPAND_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SPAND_OK = 0,
    SPAND_ERR_ARG = 1,
    SPAND_ERR_NOT_SPD = 2,        /* non-SPD pivot: "Error: Non-SPD Pivot" (src/tree.cpp:587-590) */
    SPAND_ERR_SINGULAR = 3,       /* singular pivot (src/tree.cpp:625-628) */
    SPAND_ERR_HIP = 4,            /* HIP runtime failure */
    SPAND_ERR_STATE = 5,          /* call out of order (partition -> assemble -> factorize -> solve) */
    SPAND_ERR_NOMEM = 6,          /* device allocation failed */
    SPAND_ERR_NOT_CONVERGED = 7,  /* Krylov iteration cap reached before the tolerance */
    SPAND_ERR_UNSUPPORTED = 8     /* option not built (algebraic partitioning, GEN/PLU, ...) */
} spand_status;

enum { SPAND_SYMM_SPD = 0, SPAND_SYMM_SYM = 1, SPAND_SYMM_GEN = 2 }; /* SymmKind, include/spaND.h:21 */
enum { SPAND_F64 = 0, SPAND_F32 = 1 };                               /* factor storage / arithmetic */

typedef struct spand_tree spand_tree;
typedef struct spand_matrix spand_matrix;

/* Tree::Tree(int nlevels)                          include/tree.h:163 */
spand_status spand_create(int nlevels, spand_tree** out);
void spand_destroy(spand_tree* t);
/* Message of the last failing call on t (or of the last failing spand_create/matrix call
 * when t is NULL). */
const char* spand_last_error(const spand_tree* t);

/* Tree::set_tol / set_skip / set_verb / set_symm_kind / set_use_geo / set_Xcoo /
 * set_use_sparsify                                  include/tree.h:133-147 */
spand_status spand_set_tol(spand_tree* t, double tol);
spand_status spand_set_skip(spand_tree* t, int skip);
spand_status spand_set_verb(spand_tree* t, int verb);
/* symm_kind: 0 SPD (ScalingKind::LLT), 2 GEN (ScalingKind::PLU); 1 (SYM) is not built. */
spand_status spand_set_symm_kind(spand_tree* t, int symm_kind);
spand_status spand_set_use_geo(spand_tree* t, int use_geo);
/* X is dim x N, column-major (the reference's Eigen::MatrixXd Xcoo). Copied. */
spand_status spand_set_Xcoo(spand_tree* t, int dim, int N, const double* X);
spand_status spand_set_use_sparsify(spand_tree* t, int use);
/* Tree::set_monitor_flops(bool)   include/tree.h:147, src/tree.cpp:129: record the shape of
 * every dense call of the next factorizations (spand_get_flops_log). Off by default. */
spand_status spand_set_monitor_flops(spand_tree* t, int on);
/* No reference counterpart: element type of the factorization (SPAND_F64 / SPAND_F32).
 * FP32 factors are corrected by the FP64 Krylov iteration. */
spand_status spand_set_precision(spand_tree* t, int precision);

/* Tree::partition(SpMat&) (include/tree.h:169, src/tree.cpp:306-419). The pattern must be
 * symmetric (pass |A|+|A^T|+I as tests/spaND.cpp:237 does). part_out (nullable) receives
 * 6 ints per dof: self.lvl, self.sep, left.lvl, left.sep, right.lvl, right.sep
 * (the returned std::vector<ClusterID>). */
spand_status spand_partition(spand_tree* t, int N, const int* colptr, const int* rowval, int* part_out);
/* Tree::assemble(SpMat&)                           include/tree.h:175, src/tree.cpp:505-575
 * May also be called after an assembly or a factorization, with new values of the
 * partitioned pattern (re-assembly; then factorize again). */
spand_status spand_assemble(spand_tree* t, int N, const int* colptr, const int* rowval, const double* vals);
/* Tree::assemble(SpMat&) with new values only (include/tree.h:175), read from device memory:
 * d_vals holds the nnz stored entries in the CSC order given to the last spand_assemble
 * (same pattern). The copy is ordered on `stream` and complete on return; the next
 * spand_factorize uses the new values. For callers whose values live on the GPU (a
 * differentiable solve's parameters). */
spand_status spand_assemble_values_device(spand_tree* t, const double* d_vals, long long nnz, void* stream);
/* Tree::factorize()                                include/tree.h:179, src/tree.cpp:1447-1551
 * May be called again after a completed factorization (refactorization of the assembled
 * matrix; no device allocation once the arenas have grown). */
spand_status spand_factorize(spand_tree* t);
/* Tree::solve(VectorXd&) const                     include/tree.h:185, src/tree.cpp:1610-1635
 * x <- M^{-1} x in place for nrhs columns of a host array (ld >= N). */
spand_status spand_solve(spand_tree* t, double* x, int nrhs, int ld);
/* Same on device memory, ordered on `stream` (a hipStream_t; NULL = default stream). */
spand_status spand_solve_device(spand_tree* t, double* d_x, int nrhs, int ld, void* stream);
/* No counterpart in the reference: spand_solve_device for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The apply's workspace (segments and partial sums, 8 * nrhs * (N + 64 * pieces)
 * bytes) lives in a static device array of 64 MiB that every GPU the library runs on holds
 * until the process ends; a call needing more returns SPAND_ERR_ARG and does nothing. No
 * device allocation and no event creation happen in the call. Calls that overlap on one
 * GPU share the array and race: serialize them (one stream, or host synchronization). */
spand_status spand_solve_device_noalloc(spand_tree* t, double* d_x, int nrhs, int ld, void* stream);
/* x <- M^{-T} x, the adjoint apply of a differentiable solve. No counterpart in the
 * reference (its Tree has no transposed solve); for callers that differentiate through a
 * solve. The SPD (LLT) preconditioner is symmetric, so this equals spand_solve; a GEN
 * (PLU) tree applies the transposed ops in reverse order. */
spand_status spand_solve_transpose(spand_tree* t, double* x, int nrhs, int ld);
spand_status spand_solve_transpose_device(spand_tree* t, double* d_x, int nrhs, int ld, void* stream);

/* Tree::get_N / get_nlevels / get_stop / nnz / get_assembly_perm   include/tree.h:150-160 */
spand_status spand_get_N(const spand_tree* t, int* N);
spand_status spand_get_nlevels(const spand_tree* t, int* nlevels);
spand_status spand_get_stop(const spand_tree* t, int* stop);
spand_status spand_nnz(const spand_tree* t, long long* nnz);
spand_status spand_get_assembly_perm(const spand_tree* t, int* perm);
/* Tree::get_clusters   include/tree.h:156, src/tree.cpp:265-273 (what write_clusters,
 * write_merging and write_stats of src/tree.cpp:29-55 print). Rows of 7 ints, one per cluster
 * of every level's bottom in the reference's list order: order, level, merge level (level -
 * depth, src/cluster.cpp:178-193), parent's order (-1: none), original size, size (the rank
 * kept), start. Every level's bottom is listed from partition on (the reference builds them
 * there); sizes are the factorization's once it has run. Call with out == NULL to get the row
 * count in *count first. */
spand_status spand_get_clusters(const spand_tree* t, int* out, int* count);
/* Tree::get_dof2ID   include/tree.h:157, src/tree.cpp:275-284: ids[dof] = order of the level-0
 * cluster holding dof (N ints). Valid from partition on. */
spand_status spand_get_dof2ID(const spand_tree* t, int* ids);
/* Tree::get_trailing_mat (src/tree.cpp:1730-1763) as COO of the lower triangle (SPD).
 * Call with I == NULL to get the count in *nnz first. */
spand_status spand_get_trailing_coo(const spand_tree* t, long long* nnz, int* I, int* J, double* V);

/* Per-level log (Tree::log, include/util.h:366-401 and print_log, src/tree.cpp:1645-1727):
 * 20 doubles per level:
 *  0 dofs_nd, 1 dofs_left_nd, 2 dofs_left_elim, 3 dofs_left_spars, 4 fact_nnz,
 *  5 rank_before (sum), 6 rank_after (sum), 7 #sparsified, 8 #ignored, 9 #dependency waves,
 * 10 t_plan, 11 t_elim+scale+rrqr (device), 12 t_rank_apply, 13 t_merge (seconds),
 * 14 potrf flops, 15 trsm flops, 16 schur flops, 17 rrqr flops, 18 t_solve_plan (seconds,
 * host build of the level's preconditioner-apply plan), 19 reserved */
spand_status spand_get_level_stats(const spand_tree* t, double* out);
/* Tree::tprof_flops (include/tree.h:165, include/util.h:311-342; written by write_log_flops,
 * src/tree.cpp:60-77): the shape of every dense call of the last factorization, recorded when
 * spand_set_monitor_flops is on, in the reference's call order. 5 ints per record: level,
 * kind (0 Cholesky pivot, 1 panel triangular solve, 2 gemm, 3 syrk, 4 geqp3, 5 LU pivot:
 * the reference's pivot = 0 and 5, panel = 1, gemm = 2 and 3, rrqr = 4), then the
 * reference's (rows, cols, inner): pivot (n, 0, 0), panel (neighbour size, pivot size, 0),
 * gemm / syrk (target rows, target cols, pivot size), rrqr (Asn rows, Asn cols, 0). The
 * reference's per-call wall time has no counterpart: the calls run batched per level
 * (spand_kernel_profile gives the kernel families' times). Call with out == NULL to get the
 * count first. */
spand_status spand_get_flops_log(const spand_tree* t, long long* count, int* out);
/* No counterpart in the reference: fallbacks the library took at run time, as bits of *flags.
 * Bit 0: the stepped RRQR's launch throttle (the host keeps at most 12 step launches ahead of
 * the device by polling a progress counter) gave up after a wait longer than two seconds;
 * later steps were launched without it (same results, possibly more empty step launches). */
spand_status spand_get_status_flags(const spand_tree* t, unsigned* flags);
/* Wall seconds of the last spand_factorize, device bytes of the stored factor, of the arena. */
spand_status spand_get_factor_info(const spand_tree* t, double* seconds, long long* factor_bytes,
                                   long long* arena_bytes);

/* No counterpart in the reference: the launch configuration of every level of the last
 * factorization, chosen from the level's size histogram by a table measured on gfx950 (or
 * set through the environment). 13 ints per level: the shape (pivots eliminated, median and
 * largest pivot, clusters sparsified, median and largest Asn rows, median and largest Asn
 * columns), then the configuration (register-path RRQR on, Asn elements up to which the
 * one-workgroup RRQR runs, wavefronts per stepped slab (0: by slab count), trsm chunk policy,
 * source: 0 table, 1 environment). No choice changes the arithmetic. */
spand_status spand_get_level_tune(const spand_tree* t, int* out);
/* The table's configuration (5 ints as above) for a level shape (8 ints as above) and a
 * factor element size of elem_bytes (4 or 8). Host only. */
spand_status spand_tune_choose(int elem_bytes, const int* shape, int* out);

/* Kernel timing (the reference's timer log, include/util.h:366-401, at kernel granularity):
 * for every kernel family whose bit is set in family_mask (bit f = family f; 0 turns timing
 * off, the default), each launch of the tree (factorization, preconditioner apply) and of
 * spand_pcg is bracketed by two HIP events on its stream; the stepped RRQR's steps are
 * bracketed per dependency wave (one event pair around all steps of the wave).
 * spand_kernel_profile waits for the recorded events and returns, for kernel family
 * `family` (0 <= family < spand_kernel_family_count()): launches, summed event time (ms),
 * algorithmic bytes and flops of those launches. */
spand_status spand_set_profiling(spand_tree* t, unsigned long long family_mask);
spand_status spand_profile_reset(spand_tree* t);
int spand_kernel_family_count(void);
const char* spand_kernel_family_name(int family);
spand_status spand_kernel_profile(spand_tree* t, int family, long long* launches, double* ms, double* bytes,
                                  double* flops);

/* Recorded operations (the Operation list, include/operations.h), for op-by-op comparison.
 * 8 ints per op: type (0 ScalingLLT, 1 ScalingPLUQ, 2 GemmSymmOut, 3 GemmSymmIn, 4 GemmOut,
 * 5 GemmIn, 6 Orthogonal, 7 Merge, 8 Split), order(a), order(b), size(a), size(b), rows,
 * cols, level. */
spand_status spand_num_ops(const spand_tree* t, int* nops);
spand_status spand_ops_info(const spand_tree* t, int* info);
/* Copies op i's matrix (rows x cols, column-major, as FP64) and its tau (Orthogonal). */
spand_status spand_op_data(const spand_tree* t, int i, double* M, double* tau);
/* ScalingPLUQ op i (include/operations.h, ScalingPLUQ): U (size x size, column-major, as
 * FP64) and the row permutation p (P^T A = L U, row i of P^T A is row p[i] of A); its L is
 * the op's matrix of spand_op_data. */
spand_status spand_op_plu(const spand_tree* t, int i, double* U, int* perm);

/* Device copy of a square CSC matrix, used as the Krylov operator (the SpMat argument of
 * cg / gmres, include/is.h:12-14). */
spand_status spand_matrix_create(int N, const int* colptr, const int* rowval, const double* vals,
                                 spand_matrix** out);
void spand_matrix_destroy(spand_matrix* A);
/* New values of the matrix's pattern from device memory (nnz entries, the CSC order given to
 * spand_matrix_create), ordered on `stream` and complete on return. No counterpart in the
 * reference (its SpMat is rebuilt on the host). */
spand_status spand_matrix_set_values_device(spand_matrix* A, const double* d_vals, long long nnz, void* stream);
/* vals_out[k] = alpha * sum_r u[i_k + r*ld] * v[j_k + r*ld] for every stored entry k = (i_k, j_k)
 * of A, in the CSC order given to spand_matrix_create: with alpha = -1, u the adjoint and v
 * the solution of A x = b, the gradient of a loss with respect to A's stored values. No
 * counterpart in the reference; for callers that differentiate through a solve. u, v,
 * vals_out are device pointers when on_device != 0. */
spand_status spand_matrix_sampled_outer(const spand_matrix* A, const double* u, const double* v, int nrhs, int ld,
                                        double alpha, double* vals_out, int on_device);

typedef struct {
    int iterations;       /* max over right-hand sides of the reference's count (i+1) */
    int converged;        /* number of columns that met the tolerance */
    double relres_max;    /* max over columns of the true ||b - A x|| / ||b|| (FP64) */
    double seconds;       /* wall time of the solve */
    double seconds_precond;
    double seconds_spmv;
    int applies;          /* preconditioner applications */
} spand_krylov_stats;

/* cg(mat, rhs, x, precond, iters, tol)             include/is.h:13, src/is.cpp:39-121
 * Preconditioned CG on nrhs right-hand sides at once, FP64 throughout (the
 * preconditioner's own precision is the tree's). Stops a column when its recursive
 * residual satisfies ||r||^2 < tol^2 ||b||^2, as the reference does. b and x (in: x0,
 * out: solution) are column-major with leading dimension ld, on the host (on_device = 0)
 * or on the device (on_device = 1). iters_out (nullable, nrhs ints) receives each
 * column's count. */
spand_status spand_pcg(spand_tree* t, const spand_matrix* A, const double* b, double* x, int nrhs, int ld,
                       int maxit, double tol, int on_device, int* iters_out, spand_krylov_stats* stats);

/* gmres(mat, rhs, x, precond, iters, restart, tol, verb)   include/is.h:13, src/is.cpp:123-300
 * Left-preconditioned GMRES(restart) with Householder Arnoldi, FP64, on nrhs right-hand
 * sides in lockstep. A column stops when |w_k| / ||M^{-1}(b - A x0)|| < tol (the
 * reference's preconditioned-residual test, with the initial norm kept across restarts)
 * or after maxit iterations; stats->relres_max is the true ||b - A x|| / ||b||. Same
 * argument conventions as spand_pcg. */
spand_status spand_gmres(spand_tree* t, const spand_matrix* A, const double* b, double* x, int nrhs, int ld,
                         int maxit, int restart, double tol, int on_device, int* iters_out,
                         spand_krylov_stats* stats);

/* The adjoint solve A^T x = b: gmres above on the transposed operator with the transposed
 * preconditioner M^{-T} (spand_solve_transpose). No counterpart in the reference; for the
 * backward pass of a differentiable nonsymmetric solve. Same arguments as spand_gmres. */
spand_status spand_gmres_transpose(spand_tree* t, const spand_matrix* A, const double* b, double* x, int nrhs, int ld,
                                   int maxit, int restart, double tol, int on_device, int* iters_out,
                                   spand_krylov_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* SPAND_H */
