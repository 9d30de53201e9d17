// Host-side launchers of the HIP kernels (kernels_factor.hip, kernels_solve.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "tasks.hpp"

namespace spand {

// Dynamic LDS granted to the staged factorization kernels (bytes). Static LDS of the
// RRQR kernel (about 27 KB) comes on top; the total stays below 64 KB so that two
// workgroups can share a CU.
constexpr int LDS_BYTES = 32768;

template <typename T> hipError_t launch_fill(T* p, size_t n, T v, hipStream_t st);
// base[off[k]] = (T) val[k]: level-0 assembly of the permuted matrix into its blocks.
template <typename T>
hipError_t launch_scatter(const long long* off, const double* val, long long n, T* base, hipStream_t st);
template <typename T> hipError_t launch_potrf(const PotrfTask* d_tasks, int ntasks, int* d_info, hipStream_t st);
// Pivots larger than POTRF_SMALL_MAX: one panel of 32 columns per call pair, kb = 0, 32, ...
// (both grids are sized by nmax, the largest pivot of the batch).
constexpr int POTRF_SMALL_MAX = 128;
// Pivots of at most 64 rows: one wavefront per pivot.
template <typename T> hipError_t launch_potrf64(const PotrfTask* d_tasks, int ntasks, int* d_info, hipStream_t st);
template <typename T>
hipError_t launch_potrf_panel(const PotrfTask* d_tasks, int ntasks, int kb, int nmax, int* d_info,
                              hipStream_t st);
template <typename T>
hipError_t launch_potrf_trail(const PotrfTask* d_tasks, int ntasks, int kb, int nmax, hipStream_t st);
// FP32 tasks from index mfma_from on (triangles longer than TRSM_MFMA_MIN_N, placed last by
// the caller) run their trailing updates on the matrix cores; the others, and every FP64
// task, run the ordered-FMA vector kernel.
constexpr int TRSM_MFMA_MIN_N = 64;
template <typename T> hipError_t launch_trsm(const TrsmTask* d_tasks, int ntasks, int mfma_from, hipStream_t st);
// LU pivots (GEN / PLU): in_lds = 1 runs the batch in LDS (every task must satisfy
// n (n + 1) sizeof(T) <= GETRF_LDS_BYTES, nmax the largest n), in_lds = 0 in global memory.
constexpr int GETRF_LDS_BYTES = 65536;
template <typename T>
hipError_t launch_getrf(const GetrfTask* d_tasks, int ntasks, int in_lds, int nmax, int* d_info, hipStream_t st);
template <typename T>
hipError_t launch_schur(const SchurTile* d_tiles, int ntiles, const SchurContrib* d_contribs, hipStream_t st);
// Small clusters, one workgroup each. lds_bytes (<= RRQR_LDS_MAX) holds a task's Asn when
// rows * cols fits; larger tasks carry a global scratch.
constexpr int RRQR_LDS_MAX = 98304;
template <typename T>
hipError_t launch_rrqr(const RrqrTask* d_tasks, int ntasks, const RrqrSeg* d_segs, int* d_size, double tol,
                       int lds_bytes, hipStream_t st);
// Register-resident RRQR (k_rrqr_reg): one workgroup per cluster, the cluster's Asn held in
// registers for the whole factorization. rrqr_reg_variant gives the kernel form that holds
// a rows x cols Asn of elements of esz bytes (-1: none does); a batch is launched with a
// variant that holds each of its tasks.
int rrqr_reg_variant(size_t esz, int rows, int cols);
template <typename T>
hipError_t launch_rrqr_reg(const RrqrTask* d_tasks, int ntasks, const RrqrSeg* d_segs, int* d_size, double tol,
                           int variant, hipStream_t st);
// Stepped RRQR of the large clusters of one wave: init, then steps 0..nsteps-1, then finish.
// lds_bytes is the dynamic LDS of the pivot-column / reflector buffers (2 x max_rows
// elements for a step, max_rows for finish), or 0 when the tasks carry global buffers
// (RrqrBig::vbuf) because they exceed STEP_LDS_MAX.
constexpr int STEP_LDS_MAX = 40960;
template <typename T>
hipError_t launch_rrqr_init(const RrqrBig* d_bigs, const RrqrSlab* d_slabs, int nslabs, const RrqrSeg* d_segs,
                            const int* d_size, double tol, hipStream_t st);
// All steps of a set of whole clusters in one cooperative launch (clusters of at most
// WAVE_ROWS_MAX rows, RrqrBig::vbuf holding rrqr_relay_bytes of relay words, relay = 1);
// lds_bytes = max_cols ints (the position map; the rest of the kernel's LDS is static).
// The slabs of each cluster exchange their candidates through the relay words (waits of at
// most spin_max polls, RRQR_SPIN_MAX by default: *fail is set to 1 if one gives up).
// Returns hipErrorCooperativeLaunchTooLarge, launching nothing, when the slabs do not fit
// the device at once; rrqr_wave_capacity gives how many do.
constexpr int WAVE_ROWS_MAX = 512;
constexpr long long RRQR_SPIN_MAX = 1ll << 21;
constexpr int WAVE_LDS_MAX = 49152;
inline size_t rrqr_relay_bytes(size_t esz, int nslab, int max_rows) {
    return ((size_t)2 * nslab * 4 + (size_t)2 * nslab * max_rows * (esz / 4)) * 8;
}
template <typename T>
hipError_t launch_rrqr_wave(const RrqrBig* d_bigs, const RrqrSlab* d_slabs, int nslabs, int nsteps, int max_rows,
                            double tol, size_t lds_bytes, int* fail, long long spin_max, hipStream_t st);
template <typename T>
hipError_t rrqr_wave_capacity(int max_rows, size_t lds_bytes, int* cap);
// One pivot step per launch. waves: wavefronts per slab workgroup (4, 8 or 16; 0 picks by
// the slab count, with at most 64 cached rows per thread). prog (host-visible, or nullptr)
// receives seq when the launch starts; each cluster's RrqrBig::done is set by its last
// step that does work.
constexpr int RRQR_STEP_SLABS16 = 384;  // up to this many slabs: 16 wavefronts each
constexpr int RRQR_STEP_SLABS8 = 768;   // up to this many: 8, above: 4
int rrqr_step_waves(int nslabs, int max_rows, int waves);
template <typename T>
hipError_t launch_rrqr_step(const RrqrBig* d_bigs, const RrqrSlab* d_slabs, int nslabs, int step, int max_rows,
                            double tol, size_t lds_bytes, int waves, long long* prog, long long seq, hipStream_t st);
template <typename T>
hipError_t launch_rrqr_finish(const RrqrBig* d_bigs, const RrqrSlab* d_slabs, int nslabs, const RrqrSeg* d_segs,
                              int* d_size, double tol, size_t lds_bytes, hipStream_t st);
template <typename T> hipError_t launch_copy(const CopyTask* d_tasks, int ntasks, hipStream_t st);

// Preconditioner sweeps. ws holds the per-cluster segments, interleaved over nrhs.
// Element (i, r) of the caller's block is at b[i * sx + r * sr].
// The segments are FP64 for either factor type T (T selects the factor blocks' type).
template <typename T>
hipError_t launch_perm_in(const double* b, size_t sx, size_t sr, int nrhs, const int* perm, int N, double* ws,
                          hipStream_t st);
template <typename T>
hipError_t launch_perm_out(const double* ws, const int* perm, int N, int nrhs, double* x, size_t sx, size_t sr,
                           hipStream_t st);
hipError_t launch_cm2il(const double* src, int ld, int N, int nrhs, double* dst, hipStream_t st);
hipError_t launch_il2cm(const double* src, int N, int nrhs, double* dst, int ld, hipStream_t st);
// mv phases of the preconditioner apply: SOLVE_CHUNK outputs per chunk, pieces of at most
// MV_PIECE_WORK inner elements and MV_PIECE_TERMS terms; partial holds npieces x
// SOLVE_CHUNK x nrhs doubles.
constexpr int SOLVE_CHUNK = 64;
constexpr int MV_PIECE_WORK = 64;
constexpr int MV_PIECE_TERMS = 64;
template <typename T>
hipError_t launch_mv(const MvGroup* d_groups, const MvChunk* d_chunks, const MvPiece* d_pieces, int npieces,
                     const MvTerm* d_terms, const int* d_rchunks, int nreduce, double* ws, double* partial, int nrhs,
                     hipStream_t st);
// Explicit V, S and the identity X of the Orthogonal ops of one level (WyTask); tiles are
// the (task, a0, c0) 32 x 32 tiles of the Gram matrices (a0 <= c0). X = S^{-1} is left to
// launch_trsm.
template <typename T>
hipError_t launch_wy(const WyTask* d_tasks, int ntasks, const int4* d_tiles, int ntiles, hipStream_t st);
template <typename T>
hipError_t launch_merge(const MergeTask* d_tasks, int ntasks, const int2* d_children, double* ws, int nrhs, int fwd,
                        hipStream_t st);

// FP64 Krylov vector kernels (interleaved N x nrhs).
hipError_t launch_spmv(int N, const int* rowptr, const int* colidx, const double* val, const double* x, double* y,
                       int nrhs, hipStream_t st);
// out[csc_of[p]] = alpha * sum_r u(i_p, r) v(j_p, r) over the stored entries p of a CSR matrix.
hipError_t launch_sampled_outer(int N, const int* rowptr, const int* colidx, const int* csc_of, const double* u,
                                const double* v, int nrhs, int ld, double alpha, double* out, hipStream_t st);
// partial[b * nrhs + r] = sum over block b of a[i,r] * b[i,r]; out[r] = sum_b partial
hipError_t launch_dot(int N, int nrhs, const double* a, const double* b, double* partial, double* out,
                      hipStream_t st);
// y[:, r] = x[:, r] * sx + y[:, r] * sy with per-column device scalars (nullptr => 1).
hipError_t launch_axpby(int N, int nrhs, const double* x, const double* ax, double sgn_ax, double* y,
                        const double* by, hipStream_t st);
// alpha = active ? num / den : 0
hipError_t launch_ratio(int nrhs, const double* num, const double* den, const int* active, double* out,
                        hipStream_t st);
hipError_t launch_copy_f64(double* dst, const double* src, size_t n, hipStream_t st);
// dst[k] = src[idx[k]] for k < n (FP64 values of a fixed pattern, reordered)
hipError_t launch_gather_f64(const long long* idx, const double* src, long long n, double* dst, hipStream_t st);
hipError_t launch_gather_f64(const int* idx, const double* src, long long n, double* dst, hipStream_t st);
// GMRES: out = base + [ye; 0] - sum_{j<k} V_j z_j; W = V^T v for k blocks (partial holds
// 256 * k * nrhs doubles); Householder vector of step k.
hipError_t launch_gm_comb(int N, int nrhs, const double* base, const double* ye, int ny, const double* V, int k,
                          const double* z, double* out, hipStream_t st);
hipError_t launch_gm_multidot(int N, int nrhs, int k, const double* V, const double* v, double* partial, double* out,
                              hipStream_t st);
hipError_t launch_gm_householder(int N, int nrhs, int k, const double* v, const double* d, double* u, hipStream_t st);

}  // namespace spand
