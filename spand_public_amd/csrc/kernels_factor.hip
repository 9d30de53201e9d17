// Level-batched factorization kernels for gfx950 (CDNA4): one workgroup of 256 threads
// (four 64-wide wavefronts) per task; a launch covers every cluster or edge of one ND
// level, of uneven sizes. Task layouts are in tasks.hpp.
//
// Reference kernels replaced (all FP64 LAPACK/BLAS calls on column-major Eigen blocks):
//   potrf  -> LAPACKE_dpotrf('L')             src/util.cpp:158-165
//   trsm   -> cblas_dtrsm Left/Lower/NoTrans,  src/util.cpp:266-274
//             cblas_dtrsm Right/Lower/Trans    src/util.cpp:256-264
//   schur  -> cblas_dgemm / cblas_dsyrk('L')   src/util.cpp:122-156 via gemm_edges
//             (src/tree.cpp:748-793)
//   rrqr   -> LAPACKE_dgeqp3 + choose_rank     src/util.cpp:381-394, 434-452, fused with
//             assemble_Asn and the scatter of shrink_split_scatter_phi
//             (src/tree.cpp:1189-1224, 1292-1347, 980-1104)
//   copy   -> merge block copies                src/tree.cpp:1133-1184
#include <hip/hip_cooperative_groups.h>

#include <cstdlib>
#include <hip/hip_runtime.h>

#include <cfloat>
#include <algorithm>
#include <cstdint>
#include <limits>

#include "kernels.hpp"
#include "tasks.hpp"

namespace spand {

// Device-memory pointers that the task records carry as plain pointers: accessed through a
// pointer in the global address space, the compiler emits global_* instead of flat_*
// instructions (a flat access also holds the LDS wait counter, so an LDS wait would wait
// for every global load in flight as well).
template <typename X>
using gptr = __attribute__((address_space(1))) X*;
template <typename X>
__device__ __attribute__((always_inline)) inline gptr<X> gp(X* p) {
    return (gptr<X>)p;
}

constexpr int NT = 256;

// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(NT) k_fill(T* p, size_t n, T v) {
    size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
    size_t stride = (size_t)gridDim.x * NT;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
__global__ void __launch_bounds__(NT) k_scatter(const long long* off, const double* val, long long n, T* base) {
    for (long long k = (long long)blockIdx.x * NT + threadIdx.x; k < n; k += (long long)gridDim.x * NT)
        base[off[k]] = (T)val[k];
}

// ---------------------------------------------------------------------------------------
// potrf: right-looking Cholesky of the lower triangle in 32-column blocks. Every element
// receives its updates in k order as A(a,b) = fma(-L(a,k), L(b,k), A(a,b)), column j is
// scaled by 1/L(j,j) after its own updates: the element recurrence of the oracle, so FP64
// results are bit-identical to it. The diagonal block lives in LDS; panel and trailing
// updates read the block from global memory (L2 resident).
constexpr int PB = 32;

// Rows i >= kb + nb of panel kb, one row per thread: column k is scaled by 1/L(k,k), then
// the later panel columns receive fma(-a_k, L(k2,k), a_k2), in k order: the element
// recurrence of the unblocked right-looking factorization. D has 2 PB rows, the last PB of
// them zero.
template <typename T, typename P>
__device__ inline void panel_rows(P A, int n, int kb, int nb, const T (*D)[PB + 1], const T* rdiag, int i0, int i1) {
#pragma clang fp contract(off)
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const auto Ai = A + i + (size_t)kb * n;  // row i of the panel, held in registers
        T x[PB];
#pragma unroll
        for (int k = 0; k < PB; k++) x[k] = k < nb ? Ai[(size_t)k * n] : T(0);
        if (nb == PB) {
            // a full panel (every panel with rows below it is one): column k is always x[0],
            // and each step shifts the row down by one register (x[q - 1] takes the updated
            // x[q]), so every register index is a constant while the loop over the columns
            // stays rolled. Column k + q reads D[k + q][k]; for k + q >= PB that is a zero
            // pad row of D, and the slot it updates is never stored.
#pragma clang loop unroll(disable)
            for (int k = 0; k < PB; k++) {
                const T a = x[0] * rdiag[k];
                Ai[(size_t)k * n] = a;
                const T* Dk = &D[k][k];  // D[k + q][k] = Dk[q * (PB + 1)]
#pragma unroll
                for (int q = 1; q < PB; q++) x[q - 1] = fma(-a, Dk[q * (PB + 1)], x[q]);
                x[PB - 1] = T(0);
            }
            continue;
        }
        // a partial panel: a rolled loop over the columns, x indexed by the unrolled q / k2
        // only
#pragma clang loop unroll(disable)
        for (int k = 0; k < nb; k++) {
            const T r = rdiag[k];
            T a = T(0);
#pragma unroll
            for (int q = 0; q < PB; q++) {
                const T xs = x[q] * r;
                a = q == k ? xs : a;
                x[q] = q == k ? xs : x[q];
            }
            // every block entry of column k is read (in bounds) and selected, no branches:
            // the reads then issue together instead of one round trip per row
#pragma unroll
            for (int k2 = 1; k2 < PB; k2++) {
                const T d = D[k2][k];
                const T u = fma(-a, d, x[k2]);
                x[k2] = (k2 > k && k2 < nb) ? u : x[k2];
            }
        }
#pragma unroll
        for (int k = 0; k < PB; k++)
            if (k < nb) Ai[(size_t)k * n] = x[k];
    }
}

// Unblocked right-looking Cholesky of the nb x nb diagonal block D (nb <= 32) by wavefront
// 0, lane a holding row a in registers and column j broadcast by v_readlane: L(j,j) =
// sqrt(d), rdiag[j] = 1 / L(j,j), L(a,j) = A(a,j) * rdiag[j], then A(a,b) = fma(-L(a,j),
// L(b,j), A(a,b)) for a >= b > j, in j order. On a non-positive pivot at column j, *bad
// becomes kb + j + 1 and the block is left as it stands. Every thread calls; ends with a
// barrier.
// Value of v in lane `lane` (a constant after unrolling) of the calling wavefront, broadcast
// through v_readlane (no LDS round trip, unlike a general shuffle). Every lane of the
// wavefront must be active.
__device__ __attribute__((always_inline)) inline float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __attribute__((always_inline)) inline double lane_value(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Always inlined: called out of line, its LDS accesses go through flat pointers.
template <typename T>
__device__ __attribute__((always_inline)) inline void chol_diag_wave(T (*D)[PB + 1], T* rdiag, int nb, int kb,
                                                                  int* bad) {
#pragma clang fp contract(off)
    if (threadIdx.x < 64) {
        const int a = threadIdx.x;
        const bool row = a < nb;
        T x[PB];
#pragma unroll
        for (int b = 0; b < PB; b++) x[b] = (row && b < nb) ? D[a][b] : T(0);
        int fail = 0;
#pragma unroll
        for (int j = 0; j < PB; j++) {
            if (j < nb) {  // uniform
                const T d = lane_value(x[j], j);
                const bool ok = d > T(0);
                if (!ok && !fail) fail = kb + j + 1;  // the block is not used after a failure
                const T ljj = ok ? T(sqrt(d)) : T(1);
                const T rd = T(1) / ljj;
                if (a == j) rdiag[j] = rd;
                x[j] = a == j ? ljj : (a > j ? x[j] * rd : x[j]);
                const T laj = x[j];
                // all broadcasts issue together (no branch per b); rows a >= b < nb update
#pragma unroll
                for (int b = j + 1; b < PB; b++) {
                    const T lbj = lane_value(laj, b);
                    const T u = fma(-laj, lbj, x[b]);
                    x[b] = (row && a >= b && b < nb) ? u : x[b];
                }
            }
        }
        if (row)
#pragma unroll
            for (int b = 0; b < PB; b++)
                if (b < nb) D[a][b] = x[b];
        if (a == 0 && fail) *bad = fail;
    }
    __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(NT) k_potrf(const PotrfTask* tasks, int* info) {
#pragma clang fp contract(off)
    static_assert(POTRF_SMALL_MAX <= 128 + PB, "panel staging holds n - 32 rows");
    __shared__ T P[128][PB + 1];  // rows below the diagonal block of the current panel
    __shared__ T D[2 * PB][PB + 1];  // the diagonal block, then panel_rows' zero pad
    __shared__ T rdiag[PB];
    __shared__ int bad;
    const PotrfTask t = tasks[blockIdx.x];
    const int n = t.n;
    const int tid = threadIdx.x;
    if (n == 0) return;
    const gptr<T> A = gp(reinterpret_cast<T*>(t.A));
    if (tid == 0) bad = 0;
    for (int idx = tid; idx < PB * (PB + 1); idx += NT) D[PB + idx / (PB + 1)][idx % (PB + 1)] = T(0);
    for (int kb = 0; kb < n; kb += PB) {
        const int nb = min(PB, n - kb);
        for (int idx = tid; idx < nb * nb; idx += NT) {
            int i = idx % nb, k = idx / nb;
            D[i][k] = A[(kb + i) + (size_t)(kb + k) * n];
        }
        __syncthreads();
        // diagonal block: unblocked right-looking
        chol_diag_wave<T>(D, rdiag, nb, kb, &bad);
        if (bad) break;
        for (int idx = tid; idx < nb * nb; idx += NT) {
            int i = idx % nb, k = idx / nb;
            if (i >= k) A[(kb + i) + (size_t)(kb + k) * n] = D[i][k];
        }
        // panel rows i >= kb + nb: scale / update across the block's columns in order
        panel_rows<T>(A, n, kb, nb, D, rdiag, kb + nb, n);
        __syncthreads();
        // trailing lower update by the block's columns, k in order, the panel staged in LDS
        const int mt = n - kb - nb;
        for (int idx = tid; idx < mt * nb; idx += NT) {
            const int i = idx % mt, k = idx / mt;
            P[i][k] = A[(kb + nb + i) + (size_t)(kb + k) * n];
        }
        __syncthreads();
        for (int idx = tid; idx < mt * mt; idx += NT) {
            const int i = idx % mt, j = idx / mt;
            if (i < j) continue;
            const auto a = A + (kb + nb + i) + (size_t)(kb + nb + j) * n;
            T acc = *a;
            if (nb == PB) {  // a full panel: unrolled, so the LDS reads run ahead of the FMAs
#pragma unroll
                for (int k = 0; k < PB; k++) acc = fma(-P[i][k], P[j][k], acc);
            } else {
                for (int k = 0; k < nb; k++) acc = fma(-P[i][k], P[j][k], acc);
            }
            *a = acc;
        }
        __syncthreads();
    }
    if (bad && tid == 0) atomicCAS(info, 0, t.cluster + 1);
}

// Pivots of at most 64 rows: one wavefront each, lane a owning row a (in registers). Same
// element recurrence as k_potrf (column j scaled by 1/L(j,j), then
// A(a,b) = fma(-L(a,j), L(b,j), A(a,b)) for a >= b > j, in j order).
template <typename T>
__global__ void __launch_bounds__(64) k_potrf64(const PotrfTask* tasks, int* info) {
#pragma clang fp contract(off)
    // r[b] = A[a][b]; step j is unrolled, so every register index is a constant. The
    // diagonal's lane takes the square root, lane a > j scales its L[a][j] and publishes it
    // in LDS, and every lane subtracts L[a][j] L[b][j] from its r[b], b > j (entries above a
    // lane's diagonal are updated too, never read, and not stored).
    __shared__ T col[64];
    __shared__ T rd;
    __shared__ int bad;
    const PotrfTask t = tasks[blockIdx.x];
    const int n = t.n, a = threadIdx.x;
    if (n == 0) return;
    const gptr<T> A = gp(reinterpret_cast<T*>(t.A));
    T r[64];
#pragma unroll
    for (int b = 0; b < 64; b++) r[b] = (a < n && b < n) ? A[a + (size_t)b * n] : T(0);
    if (a == 0) bad = 0;
    __syncthreads();
    bool stopped = false;  // uniform: a failed pivot or the last column
#pragma clang loop unroll(full)
    for (int j = 0; j < 64; j++) {
        if (stopped || j >= n) {
            stopped = true;
        } else {
            if (a == j) {
                const T d = r[j];
                if (!(d > T(0))) {
                    bad = j + 1;
                } else {
                    r[j] = sqrt(d);
                }
                rd = T(1) / r[j];
            }
            __syncthreads();
            if (bad) {
                stopped = true;
            } else {
                if (a > j) r[j] *= rd;
                col[a] = r[j];
                __syncthreads();
                const T laj = r[j];
#pragma unroll
                for (int b = j + 1; b < 64; b++)
                    if (b < n) r[b] = fma(-laj, col[b], r[b]);  // n is uniform: a scalar test
                __syncthreads();  // col is rewritten in the next step
            }
        }
    }
    if (bad) {
        if (a == 0) atomicCAS(info, 0, t.cluster + 1);
        return;
    }
    if (a < n) {
#pragma unroll
        for (int b = 0; b < 64; b++)
            if (b <= a && b < n) A[a + (size_t)b * n] = r[b];
    }
}

// Large pivots take the same blocked algorithm one panel per launch pair, so the work of
// every panel spreads over many workgroups: k_potrf_panel (grid: blocks of NT panel rows x
// pivots) factors the diagonal block kb in every workgroup and scales / updates its NT
// rows below it, then k_potrf_trail stores the factored diagonal block (from dbuf) and applies the
// panel to the trailing lower triangle, one 32 x 32 tile per workgroup. Every element sees
// the same operations in the same order as in k_potrf.
template <typename T>
__global__ void __launch_bounds__(NT) k_potrf_panel(const PotrfTask* tasks, int kb, int* info) {
#pragma clang fp contract(off)
    __shared__ T D[2 * PB][PB + 1];  // the diagonal block, then panel_rows' zero pad
    __shared__ T rdiag[PB];
    __shared__ int bad;
    const PotrfTask t = tasks[blockIdx.y];
    const int n = t.n;
    const int tid = threadIdx.x;
    if (kb >= n) return;
    const int nb = min(PB, n - kb);
    const int r0 = kb + nb + blockIdx.x * NT;  // this workgroup's panel rows [r0, r0 + NT)
    if (blockIdx.x > 0 && r0 >= n) return;
    const gptr<T> A = gp(reinterpret_cast<T*>(t.A));
    if (tid == 0) bad = 0;
    for (int idx = tid; idx < PB * (PB + 1); idx += NT) D[PB + idx / (PB + 1)][idx % (PB + 1)] = T(0);
    for (int idx = tid; idx < nb * nb; idx += NT) {
        int i = idx % nb, k = idx / nb;
        D[i][k] = A[(kb + i) + (size_t)(kb + k) * n];
    }
    __syncthreads();
    // every workgroup of the task factors the block itself; workgroup 0 leaves it in dbuf
    // for k_potrf_trail to store (a store into A here could reach another workgroup of this
    // launch before it reads the block)
    chol_diag_wave<T>(D, rdiag, nb, kb, &bad);
    if (bad) {
        if (blockIdx.x == 0 && tid == 0) atomicCAS(info, 0, t.cluster + 1);
        return;
    }
    if (blockIdx.x == 0) {
        const gptr<T> db = gp(reinterpret_cast<T*>(t.dbuf));
        for (int idx = tid; idx < nb * nb; idx += NT) db[idx] = D[idx % nb][idx / nb];
    }
    panel_rows<T>(A, n, kb, nb, D, rdiag, r0, min(n, r0 + NT));
}

// grid (tiles, tasks): tile index -> (bi, bj), bj <= bi, over the trailing lower triangle
template <typename T>
__global__ void __launch_bounds__(NT) k_potrf_trail(const PotrfTask* tasks, int kb) {
#pragma clang fp contract(off)
    __shared__ T As[PB][PB + 1];
    __shared__ T Bs[PB][PB + 1];
    const PotrfTask t = tasks[blockIdx.y];
    const int n = t.n, tid = threadIdx.x;
    if (kb >= n) return;
    if (blockIdx.x == 0) {  // the panel's factored diagonal block, left in dbuf by k_potrf_panel
        const int nb = min(PB, n - kb);
        const gptr<T> A = gp(reinterpret_cast<T*>(t.A));
        const gptr<const T> db = gp(reinterpret_cast<const T*>(t.dbuf));
        for (int idx = tid; idx < nb * nb; idx += NT) {
            int i = idx % nb, k = idx / nb;
            if (i >= k) A[(kb + i) + (size_t)(kb + k) * n] = db[idx];
        }
    }
    const int r0 = kb + PB;  // a full panel precedes any trailing rows
    if (r0 >= n) return;
    const int nt = (n - r0 + PB - 1) / PB;
    int p = blockIdx.x, bi = 0;
    while (p > bi) {
        p -= bi + 1;
        bi++;
    }
    if (bi >= nt) return;
    const int bj = p;
    const int a0 = r0 + bi * PB, b0 = r0 + bj * PB;
    const int na = min(PB, n - a0), nbb = min(PB, n - b0);
    const gptr<T> A = gp(reinterpret_cast<T*>(t.A));
    for (int idx = tid; idx < PB * PB; idx += NT) {
        const int i = idx % PB, k = idx / PB;
        As[k][i] = i < na ? A[(a0 + i) + (size_t)(kb + k) * n] : T(0);
        Bs[k][i] = i < nbb ? A[(b0 + i) + (size_t)(kb + k) * n] : T(0);
    }
    __syncthreads();
    const int ti = tid % PB, tj0 = (tid / PB) * 4;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int a = a0 + ti, b = b0 + tj0 + q;
        if (ti < na && tj0 + q < nbb && a >= b) {
            T acc = A[a + (size_t)b * n];
            for (int k = 0; k < PB; k++) acc = fma(-As[k][ti], Bs[k][tj0 + q], acc);
            A[a + (size_t)b * n] = acc;
        }
    }
}

// ---------------------------------------------------------------------------------------
// trsm on a chunk of <= 64 independent columns (left) or rows (right), in 32-row blocks of
// L: the diagonal triangle is solved by one thread per independent index, the rows below
// by the whole workgroup as an ordered-FMA update. Every element sees the reference BLAS
// operations in the same order (FP64 bit-identical to the oracle).
template <typename T>
__global__ void __launch_bounds__(NT) k_trsm(const TrsmTask* tasks) {
#pragma clang fp contract(off)
    __shared__ T Bk[PB][64 + 1];
    __shared__ T Lk[PB][PB + 1];
    __shared__ T Lr[PB][64 + 1];
    const TrsmTask t = tasks[blockIdx.x];
    const int n = t.n, m = t.m, ld = t.ldb;
    const int tid = threadIdx.x;
    if (n == 0 || m == 0) return;
    const gptr<const T> L = gp(reinterpret_cast<const T*>(t.L));
    const gptr<T> G = gp(reinterpret_cast<T*>(t.B));
    const bool left = t.side == 0;
    // element (triangular index i, independent index c)
    auto at = [&](int i, int c) -> decltype(G[0]) { return left ? G[i + (size_t)c * ld] : G[c + (size_t)i * ld]; };
    // a left solve from another block reads row perm[i] of it at the first touch of row i
    // (every row is first read in the kb = 0 pass), and its own block afterwards
    const gptr<const T> S = gp(reinterpret_cast<const T*>(t.src));
    const gptr<const int> P = gp(t.perm);
    auto first = [&](int i, int c) -> T {
        return S ? S[(P ? P[i] : i) + (size_t)c * ld] : at(i, c);
    };
    for (int kb = 0; kb < n; kb += PB) {
        const int nb = min(PB, n - kb);
        for (int idx = tid; idx < nb * m; idx += NT) {
            int k, c;
            if (left) { k = idx % nb; c = idx / nb; } else { c = idx % m; k = idx / m; }
            Bk[k][c] = kb == 0 ? first(kb + k, c) : at(kb + k, c);
        }
        for (int idx = tid; idx < nb * nb; idx += NT) {
            int i = idx % nb, k = idx / nb;
            Lk[i][k] = L[(kb + i) + (size_t)(kb + k) * n];
        }
        __syncthreads();
        if (tid < m) {  // column c of the diagonal block, held in registers
            const int c = tid;
            T x[PB];
#pragma unroll
            for (int k = 0; k < PB; k++) x[k] = k < nb ? Bk[k][c] : T(0);
#pragma unroll
            for (int k = 0; k < PB; k++) {
                if (k < nb) {
                    T b;
                    if (left) {
                        b = x[k] / Lk[k][k];
                    } else {
                        const T r = T(1) / Lk[k][k];
                        b = x[k] * r;
                    }
                    x[k] = b;
#pragma unroll
                    for (int i = k + 1; i < PB; i++)
                        if (i < nb) x[i] = fma(-b, Lk[i][k], x[i]);
                }
            }
#pragma unroll
            for (int k = 0; k < PB; k++)
                if (k < nb) Bk[k][c] = x[k];
        }
        __syncthreads();
        for (int idx = tid; idx < nb * m; idx += NT) {
            int k, c;
            if (left) { k = idx % nb; c = idx / nb; } else { c = idx % m; k = idx / m; }
            at(kb + k, c) = Bk[k][c];
        }
        // rows below the block, 64 at a time: L[i0:i0+64, kb:kb+nb] staged in LDS
        for (int i0 = kb + nb; i0 < n; i0 += 64) {
            const int ni = min(64, n - i0);
            for (int idx = tid; idx < ni * nb; idx += NT) {
                const int ii = idx % ni, k = idx / ni;
                Lr[k][ii] = L[(i0 + ii) + (size_t)(kb + k) * n];
            }
            __syncthreads();
            for (int idx = tid; idx < ni * m; idx += NT) {
                int ii, c;
                if (left) { ii = idx % ni; c = idx / ni; } else { c = idx % m; ii = idx / m; }
                T acc = kb == 0 ? first(i0 + ii, c) : at(i0 + ii, c);
                if (nb == PB) {  // a full block: unrolled, so the LDS reads run ahead of the FMAs
#pragma unroll
                    for (int k = 0; k < PB; k++) acc = fma(-Bk[k][c], Lr[k][ii], acc);
                } else {
                    for (int k = 0; k < nb; k++) acc = fma(-Bk[k][c], Lr[k][ii], acc);
                }
                at(i0 + ii, c) = acc;
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------------------
// FP32 trsm with the trailing updates on the matrix cores. Same blocking as k_trsm: the
// 32 x 32 diagonal block is solved by one thread per independent index (the recurrence of
// k_trsm), and each 64-row group below it is updated as C -= L[i0:i0+64, kb:kb+32] X by
// v_mfma_f32_32x32x2f32 with FP32 accumulation: four wavefronts, one 32 x 32 output tile
// each. The tile is laid out so that the lanes of a wavefront run along the contiguous
// index of the target block: left solves (C(i, c) at i + c ld) compute the tile transposed,
// D(c, i) = X^T L^T, right solves (C(c, i) at c + i ld) compute D(i, c) = L X. FP32 only:
// the FP64 path keeps the ordered-FMA recurrence of k_trsm.
typedef float mfma_f32x16 __attribute__((ext_vector_type(16)));
__global__ void __launch_bounds__(NT, 3) k_trsm_f32_mfma(const TrsmTask* tasks) {
    __shared__ float Bk[PB][64 + 1];
    __shared__ float Lk[PB][PB + 1];
    __shared__ float Lr[PB][64 + 1];
    const TrsmTask t = tasks[blockIdx.x];
    const int n = t.n, m = t.m, ld = t.ldb;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (n == 0 || m == 0) return;
    const float* L = reinterpret_cast<const float*>(t.L);
    float* G = reinterpret_cast<float*>(t.B);
    const bool left = t.side == 0;
    auto at = [&](int i, int c) -> float& { return left ? G[i + (size_t)c * ld] : G[c + (size_t)i * ld]; };
    const float* S = reinterpret_cast<const float*>(t.src);
    const int* P = t.perm;
    auto first = [&](int i, int c) -> float { return S ? S[(P ? P[i] : i) + (size_t)c * ld] : at(i, c); };
    // this wavefront's output tile: rows rh*32.. of the 64-row group, columns ch*32.. of the chunk
    const int rh = wv & 1, ch = wv >> 1;
    for (int kb = 0; kb < n; kb += PB) {
        const int nb = min(PB, n - kb);
        for (int idx = tid; idx < PB * 64; idx += NT) {  // zero padding: rows >= nb contribute nothing
            int k, c;
            if (left) { k = idx % PB; c = idx / PB; } else { c = idx % 64; k = idx / 64; }
            Bk[k][c] = (k < nb && c < m) ? (kb == 0 ? first(kb + k, c) : at(kb + k, c)) : 0.0f;
        }
        for (int idx = tid; idx < nb * nb; idx += NT) {
            int i = idx % nb, k = idx / nb;
            Lk[i][k] = L[(kb + i) + (size_t)(kb + k) * n];
        }
        __syncthreads();
        if (tid < m) {  // column c of the diagonal block, held in registers
            const int c = tid;
            float x[PB];
#pragma unroll
            for (int k = 0; k < PB; k++) x[k] = k < nb ? Bk[k][c] : 0.0f;
#pragma unroll
            for (int k = 0; k < PB; k++) {
                if (k < nb) {
                    float b;
                    if (left) {
                        b = x[k] / Lk[k][k];
                    } else {
                        const float r = 1.0f / Lk[k][k];
                        b = x[k] * r;
                    }
                    x[k] = b;
#pragma unroll
                    for (int i = k + 1; i < PB; i++)
                        if (i < nb) x[i] = fmaf(-b, Lk[i][k], x[i]);
                }
            }
#pragma unroll
            for (int k = 0; k < PB; k++)
                if (k < nb) Bk[k][c] = x[k];
        }
        __syncthreads();
        for (int idx = tid; idx < nb * m; idx += NT) {
            int k, c;
            if (left) { k = idx % nb; c = idx / nb; } else { c = idx % m; k = idx / m; }
            at(kb + k, c) = Bk[k][c];
        }
        for (int i0 = kb + nb; i0 < n; i0 += 64) {
            const int ni = min(64, n - i0);
            for (int idx = tid; idx < PB * 64; idx += NT) {
                const int ii = idx % 64, k = idx / 64;
                Lr[k][ii] = (ii < ni && k < nb) ? L[(i0 + ii) + (size_t)(kb + k) * n] : 0.0f;
            }
            __syncthreads();
            if (rh * 32 < ni && ch * 32 < m) {
                // element v of this lane: tile row (8 (v/4) + 4 (lane/32) + v%4), tile column lane%32
                const int tc = lane & 31, tr0 = 4 * (lane >> 5);
                mfma_f32x16 acc;
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int tr = 8 * (v >> 2) + tr0 + (v & 3);
                    // left: tile rows are chunk columns c, tile columns are rows i; right: the reverse
                    const int ii = rh * 32 + (left ? tc : tr), c = ch * 32 + (left ? tr : tc);
                    float val = 0.0f;
                    if (ii < ni && c < m) val = kb == 0 ? first(i0 + ii, c) : at(i0 + ii, c);
                    acc[v] = val;
                }
#pragma unroll
                for (int k = 0; k < PB; k += 2) {
                    const int kk = k + (lane >> 5);
                    const float x = -Bk[kk][ch * 32 + (lane & 31)];
                    const float l = Lr[kk][rh * 32 + (lane & 31)];
                    acc = left ? __builtin_amdgcn_mfma_f32_32x32x2f32(x, l, acc, 0, 0, 0)
                               : __builtin_amdgcn_mfma_f32_32x32x2f32(l, x, acc, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int tr = 8 * (v >> 2) + tr0 + (v & 3);
                    const int ii = rh * 32 + (left ? tc : tr), c = ch * 32 + (left ? tr : tc);
                    if (ii < ni && c < m) at(i0 + ii, c) = acc[v];
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------------------
// getrf + split_LU, one workgroup per pivot (getf_cluster, src/tree.cpp:614-635; getf,
// src/util.cpp:184-199; split_LU, src/util.cpp:213-227). The steps follow dgetf2: the pivot
// row is the first one of largest |a_ij| (idamax), rows j and jp are interchanged across
// all columns, the column is scaled by 1/a_jj (divided by a_jj when |a_jj| is below the
// smallest normal number), then a_ic = a_ic + a_ij * (-a_jc) for every trailing element with
// a_jc != 0 (dger), in j order -- the element recurrence of the oracle. The row
// permutation is accumulated as swap2perm does (src/util.cpp:76-88). Blocks of at most
// GETRF_LDS_BYTES are factored in LDS (ld n + 1), larger ones in place in global memory.
template <typename T, int NTG>
__global__ void __launch_bounds__(NTG) k_getrf(const GetrfTask* tasks, int in_lds, int* info) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char getrf_lds[];
    constexpr int NW = NTG / 64;
    __shared__ T red_v[NW];
    __shared__ int red_i[NW];
    __shared__ int s_jp, s_nz, bad;
    const GetrfTask t = tasks[blockIdx.x];
    const int n = t.n;
    const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
    if (n == 0) return;
    const gptr<T> G = gp(reinterpret_cast<T*>(t.A));
    const gptr<T> Ut = gp(reinterpret_cast<T*>(t.Ut));
    const gptr<int> perm = gp(t.perm);
    T* A = in_lds ? reinterpret_cast<T*>(getrf_lds) : (T*)G;
    const int ld = in_lds ? n + 1 : n;
    if (in_lds)
        for (int c = wv; c < n; c += NW)
            for (int i = lane; i < n; i += 64) A[i + (size_t)c * ld] = G[i + (size_t)c * n];
    for (int i = tid; i < n; i += NTG) perm[i] = i;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < n; j++) {
        // idamax over rows j..n-1 of column j: first index of the largest magnitude
        T bv = T(-1);
        int bi = n;
        for (int i = j + tid; i < n; i += NTG) {
            const T v = fabs(A[i + (size_t)j * ld]);
            if (v > bv) { bv = v; bi = i; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const T ov = __shfl_xor(bv, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if (lane == 0) { red_v[wv] = bv; red_i[wv] = bi; }
        __syncthreads();
        if (tid == 0) {
            T v = red_v[0];
            int jp = red_i[0];
            for (int w = 1; w < NW; w++)
                if (red_v[w] > v || (red_v[w] == v && red_i[w] < jp)) { v = red_v[w]; jp = red_i[w]; }
            // idamax returns the first index when a_jj is NaN (every comparison fails), and
            // an all-NaN column has no candidate: keep row j in both cases, so that jp never
            // leaves [j, n)
            const T ajj = A[j + (size_t)j * ld];
            if (jp < j || jp >= n || ajj != ajj) jp = j;
            s_jp = jp;
            s_nz = A[jp + (size_t)j * ld] != T(0);
            if (!s_nz && !bad) bad = j + 1;
            const int tmp = perm[j];
            perm[j] = perm[jp];
            perm[jp] = tmp;
        }
        __syncthreads();
        const int jp = s_jp;
        if (s_nz) {
            if (jp != j)
                for (int c = tid; c < n; c += NTG) {
                    const T a = A[j + (size_t)c * ld];
                    A[j + (size_t)c * ld] = A[jp + (size_t)c * ld];
                    A[jp + (size_t)c * ld] = a;
                }
            __syncthreads();
            const T piv = A[j + (size_t)j * ld];
            if (fabs(piv) >= std::numeric_limits<T>::min()) {
                const T r = T(1) / piv;
                for (int i = j + 1 + tid; i < n; i += NTG) A[i + (size_t)j * ld] *= r;
            } else {
                for (int i = j + 1 + tid; i < n; i += NTG) A[i + (size_t)j * ld] /= piv;
            }
        }
        __syncthreads();
        // dger: trailing columns by wavefront, rows by lane
        for (int c = j + 1 + wv; c < n; c += NW) {
            const T y = A[j + (size_t)c * ld];
            if (y != T(0)) {
                const T mt = -y;
                for (int i = j + 1 + lane; i < n; i += 64)
                    A[i + (size_t)c * ld] = A[i + (size_t)c * ld] + A[i + (size_t)j * ld] * mt;
            }
        }
        __syncthreads();
    }
    // split_LU: U(i,c) = (sign(d_i) sqrt|d_i|) * ((1 / d_i) * a_ic) above the diagonal,
    // sign(d_i) sqrt|d_i| on it; stored transposed (Ut column i = row i of U), zero above
    for (int i = wv; i < n; i += NW) {
        const T d = A[i + (size_t)i * ld];
        const T f = T((d > T(0)) - (d < T(0))) * sqrt(fabs(d));
        const T inv = T(1) / d;
        for (int c = lane; c < n; c += 64)
            Ut[c + (size_t)i * n] = c < i ? T(0) : (c == i ? f * T(1) : f * (inv * A[i + (size_t)c * ld]));
    }
    __syncthreads();
    // L(r,j) = a_rj * sqrt|d_j| below the diagonal, sqrt|d_j| on it, zero above
    for (int j = wv; j < n; j += NW) {
        const T sj = sqrt(fabs(A[j + (size_t)j * ld]));
        for (int r = lane; r < n; r += 64)
            if (r != j) G[r + (size_t)j * n] = r > j ? A[r + (size_t)j * ld] * sj : T(0);
    }
    __syncthreads();
    for (int j = tid; j < n; j += NTG) G[j + (size_t)j * n] = sqrt(fabs(A[j + (size_t)j * ld])) * T(1);
    if (bad && tid == 0) atomicCAS(info, 0, t.cluster + 1);
}

// ---------------------------------------------------------------------------------------
// Schur complement: one workgroup per 32x32 tile of a target block; the target's
// contributions are applied in recorded order (deterministic, no atomics), and every
// element accumulates C = fma(-a_l, b_l, C) in k order, the element recurrence of the
// reference dgemm/dsyrk loops. Each contribution's k-loop is staged through LDS in
// chunks of 32.
constexpr int ST = 32;   // tile edge
constexpr int SK = 32;   // k chunk
template <typename T>
__global__ void __launch_bounds__(NT) k_schur(const SchurTile* tiles, const SchurContrib* contribs) {
#pragma clang fp contract(off)
    __shared__ T As[SK][ST + 1];
    __shared__ T Bs[SK][ST + 1];
    const SchurTile t = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    // thread -> 4 outputs of the tile: row tid % 32, columns (tid / 32) * 4 + 0..3
    const int ti = tid % ST;
    const int tj0 = (tid / ST) * 4;
    const gptr<T> C = gp(reinterpret_cast<T*>(t.C));
    T acc_c[4];
    for (int q = 0; q < 4; q++) {
        int i = ti, j = tj0 + q;
        acc_c[q] = (i < t.tm && j < t.tn) ? C[(t.r0 + i) + (size_t)(t.c0 + j) * t.ldc] : T(0);
    }
    for (int kk = t.k0; kk < t.k1; kk++) {
        const SchurContrib c = contribs[kk];
        const gptr<const T> A = gp(reinterpret_cast<const T*>(c.A));
        const gptr<const T> B = gp(reinterpret_cast<const T*>(c.B));
        for (int l0 = 0; l0 < c.k; l0 += SK) {
            const int kc = min(SK, c.k - l0);
            // load opA(r0+i, l0+l) into As[l][i] and opB(l0+l, c0+j) into Bs[l][j]
            for (int idx = tid; idx < SK * ST; idx += NT) {
                int i = idx % ST, l = idx / ST;
                T va = T(0), vb = T(0);
                if (l < kc) {
                    int gi = t.r0 + i, gl = l0 + l, gj = t.c0 + i;
                    if (i < t.tm) va = c.ta ? A[gl + (size_t)gi * c.lda] : A[gi + (size_t)gl * c.lda];
                    if (i < t.tn) vb = c.tb ? B[gj + (size_t)gl * c.ldb] : B[gl + (size_t)gj * c.ldb];
                }
                As[l][i] = va;
                Bs[l][i] = vb;
            }
            __syncthreads();
            if (kc == SK) {  // a full chunk: unrolled, so the LDS reads run ahead of the FMAs
#pragma unroll
                for (int l = 0; l < SK; l++) {
                    const T a = -As[l][ti];
#pragma unroll
                    for (int q = 0; q < 4; q++) acc_c[q] = fma(a, Bs[l][tj0 + q], acc_c[q]);
                }
            } else {
                for (int l = 0; l < kc; l++) {
                    const T a = -As[l][ti];
#pragma unroll
                    for (int q = 0; q < 4; q++) acc_c[q] = fma(a, Bs[l][tj0 + q], acc_c[q]);
                }
            }
            __syncthreads();
        }
    }
    for (int q = 0; q < 4; q++) {
        int i = ti, j = tj0 + q;
        if (i < t.tm && j < t.tn) {
            int gi = t.r0 + i, gj = t.c0 + j;
            if (!t.lower || gi >= gj) C[gi + (size_t)gj * t.ldc] = acc_c[q];
        }
    }
}

// FP32 Schur complement on the matrix cores: the same 32 x 32 target tiles and staged k
// chunks as k_schur; the four wavefronts each accumulate every fourth k-pair of the tile's
// contributions with v_mfma_f32_32x32x2f32 (FP32 accumulation, the tile computed
// transposed so that a wavefront's lanes run down the target's columns), and the four
// partial sums are added in LDS and subtracted from the target once at the end.
__global__ void __launch_bounds__(NT) k_schur_f32_mfma(const SchurTile* tiles, const SchurContrib* contribs) {
    __shared__ float As[SK][ST + 1];
    __shared__ float Bs[SK][ST + 1];
    __shared__ float Ps[NT / 64][ST][ST + 1];
    const SchurTile t = tiles[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const gptr<float> C = gp(reinterpret_cast<float*>(t.C));
    mfma_f32x16 acc = {};
    for (int kk = t.k0; kk < t.k1; kk++) {
        const SchurContrib c = contribs[kk];
        const gptr<const float> A = gp(reinterpret_cast<const float*>(c.A));
        const gptr<const float> B = gp(reinterpret_cast<const float*>(c.B));
        for (int l0 = 0; l0 < c.k; l0 += SK) {
            const int kc = min(SK, c.k - l0);
            for (int idx = tid; idx < SK * ST; idx += NT) {
                int i = idx % ST, l = idx / ST;
                float va = 0.0f, vb = 0.0f;
                if (l < kc) {
                    int gi = t.r0 + i, gl = l0 + l, gj = t.c0 + i;
                    if (i < t.tm) va = c.ta ? A[gl + (size_t)gi * c.lda] : A[gi + (size_t)gl * c.lda];
                    if (i < t.tn) vb = c.tb ? B[gj + (size_t)gl * c.ldb] : B[gl + (size_t)gj * c.ldb];
                }
                As[l][i] = va;
                Bs[l][i] = vb;
            }
            __syncthreads();
            // D(j, i) += sum_l opB(l, j) opA(i, l): the A operand is opB^T, the B operand opA^T
#pragma unroll
            for (int q = 0; q < SK / 8; q++) {
                const int kp = 2 * (wv + 4 * q) + (lane >> 5);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Bs[kp][lane & 31], As[kp][lane & 31], acc, 0, 0, 0);
            }
            __syncthreads();
        }
    }
    // element v of this lane: D row j = 8 (v/4) + 4 (lane/32) + v%4, column i = lane%32
#pragma unroll
    for (int v = 0; v < 16; v++) Ps[wv][8 * (v >> 2) + 4 * (lane >> 5) + (v & 3)][lane & 31] = acc[v];
    __syncthreads();
    for (int idx = tid; idx < ST * ST; idx += NT) {
        const int i = idx % ST, j = idx / ST;
        if (i < t.tm && j < t.tn) {
            const int gi = t.r0 + i, gj = t.c0 + j;
            if (!t.lower || gi >= gj) {
                const float sum = (Ps[0][j][i] + Ps[1][j][i]) + (Ps[2][j][i] + Ps[3][j][i]);
                const auto cp = C + gi + (size_t)gj * t.ldc;
                *cp = *cp - sum;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// RRQR: gather Asn, dgeqp3 (dlaqp2 step order: first-max pivot, LAWN-176 norm downdate with
// recomputation below sqrt(eps/2)), choose_rank, Q/tau store and scatter of R[:r,:] P^T
// into the edge blocks. Column norms, the downdate and the rank test run in FP64 for both
// element types; scalar steps follow the reference LAPACK formulas one-for-one.
//
// Every sum (column norms, the dlarfg norm, w = v^T a_j) is taken in one fixed order: term
// t goes to chain t % 16, each chain accumulates in increasing t with fused multiply-adds
// from 0, and the 16 chains are combined by a halving tree (+8, +4, +2, +1). The CPU oracle
// uses the same order, so FP64 runs agree bit for bit; sixteen wavefronts (one per chain)
// compute each sum in parallel.
//
// The factorization stops at the rank: once |R_ii| / |R_00| < tol (choose_rank,
// src/util.cpp:434-452) nothing of step i or later is kept, so those steps are skipped.
constexpr int RRQR_MAXSEG = 1024;
constexpr int RRQR_LDS_COLS = 1024;
constexpr int NCH = 16;

__device__ inline double lapy2_d(double x, double y) {
    double xa = fabs(x), ya = fabs(y);
    double w = fmax(xa, ya), z = fmin(xa, ya);
    if (z == 0.0) return w;
    double q = z / w;
    return w * sqrt(1.0 + q * q);
}

// halving tree over the 16 chain values (c is overwritten)
template <typename A>
__device__ inline A tree16(A* c) {
#pragma clang fp contract(off)
#pragma unroll
    for (int w = 0; w < 8; w++) c[w] = c[w] + c[w + 8];
#pragma unroll
    for (int w = 0; w < 4; w++) c[w] = c[w] + c[w + 4];
#pragma unroll
    for (int w = 0; w < 2; w++) c[w] = c[w] + c[w + 2];
    return c[0] + c[1];
}

// sum over t < n of g(t)^2 in the chain order, by one thread
template <typename G>
__device__ inline double sumsq16_serial(int n, G g) {
#pragma clang fp contract(off)
    double c[NCH];
#pragma unroll
    for (int w = 0; w < NCH; w++) c[w] = 0.0;
    int t0 = 0;
    for (; t0 + NCH <= n; t0 += NCH) {
#pragma unroll
        for (int w = 0; w < NCH; w++) {
            const double v = g(t0 + w);
            c[w] = fma(v, v, c[w]);
        }
    }
#pragma unroll
    for (int w = 0; w < NCH; w++)
        if (t0 + w < n) {
            const double v = g(t0 + w);
            c[w] = fma(v, v, c[w]);
        }
    return tree16(c);
}

// sum over t < n of a(t) * b(t) in the chain order, in T, by one thread
template <typename T, typename GA, typename GB>
__device__ inline T dot16_serial(int n, GA a, GB b) {
#pragma clang fp contract(off)
    T c[NCH];
#pragma unroll
    for (int w = 0; w < NCH; w++) c[w] = T(0);
    int t0 = 0;
    for (; t0 + NCH <= n; t0 += NCH) {
#pragma unroll
        for (int w = 0; w < NCH; w++) c[w] = fma(a(t0 + w), b(t0 + w), c[w]);
    }
#pragma unroll
    for (int w = 0; w < NCH; w++)
        if (t0 + w < n) c[w] = fma(a(t0 + w), b(t0 + w), c[w]);
    return tree16(c);
}

// Same sum of squares by threads 0..15 of the block (one chain each); every thread calls,
// the result is valid in thread 0. s_ch: 16 doubles of LDS.
template <typename G>
__device__ inline double sumsq16_block(int n, G g, double* s_ch) {
#pragma clang fp contract(off)
    if (threadIdx.x < NCH) {
        double c = 0.0;
        int t = threadIdx.x;
        for (; t + 3 * NCH < n; t += 4 * NCH) {  // four loads in flight, the sum in order
            const double v0 = g(t), v1 = g(t + NCH), v2 = g(t + 2 * NCH), v3 = g(t + 3 * NCH);
            c = fma(v0, v0, c);
            c = fma(v1, v1, c);
            c = fma(v2, v2, c);
            c = fma(v3, v3, c);
        }
        for (; t < n; t += NCH) {
            const double v = g(t);
            c = fma(v, v, c);
        }
        s_ch[threadIdx.x] = c;
    }
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        double c[NCH];
#pragma unroll
        for (int w = 0; w < NCH; w++) c[w] = s_ch[w];
        r = tree16(c);
    }
    return r;
}

// dlarfg scalars of the column segment [alpha; x(0..n1)): tau, beta, the final scale and
// the number of safe-minimum rescalings (reference dlarfg). Every thread calls; sh[0..3] =
// tau, beta, scal, knt on return (after a barrier).
template <typename T>
__device__ void larfg_block(double alpha, const T* x, int n1, double* sh, double* s_ch) {
#pragma clang fp contract(off)
    const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
    const double rs = 1.0 / safmin;
    const double xs0 = sumsq16_block(n1, [&](int t) { return (double)x[t]; }, s_ch);
    if (threadIdx.x == 0) {
        double tau = 0.0, scal = 1.0, beta = alpha;
        int knt = 0;
        if (n1 >= 1) {
            const double xnorm = sqrt(xs0);
            if (xnorm != 0.0) {
                beta = -copysign(lapy2_d(alpha, xnorm), alpha);
                if (fabs(beta) < safmin) {
                    do {
                        knt++;
                        beta *= rs;
                        alpha *= rs;
                    } while (fabs(beta) < safmin && knt < 20);
                } else {
                    tau = (beta - alpha) / beta;
                    scal = 1.0 / (alpha - beta);
                }
            }
        }
        sh[0] = tau;
        sh[1] = beta;
        sh[2] = scal;
        sh[3] = (double)knt;
        s_ch[NCH] = alpha;  // rescaled alpha, for the rare path below
    }
    __syncthreads();
    const int knt = (int)sh[3];
    if (knt > 0) {  // |beta| below the safe minimum: norm of the rescaled vector (reference dlarfg)
        const double xs = sumsq16_block(
            n1,
            [&](int t) {
                T xv = x[t];
                for (int k = 0; k < knt; k++) xv = (T)((double)xv * rs);
                return (double)xv;
            },
            s_ch);
        if (threadIdx.x == 0) {
            const double alpha_s = s_ch[NCH];
            double beta = -copysign(lapy2_d(alpha_s, sqrt(xs)), alpha_s);
            sh[0] = (beta - alpha_s) / beta;
            sh[2] = 1.0 / (alpha_s - beta);
            for (int k = 0; k < knt; k++) beta *= safmin;
            sh[1] = beta;
        }
        __syncthreads();
    }
}

// halving tree over 16 chain values in LDS (stride between chains)
__device__ inline double tree16_lds(const double* c, int stride) {
#pragma clang fp contract(off)
    const double x0 = ((c[0] + c[8 * stride]) + (c[4 * stride] + c[12 * stride])) +
                      ((c[2 * stride] + c[10 * stride]) + (c[6 * stride] + c[14 * stride]));
    const double x1 = ((c[1 * stride] + c[9 * stride]) + (c[5 * stride] + c[13 * stride])) +
                      ((c[3 * stride] + c[11 * stride]) + (c[7 * stride] + c[15 * stride]));
    return x0 + x1;
}
__device__ inline float tree16_lds_f(const double* c, int stride) {
#pragma clang fp contract(off)
    auto v = [&](int w) { return (float)c[w * stride]; };
    const float x0 = ((v(0) + v(8)) + (v(4) + v(12))) + ((v(2) + v(10)) + (v(6) + v(14)));
    const float x1 = ((v(1) + v(9)) + (v(5) + v(13))) + ((v(3) + v(11)) + (v(7) + v(15)));
    return x0 + x1;
}
template <typename T> __device__ inline T tree16_lds_t(const double* c, int stride);
template <> __device__ inline double tree16_lds_t<double>(const double* c, int stride) { return tree16_lds(c, stride); }
template <> __device__ inline float tree16_lds_t<float>(const double* c, int stride) { return tree16_lds_f(c, stride); }

template <typename T>
__device__ inline T larfg_scale(T xv, int knt, double scal) {
    const double rs = 1.0 / (DBL_MIN / (DBL_EPSILON * 0.5));
    for (int k = 0; k < knt; k++) xv = (T)((double)xv * rs);
    return xv * (T)scal;
}

// One combining step of wave_best: the pair of the lane that the DPP control CTRL names
// (lanes of rows outside ROWMASK combine with themselves).
template <int CTRL, int ROWMASK>
__device__ inline void best_step(double& v, int& p) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xF, false);
    const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xF, false);
    const int p2 = __builtin_amdgcn_update_dpp(p, p, CTRL, ROWMASK, 0xF, false);
    const double v2 = __hiloint2double(hi2, lo2);
    if (v2 > v || (v2 == v && p2 < p)) {
        v = v2;
        p = p2;
    }
}
// (value desc, position asc) maximum over the 64 lanes of a wavefront (all active); every
// lane gets it. The order is total, so the reduction tree does not change the result: DPP
// moves within quads (xor 1, xor 2), half-rows and rows (mirrors), then row 0 -> 1, 2 -> 3
// (row_bcast15) and row 1 -> 2, 3 (row_bcast31) leave the maximum in lane 63.
__device__ inline void wave_best(double& v, int& p) {
    best_step<0xB1, 0xF>(v, p);   // quad_perm [1, 0, 3, 2]
    best_step<0x4E, 0xF>(v, p);   // quad_perm [2, 3, 0, 1]
    best_step<0x141, 0xF>(v, p);  // row_half_mirror
    best_step<0x140, 0xF>(v, p);  // row_mirror
    best_step<0x142, 0xA>(v, p);  // row_bcast15
    best_step<0x143, 0xC>(v, p);  // row_bcast31
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    v = __hiloint2double(hi, lo);
    p = __builtin_amdgcn_readlane(p, 63);
}

// choose_rank (src/util.cpp:434-452) decided during the factorization: -1 while undecided.
__device__ inline int rank_fixed(double tol, int mn) {
    if (tol == 0.0) return mn;
    if (tol >= 1.0) return 0;
    if (mn <= 1) return mn;
    return -1;
}

template <typename T>
__global__ void __launch_bounds__(NT) k_rrqr(const RrqrTask* tasks, const RrqrSeg* segs, int* dsize,
                                             double tol, int lds_elems) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_col0[RRQR_MAXSEG + 1];
    __shared__ double r_val[NT / 64];
    __shared__ int r_idx[NT / 64];
    __shared__ double s_vn1[RRQR_LDS_COLS], s_vn2[RRQR_LDS_COLS];
    __shared__ int s_jp[RRQR_LDS_COLS];
    __shared__ double sh[4], s_ch[NCH + 1];

    const RrqrTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    const int rows = dsize[t.cluster];
    const RrqrSeg* sg = segs + t.seg0;
    if (tid == 0) {
        int c = 0;
        for (int k = 0; k < t.nseg; k++) {
            s_col0[k] = c;
            c += dsize[sg[k].nbr];
        }
        s_col0[t.nseg] = c;
    }
    __syncthreads();
    const int cols = s_col0[t.nseg];
    if (rows == 0) return;
    const int mn = min(rows, cols);
    // column stride: odd in LDS, so that the threads of a wavefront (one column each) fall
    // in different banks; rows in the global scratch
    const bool in_lds = (long long)(rows | 1) * cols <= lds_elems;
    const int ld = in_lds ? (rows | 1) : rows;
    T* A = in_lds ? reinterpret_cast<T*>(smem_raw) : reinterpret_cast<T*>(t.scratch);
    double* vn1 = cols <= RRQR_LDS_COLS ? s_vn1 : t.norms;
    double* vn2 = cols <= RRQR_LDS_COLS ? s_vn2 : t.norms + t.max_cols;
    int* jp = cols <= RRQR_LDS_COLS ? s_jp : t.jpvt;

    auto seg_of = [&](int j) {
        int lo = 0, hi = t.nseg - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (s_col0[mid] <= j) lo = mid; else hi = mid - 1;
        }
        return lo;
    };
    // gather
    for (long long idx = tid; idx < (long long)rows * cols; idx += NT) {
        int i = (int)(idx % rows), j = (int)(idx / rows);
        int k = seg_of(j);
        int jj = j - s_col0[k], w = s_col0[k + 1] - s_col0[k];
        const T* E = reinterpret_cast<const T*>(sg[k].E);
        A[i + (size_t)j * ld] = sg[k].out ? E[jj + (size_t)i * w] : E[i + (size_t)jj * rows];
    }
    __syncthreads();
    for (int j = tid; j < cols; j += NT) {
        const T* a = A + (size_t)j * ld;
        jp[j] = j;
        vn1[j] = sqrt(sumsq16_serial(rows, [&](int r) { return (double)a[r]; }));
        vn2[j] = vn1[j];
    }
    __syncthreads();
    const double tol3z = sqrt(DBL_EPSILON * 0.5);
    const int fixed = rank_fixed(tol, mn);
    const int nsteps = fixed == 0 ? 0 : mn;
    int rank = fixed >= 0 ? fixed : mn;
    for (int i = 0; i < nsteps; i++) {
        // pivot = first index of the max of vn1[i:cols]
        double bv = -1.0;
        int bi = cols;
        for (int j = i + tid; j < cols; j += NT) {
            double v = fabs(vn1[j]);
            if (v > bv) { bv = v; bi = j; }
        }
        wave_best(bv, bi);
        if (tid % 64 == 0) {
            r_val[tid / 64] = bv;
            r_idx[tid / 64] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < NT / 64; w++)
                if (r_val[w] > r_val[0] || (r_val[w] == r_val[0] && r_idx[w] < r_idx[0])) {
                    r_val[0] = r_val[w];
                    r_idx[0] = r_idx[w];
                }
        }
        __syncthreads();
        // no candidate (every remaining norm NaN, after a failed pivot upstream): keep column i
        // in place, so that no index leaves [i, cols)
        const int pvt = (r_idx[0] >= i && r_idx[0] < cols) ? r_idx[0] : i;
        if (pvt != i) {
            for (int r = tid; r < rows; r += NT) {
                T tmp = A[r + (size_t)pvt * ld];
                A[r + (size_t)pvt * ld] = A[r + (size_t)i * ld];
                A[r + (size_t)i * ld] = tmp;
            }
            if (tid == 0) {
                int tj = jp[pvt]; jp[pvt] = jp[i]; jp[i] = tj;
                vn1[pvt] = vn1[i];
                vn2[pvt] = vn2[i];
            }
        }
        __syncthreads();
        // dlarfg on A(i:rows, i)
        T* ci = A + (size_t)i * ld;
        larfg_block<T>((double)ci[i], ci + i + 1, rows - i - 1, sh, s_ch);
        const double taud = sh[0];
        if (taud != 0.0) {
            const int knt = (int)sh[3];
            const double scal = sh[2];
            for (int r = i + 1 + tid; r < rows; r += NT) ci[r] = larfg_scale<T>(ci[r], knt, scal);
        }
        __syncthreads();
        if (tid == 0 && taud != 0.0) ci[i] = (T)sh[1];
        __syncthreads();
        // rank reached: |R_ii| / |R_00| < tol at i >= 1
        if (fixed < 0 && i >= 1 && !(fabs((double)ci[i]) / fabs((double)A[0]) >= tol)) {
            rank = i;
            break;
        }
        const T tau = (T)taud;
        // apply H(i) to A(i:rows, i+1:cols) (dlarf: w = v^T a_j, skip w = 0 as dger does)
        // and downdate the partial norms
        for (int j = i + 1 + tid; j < cols; j += NT) {
            T* cj = A + (size_t)j * ld;
            if (tau != T(0)) {
                const T w = dot16_serial<T>(
                    rows - i, [&](int q) { return cj[i + q]; }, [&](int q) { return q == 0 ? T(1) : ci[i + q]; });
                if (w != T(0)) {
                    const T tm = -tau * w;
                    cj[i] = cj[i] + tm;
                    for (int r = i + 1; r < rows; r++) cj[r] = fma(ci[r], tm, cj[r]);
                }
            }
            if (vn1[j] != 0.0) {
                double q = fabs((double)cj[i]) / vn1[j];
                double temp = 1.0 - q * q;
                temp = fmax(temp, 0.0);
                double rr = vn1[j] / vn2[j];
                double temp2 = temp * rr * rr;
                if (temp2 <= tol3z) {
                    if (i < rows - 1) {
                        vn1[j] = sqrt(sumsq16_serial(rows - i - 1, [&](int q) { return (double)cj[i + 1 + q]; }));
                        vn2[j] = vn1[j];
                    } else {
                        vn1[j] = 0.0;
                        vn2[j] = 0.0;
                    }
                } else {
                    vn1[j] = vn1[j] * sqrt(temp);
                }
            }
        }
        if (tid == 0) gp(reinterpret_cast<T*>(t.tau))[i] = tau;
        __syncthreads();
    }
    if (rank >= rows) return;
    // Householder vectors v = A[:, :rank] (rows x rank, ld rows)
    const gptr<T> Q = gp(reinterpret_cast<T*>(t.Q));
    for (long long idx = tid; idx < (long long)rows * rank; idx += NT)
        Q[idx] = A[(idx % rows) + (size_t)(idx / rows) * ld];
    // AsnP[:, jpvt[j]] = triu(R)[:rank, j], scattered into the shrunk edge blocks
    for (int j = tid; j < cols; j += NT) {
        const int p = jp[j];
        if (p < 0 || p >= cols) continue;  // jp is a permutation of [0, cols)
        int k = seg_of(p);
        int jj = p - s_col0[k], w = s_col0[k + 1] - s_col0[k];
        T* E = reinterpret_cast<T*>(sg[k].E);
        const T* rj = A + (size_t)j * ld;
        if (sg[k].out) {
            for (int r = 0; r < rank; r++) E[jj + (size_t)r * w] = (r <= j) ? rj[r] : T(0);
        } else {
            for (int r = 0; r < rank; r++) E[r + (size_t)jj * rank] = (r <= j) ? rj[r] : T(0);
        }
    }
    __syncthreads();
    if (tid == 0) dsize[t.cluster] = rank;
}

// ---------------------------------------------------------------------------------------
// Register-resident RRQR of a mid-size cluster: one workgroup per cluster, thread t holds
// columns t + NTR k (k < K) of Asn, rows 0 .. RR - 1 in registers, for the whole
// factorization. Columns never move: each keeps its LAPACK position (column swaps and the
// first-maximum tie-break of dgeqp3 are reproduced on positions). A pivot step:
//   1. every free column offers its partial norm; the maximum (first position on ties)
//      is found by wavefront shuffles and one pass over the per-wavefront winners;
//   2. the thread holding the pivot column computes dlarfg (reference dlarfg, as in
//      k_rrqr), stores the reflector in LDS and the scalars, and keeps beta / v in its
//      registers (they are Q's column and R's diagonal);
//   3. every free column applies the reflector (w = v^T a, a += -tau w v) and downdates
//      its norm (LAPACK dlaqp2 rule, with the same recomputation test).
// Two barriers per step and no global memory traffic between the gather and the final
// writes of Q, tau, R and the rank. Every sum follows the 16-chain order of k_rrqr (chain
// of row r = (r - start) mod 16, fma in increasing r, halving tree): rows are indexed
// statically, so the chains are accumulated by r mod 16 and rotated by the start row.

// segment of Asn column j: the last k with col0[k] <= j
__device__ inline int seg_find(const int* col0, int nseg, int j) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (col0[mid] <= j) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// CPW consecutive values of T from LDS aligned to their size (8 or 16-byte reads)
template <typename T, int CPW>
__device__ inline void lds_row_group(const T* p, T (&out)[CPW]) {
    static_assert(CPW * sizeof(T) % 8 == 0, "whole 8-byte reads");
    __builtin_memcpy(out, __builtin_assume_aligned(p, CPW * sizeof(T) >= 16 ? 16 : 8), CPW * sizeof(T));
}

template <typename T, int CPW>
__device__ inline void lds_store_group(T* p, const T (&in)[CPW]) {
    static_assert(CPW * sizeof(T) % 8 == 0, "whole 8-byte writes");
    __builtin_memcpy(__builtin_assume_aligned(p, CPW * sizeof(T) >= 16 ? 16 : 8), in, CPW * sizeof(T));
}

// (value desc, position asc) maximum over the 64 lanes of a wavefront, carrying a tag; the
// DPP steps of wave_best (positions of real candidates are distinct, so the tag of the
// maximum does not depend on the reduction tree)
template <int CTRL, int ROWMASK>
__device__ inline void best_step_tag(double& v, int& p, int& tag) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xF, false);
    const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xF, false);
    const int p2 = __builtin_amdgcn_update_dpp(p, p, CTRL, ROWMASK, 0xF, false);
    const int t2 = __builtin_amdgcn_update_dpp(tag, tag, CTRL, ROWMASK, 0xF, false);
    const double v2 = __hiloint2double(hi2, lo2);
    if (v2 > v || (v2 == v && p2 < p)) {
        v = v2;
        p = p2;
        tag = t2;
    }
}
__device__ inline void wave_best_tag(double& v, int& p, int& tag) {
    best_step_tag<0xB1, 0xF>(v, p, tag);   // quad_perm [1, 0, 3, 2]
    best_step_tag<0x4E, 0xF>(v, p, tag);   // quad_perm [2, 3, 0, 1]
    best_step_tag<0x141, 0xF>(v, p, tag);  // row_half_mirror
    best_step_tag<0x140, 0xF>(v, p, tag);  // row_mirror
    best_step_tag<0x142, 0xA>(v, p, tag);  // row_bcast15
    best_step_tag<0x143, 0xC>(v, p, tag);  // row_bcast31
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    v = __hiloint2double(hi, lo);
    p = __builtin_amdgcn_readlane(p, 63);
    tag = __builtin_amdgcn_readlane(tag, 63);
}

// tree16 of the chains rotated by S (chain k = e[(k + S) mod 16]), without moving registers
template <int S, typename A>
__device__ inline A tree16_rot_s(const A* e) {
#pragma clang fp contract(off)
    auto c = [&](int k) { return e[(k + S) & (NCH - 1)]; };
    const A x0 = ((c(0) + c(8)) + (c(4) + c(12))) + ((c(2) + c(10)) + (c(6) + c(14)));
    const A x1 = ((c(1) + c(9)) + (c(5) + c(13))) + ((c(3) + c(11)) + (c(7) + c(15)));
    return x0 + x1;
}
template <typename A>
__device__ inline A tree16_rot(const A* e, int s) {  // s uniform
    switch (s & (NCH - 1)) {
        case 0: return tree16_rot_s<0>(e);
        case 1: return tree16_rot_s<1>(e);
        case 2: return tree16_rot_s<2>(e);
        case 3: return tree16_rot_s<3>(e);
        case 4: return tree16_rot_s<4>(e);
        case 5: return tree16_rot_s<5>(e);
        case 6: return tree16_rot_s<6>(e);
        case 7: return tree16_rot_s<7>(e);
        case 8: return tree16_rot_s<8>(e);
        case 9: return tree16_rot_s<9>(e);
        case 10: return tree16_rot_s<10>(e);
        case 11: return tree16_rot_s<11>(e);
        case 12: return tree16_rot_s<12>(e);
        case 13: return tree16_rot_s<13>(e);
        case 14: return tree16_rot_s<14>(e);
        default: return tree16_rot_s<15>(e);
    }
}

// Sum of g(r, acc) over rows r0 <= r < RR of a column held in registers, chain (r - r0) mod
// 16: rows are visited by their static index (chains by r mod 16, rotated by r0 at the end).
// Padding rows beyond the cluster's rows hold zeros and add nothing; 16-row blocks wholly
// above r0 or below the last row are skipped (uniform branches).
template <int RR, typename A, typename G>
__device__ inline A chain_sum(int r0, int rows, G g) {
#pragma clang fp contract(off)
    A e[NCH];
#pragma unroll
    for (int m = 0; m < NCH; m++) e[m] = A(0);
    const int b0 = r0 / NCH, m0 = r0 % NCH;
#pragma unroll
    for (int b = 0; b < RR / NCH; b++) {
        if (b < b0 || NCH * b >= rows) continue;
        if (b == b0) {
#pragma unroll
            for (int m = 0; m < NCH; m++)
                if (m >= m0) e[m] = g(NCH * b + m, e[m]);
        } else {
#pragma unroll
            for (int m = 0; m < NCH; m++) e[m] = g(NCH * b + m, e[m]);
        }
        asm volatile("" ::: "memory");  // LDS operands of g are read block by block
    }
    return tree16_rot(e, m0);
}

// element r (uniform, < RR) of a register column
template <int RR, typename T>
__device__ inline T reg_at(const T* a, int r) {
    T x = T(0);
    const int b0 = r / NCH, m0 = r % NCH;
#pragma unroll
    for (int b = 0; b < RR / NCH; b++) {
        if (b != b0) continue;
#pragma unroll
        for (int m = 0; m < NCH; m++)
            if (m == m0) x = a[NCH * b + m];
    }
    return x;
}

template <typename T, int RR, int K, int NTR>
__global__ void __launch_bounds__(NTR) k_rrqr_reg(const RrqrTask* tasks, const RrqrSeg* segs, int* dsize, double tol) {
#pragma clang fp contract(off)
    static_assert(RR % NCH == 0, "rows in blocks of 16");
    constexpr int NWV = NTR / 64;
    __shared__ int s_col0[RRQR_MAXSEG + 1];
    __shared__ double s_bv[NWV];
    __shared__ int s_bp[NWV], s_bc[NWV];
    __shared__ T s_v[RR];
    __shared__ __align__(16) T s_x[RR];  // the pivot column as it stands, written by its owner
    __shared__ double s_tau, s_rii, s_scal;
    __shared__ int s_knt;
    __shared__ T s_taus[RR];  // every step's tau, stored to the task once at the end
    const RrqrTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rows = dsize[t.cluster];
    const RrqrSeg* sg = segs + t.seg0;
    if (tid == 0) {
        int c = 0;
        for (int k = 0; k < t.nseg; k++) {
            s_col0[k] = c;
            c += dsize[sg[k].nbr];
        }
        s_col0[t.nseg] = c;
    }
    for (int r = tid; r < RR; r += NTR) s_v[r] = T(0);  // padding rows of the reflector stay zero
    __syncthreads();
    const int cols = s_col0[t.nseg];
    if (rows == 0) return;
    const int mn = min(rows, cols);
    T a[K][RR];
    double vn1[K], vn2[K];
    int pos[K];  // LAPACK position of the column; -1: no column
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = tid + NTR * k;
        const bool jin = j < cols;
        gptr<const T> E = gp(reinterpret_cast<const T*>(sg[0].E));
        int sr = 0;
        if (jin) {
            const int sgi = seg_find(s_col0, t.nseg, j);
            const int jj = j - s_col0[sgi], w = s_col0[sgi + 1] - s_col0[sgi];
            E = gp(reinterpret_cast<const T*>(sg[sgi].E));
            if (sg[sgi].out) {  // out-edge block (n x s): row jj of it
                E += jj;
                sr = w;
            } else {  // in-edge block (s x n): column jj of it
                E += (size_t)jj * rows;
                sr = 1;
            }
        }
        // 16 rows at a time (bounded loads in flight, 32-bit offsets)
#pragma unroll
        for (int b = 0; b < RR / NCH; b++) {
#pragma unroll
            for (int m = 0; m < NCH; m++) {
                const int r = NCH * b + m;
                a[k][r] = (jin && r < rows) ? E[r * sr] : T(0);
            }
            asm volatile("" ::: "memory");
        }
        vn1[k] = jin ? sqrt(chain_sum<RR, double>(0, rows, [&](int r, double c) {
                     const double d = (double)a[k][r];
                     return fma(d, d, c);
                 }))
                     : 0.0;
        vn2[k] = vn1[k];
        pos[k] = jin ? j : -1;
    }
    const double tol3z = sqrt(DBL_EPSILON * 0.5);
    const int fixed = rank_fixed(tol, mn);
    const int nsteps = fixed == 0 ? 0 : mn;
    int rank = fixed >= 0 ? fixed : mn;
    double r00 = 0.0;
    for (int i = 0; i < nsteps; i++) {
        const int b0 = i / NCH, m0 = i % NCH;
        // 1. pivot: first maximum of |vn1| over the free positions i .. cols - 1
        double bv = -1.0;
        int bp = INT32_MAX, bc = -1;
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (pos[k] >= i) {
                const double v = fabs(vn1[k]);
                if (v > bv || (v == bv && pos[k] < bp)) {
                    bv = v;
                    bp = pos[k];
                    bc = tid + NTR * k;
                }
            }
        }
        wave_best_tag(bv, bp, bc);
        if (lane == 0) {
            s_bv[wv] = bv;
            s_bp[wv] = bp;
            s_bc[wv] = bc;
        }
        __syncthreads();
        double v0 = s_bv[0];
        int p0 = s_bp[0], c0 = s_bc[0];
#pragma unroll
        for (int w = 1; w < NWV; w++) {
            const double v1 = s_bv[w];
            const int p1 = s_bp[w];
            if (v1 > v0 || (v1 == v0 && p1 < p0)) {
                v0 = v1;
                p0 = p1;
                c0 = s_bc[w];
            }
        }
        // no candidate (every remaining norm NaN, after a failed pivot upstream): the column
        // at position i stays the pivot, so that no position leaves [i, cols)
        const bool valid = p0 >= i && p0 < cols;
        const int p = valid ? p0 : i;
        bool own[K];
#pragma unroll
        for (int k = 0; k < K; k++) own[k] = pos[k] >= 0 && (valid ? (tid + NTR * k == c0) : (pos[k] == i));
        // the column at position i moves to the pivot's position (dgeqp3's swap)
#pragma unroll
        for (int k = 0; k < K; k++)
            if (!own[k] && pos[k] == i) pos[k] = p;
        // 2. the pivot column's thread: dlarfg of rows i .. rows - 1 (reference dlarfg). It
        // publishes the column's rows (whole 16-row blocks in 16-byte writes) and the scalars;
        // the reflector is formed in LDS by every thread after the barrier.
        double beta_own = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!own[k]) continue;
#pragma unroll
            for (int b = 0; b < RR / NCH; b++) {
                if (b < b0 || NCH * b >= rows) continue;
                T x[NCH];
#pragma unroll
                for (int m = 0; m < NCH; m++) x[m] = a[k][NCH * b + m];
                lds_store_group<T, NCH>(s_x + NCH * b, x);
            }
            const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
            const double rs = 1.0 / safmin;
            const T alpha_t = s_x[i];  // written just above by this thread
            double alpha = (double)alpha_t;
            const int n1 = rows - i - 1;
            double tau = 0.0, scal = 1.0, beta = alpha;
            int knt = 0;
            if (n1 >= 1) {
                const double xs0 = chain_sum<RR, double>(i + 1, rows, [&](int r, double c) {
                    const double d = (double)a[k][r];
                    return fma(d, d, c);
                });
                const double xnorm = sqrt(xs0);
                if (xnorm != 0.0) {
                    beta = -copysign(lapy2_d(alpha, xnorm), alpha);
                    if (fabs(beta) < safmin) {
                        do {
                            knt++;
                            beta *= rs;
                            alpha *= rs;
                        } while (fabs(beta) < safmin && knt < 20);
                        // norm of the rescaled vector
                        const double xs = chain_sum<RR, double>(i + 1, rows, [&](int r, double c) {
                            T xv = a[k][r];
                            for (int q = 0; q < knt; q++) xv = (T)((double)xv * rs);
                            const double d = (double)xv;
                            return fma(d, d, c);
                        });
                        beta = -copysign(lapy2_d(alpha, sqrt(xs)), alpha);
                        tau = (beta - alpha) / beta;
                        scal = 1.0 / (alpha - beta);
                        for (int q = 0; q < knt; q++) beta *= safmin;
                    } else {
                        tau = (beta - alpha) / beta;
                        scal = 1.0 / (alpha - beta);
                    }
                }
            }
            // R_ii as stored: beta, or alpha where tau = 0 leaves the column as it is
            s_rii = fabs((double)(tau != 0.0 ? (T)beta : alpha_t));
            s_tau = tau;
            s_scal = scal;
            s_knt = knt;
            beta_own = beta;
            s_taus[i] = (T)tau;
        }
        __syncthreads();
        const double taud = s_tau, rii = s_rii;
        if (i == 0) {
            r00 = rii;
        } else if (fixed < 0 && !(rii / r00 >= tol)) {  // |R_ii| / |R_00| < tol: rank i
            rank = i;
            break;
        }
        // the reflector v (v_i = 1, v_r = x_r scaled below; used only when tau != 0), one row
        // per thread; rows past the last stay zero
        if (taud != 0.0) {
            const int knt = s_knt;
            const double scal = s_scal;
            for (int r = i + tid; r < rows; r += NTR) s_v[r] = r == i ? T(1) : larfg_scale<T>(s_x[r], knt, scal);
        }
        __syncthreads();
        // the owner's column keeps beta and v (with tau = 0 it is left as it is): formed in
        // its registers beside the other threads' updates (s_knt and s_scal are rewritten
        // only after the next step's first barrier)
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!own[k]) continue;
            pos[k] = i;
            if (taud == 0.0) continue;
            const int knt = s_knt;
            const double scal = s_scal;
#pragma unroll
            for (int b = 0; b < RR / NCH; b++) {
                if (b < b0 || NCH * b >= rows) continue;
#pragma unroll
                for (int m = 0; m < NCH; m++) {
                    const int r = NCH * b + m;
                    if (b > b0 || m > m0) a[k][r] = larfg_scale<T>(a[k][r], knt, scal);
                    else if (m == m0) a[k][r] = (T)beta_own;
                }
            }
        }
        const T tau = (T)taud;
        // 3. apply H(i) to the free columns (positions > i after the swap), downdate the norms
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!(pos[k] > i)) continue;
            if (tau != T(0)) {
                const T w = chain_sum<RR, T>(i, rows, [&](int r, T c) { return fma(a[k][r], s_v[r], c); });
                if (w != T(0)) {
                    const T tm = -tau * w;
                    asm volatile("" ::: "memory");  // re-read v from LDS instead of holding it in registers
#pragma unroll
                    for (int b = 0; b < RR / NCH; b++) {
                        if (b < b0 || NCH * b >= rows) continue;
#pragma unroll
                        for (int m = 0; m < NCH; m++) {
                            const int r = NCH * b + m;
                            if (b > b0 || m > m0) a[k][r] = fma(s_v[r], tm, a[k][r]);
                            else if (m == m0) a[k][r] = a[k][r] + tm;
                        }
                    }
                }
            }
            if (vn1[k] != 0.0) {
                const double q = fabs((double)reg_at<RR>(a[k], i)) / vn1[k];
                double temp = 1.0 - q * q;
                temp = fmax(temp, 0.0);
                const double rr = vn1[k] / vn2[k];
                const double temp2 = temp * rr * rr;
                if (temp2 <= tol3z) {
                    if (i < rows - 1) {
                        vn1[k] = sqrt(chain_sum<RR, double>(i + 1, rows, [&](int r, double c) {
                            const double d = (double)a[k][r];
                            return fma(d, d, c);
                        }));
                        vn2[k] = vn1[k];
                    } else {
                        vn1[k] = 0.0;
                        vn2[k] = 0.0;
                    }
                } else {
                    vn1[k] = vn1[k] * sqrt(temp);
                }
            }
        }
    }
    if (rank >= rows) return;
    // tau of the kept reflectors (each written to LDS in its step, before that step's barrier)
    for (int r = tid; r < rank; r += NTR) gp(reinterpret_cast<T*>(t.tau))[r] = s_taus[r];
    // Q: the reflectors of positions 0 .. rank - 1, column by column (R above, beta, v below)
    const gptr<T> Q = gp(reinterpret_cast<T*>(t.Q));
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (pos[k] < 0 || pos[k] >= rank) continue;
        const gptr<T> q = Q + (size_t)pos[k] * rows;
#pragma unroll
        for (int r = 0; r < RR; r++)
            if (r < rows) q[r] = a[k][r];
    }
    // AsnP[:, j] = triu(R)[:rank, pos(j)], scattered into the shrunk edge blocks (every
    // gather read of these blocks happened before the first barrier)
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (pos[k] < 0) continue;
        const int j = tid + NTR * k;
        const int sgi = seg_find(s_col0, t.nseg, j);
        const int jj = j - s_col0[sgi], w = s_col0[sgi + 1] - s_col0[sgi];
        const gptr<T> E = gp(reinterpret_cast<T*>(sg[sgi].E));
        const bool out = sg[sgi].out;
#pragma unroll
        for (int r = 0; r < RR; r++) {
            if (r < rank) {
                const T v = (r <= pos[k]) ? a[k][r] : T(0);
                if (out) E[jj + (size_t)r * w] = v;
                else E[r + (size_t)jj * rank] = v;
            }
        }
    }
    if (tid == 0) dsize[t.cluster] = rank;
}

// ---------------------------------------------------------------------------------------
// Stepped RRQR for clusters too large for one workgroup. Launch sequence per wave:
// k_rrqr_init (gather + norms), k_rrqr_step for every pivot step, k_rrqr_finish (rank,
// Q, scatter). The columns are split into 64-wide slabs, one 1024-thread workgroup each:
// lane l of wavefront w handles column l of the slab and chain w (rows i+w, i+w+16, ...)
// of every sum. Each step: every slab reads the candidates of all slabs and picks the
// pivot (first max by position), recomputes the pivot's dlarfg scalars redundantly
// (identical in every workgroup), and updates / downdates its own free columns. The
// pivot column is only read during its step; its owner writes the scaled reflector one
// step later, so no data crosses workgroups inside a launch.
constexpr int SLAB = 64;
constexpr int STEP_NT = SLAB * NCH;

// Stepped-path storage: 64-column slabs, row-major inside a slab, so a wavefront reading
// one row of its slab touches one contiguous segment.
__device__ inline size_t slab_idx(int rows, int r, int j) {
    return ((size_t)(j / SLAB) * rows + r) * SLAB + (j % SLAB);
}

// Relay words of the persistent RRQR: 64-bit words (tag << 32 | 32-bit payload) stored and
// loaded whole by device-scope relaxed atomics (on gfx950 these bypass the per-XCD L2
// caches, which are not coherent with each other). A reader that sees the tag it waits for
// sees the payload of the same store, so no fence orders data and flag: each word is its
// own flag. Tags are step + 1; k_rrqr_init clears the words to tag 0 before the launch.
using RelayWord = unsigned long long;
typedef __attribute__((address_space(1))) RelayWord GRelayWord;

__device__ inline void relay_put(GRelayWord* p, unsigned tag, unsigned v) {
    __hip_atomic_store(p, ((RelayWord)tag << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// a value of T as sizeof(T) / 4 relay words
template <typename T>
__device__ inline void relay_put_val(GRelayWord* p, unsigned tag, T x) {
    if constexpr (sizeof(T) == 4) {
        relay_put(p, tag, __float_as_uint(x));
    } else {
        const unsigned long long b = (unsigned long long)__double_as_longlong(x);
        relay_put(p, tag, (unsigned)b);
        relay_put(p + 1, tag, (unsigned)(b >> 32));
    }
}
// Waits until the first n of the words p[0..N) all carry `tag` (n loads in flight per poll,
// one round trip when the data is there); payloads in v. False after spin_max polls
// (RRQR_SPIN_MAX unless the caller lowers it) or once *fail is set: the launch is
// cooperative, so every slab is resident and a correct schedule never comes near the
// bound; the host turns the flag into an error.
template <int N>
__device__ inline bool relay_get_all(const GRelayWord* const (&p)[N], int n, unsigned tag, unsigned (&v)[N],
                                     int* fail, long long spin_max) {
    for (long long spins = 0;; spins++) {
        RelayWord w[N];
#pragma unroll
        for (int k = 0; k < N; k++)
            w[k] = k < n ? __hip_atomic_load(p[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ((RelayWord)tag << 32);
        bool ok = true;
#pragma unroll
        for (int k = 0; k < N; k++) ok = ok && (unsigned)(w[k] >> 32) == tag;
        if (ok) {
#pragma unroll
            for (int k = 0; k < N; k++) v[k] = (unsigned)w[k];
            return true;
        }
        if (spins >= spin_max ||
            ((spins & 63) == 63 && __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
            __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}
// relay layout of one cluster: candidates [2][nslab][4] (norm low, norm high, position),
// then candidate columns [2][nslab][max_rows * sizeof(T) / 4]
template <typename T>
__device__ inline GRelayWord* relay_cand(const RrqrBig& B, int buf, int slab) {
    return (GRelayWord*)(B.vbuf) + ((size_t)buf * B.nslab + slab) * 4;
}
template <typename T>
__device__ inline GRelayWord* relay_col(const RrqrBig& B, int buf, int slab) {
    return (GRelayWord*)(B.vbuf) + (size_t)2 * B.nslab * 4 +
           ((size_t)buf * B.nslab + slab) * B.t.max_rows * (sizeof(T) / 4);
}
__device__ inline void relay_put_cand(GRelayWord* p, unsigned tag, double v, int pos) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    relay_put(p, tag, (unsigned)b);
    relay_put(p + 1, tag, (unsigned)(b >> 32));
    relay_put(p + 2, tag, (unsigned)pos);
}

// pivot column and reflector buffers (2 x max_rows): LDS, or the task's global buffer
template <typename T>
__device__ inline T* step_buf(const RrqrBig& B, unsigned char* smem, int slab) {
    return B.vbuf ? reinterpret_cast<T*>(B.vbuf) + (size_t)slab * 2 * B.t.max_rows : reinterpret_cast<T*>(smem);
}

// Host-visible signals of the per-step launches (fine-grained pinned host memory): plain
// system-scope stores that the host polls to stop launching steps that would do nothing.
__device__ inline void signal_store(long long* p, long long v) {
    if (p) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T>
__global__ void __launch_bounds__(STEP_NT) k_rrqr_init(const RrqrBig* bigs, const RrqrSlab* slabs, const RrqrSeg* segs,
                                                       const int* dsize, double tol) {
#pragma clang fp contract(off)
    __shared__ int s_col0[RRQR_MAXSEG + 1];
    __shared__ double s_part[NCH][SLAB];
    const RrqrSlab sl = slabs[blockIdx.x];
    const RrqrBig B = bigs[sl.big];
    const RrqrSeg* sg = segs + B.t.seg0;
    const int tid = threadIdx.x, lane = tid % SLAB, wv = tid / SLAB;
    const int rows = dsize[B.t.cluster];
    if (tid == 0) {
        int c = 0;
        for (int k = 0; k < B.t.nseg; k++) {
            s_col0[k] = c;
            c += dsize[sg[k].nbr];
        }
        s_col0[B.t.nseg] = c;
    }
    __syncthreads();
    const int cols = s_col0[B.t.nseg];
    if (sl.slab == 0 && tid == 0) {
        const int mn = min(rows, cols);
        B.dims[0] = rows;
        B.dims[1] = cols;
        B.dims[2] = mn;
        B.dims[3] = rank_fixed(tol, mn) == 0 ? 0 : mn;  // steps to run (lowered at the rank)
        B.dims[4] = 2;  // 2: the steps store every kept reflector in Q; 1: the wave kernel left them in A
        if (B.dims[3] == 0) signal_store(B.done, 1);
    }
    if (B.relay) {  // the persistent launch's relay words of this slab: tag 0 (no step)
        for (int b = 0; b < 2; b++) {
            GRelayWord* c = relay_cand<T>(B, b, sl.slab);
            if (tid < 4) c[tid] = 0;
            GRelayWord* col = relay_col<T>(B, b, sl.slab);
            for (int r = tid; r < B.t.max_rows * (int)(sizeof(T) / 4); r += STEP_NT) col[r] = 0;
        }
    }
    const gptr<T> A = gp(reinterpret_cast<T*>(B.t.scratch));
    const int mc = B.t.max_cols;
    const int j = sl.slab * SLAB + lane;
    const bool active = j < cols && rows > 0;
    double c = 0.0;
    if (active) {
        const int k = seg_find(s_col0, B.t.nseg, j);
        const int jj = j - s_col0[k], w = s_col0[k + 1] - s_col0[k];
        const T* E = reinterpret_cast<const T*>(sg[k].E);
        const auto a = A + (size_t)sl.slab * rows * SLAB + lane;
        for (int i = wv; i < rows; i += NCH) {
            const T v = sg[k].out ? E[jj + (size_t)i * w] : E[i + (size_t)jj * rows];
            a[(size_t)i * SLAB] = v;
            const double dv = (double)v;
            c = fma(dv, dv, c);
        }
    }
    s_part[wv][lane] = c;
    __syncthreads();
    if (wv == 0) {
        double best = -1.0;
        int bpos = INT32_MAX;
        if (active) {
            const double nrm = sqrt(tree16_lds(&s_part[0][lane], SLAB));
            B.t.norms[j] = nrm;
            B.t.norms[mc + j] = nrm;
            B.map[j] = j;
            B.map[2 * mc + j] = j;
            best = nrm;
            bpos = j;
        }
        wave_best(best, bpos);
        if (lane == 0) {
            B.cand_v[sl.slab] = best;
            B.cand_p[sl.slab] = bpos;
        }
    }
}

// One pivot step of the stepped RRQR. RW > 0: each chain's rows of the thread's column
// (r = i + c + 16 q, q < RW; rows <= 16 RW) are loaded into registers at the start, before
// the pivot is known, and serve both the w_j sums and the update; RW = 0 re-reads them.
// NW wavefronts per workgroup: wavefront wv holds chains c = wv + NW k (k < 16 / NW), each
// with its own accumulator, so every sum runs in the same 16-chain order for any NW; fewer
// wavefronts per slab keep a wave with many slabs resident in one round on the chip.
// The pivot column's owner stores the step's reflector (R above, beta, v below) in Q.
template <typename T, int RW, int NW>
__global__ void __launch_bounds__(SLAB * NW) k_rrqr_step(const RrqrBig* bigs, const RrqrSlab* slabs, int step, double tol,
                                                         long long* prog, long long seq) {
#pragma clang fp contract(off)
    static_assert(NCH % NW == 0 && (RW > 0 || NW == NCH), "chains split evenly; RW = 0 only with 16 wavefronts");
    constexpr int NTH = SLAB * NW, CPW = NCH / NW;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double s_part[NCH][SLAB];
    __shared__ double sh[4], s_ch[NCH + 1];
    __shared__ T s_tm[SLAB];
    __shared__ int s_upd[SLAB];
    __shared__ int s_p, s_ca, s_cb, s_stop;
    const RrqrSlab sl = slabs[blockIdx.x];
    const RrqrBig B = bigs[sl.big];
    const int tid = threadIdx.x, lane = tid % SLAB, wv = tid / SLAB;
    if (blockIdx.x == 0 && tid == 0) signal_store(prog, seq);  // this launch has started
    // dims[3] may be lowered by slab 0 during this launch: read it once per workgroup
    const gptr<int> dims = gp(B.dims);
    if (tid == 0) s_stop = dims[3];
    const int rows = dims[0], cols = dims[1], mn = dims[2];
    const int i = step;
    const int j = sl.slab * SLAB + lane;
    const gptr<T> A = gp(reinterpret_cast<T*>(B.t.scratch));
    const gptr<T> cj = A + (size_t)sl.slab * rows * SLAB + lane;  // element r of column j at cj[r * SLAB]
    __syncthreads();
    const int stop = s_stop;
    if (rows == 0 || i >= mn || i >= stop) {
        if (sl.slab == 0 && tid == 0) signal_store(B.done, 1);
        return;
    }
    // this thread's rows, loaded while wavefront 0 picks the pivot
    constexpr int NR = RW > 0 ? RW : 1;
    T a[CPW][NR];
    if (RW > 0 && j < cols) {
#pragma unroll
        for (int k = 0; k < CPW; k++)
#pragma unroll
            for (int q = 0; q < NR; q++) {
                const int r = i + wv + NW * k + NCH * q;
                a[k][q] = r < rows ? cj[(size_t)r * SLAB] : T(0);
            }
    }
    const int cur = i & 1, nxt = cur ^ 1;
    const int mc = B.t.max_cols, ns = B.nslab;
    const gptr<int> map = gp(B.map);
    const gptr<const int> p2c_cur = map + cur * mc;
    const gptr<int> p2c_nxt = map + nxt * mc;
    const gptr<const int> c2p_cur = map + 2 * mc + cur * mc;
    const gptr<int> c2p_nxt = map + 2 * mc + nxt * mc;
    const gptr<double> vn1 = gp(B.t.norms);
    const gptr<double> vn2 = vn1 + mc;
    const gptr<double> cand_v = gp(B.cand_v);
    const gptr<int> cand_p = gp(B.cand_p);
    // the pivot column and its reflector (v[i] = 1): always in LDS when rows are cached in
    // registers (at most 16 RW rows; the launcher checks the LDS size)
    T* s_col = RW > 0 ? reinterpret_cast<T*>(smem_raw) : step_buf<T>(B, smem_raw, sl.slab);
    T* s_v = s_col + B.t.max_rows;
    if (wv == 0) {  // pivot: first maximum over the slabs' candidates
        double bv = -1.0;
        int bp = INT32_MAX;
        for (int k = lane; k < ns; k += SLAB) {
            const double v = cand_v[cur * ns + k];
            const int pp = cand_p[cur * ns + k];
            if (v > bv || (v == bv && pp < bp)) {
                bv = v;
                bp = pp;
            }
        }
        wave_best(bv, bp);
        if (lane == 0) {
            s_p = bp;
            s_ca = (bp >= i && bp < cols) ? p2c_cur[bp] : 0;
            s_cb = p2c_cur[i];
        }
    }
    __syncthreads();
    const int p = s_p, ca = s_ca, cb = s_cb;
    if (p < i || p >= cols) return;  // cannot happen for i < mn: some column is still free
    if (sl.slab == 0) {
        for (int jj = tid; jj < cols; jj += NTH) {
            p2c_nxt[jj] = p2c_cur[jj];
            c2p_nxt[jj] = c2p_cur[jj];
        }
        __syncthreads();
        if (tid == 0) {
            p2c_nxt[i] = ca;
            p2c_nxt[p] = cb;
            c2p_nxt[ca] = i;
            c2p_nxt[cb] = p;
        }
    }
    // this step's reflector, recomputed identically in every slab
    for (int r = i + tid; r < rows; r += NTH) s_col[r] = A[slab_idx(rows, r, ca)];
    __syncthreads();
    larfg_block<T>((double)s_col[i], s_col + i + 1, rows - i - 1, sh, s_ch);
    // R_ii as stored (tau = 0 leaves alpha on the diagonal)
    const double rii = fabs((double)(T)sh[1]);
    if (i == 0) {
        if (sl.slab == 0 && tid == 0) cand_v[2 * ns] = rii;
    } else if (rank_fixed(tol, mn) < 0 && !(rii / cand_v[2 * ns] >= tol)) {
        if (sl.slab == 0 && tid == 0) {
            dims[3] = i;  // rank = i: step i and later are not kept
            signal_store(B.done, 1);
        }
        return;
    }
    if (sl.slab == 0 && tid == 0 && i + 1 >= mn) signal_store(B.done, 1);  // the last step
    const double taud = sh[0];
    const T tau = (T)taud;
    if (taud != 0.0) {
        const int knt = (int)sh[3];
        for (int r = i + 1 + tid; r < rows; r += NTH) s_v[r] = larfg_scale<T>(s_col[r], knt, sh[2]);
    }
    if (tid == 0) s_v[i] = T(1);
    __syncthreads();
    if (ca / SLAB == sl.slab) {  // the owner stores the reflector: Q[:, i] = (R(0:i, ca), beta, v)
        const gptr<T> Qi = gp(reinterpret_cast<T*>(B.t.Q)) + (size_t)i * rows;
        for (int r = tid; r < rows; r += NTH) {
            T v;
            if (r < i) v = A[slab_idx(rows, r, ca)];
            else if (taud == 0.0) v = s_col[r];
            else v = r == i ? (T)sh[1] : s_v[r];
            Qi[r] = v;
        }
        if (tid == 0) gp(reinterpret_cast<T*>(B.t.tau))[i] = tau;
    }
    const int pos = (j == cb) ? p : (j < cols ? c2p_cur[j] : 0);
    const bool active = j < cols && j != ca && pos > i;
    // w_j = v^T a_j: chain c = wv + NW k of the sum
    if (taud != 0.0) {
        T c[CPW];
#pragma unroll
        for (int k = 0; k < CPW; k++) c[k] = T(0);
        if (active) {
            if (RW > 0) {
#pragma unroll
                for (int k = 0; k < CPW; k++)
#pragma unroll
                    for (int q = 0; q < NR; q++) {
                        const int r = i + wv + NW * k + NCH * q;
                        if (r < rows) c[k] = fma(a[k][q], s_v[r], c[k]);
                    }
            } else {
                // four rows of the chain loaded ahead of the (ordered) accumulation
                int r = i + wv;
                T c0 = T(0);
                for (; r + 3 * NCH < rows; r += 4 * NCH) {
                    const T a0 = cj[(size_t)r * SLAB], a1 = cj[(size_t)(r + NCH) * SLAB];
                    const T a2 = cj[(size_t)(r + 2 * NCH) * SLAB], a3 = cj[(size_t)(r + 3 * NCH) * SLAB];
                    c0 = fma(a0, s_v[r], c0);
                    c0 = fma(a1, s_v[r + NCH], c0);
                    c0 = fma(a2, s_v[r + 2 * NCH], c0);
                    c0 = fma(a3, s_v[r + 3 * NCH], c0);
                }
                for (; r < rows; r += NCH) c0 = fma(cj[(size_t)r * SLAB], s_v[r], c0);
                c[0] = c0;
            }
        }
#pragma unroll
        for (int k = 0; k < CPW; k++) s_part[wv + NW * k][lane] = (double)c[k];
        __syncthreads();
        if (wv == 0) {
            const T w = tree16_lds_t<T>(&s_part[0][lane], SLAB);
            s_tm[lane] = -tau * w;
            s_upd[lane] = w != T(0);  // dger skips columns with w_j = 0
        }
        __syncthreads();
    }
    // update a_j := a_j + tm v and the chains of sum over rows > i of a_j^2 (the norm
    // recomputation): row r > i of chain c belongs to chain (r - i - 1) % 16 = (c + 15) % 16
    T ai = T(0);
    double c2[CPW];
#pragma unroll
    for (int k = 0; k < CPW; k++) c2[k] = 0.0;
    if (active) {
        const bool upd = taud != 0.0 && s_upd[lane];
        const T tm = upd ? s_tm[lane] : T(0);
        if (RW > 0) {
#pragma unroll
            for (int k = 0; k < CPW; k++)
#pragma unroll
                for (int q = 0; q < NR; q++) {
                    const int r = i + wv + NW * k + NCH * q;
                    if (r < rows) {
                        T x = a[k][q];
                        if (upd) {
                            x = (r == i) ? x + tm : fma(s_v[r], tm, x);
                            cj[(size_t)r * SLAB] = x;
                        }
                        if (r == i) {
                            ai = x;
                        } else {
                            const double d = (double)x;
                            c2[k] = fma(d, d, c2[k]);
                        }
                    }
                }
        } else {
            double c20 = 0.0;
            int r = i + wv;
            if (r == i && r < rows) {  // row i, chain 0's first row
                T x = cj[(size_t)r * SLAB];
                if (upd) {
                    x = x + tm;
                    cj[(size_t)r * SLAB] = x;
                }
                ai = x;
                r += NCH;
            }
            for (; r + 3 * NCH < rows; r += 4 * NCH) {
                T x[4];
#pragma unroll
                for (int q = 0; q < 4; q++) x[q] = cj[(size_t)(r + q * NCH) * SLAB];
                if (upd) {
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        x[q] = fma(s_v[r + q * NCH], tm, x[q]);
                        cj[(size_t)(r + q * NCH) * SLAB] = x[q];
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const double d = (double)x[q];
                    c20 = fma(d, d, c20);
                }
            }
            for (; r < rows; r += NCH) {
                T x = cj[(size_t)r * SLAB];
                if (upd) {
                    x = fma(s_v[r], tm, x);
                    cj[(size_t)r * SLAB] = x;
                }
                const double d = (double)x;
                c20 = fma(d, d, c20);
            }
            c2[0] = c20;
        }
    }
#pragma unroll
    for (int k = 0; k < CPW; k++) s_part[(wv + NW * k + NCH - 1) % NCH][lane] = c2[k];
    __syncthreads();
    if (wv == 0) {
        const double tol3z = sqrt(DBL_EPSILON * 0.5);
        double best = -1.0;
        int bpos = INT32_MAX;
        if (active) {
            if (vn1[j] != 0.0) {
                const double q = fabs((double)ai) / vn1[j];
                double temp = 1.0 - q * q;
                temp = fmax(temp, 0.0);
                const double rr = vn1[j] / vn2[j];
                const double temp2 = temp * rr * rr;
                if (temp2 <= tol3z) {
                    if (i < rows - 1) {
                        vn1[j] = sqrt(tree16_lds(&s_part[0][lane], SLAB));
                        vn2[j] = vn1[j];
                    } else {
                        vn1[j] = 0.0;
                        vn2[j] = 0.0;
                    }
                } else {
                    vn1[j] = vn1[j] * sqrt(temp);
                }
            }
            best = vn1[j];
            bpos = pos;
        }
        wave_best(best, bpos);
        if (lane == 0) {
            cand_v[nxt * ns + sl.slab] = best;
            cand_p[nxt * ns + sl.slab] = bpos;
        }
    }
}

// Persistent form of the stepped RRQR: one cooperative launch runs every pivot step of the
// clusters of a wave (or of the part of it that the device holds at once), and each
// cluster's slabs pace each other only. After its update at step i, every slab publishes
// its best column (candidate for pivot i + 1: partial norm, position, rows i + 1..) in relay
// words tagged i + 2; at step i + 1 each slab waits for the candidates of its cluster,
// takes the first maximum and reads the winner's rows, then recomputes the step's dlarfg
// itself: two reads through memory per step (the candidates, then the winner's rows), one
// publication, and no fences.
// Wavefront wv of NW holds, for its lane's column, rows r = 16 q + CPW wv + k (k < CPW =
// 16 / NW, q < RW) in registers for the whole factorization. Row r belongs to chain
// (r - i) % 16 of the w_j sums at step i and to chain (r - i - 1) % 16 of the norm sums,
// each chain summed in increasing rows, so every sum runs in the 16-chain order of
// k_rrqr_step. The sums run over every cached row with the reflector zero above row i and
// past the last row (s_v) and a 0/1 mask for the norms (s_m): a zero term leaves a chain's
// sum unchanged bit for bit, and the loops test rows only per group of CPW (skipping the
// groups wholly above row i or past the last row). The pivot's owner keeps the reflector
// (beta, v) in its registers; the registers and the final maps go back to the slab storage
// at the end, where k_rrqr_finish copies the reflectors into Q.
// Wavefronts per SIMD the register budget targets: the cached values plus 96 registers
// (a few loop-invariant values may spill); 128 cached registers take the whole file (the
// accumulation registers hold what does not fit).
template <typename T, int RW, int NW>
constexpr int wave_occupancy() {
    constexpr int c = (NCH / NW) * RW * (int)(sizeof(T) / 4);
    return c > 64 ? 1 : 512 / (c + 96) > 8 ? 8 : 512 / (c + 96);
}
template <typename T, int RW, int NW>
__global__ void __launch_bounds__(SLAB * NW, (wave_occupancy<T, RW, NW>()))
    k_rrqr_wave(const RrqrBig* bigs, const RrqrSlab* slabs, int nsteps, double tol, int* fail, long long spin_max) {
#pragma clang fp contract(off)
    static_assert(NCH % NW == 0, "chains split evenly over the wavefronts");
    constexpr int NTH = SLAB * NW, CPW = NCH / NW, RMAX = NCH * RW;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double s_part[NCH][SLAB];
    __shared__ double sh[4], s_ch[NCH + 1];
    __shared__ __align__(16) T s_col[RMAX];  // the pivot column
    __shared__ __align__(16) T s_v[RMAX];    // its reflector; zero above row i and past the last row
    __shared__ __align__(16) T s_m[RMAX];    // 1 for rows i + 1 .. rows - 1, else 0
    __shared__ T s_tm[SLAB], s_rowi[SLAB];
    __shared__ int s_upd[SLAB], s_pos[SLAB], s_rc[SLAB];
    __shared__ int s_p, s_ca, s_cb, s_best, s_ok, s_ok_col, s_anyrc;
    const RrqrSlab sl = slabs[blockIdx.x];
    const RrqrBig B = bigs[sl.big];
    const int tid = threadIdx.x, lane = tid % SLAB;
    const int wv = __builtin_amdgcn_readfirstlane(tid / SLAB);  // uniform: row tests are scalar
    const int r0 = CPW * wv;                                      // first cached row
    const gptr<const int> dims = gp(reinterpret_cast<const int*>(B.dims));
    const int rows = dims[0], cols = dims[1], mn = dims[2];
    int stop = dims[3];
    const int ns = B.nslab, mc = B.t.max_cols;
    int* p2c = reinterpret_cast<int*>(smem_raw);  // position -> column, this workgroup's copy
    const gptr<T> tau_out = gp(reinterpret_cast<T*>(B.t.tau));
    const int j = sl.slab * SLAB + lane;
    const bool jin = j < cols && rows > 0;
    T a[CPW][RW];
    {
        const gptr<const T> aslab = gp(reinterpret_cast<const T*>(B.t.scratch)) + (size_t)sl.slab * rows * SLAB + lane;
#pragma unroll
        for (int q = 0; q < RW; q++)
#pragma unroll
            for (int k = 0; k < CPW; k++) {
                const int r = NCH * q + r0 + k;
                a[k][q] = (jin && r < rows) ? aslab[(size_t)r * SLAB] : T(0);
            }
    }
    for (int k = tid; k < cols; k += NTH) p2c[k] = k;
    for (int r = tid; r < RMAX; r += NTH) {
        s_v[r] = T(0);
        s_m[r] = r < rows ? T(1) : T(0);
    }
    if (tid < SLAB) s_pos[tid] = j;
    double vn1 = 0.0, vn2 = 0.0;
    if (wv == 0 && jin) {
        vn1 = B.t.norms[j];
        vn2 = B.t.norms[mc + j];
    }
    const double tol3z = sqrt(DBL_EPSILON * 0.5);
    const int fixed = rank_fixed(tol, mn);
    double sref = 0.0;
    // candidate of step 0 (k_rrqr_init left the slab's best in cand_*[slab]), tag 1; its
    // column goes through LDS as in the steps
    const bool run = rows > 0 && mn > 0 && stop > 0;
    const int bp0 = run ? B.cand_p[sl.slab] : INT32_MAX;
    if (run && tid == 0) relay_put_cand(relay_cand<T>(B, 0, sl.slab), 1u, B.cand_v[sl.slab], bp0);
    if (bp0 != INT32_MAX && lane == bp0 % SLAB) {
#pragma unroll
        for (int q = 0; q < RW; q++) {
            T x[CPW];
#pragma unroll
            for (int k = 0; k < CPW; k++) x[k] = a[k][q];
            lds_store_group<T, CPW>(s_col + NCH * q + r0, x);
        }
    }
    __syncthreads();
    if (bp0 != INT32_MAX) {
        GRelayWord* dst = relay_col<T>(B, 0, sl.slab);
        for (int r = tid; r < rows; r += NTH) relay_put_val<T>(dst + (size_t)r * (sizeof(T) / 4), 1u, s_col[r]);
    }
    bool abandoned = false;
    for (int i = 0; i < nsteps; i++) {
        if (rows == 0 || i >= mn || i >= stop) break;  // the same decision in every slab of the cluster
        const int cur = i & 1, nxt = cur ^ 1;
        const unsigned tag = (unsigned)i + 1u;
        // this wavefront's row window [i, rows) relative to r0, opaque to the compiler so
        // that no per-row test is hoisted out of the step loop (64 live scalar masks)
        int iw = i - r0, rw = rows - r0;
        asm volatile("" : "+s"(rw));
        if (wv == 0) {  // pivot: first maximum over the slabs' candidates
            double bv = -1.0;
            int bp = INT32_MAX;
            bool ok = true;
            for (int k = lane; k < ns && ok; k += SLAB) {
                const GRelayWord* c = relay_cand<T>(B, cur, k);
                const GRelayWord* const pc[3] = {c, c + 1, c + 2};
                unsigned u[3];
                ok = relay_get_all<3>(pc, 3, tag, u, fail, spin_max);
                if (ok) {
                    const double v = __longlong_as_double((long long)(((unsigned long long)u[1] << 32) | u[0]));
                    const int pp = (int)u[2];
                    if (v > bv || (v == bv && pp < bp)) {
                        bv = v;
                        bp = pp;
                    }
                }
            }
            wave_best(bv, bp);
            ok = __all(ok);
            if (lane == 0) {
                s_ok = ok;
                s_ok_col = 1;  // cleared below by a failed read of the winner's rows
                s_p = bp;
                s_ca = (bp >= i && bp < cols) ? p2c[bp] : 0;
                s_cb = p2c[i];
            }
        }
        __syncthreads();
        if (!s_ok) {
            abandoned = true;
            break;
        }
        const int p = s_p, ca = s_ca, cb = s_cb;
        if (p < i || p >= cols) {  // cannot happen for i < mn: some column is still free
            stop = i;
            break;
        }
        {  // the winner's rows i.. (published by the owner of column ca at step i - 1)
            static_assert(RMAX <= 2 * NTH, "at most two rows per thread");
            constexpr int WD = (int)(sizeof(T) / 4), NWD = 2 * WD;
            const GRelayWord* src = relay_col<T>(B, cur, ca / SLAB);
            const int r1 = i + tid, r2 = i + tid + NTH;
            const int nr = (r1 < rows) + (r2 < rows);
            if (nr > 0) {
                const GRelayWord* pw[NWD];
#pragma unroll
                for (int k = 0; k < NWD; k++) pw[k] = src + (size_t)(k < WD ? r1 : r2) * WD + k % WD;
                unsigned u[NWD];
                if (relay_get_all<NWD>(pw, nr * WD, tag, u, fail, spin_max)) {
                    T x[2];
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        if constexpr (sizeof(T) == 4) x[h] = __uint_as_float(u[h]);
                        else x[h] = __longlong_as_double((long long)(((unsigned long long)u[2 * h + 1] << 32) | u[2 * h]));
                    }
                    s_col[r1] = x[0];
                    if (nr > 1) s_col[r2] = x[1];
                } else {
                    s_ok_col = 0;
                }
            }
        }
        __syncthreads();
        if (!s_ok_col) {
            abandoned = true;
            break;
        }
        larfg_block<T>((double)s_col[i], s_col + i + 1, rows - i - 1, sh, s_ch);
        // R_ii as stored (tau = 0 leaves alpha on the diagonal); the QR stops at the rank
        const double rii = fabs((double)(T)sh[1]);
        if (i == 0) {
            sref = rii;
        } else if (fixed < 0 && !(rii / sref >= tol)) {
            stop = i;
            break;
        }
        const double taud = sh[0];
        const T tau = (T)taud;
        if (taud != 0.0) {
            const int knt = (int)sh[3];
            for (int r = i + 1 + tid; r < rows; r += NTH) s_v[r] = larfg_scale<T>(s_col[r], knt, sh[2]);
        }
        const bool own_pivot = ca / SLAB == sl.slab;
        if (tid == 0) {
            if (i > 0) s_v[i - 1] = T(0);
            s_v[i] = T(1);
            s_m[i] = T(0);
            p2c[i] = ca;
            p2c[p] = cb;
            if (own_pivot) {
                s_pos[ca % SLAB] = i;
                tau_out[i] = tau;
            }
            if (cb / SLAB == sl.slab && cb != ca) s_pos[cb % SLAB] = p;
        }
        __syncthreads();
        const int pos = s_pos[lane];
        const bool active = jin && pos > i;
        // a_j += tm_j v (v is 1 on row i and 0 above it) on this thread's rows of its column
        auto update_column = [&](T tmj) {
#pragma unroll
            for (int q = 0; q < RW; q++) {
                if (NCH * q + CPW <= iw || NCH * q >= rw) continue;
                T v[CPW];
                lds_row_group<T, CPW>(s_v + NCH * q + r0, v);
#pragma unroll
                for (int k = 0; k < CPW; k++) a[k][q] = fma(v[k], tmj, a[k][q]);
            }
        };
        // this column's update of the step, applied once: to the columns whose norm must be
        // recomputed and to the slab's candidate before the candidate is published, to the
        // others after it (off the path the other slabs wait on)
        T tm = T(0);
        bool pending = false;
        if (taud != 0.0) {  // w_j = v^T a_j, chain (r - i) % 16
            T c[CPW];
#pragma unroll
            for (int k = 0; k < CPW; k++) c[k] = T(0);
            if (active) {
#pragma unroll
                for (int q = 0; q < RW; q++) {
                    if (NCH * q + CPW <= iw || NCH * q >= rw) continue;  // rows all above i or past the last
                    T v[CPW];
                    lds_row_group<T, CPW>(s_v + NCH * q + r0, v);
#pragma unroll
                    for (int k = 0; k < CPW; k++) c[k] = fma(a[k][q], v[k], c[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < CPW; k++) s_part[(r0 + k - i % NCH + NCH) % NCH][lane] = (double)c[k];
            __syncthreads();
            if (wv == 0) {
                const T w = tree16_lds_t<T>(&s_part[0][lane], SLAB);
                s_tm[lane] = -tau * w;
                s_upd[lane] = w != T(0);  // dger skips columns with w_j = 0
            }
            __syncthreads();
            pending = active && s_upd[lane];
            tm = pending ? s_tm[lane] : T(0);
        }
        // row i of every column after the update (its new value is a_ij + tm: v_i = 1), for
        // the norm downdate
        if (iw >= 0 && iw / NCH < RW && iw % NCH < CPW) {
            const int kq = iw % NCH, qq = iw / NCH;
            T x = T(0);
#pragma unroll
            for (int q = 0; q < RW; q++)
#pragma unroll
                for (int k = 0; k < CPW; k++)
                    if (k == kq && q == qq) x = a[k][q];
            s_rowi[lane] = pending ? x + tm : x;
        }
        __syncthreads();
        // downdate (wavefront 0): the columns whose downdated norm has lost its accuracy are
        // flagged for the recomputation over rows > i of the updated column
        if (wv == 0) {
            int rc = 0;
            if (active && vn1 != 0.0) {
                const double qv = fabs((double)s_rowi[lane]) / vn1;
                double temp = 1.0 - qv * qv;
                temp = fmax(temp, 0.0);
                const double rr = vn1 / vn2;
                const double temp2 = temp * rr * rr;
                if (temp2 <= tol3z) {
                    if (i < rows - 1) {
                        rc = 1;
                    } else {
                        vn1 = 0.0;
                        vn2 = 0.0;
                    }
                } else {
                    vn1 = vn1 * sqrt(temp);
                }
            }
            s_rc[lane] = rc;
            const int any = __any(rc);
            if (lane == 0) s_anyrc = any;
        }
        __syncthreads();
        if (s_anyrc) {  // uniform: the flagged columns updated now, their sums of squares in chains (r - i - 1) % 16
            const bool rc = s_rc[lane] != 0;
            if (rc && pending) {
                update_column(tm);
                pending = false;
            }
            double c2[CPW];
#pragma unroll
            for (int k = 0; k < CPW; k++) c2[k] = 0.0;
            if (rc) {
#pragma unroll
                for (int q = 0; q < RW; q++) {
                    if (NCH * q + CPW <= iw || NCH * q >= rw) continue;
                    if (NCH * q <= iw) {  // the group holding row i: rows <= i masked out
                        T m[CPW];
                        lds_row_group<T, CPW>(s_m + NCH * q + r0, m);
#pragma unroll
                        for (int k = 0; k < CPW; k++) {
                            const double d = (double)(a[k][q] * m[k]);
                            c2[k] = fma(d, d, c2[k]);
                        }
                    } else {  // rows > i (rows past the last hold 0)
#pragma unroll
                        for (int k = 0; k < CPW; k++) {
                            const double d = (double)a[k][q];
                            c2[k] = fma(d, d, c2[k]);
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < CPW; k++) s_part[(r0 + k - i % NCH - 1 + 2 * NCH) % NCH][lane] = c2[k];
            __syncthreads();
            if (wv == 0 && rc) {
                vn1 = sqrt(tree16_lds(&s_part[0][lane], SLAB));
                vn2 = vn1;
            }
        }
        if (wv == 0) {  // the slab's best column: the candidate for step i + 1
            double best = -1.0;
            int bpos = INT32_MAX;
            if (active) {
                best = vn1;
                bpos = pos;
            }
            wave_best(best, bpos);
            if (lane == 0) {
                relay_put_cand(relay_cand<T>(B, nxt, sl.slab), tag + 1u, best, bpos);
                s_best = bpos != INT32_MAX ? p2c[bpos] : -1;
            }
        }
        __syncthreads();
        // the candidate column, updated, goes through LDS (s_col is free until the next
        // step's read), then every thread stores one or two of its rows > i
        const int best = s_best;
        if (best >= 0 && lane == best % SLAB) {
            if (pending) {
                update_column(tm);
                pending = false;
            }
#pragma unroll
            for (int q = 0; q < RW; q++) {
                if (NCH * q + CPW <= iw + 1 || NCH * q >= rw) continue;
                T x[CPW];
#pragma unroll
                for (int k = 0; k < CPW; k++) x[k] = a[k][q];
                lds_store_group<T, CPW>(s_col + NCH * q + r0, x);
            }
        }
        __syncthreads();
        if (best >= 0) {
            GRelayWord* dst = relay_col<T>(B, nxt, sl.slab);
            for (int r = i + 1 + tid; r < rows; r += NTH) relay_put_val<T>(dst + (size_t)r * (sizeof(T) / 4), tag + 1u, s_col[r]);
        }
        if (pending) update_column(tm);  // the other columns' update, after the publication
        // the pivot column is done (never active again): its owner puts the reflector in its
        // registers, off the path the other slabs wait on (s_v, s_m and sh hold until the
        // next step's dlarfg)
        if (own_pivot && lane == ca % SLAB && taud != 0.0) {  // beta on row i, v below
#pragma unroll
            for (int q = 0; q < RW; q++) {
                if (NCH * q + CPW <= iw || NCH * q >= rw) continue;
                T v[CPW], m[CPW];
                lds_row_group<T, CPW>(s_v + NCH * q + r0, v);
                lds_row_group<T, CPW>(s_m + NCH * q + r0, m);
#pragma unroll
                for (int k = 0; k < CPW; k++)
                    if (m[k] != T(0)) a[k][q] = v[k];  // rows i + 1 .. rows - 1
            }
            if (iw >= 0 && iw / NCH < RW && iw % NCH < CPW) {
                const int kq = iw % NCH, qq = iw / NCH;
#pragma unroll
                for (int q = 0; q < RW; q++)
#pragma unroll
                    for (int k = 0; k < CPW; k++)
                        if (k == kq && q == qq) a[k][q] = (T)sh[1];
            }
        }
    }
    if (abandoned) {
        // the host reports the failed wait. The cluster's maps and slab storage are not the
        // QR's (this slab's registers are dropped; its cluster's other slabs abandon too, or
        // finish without it): rows = 0 makes k_rrqr_finish leave the cluster alone
        if (tid == 0) gp(B.dims)[0] = 0;
        return;
    }
    // registers back to the slab storage; final maps after `stop` steps, as k_rrqr_finish reads them
    if (jin) {
        int rw = rows - r0;
        asm volatile("" : "+s"(rw));
        const gptr<T> wb = gp(reinterpret_cast<T*>(B.t.scratch)) + (size_t)sl.slab * rows * SLAB + lane;
#pragma unroll
        for (int q = 0; q < RW; q++)
#pragma unroll
            for (int k = 0; k < CPW; k++)
                if (NCH * q + k < rw) wb[(size_t)(NCH * q + r0 + k) * SLAB] = a[k][q];
    }
    if (sl.slab == 0 && rows > 0) {
        const int fin = stop & 1;
        const gptr<int> map = gp(B.map);
        for (int k = tid; k < cols; k += NTH) {
            map[fin * mc + k] = p2c[k];
            map[2 * mc + fin * mc + p2c[k]] = k;
        }
        if (tid == 0) {
            gp(B.dims)[3] = stop;
            gp(B.dims)[4] = 1;  // the reflectors are in the pivot columns (k_rrqr_finish copies them)
        }
    }
}

// One workgroup per slab: every slab copies the reflectors held in its own columns into Q
// (persistent path: a column at pivot position pos < rank holds reflector pos) and scatters
// R[:rank, j] of its columns into the shrunk edge blocks.
template <typename T>
__global__ void __launch_bounds__(NT) k_rrqr_finish(const RrqrBig* bigs, const RrqrSlab* slabs, const RrqrSeg* segs,
                                                    int* dsize, double tol) {
    __shared__ int s_col0[RRQR_MAXSEG + 1];
    const RrqrSlab sl = slabs[blockIdx.x];
    const RrqrBig B = bigs[sl.big];
    const RrqrSeg* sg = segs + B.t.seg0;
    const int tid = threadIdx.x;
    const int rows = B.dims[0], cols = B.dims[1], mn = B.dims[2], stop = B.dims[3];
    if (rows == 0) return;
    if (tid == 0) {
        int c = 0;
        for (int k = 0; k < B.t.nseg; k++) {
            s_col0[k] = c;
            c += dsize[sg[k].nbr];
        }
        s_col0[B.t.nseg] = c;
    }
    const int mc = B.t.max_cols;
    const int fin = stop & 1;  // maps after `stop` steps
    const int* c2p = B.map + 2 * mc + fin * mc;  // column -> pivot position
    const gptr<T> A = gp(reinterpret_cast<T*>(B.t.scratch));
    __syncthreads();
    const int fixed = rank_fixed(tol, mn);
    const int rank = fixed >= 0 ? fixed : stop;
    if (rank >= rows) return;
    const gptr<T> Q = gp(reinterpret_cast<T*>(B.t.Q));
    // the wave kernel leaves the reflectors in the pivot columns of A; the per-step path has
    // stored them in Q already, and the pivot columns' R entries come from there
    const bool in_a = B.dims[4] == 1;
    const int j0 = sl.slab * SLAB;
    if (in_a)  // column j of this slab at pivot position pos < rank holds reflector pos
        for (int idx = tid; idx < rows * SLAB; idx += NT) {
            const int r = idx % rows, j = j0 + idx / rows;
            if (j >= cols) break;  // j grows with idx
            const int pos = c2p[j];
            if (pos < rank) Q[r + (size_t)pos * rows] = A[slab_idx(rows, r, j)];
        }
    for (int idx = tid; idx < rank * SLAB; idx += NT) {
        const int j = j0 + idx % SLAB, r = idx / SLAB;
        if (j >= cols) continue;
        const int pos = c2p[j];
        const int k = seg_find(s_col0, B.t.nseg, j);
        const int jj = j - s_col0[k], w = s_col0[k + 1] - s_col0[k];
        T* E = reinterpret_cast<T*>(sg[k].E);
        const T v = (r > pos) ? T(0) : ((!in_a && pos < rank) ? Q[r + (size_t)pos * rows] : A[slab_idx(rows, r, j)]);
        if (sg[k].out) E[jj + (size_t)r * w] = v;
        else E[r + (size_t)jj * rank] = v;
    }
    if (sl.slab == 0 && tid == 0) dsize[B.t.cluster] = rank;
}

// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(NT) k_copy(const CopyTask* tasks) {
    const CopyTask t = tasks[blockIdx.x];
    const gptr<const T> S = gp(reinterpret_cast<const T*>(t.src));
    const gptr<T> D = gp(reinterpret_cast<T*>(t.dst));
    const long long n = (long long)t.rows * t.cols;
    for (long long idx = threadIdx.x; idx < n; idx += NT) {
        int i = (int)(idx % t.rows), j = (int)(idx / t.rows);
        D[i + (size_t)j * t.ldd] = S ? S[idx] : (i == j ? T(1) : T(0));
    }
}

// ---------------------------------------------------------------------------------------
// Launchers
static inline int grid_for(size_t n) {
    size_t g = (n + NT - 1) / NT;
    return (int)(g > 65536 ? 65536 : (g == 0 ? 1 : g));
}

template <typename T>
hipError_t launch_fill(T* p, size_t n, T v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill<T><<<grid_for(n), NT, 0, st>>>(p, n, v);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_scatter(const long long* off, const double* val, long long n, T* base, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_scatter<T><<<grid_for((size_t)n), NT, 0, st>>>(off, val, n, base);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_potrf(const PotrfTask* d_tasks, int ntasks, int* d_info, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    k_potrf<T><<<ntasks, NT, 0, st>>>(d_tasks, d_info);
    return hipGetLastError();
}

// FP32 updates on the matrix cores. The Schur update takes them unless SPAND_SCHUR_MFMA=0.
// The trsm's trailing updates only with SPAND_TRSM_MFMA=1: on the 64^3 workload its time is
// in the sequential diagonal-block solves and the block loads, and the matrix-core form
// measured slower (12.4 ms against 10.7 ms of trsm per FP32 factorization, kernel trace).
static bool env_flag(const char* name, bool dflt) {
    const char* e = std::getenv(name);
    if (!e || !e[0]) return dflt;
    return e[0] != '0';
}
static bool trsm_mfma_enabled() { return env_flag("SPAND_TRSM_MFMA", false); }
static bool schur_mfma_enabled() { return env_flag("SPAND_SCHUR_MFMA", true); }

template <typename T>
hipError_t launch_trsm(const TrsmTask* d_tasks, int ntasks, int mfma_from, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    int nv = ntasks;  // tasks for the vector kernel
    if constexpr (sizeof(T) == sizeof(float)) {
        if (trsm_mfma_enabled()) nv = std::max(0, std::min(mfma_from, ntasks));
    }
    if (nv > 0) {
        k_trsm<T><<<nv, NT, 0, st>>>(d_tasks);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if constexpr (sizeof(T) == sizeof(float)) {
        if (nv < ntasks) {
            k_trsm_f32_mfma<<<ntasks - nv, NT, 0, st>>>(d_tasks + nv);
            return hipGetLastError();
        }
    }
    return hipSuccess;
}

template <typename T>
hipError_t launch_getrf(const GetrfTask* d_tasks, int ntasks, int in_lds, int nmax, int* d_info, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    if (in_lds) {
        const size_t bytes = (size_t)nmax * (nmax + 1) * sizeof(T);
        if (bytes > (size_t)GETRF_LDS_BYTES) return hipErrorInvalidValue;
        k_getrf<T, NT><<<ntasks, NT, bytes, st>>>(d_tasks, 1, d_info);
    } else {
        k_getrf<T, 1024><<<ntasks, 1024, 0, st>>>(d_tasks, 0, d_info);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_schur(const SchurTile* d_tiles, int ntiles, const SchurContrib* d_contribs, hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    if constexpr (sizeof(T) == sizeof(float)) {
        if (schur_mfma_enabled()) {
            k_schur_f32_mfma<<<ntiles, NT, 0, st>>>(d_tiles, d_contribs);
            return hipGetLastError();
        }
    }
    k_schur<T><<<ntiles, NT, 0, st>>>(d_tiles, d_contribs);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_potrf64(const PotrfTask* d_tasks, int ntasks, int* d_info, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    k_potrf64<T><<<ntasks, 64, 0, st>>>(d_tasks, d_info);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_potrf_panel(const PotrfTask* d_tasks, int ntasks, int kb, int nmax, int* d_info, hipStream_t st) {
    if (ntasks == 0 || kb >= nmax) return hipSuccess;
    const int rows = nmax - kb - PB;  // panel rows of the largest pivot, NT per workgroup
    const int nblk = rows > 0 ? (rows + NT - 1) / NT : 1;
    k_potrf_panel<T><<<dim3(nblk, ntasks), NT, 0, st>>>(d_tasks, kb, d_info);
    return hipGetLastError();
}

// Always at least one workgroup per task: workgroup 0 stores the panel's diagonal block.
template <typename T>
hipError_t launch_potrf_trail(const PotrfTask* d_tasks, int ntasks, int kb, int nmax, hipStream_t st) {
    if (ntasks == 0 || kb >= nmax) return hipSuccess;
    const int nt = std::max((nmax - kb - PB + PB - 1) / PB, 0);
    k_potrf_trail<T><<<dim3(std::max(nt * (nt + 1) / 2, 1), ntasks), NT, 0, st>>>(d_tasks, kb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rrqr(const RrqrTask* d_tasks, int ntasks, const RrqrSeg* d_segs, int* d_size, double tol,
                       int lds_bytes, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    if (lds_bytes <= 0 || lds_bytes > RRQR_LDS_MAX) return hipErrorInvalidValue;
    static bool opted = false;  // dynamic LDS above 64 KiB needs the opt-in
    if (!opted) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_rrqr<T>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, RRQR_LDS_MAX);
        if (e != hipSuccess) return e;
        opted = true;
    }
    k_rrqr<T><<<ntasks, NT, lds_bytes, st>>>(d_tasks, d_segs, d_size, tol, lds_bytes / (int)sizeof(T));
    return hipGetLastError();
}

// Register path variants: (rows held per thread RR, columns per thread K, threads NTR). The
// table keeps to the forms that compile without register spills on gfx950 (1024 threads
// leave 128 VGPRs per thread, 512 threads 256): FP32 rows <= 32 up to 1024 columns, rows
// <= 64 up to 1024 (2 columns per thread above 512), rows <= 96 up to 512; FP64 rows <= 32
// up to 1024 (2 per thread above 512), rows <= 48 up to 512. Unused rows cost registers
// only: the sweeps skip 16-row blocks beyond a cluster's rows.
int rrqr_reg_variant(size_t esz, int rows, int cols) {
    if (rows < 0 || cols < 0) return -1;
    const int cc = cols <= 256 ? 0 : cols <= 512 ? 1 : cols <= 1024 ? 2 : 3;
    if (cc == 3) return -1;
    int rc;
    if (esz == sizeof(double)) {
        rc = rows <= 32 ? 0 : rows <= 48 ? 1 : -1;
        if (rc == 1 && cc == 2) return -1;
    } else {
        rc = rows <= 32 ? 0 : rows <= 64 ? 1 : rows <= 96 ? 2 : -1;
        if (rc == 2 && cc == 2) return -1;
    }
    return rc < 0 ? -1 : 3 * rc + cc;
}

template <typename T, int RR, int K, int NTR>
static hipError_t launch_reg_variant(const RrqrTask* d_tasks, int ntasks, const RrqrSeg* d_segs, int* d_size,
                                     double tol, hipStream_t st) {
    k_rrqr_reg<T, RR, K, NTR><<<ntasks, NTR, 0, st>>>(d_tasks, d_segs, d_size, tol);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rrqr_reg(const RrqrTask* d_tasks, int ntasks, const RrqrSeg* d_segs, int* d_size, double tol,
                           int variant, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
#define SPAND_REG(RR, K, NTR) launch_reg_variant<T, RR, K, NTR>(d_tasks, ntasks, d_segs, d_size, tol, st)
    if constexpr (sizeof(T) == sizeof(double)) {
        switch (variant) {
            case 0: return SPAND_REG(32, 1, 256);
            case 1: return SPAND_REG(32, 1, 512);
            case 2: return SPAND_REG(32, 2, 512);
            case 3: return SPAND_REG(48, 1, 256);
            case 4: return SPAND_REG(48, 1, 512);
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (variant) {
            case 0: return SPAND_REG(32, 1, 256);
            case 1: return SPAND_REG(32, 1, 512);
            case 2: return SPAND_REG(32, 1, 1024);
            case 3: return SPAND_REG(64, 1, 256);
            case 4: return SPAND_REG(64, 1, 512);
            case 5: return SPAND_REG(64, 2, 512);
            case 6: return SPAND_REG(96, 1, 256);
            case 7: return SPAND_REG(96, 1, 512);
            default: return hipErrorInvalidValue;
        }
    }
#undef SPAND_REG
}

template <typename T>
hipError_t launch_rrqr_init(const RrqrBig* d_bigs, const RrqrSlab* d_slabs, int nslabs, const RrqrSeg* d_segs,
                            const int* d_size, double tol, hipStream_t st) {
    if (nslabs == 0) return hipSuccess;
    k_rrqr_init<T><<<nslabs, STEP_NT, 0, st>>>(d_bigs, d_slabs, d_segs, d_size, tol);
    return hipGetLastError();
}

struct StepArgs {
    const RrqrBig* bigs;
    const RrqrSlab* slabs;
    int nslabs, step;
    double tol;
    size_t lds;
    long long* prog;
    long long seq;
    hipStream_t st;
};

template <typename T, int RW, int NW>
static void step_rw(const StepArgs& a) {
    k_rrqr_step<T, RW, NW><<<a.nslabs, SLAB * NW, a.lds, a.st>>>(a.bigs, a.slabs, a.step, a.tol, a.prog, a.seq);
}

template <typename T, int NW>
static void step_nw(const StepArgs& a, int max_rows) {
    if (max_rows <= NCH * 4) step_rw<T, 4, NW>(a);
    else if (max_rows <= NCH * 8) step_rw<T, 8, NW>(a);
    else if (max_rows <= NCH * 16) step_rw<T, 16, NW>(a);
    else step_rw<T, 32, (NW < 8 ? 8 : NW)>(a);
}

// Wavefronts per slab workgroup: 16 while the wave's slabs fit the chip in one round at
// 16 wavefronts each, fewer for waves with many slabs (each thread then holds 16 / NW
// chains, at most 64 cached rows). waves = 0 picks by the slab count.
int rrqr_step_waves(int nslabs, int max_rows, int waves) {
    if (max_rows > NCH * 32) return NCH;  // uncached rows: one chain per wavefront
    const int rw = max_rows <= NCH * 4 ? 4 : max_rows <= NCH * 8 ? 8 : max_rows <= NCH * 16 ? 16 : 32;
    int nw = waves;
    if (nw != 4 && nw != 8 && nw != 16) nw = nslabs <= RRQR_STEP_SLABS16 ? 16 : nslabs <= RRQR_STEP_SLABS8 ? 8 : 4;
    while (nw < NCH && (NCH / nw) * rw > 64) nw *= 2;
    return nw;
}

template <typename T>
hipError_t launch_rrqr_step(const RrqrBig* d_bigs, const RrqrSlab* d_slabs, int nslabs, int step, int max_rows,
                            double tol, size_t lds_bytes, int waves, long long* prog, long long seq, hipStream_t st) {
    if (nslabs == 0) return hipSuccess;
    // the variants that cache rows (max_rows <= 16 * 32) keep the pivot column and its
    // reflector in the dynamic LDS: it must hold both for the wave's largest cluster
    if (max_rows <= NCH * 32 && lds_bytes < 2 * (size_t)std::max(max_rows, 1) * sizeof(T)) return hipErrorInvalidValue;
    const StepArgs a{d_bigs, d_slabs, nslabs, step, tol, lds_bytes, prog, seq, st};
    const int nw = rrqr_step_waves(nslabs, max_rows, waves);
    if (max_rows > NCH * 32) step_rw<T, 0, NCH>(a);
    else if (nw == 4) step_nw<T, 4>(a, max_rows);
    else if (nw == 8) step_nw<T, 8>(a, max_rows);
    else step_nw<T, 16>(a, max_rows);
    return hipGetLastError();
}

// Persistent wave launch: rows per thread and wavefronts per slab by the wave's largest
// cluster (16 / NW * RW values of T per thread, at most 128 registers' worth).
template <typename T, int RW, int NW>
static hipError_t wave_capacity_rw(size_t lds_bytes, int* cap) {
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const void* f = reinterpret_cast<const void*>(k_rrqr_wave<T, RW, NW>);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, SLAB * NW, lds_bytes);
    *cap = e == hipSuccess ? per_cu * ncu : 0;
    return e;
}

template <typename T, int RW, int NW>
static hipError_t launch_wave_rw(const RrqrBig* d_bigs, const RrqrSlab* d_slabs, int nslabs, int nsteps, double tol,
                                 size_t lds_bytes, int* fail, long long spin_max, hipStream_t st) {
    int cap = 0;
    hipError_t e = wave_capacity_rw<T, RW, NW>(lds_bytes, &cap);
    if (e != hipSuccess) return e;
    if (cap < nslabs) return hipErrorCooperativeLaunchTooLarge;
    const void* f = reinterpret_cast<const void*>(k_rrqr_wave<T, RW, NW>);
    void* args[] = {(void*)&d_bigs, (void*)&d_slabs, (void*)&nsteps, (void*)&tol, (void*)&fail, (void*)&spin_max};
    return hipLaunchCooperativeKernel(f, dim3(nslabs), dim3(SLAB * NW), args, (unsigned)lds_bytes, st);
}

// One dispatch table for the capacity query and the launch: the variant for max_rows.
template <typename T, bool LAUNCH>
static hipError_t wave_variant(int max_rows, size_t lds_bytes, int* cap, const RrqrBig* d_bigs,
                               const RrqrSlab* d_slabs, int nslabs, int nsteps, double tol, int* fail,
                               long long spin_max, hipStream_t st) {
#define SPAND_WAVE(RW, NW)                                                                                             \
    return LAUNCH ? launch_wave_rw<T, RW, NW>(d_bigs, d_slabs, nslabs, nsteps, tol, lds_bytes, fail, spin_max, st) \
                  : wave_capacity_rw<T, RW, NW>(lds_bytes, cap)
    if (max_rows <= NCH * 4) SPAND_WAVE(4, 4);
    if (max_rows <= NCH * 8) SPAND_WAVE(8, 4);
    if (max_rows <= NCH * 16) SPAND_WAVE(16, 4);
    if (max_rows <= NCH * 32) SPAND_WAVE(32, 8);
#undef SPAND_WAVE
    return LAUNCH ? hipErrorCooperativeLaunchTooLarge : hipSuccess;
}

template <typename T>
hipError_t rrqr_wave_capacity(int max_rows, size_t lds_bytes, int* cap) {
    *cap = 0;
    return wave_variant<T, false>(max_rows, lds_bytes, cap, nullptr, nullptr, 0, 0, 0.0, nullptr, 0, nullptr);
}

// Persistent launch over nslabs slabs (whole clusters): returns hipErrorCooperativeLaunchTooLarge
// (and launches nothing) when they cannot all be resident at once.
template <typename T>
hipError_t launch_rrqr_wave(const RrqrBig* d_bigs, const RrqrSlab* d_slabs, int nslabs, int nsteps, int max_rows,
                            double tol, size_t lds_bytes, int* fail, long long spin_max, hipStream_t st) {
    if (nslabs == 0 || nsteps == 0) return hipSuccess;
    if (spin_max < 1) return hipErrorInvalidValue;
    return wave_variant<T, true>(max_rows, lds_bytes, nullptr, d_bigs, d_slabs, nslabs, nsteps, tol, fail, spin_max,
                                 st);
}

template <typename T>
hipError_t launch_rrqr_finish(const RrqrBig* d_bigs, const RrqrSlab* d_slabs, int nslabs, const RrqrSeg* d_segs,
                              int* d_size, double tol, size_t lds_bytes, hipStream_t st) {
    if (nslabs == 0) return hipSuccess;
    k_rrqr_finish<T><<<nslabs, NT, lds_bytes, st>>>(d_bigs, d_slabs, d_segs, d_size, tol);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_copy(const CopyTask* d_tasks, int ntasks, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    k_copy<T><<<ntasks, NT, 0, st>>>(d_tasks);
    return hipGetLastError();
}

#define SPAND_INST(T)                                                                              \
    template hipError_t launch_fill<T>(T*, size_t, T, hipStream_t);                               \
    template hipError_t launch_scatter<T>(const long long*, const double*, long long, T*, hipStream_t); \
    template hipError_t launch_potrf<T>(const PotrfTask*, int, int*, hipStream_t);                \
    template hipError_t launch_potrf64<T>(const PotrfTask*, int, int*, hipStream_t);              \
    template hipError_t launch_potrf_panel<T>(const PotrfTask*, int, int, int, int*, hipStream_t);     \
    template hipError_t launch_potrf_trail<T>(const PotrfTask*, int, int, int, hipStream_t);      \
    template hipError_t launch_trsm<T>(const TrsmTask*, int, int, hipStream_t);                   \
    template hipError_t launch_getrf<T>(const GetrfTask*, int, int, int, int*, hipStream_t);      \
    template hipError_t launch_schur<T>(const SchurTile*, int, const SchurContrib*, hipStream_t); \
    template hipError_t launch_rrqr<T>(const RrqrTask*, int, const RrqrSeg*, int*, double, int, hipStream_t); \
    template hipError_t launch_rrqr_reg<T>(const RrqrTask*, int, const RrqrSeg*, int*, double, int, hipStream_t); \
    template hipError_t launch_copy<T>(const CopyTask*, int, hipStream_t);                        \
    template hipError_t launch_rrqr_init<T>(const RrqrBig*, const RrqrSlab*, int, const RrqrSeg*, const int*,      \
                                            double, hipStream_t);                                                  \
    template hipError_t launch_rrqr_wave<T>(const RrqrBig*, const RrqrSlab*, int, int, int, double, size_t, int*, \
                                            long long, hipStream_t);                                               \
    template hipError_t rrqr_wave_capacity<T>(int, size_t, int*);                                                  \
    template hipError_t launch_rrqr_step<T>(const RrqrBig*, const RrqrSlab*, int, int, int, double, size_t, int,  \
                                            long long*, long long, hipStream_t);                                                       \
    template hipError_t launch_rrqr_finish<T>(const RrqrBig*, const RrqrSlab*, int, const RrqrSeg*, int*, double,  \
                                              size_t, hipStream_t);
SPAND_INST(float)
SPAND_INST(double)

}  // namespace spand
