// Preconditioner-apply sweeps and FP64 Krylov vector kernels for gfx950.
//
// Tree::solve (src/tree.cpp:1610-1635) replays the recorded operations forward then
// backward on per-cluster segments of one vector; here every recorded op of one ND level
// and one kind runs as a single launch (one workgroup per cluster / target segment) on a
// block of nrhs right-hand sides stored interleaved, x[(xoff + i) * nrhs + r].
//   trsv   -> ScalingLLT fwd/bwd          src/operations.cpp:95-103 (cblas_dtrsv)
//   gemv   -> GemmSymmOut/In fwd/bwd       src/operations.cpp:55-93  (cblas_dgemv)
//   qapply -> Orthogonal fwd/bwd          src/operations.cpp:149-158 (LAPACKE_dormqr)
//   merge  -> Merge fwd/bwd               src/operations.cpp:166-192
// The Krylov kernels back cg (src/is.cpp:39-121): CSR SpMV and deterministic two-pass dots.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "tasks.hpp"

namespace spand {

constexpr int NT = 256;
// Segments are FP64 whatever the factor's element type T: an FP32 factor is read as FP32 and
// applied in FP64 arithmetic, so the preconditioner stays one fixed linear operator.
using VT = double;
// Dynamic LDS of the solve kernels: a chunk of up to 8 right-hand sides of the segment (and
// for the Householder apply two k-vectors per right-hand side); the chunk width is derived
// from the segment size, and the host rejects segments that would not fit one column.

static inline int grid_for(size_t n) {
    size_t g = (n + NT - 1) / NT;
    return (int)(g > 65536 ? 65536 : (g == 0 ? 1 : g));
}

// ---------------------------------------------------------------------------------------
// Element (i, r) of the caller's vector block lives at b[i * sx + r * sr] (column-major:
// sx = 1, sr = ld; interleaved: sx = nrhs, sr = 1).
template <typename T>
__global__ void __launch_bounds__(NT) k_perm_in(const double* b, size_t sx, size_t sr, int nrhs, const int* perm,
                                                int N, VT* ws) {
    size_t n = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
        ws[idx] = b[(size_t)perm[i] * sx + (size_t)r * sr];
    }
}
template <typename T>
__global__ void __launch_bounds__(NT) k_perm_out(const VT* ws, const int* perm, int N, int nrhs, double* x, size_t sx,
                                                 size_t sr) {
    size_t n = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
        x[(size_t)perm[i] * sx + (size_t)r * sr] = ws[idx];
    }
}

// column-major (ld) <-> interleaved conversions of an N x nrhs block
__global__ void __launch_bounds__(NT) k_cm2il(const double* src, int ld, int N, int nrhs, double* dst) {
    size_t n = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
        dst[idx] = src[(size_t)i + (size_t)r * ld];
    }
}
__global__ void __launch_bounds__(NT) k_il2cm(const double* src, int N, int nrhs, double* dst, int ld) {
    size_t n = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
        dst[(size_t)i + (size_t)r * ld] = src[idx];
    }
}

// ---------------------------------------------------------------------------------------
// Block-of-right-hand-sides helpers. A segment's nrhs columns are processed in chunks of
// RG <= 8 right-hand sides, staged in LDS as S[i * RG + r].
constexpr int RG_MAX = 8;
constexpr int WAVE = 64;

template <typename T>
__device__ inline T wave_sum(T v) {
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Solve sweeps are split into chunks of SC = 64 outputs of one task (SolveChunk), so a large
// cluster spreads over many workgroups. A workgroup is 64 lanes (one output each) x 4
// partitions of the inner sum; the partitions are combined in LDS at the end.
constexpr int SC = SOLVE_CHUNK;
constexpr int NP = NT / SC;  // 4
constexpr int TR = 32;       // rows of a transposed tile

// y = Linv x[perm] (trans 0) or Lb^T x (trans 1) for chunk [o0, o0 + 64) of a scaling
// segment, written to wt (the copy-back kernel returns it to ws: other chunks of the same
// task still read the old x). Linv and Lb are zero above their diagonals, so the forward
// sum runs over all j < min(n, o0 + 64) without a mask.
template <typename T>
__global__ void __launch_bounds__(NT) k_trmv(const TrsvTask* tasks, const SolveChunk* chunks, const VT* ws, VT* wt,
                                             int nrhs, int trans) {
    __shared__ T tile[TR][SC + 1];
    __shared__ VT red[NP][SC][RG_MAX];
    const SolveChunk ch = chunks[blockIdx.x];
    const TrsvTask t = tasks[ch.task];
    const int n = t.n, o0 = ch.o0, no = min(SC, n - o0);
    const int tid = threadIdx.x, lane = tid % SC, kp = tid / SC;
    const T* L = reinterpret_cast<const T*>(trans ? t.Lb : t.L);
    const int* P = t.perm;
    const VT* x = ws + (size_t)t.xoff * nrhs;
    for (int r0 = 0; r0 < nrhs; r0 += RG_MAX) {
        const int rg = min(RG_MAX, nrhs - r0);
        VT acc[RG_MAX];
#pragma unroll
        for (int r = 0; r < RG_MAX; r++) acc[r] = VT(0);
        if (!trans) {
            const int i = o0 + lane;
            const int jmax = min(n, o0 + SC);
            if (lane < no) {
                for (int j = kp; j < jmax; j += NP) {
                    const VT l = L[i + (size_t)j * n];
                    const VT* xj = x + (size_t)(P ? P[j] : j) * nrhs + r0;
                    for (int r = 0; r < rg; r++) acc[r] += l * xj[r];
                }
            }
        } else {
            for (int i0 = o0; i0 < n; i0 += TR) {
                const int ni = min(TR, n - i0);
                for (int idx = tid; idx < TR * SC; idx += NT) {
                    const int ii = idx % TR, jj = idx / TR;
                    tile[ii][jj] = (ii < ni && jj < no) ? L[(i0 + ii) + (size_t)(o0 + jj) * n] : T(0);
                }
                __syncthreads();
                for (int ii = kp; ii < ni; ii += NP) {
                    const VT l = tile[ii][lane];
                    const VT* xi = x + (size_t)(i0 + ii) * nrhs + r0;
                    for (int r = 0; r < rg; r++) acc[r] += l * xi[r];
                }
                __syncthreads();
            }
        }
        for (int r = 0; r < rg; r++) red[kp][lane][r] = acc[r];
        __syncthreads();
        if (kp == 0 && lane < no) {
            VT* y = wt + (size_t)(t.xoff + o0 + lane) * nrhs + r0;
            for (int r = 0; r < rg; r++) y[r] = red[0][lane][r] + red[1][lane][r] + red[2][lane][r] + red[3][lane][r];
        }
        __syncthreads();
    }
}

// ws[seg] = wt[seg] for the segments of a trmv launch, one workgroup per chunk
template <typename T>
__global__ void __launch_bounds__(NT) k_copyback(const TrsvTask* tasks, const SolveChunk* chunks, VT* ws, const VT* wt,
                                                 int nrhs) {
    const SolveChunk ch = chunks[blockIdx.x];
    const TrsvTask t = tasks[ch.task];
    const int no = min(SC, t.n - ch.o0);
    const size_t base = (size_t)(t.xoff + ch.o0) * nrhs;
    for (int idx = threadIdx.x; idx < no * nrhs; idx += NT) ws[base + idx] = wt[base + idx];
}

// x[dst] -= sum_k op(A_k) x[src_k] for chunk [o0, o0 + 64) of the destination segment.
// Destinations and sources of one launch are disjoint segments, so the update is in place.
template <typename T>
__global__ void __launch_bounds__(NT) k_gemv(const GemvGroup* groups, const SolveChunk* chunks, const GemvTerm* terms,
                                             VT* ws, int nrhs) {
    __shared__ T tile[TR][SC + 1];
    __shared__ VT red[NP][SC][RG_MAX];
    const SolveChunk ch = chunks[blockIdx.x];
    const GemvGroup g = groups[ch.task];
    const int o0 = ch.o0, no = min(SC, g.n - o0);
    const int tid = threadIdx.x, lane = tid % SC, kp = tid / SC;
    for (int r0 = 0; r0 < nrhs; r0 += RG_MAX) {
        const int rg = min(RG_MAX, nrhs - r0);
        VT acc[RG_MAX];
#pragma unroll
        for (int r = 0; r < RG_MAX; r++) acc[r] = VT(0);
        for (int k = g.k0; k < g.k1; k++) {
            const GemvTerm tm = terms[k];
            const T* A = reinterpret_cast<const T*>(tm.A);
            const VT* xs = ws + (size_t)tm.xsrc * nrhs + r0;
            if (!tm.trans) {
                // output i = row of A (rows = destination length), sum over columns j
                if (lane < no) {
                    const int i = o0 + lane;
                    for (int j = kp; j < tm.cols; j += NP) {
                        const VT a = A[i + (size_t)j * tm.rows];
                        const VT* xj = xs + (size_t)j * nrhs;
                        for (int r = 0; r < rg; r++) acc[r] += a * xj[r];
                    }
                }
            } else {
                // output j = column of A (cols = destination length), sum over rows i
                for (int i0 = 0; i0 < tm.rows; i0 += TR) {
                    const int ni = min(TR, tm.rows - i0);
                    for (int idx = tid; idx < TR * SC; idx += NT) {
                        const int ii = idx % TR, jj = idx / TR;
                        tile[ii][jj] = (ii < ni && jj < no) ? A[(i0 + ii) + (size_t)(o0 + jj) * tm.rows] : T(0);
                    }
                    __syncthreads();
                    for (int ii = kp; ii < ni; ii += NP) {
                        const VT a = tile[ii][lane];
                        const VT* xi = xs + (size_t)(i0 + ii) * nrhs;
                        for (int r = 0; r < rg; r++) acc[r] += a * xi[r];
                    }
                    __syncthreads();
                }
            }
        }
        for (int r = 0; r < rg; r++) red[kp][lane][r] = acc[r];
        __syncthreads();
        if (kp == 0 && lane < no) {
            VT* y = ws + (size_t)(g.xdst + o0 + lane) * nrhs + r0;
            for (int r = 0; r < rg; r++) y[r] -= red[0][lane][r] + red[1][lane][r] + red[2][lane][r] + red[3][lane][r];
        }
        __syncthreads();
    }
}

// Householder reflectors in blocked compact-WY form (the dormqr scheme): Q = Q_0 Q_1 ...
// with Q_b = I - V_b T_b V_b^T over reflectors [32b, 32b + 32). Q^T x applies Q_0^T first,
// Q x applies the last block first. V has a unit diagonal and zeros above it implicitly.
// One workgroup per op (the blocks are sequential; the segment is updated in place);
// W = V_b^T x is accumulated over 64-row tiles of V_b staged in LDS (32 reflectors x 8
// row partitions).
template <typename T>
__global__ void __launch_bounds__(NT) k_qapply(const QTask* tasks, VT* ws, int nrhs, int trans) {
    __shared__ T vt[SC][QB + 1];
    __shared__ VT wred[NT / QB][QB][RG_MAX];
    __shared__ VT W[QB][RG_MAX], U[QB][RG_MAX];
    const QTask t = tasks[blockIdx.x];
    const int m = t.m, k = t.k, tid = threadIdx.x;
    if (m == 0 || k == 0) return;
    const T* V = reinterpret_cast<const T*>(t.V);
    VT* x = ws + (size_t)t.xoff * nrhs;
    const int nblk = (k + QB - 1) / QB;
    const int c = tid % QB, part = tid / QB, nparts = NT / QB;
    for (int r0 = 0; r0 < nrhs; r0 += RG_MAX) {
        const int rg = min(RG_MAX, nrhs - r0);
        for (int bi = 0; bi < nblk; bi++) {
            const int b = trans ? bi : nblk - 1 - bi;
            const int b0 = b * QB, nb = min(QB, k - b0);
            const T* Tb = reinterpret_cast<const T*>(t.Tm) + (size_t)b * QB * QB;
            // W = V_b^T x with v_c = [1 at row b0 + c, V below]
            VT acc[RG_MAX];
#pragma unroll
            for (int r = 0; r < RG_MAX; r++) acc[r] = VT(0);
            for (int i0 = b0; i0 < m; i0 += SC) {
                const int ni = min(SC, m - i0);
                for (int idx = tid; idx < SC * QB; idx += NT) {
                    const int ii = idx % SC, cc = idx / SC, i = i0 + ii, col = b0 + cc;
                    T v = T(0);
                    if (ii < ni && cc < nb) v = (i == col) ? T(1) : (i > col ? V[i + (size_t)col * m] : T(0));
                    vt[ii][cc] = v;
                }
                __syncthreads();
                for (int ii = part; ii < ni; ii += nparts) {
                    const VT v = vt[ii][c];
                    const VT* xi = x + (size_t)(i0 + ii) * nrhs + r0;
                    for (int r = 0; r < rg; r++) acc[r] += v * xi[r];
                }
                __syncthreads();
            }
            for (int r = 0; r < rg; r++) wred[part][c][r] = acc[r];
            __syncthreads();
            if (tid < nb * rg) {
                const int cc = tid / rg, r = tid % rg;
                VT s = VT(0);
                for (int p = 0; p < nparts; p++) s += wred[p][cc][r];
                W[cc][r] = s;
            }
            __syncthreads();
            // U = T_b^T W (Q^T) or T_b W (Q); T_b is nb x nb upper triangular, ld QB
            if (tid < nb * rg) {
                const int a = tid / rg, r = tid % rg;
                VT s = VT(0);
                if (trans) {
                    for (int q = 0; q <= a; q++) s += (VT)Tb[q + (size_t)a * QB] * W[q][r];
                } else {
                    for (int q = a; q < nb; q++) s += (VT)Tb[a + (size_t)q * QB] * W[q][r];
                }
                U[a][r] = s;
            }
            __syncthreads();
            // x -= V_b U for rows i >= b0: U_{i-b0} on the unit diagonal + sum_{c < i-b0} V(i, b0+c) U_c
            for (int idx = tid; idx < (m - b0) * rg; idx += NT) {
                const int i = b0 + idx / rg, r = idx % rg;
                const int li = i - b0;
                VT s = li < nb ? U[li][r] : VT(0);
                const int cmax = min(li, nb);
                for (int cc = 0; cc < cmax; cc++) s += (VT)V[i + (size_t)(b0 + cc) * m] * U[cc][r];
                x[(size_t)i * nrhs + r0 + r] -= s;
            }
            __syncthreads();
        }
    }
}

// dlarft (forward, columnwise) for one block of at most QB reflectors [b0, b0 + nb):
// G(a, c) = v_a^T v_c is accumulated over the block's rows in one pass, then
// T(c, c) = tau_c, T(0:c, c) = T(0:c, 0:c) (-tau_c G(c, 0:c)^T).
template <typename T>
__global__ void __launch_bounds__(NT) k_larft(const LarftTask* tasks) {
    __shared__ T vs[WAVE][QB + 1];
    __shared__ T G[QB][QB + 1];
    __shared__ T Tl[QB][QB + 1];
    __shared__ T tcol[QB];
    const LarftTask t = tasks[blockIdx.x];
    const int m = t.m, nb = t.k, b0 = t.b0, tid = threadIdx.x;
    const T* V = reinterpret_cast<const T*>(t.V);
    const T* tau = reinterpret_cast<const T*>(t.tau) + b0;
    // pairs (a > c) of the block, two per thread at most (QB * (QB - 1) / 2 <= 2 * NT)
    int pa[2], pc[2];
    T acc[2] = {T(0), T(0)};
    for (int q = 0; q < 2; q++) {
        int p = tid + q * NT, a = 1;
        while (p >= a) {
            p -= a;
            a++;
        }
        pa[q] = a;
        pc[q] = p;
    }
    for (int r0 = b0; r0 < m; r0 += WAVE) {
        for (int idx = tid; idx < WAVE * QB; idx += NT) {
            const int rr = idx % WAVE, c = idx / WAVE, r = r0 + rr;
            T v = T(0);
            if (c < nb && r < m) v = (r == b0 + c) ? T(1) : (r > b0 + c ? V[r + (size_t)(b0 + c) * m] : T(0));
            vs[rr][c] = v;
        }
        __syncthreads();
        for (int q = 0; q < 2; q++) {
            if (pa[q] < nb) {
                T s = acc[q];
                for (int rr = 0; rr < WAVE; rr++) s += vs[rr][pa[q]] * vs[rr][pc[q]];
                acc[q] = s;
            }
        }
        __syncthreads();
    }
    for (int q = 0; q < 2; q++)
        if (pa[q] < nb) G[pa[q]][pc[q]] = acc[q];
    __syncthreads();
    for (int c = 0; c < nb; c++) {
        const T tc = tau[c];
        if (tid < c) tcol[tid] = -tc * G[c][tid];
        __syncthreads();
        if (tid < c) {
            T s = T(0);
            for (int q = tid; q < c; q++) s += Tl[tid][q] * tcol[q];
            Tl[tid][c] = s;
        }
        if (tid == 0) Tl[c][c] = tc;
        if (tid > c && tid < QB) Tl[tid][c] = T(0);
        __syncthreads();
    }
    T* Tm = reinterpret_cast<T*>(t.Tm);
    for (int idx = tid; idx < QB * QB; idx += NT) {
        const int a = idx % QB, c = idx / QB;
        Tm[idx] = (a < nb && c < nb) ? Tl[a][c] : T(0);
    }
}

// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(NT) k_merge(const MergeTask* tasks, const int2* children, VT* ws, int nrhs, int fwd) {
    const MergeTask t = tasks[blockIdx.x];
    int k = 0;
    for (int c = t.c0; c < t.c1; c++) {
        const int2 ch = children[c];  // (xoff, size)
        const int n = ch.y * nrhs;
        VT* P = ws + (size_t)(t.pxoff + k) * nrhs;
        VT* Cx = ws + (size_t)ch.x * nrhs;
        if (fwd)
            for (int i = threadIdx.x; i < n; i += NT) P[i] = Cx[i];
        else
            for (int i = threadIdx.x; i < n; i += NT) Cx[i] = P[i];
        k += ch.y;
    }
}

// ---------------------------------------------------------------------------------------
// FP64 Krylov kernels
__global__ void __launch_bounds__(NT) k_spmv(int N, const int* rowptr, const int* colidx, const double* val,
                                             const double* x, double* y, int nrhs) {
    size_t n = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
        double acc = 0.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) acc += val[k] * x[(size_t)colidx[k] * nrhs + r];
        y[idx] = acc;
    }
}

constexpr int DOT_BLOCKS = 1024;
__global__ void __launch_bounds__(NT) k_dot_partial(int N, int nrhs, const double* a, const double* b,
                                                    double* partial) {
    __shared__ double red[NT];
    int nrp = 1;
    while (nrp < nrhs) nrp <<= 1;
    const int groups = NT / nrp;
    const int tid = threadIdx.x, r = tid % nrp, g = tid / nrp;
    double s = 0.0;
    if (r < nrhs) {
        for (size_t i = (size_t)blockIdx.x * groups + g; i < (size_t)N; i += (size_t)gridDim.x * groups)
            s += a[i * nrhs + r] * b[i * nrhs + r];
    }
    red[tid] = s;
    __syncthreads();
    for (int off = groups / 2; off > 0; off >>= 1) {
        if (g < off) red[tid] += red[tid + off * nrp];
        __syncthreads();
    }
    if (g == 0 && r < nrhs) partial[(size_t)blockIdx.x * nrhs + r] = red[tid];
}
__global__ void __launch_bounds__(NT) k_dot_final(int nblocks, int nrhs, const double* partial, double* out) {
    __shared__ double red[NT];
    for (int r = 0; r < nrhs; r++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += NT) s += partial[(size_t)b * nrhs + r];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int off = NT / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[r] = red[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(NT) k_axpby(int N, int nrhs, const double* x, const double* ax, double sgn,
                                              double* y, const double* by) {
    size_t n = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        int r = (int)(idx % nrhs);
        double a = sgn * (ax ? ax[r] : 1.0);
        double b = by ? by[r] : 1.0;
        y[idx] = a * x[idx] + b * y[idx];
    }
}

// GMRES kernels. V holds Householder vectors as interleaved N x nrhs blocks (block j at
// V + j * N * nrhs). out = base + [ye; 0] - sum_{j<k} V_j z_j, per column (base may be
// nullptr = 0; ye holds ny rows, or nullptr).
__global__ void __launch_bounds__(NT) k_gm_comb(int N, int nrhs, const double* base, const double* ye, int ny,
                                               const double* V, int k, const double* z, double* out) {
    const size_t n = (size_t)N * nrhs, blk = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        const int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
        double acc = base ? base[idx] : 0.0;
        if (ye && i < ny) acc += ye[idx];
        for (int j = 0; j < k; j++) acc -= V[(size_t)j * blk + idx] * z[(size_t)j * nrhs + r];
        out[idx] = acc;
    }
}
// partial[(b * k + j) * nrhs + r] = sum over block b of V_j[i, r] * v[i, r]; grid (blocks, k)
__global__ void __launch_bounds__(NT) k_gm_multidot_partial(int N, int nrhs, const double* V, const double* v,
                                                            double* partial) {
    __shared__ double red[NT];
    int nrp = 1;
    while (nrp < nrhs) nrp <<= 1;
    const int groups = NT / nrp;
    const int tid = threadIdx.x, r = tid % nrp, g = tid / nrp, j = blockIdx.y;
    const double* Vj = V + (size_t)j * N * nrhs;
    double s = 0.0;
    if (r < nrhs)
        for (size_t i = (size_t)blockIdx.x * groups + g; i < (size_t)N; i += (size_t)gridDim.x * groups)
            s += Vj[i * nrhs + r] * v[i * nrhs + r];
    red[tid] = s;
    __syncthreads();
    for (int off = groups / 2; off > 0; off >>= 1) {
        if (g < off) red[tid] += red[tid + off * nrp];
        __syncthreads();
    }
    if (g == 0 && r < nrhs) partial[((size_t)blockIdx.x * gridDim.y + j) * nrhs + r] = red[tid];
}
// out[j * nrhs + r] = sum over blocks; one workgroup per (j, r)
__global__ void __launch_bounds__(NT) k_gm_multidot_final(int nblocks, int k, int nrhs, const double* partial,
                                                          double* out) {
    __shared__ double red[NT];
    const int jr = blockIdx.x;  // j * nrhs + r
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += NT) s += partial[(size_t)b * k * nrhs + jr];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = NT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[jr] = red[0];
}
// Householder vector of step k from v: u[i] = 0 (i < k), 1 (i = k), v[i] / d[r] (i > k);
// d[r] = 0 means a zero essential part (tau = 0).
__global__ void __launch_bounds__(NT) k_gm_householder(int N, int nrhs, int k, const double* v, const double* d,
                                                       double* u) {
    const size_t n = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        const int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
        u[idx] = i < k ? 0.0 : (i == k ? 1.0 : (d[r] != 0.0 ? v[idx] / d[r] : 0.0));
    }
}

__global__ void k_ratio(int nrhs, const double* num, const double* den, const int* active, double* out) {
    int r = threadIdx.x;
    if (r < nrhs) out[r] = (active == nullptr || active[r]) ? num[r] / den[r] : 0.0;
}

__global__ void __launch_bounds__(NT) k_copy_f64(double* dst, const double* src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) dst[i] = src[i];
}

// ---------------------------------------------------------------------------------------
template <typename T>
hipError_t launch_perm_in(const double* b, size_t sx, size_t sr, int nrhs, const int* perm, int N, double* ws,
                          hipStream_t st) {
    k_perm_in<T><<<grid_for((size_t)N * nrhs), NT, 0, st>>>(b, sx, sr, nrhs, perm, N, ws);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_perm_out(const double* ws, const int* perm, int N, int nrhs, double* x, size_t sx, size_t sr,
                           hipStream_t st) {
    k_perm_out<T><<<grid_for((size_t)N * nrhs), NT, 0, st>>>(ws, perm, N, nrhs, x, sx, sr);
    return hipGetLastError();
}
hipError_t launch_cm2il(const double* src, int ld, int N, int nrhs, double* dst, hipStream_t st) {
    k_cm2il<<<grid_for((size_t)N * nrhs), NT, 0, st>>>(src, ld, N, nrhs, dst);
    return hipGetLastError();
}
hipError_t launch_il2cm(const double* src, int N, int nrhs, double* dst, int ld, hipStream_t st) {
    k_il2cm<<<grid_for((size_t)N * nrhs), NT, 0, st>>>(src, N, nrhs, dst, ld);
    return hipGetLastError();
}
// LDS for one RG chunk of the largest segment in a launch is sized by the caller through
// max_n; the launchers take the largest dimension from the task arrays' host copies.
template <typename T>
hipError_t launch_trmv(const TrsvTask* d_tasks, const SolveChunk* d_chunks, int nchunks, double* ws, double* wt, int nrhs,
                       int trans, hipStream_t st) {
    if (nchunks == 0) return hipSuccess;
    k_trmv<T><<<nchunks, NT, 0, st>>>(d_tasks, d_chunks, ws, wt, nrhs, trans);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_copyback<T><<<nchunks, NT, 0, st>>>(d_tasks, d_chunks, ws, wt, nrhs);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_gemv(const GemvGroup* d_groups, const SolveChunk* d_chunks, int nchunks, const GemvTerm* d_terms, double* ws,
                       int nrhs, hipStream_t st) {
    if (nchunks == 0) return hipSuccess;
    k_gemv<T><<<nchunks, NT, 0, st>>>(d_groups, d_chunks, d_terms, ws, nrhs);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_qapply(const QTask* d_tasks, int ntasks, double* ws, int nrhs, int trans, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    k_qapply<T><<<ntasks, NT, 0, st>>>(d_tasks, ws, nrhs, trans);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_larft(const LarftTask* d_tasks, int ntasks, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    k_larft<T><<<ntasks, NT, 0, st>>>(d_tasks);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_merge(const MergeTask* d_tasks, int ntasks, const int2* d_children, double* ws, int nrhs, int fwd,
                        hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    k_merge<T><<<ntasks, NT, 0, st>>>(d_tasks, d_children, ws, nrhs, fwd);
    return hipGetLastError();
}

hipError_t launch_spmv(int N, const int* rowptr, const int* colidx, const double* val, const double* x, double* y,
                       int nrhs, hipStream_t st) {
    k_spmv<<<grid_for((size_t)N * nrhs), NT, 0, st>>>(N, rowptr, colidx, val, x, y, nrhs);
    return hipGetLastError();
}
hipError_t launch_dot(int N, int nrhs, const double* a, const double* b, double* partial, double* out,
                      hipStream_t st) {
    k_dot_partial<<<DOT_BLOCKS, NT, 0, st>>>(N, nrhs, a, b, partial);
    k_dot_final<<<1, NT, 0, st>>>(DOT_BLOCKS, nrhs, partial, out);
    return hipGetLastError();
}
hipError_t launch_axpby(int N, int nrhs, const double* x, const double* ax, double sgn_ax, double* y,
                        const double* by, hipStream_t st) {
    k_axpby<<<grid_for((size_t)N * nrhs), NT, 0, st>>>(N, nrhs, x, ax, sgn_ax, y, by);
    return hipGetLastError();
}
hipError_t launch_gm_comb(int N, int nrhs, const double* base, const double* ye, int ny, const double* V, int k,
                          const double* z, double* out, hipStream_t st) {
    k_gm_comb<<<grid_for((size_t)N * nrhs), NT, 0, st>>>(N, nrhs, base, ye, ny, V, k, z, out);
    return hipGetLastError();
}
hipError_t launch_gm_multidot(int N, int nrhs, int k, const double* V, const double* v, double* partial, double* out,
                              hipStream_t st) {
    if (k == 0) return hipSuccess;
    constexpr int MB = 256;
    k_gm_multidot_partial<<<dim3(MB, k), NT, 0, st>>>(N, nrhs, V, v, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_gm_multidot_final<<<k * nrhs, NT, 0, st>>>(MB, k, nrhs, partial, out);
    return hipGetLastError();
}
hipError_t launch_gm_householder(int N, int nrhs, int k, const double* v, const double* d, double* u, hipStream_t st) {
    k_gm_householder<<<grid_for((size_t)N * nrhs), NT, 0, st>>>(N, nrhs, k, v, d, u);
    return hipGetLastError();
}
hipError_t launch_ratio(int nrhs, const double* num, const double* den, const int* active, double* out,
                        hipStream_t st) {
    k_ratio<<<1, 64, 0, st>>>(nrhs, num, den, active, out);
    return hipGetLastError();
}
hipError_t launch_copy_f64(double* dst, const double* src, size_t n, hipStream_t st) {
    k_copy_f64<<<grid_for(n), NT, 0, st>>>(dst, src, n);
    return hipGetLastError();
}

#define SPAND_SOLVE_INST(T)                                                                                  \
    template hipError_t launch_perm_in<T>(const double*, size_t, size_t, int, const int*, int, double*, hipStream_t); \
    template hipError_t launch_perm_out<T>(const double*, const int*, int, int, double*, size_t, size_t, hipStream_t); \
    template hipError_t launch_trmv<T>(const TrsvTask*, const SolveChunk*, int, double*, double*, int, int, hipStream_t); \
    template hipError_t launch_larft<T>(const LarftTask*, int, hipStream_t);                                \
    template hipError_t launch_gemv<T>(const GemvGroup*, const SolveChunk*, int, const GemvTerm*, double*, int, \
                                       hipStream_t);                                                        \
    template hipError_t launch_qapply<T>(const QTask*, int, double*, int, int, hipStream_t);                \
    template hipError_t launch_merge<T>(const MergeTask*, int, const int2*, double*, int, int, hipStream_t);
SPAND_SOLVE_INST(float)
SPAND_SOLVE_INST(double)

}  // namespace spand
