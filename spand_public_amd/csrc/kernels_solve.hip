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

// Device-memory pointers that the task records carry as plain pointers: accessed through a
// pointer in the global address space, the compiler emits global_* instead of flat_*
// instructions (a flat access also holds the LDS wait counter).
template <typename X>
using gptr = __attribute__((address_space(1))) X*;
template <typename X>
__device__ __attribute__((always_inline)) inline gptr<X> gp(X* p) {
    return (gptr<X>)p;
}

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
// mv phases (task layouts in tasks.hpp). A workgroup computes one piece: 64 outputs of one
// chunk over at most MV_PIECE_WORK inner positions, gathered from at most MV_PIECE_TERMS
// terms. The work is staged in one round so that its global loads are in flight together:
//   1. one lane per term reads the term record, clips its inner range, and a wavefront scan
//      packs the terms' ranges into consecutive positions;
//   2. every thread loads its share of the piece's matrix entries (MV_EPT of them, all
//      issued before any is stored) into tile[position][output], and the sources of up to
//      RG_MAX right-hand sides into xl[position][r];
//   3. 4 partitions of the positions (one wavefront each, position p to partition p % 4)
//      accumulate each output, and the partitions are added in a fixed order.
// A chunk with several pieces writes partial sums, which k_mv_reduce adds in piece order
// after the launch. (Adding them in the chunk's last piece to finish, found by a counter,
// measured 3-7x slower on gfx950: the device-scope release fence each piece needs writes its
// XCD's L2 back.)
constexpr int RG_MAX = 8;
constexpr int WAVE = 64;
constexpr int SC = SOLVE_CHUNK;
constexpr int NP = NT / SC;                      // 4 partitions of the inner positions
constexpr int PW = MV_PIECE_WORK;                // positions per piece
constexpr int MV_EPT = PW * SC / NT;             // matrix entries per thread: 16
static_assert(MV_PIECE_TERMS <= WAVE, "one lane per term record");
static_assert(PW * RG_MAX <= 2 * NT, "two source loads per thread");

struct MvTermL {  // a term's record in LDS, with its clipped inner range and first position
    const void* A;
    const int* perm;
    int rows, lo, len, off, xsrc, trans;
};

// y (op)= sum: the chunk's outputs, in the group's mode, through dperm if the group has one
__device__ inline void mv_apply(VT* y, int mode, VT sum) {
    *y = mode == MV_SUB ? *y - sum : (mode == MV_SETNEG ? -sum : sum);
}

// adds the partial sums of one chunk's pieces in piece order and applies them
__device__ inline void mv_chunk_reduce(const MvChunk& ch, const MvGroup& g, const VT* partial, VT* ws, int nrhs) {
    const int no = min(SC, g.n - ch.o0);
    for (int idx = threadIdx.x; idx < no * nrhs; idx += NT) {
        const int lane = idx / nrhs, r = idx % nrhs;
        VT sum = VT(0);
        for (int p = ch.p0; p < ch.p1; p++) sum += partial[((size_t)p * SC + lane) * nrhs + r];
        const int o = ch.o0 + lane;
        mv_apply(ws + (size_t)(g.xdst + (g.dperm ? gp(g.dperm)[o] : o)) * nrhs + r, g.mode, sum);
    }
}

// RED_IN_TILE (only for nrhs <= RG_MAX, one right-hand-side group): the partitions' sums are
// written over the tile once every partition has read it. That cuts the LDS of a workgroup
// from 35.7 to 23.4 KB (FP32) and 52.3 to 40.0 KB (FP64): 7 instead of 4 (4 instead of 3)
// workgroups share a CU, so a phase of up to 1792 (1024) pieces loads in one round.
template <typename T, bool RED_IN_TILE>
__global__ void __launch_bounds__(NT) k_mv_piece(const MvPiece* pieces, const MvTerm* terms, VT* ws, VT* partial,
                                                 int nrhs) {
    constexpr size_t TILE_B = sizeof(T[PW][SC + 1]), RED_B = sizeof(VT[NP - 1][SC][RG_MAX]);
    static_assert(TILE_B % alignof(VT) == 0, "red follows the tile");
    __shared__ __align__(16) unsigned char s_tr[RED_IN_TILE ? (TILE_B > RED_B ? TILE_B : RED_B) : TILE_B + RED_B];
    // the piece's matrix: inner position x output
    T (*const tile)[SC + 1] = reinterpret_cast<T (*)[SC + 1]>(s_tr);
    // partitions 1..3 (partition 0 keeps its registers)
    VT (*const red)[SC][RG_MAX] = reinterpret_cast<VT (*)[SC][RG_MAX]>(s_tr + (RED_IN_TILE ? 0 : TILE_B));
    __shared__ VT xl[PW][RG_MAX];           // the sources of the current right-hand-side group
    __shared__ MvTermL st[MV_PIECE_TERMS];
    __shared__ unsigned char p2t[PW];       // position -> term
    __shared__ int s_total;
    const MvPiece pc = pieces[blockIdx.x];
    const int o0 = pc.o0, no = pc.no, nt = pc.k1 - pc.k0;
    const int tid = threadIdx.x, lane = tid % SC, kp = tid / SC;
    // 1. term records and their positions
    if (tid < WAVE) {
        int len = 0;
        if (tid < nt) {
            const MvTerm tm = terms[pc.k0 + tid];
            int lo = 0, hi = tm.trans ? tm.rows : tm.cols;
            if (nt == 1) {
                lo = pc.j0;
                hi = pc.j1;
            }
            if (tm.tri == 1) hi = min(hi, o0 + SC);
            if (tm.tri == 2) lo = max(lo, o0);
            len = max(0, hi - lo);
            st[tid] = MvTermL{tm.A, tm.perm, tm.rows, lo, len, 0, tm.xsrc, tm.trans};
        }
        int incl = len;  // inclusive scan over the wavefront
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const int v = __shfl_up(incl, d);
            if (tid >= d) incl += v;
        }
        if (tid < nt) st[tid].off = incl - len;
        if (tid == WAVE - 1) s_total = incl;
    }
    __syncthreads();
    const int total = s_total;  // <= PW (the planner's bound)
    for (int p = tid; p < total; p += NT) {  // the last term starting at or before p holds p
        int lo = 0, hi = nt - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (st[mid].off <= p) lo = mid;
            else hi = mid - 1;
        }
        p2t[p] = (unsigned char)lo;
    }
    __syncthreads();
    // 2. the matrix entries: entry e of the piece belongs to the term holding position e / SC;
    // within the term's block q = e - SC * off runs along the outputs of one inner index (A's
    // column, not transposed) or along the inner indices of one output (A's column, transposed),
    // so that neighbouring lanes read neighbouring addresses either way
    const int ne = total * SC;
    T v[MV_EPT];
    int dst[MV_EPT];
#pragma unroll
    for (int u = 0; u < MV_EPT; u++) {
        const int e = tid + u * NT;
        v[u] = T(0);
        dst[u] = -1;
        if (e < ne) {
            const MvTermL& q = st[p2t[e / SC]];
            const int qe = e - q.off * SC;
            int p, i;
            if (q.trans) {
                p = qe % q.len;
                i = qe / q.len;
            } else {
                p = qe / SC;
                i = qe % SC;
            }
            const gptr<const T> A = gp(reinterpret_cast<const T*>(q.A));
            if (i < no) v[u] = q.trans ? A[(q.lo + p) + (size_t)(o0 + i) * q.rows] : A[(o0 + i) + (size_t)(q.lo + p) * q.rows];
            dst[u] = (q.off + p) * (SC + 1) + i;
        }
    }
    auto stage_x = [&](int r0, int rg) {
        VT xv[2];
        int xd[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int idx = tid + u * NT;
            xd[u] = -1;
            if (idx < total * rg) {
                const int pos = idx / rg, r = idx % rg;
                const MvTermL& q = st[p2t[pos]];
                const int j = q.lo + (pos - q.off);
                const int src = (!q.trans && q.perm) ? gp(q.perm)[j] : j;
                xv[u] = ws[(size_t)(q.xsrc + src) * nrhs + r0 + r];
                xd[u] = pos * RG_MAX + r;
            }
        }
#pragma unroll
        for (int u = 0; u < 2; u++)
            if (xd[u] >= 0) (&xl[0][0])[xd[u]] = xv[u];
    };
    stage_x(0, min(RG_MAX, nrhs));
#pragma unroll
    for (int u = 0; u < MV_EPT; u++)
        if (dst[u] >= 0) (&tile[0][0])[dst[u]] = v[u];
    __syncthreads();
    // 3. accumulate, per group of up to RG_MAX right-hand sides
    for (int r0 = 0; r0 < nrhs; r0 += RG_MAX) {
        const int rg = min(RG_MAX, nrhs - r0);
        if (r0 > 0) {
            __syncthreads();  // every read of xl for the previous group is done
            stage_x(r0, rg);
            __syncthreads();
        }
        VT acc[RG_MAX];
#pragma unroll
        for (int r = 0; r < RG_MAX; r++) acc[r] = VT(0);
        // positions kp, kp + 4, ... < total (at most 16 per partition), unrolled so that the
        // LDS reads run ahead; the sum order per output is the position order either way
#pragma unroll
        for (int q = 0; q < PW / NP; q++) {
            const int p = kp + q * NP;
            if (p < total) {
                const VT a = tile[p][lane];
#pragma unroll
                for (int r = 0; r < RG_MAX; r++)
                    if (r < rg) acc[r] += a * xl[p][r];
            }
        }
        if (RED_IN_TILE) __syncthreads();  // every partition has read the tile
        if (kp > 0)
#pragma unroll
            for (int r = 0; r < RG_MAX; r++) red[kp - 1][lane][r] = acc[r];
        __syncthreads();
        if (kp == 0 && lane < no) {
            if (pc.direct) {
                VT* y = ws + (size_t)(pc.xdst + o0 + lane) * nrhs + r0;
                for (int r = 0; r < rg; r++)
                    mv_apply(y + r, pc.mode, ((acc[r] + red[0][lane][r]) + red[1][lane][r]) + red[2][lane][r]);
            } else {
                VT* y = partial + ((size_t)blockIdx.x * SC + lane) * nrhs + r0;
                for (int r = 0; r < rg; r++) y[r] = ((acc[r] + red[0][lane][r]) + red[1][lane][r]) + red[2][lane][r];
            }
        }
        __syncthreads();  // red is rewritten by the next group
    }
}

// Chunks with several pieces: add the pieces' partial sums in piece order, then apply.
__global__ void __launch_bounds__(NT) k_mv_reduce(const MvGroup* groups, const MvChunk* chunks, const int* rchunks,
                                                  const VT* partial, VT* ws, int nrhs) {
    const MvChunk ch = chunks[rchunks[blockIdx.x]];
    mv_chunk_reduce(ch, groups[ch.group], partial, ws, nrhs);
}

// ---------------------------------------------------------------------------------------
// Compact-WY factors of the Orthogonal ops (task layout in tasks.hpp): Ve = V with its unit
// diagonal and zeros above; S = diag(1 / tau) + stril(Ve^T Ve); X = identity (the caller
// then solves S X = I with the batched trsm, so that X = T^T).
template <typename T>
__global__ void __launch_bounds__(NT) k_wy_explicit(const WyTask* tasks) {
    const WyTask t = tasks[blockIdx.x];
    const T* V = reinterpret_cast<const T*>(t.V);
    T* Ve = reinterpret_cast<T*>(t.Ve);
    const size_t n = (size_t)t.m * t.k;
    for (size_t idx = (size_t)blockIdx.y * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.y * NT) {
        const int i = (int)(idx % t.m), c = (int)(idx / t.m);
        Ve[idx] = i == c ? T(1) : (i > c ? V[idx] : T(0));
    }
    // S on and above the diagonal, X
    const T* tau = reinterpret_cast<const T*>(t.tau);
    T* S = reinterpret_cast<T*>(t.S);
    T* X = reinterpret_cast<T*>(t.X);
    const size_t kk = (size_t)t.k * t.k;
    for (size_t idx = (size_t)blockIdx.y * NT + threadIdx.x; idx < kk; idx += (size_t)gridDim.y * NT) {
        const int a = (int)(idx % t.k), c = (int)(idx / t.k);
        if (a <= c) S[idx] = a == c ? T(1) / tau[a] : T(0);
        X[idx] = a == c ? T(1) : T(0);
    }
}

// one workgroup per 32 x 32 tile (a0, c0), a0 <= c0, of one op's Gram matrix; entry
// (a, c), a < c, goes to S(c, a)
template <typename T>
__global__ void __launch_bounds__(NT) k_wy_gram(const WyTask* tasks, const int4* tiles) {
    constexpr int GT = 32;  // tile edge (the host plans 32 x 32 tiles)
    __shared__ T As[GT][GT + 1], Cs[GT][GT + 1];
    const int4 tl = tiles[blockIdx.x];  // (task, a0, c0, unused)
    const WyTask t = tasks[tl.x];
    const int a0 = tl.y, c0 = tl.z, m = t.m, k = t.k;
    const T* Ve = reinterpret_cast<const T*>(t.Ve);
    const int tid = threadIdx.x;
    // thread -> 4 outputs: a = a0 + tid % 32, c = c0 + (tid / 32) * 4 + q
    const int ta = tid % GT, tc0 = (tid / GT) * 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i0 = c0; i0 < m; i0 += GT) {  // Ve(i, c) = 0 for i < c
        const int ni = min(GT, m - i0);
        for (int idx = tid; idx < GT * GT; idx += NT) {
            const int ii = idx % GT, cc = idx / GT;
            As[ii][cc] = (ii < ni && a0 + cc < k) ? Ve[(i0 + ii) + (size_t)(a0 + cc) * m] : T(0);
            Cs[ii][cc] = (ii < ni && c0 + cc < k) ? Ve[(i0 + ii) + (size_t)(c0 + cc) * m] : T(0);
        }
        __syncthreads();
        for (int ii = 0; ii < ni; ii++) {
            const double a = As[ii][ta];
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] += a * (double)Cs[ii][tc0 + q];
        }
        __syncthreads();
    }
    T* S = reinterpret_cast<T*>(t.S);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int a = a0 + ta, c = c0 + tc0 + q;
        if (a < k && c < k && a < c) S[c + (size_t)a * k] = (T)acc[q];
    }
}

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

// Sampled outer product on A's pattern: for stored entry p = (i, j) (CSR position, row i),
// out[csc_of[p]] = alpha * sum over r of u(i, r) v(j, r), the right-hand sides in
// increasing r; u, v column-major with leading dimension ld. One thread per row.
__global__ void __launch_bounds__(NT) k_sampled_outer(int N, const int* rowptr, const int* colidx, const int* csc_of,
                                                      const double* u, const double* v, int nrhs, int ld,
                                                      double alpha, double* out) {
    for (int i = blockIdx.x * NT + threadIdx.x; i < N; i += gridDim.x * NT) {
        for (int p = rowptr[i]; p < rowptr[i + 1]; p++) {
            const int j = colidx[p];
            double acc = 0.0;
            for (int r = 0; r < nrhs; r++) acc = fma(u[i + (size_t)r * ld], v[j + (size_t)r * ld], acc);
            out[csc_of[p]] = alpha * acc;
        }
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

// dst[k] = src[idx[k]]: values of a fixed pattern moved from the caller's order to a stored one
template <typename I>
__global__ void __launch_bounds__(NT) k_gather_f64(const I* idx, const double* src, long long n, double* dst) {
    for (long long k = (long long)blockIdx.x * NT + threadIdx.x; k < n; k += (long long)gridDim.x * NT)
        dst[k] = src[idx[k]];
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
template <typename T>
hipError_t launch_mv(const MvGroup* d_groups, const MvChunk* d_chunks, const MvPiece* d_pieces, int npieces,
                     const MvTerm* d_terms, const int* d_rchunks, int nreduce, double* ws, double* partial, int nrhs,
                     hipStream_t st) {
    if (npieces == 0) return hipSuccess;
    if (nrhs <= RG_MAX) k_mv_piece<T, true><<<npieces, NT, 0, st>>>(d_pieces, d_terms, ws, partial, nrhs);
    else k_mv_piece<T, false><<<npieces, NT, 0, st>>>(d_pieces, d_terms, ws, partial, nrhs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nreduce == 0) return e;
    k_mv_reduce<<<nreduce, NT, 0, st>>>(d_groups, d_chunks, d_rchunks, partial, ws, nrhs);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_wy(const WyTask* d_tasks, int ntasks, const int4* d_tiles, int ntiles, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    k_wy_explicit<T><<<dim3(ntasks, 16), NT, 0, st>>>(d_tasks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ntiles == 0) return e;
    k_wy_gram<T><<<ntiles, NT, 0, st>>>(d_tasks, d_tiles);
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
hipError_t launch_sampled_outer(int N, const int* rowptr, const int* colidx, const int* csc_of, const double* u,
                                const double* v, int nrhs, int ld, double alpha, double* out, hipStream_t st) {
    if (N == 0) return hipSuccess;
    k_sampled_outer<<<grid_for((size_t)N), NT, 0, st>>>(N, rowptr, colidx, csc_of, u, v, nrhs, ld, alpha, out);
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
hipError_t launch_gather_f64(const long long* idx, const double* src, long long n, double* dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_gather_f64<long long><<<grid_for((size_t)n), NT, 0, st>>>(idx, src, n, dst);
    return hipGetLastError();
}
hipError_t launch_gather_f64(const int* idx, const double* src, long long n, double* dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_gather_f64<int><<<grid_for((size_t)n), NT, 0, st>>>(idx, src, n, dst);
    return hipGetLastError();
}

#define SPAND_SOLVE_INST(T)                                                                                  \
    template hipError_t launch_perm_in<T>(const double*, size_t, size_t, int, const int*, int, double*, hipStream_t); \
    template hipError_t launch_perm_out<T>(const double*, const int*, int, int, double*, size_t, size_t, hipStream_t); \
    template hipError_t launch_mv<T>(const MvGroup*, const MvChunk*, const MvPiece*, int, const MvTerm*, const int*, \
                                     int, double*, double*, int, hipStream_t);                             \
    template hipError_t launch_wy<T>(const WyTask*, int, const int4*, int, hipStream_t);                    \
    template hipError_t launch_merge<T>(const MergeTask*, int, const int2*, double*, int, int, hipStream_t);
SPAND_SOLVE_INST(float)
SPAND_SOLVE_INST(double)

}  // namespace spand
