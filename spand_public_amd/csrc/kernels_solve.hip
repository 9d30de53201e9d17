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

#include "kernels.hpp"
#include "tasks.hpp"

namespace spand {

constexpr int NT = 256;
constexpr int SOLVE_LDS_ELEMS_BYTES = 32768;

static inline int grid_for(size_t n) {
    size_t g = (n + NT - 1) / NT;
    return (int)(g > 65536 ? 65536 : (g == 0 ? 1 : g));
}

// ---------------------------------------------------------------------------------------
// Element (i, r) of the caller's vector block lives at b[i * sx + r * sr] (column-major:
// sx = 1, sr = ld; interleaved: sx = nrhs, sr = 1).
template <typename T>
__global__ void __launch_bounds__(NT) k_perm_in(const double* b, size_t sx, size_t sr, int nrhs, const int* perm,
                                                int N, T* ws) {
    size_t n = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
        ws[idx] = (T)b[(size_t)perm[i] * sx + (size_t)r * sr];
    }
}
template <typename T>
__global__ void __launch_bounds__(NT) k_perm_out(const T* ws, const int* perm, int N, int nrhs, double* x, size_t sx,
                                                 size_t sr) {
    size_t n = (size_t)N * nrhs;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
        x[(size_t)perm[i] * sx + (size_t)r * sr] = (double)ws[idx];
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
// x := L^{-1} x (trans 0) or L^{-T} x (trans 1), both as column sweeps (the second on
// U = L^T) with explicit fused multiply-adds; the segment is staged in LDS when it fits.
template <typename T>
__global__ void __launch_bounds__(NT) k_trsv(const TrsvTask* tasks, T* ws, int nrhs, int trans, int lds_elems) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* S = reinterpret_cast<T*>(smem_raw);
    const TrsvTask t = tasks[blockIdx.x];
    const int n = t.n, tid = threadIdx.x;
    if (n == 0) return;
    const T* L = reinterpret_cast<const T*>(t.L);
    T* G = ws + (size_t)t.xoff * nrhs;
    const int nx = n * nrhs;
    const bool lds = nx <= lds_elems;
    T* x = lds ? S : G;
    if (lds) {
        for (int i = tid; i < nx; i += NT) S[i] = G[i];
        __syncthreads();
    }
    if (!trans) {
        for (int j = 0; j < n; j++) {
            const T ljj = L[j + (size_t)j * n];
            for (int r = tid; r < nrhs; r += NT) x[j * nrhs + r] /= ljj;
            __syncthreads();
            const int rem = (n - j - 1) * nrhs;
            for (int idx = tid; idx < rem; idx += NT) {
                int i = j + 1 + idx / nrhs, r = idx % nrhs;
                x[i * nrhs + r] = fma(-x[j * nrhs + r], L[i + (size_t)j * n], x[i * nrhs + r]);
            }
            __syncthreads();
        }
    } else {
        for (int j = n - 1; j >= 0; j--) {
            const T ljj = L[j + (size_t)j * n];
            for (int r = tid; r < nrhs; r += NT) x[j * nrhs + r] /= ljj;
            __syncthreads();
            const int rem = j * nrhs;
            for (int idx = tid; idx < rem; idx += NT) {
                int i = idx / nrhs, r = idx % nrhs;
                x[i * nrhs + r] = fma(-x[j * nrhs + r], L[j + (size_t)i * n], x[i * nrhs + r]);
            }
            __syncthreads();
        }
    }
    if (lds) {
        for (int i = tid; i < nx; i += NT) G[i] = S[i];
    }
}

// ---------------------------------------------------------------------------------------
// x[dst] -= sum_k op(A_k) x[src_k], terms applied in recorded order by the thread that owns
// each output element.
template <typename T>
__global__ void __launch_bounds__(NT) k_gemv(const GemvGroup* groups, const GemvTerm* terms, T* ws, int nrhs) {
#pragma clang fp contract(off)
    const GemvGroup g = groups[blockIdx.x];
    const int tid = threadIdx.x;
    const int nout = g.n * nrhs;
    T* xd = ws + (size_t)g.xdst * nrhs;
    for (int idx = tid; idx < nout; idx += NT) {
        const int i = idx / nrhs, r = idx % nrhs;
        T acc_out = xd[idx];
        for (int k = g.k0; k < g.k1; k++) {
            const GemvTerm tm = terms[k];
            const T* A = reinterpret_cast<const T*>(tm.A);
            const T* xs = ws + (size_t)tm.xsrc * nrhs + r;
            if (!tm.trans) {
                for (int j = 0; j < tm.cols; j++) acc_out = fma(-A[i + (size_t)j * tm.rows], xs[(size_t)j * nrhs], acc_out);
            } else {
                const T* a = A + (size_t)i * tm.rows;
                T acc = T(0);
                for (int j = 0; j < tm.rows; j++) acc = fma(a[j], xs[(size_t)j * nrhs], acc);
                acc_out -= acc;
            }
        }
        xd[idx] = acc_out;
    }
}

// ---------------------------------------------------------------------------------------
// Householder sequence: x := H(k-1)..H(0) x (trans 1, Q^T x) or H(0)..H(k-1) x (trans 0).
template <typename T>
__global__ void __launch_bounds__(NT) k_qapply(const QTask* tasks, T* ws, int nrhs, int trans, int lds_elems) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* S = reinterpret_cast<T*>(smem_raw);
    __shared__ T red[NT];
    const QTask t = tasks[blockIdx.x];
    const int m = t.m, k = t.k, tid = threadIdx.x;
    if (m == 0 || k == 0) return;
    const T* V = reinterpret_cast<const T*>(t.V);
    const T* tau = reinterpret_cast<const T*>(t.tau);
    T* G = ws + (size_t)t.xoff * nrhs;
    const int nx = m * nrhs;
    const bool lds = nx <= lds_elems;
    T* x = lds ? S : G;
    if (lds) {
        for (int i = tid; i < nx; i += NT) S[i] = G[i];
        __syncthreads();
    }
    int nrp = 1;
    while (nrp < nrhs) nrp <<= 1;
    const int groups = NT / nrp;
    const int r = tid % nrp, gidx = tid / nrp;
    for (int q = 0; q < k; q++) {
        const int i = trans ? q : k - 1 - q;
        const T* v = V + (size_t)i * m;
        T part = T(0);
        if (r < nrhs) {
            for (int j = i + 1 + gidx; j < m; j += groups) part += v[j] * x[j * nrhs + r];
        }
        red[tid] = part;
        __syncthreads();
        for (int off = groups / 2; off > 0; off >>= 1) {
            if (gidx < off) red[tid] += red[tid + off * nrp];
            __syncthreads();
        }
        // w_r = x_i + red[r];  x_i -= tau w ; x_j -= tau v_j w
        const T ti = tau[i];
        for (int idx = tid; idx < (m - i) * nrhs; idx += NT) {
            int j = i + idx / nrhs, rr = idx % nrhs;
            T w = x[i * nrhs + rr] + red[rr];
            T vj = (j == i) ? T(1) : v[j];
            (void)w;
            // defer the write of row i until every thread has read it
            if (j != i) x[j * nrhs + rr] -= ti * vj * w;
        }
        __syncthreads();
        for (int rr = tid; rr < nrhs; rr += NT) {
            T w = x[i * nrhs + rr] + red[rr];
            x[i * nrhs + rr] -= ti * w;
        }
        __syncthreads();
    }
    if (lds) {
        for (int i = tid; i < nx; i += NT) G[i] = S[i];
    }
}

// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(NT) k_merge(const MergeTask* tasks, const int2* children, T* ws, int nrhs, int fwd) {
    const MergeTask t = tasks[blockIdx.x];
    int k = 0;
    for (int c = t.c0; c < t.c1; c++) {
        const int2 ch = children[c];  // (xoff, size)
        const int n = ch.y * nrhs;
        T* P = ws + (size_t)(t.pxoff + k) * nrhs;
        T* Cx = ws + (size_t)ch.x * nrhs;
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

__global__ void k_ratio(int nrhs, const double* num, const double* den, const int* active, double* out) {
    int r = threadIdx.x;
    if (r < nrhs) out[r] = (active == nullptr || active[r]) ? num[r] / den[r] : 0.0;
}

__global__ void __launch_bounds__(NT) k_copy_f64(double* dst, const double* src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) dst[i] = src[i];
}

// ---------------------------------------------------------------------------------------
template <typename T>
hipError_t launch_perm_in(const double* b, size_t sx, size_t sr, int nrhs, const int* perm, int N, T* ws,
                          hipStream_t st) {
    k_perm_in<T><<<grid_for((size_t)N * nrhs), NT, 0, st>>>(b, sx, sr, nrhs, perm, N, ws);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_perm_out(const T* ws, const int* perm, int N, int nrhs, double* x, size_t sx, size_t sr,
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
hipError_t launch_trsv(const TrsvTask* d_tasks, int ntasks, T* ws, int nrhs, int trans, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    k_trsv<T><<<ntasks, NT, SOLVE_LDS_ELEMS_BYTES, st>>>(d_tasks, ws, nrhs, trans,
                                                         SOLVE_LDS_ELEMS_BYTES / (int)sizeof(T));
    return hipGetLastError();
}
template <typename T>
hipError_t launch_gemv(const GemvGroup* d_groups, int ngroups, const GemvTerm* d_terms, T* ws, int nrhs,
                       hipStream_t st) {
    if (ngroups == 0) return hipSuccess;
    k_gemv<T><<<ngroups, NT, 0, st>>>(d_groups, d_terms, ws, nrhs);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_qapply(const QTask* d_tasks, int ntasks, T* ws, int nrhs, int trans, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    k_qapply<T><<<ntasks, NT, SOLVE_LDS_ELEMS_BYTES, st>>>(d_tasks, ws, nrhs, trans,
                                                           SOLVE_LDS_ELEMS_BYTES / (int)sizeof(T));
    return hipGetLastError();
}
template <typename T>
hipError_t launch_merge(const MergeTask* d_tasks, int ntasks, const int2* d_children, T* ws, int nrhs, int fwd,
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
    template hipError_t launch_perm_in<T>(const double*, size_t, size_t, int, const int*, int, T*, hipStream_t); \
    template hipError_t launch_perm_out<T>(const T*, const int*, int, int, double*, size_t, size_t, hipStream_t); \
    template hipError_t launch_trsv<T>(const TrsvTask*, int, T*, int, int, hipStream_t);                    \
    template hipError_t launch_gemv<T>(const GemvGroup*, int, const GemvTerm*, T*, int, hipStream_t);       \
    template hipError_t launch_qapply<T>(const QTask*, int, T*, int, int, hipStream_t);                     \
    template hipError_t launch_merge<T>(const MergeTask*, int, const int2*, T*, int, int, hipStream_t);
SPAND_SOLVE_INST(float)
SPAND_SOLVE_INST(double)

}  // namespace spand
