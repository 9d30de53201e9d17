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
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "kernels.hpp"
#include "tasks.hpp"

namespace spand {

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
// potrf: right-looking column Cholesky of the lower triangle; every element receives its
// updates in k order as A(a,b) = fma(-L(a,k), L(b,k), A(a,b)). The block is staged in LDS
// when it fits, otherwise the same loop runs on the global copy (L2 resident).
template <typename T>
__global__ void __launch_bounds__(NT) k_potrf(const PotrfTask* tasks, int* info, int lds_elems) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* S = reinterpret_cast<T*>(smem_raw);
    __shared__ int bad;
    const PotrfTask t = tasks[blockIdx.x];
    const int n = t.n;
    const int tid = threadIdx.x;
    if (n == 0) return;
    T* G = reinterpret_cast<T*>(t.A);
    const bool lds = (long long)n * n <= lds_elems;
    T* M = lds ? S : G;
    if (lds) {
        for (int i = tid; i < n * n; i += NT) S[i] = G[i];
    }
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < n; j++) {
        if (tid == 0) {
            T ajj = M[j + (size_t)j * n];
            if (!(ajj > T(0))) {
                bad = j + 1;
            } else {
                M[j + (size_t)j * n] = sqrt(ajj);
            }
        }
        __syncthreads();
        if (bad) break;
        const T r = T(1) / M[j + (size_t)j * n];
        for (int i = j + 1 + tid; i < n; i += NT) M[i + (size_t)j * n] *= r;
        __syncthreads();
        const int m = n - j - 1;
        // trailing lower update: M(a,b) -= M(a,j) M(b,j), a >= b > j
        for (int idx = tid; idx < m * m; idx += NT) {
            int a = idx % m, b = idx / m;
            if (a >= b) {
                int ra = j + 1 + a, cb = j + 1 + b;
                M[ra + (size_t)cb * n] = fma(-M[ra + (size_t)j * n], M[cb + (size_t)j * n], M[ra + (size_t)cb * n]);
            }
        }
        __syncthreads();
    }
    if (bad) {
        if (tid == 0) atomicCAS(info, 0, t.cluster + 1);
        return;
    }
    if (lds) {
        for (int i = tid; i < n * n; i += NT) G[i] = S[i];
    }
}

// ---------------------------------------------------------------------------------------
// trsm: substitution in the reference BLAS element order, with explicit fused
// multiply-adds (the FP64 result is bit-identical to the oracle's). B is staged in LDS
// when it fits.
template <typename T>
__global__ void __launch_bounds__(NT) k_trsm(const TrsmTask* tasks, int lds_elems) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* S = reinterpret_cast<T*>(smem_raw);
    const TrsmTask t = tasks[blockIdx.x];
    const int n = t.n, m = t.m;
    const int tid = threadIdx.x;
    if (n == 0 || m == 0) return;
    const T* L = reinterpret_cast<const T*>(t.L);
    T* G = reinterpret_cast<T*>(t.B);
    const long long nb = (long long)n * m;
    const bool lds = nb <= lds_elems;
    T* B = lds ? S : G;
    if (lds) {
        for (int i = tid; i < nb; i += NT) S[i] = G[i];
        __syncthreads();
    }
    if (t.side == 0) {
        // B (n x m, ld n) := L^{-1} B
        for (int k = 0; k < n; k++) {
            const T lkk = L[k + (size_t)k * n];
            for (int c = tid; c < m; c += NT) B[k + (size_t)c * n] /= lkk;
            __syncthreads();
            const int rem = n - k - 1;
            for (int idx = tid; idx < rem * m; idx += NT) {
                int i = k + 1 + idx % rem, c = idx / rem;
                B[i + (size_t)c * n] = fma(-B[k + (size_t)c * n], L[i + (size_t)k * n], B[i + (size_t)c * n]);
            }
            __syncthreads();
        }
    } else {
        // B (m x n, ld m) := B L^{-T}
        for (int k = 0; k < n; k++) {
            const T r = T(1) / L[k + (size_t)k * n];
            for (int i = tid; i < m; i += NT) B[i + (size_t)k * m] *= r;
            __syncthreads();
            const int rem = n - k - 1;
            for (int idx = tid; idx < rem * m; idx += NT) {
                int i = idx % m, j = k + 1 + idx / m;
                B[i + (size_t)j * m] = fma(-L[j + (size_t)k * n], B[i + (size_t)k * m], B[i + (size_t)j * m]);
            }
            __syncthreads();
        }
    }
    if (lds) {
        for (int i = tid; i < nb; i += NT) G[i] = S[i];
    }
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
    T* C = reinterpret_cast<T*>(t.C);
    T acc_c[4];
    for (int q = 0; q < 4; q++) {
        int i = ti, j = tj0 + q;
        acc_c[q] = (i < t.tm && j < t.tn) ? C[(t.r0 + i) + (size_t)(t.c0 + j) * t.ldc] : T(0);
    }
    for (int kk = t.k0; kk < t.k1; kk++) {
        const SchurContrib c = contribs[kk];
        const T* A = reinterpret_cast<const T*>(c.A);
        const T* B = reinterpret_cast<const T*>(c.B);
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
            for (int l = 0; l < kc; l++) {
                const T a = -As[l][ti];
#pragma unroll
                for (int q = 0; q < 4; q++) acc_c[q] = fma(a, Bs[l][tj0 + q], acc_c[q]);
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

// ---------------------------------------------------------------------------------------
// RRQR: gather Asn, dgeqp3 (dlaqp2 order: first-max pivot, LAWN-176 norm downdate with
// recomputation below sqrt(eps)), choose_rank, Q/tau store and scatter of R[:r,:] P^T into
// the edge blocks. Column norms, the downdate and the rank test run in FP64 for both
// element types. Scalar steps follow the reference LAPACK formulas one-for-one so that an
// FP64 run on the same input reproduces LAPACK's pivot sequence.
constexpr int RRQR_MAXSEG = 1024;
constexpr int RRQR_LDS_COLS = 1024;

__device__ inline double lapy2_d(double x, double y) {
    double xa = fabs(x), ya = fabs(y);
    double w = fmax(xa, ya), z = fmin(xa, ya);
    if (z == 0.0) return w;
    double q = z / w;
    return w * sqrt(1.0 + q * q);
}

template <typename T>
__global__ void __launch_bounds__(NT) k_rrqr(const RrqrTask* tasks, const RrqrSeg* segs, int* dsize,
                                             double tol, int lds_elems) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_col0[RRQR_MAXSEG + 1];
    __shared__ double r_val[NT];
    __shared__ int r_idx[NT];
    __shared__ double s_vn1[RRQR_LDS_COLS], s_vn2[RRQR_LDS_COLS];
    __shared__ int s_jp[RRQR_LDS_COLS];
    __shared__ int s_rank;
    __shared__ double s_tau, s_scal;

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
    T* A = ((long long)rows * cols <= lds_elems) ? reinterpret_cast<T*>(smem_raw) : reinterpret_cast<T*>(t.scratch);
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
        A[idx] = sg[k].out ? E[jj + (size_t)i * w] : E[i + (size_t)jj * rows];
    }
    __syncthreads();
    for (int j = tid; j < cols; j += NT) {
        double s = 0.0;
        const T* a = A + (size_t)j * rows;
        for (int i = 0; i < rows; i++) {
            double v = (double)a[i];
            s = fma(v, v, s);
        }
        jp[j] = j;
        vn1[j] = sqrt(s);
        vn2[j] = vn1[j];
    }
    __syncthreads();
    const double tol3z = sqrt(DBL_EPSILON * 0.5);
    const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
    for (int i = 0; i < mn; i++) {
        // pivot = first index of the max of vn1[i:cols]
        double bv = -1.0;
        int bi = cols;
        for (int j = i + tid; j < cols; j += NT) {
            double v = fabs(vn1[j]);
            if (v > bv) { bv = v; bi = j; }
        }
        r_val[tid] = bv;
        r_idx[tid] = bi;
        __syncthreads();
        for (int off = NT / 2; off > 0; off >>= 1) {
            if (tid < off) {
                double v2 = r_val[tid + off];
                int i2 = r_idx[tid + off];
                if (v2 > r_val[tid] || (v2 == r_val[tid] && i2 < r_idx[tid])) {
                    r_val[tid] = v2;
                    r_idx[tid] = i2;
                }
            }
            __syncthreads();
        }
        const int pvt = r_idx[0];
        if (pvt != i) {
            for (int r = tid; r < rows; r += NT) {
                T tmp = A[r + (size_t)pvt * rows];
                A[r + (size_t)pvt * rows] = A[r + (size_t)i * rows];
                A[r + (size_t)i * rows] = tmp;
            }
            if (tid == 0) {
                int tj = jp[pvt]; jp[pvt] = jp[i]; jp[i] = tj;
                vn1[pvt] = vn1[i];
                vn2[pvt] = vn2[i];
            }
        }
        __syncthreads();
        // dlarfg on A(i:rows, i)
        T* ci = A + (size_t)i * rows;
        if (tid == 0) {
            double tau = 0.0, scal = 1.0;
            const int nh = rows - i;
            if (nh > 1) {
                double xs = 0.0;
                for (int r = i + 1; r < rows; r++) {
                    double v = (double)ci[r];
                    xs = fma(v, v, xs);
                }
                double xnorm = sqrt(xs);
                if (xnorm != 0.0) {
                    double alpha = (double)ci[i];
                    double beta = -copysign(lapy2_d(alpha, xnorm), alpha);
                    int knt = 0;
                    if (fabs(beta) < safmin) {
                        const double rs = 1.0 / safmin;
                        do {
                            knt++;
                            for (int r = i + 1; r < rows; r++) ci[r] = (T)((double)ci[r] * rs);
                            beta *= rs;
                            alpha *= rs;
                        } while (fabs(beta) < safmin && knt < 20);
                        xs = 0.0;
                        for (int r = i + 1; r < rows; r++) {
                            double v = (double)ci[r];
                            xs = fma(v, v, xs);
                        }
                        beta = -copysign(lapy2_d(alpha, sqrt(xs)), alpha);
                    }
                    tau = (beta - alpha) / beta;
                    scal = 1.0 / (alpha - beta);
                    for (int j = 0; j < knt; j++) beta *= safmin;
                    ci[i] = (T)beta;
                }
            }
            s_tau = tau;
            s_scal = scal;
        }
        __syncthreads();
        const T tau = (T)s_tau;
        if (s_tau != 0.0) {
            const T scal = (T)s_scal;
            for (int r = i + 1 + tid; r < rows; r += NT) ci[r] *= scal;
        }
        __syncthreads();
        // apply H(i) to A(i:rows, i+1:cols) (dlarf) and downdate the partial norms
        for (int j = i + 1 + tid; j < cols; j += NT) {
            T* cj = A + (size_t)j * rows;
            if (tau != T(0)) {
                T w = cj[i];
                for (int r = i + 1; r < rows; r++) w = fma(cj[r], ci[r], w);
                T tm = -tau * w;
                cj[i] = cj[i] + tm;
                for (int r = i + 1; r < rows; r++) cj[r] = fma(ci[r], tm, cj[r]);
            }
            if (vn1[j] != 0.0) {
                double q = fabs((double)cj[i]) / vn1[j];
                double temp = 1.0 - q * q;
                temp = fmax(temp, 0.0);
                double rr = vn1[j] / vn2[j];
                double temp2 = temp * rr * rr;
                if (temp2 <= tol3z) {
                    if (i < rows - 1) {
                        double s = 0.0;
                        for (int r = i + 1; r < rows; r++) {
                            double v = (double)cj[r];
                            s = fma(v, v, s);
                        }
                        vn1[j] = sqrt(s);
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
        if (tid == 0) reinterpret_cast<T*>(t.tau)[i] = tau;
        __syncthreads();
    }
    // choose_rank (src/util.cpp:434-452) on diag(R), FP64
    if (tid == 0) {
        int rank;
        if (tol == 0.0) rank = mn;
        else if (tol >= 1.0) rank = 0;
        else if (mn <= 1) rank = mn;
        else {
            double sref = fabs((double)A[0]);
            rank = 1;
            while (rank < mn && fabs((double)A[rank + (size_t)rank * rows]) / sref >= tol) rank++;
        }
        s_rank = rank;
    }
    __syncthreads();
    const int rank = s_rank;
    if (rank >= rows) return;
    // Householder vectors v = A[:, :rank] (rows x rank, ld rows)
    T* Q = reinterpret_cast<T*>(t.Q);
    for (long long idx = tid; idx < (long long)rows * rank; idx += NT) Q[idx] = A[idx];
    // AsnP[:, jpvt[j]] = triu(R)[:rank, j], scattered into the shrunk edge blocks
    for (int j = tid; j < cols; j += NT) {
        int p = jp[j];
        int k = seg_of(p);
        int jj = p - s_col0[k], w = s_col0[k + 1] - s_col0[k];
        T* E = reinterpret_cast<T*>(sg[k].E);
        const T* rj = A + (size_t)j * rows;
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
template <typename T>
__global__ void __launch_bounds__(NT) k_copy(const CopyTask* tasks) {
    const CopyTask t = tasks[blockIdx.x];
    const T* S = reinterpret_cast<const T*>(t.src);
    T* D = reinterpret_cast<T*>(t.dst);
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
    const int bytes = LDS_BYTES;
    k_potrf<T><<<ntasks, NT, bytes, st>>>(d_tasks, d_info, bytes / (int)sizeof(T));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_trsm(const TrsmTask* d_tasks, int ntasks, hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    const int bytes = LDS_BYTES;
    k_trsm<T><<<ntasks, NT, bytes, st>>>(d_tasks, bytes / (int)sizeof(T));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_schur(const SchurTile* d_tiles, int ntiles, const SchurContrib* d_contribs, hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    k_schur<T><<<ntiles, NT, 0, st>>>(d_tiles, d_contribs);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rrqr(const RrqrTask* d_tasks, int ntasks, const RrqrSeg* d_segs, int* d_size, double tol,
                       hipStream_t st) {
    if (ntasks == 0) return hipSuccess;
    const int bytes = LDS_BYTES;
    k_rrqr<T><<<ntasks, NT, bytes, st>>>(d_tasks, d_segs, d_size, tol, bytes / (int)sizeof(T));
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
    template hipError_t launch_trsm<T>(const TrsmTask*, int, hipStream_t);                        \
    template hipError_t launch_schur<T>(const SchurTile*, int, const SchurContrib*, hipStream_t); \
    template hipError_t launch_rrqr<T>(const RrqrTask*, int, const RrqrSeg*, int*, double, hipStream_t); \
    template hipError_t launch_copy<T>(const CopyTask*, int, hipStream_t);
SPAND_INST(float)
SPAND_INST(double)

}  // namespace spand
