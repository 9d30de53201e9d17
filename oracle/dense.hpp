// ORACLE — test infrastructure only. CPU FP64 restatement of the dense LAPACK/BLAS
// kernels that spaND calls through src/util.cpp:122-452 of the reference. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code, and only as
// the checker; the product library never links or calls it.
//
// All matrices are column-major with an explicit leading dimension. Each routine follows
// the operation order of the reference-BLAS / reference-LAPACK routine named next to it,
// so results can be pinned against scipy.linalg.lapack / scipy.linalg.blas golden vectors.
#pragma once
#include <vector>

namespace oracle {

// dnrm2-equivalent: sequential sum of squares in index order with fused multiply-add,
// then sqrt.
double nrm2(int n, const double* x, int incx);

// Reduction order of the column-pivoted QR (geqp3 below; the GPU RRQR kernels follow it
// term for term). The n terms are dealt round-robin to 16 chains (term t to chain t % 16);
// each chain accumulates its terms in increasing t with fused multiply-adds, starting from
// 0; the 16 chain values are combined by the halving tree
//   c[w] += c[w + 8] (w < 8), c[w] += c[w + 4] (w < 4), c[w] += c[w + 2] (w < 2), c[0] + c[1].
// BLAS leaves the order of dnrm2 / ddot / dgemv sums to the implementation; this one is
// chosen so that a 16-wavefront workgroup computes every sum in parallel.
constexpr int kChains = 16;
double tree16(double c[kChains]);
double sumsq16(int n, const double* x, int incx);           // sum x[t]^2
double dot16(int n, const double* a, const double* b);      // sum a[t] * b[t]

// dpotrf('L') semantics: A = L L^T in place (lower triangle), returns info (0 or j+1 of
// the first non-positive pivot). Reference call site: src/util.cpp:158-165.
int potrf_lower(int n, double* A, int lda);

// B := L^{-1} B (dtrsm Left/Lower/NoTrans/NonUnit), B is m x n, L is m x m.
// Reference call site: src/util.cpp:266-274 (trsm_left).
void trsm_left_lower_notrans(int m, int n, const double* L, int ldl, double* B, int ldb);
// B := B L^{-T} (dtrsm Right/Lower/Trans/NonUnit), B is m x n, L is n x n.
// Reference call site: src/util.cpp:256-264 (trsm_right).
void trsm_right_lower_trans(int m, int n, const double* L, int ldl, double* B, int ldb);
// B := B U^{-1} (dtrsm Right/Upper/NoTrans/NonUnit), B is m x n, U is n x n.
void trsm_right_upper_notrans(int m, int n, const double* U, int ldu, double* B, int ldb);

// C := alpha A A^T + beta C (trans=false, A is n x k) or alpha A^T A + beta C
// (trans=true, A is k x n); lower triangle only (dsyrk 'L'). src/util.cpp:148-156.
void syrk_lower(bool trans, int n, int k, double alpha, const double* A, int lda,
                double beta, double* C, int ldc);

// Schur-complement update C -= op(A) op(B) (lower triangle only if `lower`), each element
// accumulated in k order with fused multiply-adds: the gemm/syrk calls of gemm_edges
// (src/tree.cpp:778-781) with alpha = -1, beta = 1.
void schur_update(bool tA, bool tB, bool lower, int m, int n, int k, const double* A, int lda,
                  const double* B, int ldb, double* C, int ldc);

// C := alpha op(A) op(B) + beta C (dgemm). src/util.cpp:122-138.
void gemm(bool tA, bool tB, int m, int n, int k, double alpha, const double* A, int lda,
          const double* B, int ldb, double beta, double* C, int ldc);

// x := L^{-1} x, L^{-T} x, U^{-1} x (dtrsv NonUnit). src/util.cpp:276-283.
void trsv_lower_notrans(int n, const double* A, int lda, double* x);
void trsv_lower_trans(int n, const double* A, int lda, double* x);
void trsv_upper_notrans(int n, const double* A, int lda, double* x);

// y := y - A x  (dgemv NoTrans, alpha=-1, beta=1), A is m x n. src/util.cpp:306-315.
void gemv_sub_notrans(int m, int n, const double* A, int lda, const double* x, double* y);
// y := y - A^T x (dgemv Trans, alpha=-1, beta=1), A is m x n. src/util.cpp:317-326.
void gemv_sub_trans(int m, int n, const double* A, int lda, const double* x, double* y);

// dlarfg: generates H = I - tau v v^T with v(0)=1 such that H [alpha; x] = [beta; 0];
// ||x|| by sumsq16.
void larfg(int n, double* alpha, double* x, int incx, double* tau);
// dlarf 'Left': C := (I - tau v v^T) C, C is m x n, v has v(0) explicitly stored;
// w = C^T v by dot16.
void larf_left(int m, int n, const double* v, double tau, double* C, int ldc, double* work);

// dgeqp3 (all columns free): column-pivoted Householder QR in the unblocked dlaqp2 step
// order (pivot = first column of largest partial norm, LAWN 176 norm downdating with
// recomputation below tol3z = sqrt(eps/2)), every sum in the sumsq16 / dot16 order.
// jpvt is returned 0-based. tau has min(m,n) entries. src/util.cpp:381-394.
void geqp3(int m, int n, double* A, int lda, int* jpvt, double* tau);

// dormqr('L','T') / ('L','N') applied to one vector x of length m with k reflectors
// stored below the diagonal of V (m x k, ld ldv). src/util.cpp:328-350.
void ormqr_vec(bool trans, int m, int k, const double* V, int ldv, const double* tau,
               double* x);

// dgetrf with partial pivoting (dgetf2 order): ipiv is 0-based row swaps; returns info.
// src/util.cpp:183-200.
int getrf(int n, double* A, int lda, int* ipiv);

// swap2perm: src/util.cpp:76-88.
void swap2perm(int n, const int* swap, int* perm);

// choose_rank: src/util.cpp:434-452.
int choose_rank(const double* diag, int n, double tol);

}  // namespace oracle
