// ORACLE — test infrastructure only (see dense.hpp). FP64 reference-BLAS/LAPACK
// restatements used by the CPU oracle of spaND's factorization.
#include "dense.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace oracle {

#define IDX(A, i, j, ld) (A)[(size_t)(i) + (size_t)(j) * (size_t)(ld)]

double nrm2(int n, const double* x, int incx) {
    double s = 0.0;
    for (int i = 0; i < n; i++) {
        double a = x[(size_t)i * incx];
        s = std::fma(a, a, s);
    }
    return std::sqrt(s);
}

double tree16(double c[kChains]) {
    for (int w = 0; w < 8; w++) c[w] = c[w] + c[w + 8];
    for (int w = 0; w < 4; w++) c[w] = c[w] + c[w + 4];
    for (int w = 0; w < 2; w++) c[w] = c[w] + c[w + 2];
    return c[0] + c[1];
}

double sumsq16(int n, const double* x, int incx) {
    double c[kChains] = {0.0};
    for (int t = 0; t < n; t++) {
        const double a = x[(size_t)t * incx];
        c[t % kChains] = std::fma(a, a, c[t % kChains]);
    }
    return tree16(c);
}

double dot16(int n, const double* a, const double* b) {
    double c[kChains] = {0.0};
    for (int t = 0; t < n; t++) c[t % kChains] = std::fma(a[t], b[t], c[t % kChains]);
    return tree16(c);
}

// Right-looking column Cholesky of the lower triangle with explicit fused multiply-adds:
// column j is scaled by 1/L(j,j), then every trailing lower element receives
// A(a,b) = fma(-L(a,j), L(b,j), A(a,b)). Each element therefore sees its updates in the
// same k order as dpotf2; the GPU kernel uses the same element recurrence.
int potrf_lower(int n, double* A, int lda) {
    for (int j = 0; j < n; j++) {
        double ajj = IDX(A, j, j, lda);
        if (!(ajj > 0.0)) return j + 1;
        ajj = std::sqrt(ajj);
        IDX(A, j, j, lda) = ajj;
        double r = 1.0 / ajj;
        for (int i = j + 1; i < n; i++) IDX(A, i, j, lda) *= r;
        for (int b = j + 1; b < n; b++) {
            double lb = IDX(A, b, j, lda);
            for (int a = b; a < n; a++) IDX(A, a, b, lda) = std::fma(-IDX(A, a, j, lda), lb, IDX(A, a, b, lda));
        }
    }
    return 0;
}

void trsm_left_lower_notrans(int m, int n, const double* L, int ldl, double* B, int ldb) {
    if (m == 0 || n == 0) return;
    for (int j = 0; j < n; j++) {
        double* b = B + (size_t)j * ldb;
        for (int k = 0; k < m; k++) {
            b[k] /= IDX(L, k, k, ldl);
            double t = b[k];
            for (int i = k + 1; i < m; i++) b[i] = std::fma(-t, IDX(L, i, k, ldl), b[i]);
        }
    }
}

void trsm_right_lower_trans(int m, int n, const double* L, int ldl, double* B, int ldb) {
    if (m == 0 || n == 0) return;
    for (int k = 0; k < n; k++) {
        double t = 1.0 / IDX(L, k, k, ldl);
        double* bk = B + (size_t)k * ldb;
        for (int i = 0; i < m; i++) bk[i] *= t;
        for (int j = k + 1; j < n; j++) {
            double a = IDX(L, j, k, ldl);
            double* bj = B + (size_t)j * ldb;
            for (int i = 0; i < m; i++) bj[i] = std::fma(-a, bk[i], bj[i]);
        }
    }
}

// B := B U^{-1}: column j receives fma(-U(k,j), b_k, b_j) for k < j in increasing k, then
// the scaling by 1 / U(j,j) -- the element recurrence of trsm_right_lower_trans with
// L = U^T, which is how the GPU applies it.
void trsm_right_upper_notrans(int m, int n, const double* U, int ldu, double* B, int ldb) {
    if (m == 0 || n == 0) return;
    for (int j = 0; j < n; j++) {
        double* bj = B + (size_t)j * ldb;
        for (int k = 0; k < j; k++) {
            const double a = IDX(U, k, j, ldu);
            const double* bk = B + (size_t)k * ldb;
            for (int i = 0; i < m; i++) bj[i] = std::fma(-a, bk[i], bj[i]);
        }
        double t = 1.0 / IDX(U, j, j, ldu);
        for (int i = 0; i < m; i++) bj[i] *= t;
    }
}

void syrk_lower(bool trans, int n, int k, double alpha, const double* A, int lda,
                double beta, double* C, int ldc) {
    if (n == 0) return;
    if (!trans) {
        for (int j = 0; j < n; j++) {
            if (beta != 1.0)
                for (int i = j; i < n; i++) IDX(C, i, j, ldc) *= beta;
            for (int l = 0; l < k; l++) {
                double a = IDX(A, j, l, lda);
                if (a != 0.0) {
                    double t = alpha * a;
                    for (int i = j; i < n; i++) IDX(C, i, j, ldc) += t * IDX(A, i, l, lda);
                }
            }
        }
    } else {
        for (int j = 0; j < n; j++) {
            for (int i = j; i < n; i++) {
                double t = 0.0;
                for (int l = 0; l < k; l++) t += IDX(A, l, i, lda) * IDX(A, l, j, lda);
                IDX(C, i, j, ldc) = (beta == 0.0 ? 0.0 : beta * IDX(C, i, j, ldc)) + alpha * t;
            }
        }
    }
}

void gemm(bool tA, bool tB, int m, int n, int k, double alpha, const double* A, int lda,
          const double* B, int ldb, double beta, double* C, int ldc) {
    if (m == 0 || n == 0 || k == 0) return;
    for (int j = 0; j < n; j++) {
        double* c = C + (size_t)j * ldc;
        if (beta != 1.0)
            for (int i = 0; i < m; i++) c[i] = (beta == 0.0 ? 0.0 : beta * c[i]);
        if (!tA) {
            for (int l = 0; l < k; l++) {
                double b = tB ? IDX(B, j, l, ldb) : IDX(B, l, j, ldb);
                if (b != 0.0) {
                    double t = alpha * b;
                    const double* a = A + (size_t)l * lda;
                    for (int i = 0; i < m; i++) c[i] += t * a[i];
                }
            }
        } else {
            for (int i = 0; i < m; i++) {
                const double* a = A + (size_t)i * lda;
                double t = 0.0;
                if (!tB)
                    for (int l = 0; l < k; l++) t += a[l] * IDX(B, l, j, ldb);
                else
                    for (int l = 0; l < k; l++) t += a[l] * IDX(B, j, l, ldb);
                c[i] += alpha * t;
            }
        }
    }
}

void schur_update(bool tA, bool tB, bool lower, int m, int n, int k, const double* A, int lda,
                  const double* B, int ldb, double* C, int ldc) {
    for (int j = 0; j < n; j++)
        for (int i = lower ? j : 0; i < m; i++) {
            double c = IDX(C, i, j, ldc);
            for (int l = 0; l < k; l++) {
                double a = tA ? IDX(A, l, i, lda) : IDX(A, i, l, lda);
                double b = tB ? IDX(B, j, l, ldb) : IDX(B, l, j, ldb);
                c = std::fma(-a, b, c);
            }
            IDX(C, i, j, ldc) = c;
        }
}

void trsv_lower_notrans(int n, const double* A, int lda, double* x) {
    for (int j = 0; j < n; j++) {
        x[j] /= IDX(A, j, j, lda);
        double t = x[j];
        for (int i = j + 1; i < n; i++) x[i] = std::fma(-t, IDX(A, i, j, lda), x[i]);
    }
}

// x := L^{-T} x as the column sweep of U = L^T: x_j /= L(j,j), then x_i -= x_j L(j,i), i < j.
void trsv_lower_trans(int n, const double* A, int lda, double* x) {
    for (int j = n - 1; j >= 0; j--) {
        x[j] /= IDX(A, j, j, lda);
        double t = x[j];
        for (int i = 0; i < j; i++) x[i] = std::fma(-t, IDX(A, j, i, lda), x[i]);
    }
}

void trsv_upper_notrans(int n, const double* A, int lda, double* x) {
    for (int j = n - 1; j >= 0; j--) {
        if (x[j] != 0.0) {
            x[j] /= IDX(A, j, j, lda);
            double t = x[j];
            for (int i = j - 1; i >= 0; i--) x[i] -= t * IDX(A, i, j, lda);
        }
    }
}

void gemv_sub_notrans(int m, int n, const double* A, int lda, const double* x, double* y) {
    if (m == 0 || n == 0) return;
    for (int i = 0; i < m; i++) {
        double acc = y[i];
        for (int j = 0; j < n; j++) acc = std::fma(-IDX(A, i, j, lda), x[j], acc);
        y[i] = acc;
    }
}

void gemv_sub_trans(int m, int n, const double* A, int lda, const double* x, double* y) {
    if (m == 0 || n == 0) return;
    for (int j = 0; j < n; j++) {
        const double* a = A + (size_t)j * lda;
        double t = 0.0;
        for (int i = 0; i < m; i++) t = std::fma(a[i], x[i], t);
        y[j] -= t;
    }
}

static double lapy2(double x, double y) {
    double xa = std::fabs(x), ya = std::fabs(y);
    double w = std::max(xa, ya), z = std::min(xa, ya);
    if (z == 0.0) return w;
    double q = z / w;
    return w * std::sqrt(1.0 + q * q);
}

void larfg(int n, double* alpha, double* x, int incx, double* tau) {
    if (n <= 1) {
        *tau = 0.0;
        return;
    }
    double xnorm = std::sqrt(sumsq16(n - 1, x, incx));
    if (xnorm == 0.0) {
        *tau = 0.0;
        return;
    }
    double beta = -std::copysign(lapy2(*alpha, xnorm), *alpha);
    const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
    int knt = 0;
    if (std::fabs(beta) < safmin) {
        const double rsafmn = 1.0 / safmin;
        do {
            knt++;
            for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] *= rsafmn;
            beta *= rsafmn;
            *alpha *= rsafmn;
        } while (std::fabs(beta) < safmin && knt < 20);
        xnorm = std::sqrt(sumsq16(n - 1, x, incx));
        beta = -std::copysign(lapy2(*alpha, xnorm), *alpha);
    }
    *tau = (beta - *alpha) / beta;
    double scal = 1.0 / (*alpha - beta);
    for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] *= scal;
    for (int j = 0; j < knt; j++) beta *= safmin;
    *alpha = beta;
}

void larf_left(int m, int n, const double* v, double tau, double* C, int ldc, double* work) {
    if (tau == 0.0 || m == 0 || n == 0) return;
    // w := C^T v  (dgemv Trans, beta = 0)
    for (int j = 0; j < n; j++) work[j] = dot16(m, C + (size_t)j * ldc, v);
    // C := C - tau v w^T  (dger, alpha = -tau; reference dger skips columns with w_j = 0)
    for (int j = 0; j < n; j++) {
        if (work[j] != 0.0) {
            double t = -tau * work[j];
            double* c = C + (size_t)j * ldc;
            for (int i = 0; i < m; i++) c[i] = std::fma(v[i], t, c[i]);
        }
    }
}

void geqp3(int m, int n, double* A, int lda, int* jpvt, double* tau) {
    if (m == 0 || n == 0) return;
    const int mn = std::min(m, n);
    const double tol3z = std::sqrt(DBL_EPSILON * 0.5);
    std::vector<double> vn1(n), vn2(n), work(n), vcol(m);
    for (int j = 0; j < n; j++) {
        jpvt[j] = j;
        vn1[j] = std::sqrt(sumsq16(m, A + (size_t)j * lda, 1));
        vn2[j] = vn1[j];
    }
    for (int i = 0; i < mn; i++) {
        // pivot: first index of max vn1[i:n]
        int pvt = i;
        double best = std::fabs(vn1[i]);
        for (int j = i + 1; j < n; j++) {
            if (std::fabs(vn1[j]) > best) {
                best = std::fabs(vn1[j]);
                pvt = j;
            }
        }
        if (pvt != i) {
            for (int r = 0; r < m; r++) std::swap(IDX(A, r, pvt, lda), IDX(A, r, i, lda));
            std::swap(jpvt[pvt], jpvt[i]);
            vn1[pvt] = vn1[i];
            vn2[pvt] = vn2[i];
        }
        double* aii = &IDX(A, i, i, lda);
        if (i < m - 1)
            larfg(m - i, aii, aii + 1, 1, &tau[i]);
        else
            larfg(1, aii, aii, 1, &tau[i]);
        if (i < n - 1) {
            double save = *aii;
            *aii = 1.0;
            larf_left(m - i, n - i - 1, aii, tau[i], &IDX(A, i, i + 1, lda), lda, work.data());
            *aii = save;
        }
        for (int j = i + 1; j < n; j++) {
            if (vn1[j] != 0.0) {
                double q = std::fabs(IDX(A, i, j, lda)) / vn1[j];
                double temp = 1.0 - q * q;
                temp = std::max(temp, 0.0);
                double r = vn1[j] / vn2[j];
                double temp2 = temp * r * r;
                if (temp2 <= tol3z) {
                    if (i < m - 1) {
                        vn1[j] = std::sqrt(sumsq16(m - i - 1, &IDX(A, i + 1, j, lda), 1));
                        vn2[j] = vn1[j];
                    } else {
                        vn1[j] = 0.0;
                        vn2[j] = 0.0;
                    }
                } else {
                    vn1[j] = vn1[j] * std::sqrt(temp);
                }
            }
        }
    }
}

void ormqr_vec(bool trans, int m, int k, const double* V, int ldv, const double* tau,
               double* x) {
    if (m == 0 || k == 0) return;
    auto apply = [&](int i) {
        // H(i) = I - tau v v^T, v = [1; V(i+1:m, i)], acting on x(i:m)
        if (tau[i] == 0.0) return;
        const double* v = &IDX(V, i, i, ldv);
        double w = x[i];
        for (int r = i + 1; r < m; r++) w = std::fma(x[r], v[r - i], w);
        double t = -tau[i] * w;
        x[i] += t;
        for (int r = i + 1; r < m; r++) x[r] = std::fma(v[r - i], t, x[r]);
    };
    if (trans)
        for (int i = 0; i < k; i++) apply(i);
    else
        for (int i = k - 1; i >= 0; i--) apply(i);
}

int getrf(int n, double* A, int lda, int* ipiv) {
    int info = 0;
    for (int j = 0; j < n; j++) {
        int jp = j;
        double best = std::fabs(IDX(A, j, j, lda));
        for (int i = j + 1; i < n; i++) {
            double v = std::fabs(IDX(A, i, j, lda));
            if (v > best) {
                best = v;
                jp = i;
            }
        }
        ipiv[j] = jp;
        if (IDX(A, jp, j, lda) != 0.0) {
            if (jp != j)
                for (int c = 0; c < n; c++) std::swap(IDX(A, j, c, lda), IDX(A, jp, c, lda));
            double piv = IDX(A, j, j, lda);
            if (std::fabs(piv) >= DBL_MIN) {
                double r = 1.0 / piv;
                for (int i = j + 1; i < n; i++) IDX(A, i, j, lda) *= r;
            } else {
                for (int i = j + 1; i < n; i++) IDX(A, i, j, lda) /= piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        // rank-1 update of the trailing block (dger, alpha = -1)
        for (int c = j + 1; c < n; c++) {
            double y = IDX(A, j, c, lda);
            if (y != 0.0) {
                double t = -y;
                for (int i = j + 1; i < n; i++) IDX(A, i, c, lda) += IDX(A, i, j, lda) * t;
            }
        }
    }
    return info;
}

void swap2perm(int n, const int* swap, int* perm) {
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = 0; i < n; i++) {
        int ipiv = swap[i];
        std::swap(perm[ipiv], perm[i]);
    }
}

int choose_rank(const double* s, int n, double tol) {
    if (tol == 0.0) return n;
    if (tol >= 1.0) return 0;
    if (n <= 1) return n;
    double sref = std::fabs(s[0]);
    int rank = 1;
    while (rank < n && std::fabs(s[rank]) / sref >= tol) rank++;
    return rank;
}

}  // namespace oracle
