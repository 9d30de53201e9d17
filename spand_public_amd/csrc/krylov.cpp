// Device FP64 preconditioned CG on a block of right-hand sides (cg, src/is.cpp:39-121).
// Every column follows the reference recurrence exactly and stops on its own recursive
// residual; columns run in lockstep so one preconditioner apply serves all of them.
#include "krylov.hpp"

#include <cfloat>
#include <chrono>
#include <cmath>
#include <vector>

#include "kernels.hpp"

namespace spand {

static double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

DevMatrix::DevMatrix(int n, const int* colptr, const int* rowval, const double* vals) : N(n) {
    nnz = colptr[n];
    // CSR of A = CSC of A^T: transpose the CSC arrays
    std::vector<int> rp(n + 1, 0), ci(nnz), co(nnz);
    std::vector<double> v(nnz);
    for (long long k = 0; k < nnz; k++) rp[rowval[k] + 1]++;
    for (int i = 0; i < n; i++) rp[i + 1] += rp[i];
    std::vector<int> fill(rp.begin(), rp.end() - 1);
    for (int j = 0; j < n; j++)
        for (int k = colptr[j]; k < colptr[j + 1]; k++) {
            int p = fill[rowval[k]]++;
            ci[p] = j;
            v[p] = vals[k];
            co[p] = k;
        }
    hip_check(hipMalloc((void**)&rowptr, (n + 1) * sizeof(int)), "hipMalloc(csr)");
    hip_check(hipMalloc((void**)&colidx, std::max<long long>(nnz, 1) * sizeof(int)), "hipMalloc(csr)");
    hip_check(hipMalloc((void**)&val, std::max<long long>(nnz, 1) * sizeof(double)), "hipMalloc(csr)");
    hip_check(hipMalloc((void**)&csc_of, std::max<long long>(nnz, 1) * sizeof(int)), "hipMalloc(csr)");
    hip_check(hipMalloc((void**)&tptr, (n + 1) * sizeof(int)), "hipMalloc(csc)");
    hip_check(hipMalloc((void**)&tidx, std::max<long long>(nnz, 1) * sizeof(int)), "hipMalloc(csc)");
    hip_check(hipMalloc((void**)&tval, std::max<long long>(nnz, 1) * sizeof(double)), "hipMalloc(csc)");
    hip_check(hipMemcpy(rowptr, rp.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice), "upload csr");
    hip_check(hipMemcpy(tptr, colptr, (n + 1) * sizeof(int), hipMemcpyHostToDevice), "upload csc");
    if (nnz) {
        hip_check(hipMemcpy(colidx, ci.data(), nnz * sizeof(int), hipMemcpyHostToDevice), "upload csr");
        hip_check(hipMemcpy(val, v.data(), nnz * sizeof(double), hipMemcpyHostToDevice), "upload csr");
        hip_check(hipMemcpy(csc_of, co.data(), nnz * sizeof(int), hipMemcpyHostToDevice), "upload csr");
        hip_check(hipMemcpy(tidx, rowval, nnz * sizeof(int), hipMemcpyHostToDevice), "upload csc");
        hip_check(hipMemcpy(tval, vals, nnz * sizeof(double), hipMemcpyHostToDevice), "upload csc");
    }
}

void DevMatrix::set_values_device(const double* vals, hipStream_t st) {
    if (nnz == 0) return;
    if (!vals) throw Error(ERR_ARG, "set_values: null values");
    hip_check(launch_gather_f64(csc_of, vals, nnz, val, st), "gather values");
    hip_check(launch_copy_f64(tval, vals, (size_t)nnz, st), "copy values");
    hip_check(hipStreamSynchronize(st), "set values");
}

void DevMatrix::sampled_outer(const double* u, const double* v, int nrhs, int ld, double alpha, double* vals_out,
                              bool on_device) const {
    if (nrhs <= 0 || ld < N) throw Error(ERR_ARG, "sampled_outer: need nrhs >= 1 and ld >= N");
    if (on_device) {
        hip_check(launch_sampled_outer(N, rowptr, colidx, csc_of, u, v, nrhs, ld, alpha, vals_out, 0),
                  "sampled outer");
        hip_check(hipStreamSynchronize(0), "sampled outer");
        return;
    }
    const size_t vb = (size_t)N * nrhs * sizeof(double);
    double *du = nullptr, *dv = nullptr, *dout = nullptr;
    hip_check(hipMalloc((void**)&du, vb), "hipMalloc(sampled outer)");
    hip_check(hipMalloc((void**)&dv, vb), "hipMalloc(sampled outer)");
    hip_check(hipMalloc((void**)&dout, std::max<long long>(nnz, 1) * sizeof(double)), "hipMalloc(sampled outer)");
    hip_check(hipMemcpy2D(du, (size_t)N * sizeof(double), u, (size_t)ld * sizeof(double), (size_t)N * sizeof(double),
                          nrhs, hipMemcpyHostToDevice),
              "upload u");
    hip_check(hipMemcpy2D(dv, (size_t)N * sizeof(double), v, (size_t)ld * sizeof(double), (size_t)N * sizeof(double),
                          nrhs, hipMemcpyHostToDevice),
              "upload v");
    hip_check(launch_sampled_outer(N, rowptr, colidx, csc_of, du, dv, nrhs, N, alpha, dout, 0), "sampled outer");
    hip_check(hipMemcpy(vals_out, dout, nnz * sizeof(double), hipMemcpyDeviceToHost), "download");
    (void)hipFree(du);
    (void)hipFree(dv);
    (void)hipFree(dout);
}

void* DevMatrix::workspace(size_t bytes) const {
    if (bytes > ws.bytes) {
        (void)hipFree(ws.p);
        ws.p = nullptr;
        ws.bytes = 0;
        hip_check(hipMalloc(&ws.p, bytes), "hipMalloc(krylov workspace)");
        ws.bytes = bytes;
    }
    return ws.p;
}

DevMatrix::~DevMatrix() {
    (void)hipFree(ws.p);
    (void)hipFree(rowptr);
    (void)hipFree(colidx);
    (void)hipFree(val);
    (void)hipFree(csc_of);
    (void)hipFree(tptr);
    (void)hipFree(tidx);
    (void)hipFree(tval);
}

KrylovResult pcg(Tree& t, const DevMatrix& A, const double* b, double* x, int nrhs, int ld, int maxit, double tol,
                 bool on_device) {
    if (A.N != t.N()) throw Error(ERR_ARG, "pcg: matrix and tree sizes differ");
    if (nrhs <= 0 || nrhs > 64 || ld < A.N) throw Error(ERR_ARG, "pcg: bad nrhs / ld");
    const double t0 = wall();
    const int N = A.N;
    const size_t nv = (size_t)N * nrhs;
    const hipStream_t st = 0;
    KrylovResult res;
    res.iters.assign(nrhs, 0);
    res.relres.assign(nrhs, 0.0);
    // interleaved work vectors (b, x, r, p, z, tmp, io), dot partials, scalars, active flags
    const size_t n_vec = 7 * nv, n_part = (size_t)1024 * nrhs, n_sc = (size_t)8 * nrhs;
    char* wsb = static_cast<char*>(A.workspace((n_vec + n_part + n_sc) * sizeof(double) + nrhs * sizeof(int)));
    double* base = reinterpret_cast<double*>(wsb);
    double *db = base, *dx = base + nv, *r = base + 2 * nv, *p = base + 3 * nv, *z = base + 4 * nv;
    double *tmp = base + 5 * nv, *io = base + 6 * nv;
    double* partial = base + n_vec;
    double* sc = partial + n_part;
    int* dact = reinterpret_cast<int*>(sc + n_sc);
    double* bn2 = sc;
    double* rn2 = sc + nrhs;
    double* absNew = sc + 2 * nrhs;
    double* absOld = sc + 3 * nrhs;
    double* pAp = sc + 4 * nrhs;
    double* alpha = sc + 5 * nrhs;
    double* beta = sc + 6 * nrhs;
    KernelProfiler& prof = t.profiler();
    const double vbytes = (double)nv * sizeof(double);
    auto vec = [&](double passes, const char* what, auto&& launch) {
        hip_check(timed(prof, KF_KRYLOV_VEC, st, passes * vbytes, 0, launch), what);
    };
    // column-major (ld) <-> interleaved, on the device
    auto to_inter = [&](const double* src, double* dst) {
        if (on_device) {
            vec(2, "cm2il", [&] { return launch_cm2il(src, ld, N, nrhs, dst, st); });
        } else {
            hip_check(hipMemcpy2D(io, (size_t)N * sizeof(double), src, (size_t)ld * sizeof(double),
                                  (size_t)N * sizeof(double), nrhs, hipMemcpyHostToDevice),
                      "pcg io");
            vec(2, "cm2il", [&] { return launch_cm2il(io, N, N, nrhs, dst, st); });
        }
    };
    to_inter(b, db);
    to_inter(x, dx);
    std::vector<double> h_bn2(nrhs), h_rn2(nrhs), thr(nrhs);
    std::vector<int> act(nrhs, 1);
    double t_prec = 0, t_spmv = 0;
    int applies = 0;
    auto spmv = [&](const double* in, double* out) {
        double a = wall();
        hip_check(timed(prof, KF_SPMV, st, 12.0 * A.nnz + 4.0 * (N + 1) + 2.0 * vbytes, 2.0 * A.nnz * nrhs,
                        [&] { return launch_spmv(N, A.rowptr, A.colidx, A.val, in, out, nrhs, st); }),
                  "spmv");
        t_spmv += wall() - a;
    };
    auto precond = [&](double* v) {
        double a = wall();
        t.solve_interleaved(v, nrhs, st);
        applies++;
        t_prec += wall() - a;
    };
    auto dot = [&](const double* a, const double* c, double* out) {
        vec(2, "dot", [&] { return launch_dot(N, nrhs, a, c, partial, out, st); });
    };
    // r = b - A x
    spmv(dx, tmp);
    vec(2, "copy", [&] { return launch_copy_f64(r, db, nv, st); });
    vec(3, "axpby", [&] { return launch_axpby(N, nrhs, tmp, nullptr, -1.0, r, nullptr, st); });
    dot(db, db, bn2);
    dot(r, r, rn2);
    hip_check(hipMemcpy(h_bn2.data(), bn2, nrhs * sizeof(double), hipMemcpyDeviceToHost), "pcg");
    hip_check(hipMemcpy(h_rn2.data(), rn2, nrhs * sizeof(double), hipMemcpyDeviceToHost), "pcg");
    int nact = 0;
    for (int c = 0; c < nrhs; c++) {
        thr[c] = tol * tol * h_bn2[c];
        if (h_bn2[c] == 0.0 || h_rn2[c] < thr[c]) {
            act[c] = 0;
            res.iters[c] = 0;
        } else {
            nact++;
        }
    }
    // columns with b == 0 get x = 0 (src/is.cpp:58-64)
    bool any_zero_b = false;
    for (int c = 0; c < nrhs; c++) any_zero_b |= h_bn2[c] == 0.0;
    if (any_zero_b) {
        std::vector<double> hx(nv);
        hip_check(hipMemcpy(hx.data(), dx, nv * sizeof(double), hipMemcpyDeviceToHost), "pcg");
        for (int c = 0; c < nrhs; c++)
            if (h_bn2[c] == 0.0)
                for (int i = 0; i < N; i++) hx[(size_t)i * nrhs + c] = 0.0;
        hip_check(hipMemcpy(dx, hx.data(), nv * sizeof(double), hipMemcpyHostToDevice), "pcg");
    }
    hip_check(hipMemcpy(dact, act.data(), nrhs * sizeof(int), hipMemcpyHostToDevice), "pcg");
    if (nact > 0) {
        vec(2, "copy", [&] { return launch_copy_f64(p, r, nv, st); });
        precond(p);
        dot(r, p, absNew);
        int i = 0;
        while (i < maxit) {
            spmv(p, tmp);
            dot(p, tmp, pAp);
            hip_check(launch_ratio(nrhs, absNew, pAp, dact, alpha, st), "ratio");
            vec(3, "axpby", [&] { return launch_axpby(N, nrhs, p, alpha, 1.0, dx, nullptr, st); });
            vec(3, "axpby", [&] { return launch_axpby(N, nrhs, tmp, alpha, -1.0, r, nullptr, st); });
            dot(r, r, rn2);
            hip_check(hipMemcpy(h_rn2.data(), rn2, nrhs * sizeof(double), hipMemcpyDeviceToHost), "pcg");
            bool changed = false;
            for (int c = 0; c < nrhs; c++)
                if (act[c] && h_rn2[c] < thr[c]) {
                    act[c] = 0;
                    res.iters[c] = i + 1;
                    nact--;
                    changed = true;
                }
            if (nact == 0) break;
            if (changed) hip_check(hipMemcpy(dact, act.data(), nrhs * sizeof(int), hipMemcpyHostToDevice), "pcg");
            vec(2, "copy", [&] { return launch_copy_f64(z, r, nv, st); });
            precond(z);
            hip_check(launch_copy_f64(absOld, absNew, nrhs, st), "copy");
            dot(r, z, absNew);
            hip_check(launch_ratio(nrhs, absNew, absOld, dact, beta, st), "ratio");
            // p = z + beta p
            vec(3, "axpby", [&] { return launch_axpby(N, nrhs, z, nullptr, 1.0, p, beta, st); });
            i++;
        }
        for (int c = 0; c < nrhs; c++)
            if (act[c]) res.iters[c] = i + 1;
    }
    // true residual
    spmv(dx, tmp);
    vec(2, "copy", [&] { return launch_copy_f64(r, db, nv, st); });
    vec(3, "axpby", [&] { return launch_axpby(N, nrhs, tmp, nullptr, -1.0, r, nullptr, st); });
    dot(r, r, rn2);
    hip_check(hipMemcpy(h_rn2.data(), rn2, nrhs * sizeof(double), hipMemcpyDeviceToHost), "pcg");
    res.converged = 0;
    for (int c = 0; c < nrhs; c++) {
        res.relres[c] = h_bn2[c] > 0 ? std::sqrt(h_rn2[c] / h_bn2[c]) : 0.0;
        if (!act[c]) res.converged++;
    }
    // interleaved -> column-major x
    if (on_device) {
        hip_check(launch_il2cm(dx, N, nrhs, x, ld, st), "il2cm");
        hip_check(hipStreamSynchronize(st), "pcg");
    } else {
        hip_check(launch_il2cm(dx, N, nrhs, io, N, st), "il2cm");
        hip_check(hipMemcpy2D(x, (size_t)ld * sizeof(double), io, (size_t)N * sizeof(double),
                              (size_t)N * sizeof(double), nrhs, hipMemcpyDeviceToHost),
                  "pcg io");
    }
    prof.collect();
    res.seconds = wall() - t0;
    res.seconds_precond = t_prec;
    res.seconds_spmv = t_spmv;
    res.applies = applies;
    return res;
}

// ---------------------------------------------------------------------------------------
// gmres (src/is.cpp:123-300) on nrhs columns in lockstep, FP64 on the device.
//
// The reference builds the Arnoldi basis with Householder reflections H_i = I - tau_i u_i
// u_i^T (Eigen's makeHouseholder / applyHouseholderOnTheLeft) and applies them one at a
// time. Here the reflections of a cycle are kept as interleaved device blocks V and applied
// together in compact-WY form, Q_k = H_0 ... H_{k-1} = I - V T V^T (T upper triangular,
// k x k, on the host): one multi-dot and one combination kernel per application instead of
// k dot/axpy pairs. The Hessenberg matrix, the Givens rotations and w live on the host (a
// few hundred numbers per column). A column stops when |w_k| / ||M^{-1}(b - A x0)|| < tol,
// with the initial preconditioned residual norm kept across restarts (src/is.cpp:161,245).
namespace {
struct Givens {
    double c = 1.0, s = 0.0;
};
Givens make_givens(double p, double q) {  // JacobiRotation::makeGivens, real case
    Givens g;
    if (q == 0.0) {
        g.c = p < 0.0 ? -1.0 : 1.0;
    } else if (p == 0.0) {
        g.c = 0.0;
        g.s = q < 0.0 ? 1.0 : -1.0;
    } else if (std::fabs(p) > std::fabs(q)) {
        const double t = q / p;
        double u = std::sqrt(1.0 + t * t);
        if (p < 0.0) u = -u;
        g.c = 1.0 / u;
        g.s = -t * g.c;
    } else {
        const double t = p / q;
        double u = std::sqrt(1.0 + t * t);
        if (q < 0.0) u = -u;
        g.s = -1.0 / u;
        g.c = -t * g.s;
    }
    return g;
}
void apply_givens_adjoint(const Givens& g, double& x, double& y) {  // applyOnTheLeft(i, i+1, G.adjoint())
    const double x0 = x, y0 = y;
    x = g.c * x0 - g.s * y0;
    y = g.s * x0 + g.c * y0;
}
}  // namespace

KrylovResult gmres(Tree& t, const DevMatrix& A, const double* b, double* x, int nrhs, int ld, int maxit, int restart,
                   double tol, bool on_device, bool transpose) {
    if (A.N != t.N()) throw Error(ERR_ARG, "gmres: matrix and tree sizes differ");
    if (nrhs <= 0 || nrhs > 64 || ld < A.N) throw Error(ERR_ARG, "gmres: bad nrhs / ld");
    if (restart < 1) throw Error(ERR_ARG, "gmres: restart must be positive");
    const double t0 = wall();
    const int N = A.N, R = std::min(restart, N);
    const size_t nv = (size_t)N * nrhs;
    const hipStream_t st = 0;
    KrylovResult res;
    res.iters.assign(nrhs, 0);
    res.relres.assign(nrhs, 0.0);
    // device: b, x, v, tmp, io, V[R + 1], multidot partials, small per-column arrays
    const size_t n_small = (size_t)(R + 2) * nrhs;
    const size_t n_part = (size_t)256 * (R + 1) * nrhs;
    char* wsb = static_cast<char*>(
        A.workspace(((5 + (size_t)(R + 1)) * nv + n_part + 4 * n_small + (size_t)1024 * nrhs + 8 * nrhs) *
                    sizeof(double)));
    double* base = reinterpret_cast<double*>(wsb);
    double *db = base, *dx = base + nv, *v = base + 2 * nv, *tmp = base + 3 * nv, *io = base + 4 * nv;
    double* V = base + 5 * nv;
    double* part = V + (size_t)(R + 1) * nv;
    double* dz = part + n_part;
    double* dye = dz + n_small;
    double* dsm = dye + n_small;
    double* dd = dsm + n_small;
    double* dpart = dd + n_small;
    double* dsc = dpart + (size_t)1024 * nrhs;
    KernelProfiler& prof = t.profiler();
    const double vbytes = (double)nv * sizeof(double);
    auto vec = [&](double passes, const char* what, auto&& launch) {
        hip_check(timed(prof, KF_KRYLOV_VEC, st, passes * vbytes, 0, launch), what);
    };
    double t_prec = 0, t_spmv = 0;
    int applies = 0;
    auto spmv = [&](const double* in, double* out) {
        double a = wall();
        const int* rp = transpose ? A.tptr : A.rowptr;
        const int* ci = transpose ? A.tidx : A.colidx;
        const double* va = transpose ? A.tval : A.val;
        hip_check(timed(prof, KF_SPMV, st, 12.0 * A.nnz + 4.0 * (N + 1) + 2.0 * vbytes, 2.0 * A.nnz * nrhs,
                        [&] { return launch_spmv(N, rp, ci, va, in, out, nrhs, st); }),
                  "spmv");
        t_spmv += wall() - a;
    };
    auto precond = [&](double* w) {
        double a = wall();
        t.solve_interleaved(w, nrhs, st, transpose);
        applies++;
        t_prec += wall() - a;
    };
    std::vector<double> hsc(nrhs);
    auto dot_host = [&](const double* a, const double* c, int n) {  // per-column dot over n rows
        vec(2, "dot", [&] { return launch_dot(n, nrhs, a, c, dpart, dsc, st); });
        hip_check(hipMemcpy(hsc.data(), dsc, nrhs * sizeof(double), hipMemcpyDeviceToHost), "gmres");
        return hsc;
    };
    auto to_inter = [&](const double* src, double* dst) {
        if (on_device) {
            vec(2, "cm2il", [&] { return launch_cm2il(src, ld, N, nrhs, dst, st); });
        } else {
            hip_check(hipMemcpy2D(io, (size_t)N * sizeof(double), src, (size_t)ld * sizeof(double),
                                  (size_t)N * sizeof(double), nrhs, hipMemcpyHostToDevice),
                      "gmres io");
            vec(2, "cm2il", [&] { return launch_cm2il(io, N, N, nrhs, dst, st); });
        }
    };
    to_inter(b, db);
    to_inter(x, dx);
    // per-column host state; column-major small matrices with leading dimension R + 1
    const int L = R + 1;
    std::vector<double> H((size_t)nrhs * L * L), T((size_t)nrhs * L * L), vtop((size_t)nrhs * L * L);
    std::vector<double> w((size_t)nrhs * (L + 1)), tau((size_t)nrhs * L), r0norm(nrhs, 0.0), rel(nrhs, 0.0);
    std::vector<Givens> G((size_t)nrhs * R);
    std::vector<int> act(nrhs, 1);
    auto h_at = [&](std::vector<double>& M, int c, int i, int j) -> double& { return M[(size_t)c * L * L + i + (size_t)j * L]; };
    std::vector<double> hz(n_small), hye(n_small), hrows((size_t)L * nrhs), hd(nrhs);
    // columns with b = 0 get x = 0 (src/is.cpp:140-144)
    {
        std::vector<double> bn2 = dot_host(db, db, N);
        bool any = false;
        for (int c = 0; c < nrhs; c++)
            if (std::sqrt(bn2[c]) <= DBL_MIN) {
                act[c] = 0;
                any = true;
            }
        if (any) {
            std::vector<double> hx(nv);
            hip_check(hipMemcpy(hx.data(), dx, nv * sizeof(double), hipMemcpyDeviceToHost), "gmres");
            for (int c = 0; c < nrhs; c++)
                if (!act[c])
                    for (int i = 0; i < N; i++) hx[(size_t)i * nrhs + c] = 0.0;
            hip_check(hipMemcpy(dx, hx.data(), nv * sizeof(double), hipMemcpyHostToDevice), "gmres");
        }
    }
    // rows 0..R of an interleaved vector
    auto read_rows = [&](const double* dv) {
        hip_check(hipMemcpy(hrows.data(), dv, (size_t)L * nrhs * sizeof(double), hipMemcpyDeviceToHost), "gmres");
    };
    // makeHouseholder of x[k:] from x[k] and ||x[k+1:]||^2; writes u_k = V_k and vtop_k
    auto householder = [&](int k, const double* dv, const std::vector<double>& tail_sq, std::vector<double>& beta_out,
                           const std::vector<int>& make) {
        read_rows(dv);
        for (int c = 0; c < nrhs; c++) {
            const double c0 = hrows[(size_t)k * nrhs + c];
            double tau_k = 0.0, beta = c0, d = 0.0;
            if (make[c] && tail_sq[c] > DBL_MIN) {
                beta = std::sqrt(c0 * c0 + tail_sq[c]);
                if (c0 >= 0.0) beta = -beta;
                d = c0 - beta;
                tau_k = (beta - c0) / beta;
            }
            hd[c] = d;
            tau[(size_t)c * L + k] = tau_k;
            beta_out[c] = beta;
            for (int i = 0; i < L; i++)
                h_at(vtop, c, i, k) = i < k ? 0.0 : (i == k ? 1.0 : (d != 0.0 ? hrows[(size_t)i * nrhs + c] / d : 0.0));
        }
        hip_check(hipMemcpy(dd, hd.data(), nrhs * sizeof(double), hipMemcpyHostToDevice), "gmres");
        vec(2, "householder", [&] { return launch_gm_householder(N, nrhs, k, dv, dd, V + (size_t)k * nv, st); });
    };
    std::vector<double> beta(nrhs), zero_tail(nrhs, 0.0);
    // a new cycle: r0 = M^{-1}(b - A x), u_0 from r0, w = beta e_0, T = tau_0
    auto start_cycle = [&](bool first) {
        spmv(dx, tmp);
        vec(2, "copy", [&] { return launch_copy_f64(v, db, nv, st); });
        vec(3, "axpby", [&] { return launch_axpby(N, nrhs, tmp, nullptr, -1.0, v, nullptr, st); });
        precond(v);
        if (first) {
            std::vector<double> n2 = dot_host(v, v, N);
            for (int c = 0; c < nrhs; c++) {
                r0norm[c] = std::sqrt(n2[c]);
                if (act[c] && r0norm[c] == 0.0) act[c] = 0;  // initial guess already exact
            }
        }
        std::vector<double> tail = N > 1 ? dot_host(v + nrhs, v + nrhs, N - 1) : zero_tail;
        std::vector<int> make(nrhs, 1);
        householder(0, v, tail, beta, make);
        for (int c = 0; c < nrhs; c++) {
            std::fill(H.begin() + (size_t)c * L * L, H.begin() + (size_t)(c + 1) * L * L, 0.0);
            std::fill(T.begin() + (size_t)c * L * L, T.begin() + (size_t)(c + 1) * L * L, 0.0);
            std::fill(w.begin() + (size_t)c * (L + 1), w.begin() + (size_t)(c + 1) * (L + 1), 0.0);
            w[(size_t)c * (L + 1)] = beta[c];
            h_at(T, c, 0, 0) = tau[(size_t)c * L];
        }
    };
    // device v := op applied with z (k x nrhs, host) and ye (ny rows, host)
    auto comb = [&](const double* basev, bool with_ye, int ny, int k, double* out) {
        hip_check(hipMemcpy(dz, hz.data(), (size_t)std::max(k, 1) * nrhs * sizeof(double), hipMemcpyHostToDevice), "gmres");
        if (with_ye)
            hip_check(hipMemcpy(dye, hye.data(), (size_t)std::max(ny, 1) * nrhs * sizeof(double), hipMemcpyHostToDevice),
                      "gmres");
        vec(1.0 + k, "gmres comb", [&] {
            return launch_gm_comb(N, nrhs, basev, with_ye ? dye : nullptr, ny, V, k, dz, out, st);
        });
    };
    auto multidot = [&](int k, const double* vv, std::vector<double>& out) {  // out[j * nrhs + c] = V_j . vv
        vec(1.0 + k, "gmres multidot", [&] { return launch_gm_multidot(N, nrhs, k, V, vv, part, dsm, st); });
        out.resize((size_t)k * nrhs);
        hip_check(hipMemcpy(out.data(), dsm, (size_t)k * nrhs * sizeof(double), hipMemcpyDeviceToHost), "gmres");
    };
    int nact = 0;
    for (int c = 0; c < nrhs; c++) nact += act[c];
    if (nact > 0) start_cycle(true);
    nact = 0;
    for (int c = 0; c < nrhs; c++) nact += act[c];
    std::vector<double> W;
    int k = 1;
    while (nact > 0) {
        for (int c = 0; c < nrhs; c++)
            if (act[c]) res.iters[c]++;
        // v = Q_k e_{k-1} = e_{k-1} - V T (V^T e_{k-1}); (V^T e_{k-1})_j = u_j[k-1]
        for (int c = 0; c < nrhs; c++) {
            for (int a = 0; a < k; a++) {
                double s = 0.0;
                for (int bb = a; bb < k; bb++) s += h_at(T, c, a, bb) * h_at(vtop, c, k - 1, bb);
                hz[(size_t)a * nrhs + c] = s;
            }
            for (int i = 0; i < k; i++) hye[(size_t)i * nrhs + c] = i == k - 1 ? 1.0 : 0.0;
        }
        comb(nullptr, true, k, k, v);
        spmv(v, tmp);
        vec(2, "copy", [&] { return launch_copy_f64(v, tmp, nv, st); });
        precond(v);
        // v = Q_k^T v = v - V T^T (V^T v)
        multidot(k, v, W);
        for (int c = 0; c < nrhs; c++)
            for (int a = 0; a < k; a++) {
                double s = 0.0;
                for (int bb = 0; bb <= a; bb++) s += h_at(T, c, bb, a) * W[(size_t)bb * nrhs + c];
                hz[(size_t)a * nrhs + c] = s;
            }
        comb(v, false, 0, k, v);
        // Householder reflection of v[k:] (src/is.cpp:209-217)
        std::vector<double> tail = (k + 1 < N) ? dot_host(v + (size_t)(k + 1) * nrhs, v + (size_t)(k + 1) * nrhs, N - k - 1)
                                               : zero_tail;
        read_rows(v);
        std::vector<double> vh((size_t)(k + 1) * nrhs);
        for (size_t q = 0; q < vh.size(); q++) vh[q] = hrows[q];
        std::vector<int> make(nrhs, 0);
        for (int c = 0; c < nrhs; c++) {
            const double c0 = k < N ? vh[(size_t)k * nrhs + c] : 0.0;
            make[c] = k < N && std::sqrt(c0 * c0 + tail[c]) != 0.0;
        }
        if (k < N && k <= R) {
            householder(k, v, tail, beta, make);
            for (int c = 0; c < nrhs; c++)
                if (make[c]) vh[(size_t)k * nrhs + c] = beta[c];  // H_k maps v[k:] to beta e_0
        }
        // T gains the column of u_k: T(0:k, k) = -tau_k T(0:k, 0:k) (V^T u_k), T(k, k) = tau_k
        if (k < R) {
            multidot(k, V + (size_t)k * nv, W);
            for (int c = 0; c < nrhs; c++) {
                const double tk = tau[(size_t)c * L + k];
                for (int a = 0; a < k; a++) {
                    double s = 0.0;
                    for (int bb = a; bb < k; bb++) s += h_at(T, c, a, bb) * W[(size_t)bb * nrhs + c];
                    h_at(T, c, a, k) = -tk * s;
                }
                h_at(T, c, k, k) = tk;
            }
        }
        // Givens rotations, Hessenberg column, residual estimate
        std::vector<int> finish(nrhs, 0);
        bool any_finish = false;
        for (int c = 0; c < nrhs; c++) {
            if (!act[c]) continue;
            auto vv = [&](int i) -> double& { return vh[(size_t)i * nrhs + c]; };
            Givens* Gc = G.data() + (size_t)c * R;
            double* wc = w.data() + (size_t)c * (L + 1);
            for (int i = 0; i + 1 < k; i++) apply_givens_adjoint(Gc[i], vv(i), vv(i + 1));
            if (k < N && vv(k) != 0.0) {
                Gc[k - 1] = make_givens(vv(k - 1), vv(k));
                apply_givens_adjoint(Gc[k - 1], vv(k - 1), vv(k));
                apply_givens_adjoint(Gc[k - 1], wc[k - 1], wc[k]);
            }
            for (int i = 0; i < k; i++) h_at(H, c, i, k - 1) = vv(i);
            rel[c] = std::fabs(wc[k]) / r0norm[c];
            const bool stop = (k == N || rel[c] < tol || res.iters[c] == maxit);
            if (stop || k == R) {
                finish[c] = stop ? 2 : 1;
                any_finish = true;
            }
        }
        if (!any_finish) {
            k++;
            continue;
        }
        // x += Q_k [y; 0] for the columns ending their cycle: y = H^{-1} w (upper triangular)
        for (int c = 0; c < nrhs; c++) {
            std::vector<double> y(k, 0.0);
            if (finish[c]) {
                const double* wc = w.data() + (size_t)c * (L + 1);
                for (int i = k - 1; i >= 0; i--) {
                    double s = wc[i];
                    for (int j = i + 1; j < k; j++) s -= h_at(H, c, i, j) * y[j];
                    y[i] = s / h_at(H, c, i, i);
                }
            }
            // e_j = (V^T [y; 0])_j = sum_{i >= j} u_j[i] y_i; z = T e
            std::vector<double> e(k, 0.0);
            for (int j = 0; j < k; j++)
                for (int i = j; i < k; i++) e[j] += h_at(vtop, c, i, j) * y[i];
            for (int a = 0; a < k; a++) {
                double s = 0.0;
                for (int bb = a; bb < k; bb++) s += h_at(T, c, a, bb) * e[bb];
                hz[(size_t)a * nrhs + c] = finish[c] ? s : 0.0;
            }
            for (int i = 0; i < k; i++) hye[(size_t)i * nrhs + c] = finish[c] ? y[i] : 0.0;
        }
        comb(dx, true, k, k, dx);
        bool restart_cycle = false;
        for (int c = 0; c < nrhs; c++) {
            if (finish[c] == 2) {
                act[c] = 0;
                nact--;
            } else if (finish[c] == 1) {
                restart_cycle = true;
            }
        }
        if (nact > 0 && restart_cycle) {
            start_cycle(false);
            k = 1;
        } else {
            k++;
        }
    }
    // true residual
    spmv(dx, tmp);
    vec(2, "copy", [&] { return launch_copy_f64(v, db, nv, st); });
    vec(3, "axpby", [&] { return launch_axpby(N, nrhs, tmp, nullptr, -1.0, v, nullptr, st); });
    std::vector<double> rn2 = dot_host(v, v, N), bn2 = dot_host(db, db, N);
    res.converged = 0;
    for (int c = 0; c < nrhs; c++) {
        res.relres[c] = bn2[c] > 0 ? std::sqrt(rn2[c] / bn2[c]) : 0.0;
        if (rel[c] < tol || res.iters[c] == 0) res.converged++;
    }
    if (on_device) {
        vec(2, "il2cm", [&] { return launch_il2cm(dx, N, nrhs, x, ld, st); });
        hip_check(hipStreamSynchronize(st), "gmres");
    } else {
        vec(2, "il2cm", [&] { return launch_il2cm(dx, N, nrhs, io, N, st); });
        hip_check(hipMemcpy2D(x, (size_t)ld * sizeof(double), io, (size_t)N * sizeof(double),
                              (size_t)N * sizeof(double), nrhs, hipMemcpyDeviceToHost),
                  "gmres io");
    }
    prof.collect();
    res.seconds = wall() - t0;
    res.seconds_precond = t_prec;
    res.seconds_spmv = t_spmv;
    res.applies = applies;
    return res;
}

}  // namespace spand
