// Device FP64 preconditioned CG on a block of right-hand sides (cg, src/is.cpp:39-121).
// Every column follows the reference recurrence exactly and stops on its own recursive
// residual; columns run in lockstep so one preconditioner apply serves all of them.
#include "krylov.hpp"

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
    std::vector<int> rp(n + 1, 0), ci(nnz);
    std::vector<double> v(nnz);
    for (long long k = 0; k < nnz; k++) rp[rowval[k] + 1]++;
    for (int i = 0; i < n; i++) rp[i + 1] += rp[i];
    std::vector<int> fill(rp.begin(), rp.end() - 1);
    for (int j = 0; j < n; j++)
        for (int k = colptr[j]; k < colptr[j + 1]; k++) {
            int p = fill[rowval[k]]++;
            ci[p] = j;
            v[p] = vals[k];
        }
    hip_check(hipMalloc((void**)&rowptr, (n + 1) * sizeof(int)), "hipMalloc(csr)");
    hip_check(hipMalloc((void**)&colidx, std::max<long long>(nnz, 1) * sizeof(int)), "hipMalloc(csr)");
    hip_check(hipMalloc((void**)&val, std::max<long long>(nnz, 1) * sizeof(double)), "hipMalloc(csr)");
    hip_check(hipMemcpy(rowptr, rp.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice), "upload csr");
    if (nnz) {
        hip_check(hipMemcpy(colidx, ci.data(), nnz * sizeof(int), hipMemcpyHostToDevice), "upload csr");
        hip_check(hipMemcpy(val, v.data(), nnz * sizeof(double), hipMemcpyHostToDevice), "upload csr");
    }
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

}  // namespace spand
