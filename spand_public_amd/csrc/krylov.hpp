// Device FP64 Krylov solvers preconditioned by the spaND tree (src/is.cpp).
#pragma once
#include <vector>

#include "tree.hpp"

namespace spand {

// Square sparse operator on the device (CSR of A built from the caller's CSC).
struct DevMatrix {
    int N = 0;
    long long nnz = 0;
    int* rowptr = nullptr;
    int* colidx = nullptr;
    double* val = nullptr;
    int* csc_of = nullptr;  // CSR position -> index in the CSC arrays the matrix was built from
    // the caller's CSC arrays, which are the CSR arrays of A^T (the transposed operator)
    int* tptr = nullptr;
    int* tidx = nullptr;
    double* tval = nullptr;
    // Krylov work vectors, kept between calls (grown on demand): repeated solves with the
    // same operator allocate nothing.
    struct Workspace {
        void* p = nullptr;
        size_t bytes = 0;
    };
    mutable Workspace ws;
    void* workspace(size_t bytes) const;
    DevMatrix(int n, const int* colptr, const int* rowval, const double* vals);
    // vals_out[k] = alpha * sum_r u(i_k, r) v(j_k, r) for the stored entries in CSC order;
    // u, v, vals_out on the device (on_device) or on the host
    void sampled_outer(const double* u, const double* v, int nrhs, int ld, double alpha, double* vals_out,
                       bool on_device) const;
    // New values of the same pattern (device array, CSC order), ordered on st.
    void set_values_device(const double* vals, hipStream_t st);
    ~DevMatrix();
    DevMatrix(const DevMatrix&) = delete;
    DevMatrix& operator=(const DevMatrix&) = delete;
};

struct KrylovResult {
    std::vector<int> iters;
    std::vector<double> relres;
    int converged = 0;
    double seconds = 0, seconds_precond = 0, seconds_spmv = 0;
    int applies = 0;
};

// cg (src/is.cpp:39-121) on nrhs columns; b, x column-major with leading dimension ld.
KrylovResult pcg(Tree& t, const DevMatrix& A, const double* b, double* x, int nrhs, int ld, int maxit, double tol,
                 bool on_device);
// gmres (src/is.cpp:123-300): left-preconditioned GMRES(restart), columns in lockstep.
// transpose: solves A^T x = b with the adjoint preconditioner M^{-T}.
KrylovResult gmres(Tree& t, const DevMatrix& A, const double* b, double* x, int nrhs, int ld, int maxit, int restart,
                   double tol, bool on_device, bool transpose = false);

}  // namespace spand
