// extern "C" boundary (include/spand.h): wraps spand::Tree and the device Krylov driver,
// turning C++ exceptions into status codes and per-handle error messages.
#include <cstring>
#include <memory>
#include <string>

#include "../../include/spand.h"
#include "krylov.hpp"
#include "tree.hpp"

struct spand_tree {
    std::unique_ptr<spand::Tree> t;
    std::string err;
};
struct spand_matrix {
    std::unique_ptr<spand::DevMatrix> m;
};

static thread_local std::string g_err;

template <class F>
static spand_status guarded(spand_tree* h, F&& f) {
    try {
        f();
        return SPAND_OK;
    } catch (const spand::Error& e) {
        (h ? h->err : g_err) = e.what();
        return (spand_status)e.code;
    } catch (const std::bad_alloc&) {
        (h ? h->err : g_err) = "host allocation failed";
        return SPAND_ERR_NOMEM;
    } catch (const std::exception& e) {
        (h ? h->err : g_err) = e.what();
        return SPAND_ERR_ARG;
    }
}

#define CHECK_T(t)                                        \
    do {                                                  \
        if (!(t) || !(t)->t) {                            \
            g_err = "null tree handle";                   \
            return SPAND_ERR_ARG;                         \
        }                                                 \
    } while (0)

extern "C" {

spand_status spand_create(int nlevels, spand_tree** out) {
    if (!out) {
        g_err = "null output pointer";
        return SPAND_ERR_ARG;
    }
    *out = nullptr;
    return guarded(nullptr, [&] {
        auto h = std::make_unique<spand_tree>();
        h->t = std::make_unique<spand::Tree>(nlevels);
        *out = h.release();
    });
}

void spand_destroy(spand_tree* t) { delete t; }

const char* spand_last_error(const spand_tree* t) { return t ? t->err.c_str() : g_err.c_str(); }

spand_status spand_set_tol(spand_tree* t, double tol) {
    CHECK_T(t);
    return guarded(t, [&] {
        if (!(tol >= 0.0)) throw spand::Error(spand::ERR_ARG, "tol must be >= 0");
        t->t->set_tol(tol);
    });
}
spand_status spand_set_skip(spand_tree* t, int skip) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->set_skip(skip); });
}
spand_status spand_set_verb(spand_tree* t, int verb) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->set_verb(verb != 0); });
}
spand_status spand_set_symm_kind(spand_tree* t, int k) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->set_symm_kind(k); });
}
spand_status spand_set_use_geo(spand_tree* t, int g) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->set_use_geo(g != 0); });
}
spand_status spand_set_Xcoo(spand_tree* t, int dim, int N, const double* X) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->set_Xcoo(dim, N, X); });
}
spand_status spand_set_use_sparsify(spand_tree* t, int use) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->set_use_sparsify(use != 0); });
}
spand_status spand_set_monitor_flops(spand_tree* t, int on) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->set_monitor_flops(on != 0); });
}
spand_status spand_set_precision(spand_tree* t, int p) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->set_precision(p); });
}

spand_status spand_partition(spand_tree* t, int N, const int* colptr, const int* rowval, int* part_out) {
    CHECK_T(t);
    return guarded(t, [&] {
        if (N <= 0 || !colptr || !rowval) throw spand::Error(spand::ERR_ARG, "partition: bad arguments");
        t->t->partition(N, colptr, rowval);
        if (part_out) std::memcpy(part_out, t->t->part().data(), t->t->part().size() * sizeof(int));
    });
}
spand_status spand_assemble(spand_tree* t, int N, const int* colptr, const int* rowval, const double* vals) {
    CHECK_T(t);
    return guarded(t, [&] {
        if (N <= 0 || !colptr || !rowval || !vals) throw spand::Error(spand::ERR_ARG, "assemble: bad arguments");
        t->t->assemble(N, colptr, rowval, vals);
    });
}
spand_status spand_assemble_values_device(spand_tree* t, const double* d_vals, long long nnz, void* stream) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->assemble_values_device(d_vals, nnz, static_cast<hipStream_t>(stream)); });
}
spand_status spand_factorize(spand_tree* t) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->factorize(); });
}
spand_status spand_solve(spand_tree* t, double* x, int nrhs, int ld) {
    CHECK_T(t);
    return guarded(t, [&] {
        if (!x) throw spand::Error(spand::ERR_ARG, "solve: null x");
        t->t->solve(x, nrhs, ld, false, 0);
    });
}
spand_status spand_solve_device(spand_tree* t, double* d_x, int nrhs, int ld, void* stream) {
    CHECK_T(t);
    return guarded(t, [&] {
        if (!d_x) throw spand::Error(spand::ERR_ARG, "solve: null x");
        t->t->solve(d_x, nrhs, ld, true, (hipStream_t)stream);
    });
}

spand_status spand_solve_device_noalloc(spand_tree* t, double* d_x, int nrhs, int ld, void* stream) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->solve_device_noalloc(d_x, nrhs, ld, (hipStream_t)stream); });
}

spand_status spand_solve_transpose(spand_tree* t, double* x, int nrhs, int ld) {
    CHECK_T(t);
    return guarded(t, [&] {
        if (!x) throw spand::Error(spand::ERR_ARG, "solve_transpose: null x");
        t->t->solve_transpose(x, nrhs, ld, false, 0);
    });
}
spand_status spand_solve_transpose_device(spand_tree* t, double* d_x, int nrhs, int ld, void* stream) {
    CHECK_T(t);
    return guarded(t, [&] {
        if (!d_x) throw spand::Error(spand::ERR_ARG, "solve_transpose: null x");
        t->t->solve_transpose(d_x, nrhs, ld, true, (hipStream_t)stream);
    });
}

spand_status spand_get_N(const spand_tree* t, int* N) {
    CHECK_T(t);
    *N = t->t->N();
    return SPAND_OK;
}
spand_status spand_get_nlevels(const spand_tree* t, int* n) {
    CHECK_T(t);
    *n = t->t->nlevels();
    return SPAND_OK;
}
spand_status spand_get_stop(const spand_tree* t, int* stop) {
    CHECK_T(t);
    *stop = t->t->get_stop();
    return SPAND_OK;
}
spand_status spand_nnz(const spand_tree* t, long long* nnz) {
    CHECK_T(t);
    *nnz = t->t->nnz();
    return SPAND_OK;
}
spand_status spand_get_assembly_perm(const spand_tree* t, int* perm) {
    CHECK_T(t);
    const auto& p = t->t->perm();
    std::memcpy(perm, p.data(), p.size() * sizeof(int));
    return SPAND_OK;
}
spand_status spand_get_clusters(const spand_tree* t, int* out, int* count) {
    CHECK_T(t);
    if (!count) return SPAND_ERR_ARG;
    return guarded(const_cast<spand_tree*>(t), [&] { *count = t->t->clusters_info(out); });
}
spand_status spand_get_dof2ID(const spand_tree* t, int* ids) {
    CHECK_T(t);
    if (!ids) return SPAND_ERR_ARG;
    return guarded(const_cast<spand_tree*>(t), [&] { t->t->dof2id(ids); });
}
spand_status spand_get_trailing_coo(const spand_tree* t, long long* nnz, int* I, int* J, double* V) {
    CHECK_T(t);
    return guarded(const_cast<spand_tree*>(t), [&] { *nnz = t->t->trailing_coo(I, J, V); });
}
spand_status spand_get_level_stats(const spand_tree* t, double* out) {
    CHECK_T(t);
    const auto& st = t->t->stats();
    for (size_t l = 0; l < st.size(); l++) {
        const spand::LevelStats& s = st[l];
        double* o = out + 20 * l;
        o[0] = s.dofs_nd; o[1] = s.dofs_left_nd; o[2] = s.dofs_left_elim; o[3] = s.dofs_left_spars;
        o[4] = (double)s.fact_nnz; o[5] = s.rank_before; o[6] = s.rank_after; o[7] = s.nsparsified;
        o[8] = s.nignored; o[9] = s.nwaves; o[10] = s.t_plan; o[11] = s.t_elim_scale; o[12] = s.t_spars;
        o[13] = s.t_merge; o[14] = s.flops_potrf; o[15] = s.flops_trsm; o[16] = s.flops_schur;
        o[17] = s.flops_rrqr; o[18] = s.t_solve_plan; o[19] = 0;
    }
    return SPAND_OK;
}
spand_status spand_get_flops_log(const spand_tree* t, long long* count, int* out) {
    CHECK_T(t);
    return guarded(const_cast<spand_tree*>(t), [&] {
        if (!count) throw spand::Error(spand::ERR_ARG, "get_flops_log: null count");
        const auto& log = t->t->flops_log();
        *count = (long long)log.size();
        if (out)
            for (size_t i = 0; i < log.size(); i++) {
                int* o = out + 5 * i;
                o[0] = log[i].level; o[1] = log[i].kind; o[2] = log[i].m; o[3] = log[i].n; o[4] = log[i].k;
            }
    });
}
spand_status spand_get_status_flags(const spand_tree* t, unsigned* flags) {
    CHECK_T(t);
    return guarded(const_cast<spand_tree*>(t), [&] {
        if (!flags) throw spand::Error(spand::ERR_ARG, "get_status_flags: null output");
        *flags = t->t->status_flags();
    });
}
spand_status spand_get_factor_info(const spand_tree* t, double* seconds, long long* fb, long long* ab) {
    CHECK_T(t);
    if (seconds) *seconds = t->t->factor_seconds();
    if (fb) *fb = (long long)t->t->factor_bytes();
    if (ab) *ab = (long long)t->t->arena_bytes();
    return SPAND_OK;
}
spand_status spand_set_profiling(spand_tree* t, unsigned long long family_mask) {
    CHECK_T(t);
    return guarded(t, [&] {
        t->t->profiler().collect();
        t->t->profiler().set_mask(family_mask);
    });
}
spand_status spand_profile_reset(spand_tree* t) {
    CHECK_T(t);
    return guarded(t, [&] { t->t->profiler().reset(); });
}
int spand_kernel_family_count(void) { return spand::KF_COUNT; }
const char* spand_kernel_family_name(int family) { return spand::kernel_family_name(family); }
spand_status spand_kernel_profile(spand_tree* t, int family, long long* launches, double* ms, double* bytes,
                                  double* flops) {
    CHECK_T(t);
    if (family < 0 || family >= spand::KF_COUNT) {
        t->err = "kernel family out of range";
        return SPAND_ERR_ARG;
    }
    return guarded(t, [&] {
        t->t->profiler().collect();
        spand::KernelTotals k = t->t->profiler().totals(family);
        if (launches) *launches = k.launches;
        if (ms) *ms = k.ms;
        if (bytes) *bytes = k.bytes;
        if (flops) *flops = k.flops;
    });
}
spand_status spand_num_ops(const spand_tree* t, int* n) {
    CHECK_T(t);
    *n = (int)t->t->ops().size();
    return SPAND_OK;
}
spand_status spand_ops_info(const spand_tree* t, int* info) {
    CHECK_T(t);
    const auto& ops = t->t->ops();
    for (size_t i = 0; i < ops.size(); i++) {
        const spand::OpRec& o = ops[i];
        int* p = info + 8 * i;
        p[0] = o.type; p[1] = o.a; p[2] = o.b; p[3] = o.sa; p[4] = o.sb; p[5] = o.rows; p[6] = o.cols; p[7] = o.level;
    }
    return SPAND_OK;
}
spand_status spand_op_data(const spand_tree* t, int i, double* M, double* tau) {
    CHECK_T(t);
    return guarded(const_cast<spand_tree*>(t), [&] { t->t->op_data(i, M, tau); });
}
spand_status spand_op_plu(const spand_tree* t, int i, double* U, int* perm) {
    CHECK_T(t);
    return guarded(const_cast<spand_tree*>(t), [&] { t->t->op_plu(i, U, perm); });
}

spand_status spand_get_level_tune(const spand_tree* t, int* out) {
    CHECK_T(t);
    return guarded(const_cast<spand_tree*>(t), [&] {
        if (!out) throw spand::Error(spand::ERR_ARG, "get_level_tune: null output");
        const auto& tu = t->t->tunes();
        const auto& sh = t->t->shapes();
        const int nl = t->t->nlevels();
        for (int l = 0; l < nl; l++) {
            int* o = out + (size_t)l * (spand::TUNE_SHAPE_INTS + spand::TUNE_INTS);
            spand::LevelShape s = l < (int)sh.size() ? sh[l] : spand::LevelShape();
            spand::LevelTune u = l < (int)tu.size() ? tu[l] : spand::LevelTune();
            const int v[] = {s.n_elim, s.elim_med, s.elim_max, s.n_spars, s.rows_med, s.rows_max, s.cols_med,
                             s.cols_max, u.rrqr_reg, u.small_elems, u.step_waves, u.trsm_cw, u.source};
            for (int k = 0; k < spand::TUNE_SHAPE_INTS + spand::TUNE_INTS; k++) o[k] = v[k];
        }
    });
}

spand_status spand_tune_choose(int elem_bytes, const int* shape, int* out) {
    if (!shape || !out || (elem_bytes != 4 && elem_bytes != 8)) {
        g_err = "tune_choose: bad arguments";
        return SPAND_ERR_ARG;
    }
    spand::LevelShape s;
    s.n_elim = shape[0];
    s.elim_med = shape[1];
    s.elim_max = shape[2];
    s.n_spars = shape[3];
    s.rows_med = shape[4];
    s.rows_max = shape[5];
    s.cols_med = shape[6];
    s.cols_max = shape[7];
    const spand::LevelTune u = spand::choose_tune((size_t)elem_bytes, s);
    out[0] = u.rrqr_reg;
    out[1] = u.small_elems;
    out[2] = u.step_waves;
    out[3] = u.trsm_cw;
    out[4] = u.source;
    return SPAND_OK;
}

spand_status spand_matrix_create(int N, const int* colptr, const int* rowval, const double* vals, spand_matrix** out) {
    if (!out || N <= 0 || !colptr || !rowval || !vals) {
        g_err = "matrix_create: bad arguments";
        return SPAND_ERR_ARG;
    }
    *out = nullptr;
    return guarded(nullptr, [&] {
        auto m = std::make_unique<spand_matrix>();
        m->m = std::make_unique<spand::DevMatrix>(N, colptr, rowval, vals);
        *out = m.release();
    });
}
void spand_matrix_destroy(spand_matrix* A) { delete A; }
spand_status spand_matrix_set_values_device(spand_matrix* A, const double* d_vals, long long nnz, void* stream) {
    if (!A || !A->m || !d_vals) {
        g_err = "matrix_set_values_device: null argument";
        return SPAND_ERR_ARG;
    }
    if (nnz != A->m->nnz) {
        g_err = "matrix_set_values_device: nnz differs from the matrix pattern";
        return SPAND_ERR_ARG;
    }
    return guarded(nullptr, [&] { A->m->set_values_device(d_vals, static_cast<hipStream_t>(stream)); });
}
spand_status spand_matrix_sampled_outer(const spand_matrix* A, const double* u, const double* v, int nrhs, int ld,
                                        double alpha, double* vals_out, int on_device) {
    if (!A || !A->m || !u || !v || !vals_out) {
        g_err = "matrix_sampled_outer: null argument";
        return SPAND_ERR_ARG;
    }
    return guarded(nullptr, [&] { A->m->sampled_outer(u, v, nrhs, ld, alpha, vals_out, on_device != 0); });
}

static spand_status krylov_out(spand_tree* t, const spand::KrylovResult& r, int nrhs, int* iters_out,
                               spand_krylov_stats* stats, const char* name) {
    if (iters_out)
        for (int c = 0; c < nrhs; c++) iters_out[c] = r.iters[c];
    int itmax = 0;
    double rr = 0;
    for (int c = 0; c < nrhs; c++) {
        itmax = std::max(itmax, r.iters[c]);
        rr = std::max(rr, r.relres[c]);
    }
    if (stats) {
        stats->iterations = itmax;
        stats->converged = r.converged;
        stats->relres_max = rr;
        stats->seconds = r.seconds;
        stats->seconds_precond = r.seconds_precond;
        stats->seconds_spmv = r.seconds_spmv;
        stats->applies = r.applies;
    }
    if (r.converged < nrhs) {
        t->err = std::string(name) + ": iteration cap reached before the tolerance";
        return SPAND_ERR_NOT_CONVERGED;
    }
    return SPAND_OK;
}

spand_status spand_pcg(spand_tree* t, const spand_matrix* A, const double* b, double* x, int nrhs, int ld, int maxit,
                       double tol, int on_device, int* iters_out, spand_krylov_stats* stats) {
    CHECK_T(t);
    if (!A || !A->m || !b || !x) {
        t->err = "pcg: null argument";
        return SPAND_ERR_ARG;
    }
    spand::KrylovResult r;
    spand_status s = guarded(t, [&] { r = spand::pcg(*t->t, *A->m, b, x, nrhs, ld, maxit, tol, on_device != 0); });
    if (s != SPAND_OK) return s;
    return krylov_out(t, r, nrhs, iters_out, stats, "pcg");
}

spand_status spand_gmres(spand_tree* t, const spand_matrix* A, const double* b, double* x, int nrhs, int ld,
                         int maxit, int restart, double tol, int on_device, int* iters_out, spand_krylov_stats* stats) {
    CHECK_T(t);
    if (!A || !A->m || !b || !x) {
        t->err = "gmres: null argument";
        return SPAND_ERR_ARG;
    }
    spand::KrylovResult r;
    spand_status s = guarded(
        t, [&] { r = spand::gmres(*t->t, *A->m, b, x, nrhs, ld, maxit, restart, tol, on_device != 0); });
    if (s != SPAND_OK) return s;
    return krylov_out(t, r, nrhs, iters_out, stats, "gmres");
}

spand_status spand_gmres_transpose(spand_tree* t, const spand_matrix* A, const double* b, double* x, int nrhs, int ld,
                                   int maxit, int restart, double tol, int on_device, int* iters_out,
                                   spand_krylov_stats* stats) {
    CHECK_T(t);
    if (!A || !A->m || !b || !x) {
        t->err = "gmres_transpose: null argument";
        return SPAND_ERR_ARG;
    }
    spand::KrylovResult r;
    spand_status s = guarded(
        t, [&] { r = spand::gmres(*t->t, *A->m, b, x, nrhs, ld, maxit, restart, tol, on_device != 0, true); });
    if (s != SPAND_OK) return s;
    return krylov_out(t, r, nrhs, iters_out, stats, "gmres_transpose");
}

}  // extern "C"
