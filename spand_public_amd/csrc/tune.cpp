// Table-driven choice of the per-level launch configuration (tune.hpp).
//
// The table was measured on one MI355X (gfx950) with the level-by-level device times of
// repeated factorizations of the 3D Laplacian 64^3 (12 levels) and the variable-kappa
// diffusion 96^3 (13 levels), each knob swept with the others at their defaults; a level
// falls in the bucket of its median Asn rows and columns (sparsification knobs) and of its
// median pivot (trsm chunks). Rows of the table: [lo, hi) of the features -> configuration,
// the first matching row wins.
#include "tune.hpp"

#include <cstdlib>

namespace spand {

namespace {
struct RrqrRule {
    int rows_lo, rows_hi;  // median Asn rows of the level's sparsified clusters
    int cols_lo, cols_hi;  // median Asn columns
    int rrqr_reg, small_elems, step_waves;
};
struct TrsmRule {
    int piv_lo, piv_hi;  // median pivot size of the level's eliminations
    int trsm_cw;
};
// Measured (tune_sweep.py, per-level device time, a choice kept when it beat the default by
// more than 3 %):
// * levels whose clusters have a median of fewer than 32 rows and 300 columns (level 0:
//   thousands of clusters of about 20 x 200) run faster on the LDS one-workgroup RRQR than
//   on the register path (11-21 % of the level); with about 320 columns (level 1 of the
//   64^3 Laplacian) the register path is faster (4.2 against 5.1 ms);
// * FP64 stepped steps with 8 wavefronts per slab at median rows 64-160 (6-13 %) and 16 at
//   160-400 (5-10 %) beat the choice by slab count; FP32 showed no gain from either;
// * the trsm chunk policies were never better than the default (64 / 32 / 16 columns).
// FP32 factor (4-byte elements)
constexpr int INF = 1 << 30;
const RrqrRule kRrqrF32[] = {
    {0, 32, 0, 300, 0, 10000, 0},
    {0, INF, 0, INF, 1, 10000, 0},
};
// FP64 factor (8-byte elements)
const RrqrRule kRrqrF64[] = {
    {0, 32, 0, 300, 0, 10000, 0},
    {0, 64, 0, INF, 1, 10000, 0},
    {64, 160, 0, INF, 1, 10000, 8},
    {160, 400, 0, INF, 1, 10000, 16},
    {0, INF, 0, INF, 1, 10000, 0},
};
const TrsmRule kTrsmF32[] = {{0, 1 << 30, 0}};
const TrsmRule kTrsmF64[] = {{0, 1 << 30, 0}};

// the first rule whose ranges hold the level's median rows and columns
template <size_t N>
const RrqrRule& pick(const RrqrRule (&rules)[N], int rows, int cols) {
    for (size_t i = 0; i < N; i++)
        if (rows >= rules[i].rows_lo && rows < rules[i].rows_hi && cols >= rules[i].cols_lo &&
            cols < rules[i].cols_hi)
            return rules[i];
    return rules[N - 1];
}
template <size_t N>
const TrsmRule& pick_trsm(const TrsmRule (&rules)[N], int x) {
    for (size_t i = 0; i < N; i++)
        if (x >= rules[i].piv_lo && x < rules[i].piv_hi) return rules[i];
    return rules[N - 1];
}

long long env_ll(const char* name, long long dflt, bool* set) {
    const char* e = std::getenv(name);
    if (!e || !*e) return dflt;
    if (set) *set = true;
    return std::atoll(e);
}
}  // namespace

LevelTune choose_tune(size_t esz, const LevelShape& s) {
    const bool f64 = esz == sizeof(double);
    const RrqrRule& r = f64 ? pick(kRrqrF64, s.rows_med, s.cols_med) : pick(kRrqrF32, s.rows_med, s.cols_med);
    const TrsmRule& q = f64 ? pick_trsm(kTrsmF64, s.elim_med) : pick_trsm(kTrsmF32, s.elim_med);
    LevelTune t;
    t.rrqr_reg = r.rrqr_reg;
    t.small_elems = r.small_elems;
    t.step_waves = r.step_waves;
    t.trsm_cw = q.trsm_cw;
    t.source = 0;
    return t;
}

void apply_tune_env(LevelTune& t) {
    bool set = false;
    t.rrqr_reg = (int)(env_ll("SPAND_RRQR_REG", t.rrqr_reg, &set) != 0);
    t.small_elems = (int)env_ll("SPAND_RRQR_STEPPED_MIN_ELEMS", t.small_elems, &set);
    t.step_waves = (int)env_ll("SPAND_RRQR_STEP_WAVES", t.step_waves, &set);
    t.trsm_cw = (int)env_ll("SPAND_TRSM_CW", t.trsm_cw, &set);
    if (set) t.source = 1;
}

int trsm_chunk_width(int policy, int n) {
    switch (policy) {
        case 1: return n >= 128 ? 16 : (n >= 64 ? 32 : 64);   // narrow: more workgroups
        case 2: return n >= 256 ? 32 : 64;                    // wide: fewer, longer tasks
        default: return n >= 256 ? 16 : (n >= 128 ? 32 : 64);
    }
}

}  // namespace spand
