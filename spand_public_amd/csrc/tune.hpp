// Per-level launch configuration of the factorization, chosen from the level's size
// histogram by a table measured on gfx950 (tune.cpp). Every choice here changes only how the
// work is split over launches and workgroups, never an operation order of the arithmetic:
// results are bitwise the same under any configuration.
#pragma once
#include <cstddef>
#include <string>

namespace spand {

// Size histogram features of one ND level.
struct LevelShape {
    int n_elim = 0;        // pivots eliminated at the level
    int elim_med = 0;      // median / largest pivot size (the trsm triangles)
    int elim_max = 0;
    int n_spars = 0;       // clusters sparsified at the level
    int rows_med = 0;      // median / largest Asn rows (cluster size)
    int rows_max = 0;
    int cols_med = 0;      // median / largest Asn columns (sum of the neighbours' sizes)
    int cols_max = 0;
};

struct LevelTune {
    int rrqr_reg = 1;          // register-resident RRQR where a kernel form holds the cluster
    int small_elems = 10000;   // one-workgroup LDS RRQR up to this many Asn elements, stepped above
    int step_waves = 0;        // wavefronts per 64-column slab of a stepped step (0: by slab count)
    int trsm_cw = 0;           // trsm chunk widths by triangle length: 0 (64 / 32 / 16), 1 narrow, 2 wide
    int source = 0;            // 0 measured table, 1 environment override
};

constexpr int TUNE_SHAPE_INTS = 8;
constexpr int TUNE_INTS = 5;

// The table's configuration for a level of this shape and element size (bytes).
LevelTune choose_tune(size_t esz, const LevelShape& s);
// Environment overrides (SPAND_RRQR_REG, SPAND_RRQR_STEPPED_MIN_ELEMS, SPAND_RRQR_STEP_WAVES,
// SPAND_TRSM_CW) applied on top of a choice.
void apply_tune_env(LevelTune& t);
// Widest trsm chunk for a triangle of n rows under a chunk-width policy.
int trsm_chunk_width(int policy, int n);

}  // namespace spand
