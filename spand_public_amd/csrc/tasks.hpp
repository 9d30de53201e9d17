// Task descriptors shared by the host level planner (tree.cpp) and the level-batched HIP
// kernels (kernels.hip). Every kernel launch processes an array of these, one workgroup per
// task (or per task tile), so one launch covers every cluster / edge of one ND level, of
// uneven sizes. Blocks are column-major with ld = rows, in the element type of the
// factorization (float or double); pointers are device addresses in the block arena.
#pragma once
#include <cstdint>

namespace spand {

// In-place lower Cholesky of an n x n pivot block (reference: potf, src/util.cpp:158-165).
struct PotrfTask {
    void* A;
    int n;
    int cluster;  // reported through the info array on a non-SPD pivot
    void* dbuf;   // pivots above POTRF_SMALL_MAX: the factored 32 x 32 diagonal block of the
                  // current panel, handed from k_potrf_panel to k_potrf_trail
};

// Triangular solve of (a 64-wide chunk of) an edge block against a pivot factor L
// (n x n, lower, ld n):
//   side 0: B (n x m, ld ldb) := L^{-1} B      (trsm_left,  src/util.cpp:266-274)
//   side 1: B (m x n, ld ldb) := B L^{-T}      (trsm_right, src/util.cpp:256-264)
// m <= 64: the host splits wide blocks into independent column (left) / row (right) chunks.
// A left solve may read its right-hand side from another block: B := L^{-1} P^T src, row i
// of P^T src being row perm[i] of src (perm nullptr: the identity), the row interchanges of
// an LU pivot applied to an in-edge (trsm_getf_edgeIn, src/tree.cpp:668-677).
struct TrsmTask {
    const void* L;
    void* B;
    int n;
    int m;
    int side;
    int ldb;
    const void* src;  // nullptr: in place
    const int* perm;
};

// LU with partial pivoting of an n x n pivot block (getf, src/util.cpp:184-199, in dgetf2
// order) followed by split_LU (src/util.cpp:213-227). On return A holds L (lower, diagonal
// sqrt|d_j|, zero above), Ut holds U^T (lower, zero above) and perm the row permutation p
// with P^T A = L U (row i of P^T A is row p[i] of A).
struct GetrfTask {
    void* A;
    void* Ut;
    int* perm;
    int n;
    int cluster;  // reported through the info array on a zero pivot
};

// One Schur-complement target tile: C[r0:r0+tm, c0:c0+tn] of a block C (ldc rows) receives
// contributions [k0, k1) in order, each C -= opA(A) opB(B) (gemm / syrk, src/util.cpp:122-156).
// lower = 1 restricts the update to the lower triangle (syrk on a diagonal block).
struct SchurTile {
    void* C;
    int ldc;
    int r0, c0, tm, tn;
    int lower;
    int k0, k1;
};
struct SchurContrib {
    const void* A;  // opA(A) is m x k: ta ? A^T (A is k x m, ld lda) : A (m x k, ld lda)
    const void* B;  // opB(B) is k x n: tb ? B^T (B is n x k, ld ldb) : B (k x n, ld ldb)
    int k;
    int lda, ldb;
    int ta, tb;
    int pad;
};

// Column-pivoted QR sparsification of one cluster (sparsify_adaptive_only,
// src/tree.cpp:1292-1347). Its Asn = [A_sn(in edges) | A_ns^T(out edges)] is gathered
// from segs[seg0:seg0+nseg] using the live sizes in dsize[], factored (dgeqp3 order),
// truncated by choose_rank and scattered back into the edge blocks.
struct RrqrTask {
    int cluster;
    int seg0, nseg;
    int max_rows, max_cols;  // capacity of the scratch (rows x cols) and Q (rows x rows)
    int pad;
    void* scratch;     // T[max_rows * max_cols]
    double* norms;     // double[2 * max_cols]
    int* jpvt;         // int[max_cols]
    void* Q;           // T[max_rows * min(max_rows, max_cols)] (Householder vectors)
    void* tau;         // T[min(max_rows, max_cols)]
};
struct RrqrSeg {
    void* E;   // edge block
    int nbr;   // neighbour cluster (its live size gives the segment width)
    int out;   // 0: in-edge, block is s x n; 1: out-edge, block is n x s (transposed)
};

// Large clusters use the stepped RRQR (one launch per pivot step, columns split into
// 64-wide slabs, one wavefront each). Columns never move: a double-buffered
// position <-> column map reproduces LAPACK's column swaps and its first-max tie-break.
struct RrqrBig {
    RrqrTask t;      // scratch holds A (rows x cols, ld rows); norms holds vn1 | vn2 by column
    int* map;        // p2c[2][max_cols] then c2p[2][max_cols]
    double* cand_v;  // [2][nslab] per-slab best partial norm for the next step, then |R_00|
    int* cand_p;     // [2][nslab] its position
    int* dims;       // rows, cols, min(rows, cols), steps to run (the rank once it is known),
                     // 1 once every kept reflector is stored
    void* vbuf;      // nullptr, or [nslab][2][max_rows] pivot column / reflector buffers
                     // (per-step launches), or the relay words of the persistent launch
                     // (relay = 1): candidates [2][nslab][4], then candidate columns
                     // [2][nslab][max_rows * sizeof(T) / 4], 64 bits each
    long long* done; // host-visible flag, or nullptr: set to 1 by the last step that does
                     // work (rank reached or all min(rows, cols) steps taken)
    int nslab;
    int relay;       // 1: vbuf holds relay words (k_rrqr_init clears them)
};
struct RrqrSlab {
    int big;   // index into the RrqrBig array
    int slab;  // columns [64 * slab, 64 * slab + 64)
};

// Merge copy of one child block into a parent block (update_edges, src/tree.cpp:1133-1184).
struct CopyTask {
    const void* src;  // nullptr => identity block
    void* dst;
    int rows, cols;
    int ldd;
    int pad;
};

// ---------------------------------------------------------------------------------------
// Preconditioner apply (Tree::solve, src/tree.cpp:1610-1635). Vectors are stored
// interleaved: element i of right-hand side r lives at x[(xoff + i) * nrhs + r].
//
// Every recorded op is a matrix-vector product on segments, applied per ND level and op
// kind in one "mv phase": a set of groups x[dst] (op)= sum_k op(A_k) x[src_k], split into
// chunks of 64 outputs and, along the inner dimension, into pieces of bounded work, one
// workgroup each. A chunk with one piece writes its destination directly; otherwise the
// pieces write partial sums that a second launch adds in piece order (deterministic).
//   ScalingLLT   (src/operations.cpp:95-103):  x := Linv x,        x := Linv^T x
//   ScalingPLUQ  (src/operations.cpp:114-121): x := Linv x[perm],  x := (U^{-T})^T x
//   Gemm*        (src/operations.cpp:31-93):   x_n -= A x_s / A^T x_s
//   Orthogonal   (src/operations.cpp:149-158): Q = I - V T V^T (compact WY over all kept
//                reflectors), X = T^T: y := -V^T x, z := -X y (fwd) or -X^T y (bwd), x -= V z
struct MvTerm {
    const void* A;     // rows x cols, ld rows, the factor's element type
    const int* perm;   // not transposed: inner index j reads source element perm[j]
    int rows, cols;
    int xsrc;          // source segment
    int trans;         // 0: output = row of A, inner = column; 1: output = column, inner = row
    int tri;           // inner-range clip of a triangular A for outputs [o0, o0 + 64):
                       // 1 inner < o0 + 64, 2 inner >= o0, 0 none
    int pad;
};
enum MvMode { MV_SUB = 0, MV_SETNEG = 1, MV_SET = 2 };  // x -= s, x = -s, x = s
struct MvGroup {
    int xdst, n;  // destination segment and length
    int k0, k1;   // term range
    int mode;
    int pad;
    const int* dperm;  // nullptr, or output o goes to xdst + dperm[o] (then always through
                       // the partial sums: a permuted destination is never written directly)
};
struct MvChunk {
    int group, o0;  // outputs [o0, o0 + 64) of the group
    int p0, p1;     // its pieces (indices into the phase's piece array)
};
struct MvPiece {
    int chunk;
    int k0, k1;    // terms; when k1 == k0 + 1 only inner indices [j0, j1) of term k0
    int j0, j1;
    int direct;    // the chunk's only piece: writes the destination itself
    // copies of its chunk / group fields, so that a workgroup reads one record
    int o0, no;    // the chunk's first output and output count
    int xdst, mode;
};
// Explicit Householder vectors V (unit diagonal, zeros above) and the compact-WY factor of
// one Orthogonal op, built at factorization. Q = H_0 H_1 ... H_{k-1} = I - V T V^T with T
// upper triangular and T^{-1} = diag(1 / tau) + striu(V^T V) (the UT transform); the solve
// stores X = T^T = S^{-1}, S = T^{-T} = diag(1 / tau) + stril(V^T V), computed by a
// triangular solve against the identity.
struct WyTask {
    const void* V;    // m x k reflectors as stored by the QR (R on and above the diagonal)
    const void* tau;
    void* Ve;         // m x k explicit V
    void* S;          // k x k lower: diag(1 / tau) + stril(V^T V)
    void* X;          // k x k, set to the identity here (then S^{-1} by the trsm)
    int m, k;
};
struct MergeTask {  // Merge (src/operations.cpp:166-192)
    int pxoff;
    int c0, c1;  // range into (child xoff, child size) pairs
    int pad;
};

}  // namespace spand
