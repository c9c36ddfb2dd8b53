// Internal declarations of the sprgpu library (host plan, device handle,
// kernel launchers).  Not part of the public C ABI (include/sprgpu.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "sprgpu.h"

namespace sprgpu {

using i64 = int64_t;
using i32 = int32_t;

// Tile and block sizes shared by the planner and the kernels.
constexpr int kTileSmall = 64;   // MFMA update tile (the default for every update)
constexpr int kTileBig = 128;    // larger tile, selectable for the rank-kOuterRank and Schur
                                 // updates (sprgpu_set_plan_params, the autotuner)
constexpr int kAutoBigTiles = 4096;  // FP32 update launches of at least this many 64-tiles
                                     // take the 128 tile (big_tile 0)
constexpr int kGemmAlign = 4;    // MFMA update tiles start on rows that are multiples of this
constexpr int kFrontLdAlign = 16; // front leading dimensions are multiples of this (64 B in FP32)
constexpr int kPanelNB = 32;     // columns per panel step of the blocked front factor
constexpr int kPanelRows = 256;  // panel rows per workgroup (one row per thread)
constexpr int kPanelOuter = 128; // outer block of the two-level panel blocking (block kernels)
constexpr int kOuterRank = 256;  // rank of the trailing updates between outer blocks
constexpr int kTrsmRows = 16;    // rows below an outer block per wavefront (k_block_trsm) ...
constexpr int kTrsmWaves = 4;    // ... and wavefronts per workgroup
// Elements of T per outer block that the fused outer-block launch hands from
// the diagonal workgroup to the rows below: Bm of the four sub-steps and
// d_k L(r, k) of the block's rows past the first sub-step.
constexpr int kPoolExactFronts = 16384;  // largest fronts placed in the pool by lifetime overlap
constexpr int kFusedMaxFronts = 64;  // outer-block launches of at most this many fronts are fused
constexpr int kPubElems = 4 * kPanelNB * kPanelNB + (kPanelOuter - kPanelNB) * (kPanelOuter - kPanelNB);
constexpr int kSmallF = 64;      // fronts with F <= kSmallF can be factored in LDS (one lane a row)
// XCDs of the device: consecutive workgroup ids are dealt to them in turn,
// each with its own L2
constexpr int kXcds = 8;
constexpr int kForestWaves = 4;  // wavefronts per workgroup of the forest schedule (one item each)
constexpr int kForestWgsPerLevel = 1024;  // at most this many workgroups per tree height
constexpr int kInitCols = 8;     // front columns per workgroup in the init kernel
constexpr int kEaCols = 8;       // child columns per workgroup in extend-add
constexpr int kSolveSmallS = 64;   // narrow supernodes: S <= this ...
constexpr int kSolveSmallF = 1024; // ... and F <= this: one workgroup, vector in LDS (taller
                                   // fronts take the wide path, whose rows below split
                                   // into parallel chunks)
constexpr int kSolveNB = 128;      // column block (and rectangle row chunk) of the wide solve
// The inverse of a wide diagonal block is kept in two zero-padded layouts of
// kInvLayout entries each (the forward and the backward sweep's), i.e. the
// lower triangle's three 64 x 64 quadrants per layout.
constexpr int kInvLayout = 3 * kSolveNB * kSolveNB / 4;
// Per wide block: the inverse in both layouts, then M = L11^-1 L(block,
// block - 1) (forward, blocks after the first) and M' = L11^-T L(block + 1,
// block)^T (backward, blocks before the last), each kSolveNB x kSolveNB
// column-major.
constexpr int kBlockSolveData = 2 * kInvLayout + 2 * kSolveNB * kSolveNB;

// A wide supernode in the dataflow solves: its diagonal part is split into
// nblk blocks of kSolveNB columns (one workgroup each, ordered so that every
// workgroup only waits for lower-indexed ones) and its F-S rectangle rows
// into nrect chunks of kSolveNB rows.  flag0 is the index of its first block
// among the blocks of all wide supernodes (block inverses, result slots),
// pbase the offset of its nblk*nrect*kSolveNB backward partial sums.
struct WideDesc {
    i32 node, nblk, nrect, pad;
    i64 flag0, pbase;
};

// One supernode as the device sees it.  Rows of the front are the supernode's
// LINDX list (sorted ascending, the S own columns first); the front is a dense
// column-major F x F block (lower triangle used) with leading dimension ld.
struct NodeInfo {
    i64 foff;      // offset of the front in the pool (elements), while it is alive
    i64 loff;      // offset of the factored panel (ld x S) in the factor store
    i64 lxoff;     // offset of the row list in lindx
    i64 xlnz;      // offset of the packed trapezoid in LNZ (0-based)
    i64 voff;      // offset of the node's vector slot in the solve workspace
    i64 mapoff;    // offset of this node's (child) relative map in relmap
    i32 first;     // first equation (0-based)
    i32 S, F, ld;
    i32 parent;    // -1 for a root
    i32 child0, nchild;  // range in the children array (SIBL order)
};

// Pivot rule arguments of the factor kernels: TOL(1), TOL(4) and the mode
// (1: SPRFC1, record a failing pivot; 3: SPRFC3, replace it; 4: SPRFC4,
// replace it or, below -TOL(1)|a_jj|, record the failure).
struct PivotArgs {
    double tol1, tol4;
    int mode;
};

// The forest schedule (every front F <= kSmallF): per supernode what its
// work item needs in one record, and its children's records (SIBL order).
struct ForestNode {
    i64 xlnz;      // packed trapezoid in LNZ (0-based)
    i64 loff;      // factored panel in the factor store
    i64 uoff;      // packed update block in the forest's update-block array
    i64 voff;      // vector slot in the solve workspace
    i64 lxoff;     // row list in lindx
    i32 first, S, F, ld;
    i32 parent;    // -1 for a root
    i32 child0, nchild;  // range in the child records
    i32 pad;
};
struct ForestChild {
    i64 uoff;      // the child's packed update block
    i64 vofs;      // its update vector: vwork[vofs .. vofs + U)
    i64 mapoff;    // its U rows' positions in the parent front (relmap)
    i32 U, pad;
};
struct FactorStatus;
// Arguments of the persistent forest kernel.
struct ForestArgs {
    const ForestNode *fnodes;
    const ForestChild *fchild;
    i32 n;          // supernodes
    i32 phases;     // bit 0: ascending pass (factor and/or forward), bit 1: backward pass
    const i32 *order;   // supernodes: the ascending pass's order, then the backward pass's
    const i32 *chunks;  // workgroup chunks over order: nasc ascending, then nbwd backward, end
    i32 nasc, nbwd;
    const i32 *relmap, *lindx;
    const double *lnz;
    void *ublk, *store;
    double *dvec, *adiag, *dcor;
    FactorStatus *st;  // factor launches
    int *err;          // solve launches: set when a wait gave up
    PivotArgs pa;
    double *b, *vwork;  // one right-hand side
    unsigned *sync;     // [0] unused, [1] a wait gave up, then per supernode: children
                        // done, ascending pass done, backward pass done
};

// Work descriptors consumed by the kernels (built once per plan).
struct ColTask { i32 node, c0, c1, pad; };        // front columns [c0,c1)
// one supernode's part of a row of L: columns 0..kend-1 of the supernode
// (first equation `first`, order F); column k's entry at base + k*(F-1) -
// k*(k-1)/2 in LNZ
struct TmmSeg { i64 base; i32 F, first, kend, pad; };
struct EaTask { i32 child, c0, c1, pad; };        // child update columns [c0,c1)
struct PanelDesc { i32 node, k0, nb, nrb; };       // panel step at column k0, nrb row blocks
// Lower-trapezoidal tile set: rows [r0,rend) x cols [r0,cend) (square tiles
// of the launch's tile edge anchored at r0, only tiles with ti>=tj), updated
// with columns
// [k0,k0+kn) of the same front.
struct GemmDesc {
    i64 foff;
    i32 ld, r0, rend, cend, k0, kn;
    i32 nti, ntj;  // tile rows / tile cols
    i32 first;     // first equation of the front (pivots d come from the D array)
    i32 pad;
};

enum LaunchKind : i32 {
    L_INIT = 0,
    L_EA = 1,
    L_SMALL = 2,
    L_PANEL = 3,
    L_GEMM = 4,
    L_BDIAG = 5,   // outer-block diagonal factor (one workgroup per front)
    L_BTRSM = 6,   // rows below an outer block (16 per wavefront)
    L_RECORD = 7,  // no kernel: record the event of stream `stream`
    L_WAIT = 8,    // no kernel: stream `stream` waits for the other stream's event
    L_BFUSED = 9,  // outer-block diagonal factor and the rows below it in one launch
    L_STORE = 10,  // a level's factored panels from the pool to the factor store
};

// A launch of one kernel over a contiguous range of a task array.
struct Launch {
    i32 kind;
    i32 level;
    i64 off, count;  // range in the task array of this kind
    i64 blocks;      // grid size
    double flops;    // algorithmic flops of the launch (multiply-add = 2)
    i32 tile;        // MFMA update tile edge (L_GEMM)
    i32 outer;       // L_GEMM: 1 for rank-kOuterRank and Schur updates, 0 for narrower ones
    i32 stream = 0;  // 0: the handle's stream; 1: the second stream of the bulk updates
};

// Tunable schedule parameters (chosen by sprgpu_autotune or left at their
// defaults).  Changing them changes launches, not the data layout.
struct PlanParams {
    i32 small_f = kSmallF;    // fronts with F <= small_f are assembled and factored in LDS
    i32 small_waves = 1;      // such fronts per workgroup (one wavefront each)
    i32 big_tile = 0;         // MFMA tile of rank-kOuterRank/Schur updates (0: per launch,
                              // 128 for large FP32 launches, else 64)
    i32 narrow_threads = 0;   // workgroup size of the narrow solves (0: by front size)
    i32 block_factor = 1;     // blocked fronts: 1 = outer blocks in one launch each, 2 = their
                              // diagonal blocks and rows below as two launches, 0 = panel steps
    i32 forest = 1;           // 1: a plan whose fronts all have F <= kSmallF is factored (and
                              // solved, one right-hand side at a time) by the persistent forest
                              // kernel; 0: always the level schedule
};

struct Plan {
    PlanParams params;
    i64 neq = 0, nsuper = 0, nofsub = 0, nzerol = 0, mfront = 0, maxsup = 0;
    i64 pool_elems = 0, store_elems = 0, vwork_elems = 0;
    i32 nlevels = 0;
    i32 big_tile = 0;               // configured tile of the rank-kOuterRank and Schur updates
                                    // (0: per launch, Launch::tile)
    i32 small_f = 0;                // fronts with F <= small_f are factored in LDS (<= kSmallF)
    std::vector<NodeInfo> nodes;
    std::vector<i32> children;
    std::vector<i32> lindx;         // 0-based equations, NOFSUB
    std::vector<i32> relmap;        // per child: its U rows -> parent local row
    std::vector<i32> level_ptr, level_nodes;  // nodes grouped by height
    // factor schedule
    std::vector<ColTask> init_tasks;   // front initialisation of the blocked fronts
    std::vector<ColTask> copy_tasks;   // the S columns of every supernode (factor <-> SM)
    std::vector<ColTask> store_tasks;  // the blocked fronts' S columns (pool -> store), by level
    std::vector<EaTask> ea_tasks;
    std::vector<i32> small_nodes;
    std::vector<PanelDesc> panel_desc;  // panel steps, or outer blocks (k0 = ob, nb = width,
                                        // nrb = k_block_trsm workgroups below the block)
    std::vector<i64> panel_prefix;  // per launch-local prefix of row blocks (exclusive)
    std::vector<GemmDesc> gemm_desc;
    std::vector<i64> gemm_prefix;   // per launch-local prefix of tiles (exclusive)
    std::vector<Launch> launches;
    double factor_flops = 0;        // RINFO(3) formula, sprtrs.f:1961-1969
    // solve schedule: narrow supernodes one workgroup each, wide ones by a
    // dataflow kernel per level
    std::vector<i32> solve_small, solve_small_ptr;    // grouped by level
    std::vector<i32> small_maxF, small_maxS;          // per level, for the LDS size
    std::vector<WideDesc> wide;                       // grouped by level
    std::vector<i64> wide_fwd_prefix, wide_bwd_prefix, wide_rect_prefix;  // launch-local
    std::vector<i64> wide_frect_prefix;  // launch-local, the forward rectangle chunks alone
    std::vector<i32> wide_ptr;                        // nlevels+1
    std::vector<i64> wide_fwd_wgs, wide_bwd_wgs, wide_rect_wgs;  // workgroups per level
    std::vector<i64> wide_frect_wgs;
    i64 nflags = 0;                                   // blocks of all wide supernodes
    i64 npart = 0;                                    // backward partial sums per RHS
    // forward assembly as a gather: entry i of the solve workspace (node
    // n's local row i - n.voff) receives vwork[upd_src[k]] for k in
    // [upd_ptr[i], upd_ptr[i+1]), the children's update entries in SIBL order
    std::vector<i64> upd_ptr;
    std::vector<i32> upd_src;
    // persistent forest schedule (forest true): see ForestNode
    bool forest = false;
    std::vector<ForestNode> forest_nodes;
    std::vector<ForestChild> forest_child;
    // the ascending pass's supernodes (deepest first), then the backward
    // pass's (highest first), and the workgroup chunks over them (starts,
    // one depth or height each): forest_nasc ascending chunks, then the
    // backward chunks, then the end
    std::vector<i32> forest_order, forest_chunks;
    i32 forest_nasc = 0;
    i64 ublk_elems = 0;             // packed update blocks of all supernodes
};

// Builds the plan from the reference symbolic arrays (1-based, INTEGER*8)
// for a factor in the given precision (SPRGPU_FP64 / SPRGPU_FP32: the MFMA
// update tiles differ).  Returns an empty string or an error message.
std::string build_plan(const i64 *mspar, const i64 *mtrees, const i64 *msifa, int precision,
                       const PlanParams &params, Plan &plan);

// Pattern of A used by the FP64 residual: CSR over all rows, with the
// position of each value in the packed LNZ array.
struct CsrPattern {
    std::vector<i64> rowptr;  // neq+1
    std::vector<i32> col;     // nnz
    std::vector<i64> src;     // nnz, index into LNZ (0-based)
    i64 nnz_lower = 0;        // non-zeros in the packed lower triangle
};
void build_csr_pattern(const Plan &plan, const double *lnz, CsrPattern &pat);
// the same from a bit set over LNZ positions (bit pos set: in the pattern)
void build_csr_pattern_marked(const Plan &plan, const std::vector<uint64_t> &marked,
                              CsrPattern &pat);

struct FactorStatus {
    unsigned long long first_fail;   // smallest failing equation (UINT64_MAX = none)
    unsigned long long first_zero;   // smallest zero diagonal before the factor
    unsigned long long nzero;        // zero diagonals before the factor
    unsigned long long nrepl;        // pivots replaced (SPRFC3 mode)
    unsigned long long nnz_seen;     // non-zeros in the packed input
    unsigned long long wait_timeout; // waits of the fused outer-block launch that gave up
    unsigned long long pad[2];
};

struct DeviceBuffers {
    NodeInfo *nodes = nullptr;
    i32 *children = nullptr, *lindx = nullptr, *relmap = nullptr;
    ColTask *init_tasks = nullptr, *copy_tasks = nullptr, *store_tasks = nullptr;
    EaTask *ea_tasks = nullptr;
    i32 *small_nodes = nullptr;
    i32 *level_nodes = nullptr;
    PanelDesc *panel_desc = nullptr;
    i64 *panel_prefix = nullptr;
    GemmDesc *gemm_desc = nullptr;
    i64 *gemm_prefix = nullptr;
    void *pool = nullptr;       // T[pool_elems]: the fronts being factored
    void *store = nullptr;      // T[store_elems]: the factored panels
    double *dvec = nullptr;     // pivots d (NEQ), FP64
    double *adiag = nullptr;    // a_jj of the assembled matrix (NEQ)
    double *dcor = nullptr;     // DCOR(2,NEQ) for SPRFC3 mode
    FactorStatus *status = nullptr;
    double *stats = nullptr;    // reduction scratch for TOL(2)/TOL(3)/NUMNEG
    void *dscr = nullptr;       // factored diagonal blocks of the panel steps, T[NEQ*kPanelNB]
    void *dinv = nullptr;       // per wide-solve column block: inverses, M, M' (T[nflags*kBlockSolveData])
    void *pub = nullptr;        // fused outer-block launches: T[pub_elems], kPubElems per descriptor
    i64 pub_elems = 0, pub_cap = 0;
    double *vwork = nullptr;    // solve workspace (vwork_elems * nrhs_cap)
    i64 vwork_cap = 0;
    // residual pattern
    i64 *csr_rowptr = nullptr;
    i32 *csr_col = nullptr;
    i64 *csr_src = nullptr;
    double *csr_val = nullptr;
    i64 csr_nnz = 0, nnz_lower = 0;
    double *irwork = nullptr;   // IR vectors
    i32 *solve_small = nullptr;
    WideDesc *wide = nullptr;
    i64 *wide_fwd_prefix = nullptr, *wide_bwd_prefix = nullptr, *wide_rect_prefix = nullptr;
    i64 *wide_frect_prefix = nullptr;
    i64 *upd_ptr = nullptr;
    i32 *upd_src = nullptr;
    double *partials = nullptr; // backward partial sums (npart * nrhs)
    i64 partials_cap = 0;
    int *spin_error = nullptr;  // set when a dataflow wait gave up
    unsigned long long *prm_count = nullptr;  // non-zero counts of a product's A (2)
    i32 *col_node = nullptr;    // supernode of each equation (adjoint pattern gradient)
    // row index of the stored lower triangle by supernode (SPRTMM's L B):
    // row r's segments tmm_segs[k], k in [tmm_rptr[r], tmm_rptr[r+1]),
    // supernodes ascending
    i64 *tmm_rptr = nullptr;
    TmmSeg *tmm_segs = nullptr;
    i32 *tmm_long = nullptr;    // columns summed by a wavefront each (SPRTMM's L^T B)
    i64 tmm_nlong = 0;
    i64 irwork_cap = 0;
    double *spmm_work = nullptr;  // many-RHS product: X and Y equation-major
    i64 spmm_cap = 0;
    // element assembly map: the positions that receive element entries
    // (ascending; LNZ positions, then NZEROL + the SV entries), the range of
    // each one's contributions, and the contributions in SPRADM order, each
    // (element << 32) | (I + J*nedof); with constraint equations also each
    // contribution's weight code (kind << 56 | IP << 28 | JP, into TTCC)
    i64 *asm_tpos = nullptr, *asm_tptr = nullptr, *asm_src = nullptr;
    i64 *asm_emoff = nullptr;   // default element-matrix offsets (consecutive)
    i64 *asm_code = nullptr;
    i64 asm_ntgt = 0, asm_ncontrib = 0, asm_nel = 0;
    i64 asm_nsv = 0;            // NSV of the map (0: SM only)
    bool asm_ceq = false;       // built by set_assembly with constraint equations
    // PERI2E (MSPAR(58) >= 1): internal -> external equation, 0-based, and
    // the work space of the permuted entry points (two blocks)
    i32 *peri2e = nullptr;
    double *permwork = nullptr;
    i64 permwork_cap = 0;
    // host: peri2e_pmax[k] = max(PERI2E(1..k)) - 1 (0-based), k <= NEQ; the
    // leading k external equations are the leading k internal ones iff
    // peri2e_pmax[k] == k - 1
    std::vector<i32> peri2e_pmax;
    // persistent forest schedule
    ForestNode *forest_nodes = nullptr;
    ForestChild *forest_child = nullptr;
    i32 *forest_chunks = nullptr, *forest_order = nullptr;
    void *ublk = nullptr;       // T[ublk_elems]
    unsigned *forest_sync = nullptr;  // 2 + 3 NSUPER counters and flags
};

}  // namespace sprgpu

struct sprgpu_handle {
    sprgpu::Plan plan;
    sprgpu::DeviceBuffers d;
    int precision = SPRGPU_FP64;
    hipStream_t stream = nullptr;
    int device = 0;
    bool factored = false;
    bool has_pattern = false;
    int mode = 1;               // 1 SPRFC1, 2 SPRFC2 (CC' from the LDL' factor), 3 SPRFC3,
                                // 4 SPRFC4 (CC' from the LDL' factor with SPRFC4's pivot rule)
    double cc_tol1 = 0.0;       // TOL(1) of the last SPRFC2-mode factorization
    double tol[4] = {0, 0, 0, 0};
    const double *last_sm_dev = nullptr;  // input of the last factorization
    sprgpu_int device_bytes = 0;
    // symbolic input, kept to re-plan (sprgpu_autotune)
    std::vector<sprgpu_int> mspar, mtrees, msifa;
    // second stream of the factorization (the rank-kOuterRank updates of the
    // columns two groups ahead and beyond run there, beside the next group's
    // diagonal blocks) and one event per stream
    hipStream_t stream2 = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // captured factor + solve (sprgpu_factor_solve_graph)
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    const double *graph_sm = nullptr;
    double *graph_b = nullptr;
    sprgpu_int graph_ldb = 0, graph_nrhs = 0;
    // Generation of the device buffers a captured graph may reference: bumped
    // whenever one of them is freed and allocated again (the solve workspace
    // growing for more right-hand sides); a graph captured under another
    // generation holds stale pointers and is captured again.
    unsigned long long buf_gen = 0, graph_gen = 0;
    // Set by sprgpu_solve_noalloc while the solve buffers point into the
    // library's static scratch: the buffer-growth paths then refuse instead
    // of allocating.
    bool noalloc = false;
};

namespace sprgpu {
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
#define SPRGPU_HIP(call)                                              \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) return ::sprgpu::hip_fail(e_, #call);   \
    } while (0)

// forest.hip: the persistent forest schedule (plan.forest)
int forest_factor(sprgpu_handle *h, const double *sm_dev, int mode, const double *tol, double *b);
int forest_solve(sprgpu_handle *h, double *b, int which);
// factor.hip
int run_factor(sprgpu_handle *h, const double *sm_dev, int mode, const double *tol);
// factor and solve of one right-hand side b in the same launch (plan.forest)
int run_factor_solve(sprgpu_handle *h, const double *sm_dev, int mode, const double *tol,
                     double *b);
int read_factor_info(sprgpu_handle *h, sprgpu_factor_info *info);
int copy_factor_to_sm(sprgpu_handle *h, double *sm_dev_out);
int load_factor_from_sm(sprgpu_handle *h, const double *sm_dev);
// Inverses of the wide-solve diagonal blocks from the factor in the pool
// (after every factorization and factor upload), on the handle's stream.
int compute_block_inverses(sprgpu_handle *h);
int profile_factor(sprgpu_handle *h, const double *sm_dev, int mode, const double *tol,
                   double *out);
// solve.hip
int run_solve(sprgpu_handle *h, double *b, i64 ldb, i64 nrhs, int which);
int prepare_solve(sprgpu_handle *h, i64 nrhs);  // allocates the solve buffers for nrhs
int check_solve_error(sprgpu_handle *h);
int run_spmv(sprgpu_handle *h, const double *x, double *y, i64 ld, i64 nrhs);
int run_solve_ir(sprgpu_handle *h, const double *b, double *x, i64 ld, i64 nrhs, double rtol,
                 int maxit, sprgpu_ir_info *info);
int upload_pattern(sprgpu_handle *h, const CsrPattern &pat);
int run_sprprm(sprgpu_handle *h, const double *a, const double *b, double *c, double *wa,
               i64 m, i64 n, i64 ksa, i64 iflag);
// adjoint.hip
int run_element_grad(sprgpu_handle *h, i64 nel, int nedof, const i64 *elem_eq, const double *k0,
                     const i64 *k0_index, const double *lam, const double *x, double *grad);
int run_pattern_grad(sprgpu_handle *h, const double *lam, const double *x, double *glnz);
int upload_col_node(sprgpu_handle *h);  // supernode of every equation, once
// cholesky.hip: the CC' form (SPRFC2/FS2/BS2) on the LDL' factor
int cc_factor_info(sprgpu_handle *h, sprgpu_factor_info *info);
int cc_scale(sprgpu_handle *h, double *b, i64 ldb, i64 nrhs, int inverse);
int cc_factor_layout(sprgpu_handle *h, double *sm_dev);
int cc_to_ldl_layout(sprgpu_handle *h, double *sm_dev);
// lanczos.hip
int run_lanczos(sprgpu_handle *h, const double *a_sm, const double *b, int ksb, double shift,
                i64 neval, i64 nevec, i64 maxlan, double tol, double *evl, double *everr,
                double *v, i64 *nconv);
// products.hip
int run_sprtmm(sprgpu_handle *h, const double *u, const double *b, double *c, double *wa,
               i64 m, i64 n, i64 ksu, i64 iflag);
// C(i, j) = C(j, i) = B(:, i) . wa for i = j..n-1 (B is m x n, C n x n,
// column-major): SPRTRA's dot products
int run_tra_dots(sprgpu_handle *h, const double *b, i64 m, const double *wa, i64 n, i64 j,
                 double *c);
// noalloc.hip: the static scratch of sprgpu_solve_noalloc, located once per
// device at handle creation
int noalloc_scratch_init();
// permute.hip: the PERI2E permutation (MSPAR(58) >= 1)
int upload_permutation(sprgpu_handle *h, const i64 *msifa);
// two work blocks of NEQ x nrhs each
int perm_work(sprgpu_handle *h, i64 nrhs, double **w0, double **w1);
// dst(i, c) = src(PERI2E(i), c) and back, i < NEQ, c < nrhs
int perm_gather(sprgpu_handle *h, const double *src, i64 lds, double *dst, i64 ldd, i64 nrhs);
int perm_scatter(sprgpu_handle *h, const double *src, i64 lds, double *dst, i64 ldd, i64 nrhs);
// the leading nlead equations (external order when PERI2E is set):
// dst(i, c) = src(e(i), c) if e(i) < nlead else 0, e = PERI2E or identity
int lead_gather(sprgpu_handle *h, const double *src, i64 lds, double *dst, i64 ldd, i64 nrhs,
                i64 nlead);
// x(i, c) = 0 where e(i) >= nlead
int lead_mask(sprgpu_handle *h, double *x, i64 ldx, i64 nrhs, i64 nlead);
// dst(e(i), c) = src(i, c) where e(i) < nlead
int lead_scatter(sprgpu_handle *h, const double *src, i64 lds, double *dst, i64 ldd, i64 nrhs,
                 i64 nlead);
// assemble.hip
// constraint equations of the assembly (SAM's MPMCEQ, MMCEQ, MEQN; 1-based)
struct CeqInput {
    i64 nceq = 0;
    const i64 *mpmceq = nullptr, *mmceq = nullptr;
    i64 ndof = 0;
    const i64 *meqn = nullptr;
    i64 nsv = 0;
};
int set_assembly(sprgpu_handle *h, i64 nel, const i64 *meen_ptr, const i64 *meen,
                 const CeqInput *ceq);
int run_assemble(sprgpu_handle *h, const double *em, const i64 *em_off, const double *escale,
                 const double *ttcc, double *sm, double *sv, int accumulate);
void free_assembly(sprgpu_handle *h);
// elements.hip: hex8 element stiffness matrices (576 doubles each, consecutive)
int hex8_stiffness(sprgpu_handle *h, i64 nel, i64 nanod, const double *xyz, const i64 *mmnpc,
                   const double *emod, double nu, double *em);
}  // namespace sprgpu
