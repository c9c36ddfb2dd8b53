// Forward + diagonal and backward substitution with the supernodal factor.
//
// Replaces SPRFS1/SPRFS3/SPRFS5/SPRFS6 (sprfs1.f:5-633, :929-1141) and
// SPRBS1/SPRBS4/SPRBS6 (sprbs1.f:5-574).  The reference walks the supernodes
// one at a time, one right-hand side at a time, gathering B(LINDX) into a
// local vector and scattering the updated entries back by assignment.  Here
// the supernodes of one tree height run together, one launch per kind:
//
// * Narrow supernodes (S <= kSolveSmallS, F <= kSolveSmallF) get one
//   workgroup each with the local vector in LDS (sized for the level) and
//   the unit-triangular part solved by one wavefront with readlane
//   broadcasts.
// * Wide supernodes are split into kSolveNB-column blocks, one workgroup per
//   block, plus (forward) one workgroup per kSolveNB rows below the diagonal
//   part.  A block's workgroup loads the L values of each block it depends
//   on before waiting for that block's result and applies them as a
//   GEMV (forward: two threads per row; backward, the transposed product:
//   eight lanes per column, so that loads stay coalesced), then solves its
//   own 128 x 128 unit-triangular block as one more GEMV with the block's
//   inverse, computed after the factorization (k_block_inverse).
//   Every workgroup only waits for lower-indexed workgroups of the same
//   launch, which the hardware dispatches first, and each wait is bounded.
//   In the backward sweep the products with the rows below the diagonal part
//   (known on entry) are taken by a separate, fully parallel launch into
//   partial sums that the block workgroups add in a fixed order.
//
// The forward sweep hands each supernode's update vector (its F-S rows below
// the diagonal block) to the parent as a gather: entry i of the parent's
// vector sums the children's entries that map to it, in SIBL order (the same
// relative map the factor's extend-add uses), so no two workgroups write the
// same entry and every sum is taken in a fixed order: results do not depend
// on scheduling.  Vectors are FP64 whatever the factor precision; the
// division by d is a true division (sprfs1.f:622-625).
#include "lanes.h"
#include "sprgpu_internal.h"

namespace sprgpu {

constexpr unsigned long long kSpinLimit = 1ull << 24;
constexpr int NB = kSolveNB;
constexpr int NH = NB / 2;  // columns (forward) or rows (backward) per thread of a block GEMV
static_assert(NB == 128, "two threads per row (column) of a block GEMV, 64 entries each");

// Block results cross workgroups (and XCDs, whose L2 caches are not kept
// coherent with each other) as agent-scope atomic stores and loads, which
// are coherent on their own.  Before a sweep every block result slot holds
// kPending (k_mark_pending); a reader spins on each value it needs until it
// differs, so a result is handed over by its own store, with no ready flag
// and no wait for the stores to complete.  kPending is a signalling NaN:
// every arithmetic result is a number or a quiet NaN, so no result equals
// it.  This avoids the agent-scope fences, which write back or invalidate a
// whole L2.
constexpr unsigned long long kPending = 0xFFF7A5A5A5A5A5A5ull;

__device__ __forceinline__ void store_shared(double *p, double x)
{
    __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_shared(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The value at p once its writer has stored it, given a first look v
// (bounded: a wait that gives up sets *err and returns 0).
__device__ __forceinline__ double await_value(double v, const double *p, int *err)
{
    unsigned long long spins = 0;
    while (__builtin_bit_cast(unsigned long long, v) == kPending) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSpinLimit) {
            atomicExch(err, 1);
            return 0.0;
        }
        v = load_shared(p);
    }
    return v;
}

// v(lane with h = 0) + v(lane with h = 1) of the lane pair (lane, lane ^ 1),
// in both lanes, with the same operands in the same order.
__device__ __forceinline__ double pair_sum(double v, int h)
{
    const double o = dpp_move<0xB1>(v);  // quad_perm [1, 0, 3, 2]
    return h == 0 ? v + o : o + v;
}

// Every block result slot of the wide supernodes to kPending, for nrhs
// vectors: workgroup (b, r) takes block b (of all supernodes' blocks, in
// flag0 order) of vector r.
__global__ __launch_bounds__(NB) void k_mark_pending(const WideDesc *__restrict__ wide, int nwide,
                                                     const NodeInfo *__restrict__ nodes,
                                                     double *vwork, i64 vstride)
{
    const i64 blkid = blockIdx.x;
    int lo = 0, hi = nwide - 1;  // the supernode whose blocks hold blkid
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wide[mid].flag0 <= blkid)
            lo = mid;
        else
            hi = mid - 1;
    }
    const WideDesc w = wide[lo];
    const NodeInfo n = nodes[w.node];
    const int c = int(blkid - w.flag0) * NB + int(threadIdx.x);
    if (c < n.S)
        vwork[i64(blockIdx.y) * vstride + n.voff + c] =
            __builtin_bit_cast(double, kPending);
}

// Layouts of a diagonal block's inverse X (zero above the diagonal and past
// the block), the lower triangle's three 64 x 64 quadrants each:
//   forward:  columns 0..63 over all rows, X(r, c) at c*NB + r, then columns
//             64..127 over rows 64..127 at NB*NB/2 + (c-64)*64 + (r-64);
//   backward: rows 0..63 over columns 0..63, X(r, c) at r*64 + c, then rows
//             64..127 over all columns at NB*NB/4 + (r-64)*NB + c.
// A block GEMV walks a uniform index (column, resp. row) while the lanes take
// consecutive rows (columns): every read is a scalar offset from the lane's
// base, with consecutive lanes on consecutive addresses, and no entry outside
// the stored quadrants is ever needed (the one wavefront whose range is
// entirely above the diagonal skips).
static_assert(kInvLayout == 3 * NB * NB / 4, "three quadrants per layout");

// sum of p[k*stride] for k = k0, k0+step, ... < n in that order; the loads
// are issued eight at a time (a loop with a run-time trip count would wait
// for each load in turn)
__device__ __forceinline__ double strided_sum(const double *p, i64 stride, int k0, int step, int n)
{
    double s = 0.0;
    int k = k0;
    for (; k + 7 * step < n; k += 8 * step) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = p[i64(k + u * step) * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += t[u];
    }
    for (; k < n; k += step) s += p[i64(k) * stride];
    return s;
}

// ---------------------------------------------------------------------------
// narrow supernodes: S <= kSolveSmallS, F <= kSolveSmallF
// ---------------------------------------------------------------------------
// A workgroup serves R right-hand sides (rhs0 = (block % nrb) * R, ...): it
// reads each L value once and applies it to all R of them, each with the
// same sequence of operations as a single right-hand side.  Right-hand
// sides past nrhs (in the last group) load the last valid one's data and
// store nothing.
//
// Dynamic LDS: v[R][maxF] (FP64), then L11 (S x S, T).
template <typename T, int R>
__global__ __launch_bounds__(256) void k_fwd_narrow(const i32 *__restrict__ list, int maxF,
                                                    const NodeInfo *__restrict__ nodes,
                                                    const i32 *__restrict__ lindx,
                                                    const i64 *__restrict__ upd_ptr,
                                                    const i32 *__restrict__ upd_src,
                                                    const T *__restrict__ store,
                                                    const double *__restrict__ dvec, double *b,
                                                    i64 ldb, double *vwork, i64 vstride, int nrhs)
{
    extern __shared__ double smem[];
    __shared__ double vx[R][64];  // x of the S columns, zero past S
    const int nrb = (nrhs + R - 1) / R;
    double *v = smem;  // v[r * maxF + i]
    T *Lb = reinterpret_cast<T *>(smem + R * maxF);  // Lb[c*S + r] = L(r,c), r > c
    const NodeInfo n = nodes[list[blockIdx.x / nrb]];
    const int rhs0 = int(blockIdx.x % nrb) * R;
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    const int S = n.S, F = n.F;
    const i32 *rows = lindx + n.lxoff;
    const T *L = store + n.loff;
    const i64 ld = n.ld;
    const i64 *up = upd_ptr + n.voff;
    for (int i = tid; i < F; i += nt) {
        const i64 k0 = up[i], k1 = up[i + 1];
        const i32 gi = (i < S) ? rows[i] : 0;
#pragma unroll 4
        for (int r = 0; r < R; ++r) {
            const i64 rh = min(rhs0 + r, nrhs - 1);
            const double *vw = vwork + rh * vstride;
            double s = (i < S) ? b[rh * ldb + gi] : 0.0;
            for (i64 k = k0; k < k1; ++k) s += vw[upd_src[k]];
            v[r * maxF + i] = s;
        }
    }
    for (int idx = tid; idx < S * S; idx += nt) {
        const int r = idx % S, c = idx / S;
        if (r > c) Lb[idx] = L[r + c * ld];
    }
    __syncthreads();
    // unit lower L11, one right-hand side per wavefront at a time: lane r
    // holds x_r; column c is final when reached.  Columns are read from LDS
    // eight at a time (in bounds for every lane) ahead of the dependent
    // readlane/FMA chain.
    for (int q = wave; q < R; q += nw) {
        double x = (lane < S) ? v[q * maxF + lane] : 0.0;
        const int lc = min(lane, S - 1);
        for (int cb = 0; cb + 1 < S; cb += 8) {
            T la[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) la[u] = Lb[min(cb + u, S - 1) * S + lc];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int c = cb + u;
                if (c + 1 < S) {
                    const double xc = lane_bcast(x, c);
                    x -= ((lane > c && lane < S) ? double(la[u]) : 0.0) * xc;
                }
            }
        }
        if (lane < S && rhs0 + q < nrhs) b[i64(rhs0 + q) * ldb + rows[lane]] = x / dvec[n.first + lane];
        vx[q][lane] = (lane < S) ? x : 0.0;
    }
    __syncthreads();
    // update vectors for the parent: s_i = v_i - sum_{c<S} L(i,c) x_c, the
    // columns in batches of 16 (clamped loads, vx is zero past S)
    constexpr int RS = R < 8 ? R : 8;  // right-hand sides accumulated at a time
    for (int i = S + tid; i < F; i += nt) {
        const T *Li = L + i;
        for (int q0 = 0; q0 < R; q0 += RS) {
            double s0[RS], s1[RS], s2[RS], s3[RS];
#pragma unroll
            for (int q = 0; q < RS; ++q) s0[q] = s1[q] = s2[q] = s3[q] = 0.0;
            for (int cb = 0; cb < S; cb += 16) {
                T lv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) lv[u] = Li[i64(min(cb + u, S - 1)) * ld];
#pragma unroll
                for (int q = 0; q < RS; ++q)
#pragma unroll
                    for (int u = 0; u < 16; u += 4) {
                        s0[q] += double(lv[u]) * vx[q0 + q][cb + u];
                        s1[q] += double(lv[u + 1]) * vx[q0 + q][cb + u + 1];
                        s2[q] += double(lv[u + 2]) * vx[q0 + q][cb + u + 2];
                        s3[q] += double(lv[u + 3]) * vx[q0 + q][cb + u + 3];
                    }
            }
#pragma unroll
            for (int q = 0; q < RS; ++q)
                if (rhs0 + q0 + q < nrhs)
                    vwork[i64(rhs0 + q0 + q) * vstride + n.voff + i] =
                        v[(q0 + q) * maxF + i] - ((s0[q] + s1[q]) + (s2[q] + s3[q]));
        }
    }
}

// Dynamic LDS: v[R][maxF + 16] (FP64), part[R][waves*64] (FP64), then L11 (S x S, T).
template <typename T, int R>
__global__ __launch_bounds__(256) void k_bwd_narrow(const i32 *__restrict__ list, int maxF,
                                                    const NodeInfo *__restrict__ nodes,
                                                    const i32 *__restrict__ lindx,
                                                    const T *__restrict__ store, double *b,
                                                    i64 ldb, int nrhs)
{
    extern __shared__ double smem[];
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    const int VS = maxF + 16, PS = nt;     // PS: partial sums per right-hand side
    double *v = smem;                      // v[r * VS + i]: F entries and 16 zeros
    double *part = smem + R * VS;          // part[r * PS + wave * 64 + lane]
    T *Lb = reinterpret_cast<T *>(part + R * PS);  // Lb[r*S + c] = L(r,c), r > c
    const int nrb = (nrhs + R - 1) / R;
    const NodeInfo n = nodes[list[blockIdx.x / nrb]];
    const int rhs0 = int(blockIdx.x % nrb) * R;
    const int S = n.S, F = n.F;
    const i32 *rows = lindx + n.lxoff;
    const T *L = store + n.loff;
    const i64 ld = n.ld;
    for (int i = tid; i < F + 16; i += nt) {
        const i32 gi = (i < F) ? rows[i] : 0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const i64 rh = min(rhs0 + q, nrhs - 1);
            v[q * VS + i] = (i < F) ? b[rh * ldb + gi] : 0.0;
        }
    }
    for (int idx = tid; idx < S * S; idx += nt) {
        const int r = idx % S, c = idx / S;
        if (r > c) Lb[r * S + c] = L[r + c * ld];
    }
    __syncthreads();
    // x1 -= L21^T x2: lane = column, the waves split the rows in multiples
    // of 16 that are read 16 at a time (clamped loads, v is zero past F)
    {
        const T *Lcol = L + i64(min(lane, S - 1)) * ld;
        const int per = ((F - S + nw - 1) / nw + 15) / 16 * 16;
        const int ra = S + wave * per, rz = min(F, ra + per);
        constexpr int RS = R < 8 ? R : 8;  // right-hand sides accumulated at a time
        for (int q0 = 0; q0 < R; q0 += RS) {
            double a0[RS], a1[RS], a2[RS], a3[RS];
#pragma unroll
            for (int q = 0; q < RS; ++q) a0[q] = a1[q] = a2[q] = a3[q] = 0.0;
            for (int r = ra; r < rz; r += 16) {
                T lv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) lv[u] = Lcol[min(r + u, F - 1)];
#pragma unroll
                for (int q = 0; q < RS; ++q)
#pragma unroll
                    for (int u = 0; u < 16; u += 4) {
                        a0[q] += double(lv[u]) * v[(q0 + q) * VS + r + u];
                        a1[q] += double(lv[u + 1]) * v[(q0 + q) * VS + r + u + 1];
                        a2[q] += double(lv[u + 2]) * v[(q0 + q) * VS + r + u + 2];
                        a3[q] += double(lv[u + 3]) * v[(q0 + q) * VS + r + u + 3];
                    }
            }
#pragma unroll
            for (int q = 0; q < RS; ++q)
                part[(q0 + q) * PS + wave * 64 + lane] = (a0[q] + a1[q]) + (a2[q] + a3[q]);
        }
    }
    __syncthreads();
    // L11^T x1 = y1 by columns from the last, one right-hand side per
    // wavefront at a time: x_c -= l_rc x_r for c < r
    const int lc = min(lane, S - 1);
    for (int q = wave; q < R; q += nw) {
        double x = 0.0;
        if (lane < S) {
            double sp = part[q * PS + lane];
            for (int k = 1; k < nw; ++k) sp += part[q * PS + k * 64 + lane];
            x = v[q * VS + lane] - sp;
        }
        for (int rb = S - 1; rb > 0; rb -= 8) {
            T la[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) la[u] = Lb[max(rb - u, 1) * S + lc];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int r = rb - u;
                if (r > 0) {
                    const double xr = lane_bcast(x, r);
                    x -= ((lane < r) ? double(la[u]) : 0.0) * xr;
                }
            }
        }
        if (lane < S && rhs0 + q < nrhs) b[i64(rhs0 + q) * ldb + rows[lane]] = x;
    }
}

// ---------------------------------------------------------------------------
// wide supernodes: block products
// ---------------------------------------------------------------------------
// Rows r..r+3 of an L column as one vector load (at any element alignment:
// gfx950 loads a dwordx4 at 4-byte alignment).  Callers keep the four rows
// inside the front: a group that starts at or past F is read at F - 4
// instead (its values are not used), and one that runs past F starts on a
// multiple of 4 (then its rows stay below the column's padded length ld) or
// lies in a column with another one after it in the front.  Every load is
// issued unconditionally, so all of a block's loads stay in flight together
// (a second, clamped load path would make the compiler wait for the loads
// of one path before those of the other).
template <typename T>
struct Rows4 {
    T v[4];
};
typedef unsigned long long lanemask;
template <typename T>
__device__ __forceinline__ Rows4<T> load_rows4(const T *Lc, int r, int F)
{
    typedef T v4u __attribute__((ext_vector_type(4), aligned(sizeof(T))));
    const v4u v = *reinterpret_cast<const v4u *>(Lc + (r < F ? r : F - 4));
    Rows4<T> o;
#pragma unroll
    for (int i = 0; i < 4; ++i) o.v[i] = v[i];
    return o;
}

// Forward: y_i -= sum_p L(i, block p) x_p over a chunk of kSolveNB rows
// (rows r0 + 4g + i of the front).  Every forward kernel takes these
// products the same way, so results do not depend on how right-hand sides
// are grouped: thread (g, cg) = (tid / 8, tid % 8) takes rows 4g..4g+3 of
// the chunk (a vector load per column) and columns 16cg..16cg+15 of the
// block, sums its 16 products per row in column order, the eight lanes of
// a row group add their sums (sum8), and lane cg = i < 4, the owner of row
// 4g + i, adds the block's sum to the row's total, blocks in ascending
// order.  Staged x is padded by 2 values per 16 (xs_at), so that the eight
// column groups of a wavefront read distinct LDS banks.
constexpr int kXsLd = NB + NB / 8;
__device__ __forceinline__ int xs_at(int c) { return c + 2 * (c >> 4); }

// (Rows past F give values that only reach rows past F, which are not
// stored.)
template <typename T>
__device__ __forceinline__ void fwd_load_block(Rows4<T> (&dst)[16], const T *L, i64 ld, int F,
                                               int S, int r, int cg, int p)
{
    const int c0 = p * NB, nc = min(NB, S - c0);
#pragma unroll
    for (int c = 0; c < 16; ++c) dst[c] = load_rows4(L + i64(c0 + min(16 * cg + c, nc - 1)) * ld, r, F);
}

template <typename T, int R>
__device__ __forceinline__ void fwd_block_sums(const Rows4<T> (&lv)[16],
                                               const double (*__restrict__ xb)[kXsLd], int cg,
                                               double (&tot)[R])
{
    constexpr int RS = R < 8 ? R : 8;  // right-hand sides per pass
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += RS) {
        double part[4][RS];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int k = 0; k < RS; ++k) part[i][k] = 0.0;
#pragma unroll
        for (int c = 0; c < 16; ++c)
#pragma unroll
            for (int k = 0; k < RS; ++k) {
                const double xv = xb[k0 + k][xs_at(16 * cg + c)];
#pragma unroll
                for (int i = 0; i < 4; ++i) part[i][k] += double(lv[c].v[i]) * xv;
            }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int k = 0; k < RS; ++k) {
                const double v = sum8(part[i][k]);
                if (cg == i) tot[k0 + k] += v;
            }
    }
}

// Backward: y_c -= sum_p L(block p, c)^T x_p over the columns c of block
// blk: thread (cq, rr) = (tid / 8, tid % 8) takes columns 4cq..4cq+3 and
// rows 4rr + 32m + i (m, i < 4) of block p (a vector load per column and
// m: the eight lanes of a column read 128 consecutive bytes), sums its 16
// products per column in row order, the eight lanes add their sums (sum8),
// and lane rr = j < 4, the owner of column 4cq + j, adds the block's sum to
// the column's total, blocks in descending order.
// Rows past F are summed with x = 0 but may hold anything (the front's
// padding rows are never written), so a wavefront that reaches F zeroes
// them first (a scalar branch around arithmetic only).
template <typename T>
__device__ __forceinline__ void bwd_load_block(Rows4<T> (&dst)[16], const T *L, i64 ld, int F,
                                               int c0, int nc, int cq, int rr, int p)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const T *Lc = L + i64(c0 + min(4 * cq + j, nc - 1)) * ld;
#pragma unroll
        for (int m = 0; m < 4; ++m) dst[4 * j + m] = load_rows4(Lc, p * NB + 4 * rr + 32 * m, F);
    }
}

template <typename T, int R>
__device__ __forceinline__ void bwd_block_sums(Rows4<T> (&lv)[16],
                                               const double (*__restrict__ xb)[NB], int rr,
                                               int F, int p, double (&tot)[R])
{
    if (__builtin_amdgcn_ballot_w64(p * NB + 4 * rr + 99 >= F) != lanemask(0)) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (p * NB + 4 * rr + 32 * m + i >= F) lv[4 * j + m].v[i] = T(0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < R; ++k) {
            double part = 0.0;
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    part += double(lv[4 * j + m].v[i]) * xb[k][4 * rr + 32 * m + i];
            const double v = sum8(part);
            if (rr == j) tot[k] += v;
        }
}

// ---------------------------------------------------------------------------
// wide supernodes, forward
// ---------------------------------------------------------------------------
// Workgroup local < nblk solves diagonal block local after applying blocks
// 0..local-1; workgroups nblk.. compute the update vector of kSolveNB rows
// below the diagonal part, applying every block as it finishes.  As in the
// narrow kernels, a workgroup serves R right-hand sides.  Each step of the
// block loop costs one barrier; the diagonal block costs one more (its
// right-hand side, complete, to LDS for the product with its inverse,
// thread (ri, h) = (tid / 2, tid % 2) taking row ri and column half h, the
// halves added with one DPP exchange).
template <typename T, int R>
__global__ __launch_bounds__(256) void k_fwd_wide(const WideDesc *__restrict__ wide, int nwide,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i32 *__restrict__ lindx,
                                                  const i64 *__restrict__ upd_ptr,
                                                  const i32 *__restrict__ upd_src,
                                                  const T *__restrict__ store,
                                                  const double *__restrict__ dvec, double *b,
                                                  i64 ldb, double *vwork, i64 vstride, int *err,
                                                  int nrhs, const T *__restrict__ dinv)
{
    __shared__ double xs[4][R][kXsLd];
    __shared__ double ys[R][NB];
    __shared__ double zs[R][NB];
    __shared__ T Lf[kInvLayout];  // L11^-1, forward layout
    const int nrb = (nrhs + R - 1) / R;
    const i64 bid = blockIdx.x / nrb;
    const int rhs0 = int(blockIdx.x % nrb) * R;
    const int q = find_desc(prefix, nwide, bid);
    const WideDesc w = wide[q];
    const int local = int(bid - prefix[q]);
    const NodeInfo n = nodes[w.node];
    const int tid = threadIdx.x;
    const int S = n.S, F = n.F;
    const i32 *rows = lindx + n.lxoff;
    const T *L = store + n.loff;
    const i64 ld = n.ld;
    const i64 *up = upd_ptr + n.voff;
    i64 vo[R];  // this group's vectors of the supernode (the last valid one past nrhs)
#pragma unroll
    for (int k = 0; k < R; ++k) vo[k] = i64(min(rhs0 + k, nrhs - 1)) * vstride + n.voff;
    const bool diag = local < w.nblk;
    const int r0 = diag ? local * NB : S + (local - w.nblk) * NB;
    const int nr = min(NB, (diag ? S : F) - r0);
    const int cg = tid & 7, g = tid >> 3;
    // the row I own (lanes cg < 4): B (own columns) plus the children's updates
    const int orow = 4 * g + cg;
    const bool owner = cg < 4 && orow < nr;
    // its destination and pivot (a diagonal block's rows; read here, before
    // any wait, for the stores at the end)
    const int orc = diag ? r0 + min(orow, nr - 1) : 0;
    const i32 dst = rows[orc];
    const double dpiv = dvec[n.first + orc];
    double y[R], tot[R];
#pragma unroll
    for (int k = 0; k < R; ++k) y[k] = tot[k] = 0.0;
    if (owner) {
        const int row = r0 + orow;
        const i64 k0 = up[row], k1 = up[row + 1];
        const i32 gi = (row < S) ? rows[row] : 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const i64 rh = min(rhs0 + k, nrhs - 1);
            double v = (row < S) ? b[rh * ldb + gi] : 0.0;
            const double *vw = vwork + rh * vstride;
            for (i64 j = k0; j < k1; ++j) v += vw[upd_src[j]];
            y[k] = v;
        }
    }
    const T *bdat = dinv + (w.flag0 + (diag ? local : 0)) * i64(kBlockSolveData);
    if (diag) {
        for (int idx = tid; idx < kInvLayout; idx += 256) Lf[idx] = bdat[idx];
    }
    // A diagonal block takes the blocks before the previous one here, and
    // the previous one through M after its own inverse product
    const int nprev = diag ? max(local - 1, 0) : w.nblk;
    // Blocks are taken two at a time while both are available ahead of the
    // last one (with one right-hand side; a group takes one at a time, two
    // blocks' L values would not fit in registers): the looks at their x and
    // the loads of their L values are issued together, so that a step costs
    // one memory round trip and one barrier.  The last block (on the chain's
    // critical path for a diagonal block) goes alone.  The staged x
    // alternate between two pairs of LDS buffers, so no barrier is needed
    // before they are rewritten.
    constexpr int PB = R == 1 ? 2 : 1;  // blocks per step
    Rows4<T> lv[PB][16];
    int it = 0;
    auto step = [&](int p, int nb) {
        double xv[PB][R];
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            const int c0 = (p + u) * NB, nc = min(NB, S - c0);
#pragma unroll
            for (int k = 0; k < R; ++k)
                xv[u][k] = (u < nb && tid < nc) ? load_shared(vwork + vo[k] + c0 + tid) : 0.0;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
        for (int u = 0; u < PB; ++u)
            fwd_load_block(lv[u], L, ld, F, S, r0 + 4 * g, cg, min(p + u, nprev - 1));
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (tid < NB) {
#pragma unroll
            for (int u = 0; u < PB; ++u) {
                const int c0 = (p + u) * NB, nc = min(NB, S - c0);
#pragma unroll
                for (int k = 0; k < R; ++k)
                    xs[2 * (it & 1) + u][k][xs_at(tid)] =
                        (u < nb && tid < nc) ? await_value(xv[u][k], vwork + vo[k] + c0 + tid, err) : 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PB; ++u)
            if (u < nb) fwd_block_sums<T, R>(lv[u], xs[2 * (it & 1) + u], cg, tot);
        ++it;
    };
    int p = 0;
    for (; p + PB < nprev; p += PB) step(p, PB);
    for (; p < nprev; ++p) step(p, 1);
    if (!diag) {
        if (owner) {  // update for the parent
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (rhs0 + k < nrhs) vwork[vo[k] + r0 + orow] = y[k] - tot[k];
        }
        return;
    }
    // Diagonal block: z = L11^-1 y with the block's inverse in LDS, thread
    // (ri, h) summing columns [64h, 64h + 64) of row ri; in the forward
    // layout, columns 64..127 of rows 0..63 are zero and not stored.
    if (cg < 4) {
#pragma unroll
        for (int k = 0; k < R; ++k) ys[k][orow] = owner ? y[k] - tot[k] : 0.0;
    }
    __syncthreads();
    const int ri = tid >> 1, h = tid & 1;
    const bool above = h == 1 && ri < NH;
    const T *Lrow = Lf + (h == 0 ? ri : NB * NB / 2 + (ri - NH));
    const int cs = h == 0 ? NB : NH;
    double z[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double a0 = 0.0, a1 = 0.0;
        if (!above) {
#pragma unroll 16
            for (int c = 0; c < NH; c += 2) {
                a0 += double(Lrow[c * cs]) * ys[k][h * NH + c];
                a1 += double(Lrow[(c + 1) * cs]) * ys[k][h * NH + c + 1];
            }
        }
        z[k] = pair_sum(a0 + a1, h);
    }
    if (local == 0) {  // the first block: x = z
        if (h == 0 && ri < nr) {
            const i32 zdst = rows[r0 + ri];
            const double zpiv = dvec[n.first + r0 + ri];
            // the result for the blocks that wait on it first, then B
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (rhs0 + k < nrhs) store_shared(vwork + vo[k] + r0 + ri, z[k]);
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (rhs0 + k < nrhs) b[i64(rhs0 + k) * ldb + zdst] = z[k] / zpiv;
        }
        return;
    }
    // x = z - M x_prev, M = L11^-1 L(block, previous block): the products
    // are taken as block products (fwd_block_sums) with M for L, and the
    // owner lane of a row stores its x
    if (h == 0) {
#pragma unroll
        for (int k = 0; k < R; ++k) zs[k][ri] = z[k];
    }
    // (the look at x_prev goes before M's loads, so that finding it stored
    // does not wait for them)
    const int cp = (local - 1) * NB;  // the previous block: full
    double xv[R];
#pragma unroll
    for (int k = 0; k < R; ++k) xv[k] = (tid < NB) ? load_shared(vwork + vo[k] + cp + tid) : 0.0;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    Rows4<T> mv[16];
    fwd_load_block(mv, bdat + 2 * kInvLayout, NB, NB, NB, 4 * g, cg, 0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (tid < NB) {
#pragma unroll
        for (int k = 0; k < R; ++k) xs[0][k][xs_at(tid)] = await_value(xv[k], vwork + vo[k] + cp + tid, err);
    }
    __syncthreads();
    double mt[R];
#pragma unroll
    for (int k = 0; k < R; ++k) mt[k] = 0.0;
    fwd_block_sums<T, R>(mv, xs[0], cg, mt);
    if (owner) {
        double x[R];
#pragma unroll
        for (int k = 0; k < R; ++k) x[k] = zs[k][orow] - mt[k];
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (rhs0 + k < nrhs) store_shared(vwork + vo[k] + r0 + orow, x[k]);
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (rhs0 + k < nrhs) b[i64(rhs0 + k) * ldb + dst] = x[k] / dpiv;
    }
}

// The rows below the diagonal part of the wide supernodes as a launch of
// their own, after the supernodes' block chains (k_fwd_wide over the
// diagonal blocks only): used for groups of right-hand sides (R = 8 or 32),
// where a group reads L21 once and nothing waits.  Each workgroup takes
// kSolveNB rows of one supernode with the block products of k_fwd_wide, so
// the update vectors are identical bit for bit.  The gathered children's
// updates go to the rows' vector slots first and are read back at the end.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_fwd_rect(const WideDesc *__restrict__ wide, int nwide,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i64 *__restrict__ upd_ptr,
                                                  const i32 *__restrict__ upd_src,
                                                  const T *__restrict__ store, double *vwork,
                                                  i64 vstride, int nrhs)
{
    __shared__ double xs[2][R][kXsLd];
    const int nrb = (nrhs + R - 1) / R;
    const i64 bid = blockIdx.x / nrb;
    const int rhs0 = int(blockIdx.x % nrb) * R;
    const int q = find_desc(prefix, nwide, bid);
    const WideDesc w = wide[q];
    const int chunk = int(bid - prefix[q]);
    const NodeInfo n = nodes[w.node];
    const int tid = threadIdx.x;
    const int S = n.S, F = n.F;
    const T *L = store + n.loff;
    const i64 ld = n.ld;
    const i64 *up = upd_ptr + n.voff;
    const int r0 = S + chunk * NB;
    const int nr = min(NB, F - r0);
    const int cg = tid & 7, g = tid >> 3;
    const int orow = 4 * g + cg;
    const bool owner = cg < 4 && orow < nr;
    if (owner) {  // the children's updates of my row (no B entries below the diagonal part)
        const i64 k0 = up[r0 + orow], k1 = up[r0 + orow + 1];
        for (int k = 0; k < R; ++k) {
            if (rhs0 + k >= nrhs) break;
            double *vw = vwork + i64(rhs0 + k) * vstride;
            double v = 0.0;
            for (i64 j = k0; j < k1; ++j) v += vw[upd_src[j]];
            vw[n.voff + r0 + orow] = v;
        }
    }
    double tot[R];
#pragma unroll
    for (int k = 0; k < R; ++k) tot[k] = 0.0;
    Rows4<T> lv[16];
    for (int p = 0; p < w.nblk; ++p) {
        const int c0 = p * NB, nc = min(NB, S - c0);
        fwd_load_block(lv, L, ld, F, S, r0 + 4 * g, cg, p);
        double(*xb)[kXsLd] = xs[p & 1];
        if (tid < NB) {
            for (int k = 0; k < R; ++k) {
                const i64 rh = min(rhs0 + k, nrhs - 1);
                xb[k][xs_at(tid)] = (tid < nc) ? vwork[rh * vstride + n.voff + c0 + tid] : 0.0;
            }
        }
        __syncthreads();
        fwd_block_sums<T, R>(lv, xb, cg, tot);
    }
    if (owner) {  // update for the parent
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (rhs0 + k >= nrhs) break;
            double *slot = vwork + i64(rhs0 + k) * vstride + n.voff + r0 + orow;
            *slot = *slot - tot[k];
        }
    }
}

// ---------------------------------------------------------------------------
// wide supernodes, backward
// ---------------------------------------------------------------------------
// Partial sums of L21^T x2: workgroup (blk, k) multiplies rows [S+k*NB, +NB)
// of the columns of block blk by the solution at those rows (final: they
// belong to ancestors).  Same mapping as the block products of k_bwd_wide:
// the thread sums columns gc + 8j (j < 4) over rows 8q + gs (q < 16), so a
// load instruction reads 8 rows of 8 columns, and the eight lanes of a
// column are added with sum8.  The L values stay in registers while the
// right-hand sides pass in groups of at most 8; every column's sum takes the
// same operations whatever R is.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_bwd_rect(const WideDesc *__restrict__ wide, int nwide,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i32 *__restrict__ lindx,
                                                  const T *__restrict__ store,
                                                  const double *__restrict__ b, i64 ldb,
                                                  double *partials, i64 npart, int nrhs)
{
    constexpr int KS = R < 8 ? R : 8;  // right-hand sides per pass
    static_assert(R % KS == 0, "whole passes");
    __shared__ double xs[R][NB];
    const int nrb = (nrhs + R - 1) / R;
    const i64 bid = blockIdx.x / nrb;
    const int rhs0 = int(blockIdx.x % nrb) * R;
    const int q = find_desc(prefix, nwide, bid);
    const WideDesc w = wide[q];
    const int t = int(bid - prefix[q]);
    const int blk = t / w.nrect, k = t % w.nrect;
    const NodeInfo n = nodes[w.node];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int S = n.S, F = n.F;
    const int c0 = blk * NB, nc = min(NB, S - c0);
    const int r0 = S + k * NB, nr = min(NB, F - r0);
    const T *L = store + n.loff;
    const i64 ld = n.ld;
    if (tid < NB) {
        const i32 gi = lindx[n.lxoff + r0 + min(tid, nr - 1)];
#pragma unroll
        for (int j = 0; j < R; ++j)
            xs[j][tid] = (tid < nr) ? b[i64(min(rhs0 + j, nrhs - 1)) * ldb + gi] : 0.0;
    }
    // unconditional loads (columns past the block read its last column and
    // are not stored; rows past the chunk read its last row, xs is zero)
    const int gs = lane & 7, gc = 32 * wave + (lane >> 3);
    T lv[4][NB / 8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const T *Lc = L + r0 + i64(c0 + min(gc + 8 * j, nc - 1)) * ld;
#pragma unroll
        for (int qq = 0; qq < NB / 8; ++qq) lv[j][qq] = Lc[min(8 * qq + gs, nr - 1)];
    }
    __syncthreads();
    double *dst = partials + w.pbase + (i64(blk) * w.nrect + k) * NB;
#pragma unroll 1
    for (int g = 0; g < R; g += KS) {
        double acc[4][KS];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < KS; ++r) acc[j][r] = 0.0;
#pragma unroll 4
        for (int qq = 0; qq < NB / 8; ++qq)
#pragma unroll
            for (int r = 0; r < KS; ++r) {
                const double xv = xs[g + r][8 * qq + gs];
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j][r] += double(lv[j][qq]) * xv;
            }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < KS; ++r) {
                const double v = sum8(acc[j][r]);
                if (gs == 0 && gc + 8 * j < nc && rhs0 + g + r < nrhs)
                    dst[i64(rhs0 + g + r) * npart + gc + 8 * j] = v;
            }
    }
}

// ---------------------------------------------------------------------------
// wide supernodes, many right-hand sides: the rectangle products on MFMA
// ---------------------------------------------------------------------------
// With kMmGroup or more right-hand sides, the rectangle products of the wide
// supernodes (forward: the update rows L21 x1; backward: the partial sums of
// L21^T x2) are GEMMs over kMmGroup right-hand sides at a time on the FP64
// 16x16x4 MFMA, so a workgroup reads its part of L once for all of them.  An
// FP32 factor is widened as it is staged: the products stay FP64, as in the
// other kernels, but are summed in another order, so these columns agree with
// a solve of fewer right-hand sides to rounding, not bit for bit.
// Both kernels form C(rhs, j) = sum_k X(k, rhs) Y(k, j) with the right-hand
// sides as MFMA operand A (As[k][rhs]) and the factor as operand B
// (Bs[k][j], j = a row of L21 forward, a column backward), so the 16 adjacent
// j of an accumulator tile are adjacent in memory when it is stored.  A
// workgroup computes kMmGroup right-hand sides x kMmCols j, wavefront w the 16
// j from 16w; K runs in stages of kMmK staged in two LDS buffers: the next
// stage's global loads are issued before this stage's MFMAs and written to
// the other buffer after them, one barrier per stage.
constexpr int kMmGroup = 64;            // right-hand sides per workgroup
constexpr int kMmCols = 64;             // rows (forward) or columns (backward) of L21 per workgroup
constexpr int kMmK = 32;                // K per stage
constexpr int kMmLd = 80;               // LDS row stride (doubles): 16-lane groups on distinct banks
constexpr int kMmTiles = kMmGroup / 16; // accumulator tiles per wavefront
static_assert(kMmCols == 64 && NB % kMmCols == 0, "four wavefronts of 16, whole halves of a chunk");

typedef double mm_acc_t __attribute__((ext_vector_type(4)));

// The stage loop shared by both kernels: fetch(k0) issues the global loads
// of the stage at k0 into registers, put(buf, k0) writes them to LDS.  The
// accumulator element e of tile t of a lane is C(16t + lane/16 + 4e, 16w +
// lane%16) (FP64 16x16x4: A[i][k] from lane i + 16k, B[k][j] from lane
// j + 16k, element e of row lane/16 + 4e).
template <typename FetchF, typename PutF>
__device__ __forceinline__ void mm_stages(double (*As)[kMmK][kMmLd], double (*Bs)[kMmK][kMmLd],
                                          int nst, FetchF &&fetch, PutF &&put,
                                          mm_acc_t (&acc)[kMmTiles])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < kMmTiles; ++t) acc[t] = mm_acc_t{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    put(0, 0);
    __syncthreads();
    for (int s = 0; s < nst; ++s) {
        const int buf = s & 1;
        const bool more = s + 1 < nst;
        if (more) fetch((s + 1) * kMmK);
#pragma unroll
        for (int kk = 0; kk < kMmK / 4; ++kk) {
            const int k = 4 * kk + (lane >> 4);
            const double bv = Bs[buf][k][16 * wave + (lane & 15)];
#pragma unroll
            for (int t = 0; t < kMmTiles; ++t)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[buf][k][16 * t + (lane & 15)], bv,
                                                              acc[t], 0, 0, 0);
        }
        if (more) put(buf ^ 1, (s + 1) * kMmK);
        __syncthreads();
    }
}

// Forward: workgroup (chunk, half, group) computes rows [r0, r0 + kMmCols)
// of a wide supernode's update vector (r0 = S + NB chunk + kMmCols half) for
// kMmGroup right-hand sides: y = (the children's entries of the row, summed
// in SIBL order) - sum_k L(r, k) x_k over the S columns.  Staging: thread
// (tid / 8, tid % 8) loads rows 8 (tid % 8) .. + 7 of L column tid / 8 (two
// vector loads, clamped as in load_rows4) and x_k, k = 4 (tid % 8) .. + 3,
// of right-hand sides tid / 8 and tid / 8 + 32.  Columns past S read
// column S - 1 against x = 0; rows past F are not stored.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_rect_mm(const WideDesc *__restrict__ wide, int nwide,
                                                     const i64 *__restrict__ prefix,
                                                     const NodeInfo *__restrict__ nodes,
                                                     const i64 *__restrict__ upd_ptr,
                                                     const i32 *__restrict__ upd_src,
                                                     const T *__restrict__ store, double *vwork,
                                                     i64 vstride, int nrhs)
{
    __shared__ double As[2][kMmK][kMmLd];
    __shared__ double Bs[2][kMmK][kMmLd];
    const int nrg = (nrhs + kMmGroup - 1) / kMmGroup;
    const i64 bid = blockIdx.x / nrg;
    const int rhs0 = int(blockIdx.x % nrg) * kMmGroup;
    const i64 cid = bid / (NB / kMmCols);
    const int q = find_desc(prefix, nwide, cid);
    const WideDesc w = wide[q];
    const NodeInfo n = nodes[w.node];
    const int S = n.S, F = n.F;
    const int r0 = S + int(cid - prefix[q]) * NB + int(bid % (NB / kMmCols)) * kMmCols;
    if (r0 >= F) return;  // the whole workgroup: a chunk's second half past F
    const T *L = store + n.loff;
    const i64 ld = n.ld;
    const int tid = threadIdx.x;
    const int sc = tid >> 3, sr = 8 * (tid & 7), sk = 4 * (tid & 7);
    Rows4<T> lv[2];
    double xv[2][4];
    auto fetch = [&](int k0) {
        const T *Lc = L + i64(min(k0 + sc, S - 1)) * ld;
        lv[0] = load_rows4(Lc, r0 + sr, F);
        lv[1] = load_rows4(Lc, r0 + sr + 4, F);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const double *xp = vwork + i64(min(rhs0 + sc + 32 * h, nrhs - 1)) * vstride + n.voff;
#pragma unroll
            for (int i = 0; i < 4; ++i) xv[h][i] = xp[min(k0 + sk + i, S - 1)];
        }
    };
    auto put = [&](int buf, int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            Bs[buf][sc][sr + i] = double(lv[0].v[i]);
            Bs[buf][sc][sr + 4 + i] = double(lv[1].v[i]);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                As[buf][sk + i][sc + 32 * h] =
                    (k0 + sk + i < S && rhs0 + sc + 32 * h < nrhs) ? xv[h][i] : 0.0;
    };
    mm_acc_t acc[kMmTiles];
    mm_stages(As, Bs, (S + kMmK - 1) / kMmK, fetch, put, acc);
    const int lane = tid & 63;
    const int r = r0 + 16 * (tid >> 6) + (lane & 15);
    if (r < F) {
        const i64 j0 = upd_ptr[n.voff + r], j1 = upd_ptr[n.voff + r + 1];
#pragma unroll
        for (int t = 0; t < kMmTiles; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rh = rhs0 + 16 * t + (lane >> 4) + 4 * e;
                if (rh < nrhs) {
                    double *vw = vwork + i64(rh) * vstride;
                    double v = 0.0;
                    for (i64 j = j0; j < j1; ++j) v += vw[upd_src[j]];
                    vw[n.voff + r] = v - acc[t][e];
                }
            }
    }
}

// Backward: workgroup (blk, chunk, half, group) computes the partial sums
// sum_r L(r, c) x_r over the NB rows of a rectangle chunk (as k_bwd_rect) for
// the kMmCols columns c from NB blk + kMmCols half and kMmGroup right-hand
// sides.  Staging: thread (tid / 4, tid % 4) loads rows 8 (tid % 4) .. + 7
// of the stage in column tid / 4 (two vector loads, written to LDS
// transposed) and thread (tid / 8, tid % 8) x at rows 4 (tid % 8) .. + 3
// (through the row list) for right-hand sides tid / 8 and tid / 8 + 32.
// Rows past the chunk are read clamped against x = 0; columns past the
// block read its last column and are not stored.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_rect_mm(const WideDesc *__restrict__ wide, int nwide,
                                                     const i64 *__restrict__ prefix,
                                                     const NodeInfo *__restrict__ nodes,
                                                     const i32 *__restrict__ lindx,
                                                     const T *__restrict__ store,
                                                     const double *__restrict__ b, i64 ldb,
                                                     double *partials, i64 npart, int nrhs)
{
    __shared__ double As[2][kMmK][kMmLd];
    __shared__ double Bs[2][kMmK][kMmLd];
    const int nrg = (nrhs + kMmGroup - 1) / kMmGroup;
    const i64 bid = blockIdx.x / nrg;
    const int rhs0 = int(blockIdx.x % nrg) * kMmGroup;
    const i64 pid = bid / (NB / kMmCols);
    const int half = int(bid % (NB / kMmCols));
    const int q = find_desc(prefix, nwide, pid);
    const WideDesc w = wide[q];
    const int t = int(pid - prefix[q]);
    const int blk = t / w.nrect, kch = t % w.nrect;
    const NodeInfo n = nodes[w.node];
    const int S = n.S, F = n.F;
    const int c0 = blk * NB + half * kMmCols, nc = min(kMmCols, S - c0);
    if (nc <= 0) return;  // the whole workgroup: a block's second half past S
    const int r0 = S + kch * NB, nr = min(NB, F - r0);
    const T *L = store + n.loff;
    const i64 ld = n.ld;
    const int tid = threadIdx.x;
    const int lc = tid >> 2, lr = 8 * (tid & 3);
    const int xc = tid >> 3, xk = 4 * (tid & 7);
    const T *Lc = L + r0 + i64(c0 + min(lc, nc - 1)) * ld;
    const i32 *rows = lindx + n.lxoff + r0;
    Rows4<T> lv[2];
    double xv[2][4];
    auto fetch = [&](int k0) {
        lv[0] = load_rows4(Lc, k0 + lr, F - r0);
        lv[1] = load_rows4(Lc, k0 + lr + 4, F - r0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const i32 gi = rows[min(k0 + xk + i, nr - 1)];
#pragma unroll
            for (int h = 0; h < 2; ++h)
                xv[h][i] = b[i64(min(rhs0 + xc + 32 * h, nrhs - 1)) * ldb + gi];
        }
    };
    auto put = [&](int buf, int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {  // rows past the chunk may be unwritten memory
            Bs[buf][lr + i][lc] = (k0 + lr + i < nr) ? double(lv[0].v[i]) : 0.0;
            Bs[buf][lr + 4 + i][lc] = (k0 + lr + 4 + i < nr) ? double(lv[1].v[i]) : 0.0;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                As[buf][xk + i][xc + 32 * h] =
                    (k0 + xk + i < nr && rhs0 + xc + 32 * h < nrhs) ? xv[h][i] : 0.0;
    };
    mm_acc_t acc[kMmTiles];
    mm_stages(As, Bs, (nr + kMmK - 1) / kMmK, fetch, put, acc);
    const int lane = tid & 63;
    const int c = 16 * (tid >> 6) + (lane & 15);
    if (c < nc) {
        double *dst = partials + w.pbase + (i64(blk) * w.nrect + kch) * NB + half * kMmCols + c;
#pragma unroll
        for (int tt = 0; tt < kMmTiles; ++tt)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rh = rhs0 + 16 * tt + (lane >> 4) + 4 * e;
                if (rh < nrhs) dst[i64(rh) * npart] = acc[tt][e];
            }
    }
}

// Workgroup local handles block nblk-1-local: y = B - partial sums of the
// rows below, minus the later blocks' products as they finish (bwd_block_sums),
// then x = L11^-T y on the block with its inverse.  Each step of the block
// loop costs one barrier (as in k_fwd_wide); the diagonal block costs one
// more.  For the product with the inverse, thread (ci, h) = (tid / 2,
// tid % 2) takes column ci and row half h, and the two halves (adjacent
// lanes) are added with one DPP exchange.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_bwd_wide(const WideDesc *__restrict__ wide, int nwide,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i32 *__restrict__ lindx,
                                                  const T *__restrict__ store, double *b, i64 ldb,
                                                  double *vwork, i64 vstride,
                                                  const double *__restrict__ partials, i64 npart,
                                                  int *err, int nrhs, const T *__restrict__ dinv)
{
    __shared__ double xs[4][R][NB];
    __shared__ double ys[R][NB];
    __shared__ double zs[R][NB];
    __shared__ T Lb[kInvLayout];  // L11^-1, backward layout
    const int nrb = (nrhs + R - 1) / R;
    const i64 bid = blockIdx.x / nrb;
    const int rhs0 = int(blockIdx.x % nrb) * R;
    const int q = find_desc(prefix, nwide, bid);
    const WideDesc w = wide[q];
    const int blk = w.nblk - 1 - int(bid - prefix[q]);
    const NodeInfo n = nodes[w.node];
    const int tid = threadIdx.x;
    const int S = n.S, F = n.F;
    const i32 *rows = lindx + n.lxoff;
    const T *L = store + n.loff;
    const i64 ld = n.ld;
    i64 vo[R];
#pragma unroll
    for (int k = 0; k < R; ++k) vo[k] = i64(min(rhs0 + k, nrhs - 1)) * vstride + n.voff;
    const int c0 = blk * NB, nc = min(NB, S - c0);
    const T *bdat = dinv + (w.flag0 + blk) * i64(kBlockSolveData);
    for (int idx = tid; idx < kInvLayout; idx += 256) Lb[idx] = bdat[kInvLayout + idx];
    const int ci = tid >> 1, h = tid & 1;
    const bool cv = ci < nc;
    const i32 dst = rows[c0 + min(ci, nc - 1)];
    const i32 odst = rows[c0 + min(4 * (tid >> 3) + (tid & 7), nc - 1)];  // of the M' step's owner column
    // products with the rows below the diagonal part: the two lanes of a
    // column sum the even and the odd row chunks' partials
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double *ps = partials + i64(min(rhs0 + k, nrhs - 1)) * npart + w.pbase +
                           i64(blk) * w.nrect * NB + min(ci, nc - 1);
        const double sp = pair_sum(strided_sum(ps, NB, h, 2, w.nrect), h);
        if (h == 0)
            ys[k][ci] = cv ? b[i64(min(rhs0 + k, nrhs - 1)) * ldb + dst] - sp : 0.0;
    }
    // products with the later blocks (bwd_block_sums), into the totals of the
    // columns' owner lanes
    const int rr = tid & 7, cq = tid >> 3;
    double tot[R];
#pragma unroll
    for (int k = 0; k < R; ++k) tot[k] = 0.0;
    // Steps as in k_fwd_wide, blocks from the last one down.
    constexpr int PB = R == 1 ? 2 : 1;  // blocks per step
    Rows4<T> lv[PB][16];
    int it = 0;
    auto step = [&](int p, int nb) {  // blocks p, p - 1, ..., p - nb + 1
        double xv[PB][R];
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            const int rp0 = (p - u) * NB, np = min(NB, S - rp0);
#pragma unroll
            for (int k = 0; k < R; ++k)
                xv[u][k] = (u < nb && tid < np) ? load_shared(vwork + vo[k] + rp0 + tid) : 0.0;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
        for (int u = 0; u < PB; ++u)
            bwd_load_block(lv[u], L, ld, F, c0, nc, cq, rr, max(p - u, blk + 1));
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (tid < NB) {
#pragma unroll
            for (int u = 0; u < PB; ++u) {
                const int rp0 = (p - u) * NB, np = min(NB, S - rp0);
#pragma unroll
                for (int k = 0; k < R; ++k)
                    xs[2 * (it & 1) + u][k][tid] =
                        (u < nb && tid < np) ? await_value(xv[u][k], vwork + vo[k] + rp0 + tid, err) : 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PB; ++u)
            if (u < nb) bwd_block_sums<T, R>(lv[u], xs[2 * (it & 1) + u], rr, F, p - u, tot);
        ++it;
    };
    // blocks after the next one here; the next one through M' after the
    // block's own inverse product
    int p = w.nblk - 1;
    for (; p - PB > blk; p -= PB) step(p, PB);
    for (; p > blk + 1; --p) step(p, 1);
    __syncthreads();  // ys complete (its initial values) before the owners update it
    if (rr < 4 && 4 * cq + rr < nc) {
#pragma unroll
        for (int k = 0; k < R; ++k) ys[k][4 * cq + rr] -= tot[k];
    }
    __syncthreads();
    // Diagonal block: z = L11^-T y with the block's inverse in LDS, thread
    // (ci, h) summing rows [64h, 64h + 64) of column ci; in the backward
    // layout, columns 64..127 of rows 0..63 are zero and not stored.
    const bool above = h == 0 && ci >= NH;
    const T *Lcol = Lb + (h == 0 ? ci : NB * NB / 4 + ci);
    const int rs = h == 0 ? NH : NB;
    double z[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double a0 = 0.0, a1 = 0.0;
        if (!above) {
#pragma unroll 16
            for (int r = 0; r < NH; r += 2) {
                a0 += double(Lcol[r * rs]) * ys[k][h * NH + r];
                a1 += double(Lcol[(r + 1) * rs]) * ys[k][h * NH + r + 1];
            }
        }
        z[k] = pair_sum(a0 + a1, h);
    }
    if (blk == w.nblk - 1) {  // the last block: x = z
        if (h == 0 && cv) {
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (rhs0 + k < nrhs) store_shared(vwork + vo[k] + c0 + ci, z[k]);
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (rhs0 + k < nrhs) b[i64(rhs0 + k) * ldb + dst] = z[k];
        }
        return;
    }
    // x = z - M' x_next, M' = L11^-T L(next block, block)^T: block products
    // as in the forward sweep (fwd_block_sums, M' for L, the rows of M' are
    // the block's columns), the owner lane of a column stores its x
    if (h == 0) {
#pragma unroll
        for (int k = 0; k < R; ++k) zs[k][ci] = z[k];
    }
    const int g = tid >> 3, cg = tid & 7, oc = 4 * g + cg;
    const bool owner = cg < 4 && oc < nc;
    const int rn = (blk + 1) * NB, nn = min(NB, S - rn);  // the next block
    double xv[R];
#pragma unroll
    for (int k = 0; k < R; ++k) xv[k] = (tid < nn) ? load_shared(vwork + vo[k] + rn + tid) : 0.0;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    Rows4<T> mv[16];
    fwd_load_block(mv, bdat + 2 * kInvLayout + NB * NB, NB, NB, NB, 4 * g, cg, 0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    double(*xm)[kXsLd] = reinterpret_cast<double(*)[kXsLd]>(&xs[0][0][0]);
    if (tid < NB) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            xm[k][xs_at(tid)] = (tid < nn) ? await_value(xv[k], vwork + vo[k] + rn + tid, err) : 0.0;
    }
    __syncthreads();
    double mt[R];
#pragma unroll
    for (int k = 0; k < R; ++k) mt[k] = 0.0;
    fwd_block_sums<T, R>(mv, xm, cg, mt);
    if (owner) {
        double x[R];
#pragma unroll
        for (int k = 0; k < R; ++k) x[k] = zs[k][oc] - mt[k];
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (rhs0 + k < nrhs) store_shared(vwork + vo[k] + c0 + oc, x[k]);
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (rhs0 + k < nrhs) b[i64(rhs0 + k) * ldb + odst] = x[k];
    }
}

// ---------------------------------------------------------------------------
// inverses of the wide diagonal blocks
// ---------------------------------------------------------------------------
// X = L11^-1 of column block blk of wide supernode w (unit lower triangular,
// nc <= NB), computed in double by block rows of 32 and stored as T in the
// forward and the backward layout (zero above the diagonal and past nc) at
// dinv + kBlockSolveData (flag0 + blk), followed by M = X L(blk, blk - 1)
// (blocks after the first) and M' = X^T L(blk + 1, blk)^T (blocks before
// the last), column-major NB x NB, zero past the block's rows (columns):
// with them a sweep applies the neighbouring block's result with one
// product after everything else of the block is done.  Block row I: A) T = -L(I, <I) X(<I, :) for the
// columns left of the block row (the identity for its own columns), B)
// X(I, :) = L(I, I)^-1 T by substitution, one column per thread.
template <typename T>
__global__ __launch_bounds__(256) void k_block_inverse(const WideDesc *__restrict__ wide, int nwide,
                                                       const NodeInfo *__restrict__ nodes,
                                                       const T *__restrict__ store, T *dinv)
{
    __shared__ double X[NB][NB + 1];  // X[c][r]; odd stride: a column walk is conflict-free
    __shared__ double Lii[32][33];    // the block row's diagonal 32 x 32 block of L
    __shared__ double Lst[16][NB + 1];  // 16 rows (columns) of L for M (M')
    const i64 b = blockIdx.x;
    int lo = 0, hi = nwide - 1;  // the supernode whose blocks hold b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wide[mid].flag0 <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const WideDesc w = wide[lo];
    const int blk = int(b - w.flag0);
    const NodeInfo n = nodes[w.node];
    const T *L = store + n.loff;
    const i64 ld = n.ld;
    const int c0 = blk * NB, nc = min(NB, n.S - c0);
    const int tid = threadIdx.x;
    for (int idx = tid; idx < NB * (NB + 1); idx += 256) (&X[0][0])[idx] = 0.0;
    const int r = tid & 31, cg = tid >> 5;  // step A: row r of the block row, columns cg + 8m
    for (int I = 0; 32 * I < nc; ++I) {
        const int rI = 32 * I + r;
        // the diagonal block of L (zero past nc and on and above the diagonal)
        for (int idx = tid; idx < 32 * 32; idx += 256) {
            const int i = idx & 31, k = idx >> 5;
            const int gi = 32 * I + i, gk = 32 * I + k;
            Lii[i][k] = (i > k && gi < nc) ? double(L[(c0 + gi) + i64(c0 + gk) * ld]) : 0.0;
        }
        __syncthreads();  // X rows above the block row and Lii complete
        if (I > 0) {
            double acc[12];
#pragma unroll
            for (int m = 0; m < 12; ++m) acc[m] = 0.0;
            const T *Lr = L + (c0 + min(rI, nc - 1)) + i64(c0) * ld;
            for (int k = 0; k < 32 * I; ++k) {
                const double a = double(Lr[i64(k) * ld]);
#pragma unroll
                for (int m = 0; m < 12; ++m) acc[m] += a * X[cg + 8 * m][k];  // zero for k < c
            }
            if (rI < nc) {
#pragma unroll
                for (int m = 0; m < 12; ++m)
                    if (cg + 8 * m < 32 * I) X[cg + 8 * m][rI] = -acc[m];
            }
        }
        __syncthreads();
        // B: columns c < 32I + 32 of the block row, one per thread
        const int c = tid;
        if (c < min(nc, 32 * I + 32)) {
            double x[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) x[i] = X[c][32 * I + i] + (c == 32 * I + i ? 1.0 : 0.0);
            // right-looking, a column of Lii per step (zero on and above its
            // diagonal, so every row takes the update); a rolled loop keeps
            // the Lii loads from being hoisted all at once
#pragma unroll 1
            for (int k = 0; k < 31; ++k) {
                double xk = 0.0;
#pragma unroll
                for (int i = 0; i < 32; ++i) xk = (i == k) ? x[i] : xk;
#pragma unroll
                for (int i = 0; i < 32; ++i) x[i] -= Lii[i][k] * xk;
            }
#pragma unroll
            for (int i = 0; i < 32; ++i)
                if (32 * I + i < nc) X[c][32 * I + i] = x[i];
        }
        __syncthreads();
    }
    // both layouts (X is zero above the diagonal; rows past nc were never
    // written: zero)
    T *out = dinv + (w.flag0 + blk) * i64(kBlockSolveData);
    for (int idx = tid; idx < kInvLayout; idx += 256) {
        int rr, cc;
        if (idx < NB * NB / 2) {
            cc = idx >> 7, rr = idx & (NB - 1);
        } else {
            const int t = idx - NB * NB / 2;
            cc = NH + (t >> 6), rr = NH + (t & (NH - 1));
        }
        out[idx] = T(X[cc][rr]);
    }
    for (int idx = tid; idx < kInvLayout; idx += 256) {
        int rr, cc;
        if (idx < NB * NB / 4) {
            rr = idx >> 6, cc = idx & (NH - 1);
        } else {
            const int t = idx - NB * NB / 4;
            rr = NH + (t >> 7), cc = t & (NB - 1);
        }
        out[kInvLayout + idx] = T(X[cc][rr]);
    }
    // M(r, c) = sum_i X(r, i) L(c0 + i, c0 - NB + c): thread (r, half) takes
    // 64 columns; 16 rows of L at a time in LDS (the previous block is full)
    const int rm = tid & (NB - 1), half = tid >> 7;
    if (blk > 0) {
        double acc[NH];
#pragma unroll
        for (int c = 0; c < NH; ++c) acc[c] = 0.0;
        for (int i0 = 0; i0 < nc; i0 += 16) {
            __syncthreads();  // Lst free
            for (int idx = tid; idx < 16 * NB; idx += 256) {
                const int ii = idx & 15, c = idx >> 4;
                Lst[ii][c] = (i0 + ii < nc) ? double(L[(c0 + i0 + ii) + i64(c0 - NB + c) * ld]) : 0.0;
            }
            __syncthreads();
#pragma unroll 4
            for (int ii = 0; ii < 16; ++ii) {
                const double xv = X[i0 + ii][rm];
#pragma unroll
                for (int c = 0; c < NH; ++c) acc[c] += xv * Lst[ii][half * NH + c];
            }
        }
        T *m = out + 2 * kInvLayout;
#pragma unroll
        for (int c = 0; c < NH; ++c) m[(half * NH + c) * NB + rm] = rm < nc ? T(acc[c]) : T(0);
    }
    // M'(c, r) = sum_i X(i, c) L(c0 + NB + r, c0 + i): thread (c, half) takes
    // 64 rows r of the next block (np of them valid)
    if (blk + 1 < w.nblk) {
        const int np = min(NB, n.S - (c0 + NB));
        double acc[NH];
#pragma unroll
        for (int r = 0; r < NH; ++r) acc[r] = 0.0;
        for (int i0 = 0; i0 < nc; i0 += 16) {
            __syncthreads();  // Lst free
            for (int idx = tid; idx < 16 * NB; idx += 256) {
                const int r = idx & (NB - 1), ii = idx >> 7;
                Lst[ii][r] = (r < np && i0 + ii < nc) ? double(L[(c0 + NB + r) + i64(c0 + i0 + ii) * ld]) : 0.0;
            }
            __syncthreads();
#pragma unroll 4
            for (int ii = 0; ii < 16; ++ii) {
                const double xv = X[rm][i0 + ii];
#pragma unroll
                for (int r = 0; r < NH; ++r) acc[r] += xv * Lst[ii][half * NH + r];
            }
        }
        T *m = out + 2 * kInvLayout + NB * NB;
#pragma unroll
        for (int r = 0; r < NH; ++r) m[(half * NH + r) * NB + rm] = rm < nc ? T(acc[r]) : T(0);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int ensure_solve_buffers(sprgpu_handle *h, i64 nrhs)
{
    DeviceBuffers &d = h->d;
    const Plan &p = h->plan;
    if (h->noalloc) {  // the buffers are views of the static scratch, sized by the caller
        if (d.vwork_cap < nrhs || d.partials_cap < std::max<i64>(p.npart, 1) * nrhs ||
            !d.spin_error) {
            set_error("solve workspace larger than the scratch view");
            return -2;
        }
        return 0;
    }
    if (d.vwork_cap < nrhs) {
        ++h->buf_gen;
        if (d.vwork) SPRGPU_HIP(hipFree(d.vwork));
        d.vwork = nullptr;
        SPRGPU_HIP(hipMalloc(&d.vwork, sizeof(double) * size_t(p.vwork_elems) * size_t(nrhs)));
        h->device_bytes += sizeof(double) * p.vwork_elems * (nrhs - d.vwork_cap);
        d.vwork_cap = nrhs;
    }
    const i64 np = std::max<i64>(p.npart, 1) * nrhs;
    if (d.partials_cap < np) {
        ++h->buf_gen;
        if (d.partials) SPRGPU_HIP(hipFree(d.partials));
        d.partials = nullptr;
        SPRGPU_HIP(hipMalloc(&d.partials, sizeof(double) * np));
        h->device_bytes += sizeof(double) * (np - d.partials_cap);
        d.partials_cap = np;
    }
    if (!d.spin_error) {
        SPRGPU_HIP(hipMalloc(&d.spin_error, sizeof(int)));
        SPRGPU_HIP(hipMemset(d.spin_error, 0, sizeof(int)));
    }
    return 0;
}

// Right-hand sides per workgroup: kSolveGroup when a call has at least that
// many (the wide kernels always, the narrow ones while their LDS stays within
// 64 KB, else 4 or 1), one otherwise.
constexpr int kSolveGroup = 8;
constexpr size_t kNarrowLds = 64 * 1024;

template <typename T>
static size_t fwd_narrow_lds(int R, int mf, int ms, int nt)
{
    (void)nt;
    return sizeof(double) * size_t(R) * mf + sizeof(T) * size_t(ms) * ms;
}
template <typename T>
static size_t bwd_narrow_lds(int R, int mf, int ms, int nt)
{
    return sizeof(double) * size_t(R) * (mf + 16 + nt) + sizeof(T) * size_t(ms) * ms;
}
// workgroup size of a level's narrow launches
static unsigned narrow_threads(const Plan &p, int l)
{
    if (p.params.narrow_threads) return unsigned(p.params.narrow_threads);
    return p.small_maxF[l] <= 128 ? 64u : 256u;
}
// Right-hand sides per workgroup of a level's narrow launches: the larger of
// kSolveGroup and 4 whose LDS fits in kNarrowLds, or 1 below kSolveGroup
// right-hand sides.  (Larger groups were measured slower, 32 on quad4 1000^2
// with 128 right-hand sides by 1.6x: a workgroup's gathers and triangular
// chains for its right-hand sides run in sequence.)
// Right-hand sides per workgroup of the wide supernodes' block chains when a
// call has at least kSolveGroup of them: 4 with an FP32 factor, kSolveGroup
// with FP64 (40^3 x 128: the FP32 + IR step 140.7 ms with 8, 136.7 ms with
// 4, 138.6 ms with 2; FP64 137.2 ms with 8, 139.0 ms with 4)
template <typename T>
constexpr int wide_group()
{
    return sizeof(T) == 4 ? 4 : kSolveGroup;
}

template <typename T>
static int narrow_group(const Plan &p, int l, i64 nrhs)
{
    if (nrhs < kSolveGroup) return 1;
    const int mf = p.small_maxF[l], ms = p.small_maxS[l], nt = int(narrow_threads(p, l));
    for (int R : {kSolveGroup, 4}) {
        if (std::max(fwd_narrow_lds<T>(R, mf, ms, nt), bwd_narrow_lds<T>(R, mf, ms, nt)) <=
            kNarrowLds)
            return R;
    }
    return 1;
}

// The inverses of the wide solve's 128 x 128 diagonal blocks (k_block_inverse).
template <typename T>
static const T *block_inverses(sprgpu_handle *h)
{
    return static_cast<const T *>(h->d.dinv);
}

template <typename T, int R>
static void fwd_narrow_level(sprgpu_handle *h, hipStream_t st, int l, double *b, i64 ldb,
                             i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    const T *store = static_cast<const T *>(d.store);
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const unsigned ns = unsigned(p.solve_small_ptr[l + 1] - p.solve_small_ptr[l]);
    const int mf = p.small_maxF[l], ms = p.small_maxS[l];
    const unsigned nt = narrow_threads(p, l);
    k_fwd_narrow<T, R><<<ns * groups, nt, fwd_narrow_lds<T>(R, mf, ms, nt), st>>>(
        d.solve_small + p.solve_small_ptr[l], mf, d.nodes, d.lindx, d.upd_ptr, d.upd_src, store,
        d.dvec, b, ldb, d.vwork, p.vwork_elems, int(nrhs));
}

// The level's wide supernodes forward: the block chains with their rectangle
// rows in the same launch (fused), or the block chains alone.
template <typename T, int R>
static void fwd_wide_level(sprgpu_handle *h, int l, double *b, i64 ldb, i64 nrhs, bool fused)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    const T *store = static_cast<const T *>(d.store);
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
    const i64 wgs = fused ? p.wide_fwd_wgs[l] : p.wide_bwd_wgs[l];
    const i64 *prefix = fused ? d.wide_fwd_prefix : d.wide_bwd_prefix;  // all blocks, or the chains'
    k_fwd_wide<T, R><<<unsigned(wgs) * groups, 256, 0, h->stream>>>(
        d.wide + w0, nw, prefix + w0, d.nodes, d.lindx, d.upd_ptr, d.upd_src, store, d.dvec, b,
        ldb, d.vwork, p.vwork_elems, d.spin_error, int(nrhs), block_inverses<T>(h));
}

template <typename T, int R>
static void fwd_rect_level(sprgpu_handle *h, int l, i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    if (p.wide_frect_wgs[l] == 0) return;
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
    k_fwd_rect<T, R><<<unsigned(p.wide_frect_wgs[l]) * groups, 256, 0, h->stream>>>(
        d.wide + w0, nw, d.wide_frect_prefix + w0, d.nodes, d.upd_ptr, d.upd_src,
        static_cast<const T *>(d.store), d.vwork, p.vwork_elems, int(nrhs));
}

template <typename T, int R>
static void bwd_narrow_level(sprgpu_handle *h, hipStream_t st, int l, double *b, i64 ldb,
                             i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    const T *store = static_cast<const T *>(d.store);
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const unsigned ns = unsigned(p.solve_small_ptr[l + 1] - p.solve_small_ptr[l]);
    const int mf = p.small_maxF[l], ms = p.small_maxS[l];
    const unsigned nt = narrow_threads(p, l);
    k_bwd_narrow<T, R><<<ns * groups, nt, bwd_narrow_lds<T>(R, mf, ms, nt), st>>>(
        d.solve_small + p.solve_small_ptr[l], mf, d.nodes, d.lindx, store, b, ldb, int(nrhs));
}

template <typename T, int R>
static void bwd_wide_level(sprgpu_handle *h, int l, double *b, i64 ldb, i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    const T *store = static_cast<const T *>(d.store);
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const i64 npart = std::max<i64>(p.npart, 1);
    const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
    k_bwd_wide<T, R><<<unsigned(p.wide_bwd_wgs[l]) * groups, 256, 0, h->stream>>>(
        d.wide + w0, nw, d.wide_bwd_prefix + w0, d.nodes, d.lindx, store, b, ldb, d.vwork,
        p.vwork_elems, d.partials, npart, d.spin_error, int(nrhs), block_inverses<T>(h));
}

template <typename T, int R>
static void bwd_rect_level(sprgpu_handle *h, int l, double *b, i64 ldb, i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    if (p.wide_rect_wgs[l] == 0) return;
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
    k_bwd_rect<T, R><<<unsigned(p.wide_rect_wgs[l]) * groups, 256, 0, h->stream>>>(
        d.wide + w0, nw, d.wide_rect_prefix + w0, d.nodes, d.lindx, static_cast<const T *>(d.store),
        b, ldb, d.partials, std::max<i64>(p.npart, 1), int(nrhs));
}

// Right-hand sides per workgroup of the rectangle launches (no waits, L21
// read once per group): kRectGroup when a call has that many, and the MFMA
// kernels (kMmGroup per workgroup) from kMmGroup right-hand sides on.
constexpr int kRectGroup = 32;

template <typename T>
static void fwd_rect_mm_level(sprgpu_handle *h, int l, i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    if (p.wide_frect_wgs[l] == 0) return;
    const unsigned groups = unsigned((nrhs + kMmGroup - 1) / kMmGroup);
    const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
    k_fwd_rect_mm<T><<<unsigned(p.wide_frect_wgs[l]) * (NB / kMmCols) * groups, 256, 0,
                       h->stream>>>(d.wide + w0, nw, d.wide_frect_prefix + w0, d.nodes, d.upd_ptr,
                                    d.upd_src, static_cast<const T *>(d.store), d.vwork,
                                    p.vwork_elems, int(nrhs));
}

template <typename T>
static void bwd_rect_mm_level(sprgpu_handle *h, int l, double *b, i64 ldb, i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    if (p.wide_rect_wgs[l] == 0) return;
    const unsigned groups = unsigned((nrhs + kMmGroup - 1) / kMmGroup);
    const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
    k_bwd_rect_mm<T><<<unsigned(p.wide_rect_wgs[l]) * (NB / kMmCols) * groups, 256, 0,
                       h->stream>>>(d.wide + w0, nw, d.wide_rect_prefix + w0, d.nodes, d.lindx,
                                    static_cast<const T *>(d.store), b, ldb, d.partials,
                                    std::max<i64>(p.npart, 1), int(nrhs));
}

// Every block result slot of the wide supernodes to kPending, before a sweep.
static void mark_pending(sprgpu_handle *h, i64 nrhs)
{
    const Plan &p = h->plan;
    if (p.nflags == 0) return;
    k_mark_pending<<<dim3(unsigned(p.nflags), unsigned(nrhs)), NB, 0, h->stream>>>(
        h->d.wide, int(p.wide.size()), h->d.nodes, h->d.vwork, p.vwork_elems);
}

// A level with both narrow and wide supernodes runs the two kinds side by
// side: the wide launch on the handle's stream, enqueued first (its block
// chains are the sweep's critical path), and the narrow launch on the
// handle's second stream, forked from the handle's stream before the wide
// launch and joined to it after, by events (the two touch different
// supernodes, and the next level waits for both).
template <typename NarrowF, typename WideF>
static int run_level(sprgpu_handle *h, bool narrow, bool wide, NarrowF &&launch_narrow,
                     WideF &&launch_wide)
{
    if (!(narrow && wide && h->stream2)) {
        if (narrow) launch_narrow(h->stream);
        if (wide) launch_wide();
        return 0;
    }
    SPRGPU_HIP(hipEventRecord(h->ev[0], h->stream));
    launch_wide();
    SPRGPU_HIP(hipStreamWaitEvent(h->stream2, h->ev[0], 0));
    launch_narrow(h->stream2);
    SPRGPU_HIP(hipEventRecord(h->ev[1], h->stream2));
    SPRGPU_HIP(hipStreamWaitEvent(h->stream, h->ev[1], 0));
    return 0;
}

template <typename T>
static int solve_chunk(sprgpu_handle *h, double *b, i64 ldb, i64 nrhs, int which)
{
    const Plan &p = h->plan;
    const bool grouped = nrhs >= kSolveGroup;
    if (which & 1) {
        mark_pending(h, nrhs);
        for (int l = 0; l < p.nlevels; ++l) {
            const bool narrow = p.solve_small_ptr[l + 1] > p.solve_small_ptr[l];
            const bool wide = p.wide_ptr[l + 1] > p.wide_ptr[l];
            auto nar = [&](hipStream_t st) {
                const int R = narrow_group<T>(p, l, nrhs);
                if (R == kSolveGroup) fwd_narrow_level<T, kSolveGroup>(h, st, l, b, ldb, nrhs);
                else if (R == 4) fwd_narrow_level<T, 4>(h, st, l, b, ldb, nrhs);
                else fwd_narrow_level<T, 1>(h, st, l, b, ldb, nrhs);
            };
            auto wid = [&]() {
                if (grouped) {  // the chains, then the rectangles in wider groups
                    fwd_wide_level<T, wide_group<T>()>(h, l, b, ldb, nrhs, false);
                    if (nrhs >= kMmGroup) fwd_rect_mm_level<T>(h, l, nrhs);
                    else if (nrhs >= kRectGroup) fwd_rect_level<T, kRectGroup>(h, l, nrhs);
                    else fwd_rect_level<T, kSolveGroup>(h, l, nrhs);
                } else {
                    fwd_wide_level<T, 1>(h, l, b, ldb, nrhs, true);
                }
            };
            if (int rc = run_level(h, narrow, wide, nar, wid)) return rc;
        }
    }
    if (which & 2) {
        mark_pending(h, nrhs);
        for (int l = p.nlevels - 1; l >= 0; --l) {
            const bool narrow = p.solve_small_ptr[l + 1] > p.solve_small_ptr[l];
            const bool wide = p.wide_ptr[l + 1] > p.wide_ptr[l];
            auto nar = [&](hipStream_t st) {
                const int R = narrow_group<T>(p, l, nrhs);
                if (R == kSolveGroup) bwd_narrow_level<T, kSolveGroup>(h, st, l, b, ldb, nrhs);
                else if (R == 4) bwd_narrow_level<T, 4>(h, st, l, b, ldb, nrhs);
                else bwd_narrow_level<T, 1>(h, st, l, b, ldb, nrhs);
            };
            auto wid = [&]() {
                if (nrhs >= kMmGroup) bwd_rect_mm_level<T>(h, l, b, ldb, nrhs);
                else if (nrhs >= kRectGroup) bwd_rect_level<T, kRectGroup>(h, l, b, ldb, nrhs);
                else if (grouped) bwd_rect_level<T, kSolveGroup>(h, l, b, ldb, nrhs);
                else bwd_rect_level<T, 1>(h, l, b, ldb, nrhs);
                if (grouped) bwd_wide_level<T, wide_group<T>()>(h, l, b, ldb, nrhs);
                else bwd_wide_level<T, 1>(h, l, b, ldb, nrhs);
            };
            if (int rc = run_level(h, narrow, wide, nar, wid)) return rc;
        }
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

// A launch holds at most 2^23 workgroups: large batches go in chunks of
// right-hand sides sized with each level's group size.
template <typename T>
static int solve_impl(sprgpu_handle *h, double *b, i64 ldb, i64 nrhs, int which)
{
    const Plan &p = h->plan;
    if (int rc = ensure_solve_buffers(h, nrhs)) return rc;
    if (p.forest && nrhs == 1) return forest_solve(h, b, which);
    constexpr i64 kMaxWgs = i64(1) << 23;
    i64 chunk = nrhs;
    for (int l = 0; l < p.nlevels; ++l) {
        const i64 ns = p.solve_small_ptr[l + 1] - p.solve_small_ptr[l];
        const i64 wmax = std::max({p.wide_fwd_wgs[l], p.wide_bwd_wgs[l], p.wide_rect_wgs[l]});
        const i64 rn = ns > 0 ? narrow_group<T>(p, l, nrhs) : 1;  // group of the narrow launches
        const i64 rw = (nrhs >= kSolveGroup) ? kSolveGroup : 1;
        if (ns > 0) chunk = std::min(chunk, std::max<i64>(1, kMaxWgs / ns) * rn);
        if (wmax > 0) chunk = std::min(chunk, std::max<i64>(1, kMaxWgs / wmax) * rw);
    }
    for (i64 r0 = 0; r0 < nrhs; r0 += chunk) {
        if (int rc = solve_chunk<T>(h, b + r0 * ldb, ldb, std::min(chunk, nrhs - r0), which))
            return rc;
    }
    return 0;
}

int prepare_solve(sprgpu_handle *h, i64 nrhs) { return ensure_solve_buffers(h, nrhs); }

int compute_block_inverses(sprgpu_handle *h)
{
    const Plan &p = h->plan;
    if (p.nflags == 0) return 0;
    if (h->precision == SPRGPU_FP32)
        k_block_inverse<float><<<unsigned(p.nflags), 256, 0, h->stream>>>(
            h->d.wide, int(p.wide.size()), h->d.nodes, static_cast<const float *>(h->d.store),
            static_cast<float *>(h->d.dinv));
    else
        k_block_inverse<double><<<unsigned(p.nflags), 256, 0, h->stream>>>(
            h->d.wide, int(p.wide.size()), h->d.nodes, static_cast<const double *>(h->d.store),
            static_cast<double *>(h->d.dinv));
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int run_solve(sprgpu_handle *h, double *b, i64 ldb, i64 nrhs, int which)
{
    if (!h->factored) {
        set_error("solve called before a successful factorization");
        return -1;
    }
    if (nrhs <= 0) return 0;
    if (nrhs > 65535) {
        set_error("at most 65535 right-hand sides per call");
        return -1;
    }
    // The group of right-hand sides is the fastest-varying workgroup index
    // (the workgroups of one supernode for all groups run together and share
    // its factor through the caches).
    return (h->precision == SPRGPU_FP32) ? solve_impl<float>(h, b, ldb, nrhs, which)
                                         : solve_impl<double>(h, b, ldb, nrhs, which);
}

int check_solve_error(sprgpu_handle *h)
{
    if (!h->d.spin_error) return 0;
    int e = 0;
    SPRGPU_HIP(hipMemcpyAsync(&e, h->d.spin_error, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    SPRGPU_HIP(hipStreamSynchronize(h->stream));
    if (e) {
        set_error("a solve dataflow wait timed out");
        return -5;
    }
    return 0;
}

}  // namespace sprgpu
