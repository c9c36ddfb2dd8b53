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
//   on before waiting for that block's ready flag, applies it as a GEMV with
//   two threads per row, and finally solves its own 128 x 128 unit-triangular
//   block as four 32 x 32 sub-blocks: a readlane chain over registers on one
//   wavefront per sub-block, the updates between sub-blocks in parallel.
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
static_assert(NB == 128, "two threads per row of a block; one 32-row sub-block per wavefront");

// Block results cross workgroups (and XCDs, whose L2 caches are not kept
// coherent with each other) as agent-scope atomic stores and loads, which
// are coherent on their own; the publishing wavefront waits for its stores
// to complete (s_waitcnt) before it stores the ready flag, and readers load
// the results only after seeing the flag.  This avoids the agent-scope
// fences, which write back or invalidate a whole L2.

// One thread waits until *flag reaches epoch (bounded), then the workgroup
// synchronises.  A wait that gives up sets *err and continues.
__device__ __forceinline__ void wait_ready(const unsigned *flag, unsigned epoch, int *err)
{
    if (threadIdx.x == 0) {
        unsigned long long spins = 0;
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpinLimit) {
                atomicExch(err, 1);
                break;
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ void store_shared(double *p, double x)
{
    __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_shared(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by every lane of the wavefront that stored the block (store_shared).
__device__ __forceinline__ void publish_ready(unsigned *flag, unsigned epoch)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);  // vmcnt = expcnt = lgkmcnt = 0: the stores have completed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if ((threadIdx.x & 63) == 0)
        __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The epoch of a sweep lives in device memory and is advanced by a kernel
// before the sweep, so that a captured sequence (hipGraph) replays with a new
// epoch each time; flags start at zero and only ever take epoch values.
__global__ void k_epoch_bump(unsigned *epoch)
{
    if (threadIdx.x == 0) *epoch += 1;
}

// Offset of column c in a packed lower-triangular NB x NB block stored by
// columns (column c holds rows c..NB-1).
__device__ __forceinline__ int packed_col(int c) { return c * NB - c * (c - 1) / 2; }
// The packed blocks in LDS carry one extra zero entry: a read of an entry
// outside the block selects this address instead of masking the loaded value
// (a select on a loaded value makes the compiler wait for each load).
constexpr int kPacked = NB * (NB + 1) / 2;
constexpr int kZero = kPacked;

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
                                                    const T *__restrict__ pool,
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
    const T *L = pool + n.foff;
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
                                                    const T *__restrict__ pool, double *b,
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
    const T *L = pool + n.foff;
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
// wide supernodes, forward
// ---------------------------------------------------------------------------
// Workgroup local < nblk solves diagonal block local after applying blocks
// 0..local-1; workgroups nblk.. compute the update vector of kSolveNB rows
// below the diagonal part, applying every block as it finishes.  As in the
// narrow kernels, a workgroup serves R right-hand sides; its ready flags are
// those of its group of right-hand sides.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_fwd_wide(const WideDesc *__restrict__ wide, int nwide,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i32 *__restrict__ lindx,
                                                  const i64 *__restrict__ upd_ptr,
                                                  const i32 *__restrict__ upd_src,
                                                  const T *__restrict__ pool,
                                                  const double *__restrict__ dvec, double *b,
                                                  i64 ldb, double *vwork, i64 vstride,
                                                  unsigned *flags, i64 nflags,
                                                  const unsigned *epoch_ptr, int *err, int nrhs,
                                                  const T *__restrict__ bmscr)
{
    __shared__ double xs[R][NB];
    __shared__ double ys[R][NB];
    __shared__ double part[2][R][NB];
    __shared__ T Ld[kPacked + 1];
    const unsigned epoch = *epoch_ptr;  // packed lower diagonal block, by columns; Ld[kZero] = 0
    const int nrb = (nrhs + R - 1) / R;
    const i64 bid = blockIdx.x / nrb;
    const int rg = int(blockIdx.x % nrb), rhs0 = rg * R;
    const int q = find_desc(prefix, nwide, bid);
    const WideDesc w = wide[q];
    const int local = int(bid - prefix[q]);
    const NodeInfo n = nodes[w.node];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int S = n.S, F = n.F;
    const i32 *rows = lindx + n.lxoff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    const i64 *up = upd_ptr + n.voff;
    unsigned *fl = flags + rg * nflags + w.flag0;
    i64 vo[R];  // this group's vectors of the supernode (the last valid one past nrhs)
#pragma unroll
    for (int k = 0; k < R; ++k) vo[k] = i64(min(rhs0 + k, nrhs - 1)) * vstride + n.voff;
    const bool diag = local < w.nblk;
    const int r0 = diag ? local * NB : S + (local - w.nblk) * NB;
    const int nr = min(NB, (diag ? S : F) - r0);
    const int ri = tid & (NB - 1), h = tid >> 7;
    const int row = r0 + ri;
    const bool rv = ri < nr;
    // right-hand side of my rows: B (own columns) plus the children's updates
    if (h == 0) {
        const i64 k0 = rv ? up[row] : 0, k1 = rv ? up[row + 1] : 0;
        const i32 gi = (rv && row < S) ? rows[row] : 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const i64 rh = min(rhs0 + k, nrhs - 1);
            double y = 0.0;
            if (rv) {
                y = (row < S) ? b[rh * ldb + gi] : 0.0;
                const double *vw = vwork + rh * vstride;
                for (i64 j = k0; j < k1; ++j) y += vw[upd_src[j]];
            }
            ys[k][ri] = y;
        }
    }
    // destinations and pivots of the two block rows wave 0 writes at the end
    i32 dst[2] = {0, 0};
    double dpiv[2] = {1.0, 1.0};
    if (diag && wave == 0) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = min(lane + 64 * k, nr - 1);
            dst[k] = rows[r0 + i];
            dpiv[k] = dvec[n.first + r0 + i];
        }
    }
    if (diag) {
        for (int idx = tid; idx < nr * nr; idx += 256) {
            const int rr = idx % nr, cc = idx / nr;
            if (rr > cc) Ld[packed_col(cc) + (rr - cc)] = L[(r0 + rr) + i64(r0 + cc) * ld];
        }
        if (tid == 0) Ld[kZero] = T(0);
    }
    // Diagonal solve operands: with one right-hand side (the latency-bound
    // case) loaded ahead of the waits, so that their latency stays off the
    // chain of blocks; with a group, after the loop (registers).  Lanes
    // 32..63 mirror lanes 0..31 (lane & 31) and do no work there.
    const int sr = 32 * wave + (lane & 31);  // block row of this lane (lanes < 32 work)
    const bool sw = diag && lane < 32 && sr < nr;
    const int nbw = min(32, nr - 32 * wave);  // rows of this wave's sub-block
    T lsub[32];
    double dsr = 0.0;
    auto load_bm = [&]() {
        // column lane of the sub-block's Bm (rows past the sub-block zero)
        const T *bm = bmscr + i64(n.first + r0 + 32 * wave) * kPanelNB + (lane & 31);
#pragma unroll
        for (int c = 0; c < 32; ++c) lsub[c] = (sw && c < nbw) ? bm[c * kPanelNB] : T(0);
        dsr = sw ? dvec[n.first + r0 + sr] : 0.0;
    };
    if (R == 1 && bmscr) load_bm();
    const int nprev = diag ? local : w.nblk;
    double acc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) acc[k] = 0.0;
    for (int p = 0; p < nprev; ++p) {
        const int c0 = p * NB;
        const int nc = min(NB, S - c0);
        // loads are unconditional and unmasked so that they all stay in
        // flight: rows past the block read row r0 (their results are not
        // stored) and columns past the block read its last column (xs is
        // zero there)
        T lv[NH];
        const T *Lr = L + (rv ? row : r0) + i64(c0) * ld;
#pragma unroll
        for (int c = 0; c < NH; ++c) lv[c] = Lr[i64(min(h * NH + c, nc - 1)) * ld];
        wait_ready(fl + p, epoch, err);
        if (tid < NB) {
#pragma unroll
            for (int k = 0; k < R; ++k) xs[k][tid] = (tid < nc) ? load_shared(vwork + vo[k] + c0 + tid) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NH; ++c)
#pragma unroll
            for (int k = 0; k < R; ++k) acc[k] += double(lv[c]) * xs[k][h * NH + c];
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < R; ++k) part[h][k][ri] = acc[k];
    __syncthreads();
    if (!diag) {
        if (h == 0 && rv) {  // update for the parent
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (rhs0 + k < nrhs) vwork[vo[k] + row] = ys[k][ri] - (part[0][k][ri] + part[1][k][ri]);
        }
        return;
    }
    // Diagonal block: four 32 x 32 sub-blocks in sequence.  Wave w solves
    // its sub-block, lane i < 32 holding row 32w + i: with the factor's
    // inverses (bmscr: Bm = L11^-T D^-1 of every 32-column sub-block, so
    // L11^-1 y = D Bm^T y) as a product x_i = d_i sum_c Bm(c, i) y_c, else by
    // a readlane chain over the row of L11 in registers.  The waves below
    // then add its products with their rows of the packed block in LDS to
    // their pending sums, registers with one right-hand side (pend), LDS for
    // a group (registers there would become hoisted loads and spills), so
    // that one barrier per sub-block orders the steps.
    if (bmscr) {
        if (R != 1) load_bm();
    } else {
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            const int cc = 32 * wave + c;
            lsub[c] = Ld[(sw && c < (lane & 31)) ? packed_col(cc) + sr - cc : kZero];
        }
    }
    if (tid < NB) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            ys[k][tid] = (tid < nr) ? ys[k][tid] - (part[0][k][tid] + part[1][k][tid]) : 0.0;
    }
    __syncthreads();
    if (tid < NB) {  // part[0] holds a group's pending sums from here
#pragma unroll
        for (int k = 0; k < R; ++k) part[0][k][tid] = 0.0;
    }
    __syncthreads();
    double pend[R];
#pragma unroll
    for (int k = 0; k < R; ++k) pend[k] = 0.0;
    for (int j = 0; j < 4 && 32 * j < nr; ++j) {
        if (wave == j) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double x;
                if (bmscr) {
                    // the wave's own LDS stores precede its loads (in order per wave)
                    if (lane < 32) ys[k][sr] -= (R == 1 ? pend[k] : part[0][k][sr]);
                    __builtin_amdgcn_wave_barrier();
                    double a0 = 0.0, a1 = 0.0;
#pragma unroll
                    for (int c = 0; c < 32; c += 2) {
                        a0 += double(lsub[c]) * ys[k][32 * j + c];
                        a1 += double(lsub[c + 1]) * ys[k][32 * j + c + 1];
                    }
                    x = dsr * (a0 + a1);
                } else {
                    x = (sw ? ys[k][sr] : 0.0) - (R == 1 ? pend[k] : part[0][k][sr]);
#pragma unroll
                    for (int c = 0; c < 32; ++c) x -= double(lsub[c]) * lane_bcast(x, c);
                }
                __builtin_amdgcn_wave_barrier();
                if (lane < 32) ys[k][sr] = x;
            }
        }
        __syncthreads();
        if (wave > j && sw) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double s0 = 0.0, s1 = 0.0;
#pragma unroll
                for (int c = 0; c < 32; c += 2) {
                    const int cc = 32 * j + c;
                    const T a0 = Ld[cc < nr ? packed_col(cc) + sr - cc : kZero];
                    const T a1 = Ld[cc + 1 < nr ? packed_col(cc + 1) + sr - cc - 1 : kZero];
                    s0 += double(a0) * ys[k][cc];
                    s1 += double(a1) * ys[k][cc + 1];
                }
                if constexpr (R == 1)
                    pend[k] += s0 + s1;
                else
                    part[0][k][sr] += s0 + s1;
            }
        }
    }
    if (wave == 0) {
        // results for the blocks that wait on this one first, then B
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (lane + 64 * i < nr) {
#pragma unroll
                for (int k = 0; k < R; ++k)
                    if (rhs0 + k < nrhs) store_shared(vwork + vo[k] + r0 + lane + 64 * i, ys[k][lane + 64 * i]);
            }
        publish_ready(fl + local, epoch);
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (lane + 64 * i < nr) {
#pragma unroll
                for (int k = 0; k < R; ++k)
                    if (rhs0 + k < nrhs) b[i64(rhs0 + k) * ldb + dst[i]] = ys[k][lane + 64 * i] / dpiv[i];
            }
    }
}

// ---------------------------------------------------------------------------
// wide supernodes, backward
// ---------------------------------------------------------------------------
// Partial sums of L21^T x2: workgroup (blk, k) multiplies rows [S+k*NB, +NB)
// of the columns of block blk by the solution at those rows (final: they
// belong to ancestors).  Two threads per column, each walking half the rows.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_bwd_rect(const WideDesc *__restrict__ wide, int nwide,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i32 *__restrict__ lindx,
                                                  const T *__restrict__ pool,
                                                  const double *__restrict__ b, i64 ldb,
                                                  double *partials, i64 npart, int nrhs)
{
    __shared__ double xs[R][NB];
    __shared__ double part[2][R][NB];
    const int nrb = (nrhs + R - 1) / R;
    const i64 bid = blockIdx.x / nrb;
    const int rhs0 = int(blockIdx.x % nrb) * R;
    const int q = find_desc(prefix, nwide, bid);
    const WideDesc w = wide[q];
    const int t = int(bid - prefix[q]);
    const int blk = t / w.nrect, k = t % w.nrect;
    const NodeInfo n = nodes[w.node];
    const int tid = threadIdx.x;
    const int S = n.S, F = n.F;
    const int c0 = blk * NB, nc = min(NB, S - c0);
    const int r0 = S + k * NB, nr = min(NB, F - r0);
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    if (tid < NB) {
        const i32 gi = lindx[n.lxoff + r0 + min(tid, nr - 1)];
#pragma unroll
        for (int j = 0; j < R; ++j)
            xs[j][tid] = (tid < nr) ? b[i64(min(rhs0 + j, nrhs - 1)) * ldb + gi] : 0.0;
    }
    __syncthreads();
    const int ci = tid & (NB - 1), h = tid >> 7;
    // unconditional loads (columns past the block read its last column and
    // are not stored; rows past the chunk read its last row, xs is zero)
    T lv[NH];
    const T *Lc = L + r0 + i64(c0 + min(ci, nc - 1)) * ld;
#pragma unroll
    for (int r = 0; r < NH; ++r) lv[r] = Lc[min(h * NH + r, nr - 1)];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
        for (int r = 0; r < NH; r += 4) {
            a0 += double(lv[r]) * xs[j][h * NH + r];
            a1 += double(lv[r + 1]) * xs[j][h * NH + r + 1];
            a2 += double(lv[r + 2]) * xs[j][h * NH + r + 2];
            a3 += double(lv[r + 3]) * xs[j][h * NH + r + 3];
        }
        part[h][j][ci] = (a0 + a1) + (a2 + a3);
    }
    __syncthreads();
    if (tid < nc) {
#pragma unroll
        for (int j = 0; j < R; ++j)
            if (rhs0 + j < nrhs)
                partials[i64(rhs0 + j) * npart + w.pbase + (i64(blk) * w.nrect + k) * NB + tid] =
                    part[0][j][tid] + part[1][j][tid];
    }
}

// Workgroup local handles block nblk-1-local: y = B - partial sums of the
// rows below, minus the later blocks' products as they finish, then
// L11^T x = y on the block by wave 0 (two columns per lane).
template <typename T, int R>
__global__ __launch_bounds__(256) void k_bwd_wide(const WideDesc *__restrict__ wide, int nwide,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i32 *__restrict__ lindx,
                                                  const T *__restrict__ pool, double *b, i64 ldb,
                                                  double *vwork, i64 vstride,
                                                  const double *__restrict__ partials, i64 npart,
                                                  unsigned *flags, i64 nflags,
                                                  const unsigned *epoch_ptr, int *err, int nrhs,
                                                  const T *__restrict__ bmscr,
                                                  const double *__restrict__ dvec)
{
    __shared__ double xs[R][NB];
    __shared__ double ys[R][NB];
    __shared__ double part[2][R][NB];
    __shared__ T Lr[kPacked + 1];
    const unsigned epoch = *epoch_ptr;  // packed lower diagonal block, by rows; Lr[kZero] = 0
    const int nrb = (nrhs + R - 1) / R;
    const i64 bid = blockIdx.x / nrb;
    const int rg = int(blockIdx.x % nrb), rhs0 = rg * R;
    const int q = find_desc(prefix, nwide, bid);
    const WideDesc w = wide[q];
    const int blk = w.nblk - 1 - int(bid - prefix[q]);
    const NodeInfo n = nodes[w.node];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int S = n.S;
    const i32 *rows = lindx + n.lxoff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    unsigned *fl = flags + rg * nflags + w.flag0;
    i64 vo[R];
#pragma unroll
    for (int k = 0; k < R; ++k) vo[k] = i64(min(rhs0 + k, nrhs - 1)) * vstride + n.voff;
    const int c0 = blk * NB, nc = min(NB, S - c0);
    for (int idx = tid; idx < nc * nc; idx += 256) {
        const int rr = idx % nc, cc = idx / nc;
        if (rr > cc) Lr[rr * (rr + 1) / 2 + cc] = L[(c0 + rr) + i64(c0 + cc) * ld];
    }
    if (tid == 0) Lr[kZero] = T(0);
    i32 dst[2] = {0, 0};
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < 2; ++k) dst[k] = rows[c0 + min(lane + 64 * k, nc - 1)];
    }
    const int ci = tid & (NB - 1), h = tid >> 7;
    // products with the rows below the diagonal part: the two halves of the
    // workgroup sum the even and the odd row chunks' partials
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double *ps = partials + i64(min(rhs0 + k, nrhs - 1)) * npart + w.pbase +
                           i64(blk) * w.nrect * NB + min(ci, nc - 1);
        part[h][k][ci] = strided_sum(ps, NB, h, 2, w.nrect);
    }
    __syncthreads();
    if (tid < NB) {
        const i32 gi = rows[c0 + min(tid, nc - 1)];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            ys[k][tid] = (tid < nc)
                             ? b[i64(min(rhs0 + k, nrhs - 1)) * ldb + gi] - (part[0][k][tid] + part[1][k][tid])
                             : 0.0;
            part[0][k][tid] = 0.0;  // a group's pending sums in the diagonal solve
        }
    }
    // Diagonal solve operands: with one right-hand side (the latency-bound
    // case) loaded ahead of the waits, so that their latency stays off the
    // chain of blocks; with a group, after the loop (registers).  Lanes
    // 32..63 mirror lanes 0..31 (lane & 31) and do no work there.
    const int sc = 32 * wave + (lane & 31);  // block column of this lane (lanes < 32 work)
    const bool sw = lane < 32 && sc < nc;
    const int nbw = min(32, nc - 32 * wave);  // columns of this wave's sub-block
    T lcol[32];
    double dsc = 0.0;  // d of this lane's column (with the inverses)
    auto load_bm = [&]() {
        // row lane of the sub-block's Bm (columns past the sub-block zero)
        const T *bm = bmscr + i64(n.first + c0 + sc) * kPanelNB;
#pragma unroll
        for (int r = 0; r < 32; ++r) lcol[r] = (sw && r < nbw) ? bm[r] : T(0);
        dsc = sw ? dvec[n.first + c0 + sc] : 0.0;
    };
    if (R == 1 && bmscr) load_bm();
    __syncthreads();
    // Products with the later blocks, L(p, blk)^T x_p: the thread sums
    // columns gc + 8j (j < 4) over rows 8q + gs (q < 16) of every block, so
    // that a load instruction reads 8 rows of 8 columns (8 cache lines where
    // a lane per column would touch 64 and be bound by address processing);
    // the eight lanes of a column are added at the end (xor 1, 2, 4).
    const int gs = lane & 7, gc = 32 * wave + (lane >> 3);
    double acc[4][R];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < R; ++k) acc[j][k] = 0.0;
    for (int p = w.nblk - 1; p > blk; --p) {
        const int rp0 = p * NB, np = min(NB, S - rp0);
        // unconditional, unmasked loads (see k_fwd_wide): columns past the
        // block read its last column (not stored), rows past block p read
        // its last row (xs is zero there)
        T lv[4][NB / 8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const T *Lc = L + rp0 + i64(c0 + min(gc + 8 * j, nc - 1)) * ld;
#pragma unroll
            for (int q = 0; q < NB / 8; ++q) lv[j][q] = Lc[min(8 * q + gs, np - 1)];
        }
        wait_ready(fl + p, epoch, err);
        if (tid < NB) {
#pragma unroll
            for (int k = 0; k < R; ++k)
                xs[k][tid] = (tid < np) ? load_shared(vwork + vo[k] + rp0 + tid) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < NB / 8; ++q)
#pragma unroll
                for (int k = 0; k < R; ++k) acc[j][k] += double(lv[j][q]) * xs[k][8 * q + gs];
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            double v = acc[j][k];
            v += __shfl_xor(v, 1);
            v += __shfl_xor(v, 2);
            v += __shfl_xor(v, 4);
            if (gs == 0 && gc + 8 * j < nc) ys[k][gc + 8 * j] -= v;
        }
    }
    __syncthreads();
    // Diagonal block, L11^T x = y: four 32 x 32 sub-blocks from the last.
    // Wave w solves its sub-block, lane i < 32 holding column 32w + i: with
    // the factor's inverses (L11^-T y = Bm D y) as a product
    // x_i = sum_r Bm(i, r) d_r y_r, else by a readlane chain over the column
    // of L11 in registers.  The waves to the left then add its products with
    // their columns of the packed block in LDS to their pending sums,
    // registers with one right-hand side (pend), LDS for a group (registers
    // there would become hoisted loads and spills), so that one barrier per
    // sub-block orders the steps.
    if (bmscr) {
        if (R != 1) load_bm();
    } else {
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const int rr = 32 * wave + r;
            lcol[r] = Lr[(sw && r > (lane & 31) && rr < nc) ? rr * (rr + 1) / 2 + sc : kZero];
        }
    }
    double pend[R];
#pragma unroll
    for (int k = 0; k < R; ++k) pend[k] = 0.0;
    for (int j = 3; j >= 0; --j) {
        if (32 * j >= nc) continue;
        if (wave == j) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double x;
                if (bmscr) {
                    // d_r y_r of the sub-block; the wave's own LDS stores
                    // precede its loads (in order per wave)
                    if (lane < 32) ys[k][sc] = dsc * (ys[k][sc] - (R == 1 ? pend[k] : part[0][k][sc]));
                    __builtin_amdgcn_wave_barrier();
                    double a0 = 0.0, a1 = 0.0;
#pragma unroll
                    for (int r = 0; r < 32; r += 2) {
                        a0 += double(lcol[r]) * ys[k][32 * j + r];
                        a1 += double(lcol[r + 1]) * ys[k][32 * j + r + 1];
                    }
                    x = a0 + a1;
                } else {
                    x = (sw ? ys[k][sc] : 0.0) - (R == 1 ? pend[k] : part[0][k][sc]);
#pragma unroll
                    for (int r = 31; r > 0; --r) x -= double(lcol[r]) * lane_bcast(x, r);
                }
                __builtin_amdgcn_wave_barrier();
                if (lane < 32) ys[k][sc] = x;
            }
        }
        __syncthreads();
        if (wave < j && sw) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double s0 = 0.0, s1 = 0.0;
#pragma unroll
                for (int r = 0; r < 32; r += 2) {
                    const int rr = 32 * j + r;
                    const T a0 = Lr[rr < nc ? rr * (rr + 1) / 2 + sc : kZero];
                    const T a1 = Lr[rr + 1 < nc ? (rr + 1) * (rr + 2) / 2 + sc : kZero];
                    s0 += double(a0) * ys[k][rr];
                    s1 += double(a1) * ys[k][rr + 1];
                }
                if constexpr (R == 1)
                    pend[k] += s0 + s1;
                else
                    part[0][k][sc] += s0 + s1;
            }
        }
    }
    if (wave == 0) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (lane + 64 * i < nc) {
#pragma unroll
                for (int k = 0; k < R; ++k)
                    if (rhs0 + k < nrhs) store_shared(vwork + vo[k] + c0 + lane + 64 * i, ys[k][lane + 64 * i]);
            }
        publish_ready(fl + blk, epoch);
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (lane + 64 * i < nc) {
#pragma unroll
                for (int k = 0; k < R; ++k)
                    if (rhs0 + k < nrhs) b[i64(rhs0 + k) * ldb + dst[i]] = ys[k][lane + 64 * i];
            }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int ensure_solve_buffers(sprgpu_handle *h, i64 nrhs)
{
    DeviceBuffers &d = h->d;
    const Plan &p = h->plan;
    if (d.vwork_cap < nrhs) {
        if (d.vwork) SPRGPU_HIP(hipFree(d.vwork));
        d.vwork = nullptr;
        SPRGPU_HIP(hipMalloc(&d.vwork, sizeof(double) * size_t(p.vwork_elems) * size_t(nrhs)));
        h->device_bytes += sizeof(double) * p.vwork_elems * (nrhs - d.vwork_cap);
        d.vwork_cap = nrhs;
    }
    const i64 nf = std::max<i64>(p.nflags, 1) * nrhs;
    if (d.flags_cap < nf) {
        if (d.flags) SPRGPU_HIP(hipFree(d.flags));
        d.flags = nullptr;
        SPRGPU_HIP(hipMalloc(&d.flags, sizeof(unsigned) * nf));
        SPRGPU_HIP(hipMemset(d.flags, 0, sizeof(unsigned) * nf));
        d.flags_cap = nf;
        if (!d.epoch) SPRGPU_HIP(hipMalloc(&d.epoch, sizeof(unsigned)));
        SPRGPU_HIP(hipMemset(d.epoch, 0, sizeof(unsigned)));  // every flag is zero again
    }
    const i64 np = std::max<i64>(p.npart, 1) * nrhs;
    if (d.partials_cap < np) {
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
// right-hand sides.  (Larger groups on the smallest fronts were measured
// slower: one wavefront's chains and gathers for them run in sequence.)
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

// The factor's 32 x 32 diagonal-block inverses, when the outer-block scheme
// computed them (in the panel-step scheme the scratch holds the factored
// blocks instead and the solves use readlane chains).
template <typename T>
static const T *solve_inverses(sprgpu_handle *h)
{
    return h->plan.params.block_factor ? static_cast<const T *>(h->d.dscr) : nullptr;
}

template <typename T, int R>
static void fwd_narrow_level(sprgpu_handle *h, int l, double *b, i64 ldb, i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    const T *pool = static_cast<const T *>(d.pool);
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const unsigned ns = unsigned(p.solve_small_ptr[l + 1] - p.solve_small_ptr[l]);
    const int mf = p.small_maxF[l], ms = p.small_maxS[l];
    const unsigned nt = narrow_threads(p, l);
    k_fwd_narrow<T, R><<<ns * groups, nt, fwd_narrow_lds<T>(R, mf, ms, nt), h->stream>>>(
        d.solve_small + p.solve_small_ptr[l], mf, d.nodes, d.lindx, d.upd_ptr, d.upd_src, pool,
        d.dvec, b, ldb, d.vwork, p.vwork_elems, int(nrhs));
}

template <typename T, int R>
static void fwd_wide_level(sprgpu_handle *h, int l, double *b, i64 ldb, i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    const T *pool = static_cast<const T *>(d.pool);
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
    k_fwd_wide<T, R><<<unsigned(p.wide_fwd_wgs[l]) * groups, 256, 0, h->stream>>>(
        d.wide + w0, nw, d.wide_fwd_prefix + w0, d.nodes, d.lindx, d.upd_ptr, d.upd_src, pool,
        d.dvec, b, ldb, d.vwork, p.vwork_elems, d.flags, std::max<i64>(p.nflags, 1), d.epoch,
        d.spin_error, int(nrhs), solve_inverses<T>(h));
}

template <typename T, int R>
static void bwd_narrow_level(sprgpu_handle *h, int l, double *b, i64 ldb, i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    const T *pool = static_cast<const T *>(d.pool);
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const unsigned ns = unsigned(p.solve_small_ptr[l + 1] - p.solve_small_ptr[l]);
    const int mf = p.small_maxF[l], ms = p.small_maxS[l];
    const unsigned nt = narrow_threads(p, l);
    k_bwd_narrow<T, R><<<ns * groups, nt, bwd_narrow_lds<T>(R, mf, ms, nt), h->stream>>>(
        d.solve_small + p.solve_small_ptr[l], mf, d.nodes, d.lindx, pool, b, ldb, int(nrhs));
}

template <typename T, int R>
static void bwd_wide_level(sprgpu_handle *h, int l, double *b, i64 ldb, i64 nrhs)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    const T *pool = static_cast<const T *>(d.pool);
    const unsigned groups = unsigned((nrhs + R - 1) / R);
    const i64 npart = std::max<i64>(p.npart, 1);
    const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
    if (p.wide_rect_wgs[l] > 0)
        k_bwd_rect<T, R><<<unsigned(p.wide_rect_wgs[l]) * groups, 256, 0, h->stream>>>(
            d.wide + w0, nw, d.wide_rect_prefix + w0, d.nodes, d.lindx, pool, b, ldb, d.partials,
            npart, int(nrhs));
    k_bwd_wide<T, R><<<unsigned(p.wide_bwd_wgs[l]) * groups, 256, 0, h->stream>>>(
        d.wide + w0, nw, d.wide_bwd_prefix + w0, d.nodes, d.lindx, pool, b, ldb, d.vwork,
        p.vwork_elems, d.partials, npart, d.flags, std::max<i64>(p.nflags, 1), d.epoch,
        d.spin_error, int(nrhs), solve_inverses<T>(h), d.dvec);
}

template <typename T>
static int solve_chunk(sprgpu_handle *h, double *b, i64 ldb, i64 nrhs, int which)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    hipStream_t s = h->stream;
    const bool grouped = nrhs >= kSolveGroup;
    if (which & 1) {
        k_epoch_bump<<<1, 64, 0, s>>>(d.epoch);
        for (int l = 0; l < p.nlevels; ++l) {
            if (p.solve_small_ptr[l + 1] > p.solve_small_ptr[l]) {
                const int R = narrow_group<T>(p, l, nrhs);
                if (R == kSolveGroup) fwd_narrow_level<T, kSolveGroup>(h, l, b, ldb, nrhs);
                else if (R == 4) fwd_narrow_level<T, 4>(h, l, b, ldb, nrhs);
                else fwd_narrow_level<T, 1>(h, l, b, ldb, nrhs);
            }
            if (p.wide_ptr[l + 1] > p.wide_ptr[l]) {
                if (grouped) fwd_wide_level<T, kSolveGroup>(h, l, b, ldb, nrhs);
                else fwd_wide_level<T, 1>(h, l, b, ldb, nrhs);
            }
        }
    }
    if (which & 2) {
        k_epoch_bump<<<1, 64, 0, s>>>(d.epoch);
        for (int l = p.nlevels - 1; l >= 0; --l) {
            if (p.wide_ptr[l + 1] > p.wide_ptr[l]) {
                if (grouped) bwd_wide_level<T, kSolveGroup>(h, l, b, ldb, nrhs);
                else bwd_wide_level<T, 1>(h, l, b, ldb, nrhs);
            }
            if (p.solve_small_ptr[l + 1] > p.solve_small_ptr[l]) {
                const int R = narrow_group<T>(p, l, nrhs);
                if (R == kSolveGroup) bwd_narrow_level<T, kSolveGroup>(h, l, b, ldb, nrhs);
                else if (R == 4) bwd_narrow_level<T, 4>(h, l, b, ldb, nrhs);
                else bwd_narrow_level<T, 1>(h, l, b, ldb, nrhs);
            }
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
