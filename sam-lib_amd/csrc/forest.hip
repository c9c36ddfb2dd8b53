// Factorization and solve of forests of small fronts in one launch.
//
// When every supernode has F <= kSmallF (batches of small systems merged
// into one forest: frames, small meshes), the level schedule spends most of
// its time in launches of few fronts each (one per tree height and kind,
// 120 heights for the 512-frame batch) and waits at every height for the
// slowest front of all systems.  Here one launch takes the whole forest as
// work items, each waiting for exactly the items it needs:
//   the ascending pass, one item per supernode: factor (SPRF11/SPRF13/
//     SPRF15/SPRF17, sprfc1.f:286-1299, for a front held in LDS and one row
//     per lane) and/or the forward and diagonal solve step (SPRFS5,
//     sprfs1.f:531-633), after all children are done;
//   the backward pass (SPRBS4/SPRBS6, sprbs1.f:251-574), one item per
//     supernode, after its parent's.
// The ascending items go by depth, deepest first, and the backward items by
// height, highest first, so that the longest chains start first and the
// rest of the forest fills the machine around them.  Workgroups take chunks
// of items of one depth (height); every wait is for items of lower-indexed
// workgroups, which the hardware dispatches first, so the launch completes
// whatever the residency (the dispatcher is the work queue: a counter that
// wavefronts claim items from serialises at its one address, about 20 ns a
// claim).  Each wait is bounded; one that gives up sets a flag that makes
// every later wait give up at once.
//
// Data handed from one item to another (update blocks, update vectors,
// factored panels, solution entries) is stored and loaded with agent-scope
// atomic accesses, which are coherent across the XCDs' L2 caches on their
// own; an item signals its consumers only after all its stores have
// completed (s_waitcnt), so no cache-wide fence is needed.
//
// The front is assembled in LDS as a packed lower triangle in the layout of
// the reference's packed trapezoids (column c from the diagonal down,
// starting at c*F - c(c-1)/2), so the trapezoid is one contiguous copy of
// LNZ; children's update blocks are stored packed the same way and added
// entry by entry through the child's relative map, children in SIBL order.
// Pivot rules and outputs (d, a_jj, DCOR, NUMNEG and TOL statistics through
// the common kernels, the factor store) are those of the level schedule's
// small-front kernel.  Results agree with the level schedule's to rounding
// (the operations per entry are the same; the forward solve is folded into
// the elimination).
#include "lanes.h"
#include "pivots.h"
#include "sprgpu_internal.h"

namespace sprgpu {

namespace {

constexpr unsigned kForestSpin = 1u << 24;       // polls of one wait (each after s_sleep 2)
constexpr int kForestTrap = kSmallF * (kSmallF + 1) / 2;  // packed front entries, F <= kSmallF
constexpr int kForestPerLane = (kForestTrap + 63) / 64;   // ... per lane
constexpr int kRound = 16;  // loads per lane in flight in the front assembly

template <typename X>
__device__ __forceinline__ void put(X *p, X v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename X>
__device__ __forceinline__ X get(const X *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Waits until every store of the wavefront has completed (the signal that
// follows may then be seen only after the data).
__device__ __forceinline__ void stores_done()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// First entry of packed column c of an order-n lower triangle
__device__ __forceinline__ int pcol(int c, int n) { return c * n - (c * (c - 1)) / 2; }

// First entry of row r of a row-packed lower triangle
__device__ __forceinline__ int prow(int r) { return (r * (r + 1)) >> 1; }

// Row of entry e of a row-packed lower triangle (e < 2^20: the float root
// is then within one of the answer, and one correction each way is enough)
__device__ __forceinline__ int prow_of(int e)
{
    int r = int((sqrtf(8.0f * e + 1.0f) - 1.0f) * 0.5f);
    r -= (prow(r) > e);
    r += (prow(r + 1) <= e);
    return r;
}

// Column of packed entry e of an order-n lower triangle (0 <= e < n(n+1)/2,
// n <= 64)
__device__ __forceinline__ int pcol_of(int e, int n)
{
    const float b = 2.0f * n + 1.0f;
    int c = int((b - sqrtf(b * b - 8.0f * e)) * 0.5f);
    c -= (c > 0 && pcol(c, n) > e);
    c += (c + 1 < n && pcol(c + 1, n) <= e);
    return c;
}

// Waits (the whole wavefront) until *p >= target.  false: the wait gave up,
// or an earlier one did (sync[1] set).
__device__ bool wait_for(const unsigned *p, unsigned target, unsigned *sync,
                         FactorStatus *st, int *err)
{
    unsigned spins = 0;
    for (;;) {
        // both looks in one memory round trip
        const unsigned v0 = get(p), d0 = get(sync + 1);
        const unsigned v = __builtin_amdgcn_readfirstlane(v0);
        if (v >= target) return true;
        const unsigned dead = __builtin_amdgcn_readfirstlane(d0);
        if (dead) return false;
        if (++spins > kForestSpin) {
            if ((threadIdx.x & 63) == 0) {
                put(sync + 1, 1u);
                if (st) atomicAdd(&st->wait_timeout, 1ull);
                if (err) atomicExch(err, 1);
            }
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

template <typename T>
struct ForestLds {
    T front[kForestTrap + kSmallF];  // row-packed front (ascending pass) or panel (backward
                                     // pass); row r at r(r+1)/2 (reads run one row past)
    double ash[kSmallF];      // a_jj of the supernode's columns
    double yl[kSmallF];       // the local vector of the forward step
    double xs[kSmallF];       // the backward step's vector entries
    T lcol[2 * kSmallF];      // multipliers of the elimination step (zeros past the front)
    int mapl[kSmallF];        // a child's rows in the front
};

template <typename T, bool FACT, bool SOLVE>
__device__ __forceinline__ void ascend(const ForestArgs &a, int id, ForestLds<T> &L)
{
    const ForestNode nd = a.fnodes[id];
    const int F = nd.F, S = nd.S, r = threadIdx.x & 63;
    const bool rv = r < F;
    unsigned *cdone = a.sync + 2, *adone = cdone + a.n;
    T *store = static_cast<T *>(a.store);
    T *ublk = static_cast<T *>(a.ublk);
    const i32 *rows = a.lindx + nd.lxoff;
    const PivotArgs pa = a.pa;

    // 1. the trapezoid (FACT) and the own rows of b (SOLVE)
    if (FACT) {
        const int tri = F * (F + 1) / 2, trap = S * F - (S * (S - 1)) / 2;
#pragma unroll
        for (int t = 0; t < kForestPerLane; ++t)
            if (r + 64 * t < tri) L.front[r + 64 * t] = T(0);
        wave_sync();
        // the trapezoid, contiguous in LNZ (column-packed), into the
        // row-packed front, kRound loads per lane in flight at a time
        unsigned long long nnz = 0;
#pragma nounroll
        for (int t0 = 0; 64 * t0 < trap; t0 += kRound) {
            double tv[kRound];
#pragma unroll
            for (int u = 0; u < kRound; ++u) {
                const int e = r + 64 * (t0 + u);
                tv[u] = (e < trap) ? a.lnz[nd.xlnz + e] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kRound; ++u) {
                const int e = r + 64 * (t0 + u);
                if (e < trap) {
                    const int c = pcol_of(e, F), rr = c + (e - pcol(c, F));
                    L.front[prow(rr) + c] = T(tv[u]);
                    nnz += (tv[u] != 0.0);
                }
            }
        }
        for (int o = 32; o > 0; o >>= 1) nnz += __shfl_down(nnz, o);
        if (r == 0 && nnz) atomicAdd(&a.st->nnz_seen, nnz);
        wave_sync();
        if (r < S) {  // the pivots' a_jj: sprfc1.f:473-480, sprfc3.f:560-575
            const int eq = nd.first + r;
            double v = a.lnz[nd.xlnz + pcol(r, F)];
            if (pa.mode >= 3) {
                a.dcor[2 * i64(eq)] = 0.0;
                a.dcor[2 * i64(eq) + 1] = 0.0;
            }
            if (v == 0.0) {
                if (pa.mode >= 3) {
                    v = 1.0;
                    a.dcor[2 * i64(eq) + 1] = 1.0;
                    L.front[prow(r) + r] = T(1);
                } else {
                    atomicAdd(&a.st->nzero, 1ull);
                    atomicMin(&a.st->first_zero, (unsigned long long)eq);
                }
            }
            a.adiag[eq] = v;
            L.ash[r] = v;
        }
    }
    if (SOLVE && r < kSmallF) L.yl[r] = (r < S) ? get(a.b + rows[r]) : 0.0;
    wave_sync();
    // (everything above is the item's own data, which no child writes: it is
    // loaded while the children may still be running, and only what follows
    // waits for them)
    ForestChild chn = (nd.nchild > 0) ? a.fchild[nd.child0] : ForestChild{};
    if (nd.nchild > 0) wait_for(cdone + id, unsigned(nd.nchild), a.sync, a.st, a.err);

    // 2. the children's update blocks and update vectors, one child at a
    //    time in SIBL order (a front entry receives at most one entry per
    //    child); a child's map, update vector and first round of its update
    //    block are loaded together, and the next child's record meanwhile
    for (int k = 0; k < nd.nchild; ++k) {
        const ForestChild ch = chn;
        if (k + 1 < nd.nchild) chn = a.fchild[nd.child0 + k + 1];
        const int U = ch.U;
        const int cnt = FACT ? U * (U + 1) / 2 : 0;
        T uv[kRound];
#pragma unroll
        for (int u = 0; u < kRound; ++u) {
            const int e = r + 64 * u;
            uv[u] = (e < cnt) ? get(ublk + ch.uoff + e) : T(0);
        }
        if (r < U) L.mapl[r] = a.relmap[ch.mapoff + r];
        double vv = 0.0;
        if (SOLVE && r < U) vv = get(a.vwork + ch.vofs + r);
        wave_sync();
        if (FACT) {  // the row-packed update block, kRound loads per lane at a time
#pragma nounroll
            for (int t0 = 0; 64 * t0 < cnt; t0 += kRound) {
                if (t0 > 0) {
#pragma unroll
                    for (int u = 0; u < kRound; ++u) {
                        const int e = r + 64 * (t0 + u);
                        uv[u] = (e < cnt) ? get(ublk + ch.uoff + e) : T(0);
                    }
                }
#pragma unroll
                for (int u = 0; u < kRound; ++u) {
                    const int e = r + 64 * (t0 + u);
                    if (e < cnt) {
                        const int i = prow_of(e), j = e - prow(i);
                        L.front[prow(L.mapl[i]) + L.mapl[j]] += uv[u];
                    }
                }
            }
        }
        if (SOLVE && r < U) L.yl[L.mapl[r]] += vv;
        wave_sync();
    }

    // 3. the elimination (FACT): lane r holds its row in registers, shifted
    //    left by one column per step so that column j is always m[0] (the
    //    steps are then a loop, not unrolled code: a fully unrolled
    //    elimination is too large for the instruction cache); entries right
    //    of the diagonal are scratch.  Each step stores its column of the
    //    factored panel.  The forward step is folded in (SOLVE): at step j
    //    every lane below j subtracts l_rj y_j from its y.
    double dr = 0.0, y = 0.0;
    if (SOLVE && rv) y = L.yl[r];
    if (FACT) {
        T m[kSmallF];
#pragma unroll
        for (int c = 0; c < kSmallF; ++c) m[c] = (rv && c <= r) ? L.front[prow(r) + c] : T(0);
        L.lcol[kSmallF + r] = T(0);  // read by the shifted update past the front
        // the pivot of each step is formed during the previous step's update
        // (its column is updated first), so that its reciprocal overlaps the
        // rest of that update
        double d_in = double(lane_bcast(m[0], 0));
        double d = pivot_value(d_in, L.ash[0], pa);
        T inv = T(1) / T(d);
        for (int j = 0; j < S; ++j) {
            if (r == 0) pivot_record(d_in, d, L.ash[j], nd.first + j, pa, a.dvec, a.dcor, a.st);
            const bool below = rv && r > j;
            const T w = below ? m[0] : T(0);
            const T l = w * inv;
            L.lcol[r] = below ? l : T(0);
            if (SOLVE) {
                const double yj = lane_bcast(y, j);
                y -= double(l) * yj;  // l = 0 on and above the pivot row
            }
            if (r == j) {
                dr = d;
                put(store + nd.loff + r + i64(j) * nd.ld, T(d));
            }
            if (below) put(store + nd.loff + r + i64(j) * nd.ld, l);
            wave_sync();
            // m[k-1] = entry (r, j+k) - w l_(j+k)j: the update and the shift
            // (straight-line code: every register is rewritten once per step),
            // column j+1 first, then the next pivot
            const T *lc = L.lcol + j;
            m[0] = m[1] - w * lc[1];
            if (j + 1 < S) {
                d_in = double(lane_bcast(m[0], j + 1));
                d = pivot_value(d_in, L.ash[j + 1], pa);
                inv = T(1) / T(d);
            }
#pragma unroll
            for (int k = 2; k < kSmallF; ++k) {
                if (k % 16 == 0) __builtin_amdgcn_sched_barrier(0);  // bounds the lc reads in flight
                m[k - 1] = m[k] - w * lc[k];
            }
            m[kSmallF - 1] = T(0);
            wave_sync();  // lcol is rewritten by the next step
        }
        // the update block (row r - S, columns S + k), packed, for the parent
        if (rv && r >= S) {
            T *ub = ublk + nd.uoff + prow(r - S);  // row r - S of the row-packed block
#pragma unroll
            for (int k = 0; k < kSmallF; ++k)
                if (S + k <= r) put(ub + k, m[k]);
        }
    } else if (SOLVE) {
        // the rows of the factored panel below the diagonal into LDS
        const T *pr = store + nd.loff + r;  // row r of the panel, column by column
#pragma unroll
        for (int c = 0; c < kSmallF; ++c, pr += nd.ld)
            if (rv && c < S && c < r) L.front[prow(r) + c] = *pr;
        if (r < S) dr = a.dvec[nd.first + r];
        wave_sync();
        for (int j = 0; j < S; ++j) {
            const double yj = lane_bcast(y, j);
            if (rv && r > j) y -= double(L.front[prow(r) + j]) * yj;
        }
    }

    // 4. z = x / d of the own rows into b and the update vector for the
    //    parent (SOLVE)
    if (SOLVE && rv) {
        if (r < S)
            put(a.b + rows[r], y / dr);
        else
            put(a.vwork + nd.voff + r, y);
    }
    stores_done();
    if (r == 0) {
        put(adone + id, 1u);  // for this item's backward step: its panel is final
        if (nd.parent >= 0)
            __hip_atomic_fetch_add(cdone + nd.parent, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename T>
__device__ __forceinline__ void descend(const ForestArgs &a, int id, ForestLds<T> &L)
{
    const ForestNode nd = a.fnodes[id];
    const int F = nd.F, S = nd.S, r = threadIdx.x & 63;
    unsigned *cdone = a.sync + 2, *adone = cdone + a.n, *bdone = adone + a.n;
    // the panel's rows below the diagonal (lane = row), once this item's own
    // ascending step has made them (when it ran in this launch), loaded
    // while the parent's backward step may still be running
    if (a.phases & 1) wait_for(adone + id, 1u, a.sync, a.st, a.err);
    const T *store = static_cast<const T *>(a.store);
    const i32 *rows = a.lindx + nd.lxoff;
    T lv[kSmallF];
    const T *pr = store + nd.loff + r;  // row r of the panel, column by column
#pragma unroll
    for (int c = 0; c < kSmallF; ++c, pr += nd.ld) lv[c] = (r < F && c < S && c < r) ? get(pr) : T(0);
    // then the vector entries of the front's rows (z of the own rows, x of
    // the ancestors' rows), once the parent's backward step is done
    if (nd.parent >= 0) wait_for(bdone + nd.parent, 1u, a.sync, a.st, a.err);
    const double xr = (r < F) ? get(a.b + rows[r]) : 0.0;
    if (r < F) L.xs[r] = xr;
#pragma unroll
    for (int c = 0; c < kSmallF; ++c)
        if (r < F && c < S && c < r) L.front[prow(r) + c] = lv[c];
    wave_sync();
    // lane = column c < S: x_c = z_c - sum_{r >= S} l_rc x_r (rows ascending),
    // then the unit upper solve with L11^T from the last column: x_c -= l_rc
    // x_r for r = S-1 .. c+1
    double acc = (r < S) ? xr : 0.0;
    if (r < S) {
        for (int q = S; q < F; ++q) acc -= double(L.front[prow(q) + r]) * L.xs[q];
    }
    for (int q = S - 1; q > 0; --q) {
        const double xq = lane_bcast(acc, q);
        if (r < q) acc -= double(L.front[prow(q) + r]) * xq;
    }
    if (r < S) put(a.b + rows[r], acc);
    stores_done();
    if (r == 0) put(bdone + id, 1u);
}

}  // namespace

// Workgroup w takes one chunk of items of one depth (ascending chunks first)
// or height (the backward chunks), its kForestWaves wavefronts the chunk's
// items round robin, each with its own LDS.  Items of one chunk do not
// depend on each other, and every item waits only for items of
// lower-indexed workgroups, which the hardware dispatches first.
template <typename T, bool FACT, bool SOLVE>
__global__ __launch_bounds__(64 * kForestWaves, 2) void k_forest(ForestArgs a)
{
    __shared__ ForestLds<T> Ls[kForestWaves];
    const int v = threadIdx.x >> 6;
    ForestLds<T> &L = Ls[v];
    const int nch1 = (a.phases & 1) ? a.nasc : 0;
    const int w = blockIdx.x;
    if (w < nch1) {
        const int i0 = a.chunks[w], i1 = a.chunks[w + 1];
        for (int i = i0 + v; i < i1; i += kForestWaves) ascend<T, FACT, SOLVE>(a, a.order[i], L);
    } else if (SOLVE) {
        const int q = a.nasc + (w - nch1);
        const int i0 = a.chunks[q], i1 = a.chunks[q + 1];
        for (int i = i0 + v; i < i1; i += kForestWaves) descend<T>(a, a.order[i], L);
    }
}

__global__ __launch_bounds__(256) void k_forest_reset(unsigned *sync, i64 count)
{
    for (i64 i = blockIdx.x * i64(256) + threadIdx.x; i < count; i += i64(gridDim.x) * 256)
        sync[i] = 0u;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T, bool FACT, bool SOLVE>
static int forest_launch(sprgpu_handle *h, const double *lnz, int mode, const double *tol,
                         double *b, int phases)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    ForestArgs a{};
    a.fnodes = d.forest_nodes;
    a.fchild = d.forest_child;
    a.n = i32(p.nsuper);
    a.phases = phases;
    a.order = d.forest_order;
    a.chunks = d.forest_chunks;
    a.nasc = p.forest_nasc;
    a.nbwd = i32(p.forest_chunks.size()) - 1 - p.forest_nasc;
    a.relmap = d.relmap;
    a.lindx = d.lindx;
    a.lnz = lnz;
    a.ublk = d.ublk;
    a.store = d.store;
    a.dvec = d.dvec;
    a.adiag = d.adiag;
    a.dcor = d.dcor;
    a.st = FACT ? d.status : nullptr;
    a.err = FACT ? nullptr : d.spin_error;
    a.pa = PivotArgs{tol ? tol[0] : 0.0, tol ? tol[3] : 0.0, mode};
    a.b = b;
    a.vwork = d.vwork;
    a.sync = d.forest_sync;
    const i64 ns = 2 + 3 * p.nsuper;
    k_forest_reset<<<unsigned(std::min<i64>((ns + 255) / 256, 1024)), 256, 0, h->stream>>>(
        d.forest_sync, ns);
    const i64 nasc = p.forest_nasc, nbwd = i64(p.forest_chunks.size()) - 1 - nasc;
    const i64 wgs = ((phases & 1) ? nasc : 0) + ((SOLVE && (phases & 2)) ? nbwd : 0);
    if (wgs > 0) k_forest<T, FACT, SOLVE><<<unsigned(wgs), 64 * kForestWaves, 0, h->stream>>>(a);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int forest_factor(sprgpu_handle *h, const double *sm_dev, int mode, const double *tol, double *b)
{
    const double *lnz = sm_dev + h->plan.neq;
    if (b) {  // factor and solve of one right-hand side in the same launch
        if (h->precision == SPRGPU_FP32) return forest_launch<float, true, true>(h, lnz, mode, tol, b, 3);
        return forest_launch<double, true, true>(h, lnz, mode, tol, b, 3);
    }
    if (h->precision == SPRGPU_FP32) return forest_launch<float, true, false>(h, lnz, mode, tol, nullptr, 1);
    return forest_launch<double, true, false>(h, lnz, mode, tol, nullptr, 1);
}

int forest_solve(sprgpu_handle *h, double *b, int which)
{
    if (h->precision == SPRGPU_FP32) return forest_launch<float, false, true>(h, nullptr, 1, nullptr, b, which);
    return forest_launch<double, false, true>(h, nullptr, 1, nullptr, b, which);
}

}  // namespace sprgpu
