// Multifrontal supernodal LDL' factorization on gfx950.
//
// Replaces the numeric work of SPRF11/SPRF13/SPRF15/SPRF17 and the
// LMVY18/LMMY18 callbacks (sprfc1.f:286-1299, lmvy18.f, lmmy18.f).  The
// reference walks the supernodes in postorder with a LIFO stack of packed
// update matrices; here every front lives in one HBM pool as a dense
// column-major block, and all supernodes of one tree height are processed by
// the same launches:
//   k_front_init   copy the assembled trapezoid (packed LNZ) into the front,
//                  zero the update block                      (sprfc1.f:518-523)
//   k_extend_add   add each child's update block into its parent through
//                  the child->parent relative map            (SPRF12/SPRF13)
//   k_small        fronts with F<=64: whole LDL' + Schur update in LDS
//   k_panel        blocked fronts: LDL' of a kPanelNB-wide diagonal block and
//                  the panel TRSM of the rows below it         (SPRF17)
//   k_gemm         MFMA update C -= L D L^T on lower tiles: the trailing
//                  update inside the panel and the Schur complement of the
//                  update block                              (LMVY18/LMMY18)
// Pivots follow the reference's rules: d is tested against TOL(1)*|a_jj|
// with a_jj the assembled diagonal (sprfc1.f:1272-1279), columns are scaled
// by the reciprocal 1/d (sprfc1.f:1288-1294), and in SPRFC3 mode small pivots
// are replaced as in sprfc3.f:1083-1094.
#include <cfloat>
#include <climits>
#include <type_traits>

#include "lanes.h"
#include "pivots.h"
#include "sprgpu_internal.h"

namespace sprgpu {

typedef double dv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA wrappers.  Both take A[i][k] from lane i + 16k and B[k][j]
// from lane j + 16k.  The 4 accumulator elements of a lane hold column
// j = lane & 15 and rows that differ between the two instructions (measured
// on gfx950): f32 element e is row 4*(lane>>4) + e, f64 element e is row
// (lane>>4) + 4*e.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef dv4 acc_t;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int e) { return (lane >> 4) + 4 * e; }
};
template <> struct Mfma<float> {
    typedef fv4 acc_t;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};

// ---------------------------------------------------------------------------
// status reset (a kernel rather than a memset so that the sequence can be
// captured in a graph)
// ---------------------------------------------------------------------------
__global__ void k_status_reset(FactorStatus *st)
{
    if (threadIdx.x == 0) {
        st->first_fail = ULLONG_MAX;
        st->wait_timeout = 0;
        st->first_zero = ULLONG_MAX;
        st->nzero = 0;
        st->nrepl = 0;
        st->nnz_seen = 0;
    }
}

// ---------------------------------------------------------------------------
// front initialisation
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_front_init(const ColTask *__restrict__ tasks,
                                                    const NodeInfo *__restrict__ nodes,
                                                    const double *__restrict__ lnz,
                                                    T *__restrict__ pool, double *adiag,
                                                    double *dcor, FactorStatus *st, int mode)
{
    const ColTask t = tasks[blockIdx.x];
    const NodeInfo n = nodes[t.node];
    unsigned long long nnz = 0;
    for (int c = t.c0; c < t.c1; ++c) {
        T *col = pool + n.foff + i64(c) * n.ld;
        if (c < n.S) {
            // packed column c of the trapezoid starts at row c:
            // offset c*F - c*(c-1)/2 (OFFSET(F,c) in sprfc1.f)
            const double *src = lnz + n.xlnz + i64(c) * n.F - (i64(c) * (c - 1)) / 2;
            int r = c + 1 + threadIdx.x;
            // rows below the diagonal, eight loads in flight per thread
            for (; r + 7 * 256 < n.F; r += 8 * 256) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = src[r + 256 * u - c];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    nnz += (v[u] != 0.0);
                    col[r + 256 * u] = T(v[u]);
                }
            }
            for (; r < n.F; r += 256) {
                const double v = src[r - c];
                nnz += (v != 0.0);
                col[r] = T(v);
            }
            if (threadIdx.x == 0) {  // the diagonal: a_jj, zero pivots, DCOR
                double v = src[0];
                nnz += (v != 0.0);
                const int eq = n.first + c;
                if (mode >= 3) {
                    dcor[2 * i64(eq)] = 0.0;
                    dcor[2 * i64(eq) + 1] = 0.0;
                }
                if (v == 0.0) {
                    if (mode >= 3) {  // sprfc3.f:560-575, sprfc4.f:459-472
                        v = 1.0;
                        dcor[2 * i64(eq) + 1] = 1.0;
                    } else {          // sprfc1.f:473-480
                        atomicAdd(&st->nzero, 1ull);
                        atomicMin(&st->first_zero, (unsigned long long)eq);
                    }
                }
                adiag[eq] = v;
                col[c] = T(v);
            }
        } else {
            for (int r = c + threadIdx.x; r < n.F; r += 256) col[r] = T(0);
        }
    }
    // block reduction of the non-zero count
    __shared__ unsigned long long red[4];
    for (int o = 32; o > 0; o >>= 1) nnz += __shfl_down(nnz, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nnz;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long s = red[0] + red[1] + red[2] + red[3];
        if (s) atomicAdd(&st->nnz_seen, s);
    }
}

// ---------------------------------------------------------------------------
// extend-add: parent(map[i], map[j]) += U_child(i, j), i >= j
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_extend_add(const EaTask *__restrict__ tasks,
                                                    const NodeInfo *__restrict__ nodes,
                                                    const i32 *__restrict__ relmap,
                                                    T *__restrict__ pool)
{
    const EaTask t = tasks[blockIdx.x];
    const NodeInfo c = nodes[t.child];
    const NodeInfo p = nodes[c.parent];
    const i32 *map = relmap + c.mapoff;
    const int U = c.F - c.S;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int jc = t.c0 + wave; jc < t.c1; jc += 4) {
        const int pj = map[jc];
        const T *src = pool + c.foff + i64(c.S + jc) * c.ld + c.S;
        T *dst = pool + p.foff + i64(pj) * p.ld;
        // eight rows per lane at a time: the map and child loads, then the
        // parent loads, then the stores, so each group costs two memory
        // round trips instead of three per row
        int i = jc + lane;
        for (; i + 7 * 64 < U; i += 8 * 64) {
            int m[8];
            T v[8], d[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                m[u] = map[i + 64 * u];
                v[u] = src[i + 64 * u];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) d[u] = dst[m[u]];
#pragma unroll
            for (int u = 0; u < 8; ++u) dst[m[u]] = d[u] + v[u];
        }
        for (; i < U; i += 64) dst[map[i]] += src[i];
    }
}

// ---------------------------------------------------------------------------
// small fronts (F <= kSmallF): assembly and factor in LDS, one wavefront each
// ---------------------------------------------------------------------------
// Each wavefront (lane = front row) builds its front in LDS: the supernode's
// trapezoid from SM (as k_front_init), then the children's update blocks
// added in SIBL order (as k_extend_add), then eliminates the S columns
// (SPRF17/LMMY18's recurrence: l = w * (1/d), a_rc -= w_r l_c) and writes the
// factored panel to the factor store, where the solves read it, and the
// update block back to the pool, where the parent's extend-add reads it.  A workgroup holds `waves` fronts, one per wavefront; the
// wavefronts synchronise only among their own lanes.  Dynamic LDS per
// wavefront: F x (F+1) elements of T (F = the launch's largest front) plus
// the pivots' a_jj and kSmallF multipliers.
__host__ __device__ __forceinline__ size_t small_front_bytes(int maxF, size_t esz)
{
    return ((size_t(maxF) * (maxF + 1) * esz + 7) / 8 * 8 + 8 * size_t(maxF) + esz * kSmallF +
            15) / 16 * 16;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__global__ __launch_bounds__(256) void k_small(const i32 *__restrict__ list, int count, int maxF,
                                               const NodeInfo *__restrict__ nodes,
                                               const i32 *__restrict__ children,
                                               const i32 *__restrict__ relmap,
                                               const double *__restrict__ lnz,
                                               T *__restrict__ pool, T *__restrict__ store,
                                               double *dvec, double *adiag,
                                               double *dcor, FactorStatus *st, PivotArgs pa)
{
    extern __shared__ double smem_small[];
    const int wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    const int q = blockIdx.x * waves + wave;
    if (q >= count) return;
    const int LDM = maxF + 1;
    char *base = reinterpret_cast<char *>(smem_small) + wave * small_front_bytes(maxF, sizeof(T));
    T *M = reinterpret_cast<T *>(base);  // M[r + c*LDM], lower triangle
    double *ash = reinterpret_cast<double *>(base + (size_t(maxF) * LDM * sizeof(T) + 7) / 8 * 8);
    T *lcol = reinterpret_cast<T *>(ash + maxF);  // the scaled column of the step
    const NodeInfo n = nodes[list[q]];
    const int F = n.F, S = n.S, r = threadIdx.x & 63;
    const bool rv = r < F;
    // 1. trapezoid from SM, zero elsewhere.  The loads go eight columns at a
    //    time, unconditionally (entries outside the trapezoid read its first
    //    entry and are not used), so that eight are in flight together.
    unsigned long long nnz = 0;
    constexpr int U8 = 8;
    for (int c0 = 0; c0 < F; c0 += U8) {
        double lv[U8];
#pragma unroll
        for (int u = 0; u < U8; ++u) {
            const int c = c0 + u;
            const bool in = rv && c < S && r >= c;
            lv[u] = lnz[n.xlnz + (in ? i64(c) * F - (i64(c) * (c - 1)) / 2 + (r - c) : 0)];
        }
#pragma unroll
        for (int u = 0; u < U8; ++u) {
            const int c = c0 + u;
            if (c >= F || !rv || r < c) continue;
            double v = 0.0;
            if (c < S) {
                v = lv[u];
                nnz += (v != 0.0);
                if (r == c) {
                    const int eq = n.first + c;
                    if (pa.mode >= 3) {
                        dcor[2 * i64(eq)] = 0.0;
                        dcor[2 * i64(eq) + 1] = 0.0;
                    }
                    if (v == 0.0) {
                        if (pa.mode >= 3) {  // sprfc3.f:560-575, sprfc4.f:459-472
                            v = 1.0;
                            dcor[2 * i64(eq) + 1] = 1.0;
                        } else {             // sprfc1.f:473-480
                            atomicAdd(&st->nzero, 1ull);
                            atomicMin(&st->first_zero, (unsigned long long)eq);
                        }
                    }
                    adiag[eq] = v;
                    ash[c] = v;
                }
            }
            M[r + c * LDM] = T(v);
        }
    }
    for (int o = 32; o > 0; o >>= 1) nnz += __shfl_down(nnz, o);
    if (r == 0 && nnz) atomicAdd(&st->nnz_seen, nnz);
    wave_sync();
    // 2. children's update blocks, one child at a time in SIBL order (each
    //    entry of the front is added to at most once per child); the child's
    //    entries and map values are loaded eight columns at a time,
    //    unconditionally (clamped into the child's update block)
    for (int k = 0; k < n.nchild; ++k) {
        const NodeInfo c = nodes[children[n.child0 + k]];
        const i32 *map = relmap + c.mapoff;
        const int U = c.F - c.S;
        if (U <= 0) continue;  // (a child always has update rows)
        const int rr = min(r, U - 1);
        const int pr = map[rr];
        const T *src = pool + c.foff + c.S + rr + i64(c.S) * c.ld;
        for (int j0 = 0; j0 < U; j0 += U8) {
            T sv[U8];
            int mj[U8];
#pragma unroll
            for (int u = 0; u < U8; ++u) {
                const int j = min(j0 + u, rr);
                sv[u] = src[i64(j) * c.ld];
                mj[u] = map[j];
            }
#pragma unroll
            for (int u = 0; u < U8; ++u)
                if (r < U && j0 + u <= r) M[pr + mj[u] * LDM] += sv[u];
        }
        wave_sync();
    }
    // 3. eliminate the S columns with the lane's row in registers (entries
    //    right of its diagonal are scratch): step j takes the pivot from lane
    //    j, the lanes below it scale their entry of column j and write the
    //    multiplier to LDS, and every lane updates its entries right of
    //    column j (w = 0 from the pivot row up).  Each entry of the lower
    //    triangle takes the same operations as with the front in LDS.
    T m[kSmallF];
#pragma unroll
    for (int c = 0; c < kSmallF; ++c) m[c] = (rv && c <= r) ? M[r + c * LDM] : T(0);
#pragma clang loop unroll(full)
    for (int j = 0; j < kSmallF; ++j) {
        if (j >= S) continue;  // (a constant trip count: every copy's j is known)
        const double d_in = double(lane_bcast(m[j], j));
        const double d = pivot_value(d_in, ash[j], pa);
        if (r == 0) pivot_record(d_in, d, ash[j], n.first + j, pa, dvec, dcor, st);
        const T inv = T(1) / T(d);
        const T w = (rv && r > j) ? m[j] : T(0);
        const T l = w * inv;
        if (rv && r > j) lcol[r] = l;
        wave_sync();
#pragma clang loop unroll(full)
        for (int c = 1; c < kSmallF; ++c)
            if (c > j) m[c] -= w * lcol[c];
        if (r > j) m[j] = l;
        if (r == j) m[j] = T(d);
        wave_sync();  // lcol is rewritten by the next step
    }
    // 4. the factored panel (columns < S) to the factor store, the update
    //    block (the parent's extend-add reads it) back to the pool
    T *fr = pool + n.foff, *fl = store + n.loff;
#pragma unroll
    for (int c = 0; c < kSmallF; ++c)
        if (rv && c <= r) (c < S ? fl : fr)[r + i64(c) * n.ld] = m[c];
}

// Reciprocal from the hardware estimate refined by Newton steps (one for
// FP32, two for FP64); zero and infinite inputs give the estimate itself
// (inf and 0, as 1/d would).
__device__ __forceinline__ float recip(float d)
{
    const float r = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, r, 1.0f);
    const float x = __builtin_fmaf(r, e, r);
    return __builtin_isfinite(x) ? x : r;
}
__device__ __forceinline__ double recip(double d)
{
    const double r = __builtin_amdgcn_rcp(d);
    double x = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
    x = __builtin_fma(x, __builtin_fma(-d, x, 1.0), x);
    return __builtin_isfinite(x) ? x : r;
}

// LDL' of an nb x nb diagonal block (nb <= kPanelNB; entry (r, c) of its
// lower triangle at src[r*srs + c*scs]) by a workgroup of kChainThreads,
// right-looking over the columns (SPRF17's recurrence
// l_rj = (a_rj - sum_k l_rk d_k l_jk) / d_j, one column at a time).  Thread
// t holds entries (r, 4g..4g+3) of the block, r = t / 8, g = t % 8, in
// registers; entries right of a row's diagonal are scratch.  At step j the
// threads holding column j write it to LDS (rows above j as zeros) and the
// threads of row j write their part of row j of X = L11^-1; after one
// barrier every thread reads the pivot, its row's entry and the four
// column-j values and X(j, :) values it needs, and updates its entries
// unmasked (the zeros leave finished columns unchanged).  Steps go two
// columns per barrier (see the loop).  X is built as the
// product of the elimination steps, X(r, :) -= l_rj X(j, :), in the same
// loop.  The pivots are recorded (when record is set) by one thread per row
// after the loop.  Bm = L11^-T D^-1 = X^T D^-1 (upper triangular, zero
// outside nb x nb) is left in LDS.  The factored rows (l_rc for c < r, d_r at
// c = r) go to out[r*rs + c*cs] when out is set, with zeros right of the
// diagonal if zero_upper; the rows of Bm go to bm_out + r*kPanelNB when
// bm_out is set (rows r < nb), as agent-scope atomic stores if bm_shared
// (read by other workgroups of the same launch); out2, when set, receives
// the factored rows as out does (the lower triangle only).  lsh holds kLshSize elements of T and then
// 2 kPanelNB doubles of LDS scratch (16-byte aligned).  Returns the pivot d_r of the thread's row (zero for
// r >= nb).  Every thread of the workgroup calls it.
constexpr int kChainThreads = 256;
constexpr int kLshSize = 11 * kPanelNB;  // LDS scratch elements of T (plus kPanelNB doubles)

template <typename T>
__device__ __forceinline__ T diag_block_factor(const T *src, i64 srs, i64 scs, int nb, int eq0,
                                               const double *__restrict__ adiag, bool record,
                                               const PivotArgs &pa, double *dvec, double *dcor,
                                               FactorStatus *st, T (*Bm)[kPanelNB + 1], T *lsh,
                                               T *out, i64 rs, i64 cs, bool zero_upper, T *bm_out,
                                               bool bm_shared = false, T *out2 = nullptr)
{
    constexpr int NB = kPanelNB, Q = 4;
    static_assert(NB * NB == kChainThreads * Q, "four entries of the block per thread");
    typedef T vec_t __attribute__((ext_vector_type(Q)));
    const int t = threadIdx.x, r = t >> 3, g = t & 7, c0 = Q * g;
    T *pd = lsh;                 // [NB] pivots d_j
    T *pi = lsh + NB;            // [NB] 1/d_j
    T *pin = lsh + 2 * NB;       // [NB] pivots before the replacement rule
    double *ash = reinterpret_cast<double *>(lsh + 11 * NB);  // [NB] a_jj
    double *pdd = ash + NB;      // [NB] pivots d_j in FP64 (recorded in D and DCOR)
    if (t < nb) ash[t] = adiag[eq0 + t];
    T a[Q], x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        a[q] = (r < nb && c0 + q <= r) ? src[r * srs + (c0 + q) * scs] : T(0);
        x[q] = (c0 + q == r) ? T(1) : T(0);
    }
    // Two columns per barrier: after columns j and j+1 and rows j and j+1 of
    // X are in LDS, every thread takes step j and then step j+1 from the
    // column j+1 and X row j+1 that step j leaves, computed with the same
    // operations as the threads holding them would (so the factor is the
    // same as one column per barrier).  Buffers alternate per pair: one is
    // rewritten two pairs later, after every thread has passed the next
    // barrier.
    // The pairs are unrolled (j is a constant in each copy, so the thread's
    // entries of columns j, j+1 are named statically) and leave at nb.
    // Outside SPRFC3/SPRFC4 mode no pivot is replaced, and the FP64 test is
    // skipped.
    const bool replace = pa.mode >= 3;
    auto pivot = [&](T dj, double a, double &dp) -> T {
        if (replace) {
            dp = pivot_value(double(dj), a, pa);
            return T(dp);
        }
        dp = double(dj);
        return dj;
    };
    int jn = 0;  // the first column not yet taken
#pragma unroll
    for (int jj = 0; jj < NB / 2; ++jj) {
        if (2 * jj + 1 >= nb) break;
        const int j = 2 * jj;
        T *wa = lsh + (3 + 4 * ((j >> 1) & 1)) * NB, *wb = wa + NB, *xa = wb + NB, *xb = xa + NB;
        if (g == (j >> 2)) {  // j is even: both columns belong to group j >> 2
            const bool hi = (j & 3) == 2;
            const T va = hi ? a[2] : a[0], vb = hi ? a[3] : a[1];
            wa[r] = (r >= j) ? va : T(0);
            wb[r] = (r >= j + 1) ? vb : T(0);
        }
        if (r == j || r == j + 1) {
            vec_t xv;
#pragma unroll
            for (int q = 0; q < Q; ++q) xv[q] = x[q];
            *reinterpret_cast<vec_t *>((r == j ? xa : xb) + c0) = xv;
        }
        __syncthreads();
        // step j
        const T dj = wa[j];
        const T aj1 = wa[j + 1];  // A(j+1, j)
        const T dj1r = wb[j + 1];
        const T wra = wa[r], wrb = wb[r];
        const vec_t wca = *reinterpret_cast<const vec_t *>(wa + c0);
        const vec_t wcb = *reinterpret_cast<const vec_t *>(wb + c0);
        const vec_t xja = *reinterpret_cast<const vec_t *>(xa + c0);
        const vec_t xjb = *reinterpret_cast<const vec_t *>(xb + c0);
        double dp;
        const T d = pivot(dj, ash[j], dp);
        const T di = recip(d);
        const T la = (r > j) ? wra * di : T(0);
        // column j+1 and row j+1 of X after step j
        const T lj1 = aj1 * di;  // l(j+1, j)
        T dj1 = dj1r;
        dj1 -= lj1 * aj1;
        T wrb1 = wrb;
        wrb1 -= la * aj1;
        vec_t wcb1, xjb1;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int c = c0 + q;
            const T lc = (c > j) ? wca[q] * di : T(0);
            T w = wcb[q];
            w -= lc * aj1;
            wcb1[q] = (c >= j + 1) ? w : T(0);
            T xv = xjb[q];
            xv -= lj1 * xja[q];
            xjb1[q] = xv;
        }
        // step j+1
        double dp1;
        const T d1 = pivot(dj1, ash[j + 1], dp1);
        const T di1 = recip(d1);
        const T lb = (r > j + 1) ? wrb1 * di1 : T(0);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            a[q] -= la * wca[q];
            x[q] -= la * xja[q];
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            a[q] -= lb * wcb1[q];
            x[q] -= lb * xjb1[q];
        }
        if (g == (j >> 2)) {
            const T fa = (r == j) ? d : la, fb = (r == j + 1) ? d1 : lb;
            if ((j & 3) == 2) {
                a[2] = fa;
                a[3] = fb;
            } else {
                a[0] = fa;
                a[1] = fb;
            }
        }
        if (t == 0) {
            pd[j] = d;
            pi[j] = di;
            pin[j] = dj;
            pdd[j] = dp;
            pd[j + 1] = d1;
            pi[j + 1] = di1;
            pin[j + 1] = dj1;
            pdd[j + 1] = dp1;
        }
        jn = j + 2;
    }
    if (const int j = jn; j < nb) {  // an odd last column, one step
        T *wcol = lsh + (3 + 4 * ((j >> 1) & 1)) * NB, *xrow = wcol + 2 * NB;
        if (g == (j >> 2)) {
            T v = a[0];
#pragma unroll
            for (int q = 1; q < Q; ++q)
                if ((j & 3) == q) v = a[q];
            wcol[r] = (r >= j) ? v : T(0);
        }
        if (r == j) {
            vec_t xv;
#pragma unroll
            for (int q = 0; q < Q; ++q) xv[q] = x[q];
            *reinterpret_cast<vec_t *>(xrow + c0) = xv;
        }
        __syncthreads();
        const T dj = wcol[j];
        const T wr = wcol[r];
        const vec_t wc = *reinterpret_cast<const vec_t *>(wcol + c0);
        const vec_t xj = *reinterpret_cast<const vec_t *>(xrow + c0);
        const double dp = pivot_value(double(dj), ash[j], pa);
        const T d = T(dp);
        const T di = recip(d);
        const T l = (r > j) ? wr * di : T(0);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            a[q] -= l * wc[q];
            x[q] -= l * xj[q];
        }
        if (g == (j >> 2)) {
#pragma unroll
            for (int q = 0; q < Q; ++q)
                if ((j & 3) == q) a[q] = (r == j) ? d : l;
        }
        if (t == 0) {
            pd[j] = d;
            pi[j] = di;
            pin[j] = dj;
            pdd[j] = dp;
        }
    }
    __syncthreads();
    const T dr = (r < nb) ? pd[r] : T(0);
    const T dir = (r < nb) ? pi[r] : T(0);
    // The pivot's test, its replacement (SPRFC3) and its value in D and DCOR
    // are FP64 whatever T is: an FP32 front uses d rounded to FP32 for its
    // own columns only
    if (record && g == 0 && r < nb)
        pivot_record(double(pin[r]), pdd[r], ash[r], eq0 + r, pa, dvec, dcor, st);
    if (out && r < nb) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int c = c0 + q;
            if (c <= r || zero_upper) out[r * rs + c * cs] = (c <= r) ? a[q] : T(0);
            if (out2 && c <= r) out2[r * rs + c * cs] = a[q];
        }
    }
    // Bm[i][c] = X[c][i] / d_c: the threads of row c write column c (zero
    // for c >= nb)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const T v = x[q] * dir;
        Bm[c0 + q][r] = v;
        if (bm_out && c0 + q < nb && r < nb) {
            if (bm_shared)
                __hip_atomic_store(bm_out + i64(c0 + q) * NB + r, v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            else
                bm_out[i64(c0 + q) * NB + r] = v;
        }
    }
    return dr;
}

// ---------------------------------------------------------------------------
// panel step: factor the nb x nb diagonal block at column k0 and solve the
// rows below it (L21 = A21 L11^-T D^-1)
// ---------------------------------------------------------------------------
// Panel step: columns [k0, k0+nb) of a blocked front, nb <= kPanelNB.
// Every workgroup factors the nb x nb diagonal block itself
// (diag_block_factor, one wavefront), and the workgroup computes its
// kPanelRows rows below the block as the product L21 = A21 (L11^-T D^-1) on
// MFMA, one 64-row slice per wavefront.  The A21 loads are issued before the diagonal factor so their
// latency overlaps it.  Workgroup 0 records the pivots and stores the factored
// block in a scratch array indexed by equation; the diagonal blocks are
// written into the factor store after the whole factorization (k_diag_writeback),
// so no workgroup ever reads a block another one writes.
template <typename T>
__global__ __launch_bounds__(256) void k_panel(const PanelDesc *__restrict__ descs, int ndesc,
                                               const i64 *__restrict__ prefix,
                                               const NodeInfo *__restrict__ nodes,
                                               T *__restrict__ pool, double *dvec,
                                               const double *__restrict__ adiag, double *dcor,
                                               FactorStatus *st, PivotArgs pa, T *dscr)
{
    constexpr int NB = kPanelNB;
    static_assert(NB == 32 && kPanelRows == 256,
                  "one lane per diagonal-block row; four 64-row MFMA slices per workgroup");
    __shared__ T Bm[NB][NB + 1];  // L11^-T D^-1 (upper triangular), zero outside nb x nb
    __shared__ __attribute__((aligned(16))) T Lsh[kLshSize + 2 * kPanelNB * sizeof(double) / sizeof(T)];
    const int q = find_desc(prefix, ndesc, blockIdx.x);
    const PanelDesc pd = descs[q];
    const int rb = int(blockIdx.x - prefix[q]);
    const NodeInfo n = nodes[pd.node];
    T *fr = pool + n.foff;
    const i64 ld = n.ld;
    const int k0 = pd.k0, nb = pd.nb, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // this wavefront's 64 rows of A21 in the MFMA A layout: a[ti][ks] holds
    // A21(rw + 16 ti + (lane & 15), 4 ks + (lane >> 4))
    const int rw = k0 + nb + rb * kPanelRows + wave * 64;
    const bool active = rw < n.F;
    T a[4][NB / 4];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int ks = 0; ks < NB / 4; ++ks) {
            const int row = rw + ti * 16 + (lane & 15), col = ks * 4 + (lane >> 4);
            a[ti][ks] = (row < n.F && col < nb) ? fr[row + (k0 + col) * ld] : T(0);
        }
    diag_block_factor<T>(fr + k0 + k0 * ld, 1, ld, nb, n.first + k0, adiag, rb == 0, pa, dvec,
                         dcor, st, Bm, Lsh, rb == 0 ? dscr + i64(n.first + k0) * NB : nullptr, NB,
                         1, true, nullptr);
    __syncthreads();
    if (!active) return;
    typename Mfma<T>::acc_t acc[4][2];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = typename Mfma<T>::acc_t{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < NB / 4; ++ks) {
        T b[2];
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) b[tj] = Bm[ks * 4 + (lane >> 4)][tj * 16 + (lane & 15)];
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = Mfma<T>::op(a[ti][ks], b[tj], acc[ti][tj]);
    }
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = rw + ti * 16 + Mfma<T>::row(lane, e);
                const int col = tj * 16 + (lane & 15);
                if (row < n.F && col < nb) fr[row + (k0 + col) * ld] = acc[ti][tj][e];
            }
}

// Writes the factored diagonal blocks of every panel step into the factored
// panels in the store (after the whole factorization: the panels are there).
template <typename T>
__global__ __launch_bounds__(256) void k_diag_writeback(const PanelDesc *__restrict__ descs,
                                                        const NodeInfo *__restrict__ nodes,
                                                        T *__restrict__ store,
                                                        const T *__restrict__ dscr)
{
    const PanelDesc pd = descs[blockIdx.x];
    const NodeInfo n = nodes[pd.node];
    T *fr = store + n.loff;
    const i64 ld = n.ld;
    for (int idx = threadIdx.x; idx < pd.nb * pd.nb; idx += 256) {
        const int r = idx % pd.nb, c = idx / pd.nb;
        if (r >= c)
            fr[(pd.k0 + r) + (pd.k0 + c) * ld] = dscr[i64(n.first + pd.k0 + r) * kPanelNB + c];
    }
}

// ---------------------------------------------------------------------------
// outer-block factor: a kPanelOuter-wide column block of a blocked front
// ---------------------------------------------------------------------------
// The block [ob, ob+nbw) (descriptor k0 = ob, nb = nbw) is factored without
// nbw/kPanelNB panel steps and their rank-kPanelNB updates:
//   diagonal      one workgroup per front factors the nbw x nbw diagonal
//                 block: kPanelNB-column sub-steps of diag_block_factor,
//                 the sub-panel rows below inside the block (MFMA with Bm)
//                 and the rank-kPanelNB update of the rest of the block.
//   rows below    16 per wavefront, left-looking over the sub-steps:
//                 L(r, jb) = (A(r, jb) - sum_{kb<jb} L(r, kb) D_kb
//                 L(jb, kb)^T) Bm_jb.
// The operations are those of the panel steps (same recurrences, same Bm
// products); only the order of the rank-kPanelNB updates of a row changes.
// k_block_fused runs both in one launch: the diagonal workgroups first (one
// per descriptor), then the row workgroups, which take sub-step jb as soon
// as its operands are made, instead of after the whole block: Bm of jb and
// d_k L(r, k) for the sub-step's rows r at the block's earlier columns k.
// The diagonal workgroup stores them in the descriptor's publication area
// as agent-scope atomic stores over a fill of pub_pending (a signalling
// NaN, which no arithmetic produces), and a row workgroup reads each value
// with agent-scope atomic loads until it differs from the fill: each value
// is handed over by its own store, with no fence (an agent-scope fence
// writes back or invalidates a whole L2, under the far updates' traffic).
// A row workgroup waits only for a diagonal workgroup, which has a lower
// index and is dispatched first; each wait is bounded (a wait that gives up
// counts in the status, and the factorization reports an error).
// k_block_diag and k_block_trsm are the two halves as separate launches
// (the rows below then read Bm from a scratch array indexed by equation,
// and the earlier columns from the front).
constexpr int kBdWaves = 4;                                      // diagonal workgroup
static_assert(64 * kBdWaves == kChainThreads, "the diagonal chain uses the whole workgroup");
constexpr int kBdTiles = (kPanelOuter / 16) * (kPanelOuter / 16 + 1) / 2;
constexpr int kBdSlots = (kBdTiles + kBdWaves - 1) / kBdWaves;  // 16x16 tiles per wavefront
constexpr int kBtWaves = kTrsmWaves;                             // row workgroup
static_assert(kBtWaves == kBdWaves, "one workgroup size for the fused launch");

// (ti, tj) of tile t, column-major over the lower triangle of nt x nt tiles
__device__ __forceinline__ void tri_tile(int t, int nt, int &ti, int &tj)
{
    tj = 0;
    while (t >= nt - tj) {
        t -= nt - tj;
        ++tj;
    }
    ti = tj + t;
}

// Publication area of one descriptor (kPubElems elements of T): Bm of each
// sub-step (kPanelNB x kPanelNB, [k][c]), then d_k L(r, k) for the rows r of
// sub-steps 1.. (row r - kPanelNB of kPubRows) at the columns k < kPanelNB
// * (r / kPanelNB) of the block.
constexpr int kPubBm = (kPanelOuter / kPanelNB) * kPanelNB * kPanelNB;
constexpr int kPubRows = kPanelOuter - kPanelNB;
static_assert(kPubElems == kPubBm + kPubRows * kPubRows, "publication area layout");
constexpr unsigned kPubSpinLimit = 1u << 22;

template <typename T> struct PubBits;
template <> struct PubBits<float> {
    typedef unsigned type;
    static constexpr type pending = 0xFFA5A5A5u;
};
template <> struct PubBits<double> {
    typedef unsigned long long type;
    static constexpr type pending = 0xFFF7A5A5A5A5A5A5ull;
};
template <typename T>
__device__ __forceinline__ void pub_store(T *p, T x)
{
    __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T pub_load(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The value at p once its writer has stored it, given a first look v
// (bounded: a wait that gives up counts in the status, sets *dead so that
// later waits of the thread give up at once, and returns 0).
template <typename T>
__device__ __forceinline__ T pub_take(const T *p, T v, FactorStatus *st, bool &dead)
{
    unsigned spins = 0;
    while (__builtin_bit_cast(typename PubBits<T>::type, v) == PubBits<T>::pending) {
        if (dead || ++spins > kPubSpinLimit) {
            if (!dead) atomicAdd(&st->wait_timeout, 1ull);
            dead = true;
            return T(0);
        }
        __builtin_amdgcn_s_sleep(2);
        v = pub_load(p);
    }
    return v;
}

// Every element of the publication areas to pub_pending, before the
// factorization.
template <typename T>
__global__ __launch_bounds__(256) void k_pub_mark(T *pub, i64 n)
{
    typedef typename PubBits<T>::type B;
    B *q = reinterpret_cast<B *>(pub);
    for (i64 i = blockIdx.x * i64(256) + threadIdx.x; i < n; i += i64(gridDim.x) * 256)
        q[i] = PubBits<T>::pending;
}

template <typename T>
struct BdLds {
    T Bm[kPanelNB][kPanelNB + 1];
    alignas(16) T Lsh[kLshSize + 2 * kPanelNB * sizeof(double) / sizeof(T)];
    T Ls[kPanelOuter * (kPanelNB + 1)];  // the sub-step's two tile columns, rows relative to k0
    T dsh[kPanelNB];
};
// row stride of the row workgroup's operands (rows padded by 16 bytes)
template <typename T>
constexpr int bt_ld = kPanelNB + 16 / int(sizeof(T));
template <typename T>
struct BtLds {
    alignas(16) T Rs[(kPanelOuter - kPanelNB) * bt_ld<T>];  // d_k L(k0 + c, k)
    alignas(16) T Bs[kPanelNB * bt_ld<T>];                  // Bm of the sub-step
};

// The block's lower triangle lives in registers as 16x16 MFMA accumulator
// tiles, tile t owned by wavefront t % kBdWaves (column-major order, so the
// tiles of every sub-step's trailing update are spread evenly).  A sub-step
// moves its two tile columns into LDS, wavefront 0 factors the diagonal
// 32 x 32 part, all wavefronts multiply the rows below by Bm (writing L to
// the front) and then update their own tiles from the L rows in LDS.  With
// PUB, each sub-step's Bm and the d_k L(r, k) of the rows below it in
// the block go to the publication area as well (bmscr is not written).
template <typename T, bool PUB>
__device__ __forceinline__ void block_diag_body(const PanelDesc &pd, const NodeInfo &n,
                                                T *__restrict__ pool, double *dvec,
                                                const double *__restrict__ adiag, double *dcor,
                                                FactorStatus *st, const PivotArgs &pa, T *bmscr,
                                                T *pub, T *__restrict__ store, BdLds<T> &S)
{
    constexpr int NB = kPanelNB, LS = NB + 1;
    typedef typename Mfma<T>::acc_t acc_t;
    T *fr = pool + n.foff + pd.k0 + pd.k0 * n.ld;  // block origin
    T *fs = store + n.loff + pd.k0 + pd.k0 * n.ld;  // the same in the factor store
    const i64 ld = n.ld;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbw = pd.nb, nt = (nbw + 15) / 16, ntiles = nt * (nt + 1) / 2;
    const int lr = lane & 15;
    int sti[kBdSlots], stj[kBdSlots];
    acc_t acc[kBdSlots];
#pragma unroll
    for (int q = 0; q < kBdSlots; ++q) {
        const int t = q * kBdWaves + wave;
        sti[q] = -1;
        stj[q] = -1;
        acc[q] = acc_t{0, 0, 0, 0};
        if (t < ntiles) {
            tri_tile(t, nt, sti[q], stj[q]);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = sti[q] * 16 + Mfma<T>::row(lane, e), col = stj[q] * 16 + lr;
                if (row < nbw && col < nbw && row >= col) acc[q][e] = fr[row + col * ld];
            }
        }
    }
    for (int j = 0; j * NB < nbw; ++j) {
        const int k0 = j * NB;  // relative to the block
        const int nb = min(NB, nbw - k0);
        const int m = nbw - k0;  // rows of the sub-panel, diagonal block included
        // lane-derived addresses are recomputed in every sub-step rather than
        // kept in registers across the loop
        const int lane = opaque(tid & 63), lr = lane & 15;
        // 1. the sub-panel's tiles (tile columns 2j, 2j+1) to LDS
#pragma unroll
        for (int q = 0; q < kBdSlots; ++q) {
            if (stj[q] == 2 * j || stj[q] == 2 * j + 1) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = sti[q] * 16 + Mfma<T>::row(lane, e) - k0;
                    const int col = (stj[q] - 2 * j) * 16 + lr;
                    S.Ls[row * LS + col] = acc[q][e];
                }
            }
        }
        __syncthreads();
        // 2. the diagonal 32 x 32 block
        {
            T *bm_out = PUB ? pub + j * NB * NB : bmscr + i64(n.first + pd.k0 + k0) * NB;
            const T dr = diag_block_factor<T>(S.Ls, LS, 1, nb, n.first + pd.k0 + k0, adiag, true,
                                              pa, dvec, dcor, st, S.Bm, S.Lsh, fr + k0 + k0 * ld,
                                              1, ld, false, bm_out, PUB, fs + k0 + k0 * ld);
            if ((tid & 7) == 0) S.dsh[tid >> 3] = dr;
        }
        __syncthreads();
        // 3. rows below it: L = A Bm, 16-row slices dealt to the wavefronts
        const int nsl = (m + 15) / 16;
        for (int u = 2 + wave; u < nsl; u += kBdWaves) {
            acc_t o[2] = {acc_t{0, 0, 0, 0}, acc_t{0, 0, 0, 0}};
#pragma unroll
            for (int ks = 0; ks < NB / 4; ++ks) {
                const int k = ks * 4 + (lane >> 4);
                const T a = S.Ls[(u * 16 + lr) * LS + k];
#pragma unroll
                for (int h = 0; h < 2; ++h) o[h] = Mfma<T>::op(a, S.Bm[k][h * 16 + lr], o[h]);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = u * 16 + Mfma<T>::row(lane, e), col = h * 16 + lr;
                    const T v = (row < m && col < nb) ? o[h][e] : T(0);
                    S.Ls[row * LS + col] = v;
                    if (row < m && col < nb) {
                        fr[(k0 + row) + (k0 + col) * ld] = v;
                        fs[(k0 + row) + (k0 + col) * ld] = v;
                        if (PUB)  // (row >= NB: a row of a later sub-step)
                            pub_store(pub + kPubBm + (k0 + row - NB) * kPubRows + k0 + col,
                                      S.dsh[col] * v);
                    }
                }
        }
        __syncthreads();
        // 4. rank-nb update of the tiles right of the sub-panel, summed apart
        //    from the tile and subtracted once (for accuracy in FP32)
#pragma unroll
        for (int q = 0; q < kBdSlots; ++q) {
            if (stj[q] >= 2 * j + 2) {
                const int ra = sti[q] * 16 - k0 + lr, rb = stj[q] * 16 - k0 + lr;
                acc_t u = acc_t{0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < NB / 4; ++ks) {
                    const int k = ks * 4 + (lane >> 4);
                    const T a = S.Ls[ra * LS + k];
                    const T b = S.Ls[rb * LS + k] * S.dsh[k];
                    u = Mfma<T>::op(a, b, u);
                }
                acc[q] -= u;
            }
        }
        __syncthreads();
    }
}

// Rows below an outer block: each wavefront owns kTrsmRows = 16 rows and
// keeps their L entries of the whole block in registers; a workgroup of four
// shares the staging of each sub-step's operands in LDS.  Each product is
// computed transposed, D(c, r) = sum_k X(c, k) Y(k, r) with the rows r in
// the MFMA column index, so a lane's accumulator elements are L(r, c) for
// the k-slots it supplies as an MFMA B operand: accumulator element e of lane
// l holds column c = Mfma::row(l, e) of a 16-column half, and the same
// element feeds the k index Mfma::row(l, e) of a later product.  The A
// operands (d_k L(jb, kb)^T and Bm, staged in LDS) follow that k assignment.
// The last sub-step of a block may be narrower than kPanelNB: its Bm rows
// k >= nb are staged as zeros and addresses are clamped into the block.
// With PUB, the operands come from the descriptor's publication area as
// the diagonal workgroup stores them (all loads of a sub-step issued, then
// each awaited).
template <typename T, bool PUB>
__device__ __forceinline__ void block_trsm_body(const PanelDesc &pd, int sb, const NodeInfo &n,
                                                T *__restrict__ pool,
                                                const double *__restrict__ dvec,
                                                const T *__restrict__ bmscr, const T *pub,
                                                FactorStatus *st, T *__restrict__ store,
                                                BtLds<T> &S)
{
    constexpr int NB = kPanelNB, NJB = kPanelOuter / kPanelNB, LR = bt_ld<T>;
    constexpr int NT = 64 * kBtWaves;
    static_assert(NB == 32 && kTrsmRows == 16, "two 16-column halves per sub-step");
    static_assert(NB * NB == 4 * NT, "four Bm entries per thread");
    typedef typename Mfma<T>::acc_t acc_t;
    T *fr = pool + n.foff;
    const i64 ld = n.ld;
    const int tid = threadIdx.x, lane = tid & 63, c = lane & 15;
    const int bend = pd.k0 + pd.nb;
    const int njb = (pd.nb + NB - 1) / NB;
    const int row = bend + sb * kTrsmRows + c;  // row r = lane & 15 of the stripe
    const bool rowok = row < n.F;
    const T *frow = fr + (rowok ? row : n.F - 1);
    bool dead = false;
    int kk[4];  // k slot (within a 16-wide half) this lane supplies at step e
#pragma unroll
    for (int e = 0; e < 4; ++e) kk[e] = Mfma<T>::row(lane, e);
    acc_t O[NJB][2];  // O[jb][h][e] = L(row, ob + 32 jb + 16 h + kk[e])
#pragma unroll
    for (int jb = 0; jb < NJB; ++jb) {
        if (jb >= njb) break;
        const int k0 = pd.k0 + jb * NB;
        const int nb = min(NB, bend - k0);
        // this wavefront's rows of the sub-step (issued before the staging)
        acc_t acc[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[h][e] = frow[i64(k0 + min(16 * h + kk[e], nb - 1)) * ld];
        // stage d_k L(k0 + cc, k) for k in [ob, k0) and Bm (rows k >= nb zero)
        __syncthreads();  // the previous sub-step's reads of Rs/Bs are done
        if (PUB) {
            // x = tid + NT u: Rs entry (kx, cc) = (x / NB, x % NB), Bm entry
            // (k, cc) likewise
            T rv[4 * (NJB - 1)], bv[4];
#pragma unroll
            for (int u = 0; u < 4 * (NJB - 1); ++u) {
                if (u >= 4 * jb) break;
                const int x = tid + NT * u, cc = x % NB, kx = x / NB;
                rv[u] = pub_load(pub + kPubBm + (k0 - pd.k0 + min(cc, nb - 1) - NB) * kPubRows + kx);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {  // (entries k, cc >= nb are not made)
                const int x = tid + NT * u, cc = x % NB, k = x / NB;
                bv[u] = pub_load(pub + jb * NB * NB + min(k, nb - 1) * NB + min(cc, nb - 1));
            }
#pragma unroll
            for (int u = 0; u < 4 * (NJB - 1); ++u) {
                if (u >= 4 * jb) break;
                const int x = tid + NT * u, cc = x % NB, kx = x / NB;
                S.Rs[kx * LR + cc] = pub_take(
                    pub + kPubBm + (k0 - pd.k0 + min(cc, nb - 1) - NB) * kPubRows + kx, rv[u], st,
                    dead);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int x = tid + NT * u, cc = x % NB, k = x / NB;
                const T v = pub_take(pub + jb * NB * NB + min(k, nb - 1) * NB + min(cc, nb - 1),
                                     bv[u], st, dead);
                S.Bs[k * LR + cc] = (k < nb && cc < nb) ? v : T(0);
            }
        } else {
            for (int x = tid; x < NB * jb * NB; x += NT) {
                const int cc = x % NB, kx = x / NB;
                const int kcol = pd.k0 + kx;
                S.Rs[kx * LR + cc] =
                    T(dvec[n.first + kcol]) * fr[(k0 + min(cc, nb - 1)) + i64(kcol) * ld];
            }
            for (int x = tid; x < NB * NB; x += NT) {
                const int cc = x % NB, k = x / NB;
                S.Bs[k * LR + cc] = (k < nb) ? bmscr[i64(n.first + k0 + k) * NB + cc] : T(0);
            }
        }
        __syncthreads();
        // acc(c, r) -= sum_k d_k L(k0 + c, kb cols) L(r, kb cols) over kb < jb,
        // the sum formed apart and subtracted once (for accuracy in FP32)
        acc_t u[2] = {acc_t{0, 0, 0, 0}, acc_t{0, 0, 0, 0}};
#pragma unroll
        for (int kb = 0; kb < NJB - 1; ++kb) {
            if (kb >= jb) break;
#pragma unroll
            for (int hk = 0; hk < 2; ++hk)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const T *rk = S.Rs + (kb * NB + 16 * hk + kk[e]) * LR;
#pragma unroll
                    for (int h = 0; h < 2; ++h)
                        u[h] = Mfma<T>::op(rk[16 * h + c], O[kb][hk][e], u[h]);
                }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) acc[h] -= u[h];
        // O(c, r) = sum_k Bm(k, c) acc(k, r)
#pragma unroll
        for (int h = 0; h < 2; ++h) O[jb][h] = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int hk = 0; hk < 2; ++hk)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const T *bk = S.Bs + (16 * hk + kk[e]) * LR;
#pragma unroll
                for (int h = 0; h < 2; ++h) O[jb][h] = Mfma<T>::op(bk[16 * h + c], acc[hk][e], O[jb][h]);
            }
        if (rowok) {
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int col = 16 * h + kk[e];
                    if (col < nb) {
                        fr[row + i64(k0 + col) * ld] = O[jb][h][e];
                        store[n.loff + row + i64(k0 + col) * ld] = O[jb][h][e];
                    }
                }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(64 * kBdWaves) void k_block_diag(const PanelDesc *__restrict__ descs,
                                                             const NodeInfo *__restrict__ nodes,
                                                             T *__restrict__ pool, double *dvec,
                                                             const double *__restrict__ adiag,
                                                             double *dcor, FactorStatus *st,
                                                             PivotArgs pa, T *bmscr, T *store)
{
    __shared__ BdLds<T> S;
    const PanelDesc pd = descs[blockIdx.x];
    const NodeInfo n = nodes[pd.node];
    block_diag_body<T, false>(pd, n, pool, dvec, adiag, dcor, st, pa, bmscr, nullptr, store, S);
}

template <typename T>
__global__ __launch_bounds__(64 * kBtWaves) void k_block_trsm(const PanelDesc *__restrict__ descs,
                                                             int ndesc,
                                                             const i64 *__restrict__ prefix,
                                                             const NodeInfo *__restrict__ nodes,
                                                             T *__restrict__ pool,
                                                             const double *__restrict__ dvec,
                                                             const T *__restrict__ bmscr,
                                                             T *__restrict__ store)
{
    __shared__ BtLds<T> S;
    const int q = find_desc(prefix, ndesc, blockIdx.x);
    const PanelDesc pd = descs[q];
    const int sb = int(blockIdx.x - prefix[q]) * kBtWaves + (threadIdx.x >> 6);
    const NodeInfo n = nodes[pd.node];
    block_trsm_body<T, false>(pd, sb, n, pool, dvec, bmscr, nullptr, nullptr, store, S);
}

// Workgroups [0, ndesc): the diagonal blocks; then the rows below them
// (prefix as k_block_trsm's).  pub: the publication areas of the launch's
// descriptors, kPubElems each, filled with pub_pending.
template <typename T>
__global__ __launch_bounds__(64 * kBdWaves) void k_block_fused(const PanelDesc *__restrict__ descs,
                                                              int ndesc,
                                                              const i64 *__restrict__ prefix,
                                                              const NodeInfo *__restrict__ nodes,
                                                              T *__restrict__ pool, double *dvec,
                                                              const double *__restrict__ adiag,
                                                              double *dcor, FactorStatus *st,
                                                              PivotArgs pa, T *pub, T *store)
{
    constexpr size_t kBytes = sizeof(BdLds<T>) > sizeof(BtLds<T>) ? sizeof(BdLds<T>) : sizeof(BtLds<T>);
    __shared__ __attribute__((aligned(16))) char lds[kBytes];
    if (int(blockIdx.x) < ndesc) {
        const PanelDesc pd = descs[blockIdx.x];
        const NodeInfo n = nodes[pd.node];
        block_diag_body<T, true>(pd, n, pool, dvec, adiag, dcor, st, pa, nullptr,
                           pub + i64(blockIdx.x) * kPubElems, store,
                           *reinterpret_cast<BdLds<T> *>(lds));
    } else {
        const i64 b = i64(blockIdx.x) - ndesc;
        const int q = find_desc(prefix, ndesc, b);
        const PanelDesc pd = descs[q];
        const int sb = int(b - prefix[q]) * kBtWaves + (threadIdx.x >> 6);
        const NodeInfo n = nodes[pd.node];
        block_trsm_body<T, true>(pd, sb, n, pool, dvec, nullptr, pub + i64(q) * kPubElems, st,
                                 store, *reinterpret_cast<BtLds<T> *>(lds));
    }
}

// ---------------------------------------------------------------------------
// MFMA update: C(i,j) -= sum_k L(i,k) d_k L(j,k) on lower tiles
// ---------------------------------------------------------------------------
// One workgroup computes a BM x BM tile of C; its four wavefronts own 2x2
// quadrants of WM = BM/2 rows and columns, each a grid of TM x TM MFMAs of
// 16x16x4.  Inside a quadrant, MFMA sub-tile ti covers the rows TM*m + ti
// (m = 0..15), so the TM operand values a lane needs at one k are adjacent
// in LDS (one 16-byte read for TM*sizeof(T) = 16) and the TM accumulator
// values it holds for one column are adjacent rows of C (one vector load and
// store).  The update is accumulated apart from C (for accuracy in FP32) and
// subtracted in the epilogue; B is scaled by the pivots as it is staged.
// The K loop moves KB columns of L per stage with 16-byte global
// loads: the next stage's loads are issued into registers before the current
// stage's MFMAs and stored into the other LDS buffer after them, so each
// stage needs a single barrier.  Consecutive workgroups are dispatched
// round-robin over the 8 XCDs; the block -> tile map gives each XCD a
// contiguous range of tiles in gemm_tile's order, so tiles that share L
// blocks share an L2.
// Tiles are anchored at r0 rounded down to kGemmAlign and front leading
// dimensions are multiples of kFrontLdAlign, so all vectors are aligned.
constexpr int kSuperCols = 8;  // tile columns per super-column of the tile order

// Tile (ti, tj) of the t-th tile of a lower trapezoid of nti x ntj tiles
// (ti >= tj).  Tiles are ordered by super-columns of kSuperCols tile columns
// and row by row inside one, so that consecutive tiles (which run together
// on one XCD) share a row block of L and cycle over few column blocks: the
// operand working set of an XCD stays within its L2.
__device__ __forceinline__ void gemm_tile(i64 t, int nti, int ntj, int &ti, int &tj)
{
    constexpr int G = kSuperCols;
    int c0 = 0;
    for (;;) {
        const int w = min(G, ntj - c0);
        const i64 tri = i64(w) * (w + 1) / 2;
        const i64 cnt = tri + i64(nti - c0 - w) * w;
        if (t < cnt || c0 + w >= ntj) {
            if (t < tri) {  // the triangle at the top: row k holds k + 1 tiles
                int k = int((sqrt(8.0 * double(t) + 1.0) - 1.0) * 0.5);
                while (i64(k) * (k + 1) / 2 > t) --k;
                while (i64(k + 1) * (k + 2) / 2 <= t) ++k;
                ti = c0 + k;
                tj = c0 + int(t - i64(k) * (k + 1) / 2);
            } else {
                const i64 u = t - tri;
                ti = c0 + w + int(u / w);
                tj = c0 + int(u % w);
            }
            return;
        }
        t -= cnt;
        c0 += w;
    }
}

template <typename T> struct GemmCfg;
template <> struct GemmCfg<float> { static constexpr int KB = 32; };
template <> struct GemmCfg<double> { static constexpr int KB = 16; };

// KIND does not change the code: 1 marks the rank-kOuterRank and Schur
// updates, 0 the narrower updates inside an outer group, so that kernel
// traces and counters report the two uses apart.
template <typename T, int BM, int KIND>
__global__ __launch_bounds__(256) void k_gemm(const GemmDesc *__restrict__ descs, int ndesc,
                                              const i64 *__restrict__ prefix,
                                              T *__restrict__ pool,
                                              const double *__restrict__ dvec)
{
    constexpr int KB = GemmCfg<T>::KB;
    constexpr int VE = 16 / sizeof(T);     // elements per 16-byte vector
    constexpr int WM = BM / 2;             // quadrant edge
    constexpr int TM = WM / 16;            // MFMA sub-tiles per quadrant edge
    constexpr int LDR = BM + (VE > TM ? VE : TM);  // padded LDS row (one k), TM-vector aligned
    constexpr int NVC = BM / VE;           // vectors per column of a stage
    constexpr int NV = NVC * KB / 256;     // vectors per thread per operand per stage
    static_assert(NV >= 1 && NVC * KB % 256 == 0, "stage must split evenly over 256 threads");
    static_assert(TM % VE == 0 || VE % TM == 0, "operand vectors");
    static_assert(kGemmAlign % VE == 0 && kFrontLdAlign % VE == 0, "vector alignment");
    typedef T vec_t __attribute__((ext_vector_type(VE)));
    typedef T tm_t __attribute__((ext_vector_type(TM)));
    __shared__ __attribute__((aligned(16))) T As[2][KB][LDR];
    __shared__ __attribute__((aligned(16))) T Bs[2][KB][LDR];

    // XCD-aware bijection of block ids onto tiles
    const i64 nb = gridDim.x, bid = blockIdx.x;
    const i64 per = nb / kXcds, rem = nb % kXcds;
    const i64 xcd = bid % kXcds, idx = bid / kXcds;
    const i64 lin = (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;

    const int q = find_desc(prefix, ndesc, lin);
    const GemmDesc g = descs[q];
    int ti, tj;
    gemm_tile(lin - prefix[q], g.nti, g.ntj, ti, tj);
    const int a0 = g.r0 / kGemmAlign * kGemmAlign;
    const int i0 = a0 + ti * BM, j0 = a0 + tj * BM;
    T *fr = pool + g.foff;
    const i64 ld = g.ld;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int kend = g.k0 + g.kn;
    // first row / column of this lane's accumulator vectors
    const int qi0 = i0 + wr * WM, qj0 = j0 + wc * WM;

    // acc[ti][tj][e] accumulates the update of
    // C(qi0 + TM*Mfma::row(lane,e) + ti, qj0 + TM*(lane&15) + tj)
    typename Mfma<T>::acc_t acc[TM][TM];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TM; ++b) acc[a][b] = typename Mfma<T>::acc_t{0, 0, 0, 0};

    // Stage loads are unconditional (an out-of-range vector reads the front's
    // first element instead) so that they all stay in flight across the
    // MFMAs; masking and the pivot scaling happen when the stage is stored.
    vec_t ra[NV], rb[NV];
    double dk[NV];
    unsigned vmask = 0;  // bit 2m: A vector m valid, bit 2m+1: B vector m valid
    auto load_stage = [&](int kk) {
        vmask = 0;
#pragma unroll
        for (int m = 0; m < NV; ++m) {
            const int v = tid + 256 * m;
            const int rv = (v % NVC) * VE, kc = v / NVC;
            const int col = kk + kc;
            const bool cv = col < kend;
            const int arow = i0 + rv, brow = j0 + rv;
            const bool va = cv && arow < g.rend, vb = cv && brow < g.cend;
            vmask |= (va ? 1u : 0u) << (2 * m);
            vmask |= (vb ? 1u : 0u) << (2 * m + 1);
            const i64 cofs = i64(cv ? col : g.k0) * ld;
            ra[m] = *reinterpret_cast<const vec_t *>(fr + (va ? arow + cofs : 0));
            rb[m] = *reinterpret_cast<const vec_t *>(fr + (vb ? brow + cofs : 0));
            dk[m] = dvec[g.first + (cv ? col : g.k0)];
        }
    };
    auto store_stage = [&](int buf) {
#pragma unroll
        for (int m = 0; m < NV; ++m) {
            const int v = tid + 256 * m;
            const int rv = (v % NVC) * VE, kc = v / NVC;
            const vec_t za = (vmask >> (2 * m)) & 1u ? ra[m] : vec_t{};
            const vec_t zb = (vmask >> (2 * m + 1)) & 1u ? rb[m] * T(dk[m]) : vec_t{};
            *reinterpret_cast<vec_t *>(&As[buf][kc][rv]) = za;
            *reinterpret_cast<vec_t *>(&Bs[buf][kc][rv]) = zb;
        }
    };

    load_stage(g.k0);
    store_stage(0);
    __syncthreads();
    int cur = 0;
    const int ar = wr * WM + TM * (lane & 15), bc = wc * WM + TM * (lane & 15);
    for (int kk = g.k0; kk < kend; kk += KB) {
        const bool more = kk + KB < kend;
        if (more) load_stage(kk + KB);
#pragma unroll
        for (int k4 = 0; k4 < KB; k4 += 4) {
            const int kidx = k4 + (lane >> 4);
            const tm_t av = *reinterpret_cast<const tm_t *>(&As[cur][kidx][ar]);
            const tm_t bv = *reinterpret_cast<const tm_t *>(&Bs[cur][kidx][bc]);
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TM; ++b) acc[a][b] = Mfma<T>::op(av[a], bv[b], acc[a][b]);
        }
        if (more) store_stage(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
    // write back the lower-trapezoid entries of rows [r0,rend) x cols [r0,cend)
#pragma unroll
    for (int b = 0; b < TM; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int row = qi0 + TM * Mfma<T>::row(lane, e);
            const int col = qj0 + TM * (lane & 15) + b;
            if (col < g.r0 || col >= g.cend || row >= g.rend || row + TM - 1 < col) continue;
            if (row >= g.r0 && row >= col && row + TM <= g.rend) {
                tm_t *pc = reinterpret_cast<tm_t *>(fr + row + col * ld);
                tm_t v = *pc;
#pragma unroll
                for (int a = 0; a < TM; ++a) v[a] -= acc[a][b][e];
                *pc = v;
            } else {
#pragma unroll
                for (int a = 0; a < TM; ++a) {
                    const int r = row + a;
                    if (r >= g.r0 && r >= col && r < g.rend) fr[r + col * ld] -= acc[a][b][e];
                }
            }
        }
}

// FP32 update on 128 x 128 tiles with the 32x32x2 MFMA.  Same tile order,
// staging, masking and pivot scaling as k_gemm; each wavefront owns a 64 x 64
// quadrant as 2 x 2 blocks of 32x32x2 MFMAs, whose operands are one float per
// lane per block (lanes 0-31 and 32-63 take k and k + 1), so the LDS reads are
// plain 4-byte reads of consecutive rows.  It moves twice the multiply-adds
// per operand read of the 16x16x4 form and runs about 10 % faster on large
// launches; accumulator element r of block (a, b) is row
// 32a + 8(r/4) + 4(lane/32) + r%4 and column 32b + lane%32 of the quadrant,
// so four consecutive elements are four consecutive rows of C.
// Launch bounds of four workgroups per CU hold it to 128 registers (the
// accumulators in VGPRs, no spills): four waves per SIMD instead of three,
// which the LDS (40 KB per workgroup) allows; about 3 % faster per launch.
typedef float fv16 __attribute__((ext_vector_type(16)));

template <int KIND>
__global__ __launch_bounds__(256, 4) void k_gemm_x32(const GemmDesc *__restrict__ descs, int ndesc,
                                                  const i64 *__restrict__ prefix,
                                                  float *__restrict__ pool,
                                                  const double *__restrict__ dvec)
{
    constexpr int BM = kTileBig, KB = 16, VE = 4;
    constexpr int WM = BM / 2;           // quadrant edge
    constexpr int MB = WM / 32;          // 32 x 32 blocks per quadrant edge
    constexpr int LDR = BM + 32;         // padded LDS row (one k)
    constexpr int NVC = BM / VE;         // vectors per column of a stage
    constexpr int NV = NVC * KB / 256;   // vectors per thread per operand per stage
    static_assert(NV >= 1 && NVC * KB % 256 == 0, "stage must split evenly over 256 threads");
    static_assert(kGemmAlign % VE == 0 && kFrontLdAlign % VE == 0, "vector alignment");
    __shared__ __attribute__((aligned(16))) float As[2][KB][LDR];
    __shared__ __attribute__((aligned(16))) float Bs[2][KB][LDR];

    const i64 nb = gridDim.x, bid = blockIdx.x;
    const i64 per = nb / kXcds, rem = nb % kXcds;
    const i64 xcd = bid % kXcds, idx = bid / kXcds;
    const i64 lin = (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;

    const int q = find_desc(prefix, ndesc, lin);
    const GemmDesc g = descs[q];
    int ti, tj;
    gemm_tile(lin - prefix[q], g.nti, g.ntj, ti, tj);
    const int a0 = g.r0 / kGemmAlign * kGemmAlign;
    const int i0 = a0 + ti * BM, j0 = a0 + tj * BM;
    float *fr = pool + g.foff;
    const i64 ld = g.ld;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int kend = g.k0 + g.kn;

    fv16 acc[MB][MB];
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
        for (int b = 0; b < MB; ++b) acc[a][b] = fv16{};

    fv4 ra[NV], rb[NV];
    double dk[NV];
    unsigned vmask = 0;  // bit 2m: A vector m valid, bit 2m+1: B vector m valid
    auto load_stage = [&](int kk) {
        vmask = 0;
#pragma unroll
        for (int m = 0; m < NV; ++m) {
            const int v = tid + 256 * m;
            const int rv = (v % NVC) * VE, kc = v / NVC;
            const int col = kk + kc;
            const bool cv = col < kend;
            const int arow = i0 + rv, brow = j0 + rv;
            const bool va = cv && arow < g.rend, vb = cv && brow < g.cend;
            vmask |= (va ? 1u : 0u) << (2 * m);
            vmask |= (vb ? 1u : 0u) << (2 * m + 1);
            const i64 cofs = i64(cv ? col : g.k0) * ld;
            ra[m] = *reinterpret_cast<const fv4 *>(fr + (va ? arow + cofs : 0));
            rb[m] = *reinterpret_cast<const fv4 *>(fr + (vb ? brow + cofs : 0));
            dk[m] = dvec[g.first + (cv ? col : g.k0)];
        }
    };
    auto store_stage = [&](int buf) {
#pragma unroll
        for (int m = 0; m < NV; ++m) {
            const int v = tid + 256 * m;
            const int rv = (v % NVC) * VE, kc = v / NVC;
            const fv4 za = (vmask >> (2 * m)) & 1u ? ra[m] : fv4{};
            const fv4 zb = (vmask >> (2 * m + 1)) & 1u ? rb[m] * float(dk[m]) : fv4{};
            *reinterpret_cast<fv4 *>(&As[buf][kc][rv]) = za;
            *reinterpret_cast<fv4 *>(&Bs[buf][kc][rv]) = zb;
        }
    };

    load_stage(g.k0);
    store_stage(0);
    __syncthreads();
    int cur = 0;
    const int l32 = lane & 31, kh = lane >> 5;
    for (int kk = g.k0; kk < kend; kk += KB) {
        const bool more = kk + KB < kend;
        if (more) load_stage(kk + KB);
#pragma unroll
        for (int k2 = 0; k2 < KB; k2 += 2) {
            float av[MB], bv[MB];
#pragma unroll
            for (int a = 0; a < MB; ++a) av[a] = As[cur][k2 + kh][wr * WM + 32 * a + l32];
#pragma unroll
            for (int b = 0; b < MB; ++b) bv[b] = Bs[cur][k2 + kh][wc * WM + 32 * b + l32];
#pragma unroll
            for (int a = 0; a < MB; ++a)
#pragma unroll
                for (int b = 0; b < MB; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        if (more) store_stage(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
    // write back the lower-trapezoid entries of rows [r0,rend) x cols [r0,cend)
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
        for (int b = 0; b < MB; ++b)
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) {
                const int row = i0 + wr * WM + 32 * a + 8 * qq + 4 * kh;
                const int col = j0 + wc * WM + 32 * b + l32;
                if (col < g.r0 || col >= g.cend || row >= g.rend || row + 3 < col) continue;
                if (row >= g.r0 && row >= col && row + 4 <= g.rend) {
                    fv4 *pc = reinterpret_cast<fv4 *>(fr + row + i64(col) * ld);
                    fv4 v = *pc;
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] -= acc[a][b][4 * qq + e];
                    *pc = v;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int r = row + e;
                        if (r >= g.r0 && r >= col && r < g.rend)
                            fr[r + i64(col) * ld] -= acc[a][b][4 * qq + e];
                    }
                }
            }
}

// ---------------------------------------------------------------------------
// statistics: NUMNEG, TOL(2), TOL(3) (sprfc1.f:686-708, :1281-1285)
// ---------------------------------------------------------------------------
// stats layout: [0] numneg, [1] max a/d, [2] sum a, [3] min d>0
constexpr int kStatBlocks = 256;

__global__ __launch_bounds__(256) void k_stats_partial(const double *dvec, const double *adiag,
                                                       i64 neq, const FactorStatus *st,
                                                       double *part)
{
    const unsigned long long lim = st->first_fail;
    double nneg = 0, mx = 0, sa = 0, mn = DBL_MAX;
    for (i64 i = blockIdx.x * 256 + threadIdx.x; i < neq; i += i64(gridDim.x) * 256) {
        const double d = dvec[i], a = adiag[i];
        sa += a;
        if (i64(i) < i64(lim < (unsigned long long)neq ? lim : neq)) {
            if (d < 0) nneg += 1;
            const double r = a / d;
            if (r > mx) mx = r;
        }
        if (d > 0 && d < mn) mn = d;
    }
    __shared__ double s[4][256];
    s[0][threadIdx.x] = nneg;
    s[1][threadIdx.x] = mx;
    s[2][threadIdx.x] = sa;
    s[3][threadIdx.x] = mn;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s[0][threadIdx.x] += s[0][threadIdx.x + o];
            s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + o]);
            s[2][threadIdx.x] += s[2][threadIdx.x + o];
            s[3][threadIdx.x] = fmin(s[3][threadIdx.x], s[3][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = s[k][0];
}

__global__ void k_stats_final(double *part, int nparts)
{
    if (threadIdx.x == 0) {
        double nneg = 0, mx = 0, sa = 0, mn = DBL_MAX;
        for (int b = 0; b < nparts; ++b) {
            nneg += part[4 * b];
            mx = fmax(mx, part[4 * b + 1]);
            sa += part[4 * b + 2];
            mn = fmin(mn, part[4 * b + 3]);
        }
        part[4 * nparts + 0] = nneg;
        part[4 * nparts + 1] = mx;
        part[4 * nparts + 2] = sa;
        part[4 * nparts + 3] = mn;
    }
}

// ---------------------------------------------------------------------------
// copy the factor into the packed SM layout
// ---------------------------------------------------------------------------
// A level's factored panels (rows c..F-1 of the columns c < S) from the
// fronts to the factor store, where the solves read them after the fronts'
// memory is reused.
template <typename T>
__global__ __launch_bounds__(256) void k_store_panel(const ColTask *__restrict__ tasks,
                                                     const NodeInfo *__restrict__ nodes,
                                                     const T *__restrict__ pool,
                                                     T *__restrict__ store)
{
    const ColTask t = tasks[blockIdx.x];
    const NodeInfo n = nodes[t.node];
    for (int c = t.c0; c < t.c1; ++c) {
        const T *src = pool + n.foff + i64(c) * n.ld;
        T *dst = store + n.loff + i64(c) * n.ld;
        int r = c + threadIdx.x;
        for (; r + 3 * 256 < n.F; r += 4 * 256) {
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = src[r + 256 * u];
#pragma unroll
            for (int u = 0; u < 4; ++u) dst[r + 256 * u] = v[u];
        }
        for (; r < n.F; r += 256) dst[r] = src[r];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_out(const ColTask *__restrict__ tasks,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const T *__restrict__ store, double *sm,
                                                  i64 neq, i32 node_limit)
{
    const ColTask t = tasks[blockIdx.x];
    if (t.node >= node_limit) return;
    const NodeInfo n = nodes[t.node];
    double *lnz = sm + neq;
    for (int c = t.c0; c < t.c1 && c < n.S; ++c) {
        double *dst = lnz + n.xlnz + i64(c) * n.F - (i64(c) * (c - 1)) / 2;
        const T *col = store + n.loff + i64(c) * n.ld;
        for (int r = c + threadIdx.x; r < n.F; r += blockDim.x) dst[r - c] = double(col[r]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_load_factor(const ColTask *__restrict__ tasks,
                                                     const NodeInfo *__restrict__ nodes,
                                                     T *__restrict__ store,
                                                     const double *__restrict__ sm, i64 neq)
{
    const ColTask t = tasks[blockIdx.x];
    const NodeInfo n = nodes[t.node];
    const double *lnz = sm + neq;
    for (int c = t.c0; c < t.c1 && c < n.S; ++c) {
        const double *src = lnz + n.xlnz + i64(c) * n.F - (i64(c) * (c - 1)) / 2;
        T *col = store + n.loff + i64(c) * n.ld;
        for (int r = c + threadIdx.x; r < n.F; r += blockDim.x) col[r] = T(src[r - c]);
    }
}

__global__ void k_copy_diag(const double *dvec, const double *adiag, double *sm, i64 neq,
                            i64 limit)
{
    for (i64 i = blockIdx.x * i64(blockDim.x) + threadIdx.x; i < neq;
         i += i64(gridDim.x) * blockDim.x)
        sm[i] = (i < limit) ? dvec[i] : adiag[i];
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
// Kernel classes for the per-class timing of sprgpu_profile_factor
// (P_PDIAG: diagonal blocks, P_PROWS: rows below them).
enum ProfClass {
    P_INIT = 0, P_EA, P_SMALL, P_PDIAG, P_PROWS, P_GEMM, P_OTHER, P_GEMM_INNER, P_GEMM_BIG, P_NCLASS
};

// Launches the whole factorization.  With ev != nullptr an event is recorded
// after every kernel and the kernel's class appended to cls, so consecutive
// events bracket one kernel each (the stream serialises them).
// With b set (a plan on the forest schedule, one right-hand side) the
// launch also solves for b.
template <typename T>
static int factor_impl(sprgpu_handle *h, const double *sm_dev, int mode, const double *tol,
                       std::vector<hipEvent_t> *ev = nullptr, std::vector<int> *cls = nullptr,
                       std::vector<double> *fl = nullptr, double *b = nullptr)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    hipStream_t s = h->stream;
    T *pool = static_cast<T *>(d.pool);
    const double *lnz = sm_dev + p.neq;
    PivotArgs pa{tol[0], tol[3], mode};
    auto mark = [&](int c, double flops) {
        if (!ev) return;
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s);
        ev->push_back(e);
        cls->push_back(c);
        fl->push_back(flops);
    };
    mark(P_OTHER, 0.0);
    k_status_reset<<<1, 64, 0, s>>>(d.status);
    mark(P_OTHER, 0.0);
    if (p.forest) {  // the whole factorization (and the solve) in one persistent launch
        if (int rc = forest_factor(h, sm_dev, mode, tol, b)) return rc;
        mark(P_SMALL, 0.0);
        k_stats_partial<<<kStatBlocks, 256, 0, s>>>(d.dvec, d.adiag, p.neq, d.status, d.stats);
        k_stats_final<<<1, 64, 0, s>>>(d.stats, kStatBlocks);
        SPRGPU_HIP(hipGetLastError());
        return 0;
    }
    if (b) {
        set_error("factor and solve in one launch needs the forest schedule");
        return -2;
    }
    if (d.pub_elems > 0) {  // the fused outer-block launches' publication areas
        k_pub_mark<T><<<unsigned(std::min<i64>((d.pub_elems + 255) / 256, 2048)), 256, 0, s>>>(
            static_cast<T *>(d.pub), d.pub_elems);
        mark(P_OTHER, 0.0);
    }
    // The plan's far updates go to a second stream, ordered by events
    // (L_RECORD, L_WAIT); the profiling pass runs everything on one stream.
    const bool two = (ev == nullptr);
    if (two) {
        SPRGPU_HIP(hipEventRecord(h->ev[0], s));  // fork
        SPRGPU_HIP(hipStreamWaitEvent(h->stream2, h->ev[0], 0));
        SPRGPU_HIP(hipEventRecord(h->ev[1], h->stream2));
    }
    for (const Launch &L : p.launches) {
        const unsigned blocks = unsigned(L.blocks);
        const hipStream_t ls = (two && L.stream) ? h->stream2 : s;
        switch (L.kind) {
        case L_RECORD:
            if (two) SPRGPU_HIP(hipEventRecord(h->ev[L.stream], ls));
            break;
        case L_WAIT:
            if (two) SPRGPU_HIP(hipStreamWaitEvent(ls, h->ev[1 - L.stream], 0));
            break;
        case L_INIT:
            k_front_init<T><<<blocks, 256, 0, ls>>>(d.init_tasks + L.off, d.nodes, lnz, pool,
                                                   d.adiag, d.dcor, d.status, mode);
            mark(P_INIT, 0.0);
            break;
        case L_EA:
            k_extend_add<T><<<blocks, 256, 0, ls>>>(d.ea_tasks + L.off, d.nodes, d.relmap, pool);
            mark(P_EA, 0.0);
            break;
        case L_SMALL: {
            // L.tile holds the launch's largest front
            const int mf = L.tile, waves = p.params.small_waves;
            const size_t per = small_front_bytes(mf, sizeof(T));
            const unsigned nb = unsigned((L.count + waves - 1) / waves);
            k_small<T><<<nb, 64 * waves, per * waves, ls>>>(
                d.small_nodes + L.off, int(L.count), mf, d.nodes, d.children, d.relmap, lnz, pool,
                static_cast<T *>(d.store), d.dvec, d.adiag, d.dcor, d.status, pa);
            mark(P_SMALL, 0.0);
            break;
        }
        case L_PANEL:
            k_panel<T><<<blocks, 256, 0, ls>>>(d.panel_desc + L.off, int(L.count),
                                              d.panel_prefix + L.off, d.nodes, pool, d.dvec,
                                              d.adiag, d.dcor, d.status, pa,
                                              static_cast<T *>(d.dscr));
            mark(P_PROWS, 0.0);
            break;
        case L_BDIAG:
            k_block_diag<T><<<blocks, 64 * kBdWaves, 0, ls>>>(d.panel_desc + L.off, d.nodes, pool,
                                                             d.dvec,
                                                   d.adiag, d.dcor, d.status, pa,
                                                   static_cast<T *>(d.dscr), static_cast<T *>(d.store));
            mark(P_PDIAG, 0.0);
            break;
        case L_BTRSM:
            k_block_trsm<T><<<blocks, 64 * kBtWaves, 0, ls>>>(d.panel_desc + L.off, int(L.count),
                                                  d.panel_prefix + L.off, d.nodes, pool, d.dvec,
                                                  static_cast<const T *>(d.dscr),
                                                  static_cast<T *>(d.store));
            mark(P_PROWS, 0.0);
            break;
        case L_STORE:
            k_store_panel<T><<<blocks, 256, 0, ls>>>(d.store_tasks + L.off, d.nodes, pool,
                                                     static_cast<T *>(d.store));
            mark(P_OTHER, 0.0);
            break;
        case L_BFUSED:
            // publication area q of the launch: the descriptor's index in the plan
            k_block_fused<T><<<blocks, 64 * kBdWaves, 0, ls>>>(
                d.panel_desc + L.off, int(L.count), d.panel_prefix + L.off, d.nodes, pool, d.dvec,
                d.adiag, d.dcor, d.status, pa, static_cast<T *>(d.pub) + L.off * kPubElems,
                static_cast<T *>(d.store));
            mark(P_PDIAG, 0.0);
            break;
        case L_GEMM:
            if (L.tile == kTileBig) {
                if constexpr (std::is_same<T, float>::value)
                    k_gemm_x32<1><<<blocks, 256, 0, ls>>>(d.gemm_desc + L.off, int(L.count),
                                                          d.gemm_prefix + L.off, pool, d.dvec);
                else
                    k_gemm<T, kTileBig, 1><<<blocks, 256, 0, ls>>>(
                        d.gemm_desc + L.off, int(L.count), d.gemm_prefix + L.off, pool, d.dvec);
            }
            else if (L.outer)
                k_gemm<T, kTileSmall, 1><<<blocks, 256, 0, ls>>>(
                    d.gemm_desc + L.off, int(L.count), d.gemm_prefix + L.off, pool, d.dvec);
            else
                k_gemm<T, kTileSmall, 0><<<blocks, 256, 0, ls>>>(
                    d.gemm_desc + L.off, int(L.count), d.gemm_prefix + L.off, pool, d.dvec);
            mark(!L.outer ? P_GEMM_INNER : L.tile == kTileBig ? P_GEMM_BIG : P_GEMM, L.flops);
            break;
        }
    }
    if (two) {  // join
        SPRGPU_HIP(hipEventRecord(h->ev[1], h->stream2));
        SPRGPU_HIP(hipStreamWaitEvent(s, h->ev[1], 0));
    }
    if (!p.params.block_factor && !p.panel_desc.empty()) {
        k_diag_writeback<T><<<unsigned(p.panel_desc.size()), 256, 0, s>>>(
            d.panel_desc, d.nodes, static_cast<T *>(d.store), static_cast<const T *>(d.dscr));
        mark(P_PDIAG, 0.0);
    }
    if (int rc = compute_block_inverses(h)) return rc;
    mark(P_OTHER, 0.0);
    k_stats_partial<<<kStatBlocks, 256, 0, s>>>(d.dvec, d.adiag, p.neq, d.status, d.stats);
    k_stats_final<<<1, 64, 0, s>>>(d.stats, kStatBlocks);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int run_factor(sprgpu_handle *h, const double *sm_dev, int mode, const double *tol)
{
    h->mode = mode;
    for (int k = 0; k < 4; ++k) h->tol[k] = tol[k];
    h->last_sm_dev = sm_dev;
    int rc = (h->precision == SPRGPU_FP32) ? factor_impl<float>(h, sm_dev, mode, tol)
                                           : factor_impl<double>(h, sm_dev, mode, tol);
    if (rc == 0) h->factored = true;
    return rc;
}

int run_factor_solve(sprgpu_handle *h, const double *sm_dev, int mode, const double *tol,
                     double *b)
{
    if (!h->plan.forest) {
        set_error("factor and solve in one launch needs the forest schedule");
        return -2;
    }
    if (int rc = prepare_solve(h, 1)) return rc;
    h->mode = mode;
    for (int k = 0; k < 4; ++k) h->tol[k] = tol[k];
    h->last_sm_dev = sm_dev;
    int rc = (h->precision == SPRGPU_FP32)
                 ? factor_impl<float>(h, sm_dev, mode, tol, nullptr, nullptr, nullptr, b)
                 : factor_impl<double>(h, sm_dev, mode, tol, nullptr, nullptr, nullptr, b);
    if (rc == 0) h->factored = true;
    return rc;
}

int profile_factor(sprgpu_handle *h, const double *sm_dev, int mode, const double *tol,
                   double *out)
{
    std::vector<hipEvent_t> ev;
    std::vector<int> cls;
    std::vector<double> fl;
    h->mode = mode;
    h->last_sm_dev = sm_dev;
    int rc = (h->precision == SPRGPU_FP32) ? factor_impl<float>(h, sm_dev, mode, tol, &ev, &cls, &fl)
                                           : factor_impl<double>(h, sm_dev, mode, tol, &ev, &cls, &fl);
    if (rc == 0) h->factored = true;
    for (int k = 0; k < 3 * P_NCLASS; ++k) out[k] = 0.0;
    if (!ev.empty()) {
        SPRGPU_HIP(hipEventSynchronize(ev.back()));
        for (size_t i = 1; i < ev.size(); ++i) {
            float ms = 0.f;
            SPRGPU_HIP(hipEventElapsedTime(&ms, ev[i - 1], ev[i]));
            const int c = cls[i];
            out[3 * c + 0] += ms;
            out[3 * c + 1] += 1.0;
            out[3 * c + 2] += fl[i];
        }
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    return rc;
}

int read_factor_info(sprgpu_handle *h, sprgpu_factor_info *info)
{
    FactorStatus st;
    double agg[4];
    SPRGPU_HIP(hipMemcpyAsync(&st, h->d.status, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    SPRGPU_HIP(hipMemcpyAsync(agg, h->d.stats + 4 * kStatBlocks, sizeof(agg),
                              hipMemcpyDeviceToHost, h->stream));
    SPRGPU_HIP(hipStreamSynchronize(h->stream));
    if (st.wait_timeout) {
        set_error("a wait of the fused outer-block factorization timed out");
        return -5;
    }
    if (h->mode == 2) return cc_factor_info(h, info);
    const i64 neq = h->plan.neq;
    // modes 3 and 4 replace zero diagonals (no failure); modes 1, 2 and 4
    // fail on a pivot (4: only below -TOL(1)|a_jj|)
    const bool zero_fails = h->mode <= 2, pivot_fails = h->mode != 3;
    info->nzero_diag = i64(st.nzero);
    info->ierr = 0;
    info->numneg = 0;
    info->supeqt = 0;
    info->tol2 = 0;
    info->tol3 = 0;
    if (zero_fails && st.nzero > 0) {
        // zero diagonals before the factorization: sprfc1.f:463-488
        info->ierr = -3;
        info->numneg = i64(st.nzero);
        info->supeqt = i64(st.first_zero) + 1;
        info->tol2 = 0;
        info->tol3 = agg[2];
        return 0;
    }
    if (pivot_fails && st.first_fail != ULLONG_MAX) {
        // pivot reduced to zero: sprfc1.f:1274-1279, :604-611; SPRFC4's
        // negative pivot: sprfc4.f:1016-1025, :611-614 (MSPAR(26) stays 0)
        const i64 eq = i64(st.first_fail);
        double dfail = 0, afail = 0;
        SPRGPU_HIP(hipMemcpy(&dfail, h->d.dvec + eq, sizeof(double), hipMemcpyDeviceToHost));
        SPRGPU_HIP(hipMemcpy(&afail, h->d.adiag + eq, sizeof(double), hipMemcpyDeviceToHost));
        info->ierr = -3;
        info->numneg = h->mode == 4 ? 0 : i64(agg[0]);
        info->supeqt = eq + 1;
        info->tol2 = dfail;
        info->tol3 = afail;
        return 0;
    }
    info->numneg = i64(agg[0]);
    info->tol2 = (agg[1] != 0.0) ? 1.0 / agg[1] : 0.0;
    const double mdiag = (agg[3] == DBL_MAX) ? 0.0 : agg[3];
    info->tol3 = (mdiag != 0.0) ? agg[2] / mdiag : 0.0;
    if (h->mode == 4) info->numneg = 0;  // SPRFC4 counts no negative pivots (sprfc4.f:438)
    if (h->mode >= 3 && st.nrepl > 0) {
        info->ierr = -3;
        info->supeqt = -i64(st.nrepl);
    } else if (info->numneg > 0) {
        info->ierr = 3;
    }
    (void)neq;
    return 0;
}

int copy_factor_to_sm(sprgpu_handle *h, double *sm_out)
{
    // supernodes before the failing one only (the reference stops there)
    FactorStatus st;
    SPRGPU_HIP(hipMemcpyAsync(&st, h->d.status, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    SPRGPU_HIP(hipStreamSynchronize(h->stream));
    const Plan &p = h->plan;
    i32 node_limit = i32(p.nsuper);
    i64 diag_limit = p.neq;
    if (h->mode <= 2 && st.nzero > 0) {
        node_limit = 0;
        diag_limit = 0;
    } else if (h->mode != 3 && st.first_fail != ULLONG_MAX) {
        diag_limit = i64(st.first_fail);
        node_limit = 0;
        while (node_limit < i32(p.nsuper) &&
               p.nodes[node_limit].first + p.nodes[node_limit].S <= diag_limit)
            ++node_limit;
    }
    const unsigned nt = unsigned(p.copy_tasks.size());
    if (nt) {
        if (h->precision == SPRGPU_FP32)
            k_copy_out<float><<<nt, 256, 0, h->stream>>>(
                h->d.copy_tasks, h->d.nodes, static_cast<const float *>(h->d.store), sm_out,
                p.neq, node_limit);
        else
            k_copy_out<double><<<nt, 256, 0, h->stream>>>(
                h->d.copy_tasks, h->d.nodes, static_cast<const double *>(h->d.store), sm_out,
                p.neq, node_limit);
    }
    k_copy_diag<<<256, 256, 0, h->stream>>>(h->d.dvec, h->d.adiag, sm_out, p.neq, diag_limit);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int load_factor_from_sm(sprgpu_handle *h, const double *sm_dev)
{
    const Plan &p = h->plan;
    const unsigned nt = unsigned(p.copy_tasks.size());
    if (nt) {
        if (h->precision == SPRGPU_FP32)
            k_load_factor<float><<<nt, 256, 0, h->stream>>>(
                h->d.copy_tasks, h->d.nodes, static_cast<float *>(h->d.store), sm_dev, p.neq);
        else
            k_load_factor<double><<<nt, 256, 0, h->stream>>>(
                h->d.copy_tasks, h->d.nodes, static_cast<double *>(h->d.store), sm_dev, p.neq);
    }
    SPRGPU_HIP(hipMemcpyAsync(h->d.dvec, sm_dev, sizeof(double) * p.neq,
                              hipMemcpyDeviceToDevice, h->stream));
    SPRGPU_HIP(hipGetLastError());
    if (int rc = compute_block_inverses(h)) return rc;
    h->factored = true;
    return 0;
}

}  // namespace sprgpu
