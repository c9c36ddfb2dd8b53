// Forward + diagonal and backward substitution with the supernodal factor.
//
// Replaces SPRFS1/SPRFS3/SPRFS5/SPRFS6 (sprfs1.f:5-633, :929-1141) and
// SPRBS1/SPRBS4/SPRBS6 (sprbs1.f:5-574).  The reference walks the supernodes
// one at a time, one right-hand side at a time, gathering B(LINDX) into a
// local vector and scattering the updated entries back by assignment.  Here
// the supernodes of one tree height run together:
//
// * Narrow supernodes (S <= kSolveSmallS, F <= kSolveSmallF) get one
//   workgroup each with the local vector in LDS and the unit-triangular part
//   solved by a single wavefront with lane shuffles.
// * Wide supernodes are split into kSolveNB-column blocks, one workgroup per
//   block, plus workgroups for the rows below the diagonal part (forward).  A
//   block's workgroup applies the finished blocks it depends on as their
//   ready flags appear, then solves its diagonal block and publishes it.
//   Every workgroup only waits for lower-indexed workgroups of the same
//   launch, which the hardware dispatches first, and each wait is bounded.
//
// The forward sweep hands each supernode's update vector (its F-S rows below
// the diagonal block) to the parent through the child->parent relative map,
// the same map the factor's extend-add uses, so no two workgroups write the
// same entry and every sum is taken in a fixed order: results do not depend
// on scheduling.  Vectors are FP64 whatever the factor precision; the
// division by d is a true division (sprfs1.f:622-625).
#include "sprgpu_internal.h"

namespace sprgpu {

__device__ __forceinline__ double wave_sum(double t)
{
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    return t;
}

constexpr unsigned long long kSpinLimit = 1ull << 24;

// One thread waits until *flag reaches epoch (bounded); then the workgroup
// acquires.  Returns false if the wait gave up.
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
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Called by every lane of the wavefront that wrote the block.
__device__ __forceinline__ void publish_ready(unsigned *flag, unsigned epoch)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if ((threadIdx.x & 63) == 0)
        __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int find_wide(const i64 *prefix, int n, i64 b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// narrow supernodes: S <= kSolveSmallS (<= 64), F <= kSolveSmallF
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_small(const i32 *__restrict__ list,
                                                   const NodeInfo *__restrict__ nodes,
                                                   const i32 *__restrict__ children,
                                                   const i32 *__restrict__ lindx,
                                                   const i32 *__restrict__ relmap,
                                                   const T *__restrict__ pool,
                                                   const double *__restrict__ dvec, double *b,
                                                   i64 ldb, double *vwork, i64 vstride)
{
    __shared__ double v[kSolveSmallF];
    __shared__ T Lb[kSolveNB][kSolveNB + 1];
    const NodeInfo n = nodes[list[blockIdx.x]];
    const int rhs = blockIdx.y, tid = threadIdx.x, lane = tid & 63;
    const int S = n.S, F = n.F;
    double *vg = vwork + rhs * vstride + n.voff;
    double *bb = b + rhs * ldb;
    const i32 *rows = lindx + n.lxoff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    for (int i = tid; i < F; i += 256) v[i] = (i < S) ? bb[rows[i]] : 0.0;
    for (int idx = tid; idx < S * S; idx += 256) {
        const int r = idx % S, c = idx / S;
        if (r > c) Lb[r][c] = L[r + c * ld];
    }
    __syncthreads();
    for (int ci = 0; ci < n.nchild; ++ci) {
        const NodeInfo c = nodes[children[n.child0 + ci]];
        const double *vc = vwork + rhs * vstride + c.voff + c.S;
        const i32 *map = relmap + c.mapoff;
        for (int i = tid; i < c.F - c.S; i += 256) v[map[i]] += vc[i];
        __syncthreads();
    }
    if (tid < 64) {
        // unit lower L11, row lane in a register
        double x = (lane < S) ? v[lane] : 0.0;
        for (int c = 0; c < S - 1; ++c) {
            const double xc = __shfl(x, c);
            if (lane > c && lane < S) x -= double(Lb[lane][c]) * xc;
        }
        if (lane < S) v[lane] = x;
    }
    __syncthreads();
    for (int i = S + tid; i < F; i += 256) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int c = 0;
        for (; c + 3 < S; c += 4) {
            s0 += double(L[i + c * ld]) * v[c];
            s1 += double(L[i + (c + 1) * ld]) * v[c + 1];
            s2 += double(L[i + (c + 2) * ld]) * v[c + 2];
            s3 += double(L[i + (c + 3) * ld]) * v[c + 3];
        }
        for (; c < S; ++c) s0 += double(L[i + c * ld]) * v[c];
        vg[i] = v[i] - ((s0 + s1) + (s2 + s3));  // update vector handed to the parent
    }
    for (int c = tid; c < S; c += 256) bb[rows[c]] = v[c] / dvec[n.first + c];
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_small(const i32 *__restrict__ list,
                                                   const NodeInfo *__restrict__ nodes,
                                                   const i32 *__restrict__ lindx,
                                                   const T *__restrict__ pool, double *b,
                                                   i64 ldb)
{
    __shared__ double v[kSolveSmallF];
    __shared__ T Lb[kSolveNB][kSolveNB + 1];
    __shared__ double part[4][64];
    const NodeInfo n = nodes[list[blockIdx.x]];
    const int rhs = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int S = n.S, F = n.F;
    double *bb = b + rhs * ldb;
    const i32 *rows = lindx + n.lxoff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    for (int i = tid; i < F; i += 256) v[i] = bb[rows[i]];
    for (int idx = tid; idx < S * S; idx += 256) {
        const int r = idx % S, c = idx / S;
        if (r > c) Lb[r][c] = L[r + c * ld];
    }
    __syncthreads();
    // x1 -= L21^T x2: lane = column, the four waves interleave the rows
    double acc = 0.0;
    if (lane < S) {
        const T *Lcol = L + i64(lane) * ld;
        const int nrw = F - S, per = (nrw + 3) / 4;
        const int ra = S + wave * per, rz = min(F, ra + per);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int r = ra;
        for (; r + 3 < rz; r += 4) {
            a0 += double(Lcol[r]) * v[r];
            a1 += double(Lcol[r + 1]) * v[r + 1];
            a2 += double(Lcol[r + 2]) * v[r + 2];
            a3 += double(Lcol[r + 3]) * v[r + 3];
        }
        for (; r < rz; ++r) a0 += double(Lcol[r]) * v[r];
        acc = (a0 + a1) + (a2 + a3);
    }
    part[wave][lane] = acc;
    __syncthreads();
    if (tid < 64) {
        // L11^T x1 = x1 by columns from the last: x_c -= l_rc x_r for c < r
        double x = (lane < S)
                       ? v[lane] - (((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane])
                       : 0.0;
        for (int r = S - 1; r > 0; --r) {
            const double xr = __shfl(x, r);
            if (lane < r) x -= double(Lb[r][lane]) * xr;
        }
        if (lane < S) bb[rows[lane]] = x;
    }
}

// ---------------------------------------------------------------------------
// wide supernodes
// ---------------------------------------------------------------------------
// v = [B(rows 0..S-1); 0] + sum_children extend(u_child)
__global__ __launch_bounds__(256) void k_fwd_assemble(const i32 *__restrict__ list,
                                                      const NodeInfo *__restrict__ nodes,
                                                      const i32 *__restrict__ children,
                                                      const i32 *__restrict__ lindx,
                                                      const i32 *__restrict__ relmap,
                                                      const double *b, i64 ldb, double *vwork,
                                                      i64 vstride)
{
    const NodeInfo n = nodes[list[blockIdx.x]];
    const int rhs = blockIdx.y, tid = threadIdx.x;
    double *v = vwork + rhs * vstride + n.voff;
    const double *bb = b + rhs * ldb;
    const i32 *rows = lindx + n.lxoff;
    for (int i = tid; i < n.F; i += 256) v[i] = (i < n.S) ? bb[rows[i]] : 0.0;
    __syncthreads();
    for (int ci = 0; ci < n.nchild; ++ci) {
        const NodeInfo c = nodes[children[n.child0 + ci]];
        const double *vc = vwork + rhs * vstride + c.voff + c.S;
        const i32 *map = relmap + c.mapoff;
        for (int i = tid; i < c.F - c.S; i += 256) v[map[i]] += vc[i];
        __syncthreads();
    }
}

// Forward dataflow: workgroup q < nblk solves diagonal block q after applying
// blocks 0..q-1; workgroups nblk.. reduce the rectangle rows.  The L values a
// workgroup needs are loaded before it waits, so only the dependent
// arithmetic is on the critical path.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_wide(const WideDesc *__restrict__ wide, int nwide,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i32 *__restrict__ lindx,
                                                  const T *__restrict__ pool,
                                                  const double *__restrict__ dvec, double *b,
                                                  i64 ldb, double *vwork, i64 vstride,
                                                  unsigned *flags, i64 nflags, unsigned epoch,
                                                  int *err)
{
    __shared__ double xs[kSolveNB];
    __shared__ double part[4][64];
    __shared__ T Lb[kSolveNB][kSolveNB + 1];
    const int q = find_wide(prefix, nwide, blockIdx.x);
    const WideDesc w = wide[q];
    const int local = int(blockIdx.x - prefix[q]);
    const NodeInfo n = nodes[w.node];
    const int rhs = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int S = n.S, F = n.F;
    double *v = vwork + rhs * vstride + n.voff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    unsigned *fl = flags + rhs * nflags + w.flag0;
    if (local < w.nblk) {
        const int r0 = local * kSolveNB;
        const int nr = min(kSolveNB, S - r0);
        const int r = r0 + lane;
        for (int idx = tid; idx < nr * nr; idx += 256) {
            const int rr = idx % nr, cc = idx / nr;
            if (rr > cc) Lb[rr][cc] = L[(r0 + rr) + i64(r0 + cc) * ld];
        }
        double acc = 0.0;
        for (int blk = 0; blk < local; ++blk) {
            T lv[16];
            const T *Lc = L + r + i64(blk * kSolveNB + wave * 16) * ld;
#pragma unroll
            for (int c = 0; c < 16; ++c) lv[c] = (lane < nr) ? Lc[c * ld] : T(0);
            wait_ready(fl + blk, epoch, err);
            if (tid < kSolveNB) xs[tid] = v[blk * kSolveNB + tid];  // full block: blk < nblk-1
            __syncthreads();
#pragma unroll
            for (int c = 0; c < 16; ++c) acc += double(lv[c]) * xs[wave * 16 + c];
            __syncthreads();
        }
        part[wave][lane] = acc;
        __syncthreads();
        if (wave == 0) {
            double x = 0.0;
            if (lane < nr)
                x = v[r] - (((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane]);
            for (int c = 0; c < nr - 1; ++c) {
                const double xc = __shfl(x, c);
                if (lane > c && lane < nr) x -= double(Lb[lane][c]) * xc;
            }
            if (lane < nr) {
                v[r] = x;
                b[rhs * ldb + lindx[n.lxoff + r]] = x / dvec[n.first + r];
            }
            publish_ready(fl + local, epoch);
        }
    } else {
        const int r = S + (local - w.nblk) * kSolveRows + tid;
        double s = (r < F) ? v[r] : 0.0;
        for (int blk = 0; blk < w.nblk; ++blk) {
            wait_ready(fl + blk, epoch, err);
            if (tid < kSolveNB)
                xs[tid] = (blk * kSolveNB + tid < S) ? v[blk * kSolveNB + tid] : 0.0;
            __syncthreads();
            const int nc = min(kSolveNB, S - blk * kSolveNB);
            if (r < F) {
                const T *Lc = L + r + i64(blk * kSolveNB) * ld;
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                int c = 0;
                for (; c + 3 < nc; c += 4) {
                    s0 += double(Lc[c * ld]) * xs[c];
                    s1 += double(Lc[(c + 1) * ld]) * xs[c + 1];
                    s2 += double(Lc[(c + 2) * ld]) * xs[c + 2];
                    s3 += double(Lc[(c + 3) * ld]) * xs[c + 3];
                }
                for (; c < nc; ++c) s0 += double(Lc[c * ld]) * xs[c];
                s -= (s0 + s1) + (s2 + s3);
            }
            __syncthreads();
        }
        if (r < F) v[r] = s;
    }
}

__global__ __launch_bounds__(256) void k_bwd_assemble(const i32 *__restrict__ list,
                                                      const NodeInfo *__restrict__ nodes,
                                                      const i32 *__restrict__ lindx,
                                                      const double *b, i64 ldb, double *vwork,
                                                      i64 vstride)
{
    const NodeInfo n = nodes[list[blockIdx.x]];
    const int rhs = blockIdx.y;
    double *v = vwork + rhs * vstride + n.voff;
    const double *bb = b + rhs * ldb;
    const i32 *rows = lindx + n.lxoff;
    for (int i = threadIdx.x; i < n.F; i += 256) v[i] = bb[rows[i]];
}

// Backward dataflow: workgroup q handles block nblk-1-q; it first applies the
// rectangle rows (known on entry), then the later blocks as they finish.
// Lane = column of the block; the four waves interleave the rows and their
// partial sums are combined in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_wide(const WideDesc *__restrict__ wide, int nwide,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i32 *__restrict__ lindx,
                                                  const T *__restrict__ pool, double *b, i64 ldb,
                                                  double *vwork, i64 vstride, unsigned *flags,
                                                  i64 nflags, unsigned epoch, int *err)
{
    __shared__ double part[4][64];
    __shared__ double xr[kSolveNB];
    __shared__ T Lb[kSolveNB][kSolveNB + 1];
    const int q = find_wide(prefix, nwide, blockIdx.x);
    const WideDesc w = wide[q];
    const int blkid = w.nblk - 1 - int(blockIdx.x - prefix[q]);
    const NodeInfo n = nodes[w.node];
    const int rhs = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int S = n.S, F = n.F;
    double *v = vwork + rhs * vstride + n.voff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    unsigned *fl = flags + rhs * nflags + w.flag0;
    const int c0 = blkid * kSolveNB;
    const int nc = min(kSolveNB, S - c0);
    for (int idx = tid; idx < nc * nc; idx += 256) {
        const int rr = idx % nc, cc = idx / nc;
        if (rr > cc) Lb[rr][cc] = L[(c0 + rr) + i64(c0 + cc) * ld];
    }
    const bool cv = lane < nc;
    const T *Lcol = L + i64(c0 + lane) * ld;
    double acc = 0.0;
    if (cv) {
        // wave w takes the rows of its quarter; 4 independent partial sums
        const int nrw = F - S, per = (nrw + 3) / 4;
        const int ra = S + wave * per, rz = min(F, ra + per);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int r = ra;
        for (; r + 3 < rz; r += 4) {
            a0 += double(Lcol[r]) * v[r];
            a1 += double(Lcol[r + 1]) * v[r + 1];
            a2 += double(Lcol[r + 2]) * v[r + 2];
            a3 += double(Lcol[r + 3]) * v[r + 3];
        }
        for (; r < rz; ++r) a0 += double(Lcol[r]) * v[r];
        acc = (a0 + a1) + (a2 + a3);
    }
    for (int blk = w.nblk - 1; blk > blkid; --blk) {
        T lv[16];
        const int rb = blk * kSolveNB + wave * 16;
#pragma unroll
        for (int k = 0; k < 16; ++k) lv[k] = (cv && rb + k < S) ? Lcol[rb + k] : T(0);
        wait_ready(fl + blk, epoch, err);
        if (tid < kSolveNB) xr[tid] = (blk * kSolveNB + tid < S) ? v[blk * kSolveNB + tid] : 0.0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) acc += double(lv[k]) * xr[wave * 16 + k];
        __syncthreads();
    }
    part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0) {
        const int r = c0 + lane;
        double x = cv ? v[r] - (((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane])
                      : 0.0;
        for (int rr = nc - 1; rr > 0; --rr) {
            const double xv = __shfl(x, rr);
            if (lane < rr) x -= double(Lb[rr][lane]) * xv;
        }
        if (cv) {
            v[r] = x;
            b[rhs * ldb + lindx[n.lxoff + r]] = x;
        }
        publish_ready(fl + blkid, epoch);
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
        d.epoch = 0;  // every flag is zero again
    }
    if (!d.spin_error) {
        SPRGPU_HIP(hipMalloc(&d.spin_error, sizeof(int)));
        SPRGPU_HIP(hipMemset(d.spin_error, 0, sizeof(int)));
    }
    return 0;
}

template <typename T>
static int solve_impl(sprgpu_handle *h, double *b, i64 ldb, i64 nrhs, int which)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    hipStream_t s = h->stream;
    if (int rc = ensure_solve_buffers(h, nrhs)) return rc;
    const T *pool = static_cast<const T *>(d.pool);
    const i64 vstride = p.vwork_elems;
    const unsigned ur = unsigned(nrhs);
    const i64 nflags = std::max<i64>(p.nflags, 1);
    if (which & 1) {
        const unsigned ep = ++d.epoch;
        for (int l = 0; l < p.nlevels; ++l) {
            const unsigned ns = unsigned(p.solve_small_ptr[l + 1] - p.solve_small_ptr[l]);
            if (ns)
                k_fwd_small<T><<<dim3(ns, ur), 256, 0, s>>>(
                    d.solve_small + p.solve_small_ptr[l], d.nodes, d.children, d.lindx, d.relmap,
                    pool, d.dvec, b, ldb, d.vwork, vstride);
            const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
            if (nw) {
                k_fwd_assemble<<<dim3(unsigned(nw), ur), 256, 0, s>>>(
                    d.wide_nodes + w0, d.nodes, d.children, d.lindx, d.relmap, b, ldb, d.vwork,
                    vstride);
                k_fwd_wide<T><<<dim3(unsigned(p.wide_fwd_wgs[l]), ur), 256, 0, s>>>(
                    d.wide + w0, nw, d.wide_fwd_prefix + w0, d.nodes, d.lindx, pool, d.dvec, b,
                    ldb, d.vwork, vstride, d.flags, nflags, ep, d.spin_error);
            }
        }
    }
    if (which & 2) {
        const unsigned ep = ++d.epoch;
        for (int l = p.nlevels - 1; l >= 0; --l) {
            const unsigned ns = unsigned(p.solve_small_ptr[l + 1] - p.solve_small_ptr[l]);
            if (ns)
                k_bwd_small<T><<<dim3(ns, ur), 256, 0, s>>>(d.solve_small + p.solve_small_ptr[l],
                                                            d.nodes, d.lindx, pool, b, ldb);
            const int w0 = p.wide_ptr[l], nw = p.wide_ptr[l + 1] - w0;
            if (nw) {
                k_bwd_assemble<<<dim3(unsigned(nw), ur), 256, 0, s>>>(
                    d.wide_nodes + w0, d.nodes, d.lindx, b, ldb, d.vwork, vstride);
                k_bwd_wide<T><<<dim3(unsigned(p.wide_bwd_wgs[l]), ur), 256, 0, s>>>(
                    d.wide + w0, nw, d.wide_bwd_prefix + w0, d.nodes, d.lindx, pool, b, ldb,
                    d.vwork, vstride, d.flags, nflags, ep, d.spin_error);
            }
        }
    }
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
