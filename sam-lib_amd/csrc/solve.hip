// Triangular solves, FP64 residual SpMV and iterative refinement.
//
// Forward + diagonal solve replaces SPRFS1/SPRFS3/SPRFS5/SPRFS6
// (sprfs1.f:5-633, :929-1141) and the backward solve replaces
// SPRBS1/SPRBS4/SPRBS6 (sprbs1.f:5-574).  The reference walks supernodes one
// at a time, gathering B(LINDX) into a local vector and scattering the
// updated entries back by assignment.  Here all supernodes of one tree height
// run in one launch; the forward sweep passes each supernode's update vector
// (the F-S rows below its diagonal block) to its parent through the same
// child->parent relative map as the factor's extend-add, so siblings never
// write the same entry and results do not depend on scheduling.  The factor
// may be FP32 or FP64; vectors are always FP64 and the pivots come from the
// FP64 array D.  The division by d is a true division (sprfs1.f:622-625).
#include <cfloat>
#include <cmath>

#include "sprgpu_internal.h"

namespace sprgpu {

// ---------------------------------------------------------------------------
// forward: v = [B(rows 0..S-1); 0] + sum_children extend(u_child);
//          x1 = L11^-1 v1;  u = v2 - L21 x1;  B(rows 0..S-1) = x1 / d
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_fwd(const i32 *__restrict__ lnodes,
                                             const NodeInfo *__restrict__ nodes,
                                             const i32 *__restrict__ children,
                                             const i32 *__restrict__ lindx,
                                             const i32 *__restrict__ relmap,
                                             const T *__restrict__ pool,
                                             const double *__restrict__ dvec, double *b, i64 ldb,
                                             double *vwork, i64 vstride)
{
    const NodeInfo n = nodes[lnodes[blockIdx.x]];
    const int rhs = blockIdx.y, tid = threadIdx.x;
    double *v = vwork + rhs * vstride + n.voff;
    double *bb = b + rhs * ldb;
    const i32 *rows = lindx + n.lxoff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    const int S = n.S, F = n.F;
    for (int i = tid; i < F; i += 256) v[i] = (i < S) ? bb[rows[i]] : 0.0;
    __syncthreads();
    for (int ci = 0; ci < n.nchild; ++ci) {
        const NodeInfo c = nodes[children[n.child0 + ci]];
        const double *vc = vwork + rhs * vstride + c.voff + c.S;
        const i32 *map = relmap + c.mapoff;
        for (int i = tid; i < c.F - c.S; i += 256) v[map[i]] += vc[i];
        __syncthreads();
    }
    for (int c = 0; c < S - 1; ++c) {
        const double xc = v[c];
        for (int r = c + 1 + tid; r < S; r += 256) v[r] -= double(L[r + c * ld]) * xc;
        __syncthreads();
    }
    for (int i = tid; i < F - S; i += 256) {
        double s = v[S + i];
        for (int c = 0; c < S; ++c) s -= double(L[S + i + c * ld]) * v[c];
        v[S + i] = s;
    }
    for (int c = tid; c < S; c += 256) bb[rows[c]] = v[c] / dvec[n.first + c];
}

// ---------------------------------------------------------------------------
// backward: x1 = y1 - L21^T x2, then L11^T x1 = x1 (unit upper)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_bwd(const i32 *__restrict__ lnodes,
                                             const NodeInfo *__restrict__ nodes,
                                             const i32 *__restrict__ lindx,
                                             const T *__restrict__ pool, double *b, i64 ldb,
                                             double *vwork, i64 vstride)
{
    const NodeInfo n = nodes[lnodes[blockIdx.x]];
    const int rhs = blockIdx.y, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    double *v = vwork + rhs * vstride + n.voff;
    double *bb = b + rhs * ldb;
    const i32 *rows = lindx + n.lxoff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    const int S = n.S, F = n.F;
    for (int i = tid; i < F; i += 256) v[i] = bb[rows[i]];
    __syncthreads();
    for (int c = wave; c < S; c += 4) {
        double s = 0.0;
        for (int i = S + lane; i < F; i += 64) s += double(L[i + c * ld]) * v[i];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
        if (lane == 0) v[c] -= s;
    }
    __syncthreads();
    if (wave == 0) {
        for (int c = S - 2; c >= 0; --c) {
            double s = 0.0;
            for (int r = c + 1 + lane; r < S; r += 64) s += double(L[r + c * ld]) * v[r];
            for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
            if (lane == 0) v[c] -= s;
            __builtin_amdgcn_wave_barrier();
            __threadfence_block();
        }
    }
    __syncthreads();
    for (int c = tid; c < S; c += 256) bb[rows[c]] = v[c];
}

// ---------------------------------------------------------------------------
// wide supernodes (S > kSolveBigS): blocked forward and backward sweeps
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_sdesc(const i64 *prefix, int n, i64 b)
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

// v = [B(rows 0..S-1); 0] + sum_children extend(u_child)
__global__ __launch_bounds__(256) void k_fwd_assemble(const i32 *__restrict__ big,
                                                      const NodeInfo *__restrict__ nodes,
                                                      const i32 *__restrict__ children,
                                                      const i32 *__restrict__ lindx,
                                                      const i32 *__restrict__ relmap,
                                                      const double *b, i64 ldb, double *vwork,
                                                      i64 vstride)
{
    const NodeInfo n = nodes[big[blockIdx.x]];
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

// x_block = L(block,block)^-1 v_block (unit lower), in place in v
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_diag(const SolveDesc *__restrict__ desc,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const T *__restrict__ pool, double *vwork,
                                                  i64 vstride)
{
    __shared__ double x[kSolveNB];
    __shared__ T Lb[kSolveNB][kSolveNB + 1];
    const SolveDesc sd = desc[blockIdx.x];
    const NodeInfo n = nodes[sd.node];
    const int rhs = blockIdx.y, tid = threadIdx.x, nb = sd.nb, c0 = sd.c0;
    double *v = vwork + rhs * vstride + n.voff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    if (tid < nb) x[tid] = v[c0 + tid];
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int r = idx % nb, c = idx / nb;
        if (r > c) Lb[r][c] = L[(c0 + r) + (c0 + c) * ld];
    }
    __syncthreads();
    for (int c = 0; c < nb - 1; ++c) {
        const double xc = x[c];
        if (tid > c && tid < nb) x[tid] -= double(Lb[tid][c]) * xc;
        __syncthreads();
    }
    if (tid < nb) v[c0 + tid] = x[tid];
}

// v(rows below the block) -= L(rows, block) x_block
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_upd(const SolveDesc *__restrict__ desc, int ndesc,
                                                 const i64 *__restrict__ prefix,
                                                 const NodeInfo *__restrict__ nodes,
                                                 const T *__restrict__ pool, double *vwork,
                                                 i64 vstride)
{
    __shared__ double x[kSolveNB];
    const int q = find_sdesc(prefix, ndesc, blockIdx.x);
    const SolveDesc sd = desc[q];
    const int chunk = int(blockIdx.x - prefix[q]);
    const NodeInfo n = nodes[sd.node];
    const int rhs = blockIdx.y, tid = threadIdx.x, nb = sd.nb, c0 = sd.c0;
    double *v = vwork + rhs * vstride + n.voff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    if (tid < nb) x[tid] = v[c0 + tid];
    __syncthreads();
    const int r = c0 + nb + chunk * kSolveRows + tid;
    if (r < n.F) {
        double s = v[r];
        for (int c = 0; c < nb; ++c) s -= double(L[r + (c0 + c) * ld]) * x[c];
        v[r] = s;
    }
}

// B(rows 0..S-1) = x / d
__global__ __launch_bounds__(256) void k_fwd_finish(const i32 *__restrict__ big,
                                                    const NodeInfo *__restrict__ nodes,
                                                    const i32 *__restrict__ lindx,
                                                    const double *__restrict__ dvec, double *b,
                                                    i64 ldb, const double *vwork, i64 vstride)
{
    const NodeInfo n = nodes[big[blockIdx.x]];
    const int rhs = blockIdx.y;
    const double *v = vwork + rhs * vstride + n.voff;
    double *bb = b + rhs * ldb;
    const i32 *rows = lindx + n.lxoff;
    for (int c = threadIdx.x; c < n.S; c += 256) bb[rows[c]] = v[c] / dvec[n.first + c];
}

__global__ __launch_bounds__(256) void k_bwd_assemble(const i32 *__restrict__ big,
                                                      const NodeInfo *__restrict__ nodes,
                                                      const i32 *__restrict__ lindx,
                                                      const double *b, i64 ldb, double *vwork,
                                                      i64 vstride)
{
    const NodeInfo n = nodes[big[blockIdx.x]];
    const int rhs = blockIdx.y;
    double *v = vwork + rhs * vstride + n.voff;
    const double *bb = b + rhs * ldb;
    const i32 *rows = lindx + n.lxoff;
    for (int i = threadIdx.x; i < n.F; i += 256) v[i] = bb[rows[i]];
}

// partial[c] = sum over one chunk of rows below the block of L(r, c) v(r)
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_partial(const SolveDesc *__restrict__ desc,
                                                     int ndesc, const i64 *__restrict__ prefix,
                                                     const NodeInfo *__restrict__ nodes,
                                                     const T *__restrict__ pool,
                                                     const double *vwork, i64 vstride,
                                                     double *partial, i64 pstride)
{
    const int q = find_sdesc(prefix, ndesc, blockIdx.x);
    const SolveDesc sd = desc[q];
    const int chunk = int(blockIdx.x - prefix[q]);
    const NodeInfo n = nodes[sd.node];
    const int rhs = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double *v = vwork + rhs * vstride + n.voff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    const int r0 = sd.c0 + sd.nb + chunk * kSolveRows;
    const int r1 = min(n.F, r0 + kSolveRows);
    double *out = partial + rhs * pstride + (prefix[q] + chunk) * kSolveNB;
    for (int c = wave; c < sd.nb; c += 4) {
        double s = 0.0;
        for (int r = r0 + lane; r < r1; r += 64) s += double(L[r + (sd.c0 + c) * ld]) * v[r];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
        if (lane == 0) out[c] = s;
    }
}

// x_block = L(block,block)^-T (v_block - sum of partials), unit upper
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_diag(const SolveDesc *__restrict__ desc,
                                                  const i64 *__restrict__ prefix,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const T *__restrict__ pool, double *vwork,
                                                  i64 vstride, const double *partial,
                                                  i64 pstride)
{
    __shared__ double x[kSolveNB];
    __shared__ T Lb[kSolveNB][kSolveNB + 1];
    const SolveDesc sd = desc[blockIdx.x];
    const NodeInfo n = nodes[sd.node];
    const int rhs = blockIdx.y, tid = threadIdx.x, nb = sd.nb, c0 = sd.c0;
    double *v = vwork + rhs * vstride + n.voff;
    const T *L = pool + n.foff;
    const i64 ld = n.ld;
    if (tid < nb) {
        double s = v[c0 + tid];
        const double *pp = partial + rhs * pstride + prefix[blockIdx.x] * kSolveNB + tid;
        for (int k = 0; k < sd.nchunk; ++k) s -= pp[k * kSolveNB];
        x[tid] = s;
    }
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int r = idx % nb, c = idx / nb;
        if (r > c) Lb[r][c] = L[(c0 + r) + (c0 + c) * ld];
    }
    __syncthreads();
    for (int r = nb - 1; r > 0; --r) {
        const double xr = x[r];
        if (tid < r) x[tid] -= double(Lb[r][tid]) * xr;
        __syncthreads();
    }
    if (tid < nb) v[c0 + tid] = x[tid];
}

__global__ __launch_bounds__(256) void k_bwd_finish(const i32 *__restrict__ big,
                                                    const NodeInfo *__restrict__ nodes,
                                                    const i32 *__restrict__ lindx, double *b,
                                                    i64 ldb, const double *vwork, i64 vstride)
{
    const NodeInfo n = nodes[big[blockIdx.x]];
    const int rhs = blockIdx.y;
    const double *v = vwork + rhs * vstride + n.voff;
    double *bb = b + rhs * ldb;
    const i32 *rows = lindx + n.lxoff;
    for (int c = threadIdx.x; c < n.S; c += 256) bb[rows[c]] = v[c];
}

// ---------------------------------------------------------------------------
// FP64 residual: y = alpha*b + beta*A x (CSR, 16 lanes per row)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_spmv(const i64 *__restrict__ rowptr,
                                              const i32 *__restrict__ col,
                                              const double *__restrict__ val,
                                              const double *__restrict__ x, double *y,
                                              const double *__restrict__ bvec, i64 n, i64 ld,
                                              double beta)
{
    const i64 row = (i64(blockIdx.x) * 256 + threadIdx.x) >> 4;
    const int sub = threadIdx.x & 15;
    const int rhs = blockIdx.y;
    const double *xx = x + rhs * ld;
    double s = 0.0;
    if (row < n) {
        for (i64 k = rowptr[row] + sub; k < rowptr[row + 1]; k += 16) s += val[k] * xx[col[k]];
    }
    for (int o = 8; o > 0; o >>= 1) s += __shfl_down(s, o, 16);
    if (row < n && sub == 0) {
        const double b0 = bvec ? bvec[rhs * ld + row] : 0.0;
        y[rhs * ld + row] = b0 + beta * s;
    }
}

// gathers the CSR values from the packed assembled matrix and counts the
// non-zeros found on or below the diagonal (each packed entry once)
__global__ __launch_bounds__(256) void k_gather_vals(const i64 *__restrict__ rowptr,
                                                     const i32 *__restrict__ col,
                                                     const i64 *__restrict__ src,
                                                     const double *__restrict__ lnz,
                                                     double *val, i64 n,
                                                     unsigned long long *count)
{
    unsigned long long c = 0;
    for (i64 row = i64(blockIdx.x) * 256 + threadIdx.x; row < n; row += i64(gridDim.x) * 256) {
        for (i64 k = rowptr[row]; k < rowptr[row + 1]; ++k) {
            const double v = lnz[src[k]];
            val[k] = v;
            c += (v != 0.0 && col[k] <= row);
        }
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

// per-column sums of squares, deterministic two-pass reduction
constexpr int kNormBlocks = 128;
__global__ __launch_bounds__(256) void k_sumsq_partial(const double *x, i64 n, i64 ld,
                                                       double *part)
{
    const int rhs = blockIdx.y;
    const double *xx = x + rhs * ld;
    double s = 0.0;
    for (i64 i = i64(blockIdx.x) * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256)
        s += xx[i] * xx[i];
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[rhs * kNormBlocks + blockIdx.x] = sh[0];
}

__global__ void k_sumsq_final(const double *part, double *out, int nrhs)
{
    const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
    if (rhs < nrhs) {
        double s = 0.0;
        for (int b = 0; b < kNormBlocks; ++b) s += part[rhs * kNormBlocks + b];
        out[rhs] = s;
    }
}

__global__ void k_axpy_cols(double *x, const double *dx, i64 n, i64 ld)
{
    const int rhs = blockIdx.y;
    for (i64 i = i64(blockIdx.x) * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256)
        x[rhs * ld + i] += dx[rhs * ld + i];
}

__global__ void k_copy_cols(double *dst, i64 ldd, const double *src, i64 lds, i64 n)
{
    const int rhs = blockIdx.y;
    for (i64 i = i64(blockIdx.x) * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256)
        dst[rhs * ldd + i] = src[rhs * lds + i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int ensure_vwork(sprgpu_handle *h, i64 nrhs)
{
    DeviceBuffers &d = h->d;
    if (d.vwork_cap >= nrhs) return 0;
    if (d.vwork) SPRGPU_HIP(hipFree(d.vwork));
    d.vwork = nullptr;
    SPRGPU_HIP(hipMalloc(&d.vwork, sizeof(double) * size_t(h->plan.vwork_elems) * size_t(nrhs)));
    h->device_bytes += sizeof(double) * h->plan.vwork_elems * (nrhs - d.vwork_cap);
    d.vwork_cap = nrhs;
    return 0;
}

template <typename T>
static int solve_impl(sprgpu_handle *h, double *b, i64 ldb, i64 nrhs, int which)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    hipStream_t s = h->stream;
    if (int rc = ensure_vwork(h, nrhs)) return rc;
    const T *pool = static_cast<const T *>(d.pool);
    const i64 vstride = p.vwork_elems;
    const unsigned ur = unsigned(nrhs);
    const i64 pstride = std::max<i64>(p.max_partial_chunks, 1) * kSolveNB;
    if ((which & 2) && p.max_partial_chunks > 0 && d.partial_cap < pstride * nrhs) {
        if (d.partial) SPRGPU_HIP(hipFree(d.partial));
        d.partial = nullptr;
        SPRGPU_HIP(hipMalloc(&d.partial, sizeof(double) * pstride * nrhs));
        d.partial_cap = pstride * nrhs;
    }
    if (which & 1) {
        size_t si = 0;
        for (int l = 0; l < p.nlevels; ++l) {
            const unsigned ns = unsigned(p.solve_small_ptr[l + 1] - p.solve_small_ptr[l]);
            if (ns)
                k_fwd<T><<<dim3(ns, ur), 256, 0, s>>>(d.solve_small + p.solve_small_ptr[l],
                                                      d.nodes, d.children, d.lindx, d.relmap,
                                                      pool, d.dvec, b, ldb, d.vwork, vstride);
            for (; si < p.fwd_steps.size() && p.fwd_steps[si].level == l; ++si) {
                const SolveStep &st = p.fwd_steps[si];
                const dim3 g(unsigned(st.blocks), ur);
                switch (st.kind) {
                case 0:
                    k_fwd_assemble<<<g, 256, 0, s>>>(d.solve_big + st.off, d.nodes, d.children,
                                                     d.lindx, d.relmap, b, ldb, d.vwork, vstride);
                    break;
                case 1:
                    k_fwd_diag<T><<<g, 256, 0, s>>>(d.fwd_desc + st.off, d.nodes, pool, d.vwork,
                                                    vstride);
                    break;
                case 2:
                    k_fwd_upd<T><<<g, 256, 0, s>>>(d.fwd_desc + st.off, int(st.count),
                                                   d.fwd_prefix + st.off, d.nodes, pool, d.vwork,
                                                   vstride);
                    break;
                case 3:
                    k_fwd_finish<<<g, 256, 0, s>>>(d.solve_big + st.off, d.nodes, d.lindx,
                                                   d.dvec, b, ldb, d.vwork, vstride);
                    break;
                }
            }
        }
    }
    if (which & 2) {
        size_t si = 0;
        for (int l = p.nlevels - 1; l >= 0; --l) {
            const unsigned ns = unsigned(p.solve_small_ptr[l + 1] - p.solve_small_ptr[l]);
            if (ns)
                k_bwd<T><<<dim3(ns, ur), 256, 0, s>>>(d.solve_small + p.solve_small_ptr[l],
                                                      d.nodes, d.lindx, pool, b, ldb, d.vwork,
                                                      vstride);
            for (; si < p.bwd_steps.size() && p.bwd_steps[si].level == l; ++si) {
                const SolveStep &st = p.bwd_steps[si];
                const dim3 g(unsigned(st.blocks), ur);
                switch (st.kind) {
                case 0:
                    k_bwd_assemble<<<g, 256, 0, s>>>(d.solve_big + st.off, d.nodes, d.lindx, b,
                                                     ldb, d.vwork, vstride);
                    break;
                case 1:
                    k_bwd_diag<T><<<g, 256, 0, s>>>(d.bwd_desc + st.off, d.bwd_prefix + st.off,
                                                    d.nodes, pool, d.vwork, vstride, d.partial,
                                                    pstride);
                    break;
                case 2:
                    k_bwd_partial<T><<<g, 256, 0, s>>>(d.bwd_desc + st.off, int(st.count),
                                                       d.bwd_prefix + st.off, d.nodes, pool,
                                                       d.vwork, vstride, d.partial, pstride);
                    break;
                case 3:
                    k_bwd_finish<<<g, 256, 0, s>>>(d.solve_big + st.off, d.nodes, d.lindx, b,
                                                   ldb, d.vwork, vstride);
                    break;
                }
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

int upload_pattern(sprgpu_handle *h, const CsrPattern &pat)
{
    DeviceBuffers &d = h->d;
    const i64 n = h->plan.neq, nnz = pat.rowptr[n];
    if (d.csr_rowptr) SPRGPU_HIP(hipFree(d.csr_rowptr));
    if (d.csr_col) SPRGPU_HIP(hipFree(d.csr_col));
    if (d.csr_src) SPRGPU_HIP(hipFree(d.csr_src));
    if (d.csr_val) SPRGPU_HIP(hipFree(d.csr_val));
    d.csr_rowptr = nullptr;
    d.csr_col = nullptr;
    d.csr_src = nullptr;
    d.csr_val = nullptr;
    SPRGPU_HIP(hipMalloc(&d.csr_rowptr, sizeof(i64) * (n + 1)));
    SPRGPU_HIP(hipMalloc(&d.csr_col, sizeof(i32) * std::max<i64>(nnz, 1)));
    SPRGPU_HIP(hipMalloc(&d.csr_src, sizeof(i64) * std::max<i64>(nnz, 1)));
    SPRGPU_HIP(hipMalloc(&d.csr_val, sizeof(double) * std::max<i64>(nnz, 1)));
    SPRGPU_HIP(hipMemcpy(d.csr_rowptr, pat.rowptr.data(), sizeof(i64) * (n + 1),
                         hipMemcpyHostToDevice));
    if (nnz) {
        SPRGPU_HIP(hipMemcpy(d.csr_col, pat.col.data(), sizeof(i32) * nnz, hipMemcpyHostToDevice));
        SPRGPU_HIP(hipMemcpy(d.csr_src, pat.src.data(), sizeof(i64) * nnz, hipMemcpyHostToDevice));
    }
    h->device_bytes += (sizeof(i64) * 2 + sizeof(i32) + sizeof(double)) * nnz;
    d.csr_nnz = nnz;
    d.nnz_lower = pat.nnz_lower;
    h->has_pattern = true;
    return 0;
}

// gathers A's values for the last factorization's input and verifies that the
// pattern still covers every non-zero
static int refresh_values(sprgpu_handle *h)
{
    DeviceBuffers &d = h->d;
    if (!h->has_pattern) {
        set_error("no residual pattern: call sprgpu_set_pattern with the assembled SM first");
        return -1;
    }
    if (!h->last_sm_dev) {
        set_error("no factorization input recorded");
        return -1;
    }
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(&d.status->pad[0]);
    SPRGPU_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), h->stream));
    k_gather_vals<<<1024, 256, 0, h->stream>>>(d.csr_rowptr, d.csr_col, d.csr_src,
                                               h->last_sm_dev + h->plan.neq, d.csr_val,
                                               h->plan.neq, cnt);
    SPRGPU_HIP(hipGetLastError());
    FactorStatus st;
    SPRGPU_HIP(hipMemcpyAsync(&st, d.status, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    SPRGPU_HIP(hipStreamSynchronize(h->stream));
    if (st.pad[0] != st.nnz_seen) {
        set_error("the assembled matrix has non-zeros outside the residual pattern; "
                  "call sprgpu_set_pattern again");
        h->has_pattern = false;
        return -1;
    }
    return 0;
}

int run_spmv(sprgpu_handle *h, const double *x, double *y, i64 ld, i64 nrhs)
{
    if (int rc = refresh_values(h)) return rc;
    const i64 n = h->plan.neq;
    dim3 grid(unsigned((n * 16 + 255) / 256), unsigned(nrhs));
    k_spmv<<<grid, 256, 0, h->stream>>>(h->d.csr_rowptr, h->d.csr_col, h->d.csr_val, x, y,
                                        nullptr, n, ld, 1.0);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

static int col_norms(sprgpu_handle *h, const double *x, i64 n, i64 ld, i64 nrhs, double *part,
                     double *out_dev, std::vector<double> &out)
{
    dim3 g1(kNormBlocks, unsigned(nrhs));
    k_sumsq_partial<<<g1, 256, 0, h->stream>>>(x, n, ld, part);
    k_sumsq_final<<<unsigned((nrhs + 63) / 64), 64, 0, h->stream>>>(part, out_dev, int(nrhs));
    out.resize(nrhs);
    SPRGPU_HIP(hipMemcpyAsync(out.data(), out_dev, sizeof(double) * nrhs, hipMemcpyDeviceToHost,
                              h->stream));
    SPRGPU_HIP(hipStreamSynchronize(h->stream));
    for (auto &v : out) v = std::sqrt(v);
    return 0;
}

int run_solve_ir(sprgpu_handle *h, const double *b, double *x, i64 ld, i64 nrhs, double rtol,
                 int maxit, sprgpu_ir_info *info)
{
    if (int rc = refresh_values(h)) return rc;
    DeviceBuffers &d = h->d;
    const i64 n = h->plan.neq;
    hipStream_t s = h->stream;
    // workspace: r (n*nrhs), partials, norms
    const i64 need = n * nrhs + i64(kNormBlocks) * nrhs + nrhs;
    if (d.irwork_cap < need) {
        if (d.irwork) SPRGPU_HIP(hipFree(d.irwork));
        d.irwork = nullptr;
        SPRGPU_HIP(hipMalloc(&d.irwork, sizeof(double) * need));
        d.irwork_cap = need;
    }
    double *r = d.irwork;
    double *part = r + n * nrhs;
    double *nrm = part + i64(kNormBlocks) * nrhs;
    const unsigned gx = unsigned(std::min<i64>((n + 255) / 256, 1024));
    dim3 gcol(gx, unsigned(nrhs));
    std::vector<double> bnorm, rnorm;
    if (int rc = col_norms(h, b, n, ld, nrhs, part, nrm, bnorm)) return rc;
    // x = A^-1 b
    k_copy_cols<<<gcol, 256, 0, s>>>(x, ld, b, ld, n);
    if (int rc = run_solve(h, x, ld, nrhs, 3)) return rc;
    dim3 gsp(unsigned((n * 16 + 255) / 256), unsigned(nrhs));
    int it = 0;
    double worst = 0.0;
    for (;;) {
        // r = b - A x
        k_spmv<<<gsp, 256, 0, s>>>(d.csr_rowptr, d.csr_col, d.csr_val, x, r, b, n, n, -1.0);
        if (ld != n) {
            set_error("solve_ir requires ld == neq");
            return -1;
        }
        if (int rc = col_norms(h, r, n, n, nrhs, part, nrm, rnorm)) return rc;
        worst = 0.0;
        for (i64 k = 0; k < nrhs; ++k) {
            const double rel = bnorm[k] > 0 ? rnorm[k] / bnorm[k] : rnorm[k];
            worst = std::max(worst, rel);
        }
        if (!(worst > rtol) || it >= maxit) break;
        if (int rc = run_solve(h, r, n, nrhs, 3)) return rc;
        k_axpy_cols<<<gcol, 256, 0, s>>>(x, r, n, ld);
        ++it;
    }
    SPRGPU_HIP(hipGetLastError());
    if (info) {
        info->iterations = it;
        info->relres = worst;
        info->converged = (worst <= rtol) ? 1 : 0;
    }
    return 0;
}

}  // namespace sprgpu
