// SPRTMM on the device: products with the stored lower triangle of an SPR
// matrix (sprtmm.f:5-241, its product loop SPRTMQ at sprtmm.f:242-344).
//
// With L the stored lower triangle (diagonal included) of U in the LNZ
// layout, IFLAG 1 gives C = L^T B and IFLAG 2 gives C = L B.  SPRTMQ walks
// the packed columns in order from C = 0:
// * IFLAG 1 adds column J's entries to C(J), the diagonal first and then the
//   rows below in LINDX order: each C(J) is a sum over one packed column,
//   which is contiguous in LNZ;
// * IFLAG 2 adds L(r,J) B(J) to C(r) as it meets column J: each C(r) is a
//   sum over row r of L in ascending column order.  Row r's entries lie in
//   the supernodes whose row lists hold r, columns 0..p of supernode s when r
//   is at position p of its list; a row index at supernode granularity
//   (NOFSUB segments, built on the host once) gives them in that order, each
//   segment with its slot base, F, first column and length, so that a lane
//   needs one memory round trip per segment.
// One lane per column or row adds the products, each rounded before it is
// added, in the reference's order, so C equals SPRTMM's bit for bit; every
// stored entry is used, fill included.  The loads of eight entries are issued
// before their products are added.
#include <algorithm>

#include "sprgpu_internal.h"

namespace sprgpu {

constexpr int kTmmBatch = 8;
constexpr int kTmmLongCol = 256;  // columns at least this long: a wavefront each

__global__ __launch_bounds__(256) void k_tmm_cols(const i32 *__restrict__ col_node,
                                                  const NodeInfo *__restrict__ nodes,
                                                  const i32 *__restrict__ lindx,
                                                  const double *__restrict__ u,
                                                  const double *__restrict__ b, double *c,
                                                  double *wa, i64 n, i64 ld, int write_c,
                                                  int nrhs)
{
#pragma clang fp contract(off)
    const i64 j = i64(blockIdx.x) * 256 + threadIdx.x;
    const int rhs = blockIdx.y;
    if (j >= n) return;
    const NodeInfo nd = nodes[col_node[j]];
    const i64 k = j - nd.first;
    const i64 len = nd.F - k;
    if (len >= kTmmLongCol) return;  // a wave of k_tmm_cols_long sums it
    const double *uc = u + nd.xlnz + k * nd.F - k * (k - 1) / 2;
    const i32 *rows = lindx + nd.lxoff + k;
    const double *bb = b + rhs * ld;
    double s = 0.0;
    for (i64 q0 = 0; q0 < len; q0 += kTmmBatch) {
        double uv[kTmmBatch], bv[kTmmBatch];
#pragma unroll
        for (int q = 0; q < kTmmBatch; ++q) {
            const i64 t = min(q0 + q, len - 1);
            uv[q] = uc[t];
            bv[q] = bb[rows[t]];
        }
#pragma unroll
        for (int q = 0; q < kTmmBatch; ++q)
            if (q0 + q < len) {
                const double p = uv[q] * bv[q];
                s = s + p;
            }
    }
    if (write_c) c[rhs * ld + j] = s;
    if (wa && rhs == nrhs - 1) wa[j] = s;
}

// Columns of kTmmLongCol entries or more, one wavefront each: the lanes load
// 64 consecutive entries (coalesced) and form their products, the next 64 in
// flight, and the running sum takes the 64 products one by one in column
// order (read from their lanes), so it is the lane kernel's sum exactly.
__global__ __launch_bounds__(256) void k_tmm_cols_long(const i32 *__restrict__ cols, i64 ncols,
                                                       const i32 *__restrict__ col_node,
                                                       const NodeInfo *__restrict__ nodes,
                                                       const i32 *__restrict__ lindx,
                                                       const double *__restrict__ u,
                                                       const double *__restrict__ b, double *c,
                                                       double *wa, i64 ld, int write_c, int nrhs)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const i64 idx = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int rhs = blockIdx.y;
    if (idx >= ncols) return;  // whole wavefront
    const i64 j = cols[idx];
    const NodeInfo nd = nodes[col_node[j]];
    const i64 k = j - nd.first;
    const i64 len = nd.F - k;
    const double *uc = u + nd.xlnz + k * nd.F - k * (k - 1) / 2;
    const i32 *rows = lindx + nd.lxoff + k;
    const double *bb = b + rhs * ld;
    auto product = [&](i64 t) { return t < len ? uc[t] * bb[rows[t]] : 0.0; };
    double nxt = product(lane);
    double s = 0.0;
    for (i64 q0 = 0; q0 < len; q0 += 64) {
        const double pr = nxt;
        if (q0 + 64 < len) nxt = product(q0 + 64 + lane);
        const i64 nq = min<i64>(64, len - q0);
        const unsigned lo = unsigned(__double_as_longlong(pr)),
                       hi = unsigned(__double_as_longlong(pr) >> 32);
#pragma unroll
        for (int i = 0; i < 64; ++i) {
            if (i < nq) {
                const unsigned long long v =
                    (unsigned long long)(unsigned)__builtin_amdgcn_readlane(int(hi), i) << 32 |
                    (unsigned)__builtin_amdgcn_readlane(int(lo), i);
                s = s + __longlong_as_double((long long)v);
            }
        }
    }
    if (lane == 0) {
        if (write_c) c[rhs * ld + j] = s;
        if (wa && rhs == nrhs - 1) wa[j] = s;
    }
}

__global__ __launch_bounds__(256) void k_tmm_rows(const i64 *__restrict__ rptr,
                                                  const TmmSeg *__restrict__ segs,
                                                  const double *__restrict__ u,
                                                  const double *__restrict__ b, double *c,
                                                  double *wa, i64 n, i64 ld, int write_c,
                                                  int nrhs)
{
#pragma clang fp contract(off)
    const i64 r = i64(blockIdx.x) * 256 + threadIdx.x;
    const int rhs = blockIdx.y;
    if (r >= n) return;
    const double *bb = b + rhs * ld;
    double s = 0.0;
    const i64 e1 = rptr[r + 1];
    i64 e = rptr[r];
    TmmSeg nxt = e < e1 ? segs[e] : TmmSeg{};
    for (; e < e1; ++e) {
        const TmmSeg sg = nxt;
        if (e + 1 < e1) nxt = segs[e + 1];  // the next segment's descriptor in flight
        // column k's entry of this row: base + k*(F-1) - k*(k-1)/2
        for (i64 k0 = 0; k0 < sg.kend; k0 += kTmmBatch) {
            double uv[kTmmBatch], bv[kTmmBatch];
#pragma unroll
            for (int q = 0; q < kTmmBatch; ++q) {
                const i64 k = min<i64>(k0 + q, sg.kend - 1);
                uv[q] = u[sg.base + k * (sg.F - 1) - k * (k - 1) / 2];
                bv[q] = bb[sg.first + k];
            }
#pragma unroll
            for (int q = 0; q < kTmmBatch; ++q)
                if (k0 + q < sg.kend) {
                    const double pr = uv[q] * bv[q];
                    s = s + pr;
                }
        }
    }
    if (write_c) c[rhs * ld + r] = s;
    if (wa && rhs == nrhs - 1) wa[r] = s;
}

// KSU <= 0: U is a diagonal held as a vector, C(J) = U(J)*B(J)
__global__ __launch_bounds__(256) void k_tmm_diag(const double *__restrict__ u,
                                                  const double *__restrict__ b, double *c,
                                                  double *wa, i64 n, i64 ld, int write_c,
                                                  int nrhs)
{
    const i64 j = i64(blockIdx.x) * 256 + threadIdx.x;
    const int rhs = blockIdx.y;
    if (j >= n) return;
    const double s = u[j] * b[rhs * ld + j];
    if (write_c) c[rhs * ld + j] = s;
    if (wa && rhs == nrhs - 1) wa[j] = s;
}

static int upload_row_index(sprgpu_handle *h)
{
    DeviceBuffers &d = h->d;
    if (d.tmm_rptr) return 0;
    const Plan &p = h->plan;
    std::vector<i64> rptr(p.neq + 1, 0);
    for (i64 k = 0; k < p.nofsub; ++k) rptr[p.lindx[k] + 1]++;
    for (i64 r = 0; r < p.neq; ++r) rptr[r + 1] += rptr[r];
    std::vector<TmmSeg> segs(std::max<i64>(p.nofsub, 1));
    std::vector<i64> cur(rptr.begin(), rptr.end() - 1);
    for (i64 s = 0; s < p.nsuper; ++s) {  // supernodes ascending: each row's list in order
        const NodeInfo &n = p.nodes[s];
        for (i32 q = 0; q < n.F; ++q) {
            // row lindx[q] meets columns 0..min(S-1, q) of this supernode
            segs[cur[p.lindx[n.lxoff + q]]++] =
                TmmSeg{n.xlnz + q, n.F, n.first, std::min(n.S, q + 1), 0};
        }
    }
    SPRGPU_HIP(hipMalloc(&d.tmm_rptr, sizeof(i64) * rptr.size()));
    SPRGPU_HIP(hipMalloc(&d.tmm_segs, sizeof(TmmSeg) * segs.size()));
    SPRGPU_HIP(hipMemcpy(d.tmm_rptr, rptr.data(), sizeof(i64) * rptr.size(),
                         hipMemcpyHostToDevice));
    SPRGPU_HIP(hipMemcpy(d.tmm_segs, segs.data(), sizeof(TmmSeg) * segs.size(),
                         hipMemcpyHostToDevice));
    h->device_bytes += sizeof(i64) * rptr.size() + sizeof(TmmSeg) * p.nofsub;
    return 0;
}

int run_sprtmm(sprgpu_handle *h, const double *u, const double *b, double *c, double *wa,
               i64 m, i64 n, i64 ksu, i64 iflag)
{
    const Plan &p = h->plan;
    const i64 neq = p.neq;
    const bool only_wa = (iflag == 11 || iflag == 22) && n == 1;
    if (!only_wa && iflag != 1 && iflag != 2) {
        set_error("sprtmm: IFLAG must be 1, 2, or 11/22 with N = 1");
        return -2;
    }
    if (m < neq) {
        set_error("sprtmm: M < NEQ (retained equations are not supported)");
        return -3;
    }
    if (n < 1 || (only_wa && !wa) || (!only_wa && !c)) {
        set_error("sprtmm: N must be positive; C (or WA for IFLAG 11/22) is required");
        return -2;
    }
    const bool upper = (iflag == 1 || iflag == 11);
    DeviceBuffers &d = h->d;
    if (ksu > 0) {
        if (upper) {
            if (int rc = upload_col_node(h)) return rc;
            if (!d.tmm_long) {  // the columns of kTmmLongCol entries or more
                std::vector<i32> lc;
                for (i64 sn = 0; sn < p.nsuper; ++sn)
                    for (i32 k = 0; k < p.nodes[sn].S; ++k)
                        if (p.nodes[sn].F - k >= kTmmLongCol) lc.push_back(p.nodes[sn].first + k);
                SPRGPU_HIP(hipMalloc(&d.tmm_long, sizeof(i32) * std::max<size_t>(lc.size(), 1)));
                if (!lc.empty())
                    SPRGPU_HIP(hipMemcpy(d.tmm_long, lc.data(), sizeof(i32) * lc.size(),
                                         hipMemcpyHostToDevice));
                d.tmm_nlong = i64(lc.size());
                h->device_bytes += sizeof(i32) * lc.size();
            }
        } else if (int rc = upload_row_index(h)) {
            return rc;
        }
    }
    hipStream_t s = h->stream;
    const unsigned gx = unsigned(std::max<i64>(1, (neq + 255) / 256));
    for (i64 r0 = 0; r0 < n; r0 += 65535) {
        const i64 nr = std::min<i64>(65535, n - r0);
        const dim3 g(gx, unsigned(nr));
        double *cc = only_wa ? nullptr : c + r0 * m;
        double *wc = (r0 + nr == n) ? wa : nullptr;
        const int wr = only_wa ? 0 : 1;
        if (ksu <= 0)
            k_tmm_diag<<<g, 256, 0, s>>>(u, b + r0 * m, cc, wc, neq, m, wr, int(nr));
        else if (upper) {
            k_tmm_cols<<<g, 256, 0, s>>>(d.col_node, d.nodes, d.lindx, u, b + r0 * m, cc, wc,
                                         neq, m, wr, int(nr));
            if (d.tmm_nlong > 0)
                k_tmm_cols_long<<<dim3(unsigned((d.tmm_nlong + 3) / 4), unsigned(nr)), 256, 0,
                                  s>>>(d.tmm_long, d.tmm_nlong, d.col_node, d.nodes, d.lindx, u,
                                       b + r0 * m, cc, wc, m, wr, int(nr));
        }
        else
            k_tmm_rows<<<g, 256, 0, s>>>(d.tmm_rptr, d.tmm_segs, u, b + r0 * m, cc, wc, neq, m,
                                         wr, int(nr));
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// SPRTRA's dot products (sprtra.f:40-46): C(I, J) = C(J, I) = B(:, I) . WA
// for I = J..N-1, one workgroup per I.  Each thread sums its strided rows in
// order (eight loads in flight), the workgroup's 256 sums are added by a
// fixed tree: the result does not depend on scheduling.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tra_dots(const double *__restrict__ b, i64 ldb,
                                                  const double *__restrict__ wa, i64 m, i64 j,
                                                  double *c, i64 ldc)
{
    __shared__ double part[256];
    const i64 i = j + blockIdx.x;
    const double *bi = b + i * ldb;
    const int t = threadIdx.x;
    double s = 0.0;
    i64 r = t;
    for (; r + 7 * 256 < m; r += 8 * 256) {
        double x[8], y[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            x[u] = bi[r + u * 256];
            y[u] = wa[r + u * 256];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s += x[u] * y[u];
    }
    for (; r < m; r += 256) s += bi[r] * wa[r];
    part[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) part[t] += part[t + w];
        __syncthreads();
    }
    if (t == 0) {
        c[i + j * ldc] = part[0];
        c[j + i * ldc] = part[0];
    }
}

int run_tra_dots(sprgpu_handle *h, const double *b, i64 m, const double *wa, i64 n, i64 j,
                 double *c)
{
    k_tra_dots<<<unsigned(n - j), 256, 0, h->stream>>>(b, m, wa, m, j, c, n);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

}  // namespace sprgpu
