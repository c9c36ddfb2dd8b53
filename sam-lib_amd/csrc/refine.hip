// FP64 residual SpMV and iterative refinement.
//
// The FP64 SpMV plays the role of the reference's SPRPRM product
// (sprprm.f:243-324: C = A B on the unfactored matrix) for the refinement
// residual r = b - A x; it runs on a CSR copy of the assembled matrix's
// non-zeros (the pattern is built once per handle, the values are gathered
// from SM at each factorization).  Refinement: x = A^-1 b with the current
// factor (FP32 or FP64), then r = b - A x in FP64 and x += A^-1 r until the
// relative residual of every column is at most the requested tolerance.
#include <cfloat>
#include <cmath>

#include "sprgpu_internal.h"

namespace sprgpu {

// ---------------------------------------------------------------------------
// FP64 residual: y = alpha*b + beta*A x (CSR, 16 lanes per row)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_spmv(const i64 *__restrict__ rowptr,
                                              const i32 *__restrict__ col,
                                              const double *__restrict__ val,
                                              const double *__restrict__ x, double *y,
                                              const double *__restrict__ bvec, i64 n, i64 ld,
                                              double beta, int nrhs)
{
    // the right-hand side varies fastest over the workgroups, so the rows'
    // matrix entries are read once from memory for all of them
    const i64 row = (i64(blockIdx.x / nrhs) * 256 + threadIdx.x) >> 4;
    const int sub = threadIdx.x & 15;
    const int rhs = blockIdx.x % nrhs;
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

// Many right-hand sides: y = b + beta A X as a product with X stored
// equation-major (row i holds x_i of every right-hand side, contiguous), so
// that every matrix entry is read once per row for all right-hand sides and
// every x row read is one coalesced run.  The caller's column-major vectors
// are transposed in and out through LDS tiles.
constexpr int kTr = 32;  // transpose tile edge
// dst(i, r) = src[r * lds + i] into dst[i * nrhs + r], i < n, r < nrhs
__global__ __launch_bounds__(256) void k_to_eq_major(const double *__restrict__ src, i64 lds,
                                                     double *__restrict__ dst, i64 n, int nrhs)
{
    __shared__ double t[kTr][kTr + 1];
    const i64 i0 = i64(blockIdx.x) * kTr;
    const int r0 = blockIdx.y * kTr;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
    for (int k = 0; k < kTr; k += 8) {
        const int r = r0 + ty + k;
        const i64 i = i0 + tx;
        t[ty + k][tx] = (r < nrhs && i < n) ? src[r * lds + i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTr; k += 8) {
        const i64 i = i0 + ty + k;
        const int r = r0 + tx;
        if (i < n && r < nrhs) dst[i * nrhs + r] = t[tx][ty + k];
    }
}

// y[r * ld + i] = b[r * ld + i] + beta * src[i * nrhs + r] (b may be null)
__global__ __launch_bounds__(256) void k_from_eq_major(const double *__restrict__ src,
                                                       double *__restrict__ y,
                                                       const double *__restrict__ b, i64 ld,
                                                       double beta, i64 n, int nrhs)
{
    __shared__ double t[kTr][kTr + 1];
    const i64 i0 = i64(blockIdx.x) * kTr;
    const int r0 = blockIdx.y * kTr;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int k = 0; k < kTr; k += 8) {
        const i64 i = i0 + ty + k;
        const int r = r0 + tx;
        t[ty + k][tx] = (i < n && r < nrhs) ? src[i * nrhs + r] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTr; k += 8) {
        const int r = r0 + ty + k;
        const i64 i = i0 + tx;
        if (r < nrhs && i < n) {
            const double b0 = b ? b[r * ld + i] : 0.0;
            y[r * ld + i] = b0 + beta * t[tx][ty + k];
        }
    }
}

// yt[i * nrhs + r] = sum_k val[k] xt[col[k] * nrhs + r] over row i: one
// wavefront per row, the lanes over the right-hand sides (two each, 128 at a
// time, when nrhs is even; else one each, 64 at a time), the row's entries in
// order with eight loads in flight.  Rows go to the XCDs in contiguous
// ranges (block ids are dealt to the XCDs in turn), so that the rows that
// share x rows (a mesh's neighbours) share an L2.
constexpr int kSpmmBatch = 8;
__global__ __launch_bounds__(256) void k_spmm_rows(const i64 *__restrict__ rowptr,
                                                   const i32 *__restrict__ col,
                                                   const double *__restrict__ val,
                                                   const double *__restrict__ xt,
                                                   double *__restrict__ yt, i64 n, int nrhs)
{
    const i64 nb = gridDim.x, bid = blockIdx.x;
    const i64 per = nb / kXcds, rem = nb % kXcds;
    const i64 xcd = bid % kXcds, idx = bid / kXcds;
    const i64 lin = (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    const i64 row = lin * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const i64 k0 = rowptr[row], k1 = rowptr[row + 1];
    if ((nrhs & 1) == 0) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        for (int r0 = 0; r0 < nrhs; r0 += 128) {
            const int r = min(r0 + 2 * lane, nrhs - 2);
            d2 s = d2{0.0, 0.0};
            for (i64 kb = k0; kb < k1; kb += kSpmmBatch) {
                double a[kSpmmBatch];
                d2 xv[kSpmmBatch];
#pragma unroll
                for (int q = 0; q < kSpmmBatch; ++q) {
                    const i64 k = min(kb + q, k1 - 1);
                    a[q] = (kb + q < k1) ? val[k] : 0.0;
                    xv[q] = *reinterpret_cast<const d2 *>(xt + i64(col[k]) * nrhs + r);
                }
#pragma unroll
                for (int q = 0; q < kSpmmBatch; ++q) s += a[q] * xv[q];
            }
            if (r0 + 2 * lane < nrhs) *reinterpret_cast<d2 *>(yt + row * nrhs + r) = s;
        }
        return;
    }
    for (int r0 = 0; r0 < nrhs; r0 += 64) {
        const int r = min(r0 + lane, nrhs - 1);
        double s = 0.0;
        for (i64 kb = k0; kb < k1; kb += kSpmmBatch) {
            double a[kSpmmBatch], xv[kSpmmBatch];
#pragma unroll
            for (int q = 0; q < kSpmmBatch; ++q) {
                const i64 k = min(kb + q, k1 - 1);
                a[q] = (kb + q < k1) ? val[k] : 0.0;
                xv[q] = xt[i64(col[k]) * nrhs + r];
            }
#pragma unroll
            for (int q = 0; q < kSpmmBatch; ++q) s += a[q] * xv[q];
        }
        if (r0 + lane < nrhs) yt[row * nrhs + r0 + lane] = s;
    }
}

// gathers the CSR values from the packed assembled matrix and counts the
// non-zeros found on or below the diagonal (each packed entry once); sixteen
// lanes per row, so that a row's index and value reads are coalesced and
// many gathers are in flight
constexpr int kGatherLanes = 16;
__global__ __launch_bounds__(256) void k_gather_vals(const i64 *__restrict__ rowptr,
                                                     const i32 *__restrict__ col,
                                                     const i64 *__restrict__ src,
                                                     const double *__restrict__ lnz,
                                                     double *val, i64 n,
                                                     unsigned long long *count)
{
    unsigned long long c = 0;
    const int g = threadIdx.x % kGatherLanes;
    const i64 stride = i64(gridDim.x) * (256 / kGatherLanes);
    for (i64 row = i64(blockIdx.x) * (256 / kGatherLanes) + threadIdx.x / kGatherLanes; row < n;
         row += stride) {
        const i64 k1 = rowptr[row + 1];
        for (i64 k = rowptr[row] + g; k < k1; k += kGatherLanes) {
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
    const i64 rows_per_wg = 256 / kGatherLanes;
    const unsigned gwg =
        unsigned(std::max<i64>(1, std::min<i64>((h->plan.neq + rows_per_wg - 1) / rows_per_wg, 16384)));
    k_gather_vals<<<gwg, 256, 0, h->stream>>>(d.csr_rowptr, d.csr_col, d.csr_src,
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

constexpr i64 kSpmmMinRhs = 16;  // from here on the equation-major product

// y = b + beta A x for nrhs columns (b may be null): up to kSpmmMinRhs - 1
// columns one product per column (in chunks of right-hand sides that keep a
// launch within 2^23 workgroups), from kSpmmMinRhs on the equation-major
// product (a work space of 2 n nrhs doubles, grown on demand)
static int launch_spmv(sprgpu_handle *h, const double *x, double *y, const double *b, i64 ld,
                       i64 nrhs, double beta)
{
    const i64 n = h->plan.neq;
    DeviceBuffers &d = h->d;
    if (nrhs >= kSpmmMinRhs && nrhs < 65535 * i64(kTr) && !h->noalloc) {
        const i64 need = 2 * n * nrhs;
        if (d.spmm_cap < need) {
            ++h->buf_gen;
            if (d.spmm_work) SPRGPU_HIP(hipFree(d.spmm_work));
            d.spmm_work = nullptr;
            SPRGPU_HIP(hipMalloc(&d.spmm_work, sizeof(double) * need));
            h->device_bytes += sizeof(double) * (need - d.spmm_cap);
            d.spmm_cap = need;
        }
        double *xt = d.spmm_work, *yt = d.spmm_work + n * nrhs;
        const dim3 gt(unsigned((n + kTr - 1) / kTr), unsigned((nrhs + kTr - 1) / kTr));
        k_to_eq_major<<<gt, 256, 0, h->stream>>>(x, ld, xt, n, int(nrhs));
        k_spmm_rows<<<unsigned((n + 3) / 4), 256, 0, h->stream>>>(d.csr_rowptr, d.csr_col,
                                                                 d.csr_val, xt, yt, n, int(nrhs));
        k_from_eq_major<<<gt, 256, 0, h->stream>>>(yt, y, b, ld, beta, n, int(nrhs));
        SPRGPU_HIP(hipGetLastError());
        return 0;
    }
    const i64 nb = (n * 16 + 255) / 256;
    const i64 chunk = std::max<i64>(1, std::min<i64>(nrhs, (i64(1) << 23) / nb));
    for (i64 r0 = 0; r0 < nrhs; r0 += chunk) {
        const i64 nr = std::min(chunk, nrhs - r0);
        k_spmv<<<unsigned(nb * nr), 256, 0, h->stream>>>(
            h->d.csr_rowptr, h->d.csr_col, h->d.csr_val, x + r0 * ld, y + r0 * ld,
            b ? b + r0 * ld : nullptr, n, ld, beta, int(nr));
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// SPRPRM on the device (sprprm.f:5-238, its product loop SPRPRQ at :243-326)
// ---------------------------------------------------------------------------
// SPRPRQ walks the packed columns in order and adds ALFA*(a_ij*b) to C(i)
// and C(j) as it meets each entry, starting from C = 0.  Each C(r) therefore
// receives the entries of row r of the symmetric A in ascending column order:
// a_rj for j < r from the columns before r, then a_rr and the a_ir, i > r,
// of column r.  The CSR pattern holds every row in that order, so one lane
// per row reproduces the reference's sums bit for bit (entries outside the
// pattern are zero and add nothing).
constexpr int kPrmBatch = 8;  // entries whose loads are issued together
__global__ __launch_bounds__(256) void k_prm_rows(const i64 *__restrict__ rowptr,
                                                  const i32 *__restrict__ col,
                                                  const i64 *__restrict__ src,
                                                  const double *__restrict__ a,
                                                  const double *__restrict__ b, double *c,
                                                  double *wa, i64 n, i64 ld, double alfa,
                                                  int accumulate, int write_c, int nrhs,
                                                  unsigned long long *count)
{
    // product rounded, then added, as the reference's separate operations
#pragma clang fp contract(off)
    const i64 row = i64(blockIdx.x) * 256 + threadIdx.x;
    const int rhs = blockIdx.y;
    unsigned long long nz = 0;
    if (row < n) {
        const double *bb = b + rhs * ld;
        double s = 0.0;
        const i64 k0 = rowptr[row], k1 = rowptr[row + 1];
        for (i64 kb = k0; kb < k1; kb += kPrmBatch) {
            // the batch's loads first (clamped to the row's last entry), then
            // the products added one by one in the row's order
            double av[kPrmBatch], bv[kPrmBatch];
            i32 cv[kPrmBatch];
#pragma unroll
            for (int q = 0; q < kPrmBatch; ++q) {
                const i64 k = min(kb + q, k1 - 1);
                cv[q] = col[k];
                av[q] = a[src[k]];
                bv[q] = bb[cv[q]];
            }
#pragma unroll
            for (int q = 0; q < kPrmBatch; ++q) {
                if (kb + q < k1) {
                    const double p = av[q] * bv[q];
                    s = s + alfa * p;
                    nz += (cv[q] <= row && av[q] != 0.0);
                }
            }
        }
        if (write_c) {
            double *cc = c + rhs * ld + row;
            *cc = accumulate ? *cc + s : s;
        }
        if (wa && rhs == nrhs - 1) wa[row] = s;
    }
    // non-zeros of A on the pattern's lower entries (first column's blocks)
    if (count && rhs == 0) {
        for (int o = 32; o > 0; o >>= 1) nz += __shfl_down(nz, o);
        if ((threadIdx.x & 63) == 0 && nz) atomicAdd(count, nz);
    }
}

// KSA <= 0: A is a diagonal held as a vector, C(j) = ALFA*(A(j)*B(j))
__global__ __launch_bounds__(256) void k_prm_diag(const double *__restrict__ a,
                                                  const double *__restrict__ b, double *c,
                                                  double *wa, i64 n, i64 ld, double alfa,
                                                  int accumulate, int write_c, int nrhs)
{
#pragma clang fp contract(off)
    const i64 row = i64(blockIdx.x) * 256 + threadIdx.x;
    const int rhs = blockIdx.y;
    if (row >= n) return;
    const double p = a[row] * b[rhs * ld + row];
    const double s = alfa * p;
    if (write_c) {
        double *cc = c + rhs * ld + row;
        *cc = accumulate ? *cc + s : s;
    }
    if (wa && rhs == nrhs - 1) wa[row] = s;
}

// non-zeros of a packed matrix, 16-byte loads over its aligned part
__global__ __launch_bounds__(256) void k_count_nonzero(const double *__restrict__ a, i64 n,
                                                       unsigned long long *count)
{
    unsigned long long c = 0;
    const i64 t = i64(blockIdx.x) * 256 + threadIdx.x, stride = i64(gridDim.x) * 256;
    const i64 head = (reinterpret_cast<uintptr_t>(a) & 15) ? 1 : 0;
    const i64 n2 = (n - head) / 2;
    const double2 *q = reinterpret_cast<const double2 *>(a + head);
    for (i64 i = t; i < n2; i += stride) {
        const double2 v = q[i];
        c += (v.x != 0.0) + (v.y != 0.0);
    }
    if (t == 0) {
        if (head && n > 0) c += (a[0] != 0.0);
        if (head + 2 * n2 < n) c += (a[n - 1] != 0.0);
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

int run_sprprm(sprgpu_handle *h, const double *a, const double *b, double *c, double *wa,
               i64 m, i64 n, i64 ksa, i64 iflag)
{
    DeviceBuffers &d = h->d;
    const i64 neq = h->plan.neq;
    const bool only_wa = (iflag == 88 && n == 1);
    if (!only_wa && iflag != 1 && iflag != -1 && iflag != 2 && iflag != -2) {
        set_error("sprprm: IFLAG must be +-1, +-2, or 88 with N = 1");
        return -2;
    }
    const bool peri2e = h->mspar.size() >= 58 && h->mspar[57] >= 1;
    if (ksa <= 0 && peri2e) {
        set_error("sprprm: a diagonal A (KSA <= 0) with the PERI2E permutation");
        return -3;
    }
    if (m < neq) {
        set_error("sprprm: M < NEQ (retained equations are not supported)");
        return -4;
    }
    if (n < 1 || (only_wa && !wa) || (!only_wa && !c)) {
        set_error("sprprm: N must be positive; C (or WA for IFLAG 88) is required");
        return -2;
    }
    hipStream_t s = h->stream;
    const double alfa = iflag < 0 ? -1.0 : 1.0;
    const int accumulate = (iflag == 2 || iflag == -2);
    if (ksa > 0) {
        if (!h->has_pattern) {
            set_error("sprprm: no pattern; call sprgpu_set_pattern or sprgpu_set_assembly");
            return -1;
        }
        // every non-zero of A must lie on the pattern: A's non-zeros are
        // counted here and those on the pattern by the product kernel
        if (!d.prm_count) SPRGPU_HIP(hipMalloc(&d.prm_count, 2 * sizeof(unsigned long long)));
        SPRGPU_HIP(hipMemsetAsync(d.prm_count, 0, 2 * sizeof(unsigned long long), s));
        const i64 nz = h->plan.nzerol;
        k_count_nonzero<<<unsigned(std::max<i64>(1, std::min<i64>((nz / 2 + 255) / 256, 8192))),
                          256, 0, s>>>(a, nz, d.prm_count);
    }
    const unsigned gx = unsigned((neq + 255) / 256);
    for (i64 r0 = 0; r0 < n; r0 += 65535) {
        const i64 nr = std::min<i64>(65535, n - r0);
        dim3 g(gx, unsigned(nr));
        double *wa_chunk = (r0 + nr == n) ? wa : nullptr;
        if (ksa > 0)
            k_prm_rows<<<g, 256, 0, s>>>(d.csr_rowptr, d.csr_col, d.csr_src, a, b + r0 * m,
                                         only_wa ? nullptr : c + r0 * m, wa_chunk, neq, m, alfa,
                                         accumulate, only_wa ? 0 : 1, int(nr),
                                         r0 == 0 ? d.prm_count + 1 : nullptr);
        else
            k_prm_diag<<<g, 256, 0, s>>>(a, b + r0 * m, only_wa ? nullptr : c + r0 * m,
                                         wa_chunk, neq, m, alfa, accumulate, only_wa ? 0 : 1,
                                         int(nr));
    }
    SPRGPU_HIP(hipGetLastError());
    if (ksa > 0) {
        unsigned long long cnt[2];
        SPRGPU_HIP(hipMemcpyAsync(cnt, d.prm_count, sizeof(cnt), hipMemcpyDeviceToHost, s));
        SPRGPU_HIP(hipStreamSynchronize(s));
        if (cnt[0] != cnt[1]) {
            set_error("sprprm: A has non-zeros outside the pattern (C is not valid); call "
                      "sprgpu_set_pattern with A");
            return -1;
        }
    }
    return 0;
}

int run_spmv(sprgpu_handle *h, const double *x, double *y, i64 ld, i64 nrhs)
{
    if (int rc = refresh_values(h)) return rc;
    if (int rc = launch_spmv(h, x, y, nullptr, ld, nrhs, 1.0)) return rc;
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

    int it = 0;
    double worst = 0.0;
    for (;;) {
        // r = b - A x
        if (ld != n) {
            set_error("solve_ir requires ld == neq");
            return -1;
        }
        if (int rc = launch_spmv(h, x, r, b, n, nrhs, -1.0)) return rc;
        if (int rc = col_norms(h, r, n, n, nrhs, part, nrm, rnorm)) return rc;
        if (int rc = check_solve_error(h)) return rc;
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
