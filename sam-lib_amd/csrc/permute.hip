// The PERI2E permutation of the SPR layout on the device.
//
// When the symbolic phase reserved it (MSPAR(58) >= 1, sprsmb.f:181) and
// SPRABC filled it (sprabc.f:5-66), MSIFA(1:NEQ) = PERI2E maps each internal
// (factor) equation to the external equation the caller numbers its vectors
// by.  The reference's permuted routines read and write the caller's vectors
// through it: SPRFB3/SPRFB4 in the solves (sprfs1.f:777-926, used by SPRFS4
// and SPRBS5, sprfs1.f:396-528, sprbs1.f:371-495), SPRPRR (sprprm.f:327-397)
// and SPRTMR (sprtmm.f:345-480).  Each of them sums every output with the
// same terms in the same order as its unpermuted sibling, only addressed
// through PERI2E, so here the caller's vectors are gathered into internal
// order, the unpermuted kernels run, and the results are scattered back:
// the results equal the permuted routines' bit for bit.
//
// Work space: the handle's permwork buffer, two blocks of NEQ x NRHS
// (grown on demand; a growth bumps the handle's buffer generation).
#include "sprgpu_internal.h"

namespace sprgpu {

// dst(i, c) = src(perm[i], c), i < n, c < nrhs
__global__ __launch_bounds__(256) void k_perm_gather(const i32 *__restrict__ perm,
                                                     const double *__restrict__ src, i64 lds,
                                                     double *__restrict__ dst, i64 ldd, i64 n)
{
    const i64 c = blockIdx.y;
    for (i64 i = blockIdx.x * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256)
        dst[c * ldd + i] = src[c * lds + perm[i]];
}

// dst(perm[i], c) = src(i, c)
__global__ __launch_bounds__(256) void k_perm_scatter(const i32 *__restrict__ perm,
                                                      const double *__restrict__ src, i64 lds,
                                                      double *__restrict__ dst, i64 ldd, i64 n)
{
    const i64 c = blockIdx.y;
    for (i64 i = blockIdx.x * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256)
        dst[c * ldd + perm[i]] = src[c * lds + i];
}

// The leading equations of a partial solve (MSPAR(57) = 1, LDB < NEQ): the
// reference gathers zero for an equation beyond LDB and never writes it back
// (SPRFB1/SPRFB3 and SPRFB2/SPRFB4, sprfs1.f:636-926).  e(i) is PERI2E(i)
// when perm is set, else i.
__global__ __launch_bounds__(256) void k_lead_gather(const i32 *__restrict__ perm,
                                                     const double *__restrict__ src, i64 lds,
                                                     double *__restrict__ dst, i64 ldd, i64 n,
                                                     i64 nlead)
{
    const i64 c = blockIdx.y;
    for (i64 i = blockIdx.x * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256) {
        const i64 e = perm ? perm[i] : i;
        dst[c * ldd + i] = e < nlead ? src[c * lds + e] : 0.0;
    }
}

__global__ __launch_bounds__(256) void k_lead_mask(const i32 *__restrict__ perm,
                                                   double *__restrict__ x, i64 ldx, i64 n,
                                                   i64 nlead)
{
    const i64 c = blockIdx.y;
    for (i64 i = blockIdx.x * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256) {
        const i64 e = perm ? perm[i] : i;
        if (e >= nlead) x[c * ldx + i] = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_lead_scatter(const i32 *__restrict__ perm,
                                                      const double *__restrict__ src, i64 lds,
                                                      double *__restrict__ dst, i64 ldd, i64 n,
                                                      i64 nlead)
{
    const i64 c = blockIdx.y;
    for (i64 i = blockIdx.x * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256) {
        const i64 e = perm ? perm[i] : i;
        if (e < nlead) dst[c * ldd + e] = src[c * lds + i];
    }
}

int upload_permutation(sprgpu_handle *h, const i64 *msifa)
{
    const i64 neq = h->plan.neq;
    std::vector<i32> perm(neq);
    std::vector<char> seen(neq, 0);
    for (i64 i = 0; i < neq; ++i) {
        const i64 e = msifa[i];
        if (e < 1 || e > neq || seen[e - 1]) {
            set_error("MSIFA(1:NEQ) is not a permutation (MSPAR(58) >= 1 needs SPRABC's PERI2E)");
            return -2;
        }
        seen[e - 1] = 1;
        perm[i] = i32(e - 1);
    }
    h->d.peri2e_pmax.assign(neq + 1, -1);
    for (i64 i = 0; i < neq; ++i) h->d.peri2e_pmax[i + 1] = std::max(h->d.peri2e_pmax[i], perm[i]);
    SPRGPU_HIP(hipMalloc(&h->d.peri2e, sizeof(i32) * neq));
    SPRGPU_HIP(hipMemcpy(h->d.peri2e, perm.data(), sizeof(i32) * neq, hipMemcpyHostToDevice));
    h->device_bytes += sizeof(i32) * neq;
    return 0;
}

int perm_work(sprgpu_handle *h, i64 nrhs, double **w0, double **w1)
{
    DeviceBuffers &d = h->d;
    const i64 need = 2 * h->plan.neq * nrhs;
    if (h->noalloc && d.permwork_cap < need) {
        set_error("permutation work space larger than the scratch view");
        return -2;
    }
    if (d.permwork_cap < need) {
        ++h->buf_gen;
        if (d.permwork) SPRGPU_HIP(hipFree(d.permwork));
        d.permwork = nullptr;
        SPRGPU_HIP(hipMalloc(&d.permwork, sizeof(double) * need));
        h->device_bytes += sizeof(double) * (need - d.permwork_cap);
        d.permwork_cap = need;
    }
    *w0 = d.permwork;
    *w1 = d.permwork + h->plan.neq * nrhs;
    return 0;
}

static dim3 perm_grid(i64 n, i64 nrhs)
{
    return dim3(unsigned(std::max<i64>(1, std::min<i64>((n + 255) / 256, 2048))),
                unsigned(nrhs));
}

int perm_gather(sprgpu_handle *h, const double *src, i64 lds, double *dst, i64 ldd, i64 nrhs)
{
    const i64 n = h->plan.neq;
    for (i64 c0 = 0; c0 < nrhs; c0 += 65535) {
        const i64 nc = std::min<i64>(65535, nrhs - c0);
        k_perm_gather<<<perm_grid(n, nc), 256, 0, h->stream>>>(h->d.peri2e, src + c0 * lds, lds,
                                                               dst + c0 * ldd, ldd, n);
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int lead_gather(sprgpu_handle *h, const double *src, i64 lds, double *dst, i64 ldd, i64 nrhs,
                i64 nlead)
{
    const i64 n = h->plan.neq;
    for (i64 c0 = 0; c0 < nrhs; c0 += 65535) {
        const i64 nc = std::min<i64>(65535, nrhs - c0);
        k_lead_gather<<<perm_grid(n, nc), 256, 0, h->stream>>>(h->d.peri2e, src + c0 * lds, lds,
                                                               dst + c0 * ldd, ldd, n, nlead);
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int lead_mask(sprgpu_handle *h, double *x, i64 ldx, i64 nrhs, i64 nlead)
{
    const i64 n = h->plan.neq;
    for (i64 c0 = 0; c0 < nrhs; c0 += 65535) {
        const i64 nc = std::min<i64>(65535, nrhs - c0);
        k_lead_mask<<<perm_grid(n, nc), 256, 0, h->stream>>>(h->d.peri2e, x + c0 * ldx, ldx, n,
                                                             nlead);
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int lead_scatter(sprgpu_handle *h, const double *src, i64 lds, double *dst, i64 ldd, i64 nrhs,
                 i64 nlead)
{
    const i64 n = h->plan.neq;
    for (i64 c0 = 0; c0 < nrhs; c0 += 65535) {
        const i64 nc = std::min<i64>(65535, nrhs - c0);
        k_lead_scatter<<<perm_grid(n, nc), 256, 0, h->stream>>>(h->d.peri2e, src + c0 * lds,
                                                                lds, dst + c0 * ldd, ldd, n,
                                                                nlead);
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int perm_scatter(sprgpu_handle *h, const double *src, i64 lds, double *dst, i64 ldd, i64 nrhs)
{
    const i64 n = h->plan.neq;
    for (i64 c0 = 0; c0 < nrhs; c0 += 65535) {
        const i64 nc = std::min<i64>(65535, nrhs - c0);
        k_perm_scatter<<<perm_grid(n, nc), 256, 0, h->stream>>>(h->d.peri2e, src + c0 * lds, lds,
                                                                dst + c0 * ldd, ldd, n);
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

}  // namespace sprgpu
