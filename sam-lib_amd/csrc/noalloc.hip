// sprgpu_solve_noalloc: the forward and/or backward substitution of
// sprgpu_solve with its work space in a static device array instead of
// buffers the handle allocates.
//
// For callers that cannot allocate device memory during a call: code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
// The solve kernels use, besides the factor and the caller's B, a vector
// slot per supernode per right-hand side (the forward sweep's update
// vectors and the wide supernodes' block results), the backward sweep's
// partial sums, a spin-error word and, with PERI2E, the permuted copy of B.  Here all of them are views of `scratch`, a file-scope array
// that every GPU the library runs on holds until the process ends
// (kScratchBytes); its device address is looked up once per device when a
// handle is created there.  Calls that do not fit return -2 and allocate
// nothing.
//
// All calls on one GPU share the array: two of them that overlap on that
// GPU (different streams, threads or handles) race.  The array keeps the
// last call's contents; every sweep marks its block result slots pending
// before it starts, and the spin-error word is cleared by a kernel at the
// start of every call.
#include <mutex>

#include "sprgpu_internal.h"

namespace sprgpu {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr size_t kScratchAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch_init()
{
    int dev = 0;
    SPRGPU_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index beyond the scratch table");
        return -2;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return 0;
    void *ptr = nullptr;
    SPRGPU_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
    return 0;
}

__global__ void k_clear_word(int *w)
{
    if (threadIdx.x == 0) *w = 0;
}

static size_t align_up(size_t x) { return (x + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }

}  // namespace sprgpu

using namespace sprgpu;

extern "C" int sprgpu_solve_noalloc(sprgpu_handle *h, double *b, sprgpu_int ldb, sprgpu_int nrhs,
                                    int which)
{
    if (!h || !b || ldb < h->plan.neq || which < 1 || which > 3 || nrhs < 0) return -2;
    if (nrhs == 0) return 0;
    if (!h->factored) {
        set_error("solve called before a successful factorization");
        return -1;
    }
    if (nrhs > 65535) {
        set_error("at most 65535 right-hand sides per call");
        return -2;
    }
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (h->device >= 0 && h->device < kMaxDevices) base = g_scratch_ptr[h->device];
    }
    if (!base) {
        set_error("no scratch address for this device (handle created elsewhere)");
        return -5;
    }
    const Plan &p = h->plan;
    // layout: spin error | partial sums | vector slots | PERI2E copies
    const size_t o_err = 0;
    const size_t o_part = align_up(o_err + sizeof(int));
    const i64 npart = std::max<i64>(p.npart, 1) * nrhs;
    const size_t o_vw = align_up(o_part + sizeof(double) * size_t(npart));
    const i64 nvw = p.vwork_elems * nrhs;
    const size_t o_perm = align_up(o_vw + sizeof(double) * size_t(nvw));
    const i64 nperm = h->d.peri2e ? 2 * p.neq * nrhs : 0;
    const size_t total = o_perm + sizeof(double) * size_t(nperm);
    if (total > kScratchBytes) {
        set_error("sprgpu_solve_noalloc: this solve needs " + std::to_string(total) +
                  " bytes of work space, the scratch holds " + std::to_string(kScratchBytes) +
                  " (use fewer right-hand sides per call, or sprgpu_solve)");
        return -2;
    }
    DeviceBuffers &d = h->d;
    DeviceBuffers saved = d;
    d.spin_error = reinterpret_cast<int *>(base + o_err);
    d.partials = reinterpret_cast<double *>(base + o_part);
    d.partials_cap = npart;
    d.vwork = reinterpret_cast<double *>(base + o_vw);
    d.vwork_cap = nrhs;
    d.permwork = nperm ? reinterpret_cast<double *>(base + o_perm) : nullptr;
    d.permwork_cap = nperm;
    h->noalloc = true;
    k_clear_word<<<1, 64, 0, h->stream>>>(d.spin_error);
    const hipError_t e = hipGetLastError();
    int rc = (e == hipSuccess) ? sprgpu_solve(h, b, ldb, nrhs, which) : hip_fail(e, "k_clear_word");
    h->noalloc = false;
    d.spin_error = saved.spin_error;
    d.partials = saved.partials;
    d.partials_cap = saved.partials_cap;
    d.vwork = saved.vwork;
    d.vwork_cap = saved.vwork_cap;
    d.permwork = saved.permwork;
    d.permwork_cap = saved.permwork_cap;
    return rc;
}
