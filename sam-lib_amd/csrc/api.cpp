// C ABI of libsprgpu: the device-resident handle API and the drop-in
// replacements of the reference's numeric SPR routines.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "sprgpu_internal.h"

namespace sprgpu {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return -5;
}

template <typename X>
static int upload(X *&dst, const std::vector<X> &src, sprgpu_int &bytes)
{
    const size_t n = std::max<size_t>(src.size(), 1);
    SPRGPU_HIP(hipMalloc(&dst, sizeof(X) * n));
    if (!src.empty())
        SPRGPU_HIP(hipMemcpy(dst, src.data(), sizeof(X) * src.size(), hipMemcpyHostToDevice));
    bytes += sizeof(X) * n;
    return 0;
}

static void free_all(sprgpu_handle *h)
{
    DeviceBuffers &d = h->d;
    void *ptrs[] = {d.nodes,    d.children,   d.lindx,       d.relmap,     d.init_tasks,
                    d.copy_tasks,
                    d.ea_tasks, d.small_nodes, d.panel_desc, d.panel_prefix, d.gemm_desc,
                    d.gemm_prefix, d.pool,     d.dvec,       d.adiag,      d.dcor,
                    d.status,   d.stats,      d.vwork,       d.csr_rowptr, d.csr_col,
                    d.dscr,
                    d.csr_src,  d.csr_val,    d.irwork,      d.level_nodes,
                    d.solve_small, d.wide,     d.wide_fwd_prefix, d.wide_bwd_prefix,
                    d.wide_rect_prefix, d.upd_ptr, d.upd_src, d.flags, d.partials,
                    d.spin_error, d.col_node};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    d = DeviceBuffers{};
}

}  // namespace sprgpu

using namespace sprgpu;

extern "C" {

const char *sprgpu_last_error(void) { return g_last_error.c_str(); }

int sprgpu_create(const sprgpu_int *mspar, const sprgpu_int *mtrees, const sprgpu_int *msifa,
                  int precision, sprgpu_handle **handle)
{
    if (!mspar || !mtrees || !msifa || !handle) {
        set_error("null argument");
        return -2;
    }
    if (precision != SPRGPU_FP64 && precision != SPRGPU_FP32) {
        set_error("precision must be SPRGPU_FP64 or SPRGPU_FP32");
        return -2;
    }
    std::unique_ptr<sprgpu_handle> h(new sprgpu_handle);
    h->precision = precision;
    std::string err = build_plan(mspar, mtrees, msifa, precision, h->plan);
    if (!err.empty()) {
        set_error(err);
        return -2;
    }
    SPRGPU_HIP(hipGetDevice(&h->device));
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    sprgpu_int &B = h->device_bytes;
    int rc = 0;
    if ((rc = upload(d.nodes, p.nodes, B)) || (rc = upload(d.children, p.children, B)) ||
        (rc = upload(d.lindx, p.lindx, B)) || (rc = upload(d.relmap, p.relmap, B)) ||
        (rc = upload(d.init_tasks, p.init_tasks, B)) || (rc = upload(d.ea_tasks, p.ea_tasks, B)) ||
        (rc = upload(d.copy_tasks, p.copy_tasks, B)) ||
        (rc = upload(d.small_nodes, p.small_nodes, B)) ||
        (rc = upload(d.panel_desc, p.panel_desc, B)) ||
        (rc = upload(d.panel_prefix, p.panel_prefix, B)) ||
        (rc = upload(d.gemm_desc, p.gemm_desc, B)) ||
        (rc = upload(d.gemm_prefix, p.gemm_prefix, B)) ||
        (rc = upload(d.level_nodes, p.level_nodes, B)) ||
        (rc = upload(d.solve_small, p.solve_small, B)) ||
        (rc = upload(d.wide, p.wide, B)) ||
        (rc = upload(d.wide_fwd_prefix, p.wide_fwd_prefix, B)) ||
        (rc = upload(d.wide_bwd_prefix, p.wide_bwd_prefix, B)) ||
        (rc = upload(d.wide_rect_prefix, p.wide_rect_prefix, B)) ||
        (rc = upload(d.upd_ptr, p.upd_ptr, B)) || (rc = upload(d.upd_src, p.upd_src, B))) {
        free_all(h.get());
        return rc;
    }
    const size_t esz = precision == SPRGPU_FP32 ? sizeof(float) : sizeof(double);
    hipError_t e = hipMalloc(&d.pool, esz * size_t(p.pool_elems));
    if (e == hipSuccess) e = hipMalloc(&d.dvec, sizeof(double) * p.neq);
    if (e == hipSuccess) e = hipMalloc(&d.adiag, sizeof(double) * p.neq);
    if (e == hipSuccess) e = hipMalloc(&d.dcor, sizeof(double) * 2 * p.neq);
    if (e == hipSuccess) e = hipMalloc(&d.status, sizeof(FactorStatus));
    if (e == hipSuccess) e = hipMalloc(&d.stats, sizeof(double) * 4 * 257);
    if (e == hipSuccess) e = hipMalloc(&d.dscr, esz * size_t(p.neq) * kPanelNB);
    if (e != hipSuccess) {
        free_all(h.get());
        return hip_fail(e, "device allocation");
    }
    if (hipMemset(d.dcor, 0, sizeof(double) * 2 * p.neq) != hipSuccess ||
        hipMemset(d.status, 0, sizeof(FactorStatus)) != hipSuccess) {
        free_all(h.get());
        return hip_fail(hipGetLastError(), "device memset");
    }
    B += esz * p.pool_elems + sizeof(double) * 4 * p.neq + sizeof(FactorStatus) + 8 * 4 * 257;
    *handle = h.release();
    return 0;
}

int sprgpu_destroy(sprgpu_handle *h)
{
    if (!h) return 0;
    (void)hipDeviceSynchronize();
    free_all(h);
    delete h;
    return 0;
}

int sprgpu_set_stream(sprgpu_handle *h, void *stream)
{
    if (!h) return -2;
    h->stream = static_cast<hipStream_t>(stream);
    return 0;
}

int sprgpu_set_pattern(sprgpu_handle *h, const double *sm_host)
{
    if (!h || !sm_host) return -2;
    CsrPattern pat;
    build_csr_pattern(h->plan, sm_host + h->plan.neq, pat);
    return upload_pattern(h, pat);
}

int sprgpu_factor(sprgpu_handle *h, const double *sm_dev, const double *tol, int mode,
                  sprgpu_factor_info *info)
{
    if (!h || !sm_dev || !tol) return -2;
    if (mode != 1 && mode != 3) {
        set_error("mode must be 1 (SPRFC1) or 3 (SPRFC3)");
        return -2;
    }
    double t4[4] = {tol[0], 0.0, 0.0, mode == 3 ? tol[3] : 0.0};
    int rc = run_factor(h, sm_dev, mode, t4);
    if (rc) return rc;
    if (info) return read_factor_info(h, info);
    return 0;
}

int sprgpu_profile_factor(sprgpu_handle *h, const double *sm_dev, const double *tol, int mode,
                          double *out)
{
    if (!h || !sm_dev || !tol || !out || (mode != 1 && mode != 3)) return -2;
    double t4[4] = {tol[0], 0.0, 0.0, mode == 3 ? tol[3] : 0.0};
    return profile_factor(h, sm_dev, mode, t4, out);
}

int sprgpu_factor_info_get(sprgpu_handle *h, sprgpu_factor_info *info)
{
    if (!h || !info) return -2;
    if (!h->factored) {
        set_error("no factorization");
        return -1;
    }
    return read_factor_info(h, info);
}

int sprgpu_solve(sprgpu_handle *h, double *b, sprgpu_int ldb, sprgpu_int nrhs, int which)
{
    if (!h || !b || ldb < h->plan.neq || which < 1 || which > 3) return -2;
    return run_solve(h, b, ldb, nrhs, which);
}

int sprgpu_spmv(sprgpu_handle *h, const double *x, double *y, sprgpu_int ld, sprgpu_int nrhs)
{
    if (!h || !x || !y || ld < h->plan.neq) return -2;
    return run_spmv(h, x, y, ld, nrhs);
}

int sprgpu_solve_ir(sprgpu_handle *h, const double *b, double *x, sprgpu_int ld,
                    sprgpu_int nrhs, double rtol, int maxit, sprgpu_ir_info *info)
{
    if (!h || !b || !x || ld < h->plan.neq) return -2;
    return run_solve_ir(h, b, x, ld, nrhs, rtol, maxit, info);
}

int sprgpu_element_grad(sprgpu_handle *h, sprgpu_int nel, sprgpu_int nedof,
                        const sprgpu_int *elem_eq, const double *k0, const sprgpu_int *k0_index,
                        const double *lambda, const double *x, double *grad)
{
    if (!h || !elem_eq || !k0 || !lambda || !x || !grad || nel < 0) return -2;
    return run_element_grad(h, nel, int(nedof), elem_eq, k0, k0_index, lambda, x, grad);
}

int sprgpu_pattern_grad(sprgpu_handle *h, const double *lambda, const double *x, double *grad_sm)
{
    if (!h || !lambda || !x || !grad_sm) return -2;
    return run_pattern_grad(h, lambda, x, grad_sm + h->plan.neq);
}

int sprgpu_get_factor(sprgpu_handle *h, double *sm_out)
{
    if (!h || !sm_out) return -2;
    if (!h->factored) {
        set_error("no factorization");
        return -1;
    }
    hipPointerAttribute_t attr;
    bool on_device = false;
    if (hipPointerGetAttributes(&attr, sm_out) == hipSuccess)
        on_device = (attr.type == hipMemoryTypeDevice);
    (void)hipGetLastError();
    const Plan &p = h->plan;
    const size_t bytes = sizeof(double) * size_t(p.neq + p.nzerol);
    if (on_device) return copy_factor_to_sm(h, sm_out);
    double *tmp = nullptr;
    SPRGPU_HIP(hipMalloc(&tmp, bytes));
    // start from the input so that entries the factor does not own keep it
    if (h->last_sm_dev)
        SPRGPU_HIP(hipMemcpyAsync(tmp, h->last_sm_dev, bytes, hipMemcpyDeviceToDevice, h->stream));
    int rc = copy_factor_to_sm(h, tmp);
    if (rc == 0) {
        hipError_t e = hipMemcpyAsync(sm_out, tmp, bytes, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = hip_fail(e, "copy factor to host");
    }
    (void)hipFree(tmp);
    return rc;
}

int sprgpu_plan_summary(const sprgpu_int *mspar, const sprgpu_int *mtrees,
                        const sprgpu_int *msifa, double *out)
{
    if (!mspar || !mtrees || !msifa || !out) return -2;
    Plan p;
    std::string err = build_plan(mspar, mtrees, msifa, SPRGPU_FP64, p);
    if (!err.empty()) {
        set_error(err);
        return -2;
    }
    i64 tiles = 0, steps = 0;
    for (const Launch &L : p.launches) {
        if (L.kind == L_GEMM) tiles += L.blocks;
        if (L.kind == L_PANEL) steps += L.count;
    }
    out[0] = double(p.neq);
    out[1] = double(p.nsuper);
    out[2] = double(p.nlevels);
    out[3] = double(p.launches.size());
    out[4] = double(p.pool_elems);
    out[5] = double(p.vwork_elems);
    out[6] = double(p.small_nodes.size());
    out[7] = double(steps);
    out[8] = double(tiles);
    out[9] = p.factor_flops;
    return 0;
}

sprgpu_int sprgpu_neq(const sprgpu_handle *h) { return h ? h->plan.neq : 0; }
sprgpu_int sprgpu_nzerol(const sprgpu_handle *h) { return h ? h->plan.nzerol : 0; }
sprgpu_int sprgpu_device_bytes(const sprgpu_handle *h) { return h ? h->device_bytes : 0; }

}  // extern "C"
