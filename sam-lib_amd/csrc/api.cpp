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
                    d.copy_tasks, d.store_tasks,
                    d.ea_tasks, d.small_nodes, d.panel_desc, d.panel_prefix, d.gemm_desc,
                    d.gemm_prefix, d.pool,     d.store,      d.dvec,       d.adiag,      d.dcor,
                    d.status,   d.stats,      d.vwork,       d.csr_rowptr, d.csr_col,
                    d.dscr,     d.dinv,       d.pub,
                    d.csr_src,  d.csr_val,    d.irwork,      d.level_nodes,
                    d.solve_small, d.wide,     d.wide_fwd_prefix, d.wide_bwd_prefix,
                    d.wide_rect_prefix, d.wide_frect_prefix, d.upd_ptr, d.upd_src, d.partials,
                    d.spin_error, d.col_node, d.prm_count,
                    d.tmm_rptr, d.tmm_segs, d.tmm_long, d.peri2e, d.permwork,
                    d.forest_nodes, d.forest_child, d.forest_chunks, d.forest_order, d.ublk,
                    d.forest_sync, d.spmm_work};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    free_assembly(h);
    d = DeviceBuffers{};
}


// Device copies of everything the planner derives (descriptor and index
// arrays); replaced as a whole when the plan is rebuilt.
static void free_plan_buffers(sprgpu_handle *h)
{
    DeviceBuffers &d = h->d;
    void **ptrs[] = {(void **)&d.nodes,         (void **)&d.children,
                     (void **)&d.lindx,         (void **)&d.relmap,
                     (void **)&d.init_tasks,    (void **)&d.copy_tasks,
                     (void **)&d.store_tasks,
                     (void **)&d.ea_tasks,      (void **)&d.small_nodes,
                     (void **)&d.panel_desc,    (void **)&d.panel_prefix,
                     (void **)&d.gemm_desc,     (void **)&d.gemm_prefix,
                     (void **)&d.level_nodes,   (void **)&d.solve_small,
                     (void **)&d.wide,          (void **)&d.wide_fwd_prefix,
                     (void **)&d.wide_bwd_prefix, (void **)&d.wide_rect_prefix,
                     (void **)&d.wide_frect_prefix,
                     (void **)&d.upd_ptr,       (void **)&d.upd_src,
                     (void **)&d.col_node,      (void **)&d.tmm_rptr,
                     (void **)&d.tmm_segs,      (void **)&d.tmm_long,
                     (void **)&d.forest_nodes,  (void **)&d.forest_child,
                     (void **)&d.forest_chunks, (void **)&d.forest_order};
    for (void **q : ptrs) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
}

// The fused outer-block launches' publication areas (kPubElems per
// descriptor of the plan), sized for the current plan.
static int ensure_pub(sprgpu_handle *h)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    bool fused = false;
    for (const Launch &L : p.launches) fused |= (L.kind == L_BFUSED);
    const i64 need = fused ? i64(p.panel_desc.size()) * kPubElems : 0;
    const i64 esz = h->precision == SPRGPU_FP32 ? i64(sizeof(float)) : i64(sizeof(double));
    if (need > d.pub_cap) {
        if (d.pub) SPRGPU_HIP(hipFree(d.pub));
        h->device_bytes -= esz * d.pub_cap;
        d.pub = nullptr;
        d.pub_cap = 0;
        SPRGPU_HIP(hipMalloc(&d.pub, size_t(esz * need)));
        h->device_bytes += esz * need;
        d.pub_cap = need;
    }
    d.pub_elems = need;
    return 0;
}

// The persistent forest schedule's update blocks and counters, sized for the
// current plan (kept when a later plan needs no more).
static int ensure_forest(sprgpu_handle *h)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    if (!p.forest) return 0;
    const i64 esz = h->precision == SPRGPU_FP32 ? i64(sizeof(float)) : i64(sizeof(double));
    if (!d.ublk) {
        SPRGPU_HIP(hipMalloc(&d.ublk, size_t(esz * std::max<i64>(p.ublk_elems, 1))));
        SPRGPU_HIP(hipMalloc(&d.forest_sync, sizeof(unsigned) * size_t(2 + 3 * p.nsuper)));
        h->device_bytes += esz * std::max<i64>(p.ublk_elems, 1) + 4 * (2 + 3 * p.nsuper);
    }
    return 0;
}

static int upload_plan(sprgpu_handle *h)
{
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    sprgpu_int &B = h->device_bytes;
    int rc = 0;
    if ((rc = upload(d.nodes, p.nodes, B)) || (rc = upload(d.children, p.children, B)) ||
        (rc = upload(d.lindx, p.lindx, B)) || (rc = upload(d.relmap, p.relmap, B)) ||
        (rc = upload(d.init_tasks, p.init_tasks, B)) || (rc = upload(d.ea_tasks, p.ea_tasks, B)) ||
        (rc = upload(d.copy_tasks, p.copy_tasks, B)) ||
        (rc = upload(d.store_tasks, p.store_tasks, B)) ||
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
        (rc = upload(d.wide_frect_prefix, p.wide_frect_prefix, B)) ||
        (rc = upload(d.upd_ptr, p.upd_ptr, B)) || (rc = upload(d.upd_src, p.upd_src, B)) ||
        (rc = upload(d.forest_nodes, p.forest_nodes, B)) ||
        (rc = upload(d.forest_child, p.forest_child, B)) ||
        (rc = upload(d.forest_chunks, p.forest_chunks, B)) ||
        (rc = upload(d.forest_order, p.forest_order, B)))
        return rc;
    return 0;
}

static void drop_graph(sprgpu_handle *h)
{
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    h->graph_exec = nullptr;
    h->graph_sm = nullptr;
    h->graph_b = nullptr;
}

// Rebuilds the plan with new parameters (same symbolic input, same data
// layout) and replaces the device descriptor arrays.
static int replan(sprgpu_handle *h, const PlanParams &params)
{
    Plan np;
    std::string err = build_plan(h->mspar.data(), h->mtrees.data(), h->msifa.data(), h->precision,
                                 params, np);
    if (!err.empty()) {
        set_error(err);
        return -2;
    }
    if (np.pool_elems != h->plan.pool_elems || np.store_elems != h->plan.store_elems) {
        set_error("re-planning changed the front pool or the factor store");  // (not expected)
        return -2;
    }
    SPRGPU_HIP(hipStreamSynchronize(h->stream));
    drop_graph(h);
    free_plan_buffers(h);
    h->plan = std::move(np);
    if (int rc = upload_plan(h)) return rc;
    if (int rc = ensure_forest(h)) return rc;
    return ensure_pub(h);
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
    std::string err = build_plan(mspar, mtrees, msifa, precision, PlanParams{}, h->plan);
    if (!err.empty()) {
        set_error(err);
        return -2;
    }
    {
        // the parts of the symbolic arrays the planner reads (to re-plan later)
        auto M = [&](int k) { return mspar[k - 1]; };
        const sprgpu_int ns = M(11);
        const sprgpu_int nt = std::max(M(45) - 1 + 3 * (ns + 1), M(46) - 1 + 2 * ns + 1);
        const sprgpu_int nf = std::max(M(47) - 1 + M(15), M(48) - 1 + ns + 1);
        h->mspar.assign(mspar, mspar + 58);
        h->mtrees.assign(mtrees, mtrees + nt);
        h->msifa.assign(msifa, msifa + nf);
    }
    SPRGPU_HIP(hipGetDevice(&h->device));
    if (int rc = noalloc_scratch_init()) return rc;  // once per device
    const Plan &p = h->plan;
    DeviceBuffers &d = h->d;
    if (int rc = upload_plan(h.get())) {
        free_all(h.get());
        return rc;
    }
    if (mspar[57] >= 1) {  // PERI2E = MSIFA(1:NEQ) (sprsmb.f:181, sprabc.f:45-66)
        if (int rc = upload_permutation(h.get(), msifa)) {
            free_all(h.get());
            return rc;
        }
    }
    sprgpu_int &B = h->device_bytes;
    const size_t esz = precision == SPRGPU_FP32 ? sizeof(float) : sizeof(double);
    hipError_t e = hipMalloc(&d.pool, esz * size_t(p.pool_elems));
    if (e == hipSuccess) e = hipMalloc(&d.store, esz * size_t(p.store_elems));
    if (e == hipSuccess) e = hipMalloc(&d.dvec, sizeof(double) * p.neq);
    if (e == hipSuccess) e = hipMalloc(&d.adiag, sizeof(double) * p.neq);
    if (e == hipSuccess) e = hipMalloc(&d.dcor, sizeof(double) * 2 * p.neq);
    if (e == hipSuccess) e = hipMalloc(&d.status, sizeof(FactorStatus));
    if (e == hipSuccess) e = hipMalloc(&d.stats, sizeof(double) * 4 * 257);
    if (e == hipSuccess) e = hipMalloc(&d.dscr, esz * size_t(p.neq) * kPanelNB);
    const size_t ninv = size_t(std::max<i64>(p.nflags, 1)) * kBlockSolveData;
    if (e == hipSuccess) e = hipMalloc(&d.dinv, esz * ninv);
    if (e != hipSuccess) {
        free_all(h.get());
        return hip_fail(e, "device allocation");
    }
    if (int rc = ensure_pub(h.get())) {
        free_all(h.get());
        return rc;
    }
    if (int rc = ensure_forest(h.get())) {
        free_all(h.get());
        return rc;
    }
    if (hipMemset(d.dcor, 0, sizeof(double) * 2 * p.neq) != hipSuccess ||
        hipMemset(d.status, 0, sizeof(FactorStatus)) != hipSuccess) {
        free_all(h.get());
        return hip_fail(hipGetLastError(), "device memset");
    }
    B += esz * (p.pool_elems + p.store_elems) + sizeof(double) * 4 * p.neq + sizeof(FactorStatus) + 8 * 4 * 257 +
         esz * (size_t(p.neq) * kPanelNB + ninv);
    // the factorization's second stream and its two events (default priority)
    if (hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev[1], hipEventDisableTiming) != hipSuccess) {
        const hipError_t e = hipGetLastError();
        for (hipEvent_t &x : h->ev)
            if (x) (void)hipEventDestroy(x);
        if (h->stream2) (void)hipStreamDestroy(h->stream2);
        free_all(h.get());
        return hip_fail(e, "stream/event creation");
    }
    *handle = h.release();
    return 0;
}

int sprgpu_destroy(sprgpu_handle *h)
{
    if (!h) return 0;
    (void)hipDeviceSynchronize();
    drop_graph(h);
    if (h->capture_stream) (void)hipStreamDestroy(h->capture_stream);
    for (hipEvent_t &x : h->ev)
        if (x) (void)hipEventDestroy(x);
    if (h->stream2) (void)hipStreamDestroy(h->stream2);
    free_all(h);
    delete h;
    return 0;
}

int sprgpu_factor_solve_graph(sprgpu_handle *h, const double *sm_dev, double *b_dev,
                              sprgpu_int ldb, sprgpu_int nrhs)
{
    if (!h || !sm_dev || !b_dev || nrhs < 1 || ldb < h->plan.neq) return -2;
    const bool same = h->graph_exec && h->graph_sm == sm_dev && h->graph_b == b_dev &&
                      h->graph_ldb == ldb && h->graph_nrhs == nrhs &&
                      h->graph_gen == h->buf_gen;
    if (!same) {
        drop_graph(h);
        const double tol[4] = {0.0, 0.0, 0.0, 0.0};
        // every buffer the sequence needs exists before capturing (the
        // factor's are allocated with the handle)
        if (int rc = prepare_solve(h, nrhs)) return rc;
        double *w0 = nullptr, *w1 = nullptr;
        if (h->d.peri2e)
            if (int rc = perm_work(h, nrhs, &w0, &w1)) return rc;
        SPRGPU_HIP(hipStreamSynchronize(h->stream));
        if (!h->capture_stream)
            SPRGPU_HIP(hipStreamCreateWithFlags(&h->capture_stream, hipStreamNonBlocking));
        hipStream_t user = h->stream;
        h->stream = h->capture_stream;
        hipGraph_t g = nullptr;
        SPRGPU_HIP(hipStreamBeginCapture(h->capture_stream, hipStreamCaptureModeThreadLocal));
        // a forest of small fronts: factor and solve in one persistent launch
        const bool fused = h->plan.forest && nrhs == 1 && !h->d.peri2e;
        int rc = fused ? run_factor_solve(h, sm_dev, 1, tol, b_dev) : run_factor(h, sm_dev, 1, tol);
        if (rc == 0 && !fused && !h->d.peri2e) rc = run_solve(h, b_dev, ldb, nrhs, 3);
        if (rc == 0 && h->d.peri2e) {  // B in external order
            if (!(rc = perm_gather(h, b_dev, ldb, w0, h->plan.neq, nrhs)) &&
                !(rc = run_solve(h, w0, h->plan.neq, nrhs, 3)))
                rc = perm_scatter(h, w0, h->plan.neq, b_dev, ldb, nrhs);
        }
        const hipError_t e = hipStreamEndCapture(h->capture_stream, &g);
        h->stream = user;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        const hipError_t ei = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            h->graph_exec = nullptr;
            return hip_fail(ei, "hipGraphInstantiate");
        }
        h->graph_sm = sm_dev;
        h->graph_b = b_dev;
        h->graph_ldb = ldb;
        h->graph_nrhs = nrhs;
        h->graph_gen = h->buf_gen;
    }
    SPRGPU_HIP(hipGraphLaunch(h->graph_exec, h->stream));
    h->factored = true;
    h->last_sm_dev = sm_dev;
    return 0;
}

int sprgpu_autotune(sprgpu_handle *h, const double *sm_dev, double *b_dev, sprgpu_int ldb,
                    sprgpu_int nrhs, double *report)
{
    if (!h || !sm_dev || !b_dev || nrhs < 1 || ldb < h->plan.neq) return -2;
    std::vector<PlanParams> cand;
    // The level schedule's parameters (FP32 tries the per-launch tile choice
    // and the 128 tile everywhere), then, for a structure whose fronts all
    // have F <= kSmallF, the persistent forest schedule (at most 37
    // candidates: the report's 256 doubles hold 42)
    const int tiles[2] = {0, kTileBig};
    for (int sf : {32, 48, 64})
        for (int w : {1, 2, 4})
            for (int nt : {64, 256})
                for (int t = 0; t < (h->precision == SPRGPU_FP32 ? 2 : 1); ++t) {
                    PlanParams pp;
                    pp.small_f = sf;
                    pp.small_waves = w;
                    pp.narrow_threads = nt;
                    pp.big_tile = (h->precision == SPRGPU_FP32) ? tiles[t] : 0;
                    pp.forest = 0;
                    cand.push_back(pp);
                }
    bool small_fronts = h->plan.mfront <= kSmallF;
    for (const NodeInfo &n : h->plan.nodes) small_fronts &= n.F <= kSmallF;
    if (small_fronts) cand.push_back(PlanParams{});
    const double tol[4] = {0.0, 0.0, 0.0, 0.0};
    hipEvent_t e0, e1;
    SPRGPU_HIP(hipEventCreate(&e0));
    SPRGPU_HIP(hipEventCreate(&e1));
    double best = 1e300;
    size_t ibest = 0;
    int rc = 0;
    for (size_t i = 0; i < cand.size() && rc == 0; ++i) {
        if ((rc = replan(h, cand[i]))) break;
        if ((rc = run_factor(h, sm_dev, 1, tol)) || (rc = run_solve(h, b_dev, ldb, nrhs, 3))) break;
        float ms = 0.f;
        (void)hipEventRecord(e0, h->stream);
        for (int k = 0; k < 3 && rc == 0; ++k)
            if (!(rc = run_factor(h, sm_dev, 1, tol))) rc = run_solve(h, b_dev, ldb, nrhs, 3);
        (void)hipEventRecord(e1, h->stream);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = hip_fail(hipGetLastError(), "autotune timing");
            break;
        }
        const double t = ms / 3.0;
        if (report && 6 * i + 6 < 256) {
            report[1 + 6 * i] = cand[i].small_f;
            report[2 + 6 * i] = cand[i].small_waves;
            report[3 + 6 * i] = cand[i].narrow_threads;
            report[4 + 6 * i] = cand[i].big_tile;
            report[5 + 6 * i] = cand[i].forest;
            report[6 + 6 * i] = t;
        }
        if (t < best) {
            best = t;
            ibest = i;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (report) report[0] = double(cand.size());
    if (rc) return rc;
    return replan(h, cand[ibest]);
}

int sprgpu_plan_params(const sprgpu_handle *h, int *out)
{
    if (!h || !out) return -2;
    out[0] = h->plan.params.small_f;
    out[1] = h->plan.params.small_waves;
    out[2] = h->plan.params.narrow_threads;
    out[3] = h->plan.big_tile;
    out[4] = h->plan.params.block_factor;
    out[5] = h->plan.forest ? 1 : 0;  // the forest schedule in use
    return 0;
}

int sprgpu_set_plan_params(sprgpu_handle *h, const int *params)
{
    if (!h || !params) return -2;
    PlanParams pp;
    pp.small_f = params[0];
    pp.small_waves = params[1];
    pp.narrow_threads = params[2];
    pp.big_tile = params[3];
    pp.block_factor = params[4];
    pp.forest = params[5];
    if (pp.forest != 0 && pp.forest != 1) {
        set_error("forest must be 0 or 1");
        return -2;
    }
    return replan(h, pp);
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

int sprgpu_set_assembly(sprgpu_handle *h, sprgpu_int nel, const sprgpu_int *meen_ptr,
                        const sprgpu_int *meen)
{
    if (!h || nel < 0 || (nel > 0 && (!meen_ptr || !meen))) return -2;
    return set_assembly(h, nel, meen_ptr, meen, nullptr);
}

int sprgpu_set_assembly_ceq(sprgpu_handle *h, sprgpu_int nel, const sprgpu_int *meen_ptr,
                            const sprgpu_int *meen, sprgpu_int nceq, const sprgpu_int *mpmceq,
                            const sprgpu_int *mmceq, sprgpu_int ndof, const sprgpu_int *meqn,
                            sprgpu_int nsv)
{
    if (!h || nel < 0 || (nel > 0 && (!meen_ptr || !meen))) return -2;
    CeqInput c;
    c.nceq = nceq;
    c.mpmceq = mpmceq;
    c.mmceq = mmceq;
    c.ndof = ndof;
    c.meqn = meqn;
    c.nsv = nsv;
    return set_assembly(h, nel, meen_ptr, meen, &c);
}

int sprgpu_assemble(sprgpu_handle *h, const double *em_dev, const sprgpu_int *em_off_dev,
                    const double *escale_dev, double *sm_dev, int accumulate)
{
    if (!h || !sm_dev || (!em_dev && h->d.asm_nel > 0)) return -2;
    return run_assemble(h, em_dev, em_off_dev, escale_dev, nullptr, sm_dev, nullptr, accumulate);
}

int sprgpu_hex8_stiffness(sprgpu_handle *h, sprgpu_int nel, sprgpu_int nanod,
                          const double *xyz_dev, const sprgpu_int *mmnpc_dev,
                          const double *emod_dev, double nu, double *em_dev)
{
    if (!h) return -2;
    return hex8_stiffness(h, nel, nanod, xyz_dev, mmnpc_dev, emod_dev, nu, em_dev);
}

int sprgpu_assemble_ceq(sprgpu_handle *h, const double *em_dev, const sprgpu_int *em_off_dev,
                        const double *escale_dev, const double *ttcc_dev, double *sm_dev,
                        double *sv_dev, int accumulate)
{
    if (!h || (!em_dev && h->d.asm_nel > 0)) return -2;
    return run_assemble(h, em_dev, em_off_dev, escale_dev, ttcc_dev, sm_dev, sv_dev, accumulate);
}

int sprgpu_factor(sprgpu_handle *h, const double *sm_dev, const double *tol, int mode,
                  sprgpu_factor_info *info)
{
    if (!h || !sm_dev || !tol) return -2;
    if (mode < 1 || mode > 4) {
        set_error("mode must be 1 (SPRFC1), 2 (SPRFC2), 3 (SPRFC3) or 4 (SPRFC4)");
        return -2;
    }
    // SPRFC2: the LDL' factorization with every pivot computed, then its
    // test; SPRFC4: the LDL' factorization with its pivot rule, C = L sqrt(D)
    double t4[4] = {mode == 2 ? 0.0 : tol[0], 0.0, 0.0, mode >= 3 ? tol[3] : 0.0};
    int rc = run_factor(h, sm_dev, mode == 2 ? 1 : mode, t4);
    if (rc) return rc;
    if (mode == 2) {
        h->mode = 2;
        h->cc_tol1 = tol[0];
    }
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

int sprgpu_get_dcor(sprgpu_handle *h, double *dcor_out)
{
    if (!h || !dcor_out) return -2;
    if (!h->factored || h->mode < 3) {
        set_error("DCOR exists after a mode-3 (SPRFC3) or mode-4 (SPRFC4) factorization only");
        return -1;
    }
    SPRGPU_HIP(hipMemcpyAsync(dcor_out, h->d.dcor, sizeof(double) * 2 * size_t(h->plan.neq),
                              hipMemcpyDefault, h->stream));
    SPRGPU_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

// The solves in internal equation order.
static int solve_internal(sprgpu_handle *h, double *b, sprgpu_int ldb, sprgpu_int nrhs,
                          int which)
{
    if ((h->mode == 2 || h->mode == 4) && which != 3) {
        // C^-1 b = sqrt(D) D^-1 L^-1 b; C^-T y = L^-T D^-1/2 y
        if (which == 1) {
            if (int rc = run_solve(h, b, ldb, nrhs, 1)) return rc;
            return cc_scale(h, b, ldb, nrhs, 0);
        }
        if (int rc = cc_scale(h, b, ldb, nrhs, 1)) return rc;
        return run_solve(h, b, ldb, nrhs, 2);
    }
    return run_solve(h, b, ldb, nrhs, which);
}

int sprgpu_solve(sprgpu_handle *h, double *b, sprgpu_int ldb, sprgpu_int nrhs, int which)
{
    if (!h || !b || ldb < h->plan.neq || which < 1 || which > 3) return -2;
    if (!h->d.peri2e || nrhs <= 0) return solve_internal(h, b, ldb, nrhs, which);
    // B in external order (SPRFS4 / SPRBS5 with SPRFB3 / SPRFB4)
    if (!h->factored) {
        set_error("solve called before a successful factorization");
        return -1;
    }
    const i64 neq = h->plan.neq;
    double *w0, *w1;
    if (int rc = perm_work(h, nrhs, &w0, &w1)) return rc;
    if (int rc = perm_gather(h, b, ldb, w0, neq, nrhs)) return rc;
    if (int rc = solve_internal(h, w0, neq, nrhs, which)) return rc;
    return perm_scatter(h, w0, neq, b, ldb, nrhs);
}

int sprgpu_solve_partial(sprgpu_handle *h, double *b, sprgpu_int ldb, sprgpu_int nrhs,
                         int which)
{
    if (!h || !b || ldb < 0 || which < 1 || which > 3) return -2;
    const i64 neq = h->plan.neq;
    if (ldb >= neq) return sprgpu_solve(h, b, ldb, nrhs, which);
    if (nrhs <= 0 || ldb == 0) return 0;
    if (!h->factored) {
        set_error("solve called before a successful factorization");
        return -1;
    }
    // The equations beyond LDB enter as zero, and results for them are
    // dropped.  In the backward sweep the reference also reads them as zero
    // where other equations use them; a zero right-hand side on the trailing
    // internal equations gives exactly that, so with PERI2E the leading
    // external equations must be the leading internal ones.
    if (h->d.peri2e && h->d.peri2e_pmax[ldb] != ldb - 1) {
        set_error("partial solve: with PERI2E the first LDB external equations must be the first "
                  "LDB internal ones");
        return -2;
    }
    double *w0, *w1;
    if (int rc = perm_work(h, nrhs, &w0, &w1)) return rc;
    if (int rc = lead_gather(h, b, ldb, w0, neq, nrhs, ldb)) return rc;
    if (which & 1)
        if (int rc = solve_internal(h, w0, neq, nrhs, 1)) return rc;
    if (which == 3)
        if (int rc = lead_mask(h, w0, neq, nrhs, ldb)) return rc;
    if (which & 2)
        if (int rc = solve_internal(h, w0, neq, nrhs, 2)) return rc;
    return lead_scatter(h, w0, neq, b, ldb, nrhs, ldb);
}

int sprgpu_spmv(sprgpu_handle *h, const double *x, double *y, sprgpu_int ld, sprgpu_int nrhs)
{
    if (!h || !x || !y || ld < h->plan.neq) return -2;
    if (!h->d.peri2e || nrhs <= 0) return run_spmv(h, x, y, ld, nrhs);
    const i64 neq = h->plan.neq;
    double *w0, *w1;
    if (int rc = perm_work(h, nrhs, &w0, &w1)) return rc;
    if (int rc = perm_gather(h, x, ld, w0, neq, nrhs)) return rc;
    if (int rc = run_spmv(h, w0, w1, neq, nrhs)) return rc;
    return perm_scatter(h, w1, neq, y, ld, nrhs);
}

int sprgpu_sprprm(sprgpu_handle *h, const double *a_dev, const double *b_dev, double *c_dev,
                  double *wa_dev, sprgpu_int m, sprgpu_int n, sprgpu_int ksa, sprgpu_int iflag)
{
    if (!h || !a_dev || !b_dev) return -2;
    if (!h->d.peri2e || ksa <= 0 || m < h->plan.neq || n < 1)
        return run_sprprm(h, a_dev, b_dev, c_dev, wa_dev, m, n, ksa, iflag);
    // SPRPRR (sprprm.f:327-397): B, C and WA in external order
    const i64 neq = h->plan.neq;
    const bool only_wa = (iflag == 88 && n == 1);
    if ((only_wa && !wa_dev) || (!only_wa && !c_dev))
        return run_sprprm(h, a_dev, b_dev, c_dev, wa_dev, m, n, ksa, iflag);  // its error
    double *w0, *w1;
    if (int rc = perm_work(h, n + 1, &w0, &w1)) return rc;
    double *wa_int = w1 + neq * n;
    if (int rc = perm_gather(h, b_dev, m, w0, neq, n)) return rc;
    if (!only_wa && (iflag == 2 || iflag == -2))
        if (int rc = perm_gather(h, c_dev, m, w1, neq, n)) return rc;
    if (int rc = run_sprprm(h, a_dev, w0, only_wa ? nullptr : w1, wa_int, neq, n, ksa, iflag))
        return rc;
    if (!only_wa)
        if (int rc = perm_scatter(h, w1, neq, c_dev, m, n)) return rc;
    if (wa_dev) return perm_scatter(h, wa_int, neq, wa_dev, m, 1);
    return 0;
}

int sprgpu_sprtra(sprgpu_handle *h, const double *a_dev, const double *b_dev, double *c_dev,
                  double *wa_dev, sprgpu_int m, sprgpu_int n, sprgpu_int ksa)
{
    // sprtra.f:38-50: per column J, WA = A B(:, J) (SPRPRM, IFLAG 88), then
    // C(I, J) = C(J, I) = B(:, I) . WA for I >= J
    if (!h || !a_dev || !b_dev || !c_dev || !wa_dev || n < 1) return -2;
    for (i64 j = 0; j < n; ++j) {
        if (int rc = sprgpu_sprprm(h, a_dev, b_dev + j * m, nullptr, wa_dev, m, 1, ksa, 88))
            return rc;
        if (int rc = run_tra_dots(h, b_dev, m, wa_dev, n, j, c_dev)) return rc;
    }
    return 0;
}

int sprgpu_sprtmm(sprgpu_handle *h, const double *u_dev, const double *b_dev, double *c_dev,
                  double *wa_dev, sprgpu_int m, sprgpu_int n, sprgpu_int ksu, sprgpu_int iflag)
{
    if (!h || !u_dev || !b_dev) return -2;
    if (!h->d.peri2e || m < h->plan.neq || n < 1)
        return run_sprtmm(h, u_dev, b_dev, c_dev, wa_dev, m, n, ksu, iflag);
    // SPRTMR (sprtmm.f:345-480): B, C and WA in external order; U is read as
    // a stored lower triangle whatever KSU says, as SPRTMR reads it
    const i64 neq = h->plan.neq;
    const bool only_wa = (iflag == 11 || iflag == 22) && n == 1;
    if ((only_wa && !wa_dev) || (!only_wa && !c_dev) ||
        (!only_wa && iflag != 1 && iflag != 2))
        return run_sprtmm(h, u_dev, b_dev, c_dev, wa_dev, m, n, 1, iflag);  // its error
    double *w0, *w1;
    if (int rc = perm_work(h, n + 1, &w0, &w1)) return rc;
    double *wa_int = w1 + neq * n;
    if (int rc = perm_gather(h, b_dev, m, w0, neq, n)) return rc;
    if (int rc = run_sprtmm(h, u_dev, w0, only_wa ? nullptr : w1, wa_int, neq, n, 1, iflag))
        return rc;
    if (!only_wa)
        if (int rc = perm_scatter(h, w1, neq, c_dev, m, n)) return rc;
    if (wa_dev) return perm_scatter(h, wa_int, neq, wa_dev, m, 1);
    return 0;
}

int sprgpu_lanczos(sprgpu_handle *h, const double *a_sm_dev, const double *b_dev, int ksb,
                   double shift, sprgpu_int neval, sprgpu_int nevec, sprgpu_int maxlan, double tol,
                   double *evl, double *everr, double *v_dev, sprgpu_int *nconv)
{
    if (!h || !a_sm_dev || !b_dev || !evl || !everr || !nconv || (nevec > 0 && !v_dev))
        return -2;
    return run_lanczos(h, a_sm_dev, b_dev, ksb, shift, neval, nevec, maxlan, tol, evl, everr,
                       v_dev, nconv);
}

int sprgpu_solve_ir(sprgpu_handle *h, const double *b, double *x, sprgpu_int ld,
                    sprgpu_int nrhs, double rtol, int maxit, sprgpu_ir_info *info)
{
    if (!h || !b || !x || ld < h->plan.neq) return -2;
    if (!h->d.peri2e || nrhs <= 0) return run_solve_ir(h, b, x, ld, nrhs, rtol, maxit, info);
    const i64 neq = h->plan.neq;
    double *w0, *w1;
    if (int rc = perm_work(h, nrhs, &w0, &w1)) return rc;
    if (int rc = perm_gather(h, b, ld, w0, neq, nrhs)) return rc;
    if (int rc = run_solve_ir(h, w0, w1, neq, nrhs, rtol, maxit, info)) return rc;
    return perm_scatter(h, w1, neq, x, ld, nrhs);
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
    if (on_device) {
        int rc = copy_factor_to_sm(h, sm_out);
        if (rc == 0 && (h->mode == 2 || h->mode == 4)) rc = cc_factor_layout(h, sm_out);
        return rc;
    }
    double *tmp = nullptr;
    SPRGPU_HIP(hipMalloc(&tmp, bytes));
    // start from the input so that entries the factor does not own keep it
    if (h->last_sm_dev)
        SPRGPU_HIP(hipMemcpyAsync(tmp, h->last_sm_dev, bytes, hipMemcpyDeviceToDevice, h->stream));
    int rc = copy_factor_to_sm(h, tmp);
    if (rc == 0 && (h->mode == 2 || h->mode == 4)) rc = cc_factor_layout(h, tmp);
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
    std::string err = build_plan(mspar, mtrees, msifa, SPRGPU_FP64, PlanParams{}, p);
    if (!err.empty()) {
        set_error(err);
        return -2;
    }
    i64 tiles = 0, steps = 0, kernels = 0;
    for (const Launch &L : p.launches) {
        if (L.kind == L_GEMM) tiles += L.blocks;
        if (L.kind == L_PANEL || L.kind == L_BDIAG || L.kind == L_BFUSED) steps += L.count;
        if (L.kind != L_RECORD && L.kind != L_WAIT) ++kernels;
    }
    out[0] = double(p.neq);
    out[1] = double(p.nsuper);
    out[2] = double(p.nlevels);
    out[3] = double(kernels);
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
