// Drop-in replacements of the reference numeric entry points.
//
// Each follows the reference driver's control flow: the MSPAR(1) stage
// protocol, the MSPAR-derived array-length checks, and the IERR / MSPAR(26)
// / MSPAR(27) / TOL(2:3) / DCOR outputs:
//   sprgpu_sprfc1  <- SPRFC1  sprfc1.f:5-283
//   sprgpu_sprfc3  <- SPRFC3  sprfc3.f:5-280
//   sprgpu_sprfc4  <- SPRFC4  sprfc4.f:5-265 (the CC' factor with SPRFC3-style replacement)
//   sprgpu_sprfs1  <- SPRFS1  sprfs1.f:5-263 (+ SPRFS4 permutation, :396-528)
//   sprgpu_sprbs1  <- SPRBS1  sprbs1.f:5-248 (+ SPRBS5 permutation, :371-495)
//   sprgpu_sprsol  <- SPRSOL  sprsol.f:5-52
// The factor stays on the GPU between calls, keyed by the SM address; SM on
// the host always receives the factor too, so reference routines that read
// it afterwards (SPREXD, SPRDAD, ...) still work.  Each cached factor keeps a
// signature of the host SM it was written to (a hash of SM(1:NEQ), the
// pivots, and of a strided sample of LNZ); a solve whose SM no longer
// matches it (a factor computed elsewhere, a new SM at a reused address)
// takes the factor from SM again.  At most kMaxEntries SM arrays are cached;
// the least recently used one is released first.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

#include "sprgpu_internal.h"

using namespace sprgpu;

namespace {

struct Entry {
    sprgpu_handle *h = nullptr;
    double *d_sm = nullptr;   // device copy of SM (NEQ+NZEROL)
    double *d_b = nullptr;    // device right-hand sides
    i64 b_cap = 0;
    std::vector<i64> key;     // fingerprint of the symbolic structure
    unsigned long long sig = 0;  // signature of the host SM holding the device factor
    bool sig_valid = false;
    unsigned long long last_use = 0;
};

constexpr size_t kMaxEntries = 4;
constexpr i64 kSigSamples = 4096;

std::mutex g_mu;
std::map<const double *, Entry> g_cache;
unsigned long long g_clock = 0;
// LIERR of the last drop-in factorization on this thread: the reference's
// internal error code behind IERR = -3 (sprfc1.f:250-262, sprfc2.f:244-255,
// sprfc3.f:250-262, sprfc4.f:240-256): -30 a pivot at or below
// TOL(1)|a_jj| (SPRFC1), -32 the same in SPRFC2 and a pivot below
// -TOL(1)|a_jj| in SPRFC4, -31 pivots replaced (SPRFC3), -33 pivots
// replaced (SPRFC4), -35 a zero diagonal before the factorization; 0
// otherwise.  The Fortran layer passes it to SPRER1.
thread_local i64 g_lierr = 0;

// Signature of a factor in SM: FNV-1a over the bits of SM(1:NEQ) and of
// kSigSamples entries of LNZ spread evenly over it.
unsigned long long factor_signature(const double *sm, i64 neq, i64 nzerol)
{
    unsigned long long h = 1469598103934665603ull;
    auto mix = [&](double v) {
        unsigned long long u;
        std::memcpy(&u, &v, sizeof(u));
        h ^= u;
        h *= 1099511628211ull;
    };
    for (i64 i = 0; i < neq; ++i) mix(sm[i]);
    const i64 step = std::max<i64>(1, nzerol / kSigSamples);
    for (i64 k = 0; k < nzerol; k += step) mix(sm[neq + k]);
    if (nzerol > 0) mix(sm[neq + nzerol - 1]);
    return h;
}

std::vector<i64> fingerprint(const i64 *mspar, const i64 *mtrees, const i64 *msifa)
{
    auto M = [&](int k) { return mspar[k - 1]; };
    std::vector<i64> key = {M(8), M(11), M(14), M(15), M(16), M(45), M(46), M(47), M(48)};
    // hash of XSUPER, XLINDX and LINDX (FNV-1a over 64-bit words)
    unsigned long long hsh = 1469598103934665603ull;
    auto mix = [&](i64 v) {
        hsh ^= (unsigned long long)v;
        hsh *= 1099511628211ull;
    };
    const i64 ns = M(11);
    for (i64 j = 0; j <= ns; ++j) mix(mtrees[M(45) - 1 + j]);
    for (i64 j = 0; j <= ns; ++j) mix(mtrees[M(46) - 1 + j]);
    for (i64 k = 0; k < M(15); ++k) mix(msifa[M(47) - 1 + k]);
    key.push_back(i64(hsh));
    return key;
}

void release_entry(Entry &e)
{
    if (e.h) sprgpu_destroy(e.h);
    if (e.d_sm) (void)hipFree(e.d_sm);
    if (e.d_b) (void)hipFree(e.d_b);
    e = Entry{};
}

// Finds or creates the device state for SM.  Returns 0, -2 (structure not
// supported) or -5 (HIP error).
int get_entry(const double *sm, const i64 *mspar, const i64 *mtrees, const i64 *msifa,
              Entry *&out, bool *fresh)
{
    auto key = fingerprint(mspar, mtrees, msifa);
    auto it = g_cache.find(sm);
    *fresh = false;
    if (it != g_cache.end() && it->second.key == key) {
        out = &it->second;
        out->last_use = ++g_clock;
        return 0;
    }
    if (it != g_cache.end()) {
        release_entry(it->second);
        g_cache.erase(it);
    }
    while (g_cache.size() >= kMaxEntries) {  // evict the least recently used
        auto lru = g_cache.begin();
        for (auto q = g_cache.begin(); q != g_cache.end(); ++q)
            if (q->second.last_use < lru->second.last_use) lru = q;
        release_entry(lru->second);
        g_cache.erase(lru);
    }
    Entry e;
    e.key = key;
    int rc = sprgpu_create(mspar, mtrees, msifa, SPRGPU_FP64, &e.h);
    if (rc) return rc == -5 ? -5 : -2;
    const i64 n = e.h->plan.neq + e.h->plan.nzerol;
    if (hipMalloc(&e.d_sm, sizeof(double) * n) != hipSuccess) {
        release_entry(e);
        set_error("device allocation of SM failed");
        return -5;
    }
    e.last_use = ++g_clock;
    g_cache[sm] = e;
    out = &g_cache[sm];
    *fresh = true;
    return 0;
}

// MTREES / MSIFA length checks as carved by the reference drivers
// (sprfc1.f:175-236, sprfs1.f:178-230).
bool lengths_ok_factor(const i64 *mspar)
{
    auto M = [&](int k) { return mspar[k - 1]; };
    const i64 nsuper = M(11), nb = M(12), nelact = M(19), nblock = M(22);
    const i64 xlindx = M(46);
    const i64 supsup = xlindx + nsuper + 1, elseq = supsup + nsuper;
    const i64 xelseq = elseq + nelact, xblock = xelseq + nb + 1, mtused = xblock + nb;
    const i64 xlnz = M(48), xsplit = xlnz + nsuper + 1, split = xsplit + nsuper;
    const i64 mlused = split + nblock;
    return !(mtused > M(4) || mlused > M(3));
}

bool lengths_ok_solve(const i64 *mspar)
{
    auto M = [&](int k) { return mspar[k - 1]; };
    const i64 nsuper = M(11), nb = M(12), nelact = M(19);
    const i64 xlindx = M(46);
    const i64 supsup = xlindx + nsuper + 1, elseq = supsup + nsuper;
    const i64 xelseq = elseq + nelact, xblock = xelseq + nb + 1, mtused = xblock + nb;
    const i64 mlused = M(47) + M(15) - 1;
    return !(mtused > M(4) || mlused > M(3));
}

// Messages go to Fortran unit LPU from the Fortran layer; the C layer only
// records the reason (sprgpu_last_error).
void report(const sprgpu_int *, const char *, i64) {}

void factor_dropin(int mode, sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa,
                   double *sm, double *tol, double *dcor, const sprgpu_int *lpu,
                   sprgpu_int *ierr, const char *name)
{
    // stage protocol (sprfc1.f:140-148; SPRFC2 and SPRFC4 use 6, sprfc2.f,
    // sprfc4.f:133-141)
    const bool cc = mode == 2 || mode == 4;
    const sprgpu_int stage = cc ? 6 : 5;
    g_lierr = 0;
    if (mspar[0] < 0) {
        mspar[0] = -stage;
        *ierr = -1;
        set_error("error from a previous SPR stage is not cleared");
        report(lpu, name, *ierr);
        return;
    }
    mspar[0] = stage;
    *ierr = 0;
    if (!lengths_ok_factor(mspar)) {
        mspar[0] = -2;
        *ierr = -2;
        set_error("MTREES and/or MSIFA are not great enough");
        report(lpu, name, *ierr);
        return;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    Entry *e = nullptr;
    bool fresh = false;
    int rc = get_entry(sm, mspar, mtrees, msifa, e, &fresh);
    if (rc) {
        *ierr = rc == -5 ? -5 : -2;
        if (*ierr == -2) mspar[0] = -2;
        report(lpu, name, *ierr);
        return;
    }
    sprgpu_handle *h = e->h;
    const i64 neq = h->plan.neq, n = neq + h->plan.nzerol;
    sprgpu_factor_info info{};
    double t4[4] = {tol[0], 0.0, 0.0, mode >= 3 ? tol[3] : 0.0};
    hipError_t he = hipMemcpy(e->d_sm, sm, sizeof(double) * n, hipMemcpyHostToDevice);
    if (he != hipSuccess) rc = hip_fail(he, "upload SM");
    if (!rc) rc = sprgpu_factor(h, e->d_sm, t4, mode, &info);
    if (!rc) rc = copy_factor_to_sm(h, e->d_sm);
    // C = L sqrt(D) in SM once every pivot is taken (SPRFC4: also with
    // replaced pivots, MSPAR(27) < 0)
    if (!rc && cc && (info.ierr == 0 || (mode == 4 && info.supeqt < 0)))
        rc = cc_factor_layout(h, e->d_sm);
    if (!rc) {
        he = hipMemcpy(sm, e->d_sm, sizeof(double) * n, hipMemcpyDeviceToHost);
        if (he != hipSuccess) rc = hip_fail(he, "download SM");
    }
    if (!rc) {
        e->sig = factor_signature(sm, neq, h->plan.nzerol);
        e->sig_valid = true;
    }
    if (!rc && mode >= 3 && dcor) {
        he = hipMemcpy(dcor, h->d.dcor, sizeof(double) * 2 * neq, hipMemcpyDeviceToHost);
        if (he != hipSuccess) rc = hip_fail(he, "download DCOR");
    }
    if (rc) {
        *ierr = -5;
        report(lpu, name, *ierr);
        return;
    }
    // outputs (sprfc1.f:245-267, sprfc3.f:245-266)
    tol[1] = info.tol2;
    tol[2] = info.tol3;
    if (info.ierr == -3) {
        *ierr = -3;
        mspar[26] = info.supeqt;
        g_lierr = (mode <= 2 && info.nzero_diag > 0) ? -35
                  : mode == 3                        ? -31
                  : mode == 4                        ? (info.supeqt < 0 ? -33 : -32)
                  : mode == 2                        ? -32
                                                     : -30;
    } else if (info.ierr == 3) {
        *ierr = 3;
    }
    mspar[25] = info.numneg;
    report(lpu, name, *ierr);
}

// Solve driver shared by SPRFS1/SPRBS1 and, with cc, SPRFS2/SPRBS2 (the CC'
// factor C in SM).  which = 1 forward, 2 backward.
void solve_dropin(int which, bool cc, sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa,
                  double *sm, double *b, const sprgpu_int *ldb, const sprgpu_int *nrhs,
                  const sprgpu_int *lpu, sprgpu_int *ierr, const char *name)
{
    const sprgpu_int stage = cc ? 8 : 7;  // sprfs1.f, sprfs2.f:140-145
    if (mspar[0] < 0) {
        mspar[0] = -stage;
        *ierr = -1;
        set_error("error from a previous SPR stage is not cleared");
        report(lpu, name, *ierr);
        return;
    }
    mspar[0] = stage;
    *ierr = 0;
    const i64 neq = mspar[7];
    // MSPAR(57) = 1 and LDB < NEQ: only the leading LDB equations
    // (sprfs1.f:200-201, sprbs1.f:189)
    const bool partial = mspar[56] == 1 && *ldb < neq;
    if (!lengths_ok_solve(mspar) || (*ldb < neq && !partial) || *ldb < 0) {
        mspar[0] = -2;
        *ierr = -2;
        set_error("MTREES, MSIFA or LDB are not great enough");
        report(lpu, name, *ierr);
        return;
    }
    const i64 k = *nrhs;
    if (k <= 0) return;
    std::lock_guard<std::mutex> lock(g_mu);
    Entry *e = nullptr;
    bool fresh = false;
    int rc = get_entry(sm, mspar, mtrees, msifa, e, &fresh);
    if (rc) {
        *ierr = rc == -5 ? -5 : -2;
        if (*ierr == -2) mspar[0] = -2;
        report(lpu, name, *ierr);
        return;
    }
    sprgpu_handle *h = e->h;
    hipError_t he = hipSuccess;
    const unsigned long long sig = factor_signature(sm, neq, h->plan.nzerol);
    if (fresh || !h->factored || !e->sig_valid || e->sig != sig) {
        // the factor was computed elsewhere (e.g. by the reference SPRFC1),
        // or SM changed since the device factor was written to it: take the
        // factor from SM
        const i64 n = neq + h->plan.nzerol;
        e->sig_valid = false;
        he = hipMemcpy(e->d_sm, sm, sizeof(double) * n, hipMemcpyHostToDevice);
        if (he != hipSuccess) rc = hip_fail(he, "upload SM");
        if (!rc && cc) rc = cc_to_ldl_layout(h, e->d_sm);  // C = L sqrt(D) to L, D
        if (!rc) rc = load_factor_from_sm(h, e->d_sm);
        if (!rc) {
            h->mode = cc ? 2 : 1;
            e->sig = sig;
            e->sig_valid = true;
        }
    }
    if (!rc && e->b_cap < neq * k) {
        if (e->d_b) (void)hipFree(e->d_b);
        e->d_b = nullptr;
        he = hipMalloc(&e->d_b, sizeof(double) * neq * k);
        if (he != hipSuccess) rc = hip_fail(he, "allocate B");
        else e->b_cap = neq * k;
    }
    // With MSPAR(58) >= 1 B is in the caller's external order; the handle
    // applies PERI2E on the device (SPRFS4 / SPRBS5 through SPRFB3 / SPRFB4,
    // sprfs1.f:396-528, sprbs1.f:371-495)
    const i64 rows = partial ? *ldb : neq;  // rows of B held on the device
    if (!rc && rows > 0) {
        he = hipMemcpy2D(e->d_b, sizeof(double) * rows, b, sizeof(double) * (*ldb),
                         sizeof(double) * rows, k, hipMemcpyHostToDevice);
        if (he != hipSuccess) rc = hip_fail(he, "upload B");
    }
    if (!rc)
        rc = partial ? sprgpu_solve_partial(h, e->d_b, rows, k, which)
                     : sprgpu_solve(h, e->d_b, neq, k, which);
    if (!rc) rc = check_solve_error(h);
    if (!rc && rows > 0) {
        he = hipMemcpy2D(b, sizeof(double) * (*ldb), e->d_b, sizeof(double) * rows,
                         sizeof(double) * rows, k, hipMemcpyDeviceToHost);
        if (he != hipSuccess) rc = hip_fail(he, "download B");
    }
    if (rc) {
        // -2: a partial solve the device cannot do (PERI2E not leading)
        *ierr = rc == -2 ? -2 : -5;
        if (rc == -2) mspar[0] = -2;
        report(lpu, name, *ierr);
    }
}

}  // namespace

extern "C" {

void sprgpu_sprfc1(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *tol, sprgpu_int *iwork, double *rwork, const sprgpu_int *lpu,
                   sprgpu_int *ierr, void (*mvmult)(void), void (*mmmult)(void))
{
    (void)iwork;
    (void)rwork;
    (void)mvmult;
    (void)mmmult;
    factor_dropin(1, mspar, mtrees, msifa, sm, tol, nullptr, lpu, ierr, "SPRFC1");
}

void sprgpu_sprfc3(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *tol, sprgpu_int *iwork, double *rwork, double *dcor,
                   const sprgpu_int *lpu, sprgpu_int *ierr, void (*mvmult)(void),
                   void (*mmmult)(void))
{
    (void)iwork;
    (void)rwork;
    (void)mvmult;
    (void)mmmult;
    factor_dropin(3, mspar, mtrees, msifa, sm, tol, dcor, lpu, ierr, "SPRFC3");
}

void sprgpu_sprfs1(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *b, const sprgpu_int *ldb, const sprgpu_int *nrhs, sprgpu_int *iwork,
                   double *rwork, const sprgpu_int *lpu, sprgpu_int *ierr)
{
    (void)iwork;
    (void)rwork;
    solve_dropin(1, false, mspar, mtrees, msifa, sm, b, ldb, nrhs, lpu, ierr, "SPRFS1");
}

void sprgpu_sprbs1(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *b, const sprgpu_int *ldb, const sprgpu_int *nrhs, double *rwork,
                   const sprgpu_int *lpu, sprgpu_int *ierr)
{
    (void)rwork;
    solve_dropin(2, false, mspar, mtrees, msifa, sm, b, ldb, nrhs, lpu, ierr, "SPRBS1");
}

void sprgpu_sprfc2(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *tol, sprgpu_int *iwork, double *rwork, const sprgpu_int *lpu,
                   sprgpu_int *ierr, void (*mvmult)(void), void (*mmmult)(void))
{
    (void)iwork;
    (void)rwork;
    (void)mvmult;
    (void)mmmult;
    factor_dropin(2, mspar, mtrees, msifa, sm, tol, nullptr, lpu, ierr, "SPRFC2");
}

void sprgpu_sprfc4(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *tol, sprgpu_int *iwork, double *rwork, double *dcor,
                   const sprgpu_int *lpu, sprgpu_int *ierr, void (*mvmult)(void),
                   void (*mmmult)(void))
{
    (void)iwork;
    (void)rwork;
    (void)mvmult;
    (void)mmmult;
    factor_dropin(4, mspar, mtrees, msifa, sm, tol, dcor, lpu, ierr, "SPRFC4");
}

void sprgpu_sprfs2(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *b, const sprgpu_int *ldb, const sprgpu_int *nrhs, sprgpu_int *iwork,
                   double *rwork, const sprgpu_int *lpu, sprgpu_int *ierr)
{
    (void)iwork;
    (void)rwork;
    solve_dropin(1, true, mspar, mtrees, msifa, sm, b, ldb, nrhs, lpu, ierr, "SPRFS2");
}

void sprgpu_sprbs2(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *b, const sprgpu_int *ldb, const sprgpu_int *nrhs, double *rwork,
                   const sprgpu_int *lpu, sprgpu_int *ierr)
{
    (void)rwork;
    solve_dropin(2, true, mspar, mtrees, msifa, sm, b, ldb, nrhs, lpu, ierr, "SPRBS2");
}

void sprgpu_sprsol(const sprgpu_int *iop, sprgpu_int *mspar, sprgpu_int *mtrees,
                   sprgpu_int *msifa, double *sm, double *b, const sprgpu_int *ldb,
                   const sprgpu_int *nrhs, double *tol, sprgpu_int *iwork, double *rwork,
                   const sprgpu_int *lpu, sprgpu_int *ierr)
{
    // sprsol.f:29-50
    *ierr = 0;
    if (*iop <= 3)
        sprgpu_sprfc1(mspar, mtrees, msifa, sm, tol, iwork, rwork, lpu, ierr, nullptr, nullptr);
    if (*iop >= 2 && *iop <= 5 && *ierr >= 0)
        sprgpu_sprfs1(mspar, mtrees, msifa, sm, b, ldb, nrhs, iwork, rwork, lpu, ierr);
    if (*iop >= 3 && *iop != 5 && *ierr >= 0)
        sprgpu_sprbs1(mspar, mtrees, msifa, sm, b, ldb, nrhs, rwork, lpu, ierr);
}

sprgpu_int sprgpu_dropin_lierr(void) { return g_lierr; }

void sprgpu_release(const double *sm)
{
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_cache.find(sm);
    if (it != g_cache.end()) {
        release_entry(it->second);
        g_cache.erase(it);
    }
}

}  // extern "C"
