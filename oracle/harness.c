/*
 * Loop drivers linked into the reference SPR build (oracle/_ref/libspr_ref.so).
 *
 * Test infrastructure only.  The reference is a Fortran-77 library whose
 * public routines take every argument by reference (INTEGER*8 in the default
 * build, src/SPR/CMakeLists.txt:12-29).  Python reaches them through ctypes;
 * these helpers exist so that per-element and per-call loops do not cross
 * the ctypes boundary tens of thousands of times, and so that the reference
 * numeric path can be timed with a monotonic clock around each call.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef int64_t fint;

/* src/SPR/spradm.f:5-13 */
extern void spradm_(const double *em, const double *ttcc, fint *mpar, fint *mspar,
                    fint *madof, fint *meqn, fint *mpmnpc, fint *mmnpc, fint *mpmceq,
                    fint *mmceq, fint *msica, fint *mtrees, fint *msifa, fint *mdofnc,
                    double *sm, double *sv, fint *iwork, const fint *iel,
                    const fint *nedof, const fint *lpu, const fint *nsv, fint *ierr);

/* src/SPR/sprsol.f:5-6 */
extern void sprsol_(const fint *iop, fint *mspar, fint *mtrees, fint *msifa, double *sm,
                    double *b, const fint *ldb, const fint *nrhs, double *tol,
                    fint *iwork, double *rwork, const fint *lpu, fint *ierr);

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/*
 * Assemble nel element matrices into SM (and optionally SV) by calling the
 * reference SPRADM once per element, in element order.
 *   em      element matrices, column-major nedof[i] x nedof[i] each, element i
 *           starting at em + em_off[i] (em_off may repeat offsets so that many
 *           elements share one base matrix)
 *   scale   optional per-element multiplier applied to the base matrix
 *           (NULL: no scaling)
 * Returns 0 or the first non-zero IERR from SPRADM (element number in *bad).
 */
fint spr_assemble_all(fint nel, const fint *nedof, const double *em, const fint *em_off,
                      const double *scale, const double *ttcc, fint *mpar, fint *mspar,
                      fint *madof, fint *meqn, fint *mpmnpc, fint *mmnpc, fint *mpmceq,
                      fint *mmceq, fint *msica, fint *mtrees, fint *msifa, fint *mdofnc,
                      double *sm, double *sv, fint nsv, fint *iwork, fint lpu, fint *bad)
{
    fint maxd = 0;
    for (fint i = 0; i < nel; ++i)
        if (nedof[i] > maxd) maxd = nedof[i];
    double *buf = NULL;
    if (scale) {
        buf = (double *)malloc((size_t)(maxd * maxd) * sizeof(double));
        if (!buf) return -99;
    }
    fint ierr = 0;
    for (fint i = 0; i < nel; ++i) {
        const double *e = em + em_off[i];
        fint nd = nedof[i];
        if (scale) {
            for (fint k = 0; k < nd * nd; ++k) buf[k] = scale[i] * e[k];
            e = buf;
        }
        fint iel = i + 1;
        spradm_(e, ttcc, mpar, mspar, madof, meqn, mpmnpc, mmnpc, mpmceq, mmceq, msica,
                mtrees, msifa, mdofnc, sm, sv, iwork, &iel, &nd, &lpu, &nsv, &ierr);
        if (ierr != 0) {
            if (bad) *bad = iel;
            break;
        }
    }
    free(buf);
    return ierr;
}

/* SPRSOL with a monotonic wall-clock measurement around the call. */
fint spr_sol_timed(fint iop, fint *mspar, fint *mtrees, fint *msifa, double *sm, double *b,
                   fint ldb, fint nrhs, double *tol, fint *iwork, double *rwork, fint lpu,
                   double *seconds)
{
    fint ierr = 0;
    double t0 = now_s();
    sprsol_(&iop, mspar, mtrees, msifa, sm, b, &ldb, &nrhs, tol, iwork, rwork, &lpu, &ierr);
    *seconds = now_s() - t0;
    return ierr;
}
