/*
 * Timing driver linked into the reference SPR build (oracle/_ref/libspr_ref.so).
 *
 * Test infrastructure only.  The reference is a Fortran-77 library whose
 * public routines take every argument by reference (INTEGER*8 in the default
 * build, src/SPR/CMakeLists.txt:12-29).  Python reaches them through ctypes;
 * this helper times the reference numeric path with a monotonic clock around
 * each call.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdint.h>
#include <time.h>

typedef int64_t fint;

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

/* SPRFC2 (sprfc2.f:5-7) with the CC' kernels SPRSOL-style callers pass:
 * CMVY18 / CMMY18 (cmvy18.f:5, cmmy18.f:5). */
extern void cmvy18_(void);
extern void cmmy18_(void);
extern void sprfc2_(fint *mspar, fint *mtrees, fint *msifa, double *sm, double *tol,
                    fint *iwork, double *rwork, const fint *lpu, fint *ierr,
                    void (*mvmult)(void), void (*mmmult)(void));

fint spr_fc2(fint *mspar, fint *mtrees, fint *msifa, double *sm, double *tol, fint *iwork,
             double *rwork, fint lpu)
{
    fint ierr = 0;
    sprfc2_(mspar, mtrees, msifa, sm, tol, iwork, rwork, &lpu, &ierr, cmvy18_, cmmy18_);
    return ierr;
}

/* SPRFC4 (sprfc4.f:5-7), the CC' factorization with replacement, with the
 * same CC' kernels. */
extern void sprfc4_(fint *mspar, fint *mtrees, fint *msifa, double *sm, double *tol,
                    fint *iwork, double *rwork, double *dcor, const fint *lpu, fint *ierr,
                    void (*mvmult)(void), void (*mmmult)(void));

fint spr_fc4(fint *mspar, fint *mtrees, fint *msifa, double *sm, double *tol, fint *iwork,
             double *rwork, double *dcor, fint lpu)
{
    fint ierr = 0;
    sprfc4_(mspar, mtrees, msifa, sm, tol, iwork, rwork, dcor, &lpu, &ierr, cmvy18_, cmmy18_);
    return ierr;
}
