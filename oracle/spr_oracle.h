/*
 * CPU restatement of the SPR numeric path (factor and solves).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker that the HIP library is
 * compared against.  It is never linked into, or called by, the product.
 *
 * Entry points mirror the reference's Fortran-77 interfaces argument by
 * argument (every argument by reference, INTEGER*8, 1-based content):
 *   oracle_sprfc1  <- SPRFC1  src/SPR/sprfc1.f:5-7
 *   oracle_sprfc3  <- SPRFC3  src/SPR/sprfc3.f:5-7
 *   oracle_sprfs1  <- SPRFS1  src/SPR/sprfs1.f:5-7
 *   oracle_sprbs1  <- SPRBS1  src/SPR/sprbs1.f:5-6
 *   oracle_sprsol  <- SPRSOL  src/SPR/sprsol.f:5-6
 *   oracle_sprprm  <- SPRPRM  src/SPR/sprprm.f:5-7 (C = +-A*B on the unfactored
 *                              SPR layout; used as the residual checker)
 * The MVMULT/MMMULT callback arguments of SPRFC1/SPRFC3 are fixed to the
 * restated LMVY18/LMMY18 (what SPRSOL passes, sprsol.f:29-35).
 */
#ifndef SPR_ORACLE_H
#define SPR_ORACLE_H
#include <stdint.h>

typedef int64_t fint;

void oracle_sprfc1(fint *mspar, const fint *mtrees, const fint *msifa, double *sm, double *tol,
                   fint *iwork, double *rwork, const fint *lpu, fint *ierr);
void oracle_sprfc3(fint *mspar, const fint *mtrees, const fint *msifa, double *sm, double *tol,
                   fint *iwork, double *rwork, double *dcor, const fint *lpu, fint *ierr);
void oracle_sprfs1(fint *mspar, const fint *mtrees, const fint *msifa, const double *sm,
                   double *b, const fint *ldb, const fint *nrhs, fint *iwork, double *rwork,
                   const fint *lpu, fint *ierr);
void oracle_sprbs1(fint *mspar, const fint *mtrees, const fint *msifa, const double *sm,
                   double *b, const fint *ldb, const fint *nrhs, double *rwork,
                   const fint *lpu, fint *ierr);
void oracle_sprsol(const fint *iop, fint *mspar, const fint *mtrees, const fint *msifa,
                   double *sm, double *b, const fint *ldb, const fint *nrhs, double *tol,
                   fint *iwork, double *rwork, const fint *lpu, fint *ierr);
/* SPRPRM(A,B,C,WA,MSPAR,MTREES,MSIFA,M,N,KSA,IFLAG,LPU,IERR) with A the
 * unfactored matrix SM(NEQ+1:) (KSA>0) or a diagonal (KSA<=0); IFLAG +-1:
 * C = +-AB, +-2: C = C +- AB.  MSPAR(58)>=1 (PERI2E) is not restated. */
void oracle_sprprm(const double *a, const double *b, double *c, double *wa, fint *mspar,
                   const fint *mtrees, const fint *msifa, const fint *m, const fint *n,
                   const fint *ksa, const fint *iflag, const fint *lpu, fint *ierr);

#endif
