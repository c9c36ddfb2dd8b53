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
 *   oracle_sprtmm  <- SPRTMM  src/SPR/sprtmm.f:5-241 (C = U B or U^T B with the
 *                              stored lower triangle)
 *   oracle_assemble_free <- SPRADM called once per element (spradm.f:5-166),
 *                              its SPRAD2 free-dof loop (spradm.f:818-859)
 *   oracle_assemble_ceq  <- the same with SPRAD2's constrained part
 *                              (spradm.f:861-1002) and the SV vectors
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
 * C = +-AB, +-2: C = C +- AB.  MSPAR(58)>=1: SPRPRR (sprprm.f:327-397). */
void oracle_sprprm(const double *a, const double *b, double *c, double *wa, fint *mspar,
                   const fint *mtrees, const fint *msifa, const fint *m, const fint *n,
                   const fint *ksa, const fint *iflag, const fint *lpu, fint *ierr);

/* SM(NEQ+1:) += EM_e(I,J) for every element e = 0..nel-1 in order, I <= J
 * with MEEN(I), MEEN(J) >= 1 (SPRAD2's loop, spradm.f:818-859), at the LNZ
 * slot of (max(IEQ,JEQ), min(IEQ,JEQ)) (SPRAD1's PSM table,
 * spradm.f:514-740).  Element e's dofs have the 1-based equations
 * meen[meen_ptr[e] .. meen_ptr[e+1]) (SPREEQ's MEEN: 0 = no equation; the
 * constrained dofs, MEEN < 0, are skipped as the free loop skips them).
 * EM_e = escale[e] * em[em_off[e] ..] (escale NULL: no scaling),
 * column-major.  Returns 0, -(e+1) if element e couples two equations outside
 * the symbolic pattern, or -1 on allocation failure. */
fint oracle_assemble_free(const fint *mspar, const fint *mtrees, const fint *msifa, fint nel,
                          const fint *meen_ptr, const fint *meen, const double *em,
                          const fint *em_off, const double *escale, double *sm);
/* SPRADM per element with constraint equations (SPRAD2, spradm.f:743-1002):
 * with nsv >= 0 the free loop above, then for every constrained dof J
 * (MEEN(J) = -JCEQ) the weighted columns: SV(NEQ*|nsv|) -= C0*EM(:,J) (for
 * nsv != 0) and SM += TTCC*EM(:,J) through J's masters (for nsv >= 0), in
 * SPRAD2's loop order.  meqn: MEQN(NDOF), to find the masters' equations.
 * Returns 0, -(e+1) for element e outside the pattern, -(nel+1) if a
 * prescribed dof's vector lies beyond |nsv|, -1 on allocation failure. */
fint oracle_assemble_ceq(const fint *mspar, const fint *mtrees, const fint *msifa, fint nel,
                         const fint *meen_ptr, const fint *meen, const double *em,
                         const fint *em_off, const double *escale, const fint *mpmceq,
                         const fint *mmceq, const double *ttcc, const fint *meqn, fint nsv,
                         double *sm, double *sv);
/* SPRTMM(U,B,C,WA,MSPAR,MTREES,MSIFA,M,N,KSU,IFLAG,LPU,IERR): with L the
 * stored lower triangle of U (diagonal included), IFLAG 1: C = L^T B, 2:
 * C = L B; 11/22 (N = 1): the same into WA only (SPRTMQ, sprtmm.f:242-344).
 * KSU <= 0: U is a diagonal.  MSPAR(58)>=1: SPRTMR (sprtmm.f:345-480). */
void oracle_sprtmm(const double *u, const double *b, double *c, double *wa, fint *mspar,
                   const fint *mtrees, const fint *msifa, const fint *m, const fint *n,
                   const fint *ksu, const fint *iflag, const fint *lpu, fint *ierr);
#endif
