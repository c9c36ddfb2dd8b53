/*
 * sprgpu -- MI355X (gfx950) numeric factorization and solves for the SPR
 * sparse LDL' solver of sam-lib.
 *
 * Two layers of entry points, all plain C, all integers INTEGER*8:
 *
 * 1. Drop-in replacements for the reference's Fortran-77 numeric routines.
 *    Same argument lists, every argument by reference, same array layouts
 *    (MSPAR/MTREES/MSIFA from the reference symbolic phase, SM from SPRADM),
 *    same IERR codes and MSPAR/TOL outputs.  The Fortran module in
 *    sam-lib_amd/fortran/sprgpu_dropin.f90 re-exports them under the
 *    reference names SPRFC1, SPRFC3, SPRFS1, SPRBS1, SPRSOL.  One extra
 *    IERR value is added: -5 = a HIP runtime error.
 *
 * 2. A device-resident handle API: the factor lives in HBM between calls,
 *    inputs and outputs are device pointers, and the FP32-factor + FP64
 *    iterative-refinement mode, multiple right-hand sides, and the adjoint
 *    are available.  The drop-in layer is built on it.
 */
#ifndef SPRGPU_H
#define SPRGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t sprgpu_int;

/* ------------------------------------------------------------------------ */
/* 1. Drop-in entry points                                                  */
/* ------------------------------------------------------------------------ */

/* Replaces SPRFC1 (src/SPR/sprfc1.f:5-7): SM(1:NEQ+NZEROL) in, L and D out in
 * the same packed supernodal layout; TOL(1) in, TOL(2:3) out; MSPAR(1),
 * MSPAR(26) (NUMNEG), MSPAR(27) (SUPEQT) out.  The factor is also kept on the
 * GPU, keyed by the SM address, for the SPRFS1/SPRBS1 replacements.  MVMULT
 * and MMMULT are accepted for signature compatibility and ignored: the dense
 * front kernels replace LMVY18/LMMY18 (lmvy18.f:5, lmmy18.f:5). */
void sprgpu_sprfc1(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *tol, sprgpu_int *iwork, double *rwork, const sprgpu_int *lpu,
                   sprgpu_int *ierr, void (*mvmult)(void), void (*mmmult)(void));

/* Replaces SPRFC3 (src/SPR/sprfc3.f:5-7): as SPRFC1, but a zero diagonal is
 * replaced by one and a pivot with |d| <= TOL(1)*|a_jj| by
 * sign(d)*TOL(1)*max(1,|TOL(4)*a_jj|), recorded in DCOR(2,NEQ). */
void sprgpu_sprfc3(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *tol, sprgpu_int *iwork, double *rwork, double *dcor,
                   const sprgpu_int *lpu, sprgpu_int *ierr, void (*mvmult)(void),
                   void (*mmmult)(void));

/* Replaces SPRFC4 (src/SPR/sprfc4.f:5-7), the CC' factorization with
 * SPRFC3-style replacement (pivot step sprfc4.f:1016-1048): a zero diagonal
 * is replaced by one (DCOR(2,j) = 1); a pivot d with
 * -TOL(1)*|a_jj| <= d <= TOL(1)*|a_jj| by TOL(1)*max(1,|TOL(4)*a_jj|),
 * counted (IERR = -3, MSPAR(27) = -count) and recorded in DCOR(1:2,j);
 * a pivot below -TOL(1)*|a_jj| fails (IERR = -3, MSPAR(27) = its equation,
 * TOL(2:3) = d and a_jj, DCOR(1:2,j) = d).  MSPAR(26) = 0, MSPAR(1) = 6;
 * SM out holds C = L sqrt(D) (c_jj on the diagonal and in SM(1:NEQ)), for
 * SPRFS2/SPRBS2.  C agrees with the reference's to rounding. */
void sprgpu_sprfc4(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *tol, sprgpu_int *iwork, double *rwork, double *dcor,
                   const sprgpu_int *lpu, sprgpu_int *ierr, void (*mvmult)(void),
                   void (*mmmult)(void));

/* Replaces SPRFS1 (src/SPR/sprfs1.f:5-7): B(LDB,NRHS) <- D^-1 L^-1 B. */
void sprgpu_sprfs1(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *b, const sprgpu_int *ldb, const sprgpu_int *nrhs, sprgpu_int *iwork,
                   double *rwork, const sprgpu_int *lpu, sprgpu_int *ierr);

/* Replaces SPRBS1 (src/SPR/sprbs1.f:5-6): B(LDB,NRHS) <- L^-T B. */
void sprgpu_sprbs1(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *b, const sprgpu_int *ldb, const sprgpu_int *nrhs, double *rwork,
                   const sprgpu_int *lpu, sprgpu_int *ierr);

/* Replace SPRFC2 (src/SPR/sprfc2.f:5-7), SPRFS2 (sprfs2.f:5-7) and SPRBS2
 * (sprbs2.f:5-6), the CC' siblings: SM in, C out in the packed layout (c_jj
 * on the diagonal and in SM(1:NEQ)); B <- C^-1 B; B <- C^-T B.  A pivot
 * fails when d <= TOL(1)*|a_jj| (sprfc2.f:934-1037).  MSPAR(1) = 6 after the
 * factor and 8 after a solve, as the reference sets them.  C is L sqrt(D) of
 * the GPU's LDL' factorization: equal to the reference's to rounding.  A C
 * in SM computed elsewhere is taken by SPRFS2/SPRBS2 as it is. */
void sprgpu_sprfc2(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *tol, sprgpu_int *iwork, double *rwork, const sprgpu_int *lpu,
                   sprgpu_int *ierr, void (*mvmult)(void), void (*mmmult)(void));
void sprgpu_sprfs2(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *b, const sprgpu_int *ldb, const sprgpu_int *nrhs, sprgpu_int *iwork,
                   double *rwork, const sprgpu_int *lpu, sprgpu_int *ierr);
void sprgpu_sprbs2(sprgpu_int *mspar, sprgpu_int *mtrees, sprgpu_int *msifa, double *sm,
                   double *b, const sprgpu_int *ldb, const sprgpu_int *nrhs, double *rwork,
                   const sprgpu_int *lpu, sprgpu_int *ierr);

/* Replaces SPRSOL (src/SPR/sprsol.f:5-6): IOP<=3 factor, 2<=IOP<=5 forward,
 * IOP>=3 and IOP!=5 backward. */
void sprgpu_sprsol(const sprgpu_int *iop, sprgpu_int *mspar, sprgpu_int *mtrees,
                   sprgpu_int *msifa, double *sm, double *b, const sprgpu_int *ldb,
                   const sprgpu_int *nrhs, double *tol, sprgpu_int *iwork, double *rwork,
                   const sprgpu_int *lpu, sprgpu_int *ierr);

/* Releases the device factor that the drop-in layer keeps for SM. */
void sprgpu_release(const double *sm);

/* LIERR of the last drop-in factorization on the calling thread: the
 * reference's internal error code behind IERR = -3 (sprfc1.f:250-262,
 * sprfc2.f:244-255, sprfc3.f:250-262, sprfc4.f:240-256): -30 a pivot at or
 * below TOL(1)|a_jj| (SPRFC1), -32 the same in SPRFC2 and a pivot below
 * -TOL(1)|a_jj| in SPRFC4, -31 pivots replaced (SPRFC3), -33 pivots
 * replaced (SPRFC4), -35 a zero diagonal found before the factorization; 0
 * otherwise.  The Fortran layer hands it to SPRER1 as the reference drivers
 * do. */
sprgpu_int sprgpu_dropin_lierr(void);

/* ------------------------------------------------------------------------ */
/* 2. Device-resident handle API                                            */
/* ------------------------------------------------------------------------ */

typedef struct sprgpu_handle sprgpu_handle;

enum sprgpu_precision { SPRGPU_FP64 = 0, SPRGPU_FP32 = 1 };

/* Factor statistics, the reference's outputs (sprfc1.f:60-80). */
typedef struct {
    sprgpu_int ierr;     /* 0, 3 (NUMNEG>0), -3 singular, -5 HIP error */
    sprgpu_int numneg;   /* MSPAR(26) */
    sprgpu_int supeqt;   /* MSPAR(27): failing equation, or -#replacements (SPRFC3) */
    double tol2;         /* TOL(2) */
    double tol3;         /* TOL(3) */
    sprgpu_int nzero_diag; /* zero diagonals found before factorization */
} sprgpu_factor_info;

typedef struct {
    sprgpu_int iterations; /* refinement steps taken */
    double relres;         /* final max over columns of ||b - A x||_2 / ||b||_2 */
    sprgpu_int converged;  /* 1 if relres <= requested tolerance */
} sprgpu_ir_info;

/* Builds the device plan from the symbolic arrays.  precision selects the
 * factor storage/arithmetic.  Returns 0 or a negative HIP/argument error. */
int sprgpu_create(const sprgpu_int *mspar, const sprgpu_int *mtrees, const sprgpu_int *msifa,
                  int precision, sprgpu_handle **handle);
int sprgpu_destroy(sprgpu_handle *handle);

/* hipStream_t to run on (NULL = default stream). */
int sprgpu_set_stream(sprgpu_handle *handle, void *stream);

/* Device-memory pattern of A used by the FP64 residual.  Must be called once
 * with a host copy of SM (assembled, before factorization) before
 * sprgpu_solve_ir or sprgpu_spmv; later factorizations check that no entry
 * outside this pattern became non-zero and rebuild it if one did. */
int sprgpu_set_pattern(sprgpu_handle *handle, const double *sm_host);

/* Element assembly on the device, replacing a host loop that calls SPRADM
 * (src/SPR/spradm.f:5-166) once per element with the free-dof loop of its
 * SPRAD2 (spradm.f:818-859) and SPRAD1's slot table (spradm.f:514-740).
 *
 * sprgpu_set_assembly (host arrays, once per connectivity): element e
 * (0-based, in the order SPRADM would be called) has the dofs whose 1-based
 * SPR equations are meen[meen_ptr[e] .. meen_ptr[e+1]) in element-matrix
 * order, 0 for a dof without one: SPREEQ's MEEN (spradm.f:169-221).  Builds
 * the map from element entries to LNZ slots and sets the residual pattern
 * to the slots that receive an entry (so sprgpu_set_pattern is not needed).
 * Returns -2 for a constrained dof (MEEN < 0: use sprgpu_set_assembly_ceq
 * below), an equation beyond NEQ, or an element coupling two equations
 * outside the symbolic pattern.
 *
 * sprgpu_assemble (device arrays, asynchronous on the handle's stream):
 * SM(NEQ+1:) receives sum_e EM_e(I,J) over I <= J, in SPRADM's order (by
 * element, then J, then I), so the result equals the host loop's bit for
 * bit.  EM_e is column-major nedof_e x nedof_e at em_dev + em_off_dev[e]
 * (em_off_dev NULL: the matrices follow each other), multiplied by
 * escale_dev[e] when escale_dev is not NULL, as the host loop does before it
 * calls SPRADM.  accumulate 0: SM(1:NEQ+NZEROL) is cleared first (the
 * caller's zeroing before the loop); 1: added to the SM given. */
int sprgpu_set_assembly(sprgpu_handle *handle, sprgpu_int nel, const sprgpu_int *meen_ptr,
                        const sprgpu_int *meen);
int sprgpu_assemble(sprgpu_handle *handle, const double *em_dev, const sprgpu_int *em_off_dev,
                    const double *escale_dev, double *sm_dev, int accumulate);

/* Element stiffness matrices of trilinear hexahedra on the device, the
 * matrices a host loop computes before each SPRADM call (spradm.f:5-13 takes
 * EM from its caller): isotropic linear elasticity, 2x2x2 Gauss quadrature,
 * K_e = sum_g det(J_g) B_g' D B_g with engineering shear strains.  Element e
 * has the 1-based nodes mmnpc_dev[8e .. 8e+7] (local order (-1,-1,-1)
 * (1,-1,-1) (1,1,-1) (-1,1,-1) (-1,-1,1) (1,-1,1) (1,1,1) (-1,1,1)), node k
 * the coordinates xyz_dev[3(k-1) .. 3(k-1)+2], modulus emod_dev[e] (NULL:
 * 1) and Poisson's ratio nu; its dofs are its nodes in that order, each
 * (ux, uy, uz).  em_dev receives nel column-major 24x24 matrices one after
 * the other (sprgpu_assemble with em_off_dev NULL takes them so).  A node
 * number outside 1..nanod gives a NaN matrix.  Asynchronous on the handle's
 * stream; -2 for bad arguments or nu outside (-1, 0.5).  No counterpart in
 * SPR itself (element routines are the caller's); it completes the device
 * assembly path. */
int sprgpu_hex8_stiffness(sprgpu_handle *handle, sprgpu_int nel, sprgpu_int nanod,
                          const double *xyz_dev, const sprgpu_int *mmnpc_dev,
                          const double *emod_dev, double nu, double *em_dev);

/* Element assembly with constraint equations: the whole of SPRAD2
 * (src/SPR/spradm.f:743-1002) for a host loop that calls SPRADM with
 * NSV = nsv, TTCC, MPMCEQ, MMCEQ and MEQN (spradm.f:5-13).
 *
 * sprgpu_set_assembly_ceq (host arrays): as sprgpu_set_assembly, where
 * meen may also hold -ICEQ for a constrained dof (SPREEQ's MEEN), ICEQ
 * naming equation ICEQ of mpmceq(nceq+1)/mmceq (SAM's MPMCEQ and MMCEQ:
 * equation I holds its constrained dof at MPMCEQ(I) and its masters after
 * it), meqn(ndof) SAM's MEQN (the masters' equations).  nsv: SPRADM's NSV
 * -- 0: SM only; 1: SM and one vector SV(NEQ); -1: SV only; |nsv| > 1: SV
 * holds |nsv| vectors, one per prescribed equation (no masters) in equation
 * order, and slaves add nothing to SV (spradm.f:866-881).  Returns -2 for
 * the errors of sprgpu_set_assembly, MPMCEQ(1) != 1 or an equation without
 * its constrained dof, a master with MMCEQ in 1..NDOF whose MEQN is not in
 * 1..NEQ, -ICEQ beyond NCEQ, or (|nsv| > 1) an element prescribed dof
 * whose vector lies beyond |nsv|.  The map depends on MPMCEQ/MMCEQ/MEQN only;
 * TTCC is read at every assembly, so C0 and the coefficients may change.
 *
 * sprgpu_assemble_ceq (device arrays, asynchronous): for nsv >= 0, SM as
 * sprgpu_assemble plus, per constrained dof J and master JP of it,
 * TTCC(JP)*EM(I,J) at (MEEN(I), MEQN(master)) (2*TTCC(JP) on the master's
 * own diagonal) and TTCC(IP)*TTCC(JP)*EM(I,J) for each master IP of a
 * constrained I with MEQN(IP's dof) <= MEQN(JP's dof); for nsv != 0,
 * SV(NEQ*|nsv|) -= C0*EM(I,J) (C0 = TTCC(MPMCEQ(JCEQ))) at the free I and,
 * weighted by TTCC(IP), at the masters of a constrained I.  Every position
 * sums its terms in SPRADM's order, so SM and SV equal the host loop's bit
 * for bit when the element matrices are finite and the output is cleared
 * (accumulate 0).  The host loop skips a term whose TTCC entry is zero; the
 * device adds it (a signed zero), which changes no bits of a sum that starts
 * at +0, but can turn a -0 of an accumulated input into +0.
 * ttcc_dev: TTCC(NMMCEQ) in device memory (may be NULL if no element has a
 * constrained dof); sm_dev may be NULL for nsv < 0, sv_dev for nsv = 0.
 * accumulate 0 clears SM (nsv >= 0) and SV (nsv != 0) first. */
int sprgpu_set_assembly_ceq(sprgpu_handle *handle, sprgpu_int nel, const sprgpu_int *meen_ptr,
                            const sprgpu_int *meen, sprgpu_int nceq, const sprgpu_int *mpmceq,
                            const sprgpu_int *mmceq, sprgpu_int ndof, const sprgpu_int *meqn,
                            sprgpu_int nsv);
int sprgpu_assemble_ceq(sprgpu_handle *handle, const double *em_dev, const sprgpu_int *em_off_dev,
                        const double *escale_dev, const double *ttcc_dev, double *sm_dev,
                        double *sv_dev, int accumulate);

/* Numeric factorization of SM given in device memory (NEQ+NZEROL doubles, the
 * SPRADM layout; not modified).  tol points to host TOL(1..4) (TOL(4) is read
 * only when mode is 3 or 4).  mode 1 = SPRFC1 semantics, 3 = SPRFC3
 * semantics, 2 = SPRFC2 (src/SPR/sprfc2.f:5-7, the CC' form: a pivot fails
 * when d <= TOL(1)*|a_jj|, sign included, sprfc2.f:934-1037; IERR, SUPEQT,
 * TOL(2:3) and MSPAR(26) as SPRFC2 reports them), 4 = SPRFC4 (sprfc4.f:5-7,
 * the CC' form with replacement, see sprgpu_sprfc4; DCOR from
 * sprgpu_get_dcor).  After a mode-2 or mode-4 factorization, sprgpu_solve
 * with which 1 is SPRFS2 (C^-1 B), 2 is SPRBS2 (C^-T B), 3 both, and
 * sprgpu_get_factor gives C in SM layout.  The CC' factor is C = L sqrt(D)
 * of the LDL' factorization, so it agrees with the reference's to rounding,
 * not bit for bit.
 * info may be NULL (then the call returns without synchronizing). */
int sprgpu_factor(sprgpu_handle *handle, const double *sm_dev, const double *tol, int mode,
                  sprgpu_factor_info *info);

/* Reads the factor statistics of the last factorization (synchronizes). */
int sprgpu_factor_info_get(sprgpu_handle *handle, sprgpu_factor_info *info);

/* DCOR(2,NEQ) of the last mode-3 (SPRFC3) or mode-4 (SPRFC4) factorization,
 * the array SPRFC3/SPRFC4 return (sprfc3.f:5-7, filled at sprfc3.f:560-575
 * and :1083-1099; sprfc4.f:459-472, :1016-1037): DCOR(1,j) the pivot before
 * its replacement (or a negative pivot: SPRFC3's kept one, SPRFC4's failing
 * one), DCOR(2,j) the replacement (1 for a zero diagonal; SPRFC4's failing
 * pivot again), zero elsewhere.  dcor_out (2*NEQ doubles) may be host or
 * device memory.  Returns -1 if the last factorization was not in mode 3
 * or 4. */
int sprgpu_get_dcor(sprgpu_handle *handle, double *dcor_out);

/* Forward (D^-1 L^-1) and/or backward (L^-T) substitution in place on device
 * B(LDB,NRHS) in FP64 (the factor may be FP32).  which: 1 forward, 2
 * backward, 3 both. */
int sprgpu_solve(sprgpu_handle *handle, double *b_dev, sprgpu_int ldb, sprgpu_int nrhs,
                 int which);

/* Partial solve, the reference's MSPAR(57) = 1 with LDB < NEQ
 * (src/SPR/sprfs1.f:200-201, sprbs1.f:189, sprfs2.f:203, sprbs2.f:191): B
 * holds only the leading LDB equations; the others enter the sweeps as zero
 * and their results are dropped (SPRFB1/SPRFB2, sprfs1.f:636-848).  The
 * forward result is the leading part of L^-1 b padded with zeros, the
 * backward one solves with the trailing equations held at zero, so which 3
 * gives A11^-1 b1 for the leading block A11.  ldb >= NEQ is sprgpu_solve.
 * With PERI2E (MSPAR(58) >= 1) the leading equations are external ones and
 * must be the leading internal ones (-2 otherwise).  Structures with
 * retained supernodes (SUPSUP > 0) are rejected when the handle is created. */
int sprgpu_solve_partial(sprgpu_handle *handle, double *b_dev, sprgpu_int ldb, sprgpu_int nrhs,
                         int which);

/* sprgpu_solve with no allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory during a call -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments, semantics and results as sprgpu_solve (bit for bit), but
 * the solve's work space (the per-supernode vector slots, partial sums, a
 * spin-error word and the PERI2E copy of B) is taken from a
 * static device array of 64 MiB that each GPU holds for the life of the
 * process, instead of buffers the handle grows.  A call whose work space
 * does not fit (about 50 MB per right-hand side at 40^3 hex8) returns -2
 * with the sizes in sprgpu_last_error and touches nothing.  Every call on a
 * GPU shares its array: calls that overlap on one GPU (other streams,
 * threads or handles) race. */
int sprgpu_solve_noalloc(sprgpu_handle *handle, double *b_dev, sprgpu_int ldb, sprgpu_int nrhs,
                         int which);

/* Y = A X with the assembled matrix of the last factorization, FP64. */
int sprgpu_spmv(sprgpu_handle *handle, const double *x_dev, double *y_dev, sprgpu_int ld,
                sprgpu_int nrhs);

/* Replaces SPRPRM (src/SPR/sprprm.f:5-7) on device arrays: for each of the N
 * columns of B(M,N), IFLAG +-1: C = +-A B; +-2: C = C +- A B; 88 (N = 1):
 * WA = A B and C is not referenced.  KSA > 0: A is a symmetric matrix in the
 * SPR layout, the NZEROL doubles of LNZ (SM(NEQ+1:), as SPRSMM passes it);
 * its non-zeros must lie on the handle's pattern (sprgpu_set_pattern or
 * sprgpu_set_assembly; checked after the product: -1, and C not valid, if
 * not).  KSA <= 0: A is a diagonal of NEQ doubles.  wa_dev (M doubles, may be NULL unless IFLAG = 88) receives
 * the last column's +-A b, as the reference's WA does.  Each C(r) is summed
 * in the reference's order (row r in ascending column order, every product
 * rounded before it is added), so the result equals SPRPRM's, up to the sign
 * of a zero sum.  With MSPAR(58) >= 1 B, C and WA are in the external
 * (PERI2E) order, as SPRPRR takes them (sprprm.f:327-397).  Returns 0, -2
 * (IFLAG, N, missing C/WA), -3 (KSA <= 0 with PERI2E), -4 (M < NEQ) as the
 * reference's IERR, or -1 (pattern). */
int sprgpu_sprprm(sprgpu_handle *handle, const double *a_dev, const double *b_dev,
                  double *c_dev, double *wa_dev, sprgpu_int m, sprgpu_int n, sprgpu_int ksa,
                  sprgpu_int iflag);

/* Replaces SPRTRA (src/SPR/sprtra.f:5-56) on device arrays: C = B' A B for
 * B(M,N) and A as in sprgpu_sprprm (KSA > 0 the SPR layout, KSA <= 0 a
 * diagonal); C is N x N, column-major, symmetric.  As the reference: per
 * column J, WA = A B(:,J) by SPRPRM (IFLAG 88; wa_dev, M doubles, holds the
 * last column's on return), then C(I,J) = C(J,I) = B(:,I) . WA for I >= J.
 * The dot products are summed in a fixed order of this library's own (the
 * reference calls the host BLAS DDOT, whose order is the BLAS build's), so
 * C agrees with the reference's to rounding.  Returns 0, -2 (arguments) or
 * sprgpu_sprprm's error. */
int sprgpu_sprtra(sprgpu_handle *handle, const double *a_dev, const double *b_dev,
                  double *c_dev, double *wa_dev, sprgpu_int m, sprgpu_int n, sprgpu_int ksa);

/* Replaces SPRTMM (src/SPR/sprtmm.f:5-7) on device arrays.  With L the stored
 * lower triangle, diagonal included, of U (NZEROL doubles in the LNZ layout:
 * an assembled matrix or a factor), for each of the N columns of B(M,N):
 * IFLAG 1: C = L^T B; 2: C = L B; 11 / 22 (N = 1): the same into WA only.
 * KSU <= 0: U is a diagonal of NEQ doubles.  wa_dev (M doubles; may be NULL
 * unless IFLAG is 11 or 22) receives the last column, as the reference's WA.
 * Every stored entry is used, in the reference's order, so C equals
 * SPRTMM's bit for bit.  With MSPAR(58) >= 1 B, C and WA are in the
 * external (PERI2E) order, as SPRTMR takes them (sprtmm.f:345-480).  Returns
 * 0, -2 (IFLAG, N, missing C/WA) or -3 (M < NEQ) as the reference's IERR. */
int sprgpu_sprtmm(sprgpu_handle *handle, const double *u_dev, const double *b_dev,
                  double *c_dev, double *wa_dev, sprgpu_int m, sprgpu_int n, sprgpu_int ksu,
                  sprgpu_int iflag);

/* The task of SPRLAN (src/SPR/sprlan.f:5-945) on the device: the NEVAL
 * eigenvalues of (A - lambda B) q = 0 closest to SHIFT and, for the first
 * NEVEC of them, B-orthonormal eigenvectors.  a_sm_dev: A as SM
 * (NEQ+NZEROL, assembled); b_dev: B as LNZ (NZEROL, ksb > 0; its non-zeros
 * on the handle's pattern) or a diagonal (NEQ, ksb <= 0).  The handle (FP64)
 * factors A - SHIFT B and keeps that factor afterwards.  Lanczos on
 * (A - SHIFT B)^-1 B with full reorthogonalisation in the B inner product,
 * at most MAXLAN steps (NEVAL < MAXLAN <= NEQ), stopping when the NEVAL
 * estimates EVERR(i) = |beta_m s_mi| / theta_i^2 are all at most
 * TOL |EVL(i)|.  evl, everr: host arrays of NEVAL, nearest to SHIFT first;
 * v_dev: NEQ x NEVEC column-major, device; nconv: how many converged.
 * SPRLAN's option arrays (MIP, MOP), restarts and other reorthogonalisation
 * schemes are not offered.  Returns 0, -2 (arguments, FP32 handle), -1 (no
 * pattern for a sparse B) or -3 (A - SHIFT B singular). */
int sprgpu_lanczos(sprgpu_handle *handle, const double *a_sm_dev, const double *b_dev, int ksb,
                   double shift, sprgpu_int neval, sprgpu_int nevec, sprgpu_int maxlan,
                   double tol, double *evl, double *everr, double *v_dev, sprgpu_int *nconv);

/* Solves A X = B with the current factor plus FP64 iterative refinement
 * until max_col ||B - A X|| / ||B|| <= rtol or maxit steps. */
int sprgpu_solve_ir(sprgpu_handle *handle, const double *b_dev, double *x_dev, sprgpu_int ld,
                    sprgpu_int nrhs, double rtol, int maxit, sprgpu_ir_info *info);

/* Adjoint gradients (no counterpart in the reference, which has no adjoint;
 * lambda = A^-1 dL/du comes from sprgpu_solve / sprgpu_solve_ir with the same
 * factor, A being symmetric).  All arrays are device memory, vectors in SPR
 * equation order.
 *
 * sprgpu_element_grad: for K = sum_e p_e K0_e, grad[e] = -lambda_e^T K0_e x_e
 * for nel elements of nedof (<= 64) dofs each.  elem_eq[e*nedof + i] is the
 * 0-based equation of element dof i or -1 (no equation: a fixed dof).  K0 is
 * k0 + k0_index[e]*nedof*nedof, column-major (k0_index NULL: one matrix
 * shared by all elements). */
int sprgpu_element_grad(sprgpu_handle *handle, sprgpu_int nel, sprgpu_int nedof,
                        const sprgpu_int *elem_eq, const double *k0,
                        const sprgpu_int *k0_index, const double *lambda, const double *x,
                        double *grad);

/* sprgpu_pattern_grad: dL/dA on the stored lower triangle in SM layout:
 * grad_sm(NEQ + LNZ position of (i,j)) = -(lambda_i x_j + lambda_j x_i) for
 * i > j and -lambda_i x_i on the diagonal; grad_sm(1:NEQ) is not written.
 * grad_sm holds NEQ+NZEROL doubles. */
int sprgpu_pattern_grad(sprgpu_handle *handle, const double *lambda, const double *x,
                        double *grad_sm);

/* Copies the factor into SM layout (NEQ+NZEROL doubles): d in SM(1:NEQ) and
 * the packed trapezoids with d on the diagonal and l_ij below.  sm_out may be
 * host or device memory. */
int sprgpu_get_factor(sprgpu_handle *handle, double *sm_out);

/* Size queries. */
sprgpu_int sprgpu_neq(const sprgpu_handle *handle);
sprgpu_int sprgpu_nzerol(const sprgpu_handle *handle);
/* Bytes of device memory held by the handle. */
sprgpu_int sprgpu_device_bytes(const sprgpu_handle *handle);

/* sprgpu_factor with a HIP event after every kernel (slower; for
 * measurement).  out[3*c+0..2] (c < 9, so out holds 27 values) =
 * milliseconds, kernel launches, algorithmic flops for kernel class c:
 * 0 front init, 1 extend-add, 2 LDS small fronts, 3 diagonal blocks
 * (128-column outer blocks, with the rows below them when one launch does
 * both; with panel steps, the write-back), 4 rows below
 * them (outer-block TRSM; or the panel steps), 5 MFMA rank-256 and Schur
 * updates on 64-row tiles, 6 other, 7 MFMA rank-128 updates inside a
 * 256-column group, 8 MFMA rank-256 and Schur updates on 128-row tiles. */
int sprgpu_profile_factor(sprgpu_handle *handle, const double *sm_dev, const double *tol,
                          int mode, double *out);

/* Host-only: builds the device schedule for a symbolic structure without
 * touching a GPU and reports it, for checking and sizing.  out[0..9] =
 * NEQ, NSUPER, tree heights, kernel launches per factorization, front-pool
 * elements (the fronts alive at once; the factored panels are kept in a
 * separate store of sum ld*S elements), solve-workspace elements per RHS, fronts factored in LDS,
 * outer blocks (or panel steps), MFMA update tiles, factor flops (RINFO(3) formula,
 * sprtrs.f:1961-1969).  Returns 0 or -2 (structure not supported). */
int sprgpu_plan_summary(const sprgpu_int *mspar, const sprgpu_int *mtrees,
                        const sprgpu_int *msifa, double *out);

/* Factor (SPRFC1 pivot rules, the handle's precision) followed by the
 * forward and backward solves of B(LDB,NRHS), as one hipGraph: captured on
 * the first call and replayed by later calls with the same sm_dev, b_dev, ldb
 * and nrhs (their contents may change; other pointers capture a new graph).
 * No counterpart in the reference; for callers that repeat factor+solve on
 * new values of one structure (batches of small systems, time stepping).
 * Errors of the factor are read afterwards with sprgpu_factor_info_get.
 * A structure whose fronts all have F <= 64 (a forest of small systems) on
 * the forest schedule (sprgpu_set_plan_params) with one right-hand side is
 * factored and solved by one persistent launch. */
int sprgpu_factor_solve_graph(sprgpu_handle *handle, const double *sm_dev, double *b_dev,
                              sprgpu_int ldb, sprgpu_int nrhs);

/* Chooses the schedule parameters (small-front threshold, fronts per
 * workgroup, narrow-solve workgroup size, FP32 update tile, forest schedule)
 * for this structure by timing factor + solve on sm_dev / b_dev for each
 * candidate, and keeps the fastest.  b_dev is overwritten.  report (NULL or
 * 256 doubles): report[0] = candidates, then per candidate small_f, waves,
 * narrow threads, tile, forest, milliseconds.  No counterpart in the
 * reference. */
int sprgpu_autotune(sprgpu_handle *handle, const double *sm_dev, double *b_dev, sprgpu_int ldb,
                    sprgpu_int nrhs, double *report);
/* The parameters in use: out[0..5] = small_f, fronts per workgroup, narrow
 * threads (0 = by size), update tile, blocked-front scheme (1 = 128-column
 * outer blocks, the diagonal block and the rows below it in one launch;
 * 2 = the same as two launches; 0 = 32-column panel steps), forest schedule
 * (requested by default: a structure whose fronts all have F <= 64 is
 * factored by one persistent launch over all supernodes, and its
 * one-right-hand-side solves too; out[5] = 1 when the schedule is in use, 0
 * for the level schedule: one launch per tree height and kernel kind). */
int sprgpu_plan_params(const sprgpu_handle *handle, int *out);
/* Re-plans with the given parameters, in the order of sprgpu_plan_params
 * (update tile 0 = per launch: 128 for FP32 launches of at least 4096 64-tiles,
 * else 64).  Returns 0 or -2 (a value out of range).
 * No counterpart in the reference; for tuning and comparisons. */
int sprgpu_set_plan_params(sprgpu_handle *handle, const int *params);

/* Last error message (thread-local, empty if none). */
const char *sprgpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
