/*
 * CPU restatement of the SPR numeric path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker for the HIP library: tests compare the GPU
 * results with these functions on the same inputs.  Nothing in the product
 * calls or links it.
 *
 * Every routine follows the operation order of the reference Fortran it
 * cites (loop order, remainder-first x8 unrolling, left-to-right sums,
 * reciprocal scaling versus true division), so that with FMA contraction
 * off it reproduces the reference bit for bit.  Arrays keep the reference's
 * 1-based addressing through the F1() macro to make the correspondence with
 * the Fortran easy to audit.
 *
 * Not restated: the multifrontal stack overflow checks (-10/-20/-40 in
 * sprfc1.f:508-513,613-618,672-680; the stack is not modelled, update
 * matrices are kept per supernode instead) and the PERI2E-permuted solve
 * variants SPRFS4/SPRBS5 (MSPAR(58)>=1).
 */
#include "spr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F1(a, k) ((a)[(k) - 1])
/* OFFSET(I,J) = I*J - J*(J-1)/2  (sprfc1.f:861, :1092) */
#define OFFSET(i, j) ((i) * (j) - ((j) * ((j) - 1)) / 2)

/* ------------------------------------------------------------------------ */
/* LMVY18: y <- y - sum_k (a_{r,k} d_k) a(:,k), x8 unrolled (lmvy18.f:5-173) */
/* ------------------------------------------------------------------------ */
static void lmvy18(fint m, fint n, double *y, const double *d, const fint *apnt, const double *a)
{
    const fint m1 = m + 1;
    const fint rem = n % 8;
    if (rem > 0) {
        fint ii[8];
        double aa[8];
        for (fint k = 1; k <= rem; ++k) {
            ii[k - 1] = F1(apnt, k + 1) - m1;
            aa[k - 1] = -(F1(a, ii[k - 1] + 1) * F1(d, k));
        }
        for (fint i = 1; i <= m; ++i) {
            double t = F1(y, i);
            for (fint k = 0; k < rem; ++k) t = t + aa[k] * F1(a, ii[k] + i);
            F1(y, i) = t;
        }
    }
    for (fint j = rem + 1; j <= n; j += 8) {
        fint ii[8];
        double aa[8];
        for (fint k = 0; k < 8; ++k) {
            ii[k] = F1(apnt, j + k + 1) - m1;
            aa[k] = -(F1(a, ii[k] + 1) * F1(d, j + k));
        }
        for (fint i = 1; i <= m; ++i) {
            double t = F1(y, i);
            for (fint k = 0; k < 8; ++k) t = t + aa[k] * F1(a, ii[k] + i);
            F1(y, i) = t;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* LMMY18: packed trailing triangle Y <- Y - X D X^T (lmmy18.f:5-192)       */
/* ------------------------------------------------------------------------ */
static void lmmy18(fint m, fint n, fint q, const fint *xpnt, const double *x, const double *d,
                   double *y, fint ldy)
{
    fint m1 = m, ly = ldy, iyl = 0;
    const fint rem = n % 8;
    for (fint ycol = 1; ycol <= q; ++ycol) {
        const fint iy1 = iyl + 1;
        const fint iy2 = iy1 + m1 - 1;
        iyl += ly;
        const fint m2 = iy1 + m1;
        if (rem > 0) {
            fint ii[8];
            double aa[8];
            for (fint k = 1; k <= rem; ++k) {
                ii[k - 1] = F1(xpnt, k + 1) - m2;
                aa[k - 1] = -(F1(x, ii[k - 1] + iy1) * F1(d, k));
            }
            for (fint i = iy1; i <= iy2; ++i) {
                double t = F1(y, i);
                for (fint k = 0; k < rem; ++k) t = t + aa[k] * F1(x, ii[k] + i);
                F1(y, i) = t;
            }
        }
        for (fint j = rem + 1; j <= n; j += 8) {
            fint ii[8];
            double aa[8];
            for (fint k = 0; k < 8; ++k) {
                ii[k] = F1(xpnt, j + k + 1) - m2;
                aa[k] = -(F1(x, ii[k] + iy1) * F1(d, j + k));
            }
            for (fint i = iy1; i <= iy2; ++i) {
                double t = F1(y, i);
                for (fint k = 0; k < 8; ++k) t = t + aa[k] * F1(x, ii[k] + i);
                F1(y, i) = t;
            }
        }
        m1 -= 1;
        ly -= 1;
    }
}

/* ------------------------------------------------------------------------ */
/* SPRF17 / SPRF37: left-looking LDL' of the first n columns of a packed    */
/* front (sprfc1.f:1204-1299; sprfc3.f:1011-1118 when dcor != NULL).         */
/* Returns the SPRF17 error (-jcol) or, for SPRF37, adds to *nrepl.          */
/* ------------------------------------------------------------------------ */
static fint sprf17(fint m, fint n, const fint *xpnt, double *x, double *lzdiag, double *dcor,
                   double *tol, fint *numneg, fint *nrepl)
{
    fint mm = m;
    fint jpnt = F1(xpnt, 1);
    for (fint jcol = 1; jcol <= n; ++jcol) {
        if (jcol > 1) lmvy18(mm, jcol - 1, &F1(x, jpnt), lzdiag, xpnt, x);
        const double adiag = F1(lzdiag, jcol);
        double diag = F1(x, jpnt);
        if (fabs(diag) <= tol[0] * fabs(adiag)) {
            if (!dcor) {
                tol[1] = diag;
                tol[2] = adiag;
                return -jcol;
            }
            /* sprfc3.f:1083-1094: insert a small pivot with the sign of d */
            *nrepl += 1;
            dcor[2 * (jcol - 1)] = diag;
            double mag = tol[0] * fmax(1.0, fabs(tol[3] * adiag));
            diag = copysign(mag, diag);
            F1(x, jpnt) = diag;
            dcor[2 * (jcol - 1) + 1] = diag;
        }
        if (diag < 0.0) {
            *numneg += 1;
            if (dcor && dcor[2 * (jcol - 1)] == 0.0) dcor[2 * (jcol - 1)] = diag;
        }
        F1(lzdiag, jcol) = diag;
        /* TOL(2) = MAX(TOL(2), ADIAG/DIAG) */
        const double ratio = adiag / diag;
        if (ratio > tol[1]) tol[1] = ratio;
        diag = 1.0 / diag;
        mm -= 1;
        jpnt += 1;
        for (fint i = 1; i <= mm; ++i) {
            F1(x, jpnt) = F1(x, jpnt) * diag;
            jpnt += 1;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* SPRF15 / SPRF35: eliminate n columns of an order-m packed front and form  */
/* the update matrix (sprfc1.f:985-1185; sprfc3.f:809-1008).                 */
/* ------------------------------------------------------------------------ */
static fint sprf15(fint m, fint n, int panel, const fint *split, double *x, double *diag,
                   double *dcor, double *tol, fint *numneg, fint *nrepl, fint *xpnt)
{
    if (m <= 0 || n <= 0) return 0;
    for (fint i = 1; i <= m; ++i) F1(xpnt, i) = OFFSET(m, i - 1) + 1;
    F1(xpnt, m + 1) = F1(xpnt, m) + 1;
    fint err = 0;
    if (panel) {
        fint jblk = 1, fstcol = 1, mm = m;
        fint fstblk = F1(split, jblk);
        while (fstcol <= n) {
            jblk += 1;
            const fint nn = F1(split, jblk) - fstblk;
            fstblk = F1(split, jblk);
            err = sprf17(mm, nn, &F1(xpnt, fstcol), x, &F1(diag, fstcol),
                         dcor ? &dcor[2 * (fstcol - 1)] : NULL, tol, numneg, nrepl);
            if (err != 0) return err;
            const fint nxtcol = fstcol + nn;
            const fint q = m - nxtcol + 1;
            mm -= nn;
            const fint jpnt = F1(xpnt, nxtcol);
            if (q > 0) lmmy18(mm, nn, q, &F1(xpnt, fstcol), x, &F1(diag, fstcol), &F1(x, jpnt), mm);
            fstcol = nxtcol;
        }
    } else {
        err = sprf17(m, n, xpnt, x, diag, dcor, tol, numneg, nrepl);
        if (err != 0) return err;
        const fint nxtcol = 1 + n;
        const fint q = m - nxtcol + 1;
        const fint mm = m - n;
        const fint jpnt = F1(xpnt, nxtcol);
        if (q > 0) lmmy18(mm, n, q, xpnt, x, diag, &F1(x, jpnt), mm);
    }
    return 0;
}

/* SPRF12: child -> parent local indices by a monotone merge (sprfc1.f:716-790) */
static void sprf12(fint suplen, fint sublen, const fint *supvec, const fint *subvec, fint *lindex)
{
    fint supnod = 0;
    for (fint isub = 1; isub <= sublen; ++isub) {
        const fint subval = F1(subvec, isub);
        do {
            supnod += 1;
        } while (supnod <= suplen && F1(supvec, supnod) < subval);
        F1(lindex, isub) = supnod;
    }
}

/* SPRF13/SPRF14: extend-add a packed child update matrix into the packed
 * parent front (sprfc1.f:793-982).  Each parent entry receives one addition
 * per child, so the two-columns-at-a-time traversal of the reference and the
 * split into NEWLNZ/ANCLNZ (contiguous in the front) do not change values. */
static void sprf13(fint lenvar, fint lenldx, const fint *lindex, const double *update,
                   double *front)
{
    for (fint j = 1; j <= lenldx; ++j) {
        const fint jind = F1(lindex, j);
        const fint jpnt = 1 - jind + OFFSET(lenvar, jind - 1);
        fint jchd = 1 + OFFSET(lenldx, j - 1);
        for (fint i = j; i <= lenldx; ++i) {
            F1(front, jpnt + F1(lindex, i)) += F1(update, jchd);
            jchd += 1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* SPRF11 / SPRF31: the supernodal multifrontal loop in postorder            */
/* (sprfc1.f:286-713; sprfc3.f:372-806).                                     */
/* ------------------------------------------------------------------------ */
static fint sprf11(fint nsuper, fint neq, fint nblock, fint lpar57, fint *numneg,
                   const fint *xsuper, const fint *xsibl, const fint *sibl, const fint *xlindx,
                   const fint *supsup, const fint *lindx, const fint *xsplit, const fint *split,
                   const fint *xlnz, double *diag, double *lnz, double *dcor, double *tol, fint mfront,
                   fint *supeqt)
{
    fint error = 0, nrepl = 0;
    *supeqt = 0;
    *numneg = 0;
    fint iplnz = 1;
    const int panel = nblock > 0;
    tol[1] = 0.0;
    tol[2] = 0.0;

    if (lpar57 > 0) {
        for (fint i = 1; i <= neq; ++i) F1(diag, i) = 0.0;
        if (dcor)
            for (fint i = 0; i < 2 * neq; ++i) dcor[i] = 0.0;
    }
    /* Initialise DIAG with a_jj; zero-pivot scan (sprfc1.f:463-488) or
     * zero-pivot replacement by one (sprfc3.f:548-575). */
    {
        for (fint js = 1; js <= nsuper; ++js) {
            if (!(F1(supsup, js) > 0 && lpar57 > 0)) {
                fint chdlen = F1(xlindx, js + 1) - F1(xlindx, js);
                fint ipjs = F1(xlnz, js);
                for (fint j = F1(xsuper, js); j < F1(xsuper, js + 1); ++j) {
                    F1(diag, j) = F1(lnz, ipjs);
                    if (dcor) {
                        dcor[2 * (j - 1)] = 0.0;
                        dcor[2 * (j - 1) + 1] = 0.0;
                        if (F1(diag, j) == 0.0) {
                            F1(diag, j) = 1.0;
                            dcor[2 * (j - 1) + 1] = 1.0;
                            F1(lnz, ipjs) = 1.0;
                        }
                    } else if (F1(diag, j) == 0.0) {
                        *numneg += 1;
                        if (*numneg == 1) {
                            *supeqt = j;
                            error = -35;
                        }
                    }
                    tol[2] = tol[2] + F1(diag, j);
                    ipjs += chdlen;
                    chdlen -= 1;
                }
            }
        }
    }
    if (error < 0) return error;

    /* Update matrices of finished children, packed (LENLDX*(LENLDX+1)/2). */
    double **upd = (double **)calloc((size_t)nsuper + 1, sizeof(double *));
    fint *updlen = (fint *)calloc((size_t)nsuper + 1, sizeof(fint));
    fint *first = (fint *)malloc((size_t)(mfront + 2) * sizeof(fint));
    fint *lindex = (fint *)malloc((size_t)(mfront + 2) * sizeof(fint));

    for (fint js = 1; js <= nsuper; ++js) {
        const fint fstvar = F1(xsuper, js);
        const fint supvar = F1(xsuper, js + 1) - fstvar;
        const fint lenvar = F1(xlindx, js + 1) - F1(xlindx, js);
        const fint fsize = (lenvar * (lenvar + 1)) / 2;
        const fint updsze = ((lenvar - supvar) * (lenvar - supvar + 1)) / 2;
        const fint supsze = fsize - updsze;
        double *front = (double *)malloc((size_t)(fsize > 0 ? fsize : 1) * sizeof(double));
        /* copy the fully summed trapezoid, zero the update part (:518-523) */
        for (fint i = 1; i <= supsze; ++i) F1(front, i) = F1(lnz, iplnz + i - 1);
        for (fint i = 1; i <= updsze; ++i) F1(front, supsze + i) = 0.0;
        /* children, last first (LIFO stack order, :534-584) */
        const fint istart = F1(xsibl, js), istop = F1(xsibl, js + 1) - 1;
        const fint pstrt = F1(xlindx, js);
        for (fint ic = istop; ic >= istart; --ic) {
            const fint child = F1(sibl, ic);
            const fint cskip = F1(xsuper, child + 1) - F1(xsuper, child);
            const fint cstrt = F1(xlindx, child) + cskip;
            const fint chdlen = F1(xlindx, child + 1) - F1(xlindx, child) - cskip;
            if (chdlen > 0) {
                sprf12(lenvar, chdlen, &F1(lindx, pstrt), &F1(lindx, cstrt), lindex);
                sprf13(lenvar, chdlen, lindex, upd[child], front);
            }
            free(upd[child]);
            upd[child] = NULL;
        }
        if (F1(supsup, js) <= 0 || lpar57 <= 0) {
            fint err;
            if (panel) {
                const fint j = F1(xsplit, js);
                err = sprf15(lenvar, supvar, 1, &F1(split, j), front, &F1(diag, fstvar),
                             dcor ? &dcor[2 * (fstvar - 1)] : NULL, tol, numneg, &nrepl, first);
            } else {
                err = sprf15(lenvar, supvar, 0, NULL, front, &F1(diag, fstvar),
                             dcor ? &dcor[2 * (fstvar - 1)] : NULL, tol, numneg, &nrepl, first);
            }
            if (err < 0) {
                *supeqt = fstvar - err - 1;
                error = -30;
                free(front);
                goto done;
            }
        }
        if (supsze > 0) {
            for (fint i = 1; i <= supsze; ++i) F1(lnz, iplnz + i - 1) = F1(front, i);
            iplnz += supsze;
        }
        if (updsze > 0) {
            upd[js] = (double *)malloc((size_t)updsze * sizeof(double));
            memcpy(upd[js], &F1(front, supsze + 1), (size_t)updsze * sizeof(double));
            updlen[js] = updsze;
        }
        free(front);
    }
    {
        /* finalize TOL (:686-708) */
        double mdiag = 0.0;
        for (fint i = 1; i <= neq; ++i)
            if (F1(diag, i) > 0.0)
                if (mdiag == 0.0 || mdiag > F1(diag, i)) mdiag = F1(diag, i);
        tol[1] = (tol[1] != 0.0) ? 1.0 / tol[1] : 0.0;
        tol[2] = (mdiag != 0.0) ? tol[2] / mdiag : 0.0;
    }
    if (dcor && nrepl > 0) {
        *supeqt = nrepl;
        error = -31;
    }
done:
    for (fint js = 0; js <= nsuper; ++js) free(upd[js]);
    free(upd);
    free(updlen);
    free(first);
    free(lindex);
    return error;
}

/* MSPAR pointers into MTREES/MSIFA, as SPRFC1 carves them (sprfc1.f:161-203) */
typedef struct {
    fint neq, nsuper, nb, maxsup, mfront, nofsub, nzerol, wstack, nelact, nblock;
    const fint *xsuper, *xsibl, *sibl, *xlindx, *supsup, *lindx, *xlnz, *xsplit, *split;
} spr_view;

static void make_view(const fint *mspar, const fint *mtrees, const fint *msifa, spr_view *v)
{
    v->neq = F1(mspar, 8);
    v->nsuper = F1(mspar, 11);
    v->nb = F1(mspar, 12);
    v->maxsup = F1(mspar, 13);
    v->mfront = F1(mspar, 14);
    v->nofsub = F1(mspar, 15);
    v->nzerol = F1(mspar, 16);
    v->wstack = F1(mspar, 17);
    v->nelact = F1(mspar, 19);
    v->nblock = F1(mspar, 22);
    const fint xsuper = F1(mspar, 45);
    v->xsuper = &F1(mtrees, xsuper);
    v->xsibl = &F1(mtrees, xsuper + v->nsuper + 1);
    v->sibl = &F1(mtrees, xsuper + 2 * (v->nsuper + 1));
    const fint xlindx = F1(mspar, 46);
    v->xlindx = &F1(mtrees, xlindx);
    v->supsup = &F1(mtrees, xlindx + v->nsuper + 1);
    v->lindx = &F1(msifa, F1(mspar, 47));
    const fint xlnz = F1(mspar, 48);
    v->xlnz = &F1(msifa, xlnz);
    v->xsplit = &F1(msifa, xlnz + v->nsuper + 1);
    v->split = &F1(msifa, xlnz + 2 * v->nsuper + 1);
}

static void factor_common(fint *mspar, const fint *mtrees, const fint *msifa, double *sm,
                          double *tol, double *dcor, fint *ierr)
{
    if (F1(mspar, 1) < 0) {
        F1(mspar, 1) = -5;
        *ierr = -1;
        return;
    }
    F1(mspar, 1) = 5;
    *ierr = 0;
    spr_view v;
    make_view(mspar, mtrees, msifa, &v);
    fint numneg = 0, supeqt = 0;
    const fint lierr = sprf11(v.nsuper, v.neq, v.nblock, F1(mspar, 57), &numneg, v.xsuper,
                              v.xsibl, v.sibl, v.xlindx, v.supsup, v.lindx, v.xsplit, v.split,
                              v.xlnz, &F1(sm, 1), &F1(sm, v.neq + 1), dcor, tol, v.mfront, &supeqt);
    if (lierr != 0) {
        *ierr = (lierr / 10 == -3) ? -3 : -4;
        F1(mspar, 27) = (dcor && lierr == -31) ? -supeqt : supeqt;
    } else if (numneg > 0) {
        *ierr = 3;
    }
    F1(mspar, 26) = numneg;
}

void oracle_sprfc1(fint *mspar, const fint *mtrees, const fint *msifa, double *sm, double *tol,
                   fint *iwork, double *rwork, const fint *lpu, fint *ierr)
{
    (void)iwork;
    (void)rwork;
    (void)lpu;
    factor_common(mspar, mtrees, msifa, sm, tol, NULL, ierr);
}

void oracle_sprfc3(fint *mspar, const fint *mtrees, const fint *msifa, double *sm, double *tol,
                   fint *iwork, double *rwork, double *dcor, const fint *lpu, fint *ierr)
{
    (void)iwork;
    (void)rwork;
    (void)lpu;
    factor_common(mspar, mtrees, msifa, sm, tol, dcor, ierr);
}

/* ------------------------------------------------------------------------ */
/* Forward and diagonal solve (sprfs1.f:266-393, :531-633, :636-926, :929)   */
/* ------------------------------------------------------------------------ */

/* SPRFS6: y <- y - A x over the rectangle, remainder first then groups of 8
 * taken in descending column order (sprfs1.f:929-1141). */
static void sprfs6(fint rowsze, fint colsze, const double *lnz, const fint *first,
                   const double *x, double *y)
{
    if (rowsze <= 0 || colsze <= 0) return;
    const fint lftovr = colsze % 8;
    if (lftovr > 0) {
        for (fint i = 1; i <= rowsze; ++i) {
            double t = F1(y, i);
            for (fint j = 1; j <= lftovr; ++j) t = t - F1(x, j) * F1(lnz, F1(first, j) + i - 1);
            F1(y, i) = t;
        }
    }
    for (fint j = lftovr + 8; j <= colsze; j += 8) {
        for (fint i = 1; i <= rowsze; ++i) {
            double t = F1(y, i);
            for (fint k = 0; k < 8; ++k)
                t = t - F1(x, j - k) * F1(lnz, F1(first, j - k) + i - 1);
            F1(y, i) = t;
        }
    }
}

/* SPRFS5: local forward step on one supernode (sprfs1.f:531-633) */
static void sprfs5(fint supvar, fint lenvar, const double *lnz, double *locrhs, fint *first)
{
#define OFS(i, j) ((i) * ((j) - 1) - (((j) - 1) * ((j) - 2)) / 2)
    for (fint j = 1; j <= supvar; ++j) {
        fint iplnz = 1 + OFS(lenvar, j);
        const double rhsj = F1(locrhs, j);
        for (fint i = j + 1; i <= supvar; ++i) {
            iplnz += 1;
            F1(locrhs, i) = F1(locrhs, i) - rhsj * F1(lnz, iplnz);
        }
        F1(first, j) = iplnz + 1;
    }
    const fint nrows = lenvar - supvar;
    if (nrows > 0) sprfs6(nrows, supvar, lnz, first, locrhs, &F1(locrhs, supvar + 1));
    for (fint j = 1; j <= supvar; ++j) {
        const fint iplnz = 1 + OFS(lenvar, j);
        F1(locrhs, j) = F1(locrhs, j) / F1(lnz, iplnz);
    }
}

/* SPRBS6: local backward step, dot products per column (sprbs1.f:498-574) */
static void sprbs6(fint supvar, fint lenvar, const double *lnz, double *locrhs)
{
    for (fint j = supvar; j >= 1; --j) {
        fint iplnz2 = 1 + OFS(lenvar, j);
        double result = 0.0;
        for (fint i = j + 1; i <= lenvar; ++i) {
            iplnz2 += 1;
            result = result + F1(locrhs, i) * F1(lnz, iplnz2);
        }
        F1(locrhs, j) = F1(locrhs, j) - result;
    }
#undef OFS
}

/* SPRFB1 gather (zero beyond NEQ) and SPRFB2 scatter by assignment
 * (sprfs1.f:636-774, :777-926) */
/* SPRFB1 / SPRFB3 (sprfs1.f:777-926): gather, through PERI2E when given */
static void sprfb1(fint neq, fint jslen, const fint *lindx, const fint *peri2e,
                   const double *glb, double *loc)
{
    for (fint k = 1; k <= jslen; ++k) {
        const fint r = peri2e ? F1(peri2e, F1(lindx, k)) : F1(lindx, k);
        F1(loc, k) = (r <= neq) ? F1(glb, r) : 0.0;
    }
}

/* SPRFB2 / SPRFB4 (sprfs1.f:636-774): scatter by assignment */
static void sprfb2(fint neq, fint jslen, const fint *lindx, const fint *peri2e, double *glb,
                   const double *loc)
{
    for (fint k = 1; k <= jslen; ++k) {
        const fint r = peri2e ? F1(peri2e, F1(lindx, k)) : F1(lindx, k);
        if (r <= neq) F1(glb, r) = F1(loc, k);
    }
}

static int solve_checks(fint *mspar, const fint *ldb, fint *neq_out, fint *ierr)
{
    if (F1(mspar, 1) < 0) {
        F1(mspar, 1) = -7;
        *ierr = -1;
        return 0;
    }
    F1(mspar, 1) = 7;
    *ierr = 0;
    fint neq = F1(mspar, 8);
    if (F1(mspar, 57) == 1 && *ldb < neq) neq = *ldb;
    if (*ldb < neq) {
        F1(mspar, 1) = -2;
        *ierr = -2;
        return 0;
    }
    *neq_out = neq;
    return 1;
}

void oracle_sprfs1(fint *mspar, const fint *mtrees, const fint *msifa, const double *sm,
                   double *b, const fint *ldb, const fint *nrhs, fint *iwork, double *rwork,
                   const fint *lpu, fint *ierr)
{
    (void)iwork;
    (void)rwork;
    (void)lpu;
    fint neq;
    if (!solve_checks(mspar, ldb, &neq, ierr)) return;
    spr_view v;
    make_view(mspar, mtrees, msifa, &v);
    const double *lnz = &F1(sm, v.neq + 1);
    double *loc = (double *)malloc((size_t)(v.mfront + 1) * sizeof(double));
    fint *first = (fint *)malloc((size_t)(v.maxsup + 1) * sizeof(fint));
    const fint lpar57 = F1(mspar, 57);
    /* SPRFS4 (sprfs1.f:396-528) when MSPAR(58) >= 1 */
    const fint *peri2e = (F1(mspar, 58) >= 1) ? msifa : NULL;
    for (fint irhs = 0; irhs < *nrhs; ++irhs) {
        double *glb = b + irhs * (*ldb);
        fint iplnz = 1;
        for (fint js = 1; js <= v.nsuper; ++js) {
            const fint supvar = F1(v.xsuper, js + 1) - F1(v.xsuper, js);
            const fint jsstrt = F1(v.xlindx, js);
            const fint jslen = F1(v.xlindx, js + 1) - jsstrt;
            const fint supsze =
                (jslen * (jslen + 1) - (jslen - supvar) * (jslen - supvar + 1)) / 2;
            if (F1(v.supsup, js) <= 0 || lpar57 <= 0) {
                sprfb1(neq, jslen, &F1(v.lindx, jsstrt), peri2e, glb, loc);
                sprfs5(supvar, jslen, &F1(lnz, iplnz), loc, first);
                sprfb2(neq, jslen, &F1(v.lindx, jsstrt), peri2e, glb, loc);
            }
            iplnz += supsze;
        }
    }
    free(loc);
    free(first);
}

void oracle_sprbs1(fint *mspar, const fint *mtrees, const fint *msifa, const double *sm,
                   double *b, const fint *ldb, const fint *nrhs, double *rwork,
                   const fint *lpu, fint *ierr)
{
    (void)rwork;
    (void)lpu;
    fint neq;
    if (!solve_checks(mspar, ldb, &neq, ierr)) return;
    spr_view v;
    make_view(mspar, mtrees, msifa, &v);
    const double *lnz = &F1(sm, v.neq + 1);
    double *loc = (double *)malloc((size_t)(v.mfront + 1) * sizeof(double));
    const fint lpar57 = F1(mspar, 57);
    /* SPRBS5 (sprbs1.f:371-495) when MSPAR(58) >= 1 */
    const fint *peri2e = (F1(mspar, 58) >= 1) ? msifa : NULL;
    for (fint irhs = 0; irhs < *nrhs; ++irhs) {
        double *glb = b + irhs * (*ldb);
        fint iplnz = v.nzerol + 1;
        for (fint js = v.nsuper; js >= 1; --js) {
            const fint supvar = F1(v.xsuper, js + 1) - F1(v.xsuper, js);
            const fint jsstrt = F1(v.xlindx, js);
            const fint jslen = F1(v.xlindx, js + 1) - jsstrt;
            const fint supsze =
                (jslen * (jslen + 1) - (jslen - supvar) * (jslen - supvar + 1)) / 2;
            iplnz -= supsze;
            if (F1(v.supsup, js) <= 0 || lpar57 <= 0) {
                sprfb1(neq, jslen, &F1(v.lindx, jsstrt), peri2e, glb, loc);
                sprbs6(supvar, jslen, &F1(lnz, iplnz), loc);
                sprfb2(neq, jslen, &F1(v.lindx, jsstrt), peri2e, glb, loc);
            }
        }
    }
    free(loc);
}

void oracle_sprsol(const fint *iop, fint *mspar, const fint *mtrees, const fint *msifa,
                   double *sm, double *b, const fint *ldb, const fint *nrhs, double *tol,
                   fint *iwork, double *rwork, const fint *lpu, fint *ierr)
{
    /* sprsol.f:29-50 */
    *ierr = 0;
    if (*iop <= 3) oracle_sprfc1(mspar, mtrees, msifa, sm, tol, iwork, rwork, lpu, ierr);
    if (*iop >= 2 && *iop <= 5 && *ierr >= 0)
        oracle_sprfs1(mspar, mtrees, msifa, sm, b, ldb, nrhs, iwork, rwork, lpu, ierr);
    if (*iop >= 3 && *iop != 5 && *ierr >= 0)
        oracle_sprbs1(mspar, mtrees, msifa, sm, b, ldb, nrhs, rwork, lpu, ierr);
}

/* ------------------------------------------------------------------------ */
/* SPRPRM / SPRPRQ: C = +-A*B on the compressed layout (sprprm.f:5-324)      */
/* ------------------------------------------------------------------------ */
void oracle_sprprm(const double *a, const double *b, double *c, double *wa, fint *mspar,
                   const fint *mtrees, const fint *msifa, const fint *m, const fint *n,
                   const fint *ksa, const fint *iflag, const fint *lpu, fint *ierr)
{
    (void)lpu;
    if (F1(mspar, 1) < 0) {
        *ierr = -1;
        return;
    }
    *ierr = 0;
    const int avoidc = (*iflag == 88 && *n == 1);
    if (!avoidc && labs(*iflag) != 1 && labs(*iflag) != 2) {
        F1(mspar, 1) = -9;
        *ierr = -2;
        return;
    }
    /* SPRPRR (sprprm.f:327-397) addresses B and C through PERI2E */
    const fint *peri2e = (F1(mspar, 58) >= 1) ? msifa : NULL;
    if (*ksa <= 0 && peri2e) {
        F1(mspar, 1) = -9;
        *ierr = -3;
        return;
    }
    spr_view v;
    make_view(mspar, mtrees, msifa, &v);
    fint neq = v.neq;
    if (*m < neq) {
        if (F1(mspar, 57) == 1) {
            neq = *m;
        } else {
            F1(mspar, 1) = -9;
            *ierr = -4;
            return;
        }
    }
    const double alfa = (*iflag < 0) ? -1.0 : 1.0;
    for (fint in = 0; in < *n; ++in) {
        const double *bb = b + in * (*m);
        if (*ksa > 0) {
            for (fint i = 1; i <= neq; ++i) F1(wa, i) = 0.0;
            for (fint js = 1; js <= v.nsuper; ++js) {
                const fint fstvar = F1(v.xsuper, js), lstvar = F1(v.xsuper, js + 1) - 1;
                fint istrt = F1(v.xlindx, js);
                const fint istop = F1(v.xlindx, js + 1) - 1;
                fint iplnz = F1(v.xlnz, js) - 1;
                for (fint j = fstvar; j <= lstvar; ++j) {
                    iplnz += 1;
                    const fint jj = peri2e ? F1(peri2e, j) : j;
                    if (jj <= neq) {
                        F1(wa, jj) = F1(wa, jj) + alfa * (F1(a, iplnz) * F1(bb, jj));
                        for (fint i = istrt + 1; i <= istop; ++i) {
                            iplnz += 1;
                            const fint ii = peri2e ? F1(peri2e, F1(v.lindx, i)) : F1(v.lindx, i);
                            if (ii <= neq) {
                                F1(wa, jj) = F1(wa, jj) + alfa * (F1(a, iplnz) * F1(bb, ii));
                                F1(wa, ii) = F1(wa, ii) + alfa * (F1(a, iplnz) * F1(bb, jj));
                            }
                        }
                    } else if (peri2e && istop > istrt) {
                        iplnz += istop - istrt;
                    }
                    istrt += 1;
                }
            }
        } else {
            for (fint j = 1; j <= neq; ++j) F1(wa, j) = alfa * (F1(a, j) * F1(bb, j));
        }
        if (avoidc) continue;
        double *cc = c + in * (*m);
        if (labs(*iflag) == 2) {
            for (fint j = 1; j <= *m; ++j) F1(cc, j) = F1(cc, j) + F1(wa, j);
        } else {
            for (fint j = 1; j <= *m; ++j) F1(cc, j) = F1(wa, j);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Element assembly: SPRAD2's free-dof loop (spradm.f:818-859), element by    */
/* element as a host loop over SPRADM (spradm.f:5-166) leaves it.             */
/* ------------------------------------------------------------------------- */
/* 1-based LNZ position of entry (row r, column c), r >= c, 1-based SPR
 * equations: the slot that SPRAD1's PSM table assigns (spradm.f:514-740).
 * Column c of supernode js is column k = c - XSUPER(js) of its packed
 * trapezoid, which starts at XLNZ(js) + OFFSET(F,k) and holds the rows
 * LINDX(XLINDX(js)+k ..) (sprsmb.f:512-575).  Returns 0 if r is not in the
 * column's row list. */
static fint lnz_position(const spr_view *v, const fint *col_sup, fint r, fint c)
{
    const fint js = col_sup[c - 1];
    const fint k = c - F1(v->xsuper, js);
    const fint f = F1(v->xlindx, js + 1) - F1(v->xlindx, js);
    fint lo = F1(v->xlindx, js) + k, hi = F1(v->xlindx, js + 1) - 1;
    while (lo <= hi) {
        const fint mid = lo + (hi - lo) / 2;
        const fint e = F1(v->lindx, mid);
        if (e == r) return F1(v->xlnz, js) + OFFSET(f, k) + (mid - F1(v->xlindx, js) - k);
        if (e < r)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return 0;
}

/* SPRAD2's free loop for element e (spradm.f:818-859): 0 or -(e+1). */
static fint free_element(const spr_view *v, const fint *col_sup, const fint *meen,
                         const fint *meen_ptr, const double *em, const fint *em_off,
                         const double *escale, fint e, double *lnz)
{
    const fint *mn = meen + meen_ptr[e];
    const fint nd = meen_ptr[e + 1] - meen_ptr[e];
    const double *ek = em + em_off[e];
    for (fint j = 1; j <= nd; ++j) {
        const fint jeq = F1(mn, j);
        if (jeq < 1) continue;
        for (fint i = 1; i <= j; ++i) {
            const fint ieq = F1(mn, i);
            if (ieq < 1) continue;
            const fint r = ieq > jeq ? ieq : jeq, c = ieq > jeq ? jeq : ieq;
            const fint k = lnz_position(v, col_sup, r, c);
            if (k == 0) return -(e + 1);
            /* the host loop scales the base matrix before the call */
            const double x = ek[(i - 1) + (j - 1) * nd];
            const double emij = escale ? escale[e] * x : x;
            F1(lnz, k) = F1(lnz, k) + emij;
        }
    }
    return 0;
}

fint oracle_assemble_free(const fint *mspar, const fint *mtrees, const fint *msifa, fint nel,
                          const fint *meen_ptr, const fint *meen, const double *em,
                          const fint *em_off, const double *escale, double *sm)
{
    spr_view v;
    make_view(mspar, mtrees, msifa, &v);
    fint *col_sup = (fint *)malloc(sizeof(fint) * (size_t)(v.neq > 0 ? v.neq : 1));
    if (!col_sup) return -1;
    for (fint js = 1; js <= v.nsuper; ++js)
        for (fint c = F1(v.xsuper, js); c < F1(v.xsuper, js + 1); ++c) col_sup[c - 1] = js;
    double *lnz = sm + v.neq;  /* SPRADM passes SM(NEQ+1) (spradm.f:161) */
    fint rc = 0;
    for (fint e = 0; e < nel && rc == 0; ++e)
        rc = free_element(&v, col_sup, meen, meen_ptr, em, em_off, escale, e, lnz);
    free(col_sup);
    return rc;
}

/* SPRAD2 complete (spradm.f:743-1002), element by element as a host loop
 * over SPRADM leaves it: the free part (above) when nsv >= 0, then the
 * constrained dofs (MEEN(J) = -JCEQ < 0).  SV(NEQ*|nsv|) receives
 * SV(KEQ) - C0*EM(I,J) for each constrained J with C0 = TTCC(MPMCEQ(JCEQ))
 * != 0 (spradm.f:866-908; for |nsv| > 1 only prescribed J, into vector K,
 * the rank of JCEQ among the equations without masters), and, when nsv >= 0,
 * SM receives TTCC(JP)*EM(I,J) per master JP of J (doubled on the master's
 * own diagonal) and TTCC(IP)*TTCC(JP)*EM(I,J) per master pair with
 * MEQN(master I) <= MEQN(master J) (spradm.f:910-999).  Masters with
 * MMCEQ <= 0 or TTCC = 0 are skipped.  Returns 0, -(e+1) for an element
 * entry outside the pattern, or -(nel+1) for a vector index beyond |nsv|. */
fint oracle_assemble_ceq(const fint *mspar, const fint *mtrees, const fint *msifa, fint nel,
                         const fint *meen_ptr, const fint *meen, const double *em,
                         const fint *em_off, const double *escale, const fint *mpmceq,
                         const fint *mmceq, const double *ttcc, const fint *meqn, fint nsv,
                         double *sm, double *sv)
{
    spr_view v;
    make_view(mspar, mtrees, msifa, &v);
    const fint neq = v.neq;
    fint *col_sup = (fint *)malloc(sizeof(fint) * (size_t)(neq > 0 ? neq : 1));
    if (!col_sup) return -1;
    for (fint js = 1; js <= v.nsuper; ++js)
        for (fint c = F1(v.xsuper, js); c < F1(v.xsuper, js + 1); ++c) col_sup[c - 1] = js;
    double *lnz = sm + neq;
    const fint nvec = nsv < 0 ? -nsv : nsv;
    fint rc = 0;
#define SM_ADD(r_, c_, val_)                                                  \
    do {                                                                      \
        const fint k_ = lnz_position(&v, col_sup, (r_), (c_));                \
        if (k_ == 0) {                                                        \
            rc = -(e + 1);                                                    \
            goto done;                                                        \
        }                                                                     \
        F1(lnz, k_) = F1(lnz, k_) + (val_);                                   \
    } while (0)
    for (fint e = 0; e < nel; ++e) {
        const fint *mn = meen + meen_ptr[e];
        const fint nd = meen_ptr[e + 1] - meen_ptr[e];
        const double *ek = em + em_off[e];
        if (nsv >= 0 && (rc = free_element(&v, col_sup, meen, meen_ptr, em, em_off, escale, e,
                                            lnz)) != 0)
            goto done;
        for (fint j = 1; j <= nd; ++j) {
            if (F1(mn, j) >= 0) continue;
            const fint jceq = -F1(mn, j);
            fint jp = F1(mpmceq, jceq);
            if (nsv != 0) {
                const double c0 = F1(ttcc, jp);
                fint kp = -1;
                if (c0 != 0.0) {
                    if (nvec == 1) {
                        kp = 0;
                    } else if (F1(mpmceq, jceq + 1) - F1(mpmceq, jceq) == 1) {
                        fint k = 0;
                        for (fint i = 1; i <= jceq; ++i)
                            if (F1(mpmceq, i + 1) - F1(mpmceq, i) == 1) ++k;
                        if (k > nvec) {
                            rc = -(nel + 1);
                            goto done;
                        }
                        kp = (k - 1) * neq;
                    }
                }
                for (fint i = 1; kp >= 0 && i <= nd; ++i) {
                    const fint ieq = F1(mn, i);
                    if (ieq == 0) continue;
                    const double x = ek[(i - 1) + (j - 1) * nd];
                    const double emij = escale ? escale[e] * x : x;
                    if (ieq > 0) {
                        F1(sv, kp + ieq) = F1(sv, kp + ieq) - c0 * emij;
                        continue;
                    }
                    const fint iceq = -ieq;
                    for (fint ip = F1(mpmceq, iceq) + 1; ip < F1(mpmceq, iceq + 1); ++ip) {
                        if (F1(mmceq, ip) <= 0 || F1(ttcc, ip) == 0.0) continue;
                        const fint keq = kp + F1(meqn, F1(mmceq, ip));
                        F1(sv, keq) = F1(sv, keq) - c0 * F1(ttcc, ip) * emij;
                    }
                }
                if (nsv < 0) continue;
            }
            const fint nmstj = F1(mpmceq, jceq + 1) - F1(mpmceq, jceq) - 1;
            for (fint jj = 1; jj <= nmstj; ++jj) {
                ++jp;
                if (F1(mmceq, jp) <= 0 || F1(ttcc, jp) == 0.0) continue;
                const fint jeq = F1(meqn, F1(mmceq, jp));
                for (fint i = 1; i <= nd; ++i) {
                    const fint ieq = F1(mn, i);
                    if (ieq == 0) continue;
                    const double x = ek[(i - 1) + (j - 1) * nd];
                    const double emij = escale ? escale[e] * x : x;
                    if (ieq > 0) {
                        if (ieq != jeq)
                            SM_ADD(ieq > jeq ? ieq : jeq, ieq > jeq ? jeq : ieq,
                                   F1(ttcc, jp) * emij);
                        else
                            SM_ADD(ieq, ieq, (F1(ttcc, jp) + F1(ttcc, jp)) * emij);
                        continue;
                    }
                    const fint iceq = -ieq;
                    for (fint ip = F1(mpmceq, iceq) + 1; ip < F1(mpmceq, iceq + 1); ++ip) {
                        if (F1(mmceq, ip) <= 0 || F1(ttcc, ip) == 0.0) continue;
                        const fint meq = F1(meqn, F1(mmceq, ip));
                        if (meq <= jeq) SM_ADD(jeq, meq, F1(ttcc, ip) * F1(ttcc, jp) * emij);
                    }
                }
            }
        }
    }
#undef SM_ADD
done:
    free(col_sup);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* SPRTMM (sprtmm.f:5-241) with its product loop SPRTMQ (sprtmm.f:242-344).   */
/* ------------------------------------------------------------------------- */
void oracle_sprtmm(const double *u, const double *b, double *c, double *wa, fint *mspar,
                   const fint *mtrees, const fint *msifa, const fint *m, const fint *n,
                   const fint *ksu, const fint *iflag, const fint *lpu, fint *ierr)
{
    (void)lpu;
    if (F1(mspar, 1) < 0) {
        F1(mspar, 1) = -10;
        *ierr = -1;
        return;
    }
    *ierr = 0;
    const int avoidc = (*iflag == 11 || *iflag == 22) && *n == 1;
    if (!avoidc && *iflag != 1 && *iflag != 2) {
        F1(mspar, 1) = -10;
        *ierr = -2;
        return;
    }
    spr_view v;
    make_view(mspar, mtrees, msifa, &v);
    fint neq = v.neq;
    if (*m < neq) {
        if (F1(mspar, 57) == 1) {
            neq = *m;
        } else {
            F1(mspar, 1) = -10;
            *ierr = -3;
            return;
        }
    }
    const int upper = (*iflag == 1 || *iflag == 11);
    /* SPRTMR (sprtmm.f:345-480): through PERI2E, U read as LNZ whatever KSU */
    const fint *peri2e = (F1(mspar, 58) >= 1) ? msifa : NULL;
    for (fint in = 0; in < *n; ++in) {
        const double *bb = b + in * (*m);
        if (*ksu > 0 || peri2e) {
            for (fint i = 1; i <= neq; ++i) F1(wa, i) = 0.0;
            for (fint js = 1; js <= v.nsuper; ++js) {
                const fint fstvar = F1(v.xsuper, js), lstvar = F1(v.xsuper, js + 1) - 1;
                fint istrt = F1(v.xlindx, js);
                const fint istop = F1(v.xlindx, js + 1) - 1;
                fint iplnz = F1(v.xlnz, js) - 1;
                for (fint j = fstvar; j <= lstvar; ++j) {
                    iplnz += 1;
                    const fint jj = peri2e ? F1(peri2e, j) : j;
                    if (jj <= neq) {
                        F1(wa, jj) = F1(wa, jj) + F1(u, iplnz) * F1(bb, jj);
                        for (fint i = istrt + 1; i <= istop; ++i) {
                            iplnz += 1;
                            const fint ii = peri2e ? F1(peri2e, F1(v.lindx, i)) : F1(v.lindx, i);
                            if (ii <= neq) {
                                if (upper)
                                    F1(wa, jj) = F1(wa, jj) + F1(u, iplnz) * F1(bb, ii);
                                else
                                    F1(wa, ii) = F1(wa, ii) + F1(u, iplnz) * F1(bb, jj);
                            }
                        }
                    } else if (peri2e && istop > istrt) {
                        iplnz += istop - istrt;
                    }
                    istrt += 1;
                }
            }
        } else {
            for (fint j = 1; j <= neq; ++j) F1(wa, j) = F1(u, j) * F1(bb, j);
        }
        if (!avoidc)
            for (fint jn = 1; jn <= *m; ++jn) c[in * (*m) + jn - 1] = F1(wa, jn);
    }
}
