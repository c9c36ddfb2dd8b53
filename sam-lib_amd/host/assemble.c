/*
 * Host side of the drop-in: a loop over SPRADM for the element assembly.
 *
 * Linked into lib/libspr_host.so together with the reference's own symbolic
 * phase and assembly routines (compiled unchanged from src/SPR, see the
 * package Makefile).  The host application calls SPRADM once per element
 * (src/SPR/spradm.f:5-13); this helper runs that loop in C so that Python
 * does not cross the ctypes boundary once per element.
 */
#include <stdint.h>
#include <stdlib.h>

typedef int64_t fint;

/* src/SPR/spradm.f:5-13 */
extern void spradm_(const double *em, const double *ttcc, fint *mpar, fint *mspar,
                    fint *madof, fint *meqn, fint *mpmnpc, fint *mmnpc, fint *mpmceq,
                    fint *mmceq, fint *msica, fint *mtrees, fint *msifa, fint *mdofnc,
                    double *sm, double *sv, fint *iwork, const fint *iel,
                    const fint *nedof, const fint *lpu, const fint *nsv, fint *ierr);

/*
 * Assemble nel element matrices into SM (and optionally SV) by calling
 * SPRADM once per element, in element order.
 *   em      element matrices, column-major nedof[i] x nedof[i] each, element i
 *           starting at em + em_off[i] (offsets may repeat so that many
 *           elements share one base matrix)
 *   scale   optional per-element multiplier of the base matrix (NULL: none)
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
