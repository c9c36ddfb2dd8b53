// Pivot rules shared by the factor kernels (included by .hip files only).
#pragma once

#include "sprgpu_internal.h"

namespace sprgpu {

// Value of the pivot after the replacement rules: SPRFC3 (mode 3,
// sprfc3.f:1083-1094) replaces a pivot with |d| <= TOL(1)|a_jj| by
// sign(d) TOL(1) max(1, |TOL(4) a_jj|); SPRFC4 (mode 4, sprfc4.f:1016-1037)
// replaces one in [-TOL(1)|a_jj|, TOL(1)|a_jj|] by the positive
// TOL(1) max(1, |TOL(4) a_jj|) (a pivot below that range fails, and is
// kept).  The identity in modes 1 and 2.
__device__ __forceinline__ double pivot_value(double d, double a, const PivotArgs &pa)
{
    const double lim = pa.tol1 * fabs(a);
    if (pa.mode == 3 && fabs(d) <= lim) return copysign(pa.tol1 * fmax(1.0, fabs(pa.tol4 * a)), d);
    if (pa.mode == 4 && d >= -lim && d <= lim) return pa.tol1 * fmax(1.0, fabs(pa.tol4 * a));
    return d;
}

// Book-keeping for pivot eq (one thread, once per pivot): d, and by mode
// the failure (the smallest failing equation), the replacement count and
// DCOR(1:2, eq) = (d before, d after), or (d, d) for SPRFC4's failing
// negative pivot (sprfc4.f:1016-1025).
__device__ __forceinline__ void pivot_record(double d_in, double d_out, double a, int eq,
                                             const PivotArgs &pa, double *dvec, double *dcor,
                                             FactorStatus *st)
{
    dvec[eq] = d_out;
    const double lim = pa.tol1 * fabs(a);
    if (pa.mode == 4) {
        if (d_in < -lim) {
            atomicMin(&st->first_fail, (unsigned long long)eq);
            dcor[2 * i64(eq)] = d_in;
            dcor[2 * i64(eq) + 1] = d_in;
        } else if (d_in <= lim) {
            atomicAdd(&st->nrepl, 1ull);
            dcor[2 * i64(eq)] = d_in;
            dcor[2 * i64(eq) + 1] = d_out;
        }
        return;
    }
    if (fabs(d_in) <= lim) {
        if (pa.mode == 3) {
            atomicAdd(&st->nrepl, 1ull);
            dcor[2 * i64(eq)] = d_in;
            dcor[2 * i64(eq) + 1] = d_out;
        } else {
            atomicMin(&st->first_fail, (unsigned long long)eq);
        }
    }
    if (pa.mode == 3 && d_out < 0.0 && dcor[2 * i64(eq)] == 0.0) dcor[2 * i64(eq)] = d_out;
}

}  // namespace sprgpu
