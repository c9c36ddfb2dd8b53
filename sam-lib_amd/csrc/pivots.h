// Pivot rules shared by the factor kernels (included by .hip files only).
#pragma once

#include "sprgpu_internal.h"

namespace sprgpu {

// Value of the pivot after the SPRFC3 replacement rule (identity in mode 1).
__device__ __forceinline__ double pivot_value(double d, double a, const PivotArgs &pa)
{
    if (pa.mode == 3 && fabs(d) <= pa.tol1 * fabs(a)) {
        const double mag = pa.tol1 * fmax(1.0, fabs(pa.tol4 * a));
        return copysign(mag, d);
    }
    return d;
}

// Book-keeping for pivot eq (one thread, once per pivot).
__device__ __forceinline__ void pivot_record(double d_in, double d_out, double a, int eq,
                                             const PivotArgs &pa, double *dvec, double *dcor,
                                             FactorStatus *st)
{
    dvec[eq] = d_out;
    if (fabs(d_in) <= pa.tol1 * fabs(a)) {
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
