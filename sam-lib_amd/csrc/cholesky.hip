// The CC' siblings (SPRFC2, SPRFS2, SPRBS2) on the LDL' kernels.
//
// SPRFC2 (sprfc2.f:5-275) factors A = C C' with the pivot step of SPRF27
// (sprfc2.f:934-1037): the pivot d (the Schur complement's diagonal, as in
// LDL') fails when d <= TOL(1)*|a_jj|, signed, so a negative pivot fails
// too; otherwise c_jj = sqrt(d) and the column below is scaled by 1/c_jj.
// In exact arithmetic C = L sqrt(D) with L D L' the SPRFC1 factor, so:
// * the factor is the LDL' factorization with TOL(1) = 0 (every pivot is
//   computed), followed by SPRFC2's test on the pivots in equation order,
//   which is the reference's elimination order;
// * SPRFS2's C^-1 b = sqrt(D) (D^-1 L^-1 b), SPRBS2's C^-T y = L^-T
//   (D^-1/2 y), and both together are the LDL' solve;
// * the factor in SM layout has sqrt(d_j) on the diagonal and in SM(1:NEQ)
//   and l_ij sqrt(d_j) below it.
// Results agree with the reference's to rounding, not bit for bit.
#include <cmath>

#include "sprgpu_internal.h"

namespace sprgpu {

__global__ __launch_bounds__(256) void k_cc_scale(const double *__restrict__ d, double *b,
                                                  i64 ldb, i64 n, int inverse)
{
    const i64 j = i64(blockIdx.x) * 256 + threadIdx.x;
    if (j >= n) return;
    const double c = sqrt(d[j]);
    double *x = b + i64(blockIdx.y) * ldb + j;
    *x = inverse ? *x / c : *x * c;
}

// LDL' layout (d on the diagonal and in SM(1:NEQ), l below) to C = L sqrt(D)
__global__ __launch_bounds__(256) void k_cc_factor_layout(const i32 *__restrict__ col_node,
                                                          const NodeInfo *__restrict__ nodes,
                                                          double *sm, i64 n)
{
    const i64 j = i64(blockIdx.x) * 256 + threadIdx.x;
    if (j >= n) return;
    const NodeInfo nd = nodes[col_node[j]];
    const i64 k = j - nd.first;
    double *col = sm + n + nd.xlnz + k * nd.F - k * (k - 1) / 2;
    const double c = sqrt(sm[j]);
    sm[j] = c;
    col[0] = c;
    for (i64 t = 1; t < nd.F - k; ++t) col[t] *= c;
}

// C in SM layout (a CC' factor computed elsewhere) to the LDL' layout:
// d_j = c_jj^2, l_ij = c_ij / c_jj
__global__ __launch_bounds__(256) void k_cc_to_ldl_layout(const i32 *__restrict__ col_node,
                                                          const NodeInfo *__restrict__ nodes,
                                                          double *sm, i64 n)
{
    const i64 j = i64(blockIdx.x) * 256 + threadIdx.x;
    if (j >= n) return;
    const NodeInfo nd = nodes[col_node[j]];
    const i64 k = j - nd.first;
    double *col = sm + n + nd.xlnz + k * nd.F - k * (k - 1) / 2;
    const double c = col[0];
    sm[j] = c * c;
    col[0] = c * c;
    for (i64 t = 1; t < nd.F - k; ++t) col[t] /= c;
}

int cc_to_ldl_layout(sprgpu_handle *h, double *sm_dev)
{
    if (int rc = upload_col_node(h)) return rc;
    const i64 n = h->plan.neq;
    k_cc_to_ldl_layout<<<unsigned((n + 255) / 256), 256, 0, h->stream>>>(h->d.col_node,
                                                                          h->d.nodes, sm_dev, n);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int cc_factor_info(sprgpu_handle *h, sprgpu_factor_info *info)
{
    // the LDL' statistics of the TOL(1) = 0 factorization
    h->mode = 1;
    const int rc = read_factor_info(h, info);
    h->mode = 2;
    if (rc) return rc;
    if (info->ierr == -3 && info->nzero_diag > 0) return 0;  // zero diagonals: as SPRFC1
    // SPRF27's test in equation order, up to an exactly zero pivot if one
    // stopped the LDL' factorization
    const i64 neq = h->plan.neq;
    const i64 last = (info->ierr == -3) ? info->supeqt - 1 : neq - 1;
    std::vector<double> d(neq), a(neq);
    SPRGPU_HIP(hipMemcpy(d.data(), h->d.dvec, sizeof(double) * neq, hipMemcpyDeviceToHost));
    SPRGPU_HIP(hipMemcpy(a.data(), h->d.adiag, sizeof(double) * neq, hipMemcpyDeviceToHost));
    const double tol1 = h->cc_tol1;
    for (i64 j = 0; j <= last; ++j)
        if (d[j] <= tol1 * std::fabs(a[j])) {
            info->ierr = -3;
            info->supeqt = j + 1;
            info->tol2 = d[j];
            info->tol3 = a[j];
            info->numneg = 0;
            return 0;
        }
    info->ierr = 0;
    info->numneg = 0;  // MSPAR(26) counts zero diagonals in SPRFC2
    return 0;
}

int cc_scale(sprgpu_handle *h, double *b, i64 ldb, i64 nrhs, int inverse)
{
    const i64 n = h->plan.neq;
    for (i64 r0 = 0; r0 < nrhs; r0 += 65535) {
        const i64 nr = std::min<i64>(65535, nrhs - r0);
        k_cc_scale<<<dim3(unsigned((n + 255) / 256), unsigned(nr)), 256, 0, h->stream>>>(
            h->d.dvec, b + r0 * ldb, ldb, n, inverse);
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int cc_factor_layout(sprgpu_handle *h, double *sm_dev)
{
    if (int rc = upload_col_node(h)) return rc;
    const i64 n = h->plan.neq;
    k_cc_factor_layout<<<unsigned((n + 255) / 256), 256, 0, h->stream>>>(h->d.col_node,
                                                                          h->d.nodes, sm_dev, n);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

}  // namespace sprgpu
