// Adjoint kernels: gradients of a scalar objective through K u = f.
//
// The reference has no adjoint.  With lambda = K^-T dL/du (for the
// symmetric K of SPR, K^-T = K^-1, so lambda comes from the same factor and
// solves), the gradient with respect to the matrix entries is
// dL/dK_ij = -lambda_i u_j; on SPR's symmetric lower storage one stored
// off-diagonal entry stands for K_ij and K_ji, so its gradient is
// -(lambda_i u_j + lambda_j u_i), and a diagonal entry's is -lambda_i u_i.
// For element matrices linear in one parameter each (K = sum_e p_e K0_e,
// e.g. Young's moduli), dL/dp_e = -lambda_e^T K0_e u_e.
#include "lanes.h"
#include "sprgpu_internal.h"

namespace sprgpu {

// One wavefront per element (four per workgroup); lane i < nedof holds
// element dof i.  eq < 0 marks a dof without an equation (fixed).
__global__ __launch_bounds__(256) void k_element_grad(i64 nel, int nedof,
                                                      const i64 *__restrict__ elem_eq,
                                                      const double *__restrict__ k0,
                                                      const i64 *__restrict__ k0_index,
                                                      const double *__restrict__ lam,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ grad)
{
    const i64 e = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (e >= nel) return;
    const i64 *eq = elem_eq + e * nedof;
    const double *K = k0 + (k0_index ? k0_index[e] : 0) * i64(nedof) * nedof;
    const i64 q = (lane < nedof) ? eq[lane] : -1;
    const double xi = (q >= 0) ? x[q] : 0.0;
    const double li = (q >= 0) ? lam[q] : 0.0;
    // y_i = sum_j K0(i,j) x_j with x_j broadcast from lane j
    double y = 0.0;
    for (int j = 0; j < nedof; ++j) {
        const double xj = lane_bcast(xi, j);
        if (lane < nedof) y += K[lane + i64(j) * nedof] * xj;
    }
    double s = li * y;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) grad[e] = -s;
}

// Gradient on the packed LNZ layout: one workgroup per supernode column.
__global__ __launch_bounds__(256) void k_pattern_grad(const i32 *__restrict__ col_node,
                                                      const NodeInfo *__restrict__ nodes,
                                                      const i32 *__restrict__ lindx,
                                                      const double *__restrict__ lam,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ glnz)
{
    const int c = blockIdx.x;  // equation = column
    const NodeInfo n = nodes[col_node[c]];
    const int k = c - n.first;  // local column
    const i32 *rows = lindx + n.lxoff;
    const i64 base = n.xlnz + i64(k) * n.F - i64(k) * (k - 1) / 2;  // OFFSET(F,k)
    const double lc = lam[c], xc = x[c];
    for (int p = k + threadIdx.x; p < n.F; p += blockDim.x) {
        const int r = rows[p];
        glnz[base + (p - k)] = (p == k) ? -lc * xc : -(lam[r] * xc + lc * x[r]);
    }
}

int run_element_grad(sprgpu_handle *h, i64 nel, int nedof, const i64 *elem_eq, const double *k0,
                     const i64 *k0_index, const double *lam, const double *x, double *grad)
{
    if (nel <= 0) return 0;
    if (nedof <= 0 || nedof > 64) {
        set_error("element_grad: 1 <= nedof <= 64 required");
        return -2;
    }
    const i64 nb = (nel + 3) / 4;
    if (nb >= (i64(1) << 31)) {
        set_error("element_grad: too many elements");
        return -2;
    }
    k_element_grad<<<unsigned(nb), 256, 0, h->stream>>>(nel, nedof, elem_eq, k0, k0_index, lam,
                                                         x, grad);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

int upload_col_node(sprgpu_handle *h)
{
    const Plan &p = h->plan;
    if (!h->d.col_node) {
        std::vector<i32> cn(std::max<i64>(p.neq, 1));
        for (i64 j = 0; j < p.nsuper; ++j)
            for (i32 k = 0; k < p.nodes[j].S; ++k) cn[p.nodes[j].first + k] = i32(j);
        SPRGPU_HIP(hipMalloc(&h->d.col_node, sizeof(i32) * cn.size()));
        SPRGPU_HIP(hipMemcpy(h->d.col_node, cn.data(), sizeof(i32) * cn.size(),
                             hipMemcpyHostToDevice));
        h->device_bytes += sizeof(i32) * p.neq;
    }
    return 0;
}

int run_pattern_grad(sprgpu_handle *h, const double *lam, const double *x, double *glnz)
{
    const Plan &p = h->plan;
    if (int rc = upload_col_node(h)) return rc;
    k_pattern_grad<<<unsigned(p.neq), 256, 0, h->stream>>>(h->d.col_node, h->d.nodes, h->d.lindx,
                                                            lam, x, glnz);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

}  // namespace sprgpu
