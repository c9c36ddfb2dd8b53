// Element stiffness matrices on the GPU: trilinear hexahedra (8 nodes, 3 dofs
// per node), isotropic linear elasticity, 2 x 2 x 2 Gauss quadrature.
//
// The host application computes an element matrix before each SPRADM call
// (spradm.f:5-13 takes EM from its caller); with the matrices made on the
// device, element generation, assembly (assemble.hip), factorization and
// solves run without host copies of EM or SM.
//
// K_e(3i+p, 3j+q) = sum_g det J_g [lambda dNi/dx_p dNj/dx_q + mu dNi/dx_q dNj/dx_p
//                                  + mu delta_pq grad Ni . grad Nj]
// over the Gauss points g = (+-1/sqrt3)^3 (weights 1), with the shape function
// gradients dN/dx = J^-1 dN/dxi, J = dN/dxi X the Jacobian of the node
// coordinates X, lambda = E nu / ((1 + nu)(1 - 2 nu)) and mu = E / (2(1 + nu)):
// B' D B with engineering shear strains, written out per entry.  Local node
// order (-1,-1,-1) (1,-1,-1) (1,1,-1) (-1,1,-1) (-1,-1,1) (1,-1,1) (1,1,1)
// (-1,1,1); the element's dofs are its nodes in connectivity order, each
// (ux, uy, uz), the order SPRADM expects with three dofs per node.
//
// One wavefront per element: lane (g, a) forms node a's gradient at Gauss
// point g (the Jacobian summed over the point's eight lanes by shuffles),
// into LDS; then every lane accumulates nine of the 576 entries over the
// Gauss points and stores them column-major (entry lane + 64 t), so an
// element's stores are contiguous.  HBM-bound by the 4.6 KB written per
// element.
#include "sprgpu_internal.h"

namespace sprgpu {

__global__ __launch_bounds__(256) void k_hex8_stiffness(i64 nel, i64 nanod,
                                                        const double *__restrict__ xyz,
                                                        const i64 *__restrict__ mmnpc,
                                                        const double *__restrict__ emod,
                                                        double nu, double *__restrict__ em)
{
    // per wavefront: dN/dx at the eight Gauss points (g, component, node) and
    // det J (g)
    __shared__ double dxs[4][8][3][8];
    __shared__ double dets[4][8];
    const int v = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const i64 e = i64(blockIdx.x) * 4 + v;
    const bool live = e < nel;
    const i64 ec = live ? e : nel - 1;
    // lane (g, a) = (lane / 8, lane % 8): Gauss point g, node a
    const int g = lane >> 3, a = lane & 7;
    const double sx = (a == 1 || a == 2 || a == 5 || a == 6) ? 1.0 : -1.0;
    const double sy = (a == 2 || a == 3 || a == 6 || a == 7) ? 1.0 : -1.0;
    const double sz = a >= 4 ? 1.0 : -1.0;
    const double gp = 0.57735026918962576451;  // 1/sqrt(3)
    const double xi = (g & 1) ? gp : -gp, et = (g & 2) ? gp : -gp, ze = (g & 4) ? gp : -gp;
    const i64 node = mmnpc[8 * ec + a];
    const bool bad_node = node < 1 || node > nanod;
    const i64 nd = bad_node ? 0 : node - 1;
    const double X0 = xyz[3 * nd], X1 = xyz[3 * nd + 1], X2 = xyz[3 * nd + 2];
    const double dn0 = 0.125 * sx * (1.0 + et * sy) * (1.0 + ze * sz);
    const double dn1 = 0.125 * sy * (1.0 + xi * sx) * (1.0 + ze * sz);
    const double dn2 = 0.125 * sz * (1.0 + xi * sx) * (1.0 + et * sy);
    // J = dN/dxi X: sums over the eight lanes of the Gauss point, in node order
    // on every lane (a butterfly of xor 1, 2, 4 adds the same pairs in the
    // same order on each lane of a group)
    double J[3][3];
    const double dn[3] = {dn0, dn1, dn2}, Xa[3] = {X0, X1, X2};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double t = dn[r] * Xa[c];
#pragma unroll
            for (int m = 1; m < 8; m <<= 1) {
                const double o = __shfl_xor(t, m, 64);
                t = (a & m) ? o + t : t + o;
            }
            J[r][c] = t;
        }
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double id = 1.0 / det;
    const double Ji00 = c00 * id, Ji10 = c01 * id, Ji20 = c02 * id;
    const double Ji01 = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    const double Ji11 = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    const double Ji21 = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    const double Ji02 = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    const double Ji12 = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    const double Ji22 = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    dxs[v][g][0][a] = Ji00 * dn0 + Ji01 * dn1 + Ji02 * dn2;
    dxs[v][g][1][a] = Ji10 * dn0 + Ji11 * dn1 + Ji12 * dn2;
    dxs[v][g][2][a] = Ji20 * dn0 + Ji21 * dn1 + Ji22 * dn2;
    if (a == 0) dets[v][g] = det;
    const bool bad = __builtin_amdgcn_ballot_w64(bad_node) != 0ull;
    __syncthreads();
    const double E = emod ? emod[ec] : 1.0;
    const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = E / (2.0 * (1.0 + nu));
    // this lane's entries: (r, c) = ((lane + 64t) % 24, (lane + 64t) / 24)
    double *out = em + 576 * ec;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int idx = lane + 64 * t, r = idx % 24, c = idx / 24;
        const int ni = r / 3, pi = r % 3, nj = c / 3, pj = c % 3;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double i0 = dxs[v][q][0][ni], i1 = dxs[v][q][1][ni], i2 = dxs[v][q][2][ni];
            const double j0 = dxs[v][q][0][nj], j1 = dxs[v][q][1][nj], j2 = dxs[v][q][2][nj];
            const double dot = i0 * j0 + i1 * j1 + i2 * j2;
            const double ip = dxs[v][q][pi][ni], iq = dxs[v][q][pj][ni];
            const double jp = dxs[v][q][pi][nj], jq = dxs[v][q][pj][nj];
            const double k = lam * ip * jq + mu * iq * jp + (pi == pj ? mu * dot : 0.0);
            acc += dets[v][q] * k;
        }
        if (live) out[idx] = bad ? __builtin_nan("") : acc;
    }
}

int hex8_stiffness(sprgpu_handle *h, i64 nel, i64 nanod, const double *xyz, const i64 *mmnpc,
                   const double *emod, double nu, double *em)
{
    if (nel < 0 || nanod < 0 || (nel > 0 && (!xyz || !mmnpc || !em))) {
        set_error("hex8_stiffness: bad arguments");
        return -2;
    }
    if (!(nu > -1.0 && nu < 0.5)) {
        set_error("hex8_stiffness: Poisson's ratio outside (-1, 0.5)");
        return -2;
    }
    if (nel == 0) return 0;
    const i64 blocks = (nel + 3) / 4;
    if (blocks >= (i64(1) << 31)) {
        set_error("hex8_stiffness: too many elements for one launch");
        return -2;
    }
    k_hex8_stiffness<<<unsigned(blocks), 256, 0, h->stream>>>(nel, nanod, xyz, mmnpc, emod, nu,
                                                              em);
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

}  // namespace sprgpu
