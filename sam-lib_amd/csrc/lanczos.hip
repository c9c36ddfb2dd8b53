// Shift-invert Lanczos on the device: the task of SPRLAN (sprlan.f:5-945).
//
// SPRLAN finds the NEVAL eigenvalues of (A - lambda B) q = 0 closest to a
// SHIFT, with B-orthonormal eigenvectors, by a truncated Lanczos method on
// (A - SHIFT B)^-1 B whose tridiagonal matrix is then diagonalised
// (sprlan.f:14-31).  Here every vector operation runs in HBM:
// * A - SHIFT B is formed on the device and factored by the handle (FP64
//   LDL'); a pivot failure means SHIFT is (close to) an eigenvalue;
// * each step is one forward/backward solve, one SPRPRM product with B
//   (B sparse on the handle's pattern, or diagonal), and full
//   reorthogonalisation in the B inner product (twice classical
//   Gram-Schmidt against all previous vectors, kept with their B products);
// * the tridiagonal eigenproblem (at most MAXLAN) is solved on the host by
//   cyclic Jacobi rotations, which also give the Ritz vectors' coefficients.
// Eigenvalue lambda = SHIFT + 1/theta for a Ritz value theta; its error
// estimate is |beta_m s_mi| / theta^2, and it counts as converged when that
// is at most TOL |lambda|.  The start vector is a fixed pseudo-random vector,
// so results are reproducible.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "sprgpu_internal.h"

namespace sprgpu {

constexpr int kLanDotBlocks = 256;

__global__ __launch_bounds__(256) void k_lan_start(double *q, i64 n)
{
    const i64 i = i64(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    q[i] = double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;  // uniform in [-1, 1)
}

// out = a - sigma * b over n entries
__global__ __launch_bounds__(256) void k_lan_shift(const double *a, const double *b, double sigma,
                                                   double *out, i64 n)
{
    for (i64 i = i64(blockIdx.x) * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256)
        out[i] = a[i] - sigma * b[i];
}

// diagonal B: the diagonal slot of column j of the LNZ part loses sigma*b_j
__global__ __launch_bounds__(256) void k_lan_shift_diag(const i32 *__restrict__ col_node,
                                                        const NodeInfo *__restrict__ nodes,
                                                        const double *b, double sigma,
                                                        double *lnz, i64 n)
{
    const i64 j = i64(blockIdx.x) * 256 + threadIdx.x;
    if (j >= n) return;
    const NodeInfo nd = nodes[col_node[j]];
    const i64 k = j - nd.first;
    lnz[nd.xlnz + k * nd.F - k * (k - 1) / 2] -= sigma * b[j];
}

// part[c*kLanDotBlocks + block] = partial sums of X(:,c) . y
__global__ __launch_bounds__(256) void k_lan_dots(const double *X, i64 ldx, const double *y,
                                                  i64 n, double *part)
{
    const int c = blockIdx.y;
    double s = 0.0;
    for (i64 i = i64(blockIdx.x) * 256 + threadIdx.x; i < n; i += i64(gridDim.x) * 256)
        s += X[c * ldx + i] * y[i];
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[c * kLanDotBlocks + blockIdx.x] = sh[0];
}

__global__ void k_lan_dots_final(const double *part, double *out, int ncol)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    double s = 0.0;
    for (int b = 0; b < kLanDotBlocks; ++b) s += part[c * kLanDotBlocks + b];
    out[c] = s;
}

// w -= X(:, 0..ncol-1) c
__global__ __launch_bounds__(256) void k_lan_sub_gemv(const double *X, i64 ldx,
                                                      const double *c, int ncol, double *w,
                                                      i64 n)
{
    const i64 i = i64(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < ncol; ++k) s += X[k * ldx + i] * c[k];
    w[i] -= s;
}

// y = a*x + b*y
__global__ __launch_bounds__(256) void k_lan_axpby(double a, const double *x, double b, double *y,
                                                   i64 n)
{
    const i64 i = i64(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

// V(:, e) = X(:, 0..m-1) S(:, e), S column-major m x nev
__global__ __launch_bounds__(256) void k_lan_ritz(const double *X, i64 ldx, const double *S,
                                                  int m, double *V, i64 n)
{
    const i64 i = i64(blockIdx.x) * 256 + threadIdx.x;
    const int e = blockIdx.y;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < m; ++k) s += X[k * ldx + i] * S[k + i64(e) * m];
    V[e * n + i] = s;
}

// Eigenvalues and vectors of the symmetric tridiagonal T (alpha, beta) of
// order m by cyclic Jacobi rotations (Golub and Van Loan, Matrix
// Computations, section 8.5): T = S diag(theta) S'.
static void tridiag_eigen(int m, const std::vector<double> &alpha, const std::vector<double> &beta,
                          std::vector<double> &theta, std::vector<double> &S)
{
    std::vector<double> A(size_t(m) * m, 0.0);
    for (int i = 0; i < m; ++i) {
        A[i + i * m] = alpha[i];
        if (i + 1 < m) A[i + (i + 1) * m] = A[(i + 1) + i * m] = beta[i];
    }
    S.assign(size_t(m) * m, 0.0);
    for (int i = 0; i < m; ++i) S[i + i * m] = 1.0;
    double fro = 0.0;
    for (double v : A) fro += v * v;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int q = 0; q < m; ++q)
            for (int p = 0; p < q; ++p) off += A[p + q * m] * A[p + q * m];
        if (off <= 1e-32 * fro) break;
        for (int q = 1; q < m; ++q)
            for (int p = 0; p < q; ++p) {
                const double apq = A[p + q * m];
                if (apq == 0.0) continue;
                const double tau = (A[q + q * m] - A[p + p * m]) / (2.0 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < m; ++k) {  // A <- A J (columns p, q)
                    const double akp = A[k + p * m], akq = A[k + q * m];
                    A[k + p * m] = c * akp - s * akq;
                    A[k + q * m] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {  // A <- J' A (rows p, q)
                    const double apk = A[p + k * m], aqk = A[q + k * m];
                    A[p + k * m] = c * apk - s * aqk;
                    A[q + k * m] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {  // S <- S J
                    const double skp = S[k + p * m], skq = S[k + q * m];
                    S[k + p * m] = c * skp - s * skq;
                    S[k + q * m] = s * skp + c * skq;
                }
            }
    }
    theta.resize(m);
    for (int i = 0; i < m; ++i) theta[i] = A[i + i * m];
}

int run_lanczos(sprgpu_handle *h, const double *a_sm, const double *b, int ksb, double shift,
                i64 neval, i64 nevec, i64 maxlan, double tol, double *evl, double *everr,
                double *v, i64 *nconv)
{
    const Plan &p = h->plan;
    const i64 n = p.neq, nz = p.nzerol;
    if (h->precision != SPRGPU_FP64) {
        set_error("lanczos: needs an FP64 handle (the operator is the factor's inverse)");
        return -2;
    }
    if (neval < 1 || nevec < 0 || nevec > neval || maxlan < neval + 1 || maxlan > n) {
        set_error("lanczos: needs 1 <= NEVAL < MAXLAN <= NEQ and 0 <= NEVEC <= NEVAL");
        return -2;
    }
    if (ksb > 0 && !h->has_pattern) {
        set_error("lanczos: a sparse B needs the handle's pattern (sprgpu_set_pattern or "
                  "sprgpu_set_assembly) to cover it");
        return -1;
    }
    hipStream_t s = h->stream;
    const unsigned gn = unsigned((n + 255) / 256);
    const i64 m1 = maxlan + 1;
    // device work space: A - shift B, Q and B Q (n x (maxlan+1) each), w, bw,
    // dot partials and coefficients
    double *sm_s = nullptr, *Q = nullptr, *BQ = nullptr, *w = nullptr, *bw = nullptr,
           *part = nullptr, *coef = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(s);
        // the handle keeps the factor of A - SHIFT B, not its freed input
        if (h->last_sm_dev == sm_s) h->last_sm_dev = nullptr;
        for (double *q : {sm_s, Q, BQ, w, bw, part, coef})
            if (q) (void)hipFree(q);
    };
#define LAN_HIP(call)                                                  \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) {                                        \
            release();                                                 \
            return hip_fail(e_, #call);                                \
        }                                                              \
    } while (0)
    LAN_HIP(hipMalloc(&sm_s, sizeof(double) * (n + nz)));
    LAN_HIP(hipMalloc(&Q, sizeof(double) * n * m1));
    LAN_HIP(hipMalloc(&BQ, sizeof(double) * n * m1));
    LAN_HIP(hipMalloc(&w, sizeof(double) * n));
    LAN_HIP(hipMalloc(&bw, sizeof(double) * n));
    LAN_HIP(hipMalloc(&part, sizeof(double) * kLanDotBlocks * m1));
    LAN_HIP(hipMalloc(&coef, sizeof(double) * std::max<i64>(m1, maxlan * nevec + 1)));

    // A - shift B and its factor
    if (ksb > 0) {
        LAN_HIP(hipMemcpyAsync(sm_s, a_sm, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        k_lan_shift<<<unsigned(std::min<i64>((nz + 255) / 256, 65536)), 256, 0, s>>>(
            a_sm + n, b, shift, sm_s + n, nz);
    } else {
        LAN_HIP(hipMemcpyAsync(sm_s, a_sm, sizeof(double) * (n + nz), hipMemcpyDeviceToDevice, s));
        if (int rc = upload_col_node(h)) {
            release();
            return rc;
        }
        k_lan_shift_diag<<<gn, 256, 0, s>>>(h->d.col_node, h->d.nodes, b, shift, sm_s + n, n);
    }
    {
        const double t4[4] = {0.0, 0.0, 0.0, 0.0};
        sprgpu_factor_info info{};
        int rc = run_factor(h, sm_s, 1, t4);
        if (!rc) rc = read_factor_info(h, &info);
        if (rc || info.ierr < 0) {
            release();
            if (!rc) set_error("lanczos: A - SHIFT*B is singular (SHIFT is an eigenvalue)");
            return rc ? rc : -3;
        }
    }
    auto bprod = [&](const double *x, double *y) {  // y = B x
        return run_sprprm(h, b, x, y, nullptr, n, 1, ksb > 0 ? 1 : 0, 1);
    };
    // dots[c] = X(:,c) . y for c < ncol, to the host
    std::vector<double> hdots(m1);
    auto dots = [&](const double *X, int ncol, const double *y) -> int {
        k_lan_dots<<<dim3(kLanDotBlocks, unsigned(ncol)), 256, 0, s>>>(X, n, y, n, part);
        k_lan_dots_final<<<unsigned((ncol + 63) / 64), 64, 0, s>>>(part, coef, ncol);
        // errors are returned; the caller releases the work space
        hipError_t e = hipMemcpyAsync(hdots.data(), coef, sizeof(double) * ncol,
                                      hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        return e == hipSuccess ? 0 : hip_fail(e, "lanczos: dot products");
    };
    // start: q0 = r / |r|_B
    k_lan_start<<<gn, 256, 0, s>>>(Q, n);
    if (int rc = bprod(Q, BQ)) {
        release();
        return rc;
    }
    if (int rc = dots(Q, 1, BQ)) {
        release();
        return rc;
    }
    {
        const double nrm = std::sqrt(hdots[0]);
        k_lan_axpby<<<gn, 256, 0, s>>>(0.0, Q, 1.0 / nrm, Q, n);
        k_lan_axpby<<<gn, 256, 0, s>>>(0.0, BQ, 1.0 / nrm, BQ, n);
    }
    std::vector<double> alpha, beta, theta, S;
    std::vector<int> order;
    i64 m = 0, good = 0;
    double beta_last = 0.0;
    for (i64 j = 0; j < maxlan; ++j) {
        double *qj = Q + j * n, *bqj = BQ + j * n;
        // w = (A - shift B)^-1 B q_j
        LAN_HIP(hipMemcpyAsync(w, bqj, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        if (int rc = run_solve(h, w, n, 1, 3)) {
            release();
            return rc;
        }
        // full reorthogonalisation, twice: w -= Q (BQ' w); the first pass's
        // coefficient on q_j is alpha_j
        double a_j = 0.0;
        for (int pass = 0; pass < 2; ++pass) {
            if (int rc = dots(BQ, int(j + 1), w)) {
                release();
                return rc;
            }
            if (pass == 0) a_j = hdots[j];
            k_lan_sub_gemv<<<gn, 256, 0, s>>>(Q, n, coef, int(j + 1), w, n);
        }
        if (int rc = bprod(w, bw)) {
            release();
            return rc;
        }
        if (int rc = dots(w, 1, bw)) {
            release();
            return rc;
        }
        const double b_j = std::sqrt(std::max(hdots[0], 0.0));
        alpha.push_back(a_j);
        m = j + 1;
        beta_last = b_j;
        const bool breakdown = !(b_j > 1e-14 * std::fabs(a_j));
        if (!breakdown) {
            beta.push_back(b_j);
            k_lan_axpby<<<gn, 256, 0, s>>>(1.0 / b_j, w, 0.0, Q + (j + 1) * n, n);
            k_lan_axpby<<<gn, 256, 0, s>>>(1.0 / b_j, bw, 0.0, BQ + (j + 1) * n, n);
        }
        // Ritz values every few steps once there can be NEVAL of them
        if (m >= neval && (breakdown || m == maxlan || (m - neval) % 4 == 0)) {
            tridiag_eigen(int(m), alpha, beta, theta, S);
            order.resize(m);
            std::iota(order.begin(), order.end(), 0);
            std::sort(order.begin(), order.end(),
                      [&](int x, int y) { return std::fabs(theta[x]) > std::fabs(theta[y]); });
            good = 0;
            for (i64 e = 0; e < neval; ++e) {
                const int r = order[e];
                const double th = theta[r];
                const double lam = shift + 1.0 / th;
                const double err = std::fabs((breakdown ? 0.0 : beta_last) * S[(m - 1) + r * m]) /
                                   (th * th);
                evl[e] = lam;
                everr[e] = err;
                good += (err <= tol * std::fabs(lam));
            }
            if (good == neval || breakdown) break;
        }
    }
    *nconv = good;
    if (nevec > 0) {  // V(:, e) = Q S(:, order[e])
        std::vector<double> Ssel(size_t(m) * nevec);
        for (i64 e = 0; e < nevec; ++e)
            for (i64 k = 0; k < m; ++k) Ssel[k + e * m] = S[k + size_t(order[e]) * m];
        LAN_HIP(hipMemcpyAsync(coef, Ssel.data(), sizeof(double) * Ssel.size(),
                               hipMemcpyHostToDevice, s));
        k_lan_ritz<<<dim3(gn, unsigned(nevec)), 256, 0, s>>>(Q, n, coef, int(m), v, n);
        LAN_HIP(hipStreamSynchronize(s));
    }
    LAN_HIP(hipGetLastError());
    release();
#undef LAN_HIP
    return 0;
}

}  // namespace sprgpu
