// Element assembly on the GPU.
//
// A host application builds SM by calling SPRADM once per element
// (spradm.f:5-166).  For the free dofs, each call's SPRAD2 loop
// (spradm.f:818-859) runs over the element matrix column by column, J = 1..NEDOF
// and I = 1..J, and adds EM(I,J) to the LNZ slot of the equation pair
// (MEEN(I), MEEN(J)), the slot SPRAD1's pointer table PSM holds
// (spradm.f:514-740).  Every slot therefore ends up as the sum of its
// contributions in a fixed order: by element, then by (J, I).
//
// Here the order is computed once per connectivity on the host (the
// assembly map): the LNZ positions that receive anything, in ascending
// order, and for each one its contributions in SPRADM's order.  Assembly is
// then a gather: one thread per position adds its contributions, in that
// order, to zero (or to the value already in SM), so the result equals the
// host loop's bit for bit.  There are no atomics and the result does not
// depend on the launch shape.  The element matrices stay where the caller
// computed them, in device memory; elements may share one base matrix
// (offsets) and carry a per-element factor (EM_e = escale[e] * base), which
// is what a parameter study or the adjoint changes between assemblies.
//
// The map also gives the residual's CSR pattern directly (the positions with
// a contribution), so no host copy of SM is needed.
//
// Out of scope: the constrained-dof part of SPRAD2 (spradm.f:861-1002, MEEN <
// 0 with TTCC couplings and the SV right-hand side); a connectivity with such
// dofs is rejected.
#include <algorithm>
#include <thread>

#include "sprgpu_internal.h"

namespace sprgpu {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_zero_fill(double *p, i64 n)
{
    const i64 stride = i64(gridDim.x) * 256;
    const i64 t = i64(blockIdx.x) * 256 + threadIdx.x;
    // 16-byte stores over the aligned part, single doubles around it
    const i64 head = (reinterpret_cast<uintptr_t>(p) & 15) ? 1 : 0;
    const i64 n2 = (n - head) / 2;
    double2 *q = reinterpret_cast<double2 *>(p + head);
    for (i64 i = t; i < n2; i += stride) q[i] = double2{0.0, 0.0};
    if (t == 0) {
        if (head && n > 0) p[0] = 0.0;
        if (head + 2 * n2 < n) p[n - 1] = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_assemble(i64 ntgt, const i64 *__restrict__ tpos,
                                                  const i64 *__restrict__ tptr,
                                                  const i64 *__restrict__ src,
                                                  const double *__restrict__ em,
                                                  const i64 *__restrict__ em_off,
                                                  const double *__restrict__ escale,
                                                  double *__restrict__ lnz, int accumulate)
{
    // the host loop scales the base matrix (one rounding) and then adds (a
    // second one): no fused multiply-add here
#pragma clang fp contract(off)
    const i64 stride = i64(gridDim.x) * 256;
    for (i64 t = i64(blockIdx.x) * 256 + threadIdx.x; t < ntgt; t += stride) {
        const i64 pos = tpos[t];
        double s = accumulate ? lnz[pos] : 0.0;
        const i64 c1 = tptr[t + 1];
        for (i64 c = tptr[t]; c < c1; ++c) {
            const i64 v = src[c];
            const i64 e = v >> 32;
            const double x = em[em_off[e] + (v & 0xffffffffll)];
            const double y = escale ? escale[e] * x : x;
            s = s + y;
        }
        lnz[pos] = s;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void free_assembly(sprgpu_handle *h)
{
    DeviceBuffers &d = h->d;
    for (i64 **p : {&d.asm_tpos, &d.asm_tptr, &d.asm_src, &d.asm_emoff}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    d.asm_ntgt = d.asm_ncontrib = d.asm_nel = 0;
}

// stable LSD radix sort of (key, val) pairs by key, 16-bit digits
static void radix_sort_pairs(std::vector<i64> &key, std::vector<i64> &val, i64 max_key)
{
    const size_t n = key.size();
    std::vector<i64> k2(n), v2(n);
    std::vector<size_t> cnt(65537);
    for (int shift = 0; shift < 63 && (max_key >> shift) > 0; shift += 16) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (size_t i = 0; i < n; ++i) cnt[((key[i] >> shift) & 0xffff) + 1]++;
        for (size_t b = 0; b < 65536; ++b) cnt[b + 1] += cnt[b];
        for (size_t i = 0; i < n; ++i) {
            const size_t o = cnt[(key[i] >> shift) & 0xffff]++;
            k2[o] = key[i];
            v2[o] = val[i];
        }
        key.swap(k2);
        val.swap(v2);
    }
}

int set_assembly(sprgpu_handle *h, i64 nel, const i64 *meen_ptr, const i64 *meen)
{
    const Plan &p = h->plan;
    const i64 neq = p.neq;
    if (nel < 0 || nel >= (i64(1) << 31)) {
        set_error("set_assembly: element count out of range");
        return -2;
    }
    if (nel > 0 && meen_ptr[0] != 0) {
        set_error("set_assembly: meen_ptr[0] must be 0");
        return -2;
    }
    std::vector<i64> cstart(nel + 1, 0), emoff(std::max<i64>(nel, 1), 0);
    i64 em_total = 0;
    for (i64 e = 0; e < nel; ++e) {
        const i64 nd = meen_ptr[e + 1] - meen_ptr[e];
        if (nd < 0 || nd > 46340) {
            set_error("set_assembly: element " + std::to_string(e) +
                      " has a dof count outside 0..46340");
            return -2;
        }
        i64 nf = 0;
        for (i64 i = 0; i < nd; ++i) {
            const i64 q = meen[meen_ptr[e] + i];
            if (q < 0) {
                set_error("set_assembly: element " + std::to_string(e) +
                          " has a constrained dof (MEEN < 0); constraint equations are "
                          "assembled by the host SPRADM");
                return -2;
            }
            if (q > neq) {
                set_error("set_assembly: element " + std::to_string(e) +
                          " names an equation beyond NEQ");
                return -2;
            }
            nf += (q >= 1);
        }
        cstart[e + 1] = cstart[e] + nf * (nf + 1) / 2;
        emoff[e] = em_total;
        em_total += nd * nd;
    }
    const i64 total = cstart[nel];

    // supernode of every equation
    std::vector<i32> col_sup(std::max<i64>(neq, 1));
    for (i64 j = 0; j < p.nsuper; ++j)
        for (i32 c = 0; c < p.nodes[j].S; ++c) col_sup[p.nodes[j].first + c] = i32(j);

    // LNZ position of every contribution, generated in SPRADM's order
    std::vector<i64> key(total), val(total);
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<i64> bad(nt, -1);
    auto work = [&](unsigned t) {
        const i64 e0 = nel * t / nt, e1 = nel * (t + 1) / nt;
        for (i64 e = e0; e < e1; ++e) {
            const i64 *mn = meen + meen_ptr[e];
            const i64 nd = meen_ptr[e + 1] - meen_ptr[e];
            i64 k = cstart[e];
            for (i64 j = 0; j < nd; ++j) {
                const i64 jeq = mn[j];
                if (jeq < 1) continue;
                for (i64 i = 0; i <= j; ++i) {
                    const i64 ieq = mn[i];
                    if (ieq < 1) continue;
                    const i32 r = i32(std::max(ieq, jeq) - 1), c = i32(std::min(ieq, jeq) - 1);
                    const NodeInfo &n = p.nodes[col_sup[c]];
                    const i32 kc = c - n.first;
                    const i32 *rows = &p.lindx[n.lxoff];
                    const i32 *it = std::lower_bound(rows + kc, rows + n.F, r);
                    if (it == rows + n.F || *it != r) {
                        bad[t] = e;
                        return;
                    }
                    const i64 idx = it - rows;
                    key[k] = n.xlnz + i64(kc) * n.F - i64(kc) * (kc - 1) / 2 + (idx - kc);
                    val[k] = (e << 32) | (i + j * nd);
                    ++k;
                }
            }
        }
    };
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, t);
        for (auto &x : th) x.join();
    }
    for (unsigned t = 0; t < nt; ++t)
        if (bad[t] >= 0) {
            set_error("set_assembly: element " + std::to_string(bad[t]) +
                      " couples two equations outside the symbolic pattern");
            return -2;
        }
    radix_sort_pairs(key, val, std::max<i64>(p.nzerol - 1, 0));

    // positions and the range of each one's contributions
    std::vector<i64> tpos, tptr;
    tpos.reserve(total / 2 + 1);
    tptr.reserve(total / 2 + 2);
    for (i64 k = 0; k < total; ++k)
        if (k == 0 || key[k] != key[k - 1]) {
            tpos.push_back(key[k]);
            tptr.push_back(k);
        }
    tptr.push_back(total);
    key.clear();
    key.shrink_to_fit();

    // the residual pattern: every position that receives an element entry
    {
        std::vector<uint64_t> marked(size_t(p.nzerol + 63) / 64 + 1, 0);
        for (i64 q : tpos) marked[size_t(q) >> 6] |= uint64_t(1) << (q & 63);
        CsrPattern pat;
        build_csr_pattern_marked(p, marked, pat);
        if (int rc = upload_pattern(h, pat)) return rc;
    }

    free_assembly(h);
    DeviceBuffers &d = h->d;
    auto put = [&](i64 *&dst, const std::vector<i64> &v) -> int {
        SPRGPU_HIP(hipMalloc(&dst, sizeof(i64) * std::max<size_t>(v.size(), 1)));
        if (!v.empty())
            SPRGPU_HIP(hipMemcpy(dst, v.data(), sizeof(i64) * v.size(), hipMemcpyHostToDevice));
        h->device_bytes += sizeof(i64) * v.size();
        return 0;
    };
    if (int rc = put(d.asm_tpos, tpos)) return rc;
    if (int rc = put(d.asm_tptr, tptr)) return rc;
    if (int rc = put(d.asm_src, val)) return rc;
    if (int rc = put(d.asm_emoff, emoff)) return rc;
    d.asm_ntgt = i64(tpos.size());
    d.asm_ncontrib = total;
    d.asm_nel = nel;
    return 0;
}

int run_assemble(sprgpu_handle *h, const double *em, const i64 *em_off, const double *escale,
                 double *sm, int accumulate)
{
    DeviceBuffers &d = h->d;
    if (!d.asm_tpos) {
        set_error("assemble: no assembly map; call sprgpu_set_assembly first");
        return -1;
    }
    const Plan &p = h->plan;
    hipStream_t s = h->stream;
    if (!accumulate) {
        const i64 n = p.neq + p.nzerol;
        const unsigned g = unsigned(std::max<i64>(1, std::min<i64>((n / 2 + 255) / 256, 8192)));
        k_zero_fill<<<g, 256, 0, s>>>(sm, n);
    }
    if (d.asm_ntgt > 0) {
        const unsigned g =
            unsigned(std::max<i64>(1, std::min<i64>((d.asm_ntgt + 255) / 256, 65536)));
        k_assemble<<<g, 256, 0, s>>>(d.asm_ntgt, d.asm_tpos, d.asm_tptr, d.asm_src, em,
                                     em_off ? em_off : d.asm_emoff, escale, sm + p.neq,
                                     accumulate);
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

}  // namespace sprgpu
