// Element assembly on the GPU.
//
// A host application builds SM (and the right-hand side SV) by calling SPRADM
// once per element (spradm.f:5-166).  Each call's SPRAD2 (spradm.f:743-1002)
// first runs over the free dofs, column by column, J = 1..NEDOF and I = 1..J,
// adding EM(I,J) to the LNZ slot of the equation pair (MEEN(I), MEEN(J)), the
// slot SPRAD1's pointer table PSM holds (spradm.f:514-740).  Then, for every
// constrained dof J (MEEN(J) = -JCEQ, equation JCEQ of MPMCEQ/MMCEQ/TTCC), it
// adds the weighted column EM(:,J): to SV, -C0*EM(I,J) (C0 = TTCC of the
// constrained dof itself), and to SM, through each master JP of J,
// TTCC(JP)*EM(I,J) at (MEEN(I), MEQN(master)) -- doubled where the two are
// the same equation -- and TTCC(IP)*TTCC(JP)*EM(I,J) for each master IP of a
// constrained I (spradm.f:861-999).  Every slot therefore ends up as the sum
// of its contributions in a fixed order: by element, then in that loop order.
//
// Here the order is computed once per connectivity on the host (the
// assembly map): the positions that receive anything, in ascending order,
// and for each one its contributions in SPRADM's order.  Assembly is then a
// gather: one thread per position adds its contributions, in that order, to
// zero (or to the value already there), so the result equals the host
// loop's bit for bit.  There are no atomics and the result does not depend
// on the launch shape.  The element matrices stay where the caller computed
// them, in device memory; elements may share one base matrix (offsets) and
// carry a per-element factor (EM_e = escale[e] * base), which is what a
// parameter study or the adjoint changes between assemblies.
//
// Constraint weights are not baked into the map: each contribution carries a
// code naming which TTCC entries form its weight, and TTCC is read from
// device memory at every assembly, so prescribed values (C0) and coupling
// coefficients may change between assemblies without a new map.  The map
// therefore keeps the contributions whose TTCC entry is zero, which SPRAD2
// skips: adding the product 0*EM (a signed zero) leaves every partial sum
// unchanged, and a sum that starts at +0 never becomes -0 in round-to-nearest,
// so with finite EM and the output cleared first the bits are those of the
// host loop.
//
// The map also gives the residual's CSR pattern directly (the LNZ positions
// with a contribution), so no host copy of SM is needed.
#include <algorithm>
#include <thread>

#include "sprgpu_internal.h"

namespace sprgpu {

// weight codes: kind << 56 | IP << 28 | JP (0-based TTCC indices)
enum : int {
    kW1 = 0,      // free pair: EM(I,J)
    kWT = 1,      // TTCC(JP) * EM
    kW2T = 2,     // (TTCC(JP) + TTCC(JP)) * EM, the master's own diagonal
    kWTT = 3,     // (TTCC(IP) * TTCC(JP)) * EM, two masters
    kWC0 = 4,     // -C0 * EM into SV, C0 = TTCC(JP)
    kWC0T = 5,    // -(C0 * TTCC(IP)) * EM into SV
};
constexpr i64 kCodeIdxBits = 28;
constexpr i64 kCodeIdxMask = (i64(1) << kCodeIdxBits) - 1;

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

__device__ __forceinline__ double weight(i64 q, const double *__restrict__ ttcc)
{
#pragma clang fp contract(off)
    const int kind = int(q >> 56);
    const double tj = ttcc[q & kCodeIdxMask];
    switch (kind) {
    case kWT: return tj;
    case kW2T: return tj + tj;
    case kWTT: return ttcc[(q >> kCodeIdxBits) & kCodeIdxMask] * tj;
    case kWC0: return -tj;
    default: return -(tj * ttcc[(q >> kCodeIdxBits) & kCodeIdxMask]);
    }
}

__global__ __launch_bounds__(256) void k_assemble(i64 ntgt, const i64 *__restrict__ tpos,
                                                  const i64 *__restrict__ tptr,
                                                  const i64 *__restrict__ src,
                                                  const i64 *__restrict__ code,
                                                  const double *__restrict__ ttcc,
                                                  const double *__restrict__ em,
                                                  const i64 *__restrict__ em_off,
                                                  const double *__restrict__ escale,
                                                  double *__restrict__ lnz, i64 nzerol,
                                                  double *__restrict__ sv, int accumulate)
{
    // the host loop scales the base matrix (one rounding), weights it (a
    // second one) and then adds or subtracts (a third): no fused multiply-add
    // here.  s - w*x equals s + (-w)*x exactly, so SV's subtractions are
    // additions of the negated weight.
#pragma clang fp contract(off)
    const i64 stride = i64(gridDim.x) * 256;
    for (i64 t = i64(blockIdx.x) * 256 + threadIdx.x; t < ntgt; t += stride) {
        const i64 pos = tpos[t];
        double *dst = pos < nzerol ? lnz + pos : sv + (pos - nzerol);
        double s = accumulate ? *dst : 0.0;
        const i64 c1 = tptr[t + 1];
        for (i64 c = tptr[t]; c < c1; ++c) {
            const i64 v = src[c];
            const i64 e = v >> 32;
            const double x = em[em_off[e] + (v & 0xffffffffll)];
            double y = escale ? escale[e] * x : x;
            if (code) {
                const i64 q = code[c];
                if (q != 0) y = weight(q, ttcc) * y;
            }
            s = s + y;
        }
        *dst = s;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void free_assembly(sprgpu_handle *h)
{
    DeviceBuffers &d = h->d;
    for (i64 **p : {&d.asm_tpos, &d.asm_tptr, &d.asm_src, &d.asm_emoff, &d.asm_code}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    d.asm_ntgt = d.asm_ncontrib = d.asm_nel = d.asm_nsv = 0;
    d.asm_ceq = false;
}

// stable LSD radix sort of (key, val[, code]) by key, 16-bit digits
static void radix_sort_pairs(std::vector<i64> &key, std::vector<i64> &val,
                             std::vector<i64> &code, i64 max_key)
{
    const size_t n = key.size();
    const bool with_code = !code.empty();
    std::vector<i64> k2(n), v2(n), c2(with_code ? n : 0);
    std::vector<size_t> cnt(65537);
    for (int shift = 0; shift < 63 && (max_key >> shift) > 0; shift += 16) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (size_t i = 0; i < n; ++i) cnt[((key[i] >> shift) & 0xffff) + 1]++;
        for (size_t b = 0; b < 65536; ++b) cnt[b + 1] += cnt[b];
        for (size_t i = 0; i < n; ++i) {
            const size_t o = cnt[(key[i] >> shift) & 0xffff]++;
            k2[o] = key[i];
            v2[o] = val[i];
            if (with_code) c2[o] = code[i];
        }
        key.swap(k2);
        val.swap(v2);
        if (with_code) code.swap(c2);
    }
}

namespace {

// The contributions of one element in SPRAD2's order (spradm.f:818-999):
// emit_sm(row, col, entry, code) with 1-based equations row >= col, and
// emit_sv(index, entry, code) with a 0-based index into SV(NEQ*|NSV|).
// entry = I + J*nedof (0-based).  Returns false if a prescribed dof's
// vector lies beyond |NSV|.
struct ElementGen {
    const i64 *meen_ptr, *meen;
    const CeqInput *ceq;         // null: free dofs only
    const std::vector<i64> *prank;  // rank of each equation without masters
    i64 neq;

    template <class SM, class SV>
    bool operator()(i64 e, SM &&emit_sm, SV &&emit_sv) const
    {
        const i64 *mn = meen + meen_ptr[e];
        const i64 nd = meen_ptr[e + 1] - meen_ptr[e];
        const i64 nsv = ceq ? ceq->nsv : 0;
        if (nsv >= 0)
            for (i64 j = 0; j < nd; ++j) {
                const i64 jeq = mn[j];
                if (jeq < 1) continue;
                for (i64 i = 0; i <= j; ++i) {
                    const i64 ieq = mn[i];
                    if (ieq < 1) continue;
                    emit_sm(std::max(ieq, jeq), std::min(ieq, jeq), i + j * nd, i64(0));
                }
            }
        if (!ceq) return true;
        const i64 *mp = ceq->mpmceq, *mm = ceq->mmceq, *meqn = ceq->meqn;
        const i64 nvec = nsv < 0 ? -nsv : nsv;
        auto code = [](int kind, i64 ip, i64 jp) {
            return (i64(kind) << 56) | (ip << kCodeIdxBits) | jp;
        };
        for (i64 j = 0; j < nd; ++j) {
            if (mn[j] >= 0) continue;
            const i64 jq = -mn[j] - 1;                  // 0-based equation
            const i64 jb = mp[jq] - 1, je = mp[jq + 1] - 1;  // its TTCC range
            if (nsv != 0) {
                // SV(KP + ...) -= C0 * EM(:,J) (spradm.f:866-908)
                i64 kp = -1;
                if (nvec == 1) {
                    kp = 0;
                } else if (je - jb == 1) {
                    const i64 k = (*prank)[jq];
                    if (k > nvec) return false;
                    kp = (k - 1) * neq;
                }
                for (i64 i = 0; kp >= 0 && i < nd; ++i) {
                    const i64 ieq = mn[i];
                    if (ieq == 0) continue;
                    if (ieq > 0) {
                        emit_sv(kp + ieq - 1, i + j * nd, code(kWC0, 0, jb));
                        continue;
                    }
                    const i64 iq = -ieq - 1;
                    for (i64 ip = mp[iq]; ip < mp[iq + 1] - 1; ++ip) {
                        if (mm[ip] <= 0) continue;
                        emit_sv(kp + meqn[mm[ip] - 1] - 1, i + j * nd, code(kWC0T, ip, jb));
                    }
                }
                if (nsv < 0) continue;
            }
            // SM through the masters of J (spradm.f:910-999)
            for (i64 jp = jb + 1; jp < je; ++jp) {
                if (mm[jp] <= 0) continue;
                const i64 jeq = meqn[mm[jp] - 1];
                for (i64 i = 0; i < nd; ++i) {
                    const i64 ieq = mn[i];
                    if (ieq == 0) continue;
                    if (ieq > 0) {
                        if (ieq != jeq)
                            emit_sm(std::max(ieq, jeq), std::min(ieq, jeq), i + j * nd,
                                    code(kWT, 0, jp));
                        else
                            emit_sm(ieq, ieq, i + j * nd, code(kW2T, 0, jp));
                        continue;
                    }
                    const i64 iq = -ieq - 1;
                    for (i64 ip = mp[iq]; ip < mp[iq + 1] - 1; ++ip) {
                        if (mm[ip] <= 0) continue;
                        const i64 meq = meqn[mm[ip] - 1];
                        if (meq <= jeq) emit_sm(jeq, meq, i + j * nd, code(kWTT, ip, jp));
                    }
                }
            }
        }
        return true;
    }
};

}  // namespace

int set_assembly(sprgpu_handle *h, i64 nel, const i64 *meen_ptr, const i64 *meen,
                 const CeqInput *ceq)
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
    // the constraint equations: MPMCEQ(1) = 1, at least the constrained dof
    // per equation, masters free (SPRSEQ's checks, sprseq.f:108-140)
    std::vector<i64> prank;
    if (ceq) {
        const i64 nceq = ceq->nceq;
        if (nceq < 0 || (nceq > 0 && (!ceq->mpmceq || !ceq->mmceq || !ceq->meqn)) ||
            ceq->ndof < 0) {
            set_error("set_assembly: bad constraint-equation arguments");
            return -2;
        }
        if (nceq > 0 && ceq->mpmceq[0] != 1) {
            set_error("set_assembly: MPMCEQ(1) must be 1");
            return -2;
        }
        prank.assign(std::max<i64>(nceq, 1), 0);
        i64 k = 0;
        for (i64 q = 0; q < nceq; ++q) {
            const i64 n = ceq->mpmceq[q + 1] - ceq->mpmceq[q];
            if (n < 1) {
                set_error("set_assembly: constraint equation " + std::to_string(q + 1) +
                          " has no constrained dof (MPMCEQ not increasing)");
                return -2;
            }
            if (n == 1) ++k;
            prank[q] = k;
        }
        const i64 nmmceq = nceq > 0 ? ceq->mpmceq[nceq] - 1 : 0;
        if (nmmceq > kCodeIdxMask) {
            set_error("set_assembly: more than 2^28 constraint coefficients");
            return -2;
        }
        for (i64 q = 0; q < nceq; ++q)
            for (i64 ip = ceq->mpmceq[q]; ip < ceq->mpmceq[q + 1] - 1; ++ip) {
                const i64 md = ceq->mmceq[ip];
                if (md <= 0) continue;  // SPRAD2 skips it
                if (md > ceq->ndof || ceq->meqn[md - 1] < 1 || ceq->meqn[md - 1] > neq) {
                    set_error("set_assembly: a master dof of constraint equation " +
                              std::to_string(q + 1) + " has no equation (MEQN not in 1..NEQ)");
                    return -2;
                }
            }
    }
    const i64 nvec = ceq ? std::abs(ceq->nsv) : 0;
    std::vector<i64> emoff(std::max<i64>(nel, 1), 0);
    i64 em_total = 0;
    bool any_ceq = false;
    for (i64 e = 0; e < nel; ++e) {
        const i64 nd = meen_ptr[e + 1] - meen_ptr[e];
        if (nd < 0 || nd > 46340) {
            set_error("set_assembly: element " + std::to_string(e) +
                      " has a dof count outside 0..46340");
            return -2;
        }
        for (i64 i = 0; i < nd; ++i) {
            const i64 q = meen[meen_ptr[e] + i];
            if (q < 0 && !ceq) {
                set_error("set_assembly: element " + std::to_string(e) +
                          " has a constrained dof (MEEN < 0); give its constraint equations "
                          "(sprgpu_set_assembly_ceq)");
                return -2;
            }
            if (q < 0 && -q > ceq->nceq) {
                set_error("set_assembly: element " + std::to_string(e) +
                          " names a constraint equation beyond NCEQ");
                return -2;
            }
            if (q > neq) {
                set_error("set_assembly: element " + std::to_string(e) +
                          " names an equation beyond NEQ");
                return -2;
            }
            any_ceq |= q < 0;
        }
        emoff[e] = em_total;
        em_total += nd * nd;
    }
    const ElementGen gen{meen_ptr, meen, (ceq && (any_ceq || ceq->nsv != 0)) ? ceq : nullptr,
                         &prank, neq};
    const bool with_code = gen.ceq != nullptr && any_ceq;

    // contributions per element, then their positions, in SPRADM's order
    std::vector<i64> cstart(nel + 1, 0);
    for (i64 e = 0; e < nel; ++e) {
        i64 n = 0;
        if (!gen(e, [&](i64, i64, i64, i64) { ++n; }, [&](i64, i64, i64) { ++n; })) {
            set_error("set_assembly: element " + std::to_string(e) +
                      " has a prescribed dof whose vector lies beyond |NSV|");
            return -2;
        }
        cstart[e + 1] = cstart[e] + n;
    }
    const i64 total = cstart[nel];

    // supernode of every equation
    std::vector<i32> col_sup(std::max<i64>(neq, 1));
    for (i64 j = 0; j < p.nsuper; ++j)
        for (i32 c = 0; c < p.nodes[j].S; ++c) col_sup[p.nodes[j].first + c] = i32(j);

    std::vector<i64> key(total), val(total), code(with_code ? total : 0);
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<i64> bad(nt, -1);
    auto work = [&](unsigned t) {
        const i64 e0 = nel * t / nt, e1 = nel * (t + 1) / nt;
        for (i64 e = e0; e < e1; ++e) {
            i64 k = cstart[e];
            bool ok = true;
            auto sm = [&](i64 row, i64 col, i64 entry, i64 w) {
                const i32 r = i32(row - 1), c = i32(col - 1);
                const NodeInfo &n = p.nodes[col_sup[c]];
                const i32 kc = c - n.first;
                const i32 *rows = &p.lindx[n.lxoff];
                const i32 *it = std::lower_bound(rows + kc, rows + n.F, r);
                if (it == rows + n.F || *it != r) {
                    ok = false;
                    key[k] = 0;
                } else {
                    key[k] = n.xlnz + i64(kc) * n.F - i64(kc) * (kc - 1) / 2 + ((it - rows) - kc);
                }
                val[k] = (e << 32) | entry;
                if (with_code) code[k] = w;
                ++k;
            };
            auto sv = [&](i64 idx, i64 entry, i64 w) {
                key[k] = p.nzerol + idx;
                val[k] = (e << 32) | entry;
                if (with_code) code[k] = w;
                ++k;
            };
            gen(e, sm, sv);
            if (!ok) {
                bad[t] = e;
                return;
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
    radix_sort_pairs(key, val, code, std::max<i64>(p.nzerol + nvec * neq - 1, 0));

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

    // the residual pattern: every LNZ position that receives an element entry
    if (!ceq || ceq->nsv >= 0) {
        std::vector<uint64_t> marked(size_t(p.nzerol + 63) / 64 + 1, 0);
        for (i64 q : tpos)
            if (q < p.nzerol) marked[size_t(q) >> 6] |= uint64_t(1) << (q & 63);
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
    if (with_code)
        if (int rc = put(d.asm_code, code)) return rc;
    d.asm_ntgt = i64(tpos.size());
    d.asm_ncontrib = total;
    d.asm_nel = nel;
    d.asm_nsv = ceq ? ceq->nsv : 0;
    d.asm_ceq = ceq != nullptr;
    return 0;
}

int run_assemble(sprgpu_handle *h, const double *em, const i64 *em_off, const double *escale,
                 const double *ttcc, double *sm, double *sv, int accumulate)
{
    DeviceBuffers &d = h->d;
    if (!d.asm_tpos) {
        set_error("assemble: no assembly map; call sprgpu_set_assembly first");
        return -1;
    }
    if (d.asm_code && !ttcc) {
        set_error("assemble: the map has constraint equations; TTCC is required");
        return -2;
    }
    if (d.asm_nsv >= 0 && !sm) {
        set_error("assemble: SM is required (NSV >= 0)");
        return -2;
    }
    if (d.asm_nsv != 0 && !sv) {
        set_error("assemble: SV is required (NSV != 0)");
        return -2;
    }
    const Plan &p = h->plan;
    hipStream_t s = h->stream;
    if (!accumulate) {
        auto fill = [&](double *x, i64 n) {
            const unsigned g = unsigned(std::max<i64>(1, std::min<i64>((n / 2 + 255) / 256, 8192)));
            k_zero_fill<<<g, 256, 0, s>>>(x, n);
        };
        if (d.asm_nsv >= 0) fill(sm, p.neq + p.nzerol);
        if (d.asm_nsv != 0) fill(sv, std::abs(d.asm_nsv) * p.neq);
    }
    if (d.asm_ntgt > 0) {
        const unsigned g =
            unsigned(std::max<i64>(1, std::min<i64>((d.asm_ntgt + 255) / 256, 65536)));
        k_assemble<<<g, 256, 0, s>>>(d.asm_ntgt, d.asm_tpos, d.asm_tptr, d.asm_src, d.asm_code,
                                     ttcc, em, em_off ? em_off : d.asm_emoff, escale,
                                     sm ? sm + p.neq : nullptr, p.nzerol, sv, accumulate);
    }
    SPRGPU_HIP(hipGetLastError());
    return 0;
}

}  // namespace sprgpu
