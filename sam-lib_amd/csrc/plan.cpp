// Host planner: turns the reference symbolic output into the device schedule.
//
// Inputs follow the reference layout exactly (sprfc1.f:161-203):
//   MTREES(MSPAR(45)...) = XSUPER(NSUPER+1), XSIBL(NSUPER+1), SIBL(NSUPER+1)
//   MTREES(MSPAR(46)...) = XLINDX(NSUPER+1), SUPSUP(NSUPER), ...
//   MSIFA(MSPAR(47)...)  = LINDX(NOFSUB);  MSIFA(MSPAR(48)...) = XLNZ(NSUPER+1)
// Supernodes are in postorder (children before parents), LINDX lists are
// sorted ascending and start with the supernode's own columns
// (sprsmb.f:512-575).
//
// The schedule groups supernodes by height in the elimination tree (leaves
// at height 0): every supernode of one height is independent of the others,
// so each kernel launch covers a whole height.  Within a height, children of
// one parent are extend-added in rounds by sibling index so that no two
// work items of one launch write the same front entry.
#include <algorithm>
#include <map>
#include <numeric>
#include <thread>

#include "sprgpu_internal.h"

namespace sprgpu {

namespace {

inline i64 tri_tiles(i64 nti, i64 ntj)
{
    // tiles (ti,tj) with 0<=tj<ntj, tj<=ti<nti
    return ntj * nti - ntj * (ntj - 1) / 2;
}

inline i64 round_up(i64 x, i64 m) { return (x + m - 1) / m * m; }

}  // namespace

std::string build_plan(const i64 *mspar, const i64 *mtrees, const i64 *msifa, int precision,
                       const PlanParams &params, Plan &p)
{
    p.params = params;
    if (params.small_f < 0 || params.small_f > kSmallF) return "small_f must be in [0, 64]";
    if (params.small_waves < 1 || params.small_waves > 4) return "small_waves must be in [1, 4]";
    if (params.big_tile != 0 && params.big_tile != kTileSmall && params.big_tile != kTileBig)
        return "big_tile must be 0, 64 or 128";
    if (params.narrow_threads != 0 && params.narrow_threads != 64 &&
        params.narrow_threads != 128 && params.narrow_threads != 256)
        return "narrow_threads must be 0, 64, 128 or 256";
    if (params.block_factor < 0 || params.block_factor > 2) return "block_factor must be 0, 1 or 2";
    p.big_tile = params.big_tile;  // 0: chosen per launch (below)
    // Tile of a rank-kOuterRank / Schur update launch: the configured one,
    // or with big_tile 0 in FP32 the 128 tile (32x32x2 MFMA kernel) for
    // launches of at least kAutoBigTiles 64-tiles, where it runs faster, and
    // the 64 tile for smaller ones, where it leaves fewer workgroups idle
    auto outer_tile = [&](i64 tiles64) -> i32 {
        if (p.big_tile) return p.big_tile;
        return (precision == SPRGPU_FP32 && tiles64 >= kAutoBigTiles) ? kTileBig : kTileSmall;
    };
    p.small_f = params.small_f;
    auto M = [&](int k) { return mspar[k - 1]; };
    if (M(1) < 0) return "MSPAR(1) < 0: the symbolic phase did not complete";
    p.neq = M(8);
    p.nsuper = M(11);
    p.maxsup = M(13);
    p.mfront = M(14);
    p.nofsub = M(15);
    p.nzerol = M(16);
    if (p.neq <= 0 || p.nsuper <= 0) return "empty system";
    if (p.neq >= (i64(1) << 31)) return "NEQ too large for 32-bit row indices";
    const i64 ns = p.nsuper;
    const i64 *xsuper = mtrees + (M(45) - 1);
    const i64 *xsibl = xsuper + (ns + 1);
    const i64 *sibl = xsibl + (ns + 1);
    const i64 *xlindx = mtrees + (M(46) - 1);
    const i64 *supsup = xlindx + (ns + 1);
    const i64 *lindx = msifa + (M(47) - 1);
    const i64 *xlnz = msifa + (M(48) - 1);
    for (i64 j = 0; j < ns; ++j)
        if (supsup[j] > 0)
            return "retained supernodes (SUPSUP>0, a superelement separator) are not supported";

    p.lindx.resize(p.nofsub);
    for (i64 k = 0; k < p.nofsub; ++k) p.lindx[k] = i32(lindx[k] - 1);

    p.nodes.resize(ns);
    p.children.clear();
    p.children.reserve(ns);
    std::vector<i32> parent(ns, -1);
    for (i64 j = 0; j < ns; ++j) {
        NodeInfo &n = p.nodes[j];
        n.first = i32(xsuper[j] - 1);
        n.S = i32(xsuper[j + 1] - xsuper[j]);
        n.F = i32(xlindx[j + 1] - xlindx[j]);
        n.lxoff = xlindx[j] - 1;
        n.xlnz = xlnz[j] - 1;
        n.child0 = i32(p.children.size());
        for (i64 c = xsibl[j]; c < xsibl[j + 1]; ++c) {
            const i32 ch = i32(sibl[c - 1] - 1);
            if (ch < 0 || ch >= j) return "SIBL does not describe a postordered tree";
            p.children.push_back(ch);
            parent[ch] = i32(j);
        }
        n.nchild = i32(p.children.size()) - n.child0;
        if (n.S <= 0 || n.F < n.S) return "inconsistent supernode partition";
    }
    for (i64 j = 0; j < ns; ++j) p.nodes[j].parent = parent[j];

    // heights
    std::vector<i32> height(ns, 0);
    for (i64 j = 0; j < ns; ++j)
        if (parent[j] >= 0) height[parent[j]] = std::max(height[parent[j]], height[j] + 1);
    p.nlevels = ns ? 1 + *std::max_element(height.begin(), height.end()) : 0;
    p.level_ptr.assign(p.nlevels + 1, 0);
    for (i64 j = 0; j < ns; ++j) p.level_ptr[height[j] + 1]++;
    for (int l = 0; l < p.nlevels; ++l) p.level_ptr[l + 1] += p.level_ptr[l];
    p.level_nodes.resize(ns);
    {
        std::vector<i32> fill(p.level_ptr.begin(), p.level_ptr.end() - 1);
        for (i64 j = 0; j < ns; ++j) p.level_nodes[fill[height[j]]++] = i32(j);
    }

    // factor store, vector workspace and child relative maps (SPRF12 logic,
    // sprfc1.f:716-790: monotone merge of the sorted row lists).  The store
    // holds each supernode's factored panel (its S columns, leading
    // dimension ld) for the solves.
    i64 store = 0, vw = 0, rm = 0;
    for (i64 j = 0; j < ns; ++j) {
        NodeInfo &n = p.nodes[j];
        n.ld = i32(round_up(n.F, kFrontLdAlign));
        n.loff = store;
        store += round_up(i64(n.ld) * n.S, 64);
        n.voff = vw;
        vw += n.F;
        n.mapoff = rm;
        rm += n.F - n.S;
    }
    p.store_elems = store;
    p.vwork_elems = vw;
    // Front pool: a front (ld x F) lives from its level's initialisation to
    // its parent's extend-add, in the parent's level; its panel goes to the
    // store at the end of its level (on the second stream, which the next
    // level's end waits for: a root's memory lives one level more).  Two
    // placements are made and the smaller kept:
    //  - level by level, the level's fronts (those living longest first,
    //    then the largest) by first fit over the free ranges, after which the
    //    fronts whose parents are in the level are released;
    //  - the kPoolExactFronts largest fronts first, largest first, each at
    //    the lowest offset that no placed front overlaps during its lifetime,
    //    and the others above them as in the first placement.
    // (The first reaches the most front memory alive at once on 2-D grids,
    // the second on 3-D ones, whose large separators live long.)
    {
        auto fsize = [&](i64 j) { return round_up(i64(p.nodes[j].ld) * p.nodes[j].F, 64); };
        // the last level of a front's memory: its parent's level (a root:
        // the level after its own, as its panel is stored during that level)
        auto last = [&](i64 j) { return parent[j] >= 0 ? height[parent[j]] : height[j] + 1; };
        // first fit, level by level, over the fronts with skip[j] == 0;
        // offsets from base; returns the top
        auto by_levels = [&](const std::vector<char> &skip, i64 base, std::vector<i64> &off) {
            std::vector<std::vector<i32>> die(p.nlevels + 1);
            for (i64 j = 0; j < ns; ++j)
                if (!skip[j]) die[last(j)].push_back(i32(j));
            std::map<i64, i64> freel;  // offset -> length of the free ranges below top
            i64 top = 0;
            auto take = [&](i64 len) {
                for (auto it = freel.begin(); it != freel.end(); ++it) {
                    if (it->second < len) continue;
                    const i64 o = it->first, rest = it->second - len;
                    freel.erase(it);
                    if (rest > 0) freel[o + len] = rest;
                    return o;
                }
                const i64 o = top;
                top += len;
                return o;
            };
            auto give = [&](i64 o, i64 len) {
                auto nx = freel.lower_bound(o);
                if (nx != freel.end() && nx->first == o + len) {
                    len += nx->second;
                    freel.erase(nx);
                }
                auto pv = freel.lower_bound(o);
                if (pv != freel.begin()) {
                    --pv;
                    if (pv->first + pv->second == o) {
                        pv->second += len;
                        return;
                    }
                }
                freel[o] = len;
            };
            for (int l = 0; l < p.nlevels; ++l) {
                std::vector<i32> v;
                for (i32 q = p.level_ptr[l]; q < p.level_ptr[l + 1]; ++q)
                    if (!skip[p.level_nodes[q]]) v.push_back(p.level_nodes[q]);
                std::stable_sort(v.begin(), v.end(), [&](i32 a, i32 b) {
                    return last(a) != last(b) ? last(a) > last(b) : fsize(a) > fsize(b);
                });
                for (i32 j : v) off[j] = base + take(fsize(j));
                for (i32 j : die[l]) give(off[j] - base, fsize(j));
            }
            return top;
        };
        std::vector<i64> off1(ns), off2(ns);
        const i64 top1 = by_levels(std::vector<char>(ns, 0), 0, off1);
        std::vector<i32> order(ns);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](i32 a, i32 b) { return fsize(a) > fsize(b); });
        const i64 nbig = std::min<i64>(ns, kPoolExactFronts);
        std::vector<char> big(ns, 0);
        struct Placed {
            i64 off, end;
            i32 a, b;  // levels of the lifetime
        };
        std::vector<Placed> placed;  // ascending offset
        placed.reserve(nbig);
        i64 topb = 0;
        for (i64 q = 0; q < nbig; ++q) {
            const i32 j = order[q], a = height[j], b = last(j);
            const i64 len = fsize(j);
            i64 cand = 0;  // scanning by offset: past every overlapping front that starts below
            for (const Placed &u : placed) {
                if (u.b < a || b < u.a) continue;
                if (u.off >= cand + len) break;
                cand = std::max(cand, u.end);
            }
            const Placed np{cand, cand + len, a, b};
            placed.insert(std::upper_bound(placed.begin(), placed.end(), np,
                                           [](const Placed &x, const Placed &y) { return x.off < y.off; }),
                          np);
            off2[j] = cand;
            topb = std::max(topb, cand + len);
            big[j] = 1;
        }
        const i64 top2 = topb + by_levels(big, topb, off2);
        const std::vector<i64> &off = top1 <= top2 ? off1 : off2;
        for (i64 j = 0; j < ns; ++j) p.nodes[j].foff = off[j];
        p.pool_elems = std::min(top1, top2);
        // check: the fronts alive in each level are disjoint and inside the pool
        std::vector<std::vector<std::pair<i64, i64>>> alive(p.nlevels + 1);
        for (i64 j = 0; j < ns; ++j)
            for (i32 l = height[j]; l <= last(j); ++l) alive[l].push_back({off[j], off[j] + fsize(j)});
        for (auto &v : alive) {
            std::sort(v.begin(), v.end());
            for (size_t k = 0; k < v.size(); ++k)
                if (v[k].second > p.pool_elems || (k > 0 && v[k].first < v[k - 1].second))
                    return "front pool placement overlaps (internal error)";
        }
    }
    p.relmap.assign(rm, -1);
    for (i64 c = 0; c < ns; ++c) {
        const NodeInfo &n = p.nodes[c];
        const i32 U = n.F - n.S;
        if (U == 0) continue;
        if (n.parent < 0) return "a root supernode has off-diagonal rows";
        const NodeInfo &par = p.nodes[n.parent];
        const i32 *sub = &p.lindx[n.lxoff + n.S];
        const i32 *sup = &p.lindx[par.lxoff];
        i32 q = 0;
        for (i32 i = 0; i < U; ++i) {
            while (q < par.F && sup[q] < sub[i]) ++q;
            if (q >= par.F || sup[q] != sub[i]) return "child row list is not contained in the parent's";
            p.relmap[n.mapoff + i] = q;
        }
    }

    // flop count (RINFO(3), sprtrs.f:1961-1969)
    p.factor_flops = 0;
    for (i64 j = 0; j < ns; ++j) {
        const double F = p.nodes[j].F, S = p.nodes[j].S;
        p.factor_flops += (F * F + 2 * F) * S - 2 * (1 + F) * S * (S + 1) / 2 +
                          S * (S + 1) * (2 * S + 1) / 6;
    }

    // ----------------------------------------------------------------------
    // schedule
    // ----------------------------------------------------------------------
    p.init_tasks.clear();
    p.copy_tasks.clear();
    for (i64 j = 0; j < ns; ++j)
        for (i32 c = 0; c < p.nodes[j].S; c += kInitCols)
            p.copy_tasks.push_back(ColTask{i32(j), c, std::min(p.nodes[j].S, c + kInitCols), 0});
    p.store_tasks.clear();
    p.ea_tasks.clear();
    p.small_nodes.clear();
    p.panel_desc.clear();
    p.panel_prefix.clear();
    p.gemm_desc.clear();
    p.gemm_prefix.clear();
    p.launches.clear();

    auto add_launch = [&](LaunchKind kind, int level, i64 off, i64 count, i64 blocks,
                          double flops = 0.0, i32 tile = 0, i32 outer = 0, i32 stream = 0) {
        if (count > 0 && blocks > 0)
            p.launches.push_back(
                Launch{kind, level, off, count, blocks, flops, tile, outer, stream});
    };
    auto add_sync = [&](LaunchKind kind, int level, i32 stream) {
        p.launches.push_back(Launch{kind, level, 0, 0, 0, 0.0, 0, 0, stream});
    };
    // useful flops of a GEMM descriptor: 2*kn per entry (i,j) of the lower
    // trapezoid rows [r0,rend) x cols [r0,cend), i >= j
    auto gemm_flops = [](const GemmDesc &g) {
        const double nc = double(g.cend - g.r0);
        const double pairs = nc * g.rend - (double(g.r0) + g.cend - 1) * nc / 2.0;
        return 2.0 * g.kn * pairs;
    };
    // Update of rows [r0,rend) x cols [r0,cend) of front n with its columns
    // [k0,k0+kn).  The tile grid is anchored at r0 rounded down to a multiple
    // of kGemmAlign so that every tile starts on a 16-byte boundary.
    auto make_gemm = [&](const NodeInfo &n, i32 r0, i32 rend, i32 cend, i32 k0, i32 kn,
                         i32 tile) {
        GemmDesc g;
        g.foff = n.foff;
        g.first = n.first;
        g.pad = 0;
        g.ld = n.ld;
        g.r0 = r0;
        g.rend = rend;
        g.cend = cend;
        g.k0 = k0;
        g.kn = kn;
        const i32 a0 = r0 / kGemmAlign * kGemmAlign;
        g.nti = (rend - a0 + tile - 1) / tile;
        g.ntj = (cend - a0 + tile - 1) / tile;
        return g;
    };

    bool store_pending = false;  // the previous level's panels are being stored
    for (int l = 0; l < p.nlevels; ++l) {
        const i32 *lv = &p.level_nodes[p.level_ptr[l]];
        const i64 nl = p.level_ptr[l + 1] - p.level_ptr[l];
        // Fronts with F <= small_f are assembled and factored in LDS by one
        // launch (k_small); the others are initialised, extend-added and
        // factored in blocks.
        std::vector<i32> small, big;
        for (i64 q = 0; q < nl; ++q)
            (p.nodes[lv[q]].F <= p.small_f ? small : big).push_back(lv[q]);
        // 1. front initialisation
        i64 off = i64(p.init_tasks.size());
        for (i32 j : big) {
            const NodeInfo &n = p.nodes[j];
            for (i32 c = 0; c < n.F; c += kInitCols)
                p.init_tasks.push_back(ColTask{j, c, std::min(n.F, c + kInitCols), 0});
        }
        add_launch(L_INIT, l, off, i64(p.init_tasks.size()) - off, i64(p.init_tasks.size()) - off);
        // 2. extend-add into the blocked fronts, one round per sibling index
        i32 maxch = 0;
        for (i32 j : big) maxch = std::max(maxch, p.nodes[j].nchild);
        for (i32 r = 0; r < maxch; ++r) {
            off = i64(p.ea_tasks.size());
            for (i32 j : big) {
                const NodeInfo &n = p.nodes[j];
                if (r >= n.nchild) continue;
                const i32 ch = p.children[n.child0 + r];
                const i32 U = p.nodes[ch].F - p.nodes[ch].S;
                for (i32 c = 0; c < U; c += kEaCols)
                    p.ea_tasks.push_back(EaTask{ch, c, std::min(U, c + kEaCols), 0});
            }
            add_launch(L_EA, l, off, i64(p.ea_tasks.size()) - off, i64(p.ea_tasks.size()) - off);
        }
        // 3. small fronts: assembly and factor in LDS (tile = largest F)
        off = i64(p.small_nodes.size());
        i32 maxf_small = 0;
        for (i32 j : small) {
            p.small_nodes.push_back(j);
            maxf_small = std::max(maxf_small, p.nodes[j].F);
        }
        add_launch(L_SMALL, l, off, i64(p.small_nodes.size()) - off,
                   i64(p.small_nodes.size()) - off, 0.0, maxf_small);
        // 4. blocked fronts: kPanelOuter-wide outer blocks, each factored by
        //    k_block_diag + k_block_trsm and followed by a rank-kPanelOuter
        //    update of the columns right of it; the update (U) block takes
        //    each kOuterRank-wide group's update with the group's far update,
        //    and a last partial group's at the end of the level.
        if (p.params.block_factor) {
            i32 maxblk = 0;
            for (i32 j : big) maxblk = std::max(maxblk, (p.nodes[j].S + kPanelOuter - 1) / kPanelOuter);
            for (i32 b = 0; b < maxblk; ++b) {
                const i32 ob = b * kPanelOuter, oend = ob + kPanelOuter;
                off = i64(p.panel_desc.size());
                i64 nstripes = 0;
                for (i32 j : big) {
                    const NodeInfo &n = p.nodes[j];
                    if (ob >= n.S) continue;
                    const i32 nbw = std::min(kPanelOuter, n.S - ob);
                    const i32 rows = kTrsmRows * kTrsmWaves;
                    const i32 ns = (n.F - (ob + nbw) + rows - 1) / rows;
                    p.panel_desc.push_back(PanelDesc{j, ob, nbw, ns});
                    p.panel_prefix.push_back(nstripes);
                    nstripes += ns;
                }
                const i64 cnt = i64(p.panel_desc.size()) - off;
                // One launch for the diagonal blocks and the rows below
                // them where the launch is short of work (few fronts: the
                // rows below then overlap the diagonal chains), two where
                // many fronts fill the device anyway (the separate kernels
                // keep more workgroups resident).
                if (p.params.block_factor == 1 && cnt <= kFusedMaxFronts) {
                    add_launch(L_BFUSED, l, off, cnt, cnt + nstripes);
                } else {
                    add_launch(L_BDIAG, l, off, cnt, cnt);
                    if (nstripes > 0) add_launch(L_BTRSM, l, off, cnt, nstripes);
                }
                // Updates by the blocks of a kOuterRank-wide group.  After an
                // earlier block of the group: a rank-kPanelOuter update of the
                // group's remaining columns.  After its last block: the
                // rank-kOuterRank update of the next group's columns ("near",
                // on the handle's stream) and of the columns beyond it,
                // through the U block ("far", on the second stream).  The far
                // update runs beside the next group's diagonal blocks, which
                // do not touch its columns; the next near update waits for
                // the part of it that covers the same columns, and nothing
                // else reads the U block before the level's end.
                const i32 g0 = ob / kOuterRank * kOuterRank, gend = g0 + kOuterRank;
                const bool last_in_group = oend >= gend;
                auto gemm_launch = [&](i32 stream, auto range) {
                    // the launch's size in 64-tiles decides its tile
                    i64 tiles64 = 0;
                    for (i32 j : big) {
                        i32 r0, cend;
                        if (!range(p.nodes[j], r0, cend)) continue;
                        const GemmDesc g = make_gemm(p.nodes[j], r0, p.nodes[j].F, cend, 0, 1,
                                                     kTileSmall);
                        tiles64 += tri_tiles(g.nti, g.ntj);
                    }
                    const i32 tile = last_in_group ? outer_tile(tiles64) : kTileSmall;
                    const i64 o = i64(p.gemm_desc.size());
                    i64 ntiles = 0;
                    double fl = 0.0;
                    for (i32 j : big) {
                        const NodeInfo &n = p.nodes[j];
                        i32 r0, cend;
                        if (!range(n, r0, cend)) continue;
                        // a group that runs past S ends at S
                        const i32 kn = last_in_group ? std::min(oend, n.S) - g0 : kPanelOuter;
                        const GemmDesc g = make_gemm(n, r0, n.F, cend, last_in_group ? g0 : ob, kn, tile);
                        p.gemm_desc.push_back(g);
                        p.gemm_prefix.push_back(ntiles);
                        ntiles += tri_tiles(g.nti, g.ntj);
                        fl += gemm_flops(g);
                    }
                    add_launch(L_GEMM, l, o, i64(p.gemm_desc.size()) - o, ntiles, fl, tile,
                               last_in_group ? 1 : 0, stream);
                };
                if (!last_in_group) {
                    gemm_launch(0, [&](const NodeInfo &n, i32 &r0, i32 &cend) {
                        r0 = oend;
                        cend = std::min(n.S, gend);
                        return oend < n.S;
                    });
                } else {
                    add_sync(L_RECORD, l, 0);  // the group's L columns are final
                    add_sync(L_WAIT, l, 0);    // the previous far update is done
                    gemm_launch(0, [&](const NodeInfo &n, i32 &r0, i32 &cend) {
                        r0 = oend;
                        cend = std::min(n.S, oend + kOuterRank);
                        return oend < n.S;
                    });
                    add_sync(L_WAIT, l, 1);
                    // far, in two launches: the group after next (which the
                    // next near update overwrites, so it waits for this part
                    // alone), then everything beyond it
                    gemm_launch(1, [&](const NodeInfo &n, i32 &r0, i32 &cend) {
                        r0 = oend + kOuterRank;
                        cend = std::min(n.S, oend + 2 * kOuterRank);
                        return r0 < n.S;
                    });
                    add_sync(L_RECORD, l, 1);
                    gemm_launch(1, [&](const NodeInfo &n, i32 &r0, i32 &cend) {
                        r0 = std::min(oend + 2 * kOuterRank, n.S);
                        cend = n.F;
                        return ob < n.S && r0 < n.F;
                    });
                }
            }
            add_sync(L_RECORD, l, 1);
            add_sync(L_WAIT, l, 0);  // every far update of the level is done
        }
        //    (or: panel steps with trailing updates inside the outer block)
        i32 maxsteps = 0;
        if (!p.params.block_factor)
            for (i32 j : big) maxsteps = std::max(maxsteps, (p.nodes[j].S + kPanelNB - 1) / kPanelNB);
        for (i32 s = 0; s < maxsteps; ++s) {
            off = i64(p.panel_desc.size());
            i64 nblk = 0;
            for (i32 j : big) {
                const NodeInfo &n = p.nodes[j];
                const i32 k0 = s * kPanelNB;
                if (k0 >= n.S) continue;
                const i32 nb = std::min(kPanelNB, n.S - k0);
                const i32 rows = n.F - (k0 + nb);
                const i32 nrb = std::max(1, (rows + kPanelRows - 1) / kPanelRows);
                p.panel_desc.push_back(PanelDesc{j, k0, nb, nrb});
                p.panel_prefix.push_back(nblk);
                nblk += nrb;
            }
            add_launch(L_PANEL, l, off, i64(p.panel_desc.size()) - off, nblk);
            // Two-level blocking of the panel: the rank-nb update of a step
            // stops at the end of its kPanelOuter-wide outer block; the
            // columns beyond get one rank-kPanelOuter update per outer block.
            const i32 k0 = s * kPanelNB;
            const i32 oend = (k0 / kPanelOuter + 1) * kPanelOuter;
            off = i64(p.gemm_desc.size());
            i64 ntiles = 0;
            double fl = 0.0;
            for (i32 j : big) {
                const NodeInfo &n = p.nodes[j];
                if (k0 >= n.S) continue;
                const i32 nb = std::min(kPanelNB, n.S - k0);
                const i32 r0 = k0 + nb;
                const i32 cend = std::min(n.S, oend);
                if (r0 >= cend) continue;
                const GemmDesc g = make_gemm(n, r0, n.F, cend, k0, nb, kTileSmall);
                p.gemm_desc.push_back(g);
                p.gemm_prefix.push_back(ntiles);
                ntiles += tri_tiles(g.nti, g.ntj);
                fl += gemm_flops(g);
            }
            add_launch(L_GEMM, l, off, i64(p.gemm_desc.size()) - off, ntiles, fl, kTileSmall);
            if (k0 + kPanelNB == oend) {
                const i32 ob = oend - kPanelOuter;
                off = i64(p.gemm_desc.size());
                ntiles = 0;
                fl = 0.0;
                const i32 ptile = p.big_tile ? p.big_tile : kTileSmall;
                for (i32 j : big) {
                    const NodeInfo &n = p.nodes[j];
                    if (oend >= n.S) continue;
                    const GemmDesc g = make_gemm(n, oend, n.F, n.S, ob, kPanelOuter, ptile);
                    p.gemm_desc.push_back(g);
                    p.gemm_prefix.push_back(ntiles);
                    ntiles += tri_tiles(g.nti, g.ntj);
                    fl += gemm_flops(g);
                }
                add_launch(L_GEMM, l, off, i64(p.gemm_desc.size()) - off, ntiles, fl, ptile, 1);
            }
        }
        // Schur complement of every blocked front of this level: with the
        // panel steps in one update with K = S; with outer blocks only the
        // columns of a last group that ended before its kOuterRank-th column
        // (the complete groups' far updates took the rest)
        off = i64(p.gemm_desc.size());
        i64 ntiles = 0;
        double fl = 0.0;
        auto schur_k0 = [&](const NodeInfo &n, i32 &k0) {
            k0 = 0;
            if (n.F == n.S) return false;
            if (p.params.block_factor) {
                const i32 ob = (n.S - 1) / kPanelOuter * kPanelOuter;
                const i32 g0 = ob / kOuterRank * kOuterRank;
                if (ob + kPanelOuter >= g0 + kOuterRank) return false;  // ended with a group update
                k0 = g0;
            }
            return true;
        };
        i64 stiles64 = 0;
        for (i32 j : big) {
            i32 k0;
            if (!schur_k0(p.nodes[j], k0)) continue;
            const NodeInfo &n = p.nodes[j];
            const GemmDesc g = make_gemm(n, n.S, n.F, n.F, k0, 1, kTileSmall);
            stiles64 += tri_tiles(g.nti, g.ntj);
        }
        const i32 stile = outer_tile(stiles64);
        for (i32 j : big) {
            const NodeInfo &n = p.nodes[j];
            i32 k0;
            if (!schur_k0(n, k0)) continue;
            const GemmDesc g = make_gemm(n, n.S, n.F, n.F, k0, n.S - k0, stile);
            p.gemm_desc.push_back(g);
            p.gemm_prefix.push_back(ntiles);
            ntiles += tri_tiles(g.nti, g.ntj);
            fl += gemm_flops(g);
        }
        add_launch(L_GEMM, l, off, i64(p.gemm_desc.size()) - off, ntiles, fl, stile, 1);
        // 5. with panel steps, the blocked fronts' factored panels to the
        //    store (the small fronts' and the outer blocks' kernels write
        //    theirs there as they make them), on the second stream beside
        //    the next level; the handle's stream waits for it at the end of
        //    the next level, before any front memory of this level is
        //    reused (two levels on at the earliest)
        off = i64(p.store_tasks.size());
        if (!p.params.block_factor)
            for (i32 j : big)
                for (i32 c = 0; c < p.nodes[j].S; c += kInitCols)
                    p.store_tasks.push_back(ColTask{j, c, std::min(p.nodes[j].S, c + kInitCols), 0});
        if (store_pending) add_sync(L_WAIT, l, 0);
        store_pending = i64(p.store_tasks.size()) > off;
        if (store_pending) {
            add_sync(L_RECORD, l, 0);
            add_sync(L_WAIT, l, 1);
            add_launch(L_STORE, l, off, i64(p.store_tasks.size()) - off,
                       i64(p.store_tasks.size()) - off, 0.0, 0, 0, 1);
            add_sync(L_RECORD, l, 1);
        }
    }

    // ----------------------------------------------------------------------
    // solve schedule
    // ----------------------------------------------------------------------
    p.solve_small.clear();
    p.wide.clear();
    p.wide_fwd_prefix.clear();
    p.wide_bwd_prefix.clear();
    p.wide_rect_prefix.clear();
    p.wide_frect_prefix.clear();
    p.solve_small_ptr.assign(p.nlevels + 1, 0);
    p.small_maxF.assign(p.nlevels, 0);
    p.small_maxS.assign(p.nlevels, 0);
    p.wide_ptr.assign(p.nlevels + 1, 0);
    p.wide_fwd_wgs.assign(p.nlevels, 0);
    p.wide_bwd_wgs.assign(p.nlevels, 0);
    p.wide_rect_wgs.assign(p.nlevels, 0);
    p.wide_frect_wgs.assign(p.nlevels, 0);
    p.nflags = 0;
    p.npart = 0;
    for (int l = 0; l < p.nlevels; ++l) {
        i64 wf = 0, wb = 0, wr = 0, wfr = 0;
        for (i32 q = p.level_ptr[l]; q < p.level_ptr[l + 1]; ++q) {
            const i32 j = p.level_nodes[q];
            const NodeInfo &n = p.nodes[j];
            if (n.S <= kSolveSmallS && n.F <= kSolveSmallF) {
                p.solve_small.push_back(j);
                p.small_maxF[l] = std::max(p.small_maxF[l], n.F);
                p.small_maxS[l] = std::max(p.small_maxS[l], n.S);
                continue;
            }
            WideDesc w;
            w.node = j;
            w.nblk = (n.S + kSolveNB - 1) / kSolveNB;
            w.nrect = (n.F - n.S + kSolveNB - 1) / kSolveNB;
            w.pad = 0;
            w.flag0 = p.nflags;
            w.pbase = p.npart;
            p.wide.push_back(w);
            p.wide_fwd_prefix.push_back(wf);
            p.wide_bwd_prefix.push_back(wb);
            p.wide_rect_prefix.push_back(wr);
            p.wide_frect_prefix.push_back(wfr);
            wf += w.nblk + w.nrect;
            wb += w.nblk;
            wr += i64(w.nblk) * w.nrect;
            wfr += w.nrect;
            p.nflags += w.nblk;
            p.npart += i64(w.nblk) * w.nrect * kSolveNB;
        }
        p.solve_small_ptr[l + 1] = i32(p.solve_small.size());
        p.wide_ptr[l + 1] = i32(p.wide.size());
        p.wide_fwd_wgs[l] = wf;
        p.wide_bwd_wgs[l] = wb;
        p.wide_rect_wgs[l] = wr;
        p.wide_frect_wgs[l] = wfr;
    }
    // gather form of the forward assembly (inverse of the child maps)
    if (p.vwork_elems >= (i64(1) << 31)) return "solve workspace too large for 32-bit indices";
    p.upd_ptr.assign(p.vwork_elems + 1, 0);
    for (i64 c = 0; c < ns; ++c) {
        const NodeInfo &n = p.nodes[c];
        if (n.parent < 0) continue;
        const i64 pv = p.nodes[n.parent].voff;
        for (i32 i = 0; i < n.F - n.S; ++i) p.upd_ptr[pv + p.relmap[n.mapoff + i] + 1]++;
    }
    for (i64 i = 0; i < p.vwork_elems; ++i) p.upd_ptr[i + 1] += p.upd_ptr[i];
    p.upd_src.assign(p.upd_ptr[p.vwork_elems], 0);
    {
        std::vector<i64> fill(p.upd_ptr.begin(), p.upd_ptr.end() - 1);
        for (i64 j = 0; j < ns; ++j) {  // parents in order, children in SIBL order
            const NodeInfo &par = p.nodes[j];
            for (i32 k = 0; k < par.nchild; ++k) {
                const NodeInfo &n = p.nodes[p.children[par.child0 + k]];
                for (i32 i = 0; i < n.F - n.S; ++i)
                    p.upd_src[fill[par.voff + p.relmap[n.mapoff + i]]++] = i32(n.voff + n.S + i);
            }
        }
    }
    // forest schedule: every update block packed at its own offset; the
    // ascending pass takes the supernodes by depth in the tree, deepest
    // first (a child is deeper than its parent), so that the longest chains
    // to a root start first; the backward pass by height, highest first (a
    // parent is higher than its child), so that the longest chains to a leaf
    // start first.  Items of one depth (height) are independent; each is
    // split into workgroup chunks.
    p.forest = params.forest != 0 && p.mfront <= kSmallF;
    for (i64 j = 0; p.forest && j < ns; ++j) p.forest = p.nodes[j].F <= kSmallF;
    p.forest_nodes.clear();
    p.forest_child.clear();
    p.forest_order.clear();
    p.forest_chunks.clear();
    p.forest_nasc = 0;
    p.ublk_elems = 0;
    if (p.forest) {
        std::vector<i64> uoff(ns);
        for (i64 j = 0; j < ns; ++j) {
            const i64 U = p.nodes[j].F - p.nodes[j].S;
            uoff[j] = p.ublk_elems;
            p.ublk_elems += round_up(U * (U + 1) / 2, 16);
        }
        p.forest_nodes.resize(ns);
        for (i64 j = 0; j < ns; ++j) {
            const NodeInfo &n = p.nodes[j];
            ForestNode &f = p.forest_nodes[j];
            f.xlnz = n.xlnz;
            f.loff = n.loff;
            f.uoff = uoff[j];
            f.voff = n.voff;
            f.lxoff = n.lxoff;
            f.first = n.first;
            f.S = n.S;
            f.F = n.F;
            f.ld = n.ld;
            f.parent = n.parent;
            f.child0 = i32(p.forest_child.size());
            f.nchild = n.nchild;
            f.pad = 0;
            for (i32 c = 0; c < n.nchild; ++c) {
                const i32 cj = p.children[n.child0 + c];
                const NodeInfo &cn = p.nodes[cj];
                ForestChild fc;
                fc.uoff = uoff[cj];
                fc.vofs = cn.voff + cn.S;
                fc.mapoff = cn.mapoff;
                fc.U = cn.F - cn.S;
                fc.pad = 0;
                p.forest_child.push_back(fc);
            }
        }
        std::vector<i32> depth(ns, 0);
        for (i64 j = ns - 1; j >= 0; --j)  // parents come after their children
            depth[j] = parent[j] >= 0 ? depth[parent[j]] + 1 : 0;
        const i32 maxd = *std::max_element(depth.begin(), depth.end());
        // chunks of a group of independent supernodes [a0, a1) of forest_order
        auto chunk = [&](i64 a0, i64 a1) {
            const i64 per = std::max<i64>(
                kForestWaves,
                round_up((a1 - a0 + kForestWgsPerLevel - 1) / kForestWgsPerLevel, kForestWaves));
            for (i64 k = a0; k < a1; k += per) p.forest_chunks.push_back(i32(k));
        };
        std::vector<std::vector<i32>> bydepth(maxd + 1);
        for (i64 j = 0; j < ns; ++j) bydepth[depth[j]].push_back(i32(j));
        for (i32 dd = maxd; dd >= 0; --dd) {
            const i64 a0 = i64(p.forest_order.size());
            p.forest_order.insert(p.forest_order.end(), bydepth[dd].begin(), bydepth[dd].end());
            chunk(a0, i64(p.forest_order.size()));
        }
        p.forest_nasc = i32(p.forest_chunks.size());
        for (int l = p.nlevels - 1; l >= 0; --l) {
            const i64 a0 = i64(p.forest_order.size());
            p.forest_order.insert(p.forest_order.end(), p.level_nodes.begin() + p.level_ptr[l],
                                  p.level_nodes.begin() + p.level_ptr[l + 1]);
            chunk(a0, i64(p.forest_order.size()));
        }
        p.forest_chunks.push_back(i32(p.forest_order.size()));
        // every item waits only for items of earlier chunks (its workgroup's
        // predecessors): a child's ascending item, a parent's backward item
        std::vector<i32> achunk(ns, -1), bchunk(ns, -1);
        for (size_t q = 0; q + 1 < p.forest_chunks.size(); ++q)
            for (i32 k = p.forest_chunks[q]; k < p.forest_chunks[q + 1]; ++k)
                (i32(q) < p.forest_nasc ? achunk : bchunk)[p.forest_order[k]] = i32(q);
        for (i64 j = 0; j < ns; ++j) {
            if (achunk[j] < 0 || bchunk[j] < 0) return "forest order misses a supernode (internal error)";
            if (parent[j] >= 0 && (achunk[parent[j]] <= achunk[j] || bchunk[parent[j]] >= bchunk[j]))
                return "forest order breaks a dependency (internal error)";
        }
    }
    return "";
}

// CSR pattern of the assembled matrix from the non-zeros of the packed LNZ.
// Rows are sorted by column: columns are visited in ascending order, each
// contributing its sub-diagonal entries to later rows (their lower parts)
// and, in its own row, the diagonal and the upper part.  Threads own
// contiguous column ranges, so per-thread counts keep the order.
// keep(pos) selects the packed positions that enter the pattern; they are
// visited in increasing order within each thread's range of supernodes
template <class Keep>
static void csr_pattern_impl(const Plan &p, Keep keep, CsrPattern &pat)
{
    const i64 neq = p.neq, ns = p.nsuper;
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    // split supernodes into nt ranges of similar LNZ size
    std::vector<i64> cut(nt + 1, ns);
    cut[0] = 0;
    {
        const i64 total = p.nzerol;
        i64 acc = 0;
        unsigned t = 1;
        for (i64 j = 0; j < ns && t < nt; ++j) {
            const i64 F = p.nodes[j].F, S = p.nodes[j].S;
            acc += F * S - S * (S - 1) / 2;
            if (acc >= total * i64(t) / i64(nt)) cut[t++] = j + 1;
        }
        for (; t < nt; ++t) cut[t] = ns;
    }
    std::vector<std::vector<i32>> cnt(nt, std::vector<i32>(neq, 0));
    std::vector<i64> nzl(nt, 0);
    auto count = [&](unsigned t) {
        auto &c = cnt[t];
        for (i64 j = cut[t]; j < cut[t + 1]; ++j) {
            const NodeInfo &n = p.nodes[j];
            const i32 *rows = &p.lindx[n.lxoff];
            i64 pos = n.xlnz;
            for (i32 col = 0; col < n.S; ++col) {
                const i32 ceq = n.first + col;
                for (i32 r = col; r < n.F; ++r, ++pos) {
                    if (!keep(pos)) continue;
                    nzl[t]++;
                    c[ceq]++;
                    if (r != col) c[rows[r]]++;
                }
            }
        }
    };
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) th.emplace_back(count, t);
        for (auto &x : th) x.join();
    }
    pat.rowptr.assign(neq + 1, 0);
    for (i64 i = 0; i < neq; ++i) {
        i64 s = 0;
        for (unsigned t = 0; t < nt; ++t) s += cnt[t][i];
        pat.rowptr[i + 1] = pat.rowptr[i] + s;
    }
    pat.nnz_lower = 0;
    for (unsigned t = 0; t < nt; ++t) pat.nnz_lower += nzl[t];
    const i64 nnz = pat.rowptr[neq];
    pat.col.assign(nnz, 0);
    pat.src.assign(nnz, 0);
    // per-thread start cursor in each row = rowptr + counts of earlier threads
    std::vector<std::vector<i64>> cur(nt);
    {
        std::vector<i64> acc(pat.rowptr.begin(), pat.rowptr.end() - 1);
        for (unsigned t = 0; t < nt; ++t) {
            cur[t].assign(acc.begin(), acc.end());
            for (i64 i = 0; i < neq; ++i) acc[i] += cnt[t][i];
        }
    }
    auto fill = [&](unsigned t) {
        auto &cu = cur[t];
        for (i64 j = cut[t]; j < cut[t + 1]; ++j) {
            const NodeInfo &n = p.nodes[j];
            const i32 *rows = &p.lindx[n.lxoff];
            i64 pos = n.xlnz;
            for (i32 col = 0; col < n.S; ++col) {
                const i32 ceq = n.first + col;
                for (i32 r = col; r < n.F; ++r, ++pos) {
                    if (!keep(pos)) continue;
                    // row ceq, column rows[r] (upper part incl. diagonal)
                    i64 k = cu[ceq]++;
                    pat.col[k] = rows[r];
                    pat.src[k] = pos;
                    if (r != col) {
                        // row rows[r], column ceq (lower part)
                        k = cu[rows[r]]++;
                        pat.col[k] = ceq;
                        pat.src[k] = pos;
                    }
                }
            }
        }
    };
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) th.emplace_back(fill, t);
        for (auto &x : th) x.join();
    }
}

void build_csr_pattern(const Plan &p, const double *lnz, CsrPattern &pat)
{
    csr_pattern_impl(p, [lnz](i64 pos) { return lnz[pos] != 0.0; }, pat);
}

void build_csr_pattern_marked(const Plan &p, const std::vector<uint64_t> &marked, CsrPattern &pat)
{
    csr_pattern_impl(
        p, [&marked](i64 pos) { return ((marked[size_t(pos) >> 6] >> (pos & 63)) & 1u) != 0; },
        pat);
}

}  // namespace sprgpu
