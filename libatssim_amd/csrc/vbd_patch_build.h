// vbd_patch_build.h -- host-side construction of the patch layout (see
// vbd_patch.h) from a cloth's colour-major single-cloth layout.  Host code
// only; the library builds it from the device arrays of a context, the CPU
// test program from a layout built on the host.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "vbd_kernels.h"
#include "vbd_patch.h"

namespace lats {

// The single-cloth layout the colour passes read (slot order).
struct PatchSource {
    uint32_t V = 0, C = 0;
    const ColourInfo* colours = nullptr;   // C entries
    const uint32_t* meta = nullptr;        // edge degree | bending degree << 8 | pinned << 16
    const float* x = nullptr;              // 3 per slot: positions the tiling follows
    const uint2* ell_e = nullptr;
    const float4* ell_bq = nullptr;
    const PairIds* ell_bi = nullptr;
};

// The patch layout, host copies of the arrays the patch launches read.
struct PatchHost {
    std::vector<PatchHead> heads;     // group g's P heads at [g * P, (g + 1) * P)
    std::vector<PatchStep> steps;     // patch q of group g: F records at (g * P + q) * F
    std::vector<uint32_t> region;     // patch q of group g: group_roff[g] + q * group_rstride[g]
    std::vector<uint2> entries, ell_e;
    std::vector<float4> ell_bq;
    std::vector<PatchPair> ell_bi;
    std::vector<uint32_t> group_lds;  // per group: most LDS bytes of one workgroup
    std::vector<uint32_t> group_c0;   // per group: first colour; group g runs colours [c0[g], c0[g + 1])
    // per group: where its regions start and the fixed stride between two
    // patches' regions (its largest region; shorter ones are padded with
    // their first slot), so a launch computes every region's address and
    // every step's from the patch number alone
    std::vector<uint32_t> group_roff, group_rstride;
    std::vector<uint32_t> group_nsmax;   // per group: most slots with frame state in one region
    PatchStats stats;
};

// Tile grid px x py of P tiles whose aspect follows the extents e1 >= e2.
inline void patch_tile_grid(uint32_t P, double e1, double e2, uint32_t& px, uint32_t& py) {
    const double want = e2 > 0 ? e1 / e2 : 1e9;
    px = P;
    py = 1;
    double best = 1e300;
    for (uint32_t d = 1; d <= P; d++) {
        if (P % d) continue;
        const double r = (double)d / (double)(P / d);
        const double cost = std::fabs(std::log(r / want));
        if (cost < best) {
            best = cost;
            px = d;
            py = P / d;
        }
    }
}

// Builds the layout for P patches (a multiple of 8, at most V) and F colours
// per launch (0 or more than C: C; at most kMaxPatchSteps).  A workgroup
// needing more than max_lds bytes of LDS fails.  On failure err says why.
inline bool patch_build_host(const PatchSource& src, uint32_t P, uint32_t F, uint32_t max_lds, PatchHost& out,
                             std::string& err) {
    const uint32_t V = src.V, C = src.C;
    const uint32_t* meta = src.meta;
    if (P == 0 || P % kXcds) { err = "patches must be a positive multiple of 8"; return false; }
    if (P > V) { err = "more patches than vertices"; return false; }
    if (F == 0 || F > C) F = C;
    if (F > kMaxPatchSteps) F = kMaxPatchSteps;
    out = PatchHost{};

    // colour of each slot and the slot-space adjacency the colouring respects
    // (spring neighbours, the other three vertices of each bending pair)
    std::vector<uint32_t> col(V), nb_off(V + 1, 0);
    for (uint32_t k = 0; k < C; k++) {
        const ColourInfo& ci = src.colours[k];
        for (uint32_t i = 0; i < ci.count; i++) {
            col[ci.start + i] = k;
            nb_off[ci.start + i + 1] = (meta[ci.start + i] & 0xffu) + 3u * ((meta[ci.start + i] >> 8) & 0xffu);
        }
    }
    for (uint32_t s = 0; s < V; s++) nb_off[s + 1] += nb_off[s];
    std::vector<uint32_t> nb(nb_off[V]);
    for (uint32_t k = 0; k < C; k++) {
        const ColourInfo& ci = src.colours[k];
        for (uint32_t i = 0; i < ci.count; i++) {
            const uint32_t s = ci.start + i, ed = meta[s] & 0xffu, bd = (meta[s] >> 8) & 0xffu;
            uint32_t* o = &nb[nb_off[s]];
            for (uint32_t j = 0; j < ed; j++) *o++ = src.ell_e[ci.eoff + (size_t)j * ci.count + i].x;
            for (uint32_t j = 0; j < bd; j++) {
                const PairIds& q = src.ell_bi[ci.boff + (size_t)j * ci.count + i];
                *o++ = q.a & 0x3fffffffu;
                *o++ = q.b;
                *o++ = q.c;
            }
        }
    }

    // patches: tiles of the shape over its two longest extents, rows of equal
    // slot count, then columns of equal count inside a row; patch = row * px +
    // column, so consecutive patches are neighbours
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t s = 0; s < V; s++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], (double)src.x[3 * (size_t)s + d]);
            hi[d] = std::max(hi[d], (double)src.x[3 * (size_t)s + d]);
        }
    int ax[3] = {0, 1, 2};
    std::stable_sort(ax, ax + 3, [&](int p0, int p1) { return hi[p0] - lo[p0] > hi[p1] - lo[p1]; });
    const int a1 = ax[0], a2 = ax[1];
    uint32_t px = 0, py = 0;
    patch_tile_grid(P, hi[a1] - lo[a1], hi[a2] - lo[a2], px, py);
    std::vector<uint32_t> order(V);
    std::iota(order.begin(), order.end(), 0u);
    auto coord = [&](uint32_t s, int d) { return src.x[3 * (size_t)s + d]; };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t u, uint32_t w) { return coord(u, a2) < coord(w, a2); });
    std::vector<uint32_t> owner(V);
    for (uint32_t ry = 0; ry < py; ry++) {
        const size_t r0 = (size_t)V * ry / py, r1 = (size_t)V * (ry + 1) / py;
        std::stable_sort(order.begin() + r0, order.begin() + r1,
                         [&](uint32_t u, uint32_t w) { return coord(u, a1) < coord(w, a1); });
        for (uint32_t cx = 0; cx < px; cx++) {
            const size_t c0 = r0 + (r1 - r0) * cx / px, c1 = r0 + (r1 - r0) * (cx + 1) / px;
            for (size_t q = c0; q < c1; q++) owner[order[q]] = ry * px + cx;
        }
    }
    std::vector<uint32_t> own_off(P + 1, 0), own(V);
    for (uint32_t s = 0; s < V; s++) own_off[owner[s] + 1]++;
    for (uint32_t q = 0; q < P; q++) own_off[q + 1] += own_off[q];
    {
        std::vector<uint32_t> fill(own_off.begin(), own_off.end() - 1);
        for (uint32_t s = 0; s < V; s++) own[fill[owner[s]]++] = s;   // ascending slots per patch
    }

    // groups of consecutive colours: ceil(C / F) launches per sweep, sizes
    // as equal as possible
    const uint32_t ngroups = (C + F - 1) / F;
    out.group_c0.resize(ngroups + 1);
    for (uint32_t g = 0; g <= ngroups; g++) out.group_c0[g] = (uint32_t)((uint64_t)C * g / ngroups);
    out.group_lds.assign(ngroups, 0);
    PatchStats& stats = out.stats;
    stats.patches = P;
    stats.colours_per_launch = F;
    stats.launches_per_sweep = ngroups;

    out.group_roff.assign(ngroups, 0);
    out.group_rstride.assign(ngroups, 0);
    out.group_nsmax.assign(ngroups, 0);
    std::vector<uint32_t> mark_need(V, 0), mark_reg(V, 0), loc(V, 0);
    uint32_t token = 0;
    std::vector<std::vector<uint32_t>> need(kMaxPatchSteps), regs(P);
    std::vector<uint32_t> extra;
    for (uint32_t g = 0; g < ngroups; g++) {
        const uint32_t c0 = out.group_c0[g], nst = out.group_c0[g + 1] - c0;
        const size_t head0 = out.heads.size();
        auto step_of = [&](uint32_t s) -> int {
            const uint32_t k = col[s];
            return (k >= c0 && k < c0 + nst) ? (int)(k - c0) : -1;
        };
        for (uint32_t q = 0; q < P; q++) {
            token++;
            for (uint32_t j = 0; j < nst; j++) need[j].clear();
            for (uint32_t r = own_off[q]; r < own_off[q + 1]; r++) {
                const uint32_t s = own[r];
                const int j = step_of(s);
                if (j >= 0) {
                    need[j].push_back(s);
                    mark_need[s] = token;
                }
            }
            // backwards: every vertex that a later step reads and an earlier
            // step of the group updates (need[jj] is complete when reached)
            for (int jj = (int)nst - 1; jj >= 1; jj--)
                for (size_t w = 0; w < need[jj].size(); w++) {
                    const uint32_t s = need[jj][w];
                    for (uint32_t r = nb_off[s]; r < nb_off[s + 1]; r++) {
                        const uint32_t u = nb[r];
                        const int j = step_of(u);
                        if (j >= 0 && j < jj && mark_need[u] != token) {
                            mark_need[u] = token;
                            need[j].push_back(u);
                        }
                    }
                }
            for (uint32_t j = 0; j < nst; j++) std::sort(need[j].begin(), need[j].end());
            // region: owned slots first, then every other slot read, ascending
            std::vector<uint32_t>& reg = regs[q];
            reg.clear();
            PatchHead hd{};
            hd.owned = own_off[q + 1] - own_off[q];
            hd.nsteps = nst;
            hd.step_off = (uint32_t)out.steps.size();
            for (uint32_t r = own_off[q]; r < own_off[q + 1]; r++) {
                mark_reg[own[r]] = token;
                loc[own[r]] = (uint32_t)reg.size();
                reg.push_back(own[r]);
            }
            // then the halo the patch updates (frame state staged too), then
            // the slots only read
            extra.clear();
            for (uint32_t j = 0; j < nst; j++)
                for (uint32_t s : need[j])
                    if (mark_reg[s] != token) {
                        mark_reg[s] = token;
                        extra.push_back(s);
                    }
            std::sort(extra.begin(), extra.end());
            for (uint32_t s : extra) {
                loc[s] = (uint32_t)reg.size();
                reg.push_back(s);
            }
            hd.nstate = (uint32_t)reg.size();
            extra.clear();
            for (uint32_t j = 0; j < nst; j++)
                for (uint32_t s : need[j])
                    for (uint32_t r = nb_off[s]; r < nb_off[s + 1]; r++)
                        if (mark_reg[nb[r]] != token) {
                            mark_reg[nb[r]] = token;
                            extra.push_back(nb[r]);
                        }
            std::sort(extra.begin(), extra.end());
            for (uint32_t s : extra) {
                loc[s] = (uint32_t)reg.size();
                reg.push_back(s);
            }
            hd.rsize = (uint32_t)reg.size();
            if (hd.rsize > 65536u) {   // (local pair entries hold 16-bit LDS indices)
                err = "patch strategy: a region of " + std::to_string(hd.rsize) + " slots, above 65536";
                return false;
            }
            if (patch_lds_bytes(hd.rsize, hd.nstate, hd.owned) > max_lds) {
                err = "patch strategy: a region of " + std::to_string(hd.rsize) + " slots (" +
                      std::to_string(hd.nstate) + " updated) needs " +
                      std::to_string(patch_lds_bytes(hd.rsize, hd.nstate, hd.owned)) + " bytes of LDS, more than one workgroup's " +
                      std::to_string(max_lds) + "; use more patches or fewer colours per launch";
                return false;
            }
            out.group_lds[g] = std::max(out.group_lds[g], patch_lds_bytes(hd.rsize, hd.nstate, hd.owned));
            out.group_nsmax[g] = std::max(out.group_nsmax[g], hd.nstate);
            stats.max_region = std::max(stats.max_region, hd.rsize);
            stats.region_slots += hd.rsize;
            // steps: update lists and local ELLs (entries hold LDS indices;
            // padding entries hold index 0, a valid slot of every region)
            for (uint32_t j = 0; j < nst; j++) {
                PatchStep ps{};
                ps.off = (uint32_t)out.entries.size();
                ps.n = (uint32_t)need[j].size();
                uint32_t ew = 0, bw = 0;
                for (uint32_t s : need[j]) {
                    ew = std::max(ew, meta[s] & 0xffu);
                    bw = std::max(bw, (meta[s] >> 8) & 0xffu);
                }
                ps.ew = ew;
                ps.bw = bw;
                ps.eoff = (uint32_t)out.ell_e.size();
                ps.boff = (uint32_t)out.ell_bq.size();
                if ((uint64_t)out.ell_e.size() + (uint64_t)ew * ps.n >= (1ull << 32) ||
                    (uint64_t)out.ell_bq.size() + (uint64_t)bw * ps.n >= (1ull << 32) ||
                    (uint64_t)out.entries.size() + ps.n >= (1ull << 32)) {
                    err = "patch strategy: local adjacency above 2^32 entries";
                    return false;
                }
                out.ell_e.resize(out.ell_e.size() + (size_t)ew * ps.n, uint2{0u, 0u});
                out.ell_bq.resize(out.ell_bq.size() + (size_t)bw * ps.n, float4{0.f, 0.f, 0.f, 0.f});
                out.ell_bi.resize(out.ell_bi.size() + (size_t)bw * ps.n, PatchPair{0u, 0u, 0u});
                for (uint32_t i = 0; i < ps.n; i++) {
                    const uint32_t s = need[j][i];
                    out.entries.push_back(uint2{s, loc[s]});
                    const ColourInfo& ci = src.colours[col[s]];
                    const uint32_t ig = s - ci.start, ed = meta[s] & 0xffu, bd = (meta[s] >> 8) & 0xffu;
                    for (uint32_t row = 0; row < ed; row++) {
                        const uint2 e = src.ell_e[ci.eoff + (size_t)row * ci.count + ig];
                        out.ell_e[ps.eoff + (size_t)row * ps.n + i] = uint2{loc[e.x], e.y};
                    }
                    for (uint32_t row = 0; row < bd; row++) {
                        const size_t at = ci.boff + (size_t)row * ci.count + ig;
                        const PairIds& id = src.ell_bi[at];
                        const size_t to = ps.boff + (size_t)row * ps.n + i;
                        const float4 q = src.ell_bq[at];
                        out.ell_bq[to] = q;
                        // the pair's four points in pair order, the vertex at o
                        const uint32_t o = id.a >> 30;
                        const uint32_t g0 = loc[id.a & 0x3fffffffu], g1 = loc[id.b], g2 = loc[id.c], me = loc[s];
                        const uint32_t pt[4] = {o == 0 ? me : g0, o == 1 ? me : (o == 0 ? g0 : g1),
                                                o == 2 ? me : (o <= 1 ? g1 : g2), o == 3 ? me : g2};
                        const float qo[4] = {q.x, q.y, q.z, q.w};
                        uint32_t qoo;
                        std::memcpy(&qoo, &qo[o], 4);
                        out.ell_bi[to] = PatchPair{pt[0] | (pt[1] << 16), pt[2] | (pt[3] << 16), qoo};
                    }
                }
                stats.updates_per_sweep += ps.n;
                stats.ell_entries += (uint64_t)ps.n * (ew + bw);
                out.steps.push_back(ps);
            }
            // F records per patch: a group of fewer colours ends in empty
            // steps, which no launch reaches (it runs the group's nst steps)
            out.steps.resize(out.steps.size() + (F - nst), PatchStep{0u, 0u, 0u, 0u, 0u, 0u});
            out.heads.push_back(hd);
        }
        // the group's regions at a fixed stride, padded with their first slot
        uint32_t rs = 1;
        for (uint32_t q = 0; q < P; q++) rs = std::max(rs, (uint32_t)regs[q].size());
        if ((uint64_t)out.region.size() + (uint64_t)P * rs >= (1ull << 32)) {
            err = "patch strategy: regions above 2^32 slots";
            return false;
        }
        out.group_roff[g] = (uint32_t)out.region.size();
        out.group_rstride[g] = rs;
        for (uint32_t q = 0; q < P; q++) {
            PatchHead& hd = out.heads[head0 + q];
            hd.roff = out.group_roff[g] + q * rs;
            out.region.insert(out.region.end(), regs[q].begin(), regs[q].end());
            out.region.resize((size_t)hd.roff + rs, regs[q].empty() ? 0u : regs[q][0]);
        }
    }
    // (a zero-width ELL reads entry 0: keep every array non-empty)
    if (out.ell_e.empty()) out.ell_e.push_back(uint2{0u, 0u});
    if (out.ell_bq.empty()) {
        out.ell_bq.push_back(float4{0.f, 0.f, 0.f, 0.f});
        out.ell_bi.push_back(PatchPair{0u, 0u, 0u});
    }
    // a sentinel entry: every step's first entry is readable, empty steps too
    out.entries.push_back(uint2{0u, 0u});
    return true;
}

}  // namespace lats
