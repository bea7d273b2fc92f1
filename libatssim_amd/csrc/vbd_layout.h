// vbd_layout.h -- host-side construction of the HBM layout of one or more
// cloths (see the layout note at the top of vbd_device.hip).  A cloth's
// vertices are appended in its colour order; every index stored on the
// device is a global slot, so several cloths can share one set of arrays.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "vbd_kernels.h"

namespace lats {

struct ClothLayout {
    std::vector<float> x;              // 3 per slot
    std::vector<float> mass;
    std::vector<uint32_t> meta;        // edge deg | bending deg << 8 | pinned << 16
    std::vector<uint32_t> perm;        // slot -> vertex id inside its cloth
    std::vector<ColourInfo> colours;   // global slots and ELL bases
    std::vector<uint2> ell_e;
    std::vector<float4> ell_bq;
    std::vector<PairIds> ell_bi;
    std::vector<uint2> e_slots;        // per edge (energy)
    std::vector<float> rest;
    std::vector<uint4> b_slots;        // per bending pair (energy)
    std::vector<float4> Q;             // 4 per pair (energy)
    size_t num_slots() const { return mass.size(); }
};

inline bool validate_mesh(const lats_mesh_desc* m, std::string& err) {
    if (!m) { err = "mesh is null"; return false; }
    if (m->num_verts == 0) { err = "mesh has no vertices"; return false; }
    if (!m->rest_x || !m->vert_mass || !m->is_fixed || !m->clusters_csr || !m->vert_adj_edges_csr ||
        !m->vert_adj_bending_csr) { err = "mesh array is null"; return false; }
    if (m->num_edges && (!m->edges || !m->edges_rest_length)) { err = "edge arrays are null"; return false; }
    if (m->num_bending_edges && (!m->bending_edges || !m->bending_Q)) { err = "bending arrays are null"; return false; }
    const uint32_t V = m->num_verts, C = m->num_clusters;
    if (m->clusters_csr[0] != C + 1 || m->clusters_csr[C] != C + 1 + V) {
        err = "clusters CSR does not partition the vertices";
        return false;
    }
    if (m->vert_adj_edges_csr[0] != V + 1 || m->vert_adj_bending_csr[0] != V + 1) {
        err = "adjacency CSR header mismatch";
        return false;
    }
    return true;
}

// Appends one cloth.  inv[vid] receives the global slot of each vertex.
inline bool append_cloth(const lats_mesh_desc& m, ClothLayout& L, std::vector<uint32_t>& inv, std::string& err) {
    if (!validate_mesh(&m, err)) return false;
    const uint32_t V = m.num_verts, E = m.num_edges, B = m.num_bending_edges, C = m.num_clusters;
    const uint32_t* cl = m.clusters_csr;
    const uint32_t* pe = m.vert_adj_edges_csr;
    const uint32_t* pb = m.vert_adj_bending_csr;
    const uint64_t base = L.num_slots();
    if (base + V >= (1ull << 30)) { err = "more than 2^30 vertices"; return false; }
    inv.assign(V, 0xffffffffu);
    for (uint32_t s = 0; s < V; s++) {
        const uint32_t vid = cl[C + 1 + s];
        if (vid >= V || inv[vid] != 0xffffffffu) { err = "clusters CSR is not a permutation of the vertices"; return false; }
        inv[vid] = (uint32_t)(base + s);
    }
    // a colour pass updates its vertices concurrently, so a cluster must be an
    // independent set of the graph the reference colours
    // (vert_adj_verts_with_bending, example2_async_iter.cpp:287-303): no
    // spring and no bending pair may hold two vertices of one cluster
    {
        std::vector<uint32_t> colour(V);
        for (uint32_t k = 0; k < C; k++)
            for (uint32_t r = cl[k]; r < cl[k + 1]; r++) colour[cl[r]] = k;
        for (uint32_t e = 0; e < E; e++) {
            const uint32_t a0 = (uint32_t)m.edges[2 * (size_t)e], a1 = (uint32_t)m.edges[2 * (size_t)e + 1];
            if (a0 >= V || a1 >= V) { err = "edge vertex out of range"; return false; }
            if (colour[a0] == colour[a1]) {
                err = "clusters are not independent: spring " + std::to_string(e) + " joins vertices " +
                      std::to_string(a0) + " and " + std::to_string(a1) + " of cluster " + std::to_string(colour[a0]);
                return false;
            }
        }
        for (uint32_t b = 0; b < B; b++) {
            const int32_t* q4 = &m.bending_edges[4 * (size_t)b];
            for (int t = 0; t < 4; t++)
                if ((uint32_t)q4[t] >= V) { err = "bending vertex out of range"; return false; }
            for (int t = 0; t < 4; t++)
                for (int u = t + 1; u < 4; u++)
                    if (colour[q4[t]] == colour[q4[u]]) {
                        err = "clusters are not independent: bending pair " + std::to_string(b) + " holds vertices " +
                              std::to_string(q4[t]) + " and " + std::to_string(q4[u]) + " of cluster " +
                              std::to_string(colour[q4[t]]);
                        return false;
                    }
        }
    }
    // per-slot arrays
    for (uint32_t s = 0; s < V; s++) {
        const uint32_t vid = cl[C + 1 + s];
        for (int k = 0; k < 3; k++) L.x.push_back(m.rest_x[3 * (size_t)vid + k]);
        L.mass.push_back(m.vert_mass[vid]);
        const uint32_t ed = pe[vid + 1] - pe[vid], bd = pb[vid + 1] - pb[vid];
        if (ed > 255 || bd > 255) { err = "vertex degree above 255"; return false; }
        L.meta.push_back(ed | (bd << 8) | ((m.is_fixed[vid] ? 1u : 0u) << 16));
        L.perm.push_back(vid);
    }
    // colour-major ELL per colour
    for (uint32_t k = 0; k < C; k++) {
        const uint32_t start = (uint32_t)base + (cl[k] - (C + 1)), n = cl[k + 1] - cl[k];
        if (n == 0) { err = "empty colour cluster"; return false; }
        uint32_t emax = 0, bmax = 0;
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t mt = L.meta[start + i];
            emax = std::max(emax, mt & 0xffu);
            bmax = std::max(bmax, (mt >> 8) & 0xffu);
        }
        const size_t eoff = L.ell_e.size(), boff = L.ell_bq.size();
        if (eoff + (size_t)emax * n >= (1ull << 32) || boff + (size_t)bmax * n >= (1ull << 32)) {
            err = "adjacency too large";
            return false;
        }
        L.colours.push_back(ColourInfo{start, n, (uint32_t)eoff, (uint32_t)boff, emax, bmax});
        L.ell_e.resize(eoff + (size_t)emax * n, uint2{0, 0});
        L.ell_bq.resize(boff + (size_t)bmax * n, float4{0, 0, 0, 0});
        L.ell_bi.resize(boff + (size_t)bmax * n, PairIds{0, 0, 0});
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t vid = cl[C + 1 + (start - (uint32_t)base) + i];
            for (uint32_t j = 0, r = pe[vid]; r < pe[vid + 1]; r++, j++) {
                const uint32_t e = pe[r];
                if (e >= E) { err = "edge id out of range"; return false; }
                const uint32_t a0 = (uint32_t)m.edges[2 * (size_t)e], a1 = (uint32_t)m.edges[2 * (size_t)e + 1];
                if (a0 >= V || a1 >= V) { err = "edge vertex out of range"; return false; }
                const uint32_t other = (a0 == vid) ? a1 : a0;
                uint32_t rl;
                std::memcpy(&rl, &m.edges_rest_length[e], 4);
                L.ell_e[eoff + (size_t)j * n + i] = uint2{inv[other], rl};
            }
            for (uint32_t j = 0, r = pb[vid]; r < pb[vid + 1]; r++, j++) {
                const uint32_t b = pb[r];
                if (b >= B) { err = "bending id out of range"; return false; }
                const int32_t* q4 = &m.bending_edges[4 * (size_t)b];
                for (int t = 0; t < 4; t++)
                    if ((uint32_t)q4[t] >= V) { err = "bending vertex out of range"; return false; }
                uint32_t o = 0;
                while (o < 4 && (uint32_t)q4[o] != vid) o++;
                if (o == 4) { err = "bending adjacency mismatch"; return false; }
                const float* Q = &m.bending_Q[16 * (size_t)b + 4 * o];
                const size_t at = boff + (size_t)j * n + i;
                L.ell_bq[at] = float4{Q[0], Q[1], Q[2], Q[3]};
                uint32_t others[3], no = 0;
                for (uint32_t t = 0; t < 4; t++)
                    if (t != o) others[no++] = inv[q4[t]];
                L.ell_bi[at] = PairIds{others[0] | (o << 30), others[1], others[2]};
            }
        }
    }
    // element lists (energy)
    for (uint32_t e = 0; e < E; e++) {
        L.e_slots.push_back(uint2{inv[m.edges[2 * (size_t)e]], inv[m.edges[2 * (size_t)e + 1]]});
        L.rest.push_back(m.edges_rest_length[e]);
    }
    for (uint32_t b = 0; b < B; b++) {
        const int32_t* q4 = &m.bending_edges[4 * (size_t)b];
        L.b_slots.push_back(uint4{inv[q4[0]], inv[q4[1]], inv[q4[2]], inv[q4[3]]});
        for (int col = 0; col < 4; col++) {
            const float* Q = &m.bending_Q[16 * (size_t)b + 4 * col];
            L.Q.push_back(float4{Q[0], Q[1], Q[2], Q[3]});
        }
    }
    return true;
}

}  // namespace lats
