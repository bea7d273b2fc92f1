// topology.cpp -- see topology.h.  Compiled with -ffp-contract=off: the rest
// lengths and Q must be bit-identical to the reference's float arithmetic.
#include "topology.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace lats {
namespace {

struct F3 { float x, y, z; };
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 neg(F3 a) { return {-a.x, -a.y, -a.z}; }
// glm semantics: dot = (x*x' + y*y') + z*z'; cross = (y z' - y' z, z x' - z' x, x y' - x' y)
inline float dot3(F3 a, F3 b) {
    const float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
    return (px + py) + pz;
}
inline F3 cross3(F3 a, F3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float len3(F3 a) { return std::sqrt(dot3(a, a)); }

// Packs per-row counts + a fill callback into the reference's packed CSR.
template <typename Fill>
std::vector<uint32_t> packed_from_counts(const std::vector<uint32_t>& counts, Fill fill) {
    const uint32_t n = (uint32_t)counts.size();
    std::vector<uint32_t> p(n + 1);
    uint32_t off = n + 1;
    for (uint32_t i = 0; i < n; i++) { p[i] = off; off += counts[i]; }
    p[n] = off;
    p.resize(off);
    std::vector<uint32_t> cursor(p.begin(), p.begin() + n);
    fill(p, cursor);
    return p;
}

}  // namespace

bool build_topology(const float* xp, uint32_t nv, const int32_t* faces, uint32_t nf, Topology& t, std::string& err) {
    t = Topology();
    t.num_verts = nv;
    t.faces.assign(faces, faces + 3 * (size_t)nf);
    for (size_t i = 0; i < 3 * (size_t)nf; i++) {
        if (faces[i] < 0 || (uint32_t)faces[i] >= nv) { err = "face index out of range"; return false; }
    }

    // --- edges and bending pairs: sort half-edges on the 96-bit key (lo, hi, face id).
    struct HalfEdge { uint64_t key; uint32_t fid; };
    std::vector<HalfEdge> he(3 * (size_t)nf);
    for (uint32_t f = 0; f < nf; f++) {
        uint32_t a = (uint32_t)faces[3 * f], b = (uint32_t)faces[3 * f + 1], c = (uint32_t)faces[3 * f + 2];
        if (a > b) std::swap(a, b);
        if (b > c) std::swap(b, c);
        if (a > b) std::swap(a, b);
        he[3 * f + 0] = {((uint64_t)a << 32) | b, f};
        he[3 * f + 1] = {((uint64_t)a << 32) | c, f};
        he[3 * f + 2] = {((uint64_t)b << 32) | c, f};
    }
    std::sort(he.begin(), he.end(), [](const HalfEdge& l, const HalfEdge& r) {
        return l.key != r.key ? l.key < r.key : l.fid < r.fid;
    });
    // Runs of equal keys: a run of length 1 is a boundary edge; in a run of
    // length r > 1 the first entry starts a bending pair with the second and
    // entries 1..r-1 are emitted as edges (the reference's type 1 / type 2).
    for (size_t i = 0; i < he.size();) {
        size_t j = i + 1;
        while (j < he.size() && he[j].key == he[i].key) j++;
        const int32_t a = (int32_t)(he[i].key >> 32), b = (int32_t)(he[i].key & 0xffffffffu);
        if (j - i == 1) {
            t.edges.push_back(a);
            t.edges.push_back(b);
        } else {
            for (size_t k = i + 1; k < j; k++) { t.edges.push_back(a); t.edges.push_back(b); }
        }
        i = j;
    }
    for (size_t i = 0; i + 1 < he.size(); i++) {
        if (he[i + 1].key != he[i].key) continue;
        if (i > 0 && he[i - 1].key == he[i].key) continue;   // only the first of a run
        const int32_t a = (int32_t)(he[i].key >> 32), b = (int32_t)(he[i].key & 0xffffffffu);
        const int32_t* f0 = faces + 3 * (size_t)he[i].fid;
        const int32_t* f1 = faces + 3 * (size_t)he[i + 1].fid;
        t.bending.push_back(a);
        t.bending.push_back(b);
        t.bending.push_back((f0[0] + f0[1] + f0[2]) - (a + b));
        t.bending.push_back((f1[0] + f1[1] + f1[2]) - (a + b));
    }
    const uint32_t ne = t.num_edges(), nb = t.num_bending();

    // --- vertex -> element packed CSRs (ascending element id per row).
    {
        std::vector<uint32_t> cnt(nv, 0);
        for (uint32_t e = 0; e < 2 * ne; e++) cnt[t.edges[e]]++;
        t.adj_edges_csr = packed_from_counts(cnt, [&](std::vector<uint32_t>& p, std::vector<uint32_t>& cur) {
            for (uint32_t e = 0; e < ne; e++)
                for (int j = 0; j < 2; j++) p[cur[t.edges[2 * e + j]]++] = e;
        });
    }
    {
        std::vector<uint32_t> cnt(nv, 0);
        for (uint32_t e = 0; e < 4 * nb; e++) cnt[t.bending[e]]++;
        t.adj_bending_csr = packed_from_counts(cnt, [&](std::vector<uint32_t>& p, std::vector<uint32_t>& cur) {
            for (uint32_t b = 0; b < nb; b++)
                for (int j = 0; j < 4; j++) p[cur[t.bending[4 * b + j]]++] = b;
        });
    }

    // --- colouring.  The reference's round-based greedy (a round takes every
    // unmarked vertex in ascending id and marks its neighbours) gives vertex v
    // the smallest colour not held by a lower-id neighbour: first-fit in id
    // order.  Neighbours = edge neighbours + the other members of each
    // bending pair of v.
    t.colour.assign(nv, 0);
    {
        const auto& pe = t.adj_edges_csr;
        const auto& pb = t.adj_bending_csr;
        uint64_t used[4];
        uint32_t maxc = 0;
        for (uint32_t v = 0; v < nv; v++) {
            used[0] = used[1] = used[2] = used[3] = 0;
            auto mark = [&](uint32_t u) {
                if (u < v) used[t.colour[u] >> 6] |= 1ull << (t.colour[u] & 63);
            };
            for (uint32_t k = pe[v]; k < pe[v + 1]; k++) {
                const uint32_t e = pe[k];
                mark((uint32_t)t.edges[2 * e] == v ? (uint32_t)t.edges[2 * e + 1] : (uint32_t)t.edges[2 * e]);
            }
            for (uint32_t k = pb[v]; k < pb[v + 1]; k++)
                for (int j = 0; j < 4; j++) mark((uint32_t)t.bending[4 * pb[k] + j]);
            uint32_t c = 0;
            while (c < 256 && (used[c >> 6] >> (c & 63) & 1)) c++;
            if (c >= 255) { err = "more than 255 colours"; return false; }
            t.colour[v] = (uint8_t)c;
            maxc = std::max(maxc, c + 1);
        }
        t.num_clusters = nv ? maxc : 0;
        std::vector<uint32_t> cnt(t.num_clusters, 0);
        for (uint32_t v = 0; v < nv; v++) cnt[t.colour[v]]++;
        t.clusters_csr = packed_from_counts(cnt, [&](std::vector<uint32_t>& p, std::vector<uint32_t>& cur) {
            for (uint32_t v = 0; v < nv; v++) p[cur[t.colour[v]]++] = v;
        });
    }

    // --- rest state (float, reference operation order).
    auto X = [&](int32_t i) { return F3{xp[3 * (size_t)i], xp[3 * (size_t)i + 1], xp[3 * (size_t)i + 2]}; };
    t.rest_length.resize(ne);
    for (uint32_t e = 0; e < ne; e++) t.rest_length[e] = len3(sub(X(t.edges[2 * e]), X(t.edges[2 * e + 1])));
    t.bending_Q.resize(16 * (size_t)nb);
    for (uint32_t b = 0; b < nb; b++) {
        const F3 p0 = X(t.bending[4 * b]), p1 = X(t.bending[4 * b + 1]);
        const F3 p2 = X(t.bending[4 * b + 2]), p3 = X(t.bending[4 * b + 3]);
        const F3 e0 = sub(p1, p0), e1 = sub(p2, p0), e2 = sub(p3, p0), e3 = sub(p2, p1), e4 = sub(p3, p1);
        auto cot = [](F3 u, F3 w) { return dot3(u, w) / len3(cross3(u, w)); };
        const float c01 = cot(e0, neg(e1)), c02 = cot(e0, neg(e2)), c03 = cot(e0, e3), c04 = cot(e0, e4);
        const float K[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
        const float A0 = 0.5f * len3(cross3(e0, e1));
        const float A1 = 0.5f * len3(cross3(e0, e2));
        const float s = 3.f / (A0 + A1);
        float* Q = &t.bending_Q[16 * (size_t)b];
        for (int col = 0; col < 4; col++)
            for (int row = 0; row < 4; row++) Q[col * 4 + row] = s * (K[row] * K[col]);
    }
    return true;
}

}  // namespace lats
