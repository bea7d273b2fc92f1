// test_patch_layout.hip -- CPU check of the patch strategy's layout
// (libatssim_amd/csrc/vbd_patch_build.h).  Host code only; runs without a GPU.
//
// For grid cloths (even, odd and narrow sizes) and an irregular mesh with
// holes, at several patch counts and colours per launch:
//  * every index a patch launch computes -- heads and steps, region slots,
//    update entries, local ELL rows for every lane of every round (lanes past
//    a step's end, rows clamped to the width, zero widths, the tail rows past
//    the prefetched rounds) and every LDS index gathered through them -- lies
//    inside the arrays the launch reads, for 256 / 512 / 1024 threads (8
//    lanes per update), and the launch's LDS sections fit its allocation;
//  * every local pair entry names the pair's four points in pair order (the
//    vertex's own LDS word at its position o) and carries Q[o][o];
//  * workgroup b -> patch (b % 8) * (P / 8) + b / 8 covers every patch once
//    and the owned slots cover every slot once per launch;
//  * emulating the launches (region -> LDS, colour steps on LDS, owned slots
//    written to the other buffer, buffers swapped) with a deterministic stand-in
//    for the vertex update that reads exactly the data the kernel reads gives
//    the same values, bit for bit, as the plain colour-by-colour sweep over the
//    global layout.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "../../libatssim_amd/csrc/vbd_layout.h"
#include "../../libatssim_amd/csrc/vbd_patch_build.h"

using namespace lats;

static int g_fail = 0;
#define CHECK(cond, ...)                                \
    do {                                                \
        if (!(cond)) {                                  \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                   \
            std::printf("\n");                          \
            g_fail++;                                   \
            return false;                               \
        }                                               \
    } while (0)

struct Cloth {
    std::vector<float> x;
    std::vector<int32_t> f;
};

static Cloth grid(uint32_t nx, uint32_t ny, uint64_t seed) {
    Cloth c;
    c.x.resize(3 * (size_t)nx * ny);
    c.f.resize(6 * (size_t)(nx - 1) * (ny - 1));
    if (lats_scene_grid(nx, ny, 0.01f, 1e-4f, seed, c.x.data(), c.f.data()) != LATS_OK) std::exit(2);
    return c;
}

// A grid with about one triangle in eight removed (holes, boundary vertices
// of every degree, isolated vertices).
static Cloth holey(uint32_t nx, uint32_t ny, uint64_t seed) {
    Cloth c = grid(nx, ny, seed);
    std::vector<int32_t> kept;
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    for (size_t t = 0; t < c.f.size() / 3; t++) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        if ((s >> 61) == 0) continue;
        kept.insert(kept.end(), c.f.begin() + 3 * t, c.f.begin() + 3 * t + 3);
    }
    c.f = kept;
    return c;
}

struct Layout {
    ClothLayout L;
    uint32_t V = 0, C = 0;
};

static bool make_layout(const Cloth& cl, Layout& out) {
    const uint32_t V = (uint32_t)(cl.x.size() / 3), F = (uint32_t)(cl.f.size() / 3);
    lats_topology* t = nullptr;
    if (lats_topology_build(cl.x.data(), V, cl.f.data(), F, &t) != LATS_OK) return false;
    uint32_t n[4];
    lats_topology_counts(t, n);
    auto arr = [&](int which, auto& vec) {
        uint64_t k = 0;
        lats_topology_size(t, which, &k);
        vec.resize(k);
        lats_topology_copy(t, which, vec.data());
    };
    std::vector<int32_t> edges, bend;
    std::vector<float> rest, Q;
    std::vector<uint32_t> ae, ab, cl_csr;
    arr(LATS_TOPO_EDGES, edges);
    arr(LATS_TOPO_BENDING, bend);
    arr(LATS_TOPO_REST_LENGTH, rest);
    arr(LATS_TOPO_BENDING_Q, Q);
    arr(LATS_TOPO_ADJ_EDGES_CSR, ae);
    arr(LATS_TOPO_ADJ_BENDING_CSR, ab);
    arr(LATS_TOPO_CLUSTERS_CSR, cl_csr);
    lats_topology_destroy(t);
    std::vector<float> mass(V, 1e-3f);
    std::vector<uint8_t> fixed(V, 0);
    lats_mesh_desc m{};
    m.num_verts = n[0];
    m.num_edges = n[1];
    m.num_bending_edges = n[2];
    m.num_clusters = n[3];
    m.rest_x = cl.x.data();
    m.edges = edges.data();
    m.bending_edges = bend.data();
    m.edges_rest_length = rest.data();
    m.bending_Q = Q.data();
    m.vert_adj_edges_csr = ae.data();
    m.vert_adj_bending_csr = ab.data();
    m.clusters_csr = cl_csr.data();
    m.vert_mass = mass.data();
    m.is_fixed = fixed.data();
    std::vector<uint32_t> inv;
    std::string err;
    if (!append_cloth(m, out.L, inv, err)) {
        std::printf("append_cloth: %s\n", err.c_str());
        return false;
    }
    out.V = V;
    out.C = n[3];
    return true;
}

// Bounds of every index one launch of group g computes (k_vbd_patch with B
// threads, 8 lanes per update).
static bool check_launch(const PatchHost& h, uint32_t g, uint32_t P, uint32_t V, uint32_t B) {
    constexpr int G = kPatchLanes;
    constexpr uint32_t PE = GroupLoads<G>::PE, PB = GroupLoads<G>::PB;
    const uint32_t S = B / G;
    std::vector<int> seen(P, 0);
    for (uint32_t b = 0; b < P; b++) {
        const uint32_t p = (b % kXcds) * (P / kXcds) + b / kXcds;
        CHECK(p < P, "block %u -> patch %u of %u", b, p, P);
        seen[p]++;
        const size_t hi = (size_t)g * P + p;
        CHECK(hi < h.heads.size(), "head %zu of %zu", hi, h.heads.size());
        const PatchHead& hd = h.heads[hi];
        // the kernel addresses the region and the steps by stride: patch p's
        // region entries [0, rstride) and step records [0, F) are read
        const uint32_t F = h.stats.colours_per_launch, rs = h.group_rstride[g];
        const uint32_t nst = h.group_c0[g + 1] - h.group_c0[g];
        CHECK(hd.roff == h.group_roff[g] + p * rs, "region address %u", hd.roff);
        CHECK(hd.rsize <= rs && rs >= 1, "region stride %u below its size %u", rs, hd.rsize);
        // frame state is read for entries [0, nsmax) of the region (in region order)
        CHECK(hd.nstate <= h.group_nsmax[g] && h.group_nsmax[g] <= rs && h.group_nsmax[g] >= 1, "state bound");
        CHECK((size_t)hd.roff + rs <= h.region.size(), "region range");
        CHECK(hd.owned <= hd.rsize, "owned > region");
        for (uint32_t r = 0; r < rs; r++) CHECK(h.region[hd.roff + r] < V, "region slot out of range");
        CHECK(hd.step_off == (uint32_t)(hi * F), "step address %u", hd.step_off);
        CHECK(hd.nsteps == nst && nst <= F, "steps %u of the group's %u", hd.nsteps, nst);
        CHECK((size_t)hd.step_off + F <= h.steps.size(), "step range");
        CHECK(hd.nsteps >= 1, "a launch without steps");
        CHECK(hd.nstate <= hd.rsize && hd.owned <= hd.nstate, "state prefix");
        CHECK(patch_lds_bytes(hd.rsize, hd.nstate, hd.owned) <= h.group_lds[g], "LDS above the launch's size");
        // the kernel's LDS sections: staged sums [0, pos) of exactly the
        // staged size, then positions, state, slots in order, disjoint, and
        // inside what the launch allocates (group_lds + staged bytes)
        {
            const PatchLds l = patch_lds(B, hd.rsize, hd.nstate, hd.owned);
            const uint32_t staged = patch_staged_bytes(B);
            CHECK(l.pos == staged && staged == B / 8 * kStagedWords * 4, "staged section");
            CHECK(l.x0m == l.pos + 16 * hd.rsize && l.vm == l.x0m + 16 * hd.nstate &&
                      l.slot == l.vm + 16 * hd.nstate && l.end == l.slot + 4 * hd.owned,
                  "LDS sections out of order");
            CHECK(l.pos % 16 == 0, "positions not 16-byte aligned");
            CHECK(l.end <= h.group_lds[g] + staged, "LDS end %u above the allocation %u", l.end,
                  h.group_lds[g] + staged);
        }
        for (uint32_t j = 0; j < hd.nsteps; j++) {
            const PatchStep& st = h.steps[hd.step_off + j];
            // every step's first entry and ELL entry 0 are loaded, empty or
            // not (an empty step's values are never used)
            CHECK(st.off < h.entries.size(), "step entry offset %u of %zu", st.off, h.entries.size());
            if (st.n == 0) {
                CHECK(st.ew == 0 && st.bw == 0, "empty step with an ELL width");
                continue;
            }
            for (uint32_t base = 0; base < st.n; base += S)
                for (uint32_t gi = 0; gi < S; gi++)
                    for (uint32_t lane = 0; lane < (uint32_t)G; lane++) {
                        const uint32_t i = base + gi, ii = i < st.n ? i : 0u;
                        const size_t e = (size_t)st.off + ii;
                        CHECK(e < h.entries.size(), "entry %zu of %zu", e, h.entries.size());
                        const uint2 ent = h.entries[e];
                        CHECK(ent.x < V && ent.y < hd.nstate, "entry slot %u / LDS %u (state %u)", ent.x, ent.y,
                              hd.nstate);
                        const uint32_t e_on = st.ew != 0u, b_on = st.bw != 0u;
                        for (uint32_t k = 0; k < PE; k++) {
                            const uint32_t jj = k * G + lane, row = jj < st.ew ? jj : st.ew - 1;
                            const size_t at = (size_t)(st.eoff + row * st.n + ii) * e_on;
                            CHECK(at < h.ell_e.size(), "spring ELL %zu of %zu", at, h.ell_e.size());
                            CHECK(h.ell_e[at].x < hd.rsize, "spring LDS index %u of %u", h.ell_e[at].x, hd.rsize);
                        }
                        for (uint32_t k = 0; k < PB; k++) {
                            const uint32_t jj = k * G + lane, row = jj < st.bw ? jj : st.bw - 1;
                            const size_t at = (size_t)(st.boff + row * st.n + ii) * b_on;
                            CHECK(at < h.ell_bq.size() && at < h.ell_bi.size(), "pair ELL %zu", at);
                            const PatchPair q = h.ell_bi[at];
                            CHECK((q.i01 & 0xffffu) < hd.rsize && (q.i01 >> 16) < hd.rsize &&
                                      (q.i23 & 0xffffu) < hd.rsize && (q.i23 >> 16) < hd.rsize,
                                  "pair LDS index");
                        }
                        if (i >= st.n) continue;
                        // tail rows past the prefetched rounds (valid lanes only)
                        for (uint32_t jj = PE * G + lane; jj < st.ew; jj += G) {
                            const size_t at = (size_t)st.eoff + jj * st.n + i;
                            CHECK(at < h.ell_e.size() && h.ell_e[at].x < hd.rsize, "tail spring");
                        }
                        for (uint32_t jj = PB * G + lane; jj < st.bw; jj += G) {
                            const size_t at = (size_t)st.boff + jj * st.n + i;
                            CHECK(at < h.ell_bi.size(), "tail pair");
                            const PatchPair q = h.ell_bi[at];
                            CHECK((q.i01 & 0xffffu) < hd.rsize && (q.i01 >> 16) < hd.rsize &&
                                      (q.i23 & 0xffffu) < hd.rsize && (q.i23 >> 16) < hd.rsize,
                                  "tail pair LDS index");
                        }
                    }
        }
    }
    for (uint32_t p = 0; p < P; p++) CHECK(seen[p] == 1, "patch %u run %d times", p, seen[p]);
    std::vector<int> owned(V, 0);
    for (uint32_t p = 0; p < P; p++) {
        const PatchHead& hd = h.heads[(size_t)g * P + p];
        for (uint32_t r = 0; r < hd.owned; r++) owned[h.region[hd.roff + r]]++;
    }
    for (uint32_t s = 0; s < V; s++) CHECK(owned[s] == 1, "slot %u owned %d times", s, owned[s]);
    return true;
}

// Stand-in for the vertex update: reads the own position, every incident
// spring (other position, rest length) and every incident pair (Q row, the
// pair's four points in pair order, Q[o][o]) in incidence order, as the
// kernel does.  ell.pair gives the points and Q[o][o] from the layout's own
// entry format: the global layout's (three other slots, the vertex at o) or
// the patch's (four LDS indices, Q[o][o] stored).
using PosFn = std::function<const double*(uint32_t)>;
template <class Pos, class Ell>
static void update(uint32_t meta, double* own, Pos pos, Ell ell) {
    const uint32_t ed = meta & 0xffu, bd = (meta >> 8) & 0xffu;
    double acc[3];
    for (int d = 0; d < 3; d++) acc[d] = 0.37 * own[d];
    for (uint32_t row = 0; row < ed; row++) {
        uint2 e;
        ell.spring(row, e);
        float rl;
        std::memcpy(&rl, &e.y, 4);
        const double* xo = pos(e.x);
        for (int d = 0; d < 3; d++) acc[d] = acc[d] + 0.11 * xo[d] + 1e-3 * rl;
    }
    for (uint32_t row = 0; row < bd; row++) {
        float4 q;
        const double* pt[4];
        float qoo;
        ell.pair(row, own, PosFn(pos), q, pt, qoo);
        for (int d = 0; d < 3; d++)
            acc[d] = acc[d] + 1e-4 * (q.x * pt[0][d] + q.y * pt[1][d] + q.z * pt[2][d] + q.w * pt[3][d]) + 1e-5 * qoo;
    }
    for (int d = 0; d < 3; d++) own[d] = acc[d];
}

static bool emulate(const Layout& lay, const PatchHost& h, uint32_t sweeps) {
    const uint32_t V = lay.V, P = h.stats.patches, ngroups = (uint32_t)h.group_lds.size();
    const ClothLayout& L = lay.L;
    std::vector<double> plain(3 * (size_t)V), a(3 * (size_t)V), b(3 * (size_t)V);
    for (size_t k = 0; k < plain.size(); k++) plain[k] = a[k] = L.x[k];
    // plain: colour by colour, in place, global ELL
    for (uint32_t sw = 0; sw < sweeps; sw++)
        for (uint32_t c = 0; c < lay.C; c++) {
            const ColourInfo& ci = L.colours[c];
            for (uint32_t i = 0; i < ci.count; i++) {
                const uint32_t s = ci.start + i;
                struct {
                    const ClothLayout& L;
                    const ColourInfo& ci;
                    uint32_t i;
                    void spring(uint32_t row, uint2& e) const { e = L.ell_e[ci.eoff + (size_t)row * ci.count + i]; }
                    void pair(uint32_t row, const double* own, const PosFn& pos, float4& q, const double* (&pt)[4],
                              float& qoo) const {
                        q = L.ell_bq[ci.boff + (size_t)row * ci.count + i];
                        const PairIds id = L.ell_bi[ci.boff + (size_t)row * ci.count + i];
                        const uint32_t o = id.a >> 30;
                        const double* g0 = pos(id.a & 0x3fffffffu);
                        const double* g1 = pos(id.b);
                        const double* g2 = pos(id.c);
                        pt[0] = o == 0 ? own : g0;
                        pt[1] = o == 1 ? own : (o == 0 ? g0 : g1);
                        pt[2] = o == 2 ? own : (o <= 1 ? g1 : g2);
                        pt[3] = o == 3 ? own : g2;
                        const float qs[4] = {q.x, q.y, q.z, q.w};
                        qoo = qs[o];
                    }
                } ell{L, ci, i};
                update(L.meta[s], &plain[3 * (size_t)s], [&](uint32_t u) { return &plain[3 * (size_t)u]; }, ell);
            }
        }
    // patch launches: read one buffer, write the other
    double* xin = a.data();
    double* xout = b.data();
    std::vector<double> lds;
    for (uint32_t sw = 0; sw < sweeps; sw++)
        for (uint32_t g = 0; g < ngroups; g++) {
            for (uint32_t p = 0; p < P; p++) {
                const PatchHead& hd = h.heads[(size_t)g * P + p];
                lds.assign(3 * (size_t)hd.rsize, 0.0);
                for (uint32_t r = 0; r < hd.rsize; r++)
                    for (int d = 0; d < 3; d++) lds[3 * r + d] = xin[3 * (size_t)h.region[hd.roff + r] + d];
                for (uint32_t j = 0; j < hd.nsteps; j++) {
                    const PatchStep& st = h.steps[hd.step_off + j];
                    for (uint32_t i = 0; i < st.n; i++) {
                        const uint2 ent = h.entries[st.off + i];
                        struct {
                            const PatchHost& h;
                            const PatchStep& st;
                            uint32_t i;
                            void spring(uint32_t row, uint2& e) const { e = h.ell_e[st.eoff + (size_t)row * st.n + i]; }
                            void pair(uint32_t row, const double*, const PosFn& pos, float4& q, const double* (&pt)[4],
                                      float& qoo) const {
                                q = h.ell_bq[st.boff + (size_t)row * st.n + i];
                                const PatchPair e = h.ell_bi[st.boff + (size_t)row * st.n + i];
                                pt[0] = pos(e.i01 & 0xffffu);
                                pt[1] = pos(e.i01 >> 16);
                                pt[2] = pos(e.i23 & 0xffffu);
                                pt[3] = pos(e.i23 >> 16);
                                std::memcpy(&qoo, &e.qoo, 4);
                            }
                        } ell{h, st, i};
                        update(L.meta[ent.x], &lds[3 * (size_t)ent.y], [&](uint32_t r) { return &lds[3 * (size_t)r]; },
                               ell);
                    }
                }
                for (uint32_t r = 0; r < hd.owned; r++)
                    for (int d = 0; d < 3; d++) xout[3 * (size_t)h.region[hd.roff + r] + d] = lds[3 * r + d];
            }
            std::swap(xin, xout);
        }
    for (size_t k = 0; k < plain.size(); k++)
        CHECK(std::memcmp(&plain[k], &xin[k], sizeof(double)) == 0, "position %zu: plain %.17g patch %.17g", k,
              plain[k], xin[k]);
    return true;
}

static bool run_case(const char* name, const Cloth& cl, std::vector<std::pair<uint32_t, uint32_t>> geoms) {
    Layout lay;
    if (!make_layout(cl, lay)) {
        std::printf("FAIL %s: layout\n", name);
        g_fail++;
        return false;
    }
    PatchSource src;
    src.V = lay.V;
    src.C = lay.C;
    src.colours = lay.L.colours.data();
    src.meta = lay.L.meta.data();
    src.x = lay.L.x.data();
    src.ell_e = lay.L.ell_e.data();
    src.ell_bq = lay.L.ell_bq.data();
    src.ell_bi = lay.L.ell_bi.data();
    for (auto [P, F] : geoms) {
        PatchHost h;
        std::string err;
        if (!patch_build_host(src, P, F, 160 * 1024, h, err)) {
            std::printf("FAIL %s P=%u F=%u: %s\n", name, P, F, err.c_str());
            g_fail++;
            continue;
        }
        bool ok = true;
        for (uint32_t g = 0; g < h.group_lds.size() && ok; g++)
            for (uint32_t B : {256u, 512u, 1024u})
                ok = ok && check_launch(h, g, P, lay.V, B);
        ok = ok && emulate(lay, h, 3);
        std::printf("%s %s V=%u C=%u P=%u F=%u launches/sweep=%u max_region=%u updates/sweep=%llu (%.2fx)\n",
                    ok ? "ok  " : "FAIL", name, lay.V, lay.C, P, h.stats.colours_per_launch,
                    h.stats.launches_per_sweep, h.stats.max_region, (unsigned long long)h.stats.updates_per_sweep,
                    (double)h.stats.updates_per_sweep / lay.V);
    }
    return true;
}

int main() {
    run_case("grid32", grid(32, 32, 1), {{8, 0}, {16, 0}, {16, 2}, {64, 3}, {128, 1}, {8, 4}});
    run_case("grid23", grid(23, 23, 2), {{8, 0}, {24, 0}, {24, 3}, {64, 2}});
    run_case("grid61x7", grid(61, 7, 3), {{8, 0}, {16, 2}, {48, 0}});
    run_case("grid96", grid(96, 96, 4), {{32, 0}, {64, 0}, {64, 4}, {128, 0}, {256, 2}});
    run_case("holey40", holey(40, 40, 5), {{8, 0}, {32, 0}, {32, 3}, {64, 1}});
    // the 256x256 cloth at its bench geometries (bounds and coverage only)
    {
        Layout lay;
        if (!make_layout(grid(256, 256, 7), lay)) return 2;
        PatchSource src;
        src.V = lay.V;
        src.C = lay.C;
        src.colours = lay.L.colours.data();
        src.meta = lay.L.meta.data();
        src.x = lay.L.x.data();
        src.ell_e = lay.L.ell_e.data();
        src.ell_bq = lay.L.ell_bq.data();
        src.ell_bi = lay.L.ell_bi.data();
        for (auto [P, F] : std::vector<std::pair<uint32_t, uint32_t>>{{256, 0}, {128, 0}, {256, 4}, {512, 2}}) {
            PatchHost h;
            std::string err;
            if (!patch_build_host(src, P, F, 160 * 1024, h, err)) {
                std::printf("FAIL grid256 P=%u F=%u: %s\n", P, F, err.c_str());
                g_fail++;
                continue;
            }
            bool ok = true;
            for (uint32_t g = 0; g < h.group_lds.size() && ok; g++)
                ok = check_launch(h, g, P, lay.V, 1024) && check_launch(h, g, P, lay.V, 512);
            std::printf("%s grid256 P=%u F=%u launches/sweep=%u max_region=%u updates/sweep=%llu (%.2fx) regions "
                        "%.2fx\n",
                        ok ? "ok  " : "FAIL", P, h.stats.colours_per_launch, h.stats.launches_per_sweep,
                        h.stats.max_region, (unsigned long long)h.stats.updates_per_sweep,
                        (double)h.stats.updates_per_sweep / lay.V, (double)h.stats.region_slots / lay.V);
        }
    }
    if (g_fail) {
        std::printf("%d failure(s)\n", g_fail);
        return 1;
    }
    std::printf("PASS\n");
    return 0;
}
