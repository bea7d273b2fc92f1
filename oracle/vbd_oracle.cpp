// vbd_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's Vertex Block Descent (VBD) cloth path,
// used to check the HIP product path.  Only tests/, the smoke()
// entry and bench.py's cpu_baseline leg may load this library.  Nothing in
// the product links it or falls back to it.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference checkout).  Arithmetic is written operation by operation in
// the reference's evaluation order (glm semantics on Linux), compiled with
// -ffp-contract=off, so FP32 results are reproducible bit for bit.
// The same code is instantiated in double for the FP64 error statement.
//
// Third-party arithmetic restated here (glm, version unpinned by the
// reference's CMake `find_package(glm CONFIG REQUIRED)`):
//   glm::dot(vec3)         = (a.x*b.x + a.y*b.y) + a.z*b.z
//   glm::length(vec3)      = sqrt(dot(v, v))
//   glm::cross(x, y)       = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)
//   glm::determinant(mat3) = + m00*(m11*m22 - m21*m12) - m10*(m01*m22 - m21*m02)
//                            + m20*(m01*m12 - m11*m02)     (m[col][row])
//   mat3 * vec3            = row i: (m[0][i]*v.x + m[1][i]*v.y) + m[2][i]*v.z
//
// Pinning (how each part is checked) lives in tests/test_oracle_*.py:
// published energy history of the reference run on square8K, the published
// colour count / cluster sizes, closed forms, finite differences, FP64 solves
// and brute force on tiny meshes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint32_t u32;

namespace orc {

template <typename T> struct V3 {
    T x, y, z;
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename T> static inline V3<T> mk(T a, T b, T c) { return V3<T>{a, b, c}; }
template <typename T> static inline V3<T> operator+(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> static inline V3<T> operator-(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> static inline V3<T> operator-(V3<T> a) { return mk(-a.x, -a.y, -a.z); }
template <typename T> static inline V3<T> operator*(T s, V3<T> a) { return mk(s * a.x, s * a.y, s * a.z); }
template <typename T> static inline V3<T> operator*(V3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <typename T> static inline V3<T> operator/(V3<T> a, T s) { return mk(a.x / s, a.y / s, a.z / s); }
// glm compute_dot<vec3>: tmp = a*b; (tmp.x + tmp.y) + tmp.z
template <typename T> static inline T dot(V3<T> a, V3<T> b) {
    T px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
    return (px + py) + pz;
}
template <typename T> static inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <typename T> static inline V3<T> cross(V3<T> x, V3<T> y) {
    return mk(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}

// Float4x3 of the reference (xpbd_constraints.h:4017-4025): columns 0..2 = H,
// column 3 = f.  Stored as h[col][row] and f[row].
template <typename T> struct Hf {
    T h[3][3];
    T f[3];
};
template <typename T> static inline Hf<T> hf_zero() {
    Hf<T> r;
    for (int c = 0; c < 3; c++) {
        for (int i = 0; i < 3; i++) r.h[c][i] = T(0);
        r.f[c] = T(0);
    }
    return r;
}
template <typename T> static inline void hf_add(Hf<T>& a, const Hf<T>& b) {
    for (int c = 0; c < 3; c++) {
        for (int i = 0; i < 3; i++) a.h[c][i] = a.h[c][i] + b.h[c][i];
        a.f[c] = a.f[c] + b.f[c];
    }
}

// ---------------------------------------------------------------------------
// Topology.  mesh_reader.h:206-298 (extract_edges_from_surface<true>).
// ---------------------------------------------------------------------------
struct Topology {
    u32 nv = 0;
    std::vector<int> faces;      // 3F
    std::vector<int> edges;      // 2E
    std::vector<int> bending;    // 4B (e0, e1, opposite in lower-fid face, opposite in higher-fid face)
    std::vector<std::vector<u32>> adj_edges, adj_bending, adj_verts, adj_verts_bend;
    std::vector<std::vector<u32>> clusters;
};

static void extract_edges(const std::vector<int>& faces, std::vector<int>& edges, std::vector<int>& bending) {
    // mesh_reader.h:213-229: per face sort its 3 ids, emit (a,b,fid),(a,c,fid),(b,c,fid)
    const size_t nf = faces.size() / 3;
    struct E3 { int a, b, f; };
    std::vector<E3> tmp(nf * 3);
    for (size_t fid = 0; fid < nf; fid++) {
        int t[3] = {faces[3 * fid], faces[3 * fid + 1], faces[3 * fid + 2]};
        std::sort(t, t + 3);
        tmp[3 * fid + 0] = {t[0], t[1], (int)fid};
        tmp[3 * fid + 1] = {t[0], t[2], (int)fid};
        tmp[3 * fid + 2] = {t[1], t[2], (int)fid};
    }
    // mesh_reader.h:230-236: lexicographic sort on (a, b, fid)
    std::sort(tmp.begin(), tmp.end(), [](const E3& l, const E3& r) {
        if (l.a != r.a) return l.a < r.a;
        if (l.b != r.b) return l.b < r.b;
        return l.f < r.f;
    });
    // mesh_reader.h:237-243: type 1 if next shares (a,b); type 2 if previous does (overrides)
    std::vector<unsigned char> type(tmp.size(), 0);
    for (size_t i = 0; i < tmp.size(); i++) {
        if (i + 1 != tmp.size() && tmp[i + 1].a == tmp[i].a && tmp[i + 1].b == tmp[i].b) type[i] = 1;
        if (i != 0 && tmp[i - 1].a == tmp[i].a && tmp[i - 1].b == tmp[i].b) type[i] = 2;
    }
    // mesh_reader.h:245-296: edges = type 0|2 in sorted order; bending = type 1
    edges.clear();
    bending.clear();
    for (size_t i = 0; i < tmp.size(); i++) {
        if (type[i] == 0 || type[i] == 2) {
            edges.push_back(tmp[i].a);
            edges.push_back(tmp[i].b);
        }
    }
    for (size_t i = 0; i < tmp.size(); i++) {
        if (type[i] != 1) continue;
        const int* cf = &faces[3 * tmp[i].f];
        const int* nf2 = &faces[3 * tmp[i + 1].f];
        const int s = tmp[i].a + tmp[i].b;
        bending.push_back(tmp[i].a);
        bending.push_back(tmp[i].b);
        bending.push_back((cf[0] + cf[1] + cf[2]) - s);
        bending.push_back((nf2[0] + nf2[1] + nf2[2]) - s);
    }
}

// example2_async_iter.cpp:229-304 (adjacency lists) and 572-646 (VBD colouring)
static void build_topology(Topology& t) {
    const u32 nv = t.nv;
    const u32 ne = (u32)t.edges.size() / 2, nb = (u32)t.bending.size() / 4;
    t.adj_edges.assign(nv, {});
    t.adj_bending.assign(nv, {});
    t.adj_verts.assign(nv, {});
    for (u32 e = 0; e < ne; e++)
        for (int j = 0; j < 2; j++) t.adj_edges[t.edges[2 * e + j]].push_back(e);          // :245-252
    for (u32 b = 0; b < nb; b++)
        for (int j = 0; j < 4; j++) t.adj_bending[t.bending[4 * b + j]].push_back(b);      // :255-261
    for (u32 e = 0; e < ne; e++)
        for (int j = 0; j < 2; j++) t.adj_verts[t.edges[2 * e + j]].push_back(t.edges[2 * e + 1 - j]);  // :264-275
    t.adj_verts_bend = t.adj_verts;                                                          // :278-303
    for (u32 b = 0; b < nb; b++)
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                if (i == j) continue;
                auto& l = t.adj_verts_bend[t.bending[4 * b + i]];
                const u32 v2 = (u32)t.bending[4 * b + j];
                if (std::find(l.begin(), l.end(), v2) == l.end()) l.push_back(v2);
            }
    // :581-625 greedy colouring: repeat { marks = copy(taken); for vid asc: if !marks -> take, mark neighbours }
    t.clusters.clear();
    std::vector<bool> taken(nv, false);
    u32 total = 0;
    while (total < nv) {
        std::vector<u32> cur;
        std::vector<bool> marked(taken);
        for (u32 v = 0; v < nv; v++) {
            if (marked[v]) continue;
            taken[v] = true;
            cur.push_back(v);
            marked[v] = true;
            for (u32 a : t.adj_verts_bend[v]) marked[a] = true;
        }
        total += (u32)cur.size();
        t.clusters.push_back(cur);
    }
}

// example2_async_iter.cpp:99-127 (upload_2d_csr_from): offsets start at n+1
static std::vector<u32> packed_csr(const std::vector<std::vector<u32>>& m) {
    const u32 n = (u32)m.size();
    std::vector<u32> out(n + 1);
    u32 p = n + 1;
    for (u32 i = 0; i < n; i++) { out[i] = p; p += (u32)m[i].size(); }
    out[n] = p;
    for (u32 i = 0; i < n; i++) out.insert(out.end(), m[i].begin(), m[i].end());
    return out;
}

// ---------------------------------------------------------------------------
// Rest state.  example2_async_iter.cpp:308-316 (rest length), 344-374 (Q).
// ---------------------------------------------------------------------------
template <typename T> struct Rest {
    std::vector<T> rest_len;   // E
    std::vector<T> Q;          // 16B, Q[b*16 + col*4 + row]
};

template <typename T>
static void build_rest(const Topology& t, const std::vector<V3<T>>& x, Rest<T>& r) {
    const u32 ne = (u32)t.edges.size() / 2, nb = (u32)t.bending.size() / 4;
    r.rest_len.resize(ne);
    for (u32 e = 0; e < ne; e++) r.rest_len[e] = length(x[t.edges[2 * e]] - x[t.edges[2 * e + 1]]);
    r.Q.resize((size_t)nb * 16);
    auto cot = [](V3<T> a, V3<T> b) { return dot(a, b) / length(cross(a, b)); };
    for (u32 b = 0; b < nb; b++) {
        const V3<T> p0 = x[t.bending[4 * b + 0]], p1 = x[t.bending[4 * b + 1]];
        const V3<T> p2 = x[t.bending[4 * b + 2]], p3 = x[t.bending[4 * b + 3]];
        const V3<T> e0 = p1 - p0, e1 = p2 - p0, e2 = p3 - p0, e3 = p2 - p1, e4 = p3 - p1;
        const T c01 = cot(e0, -e1), c02 = cot(e0, -e2), c03 = cot(e0, e3), c04 = cot(e0, e4);
        const T K[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
        const T A0 = T(0.5f) * length(cross(e0, e1));
        const T A1 = T(0.5f) * length(cross(e0, e2));
        const T s = T(3.f) / (A0 + A1);
        for (int col = 0; col < 4; col++)
            for (int row = 0; row < 4; row++) r.Q[(size_t)b * 16 + col * 4 + row] = s * (K[row] * K[col]);
    }
}

// ---------------------------------------------------------------------------
// Simulation state and the VBD frame.  example2_async_iter.cpp:1082-1143.
// ---------------------------------------------------------------------------
struct Params {
    float implicit_dt;     // scene_params.h:94 (1/60)
    u32 num_substep;       // scene_params.h:68
    u32 iter_count;        // constraint_iter_count
    float k_stretch;       // stiffness_stretch_spring
    float k_bending;       // get_stiffness_quadratic_bending() (already multiplied by bending_ui)
    float damping;         // damping_cloth
    float gravity[3];      // scene gravity used by inertia + energy (predict hard-codes (0,-9.8,0))
};

template <typename T> struct Sim {
    Topology* topo;
    Rest<T> rest;
    Params p;
    std::vector<T> mass;
    std::vector<unsigned char> fixed;
    std::vector<V3<T>> x, v, x_start, v_start, x_frame_start, v_frame_start;
    std::vector<Hf<T>> hf;
    V3<T> wind{T(0), T(0), T(0)};
    bool energy_bending = false;   // CPU reference drops bending energy (example2_async_iter.cpp:832-846)
    std::vector<float> energy;
    int nthreads = 1;
    // obstacle contact (row a10): static triangle mesh, per-frame detection
    std::vector<V3<T>> obs_x;
    std::vector<int> obs_f;
    std::vector<V3<T>> obs_n;              // face normals
    T thickness = T(0), stiffness = T(0), radius = T(0);
    int max_contacts = 0;
    struct Contact { int fid; V3<T> xs, n; };
    std::vector<std::vector<Contact>> contacts;  // per vertex, ascending fid
    // exact prefilter of the face loop (same contacts, see detect_contacts):
    // per face centroid and circumradius about it, the obstacle's box (FP64)
    bool prefilter = true;
    std::vector<double> obs_c, obs_r;
    double obs_lo[3] = {0, 0, 0}, obs_hi[3] = {0, 0, 0};
    // asynchronous iteration: 32 iterate + 32 begin position buffers,
    // indexed modulo 32 (example2_async_iter.cpp:1147-1157)
    std::vector<V3<T>> iter_buf[32], begin_buf[32];
    // XPBD (example2_async_iter.cpp:453-697): constraint clusters, the
    // constraints in cluster order ("merged"), their multipliers, 1/m
    bool xpbd_ready = false;
    std::vector<std::vector<u32>> cl_stretch, cl_bending;
    std::vector<int> m_edges, m_bending;
    std::vector<T> m_rest, m_Q, lambda_s, lambda_b, mass_inv;
};

template <typename T> static inline T substep_dt(const Params& p) {
    // scene_params.h:159: implicit_dt / float(num_substep)
    return T(p.implicit_dt) / T((float)p.num_substep);
}

// glm::normalize(v) = v * inversesqrt(dot(v, v)), inversesqrt(x) = 1 / sqrt(x)
template <typename T> static inline V3<T> normalize(V3<T> v) { return v * (T(1) / std::sqrt(dot(v, v))); }

// xpbd_constraints.h:2698-2718
template <typename T> static inline T point_line_distance(V3<T> v0, V3<T> v1, V3<T> v2, V3<T> e1hat, T projection) {
    const V3<T> e0 = v0 - v1;
    const V3<T> e1 = v2 - v1;
    if ((double)projection > 0.0 && projection < length(e1)) return length(e0 - projection * e1hat);
    const T d01 = length(v0 - v1);
    const T d02 = length(v0 - v2);
    return (d01 < d02) ? d01 : d02;
}
// xpbd_constraints.h:2724-2740
template <typename T> static inline void get_lerp(V3<T> v1, V3<T> v2, T projection, T out[2]) {
    const V3<T> e1 = v2 - v1;
    if ((double)projection < 0.0) { out[0] = T(1); out[1] = T(0); return; }
    if (projection >= length(e1)) { out[0] = T(0); out[1] = T(1); return; }
    const T ratio = projection / length(e1);
    out[0] = T(1.0f) - ratio;
    out[1] = ratio;
}
// NarrowPhase::template_vf<false> xpbd_constraints.h:2742-2820: closest point
// of v[0] on triangle (v[1], v[2], v[3]) as barycentric weights, with the
// reference's edge fallback when the projection leaves the triangle.
template <typename T> static bool template_vf(const V3<T> v[4], T min_distance, T& distance, T bary[3]) {
    const V3<T> e1 = v[2] - v[1], e2 = v[3] - v[2], e3 = v[1] - v[3];
    const V3<T> n = cross(e1, e2);
    const V3<T> na = cross(v[3] - v[2], v[0] - v[2]);
    const V3<T> nb = cross(v[1] - v[3], v[0] - v[3]);
    const V3<T> nc = cross(v[2] - v[1], v[0] - v[1]);
    bary[0] = dot(n, na) / dot(n, n);
    bary[1] = dot(n, nb) / dot(n, n);
    bary[2] = dot(n, nc) / dot(n, n);
    const T bsum = (std::fabs(bary[0]) + std::fabs(bary[1])) + std::fabs(bary[2]);
    if ((double)bsum - 1.0 > 1e-8) {
        const V3<T> h12 = normalize(e1), h23 = normalize(e2), h31 = normalize(e3);
        const T p01 = dot(v[0] - v[1], h12), p02 = dot(v[0] - v[2], h23), p03 = dot(v[0] - v[3], h31);
        const T d12 = point_line_distance(v[0], v[1], v[2], h12, p01);
        const T d23 = point_line_distance(v[0], v[2], v[3], h23, p02);
        const T d31 = point_line_distance(v[0], v[3], v[1], h31, p03);
        T l[2];
        if (d12 <= d23 && d12 <= d31) { get_lerp(v[1], v[2], p01, l); bary[0] = l[0]; bary[1] = l[1]; bary[2] = T(0.0f); }
        else if (d23 <= d12 && d23 <= d31) { get_lerp(v[2], v[3], p02, l); bary[0] = T(0.0f); bary[1] = l[0]; bary[2] = l[1]; }
        else { get_lerp(v[3], v[1], p03, l); bary[0] = l[1]; bary[1] = T(0.0f); bary[2] = l[0]; }
    }
    const V3<T> xs = (bary[0] * v[1] + bary[1] * v[2]) + bary[2] * v[3];
    distance = length(v[0] - xs);
    return (double)distance < (double)min_distance && (double)distance > 1e-8;
}

// Contact detection after predict: every obstacle face within `radius` of
// the predicted position (template_vf), the max_contacts nearest by
// (distance, face id), kept in ascending face id.  Brute force over faces,
// with an exact prefilter: template_vf's closest point lies on the triangle
// (the edge fallback clamps to the edge; the projection is used only with
// |b0| + |b1| + |b2| <= 1 + 1e-8), so a face it accepts has its centroid
// within radius + circumradius of the vertex (triangle inequality); faces
// farther than that plus a slack of 1 % (far above FP32 rounding of the
// distance) are skipped, as is every face for a vertex farther than the
// radius plus the slack from the obstacle's box.  The skipped faces are ones
// template_vf rejects, so the contacts are those of the plain loop (a CPU
// test compares both on random scenes).
template <typename T> static void detect_contacts(Sim<T>& s) {
    const long nv = (long)s.x.size();
    const int nf = (int)s.obs_f.size() / 3;
    s.contacts.assign(nv, {});
    double rmax = 0.0;
    for (int f = 0; f < nf; f++) rmax = std::max(rmax, s.obs_r[f]);
    const double rad = (double)s.radius, box_reach = 1.01 * (rad + rmax) + 1e-6;
#pragma omp parallel for schedule(dynamic, 64) num_threads(s.nthreads)
    for (long vid = 0; vid < nv; vid++) {
        struct Cand { T d; int fid; T b[3]; };
        std::vector<Cand> keep;
        const double p[3] = {(double)s.x[vid].x, (double)s.x[vid].y, (double)s.x[vid].z};
        if (s.prefilter) {
            bool near = true;
            for (int d = 0; d < 3; d++) near = near && p[d] >= s.obs_lo[d] - box_reach && p[d] <= s.obs_hi[d] + box_reach;
            if (!near) continue;
        }
        for (int f = 0; f < nf; f++) {
            if (s.prefilter) {
                const double dx = p[0] - s.obs_c[3 * f], dy = p[1] - s.obs_c[3 * f + 1], dz = p[2] - s.obs_c[3 * f + 2];
                const double reach = 1.01 * (rad + s.obs_r[f]) + 1e-6;
                if (dx * dx + dy * dy + dz * dz > reach * reach) continue;
            }
            const V3<T> v[4] = {s.x[vid], s.obs_x[s.obs_f[3 * f]], s.obs_x[s.obs_f[3 * f + 1]], s.obs_x[s.obs_f[3 * f + 2]]};
            T d, b[3];
            if (!template_vf(v, s.radius, d, b)) continue;
            keep.push_back(Cand{d, f, {b[0], b[1], b[2]}});
        }
        std::sort(keep.begin(), keep.end(), [](const Cand& a, const Cand& c) { return a.d != c.d ? a.d < c.d : a.fid < c.fid; });
        if ((int)keep.size() > s.max_contacts) keep.resize(s.max_contacts);
        std::sort(keep.begin(), keep.end(), [](const Cand& a, const Cand& c) { return a.fid < c.fid; });
        for (const Cand& c : keep) {
            const V3<T> p1 = s.obs_x[s.obs_f[3 * c.fid]], p2 = s.obs_x[s.obs_f[3 * c.fid + 1]], p3 = s.obs_x[s.obs_f[3 * c.fid + 2]];
            const V3<T> xs = (c.b[0] * p1 + c.b[1] * p2) + c.b[2] * p3;
            s.contacts[vid].push_back(typename Sim<T>::Contact{c.fid, xs, s.obs_n[c.fid]});
        }
    }
}

// VBD::compute_obstacle_collision xpbd_constraints.h:4238-4284 (stiff =
// stiffness * area, area = 1): f = (stiff C) n, H[col j][row i] = n_i (stiff n_j)
template <typename T> static inline Hf<T> compute_contact(const Sim<T>& s, u32 vid) {
    Hf<T> acc = hf_zero<T>();
    if (s.contacts.empty()) return acc;
    for (const auto& c : s.contacts[vid]) {
        const V3<T> diff = s.x[vid] - c.xs;
        const T proj = dot(diff, c.n);
        if (!(proj < s.thickness)) continue;
        const T C = s.thickness - proj;
        const V3<T> f = (s.stiffness * C) * c.n;
        const V3<T> sn = s.stiffness * c.n;
        Hf<T> hc;
        for (int col = 0; col < 3; col++)
            for (int row = 0; row < 3; row++) hc.h[col][row] = c.n[row] * sn[col];
        for (int r = 0; r < 3; r++) hc.f[r] = f[r];
        hf_add(acc, hc);
    }
    return acc;
}

// xpbd_constraints.h:321-384
template <typename T> static void predict_position(Sim<T>& s) {
    const T h = substep_dt<T>(s.p);
    const V3<T> g = mk(T(0), T(-9.8f), T(0));
    const long nv = (long)s.x.size();
#pragma omp parallel for schedule(static, 256) num_threads(s.nthreads)
    for (long vid = 0; vid < nv; vid++) {
        const V3<T> x_prev = s.x_start[vid];
        const V3<T> v_prev = s.v[vid];
        V3<T> v_pred = v_prev + h * g;
        if (s.fixed[vid]) v_pred = mk(T(0), T(0), T(0));
        s.x[vid] = x_prev + h * v_pred;
    }
}

// xpbd_constraints.h:535-581
template <typename T> static void update_velocity(Sim<T>& s) {
    const T h = substep_dt<T>(s.p);
    const T damp = T(std::exp(T(-s.p.damping) * h));
    const long nv = (long)s.x.size();
#pragma omp parallel for schedule(static, 256) num_threads(s.nthreads)
    for (long vid = 0; vid < nv; vid++) {
        const V3<T> dx = s.x[vid] - s.x_start[vid];
        V3<T> vel = dx / h;
        vel = vel * damp;
        s.v[vid] = vel;
        s.v_start[vid] = vel;
        s.x_start[vid] = s.x[vid];
    }
}

// VBD::compute_inertia xpbd_constraints.h:4027-4067, extended with a uniform
// per-vertex external force `wind` added to the outer force after m*g.
template <typename T> static inline Hf<T> compute_inertia(const Sim<T>& s, u32 vid, T h) {
    const T h2inv = T(1.f) / (h * h);
    const V3<T> xk = s.x[vid], x0 = s.x_start[vid], v0 = s.v[vid];
    const T m = s.mass[vid];
    const V3<T> g = mk(T(s.p.gravity[0]), T(s.p.gravity[1]), T(s.p.gravity[2]));
    V3<T> of = mk(T(0), T(0), T(0));
    of = of + m * g;
    of = of + s.wind;
    Hf<T> r = hf_zero<T>();
    const T md = (T(1) * m) * h2inv;          // Identity3x3 * mass * h_2_inv
    for (int c = 0; c < 3; c++) r.h[c][c] = md;
    if (s.fixed[vid])
        for (int c = 0; c < 3; c++) r.h[c][c] = r.h[c][c] + T(1E9f);
    const T a = (-m) * h2inv;
    const V3<T> d = (xk - x0) - (v0 * h);
    r.f[0] = a * d.x + of.x;
    r.f[1] = a * d.y + of.y;
    r.f[2] = a * d.z + of.z;
    return r;
}

// VBD::compute_stretch_mass_spring xpbd_constraints.h:4107-4159
template <typename T> static inline Hf<T> compute_stretch(const Sim<T>& s, u32 vid) {
    const T k = T(s.p.k_stretch);
    const T eps = T(1.19209290E-7f);
    Hf<T> acc = hf_zero<T>();
    for (u32 eid : s.topo->adj_edges[vid]) {
        const int e0 = s.topo->edges[2 * eid], e1 = s.topo->edges[2 * eid + 1];
        const V3<T> diff = s.x[e1] - s.x[e0];
        const T l = std::max(length(diff), eps);
        const T L = s.rest.rest_len[eid];
        const T C = l - L;
        const V3<T> dir = diff / l;
        V3<T> f0 = (k * dir) * C;
        const T xinv = T(1.f) / l;
        const T xsq = xinv * xinv;
        const T s1 = k * xsq;
        const T s2 = k * std::max(T(1) - L * xinv, T(0.0f));
        Hf<T> he;
        for (int c = 0; c < 3; c++)
            for (int r = 0; r < 3; r++) {
                const T xx = diff[r] * diff[c];             // outer_product: col c = diff * diff[c]
                const T id = (r == c) ? T(1) : T(0);
                he.h[c][r] = s1 * xx + s2 * (id - xsq * xx);
            }
        const V3<T> f = (vid == (u32)e0) ? f0 : -f0;
        for (int r = 0; r < 3; r++) he.f[r] = f[r];
        hf_add(acc, he);
    }
    return acc;
}

// VBD::compute_bending_quadratic xpbd_constraints.h:4161-4205
template <typename T> static inline Hf<T> compute_bending(const Sim<T>& s, u32 vid) {
    const T k = T(s.p.k_bending);
    Hf<T> acc = hf_zero<T>();
    for (u32 b : s.topo->adj_bending[vid]) {
        const int* e = &s.topo->bending[4 * b];
        const T* Q = &s.rest.Q[(size_t)b * 16];
        int o = 0;
        while ((u32)e[o] != vid) o++;
        V3<T> f = mk(T(0), T(0), T(0));
        for (int jj = 0; jj < 4; jj++) f = f - Q[o * 4 + jj] * s.x[e[jj]];
        f = k * f;
        const T hd = k * Q[o * 4 + o];
        Hf<T> hb = hf_zero<T>();
        for (int c = 0; c < 3; c++) {
            for (int r = 0; r < 3; r++) hb.h[c][r] = hd * ((r == c) ? T(1) : T(0));
            hb.f[c] = f[c];
        }
        hf_add(acc, hb);
    }
    return acc;
}

// vbd_step example2_async_iter.cpp:993-1014, determinant (glm), inverse_mat(H, det)
// float_n_n.h:342-363, Epsilon constant_value.h:9
template <typename T> static inline void vbd_step(Sim<T>& s, u32 vid) {
    const Hf<T>& a = s.hf[vid];
    const T (*m)[3] = a.h;
    const T det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2])
                - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2])
                + m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    if (std::fabs(det) > T(1.19209290E-7f)) {
        auto d2 = [](T x00, T x01, T x10, T x11) { return x00 * x11 - x01 * x10; };
        T inv[3][3];
        inv[0][0] = +(d2(m[1][1], m[1][2], m[2][1], m[2][2])) / det;
        inv[0][1] = -(d2(m[1][0], m[1][2], m[2][0], m[2][2])) / det;
        inv[0][2] = +(d2(m[1][0], m[1][1], m[2][0], m[2][1])) / det;
        inv[1][1] = +(d2(m[0][0], m[0][2], m[2][0], m[2][2])) / det;
        inv[1][2] = -(d2(m[0][0], m[0][1], m[2][0], m[2][1])) / det;
        inv[2][2] = +(d2(m[0][0], m[0][1], m[1][0], m[1][1])) / det;
        inv[1][0] = -(d2(m[0][1], m[0][2], m[2][1], m[2][2])) / det;
        inv[2][0] = +(d2(m[0][1], m[0][2], m[1][1], m[1][2])) / det;
        inv[2][1] = -(d2(m[0][0], m[0][2], m[1][0], m[1][2])) / det;
        V3<T> dx;
        for (int i = 0; i < 3; i++) dx[i] = (inv[0][i] * a.f[0] + inv[1][i] * a.f[1]) + inv[2][i] * a.f[2];
        s.x[vid] = s.x[vid] + dx;
    }
}

// CpuSolver::compute_energy example2_async_iter.cpp:785-880 with
// Energy::compute_energy_inertia xpbd_constraints.h:2079-2098,
// compute_energy_stretch_mass_spring :2100-2125, compute_energy_bending :2129-2214
// (core_bending_quadratic :1062-1083).  Per-element FP32 terms, double sum,
// stored as float.  Bending is added only when energy_bending is set (the CPU
// reference discards it; its GPU energy kernels include it).
template <typename T> static double compute_energy(const Sim<T>& s) {
    const T h = T(s.p.implicit_dt) / T((float)s.p.num_substep);
    const long nv = (long)s.x.size();
    double ei = 0.0, es = 0.0, eb = 0.0;
#pragma omp parallel for reduction(+ : ei) num_threads(s.nthreads)
    for (long vid = 0; vid < nv; vid++) {
        const T m = s.mass[vid];
        V3<T> aext = mk(T(s.p.gravity[0]), T(s.p.gravity[1]), T(s.p.gravity[2]));
        aext = aext + s.wind / m;
        if (s.fixed[vid]) aext = mk(T(0), T(0), T(0));
        const V3<T> y = (s.x_start[vid] + h * s.v_start[vid]) + (h * h) * aext;
        const V3<T> d = s.x[vid] - y;
        const T e = (dot(d, d) * m) / ((T(2) * h) * h);
        ei += (double)e;
    }
    const long ne = (long)s.rest.rest_len.size();
    const T k = T(s.p.k_stretch);
#pragma omp parallel for reduction(+ : es) num_threads(s.nthreads)
    for (long eid = 0; eid < ne; eid++) {
        const V3<T> diff = s.x[s.topo->edges[2 * eid + 1]] - s.x[s.topo->edges[2 * eid]];
        const T l2 = dot(diff, diff);
        if ((double)l2 < 1e-8) continue;    // float promoted against a double literal
        const T C = std::sqrt(l2) - s.rest.rest_len[eid];
        es += (double)((T(0.5f) * k) * (C * C));
    }
    if (s.energy_bending) {
        const long nb = (long)s.topo->bending.size() / 4;
        const T kb = T(s.p.k_bending);
#pragma omp parallel for reduction(+ : eb) num_threads(s.nthreads)
        for (long b = 0; b < nb; b++) {
            const int* e = &s.topo->bending[4 * b];
            const T* Q = &s.rest.Q[(size_t)b * 16];
            T sum = T(0);
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) sum = sum + Q[i * 4 + j] * dot(s.x[e[i]], s.x[e[j]]);
            eb += (double)((T(0.5f) * kb) * sum);
        }
    }
    double ec = 0.0;
    if (!s.contacts.empty()) {
        for (long vid = 0; vid < nv; vid++)
            for (const auto& c : s.contacts[vid]) {
                const T proj = dot(s.x[vid] - c.xs, c.n);
                if (proj < s.thickness) {
                    const T C = s.thickness - proj;
                    ec += (double)((T(0.5f) * s.stiffness) * (C * C));
                }
            }
    }
    return ei + es + eb + ec;
}

// solve_constraints_VBD example2_async_iter.cpp:1481-1509; per-colour passes
// :939-1014 (inertia grain 256, stretch/bending/step grain 32, cpu_parallel.h:83-98)
template <typename T> static void solve_vbd_iteration(Sim<T>& s, bool record, bool first) {
    const T h = substep_dt<T>(s.p);
    if (record && first) s.energy.push_back((float)compute_energy(s));
    for (const auto& cl : s.topo->clusters) {
        const long n = (long)cl.size();
#pragma omp parallel for schedule(static, 256) num_threads(s.nthreads)
        for (long i = 0; i < n; i++) s.hf[cl[i]] = compute_inertia(s, cl[i], h);
#pragma omp parallel for schedule(static, 32) num_threads(s.nthreads)
        for (long i = 0; i < n; i++) { Hf<T> a = compute_stretch(s, cl[i]); hf_add(s.hf[cl[i]], a); }
#pragma omp parallel for schedule(static, 32) num_threads(s.nthreads)
        for (long i = 0; i < n; i++) { Hf<T> a = compute_bending(s, cl[i]); hf_add(s.hf[cl[i]], a); }
        if (!s.contacts.empty()) {
#pragma omp parallel for schedule(static, 32) num_threads(s.nthreads)
            for (long i = 0; i < n; i++) { Hf<T> a = compute_contact(s, cl[i]); hf_add(s.hf[cl[i]], a); }
        }
#pragma omp parallel for schedule(static, 32) num_threads(s.nthreads)
        for (long i = 0; i < n; i++) vbd_step(s, cl[i]);
    }
    if (record) s.energy.push_back((float)compute_energy(s));
}

// SolverInterface::physics_step :1615-1648 + CpuSolver::physics_step_vbd :1082-1143
template <typename T> static void physics_step(Sim<T>& s, bool record) {
    s.x_frame_start = s.x;   // frame_end -> frame_start (x and v always hold frame end)
    s.v_frame_start = s.v;
    s.x_start = s.x_frame_start;
    s.v_start = s.v_frame_start;
    s.energy.clear();
    const u32 nsub = record ? 1u : s.p.num_substep;
    for (u32 sub = 0; sub < nsub; sub++) {
        predict_position(s);
        if (!s.obs_f.empty()) detect_contacts(s);
        for (u32 it = 0; it < s.p.iter_count; it++) solve_vbd_iteration(s, record, it == 0);
        update_velocity(s);
    }
}

// ---------------------------------------------------------------------------
// XPBD.  example2_async_iter.cpp:453-697 (constraint
// colouring, merged arrays), 698-710 (reset), 884-932 (per-cluster solves),
// 1017-1081 (frame), 1510-1534 (iteration); xpbd_constraints.h:984-1059
// (solve_stretch_mass_spring_template), 1062-1126 (core_bending_quadratic),
// 1170-1230 (solve_bending_quadratic_template), 211-236 (get_delta_xpbd_4).
// ---------------------------------------------------------------------------

// fn_graph_coloring_sequenced_constraint (example2_async_iter.cpp:458-533):
// repeat { marks = copy(taken); for every untaken constraint in id order
// whose mark is clear: take it, mark every constraint touching one of its
// vertices }.
static std::vector<std::vector<u32>> colour_constraints(u32 n, const std::vector<int>& elems, int per,
                                                         const std::vector<std::vector<u32>>& vert_adj) {
    std::vector<std::vector<u32>> out;
    std::vector<bool> taken(n, false);
    u32 total = 0;
    while (total < n) {
        std::vector<u32> cur;
        std::vector<bool> marked(taken);
        for (u32 id = 0; id < n; id++) {
            if (marked[id]) continue;
            taken[id] = true;
            cur.push_back(id);
            marked[id] = true;
            for (int j = 0; j < per; j++)
                for (u32 adj : vert_adj[elems[(size_t)per * id + j]]) marked[adj] = true;
        }
        total += (u32)cur.size();
        out.push_back(cur);
    }
    return out;
}

template <typename T> static void build_xpbd(Sim<T>& s) {
    if (s.xpbd_ready) return;
    const Topology& t = *s.topo;
    const u32 ne = (u32)t.edges.size() / 2, nb = (u32)t.bending.size() / 4;
    s.cl_stretch = colour_constraints(ne, t.edges, 2, t.adj_edges);
    s.cl_bending = colour_constraints(nb, t.bending, 4, t.adj_bending);
    s.m_edges.clear();
    s.m_rest.clear();
    for (const auto& c : s.cl_stretch)
        for (u32 e : c) {
            s.m_edges.push_back(t.edges[2 * e]);
            s.m_edges.push_back(t.edges[2 * e + 1]);
            s.m_rest.push_back(s.rest.rest_len[e]);
        }
    s.m_bending.clear();
    s.m_Q.clear();
    for (const auto& c : s.cl_bending)
        for (u32 b : c) {
            for (int j = 0; j < 4; j++) s.m_bending.push_back(t.bending[4 * b + j]);
            for (int j = 0; j < 16; j++) s.m_Q.push_back(s.rest.Q[(size_t)b * 16 + j]);
        }
    s.lambda_s.assign(ne, T(0));
    s.lambda_b.assign(nb, T(0));
    // example2_async_iter.cpp:223: 0 for pinned vertices, else 1/m
    s.mass_inv.resize(s.mass.size());
    for (size_t v = 0; v < s.mass.size(); v++) s.mass_inv[v] = s.fixed[v] ? T(0.0f) : T(1.0f) / s.mass[v];
    s.xpbd_ready = true;
}

// solve_stretch_mass_spring_template for merged constraint i
template <typename T> static void xpbd_stretch(Sim<T>& s, u32 i) {
    const int a = s.m_edges[2 * i], b = s.m_edges[2 * i + 1];
    const T lambda = s.lambda_s[i];
    const V3<T> p0 = s.x[a], p1 = s.x[b];
    const T w0 = s.mass_inv[a], w1 = s.mass_inv[b];
    const V3<T> diff = p1 - p0;
    const T lsq = dot(diff, diff);
    if ((double)lsq < 1e-8) return;
    const T l = std::sqrt(lsq);
    const T C = l - s.m_rest[i];
    const V3<T> dir = diff / l;
    const V3<T> g1 = -dir;
    const T k = T(s.p.k_stretch), h = substep_dt<T>(s.p);
    const T alpha = k < T(1.19209290E-7f) ? T(1e9f) : T(1.f) / (k * h * h);
    const T denom = alpha + (w0 + w1);
    const T dl = (-C - alpha * lambda) / denom;
    V3<T> dx0 = mk(T(0), T(0), T(0)), dx1 = dx0;
    if (denom != T(0.f)) {
        dx0 = (dl * w0) * g1;
        dx1 = ((-dl) * w1) * g1;
    }
    s.lambda_s[i] = lambda + dl;
    s.x[a] = s.x[a] + dx0;
    s.x[b] = s.x[b] + dx1;
}

// solve_bending_quadratic_template for merged constraint i (C = the pair's
// bending energy at unit stiffness, gradient Q x per vertex)
template <typename T> static void xpbd_bending(Sim<T>& s, u32 i) {
    const int* e = &s.m_bending[4 * (size_t)i];
    const T* Q = &s.m_Q[16 * (size_t)i];   // Q[col * 4 + row]
    T w[4];
    V3<T> p[4];
    for (int j = 0; j < 4; j++) {
        w[j] = s.mass_inv[e[j]];
        p[j] = s.x[e[j]];
    }
    const T lambda = s.lambda_b[i];
    T sum = T(0.f);
    for (int ii = 0; ii < 4; ii++)
        for (int jj = 0; jj < 4; jj++) sum = sum + Q[ii * 4 + jj] * dot(p[ii], p[jj]);
    const T C = (T(0.5f) * T(1.0f)) * sum;
    V3<T> g[4];
    for (int ii = 0; ii < 4; ii++) {
        V3<T> f = mk(T(0), T(0), T(0));
        for (int jj = 0; jj < 4; jj++) f = f - Q[ii * 4 + jj] * p[jj];
        f = T(1.0f) * f;
        g[ii] = -f;
    }
    const T kb = T(s.p.k_bending), h = substep_dt<T>(s.p);
    const T alpha = kb < T(1.19209290E-7f) ? T(1e12f) : T(1.f) / (kb * h * h);
    const T denom = w[0] * dot(g[0], g[0]) + w[1] * dot(g[1], g[1]) + w[2] * dot(g[2], g[2]) +
                    w[3] * dot(g[3], g[3]) + alpha;
    const T dl = (-C - alpha * lambda) / denom;
    V3<T> dx[4];
    for (int j = 0; j < 4; j++) dx[j] = mk(T(0), T(0), T(0));
    if (denom != T(0.f))
        for (int j = 0; j < 4; j++) dx[j] = (dl * w[j]) * g[j];
    s.lambda_b[i] = lambda + dl;
    for (int j = 0; j < 4; j++) s.x[e[j]] = s.x[e[j]] + dx[j];
}

// CpuSolver::physics_step_xpbd (example2_async_iter.cpp:1017-1081)
template <typename T> static void physics_step_xpbd(Sim<T>& s, bool record) {
    build_xpbd(s);
    s.x_frame_start = s.x;
    s.v_frame_start = s.v;
    s.x_start = s.x_frame_start;
    s.v_start = s.v_frame_start;
    s.energy.clear();
    const u32 nsub = record ? 1u : s.p.num_substep;
    for (u32 sub = 0; sub < nsub; sub++) {
        predict_position(s);
        std::fill(s.lambda_s.begin(), s.lambda_s.end(), T(0.0f));
        std::fill(s.lambda_b.begin(), s.lambda_b.end(), T(0.0f));
        for (u32 it = 0; it < s.p.iter_count; it++) {
            if (record && it == 0) s.energy.push_back((float)compute_energy(s));
            u32 base = 0;
            for (const auto& c : s.cl_stretch) {
                for (u32 k = 0; k < (u32)c.size(); k++) xpbd_stretch(s, base + k);
                base += (u32)c.size();
            }
            base = 0;
            for (const auto& c : s.cl_bending) {
                for (u32 k = 0; k < (u32)c.size(); k++) xpbd_bending(s, base + k);
                base += (u32)c.size();
            }
            if (record) s.energy.push_back((float)compute_energy(s));
        }
        update_velocity(s);
    }
}

// One colour pass (a2-a5) over cluster `c` on the positions in s.x
// (example2_async_iter.cpp:939-1014, as in solve_vbd_iteration).
template <typename T> static void solve_cluster(Sim<T>& s, u32 c) {
    const T h = substep_dt<T>(s.p);
    const auto& cl = s.topo->clusters[c];
    const long n = (long)cl.size();
#pragma omp parallel for schedule(static, 256) num_threads(s.nthreads)
    for (long i = 0; i < n; i++) s.hf[cl[i]] = compute_inertia(s, cl[i], h);
#pragma omp parallel for schedule(static, 32) num_threads(s.nthreads)
    for (long i = 0; i < n; i++) { Hf<T> a = compute_stretch(s, cl[i]); hf_add(s.hf[cl[i]], a); }
#pragma omp parallel for schedule(static, 32) num_threads(s.nthreads)
    for (long i = 0; i < n; i++) { Hf<T> a = compute_bending(s, cl[i]); hf_add(s.hf[cl[i]], a); }
    if (!s.contacts.empty()) {
#pragma omp parallel for schedule(static, 32) num_threads(s.nthreads)
        for (long i = 0; i < n; i++) { Hf<T> a = compute_contact(s, cl[i]); hf_add(s.hf[cl[i]], a); }
    }
#pragma omp parallel for schedule(static, 32) num_threads(s.nthreads)
    for (long i = 0; i < n; i++) vbd_step(s, cl[i]);
}

// Core::read_and_solve_conflict (xpbd_constraints.h:815-869) for one vertex:
// both iterates moved away from the common begin position => begin + w (sum
// of both moves); one moved => that one; written to self and begin.
template <typename T>
static void read_and_solve_conflict(u32 vid, std::vector<V3<T>>& begin, std::vector<V3<T>>& self,
                                    const std::vector<V3<T>>& other, T weight) {
    const V3<T> ps = self[vid], po = other[vid];
    V3<T> w;
    if (weight > T(1.19209290E-7f)) {
        const V3<T> init = begin[vid];
        const V3<T> ds = ps - init, da = po - init;
        w = init;
        const bool act_a = dot(ds, ds) != T(0), act_b = dot(da, da) != T(0);
        if (act_a && act_b) w = init + weight * (ds + da);
        else if (act_a) w = ps;
        else if (act_b) w = po;
    } else {
        w = T(0.5f) * (ps + po);
    }
    self[vid] = w;
    begin[vid] = w;
}

// Function ids of the dispatched tasks (include/libatssim_hip.h LATS_FN_*).
enum : u32 { FN_ROOT = 0, FN_TERMINAL = 1, FN_PREDICT = 2, FN_ALL_IN_ONE = 3, FN_LAST_NODE = 4, FN_UPDATE_VELOCITY = 5,
             FN_RESET = 6 };
constexpr u32 kNoBuffer = 0xFFFFFFFFu, kInputBuffer = 0x0FFFFFFFu;

struct DispatchParam {
    u32 fn, iter, cluster, buffer, left, right, nin;
    const u32* ins;
    bool main;
};

// CpuSolver::fn_dispatch (example2_async_iter.cpp:1144-1304): the task's
// colour pass runs on its iterate buffer between the prev hook (merge the
// inputs with the left buffer, or copy the last input / the predicted
// positions in) and the post hook (copy to the right buffer); energies are
// recorded where the reference's hooks record them.
template <typename T> static void dispatch(Sim<T>& s, const DispatchParam& p, bool record) {
    const size_t nv = s.x.size();
    auto iter = [&](u32 b) -> std::vector<V3<T>>& {
        auto& v = s.iter_buf[b % 32];
        if (v.size() != nv) v.assign(nv, mk(T(0), T(0), T(0)));
        return v;
    };
    auto begin = [&](u32 b) -> std::vector<V3<T>>& {
        auto& v = s.begin_buf[b % 32];
        if (v.size() != nv) v.assign(nv, mk(T(0), T(0), T(0)));
        return v;
    };
    auto copy_to_start_and_iter = [&](const std::vector<V3<T>>& src, u32 out) {
        const std::vector<V3<T>> tmp = src;   // src may be one of the outputs
        begin(out) = tmp;
        iter(out) = tmp;
    };
    auto energy_of = [&](std::vector<V3<T>>& buf) {
        std::swap(s.x, buf);
        s.energy.push_back((float)compute_energy(s));
        std::swap(s.x, buf);
    };
    auto prev = [&]() {
        if (p.nin && p.left != kNoBuffer) {
            for (u32 k = 0; k < p.nin; k++) {
                const u32 in = p.ins[k];
                std::vector<V3<T>>& bb = p.main ? begin(in) : begin(p.left);
                std::vector<V3<T>>& self = iter(in);
                const std::vector<V3<T>>& other = iter(p.left);
                for (size_t v = 0; v < nv; v++) read_and_solve_conflict((u32)v, bb, self, other, T(0.5f));
            }
        } else if (p.nin) {
            copy_to_start_and_iter(iter(p.ins[p.nin - 1]), p.buffer);
        } else if (p.left == kInputBuffer) {
            copy_to_start_and_iter(s.x, p.buffer);
        }
        if (record && p.iter == 0 && p.cluster == 0) energy_of(iter(p.buffer));
    };
    auto post = [&]() {
        if (p.right != kNoBuffer) copy_to_start_and_iter(iter(p.buffer), p.right);
        if (record && ((p.fn == FN_ALL_IN_ONE && p.cluster + 1 == (u32)s.topo->clusters.size()) || p.fn == FN_LAST_NODE))
            energy_of(iter(p.buffer));
    };
    switch (p.fn) {
        case FN_PREDICT: predict_position(s); break;
        case FN_UPDATE_VELOCITY: update_velocity(s); break;
        case FN_ALL_IN_ONE: {
            prev();
            std::vector<V3<T>>& buf = iter(p.buffer);
            std::swap(s.x, buf);
            solve_cluster(s, p.cluster);
            std::swap(s.x, buf);
            post();
            break;
        }
        case FN_LAST_NODE:
            prev();
            post();
            s.x = iter(p.buffer);
            break;
        default: break;
    }
}

// physics_step_vbd_async frame bracket (example2_async_iter.cpp:1305-1313, 1471-1479)
template <typename T> static void async_begin(Sim<T>& s) {
    s.x_frame_start = s.x;
    s.v_frame_start = s.v;
    s.x_start = s.x_frame_start;
    s.v_start = s.v_frame_start;
    s.energy.clear();
}

// Position buffers (float[3V], vertex order): 0 iterate idx, 1 begin idx,
// 2 x, 3 x_start, 4 v (the numbering of LATS_BUF_*).
template <typename T> static std::vector<V3<T>>& buffer_ref(Sim<T>& s, int which, u32 idx) {
    std::vector<V3<T>>* b = &s.x;
    if (which == 0) b = &s.iter_buf[idx % 32];
    else if (which == 1) b = &s.begin_buf[idx % 32];
    else if (which == 3) b = &s.x_start;
    else if (which == 4) b = &s.v;
    if (b->size() != s.x.size()) b->assign(s.x.size(), mk(T(0), T(0), T(0)));
    return *b;
}
template <typename T> static void buffer_get(Sim<T>& s, int which, u32 idx, float* out) {
    const auto& b = buffer_ref(s, which, idx);
    for (size_t i = 0; i < b.size(); i++)
        for (int c = 0; c < 3; c++) out[3 * i + c] = (float)b[i][c];
}
template <typename T> static void buffer_set(Sim<T>& s, int which, u32 idx, const float* in) {
    auto& b = buffer_ref(s, which, idx);
    for (size_t i = 0; i < b.size(); i++)
        for (int c = 0; c < 3; c++) b[i][c] = T(in[3 * i + c]);
}
}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI for the Python tests and the CPU baseline.
// ---------------------------------------------------------------------------
using namespace orc;

struct orc_mesh { Topology t; std::vector<float> x; };
struct orc_sim {
    orc_mesh* mesh;
    bool fp64;
    Sim<float> s32;
    Sim<double> s64;
};

template <typename T>
static void sim_init(Sim<T>& s, orc_mesh* m, const Params& p, const float* mass, const unsigned char* fixed) {
    const u32 nv = m->t.nv;
    s.topo = &m->t;
    s.p = p;
    std::vector<V3<T>> x(nv);
    for (u32 i = 0; i < nv; i++) x[i] = mk(T(m->x[3 * i]), T(m->x[3 * i + 1]), T(m->x[3 * i + 2]));
    build_rest(m->t, x, s.rest);
    s.mass.assign(mass, mass + nv);
    s.fixed.assign(fixed, fixed + nv);
    s.x = x;
    s.v.assign(nv, mk(T(0), T(0), T(0)));
    s.x_start = s.x;
    s.v_start = s.v;
    s.hf.assign(nv, hf_zero<T>());
}

template <typename T> static void get_state(const Sim<T>& s, double* x, double* v) {
    for (size_t i = 0; i < s.x.size(); i++)
        for (int c = 0; c < 3; c++) {
            if (x) x[3 * i + c] = (double)s.x[i][c];
            if (v) v[3 * i + c] = (double)s.v[i][c];
        }
}

template <typename T> static void set_state(Sim<T>& s, const double* x, const double* v) {
    for (size_t i = 0; i < s.x.size(); i++)
        for (int c = 0; c < 3; c++) {
            s.x[i][c] = (T)x[3 * i + c];
            s.v[i][c] = (T)v[3 * i + c];
        }
    s.x_start = s.x;
    s.v_start = s.v;
}

template <typename T>
static void set_obstacle(Sim<T>& s, u32 nv, const float* x, u32 nf, const int* f, float thickness, float stiffness,
                         float radius, int max_contacts) {
    s.obs_x.resize(nv);
    for (u32 i = 0; i < nv; i++) s.obs_x[i] = mk(T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2]));
    s.obs_f.assign(f, f + 3 * (size_t)nf);
    s.obs_n.resize(nf);
    for (u32 k = 0; k < nf; k++) {
        const V3<T> a = s.obs_x[f[3 * k]], b = s.obs_x[f[3 * k + 1]], c = s.obs_x[f[3 * k + 2]];
        s.obs_n[k] = normalize(cross(b - a, c - a));
    }
    s.thickness = T(thickness);
    s.stiffness = T(stiffness);
    s.radius = T(radius);
    s.max_contacts = max_contacts;
    s.obs_c.assign(3 * (size_t)nf, 0.0);
    s.obs_r.assign(nf, 0.0);
    for (int d = 0; d < 3; d++) {
        s.obs_lo[d] = 1e300;
        s.obs_hi[d] = -1e300;
    }
    for (u32 i = 0; i < nv; i++)
        for (int d = 0; d < 3; d++) {
            s.obs_lo[d] = std::min(s.obs_lo[d], (double)x[3 * i + d]);
            s.obs_hi[d] = std::max(s.obs_hi[d], (double)x[3 * i + d]);
        }
    for (u32 k = 0; k < nf; k++) {
        double c[3];
        for (int d = 0; d < 3; d++)
            c[d] = ((double)x[3 * f[3 * k] + d] + (double)x[3 * f[3 * k + 1] + d] + (double)x[3 * f[3 * k + 2] + d]) / 3.0;
        double r2 = 0.0;
        for (int q = 0; q < 3; q++) {
            double e2 = 0.0;
            for (int d = 0; d < 3; d++) e2 += ((double)x[3 * f[3 * k + q] + d] - c[d]) * ((double)x[3 * f[3 * k + q] + d] - c[d]);
            r2 = std::max(r2, e2);
        }
        for (int d = 0; d < 3; d++) s.obs_c[3 * k + d] = c[d];
        s.obs_r[k] = std::sqrt(r2);
    }
}

template <typename T> static u32 contact_list(const Sim<T>& s, u32* counts, int* fids, u32 cap) {
    u32 tot = 0;
    for (size_t v = 0; v < s.contacts.size(); v++) {
        if (counts) counts[v] = (u32)s.contacts[v].size();
        for (const auto& c : s.contacts[v]) {
            if (fids && tot < cap) fids[tot] = c.fid;
            tot++;
        }
    }
    return tot;
}

// Checker support: the positions alone (frame start and velocities kept), and
// per vertex the force vector f of the vertex solve at the current positions
// (inertia + stretch + bending + contact terms, the f vbd_step solves with).
template <typename T> static void set_positions(Sim<T>& s, const double* x) {
    for (size_t i = 0; i < s.x.size(); i++)
        for (int c = 0; c < 3; c++) s.x[i][c] = (T)x[3 * i + c];
}
template <typename T> static void solve_forces(const Sim<T>& s, double* f) {
    const T h = substep_dt<T>(s.p);
    for (u32 vid = 0; vid < (u32)s.x.size(); vid++) {
        Hf<T> a = compute_inertia(s, vid, h);
        Hf<T> b = compute_stretch(s, vid);
        hf_add(a, b);
        b = compute_bending(s, vid);
        hf_add(a, b);
        b = compute_contact(s, vid);
        hf_add(a, b);
        for (int c = 0; c < 3; c++) f[3 * (size_t)vid + c] = (double)a.f[c];
    }
}

extern "C" {

orc_mesh* orc_mesh_create(u32 nv, const float* x, u32 nf, const int* faces) {
    orc_mesh* m = new orc_mesh();
    m->t.nv = nv;
    m->x.assign(x, x + 3 * (size_t)nv);
    m->t.faces.assign(faces, faces + 3 * (size_t)nf);
    extract_edges(m->t.faces, m->t.edges, m->t.bending);
    build_topology(m->t);
    return m;
}
void orc_mesh_destroy(orc_mesh* m) { delete m; }

// counts: nv, ne, nb, nclusters
void orc_mesh_counts(const orc_mesh* m, u32* out) {
    out[0] = m->t.nv;
    out[1] = (u32)m->t.edges.size() / 2;
    out[2] = (u32)m->t.bending.size() / 4;
    out[3] = (u32)m->t.clusters.size();
}
void orc_mesh_edges(const orc_mesh* m, int* out) { std::copy(m->t.edges.begin(), m->t.edges.end(), out); }
void orc_mesh_bending(const orc_mesh* m, int* out) { std::copy(m->t.bending.begin(), m->t.bending.end(), out); }

// which: 0 adj_edges, 1 adj_bending, 2 adj_verts, 3 adj_verts_with_bending, 4 clusters
u32 orc_mesh_csr_size(const orc_mesh* m, int which) {
    const std::vector<std::vector<u32>>* l[5] = {&m->t.adj_edges, &m->t.adj_bending, &m->t.adj_verts,
                                                  &m->t.adj_verts_bend, &m->t.clusters};
    size_t n = l[which]->size() + 1;
    for (auto& v : *l[which]) n += v.size();
    return (u32)n;
}
void orc_mesh_csr(const orc_mesh* m, int which, u32* out) {
    const std::vector<std::vector<u32>>* l[5] = {&m->t.adj_edges, &m->t.adj_bending, &m->t.adj_verts,
                                                  &m->t.adj_verts_bend, &m->t.clusters};
    std::vector<u32> p = packed_csr(*l[which]);
    std::copy(p.begin(), p.end(), out);
}
void orc_mesh_rest(const orc_mesh* m, float* rest_len, float* Q) {
    std::vector<V3<float>> x(m->t.nv);
    for (u32 i = 0; i < m->t.nv; i++) x[i] = mk(m->x[3 * i], m->x[3 * i + 1], m->x[3 * i + 2]);
    Rest<float> r;
    build_rest(m->t, x, r);
    std::copy(r.rest_len.begin(), r.rest_len.end(), rest_len);
    std::copy(r.Q.begin(), r.Q.end(), Q);
}


orc_sim* orc_sim_create(orc_mesh* m, const float* params9, const u32* iparams2, const float* mass,
                        const unsigned char* fixed, int fp64) {
    // params9: implicit_dt, k_stretch, k_bending, damping, g0, g1, g2 (7 used); iparams2: num_substep, iter_count
    Params p;
    p.implicit_dt = params9[0];
    p.k_stretch = params9[1];
    p.k_bending = params9[2];
    p.damping = params9[3];
    p.gravity[0] = params9[4];
    p.gravity[1] = params9[5];
    p.gravity[2] = params9[6];
    p.num_substep = iparams2[0];
    p.iter_count = iparams2[1];
    orc_sim* s = new orc_sim();
    s->mesh = m;
    s->fp64 = fp64 != 0;
    if (s->fp64) sim_init(s->s64, m, p, mass, fixed);
    else sim_init(s->s32, m, p, mass, fixed);
    return s;
}
void orc_sim_destroy(orc_sim* s) { delete s; }

void orc_sim_set_threads(orc_sim* s, int n) { s->s32.nthreads = n; s->s64.nthreads = n; }
void orc_sim_set_energy_bending(orc_sim* s, int on) { s->s32.energy_bending = on != 0; s->s64.energy_bending = on != 0; }
void orc_sim_set_wind(orc_sim* s, const float* w) {
    s->s32.wind = mk(w[0], w[1], w[2]);
    s->s64.wind = mk((double)w[0], (double)w[1], (double)w[2]);
}

// One frame.  record != 0 => one substep with the energy history (101 values).
// Returns the number of energies written (at most cap).
u32 orc_sim_step(orc_sim* s, int record, float* energy, u32 cap) {
    std::vector<float>* e;
    if (s->fp64) { physics_step(s->s64, record != 0); e = &s->s64.energy; }
    else { physics_step(s->s32, record != 0); e = &s->s32.energy; }
    u32 n = std::min<u32>(cap, (u32)e->size());
    if (energy) std::copy(e->begin(), e->begin() + n, energy);
    return n;
}

void orc_sim_get_state(const orc_sim* s, double* x, double* v) {
    if (s->fp64) get_state(s->s64, x, v);
    else get_state(s->s32, x, v);
}
void orc_sim_set_state(orc_sim* s, const double* x, const double* v) {
    if (s->fp64) set_state(s->s64, x, v);
    else set_state(s->s32, x, v);
}
double orc_sim_energy(orc_sim* s) { return s->fp64 ? compute_energy(s->s64) : compute_energy(s->s32); }

void orc_sim_set_positions(orc_sim* s, const double* x) {
    if (s->fp64) set_positions(s->s64, x);
    else set_positions(s->s32, x);
}
void orc_sim_forces(const orc_sim* s, double* f) {
    if (s->fp64) solve_forces(s->s64, f);
    else solve_forces(s->s32, f);
}

// Static triangle obstacle (face normals = normalize(cross(b - a, c - a))).
void orc_sim_set_obstacle(orc_sim* s, u32 nv, const float* x, u32 nf, const int* f, float thickness, float stiffness,
                          float radius, int max_contacts) {
    if (s->fp64) set_obstacle(s->s64, nv, x, nf, f, thickness, stiffness, radius, max_contacts);
    else set_obstacle(s->s32, nv, x, nf, f, thickness, stiffness, radius, max_contacts);
}
// The exact prefilter of the contact query on (1, default) or off (0: the
// plain loop over every face).
void orc_sim_set_contact_prefilter(orc_sim* s, int on) {
    s->s32.prefilter = on != 0;
    s->s64.prefilter = on != 0;
}
// Contacts of the last frame: per vertex count, then (fid) lists; returns total.
u32 orc_sim_contacts(const orc_sim* s, u32* counts, int* fids, u32 cap) {
    return s->fp64 ? contact_list(s->s64, counts, fids, cap) : contact_list(s->s32, counts, fids, cap);
}
// Narrow phase alone (for tests): returns hit, writes distance and barycentric weights.
int orc_template_vf(const float* p4, float min_distance, float* distance, float* bary) {
    V3<float> v[4];
    for (int k = 0; k < 4; k++) v[k] = mk(p4[3 * k], p4[3 * k + 1], p4[3 * k + 2]);
    float d = 0.f, b[3] = {0.f, 0.f, 0.f};
    const bool hit = template_vf(v, min_distance, d, b);
    *distance = d;
    for (int k = 0; k < 3; k++) bary[k] = b[k];
    return hit ? 1 : 0;
}

// Asynchronous frame: orc_sim_async_begin, then the scheduled tasks in launch
// order through orc_sim_dispatch, then the recorded energies.
void orc_sim_async_begin(orc_sim* s) {
    if (s->fp64) async_begin(s->s64);
    else async_begin(s->s32);
}
void orc_sim_dispatch(orc_sim* s, u32 fn, u32 iter, u32 cluster, u32 buffer, u32 left, u32 right, u32 nin,
                      const u32* ins, int main_device, int record) {
    const DispatchParam p{fn, iter, cluster, buffer, left, right, nin, ins, main_device != 0};
    if (s->fp64) dispatch(s->s64, p, record != 0);
    else dispatch(s->s32, p, record != 0);
}
void orc_sim_buffer_get(orc_sim* s, int which, u32 idx, float* out) {
    if (s->fp64) buffer_get(s->s64, which, idx, out);
    else buffer_get(s->s32, which, idx, out);
}
void orc_sim_buffer_set(orc_sim* s, int which, u32 idx, const float* in) {
    if (s->fp64) buffer_set(s->s64, which, idx, in);
    else buffer_set(s->s32, which, idx, in);
}
u32 orc_sim_energies(const orc_sim* s, float* energy, u32 cap) {
    const std::vector<float>& e = s->fp64 ? s->s64.energy : s->s32.energy;
    const u32 n = std::min<u32>(cap, (u32)e.size());
    if (energy) std::copy(e.begin(), e.begin() + n, energy);
    return (u32)e.size();
}

// XPBD frame (record != 0: one substep with the energy history).
u32 orc_sim_step_xpbd(orc_sim* s, int record, float* energy, u32 cap) {
    std::vector<float>* e;
    if (s->fp64) { physics_step_xpbd(s->s64, record != 0); e = &s->s64.energy; }
    else { physics_step_xpbd(s->s32, record != 0); e = &s->s32.energy; }
    const u32 n = std::min<u32>(cap, (u32)e->size());
    if (energy) std::copy(e->begin(), e->begin() + n, energy);
    return n;
}
// XPBD constraint clusters (which: 0 springs, 1 bending pairs) as a packed CSR.
u32 orc_sim_xpbd_clusters(orc_sim* s, int which, u32* out, u32 cap) {
    std::vector<u32> p;
    if (s->fp64) { build_xpbd(s->s64); p = packed_csr(which ? s->s64.cl_bending : s->s64.cl_stretch); }
    else { build_xpbd(s->s32); p = packed_csr(which ? s->s32.cl_bending : s->s32.cl_stretch); }
    if (out) std::copy(p.begin(), p.begin() + std::min<size_t>(cap, p.size()), out);
    return (u32)p.size();
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
