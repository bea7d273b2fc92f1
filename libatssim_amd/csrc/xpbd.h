// xpbd.h -- the XPBD cloth solver behind the same context and DAG API as the
// VBD path: mass-spring distance constraints and quadratic bending
// constraints with compliance, projected cluster by cluster (Gauss-Seidel
// over constraint colours, every constraint of a cluster in parallel: no two
// share a vertex).
//
// Reference: examples/example2_async_iter.cpp:453-533 (constraint colouring),
// 638-697 (constraints in cluster order), 698-710 (multipliers reset),
// 884-932 (per-cluster solves), 1017-1081 (frame), 1510-1534 (iteration:
// every spring cluster, then every bending cluster);
// src/Solver/xpbd_constraints.h:984-1059 (solve_stretch_mass_spring_template),
// 1062-1126 (core_bending_quadratic), 1170-1230
// (solve_bending_quadratic_template), 211-236 (get_delta_xpbd_4).
// Included by vbd_device.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "vbd_common.h"
#include "vbd_kernels.h"

namespace lats {

struct XpbdArgs {
    float* x;                 // positions (slot order)
    const float* minv;        // 0 for pinned vertices, else 1/m
    const uint2* edges;       // springs in cluster order (slots)
    const float* rest;
    float* lambda_s;
    const uint4* pairs;       // bending pairs in cluster order (slots)
    const float4* Q;          // 4 float4 per pair: column c holds Q[c][0..3]
    float* lambda_b;
    float ks, kb, h;
};

// One distance constraint (solve_stretch_mass_spring_template).
__device__ __forceinline__ void xpbd_stretch(const XpbdArgs& a, uint32_t i) {
    const uint2 e = a.edges[i];
    const float lambda = a.lambda_s[i];
    const F3 p0 = ld3(a.x, e.x), p1 = ld3(a.x, e.y);
    const float w0 = a.minv[e.x], w1 = a.minv[e.y];
    const F3 diff = sub3(p1, p0);
    const float lsq = dot3(diff, diff);
    if ((double)lsq < 1e-8) return;   // a float compared with a double literal
    const float l = sqrtf(lsq);
    const float C = l - a.rest[i];
    const F3 dir = F3{diff.x / l, diff.y / l, diff.z / l};
    const F3 g1 = F3{-dir.x, -dir.y, -dir.z};
    const float alpha = a.ks < 1.19209290E-7f ? 1e9f : 1.f / ((a.ks * a.h) * a.h);
    const float denom = alpha + (w0 + w1);
    const float dl = (-C - alpha * lambda) / denom;
    F3 dx0 = F3{0.f, 0.f, 0.f}, dx1 = F3{0.f, 0.f, 0.f};
    if (denom != 0.f) {
        dx0 = scale3(dl * w0, g1);
        dx1 = scale3((-dl) * w1, g1);
    }
    a.lambda_s[i] = lambda + dl;
    st3(a.x, e.x, add3(p0, dx0));
    st3(a.x, e.y, add3(p1, dx1));
}

// One quadratic bending constraint: C = the pair's bending energy at unit
// stiffness, 1/2 sum_ij Q_ij <x_i, x_j>, gradient (Q x)_i per vertex.
__device__ __forceinline__ void xpbd_bending(const XpbdArgs& a, uint32_t i) {
    const uint4 e = a.pairs[i];
    const uint32_t id[4] = {e.x, e.y, e.z, e.w};
    float w[4];
    F3 p[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        w[j] = a.minv[id[j]];
        p[j] = ld3(a.x, id[j]);
    }
    float q[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float4 col = a.Q[4 * (size_t)i + c];
        q[c][0] = col.x;
        q[c][1] = col.y;
        q[c][2] = col.z;
        q[c][3] = col.w;
    }
    const float lambda = a.lambda_b[i];
    float sum = 0.f;
#pragma unroll
    for (int ii = 0; ii < 4; ii++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) sum = sum + q[ii][jj] * dot3(p[ii], p[jj]);
    const float C = (0.5f * 1.0f) * sum;
    F3 g[4];
#pragma unroll
    for (int ii = 0; ii < 4; ii++) {
        F3 f = F3{0.f, 0.f, 0.f};
#pragma unroll
        for (int jj = 0; jj < 4; jj++) f = sub3(f, scale3(q[ii][jj], p[jj]));
        f = scale3(1.0f, f);
        g[ii] = F3{-f.x, -f.y, -f.z};
    }
    const float alpha = a.kb < 1.19209290E-7f ? 1e12f : 1.f / ((a.kb * a.h) * a.h);
    const float denom = w[0] * dot3(g[0], g[0]) + w[1] * dot3(g[1], g[1]) + w[2] * dot3(g[2], g[2]) +
                        w[3] * dot3(g[3], g[3]) + alpha;
    const float dl = (-C - alpha * lambda) / denom;
    a.lambda_b[i] = lambda + dl;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        F3 dx = F3{0.f, 0.f, 0.f};
        if (denom != 0.f) dx = scale3(dl * w[j], g[j]);
        st3(a.x, id[j], add3(p[j], dx));
    }
}

// Greedy constraint colouring (fn_graph_coloring_sequenced_constraint):
// repeat { marks = copy(taken); for every constraint in id order whose mark
// is clear: take it and mark every constraint that touches one of its
// vertices }.  `elems` holds `per` vertex indices per constraint; returns
// the clusters in order, each in ascending constraint id.
inline std::vector<std::vector<uint32_t>> colour_constraints(const std::vector<uint32_t>& elems, uint32_t per,
                                                              uint32_t nverts) {
    const uint32_t n = per ? (uint32_t)(elems.size() / per) : 0u;
    std::vector<std::vector<uint32_t>> touching(nverts);
    for (uint32_t k = 0; k < n; k++)
        for (uint32_t j = 0; j < per; j++) touching[elems[(size_t)per * k + j]].push_back(k);
    std::vector<std::vector<uint32_t>> out;
    std::vector<char> taken(n, 0);
    uint32_t total = 0;
    while (total < n) {
        std::vector<uint32_t> cur;
        std::vector<char> marked(taken);
        for (uint32_t k = 0; k < n; k++) {
            if (marked[k]) continue;
            taken[k] = 1;
            cur.push_back(k);
            for (uint32_t j = 0; j < per; j++)
                for (uint32_t other : touching[elems[(size_t)per * k + j]]) marked[other] = 1;
        }
        total += (uint32_t)cur.size();
        out.push_back(std::move(cur));
    }
    return out;
}

}  // namespace lats
