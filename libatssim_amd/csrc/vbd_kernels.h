// vbd_kernels.h -- device-side pieces shared by the single-cloth and the
// batched VBD kernels: frame parameters, the argument block of one cloth and
// the fused per-vertex update.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "vbd_common.h"

namespace lats {

// One colour of one cloth: slots [start, start + count), ELL bases.
struct ColourInfo {
    uint32_t start, count, eoff, boff;
};

struct FrameParams {
    const float* forces;   // per-frame external force (3 floats / frame) or null
    uint32_t frame;        // frame counter, advanced by the last kernel of a frame
    uint32_t pad;
    float ext[4];          // constant external force when forces == null
    float ext_cur[4];      // force of the running frame (written by predict)
};

struct VbdArgs {
    float* x;                  // positions being iterated (slot order)
    const float* x0;           // x_start
    const float* v;            // frame-start velocity
    const float* mass;
    const uint32_t* meta;
    const uint2* ell_e;
    const float4* ell_bq;
    const uint4* ell_bi;
    const FrameParams* fp;
    float h, h2inv, ks, kb;
    float g[3];
};

__device__ __forceinline__ F3 ld3(const float* p, uint32_t s) { return F3{p[3 * s], p[3 * s + 1], p[3 * s + 2]}; }
__device__ __forceinline__ void st3(float* p, uint32_t s, F3 a) {
    p[3 * s] = a.x;
    p[3 * s + 1] = a.y;
    p[3 * s + 2] = a.z;
}

// One VBD vertex update: VBD::compute_inertia + compute_stretch_mass_spring +
// compute_bending_quadratic (xpbd_constraints.h:4027-4205) summed as
// Hf = inertia; Hf += sum(stretch); Hf += sum(bending)
// (example2_async_iter.cpp:939-992), then the guarded 3x3 Newton step of
// vbd_step (:993-1014).  i = index inside the colour, n = colour size.
__device__ __forceinline__ void vbd_vertex_ext(const VbdArgs& a, uint32_t s, uint32_t i, uint32_t n, uint32_t eoff,
                                               uint32_t boff, F3 ext) {
    const F3 xk = ld3(a.x, s);
    const F3 x0 = ld3(a.x0, s);
    const F3 v0 = ld3(a.v, s);
    const float m = a.mass[s];
    const uint32_t meta = a.meta[s];
    const uint32_t edeg = meta & 0xffu, bdeg = (meta >> 8) & 0xffu;

    // inertia
    Sym3 H;
    F3 f;
    inertia_term(xk, x0, v0, m, (meta >> 16) & 1u, a.g, ext, a.h, a.h2inv, H, f);

    // stretch: local sum from zero, then added to Hf
    Sym3 S = sym_zero();
    F3 fs = F3{0.f, 0.f, 0.f};
    for (uint32_t j = 0; j < edeg; j++) {
        const uint2 e = a.ell_e[eoff + j * n + i];
        const F3 xo = ld3(a.x, e.x);
        stretch_term(xk, xo, __uint_as_float(e.y), a.ks, S, fs);
    }
    sym_add(H, S);
    f = add3(f, fs);

    // bending: diagonal Hessian only (k * Q[o][o] * I)
    float hb = 0.f;
    F3 fb = F3{0.f, 0.f, 0.f};
    for (uint32_t j = 0; j < bdeg; j++) {
        const uint32_t k = boff + j * n + i;
        const float4 q = a.ell_bq[k];
        const uint4 id = a.ell_bi[k];
        const uint32_t o = id.x >> 30;
        const F3 p0 = (o == 0) ? xk : ld3(a.x, id.x & 0x3fffffffu);
        const F3 p1 = (o == 1) ? xk : ld3(a.x, id.y);
        const F3 p2 = (o == 2) ? xk : ld3(a.x, id.z);
        const F3 p3 = (o == 3) ? xk : ld3(a.x, id.w);
        bending_term(q, o, p0, p1, p2, p3, a.kb, hb, fb);
    }
    H.xx = H.xx + hb;
    H.yy = H.yy + hb;
    H.zz = H.zz + hb;
    f = add3(f, fb);

    F3 dx;
    if (newton_step(H, f, dx)) st3(a.x, s, add3(xk, dx));
}


}  // namespace lats
