// vbd_common.h -- per-vertex VBD arithmetic shared by every kernel variant.
//
// Each function reproduces the reference's FP32 evaluation order operation by
// operation (glm semantics), so that with -ffp-contract=off and IEEE div/sqrt
// the GPU result is bit-identical to the reference's CPU arithmetic:
//   inertia  VBD::compute_inertia            xpbd_constraints.h:4027-4067
//   stretch  VBD::compute_stretch_mass_spring xpbd_constraints.h:4107-4159
//   bending  VBD::compute_bending_quadratic  xpbd_constraints.h:4161-4205
//   step     vbd_step + inverse_mat(H, det)  example2_async_iter.cpp:993-1014,
//            float_n_n.h:342-363; glm::determinant(mat3); Epsilon constant_value.h:9
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace lats {

#define LATS_HD __host__ __device__ __forceinline__

struct F3 {
    float x, y, z;
};
LATS_HD F3 add3(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
LATS_HD F3 sub3(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
LATS_HD F3 scale3(float s, F3 a) { return F3{s * a.x, s * a.y, s * a.z}; }
// glm dot(vec3): (x*x' + y*y') + z*z'
LATS_HD float dot3(F3 a, F3 b) {
    const float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
    return (px + py) + pz;
}

// Symmetric 3x3 (every Hessian of the path is exactly symmetric: outer
// products d d^T and sums of them are computed identically for (r,c), (c,r)).
struct Sym3 {
    float xx, xy, xz, yy, yz, zz;
};
LATS_HD Sym3 sym_zero() { return Sym3{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}; }
LATS_HD void sym_add(Sym3& a, const Sym3& b) {
    a.xx = a.xx + b.xx;
    a.xy = a.xy + b.xy;
    a.xz = a.xz + b.xz;
    a.yy = a.yy + b.yy;
    a.yz = a.yz + b.yz;
    a.zz = a.zz + b.zz;
}

// --- quotients by one divisor --------------------------------------------
// The IEEE FP32 division the compiler emits for n / d is v_div_scale of d
// and of n, v_rcp, one reciprocal refinement, a product and two fma
// corrections, v_div_fmas, v_div_fixup.  When v_div_scale leaves both
// operands unscaled (it then clears VCC) -- d normal with 1/d normal, n of
// biased exponent at least 24, n / d normal and their exponents less than
// 96 apart -- v_div_fmas is a plain fma and v_div_fixup returns its input,
// so the division is exactly the steps below.  Written out, quotients by
// one divisor share its refined reciprocal.  The operands are checked
// together, by magnitude, against a box inside those conditions:
// 2^-120 <= |d| <= 2^120 and 2^-100, |d| 2^-120 <= |n| <= |d| 2^90; any
// operand outside it (zeros included) sends every quotient of the call
// through the full division.  Either way each result is the correctly
// rounded n / d.
#define LATS_D __device__ __forceinline__
struct Divisor {
    float d, y;   // the divisor and its refined reciprocal
};
LATS_D Divisor divisor(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    return Divisor{d, __builtin_fmaf(__builtin_fmaf(-d, r, 1.0f), r, r)};
}
LATS_D float quotient(float n, Divisor v) {
    const float q0 = n * v.y;
    const float q1 = __builtin_fmaf(__builtin_fmaf(-v.d, q0, n), v.y, q0);
    return __builtin_fmaf(__builtin_fmaf(-v.d, q1, n), v.y, q1);
}
// q[k] = num[k] / den for k < N.
template <int N>
LATS_D void divide_all(const float (&num)[N], float den, float (&q)[N]) {
    const Divisor v = divisor(den);
    float lo = fabsf(num[0]), hi = lo;
#pragma unroll
    for (int k = 0; k < N; k++) {
        q[k] = quotient(num[k], v);
        lo = fminf(lo, fabsf(num[k]));
        hi = fmaxf(hi, fabsf(num[k]));
    }
    const float ad = fabsf(den);
    // (NaN operands fail every comparison)
    const bool plain = ad >= 0x1p-120f && ad <= 0x1p120f && lo >= 0x1p-100f && lo >= ad * 0x1p-120f &&
                       hi <= ad * 0x1p90f;
    if (!plain) {
        // (an empty volatile asm: the block then has a side effect, so it
        // is branched over instead of being computed always and selected)
        asm volatile("");
#pragma unroll
        for (int k = 0; k < N; k++) q[k] = num[k] / den;
    }
}

// H = (m/h^2) I (+1e9 I if pinned);
// f = ((-m) h^-2) ((xk - x0) - v0 h) + ((0 + m g) + ext)
LATS_HD void inertia_term(F3 xk, F3 x0, F3 v0, float m, uint32_t pinned, const float* g, F3 ext, float h,
                          float h2inv, Sym3& H, F3& f) {
    const float md = (1.f * m) * h2inv;
    float d = md;
    if (pinned) d = d + 1E9f;
    H = Sym3{d, 0.f, 0.f, d, 0.f, d};
    const F3 of = add3(add3(F3{0.f, 0.f, 0.f}, F3{m * g[0], m * g[1], m * g[2]}), ext);
    const float a = (-m) * h2inv;
    const F3 dd = sub3(sub3(xk, x0), F3{v0.x * h, v0.y * h, v0.z * h});
    f = F3{a * dd.x + of.x, a * dd.y + of.y, a * dd.z + of.z};
}

// One incident spring.  diff = x_other - x_self equals +/- the reference's
// x[e1] - x[e0] exactly, and the force on the vertex is (k * diff/l) * C in
// both orientations (IEEE negation commutes with rounding).
// The values one incident spring adds to S and f.
LATS_D void stretch_parts(F3 xk, F3 xo, float L, float k, Sym3& t, F3& tf) {
    const F3 diff = sub3(xo, xk);
    const float len = sqrtf(dot3(diff, diff));
    const float eps = 1.19209290E-7f;
    const float l = (len < eps) ? eps : len;
    const float C = l - L;
    // diff.x / l, diff.y / l, diff.z / l, 1 / l
    const float num[4] = {diff.x, diff.y, diff.z, 1.f};
    float q[4];
    divide_all(num, l, q);
    const F3 dir = F3{q[0], q[1], q[2]};
    tf = F3{(k * dir.x) * C, (k * dir.y) * C, (k * dir.z) * C};
    const float xinv = q[3];
    const float xsq = xinv * xinv;
    const float s1 = k * xsq;
    const float r = 1.f - L * xinv;
    const float s2 = k * ((r < 0.0f) ? 0.0f : r);
    const float dxx = diff.x * diff.x, dxy = diff.x * diff.y, dxz = diff.x * diff.z;
    const float dyy = diff.y * diff.y, dyz = diff.y * diff.z, dzz = diff.z * diff.z;
    t.xx = s1 * dxx + s2 * (1.f - xsq * dxx);
    t.xy = s1 * dxy + s2 * (0.f - xsq * dxy);
    t.xz = s1 * dxz + s2 * (0.f - xsq * dxz);
    t.yy = s1 * dyy + s2 * (1.f - xsq * dyy);
    t.yz = s1 * dyz + s2 * (0.f - xsq * dyz);
    t.zz = s1 * dzz + s2 * (1.f - xsq * dzz);
}
LATS_D void stretch_term(F3 xk, F3 xo, float L, float k, Sym3& S, F3& fs) {
    Sym3 t;
    F3 tf;
    stretch_parts(xk, xo, L, k, t, tf);
    sym_add(S, t);
    fs = add3(fs, tf);
}

// One incident bending pair: f = k * (-(sum_j Q[o][j] x_j)), H += k Q[o][o] I.
// q = row o of Q in pair-local order (e0, e1, opp0, opp1).
LATS_HD F3 bending_force(float4 q, F3 p0, F3 p1, F3 p2, F3 p3, float k) {
    F3 f = F3{0.f, 0.f, 0.f};
    f = sub3(f, scale3(q.x, p0));
    f = sub3(f, scale3(q.y, p1));
    f = sub3(f, scale3(q.z, p2));
    f = sub3(f, scale3(q.w, p3));
    return scale3(k, f);
}
// The pair's terms added to the sums hb and fb.
LATS_HD void bending_term(float4 q, uint32_t o, F3 p0, F3 p1, F3 p2, F3 p3, float k, float& hb, F3& fb) {
    const F3 f = bending_force(q, p0, p1, p2, p3, k);
    const float qoo = (o == 0) ? q.x : (o == 1) ? q.y : (o == 2) ? q.z : q.w;
    hb = hb + k * qoo;
    fb = add3(fb, f);
}
// The same terms on their own (th, tf), with Q[o][o] selected without
// branches.  (Measured: the lane-group update is faster with the selects,
// the one-lane update with the compiler's branches of bending_term.)
LATS_HD void bending_parts(float4 q, uint32_t o, F3 p0, F3 p1, F3 p2, F3 p3, float k, float& th, F3& tf) {
    tf = bending_force(q, p0, p1, p2, p3, k);
    const float q01 = (o & 1u) ? q.y : q.x, q23 = (o & 1u) ? q.w : q.z;
    th = k * ((o & 2u) ? q23 : q01);
}

// vbd_step: det(H) (glm), |det| > Epsilon ? dx = adj(H)/det * f : no move.
LATS_D bool newton_step(const Sym3& H, F3 f, F3& dx) {
    // m[col][row]
    const float m00 = H.xx, m01 = H.xy, m02 = H.xz;
    const float m10 = H.xy, m11 = H.yy, m12 = H.yz;
    const float m20 = H.xz, m21 = H.yz, m22 = H.zz;
    const float det = m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02) + m20 * (m01 * m12 - m11 * m02);
    if (!(fabsf(det) > 1.19209290E-7f)) return false;
    // the adjugate's entries (negated where the inverse negates them), each / det
    const float num[9] = {(m11 * m22 - m12 * m21), -(m10 * m22 - m12 * m20), (m10 * m21 - m11 * m20),
                          (m00 * m22 - m02 * m20), -(m00 * m21 - m01 * m20), (m00 * m11 - m01 * m10),
                          -(m01 * m22 - m02 * m21), (m01 * m12 - m02 * m11), -(m00 * m12 - m02 * m10)};
    float q[9];
    divide_all(num, det, q);
    const float i00 = q[0], i01 = q[1], i02 = q[2], i11 = q[3], i12 = q[4], i22 = q[5];
    const float i10 = q[6], i20 = q[7], i21 = q[8];
    dx.x = (i00 * f.x + i10 * f.y) + i20 * f.z;
    dx.y = (i01 * f.x + i11 * f.y) + i21 * f.z;
    dx.z = (i02 * f.x + i12 * f.y) + i22 * f.z;
    return true;
}

}  // namespace lats
