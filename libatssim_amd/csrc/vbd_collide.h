// vbd_collide.h -- obstacle contact on the device: the vertex-face narrow
// phase and the per-vertex contact term, in the reference's FP32 operation
// order (glm semantics, -ffp-contract=off):
//   template_vf<false>         xpbd_constraints.h:2742-2820 (+ pointLineDistance
//                              :2698-2718, getLerp :2724-2740, glm::normalize)
//   compute_obstacle_collision xpbd_constraints.h:4238-4284
//   compute_energy_collision_vf xpbd_constraints.h:2286-2324
#pragma once
#include <hip/hip_runtime.h>

#include "vbd_common.h"

namespace lats {

constexpr int kMaxContacts = 4;

LATS_HD F3 cross3d(F3 x, F3 y) { return F3{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y}; }
LATS_HD float len3d(F3 a) { return sqrtf(dot3(a, a)); }
LATS_HD F3 normalize3d(F3 v) {
    const float s = 1.f / sqrtf(dot3(v, v));
    return F3{v.x * s, v.y * s, v.z * s};
}

LATS_HD float point_line_distance(F3 v0, F3 v1, F3 v2, F3 e1hat, float projection) {
    const F3 e0 = sub3(v0, v1);
    const F3 e1 = sub3(v2, v1);
    if ((double)projection > 0.0 && projection < len3d(e1)) return len3d(sub3(e0, scale3(projection, e1hat)));
    const float d01 = len3d(sub3(v0, v1));
    const float d02 = len3d(sub3(v0, v2));
    return (d01 < d02) ? d01 : d02;
}

LATS_HD void get_lerp(F3 v1, F3 v2, float projection, float& l0, float& l1) {
    const F3 e1 = sub3(v2, v1);
    if ((double)projection < 0.0) { l0 = 1.f; l1 = 0.f; return; }
    if (projection >= len3d(e1)) { l0 = 0.f; l1 = 1.f; return; }
    const float ratio = projection / len3d(e1);
    l0 = 1.0f - ratio;
    l1 = ratio;
}

// Closest point of p on triangle (a, b, c): barycentric weights and distance;
// hit when 1e-8 < distance < min_distance.
LATS_HD bool template_vf(F3 p, F3 a, F3 b, F3 c, float min_distance, float& distance, float bary[3]) {
    const F3 e1 = sub3(b, a), e2 = sub3(c, b), e3 = sub3(a, c);
    const F3 n = cross3d(e1, e2);
    const F3 na = cross3d(sub3(c, b), sub3(p, b));
    const F3 nb = cross3d(sub3(a, c), sub3(p, c));
    const F3 nc = cross3d(sub3(b, a), sub3(p, a));
    bary[0] = dot3(n, na) / dot3(n, n);
    bary[1] = dot3(n, nb) / dot3(n, n);
    bary[2] = dot3(n, nc) / dot3(n, n);
    const float bsum = (fabsf(bary[0]) + fabsf(bary[1])) + fabsf(bary[2]);
    if ((double)bsum - 1.0 > 1e-8) {
        const F3 h12 = normalize3d(e1), h23 = normalize3d(e2), h31 = normalize3d(e3);
        const float p01 = dot3(sub3(p, a), h12), p02 = dot3(sub3(p, b), h23), p03 = dot3(sub3(p, c), h31);
        const float d12 = point_line_distance(p, a, b, h12, p01);
        const float d23 = point_line_distance(p, b, c, h23, p02);
        const float d31 = point_line_distance(p, c, a, h31, p03);
        float l0, l1;
        if (d12 <= d23 && d12 <= d31) { get_lerp(a, b, p01, l0, l1); bary[0] = l0; bary[1] = l1; bary[2] = 0.0f; }
        else if (d23 <= d12 && d23 <= d31) { get_lerp(b, c, p02, l0, l1); bary[0] = 0.0f; bary[1] = l0; bary[2] = l1; }
        else { get_lerp(c, a, p03, l0, l1); bary[0] = l1; bary[1] = 0.0f; bary[2] = l0; }
    }
    const F3 xs = add3(add3(scale3(bary[0], a), scale3(bary[1], b)), scale3(bary[2], c));
    distance = len3d(sub3(p, xs));
    return (double)distance < (double)min_distance && (double)distance > 1e-8;
}

// One obstacle contact: f = (stiff C) n, H[col j][row i] += n_i (stiff n_j),
// active while (x - xs) . n < thickness.
LATS_HD void contact_term(F3 x, F3 xs, F3 n, float stiff, float thickness, float H[3][3], F3& f) {
    const float proj = dot3(sub3(x, xs), n);
    if (!(proj < thickness)) return;
    const float C = thickness - proj;
    const float sc = stiff * C;
    f = add3(f, F3{sc * n.x, sc * n.y, sc * n.z});
    const float sn[3] = {stiff * n.x, stiff * n.y, stiff * n.z};
    const float nn[3] = {n.x, n.y, n.z};
    for (int col = 0; col < 3; col++)
        for (int row = 0; row < 3; row++) H[col][row] = H[col][row] + nn[row] * sn[col];
}

LATS_HD float contact_energy(F3 x, F3 xs, F3 n, float stiff, float thickness) {
    const float proj = dot3(sub3(x, xs), n);
    if (!(proj < thickness)) return 0.f;
    const float C = thickness - proj;
    return (0.5f * stiff) * (C * C);
}

// vbd_step on a general 3x3 (contacts make H non-symmetric by rounding).
LATS_D bool newton_step_full(const float m[3][3], F3 f, F3& dx) {
    const float det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                      m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    if (!(fabsf(det) > 1.19209290E-7f)) return false;
    const float num[9] = {(m[1][1] * m[2][2] - m[1][2] * m[2][1]), -(m[1][0] * m[2][2] - m[1][2] * m[2][0]),
                          (m[1][0] * m[2][1] - m[1][1] * m[2][0]), (m[0][0] * m[2][2] - m[0][2] * m[2][0]),
                          -(m[0][0] * m[2][1] - m[0][1] * m[2][0]), (m[0][0] * m[1][1] - m[0][1] * m[1][0]),
                          -(m[0][1] * m[2][2] - m[0][2] * m[2][1]), (m[0][1] * m[1][2] - m[0][2] * m[1][1]),
                          -(m[0][0] * m[1][2] - m[0][2] * m[1][0])};
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
