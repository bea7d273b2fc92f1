// scene.cpp -- see scene.h.
#include "scene.h"

#include <cmath>

namespace lats {

void make_grid(uint32_t nx, uint32_t ny, float spacing, float jitter, uint64_t seed, std::vector<float>& x,
               std::vector<int32_t>& faces) {
    x.resize(3 * (size_t)nx * ny);
    SplitMix64 rng(seed);
    const double s = spacing;
    for (uint32_t r = 0; r < ny; r++)
        for (uint32_t c = 0; c < nx; c++) {
            const size_t v = (size_t)r * nx + c;
            const double u = rng.uniform();
            x[3 * v + 0] = (float)(((double)c - 0.5 * (double)(nx - 1)) * s);
            x[3 * v + 1] = (float)(-(double)r * s);
            x[3 * v + 2] = (float)((2.0 * u - 1.0) * (double)jitter);
        }
    faces.clear();
    faces.reserve(6 * (size_t)(nx - 1) * (ny - 1));
    for (uint32_t r = 0; r + 1 < ny; r++)
        for (uint32_t c = 0; c + 1 < nx; c++) {
            const int32_t a = (int32_t)(r * nx + c), b = a + 1, d = (int32_t)((r + 1) * nx + c), e = d + 1;
            if (((r + c) & 1) == 0) {
                faces.insert(faces.end(), {a, d, e, a, e, b});
            } else {
                faces.insert(faces.end(), {a, d, b, b, d, e});
            }
        }
}

std::vector<float> make_wind(uint32_t num_frames, uint32_t num_verts, float max_total, uint64_t seed) {
    std::vector<float> w(3 * (size_t)num_frames);
    SplitMix64 rng(seed);
    const double two_pi = 6.283185307179586;
    for (uint32_t f = 0; f < num_frames; f++) {
        const double mag = rng.uniform() * (double)max_total;
        const double th = rng.uniform() * two_pi;
        const double per = mag / (double)(num_verts ? num_verts : 1);
        w[3 * f + 0] = (float)(per * std::cos(th));
        w[3 * f + 1] = 0.0f;
        w[3 * f + 2] = (float)(per * std::sin(th));
    }
    return w;
}

}  // namespace lats
