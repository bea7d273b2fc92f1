// scene.h -- seeded synthetic cloth scenes (host, C++).
//
// The random streams use splitmix64 (Steele, Lea, Flood 2014) so that every
// consumer (tests, the benchmark, fixtures) regenerates identical inputs from
// a seed without depending on a standard library's distributions.
#pragma once
#include <cstdint>
#include <vector>

namespace lats {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform double in [0, 1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Regular nx x ny vertex grid in the xy plane, `spacing` apart, row 0 on top
// (y = 0, rows go to -y), x centred on 0.  Each quad is split along a
// diagonal that alternates in a checkerboard: quads with (row+col) even use
// the (r,c)-(r+1,c+1) diagonal, odd ones the (r,c+1)-(r+1,c) diagonal.
// z = jitter * (2u - 1) with u drawn per vertex (vertex-id order) from
// splitmix64(seed).
void make_grid(uint32_t nx, uint32_t ny, float spacing, float jitter, uint64_t seed, std::vector<float>& x,
               std::vector<int32_t>& faces);

// Per-frame wind: total force magnitude U(0, max_total) N, direction angle
// U(0, 2 pi) in the xz plane, spread evenly over `num_verts` vertices.
// Returns 3 floats per frame (force per vertex).
std::vector<float> make_wind(uint32_t num_frames, uint32_t num_verts, float max_total, uint64_t seed);

}  // namespace lats
