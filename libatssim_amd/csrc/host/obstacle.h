// obstacle.h -- static triangle obstacles and their spatial hash (host, C++).
//
// The hash follows the reference's SpatialHashing (xpbd_constraints.h:
// 2333-2393): integer cell coordinates floor((p - origin) / cell), key
// (ix << 20) | (iy << 10) | iz, slot = |int64(ix*92837111) ^ int64(iy*689287499)
// ^ int64(iz*283923481)| mod table size, linear probing over 32 slots.  The
// reference hashes vertices only; here every obstacle triangle is entered
// into each cell overlapped by its bounding box grown by the query radius,
// so that a cloth vertex finds every triangle within that radius in the
// single cell that contains it.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lats {

struct ObstacleHash {
    float origin[3] = {0, 0, 0};
    float inv_cell = 0.f;
    int dims[3] = {0, 0, 0};          // cells per axis (< 1024)
    uint32_t table_size = 0;
    std::vector<uint32_t> belongs;    // key per slot, 0xffffffff = empty
    std::vector<uint32_t> start;      // first entry of the slot's face list
    std::vector<uint32_t> count;      // faces in the slot
    std::vector<uint32_t> faces;      // face ids, ascending within a cell
};

// the reference's fn_hashCoord with 32-bit wrapping products
inline uint32_t hash_coord(int x, int y, int z, uint32_t table_size) {
    const int32_t hx = (int32_t)((uint32_t)x * 92837111u);
    const int32_t hy = (int32_t)((uint32_t)y * 689287499u);
    const int32_t hz = (int32_t)((uint32_t)z * 283923481u);
    int64_t v = (int64_t)hx ^ (int64_t)hy ^ (int64_t)hz;
    if (v < 0) v = -v;
    return (uint32_t)(v % (int64_t)table_size);
}
inline uint32_t cell_key(int x, int y, int z) { return ((uint32_t)x << 20) | ((uint32_t)y << 10) | (uint32_t)z; }

bool build_obstacle_hash(const float* x, uint32_t nv, const int32_t* f, uint32_t nf, float radius, float cell,
                         ObstacleHash& h, std::string& err);

// normalize(cross(b - a, c - a)) per face, glm operation order
std::vector<float> face_normals(const float* x, const int32_t* f, uint32_t nf);

// Icosahedron subdivided `level` times, projected on the sphere (centre,
// radius); faces counter-clockwise seen from outside.
void make_icosphere(uint32_t level, float radius, const float centre[3], std::vector<float>& x,
                    std::vector<int32_t>& faces);

}  // namespace lats
