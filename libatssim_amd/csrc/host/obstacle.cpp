// obstacle.cpp -- see obstacle.h.
#include "obstacle.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

namespace lats {

std::vector<float> face_normals(const float* x, const int32_t* f, uint32_t nf) {
    std::vector<float> n(3 * (size_t)nf);
    for (uint32_t k = 0; k < nf; k++) {
        const float* a = x + 3 * (size_t)f[3 * k];
        const float* b = x + 3 * (size_t)f[3 * k + 1];
        const float* c = x + 3 * (size_t)f[3 * k + 2];
        const float u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const float v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const float cr[3] = {u[1] * v[2] - v[1] * u[2], u[2] * v[0] - v[2] * u[0], u[0] * v[1] - v[0] * u[1]};
        const float px = cr[0] * cr[0], py = cr[1] * cr[1], pz = cr[2] * cr[2];
        const float inv = 1.f / std::sqrt((px + py) + pz);
        for (int d = 0; d < 3; d++) n[3 * (size_t)k + d] = cr[d] * inv;
    }
    return n;
}

bool build_obstacle_hash(const float* x, uint32_t nv, const int32_t* f, uint32_t nf, float radius, float cell,
                         ObstacleHash& h, std::string& err) {
    if (!(cell > 0.f) || !(radius >= 0.f)) { err = "cell must be > 0 and radius >= 0"; return false; }
    for (size_t i = 0; i < 3 * (size_t)nf; i++)
        if (f[i] < 0 || (uint32_t)f[i] >= nv) { err = "obstacle face index out of range"; return false; }
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t i = 0; i < nv; i++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], x[3 * (size_t)i + d]);
            hi[d] = std::max(hi[d], x[3 * (size_t)i + d]);
        }
    // grow by the radius plus a margin that covers rounding of the cell math
    const float grow = radius * 1.001f + 1e-6f;
    h = ObstacleHash();
    h.inv_cell = 1.f / cell;
    for (int d = 0; d < 3; d++) {
        h.origin[d] = lo[d] - grow - cell;
        h.dims[d] = (int)std::floor((hi[d] + grow + cell - h.origin[d]) * h.inv_cell) + 1;
        if (h.dims[d] >= 1024) { err = "obstacle spans more than 1023 cells per axis: use a larger cell"; return false; }
    }
    auto coord = [&](float p, int d) { return (int)std::floor((p - h.origin[d]) * h.inv_cell); };
    // every face goes to each cell its bounding box grown by the radius
    // touches, so the cell of a point holds every face within the radius
    // of it: counted on a dense grid over the grown bounds, then filled in
    // face order (two passes, no per-cell containers)
    std::vector<int> range(6 * (size_t)nf);
    for (uint32_t k = 0; k < nf; k++) {
        float a[3] = {3.4e38f, 3.4e38f, 3.4e38f}, b[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        for (int j = 0; j < 3; j++)
            for (int d = 0; d < 3; d++) {
                a[d] = std::min(a[d], x[3 * (size_t)f[3 * k + j] + d]);
                b[d] = std::max(b[d], x[3 * (size_t)f[3 * k + j] + d]);
            }
        for (int d = 0; d < 3; d++) {
            range[6 * (size_t)k + d] = std::max(0, coord(a[d] - grow, d));
            range[6 * (size_t)k + 3 + d] = std::min(h.dims[d] - 1, coord(b[d] + grow, d));
        }
    }
    const size_t ncell = (size_t)h.dims[0] * h.dims[1] * h.dims[2];
    if (ncell > (size_t)1 << 28) { err = "obstacle grid too fine: use a larger cell"; return false; }
    auto dense = [&](int ix, int iy, int iz) { return ((size_t)ix * h.dims[1] + iy) * h.dims[2] + iz; };
    auto each_cell = [&](uint32_t k, auto&& fn) {
        const int* r = &range[6 * (size_t)k];
        for (int ix = r[0]; ix <= r[3]; ix++)
            for (int iy = r[1]; iy <= r[4]; iy++)
                for (int iz = r[2]; iz <= r[5]; iz++) fn(ix, iy, iz);
    };
    std::vector<uint32_t> cnt(ncell + 1, 0u);
    uint64_t total = 0;
    for (uint32_t k = 0; k < nf; k++)
        each_cell(k, [&](int ix, int iy, int iz) {
            cnt[dense(ix, iy, iz)]++;
            total++;
        });
    if (total >= (1ull << 32)) { err = "obstacle hash too large: use a larger cell"; return false; }
    std::vector<uint32_t> off(ncell + 1, 0u);
    uint32_t occupied = 0;
    for (size_t c = 0; c < ncell; c++) {
        off[c + 1] = off[c] + cnt[c];
        occupied += cnt[c] != 0u;
    }
    std::vector<uint32_t> faces(total), fill(off.begin(), off.end() - 1);
    for (uint32_t k = 0; k < nf; k++) each_cell(k, [&](int ix, int iy, int iz) { faces[fill[dense(ix, iy, iz)]++] = k; });
    for (uint32_t T = std::max<uint32_t>(64, 4 * occupied);; T *= 2) {
        h.table_size = T;
        h.belongs.assign(T, 0xffffffffu);
        h.start.assign(T, 0);
        h.count.assign(T, 0);
        bool ok = true;
        for (int ix = 0; ix < h.dims[0] && ok; ix++)
            for (int iy = 0; iy < h.dims[1] && ok; iy++)
                for (int iz = 0; iz < h.dims[2]; iz++) {
                    const size_t c = dense(ix, iy, iz);
                    if (!cnt[c]) continue;
                    uint32_t s = hash_coord(ix, iy, iz, T);
                    int j = 0;
                    while (j < 32 && h.belongs[s] != 0xffffffffu) { s = (s + 1) % T; j++; }
                    if (j == 32) { ok = false; break; }
                    h.belongs[s] = cell_key(ix, iy, iz);
                    h.start[s] = off[c];
                    h.count[s] = cnt[c];
                }
        if (ok) {
            h.faces.swap(faces);
            return true;
        }
        if (T > (1u << 28)) { err = "obstacle hash does not fit"; return false; }
    }
}

void make_icosphere(uint32_t level, float radius, const float centre[3], std::vector<float>& x,
                    std::vector<int32_t>& faces) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<double> p = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                             0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    std::vector<int32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2,
                              10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5,
                              2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    auto norm = [&](size_t i) {
        const double l = std::sqrt(p[3 * i] * p[3 * i] + p[3 * i + 1] * p[3 * i + 1] + p[3 * i + 2] * p[3 * i + 2]);
        for (int d = 0; d < 3; d++) p[3 * i + d] /= l;
    };
    for (size_t i = 0; i < p.size() / 3; i++) norm(i);
    for (uint32_t l = 0; l < level; l++) {
        std::map<std::pair<int32_t, int32_t>, int32_t> mid;
        auto midpoint = [&](int32_t a, int32_t b) {
            auto k = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(k);
            if (it != mid.end()) return it->second;
            const int32_t id = (int32_t)(p.size() / 3);
            for (int d = 0; d < 3; d++) p.push_back(0.5 * (p[3 * (size_t)a + d] + p[3 * (size_t)b + d]));
            norm((size_t)id);
            mid[k] = id;
            return id;
        };
        std::vector<int32_t> nf;
        for (size_t k = 0; k < f.size() / 3; k++) {
            const int32_t a = f[3 * k], b = f[3 * k + 1], c = f[3 * k + 2];
            const int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        f.swap(nf);
    }
    x.resize(p.size());
    for (size_t i = 0; i < p.size() / 3; i++)
        for (int d = 0; d < 3; d++) x[3 * i + d] = (float)(p[3 * i + d] * radius + centre[d]);
    faces = f;
}

}  // namespace lats
