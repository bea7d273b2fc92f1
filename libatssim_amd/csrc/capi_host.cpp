// capi_host.cpp -- C ABI of the host-only parts: topology and synthetic scenes.
#include <cstring>
#include <string>

#include "../../include/libatssim_hip.h"
#include "host/errors.h"
#include "host/mesh_io.h"
#include "host/obstacle.h"
#include "host/scene.h"
#include "host/topology.h"

struct lats_topology {
    lats::Topology t;
};

struct lats_obj {
    lats::ObjMesh m;
};

namespace {
int set_err(int code, const char* msg) {
    lats::set_error(msg);
    return code;
}
}  // namespace

extern "C" {

int lats_obj_read(const char* path, lats_obj** out) {
    if (!path || !out) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    auto* o = new lats_obj();
    std::string err;
    if (!lats::read_obj(path, o->m, err)) {
        delete o;
        *out = nullptr;
        return set_err(LATS_ERR_INVALID_ARGUMENT, err.c_str());
    }
    *out = o;
    return LATS_OK;
}

int lats_obj_parse(const char* text, lats_obj** out) {
    if (!text || !out) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    auto* o = new lats_obj();
    std::string err;
    if (!lats::parse_obj(text, o->m, err)) {
        delete o;
        *out = nullptr;
        return set_err(LATS_ERR_INVALID_ARGUMENT, err.c_str());
    }
    *out = o;
    return LATS_OK;
}

int lats_obj_counts(const lats_obj* o, uint32_t counts[3]) {
    if (!o || !counts) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    counts[0] = (uint32_t)(o->m.x.size() / 3);
    counts[1] = (uint32_t)(o->m.faces.size() / 3);
    counts[2] = o->m.dropped;
    return LATS_OK;
}

int lats_obj_copy(const lats_obj* o, float* x, int32_t* faces) {
    if (!o) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (x) std::memcpy(x, o->m.x.data(), o->m.x.size() * sizeof(float));
    if (faces) std::memcpy(faces, o->m.faces.data(), o->m.faces.size() * sizeof(int32_t));
    return LATS_OK;
}

void lats_obj_destroy(lats_obj* o) { delete o; }

int lats_obj_write(const char* path, const float* x, uint32_t nv, const int32_t* faces, uint32_t nf, uint32_t mesh_id) {
    if (!path || (nv && !x) || (nf && !faces)) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    for (uint32_t k = 0; k < 3 * nf; k++)
        if (faces[k] < 0 || (uint32_t)faces[k] >= nv) return set_err(LATS_ERR_INVALID_ARGUMENT, "face index out of range");
    std::string err;
    if (!lats::write_obj(path, x, nv, faces, nf, mesh_id, err)) return set_err(LATS_ERR_INVALID_ARGUMENT, err.c_str());
    return LATS_OK;
}

int lats_topology_build(const float* x, uint32_t nv, const int32_t* faces, uint32_t nf, lats_topology** out) {
    if (!out || (nv && !x) || (nf && !faces)) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    auto* t = new lats_topology();
    std::string err;
    if (!lats::build_topology(x, nv, faces, nf, t->t, err)) {
        delete t;
        *out = nullptr;
        return set_err(LATS_ERR_INVALID_ARGUMENT, err.c_str());
    }
    *out = t;
    return LATS_OK;
}

void lats_topology_destroy(lats_topology* t) { delete t; }

int lats_topology_save(const lats_topology* t, const char* dir) {
    if (!t || !dir) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    std::string err;
    if (!lats::save_topology(t->t, dir, err)) return set_err(LATS_ERR_INVALID_ARGUMENT, err.c_str());
    return LATS_OK;
}

int lats_topology_load(const char* dir, lats_topology** out) {
    if (!dir || !out) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    auto* t = new lats_topology();
    std::string err;
    if (!lats::load_topology(dir, t->t, err)) {
        delete t;
        return set_err(LATS_ERR_INVALID_ARGUMENT, err.c_str());
    }
    *out = t;
    return LATS_OK;
}

int lats_topology_counts(const lats_topology* t, uint32_t counts[4]) {
    if (!t || !counts) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    counts[0] = t->t.num_verts;
    counts[1] = t->t.num_edges();
    counts[2] = t->t.num_bending();
    counts[3] = t->t.num_clusters;
    return LATS_OK;
}

int lats_topology_size(const lats_topology* t, int which, uint64_t* n) {
    if (!t || !n) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    const auto& T = t->t;
    switch (which) {
        case LATS_TOPO_EDGES: *n = T.edges.size(); break;
        case LATS_TOPO_BENDING: *n = T.bending.size(); break;
        case LATS_TOPO_REST_LENGTH: *n = T.rest_length.size(); break;
        case LATS_TOPO_BENDING_Q: *n = T.bending_Q.size(); break;
        case LATS_TOPO_ADJ_EDGES_CSR: *n = T.adj_edges_csr.size(); break;
        case LATS_TOPO_ADJ_BENDING_CSR: *n = T.adj_bending_csr.size(); break;
        case LATS_TOPO_CLUSTERS_CSR: *n = T.clusters_csr.size(); break;
        case LATS_TOPO_COLOUR: *n = T.colour.size(); break;
        case LATS_TOPO_FACES: *n = T.faces.size(); break;
        default: return set_err(LATS_ERR_INVALID_ARGUMENT, "unknown topology array");
    }
    return LATS_OK;
}

int lats_topology_copy(const lats_topology* t, int which, void* dst) {
    if (!t || !dst) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    const auto& T = t->t;
    auto cp = [&](const auto& v) { std::memcpy(dst, v.data(), v.size() * sizeof(v[0])); };
    switch (which) {
        case LATS_TOPO_EDGES: cp(T.edges); break;
        case LATS_TOPO_BENDING: cp(T.bending); break;
        case LATS_TOPO_REST_LENGTH: cp(T.rest_length); break;
        case LATS_TOPO_BENDING_Q: cp(T.bending_Q); break;
        case LATS_TOPO_ADJ_EDGES_CSR: cp(T.adj_edges_csr); break;
        case LATS_TOPO_ADJ_BENDING_CSR: cp(T.adj_bending_csr); break;
        case LATS_TOPO_CLUSTERS_CSR: cp(T.clusters_csr); break;
        case LATS_TOPO_COLOUR: cp(T.colour); break;
        case LATS_TOPO_FACES: cp(T.faces); break;
        default: return set_err(LATS_ERR_INVALID_ARGUMENT, "unknown topology array");
    }
    return LATS_OK;
}

int lats_scene_grid(uint32_t nx, uint32_t ny, float spacing, float jitter, uint64_t seed, float* x, int32_t* faces) {
    if (nx < 2 || ny < 2 || !x || !faces) return set_err(LATS_ERR_INVALID_ARGUMENT, "grid needs nx, ny >= 2");
    std::vector<float> hx;
    std::vector<int32_t> hf;
    lats::make_grid(nx, ny, spacing, jitter, seed, hx, hf);
    std::memcpy(x, hx.data(), hx.size() * sizeof(float));
    std::memcpy(faces, hf.data(), hf.size() * sizeof(int32_t));
    return LATS_OK;
}

int lats_scene_icosphere(uint32_t level, float radius, const float centre[3], float* x, int32_t* faces,
                         uint32_t* num_verts, uint32_t* num_faces) {
    if (!centre || level > 8) return set_err(LATS_ERR_INVALID_ARGUMENT, "bad icosphere arguments");
    std::vector<float> hx;
    std::vector<int32_t> hf;
    lats::make_icosphere(level, radius, centre, hx, hf);
    if (num_verts) *num_verts = (uint32_t)(hx.size() / 3);
    if (num_faces) *num_faces = (uint32_t)(hf.size() / 3);
    if (x) std::memcpy(x, hx.data(), hx.size() * sizeof(float));
    if (faces) std::memcpy(faces, hf.data(), hf.size() * sizeof(int32_t));
    return LATS_OK;
}

int lats_scene_wind(uint32_t nframes, uint32_t nv, float max_total, uint64_t seed, float* out) {
    if (!out && nframes) return set_err(LATS_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<float> w = lats::make_wind(nframes, nv, max_total, seed);
    std::memcpy(out, w.data(), w.size() * sizeof(float));
    return LATS_OK;
}

}  // extern "C"
