// The XPBD frame DAG (register_xpbd_dag) scheduled by HEFT and launched in
// LaunchModeGpu on the HIP slot equals the fused XPBD frame bit for bit.
// Needs a GPU.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/launcher.h"
#include "host/vbd_slot.h"

#define CHECK(x)                                                                   \
    do {                                                                           \
        int rc_ = (x);                                                             \
        if (rc_ != LATS_OK) {                                                      \
            std::printf("FAIL %s -> %d (%s)\n", #x, rc_, lats_last_error());       \
            return 1;                                                              \
        }                                                                          \
    } while (0)

int main() {
    const uint32_t n = 20, iters = 6, frames = 3;
    std::vector<float> x(3 * n * n);
    std::vector<int32_t> faces(6 * (n - 1) * (n - 1));
    CHECK(lats_scene_grid(n, n, 0.01f, 1e-4f, 11, x.data(), faces.data()));
    lats_topology* topo = nullptr;
    CHECK(lats_topology_build(x.data(), n * n, faces.data(), (uint32_t)faces.size() / 3, &topo));
    uint32_t cnt[4];
    CHECK(lats_topology_counts(topo, cnt));
    std::vector<std::vector<char>> arrays(8);
    for (int w = 0; w < 8; w++) {
        uint64_t ne = 0;
        CHECK(lats_topology_size(topo, w, &ne));
        arrays[w].resize(ne * ((w == LATS_TOPO_COLOUR) ? 1 : 4) + 4);
        CHECK(lats_topology_copy(topo, w, arrays[w].data()));
    }
    std::vector<float> mass(n * n, 1e-3f);
    std::vector<uint8_t> fixed(n * n, 0);
    fixed[0] = fixed[n - 1] = 1;
    lats_mesh_desc d{};
    d.num_verts = cnt[0];
    d.num_edges = cnt[1];
    d.num_bending_edges = cnt[2];
    d.num_clusters = cnt[3];
    d.rest_x = x.data();
    d.edges = (const int32_t*)arrays[LATS_TOPO_EDGES].data();
    d.bending_edges = (const int32_t*)arrays[LATS_TOPO_BENDING].data();
    d.edges_rest_length = (const float*)arrays[LATS_TOPO_REST_LENGTH].data();
    d.bending_Q = (const float*)arrays[LATS_TOPO_BENDING_Q].data();
    d.vert_adj_edges_csr = (const uint32_t*)arrays[LATS_TOPO_ADJ_EDGES_CSR].data();
    d.vert_adj_bending_csr = (const uint32_t*)arrays[LATS_TOPO_ADJ_BENDING_CSR].data();
    d.clusters_csr = (const uint32_t*)arrays[LATS_TOPO_CLUSTERS_CSR].data();
    d.vert_mass = mass.data();
    d.is_fixed = fixed.data();
    lats_scene_params prm{1.f / 60.f, 1, iters, 1e4f, 1e-2f, 2.0f, {0.f, -9.8f, 0.f}};
    lats_vbd *a = nullptr, *b = nullptr;
    CHECK(lats_vbd_create(&d, &prm, &a));
    CHECK(lats_vbd_create(&d, &prm, &b));
    CHECK(lats_vbd_set_solver(a, LATS_SOLVER_XPBD));
    uint64_t ns = 0, nb = 0;
    std::vector<uint32_t> hdr(1);
    CHECK(lats_vbd_xpbd_clusters(b, 0, hdr.data(), 1, &ns));
    const uint32_t nstretch = hdr[0] - 1;
    CHECK(lats_vbd_xpbd_clusters(b, 1, hdr.data(), 1, &nb));
    const uint32_t nbending = hdr[0] - 1;
    lats::VbdGpuSlot slot(b);
    for (uint32_t f = 0; f < frames; f++) {
        CHECK(lats_vbd_physics_step(a, 0));
        Launcher::Scheduler s;
        lats::register_xpbd_dag(s, {slot.implementation()}, iters, nstretch, nbending);
        s.communication_cost_matrix_uma = {{0.002f, 0.220f}, {0.145f, 0.01f}};
        s.communication_speed_matrix = {};
        s.communication_startup = {0, 0};
        if (!s.topological_sort()) { std::printf("FAIL topological sort\n"); return 1; }
        s.computation_matrix.assign(s.get_list_task().size(), {Launcher::inf, 0.01f});
        s.standardizing_dag();
        s.scheduler_dag();
        s.launch(Launcher::Scheduler::LaunchModeGpu, lats::vbd_task_to_param);
        CHECK(slot.status());
        CHECK(lats_vbd_synchronize(b));
    }
    std::vector<float> xa(3 * n * n), xb(3 * n * n), va(3 * n * n), vb(3 * n * n);
    CHECK(lats_vbd_get_state(a, xa.data(), va.data()));
    CHECK(lats_vbd_get_state(b, xb.data(), vb.data()));
    const bool same = std::memcmp(xa.data(), xb.data(), xa.size() * 4) == 0 &&
                      std::memcmp(va.data(), vb.data(), va.size() * 4) == 0;
    std::printf("XPBD: %u spring clusters, %u bending clusters; scheduled frames %s fused frames\n", nstretch, nbending,
                same ? "==" : "!=");
    lats_vbd_destroy(a);
    lats_vbd_destroy(b);
    lats_topology_destroy(topo);
    std::printf(same ? "PASS\n" : "FAIL\n");
    return same ? 0 : 1;
}
