// The host-side code a reference driver adds to use the HIP slot, as a
// program: the Launcher adapter, the hand-written GpuSolver::fn_dispatch
// lambda over the reference's own LaunchParam
// (examples/example3_hybrid_cpu_gpu.cpp:1744-1920) and the XPBD frame DAG
// (example2_async_iter.cpp:1017-1081), each run for two frames on a 24 x 24
// cloth and compared bit for bit with the fused lats_vbd_physics_step frames
// of a second context.  Needs a GPU.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/launcher.h"   // Launcher::Task / Implementation / Scheduler
#include "host/vbd_slot.h"

#define CHECK(x)                                                                   \
    do {                                                                           \
        int rc_ = (x);                                                             \
        if (rc_ != LATS_OK) {                                                      \
            std::printf("FAIL %s -> %d (%s)\n", #x, rc_, lats_last_error());       \
            return 1;                                                              \
        }                                                                          \
    } while (0)

// One frame through the shipped adapter.
static int frame_with_adapter(lats_vbd* gpu, uint32_t iters, uint32_t nclusters) {
    lats::VbdGpuSlot slot(gpu);
    Launcher::Implementation imp_gpu = slot.implementation();      // DeviceTypeGpu
    Launcher::Scheduler scheduler;
    lats::register_vbd_dag(scheduler, {imp_gpu}, iters, nclusters);  // or the driver's own Task registrations
    scheduler.communication_cost_matrix_uma = {{0.002f, 0.048f}, {0.047f, 0.01f}};
    scheduler.communication_speed_matrix = {};
    scheduler.communication_startup = {0, 0};
    scheduler.topological_sort();
    scheduler.computation_matrix.assign(scheduler.get_list_task().size(), {Launcher::inf, 0.01f});
    scheduler.standardizing_dag();
    scheduler.scheduler_dag();
    scheduler.launch(Launcher::Scheduler::LaunchModeGpu, lats::vbd_task_to_param);
    lats_vbd_synchronize(gpu);                                       // ≙ send_and_wait()
    return slot.status();
}

// One frame through a hand-written slot over LaunchParam (what the adapter does).
static int frame_with_lambda(lats_vbd* gpu, uint32_t iters, uint32_t nclusters) {
    int status = LATS_OK;
    Launcher::Implementation imp_gpu(Launcher::DeviceTypeGpu, [gpu, &status](const Launcher::LaunchParam& param) {
    lats_launch_param c{};
    switch (param.function_id) {
        case Launcher::id_xpbd_predict_position:     c.function_id = LATS_FN_PREDICT_POSITION; break;
        case Launcher::id_vbd_all_in_one:            c.function_id = LATS_FN_VBD_ALL_IN_ONE; break;
        case Launcher::id_xpbd_constraint_last_node: c.function_id = LATS_FN_CONSTRAINT_LAST_NODE; break;
        case Launcher::id_xpbd_update_velocity:      c.function_id = LATS_FN_UPDATE_VELOCITY; break;
        case Launcher::id_xpbd_reset_constrains:     c.function_id = LATS_FN_RESET_CONSTRAINS; break;
        case Launcher::id_additional_root:           c.function_id = LATS_FN_ADDITIONAL_ROOT; break;
        case Launcher::id_additional_terminal:       c.function_id = LATS_FN_ADDITIONAL_TERMINAL; break;
        case Launcher::id_xpbd_constraint_stretch_mass_spring: c.function_id = LATS_FN_XPBD_STRETCH; break;
        case Launcher::id_xpbd_constraint_bending:   c.function_id = LATS_FN_XPBD_BENDING; break;
        default: return;
    }
    c.iter_idx = param.iter_idx;            c.cluster_idx = param.cluster_idx;
    c.buffer_idx = param.buffer_idx;        c.left_buffer_idx = param.left_buffer_idx;
    c.right_buffer_idx = param.right_buffer_idx;
    c.num_input_buffers = (uint32_t)param.input_buffer_idxs.size();
    c.input_buffer_idxs = param.input_buffer_idxs.data();
    c.is_allocated_to_main_device = param.is_allocated_to_main_device;
    const int rc = lats_vbd_dispatch(gpu, &c);
    if (rc != LATS_OK && status == LATS_OK) status = rc;   // first failure kept; lats_last_error() says why
});
    Launcher::Scheduler scheduler;
    lats::register_vbd_dag(scheduler, {imp_gpu}, iters, nclusters);
    scheduler.communication_cost_matrix_uma = {{0.002f, 0.048f}, {0.047f, 0.01f}};
    scheduler.communication_speed_matrix = {};
    scheduler.communication_startup = {0, 0};
    scheduler.topological_sort();
    scheduler.computation_matrix.assign(scheduler.get_list_task().size(), {Launcher::inf, 0.01f});
    scheduler.standardizing_dag();
    scheduler.scheduler_dag();
    scheduler.launch(Launcher::Scheduler::LaunchModeGpu, lats::vbd_task_to_param);
    lats_vbd_synchronize(gpu);
    return status;
}

// One XPBD frame through the scheduler.
static int xpbd_frame(lats_vbd* gpu, uint32_t iters) {
    lats::VbdGpuSlot slot(gpu);
    std::vector<Launcher::Implementation> impls = {slot.implementation()};
    Launcher::Scheduler scheduler;
    lats_vbd_set_solver(gpu, LATS_SOLVER_XPBD);
    uint64_t ns = 0;
    lats_vbd_xpbd_clusters(gpu, 0, nullptr, 0, &ns);   // size: clusters + 1 offsets + springs
    std::vector<uint32_t> csr(ns);
    lats_vbd_xpbd_clusters(gpu, 0, csr.data(), ns, &ns);
    uint32_t n_stretch = csr[0] - 1;                   // cluster i = csr[csr[i] .. csr[i + 1])
    uint64_t nb = 0;
    lats_vbd_xpbd_clusters(gpu, 1, nullptr, 0, &nb);
    std::vector<uint32_t> bcsr(nb);
    lats_vbd_xpbd_clusters(gpu, 1, bcsr.data(), nb, &nb);
    uint32_t n_bending = bcsr[0] - 1;
    auto order = lats::register_xpbd_dag(scheduler, impls, iters, n_stretch, n_bending);
    scheduler.communication_cost_matrix_uma = {{0.002f, 0.048f}, {0.047f, 0.01f}};
    scheduler.communication_speed_matrix = {};
    scheduler.communication_startup = {0, 0};
    scheduler.topological_sort();
    scheduler.computation_matrix.assign(scheduler.get_list_task().size(), {Launcher::inf, 0.01f});
    scheduler.standardizing_dag();
    scheduler.scheduler_dag();
    scheduler.launch(Launcher::Scheduler::LaunchModeGpu, lats::vbd_task_to_param);
    lats_vbd_synchronize(gpu);
    if (order.size() != (size_t)iters * (n_stretch + n_bending)) return -1;
    return slot.status();
}

static bool same_state(lats_vbd* a, lats_vbd* b, uint32_t nv) {
    std::vector<float> xa(3 * nv), xb(3 * nv), va(3 * nv), vb(3 * nv);
    if (lats_vbd_get_state(a, xa.data(), va.data()) || lats_vbd_get_state(b, xb.data(), vb.data())) return false;
    return std::memcmp(xa.data(), xb.data(), xa.size() * 4) == 0 && std::memcmp(va.data(), vb.data(), va.size() * 4) == 0;
}

int main() {
    const uint32_t n = 24, iters = 20, frames = 2;
    std::vector<float> x(3 * n * n);
    std::vector<int32_t> faces(6 * (n - 1) * (n - 1));
    CHECK(lats_scene_grid(n, n, 0.01f, 1e-4f, 7, x.data(), faces.data()));
    lats_topology* topo = nullptr;
    CHECK(lats_topology_build(x.data(), n * n, faces.data(), (uint32_t)faces.size() / 3, &topo));
    uint32_t cnt[4];
    CHECK(lats_topology_counts(topo, cnt));
    std::vector<std::vector<char>> arrays(8);
    for (int w = 0; w < 8; w++) {
        uint64_t ne = 0;
        CHECK(lats_topology_size(topo, w, &ne));
        arrays[w].resize(ne * (w == LATS_TOPO_COLOUR ? 1 : 4) + 4);
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
    lats_vbd *fused = nullptr, *adapter = nullptr, *lambda = nullptr, *xfused = nullptr, *xsched = nullptr;
    CHECK(lats_vbd_create(&d, &prm, &fused));
    CHECK(lats_vbd_create(&d, &prm, &adapter));
    CHECK(lats_vbd_create(&d, &prm, &lambda));
    CHECK(lats_vbd_create(&d, &prm, &xfused));
    CHECK(lats_vbd_create(&d, &prm, &xsched));
    CHECK(lats_vbd_set_solver(xfused, LATS_SOLVER_XPBD));
    for (uint32_t f = 0; f < frames; f++) {
        CHECK(lats_vbd_physics_step(fused, 0));
        CHECK(frame_with_adapter(adapter, iters, cnt[3]));
        CHECK(frame_with_lambda(lambda, iters, cnt[3]));
        CHECK(lats_vbd_physics_step(xfused, 0));
        CHECK(xpbd_frame(xsched, iters));
    }
    const bool a = same_state(fused, adapter, n * n), b = same_state(fused, lambda, n * n);
    const bool c = same_state(xfused, xsched, n * n);
    std::printf("adapter %s fused, lambda %s fused, scheduled XPBD %s fused XPBD\n", a ? "==" : "!=", b ? "==" : "!=",
                c ? "==" : "!=");
    for (lats_vbd* ctx : {fused, adapter, lambda, xfused, xsched}) lats_vbd_destroy(ctx);
    lats_topology_destroy(topo);
    const bool ok = a && b && c;
    std::printf(ok ? "PASS\n" : "FAIL\n");
    return ok ? 0 : 1;
}
