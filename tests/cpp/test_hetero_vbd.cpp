// The asynchronous frame across two processors: the CPU oracle (processor 0,
// test infrastructure) and the HIP context (processor 1, the product), run by
// the Launcher in LaunchModeHetero -- the host thread iterating its clusters
// while the GPU queue runs its own behind host-value waits -- and in
// LaunchModeFakeHetero.
//
// The reference shares every position buffer between the two processors
// (unified memory).  Here each processor owns its buffers, and a buffer
// crosses at a cross-processor edge of the async schedule: the producer's
// iterate and begin buffers are copied into pinned host staging right after
// it ran (per producer task, so nothing later overwrites them), and the
// consumer copies them in before it runs; the predicted positions cross the
// same way.  A third run does exactly that with two oracle instances in
// LaunchModeFakeHetero; the heterogeneous runs must match it bit for bit.
//
// The shared-buffer replay (LaunchModeSequeceHetero on one oracle, the
// reference's semantics) is reported beside it, not matched: there a task can
// read a buffer the other processor is iterating in, and what it reads
// depends on the order the two processors happen to run in (SequeceHetero
// runs a GPU event only when the CPU waits for it, later than its scheduled
// time).  With one memory per processor a task reads exactly what its
// scheduled inputs produced, whatever the timing.
//
// Needs a GPU, except with a fourth argument "cpu": then only the two oracle
// runs are made (the harness itself, checked without a device).
// usage: test_hetero_vbd [n] [iterations] [frames] [cpu]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/launcher.h"
#include "host/vbd_slot.h"

extern "C" {
struct orc_mesh;
struct orc_sim;
orc_mesh* orc_mesh_create(uint32_t nv, const float* x, uint32_t nf, const int* faces);
void orc_mesh_destroy(orc_mesh* m);
void orc_mesh_counts(const orc_mesh* m, uint32_t* out);
orc_sim* orc_sim_create(orc_mesh* m, const float* params9, const uint32_t* iparams2, const float* mass,
                        const unsigned char* fixed, int fp64);
void orc_sim_destroy(orc_sim* s);
uint32_t orc_sim_step(orc_sim* s, int record, float* energy, uint32_t cap);
void orc_sim_async_begin(orc_sim* s);
void orc_sim_dispatch(orc_sim* s, uint32_t fn, uint32_t iter, uint32_t cluster, uint32_t buffer, uint32_t left,
                      uint32_t right, uint32_t nin, const uint32_t* ins, int main_device, int record);
void orc_sim_buffer_get(orc_sim* s, int which, uint32_t idx, float* out);
void orc_sim_buffer_set(orc_sim* s, int which, uint32_t idx, const float* in);
double orc_sim_energy(orc_sim* s);
}

#define CHECK(x)                                                                   \
    do {                                                                           \
        int rc_ = (x);                                                             \
        if (rc_ != LATS_OK) {                                                      \
            std::printf("FAIL %s -> %d (%s)\n", #x, rc_, lats_last_error());       \
            std::exit(1);                                                          \
        }                                                                          \
    } while (0)

namespace {

uint32_t c_function(Launcher::FunctionID f) {
    switch (f) {
        case Launcher::id_xpbd_predict_position: return LATS_FN_PREDICT_POSITION;
        case Launcher::id_vbd_all_in_one: return LATS_FN_VBD_ALL_IN_ONE;
        case Launcher::id_xpbd_constraint_last_node: return LATS_FN_CONSTRAINT_LAST_NODE;
        case Launcher::id_xpbd_update_velocity: return LATS_FN_UPDATE_VELOCITY;
        case Launcher::id_additional_root: return LATS_FN_ADDITIONAL_ROOT;
        case Launcher::id_additional_terminal: return LATS_FN_ADDITIONAL_TERMINAL;
        default: return LATS_FN_RESET_CONSTRAINS;
    }
}

// One processor's memory and solver.
struct Side {
    virtual ~Side() = default;
    virtual void dispatch(const Launcher::LaunchParam& p) = 0;
    virtual void put(int which, uint32_t idx, const float* host) = 0;   // host staging -> side
    virtual void get(int which, uint32_t idx, float* host) = 0;         // side -> host staging
    virtual void frame_begin() {}
};

struct OracleSide : Side {
    orc_sim* s;
    explicit OracleSide(orc_sim* sim) : s(sim) {}
    void dispatch(const Launcher::LaunchParam& p) override {
        orc_sim_dispatch(s, c_function(p.function_id), p.iter_idx, p.cluster_idx, p.buffer_idx, p.left_buffer_idx,
                         p.right_buffer_idx, (uint32_t)p.input_buffer_idxs.size(),
                         p.input_buffer_idxs.empty() ? nullptr : p.input_buffer_idxs.data(),
                         p.is_allocated_to_main_device ? 1 : 0, 0);
    }
    void put(int which, uint32_t idx, const float* host) override { orc_sim_buffer_set(s, which, idx, host); }
    void get(int which, uint32_t idx, float* host) override { orc_sim_buffer_get(s, which, idx, host); }
    void frame_begin() override { orc_sim_async_begin(s); }
};

struct GpuSide : Side {
    lats_vbd* ctx;
    lats::VbdGpuSlot slot;
    explicit GpuSide(lats_vbd* c) : ctx(c), slot(c) {}
    void dispatch(const Launcher::LaunchParam& p) override { slot.fn_dispatch(p); }
    void put(int which, uint32_t idx, const float* host) override { CHECK(lats_vbd_buffer_upload(ctx, which, idx, host)); }
    void get(int which, uint32_t idx, float* host) override { CHECK(lats_vbd_buffer_download(ctx, which, idx, host)); }
};

bool g_pinned = true;   // false without a device (the oracle-only runs)

struct Pinned {
    float* p = nullptr;
    explicit Pinned(size_t n) {
        void* q = nullptr;
        if (g_pinned) CHECK(lats_host_alloc(n * sizeof(float), &q));
        else q = std::malloc(n * sizeof(float));
        p = (float*)q;
        std::memset(p, 0, n * sizeof(float));
    }
    ~Pinned() {
        if (g_pinned) lats_host_free(p);
        else std::free(p);
    }
};

// Runs one asynchronous frame on (sides[0], sides[1]) in `mode`; buffers
// cross through staging owned by this call.
struct FrameRun {
    uint32_t V;
    Launcher::Scheduler s;
    Side* sides[2] = {nullptr, nullptr};
    const Launcher::Task* current = nullptr;
    std::vector<std::unique_ptr<Pinned>> st_iter, st_begin;
    std::unique_ptr<Pinned> st_x;
    uint32_t predict_tid = ~0u;

    FrameRun(uint32_t nv, uint32_t iters, uint32_t clusters) : V(nv) {
        Launcher::Implementation cpu(Launcher::DeviceTypeCpu, [this](const Launcher::LaunchParam& p) { run(0, p); });
        Launcher::Implementation gpu(Launcher::DeviceTypeGpu, [this](const Launcher::LaunchParam& p) { run(1, p); });
        lats::register_vbd_dag(s, {cpu, gpu}, iters, clusters);
        std::vector<std::pair<Launcher::FunctionID, Launcher::uint>> keys = {
            {Launcher::id_xpbd_predict_position, 0}, {Launcher::id_xpbd_update_velocity, 0},
            {Launcher::id_xpbd_constraint_last_node, 0}, {Launcher::id_vbd_all_in_one, 0}};
        for (uint32_t c = 0; c < clusters; c++) keys.push_back({Launcher::id_vbd_all_in_one, c});
        std::vector<std::vector<double>> costs(keys.size(), {0.1f, 0.1f});
        std::vector<double> total;
        s.profile_from(keys, costs, total);
        s.communication_cost_matrix_uma = {{0.0f, 0.1f}, {0.1f, 0.0f}};
        s.communication_speed_matrix = {};
        s.communication_startup = {0, 0};
        if (!s.topological_sort()) { std::printf("FAIL topological sort\n"); std::exit(1); }
        s.standardizing_dag();
        s.scheduler_dag();
        if (!s.make_wait_events()) { std::printf("FAIL wait events\n"); std::exit(1); }
        const auto& tasks = s.get_list_task();
        st_iter.resize(tasks.size());
        st_begin.resize(tasks.size());
        for (size_t t = 0; t < tasks.size(); t++)
            if (tasks[t].func_id == Launcher::id_xpbd_predict_position) predict_tid = (uint32_t)t;
        st_x.reset(new Pinned(3 * (size_t)V));
    }
    uint32_t proc_of(uint32_t t) const { return s.get_task_schedules()[t].proc; }

    void run(int proc, const Launcher::LaunchParam& p) {
        const auto& tasks = s.get_list_task();
        const uint32_t tid = (uint32_t)(current - tasks.data());
        const Launcher::Task& task = *current;
        Side* side = sides[proc];
        for (size_t k = 0; k < task.task_ins.size(); k++) {
            const uint32_t in = task.task_ins[k], b = task.buffer_ins[k];
            side->put(LATS_BUF_ITER, b, st_iter[in]->p);
            side->put(LATS_BUF_BEGIN, b, st_begin[in]->p);
        }
        if (task.buffer_left == Launcher::input_buffer_mask && proc_of(predict_tid) != (uint32_t)proc)
            side->put(LATS_BUF_X, 0, st_x->p);
        side->dispatch(p);
        if (!task.task_outs.empty()) {
            if (!st_iter[tid]) {
                st_iter[tid].reset(new Pinned(3 * (size_t)V));
                st_begin[tid].reset(new Pinned(3 * (size_t)V));
            }
            side->get(LATS_BUF_ITER, task.buffer_idx, st_iter[tid]->p);
            side->get(LATS_BUF_BEGIN, task.buffer_idx, st_begin[tid]->p);
        }
        if (tid == predict_tid) side->get(LATS_BUF_X, 0, st_x->p);
    }

    void launch(Launcher::Scheduler::LaunchMode mode, const Launcher::HeteroHooks* hooks) {
        // staging of every producer exists before any queued copy refers to it
        const auto& tasks = s.get_list_task();
        for (size_t t = 0; t < tasks.size(); t++)
            if (!tasks[t].task_outs.empty() && !st_iter[t]) {
                st_iter[t].reset(new Pinned(3 * (size_t)V));
                st_begin[t].reset(new Pinned(3 * (size_t)V));
            }
        sides[0]->frame_begin();
        sides[1]->frame_begin();
        s.launch(mode, [this](const Launcher::Task& t) {
            current = &t;
            return lats::vbd_task_to_param(t);
        }, hooks);
    }
};

}  // namespace

int main(int argc, char** argv) {
    const uint32_t n = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 24;
    const uint32_t iters = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 20;
    const uint32_t frames = argc > 3 ? (uint32_t)std::atoi(argv[3]) : 2;
    const bool cpu_only = argc > 4 && std::strcmp(argv[4], "cpu") == 0;
    g_pinned = !cpu_only;
    const uint32_t V = n * n;
    std::vector<float> x(3 * V);
    std::vector<int32_t> faces(6 * (n - 1) * (n - 1));
    CHECK(lats_scene_grid(n, n, 0.01f, 1e-4f, 7, x.data(), faces.data()));
    lats_topology* topo = nullptr;
    CHECK(lats_topology_build(x.data(), V, faces.data(), (uint32_t)faces.size() / 3, &topo));
    uint32_t cnt[4];
    CHECK(lats_topology_counts(topo, cnt));
    std::vector<std::vector<char>> arrays(8);
    for (int w = 0; w < 8; w++) {
        uint64_t ne = 0;
        CHECK(lats_topology_size(topo, w, &ne));
        arrays[w].resize(ne * ((w == LATS_TOPO_COLOUR) ? 1 : 4) + 4);
        CHECK(lats_topology_copy(topo, w, arrays[w].data()));
    }
    std::vector<float> mass(V, 1e-3f);
    std::vector<uint8_t> fixed(V, 0);
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

    orc_mesh* om = orc_mesh_create(V, x.data(), (uint32_t)faces.size() / 3, faces.data());
    uint32_t ocnt[4];
    orc_mesh_counts(om, ocnt);
    if (ocnt[3] != cnt[3]) { std::printf("FAIL cluster counts %u vs %u\n", ocnt[3], cnt[3]); return 1; }
    const float p9[9] = {1.f / 60.f, 1e4f, 1e-2f, 2.0f, 0.f, -9.8f, 0.f, 0.f, 0.f};
    const uint32_t i2[2] = {1, iters};
    // oracles: the shared-buffer replay, the two-oracle run, and the CPU side
    // of each heterogeneous run
    orc_sim* o_shared = orc_sim_create(om, p9, i2, mass.data(), fixed.data(), 0);
    orc_sim* o_cpu2 = orc_sim_create(om, p9, i2, mass.data(), fixed.data(), 0);
    orc_sim* o_gpu2 = orc_sim_create(om, p9, i2, mass.data(), fixed.data(), 0);
    orc_sim* o_cpu_h = orc_sim_create(om, p9, i2, mass.data(), fixed.data(), 0);
    orc_sim* o_cpu_f = orc_sim_create(om, p9, i2, mass.data(), fixed.data(), 0);
    orc_sim* o_cpu_p = orc_sim_create(om, p9, i2, mass.data(), fixed.data(), 0);
    lats_vbd *g_h = nullptr, *g_f = nullptr, *g_p = nullptr;
    if (!cpu_only) {
        CHECK(lats_vbd_create(&d, &prm, &g_h));
        CHECK(lats_vbd_create(&d, &prm, &g_f));
        CHECK(lats_vbd_create(&d, &prm, &g_p));
    }
    // a few synchronous frames first, identical everywhere
    for (int f = 0; f < 3; f++) {
        for (orc_sim* o : {o_shared, o_cpu2, o_gpu2, o_cpu_h, o_cpu_f, o_cpu_p}) orc_sim_step(o, 0, nullptr, 0);
        if (cpu_only) continue;
        CHECK(lats_vbd_physics_step(g_h, 0));
        CHECK(lats_vbd_physics_step(g_f, 0));
        CHECK(lats_vbd_physics_step(g_p, 0));
    }
    std::vector<float> a(3 * V), b(3 * V), tmp(3 * V);
    auto same_state = [&](orc_sim* o, lats_vbd* g) {
        orc_sim_buffer_get(o, LATS_BUF_X, 0, a.data());
        CHECK(lats_vbd_buffer_download(g, LATS_BUF_X, 0, b.data()));
        CHECK(lats_vbd_synchronize(g));
        return a == b;
    };
    if (!cpu_only && !same_state(o_cpu_h, g_h)) { std::printf("FAIL synchronous frames differ\n"); return 1; }

    OracleSide s_shared(o_shared), s_cpu2(o_cpu2), s_gpu2(o_gpu2), s_cpu_h(o_cpu_h), s_cpu_f(o_cpu_f), s_cpu_p(o_cpu_p);
    GpuSide s_gh(g_h), s_gf(g_f), s_gp(g_p);
    int bad = 0;
    for (uint32_t f = 0; f < frames; f++) {
        // the reference's semantics: one memory, SequeceHetero order
        {
            FrameRun r(V, iters, cnt[3]);
            r.sides[0] = r.sides[1] = &s_shared;
            r.launch(Launcher::Scheduler::LaunchModeSequeceHetero, nullptr);
        }
        // two memories, two oracles
        {
            FrameRun r(V, iters, cnt[3]);
            r.sides[0] = &s_cpu2;
            r.sides[1] = &s_gpu2;
            r.launch(Launcher::Scheduler::LaunchModeFakeHetero, nullptr);
        }
        if (cpu_only) {
            for (int which : {LATS_BUF_X, LATS_BUF_X_START, LATS_BUF_V}) {
                orc_sim_buffer_get(o_cpu2, which, 0, tmp.data());
                orc_sim_buffer_set(o_gpu2, which, 0, tmp.data());
            }
            std::vector<float> x2(3 * V), xs(3 * V);
            orc_sim_buffer_get(o_cpu2, LATS_BUF_X, 0, x2.data());
            orc_sim_buffer_get(o_shared, LATS_BUF_X, 0, xs.data());
            double dmax = 0.0, xmax = 0.0;
            for (size_t i = 0; i < xs.size(); i++) {
                dmax = std::fmax(dmax, std::fabs((double)x2[i] - xs[i]));
                xmax = std::fmax(xmax, std::fabs((double)xs[i]));
            }
            std::printf("frame %u: two-memory vs shared-buffer replay max |dx| %.3g (rel %.3g), energy %.6f vs %.6f\n", f,
                        dmax, dmax / xmax, orc_sim_energy(o_cpu2), orc_sim_energy(o_shared));
            if (dmax / xmax > 5e-2) { std::printf("FAIL far from the shared-buffer replay\n"); bad = 1; }
            continue;
        }
        // CPU oracle + GPU, concurrent
        FrameRun rh(V, iters, cnt[3]);
        if (rh.proc_of(rh.predict_tid) != 0) { std::printf("FAIL predict is not on the CPU\n"); return 1; }
        rh.sides[0] = &s_cpu_h;
        rh.sides[1] = &s_gh;
        const Launcher::HeteroHooks hooks = s_gh.slot.hetero_hooks(60.0);
        rh.launch(Launcher::Scheduler::LaunchModeHetero, &hooks);
        CHECK(s_gh.slot.status());
        // CPU oracle + GPU, one processor at a time
        FrameRun rf(V, iters, cnt[3]);
        rf.sides[0] = &s_cpu_f;
        rf.sides[1] = &s_gf;
        const Launcher::HeteroHooks hooks_f = s_gf.slot.hetero_hooks(60.0);
        rf.launch(Launcher::Scheduler::LaunchModeFakeHetero, &hooks_f);
        CHECK(s_gf.slot.status());
        CHECK(lats_vbd_synchronize(g_f));
        // CPU oracle + GPU, GPU events queued three at a time
        FrameRun rp(V, iters, cnt[3]);
        rp.sides[0] = &s_cpu_p;
        rp.sides[1] = &s_gp;
        rp.s.progressive_window = 3;
        const Launcher::HeteroHooks hooks_p = s_gp.slot.hetero_hooks(60.0);
        rp.launch(Launcher::Scheduler::LaunchModeProgressiveHetero, &hooks_p);
        CHECK(s_gp.slot.status());

        // update_velocity ran on the CPU side: hand the frame state to the GPU
        // sides for the next frame
        for (int which : {LATS_BUF_X, LATS_BUF_X_START, LATS_BUF_V}) {
            orc_sim_buffer_get(o_cpu2, which, 0, tmp.data());
            orc_sim_buffer_set(o_gpu2, which, 0, tmp.data());
            orc_sim_buffer_get(o_cpu_h, which, 0, tmp.data());
            CHECK(lats_vbd_buffer_upload(g_h, which, 0, tmp.data()));
            CHECK(lats_vbd_synchronize(g_h));
            orc_sim_buffer_get(o_cpu_f, which, 0, tmp.data());
            CHECK(lats_vbd_buffer_upload(g_f, which, 0, tmp.data()));
            CHECK(lats_vbd_synchronize(g_f));
            orc_sim_buffer_get(o_cpu_p, which, 0, tmp.data());
            CHECK(lats_vbd_buffer_upload(g_p, which, 0, tmp.data()));
            CHECK(lats_vbd_synchronize(g_p));
        }
        std::vector<float> x2(3 * V), xh(3 * V), xf(3 * V), xp(3 * V), xs(3 * V), v2(3 * V), vh(3 * V);
        orc_sim_buffer_get(o_cpu_p, LATS_BUF_X, 0, xp.data());
        orc_sim_buffer_get(o_cpu2, LATS_BUF_X, 0, x2.data());
        orc_sim_buffer_get(o_cpu_h, LATS_BUF_X, 0, xh.data());
        orc_sim_buffer_get(o_cpu_f, LATS_BUF_X, 0, xf.data());
        orc_sim_buffer_get(o_shared, LATS_BUF_X, 0, xs.data());
        orc_sim_buffer_get(o_cpu2, LATS_BUF_V, 0, v2.data());
        orc_sim_buffer_get(o_cpu_h, LATS_BUF_V, 0, vh.data());
        double dmax = 0.0, xmax = 0.0;
        for (size_t i = 0; i < xs.size(); i++) {
            dmax = std::fmax(dmax, std::fabs((double)x2[i] - xs[i]));
            xmax = std::fmax(xmax, std::fabs((double)xs[i]));
        }
        const bool eq_h = xh == x2 && vh == v2, eq_f = xf == x2, eq_p = xp == x2;
        std::printf("frame %u: hetero == two-memory replay: %s, fake-hetero: %s, progressive: %s; vs shared-buffer "
                    "replay max |dx| %.3g (rel %.3g), energy %.6f vs %.6f\n",
                    f, eq_h ? "bit-exact" : "DIFFERENT", eq_f ? "bit-exact" : "DIFFERENT",
                    eq_p ? "bit-exact" : "DIFFERENT", dmax, dmax / xmax, orc_sim_energy(o_cpu_h),
                    orc_sim_energy(o_shared));
        if (!eq_h || !eq_f || !eq_p) bad = 1;
        if (dmax / xmax > 5e-2) { std::printf("FAIL far from the shared-buffer replay\n"); bad = 1; }
        if (f == 0) {
            const auto& ev = rh.s.get_launch_events();
            std::printf("events: cpu %zu gpu %zu; refit GPU cost of the first cluster task %.4f ms\n", ev[0].size(),
                        ev[1].size(), rh.s.computation_matrix[1][1]);
        }
    }
    if (g_h) lats_vbd_destroy(g_h);
    if (g_f) lats_vbd_destroy(g_f);
    if (g_p) lats_vbd_destroy(g_p);
    for (orc_sim* o : {o_shared, o_cpu2, o_gpu2, o_cpu_h, o_cpu_f, o_cpu_p}) orc_sim_destroy(o);
    orc_mesh_destroy(om);
    lats_topology_destroy(topo);
    std::printf(bad ? "FAIL\n" : "PASS\n");
    return bad;
}
