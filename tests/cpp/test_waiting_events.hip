// Host <-> GPU event ping-pong at queue depth, after the reference's
// examples/test_1_waiting_events.cpp:48-159: one counter shared by both
// processors; 60 GPU events are queued at once, GPU event j waiting until the
// host has signalled 2j + 1 and then adding one if the counter equals 2j + 1;
// the host runs 60 events, host event k first waiting for GPU event k - 1 and
// then adding one if the counter equals 2k, and signalling 2k + 1.  Every
// loop must end with the counter at 2 * 60.  The queue side uses the
// context's host-value waits (lats_vbd_stream_wait_host) and markers; the
// host side signals with lats_vbd_host_signal and waits with a bounded poll of
// lats_vbd_query_marker.  Also checked: no GPU event completes before the
// host signal it waits for.  The counter lives in coherent pinned host memory
// (the reference's shared buffer).  Needs a GPU.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <thread>
#include <vector>

#include "../../include/libatssim_hip.h"

#define CHECK(x)                                                                   \
    do {                                                                           \
        int rc_ = (x);                                                             \
        if (rc_ != LATS_OK) {                                                      \
            std::printf("FAIL %s -> %d (%s)\n", #x, rc_, lats_last_error());       \
            return 1;                                                              \
        }                                                                          \
    } while (0)

__global__ void k_empty() {}

// the reference's test_add_1 kernel: one increment when the counter holds `desire`
__global__ void k_test_add(unsigned* counter, unsigned desire) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned v = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (v == desire) __hip_atomic_store(counter, v + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// bounded host wait for a marker: false after `seconds`
static bool wait_marker(lats_vbd* ctx, unsigned m, double seconds) {
    const auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    while (true) {
        if (lats_vbd_query_marker(ctx, m, &done) != LATS_OK) return false;
        if (done) return true;
        if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > seconds) return false;
        std::this_thread::sleep_for(std::chrono::microseconds(5));
    }
}

int main() {
    // any context provides the queue, its host flag and markers: a 4 x 4 cloth
    const uint32_t n = 4;
    std::vector<float> x(3 * n * n);
    std::vector<int32_t> faces(6 * (n - 1) * (n - 1));
    CHECK(lats_scene_grid(n, n, 0.01f, 0.f, 1, x.data(), faces.data()));
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
    lats_scene_params prm{1.f / 60.f, 1, 1, 1e4f, 1e-2f, 2.0f, {0.f, -9.8f, 0.f}};
    lats_vbd* ctx = nullptr;
    CHECK(lats_vbd_create(&d, &prm, &ctx));
    void* sp = nullptr;
    CHECK(lats_vbd_get_stream(ctx, &sp));
    hipStream_t stream = static_cast<hipStream_t>(sp);

    unsigned* counter = nullptr;
    unsigned* counter_d = nullptr;
    if (hipHostMalloc((void**)&counter, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&counter_d, counter, 0) != hipSuccess) {
        std::printf("FAIL pinned counter\n");
        return 1;
    }

    const unsigned event_length = 60, loop_time = 100;   // the reference's sizes
    unsigned good_loops = 0, early = 0;
    bool stuck = false;
    for (unsigned frame = 0; frame < loop_time && !stuck; frame++) {
        CHECK(lats_vbd_host_signal(ctx, 0));
        __atomic_store_n(counter, 0u, __ATOMIC_SEQ_CST);
        // GPU events, queued at once: GPU 1, 3, 5, ...
        for (unsigned j = 0; j < event_length; j++) {
            CHECK(lats_vbd_stream_wait_host(ctx, 2 * j + 1));
            hipLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0, stream);
            hipLaunchKernelGGL(k_test_add, dim3(1), dim3(64), 0, stream, counter_d, 2 * j + 1);
            hipLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0, stream);
            if (hipGetLastError() != hipSuccess) { std::printf("FAIL launch\n"); stuck = true; break; }
            CHECK(lats_vbd_record_marker(ctx, j));
        }
        // host events, immediate: CPU 0, 2, 4, ...
        for (unsigned k = 0; k < event_length && !stuck; k++) {
            if (k != 0 && !wait_marker(ctx, k - 1, 20.0)) { stuck = true; break; }
            int done = 0;   // GPU event k waits for this event's signal: it cannot have run
            CHECK(lats_vbd_query_marker(ctx, k, &done));
            early += done;
            const unsigned v = __atomic_load_n(counter, __ATOMIC_SEQ_CST);
            if (v == 2 * k) __atomic_store_n(counter, v + 1, __ATOMIC_SEQ_CST);
            CHECK(lats_vbd_host_signal(ctx, 2 * k + 1));
        }
        if (!stuck && !wait_marker(ctx, event_length - 1, 20.0)) stuck = true;
        const unsigned got = __atomic_load_n(counter, __ATOMIC_SEQ_CST);
        if (!stuck && got == 2 * event_length) good_loops++;
        if (frame % 25 == 0 || got != 2 * event_length)
            std::printf("In loop %3u : get value %u (desire for %u)\n", frame, got, 2 * event_length);
    }
    if (stuck) {
        std::printf("a bounded host wait gave up: releasing the queue\n");
        lats_vbd_host_signal(ctx, 0xFFFFFFFFu);
    }
    CHECK(lats_vbd_synchronize(ctx));
    (void)hipHostFree(counter);
    lats_vbd_destroy(ctx);
    lats_topology_destroy(topo);
    const bool ok = !stuck && good_loops == loop_time && early == 0;
    std::printf("%u of %u loops reached %u; %u GPU events completed before their host signal\n", good_loops,
                loop_time, 2 * event_length, early);
    std::printf(ok ? "PASS\n" : "FAIL\n");
    return ok ? 0 : 1;
}
