// Failure handling of the heterogeneous launch modes, on the host only: the
// example2 schedule (10 iterations x 10 clusters, every cost 0.1) launched in
// LaunchModeHetero and LaunchModeProgressiveHetero with synchronisation hooks
// whose k-th host wait for the queue fails (as a timed-out wait does).  The
// launch must run no host event after the failed wait, release the queue
// (signal 0xFFFFFFFF) with no lower signal after it, and drain the queue.
#include <cstdio>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/launcher.h"
#include "host/vbd_slot.h"

static int run(Launcher::Scheduler::LaunchMode mode, const char* name, unsigned fail_at) {
    Launcher::Scheduler s;
    unsigned cpu_tasks = 0, cpu_after_failure = 0, waits = 0, finishes = 0;
    bool failed = false;
    std::vector<unsigned> signals;
    Launcher::Implementation cpu(Launcher::DeviceTypeCpu, [&](const Launcher::LaunchParam&) {
        cpu_tasks++;
        if (failed) cpu_after_failure++;
    });
    Launcher::Implementation gpu(Launcher::DeviceTypeGpu, [&](const Launcher::LaunchParam&) {});
    lats::register_vbd_dag(s, {cpu, gpu}, 10, 10);
    std::vector<std::pair<Launcher::FunctionID, Launcher::uint>> keys = {
        {Launcher::id_xpbd_predict_position, 0}, {Launcher::id_xpbd_update_velocity, 0},
        {Launcher::id_xpbd_constraint_last_node, 0}, {Launcher::id_vbd_all_in_one, 0}};
    for (unsigned c = 0; c < 10; c++) keys.push_back({Launcher::id_vbd_all_in_one, c});
    std::vector<std::vector<double>> costs(keys.size(), {0.1f, 0.1f});
    std::vector<double> total;
    s.profile_from(keys, costs, total);
    s.communication_cost_matrix_uma = {{0.0f, 0.1f}, {0.1f, 0.0f}};
    s.communication_speed_matrix = {};
    s.communication_startup = {0, 0};
    if (!s.topological_sort()) { std::printf("FAIL topological sort\n"); return 1; }
    s.standardizing_dag();
    s.scheduler_dag();
    if (!s.make_wait_events()) { std::printf("FAIL wait events\n"); return 1; }
    unsigned host_waits = 0;
    for (const auto& e : s.get_launch_events()[0]) host_waits += e.wait != ~0u;
    if (host_waits <= fail_at) { std::printf("FAIL schedule has only %u host waits\n", host_waits); return 1; }

    Launcher::HeteroHooks h;
    h.cpu_wait_gpu = [&](Launcher::uint) {
        if (waits++ == fail_at) { failed = true; return false; }
        return true;
    };
    h.cpu_signal = [&](Launcher::uint v) { signals.push_back(v); };
    h.gpu_finish = [&]() { finishes++; };
    s.launch(mode, lats::vbd_task_to_param, &h);

    bool released = false, lowered = false;
    for (unsigned v : signals) {
        if (released && v != 0xFFFFFFFFu) lowered = true;
        if (v == 0xFFFFFFFFu) released = true;
    }
    const bool ok = failed && cpu_after_failure == 0 && released && !lowered && finishes >= 1 && waits == fail_at + 1;
    std::printf("%s: %u host tasks before the failed wait %u, %u after it, released %d, lowered %d, drained %u\n",
                name, cpu_tasks, fail_at, cpu_after_failure, (int)released, (int)lowered, finishes);
    return ok ? 0 : 1;
}

int main() {
    int bad = 0;
    for (unsigned k : {0u, 2u, 5u}) {
        bad += run(Launcher::Scheduler::LaunchModeHetero, "Hetero", k);
        bad += run(Launcher::Scheduler::LaunchModeProgressiveHetero, "ProgressiveHetero", k);
    }
    std::printf(bad ? "FAIL\n" : "PASS\n");
    return bad ? 1 : 0;
}
