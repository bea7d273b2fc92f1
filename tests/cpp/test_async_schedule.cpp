// The asynchronous-iteration schedule of the reference's example2
// (examples/example2_async_iter.cpp:1305-1410): the per-frame VBD DAG with
// every task costing 0.1 on both processors and 0.1 to cross between them,
// scheduled by HEFT and the async post-pass.  Prints the schedule in the
// reference's Gantt dump format (compared by tests/test_async.py with the
// dumps in documents/example2_scheduling_result.py) and the launch events,
// then runs the schedule in LaunchModeSequeceHetero and checks that every
// task ran once and after all of its predecessors (including the
// cross-processor edges the post-pass adds).
//
// usage: test_async_schedule <iterations> <clusters>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/launcher.h"
#include "host/vbd_slot.h"

int main(int argc, char** argv) {
    const unsigned iters = argc > 1 ? (unsigned)std::atoi(argv[1]) : 1;
    const unsigned clusters = argc > 2 ? (unsigned)std::atoi(argv[2]) : 10;
    Launcher::Scheduler s;
    std::vector<unsigned> ran_order;
    std::vector<Launcher::LaunchParam> params;
    auto record = [&](const Launcher::LaunchParam& p) { params.push_back(p); };
    Launcher::Implementation cpu(Launcher::DeviceTypeCpu, record), gpu(Launcher::DeviceTypeGpu, record);
    lats::register_vbd_dag(s, {cpu, gpu}, iters, clusters);

    // example2_async_iter.cpp:1375-1402
    std::vector<std::pair<Launcher::FunctionID, Launcher::uint>> keys = {
        {Launcher::id_xpbd_predict_position, 0},
        {Launcher::id_xpbd_update_velocity, 0},
        {Launcher::id_xpbd_constraint_last_node, 0},
        {Launcher::id_vbd_all_in_one, 0},
    };
    for (unsigned c = 0; c < clusters; c++) keys.push_back({Launcher::id_vbd_all_in_one, c});
    std::vector<std::vector<double>> costs(keys.size(), {0.1f, 0.1f});
    std::vector<double> total;
    s.profile_from(keys, costs, total);
    s.communication_cost_matrix_uma = {{0.0f, 0.1f}, {0.1f, 0.0f}};
    s.communication_speed_matrix = {};
    s.communication_startup = {0, 0};
    if (!s.topological_sort()) {
        std::printf("FAIL topological sort\n");
        return 1;
    }
    s.standardizing_dag();
    s.scheduler_dag();
    if (!s.make_wait_events()) {
        std::printf("FAIL wait events\n");
        return 1;
    }
    std::printf("GRAPH_BEGIN\n%sGRAPH_END\n", s.schedule_graph_string().c_str());
    const auto& evs = s.get_launch_events();
    for (unsigned p = 0; p < 2; p++) {
        std::printf("EVENTS %u:", p);
        for (const auto& e : evs[p]) std::printf(" [%u-%u w%d s%d]", e.start_idx, e.end_idx, (int)e.wait, (int)e.signal);
        std::printf("\n");
    }

    // launch order: map each recorded param back to its task through the
    // (function, iteration, cluster) key
    const auto& tasks = s.get_list_task();
    s.launch(Launcher::Scheduler::LaunchModeSequeceHetero, lats::vbd_task_to_param);
    std::vector<int> when(tasks.size(), -1);
    for (size_t k = 0; k < params.size(); k++) {
        const auto& p = params[k];
        int found = -1;
        for (size_t t = 0; t < tasks.size(); t++)
            if (tasks[t].func_id == p.function_id && tasks[t].iter_idx == p.iter_idx &&
                tasks[t].cluster_idx == p.cluster_idx && when[t] < 0) {
                found = (int)t;
                break;
            }
        if (found < 0) {
            std::printf("FAIL launched task not in the DAG (fn %u)\n", (unsigned)p.function_id);
            return 1;
        }
        when[found] = (int)k;
    }
    std::printf("ORDER:");
    for (size_t k = 0; k < params.size(); k++)
        for (size_t t = 0; t < tasks.size(); t++)
            if (when[t] == (int)k) std::printf(" %zu", t);
    std::printf("\n");
    // the launch sequence for replays: k, scheduled processor, then the
    // LaunchParam fields (C-ABI function id, iteration, cluster, buffer,
    // left, right, main device, inputs)
    auto c_fn = [](Launcher::FunctionID f) -> int {
        switch (f) {
            case Launcher::id_xpbd_predict_position: return LATS_FN_PREDICT_POSITION;
            case Launcher::id_vbd_all_in_one: return LATS_FN_VBD_ALL_IN_ONE;
            case Launcher::id_xpbd_constraint_last_node: return LATS_FN_CONSTRAINT_LAST_NODE;
            case Launcher::id_xpbd_update_velocity: return LATS_FN_UPDATE_VELOCITY;
            default: return -1;
        }
    };
    for (size_t k = 0; k < params.size(); k++)
        for (size_t t = 0; t < tasks.size(); t++) {
            if (when[t] != (int)k) continue;
            const auto& p = params[k];
            std::printf("PARAM %zu %u %d %u %u %u %u %u %d %zu", k, s.get_task_schedules()[t].proc, c_fn(p.function_id),
                        p.iter_idx, p.cluster_idx, p.buffer_idx, p.left_buffer_idx, p.right_buffer_idx,
                        p.is_allocated_to_main_device ? 1 : 0, p.input_buffer_idxs.size());
            for (unsigned b : p.input_buffer_idxs) std::printf(" %u", b);
            std::printf("\n");
        }
    for (size_t t = 0; t < tasks.size(); t++) {
        // the root and terminal the scheduler adds carry empty implementations
        if (tasks[t].func_id == Launcher::id_additional_root || tasks[t].func_id == Launcher::id_additional_terminal)
            continue;
        if (when[t] < 0) {
            std::printf("FAIL task %zu never ran\n", t);
            return 1;
        }
        for (unsigned pred : tasks[t].predecessors)
            if (when[pred] > when[t]) {
                std::printf("FAIL task %zu ran before its predecessor %u\n", t, pred);
                return 1;
            }
    }
    // the partial modes run one processor's scheduled tasks
    const size_t n_all = params.size();
    size_t on_proc[2] = {0, 0};
    for (size_t t = 0; t < tasks.size(); t++)
        if (tasks[t].func_id != Launcher::id_additional_root && tasks[t].func_id != Launcher::id_additional_terminal)
            on_proc[s.get_task_schedules()[t].proc]++;
    params.clear();
    s.launch(Launcher::Scheduler::LaunchModePartialCPU, lats::vbd_task_to_param);
    const size_t n_cpu = params.size();
    params.clear();
    s.launch(Launcher::Scheduler::LaunchModePartialGPU, lats::vbd_task_to_param);
    const size_t n_gpu = params.size();
    std::printf("PARTIAL cpu %zu of %zu, gpu %zu of %zu\n", n_cpu, on_proc[0], n_gpu, on_proc[1]);
    if (n_cpu != on_proc[0] || n_gpu != on_proc[1] || n_cpu + n_gpu != n_all) {
        std::printf("FAIL partial launch modes\n");
        return 1;
    }
    std::printf("PASS %zu tasks\n", n_all);
    return 0;
}
