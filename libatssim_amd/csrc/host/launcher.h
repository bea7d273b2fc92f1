// launcher.h -- task registration API and HEFT scheduler (host, C++).
//
// Same API surface as the reference's Launcher namespace
// (src/heterogeneous/launcher.h:15-1055): solvers register `Task`s that carry
// one `Implementation` per device type, connect them into a DAG, set the
// computation matrix [task][proc] and the communication matrices, and let
// the `Scheduler` order (topological sort, root/terminal standardisation),
// allocate (HEFT: mean upward rank, insertion-based earliest finish time) and
// launch them.  The GPU device slot (DeviceTypeGpu = 1) is filled by the HIP
// backend through the C ABI of include/libatssim_hip.h (see vbd_slot.h).
#pragma once
#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <vector>

namespace Launcher {

using uint = uint32_t;

// Task function ids used by the VBD path and the scheduler's own nodes
// (a subset of FunctionID, launcher.h:15-221).
enum FunctionID : uint {
    id_additional_root,
    id_additional_terminal,
    id_task_heft_2002,
    id_xpbd_constraint_last_node,
    id_vbd_evaluate_inertia,
    id_vbd_evaluate_stretch_mass_spring,
    id_vbd_evaluate_bending,
    id_vbd_step,
    id_vbd_all_in_one,
    id_xpbd_predict_position,
    id_xpbd_reset_constrains,
    id_xpbd_update_velocity,
    id_xpbd_constraint_stretch_mass_spring,
    id_xpbd_constraint_bending,
    id_xpbd_non,
};
extern const std::map<FunctionID, std::string> taskNames;

// launcher.h:435-539
struct LaunchParam {
    FunctionID function_id = id_xpbd_non;
    uint startIdx = 0;
    uint endIdx = 1;
    uint blockDim = 256;
    uint queue_idx = 0;
    uint fence_idx = 0;
    uint iter_idx = 0;
    uint cluster_idx = 0;
    bool launch_indirect = false;
    bool test_time = false;
    bool use_hetero = false;
    bool is_allocated_to_main_device = true;
    uint buffer_idx = 0;
    uint left_buffer_idx = ~0u;
    uint right_buffer_idx = ~0u;
    std::vector<uint> input_buffer_idxs;
    std::vector<uint> output_buffer_idxs;
    uint get_thread_num() const { return endIdx - startIdx; }
};

using DeviceType = uint;
const DeviceType DeviceTypeCpu = 0;
const DeviceType DeviceTypeGpu = 1;
const DeviceType DeviceType0 = 0;
const DeviceType DeviceType1 = 1;
const DeviceType DeviceType2 = 2;

// launcher.h:556-579
struct Implementation {
    DeviceType device = DeviceTypeCpu;
    std::function<void(const LaunchParam&)> implementation = [](const LaunchParam&) {};
    Implementation() = default;
    Implementation(DeviceType d, std::function<void(const LaunchParam&)> f) : device(d), implementation(std::move(f)) {}
    void launch_task(const LaunchParam& p) const { implementation(p); }
    bool operator==(const Implementation& r) const { return device == r.device; }
};

// launcher.h:589-767 (the fields the scheduler and the VBD slot read)
struct Task {
    FunctionID func_id = id_xpbd_non;
    uint num_threads = 0;
    uint start_idx = 0;
    uint block_dim = 256;
    uint iter_idx = ~0u;
    uint constraint_idx = ~0u;
    uint cluster_idx = 0;
    bool can_be_seperate = false;
    bool is_num_threads_constant = true;
    bool is_computation_time_constant = false;
    std::vector<Implementation> list_implementation;

    // asynchronous-iteration fields, set by the post-pass of scheduler_dag
    // (launcher.h:605-615): the position buffer the task iterates, the task
    // (and its buffer) on the same processor it continues from ("left") or
    // hands a copy to ("right"), and the tasks of the other processor whose
    // buffers it merges ("ins")
    uint buffer_left = ~0u, task_left = ~0u;
    uint buffer_right = ~0u, task_right = ~0u;
    uint buffer_idx = ~0u;
    bool is_allocated_to_main_device = true;
    bool is_first_iterative_task = false;
    std::vector<uint> buffer_ins, buffer_outs;
    std::vector<uint> task_ins, task_outs;

    std::vector<uint> predecessors;
    std::vector<uint> successors;
    std::vector<float> list_weight;

    Task() = default;
    Task(FunctionID f, uint numThreads, bool canBeSeparate, bool isNumThreadsConstant, std::vector<Implementation> l)
        : func_id(f), num_threads(numThreads), can_be_seperate(canBeSeparate),
          is_num_threads_constant(isNumThreadsConstant), list_implementation(std::move(l)) {}
    Task(FunctionID f, uint clusterIdx, std::vector<Implementation> l)
        : func_id(f), cluster_idx(clusterIdx), list_implementation(std::move(l)) {}
    Task(FunctionID f, uint iterIdx, uint constraintIdx, uint clusterIdx, std::vector<Implementation> l)
        : func_id(f), iter_idx(iterIdx), constraint_idx(constraintIdx), cluster_idx(clusterIdx),
          list_implementation(std::move(l)) {}

    const Implementation* find_implementation(DeviceType type) const {
        for (const auto& imp : list_implementation)
            if (imp.device == type) return &imp;
        return nullptr;
    }
    bool has_implementation(DeviceType type) const { return find_implementation(type) != nullptr; }
};

// launcher.h:785-814: ordered by end time, then start time
struct ScheduleEvent {
    uint task_id = 0;
    uint proc = 0;
    float start = 0.f;
    float end = 0.f;
    ScheduleEvent() = default;
    ScheduleEvent(uint t, uint p, float s, float e) : task_id(t), proc(p), start(s), end(e) {}
    bool operator<(const ScheduleEvent& r) const { return end != r.end ? end < r.end : start < r.start; }
};

// A run of consecutive tasks of one processor's schedule [start_idx,
// end_idx] that first waits for event `wait` of the other processor and, when
// done, signals (its own index in `signal`) (launcher.h:818-832).
struct LaunchEvent {
    uint start_idx = 0, end_idx = 0, wait = ~0u, signal = ~0u;
    LaunchEvent() = default;
    LaunchEvent(uint s, uint e, uint w, uint g) : start_idx(s), end_idx(e), wait(w), signal(g) {}
};

// How LaunchModeHetero synchronises the two processors (the reference uses a
// MTLSharedEvent and command buffers, launcher.cpp:918-1111).  Processor 0 is
// the calling host thread; processor 1 is an asynchronous queue:
//   gpu_begin(e, wait):  start queue event e; if wait != -1u the queue first
//                        waits until the host has signalled a value >= wait + 1
//   gpu_end(e):          close event e (its completion can be waited for)
//   cpu_wait_gpu(e):     block the host until queue event e has completed;
//                        false if the wait failed or gave up (the launch then
//                        runs no further host event and drains the queue)
//   cpu_signal(v):       host signal value v (monotonic)
//   gpu_finish():        block until every queue event has completed
//   gpu_elapsed_ms(e):   queue time of event e (refits the computation matrix)
struct HeteroHooks {
    std::function<void(uint, uint)> gpu_begin = [](uint, uint) {};
    std::function<void(uint)> gpu_end = [](uint) {};
    std::function<bool(uint)> cpu_wait_gpu = [](uint) { return true; };
    std::function<void(uint)> cpu_signal = [](uint) {};
    std::function<void()> gpu_finish = []() {};
    std::function<float(uint)> gpu_elapsed_ms = [](uint) { return 0.f; };
};

const float inf = 99999.0f;
const uint input_buffer_mask = 0x0FFFFFFF;
const uint default_buffer_mask = 0x0FFFFFFF;

class Scheduler {
public:
    using ListCost = std::vector<float>;
    using ListSchedule = std::vector<ScheduleEvent>;

    enum LaunchMode {
        LaunchModeCpu,            // every task on its CPU implementation, topological/rank order
        LaunchModeGpu,            // every task on its GPU implementation (CPU fallback if missing)
        LaunchModePartialCPU,     // the CPU processor's events only
        LaunchModePartialGPU,     // the GPU processor's events only
        LaunchModeHetero,         // both processors at once: queued GPU events behind host-value waits
        LaunchModeProgressiveHetero,   // Hetero with the GPU events queued window by window
        LaunchModeFakeHetero,     // scheduled allocation, run synchronously in schedule order
        LaunchModeSequeceHetero,  // scheduled order, every task on its CPU implementation
    };

    // ---- DAG construction (launcher.cpp:4404-4471)
    uint add_task(const Task& task);
    Task& get_task_by_tid(uint tid) { return list_task[tid]; }
    void set_connect(uint from, uint to, float weight = 1.0f);
    float delete_connect(uint from, uint to);
    void set_constraint_task_orders(const std::vector<uint>& orders) { constraint_task_orders = orders; }
    const std::vector<Task>& get_list_task() const { return list_task; }
    const std::vector<uint>& get_list_order() const { return list_order; }

    // ---- ordering and allocation
    bool topological_sort();                                     // launcher.cpp:139-192 (DFS)
    void standardizing_dag(const std::vector<std::function<void(const LaunchParam&)>>& empty_impls = {});
    void scheduler_dag();                                        // launcher.cpp:2645-2952 (HEFT)
    float fn_get_communication_cost(uint proc, uint pred_node, uint node) const;
    float fn_get_inner_communication_cost(uint proc) const;

    // Computation matrix from a (function id, cluster) -> per-processor cost
    // table; tasks without an implementation on a processor cost inf there
    // (launcher.cpp:645-704).  cost_total, if empty, receives the per-processor sums.
    void profile_from(const std::vector<std::pair<FunctionID, uint>>& map_tasks,
                      const std::vector<std::vector<double>>& map_costs, std::vector<double>& cost_total);
    // Segments both processors' schedules into LaunchEvents with cross-processor
    // waits and signals (launcher.cpp:3364-3714; two processors only).
    bool make_wait_events();
    const std::vector<std::vector<LaunchEvent>>& get_launch_events() const { return launch_events; }

    // ---- execution
    // LaunchModeHetero needs `hooks`; the other modes run on the calling thread.
    void launch(LaunchMode mode, const std::function<LaunchParam(const Task&)>& task_to_param,
                const HeteroHooks* hooks = nullptr);

    // ---- results / reporting
    float get_scheduled_end_time() const;
    std::vector<float> get_scheduled_speedups() const;   // (single-device cost - makespan) / makespan
    std::string proc_schedules_string() const;
    // The schedule in the reference's Gantt dump format
    // (print_schedule_to_graph_xpbd, launcher.cpp:3915-4126): per processor
    // (colour, task, buffer, start, end), then the cross-processor inputs and
    // the same-processor "left" links as (proc, task, proc, task, end, start).
    std::string schedule_graph_string() const;
    void print_schedule_to_graph_xpbd() const;
    void print_proc_schedules() const;
    void print_speedups_to_each_device() const;
    void update_costs_from_computation_matrix();
    const std::vector<ListSchedule>& get_proc_schedules() const { return proc_schedules; }
    const std::vector<ScheduleEvent>& get_task_schedules() const { return task_schedules; }
    const std::vector<double>& get_ranku() const { return ranku; }
    const std::vector<uint>& get_sorted_nodes() const { return sorted_nodes; }
    void set_print_scheduling_datail(bool b) { print_detail = b; }
    void set_safety_check(bool b) { use_check = b; }
    void reset_scheduler_system();

    std::vector<ListCost> computation_matrix;
    std::vector<double> summary_of_costs_each_device;
    uint progressive_window = 15;   // GPU events queued at a time in LaunchModeProgressiveHetero
    std::vector<std::vector<float>> communication_speed_matrix = {{0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
    std::vector<std::vector<float>> communication_cost_matrix_uma = {{0.002f, 0.220f}, {0.145f, 0.01f}};
    std::vector<float> communication_startup = {0, 0, 0};
    uint root_node = ~0u, terminal_node = ~0u;

private:
    void compute_ranku(uint num_procs);
    void async_post_pass();
    void refit_runtime_cost(const std::vector<float>& cpu_ms, const std::vector<float>& gpu_ms);
    ScheduleEvent compute_eft(uint node, uint proc) const;
    double avg_cost(uint tid, uint num_procs) const;

    std::vector<Task> list_task;
    std::vector<uint> list_order;
    std::vector<uint> sorted_nodes;
    std::vector<uint> constraint_task_orders;
    std::vector<ScheduleEvent> task_schedules;
    std::vector<double> ranku;
    std::vector<ListSchedule> proc_schedules;
    std::vector<std::vector<LaunchEvent>> launch_events;
    std::vector<std::vector<uint>> list_cmd_idx;
    bool print_detail = false;
    bool use_check = true;
};

}  // namespace Launcher
