// launcher.cpp -- see launcher.h.
#include "launcher.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <set>
#include <sstream>

namespace Launcher {

const std::map<FunctionID, std::string> taskNames = {
    {id_additional_root, "additional_root"},
    {id_additional_terminal, "additional_terminal"},
    {id_task_heft_2002, "task_heft_2002"},
    {id_xpbd_constraint_last_node, "xpbd_constraint_last_node"},
    {id_vbd_evaluate_inertia, "vbd_evaluate_inertia"},
    {id_vbd_evaluate_stretch_mass_spring, "vbd_evaluate_stretch_mass_spring"},
    {id_vbd_evaluate_bending, "vbd_evaluate_bending"},
    {id_vbd_step, "vbd_step"},
    {id_vbd_all_in_one, "vbd_all_in_one"},
    {id_xpbd_predict_position, "xpbd_predict_position"},
    {id_xpbd_reset_constrains, "xpbd_reset_constrains"},
    {id_xpbd_update_velocity, "xpbd_update_velocity"},
    {id_xpbd_constraint_stretch_mass_spring, "xpbd_constraint_stretch_mass_spring"},
    {id_xpbd_constraint_bending, "xpbd_constraint_bending"},
    {id_xpbd_non, "xpbd_non"},
};

uint Scheduler::add_task(const Task& task) {
    list_task.push_back(task);
    return (uint)list_task.size() - 1;
}

void Scheduler::set_connect(uint from, uint to, float weight) {
    list_task[from].successors.push_back(to);
    list_task[from].list_weight.push_back(weight);
    list_task[to].predecessors.push_back(from);
}

float Scheduler::delete_connect(uint from, uint to) {
    auto& a = list_task[from];
    auto& b = list_task[to];
    auto it = std::find(a.successors.begin(), a.successors.end(), to);
    if (it == a.successors.end()) return 0.f;
    const size_t k = (size_t)(it - a.successors.begin());
    const float w = a.list_weight[k];
    a.successors.erase(it);
    a.list_weight.erase(a.list_weight.begin() + (long)k);
    auto jt = std::find(b.predecessors.begin(), b.predecessors.end(), from);
    if (jt != b.predecessors.end()) b.predecessors.erase(jt);
    return w;
}

// Reverse post-order of a depth-first search started from every unvisited
// task in id order, successors in insertion order (launcher.cpp:29-71,
// 139-192).  Iterative, so long iteration chains do not exhaust the stack.
bool Scheduler::topological_sort() {
    const uint n = (uint)list_task.size();
    std::vector<uint8_t> state(n, 0);   // 0 new, 1 on the DFS path, 2 done
    std::vector<uint> post;
    post.reserve(n);
    std::vector<std::pair<uint, uint>> stack;
    for (uint s = 0; s < n; s++) {
        if (state[s]) continue;
        stack.push_back({s, 0});
        state[s] = 1;
        while (!stack.empty()) {
            auto& top = stack.back();
            const auto& succ = list_task[top.first].successors;
            if (top.second < succ.size()) {
                const uint nx = succ[top.second++];
                if (state[nx] == 0) {
                    state[nx] = 1;
                    stack.push_back({nx, 0});
                } else if (state[nx] == 1) {
                    std::fprintf(stderr, "Launcher: graph contains a cycle through task %u\n", nx);
                    list_order.clear();
                    return false;
                }
            } else {
                state[top.first] = 2;
                post.push_back(top.first);
                stack.pop_back();
            }
        }
    }
    list_order.assign(post.rbegin(), post.rend());
    return true;
}

// launcher.cpp:1588-1719: a zero-cost root before every source and a
// zero-cost terminal after every sink.
void Scheduler::standardizing_dag(const std::vector<std::function<void(const LaunchParam&)>>& empty_impls) {
    const uint num_procs = (uint)communication_startup.size();
    std::vector<uint> roots;
    for (uint t = 0; t < list_task.size(); t++)
        if (list_task[t].predecessors.empty()) roots.push_back(t);
    std::vector<Implementation> imps;
    for (uint p = 0; p < num_procs; p++)
        imps.emplace_back(p, p < empty_impls.size() ? empty_impls[p] : [](const LaunchParam&) {});
    root_node = add_task(Task(id_additional_root, 0, false, true, imps));
    computation_matrix.push_back(ListCost(num_procs, 0.0f));
    list_order.insert(list_order.begin(), root_node);
    for (uint t : roots)
        if (t != root_node) set_connect(root_node, t, 0.0f);
    std::vector<uint> sinks;
    for (uint t = 0; t < list_task.size(); t++)
        if (list_task[t].successors.empty()) sinks.push_back(t);
    terminal_node = add_task(Task(id_additional_terminal, 0, false, true, imps));
    computation_matrix.push_back(ListCost(num_procs, 0.0f));
    list_order.push_back(terminal_node);
    for (uint t : sinks)
        if (t != terminal_node) set_connect(t, terminal_node, 0.0f);
}

double Scheduler::avg_cost(uint tid, uint num_procs) const {
    double s = 0.0, n = 0.0;
    for (uint p = 0; p < num_procs; p++) {
        const float c = computation_matrix[tid][p];
        if (c != inf) { s += c; n += 1.0; }
    }
    return s / n;
}

// Upward rank with the mean metric (launcher.cpp:1723-2366, non-merged
// branch).  The averaged edge cost is sum(startup) + (1 + sum(startup)) for
// every edge, independent of the edge weight, as in the reference.
void Scheduler::compute_ranku(uint num_procs) {
    const uint n = (uint)list_task.size();
    double startup = 0.0;
    for (uint p = 0; p < num_procs; p++) startup += communication_startup[p];
    const double avg_comm = 1.0 + startup;
    const double edge_w = startup + avg_comm;
    ranku.assign(n, 0.0);
    ranku[terminal_node] = avg_cost(terminal_node, num_procs);
    auto from_successors = [&](uint node) {
        double best = -1.0;
        for (uint s : list_task[node].successors) best = std::max(best, edge_w + ranku[s]);
        return avg_cost(node, num_procs) + best;
    };
    for (uint i = 1; i < n; i++) {
        const uint node = list_order[n - 1 - i];
        ranku[node] = from_successors(node);
    }
    ranku[root_node] = inf;
    if (!constraint_task_orders.empty()) {
        // chain the iterative tasks in registration order so that rank-u
        // strictly decreases along it (launcher.cpp:2312-2362)
        std::vector<bool> visited(n, false);
        const uint m = (uint)constraint_task_orders.size();
        visited[constraint_task_orders.back()] = true;
        for (uint i = 1; i < m; i++) {
            const uint cur = constraint_task_orders[m - 1 - i], nxt = constraint_task_orders[m - i];
            ranku[cur] = avg_cost(cur, num_procs) + edge_w + ranku[nxt];
            visited[cur] = true;
        }
        for (uint i = 1; i < n; i++) {
            const uint node = list_order[n - 1 - i];
            if (!visited[node]) ranku[node] = from_successors(node);
        }
        ranku[root_node] = inf;
    }
}

float Scheduler::fn_get_communication_cost(uint proc, uint pred, uint node) const {
    const Task& pt = list_task[pred];
    const ScheduleEvent& pj = task_schedules[pred];
    auto it = std::find(pt.successors.begin(), pt.successors.end(), node);
    const float w = pt.list_weight[(size_t)(it - pt.successors.begin())];
    if (w == 0.f) return 0.f;
    if (!communication_cost_matrix_uma.empty()) return communication_cost_matrix_uma[pj.proc][proc];
    if (pj.proc == proc) return communication_speed_matrix[pj.proc][proc];
    return w / communication_speed_matrix[pj.proc][proc];
}

float Scheduler::fn_get_inner_communication_cost(uint proc) const {
    return !communication_cost_matrix_uma.empty() ? communication_cost_matrix_uma[proc][proc]
                                                  : communication_speed_matrix[proc][proc];
}

// Earliest finish time of `node` on `proc` with insertion into idle slots
// (launcher.cpp:2368-2404 + 2537-2604).
ScheduleEvent Scheduler::compute_eft(uint node, uint proc) const {
    float ready = 0.f;
    for (uint pred : list_task[node].predecessors)
        ready = std::max(ready, task_schedules[pred].end + fn_get_communication_cost(proc, pred, node));
    const float comp = computation_matrix[node][proc];
    const float inner = fn_get_inner_communication_cost(proc);
    const ListSchedule& s = proc_schedules[proc];
    for (size_t i = 0; i < s.size(); i++) {
        const ScheduleEvent& prev = s[i];
        if (i == 0 && prev.start > ready + comp + inner) return ScheduleEvent(node, proc, ready, ready + comp);
        if (i + 1 == s.size()) {
            const float st = std::max(ready, prev.end + inner);
            return ScheduleEvent(node, proc, st, st + comp);
        }
        const float st = std::max(ready, prev.end + inner);
        if (s[i + 1].start - st > comp) return ScheduleEvent(node, proc, st, st + comp);
    }
    return ScheduleEvent(node, proc, ready, ready + comp);
}

// HEFT (launcher.cpp:2645-2952, EFT operating mode): tasks by decreasing
// rank-u (ties in task-id order), each on the processor with the earliest
// finish time.
void Scheduler::scheduler_dag() {
    const uint num_procs = (uint)communication_startup.size();
    const uint n = (uint)list_task.size();
    compute_ranku(num_procs);
    sorted_nodes.resize(n);
    for (uint t = 0; t < n; t++) sorted_nodes[t] = t;
    std::stable_sort(sorted_nodes.begin(), sorted_nodes.end(), [&](uint a, uint b) { return ranku[a] > ranku[b]; });
    if (use_check) {
        std::vector<uint> pos(n);
        for (uint i = 0; i < n; i++) pos[sorted_nodes[i]] = i;
        for (uint i = 0; i < n; i++)
            for (uint p : list_task[sorted_nodes[i]].predecessors)
                if (pos[p] >= i)
                    std::fprintf(stderr, "Launcher: task %u ranked before its predecessor %u\n", sorted_nodes[i], p);
    }
    if (sorted_nodes[0] != root_node) {
        auto it = std::find(sorted_nodes.begin(), sorted_nodes.end(), root_node);
        std::iter_swap(sorted_nodes.begin(), it);
    }
    task_schedules.assign(n, ScheduleEvent());
    proc_schedules.assign(num_procs, ListSchedule());
    for (uint node : sorted_nodes) {
        ScheduleEvent best(node, 0, inf, inf);
        for (uint p = 0; p < num_procs; p++) {
            if (!(computation_matrix[node][p] < inf)) continue;
            const ScheduleEvent ev = compute_eft(node, p);
            if (ev.end < best.end) best = ev;
        }
        task_schedules[node] = best;
        ListSchedule& ps = proc_schedules[best.proc];
        ps.push_back(best);
        std::sort(ps.begin(), ps.end());
        if (print_detail)
            std::printf("  task %u -> proc %u [%g, %g] (rank-u %g)\n", node, best.proc, best.start, best.end, ranku[node]);
    }
    if (!constraint_task_orders.empty() && !communication_cost_matrix_uma.empty() && num_procs == 2) async_post_pass();
}

namespace {

// Task classes of the asynchronous post-pass (launcher.cpp:2957-3030),
// restricted to the function ids this library registers.
bool is_constraint_fn(FunctionID f) {
    return f == id_xpbd_constraint_last_node || f == id_xpbd_constraint_stretch_mass_spring ||
           f == id_xpbd_constraint_bending || f == id_vbd_evaluate_inertia ||
           f == id_vbd_evaluate_stretch_mass_spring || f == id_vbd_evaluate_bending || f == id_vbd_step ||
           f == id_vbd_all_in_one;
}
bool is_cloth_fn(FunctionID f) { return is_constraint_fn(f); }
bool is_tet_fn(FunctionID f) { return f == id_xpbd_constraint_last_node || f == id_vbd_all_in_one; }

constexpr uint kMainDevice = 1;   // the processor that merges every input (launcher.cpp:3093-3094)

}  // namespace

// Asynchronous iteration (launcher.cpp:2953-3334).  After HEFT, every
// iterative task is linked to the first iterative task of the other processor
// that starts after it has finished plus the transfer time; the main device
// keeps every such input (dropping one whose producer had no input of its
// own when a later one follows), the other device keeps only the last.  The
// links become DAG edges.  Then both schedules are walked together in start
// order to give every iterative task its position buffer: a task with inputs
// iterates in the buffer of its last input and merges the processor's
// previous buffer ("left") into it; a task without inputs continues in the
// previous task's buffer unless that one is handed to the other processor,
// in which case it takes a fresh buffer copied from it ("right" of the
// previous task); the first task of a processor copies the predicted
// positions (left = input_buffer_mask, task_left = the predict task).
void Scheduler::async_post_pass() {
    const uint n = (uint)list_task.size();
    const uint num_procs = 2;
    std::vector<bool> constraint(n), cloth(n), tet(n);
    uint predict_tid = ~0u;
    for (uint t = 0; t < n; t++) {
        const FunctionID f = list_task[t].func_id;
        constraint[t] = is_constraint_fn(f);
        cloth[t] = is_cloth_fn(f);
        tet[t] = is_tet_fn(f);
        if (f == id_xpbd_predict_position) predict_tid = t;
    }

    std::vector<std::vector<uint>> ins(n), outs(n);
    for (uint p = 0; p < num_procs; p++) {
        const ListSchedule& mine = proc_schedules[p];
        const ListSchedule& other = proc_schedules[1 - p];
        const float comm = communication_cost_matrix_uma[p][1 - p];
        for (const ScheduleEvent& ev : mine) {
            const uint t = ev.task_id;
            if (!constraint[t]) continue;
            for (const ScheduleEvent& ov : other) {
                const uint u = ov.task_id;
                if (!constraint[u] || !(ev.end + comm <= ov.start)) continue;
                if ((cloth[t] && cloth[u]) || (tet[t] && tet[u])) {
                    outs[t] = {u};
                    ins[u].push_back(t);
                    break;
                }
            }
        }
    }
    auto drop_out = [&](uint from, uint to) {
        auto it = std::find(outs[from].begin(), outs[from].end(), to);
        if (it != outs[from].end()) outs[from].erase(it);
    };
    for (uint p = 0; p < num_procs; p++) {
        if (p == kMainDevice) continue;
        for (const ScheduleEvent& ev : proc_schedules[p]) {
            const uint t = ev.task_id;
            if (!constraint[t]) continue;
            list_task[t].is_allocated_to_main_device = false;
            if (ins[t].size() > 1) {
                for (size_t j = 0; j + 1 < ins[t].size(); j++) drop_out(ins[t][j], t);
                ins[t] = {ins[t].back()};
            }
            if (!ins[t].empty()) set_connect(ins[t].back(), t);
        }
    }
    for (const ScheduleEvent& ev : proc_schedules[kMainDevice]) {
        const uint t = ev.task_id;
        if (!constraint[t]) continue;
        if (ins[t].size() > 1) {
            list_task[t].is_allocated_to_main_device = true;
            std::vector<uint> kept;
            for (size_t j = 0; j + 1 < ins[t].size(); j++) {
                if (ins[ins[t][j + 1]].empty()) drop_out(ins[t][j], t);
                else kept.push_back(ins[t][j]);
            }
            kept.push_back(ins[t].back());
            ins[t] = kept;
        }
        for (uint in : ins[t]) set_connect(in, t);
    }

    uint next_buffer = 0;
    std::vector<uint> buffer_of(n, ~0u);
    std::vector<uint> last_on(num_procs, ~0u);
    const ListSchedule& cpu = proc_schedules[0];
    const ListSchedule& gpu = proc_schedules[1];
    if (!cpu.empty()) list_task[cpu.front().task_id].is_first_iterative_task = true;
    if (!gpu.empty()) list_task[gpu.front().task_id].is_first_iterative_task = true;
    size_t ic = 0, ig = 0;
    while (ic < cpu.size() || ig < gpu.size()) {
        uint p;
        const ScheduleEvent* ev;
        if (ic < cpu.size() && (ig >= gpu.size() || cpu[ic].start <= gpu[ig].start)) {
            p = 0;
            ev = &cpu[ic++];
        } else {
            p = 1;
            ev = &gpu[ig++];
        }
        const uint t = ev->task_id;
        if (!constraint[t]) continue;
        Task& task = list_task[t];
        uint buf = ~0u;
        const uint left = last_on[p];
        if (!ins[t].empty()) {
            for (uint in : ins[t]) {
                const uint b = buffer_of[in];
                buf = b;
                task.buffer_ins.push_back(b);
                task.task_ins.push_back(in);
                list_task[in].buffer_outs.push_back(b);
                list_task[in].task_outs.push_back(t);
                // the main device always merges its previous buffer; the
                // other one only when that buffer was not handed over
                if (left != ~0u && (outs[left].empty() || p == 1)) {
                    task.buffer_left = buffer_of[left];
                    task.task_left = left;
                }
            }
        } else if (left != ~0u && outs[left].empty()) {
            buf = buffer_of[left];
        } else {
            buf = next_buffer++;
            if (left != ~0u) {
                task.buffer_left = buffer_of[left];
                task.task_left = left;
                list_task[left].buffer_right = buf;
                list_task[left].task_right = t;
            } else {
                task.buffer_left = input_buffer_mask;
                task.task_left = predict_tid;
            }
        }
        buffer_of[t] = buf;
        task.buffer_idx = buf;
        last_on[p] = t;
    }
}

void Scheduler::profile_from(const std::vector<std::pair<FunctionID, uint>>& map_tasks,
                             const std::vector<std::vector<double>>& map_costs, std::vector<double>& cost_total) {
    const uint num_procs = map_costs.empty() ? 0u : (uint)map_costs.back().size();
    std::vector<double> sum(num_procs, 0.0);
    computation_matrix.assign(list_task.size(), ListCost(num_procs, inf));
    for (uint t = 0; t < list_task.size(); t++) {
        const Task& task = list_task[t];
        std::vector<double> costs(num_procs, 0.0);
        bool found = false;
        for (size_t i = 0; i < map_tasks.size() && !found; i++)
            if (map_tasks[i].first == task.func_id && map_tasks[i].second == task.cluster_idx) {
                costs = map_costs[i];
                found = true;
            }
        if (!found)
            std::fprintf(stderr, "Launcher: no cost for task %s / %u\n", taskNames.at(task.func_id).c_str(),
                         task.cluster_idx);
        for (uint p = 0; p < num_procs; p++) {
            computation_matrix[t][p] = task.has_implementation(p) ? (float)costs[p] : inf;
            sum[p] += costs[p];
        }
    }
    if (cost_total.empty()) cost_total = sum;
    summary_of_costs_each_device = cost_total;
}

// launcher.cpp:3364-3714.  For every task, the latest-finishing predecessor
// on the other processor becomes a wait (unless an earlier task of this
// processor already waits for it or for a later task there); waits start a
// new event, signals end one, then wait/signal offsets become event indices.
bool Scheduler::make_wait_events() {
    const uint num_procs = (uint)communication_startup.size();
    if (num_procs != 2 || proc_schedules.size() != 2) {
        std::fprintf(stderr, "Launcher: wait events need exactly two processors\n");
        return false;
    }
    const uint n = (uint)list_task.size();
    std::vector<uint> proc_of(n, 0), offset_of(n, 0);
    for (uint p = 0; p < num_procs; p++)
        for (uint i = 0; i < proc_schedules[p].size(); i++) {
            proc_of[proc_schedules[p][i].task_id] = p;
            offset_of[proc_schedules[p][i].task_id] = i;
        }
    std::vector<std::vector<uint>> waits(num_procs), signals(num_procs);
    for (uint p = 0; p < num_procs; p++) {
        waits[p].assign(proc_schedules[p].size(), ~0u);
        signals[p].assign(proc_schedules[p].size(), ~0u);
    }
    bool ok = true;
    for (uint p = 0; p < num_procs; p++) {
        std::vector<uint> waited;
        for (uint i = 0; i < proc_schedules[p].size(); i++) {
            const ScheduleEvent& job = proc_schedules[p][i];
            uint pred_offset = ~0u;
            float last_end = 0.f;
            for (uint pred : list_task[job.task_id].predecessors) {
                if (proc_of[pred] == p) continue;
                const ScheduleEvent& pj = task_schedules[pred];
                if (pj.end > last_end) {
                    last_end = pj.end;
                    pred_offset = offset_of[pred];
                    if (pj.end > job.start) {
                        std::fprintf(stderr, "Launcher: schedule misses dependency %u -> %u\n", pred, job.task_id);
                        ok = false;
                    }
                }
            }
            if (pred_offset == ~0u || signals[1 - p][pred_offset] != ~0u) continue;
            bool redundant = false;
            for (uint w : waited)
                if (w >= pred_offset) { redundant = true; break; }
            if (redundant) continue;
            waits[p][i] = pred_offset;
            signals[1 - p][pred_offset] = i;
            waited.push_back(pred_offset);
        }
    }
    launch_events.assign(num_procs, {});
    list_cmd_idx.assign(num_procs, {});
    for (uint p = 0; p < num_procs; p++) {
        const uint m = (uint)proc_schedules[p].size();
        std::vector<LaunchEvent>& evs = launch_events[p];
        std::vector<uint>& cmd = list_cmd_idx[p];
        cmd.assign(m, 0);
        uint cmd_idx = 0, start = 0, end = 0, wait_at = ~0u;
        for (uint i = 0; i < m; i++) {
            if (waits[p][i] != ~0u) {
                if (i != 0 && end != i - 1) {   // close the open run before the wait
                    end = i - 1;
                    evs.emplace_back(start, end, wait_at, ~0u);
                    cmd_idx++;
                }
                wait_at = waits[p][i];
                start = i;
            }
            cmd[i] = cmd_idx;
            if (signals[p][i] != ~0u || i == m - 1) {
                end = i;
                evs.emplace_back(start, end, wait_at, signals[p][i]);
                cmd_idx++;
                wait_at = ~0u;
                start = i + 1;
            }
        }
    }
    for (uint p = 0; p < num_procs; p++)
        for (uint e = 0; e < launch_events[p].size(); e++) {
            const uint target = launch_events[p][e].signal;
            if (target == ~0u) continue;
            launch_events[p][e].signal = e;
            launch_events[1 - p][list_cmd_idx[1 - p][target]].wait = e;
        }
    return ok;
}

// launcher.cpp:711-793: spread the measured time of every event over its
// tasks whose computation time is not constant (new = (2 old + delta) / 2).
void Scheduler::refit_runtime_cost(const std::vector<float>& cpu_ms, const std::vector<float>& gpu_ms) {
    for (uint p = 0; p < 2; p++) {
        const std::vector<float>& ms = p == 0 ? cpu_ms : gpu_ms;
        for (uint e = 0; e < launch_events[p].size() && e < ms.size(); e++) {
            const LaunchEvent& ev = launch_events[p][e];
            float profiled = 0.f;
            uint unfixed = 0;
            for (uint i = ev.start_idx; i <= ev.end_idx; i++) {
                const uint t = proc_schedules[p][i].task_id;
                if (!list_task[t].is_computation_time_constant) unfixed++;
                profiled += computation_matrix[t][p];
                if (i != ev.end_idx) profiled += communication_cost_matrix_uma[p][p];
            }
            const float delta = (ms[e] - profiled) / (float)unfixed;
            for (uint i = ev.start_idx; i <= ev.end_idx; i++) {
                const uint t = proc_schedules[p][i].task_id;
                if (!list_task[t].is_computation_time_constant)
                    computation_matrix[t][p] = 0.5f * std::max(computation_matrix[t][p] + computation_matrix[t][p] + delta, 0.0f);
            }
        }
    }
}

void Scheduler::launch(LaunchMode mode, const std::function<LaunchParam(const Task&)>& task_to_param,
                       const HeteroHooks* hooks) {
    const std::vector<uint>& order = !sorted_nodes.empty() ? sorted_nodes : list_order;
    // runs one scheduled task on `dev`; false (and a message) if it has no such implementation
    auto run = [&](uint proc, uint i, DeviceType dev) -> bool {
        const Task& task = list_task[proc_schedules[proc][i].task_id];
        const Implementation* imp = task.find_implementation(dev);
        if (!imp) {
            std::fprintf(stderr, "Launcher: task %u (%s) has no implementation for device %u\n",
                         proc_schedules[proc][i].task_id, taskNames.at(task.func_id).c_str(), dev);
            return false;
        }
        imp->launch_task(task_to_param(task));
        return true;
    };
    auto run_event = [&](uint proc, const LaunchEvent& ev, DeviceType dev) -> bool {
        for (uint i = ev.start_idx; i <= ev.end_idx; i++)
            if (!run(proc, i, dev)) return false;
        return true;
    };
    using clock = std::chrono::steady_clock;
    auto ms_since = [](clock::time_point t0) {
        return std::chrono::duration<float, std::milli>(clock::now() - t0).count();
    };
    const bool need_events = mode != LaunchModeCpu && mode != LaunchModeGpu;
    if (need_events && (launch_events.size() != 2 || (launch_events[0].empty() && launch_events[1].empty()))) {
        std::fprintf(stderr, "Launcher: launching order is empty (make_wait_events first)\n");
        return;
    }
    switch (mode) {
        case LaunchModeCpu:
            for (uint t : order) {
                const Implementation* imp = list_task[t].find_implementation(DeviceTypeCpu);
                if (!imp) { std::fprintf(stderr, "Launcher: task %u has no CPU implementation\n", t); return; }
                imp->launch_task(task_to_param(list_task[t]));
            }
            break;
        case LaunchModeGpu:
            for (uint t : order) {
                const Implementation* imp = list_task[t].find_implementation(DeviceTypeGpu);
                if (!imp) imp = list_task[t].find_implementation(DeviceTypeCpu);
                if (imp) imp->launch_task(task_to_param(list_task[t]));
            }
            break;
        case LaunchModeFakeHetero:
        case LaunchModeSequeceHetero: {
            // launcher.cpp:1269-1473: the CPU events in order; before a CPU
            // event that waits for GPU event w, every GPU event up to w runs
            // (and, in FakeHetero, completes); the rest of the GPU events last.
            // SequeceHetero runs the GPU events with the CPU implementation.
            const bool fake = mode == LaunchModeFakeHetero;
            const DeviceType gdev = fake ? DeviceTypeGpu : DeviceTypeCpu;
            const auto& cev = launch_events[0];
            const auto& gev = launch_events[1];
            std::vector<float> cpu_ms(cev.size(), 0.f), gpu_ms(gev.size(), 0.f);
            auto run_gpu = [&](uint from, uint to) -> bool {
                for (uint e = from; e <= to && e < gev.size(); e++) {
                    if (fake && hooks) hooks->gpu_begin(e, ~0u);
                    if (!run_event(1, gev[e], gdev)) return false;
                    if (fake && hooks) {
                        hooks->gpu_end(e);
                        if (!hooks->cpu_wait_gpu(e)) return false;
                        gpu_ms[e] = hooks->gpu_elapsed_ms(e);
                    }
                }
                return true;
            };
            uint next_gpu = 0;
            for (uint e = 0; e < cev.size(); e++) {
                if (cev[e].wait != ~0u) {
                    if (!run_gpu(next_gpu, cev[e].wait)) return;
                    next_gpu = cev[e].wait + 1;
                }
                const auto t0 = clock::now();
                if (!run_event(0, cev[e], DeviceTypeCpu)) return;
                cpu_ms[e] = ms_since(t0);
            }
            if (!gev.empty() && !run_gpu(next_gpu, (uint)gev.size() - 1)) return;
            if (fake) refit_runtime_cost(cpu_ms, gpu_ms);
            break;
        }
        case LaunchModeHetero: {
            // launcher.cpp:918-1111: every GPU event is queued at once behind
            // its wait for the host; the host then runs the CPU events, each
            // after the GPU event it waits for has completed, signalling
            // (event index + 1) when done.
            if (!hooks) {
                std::fprintf(stderr, "Launcher: LaunchModeHetero needs synchronisation hooks\n");
                return;
            }
            const auto& cev = launch_events[0];
            const auto& gev = launch_events[1];
            hooks->cpu_signal(0);
            for (uint e = 0; e < gev.size(); e++) {
                hooks->gpu_begin(e, gev[e].wait);
                bool ok = run_event(1, gev[e], DeviceTypeGpu);
                hooks->gpu_end(e);
                if (!ok) {   // release every queued wait before leaving
                    hooks->cpu_signal(0xFFFFFFFFu);
                    hooks->gpu_finish();
                    return;
                }
            }
            std::vector<float> cpu_ms(cev.size(), 0.f), gpu_ms(gev.size(), 0.f);
            for (uint e = 0; e < cev.size(); e++) {
                if (cev[e].wait != ~0u && !hooks->cpu_wait_gpu(cev[e].wait)) {
                    // the GPU result this event needs never arrived: run no
                    // further host event, release every queued wait, drain
                    hooks->cpu_signal(0xFFFFFFFFu);
                    hooks->gpu_finish();
                    return;
                }
                const auto t0 = clock::now();
                const bool ok = run_event(0, cev[e], DeviceTypeCpu);
                cpu_ms[e] = ms_since(t0);
                if (!ok) {
                    hooks->cpu_signal(0xFFFFFFFFu);
                    hooks->gpu_finish();
                    return;
                }
                if (cev[e].signal != ~0u) hooks->cpu_signal(cev[e].signal + 1);
            }
            hooks->cpu_signal(0xFFFFFFFFu);   // no GPU wait can outlive the launch
            hooks->gpu_finish();
            for (uint e = 0; e < gev.size(); e++) gpu_ms[e] = hooks->gpu_elapsed_ms(e);
            refit_runtime_cost(cpu_ms, gpu_ms);
            break;
        }
        case LaunchModeProgressiveHetero: {
            // launcher.cpp:1112-1268: LaunchModeHetero with the GPU events
            // queued in windows of `progressive_window`; a window is queued
            // when the host first needs one of its events.  (The reference
            // waits for the value `wait` here, without the +1 of its Hetero
            // branch, so a GPU event waiting for CPU event 0 would not wait;
            // the Hetero convention is used.)
            if (!hooks) {
                std::fprintf(stderr, "Launcher: LaunchModeProgressiveHetero needs synchronisation hooks\n");
                return;
            }
            const auto& cev = launch_events[0];
            const auto& gev = launch_events[1];
            const uint window = std::max(1u, progressive_window);
            uint queued = 0;
            auto queue_until = [&](uint last) -> bool {   // queue whole windows up to event `last`
                while (queued < gev.size() && queued <= last) {
                    const uint end = std::min<uint>((uint)gev.size(), queued + window);
                    for (uint e = queued; e < end; e++) {
                        hooks->gpu_begin(e, gev[e].wait);
                        const bool ok = run_event(1, gev[e], DeviceTypeGpu);
                        hooks->gpu_end(e);
                        if (!ok) return false;
                    }
                    queued = end;
                }
                return true;
            };
            hooks->cpu_signal(0);
            bool ok = gev.empty() || queue_until(0);
            for (uint e = 0; ok && e < cev.size(); e++) {
                if (cev[e].wait != ~0u) {
                    ok = queue_until(cev[e].wait);
                    if (!ok) break;
                    ok = hooks->cpu_wait_gpu(cev[e].wait);
                    if (!ok) break;
                }
                ok = run_event(0, cev[e], DeviceTypeCpu);
                if (ok && cev[e].signal != ~0u) hooks->cpu_signal(cev[e].signal + 1);
            }
            if (ok && !gev.empty()) ok = queue_until((uint)gev.size() - 1);
            hooks->cpu_signal(0xFFFFFFFFu);
            hooks->gpu_finish();
            break;
        }
        case LaunchModePartialCPU: {
            // launcher.cpp:1474-1505: the CPU processor's events alone, each
            // timed against its profiled cost (reported beyond 1.3x)
            for (uint e = 0; e < launch_events[0].size(); e++) {
                const LaunchEvent& ev = launch_events[0][e];
                float profiled = 0.f;
                for (uint i = ev.start_idx; i <= ev.end_idx; i++) profiled += computation_matrix[proc_schedules[0][i].task_id][0];
                const auto t0 = clock::now();
                if (!run_event(0, ev, DeviceTypeCpu)) return;
                const float ms = ms_since(t0);
                if (profiled > 0.f && ms / profiled > 1.3f)
                    std::fprintf(stderr, "Launcher: CPU event %u took %.3f ms, profiled %.3f ms\n", e, ms, profiled);
            }
            break;
        }
        case LaunchModePartialGPU:
            // launcher.cpp:1506-1581: the GPU processor's events alone
            for (uint e = 0; e < launch_events[1].size(); e++) {
                if (hooks) hooks->gpu_begin(e, ~0u);
                const bool ok = run_event(1, launch_events[1][e], DeviceTypeGpu);
                if (hooks) hooks->gpu_end(e);
                if (!ok) break;
            }
            if (hooks) hooks->gpu_finish();
            break;
        default:
            std::fprintf(stderr, "Launcher: launch mode %d is not available in this library\n", (int)mode);
    }
}

float Scheduler::get_scheduled_end_time() const {
    float e = 0.f;
    for (const auto& ps : proc_schedules) e = std::max(ps.empty() ? 0.f : ps.back().end, e);
    return e;
}

std::vector<float> Scheduler::get_scheduled_speedups() const {
    const float e = get_scheduled_end_time();
    std::vector<float> s(proc_schedules.size(), 0.f);
    for (size_t p = 0; p < s.size() && p < summary_of_costs_each_device.size(); p++)
        s[p] = (float)((summary_of_costs_each_device[p] - e) / e);
    return s;
}

std::string Scheduler::proc_schedules_string() const {
    std::ostringstream os;
    for (size_t p = 0; p < proc_schedules.size(); p++) {
        os << "[";
        for (size_t i = 0; i < proc_schedules[p].size(); i++) {
            const auto& ev = proc_schedules[p][i];
            os << "(" << ev.task_id << ", " << p << ", " << ev.start << ", " << ev.end << ")";
            if (i + 1 != proc_schedules[p].size()) os << ", ";
        }
        os << "],\n";
    }
    return os.str();
}

std::string Scheduler::schedule_graph_string() const {
    // colour of a task in the reference's figures (launcher.cpp:3951-3978,
    // REMAP_TO_ARTICLE): a per-function base plus the cluster index
    auto colour = [](const Task& t) -> uint {
        switch (t.func_id) {
            case id_xpbd_predict_position: return 1u + t.cluster_idx;
            case id_xpbd_reset_constrains: return 13u + t.cluster_idx;
            case id_vbd_all_in_one: return 14u + t.cluster_idx;
            default: return 0u;
        }
    };
    char buf[160];
    std::ostringstream os;
    for (size_t p = 0; p < proc_schedules.size() && p < 2; p++) {
        os << "    [";
        for (size_t i = 0; i < proc_schedules[p].size(); i++) {
            const ScheduleEvent& ev = proc_schedules[p][i];
            const Task& t = list_task[ev.task_id];
            std::snprintf(buf, sizeof(buf), "(%u, %u, %u, %4.3f, %4.3f)", colour(t), ev.task_id, t.buffer_idx, ev.start,
                          ev.end);
            os << buf << (i + 1 != proc_schedules[p].size() ? ", " : "");
        }
        os << "],\n";
    }
    os << "    [";
    for (const ScheduleEvent& ev : task_schedules)
        for (uint in : list_task[ev.task_id].task_ins) {
            const ScheduleEvent& iv = task_schedules[in];
            std::snprintf(buf, sizeof(buf), "(%u, %u, %u, %u, %4.3f, %4.3f), ", iv.proc, in, ev.proc, ev.task_id, iv.end,
                          ev.start);
            os << buf;
        }
    os << "],\n";
    os << "    [";
    for (const ScheduleEvent& ev : task_schedules) {
        const uint left = list_task[ev.task_id].task_left;
        if (left == ~0u) continue;
        const ScheduleEvent& lv = task_schedules[left];
        std::snprintf(buf, sizeof(buf), "(%u, %u, %u, %u, %4.3f, %4.3f), ", lv.proc, left, ev.proc, ev.task_id, lv.end,
                      ev.start);
        os << buf;
    }
    os << "],\n";
    return os.str();
}

void Scheduler::print_schedule_to_graph_xpbd() const { std::printf("%s", schedule_graph_string().c_str()); }

void Scheduler::print_proc_schedules() const { std::printf("%s", proc_schedules_string().c_str()); }

void Scheduler::print_speedups_to_each_device() const {
    const float e = get_scheduled_end_time();
    const auto s = get_scheduled_speedups();
    for (size_t p = 0; p < s.size(); p++)
        std::printf("Speedup to proc %zu = %4.2f%% (From %4.2f to %4.2f )\n", p, s[p] * 100.f,
                    summary_of_costs_each_device[p], e);
}

// launcher.cpp:4243-4270
void Scheduler::update_costs_from_computation_matrix() {
    const uint np = (uint)proc_schedules.size();
    summary_of_costs_each_device.assign(np, 0.0);
    for (uint t = 0; t < list_task.size(); t++)
        for (uint p = 0; p < np; p++) {
            if (list_task[t].has_implementation(p))
                summary_of_costs_each_device[p] += computation_matrix[t][p] + communication_cost_matrix_uma[p][p];
            else if (p != 0)
                summary_of_costs_each_device[p] +=
                    computation_matrix[t][0] + communication_cost_matrix_uma[0][p] + communication_cost_matrix_uma[p][0];
        }
}

void Scheduler::reset_scheduler_system() {
    list_task.clear();
    list_order.clear();
    sorted_nodes.clear();
    task_schedules.clear();
    ranku.clear();
    proc_schedules.clear();
    computation_matrix.clear();
    constraint_task_orders.clear();
}

}  // namespace Launcher
