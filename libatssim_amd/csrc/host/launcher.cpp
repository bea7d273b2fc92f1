// launcher.cpp -- see launcher.h.
#include "launcher.h"

#include <algorithm>
#include <cstdio>
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
}

void Scheduler::launch(LaunchMode mode, const std::function<LaunchParam(const Task&)>& task_to_param) {
    const std::vector<uint>& order = !sorted_nodes.empty() ? sorted_nodes : list_order;
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
            // every scheduled task in start-time order (ties: processor id)
            std::vector<ScheduleEvent> all;
            for (const auto& ps : proc_schedules) all.insert(all.end(), ps.begin(), ps.end());
            std::stable_sort(all.begin(), all.end(), [](const ScheduleEvent& a, const ScheduleEvent& b) {
                return a.start != b.start ? a.start < b.start : a.proc < b.proc;
            });
            for (const auto& ev : all) {
                const Task& task = list_task[ev.task_id];
                const DeviceType dev = mode == LaunchModeSequeceHetero ? DeviceTypeCpu : ev.proc;
                const Implementation* imp = task.find_implementation(dev);
                if (!imp) { std::fprintf(stderr, "Launcher: task %u lacks device %u\n", ev.task_id, dev); return; }
                imp->launch_task(task_to_param(task));
            }
            break;
        }
        default:
            std::fprintf(stderr, "Launcher: launch mode %d is not available in this build\n", (int)mode);
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
