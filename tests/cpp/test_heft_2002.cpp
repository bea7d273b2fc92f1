// HEFT known answer: the Topcuoglu et al. 2002 DAG (10 tasks x 3 procs) as
// registered by the reference's examples/example1_heft2002.cpp.  The
// reference's scheduler reports makespan 91 and speedups 39.56 / 42.86 /
// 57.14 % against 127 / 130 / 143 on single processors (README.md:25-31).
#include <cmath>
#include <cstdio>

#include "host/launcher.h"

int main() {
    Launcher::Scheduler scheduler;
    const unsigned num_procs = 3, num_tasks = 10;
    int ran[3] = {0, 0, 0};
    Launcher::Implementation i0(0, [&](const Launcher::LaunchParam&) { ran[0]++; });
    Launcher::Implementation i1(1, [&](const Launcher::LaunchParam&) { ran[1]++; });
    Launcher::Implementation i2(2, [&](const Launcher::LaunchParam&) { ran[2]++; });
    unsigned t[11];
    for (unsigned id = 1; id <= 10; id++)
        t[id] = scheduler.add_task(Launcher::Task(Launcher::id_task_heft_2002, id, {i0, i1, i2}));
    const float w[15][3] = {{1, 2, 18}, {1, 3, 12}, {1, 4, 9},  {1, 5, 11}, {1, 6, 14},
                            {2, 8, 19}, {2, 9, 16}, {3, 7, 23}, {4, 8, 27}, {4, 9, 23},
                            {5, 9, 13}, {6, 8, 15}, {7, 10, 17}, {8, 10, 11}, {9, 10, 13}};
    for (auto& e : w) scheduler.set_connect(t[(int)e[0]], t[(int)e[1]], e[2]);
    scheduler.computation_matrix = {{14, 16, 9}, {13, 19, 18}, {11, 13, 19}, {13, 8, 17}, {12, 13, 10},
                                    {13, 16, 9}, {7, 15, 11},  {5, 11, 14},  {18, 12, 20}, {21, 7, 16}};
    scheduler.communication_speed_matrix = {{0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
    scheduler.communication_cost_matrix_uma = {};
    scheduler.communication_startup = {0, 0, 0};
    scheduler.summary_of_costs_each_device = {0, 0, 0};
    for (unsigned p = 0; p < num_procs; p++)
        for (unsigned k = 0; k < num_tasks; k++)
            scheduler.summary_of_costs_each_device[p] +=
                scheduler.computation_matrix[k][p] + scheduler.fn_get_inner_communication_cost(p);
    if (!scheduler.topological_sort()) { std::printf("FAIL cycle\n"); return 1; }
    scheduler.standardizing_dag();
    scheduler.scheduler_dag();
    scheduler.print_proc_schedules();
    scheduler.print_speedups_to_each_device();
    const float end = scheduler.get_scheduled_end_time();
    const auto sp = scheduler.get_scheduled_speedups();
    std::printf("makespan %g\n", end);
    int bad = 0;
    if (std::fabs(end - 91.f) > 1e-4f) { std::printf("FAIL makespan %g != 91\n", end); bad = 1; }
    const float want[3] = {0.3956f, 0.4286f, 0.5714f};
    for (int p = 0; p < 3; p++)
        if (std::fabs(sp[p] - want[p]) > 5e-4f) { std::printf("FAIL speedup %d: %g\n", p, sp[p]); bad = 1; }
    // three processors: no wait events (the reference builds them for two
    // only), so run every task on its CPU implementation in rank order
    scheduler.launch(Launcher::Scheduler::LaunchModeCpu, [](const Launcher::Task& task) {
        Launcher::LaunchParam p;
        p.function_id = task.func_id;
        p.cluster_idx = task.cluster_idx;
        return p;
    });
    if (ran[0] != 10) { std::printf("FAIL launched %d tasks\n", ran[0]); bad = 1; }
    std::printf(bad ? "FAIL\n" : "PASS\n");
    return bad;
}
