// vbd_slot.h -- the GPU device slot of the Launcher for the VBD cloth path.
//
// `VbdGpuSlot::fn_dispatch` is the counterpart of the reference's
// GpuSolver::fn_dispatch (examples/example3_hybrid_cpu_gpu.cpp:1744-1920):
// it turns a Launcher::LaunchParam into lats_vbd_dispatch on the HIP
// context.  `register_vbd_dag` registers the reference's per-frame VBD DAG
// (example3_hybrid_cpu_gpu.cpp:1921-1988): predict -> (iteration, cluster)
// all-in-one tasks, each cluster chained across iterations -> last node ->
// update velocity.
#pragma once
#include <vector>

#include "../../../include/libatssim_hip.h"
#include "launcher.h"

namespace lats {

class VbdGpuSlot {
public:
    explicit VbdGpuSlot(lats_vbd* ctx) : ctx_(ctx) {}
    // Enqueues the task on the context's stream; the first failure is kept
    // in status() (the reference prints and continues).
    void fn_dispatch(const Launcher::LaunchParam& param);
    Launcher::Implementation implementation() {
        return Launcher::Implementation(Launcher::DeviceTypeGpu,
                                        [this](const Launcher::LaunchParam& p) { fn_dispatch(p); });
    }
    int status() const { return status_; }
    // Synchronisation of LaunchModeHetero on this context (the reference's
    // MTLSharedEvent + command buffers, launcher.cpp:918-1111): queue event e
    // is bracketed by markers 2e and 2e + 1, queue waits are
    // lats_vbd_stream_wait_host, host signals lats_vbd_host_signal.  A host
    // wait gives up after `timeout_s` seconds: the launch then runs no further
    // host event and releases every queued wait (so the stream drains), and
    // status() reports LATS_ERR_TIMEOUT.
    Launcher::HeteroHooks hetero_hooks(double timeout_s = 60.0);

private:
    lats_vbd* ctx_;
    int status_ = LATS_OK;
};

// Maps a scheduled Task to its LaunchParam (example3_hybrid_cpu_gpu.cpp:2396-2409).
Launcher::LaunchParam vbd_task_to_param(const Launcher::Task& task);

// Registers one frame's DAG; returns the task ids of the iteration tasks in
// registration order (also set as the scheduler's constraint task order).
std::vector<Launcher::uint> register_vbd_dag(Launcher::Scheduler& scheduler,
                                             const std::vector<Launcher::Implementation>& impls,
                                             Launcher::uint constraint_iter_count, Launcher::uint num_clusters);

// Registers one XPBD frame (CpuSolver::physics_step_xpbd,
// example2_async_iter.cpp:1017-1081, with solve_constraints_XPBD :1510-1534):
// predict -> reset -> per iteration every spring cluster then every bending
// cluster (a chain: Gauss-Seidel over constraint colours) -> update velocity.
// Returns the constraint tasks in order.
std::vector<Launcher::uint> register_xpbd_dag(Launcher::Scheduler& scheduler,
                                              const std::vector<Launcher::Implementation>& impls,
                                              Launcher::uint iters, Launcher::uint num_stretch_clusters,
                                              Launcher::uint num_bending_clusters);

}  // namespace lats
