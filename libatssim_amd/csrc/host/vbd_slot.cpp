// vbd_slot.cpp -- see vbd_slot.h.
#include "vbd_slot.h"

#include <chrono>
#include <thread>

namespace lats {

namespace {
uint32_t to_c_function(Launcher::FunctionID id, bool& ok) {
    ok = true;
    switch (id) {
        case Launcher::id_additional_root: return LATS_FN_ADDITIONAL_ROOT;
        case Launcher::id_additional_terminal: return LATS_FN_ADDITIONAL_TERMINAL;
        case Launcher::id_xpbd_predict_position: return LATS_FN_PREDICT_POSITION;
        case Launcher::id_vbd_all_in_one: return LATS_FN_VBD_ALL_IN_ONE;
        case Launcher::id_xpbd_constraint_last_node: return LATS_FN_CONSTRAINT_LAST_NODE;
        case Launcher::id_xpbd_update_velocity: return LATS_FN_UPDATE_VELOCITY;
        case Launcher::id_xpbd_reset_constrains: return LATS_FN_RESET_CONSTRAINS;
        case Launcher::id_xpbd_constraint_stretch_mass_spring: return LATS_FN_XPBD_STRETCH;
        case Launcher::id_xpbd_constraint_bending: return LATS_FN_XPBD_BENDING;
        default: ok = false; return 0;
    }
}
}  // namespace

void VbdGpuSlot::fn_dispatch(const Launcher::LaunchParam& p) {
    bool ok = false;
    lats_launch_param c{};
    c.function_id = to_c_function(p.function_id, ok);
    if (!ok) {
        if (status_ == LATS_OK) status_ = LATS_ERR_INVALID_ARGUMENT;
        return;
    }
    c.iter_idx = p.iter_idx;
    c.cluster_idx = p.cluster_idx;
    c.buffer_idx = p.buffer_idx;
    c.left_buffer_idx = p.left_buffer_idx;
    c.right_buffer_idx = p.right_buffer_idx;
    c.num_input_buffers = (uint32_t)p.input_buffer_idxs.size();
    c.input_buffer_idxs = p.input_buffer_idxs.empty() ? nullptr : p.input_buffer_idxs.data();
    c.is_allocated_to_main_device = p.is_allocated_to_main_device ? 1u : 0u;
    const int rc = lats_vbd_dispatch(ctx_, &c);
    if (rc != LATS_OK && status_ == LATS_OK) status_ = rc;
}

Launcher::HeteroHooks VbdGpuSlot::hetero_hooks(double timeout_s) {
    Launcher::HeteroHooks h;
    auto keep = [this](int rc) {
        if (rc != LATS_OK && status_ == LATS_OK) status_ = rc;
    };
    h.gpu_begin = [this, keep](Launcher::uint e, Launcher::uint wait) {
        if (wait != ~0u) keep(lats_vbd_stream_wait_host(ctx_, wait + 1));
        keep(lats_vbd_record_marker(ctx_, 2 * e));
    };
    h.gpu_end = [this, keep](Launcher::uint e) { keep(lats_vbd_record_marker(ctx_, 2 * e + 1)); };
    h.cpu_wait_gpu = [this, keep, timeout_s](Launcher::uint e) -> bool {
        const auto t0 = std::chrono::steady_clock::now();
        int done = 0;
        while (true) {
            const int rc = lats_vbd_query_marker(ctx_, 2 * e + 1, &done);
            if (rc != LATS_OK) { keep(rc); return false; }
            if (done) return true;
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > timeout_s) {
                // the launch stops issuing host events and releases the queue
                keep(LATS_ERR_TIMEOUT);
                return false;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    };
    h.cpu_signal = [this, keep](Launcher::uint v) { keep(lats_vbd_host_signal(ctx_, v)); };
    h.gpu_finish = [this, keep]() { keep(lats_vbd_synchronize(ctx_)); };
    h.gpu_elapsed_ms = [this](Launcher::uint e) {
        float ms = 0.f;
        return lats_vbd_marker_elapsed_ms(ctx_, 2 * e, 2 * e + 1, &ms) == LATS_OK ? ms : 0.f;
    };
    return h;
}

Launcher::LaunchParam vbd_task_to_param(const Launcher::Task& t) {
    Launcher::LaunchParam p;
    p.function_id = t.func_id;
    p.iter_idx = t.iter_idx;
    p.cluster_idx = t.cluster_idx;
    p.is_allocated_to_main_device = t.is_allocated_to_main_device;
    p.buffer_idx = t.buffer_idx;
    p.left_buffer_idx = t.buffer_left;
    p.right_buffer_idx = t.buffer_right;
    p.input_buffer_idxs = t.buffer_ins;
    return p;
}

std::vector<Launcher::uint> register_vbd_dag(Launcher::Scheduler& s, const std::vector<Launcher::Implementation>& impls,
                                             Launcher::uint iters, Launcher::uint nclusters) {
    using namespace Launcher;
    const uint predict = s.add_task(Task(id_xpbd_predict_position, 0, impls));
    std::vector<uint> order;
    std::vector<uint> prev, cur;
    for (uint it = 0; it < iters; it++) {
        cur.clear();
        for (uint c = 0; c < nclusters; c++) {
            const uint t = s.add_task(Task(id_vbd_all_in_one, it, ~0u, c, impls));
            s.set_connect(predict, t);
            cur.push_back(t);
        }
        order.insert(order.end(), cur.begin(), cur.end());
        if (it >= 1)
            for (uint c = 0; c < nclusters; c++) s.set_connect(prev[c], cur[c]);
        prev = cur;
    }
    s.set_constraint_task_orders(order);
    const uint last = s.add_task(Task(id_xpbd_constraint_last_node, 0, impls));
    const uint upd = s.add_task(Task(id_xpbd_update_velocity, 0, impls));
    for (uint t : prev) s.set_connect(t, last);
    s.set_connect(last, upd);
    return order;
}

std::vector<Launcher::uint> register_xpbd_dag(Launcher::Scheduler& s, const std::vector<Launcher::Implementation>& impls,
                                              Launcher::uint iters, Launcher::uint nstretch, Launcher::uint nbending) {
    using namespace Launcher;
    const uint predict = s.add_task(Task(id_xpbd_predict_position, 0, impls));
    const uint reset = s.add_task(Task(id_xpbd_reset_constrains, 0, impls));
    s.set_connect(predict, reset);
    std::vector<uint> order;
    uint prev = reset;
    for (uint it = 0; it < iters; it++) {
        for (uint c = 0; c < nstretch; c++) {
            const uint t = s.add_task(Task(id_xpbd_constraint_stretch_mass_spring, it, ~0u, c, impls));
            s.set_connect(prev, t);
            order.push_back(t);
            prev = t;
        }
        for (uint c = 0; c < nbending; c++) {
            const uint t = s.add_task(Task(id_xpbd_constraint_bending, it, ~0u, c, impls));
            s.set_connect(prev, t);
            order.push_back(t);
            prev = t;
        }
    }
    const uint upd = s.add_task(Task(id_xpbd_update_velocity, 0, impls));
    s.set_connect(prev, upd);
    return order;
}

}  // namespace lats
