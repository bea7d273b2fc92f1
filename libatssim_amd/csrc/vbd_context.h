// vbd_context.h -- the single-cloth device context behind lats_vbd (shared
// by the solver and the learned warm start translation units).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/obstacle.h"
#include "vbd_kernels.h"
#include "vbd_patch.h"

namespace lats {

constexpr int kThreads = 256;

template <typename T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void**)&p, count * sizeof(T));
    }
    hipError_t upload(const std::vector<T>& h) {
        hipError_t e = alloc(h.size());
        if (e != hipSuccess || h.empty()) return e;
        e = hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
        // the copy runs on the null stream, which does not order the
        // library's non-blocking streams: wait until it has landed
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        return e;
    }
};

inline uint32_t nblocks(uint64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

}  // namespace lats

struct lats_vbd {
    uint64_t uid = 0;          // unique per created context (never reused)
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t V = 0, E = 0, B = 0, C = 0;
    lats_scene_params prm{};
    float h = 0.f, h2inv = 0.f, damp = 0.f, h_energy = 0.f;
    std::vector<uint32_t> cstart, ccount, eoff, boff;
    std::vector<lats::ColourInfo> colours;   // host copy of colours_d
    uint32_t regions = 1;
    uint32_t colour_block = 256;   // threads per block of the colour passes (64, 128, 256)
    std::vector<uint32_t> perm, inv;
    int strategy = LATS_STRATEGY_AUTO;
    int energy_bending = 1;

    lats::DBuf<float> x, x0, v, mass;
    lats::DBuf<uint32_t> meta, perm_d, inv_d;
    lats::DBuf<uint2> ell_e, e_slots;
    lats::DBuf<float4> ell_bq, Qd;
    lats::DBuf<lats::PairIds> ell_bi;
    lats::DBuf<uint4> b_slots;
    lats::DBuf<float> rest;
    lats::DBuf<lats::FrameParams> fp;
    lats::DBuf<float> forces;
    lats::DBuf<double> partial;
    lats::DBuf<float> energy;
    lats::DBuf<float> scratch3;   // 3V staging for permutations
    uint32_t energy_rows = 0, energy_cap = 0, energy_nblk = 0;
    bool record_dispatch = false;

    // multi-buffer positions of the async schedule
    lats::DBuf<float> iterbuf[LATS_NUM_ASYNC_BUFFERS], beginbuf[LATS_NUM_ASYNC_BUFFERS];

    // obstacle contact (row a10)
    bool has_obstacle = false;
    float thickness = 0.f, ob_stiffness = 0.f, ob_radius = 0.f;
    lats::ObstacleHash ob_hash;
    lats::DBuf<float> ob_x, ob_n;
    lats::DBuf<int32_t> ob_f;
    lats::DBuf<uint32_t> hb_belongs, hb_start, hb_count, hb_faces;
    lats::DBuf<uint32_t> ct_count, ct_fid;
    lats::DBuf<float4> ct_xs, ct_n;

    // persistent strategy
    lats::DBuf<lats::ColourInfo> colours_d;
    lats::DBuf<uint32_t> barrier;
    uint32_t persist_blocks = 0;      // 0 = derive from the largest colour
    // last autotune: per candidate {strategy, blocks, lanes, regions, best ms}
    std::vector<float> tune_results;
    uint32_t lanes = 1;               // lanes per vertex (1, 4, 8, 16)
    uint32_t persist_max_lanes = 0;
    uint32_t persist_max_blocks = 0;  // co-residency limit (queried on first use)
    bool cooperative_refused = false; // the last cooperative launch was refused by the runtime (nothing enqueued)

    // XPBD solver (built on first use): constraints in cluster order, their
    // multipliers, cluster bounds, 1/m per slot
    int solver = LATS_SOLVER_VBD;
    bool xpbd_built = false;
    std::vector<uint32_t> xs_bound, xb_bound;          // cluster c: [bound[c], bound[c + 1])
    std::vector<uint32_t> xs_csr, xb_csr;              // packed CSR of constraint ids (reference numbering)
    lats::DBuf<uint2> xs_edges;
    lats::DBuf<float> xs_rest, xs_lambda, xb_lambda, x_minv;
    lats::DBuf<uint4> xb_pairs;
    lats::DBuf<float4> xb_Q;

    // flow strategy: tiles, neighbour lists, per-tile phase counters
    uint32_t flow_tiles = 0;          // tiles of the built layout (0 = none)
    uint32_t flow_base = 0;           // phase every tile has completed
    lats::DBuf<uint2> flow_range;
    lats::DBuf<uint32_t> flow_nb_off, flow_nb, flow_progress, flow_error;

    // patch strategy: layout built on first use for (patches, colours per
    // launch); positions alternate between x and xalt from launch to launch
    uint32_t patch_target = 0;        // patches requested (0 = derived from the vertex count)
    uint32_t patch_colours = 0;       // colours per launch requested (0 = every colour)
    uint32_t patch_threads = 1024;    // threads per workgroup (256, 512, 1024)
    uint32_t patch_built_p = 0, patch_built_f = 0;   // built layout (0 = none)
    std::vector<uint32_t> pt_group_lds;   // per launch of a sweep: most LDS bytes of one workgroup
    // per launch of a sweep: first region entry, region stride, first colour
    // (launch g runs colours [pt_group_c0[g], pt_group_c0[g + 1]))
    std::vector<uint32_t> pt_group_roff, pt_group_rstride, pt_group_c0;
    std::vector<uint32_t> pt_group_nsmax;   // per launch of a sweep: most slots with frame state in a region
    lats::PatchStats patch_stats;
    lats::DBuf<lats::PatchHead> pt_heads;
    lats::DBuf<lats::PatchStep> pt_steps;
    lats::DBuf<uint32_t> pt_region;
    // frame state of every region entry, in region order (x_start | mass,
    // v | meta bits): gathered at the start of every run of sweeps
    lats::DBuf<float4> pt_state;
    lats::DBuf<uint2> pt_entries, pt_ell_e;
    lats::DBuf<float4> pt_ell_bq;
    lats::DBuf<lats::PatchPair> pt_ell_bi;
    lats::DBuf<float> xalt;

    // graph per (record) flag; graphs of k colour sweeps (lats_vbd_iterate)
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    std::map<uint32_t, hipGraphExec_t> sweep_graphs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;

    std::map<uint32_t, hipEvent_t> markers;
    uint32_t* host_flag = nullptr;    // pinned, mapped
    uint32_t* host_flag_d = nullptr;  // device alias
    static constexpr uint32_t kHostEnergies = 16;
    float* host_energy = nullptr;     // pinned, mapped, coherent: energies read by the host
    float* host_energy_d = nullptr;   // device alias

    // every captured graph holds kernel choices and buffer pointers of the
    // moment it was captured: drop them whenever one of those changes.  A
    // graph may still be running on the stream (launches return at once), so
    // the stream drains before any executable is destroyed.
    void drop_graphs() {
        bool any = !sweep_graphs.empty();
        for (auto& g : graph) any = any || g;
        if (any && stream) (void)hipStreamSynchronize(stream);
        for (auto& g : graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        for (auto& kv : sweep_graphs) (void)hipGraphExecDestroy(kv.second);
        sweep_graphs.clear();
    }

    ~lats_vbd() {
        drop_graphs();
        for (auto& kv : markers) (void)hipEventDestroy(kv.second);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (host_flag) (void)hipHostFree(host_flag);
        if (host_energy) (void)hipHostFree(host_energy);
        if (stream) (void)hipStreamDestroy(stream);
    }

    lats::VbdArgs args(float* xbuf) const {
        lats::VbdArgs a;
        a.x = xbuf;
        a.x0 = x0.p;
        a.v = v.p;
        a.mass = mass.p;
        a.meta = meta.p;
        a.ell_e = ell_e.p;
        a.ell_bq = ell_bq.p;
        a.ell_bi = ell_bi.p;
        a.fp = fp.p;
        a.h = h;
        a.h2inv = h2inv;
        a.ks = prm.stiffness_stretch_spring;
        a.kb = prm.stiffness_quadratic_bending;
        for (int c = 0; c < 3; c++) a.g[c] = prm.gravity[c];
        a.ccount = has_obstacle ? ct_count.p : nullptr;
        a.ct_xs = ct_xs.p;
        a.ct_n = ct_n.p;
        a.nslots = V;
        a.thickness = thickness;
        return a;
    }
};

namespace lats {
// Energies of n position buffers (slot order) of context c, with the frame
// start, velocities and contacts of its running frame: all n evaluated on
// c's stream after the work already queued there, one host round trip.
// out[k] = energy of xs[k].  Defined in vbd_device.hip.
int energies_of(lats_vbd* c, const float* const* xs, uint32_t n, float* out);
// The same without waiting: the energies land in c->host_energy[0..n) once
// the work queued on stream st (c's stream if null) has completed
// (n <= kHostEnergies).  The energy scratch rows are c's: no other energy
// evaluation of c may run until they have landed.
int energies_enqueue(lats_vbd* c, const float* const* xs, uint32_t n, hipStream_t st = nullptr);
// Patch strategy (vbd_patch.hip): builds the layout for the requested
// patches / colours per launch unless it is built already (host waits; not
// inside a capture); launches per colour sweep; one launch of group g.
int patch_prepare(lats_vbd* c);
uint32_t patch_launches_per_sweep(const lats_vbd* c);
uint32_t patch_lanes(const lats_vbd* c);
uint32_t patch_block(const lats_vbd* c);
void enqueue_patch_launch(lats_vbd* c, uint32_t g, const float* xin, float* xout);
void enqueue_patch_stage(lats_vbd* c);
}  // namespace lats
