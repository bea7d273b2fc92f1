// vbd_batch.hip -- many independent cloths of uneven sizes in one persistent
// launch per frame.
//
// All cloths share one set of HBM arrays (vbd_layout.h).  A block owns a
// whole cloth for a frame: predict, every (iteration, colour) pass with a
// block barrier between colours, update velocity; it then pulls the next
// cloth from a work counter.  Cloths are queued largest first (longest
// processing time first), so the tail of the frame is made of small cloths.
// A cloth small enough keeps its positions in the block's LDS for the whole
// frame, so its passes stream only the static ELL and the own state from
// HBM.  Block size, resident blocks, lanes per vertex and LDS staging are
// tunable (lats_batch_autotune).  The frame (counter reset + batch kernel)
// is captured once in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/errors.h"
#include "vbd_common.h"
#include "vbd_kernels.h"
#include "vbd_layout.h"

using namespace lats;

namespace {

struct ClothInfo {
    uint32_t vbase, nv, cbase, ncol;
    float ks, kb;
    uint32_t pad0, pad1;
};

struct BatchArgs {
    float* x;
    float* x0;
    float* v;
    const float* mass;
    const uint32_t* meta;
    const uint2* ell_e;
    const float4* ell_bq;
    const PairIds* ell_bi;
    const ClothInfo* cloths;
    const uint32_t* order;     // cloth ids, largest first
    const ColourInfo* colours;
    uint32_t* work;            // next queue position
    const FrameParams* fp;
    uint32_t ncloths, iters;
    uint32_t lds_slots;        // largest cloth whose positions are held in LDS (0: none)
    float h, h2inv, damp;
    float g[3];
};

__global__ void k_batch_reset(uint32_t* work) { *work = 0u; }

// One frame of one cloth by the whole block: predict, every (iteration,
// colour) pass with a block barrier between colours, update velocity.
// Mem decides where the cloth's positions live for the frame (PlainMem: in
// HBM; LdsMem: in LDS, written back to HBM by the last step).
template <int G, class Mem>
__device__ __forceinline__ void batch_cloth(const BatchArgs& p, const ClothInfo& ci, const VbdArgs& a, F3 ext,
                                            Mem mem) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t s0 = ci.vbase, s1 = ci.vbase + ci.nv;
    // Core::predict_position
    for (uint32_t s = s0 + tid; s < s1; s += nt) {
        F3 vpred = add3(ld3(p.v, s), scale3(p.h, F3{0.f, -9.8f, 0.f}));
        if ((p.meta[s] >> 16) & 1u) vpred = F3{0.f, 0.f, 0.f};
        mem.stx(p.x, s, add3(ld3(p.x0, s), scale3(p.h, vpred)));
    }
    __syncthreads();
    for (uint32_t it = 0; it < p.iters; it++) {
        for (uint32_t c = 0; c < ci.ncol; c++) {
            const ColourInfo co = p.colours[ci.cbase + c];
            if (G == 1) {
                for (uint32_t i = tid; i < co.count; i += nt)
                    vbd_vertex_ext(a, co.start + i, i, co.count, co.eoff, co.boff, co.ew, co.bw, ext, mem);
            } else {
                // groups of G lanes; a wave leaves the colour together
                for (uint32_t i0 = 0; i0 < co.count; i0 += nt / G) {
                    if (i0 + (tid / 64) * (64 / G) >= co.count) break;
                    const uint32_t i = i0 + tid / G;
                    const bool valid = i < co.count;
                    vbd_vertex_group<G>(a, co.start + (valid ? i : 0u), valid ? i : 0u, co.count, co.eoff, co.boff,
                                        co.ew, co.bw, ext, tid % G, valid, mem);
                }
            }
            __syncthreads();
        }
    }
    // Core::update_velocity
    for (uint32_t s = s0 + tid; s < s1; s += nt) {
        const F3 xk = mem.ldx(p.x, s);
        const F3 dx = sub3(xk, ld3(p.x0, s));
        F3 vel = F3{dx.x / p.h, dx.y / p.h, dx.z / p.h};
        vel = F3{vel.x * p.damp, vel.y * p.damp, vel.z * p.damp};
        st3(p.v, s, vel);
        st3(p.x0, s, xk);
        if (Mem::kStaged) st3(p.x, s, xk);
    }
    __syncthreads();
}

// G lanes per vertex (1: one thread per vertex; 4 / 8: vbd_vertex_group).
// A cloth of at most p.lds_slots vertices keeps its positions in the
// launch's dynamic LDS (12 bytes per slot) for the whole frame; a larger one
// gathers them from HBM.
template <int G>
__global__ __launch_bounds__(1024) void k_vbd_batch(BatchArgs p) {
    extern __shared__ float s_x[];
    __shared__ uint32_t s_item;
    const F3 ext = F3{p.fp->ext[0], p.fp->ext[1], p.fp->ext[2]};
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(p.work, 1u);
        __syncthreads();
        const uint32_t k = s_item;
        __syncthreads();
        if (k >= p.ncloths) break;
        const ClothInfo ci = p.cloths[p.order[k]];
        VbdArgs a;
        a.x = p.x;
        a.x0 = p.x0;
        a.v = p.v;
        a.mass = p.mass;
        a.meta = p.meta;
        a.ell_e = p.ell_e;
        a.ell_bq = p.ell_bq;
        a.ell_bi = p.ell_bi;
        a.fp = p.fp;
        a.h = p.h;
        a.h2inv = p.h2inv;
        a.ks = ci.ks;
        a.kb = ci.kb;
        a.g[0] = p.g[0];
        a.g[1] = p.g[1];
        a.g[2] = p.g[2];
        a.ccount = nullptr;
        a.ct_xs = nullptr;
        a.ct_n = nullptr;
        a.nslots = 0;
        a.thickness = 0.f;
        if (ci.nv != 0u && ci.nv <= p.lds_slots)
            batch_cloth<G>(p, ci, a, ext, LdsMem{s_x, ci.vbase, ci.nv - 1u});
        else
            batch_cloth<G>(p, ci, a, ext, PlainMem{});
    }
}

int fail(int code, const std::string& msg) {
    lats::set_error(msg);
    return code;
}

#define HIPCHK(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(LATS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

template <typename T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DBuf() { if (p) (void)hipFree(p); }
    hipError_t upload(const std::vector<T>& h) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = std::max<size_t>(h.size(), 1);
        hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
        if (e != hipSuccess || h.empty()) return e;
        e = hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
        // the copy runs on the null stream, which does not order the
        // library's non-blocking streams: wait until it has landed
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        return e;
    }
};

}  // namespace

struct lats_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t ncloths = 0;
    uint64_t total_verts = 0;
    lats_scene_params prm{};
    float h = 0.f, h2inv = 0.f, damp = 0.f;
    std::vector<ClothInfo> cloths;
    std::vector<std::vector<uint32_t>> perm;   // per cloth: local slot -> vertex id
    DBuf<float> x, x0, v, mass;
    DBuf<uint32_t> meta, order, work;
    DBuf<uint2> ell_e;
    DBuf<float4> ell_bq;
    DBuf<PairIds> ell_bi;
    DBuf<ClothInfo> cloths_d;
    DBuf<ColourInfo> colours_d;
    DBuf<FrameParams> fp;
    uint32_t block_threads = 256, blocks = 0, max_blocks_cache = 0, max_blocks_threads = 0, max_blocks_lanes = 0;
    uint32_t lanes = 1;   // lanes per vertex (1, 4, 8)
    // positions of cloths with at most lds_slots vertices stay in LDS for the
    // frame; lds_max = the most slots one block's LDS holds on this device
    uint32_t lds_slots = 0, lds_max = 0, max_nv = 0, max_blocks_lds = 0;
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    // a launched frame graph may still be running: drain the stream first
    void drop_graph() {
        if (!graph) return;
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    ~lats_batch() {
        drop_graph();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

const void* batch_kernel(uint32_t lanes) {
    switch (lanes) {
        case 4: return (const void*)k_vbd_batch<4>;
        case 8: return (const void*)k_vbd_batch<8>;
        default: return (const void*)k_vbd_batch<1>;
    }
}

size_t lds_bytes(uint32_t slots) { return 12 * (size_t)slots; }

int max_resident_blocks(lats_batch* b, uint32_t threads, uint32_t* out) {
    if (b->max_blocks_threads != threads || b->max_blocks_lanes != b->lanes || b->max_blocks_lds != b->lds_slots) {
        int per_cu = 0, ncu = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel(b->lanes), (int)threads,
                                                            lds_bytes(b->lds_slots)));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
        b->max_blocks_cache = (uint32_t)std::max(1, per_cu * ncu);
        b->max_blocks_threads = threads;
        b->max_blocks_lanes = b->lanes;
        b->max_blocks_lds = b->lds_slots;
    }
    *out = b->max_blocks_cache;
    return LATS_OK;
}

int batch_graph(lats_batch* b, hipGraphExec_t* out) {
    if (!b->graph) {
        uint32_t mx = 0;
        int rc = max_resident_blocks(b, b->block_threads, &mx);
        if (rc) return rc;
        const uint32_t g = std::min(b->blocks ? b->blocks : mx, std::max(1u, b->ncloths));
        BatchArgs p;
        p.x = b->x.p;
        p.x0 = b->x0.p;
        p.v = b->v.p;
        p.mass = b->mass.p;
        p.meta = b->meta.p;
        p.ell_e = b->ell_e.p;
        p.ell_bq = b->ell_bq.p;
        p.ell_bi = b->ell_bi.p;
        p.cloths = b->cloths_d.p;
        p.order = b->order.p;
        p.colours = b->colours_d.p;
        p.work = b->work.p;
        p.fp = b->fp.p;
        p.ncloths = b->ncloths;
        p.iters = b->prm.constraint_iter_count;
        p.lds_slots = b->lds_slots;
        p.h = b->h;
        p.h2inv = b->h2inv;
        p.damp = b->damp;
        for (int k = 0; k < 3; k++) p.g[k] = b->prm.gravity[k];
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const size_t shm = lds_bytes(b->lds_slots);
        for (uint32_t sub = 0; sub < b->prm.num_substep; sub++) {
            hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, b->stream, b->work.p);
            switch (b->lanes) {
                case 4: hipLaunchKernelGGL(k_vbd_batch<4>, dim3(g), dim3(b->block_threads), shm, b->stream, p); break;
                case 8: hipLaunchKernelGGL(k_vbd_batch<8>, dim3(g), dim3(b->block_threads), shm, b->stream, p); break;
                default: hipLaunchKernelGGL(k_vbd_batch<1>, dim3(g), dim3(b->block_threads), shm, b->stream, p); break;
            }
        }
        HIPCHK(hipStreamEndCapture(b->stream, &graph));
        hipError_t e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            b->graph = nullptr;
            return fail(LATS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    *out = b->graph;
    return LATS_OK;
}

}  // namespace

extern "C" {

int lats_batch_create(uint32_t num_cloths, const lats_mesh_desc* meshes, const float* stretch, const float* bending,
                      const lats_scene_params* prm, lats_batch** out) {
    if (!out || !meshes || !prm || !num_cloths) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (prm->num_substep == 0 || !(prm->implicit_dt > 0.f))
        return fail(LATS_ERR_INVALID_ARGUMENT, "implicit_dt must be > 0 and num_substep >= 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LATS_ERR_NO_DEVICE, "no HIP device");
    auto b = new lats_batch();
    (void)hipGetDevice(&b->device);
    b->ncloths = num_cloths;
    b->prm = *prm;
    b->h = prm->implicit_dt / (float)prm->num_substep;
    b->h2inv = 1.f / (b->h * b->h);
    b->damp = std::exp(-prm->damping_cloth * b->h);
    ClothLayout L;
    std::vector<uint32_t> inv;
    std::string err;
    for (uint32_t k = 0; k < num_cloths; k++) {
        const uint32_t vbase = (uint32_t)L.num_slots(), cbase = (uint32_t)L.colours.size();
        if (!append_cloth(meshes[k], L, inv, err)) {
            delete b;
            return fail(LATS_ERR_INVALID_ARGUMENT, "cloth " + std::to_string(k) + ": " + err);
        }
        ClothInfo ci;
        ci.vbase = vbase;
        ci.nv = meshes[k].num_verts;
        ci.cbase = cbase;
        ci.ncol = meshes[k].num_clusters;
        ci.ks = stretch ? stretch[k] : prm->stiffness_stretch_spring;
        ci.kb = bending ? bending[k] : prm->stiffness_quadratic_bending;
        ci.pad0 = ci.pad1 = 0;
        b->cloths.push_back(ci);
        b->perm.emplace_back(L.perm.begin() + vbase, L.perm.end());
        b->max_nv = std::max(b->max_nv, ci.nv);
    }
    b->total_verts = L.num_slots();
    std::vector<uint32_t> order(num_cloths);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t c) { return b->cloths[a].nv > b->cloths[c].nv; });
    FrameParams fp0;
    std::memset(&fp0, 0, sizeof(fp0));
    auto bad = [&](hipError_t e, const char* what) {
        delete b;
        return fail(e == hipErrorOutOfMemory ? LATS_ERR_OUT_OF_MEMORY : LATS_ERR_HIP,
                    std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e, "stream");
    if ((e = b->x.upload(L.x)) != hipSuccess) return bad(e, "x");
    if ((e = b->x0.upload(L.x)) != hipSuccess) return bad(e, "x0");
    if ((e = b->v.upload(std::vector<float>(L.x.size(), 0.f))) != hipSuccess) return bad(e, "v");
    if ((e = b->mass.upload(L.mass)) != hipSuccess) return bad(e, "mass");
    if ((e = b->meta.upload(L.meta)) != hipSuccess) return bad(e, "meta");
    if ((e = b->order.upload(order)) != hipSuccess) return bad(e, "order");
    if ((e = b->work.upload(std::vector<uint32_t>{0u})) != hipSuccess) return bad(e, "work");
    if ((e = b->ell_e.upload(L.ell_e)) != hipSuccess) return bad(e, "ell_e");
    if ((e = b->ell_bq.upload(L.ell_bq)) != hipSuccess) return bad(e, "ell_bq");
    if ((e = b->ell_bi.upload(L.ell_bi)) != hipSuccess) return bad(e, "ell_bi");
    if ((e = b->cloths_d.upload(b->cloths)) != hipSuccess) return bad(e, "cloths");
    if ((e = b->colours_d.upload(L.colours)) != hipSuccess) return bad(e, "colours");
    if ((e = b->fp.upload(std::vector<FrameParams>{fp0})) != hipSuccess) return bad(e, "frame params");
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess) return bad(e, "event");
    if ((e = hipEventCreate(&b->ev1)) != hipSuccess) return bad(e, "event");
    // LDS left for positions: the device's per-block limit less the largest
    // static LDS of the kernel's variants; every variant may then use it all
    int lds_limit = 0;
    if ((e = hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device)) != hipSuccess)
        return bad(e, "LDS limit");
    size_t lds_static = 0;
    for (uint32_t lanes : {1u, 4u, 8u}) {
        hipFuncAttributes fa;
        if ((e = hipFuncGetAttributes(&fa, batch_kernel(lanes))) != hipSuccess) return bad(e, "kernel attributes");
        lds_static = std::max(lds_static, fa.sharedSizeBytes);
    }
    b->lds_max = lds_limit > (int)lds_static ? (uint32_t)((lds_limit - lds_static) / 12) : 0u;
    for (uint32_t lanes : {1u, 4u, 8u})
        if ((e = hipFuncSetAttribute(batch_kernel(lanes), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds_bytes(b->lds_max))) != hipSuccess)
            return bad(e, "LDS size");
    b->lds_slots = std::min(b->max_nv, b->lds_max);
    *out = b;
    return LATS_OK;
}

int lats_batch_destroy(lats_batch* b) {
    if (!b) return LATS_OK;
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return LATS_OK;
}

int lats_batch_counts(lats_batch* b, uint32_t* num_cloths, uint64_t* total_verts) {
    if (!b) return fail(LATS_ERR_INVALID_ARGUMENT, "null batch");
    if (num_cloths) *num_cloths = b->ncloths;
    if (total_verts) *total_verts = b->total_verts;
    return LATS_OK;
}

int lats_batch_set_state(lats_batch* b, uint32_t k, const float* x, const float* v) {
    if (!b || !x || !v || k >= b->ncloths) return fail(LATS_ERR_INVALID_ARGUMENT, "bad argument");
    const ClothInfo& ci = b->cloths[k];
    std::vector<float> hx(3 * (size_t)ci.nv), hv(3 * (size_t)ci.nv);
    for (uint32_t s = 0; s < ci.nv; s++)
        for (int c = 0; c < 3; c++) {
            hx[3 * s + c] = x[3 * (size_t)b->perm[k][s] + c];
            hv[3 * s + c] = v[3 * (size_t)b->perm[k][s] + c];
        }
    const size_t off = 3 * (size_t)ci.vbase;
    HIPCHK(hipMemcpyAsync(b->x.p + off, hx.data(), hx.size() * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->x0.p + off, hx.data(), hx.size() * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->v.p + off, hv.data(), hv.size() * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return LATS_OK;
}

int lats_batch_get_state(lats_batch* b, uint32_t k, float* x, float* v) {
    if (!b || k >= b->ncloths) return fail(LATS_ERR_INVALID_ARGUMENT, "bad argument");
    const ClothInfo& ci = b->cloths[k];
    std::vector<float> hx(3 * (size_t)ci.nv), hv(3 * (size_t)ci.nv);
    const size_t off = 3 * (size_t)ci.vbase;
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(hx.data(), b->x.p + off, hx.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hv.data(), b->v.p + off, hv.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < ci.nv; s++)
        for (int c = 0; c < 3; c++) {
            if (x) x[3 * (size_t)b->perm[k][s] + c] = hx[3 * s + c];
            if (v) v[3 * (size_t)b->perm[k][s] + c] = hv[3 * s + c];
        }
    return LATS_OK;
}

int lats_batch_set_launch(lats_batch* b, uint32_t block_threads, uint32_t blocks) {
    if (!b) return fail(LATS_ERR_INVALID_ARGUMENT, "null batch");
    if (block_threads == 0 || block_threads > 1024 || block_threads % 64)
        return fail(LATS_ERR_INVALID_ARGUMENT, "block_threads must be a multiple of 64 in [64, 1024]");
    b->block_threads = block_threads;
    b->blocks = blocks;
    b->drop_graph();
    return LATS_OK;
}

int lats_batch_set_lanes_per_vertex(lats_batch* b, uint32_t lanes) {
    if (!b) return fail(LATS_ERR_INVALID_ARGUMENT, "null batch");
    if (lanes != 1 && lanes != 4 && lanes != 8) return fail(LATS_ERR_INVALID_ARGUMENT, "lanes per vertex must be 1, 4 or 8");
    if (lanes != b->lanes) {
        b->lanes = lanes;
        b->drop_graph();
    }
    return LATS_OK;
}

int lats_batch_set_lds_slots(lats_batch* b, uint32_t slots) {
    if (!b) return fail(LATS_ERR_INVALID_ARGUMENT, "null batch");
    if (slots > b->lds_max)
        return fail(LATS_ERR_INVALID_ARGUMENT, "at most " + std::to_string(b->lds_max) +
                                                   " cloth vertices fit in one block's LDS on this device");
    if (slots != b->lds_slots) {
        b->lds_slots = slots;
        b->drop_graph();
    }
    return LATS_OK;
}

int lats_batch_get_lds_slots(lats_batch* b, uint32_t* slots, uint32_t* max_slots) {
    if (!b) return fail(LATS_ERR_INVALID_ARGUMENT, "null batch");
    if (slots) *slots = b->lds_slots;
    if (max_slots) *max_slots = b->lds_max;
    return LATS_OK;
}

int lats_batch_get_lanes_per_vertex(lats_batch* b, uint32_t* lanes) {
    if (!b || !lanes) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *lanes = b->lanes;
    return LATS_OK;
}

int lats_batch_get_launch(lats_batch* b, uint32_t* block_threads, uint32_t* blocks) {
    if (!b) return fail(LATS_ERR_INVALID_ARGUMENT, "null batch");
    uint32_t mx = 0;
    int rc = max_resident_blocks(b, b->block_threads, &mx);
    if (rc) return rc;
    if (block_threads) *block_threads = b->block_threads;
    if (blocks) *blocks = std::min(b->blocks ? b->blocks : mx, std::max(1u, b->ncloths));
    return LATS_OK;
}

int lats_batch_run_frames(lats_batch* b, uint32_t nframes) {
    if (!b) return fail(LATS_ERR_INVALID_ARGUMENT, "null batch");
    hipGraphExec_t g;
    int rc = batch_graph(b, &g);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    for (uint32_t f = 0; f < nframes; f++) HIPCHK(hipGraphLaunch(g, b->stream));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    HIPCHK(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
    return LATS_OK;
}

int lats_batch_last_elapsed_ms(lats_batch* b, float* ms) {
    if (!b || !ms) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *ms = b->last_ms;
    return LATS_OK;
}

// Times `frames` frames for every (LDS staging off / on, lanes per vertex,
// block size, resident blocks) choice of a small grid and keeps the fastest;
// restores the state afterwards.
int lats_batch_autotune(lats_batch* b, uint32_t frames, uint32_t* best_threads, uint32_t* best_blocks) {
    if (!b) return fail(LATS_ERR_INVALID_ARGUMENT, "null batch");
    if (!frames) frames = 1;
    const size_t nf = 3 * (size_t)b->total_verts;
    float *sx = nullptr, *sx0 = nullptr, *sv = nullptr;
    HIPCHK(hipMalloc((void**)&sx, nf * 4));
    HIPCHK(hipMalloc((void**)&sx0, nf * 4));
    HIPCHK(hipMalloc((void**)&sv, nf * 4));
    // device-to-device copies on the batch's stream (a synchronous hipMemcpy
    // between device buffers is not ordered with a non-blocking stream)
    HIPCHK(hipMemcpyAsync(sx, b->x.p, nf * 4, hipMemcpyDeviceToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(sx0, b->x0.p, nf * 4, hipMemcpyDeviceToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(sv, b->v.p, nf * 4, hipMemcpyDeviceToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    float best = 3.4e38f;
    uint32_t bt = b->block_threads, bb = b->blocks, bl = b->lanes, bs = b->lds_slots;
    int rc = LATS_OK;
    // LDS staging off, and on for every cloth that fits
    const uint32_t staged = std::min(b->max_nv, b->lds_max);
    std::vector<uint32_t> lds_options{0u};
    if (staged) lds_options.push_back(staged);
    for (uint32_t slots : lds_options) {
        rc = lats_batch_set_lds_slots(b, slots);
        if (rc) break;
        for (uint32_t lanes : {1u, 4u, 8u}) {
            rc = lats_batch_set_lanes_per_vertex(b, lanes);
            if (rc) break;
            for (uint32_t threads : {128u, 256u, 512u, 1024u}) {
                uint32_t mx = 0;
                rc = max_resident_blocks(b, threads, &mx);
                if (rc) break;
                for (uint32_t blocks : {mx / 4, mx / 2, mx}) {
                    if (blocks == 0) continue;
                    rc = lats_batch_set_launch(b, threads, blocks);
                    if (!rc) rc = lats_batch_run_frames(b, 1);
                    if (!rc) rc = lats_batch_run_frames(b, frames);
                    if (rc) break;
                    if (b->last_ms < best) {
                        best = b->last_ms;
                        bt = threads;
                        bb = blocks;
                        bl = lanes;
                        bs = slots;
                    }
                }
                if (rc) break;
            }
            if (rc) break;
        }
        if (rc) break;
    }
    if (!rc) rc = lats_batch_set_lanes_per_vertex(b, bl);
    if (!rc) rc = lats_batch_set_lds_slots(b, bs);
    hipError_t e1 = hipMemcpyAsync(b->x.p, sx, nf * 4, hipMemcpyDeviceToDevice, b->stream);
    hipError_t e2 = hipMemcpyAsync(b->x0.p, sx0, nf * 4, hipMemcpyDeviceToDevice, b->stream);
    hipError_t e3 = hipMemcpyAsync(b->v.p, sv, nf * 4, hipMemcpyDeviceToDevice, b->stream);
    const hipError_t e4 = hipStreamSynchronize(b->stream);
    if (e1 == hipSuccess) e1 = e4;
    (void)hipFree(sx);
    (void)hipFree(sx0);
    (void)hipFree(sv);
    if (rc) return rc;
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(LATS_ERR_HIP, "state restore failed");
    rc = lats_batch_set_launch(b, bt, bb);
    if (rc) return rc;
    if (best_threads) *best_threads = bt;
    if (best_blocks) *best_blocks = bb;
    return LATS_OK;
}

}  // extern "C"
