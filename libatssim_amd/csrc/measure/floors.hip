// floors.hip -- latency floors of the GPU at the colour-pass shape: what an
// empty kernel, a kernel doing one dependent gather, and a grid-wide barrier
// inside one cooperative kernel cost, what a hand-off between neighbouring
// blocks of one persistent kernel costs, with and without fences; and a
// kernel re-reading one buffer in every launch, for the L2 counters.  Measurement support for the colour
// pass's roofline reading (a pass at 256x256 moves ~3 MB, 0.4 us of HBM time,
// so its floor is set by these latencies, not by bandwidth).  No reference
// counterpart.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/libatssim_hip.h"
#include "../host/errors.h"
#include "floors.h"

namespace {

constexpr uint32_t kB = 256;
constexpr uint32_t kGathers = 4;          // positions gathered per thread
constexpr uint32_t kSpinLimit = 1u << 24; // polls of one barrier wait before giving up

int fail(int code, const std::string& msg) {
    lats::set_error(msg);
    return code;
}

#define FCHK(call)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(LATS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Never true for the zero-filled buffer: keeps the kernel from being empty
// of side effects without touching memory on the timed path.
__global__ __launch_bounds__(kB) void k_floor_empty(float* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && p[0] > 1e30f) p[1] = 0.f;
}

// Each thread: kGathers index loads, a position load through each, one store
// (the shape of a colour pass's ELL -> neighbour chain, without the math).
__global__ __launch_bounds__(kB) void k_floor_gather(const uint32_t* idx, const float* x, float* y, uint32_t n) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= n) return;
    float s = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < kGathers; k++) s += x[idx[(size_t)k * n + t]];
    y[t] = s;
}

// Each thread reads its float4 of a buffer that the same blocks read in
// every launch (the buffer is zero-filled, so the store never happens).
// Under the PMC counters (FETCH_SIZE per dispatch) this shows whether a
// launch finds the lines the previous launch read in L2.
__global__ __launch_bounds__(kB) void k_floor_reread(const float4* x, float* y, uint32_t n) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= n) return;
    const float4 v = x[t];
    if ((v.x + v.y) + (v.z + v.w) > 1e30f) y[0] = 0.f;
}

// `rounds` neighbour hand-offs: block b starts phase q once blocks b - d and
// b + d (mod the grid) have published q - 1, then publishes q (the shape of a
// persistent colour pass that synchronises with its neighbouring tiles
// only).  fenced: an agent-scope acquire fence after the wait and a release
// fence before the publish (the C++ memory model's hand-off; on this GPU
// they write back and invalidate L2); otherwise relaxed device-scope
// accesses only.  Blocks are co-resident (cooperative launch) and every wait
// is bounded: a block that gives up records it and every block then leaves.
__global__ __launch_bounds__(kB) void k_floor_handoff(uint32_t* prog, uint32_t* timed_out, uint32_t rounds,
                                                      uint32_t d, int fenced) {
    __shared__ uint32_t s_stop;
    const uint32_t n = gridDim.x, b = blockIdx.x;
    const uint32_t lo = (b + n - d % n) % n, hi = (b + d) % n;
    for (uint32_t q = 1; q <= rounds; q++) {
        if (threadIdx.x == 0) {
            uint32_t polls = 0;
            while (__hip_atomic_load(prog + lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < q - 1 ||
                   __hip_atomic_load(prog + hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < q - 1) {
                if (++polls > kSpinLimit ||
                    __hip_atomic_load(timed_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                    atomicOr(timed_out, 1u);
                    break;
                }
            }
            if (fenced) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            s_stop = __hip_atomic_load(timed_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (s_stop) return;   // the whole block (thread 0 decided for it)
        if (threadIdx.x == 0) {
            if (fenced) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(prog + b, q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
}

// `rounds` grid barriers over a monotonically increasing arrival counter.
// The blocks are co-resident (cooperative launch); every wait is bounded
// anyway: after kSpinLimit polls the block records the failure and leaves
// the wait, so every wave reaches the end of the kernel.
__global__ __launch_bounds__(kB) void k_floor_barriers(uint32_t* bar, uint32_t* timed_out, uint32_t rounds) {
    uint32_t target = 0;
    for (uint32_t r = 0; r < rounds; r++) {
        target += gridDim.x;
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(bar, 1u);
            uint32_t polls = 0;
            while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++polls > kSpinLimit) {
                    atomicOr(timed_out, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __threadfence();
        }
        __syncthreads();
    }
}

struct Dev {
    void* p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};

struct Ev {
    hipEvent_t e = nullptr;
    ~Ev() {
        if (e) (void)hipEventDestroy(e);
    }
};

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

}  // namespace

extern "C" int lats_time_floor(int kind, uint32_t threads, uint32_t launches, float* us) {
    if (!us) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (kind < LATS_FLOOR_EMPTY || kind > LATS_FLOOR_HANDOFF_XCD)
        return fail(LATS_ERR_INVALID_ARGUMENT, "unknown floor kind");
    if (threads == 0 || threads > (1u << 24) || launches == 0 || launches > 100000)
        return fail(LATS_ERR_INVALID_ARGUMENT, "threads must be in 1..2^24 and launches in 1..100000");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LATS_ERR_NO_DEVICE, "no HIP device");
    const uint32_t blocks = (threads + kB - 1) / kB;
    const uint32_t nx = 3 * threads;
    Stream st;
    FCHK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    Dev p, x, y, idx, bar;
    FCHK(hipMalloc(&p.p, 64));
    FCHK(hipMemset(p.p, 0, 64));
    Ev e0, e1;
    FCHK(hipEventCreate(&e0.e));
    FCHK(hipEventCreate(&e1.e));
    if (kind == LATS_FLOOR_GRID_BARRIER) {
        int dev = 0, per_cu = 0;
        hipDeviceProp_t prop;
        FCHK(hipGetDevice(&dev));
        FCHK(hipGetDeviceProperties(&prop, dev));
        FCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_floor_barriers, kB, 0));
        if ((uint64_t)blocks > (uint64_t)per_cu * prop.multiProcessorCount)
            return fail(LATS_ERR_INVALID_ARGUMENT, "more blocks than can be co-resident");
        FCHK(hipMalloc(&bar.p, 8));
        uint32_t* b = static_cast<uint32_t*>(bar.p);
        uint32_t* flag = b + 1;
        uint32_t rounds = launches;
        void* kargs[] = {&b, &flag, &rounds};
        float best = 0.f;
        for (int rep = 0; rep < 2; rep++) {   // the first launch warms up; the second is timed
            FCHK(hipMemsetAsync(bar.p, 0, 8, st.s));
            FCHK(hipEventRecord(e0.e, st.s));
            FCHK(hipLaunchCooperativeKernel((const void*)k_floor_barriers, dim3(blocks), dim3(kB), kargs, 0, st.s));
            FCHK(hipEventRecord(e1.e, st.s));
            FCHK(hipStreamSynchronize(st.s));
            FCHK(hipEventElapsedTime(&best, e0.e, e1.e));
        }
        uint32_t to = 0;
        FCHK(hipMemcpy(&to, flag, 4, hipMemcpyDeviceToHost));
        if (to) return fail(LATS_ERR_TIMEOUT, "grid barrier wait exceeded its poll limit");
        *us = best * 1e3f / (float)launches;
        return LATS_OK;
    }
    if (kind == LATS_FLOOR_HANDOFF_FENCED || kind == LATS_FLOOR_HANDOFF_RELAXED || kind == LATS_FLOOR_HANDOFF_XCD) {
        int dev = 0, per_cu = 0;
        hipDeviceProp_t prop;
        FCHK(hipGetDevice(&dev));
        FCHK(hipGetDeviceProperties(&prop, dev));
        FCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_floor_handoff, kB, 0));
        if ((uint64_t)blocks > (uint64_t)per_cu * prop.multiProcessorCount)
            return fail(LATS_ERR_INVALID_ARGUMENT, "more blocks than can be co-resident");
        FCHK(hipMalloc(&bar.p, 4 * ((size_t)blocks + 1)));
        uint32_t* prog = static_cast<uint32_t*>(bar.p);
        uint32_t* flag = prog + blocks;
        uint32_t rounds = launches;
        // neighbours one block apart sit on different XCDs (workgroups go to
        // the XCDs round-robin); eight apart, on the same one
        uint32_t d = kind == LATS_FLOOR_HANDOFF_XCD ? 8u : 1u;
        int fenced = kind == LATS_FLOOR_HANDOFF_FENCED;
        void* kargs[] = {&prog, &flag, &rounds, &d, &fenced};
        float best = 0.f;
        for (int rep = 0; rep < 2; rep++) {   // the first launch warms up; the second is timed
            FCHK(hipMemsetAsync(bar.p, 0, 4 * ((size_t)blocks + 1), st.s));
            FCHK(hipEventRecord(e0.e, st.s));
            FCHK(hipLaunchCooperativeKernel((const void*)k_floor_handoff, dim3(blocks), dim3(kB), kargs, 0, st.s));
            FCHK(hipEventRecord(e1.e, st.s));
            FCHK(hipStreamSynchronize(st.s));
            FCHK(hipEventElapsedTime(&best, e0.e, e1.e));
        }
        uint32_t to = 0;
        FCHK(hipMemcpy(&to, flag, 4, hipMemcpyDeviceToHost));
        if (to) return fail(LATS_ERR_TIMEOUT, "hand-off wait exceeded its poll limit");
        *us = best * 1e3f / (float)launches;
        return LATS_OK;
    }
    if (kind == LATS_FLOOR_GATHER) {
        std::vector<uint32_t> h((size_t)kGathers * threads);
        uint64_t s = 0x9E3779B97F4A7C15ull;
        for (auto& v : h) {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            v = (uint32_t)((s >> 33) % nx);
        }
        FCHK(hipMalloc(&idx.p, h.size() * 4));
        FCHK(hipMemcpy(idx.p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        FCHK(hipMalloc(&x.p, (size_t)nx * 4));
        FCHK(hipMemset(x.p, 0, (size_t)nx * 4));
        FCHK(hipMalloc(&y.p, (size_t)threads * 4));
    }
    if (kind == LATS_FLOOR_REREAD) {
        FCHK(hipMalloc(&x.p, (size_t)threads * 16));
        FCHK(hipMemset(x.p, 0, (size_t)threads * 16));
    }
    // one graph of `launches` back-to-back kernels, as a frame graph runs its colour passes
    hipGraph_t graph = nullptr;
    FCHK(hipStreamBeginCapture(st.s, hipStreamCaptureModeThreadLocal));
    for (uint32_t l = 0; l < launches; l++) {
        if (kind == LATS_FLOOR_EMPTY)
            hipLaunchKernelGGL(k_floor_empty, dim3(blocks), dim3(kB), 0, st.s, static_cast<float*>(p.p));
        else if (kind == LATS_FLOOR_REREAD)
            hipLaunchKernelGGL(k_floor_reread, dim3(blocks), dim3(kB), 0, st.s, static_cast<const float4*>(x.p),
                               static_cast<float*>(p.p), threads);
        else
            hipLaunchKernelGGL(k_floor_gather, dim3(blocks), dim3(kB), 0, st.s, static_cast<const uint32_t*>(idx.p),
                               static_cast<const float*>(x.p), static_cast<float*>(y.p), threads);
    }
    FCHK(hipStreamEndCapture(st.s, &graph));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(LATS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    float ms = 0.f;
    for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {   // best of three timed replays after one warm-up
        if ((e = hipGraphLaunch(exec, st.s)) != hipSuccess) break;
        if ((e = hipEventRecord(e0.e, st.s)) != hipSuccess) break;
        if ((e = hipGraphLaunch(exec, st.s)) != hipSuccess) break;
        if ((e = hipEventRecord(e1.e, st.s)) != hipSuccess) break;
        if ((e = hipStreamSynchronize(st.s)) != hipSuccess) break;
        float t = 0.f;
        if ((e = hipEventElapsedTime(&t, e0.e, e1.e)) != hipSuccess) break;
        if (rep == 0 || t < ms) ms = t;
    }
    (void)hipStreamSynchronize(st.s);
    (void)hipGraphExecDestroy(exec);
    if (e != hipSuccess) return fail(LATS_ERR_HIP, std::string("floor graph: ") + hipGetErrorString(e));
    *us = ms * 1e3f / (float)launches;
    return LATS_OK;
}
