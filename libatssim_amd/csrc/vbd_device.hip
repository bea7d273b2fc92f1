// vbd_device.hip -- MI355X (gfx950) VBD cloth kernels and the single-cloth
// context behind the C ABI in include/libatssim_hip.h.
//
// Layout in HBM (all FP32 unless noted; "slot" = position in colour order,
// slots [cstart[c], cstart[c]+n[c]) hold colour c's vertices in ascending
// vertex id, i.e. the reference's packed cluster CSR flattened):
//   x, x_start, v      float[3V] AoS per slot (12-byte stride)
//   mass               float[V]
//   meta               uint32[V]: edge degree | bending degree << 8 | pinned << 16
//   ell_e              uint2[sum_c De_c * n_c]: (other slot, rest length bits),
//                      colour-major ELL, entry (c, j, i) at eoff[c] + j*n_c + i
//   ell_bq             float4[...]: row o of the pair's Q (Q[o][0..3])
//   ell_bi             uint4[...]: slots of (e0, e1, o0, o1); o in bits 30-31 of .x
// so every per-vertex load of one colour pass is one coalesced 64-lane access.
//
// Bit-exact parity with the reference's FP32 arithmetic requires this file
// to be compiled with -ffp-contract=off and IEEE div/sqrt (hipcc default).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/errors.h"
#include "host/obstacle.h"
#include "host/topology.h"
#include "vbd_common.h"
#include "vbd_kernels.h"
#include "vbd_layout.h"
#include "vbd_context.h"
#include "xpbd.h"

using namespace lats;

// ===========================================================================
// Device code
// ===========================================================================
namespace {

constexpr int kBlock = 256;

// One colour pass, G lanes per vertex (G = 1: one thread per vertex).
// The colour's index range is cut into gridDim.x bands of `band` consecutive
// indices (a colour's slots are in ascending vertex id, so a band is a band
// of the cloth) and block (r, j) works on band r.  Workgroups are handed to
// the XCDs round-robin by linear id r + j * gridDim.x, so with 8 bands each
// XCD updates the same band of the cloth in every colour pass and the
// neighbours it gathers are mostly lines its own L2 already holds; one band
// is the plain linear mapping.  B threads per block (64, 128 or 256): smaller
// blocks spread a small colour pass's waves more evenly over the CUs
// (results are identical).
// The leading scalar arguments are all the first loads (the ELL prefetch)
// need: they arrive in SGPRs with the wave (kernel argument preloading), so
// those loads issue before any read of the argument block returns.
template <int G, int B>
__global__ __launch_bounds__(B) void k_vbd_colour(const uint2* ell_e, const float4* ell_bq, const PairIds* ell_bi,
                                                  uint32_t start, uint32_t count, uint32_t eoff, uint32_t boff,
                                                  uint32_t ew, uint32_t bw, uint32_t band, VbdArgs a) {
    a.ell_e = ell_e;
    a.ell_bq = ell_bq;
    a.ell_bi = ell_bi;
    const uint32_t lo = blockIdx.x * band, j = blockIdx.y;
    const uint32_t hi = min(lo + band, count);
    const F3 ext = F3{a.fp->ext_cur[0], a.fp->ext_cur[1], a.fp->ext_cur[2]};
    if (G == 1) {
        const uint32_t i = lo + j * B + threadIdx.x;
        if (i >= hi) return;
        vbd_vertex_ext(a, start + i, i, count, eoff, boff, ew, bw, ext);
    } else {
        const uint32_t first = lo + j * (B / G) + (threadIdx.x / 64) * (64 / G);
        if (first >= hi) return;                  // the whole wave is past the band
        const uint32_t i = lo + j * (B / G) + threadIdx.x / G, lane = threadIdx.x % G;
        const bool valid = i < hi;
        if constexpr (G == 8) {
            // (8 lanes: the sums formed one value per lane through LDS)
            __shared__ float4 scr[B / 8 * kGroupStagedWords / 4];
            vbd_vertex_group_staged8(a, start + (valid ? i : lo), valid ? i : lo, count, eoff, boff, ew, bw, ext,
                                     lane, valid, reinterpret_cast<float*>(scr) + threadIdx.x / 8 * kGroupStagedWords);
        } else {
            vbd_vertex_group<G>(a, start + (valid ? i : lo), valid ? i : lo, count, eoff, boff, ew, bw, ext, lane,
                                valid);
        }
    }
}

// --- obstacle contact detection -------------------------------------------
struct DetectArgs {
    const float* x;
    uint32_t nv;
    const float* ob_x;
    const int32_t* ob_f;
    const float* ob_n;
    const uint32_t* belongs;
    const uint32_t* start;
    const uint32_t* count;
    const uint32_t* faces;
    float ox, oy, oz, inv_cell;
    int dimx, dimy, dimz;
    uint32_t table_size;
    float radius, stiffness;
    uint32_t* ccount;
    float4* ct_xs;
    float4* ct_n;
    uint32_t* ct_fid;
};

__device__ __forceinline__ uint32_t hash_coord_d(int x, int y, int z, uint32_t T) {
    const int32_t hx = (int32_t)((uint32_t)x * 92837111u);
    const int32_t hy = (int32_t)((uint32_t)y * 689287499u);
    const int32_t hz = (int32_t)((uint32_t)z * 283923481u);
    long long v = (long long)hx ^ (long long)hy ^ (long long)hz;
    if (v < 0) v = -v;
    return (uint32_t)(v % (long long)T);
}

// The hash cell of position p: on success [*f0, *f1) is its face range.
__device__ __forceinline__ bool detect_cell(const DetectArgs& d, F3 p, uint32_t* f0, uint32_t* f1) {
    const int ix = (int)floorf((p.x - d.ox) * d.inv_cell);
    const int iy = (int)floorf((p.y - d.oy) * d.inv_cell);
    const int iz = (int)floorf((p.z - d.oz) * d.inv_cell);
    if (!(ix >= 0 && iy >= 0 && iz >= 0 && ix < d.dimx && iy < d.dimy && iz < d.dimz)) return false;
    const uint32_t key = ((uint32_t)ix << 20) | ((uint32_t)iy << 10) | (uint32_t)iz;
    uint32_t h = hash_coord_d(ix, iy, iz, d.table_size);
    for (int j = 0; j < 32; j++) {
        const uint32_t b = d.belongs[h];
        if (b == key) {
            *f0 = d.start[h];
            *f1 = *f0 + d.count[h];
            return true;
        }
        if (b == 0xffffffffu) return false;
        h = (h + 1) % d.table_size;
    }
    return false;
}

// The kMaxContacts nearest candidates by (distance, face id): kd / kf / kb
// hold n of them in that order.
struct NearList {
    uint32_t n;
    float kd[kMaxContacts];
    uint32_t kf[kMaxContacts];
    float kb[kMaxContacts][3];
};

__device__ __forceinline__ bool near_before(float da, uint32_t fa, float db, uint32_t fb) {
    return da < db || (da == db && fa < fb);
}

// Inserts (ed, ef, eb) into L if it is among the nearest; returns false if
// it is not.  Static indices only, so the list stays in registers: slot q
// takes the new entry, the entry of slot q - 1, or keeps its own, decided
// from the old contents.
__device__ __forceinline__ bool insert_near(NearList& L, float ed, uint32_t ef, float eb0, float eb1, float eb2) {
    bool goes_before[kMaxContacts];
#pragma unroll
    for (int q = 0; q < kMaxContacts; q++)
        goes_before[q] = (uint32_t)q >= L.n || near_before(ed, ef, L.kd[q], L.kf[q]);
    if (!goes_before[kMaxContacts - 1]) return false;
#pragma unroll
    for (int q = kMaxContacts - 1; q >= 0; q--) {
        if (!goes_before[q]) continue;
        const bool shift = q > 0 && goes_before[q - 1];
        L.kd[q] = shift ? L.kd[q - 1] : ed;
        L.kf[q] = shift ? L.kf[q - 1] : ef;
        L.kb[q][0] = shift ? L.kb[q - 1][0] : eb0;
        L.kb[q][1] = shift ? L.kb[q - 1][1] : eb1;
        L.kb[q][2] = shift ? L.kb[q - 1][2] : eb2;
    }
    if (L.n < (uint32_t)kMaxContacts) L.n++;
    return true;
}

// Face fid against p: inserted if it passes the narrow phase within the
// radius and is among the nearest.
__device__ __forceinline__ void consider_face(const DetectArgs& d, F3 p, uint32_t fid, NearList& L) {
    const F3 A = ld3(d.ob_x, (uint32_t)d.ob_f[3 * fid]);
    const F3 B = ld3(d.ob_x, (uint32_t)d.ob_f[3 * fid + 1]);
    const F3 C = ld3(d.ob_x, (uint32_t)d.ob_f[3 * fid + 2]);
    float dist, bary[3];
    if (!template_vf(p, A, B, C, d.radius, dist, bary)) return;
    insert_near(L, dist, fid, bary[0], bary[1], bary[2]);
}

// Writes slot s's contacts in ascending face id.
__device__ __forceinline__ void order_pair(NearList& L, int a, int b) {
    if (L.kf[a] <= L.kf[b]) return;
    const float td = L.kd[a]; L.kd[a] = L.kd[b]; L.kd[b] = td;
    const uint32_t tf = L.kf[a]; L.kf[a] = L.kf[b]; L.kf[b] = tf;
#pragma unroll
    for (int c = 0; c < 3; c++) { const float tb = L.kb[a][c]; L.kb[a][c] = L.kb[b][c]; L.kb[b][c] = tb; }
}

__device__ __forceinline__ void write_contacts(const DetectArgs& d, uint32_t s, NearList& L) {
    static_assert(kMaxContacts == 4, "the sorting network below orders four entries");
    // unused entries sort last; a 4-input sorting network on the face id
#pragma unroll
    for (int q = 0; q < kMaxContacts; q++)
        if ((uint32_t)q >= L.n) L.kf[q] = 0xffffffffu;
    order_pair(L, 0, 1);
    order_pair(L, 2, 3);
    order_pair(L, 0, 2);
    order_pair(L, 1, 3);
    order_pair(L, 1, 2);
#pragma unroll
    for (int k = 0; k < kMaxContacts; k++) {
        if ((uint32_t)k >= L.n) break;
        const uint32_t fid = L.kf[k];
        const F3 A = ld3(d.ob_x, (uint32_t)d.ob_f[3 * fid]);
        const F3 B = ld3(d.ob_x, (uint32_t)d.ob_f[3 * fid + 1]);
        const F3 C = ld3(d.ob_x, (uint32_t)d.ob_f[3 * fid + 2]);
        const F3 xs = add3(add3(scale3(L.kb[k][0], A), scale3(L.kb[k][1], B)), scale3(L.kb[k][2], C));
        d.ct_xs[k * d.nv + s] = float4{xs.x, xs.y, xs.z, d.stiffness * 1.0f};
        d.ct_n[k * d.nv + s] = float4{d.ob_n[3 * fid], d.ob_n[3 * fid + 1], d.ob_n[3 * fid + 2], 0.f};
        d.ct_fid[k * d.nv + s] = fid;
    }
    d.ccount[s] = L.n;
}

// Contacts of slot s at its current (predicted) position: faces of the hash
// cell containing it that pass the narrow phase within `radius`; the
// kMaxContacts nearest by (distance, face id), stored in ascending face id.
__device__ __forceinline__ void detect_vertex(const DetectArgs& d, uint32_t s) {
    const F3 p = ld3(d.x, s);
    NearList L;
    L.n = 0;
    uint32_t f0 = 0, f1 = 0;
    if (detect_cell(d, p, &f0, &f1))
        for (uint32_t j = f0; j < f1; j++) consider_face(d, p, d.faces[j], L);
    write_contacts(d, s, L);
}

// The same with kDetectLanes lanes per vertex: lane j tests faces j, j+G, ...
// of the cell and keeps its own nearest list; the G lists are merged by
// pairwise exchanges (the nearest-k set under the total order (distance,
// face id) does not depend on how the faces were split), then lane 0
// writes.  Identical contacts to detect_vertex (used by the persistent kernel).
constexpr uint32_t kDetectLanes = 8;

// Merges O's entries into L (both ordered by (distance, face id)), keeping
// the kMaxContacts nearest; O's entries come in order, so the first one
// that does not make the list ends the merge.
__device__ __forceinline__ void merge_near(NearList& L, const NearList& O) {
#pragma unroll
    for (int e = 0; e < kMaxContacts; e++) {
        if ((uint32_t)e >= O.n) break;
        if (!insert_near(L, O.kd[e], O.kf[e], O.kb[e][0], O.kb[e][1], O.kb[e][2])) break;
    }
}

__global__ __launch_bounds__(kBlock) void k_detect_group(DetectArgs d) {
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x % kDetectLanes;
    const uint32_t s = (uint32_t)(t / kDetectLanes);
    const bool valid = s < d.nv;     // groups past the end still take part in the exchanges
    const F3 p = ld3(d.x, valid ? s : 0u);
    NearList L;
    L.n = 0;
    uint32_t f0 = 0, f1 = 0;
    if (valid && detect_cell(d, p, &f0, &f1))
        for (uint32_t j = f0 + lane; j < f1; j += kDetectLanes) consider_face(d, p, d.faces[j], L);
#pragma unroll
    for (uint32_t off = 1; off < kDetectLanes; off <<= 1) {
        NearList O;
        O.n = __shfl_xor(L.n, off);
#pragma unroll
        for (int k = 0; k < kMaxContacts; k++) {
            O.kd[k] = __shfl_xor(L.kd[k], off);
            O.kf[k] = __shfl_xor(L.kf[k], off);
            O.kb[k][0] = __shfl_xor(L.kb[k][0], off);
            O.kb[k][1] = __shfl_xor(L.kb[k][1], off);
            O.kb[k][2] = __shfl_xor(L.kb[k][2], off);
        }
        merge_near(L, O);
    }
    if (valid && lane == 0) write_contacts(d, s, L);
}

// --- persistent frame kernel ---------------------------------------------
// One cooperative launch runs whole frames: predict, every (iteration, colour)
// phase, update velocity; a grid barrier separates dependent phases.  All
// blocks are co-resident (hipLaunchCooperativeKernel), so every block reaches
// every barrier; the wait is still bounded: after kBarrierSpinLimit polls a
// block raises the error word (barrier[1]) and leaves, and every block that
// sees the word leaves at its next barrier, so the grid always drains.
struct PersistArgs {
    VbdArgs a;                 // a.x = positions, a.fp = frame parameters
    float* x0;                 // x_start (written by update velocity)
    float* v;                  // velocity (written by update velocity)
    FrameParams* fp;
    const ColourInfo* colours;
    uint32_t* barrier;         // [0] monotonically increasing arrival counter, [1] error word
    uint32_t nv, ncolours, iters, nframes;
    float damp;
    int detect;                // run contact detection after predict
    DetectArgs det;
};

constexpr uint32_t kBarrierSpinLimit = 1u << 24;

// false: a block gave up waiting (error word set); the caller leaves the kernel
__device__ __forceinline__ bool grid_barrier(uint32_t* bar, uint32_t& target) {
    __shared__ uint32_t s_fail;
    target += gridDim.x;
    __threadfence();           // every wave: its own stores reach device scope
    __syncthreads();
    if (threadIdx.x == 0) {
        s_fail = 0u;
        atomicAdd(bar, 1u);
        uint32_t spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || ++spins > kBarrierSpinLimit) {
                __hip_atomic_store(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_fail = 1u;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
    }
    __syncthreads();
    return s_fail == 0u;
}

template <int G>
__global__ __launch_bounds__(kBlock) void k_vbd_persistent(PersistArgs p) {
    uint32_t target = 0;
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t t0 = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t frame0 = p.fp->frame;
    for (uint32_t f = 0; f < p.nframes; f++) {
        // external force of this frame, read by every thread
        const float* src = p.fp->forces ? p.fp->forces + 3 * (size_t)(frame0 + f) : p.fp->ext;
        const F3 ext = F3{src[0], src[1], src[2]};
        // predict (Core::predict_position)
        for (uint32_t s = t0; s < p.nv; s += stride) {
            F3 vpred = add3(ld3(p.v, s), scale3(p.a.h, F3{0.f, -9.8f, 0.f}));
            if ((p.a.meta[s] >> 16) & 1u) vpred = F3{0.f, 0.f, 0.f};
            st3(p.a.x, s, add3(ld3(p.x0, s), scale3(p.a.h, vpred)));
        }
        if (!grid_barrier(p.barrier, target)) return;
        if (p.detect) {
            for (uint32_t s = t0; s < p.nv; s += stride) detect_vertex(p.det, s);
            if (!grid_barrier(p.barrier, target)) return;
        }
        for (uint32_t it = 0; it < p.iters; it++) {
            for (uint32_t c = 0; c < p.ncolours; c++) {
                const ColourInfo ci = p.colours[c];
                if (G == 1) {
                    for (uint32_t i = t0; i < ci.count; i += stride)
                        vbd_vertex_ext(p.a, ci.start + i, i, ci.count, ci.eoff, ci.boff, ci.ew, ci.bw, ext);
                } else {
                    // a whole wave moves through the colour together
                    for (uint32_t i = t0 / G; i - (i % (64 / G)) < ci.count; i += stride / G) {
                        const bool valid = i < ci.count;
                        vbd_vertex_group<G>(p.a, ci.start + (valid ? i : 0u), valid ? i : 0u, ci.count, ci.eoff,
                                            ci.boff, ci.ew, ci.bw, ext, t0 % G, valid);
                    }
                }
                if (!grid_barrier(p.barrier, target)) return;
            }
        }
        // update velocity (Core::update_velocity)
        for (uint32_t s = t0; s < p.nv; s += stride) {
            const F3 xk = ld3(p.a.x, s);
            const F3 dx = sub3(xk, ld3(p.x0, s));
            F3 vel = F3{dx.x / p.a.h, dx.y / p.a.h, dx.z / p.a.h};
            vel = F3{vel.x * p.damp, vel.y * p.damp, vel.z * p.damp};
            st3(p.v, s, vel);
            st3(p.x0, s, xk);
        }
        if (!grid_barrier(p.barrier, target)) return;
    }
    if (t0 == 0) {
        p.fp->frame = frame0 + p.nframes;
        const float* src = p.fp->forces ? p.fp->forces + 3 * (size_t)(frame0 + p.nframes - 1) : p.fp->ext;
        p.fp->ext_cur[0] = src[0];
        p.fp->ext_cur[1] = src[1];
        p.fp->ext_cur[2] = src[2];
    }
}

// --- frame-resident colour passes with neighbour hand-off ("flow") ---------
// One cooperative launch runs whole frames.  The cloth is cut into tiles of
// consecutive vertex ids, one block each.  Phase q of a launch (predict,
// then every (iteration, colour) pass) may start on a tile once every
// neighbouring tile (a tile holding a vertex that this tile's updates read,
// or that reads this tile's vertices) has completed phase q - 1: per-tile
// progress counters replace the grid barrier.  That is enough for the
// Gauss-Seidel order: a neighbour's vertices of an earlier colour are already
// updated, those of a later colour are not yet (a neighbour cannot start
// q + 1 before this tile completed q), so every update reads exactly what the
// colour-by-colour kernel sequence reads and the result is bit-identical.
// Positions, x_start and velocity move through device-scope (L2-coherent
// across XCDs) loads and stores; a tile publishes a phase after its stores
// have completed.  Waits are bounded: on a time-out the block raises the
// error word and leaves, and the others leave when they see it.
struct FlowArgs {
    VbdArgs a;                 // a.x positions (x0, v read through a as well)
    float* x0;
    float* v;
    const FrameParams* fp;
    const ColourInfo* colours;
    const uint2* trange;       // [tiles][colours]: index range inside the colour
    const uint32_t* nb_off;    // [tiles + 1]
    const uint32_t* nb;        // neighbouring tiles
    uint32_t* progress;        // [tiles]: last phase the tile completed
    uint32_t* error;           // 1 after a wait timed out
    uint32_t base;             // phase every tile completed before this launch
    uint32_t ncolours, iters, nframes;
    float damp;
};

constexpr uint32_t kFlowSpinLimit = 1u << 22;

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int G, class Mem>
__global__ __launch_bounds__(kBlock) void k_vbd_flow(FlowArgs p) {
    __shared__ uint32_t s_abort;
    const uint32_t t = blockIdx.x, tid = threadIdx.x, C = p.ncolours;
    if (tid == 0) s_abort = 0u;
    __syncthreads();
    const uint32_t nb0 = p.nb_off[t], nb1 = p.nb_off[t + 1];
    const uint32_t frame0 = p.fp->frame;   // advanced only after the launch (k_flow_finish)
    // every neighbouring tile has completed phase `need` (false: give up)
    auto wait_for = [&](uint32_t need) -> bool {
        for (uint32_t j = nb0 + tid; j < nb1; j += kBlock) {
            const uint32_t* flag = p.progress + p.nb[j];
            uint32_t spins = 0;
            while (ld_agent(flag) < need) {
                if (++spins > kFlowSpinLimit || ld_agent(p.error) != 0u) {
                    __hip_atomic_store(p.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_abort = 1u;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        // pairs with the neighbours' release in publish(): their position
        // stores happen before this block's reads of them
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
        return s_abort == 0u;
    };
    // this tile completed phase `q`: its coherent stores are done, then the flag
    auto publish = [&](uint32_t q) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // every wave's stores before the flag
        __syncthreads();
        if (tid == 0) __hip_atomic_store(p.progress + t, q, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    };
    uint32_t q = p.base;
    for (uint32_t f = 0; f < p.nframes; f++) {
        const float* src = p.fp->forces ? p.fp->forces + 3 * (size_t)(frame0 + f) : p.fp->ext;
        const F3 ext = F3{src[0], src[1], src[2]};
        // predict (Core::predict_position): own vertices, after the neighbours
        // finished reading them in the previous frame's last pass
        if (!wait_for(q)) return;
        for (uint32_t c = 0; c < C; c++) {
            const uint2 r = p.trange[t * C + c];
            const uint32_t cs = p.colours[c].start;
            for (uint32_t i = r.x + tid; i < r.y; i += kBlock) {
                const uint32_t s = cs + i;
                F3 vpred = add3(Mem::ld(p.v, s), scale3(p.a.h, F3{0.f, -9.8f, 0.f}));
                if ((p.a.meta[s] >> 16) & 1u) vpred = F3{0.f, 0.f, 0.f};
                Mem::st(p.a.x, s, add3(Mem::ld(p.x0, s), scale3(p.a.h, vpred)));
            }
        }
        publish(++q);
        for (uint32_t it = 0; it < p.iters; it++) {
            for (uint32_t c = 0; c < C; c++) {
                if (!wait_for(q)) return;
                const ColourInfo ci = p.colours[c];
                const uint2 r = p.trange[t * C + c];
                if (G == 1) {
                    for (uint32_t i = r.x + tid; i < r.y; i += kBlock)
                        vbd_vertex_ext<Mem>(p.a, ci.start + i, i, ci.count, ci.eoff, ci.boff, ci.ew, ci.bw, ext);
                } else {
                    // groups of G lanes; a wave leaves together
                    for (uint32_t i0 = r.x; i0 < r.y; i0 += kBlock / G) {
                        const uint32_t i = i0 + tid / G;
                        if (i0 + (tid / 64) * (64 / G) >= r.y) break;
                        const bool valid = i < r.y;
                        vbd_vertex_group<G, Mem>(p.a, ci.start + (valid ? i : r.x), valid ? i : r.x, ci.count,
                                                         ci.eoff, ci.boff, ci.ew, ci.bw, ext, tid % G, valid);
                    }
                }
                publish(++q);
            }
        }
        // update velocity (Core::update_velocity): own vertices only; no
        // other tile reads x_start or v
        for (uint32_t c = 0; c < C; c++) {
            const uint2 r = p.trange[t * C + c];
            const uint32_t cs = p.colours[c].start;
            for (uint32_t i = r.x + tid; i < r.y; i += kBlock) {
                const uint32_t s = cs + i;
                const F3 xk = Mem::ld(p.a.x, s);
                const F3 dx = sub3(xk, Mem::ld(p.x0, s));
                F3 vel = F3{dx.x / p.a.h, dx.y / p.a.h, dx.z / p.a.h};
                vel = F3{vel.x * p.damp, vel.y * p.damp, vel.z * p.damp};
                Mem::st(p.v, s, vel);
                Mem::st(p.x0, s, xk);
            }
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __syncthreads();
    }
}

// After a flow launch: advance the frame counter and publish the last
// frame's external force (as the per-frame kernels leave them).
__global__ void k_flow_finish(FrameParams* fp, uint32_t nframes) {
    if (threadIdx.x != 0) return;
    const uint32_t last = fp->frame + nframes - 1;
    const float* src = fp->forces ? fp->forces + 3 * (size_t)last : fp->ext;
    fp->ext_cur[0] = src[0];
    fp->ext_cur[1] = src[1];
    fp->ext_cur[2] = src[2];
    fp->frame = fp->frame + nframes;
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0u;
}

// Core::predict_position (xpbd_constraints.h:321-384); gravity hard-coded
// (0,-9.8,0) as the reference.  Block 0 / thread 0 also publishes this
// frame's external force.
__global__ __launch_bounds__(kBlock) void k_predict(float* x, const float* x0, const float* v, const uint32_t* meta,
                                                    FrameParams* fp, float h, uint32_t nv) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s == 0) {
        const float* src = fp->forces ? fp->forces + 3 * (size_t)fp->frame : fp->ext;
        fp->ext_cur[0] = src[0];
        fp->ext_cur[1] = src[1];
        fp->ext_cur[2] = src[2];
    }
    if (s >= nv) return;
    const F3 xp = ld3(x0, s);
    const F3 vp = ld3(v, s);
    F3 vpred = add3(vp, scale3(h, F3{0.f, -9.8f, 0.f}));
    if ((meta[s] >> 16) & 1u) vpred = F3{0.f, 0.f, 0.f};
    st3(x, s, add3(xp, scale3(h, vpred)));
}

// Core::update_velocity (xpbd_constraints.h:535-581).  damp = exp(-c h).
__global__ __launch_bounds__(kBlock) void k_update_velocity(const float* x, float* x0, float* v, FrameParams* fp,
                                                            float h, float damp, uint32_t nv, int advance_frame) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s == 0 && advance_frame) fp->frame = fp->frame + 1;
    if (s >= nv) return;
    const F3 xk = ld3(x, s);
    const F3 dx = sub3(xk, ld3(x0, s));
    F3 vel = F3{dx.x / h, dx.y / h, dx.z / h};
    vel = F3{vel.x * damp, vel.y * damp, vel.z * damp};
    st3(v, s, vel);
    st3(x0, s, xk);
}

// --- energy ---------------------------------------------------------------
struct EnergyArgs {
    const float* x;
    const float* x0;
    const float* v;
    const float* mass;
    const uint32_t* meta;
    const uint2* edges;      // slots
    const float* rest;
    const uint4* pairs;      // slots
    const float4* Q;         // 4 float4 per pair: column-major Q
    const FrameParams* fp;
    uint32_t nv, ne, nb;
    float h, ks, kb;
    float g[3];
    const uint32_t* ccount;    // obstacle contacts (null: none)
    const float4* ct_xs;
    const float4* ct_n;
    float thickness;
};

__device__ __forceinline__ double block_sum(double val) {
    __shared__ double red[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) val += __shfl_xor(val, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = val;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kBlock / 64; k++) r += red[k];
    return r;
}

// Per-element FP32 energies (Energy::compute_energy_inertia / _stretch_mass_spring /
// _bending, xpbd_constraints.h:2079-2214), FP64 block partial sums.
__global__ __launch_bounds__(kBlock) void k_energy_partial(EnergyArgs a, double* partial) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    double val = 0.0;
    if (t < a.nv) {
        const uint32_t s = t;
        const float m = a.mass[s];
        F3 aext = F3{a.g[0], a.g[1], a.g[2]};
        const F3 w = F3{a.fp->ext_cur[0], a.fp->ext_cur[1], a.fp->ext_cur[2]};
        aext = add3(aext, F3{w.x / m, w.y / m, w.z / m});
        if ((a.meta[s] >> 16) & 1u) aext = F3{0.f, 0.f, 0.f};
        const F3 y = add3(add3(ld3(a.x0, s), scale3(a.h, ld3(a.v, s))), scale3(a.h * a.h, aext));
        const F3 xv = ld3(a.x, s);
        const F3 d = sub3(xv, y);
        val = (double)((dot3(d, d) * m) / ((2.f * a.h) * a.h));
        if (a.ccount) {
            // Energy::compute_energy_collision_vf per contact
            for (uint32_t k = 0; k < a.ccount[s]; k++) {
                const float4 q = a.ct_xs[k * a.nv + s];
                const float4 nn = a.ct_n[k * a.nv + s];
                val += (double)contact_energy(xv, F3{q.x, q.y, q.z}, F3{nn.x, nn.y, nn.z}, q.w, a.thickness);
            }
        }
    } else if (t < a.nv + a.ne) {
        const uint32_t k = t - a.nv;
        const uint2 ed = a.edges[k];
        const F3 diff = sub3(ld3(a.x, ed.y), ld3(a.x, ed.x));
        const float l2 = dot3(diff, diff);
        if (!((double)l2 < 1e-8)) {
            const float C = sqrtf(l2) - a.rest[k];
            val = (double)((0.5f * a.ks) * (C * C));
        }
    } else if (t < a.nv + a.ne + a.nb) {
        const uint32_t k = t - a.nv - a.ne;
        const uint4 id = a.pairs[k];
        F3 p[4] = {ld3(a.x, id.x), ld3(a.x, id.y), ld3(a.x, id.z), ld3(a.x, id.w)};
        float sum = 0.f;
        for (int ii = 0; ii < 4; ii++) {
            const float4 qc = a.Q[4 * k + ii];
            const float qq[4] = {qc.x, qc.y, qc.z, qc.w};
            for (int jj = 0; jj < 4; jj++) sum = sum + qq[jj] * dot3(p[ii], p[jj]);
        }
        val = (double)((0.5f * a.kb) * sum);
    }
    const double r = block_sum(val);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(kBlock) void k_energy_finalize(const double* partial, uint32_t nblk, uint32_t row0,
                                                            float* out) {
    const uint32_t row = row0 + blockIdx.x;
    const double* p = partial + (size_t)row * nblk;
    double acc = 0.0;
    for (uint32_t k = threadIdx.x; k < nblk; k += kBlock) acc += p[k];
    const double r = block_sum(acc);
    if (threadIdx.x == 0) out[row] = (float)r;
}

// FP64 block partials of lats_vbd_energy_noalloc: one copy of this array per
// GPU, held until the process ends.  Host code never passes its name to a
// kernel (on the host the name is a stand-in, not a device address): it uses
// the address hipGetSymbolAddress returned for the device (scratch_address).
constexpr uint32_t kScratchPartials = 32768;   // 256 KiB: energy_nblk <= 32768
static __device__ __align__(256) double scratch[kScratchPartials];

// --- multi-buffer helpers of the heterogeneous schedule ------------------
__global__ __launch_bounds__(kBlock) void k_copy_to_2(const float* a, float* b, float* c, uint32_t n) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const float val = a[t];
    if (b) b[t] = val;
    if (c) c[t] = val;
}

// Core::read_and_solve_conflict (xpbd_constraints.h:815-869)
__global__ __launch_bounds__(kBlock) void k_read_and_solve_conflict(float* begin_self, float* iter_self,
                                                                    const float* iter_other, float weight,
                                                                    uint32_t nv) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= nv) return;
    const F3 ps = ld3(iter_self, s), po = ld3(iter_other, s);
    F3 wp;
    if (weight > 1.19209290E-7f) {
        const F3 init = ld3(begin_self, s);
        const F3 sdx = sub3(ps, init), adx = sub3(po, init);
        wp = init;
        const bool aa = dot3(sdx, sdx) != 0.f, ab = dot3(adx, adx) != 0.f;
        if (aa && ab) wp = add3(init, scale3(weight, add3(sdx, adx)));
        else if (aa) wp = ps;
        else if (ab) wp = po;
    } else {
        wp = scale3(0.5f, add3(ps, po));
    }
    st3(iter_self, s, wp);
    st3(begin_self, s, wp);
}

// --- XPBD: one thread per constraint of one cluster -------------------------
__global__ __launch_bounds__(kBlock) void k_xpbd_stretch(XpbdArgs a, uint32_t start, uint32_t count) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < count) xpbd_stretch(a, start + t);
}
__global__ __launch_bounds__(kBlock) void k_xpbd_bending(XpbdArgs a, uint32_t start, uint32_t count) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < count) xpbd_bending(a, start + t);
}
// a fill kernel rather than hipMemsetAsync: frames are captured into graphs
__global__ __launch_bounds__(kBlock) void k_fill(float* p, float v, uint32_t n) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ __launch_bounds__(kBlock) void k_gather_rows(const float* src, const uint32_t* idx, float* dst, uint32_t n) {
    // dst[t] = src[idx[t]] for float3 rows
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t k = idx[t];
    dst[3 * t] = src[3 * k];
    dst[3 * t + 1] = src[3 * k + 1];
    dst[3 * t + 2] = src[3 * k + 2];
}

}  // namespace

// ===========================================================================
// Host context
// ===========================================================================
namespace {

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


}  // namespace


namespace {

// Device address of `scratch` per device, looked up once per device (with
// that device current) when the first context is created on it.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxDevices] = {};

int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(LATS_ERR_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[device]) return LATS_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(LATS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[device] = static_cast<double*>(p);
    return LATS_OK;
}

double* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_addr[device];
}

int ensure_energy_rows(lats_vbd* c, uint32_t rows) {
    if (rows <= c->energy_cap) return LATS_OK;
    const uint32_t cap = std::max(rows, c->energy_cap * 2 + 8);
    DBuf<double> np;
    DBuf<float> ne;
    HIPCHK(np.alloc((size_t)cap * c->energy_nblk));
    HIPCHK(ne.alloc(cap));
    if (c->energy_rows) {
        // both the partial sums and the rows already finalised (the dispatch
        // path finalises one row at a time)
        HIPCHK(hipMemcpyAsync(np.p, c->partial.p, sizeof(double) * (size_t)c->energy_rows * c->energy_nblk,
                              hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(ne.p, c->energy.p, sizeof(float) * (size_t)c->energy_rows, hipMemcpyDeviceToDevice,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    std::swap(c->partial.p, np.p);
    std::swap(c->partial.n, np.n);
    std::swap(c->energy.p, ne.p);
    std::swap(c->energy.n, ne.n);
    c->energy_cap = cap;
    // graphs captured against the old buffers are stale
    c->drop_graphs();
    return LATS_OK;
}

// Enqueue the FP64 block partials (energy_nblk of them) of one energy
// evaluation of buffer xb into partial[0, energy_nblk).
void enqueue_energy_partials(lats_vbd* c, const float* xb, double* partial, hipStream_t st) {
    EnergyArgs a;
    a.x = xb;
    a.x0 = c->x0.p;
    a.v = c->v.p;
    a.mass = c->mass.p;
    a.meta = c->meta.p;
    a.edges = c->e_slots.p;
    a.rest = c->rest.p;
    a.pairs = c->b_slots.p;
    a.Q = c->Qd.p;
    a.fp = c->fp.p;
    a.nv = c->V;
    a.ne = c->E;
    a.nb = c->energy_bending ? c->B : 0;
    a.h = c->h_energy;
    a.ks = c->prm.stiffness_stretch_spring;
    a.kb = c->prm.stiffness_quadratic_bending;
    for (int k = 0; k < 3; k++) a.g[k] = c->prm.gravity[k];
    a.ccount = c->has_obstacle ? c->ct_count.p : nullptr;
    a.ct_xs = c->ct_xs.p;
    a.ct_n = c->ct_n.p;
    a.thickness = c->thickness;
    hipLaunchKernelGGL(k_energy_partial, dim3(c->energy_nblk), dim3(kBlock), 0, st ? st : c->stream, a, partial);
}

// Enqueue one energy evaluation of buffer xb into history row `row`.
void enqueue_energy(lats_vbd* c, const float* xb, uint32_t row, hipStream_t st = nullptr) {
    enqueue_energy_partials(c, xb, c->partial.p + (size_t)row * c->energy_nblk, st);
}

// The colour passes cut every colour into `n` bands of equal index count.
void split_regions(lats_vbd* c, uint32_t n) { c->regions = n; }

template <int B>
void launch_colour(lats_vbd* c, const VbdArgs& a, const ColourInfo& ci) {
    const uint32_t per_block = B / c->lanes, R = c->regions;
    const uint32_t band = (ci.count + R - 1) / R;
    const dim3 grid(R, std::max(1u, (band + per_block - 1) / per_block)), block(B);
#define LATS_COLOUR_ARGS a.ell_e, a.ell_bq, a.ell_bi, ci.start, ci.count, ci.eoff, ci.boff, ci.ew, ci.bw, band, a
    switch (c->lanes) {
        case 4: hipLaunchKernelGGL((k_vbd_colour<4, B>), grid, block, 0, c->stream, LATS_COLOUR_ARGS); break;
        case 8: hipLaunchKernelGGL((k_vbd_colour<8, B>), grid, block, 0, c->stream, LATS_COLOUR_ARGS); break;
        case 16: hipLaunchKernelGGL((k_vbd_colour<16, B>), grid, block, 0, c->stream, LATS_COLOUR_ARGS); break;
        default: hipLaunchKernelGGL((k_vbd_colour<1, B>), grid, block, 0, c->stream, LATS_COLOUR_ARGS); break;
    }
#undef LATS_COLOUR_ARGS
}

void enqueue_colour(lats_vbd* c, float* xb, uint32_t col) {
    const VbdArgs a = c->args(xb);
    const ColourInfo& ci = c->colours[col];
    switch (c->colour_block) {
        case 64: launch_colour<64>(c, a, ci); break;
        case 128: launch_colour<128>(c, a, ci); break;
        default: launch_colour<256>(c, a, ci); break;
    }
}

void enqueue_predict(lats_vbd* c) {
    hipLaunchKernelGGL(k_predict, dim3(std::max(1u, nblocks(c->V))), dim3(kBlock), 0, c->stream, c->x.p, c->x0.p,
                       c->v.p, c->meta.p, c->fp.p, c->h, c->V);
}

DetectArgs detect_args(lats_vbd* c, const float* xb) {
    DetectArgs d;
    d.x = xb;
    d.nv = c->V;
    d.ob_x = c->ob_x.p;
    d.ob_f = c->ob_f.p;
    d.ob_n = c->ob_n.p;
    d.belongs = c->hb_belongs.p;
    d.start = c->hb_start.p;
    d.count = c->hb_count.p;
    d.faces = c->hb_faces.p;
    d.ox = c->ob_hash.origin[0];
    d.oy = c->ob_hash.origin[1];
    d.oz = c->ob_hash.origin[2];
    d.inv_cell = c->ob_hash.inv_cell;
    d.dimx = c->ob_hash.dims[0];
    d.dimy = c->ob_hash.dims[1];
    d.dimz = c->ob_hash.dims[2];
    d.table_size = c->ob_hash.table_size;
    d.radius = c->ob_radius;
    d.stiffness = c->ob_stiffness;
    d.ccount = c->ct_count.p;
    d.ct_xs = c->ct_xs.p;
    d.ct_n = c->ct_n.p;
    d.ct_fid = c->ct_fid.p;
    return d;
}

void enqueue_detect(lats_vbd* c) {
    if (!c->has_obstacle) return;
    hipLaunchKernelGGL(k_detect_group, dim3(nblocks((uint64_t)c->V * kDetectLanes)), dim3(kBlock), 0, c->stream,
                       detect_args(c, c->x.p));
}

void enqueue_update(lats_vbd* c, int advance) {
    hipLaunchKernelGGL(k_update_velocity, dim3(std::max(1u, nblocks(c->V))), dim3(kBlock), 0, c->stream, c->x.p,
                       c->x0.p, c->v.p, c->fp.p, c->h, c->damp, c->V, advance);
}

XpbdArgs xpbd_args(lats_vbd* c) {
    XpbdArgs a;
    a.x = c->x.p;
    a.minv = c->x_minv.p;
    a.edges = c->xs_edges.p;
    a.rest = c->xs_rest.p;
    a.lambda_s = c->xs_lambda.p;
    a.pairs = c->xb_pairs.p;
    a.Q = c->xb_Q.p;
    a.lambda_b = c->xb_lambda.p;
    a.ks = c->prm.stiffness_stretch_spring;
    a.kb = c->prm.stiffness_quadratic_bending;
    a.h = c->h;
    return a;
}

// reset_constrains (example2_async_iter.cpp:698-710)
void enqueue_xpbd_reset(lats_vbd* c) {
    if (c->E) hipLaunchKernelGGL(k_fill, dim3(nblocks(c->E)), dim3(kBlock), 0, c->stream, c->xs_lambda.p, 0.0f, c->E);
    if (c->B) hipLaunchKernelGGL(k_fill, dim3(nblocks(c->B)), dim3(kBlock), 0, c->stream, c->xb_lambda.p, 0.0f, c->B);
}
void enqueue_xpbd_stretch(lats_vbd* c, uint32_t k) {
    const uint32_t lo = c->xs_bound[k], n = c->xs_bound[k + 1] - lo;
    if (n) hipLaunchKernelGGL(k_xpbd_stretch, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, xpbd_args(c), lo, n);
}
void enqueue_xpbd_bending(lats_vbd* c, uint32_t k) {
    const uint32_t lo = c->xb_bound[k], n = c->xb_bound[k + 1] - lo;
    if (n) hipLaunchKernelGGL(k_xpbd_bending, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, xpbd_args(c), lo, n);
}

// CpuSolver::physics_step_xpbd (example2_async_iter.cpp:1017-1081) with the
// iteration of solve_constraints_XPBD (:1510-1534).
void enqueue_xpbd_frame(lats_vbd* c, bool record) {
    const uint32_t nsub = record ? 1u : c->prm.num_substep;
    for (uint32_t sub = 0; sub < nsub; sub++) {
        enqueue_predict(c);
        enqueue_xpbd_reset(c);
        for (uint32_t it = 0; it < c->prm.constraint_iter_count; it++) {
            if (record && it == 0) enqueue_energy(c, c->x.p, 0);
            for (uint32_t k = 0; k + 1 < c->xs_bound.size(); k++) enqueue_xpbd_stretch(c, k);
            for (uint32_t k = 0; k + 1 < c->xb_bound.size(); k++) enqueue_xpbd_bending(c, k);
            if (record) enqueue_energy(c, c->x.p, it + 1);
        }
        enqueue_update(c, sub + 1 == nsub);
    }
    if (record)
        hipLaunchKernelGGL(k_energy_finalize, dim3(c->prm.constraint_iter_count + 1), dim3(kBlock), 0, c->stream,
                           c->partial.p, c->energy_nblk, 0u, c->energy.p);
}

// Constraint clusters and constraints in cluster order, from the element
// lists already on the device (the reference's edge and pair numbering).
int build_xpbd(lats_vbd* c) {
    if (c->xpbd_built) return LATS_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<uint2> es(c->E);
    std::vector<uint4> bs(c->B);
    std::vector<float> rest(c->E);
    std::vector<float4> Q(4 * (size_t)c->B);
    std::vector<float> mass(c->V);
    std::vector<uint32_t> meta(c->V);
    if (c->E) HIPCHK(hipMemcpy(es.data(), c->e_slots.p, es.size() * sizeof(uint2), hipMemcpyDeviceToHost));
    if (c->E) HIPCHK(hipMemcpy(rest.data(), c->rest.p, rest.size() * sizeof(float), hipMemcpyDeviceToHost));
    if (c->B) HIPCHK(hipMemcpy(bs.data(), c->b_slots.p, bs.size() * sizeof(uint4), hipMemcpyDeviceToHost));
    if (c->B) HIPCHK(hipMemcpy(Q.data(), c->Qd.p, Q.size() * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(mass.data(), c->mass.p, mass.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(meta.data(), c->meta.p, meta.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> ev(2 * (size_t)c->E), bv(4 * (size_t)c->B);
    for (uint32_t e = 0; e < c->E; e++) {
        ev[2 * (size_t)e] = es[e].x;
        ev[2 * (size_t)e + 1] = es[e].y;
    }
    for (uint32_t b = 0; b < c->B; b++) {
        bv[4 * (size_t)b] = bs[b].x;
        bv[4 * (size_t)b + 1] = bs[b].y;
        bv[4 * (size_t)b + 2] = bs[b].z;
        bv[4 * (size_t)b + 3] = bs[b].w;
    }
    // slots relabel the vertices; the colouring depends only on which
    // constraints share a vertex, so it equals the one on vertex ids
    const auto cs = colour_constraints(ev, 2, c->V);
    const auto cb = colour_constraints(bv, 4, c->V);
    auto pack = [](const std::vector<std::vector<uint32_t>>& cl, std::vector<uint32_t>& bound, std::vector<uint32_t>& csr) {
        const uint32_t n = (uint32_t)cl.size();
        bound.assign(1, 0u);
        csr.assign(n + 1, 0u);
        uint32_t p = n + 1;
        for (uint32_t k = 0; k < n; k++) {
            csr[k] = p;
            p += (uint32_t)cl[k].size();
            bound.push_back(bound.back() + (uint32_t)cl[k].size());
        }
        csr[n] = p;
        for (const auto& v : cl) csr.insert(csr.end(), v.begin(), v.end());
    };
    pack(cs, c->xs_bound, c->xs_csr);
    pack(cb, c->xb_bound, c->xb_csr);
    std::vector<uint2> me;
    std::vector<float> mr;
    for (const auto& cl : cs)
        for (uint32_t e : cl) {
            me.push_back(es[e]);
            mr.push_back(rest[e]);
        }
    std::vector<uint4> mb;
    std::vector<float4> mq;
    for (const auto& cl : cb)
        for (uint32_t b : cl) {
            mb.push_back(bs[b]);
            for (int k = 0; k < 4; k++) mq.push_back(Q[4 * (size_t)b + k]);
        }
    // example2_async_iter.cpp:223: 0 for pinned vertices, else 1/m
    std::vector<float> minv(c->V);
    for (uint32_t s = 0; s < c->V; s++) minv[s] = ((meta[s] >> 16) & 1u) ? 0.0f : 1.0f / mass[s];
    HIPCHK(c->xs_edges.upload(me));
    HIPCHK(c->xs_rest.upload(mr));
    HIPCHK(c->xb_pairs.upload(mb));
    HIPCHK(c->xb_Q.upload(mq));
    HIPCHK(c->x_minv.upload(minv));
    HIPCHK(c->xs_lambda.alloc(std::max(1u, c->E)));
    HIPCHK(c->xb_lambda.alloc(std::max(1u, c->B)));
    c->xpbd_built = true;
    return LATS_OK;
}

// Colour sweeps with the patch strategy: the frame state is staged in region
// order, then launches alternate between x and xalt (each reads one and
// writes the other), the result ends in x.  With `record`, the energy rows
// 0..sweeps of the frame are written as well.
void enqueue_patch_sweeps(lats_vbd* c, uint32_t sweeps, bool record) {
    enqueue_patch_stage(c);
    float* cur = c->x.p;
    float* alt = c->xalt.p;
    const uint32_t L = patch_launches_per_sweep(c);
    for (uint32_t it = 0; it < sweeps; it++) {
        if (record && it == 0) enqueue_energy(c, cur, 0);
        for (uint32_t g = 0; g < L; g++) {
            enqueue_patch_launch(c, g, cur, alt);
            std::swap(cur, alt);
        }
        if (record) enqueue_energy(c, cur, it + 1);
    }
    if (cur != c->x.p)
        hipLaunchKernelGGL(k_copy_to_2, dim3(nblocks(3 * (uint64_t)c->V)), dim3(kBlock), 0, c->stream,
                           (const float*)cur, c->x.p, (float*)nullptr, 3 * c->V);
}

bool uses_patches(const lats_vbd* c) {
    return c->solver == LATS_SOLVER_VBD && c->strategy == LATS_STRATEGY_PATCH;
}

// The whole frame as a kernel sequence (captured into a graph by the caller).
void enqueue_frame(lats_vbd* c, bool record) {
    if (c->solver == LATS_SOLVER_XPBD) return enqueue_xpbd_frame(c, record);
    const uint32_t nsub = record ? 1u : c->prm.num_substep;
    for (uint32_t sub = 0; sub < nsub; sub++) {
        enqueue_predict(c);
        enqueue_detect(c);
        if (uses_patches(c)) {
            enqueue_patch_sweeps(c, c->prm.constraint_iter_count, record);
            enqueue_update(c, sub + 1 == nsub);
            continue;
        }
        for (uint32_t it = 0; it < c->prm.constraint_iter_count; it++) {
            if (record && it == 0) enqueue_energy(c, c->x.p, 0);
            for (uint32_t col = 0; col < c->C; col++) enqueue_colour(c, c->x.p, col);
            if (record) enqueue_energy(c, c->x.p, it + 1);
        }
        enqueue_update(c, sub + 1 == nsub);
    }
    if (record)
        hipLaunchKernelGGL(k_energy_finalize, dim3(c->prm.constraint_iter_count + 1), dim3(kBlock), 0, c->stream,
                           c->partial.p, c->energy_nblk, 0u, c->energy.p);
}

const void* persistent_kernel(uint32_t lanes) {
    switch (lanes) {
        case 4: return (const void*)k_vbd_persistent<4>;
        case 8: return (const void*)k_vbd_persistent<8>;
        case 16: return (const void*)k_vbd_persistent<16>;
        default: return (const void*)k_vbd_persistent<1>;
    }
}

int persistent_grid(lats_vbd* c, uint32_t* blocks) {
    if (c->persist_max_lanes != c->lanes || !c->persist_max_blocks) {
        int per_cu = 0, ncu = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_kernel(c->lanes), kBlock, 0));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
        c->persist_max_blocks = (uint32_t)std::max(1, per_cu * ncu);
        c->persist_max_lanes = c->lanes;
    }
    uint32_t g = c->persist_blocks;
    if (!g) {
        uint32_t mx = 0;
        for (uint32_t k = 0; k < c->C; k++) mx = std::max(mx, c->ccount[k]);
        g = nblocks((uint64_t)mx * c->lanes);
    }
    *blocks = std::max(1u, std::min(g, c->persist_max_blocks));
    return LATS_OK;
}

// nframes frames in one cooperative launch (no energy recording).
int run_persistent(lats_vbd* c, uint32_t nframes) {
    if (!nframes) return LATS_OK;
    uint32_t g = 0;
    int rc = persistent_grid(c, &g);
    if (rc) return rc;
    PersistArgs pa;
    pa.a = c->args(c->x.p);
    pa.x0 = c->x0.p;
    pa.v = c->v.p;
    pa.fp = c->fp.p;
    pa.colours = c->colours_d.p;
    pa.barrier = c->barrier.p;
    pa.nv = c->V;
    pa.ncolours = c->C;
    pa.iters = c->prm.constraint_iter_count;
    pa.nframes = nframes * c->prm.num_substep;
    pa.damp = c->damp;
    pa.detect = c->has_obstacle ? 1 : 0;
    pa.det = detect_args(c, c->x.p);
    if (c->prm.num_substep != 1) return fail(LATS_ERR_UNSUPPORTED, "persistent strategy runs one substep per frame");
    // the barrier counter grows by g per phase: split long runs so it stays below 2^31
    const uint64_t per_frame = ((uint64_t)pa.iters * c->C + 2) * g;
    const uint32_t chunk = (uint32_t)std::max<uint64_t>(1, (1ull << 31) / per_frame);
    for (uint32_t done = 0; done < nframes;) {
        pa.nframes = std::min(chunk, nframes - done);
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->stream, c->barrier.p, 2u);
        void* kargs[] = {&pa};
        const hipError_t e = hipLaunchCooperativeKernel(persistent_kernel(c->lanes), dim3(g), dim3(kBlock), kargs, 0,
                                                        c->stream);
        if (e != hipSuccess) {
            c->cooperative_refused = true;
            return fail(LATS_ERR_HIP, std::string("hipLaunchCooperativeKernel (persistent): ") + hipGetErrorString(e));
        }
        done += pa.nframes;
    }
    return LATS_OK;
}

const void* flow_kernel(uint32_t lanes) {
    switch (lanes) {
        case 4: return (const void*)k_vbd_flow<4, CoherentMem>;
        case 8: return (const void*)k_vbd_flow<8, CoherentMem>;
        case 16: return (const void*)k_vbd_flow<16, CoherentMem>;
        default: return (const void*)k_vbd_flow<1, CoherentMem>;
    }
}

// Most blocks of the flow kernel that can be co-resident (cooperative launch).
int flow_capacity(lats_vbd* c, uint32_t* cap) {
    int per_cu = 0, ncu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, flow_kernel(c->lanes), kBlock, 0));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    *cap = (uint32_t)std::max(0, per_cu * ncu);
    return LATS_OK;
}

// Tiles of consecutive vertex ids: per (tile, colour) the index range inside
// the colour (a colour's slots are in ascending vertex id, so a tile's
// vertices of one colour are contiguous) and the neighbouring tiles (any
// tile sharing a spring or a bending pair with it).
int flow_setup(lats_vbd* c, uint32_t tiles) {
    const uint32_t V = c->V, C = c->C;
    tiles = std::max(1u, std::min(tiles, V));
    const uint32_t tv = (V + tiles - 1) / tiles;
    tiles = (V + tv - 1) / tv;
    if (c->flow_tiles == tiles) return LATS_OK;
    HIPCHK(hipStreamSynchronize(c->stream));   // the old counters may still be in use
    std::vector<uint2> range((size_t)tiles * C);
    for (uint32_t k = 0; k < C; k++) {
        const uint32_t* ids = c->perm.data() + c->cstart[k];
        const uint32_t n = c->ccount[k];
        for (uint32_t t = 0; t < tiles; t++) {
            const uint32_t lo = (uint32_t)(std::lower_bound(ids, ids + n, t * tv) - ids);
            const uint32_t hi = (uint32_t)(std::lower_bound(ids, ids + n, (t + 1) * tv) - ids);
            range[(size_t)t * C + k] = make_uint2(lo, hi);
        }
    }
    std::vector<uint2> es(c->E);
    std::vector<uint4> bs(c->B);
    if (c->E) HIPCHK(hipMemcpy(es.data(), c->e_slots.p, c->E * sizeof(uint2), hipMemcpyDeviceToHost));
    if (c->B) HIPCHK(hipMemcpy(bs.data(), c->b_slots.p, c->B * sizeof(uint4), hipMemcpyDeviceToHost));
    auto tile_of = [&](uint32_t slot) { return c->perm[slot] / tv; };
    std::vector<std::vector<uint32_t>> nbs(tiles);
    auto link = [&](uint32_t a, uint32_t b) {
        if (a != b) {
            nbs[a].push_back(b);
            nbs[b].push_back(a);
        }
    };
    for (const uint2& e : es) link(tile_of(e.x), tile_of(e.y));
    for (const uint4& b : bs) {
        const uint32_t q[4] = {tile_of(b.x), tile_of(b.y), tile_of(b.z), tile_of(b.w)};
        for (int i = 0; i < 4; i++)
            for (int j = i + 1; j < 4; j++) link(q[i], q[j]);
    }
    std::vector<uint32_t> off(tiles + 1, 0), nb;
    for (uint32_t t = 0; t < tiles; t++) {
        std::sort(nbs[t].begin(), nbs[t].end());
        nbs[t].erase(std::unique(nbs[t].begin(), nbs[t].end()), nbs[t].end());
        off[t + 1] = off[t] + (uint32_t)nbs[t].size();
        nb.insert(nb.end(), nbs[t].begin(), nbs[t].end());
    }
    if (nb.empty()) nb.push_back(0);
    HIPCHK(c->flow_range.upload(range));
    HIPCHK(c->flow_nb_off.upload(off));
    HIPCHK(c->flow_nb.upload(nb));
    HIPCHK(c->flow_progress.upload(std::vector<uint32_t>(tiles, 0u)));
    HIPCHK(c->flow_error.upload(std::vector<uint32_t>(1, 0u)));
    c->flow_tiles = tiles;
    c->flow_base = 0;
    return LATS_OK;
}

int run_flow(lats_vbd* c, uint32_t nframes) {
    if (!nframes) return LATS_OK;
    if (c->prm.num_substep != 1) return fail(LATS_ERR_UNSUPPORTED, "flow strategy runs one substep per frame");
    if (c->has_obstacle) return fail(LATS_ERR_UNSUPPORTED, "flow strategy runs cloths without an obstacle");
    uint32_t cap = 0;
    int rc = flow_capacity(c, &cap);
    if (rc) return rc;
    const uint32_t want = c->persist_blocks ? c->persist_blocks : std::min(256u, std::max(1u, c->V / 128));
    if (want > cap) return fail(LATS_ERR_UNSUPPORTED, "flow tiles exceed the co-resident blocks of this device");
    rc = flow_setup(c, want);
    if (rc) return rc;
    const uint64_t per_frame = 1 + (uint64_t)c->prm.constraint_iter_count * c->C;
    const uint32_t chunk = (uint32_t)std::max<uint64_t>(1, (1ull << 30) / per_frame);
    FlowArgs fa;
    fa.a = c->args(c->x.p);
    fa.x0 = c->x0.p;
    fa.v = c->v.p;
    fa.fp = c->fp.p;
    fa.colours = c->colours_d.p;
    fa.trange = c->flow_range.p;
    fa.nb_off = c->flow_nb_off.p;
    fa.nb = c->flow_nb.p;
    fa.progress = c->flow_progress.p;
    fa.error = c->flow_error.p;
    fa.ncolours = c->C;
    fa.iters = c->prm.constraint_iter_count;
    fa.damp = c->damp;
    for (uint32_t done = 0; done < nframes;) {
        fa.nframes = std::min(chunk, nframes - done);
        if ((uint64_t)c->flow_base + fa.nframes * per_frame >= (1ull << 31)) {
            hipLaunchKernelGGL(k_zero_u32, dim3(nblocks(c->flow_tiles)), dim3(kBlock), 0, c->stream,
                               c->flow_progress.p, c->flow_tiles);
            c->flow_base = 0;
        }
        fa.base = c->flow_base;
        void* kargs[] = {&fa};
        const hipError_t e = hipLaunchCooperativeKernel(flow_kernel(c->lanes), dim3(c->flow_tiles), dim3(kBlock), kargs,
                                                        0, c->stream);
        if (e != hipSuccess) {
            c->cooperative_refused = true;
            return fail(LATS_ERR_HIP, std::string("hipLaunchCooperativeKernel (flow): ") + hipGetErrorString(e));
        }
        hipLaunchKernelGGL(k_flow_finish, dim3(1), dim3(64), 0, c->stream, c->fp.p, fa.nframes);
        c->flow_base += (uint32_t)(fa.nframes * per_frame);
        done += fa.nframes;
    }
    return LATS_OK;
}

// After a synchronisation: a persistent or flow launch whose wait timed out
// leaves an error word; report it (and reset the flow tile counters).
int flow_check(lats_vbd* c) {
    if (c->strategy == LATS_STRATEGY_PERSISTENT && c->solver == LATS_SOLVER_VBD) {
        uint32_t err = 0;
        HIPCHK(hipMemcpy(&err, c->barrier.p + 1, sizeof(err), hipMemcpyDeviceToHost));
        if (!err) return LATS_OK;
        return fail(LATS_ERR_TIMEOUT, "persistent strategy: a grid barrier wait gave up (frame results invalid)");
    }
    if (c->strategy != LATS_STRATEGY_FLOW || !c->flow_tiles) return LATS_OK;
    uint32_t err = 0;
    HIPCHK(hipMemcpy(&err, c->flow_error.p, sizeof(err), hipMemcpyDeviceToHost));
    if (!err) return LATS_OK;
    c->flow_tiles = 0;   // next run rebuilds counters from zero
    return fail(LATS_ERR_HIP, "flow strategy: a tile waited too long for its neighbours (frame results invalid)");
}

int get_graph(lats_vbd* c, bool record, hipGraphExec_t* out) {
    if (uses_patches(c)) {
        int rc = patch_prepare(c);
        if (rc) return rc;
    }
    if (record) {
        int rc = ensure_energy_rows(c, c->prm.constraint_iter_count + 1);
        if (rc) return rc;
    }
    hipGraphExec_t& g = c->graph[record ? 1 : 0];
    if (!g) {
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue_frame(c, record);
        HIPCHK(hipStreamEndCapture(c->stream, &graph));
        hipError_t e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g = nullptr;
            return fail(LATS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    *out = g;
    return LATS_OK;
}

int run_frame(lats_vbd* c, bool record) {
    const bool vbd = c->solver == LATS_SOLVER_VBD;
    if (vbd && !record && c->strategy == LATS_STRATEGY_PERSISTENT) return run_persistent(c, 1);
    if (vbd && !record && c->strategy == LATS_STRATEGY_FLOW) return run_flow(c, 1);
    hipGraphExec_t g;
    int rc = get_graph(c, record, &g);
    if (rc) return rc;
    HIPCHK(hipGraphLaunch(g, c->stream));
    if (record) c->energy_rows = c->prm.constraint_iter_count + 1;
    return LATS_OK;
}

int set_frame_params(lats_vbd* c, const float* forces_dev, uint32_t frame, const float* ext) {
    FrameParams h;
    HIPCHK(hipMemcpyAsync(&h, c->fp.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    h.forces = forces_dev;
    h.frame = frame;
    if (ext) {
        h.ext[0] = ext[0];
        h.ext[1] = ext[1];
        h.ext[2] = ext[2];
    }
    HIPCHK(hipMemcpyAsync(c->fp.p, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LATS_OK;
}


}  // namespace

namespace lats {
int energies_enqueue(lats_vbd* c, const float* const* xs, uint32_t n, hipStream_t st) {
    if (!c || !xs || n == 0) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (n > lats_vbd::kHostEnergies) return fail(LATS_ERR_INVALID_ARGUMENT, "too many position buffers");
    int rc = ensure_energy_rows(c, n);
    if (rc) return rc;
    if (!st) st = c->stream;
    for (uint32_t k = 0; k < n; k++) enqueue_energy(c, xs[k], k, st);
    // the final sums go straight to pinned host memory: no copy command
    hipLaunchKernelGGL(k_energy_finalize, dim3(n), dim3(kBlock), 0, st, c->partial.p, c->energy_nblk, 0u,
                       c->host_energy_d);
    HIPCHK(hipGetLastError());
    c->energy_rows = 0;
    return LATS_OK;
}

int energies_of(lats_vbd* c, const float* const* xs, uint32_t n, float* out) {
    if (!out) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    int rc = energies_enqueue(c, xs, n);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < n; k++) out[k] = c->host_energy[k];
    return LATS_OK;
}
}  // namespace lats

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int lats_vbd_create(const lats_mesh_desc* m, const lats_scene_params* prm, lats_vbd** out) {
    if (!out || !prm || !m) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (prm->num_substep == 0 || !(prm->implicit_dt > 0.f))
        return fail(LATS_ERR_INVALID_ARGUMENT, "implicit_dt must be > 0 and num_substep >= 1");
    const uint32_t V = m->num_verts, E = m->num_edges, B = m->num_bending_edges, C = m->num_clusters;
    ClothLayout L;
    std::vector<uint32_t> inv;
    std::string err;
    if (!append_cloth(*m, L, inv, err)) return fail(LATS_ERR_INVALID_ARGUMENT, err);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LATS_ERR_NO_DEVICE, "no HIP device");

    auto c = new lats_vbd();
    c->V = V;
    c->E = E;
    c->B = B;
    c->C = C;
    c->prm = *prm;
    (void)hipGetDevice(&c->device);
    // scene_params.h:159 get_substep_dt(); the energy uses the same h
    c->h = prm->implicit_dt / (float)prm->num_substep;
    c->h_energy = prm->implicit_dt / (float)prm->num_substep;
    c->h2inv = 1.f / (c->h * c->h);
    c->damp = std::exp(-prm->damping_cloth * c->h);
    c->perm = L.perm;
    c->inv = inv;
    c->colours = L.colours;
    split_regions(c, 1);
    for (const ColourInfo& ci : L.colours) {
        c->cstart.push_back(ci.start);
        c->ccount.push_back(ci.count);
        c->eoff.push_back(ci.eoff);
        c->boff.push_back(ci.boff);
    }
    auto pad = [](auto& v, auto z) { if (v.empty()) v.push_back(z); };
    pad(L.ell_e, uint2{0, 0});
    pad(L.ell_bq, float4{0, 0, 0, 0});
    pad(L.ell_bi, PairIds{0, 0, 0});
    pad(L.e_slots, uint2{0, 0});
    pad(L.rest, 0.f);
    pad(L.b_slots, uint4{0, 0, 0, 0});
    pad(L.Q, float4{0, 0, 0, 0});
    const std::vector<float>& hx = L.x;

    auto cleanup_fail = [&](int code) { delete c; return code; };
#define UPCHK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return cleanup_fail(fail(e_ == hipErrorOutOfMemory ? LATS_ERR_OUT_OF_MEMORY : LATS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_))); } while (0)
    static std::atomic<uint64_t> next_uid{1};
    c->uid = next_uid.fetch_add(1);
    UPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    UPCHK(c->x.upload(hx));
    UPCHK(c->x0.upload(hx));
    UPCHK(c->v.upload(std::vector<float>(3 * (size_t)V, 0.f)));
    UPCHK(c->mass.upload(L.mass));
    UPCHK(c->meta.upload(L.meta));
    UPCHK(c->perm_d.upload(c->perm));
    UPCHK(c->inv_d.upload(c->inv));
    UPCHK(c->ell_e.upload(L.ell_e));
    UPCHK(c->ell_bq.upload(L.ell_bq));
    UPCHK(c->ell_bi.upload(L.ell_bi));
    UPCHK(c->e_slots.upload(L.e_slots));
    UPCHK(c->rest.upload(L.rest));
    UPCHK(c->b_slots.upload(L.b_slots));
    UPCHK(c->Qd.upload(L.Q));
    UPCHK(c->scratch3.alloc(3 * (size_t)V));
    {
        UPCHK(c->colours_d.upload(L.colours));
        UPCHK(c->barrier.alloc(2));
    }
    FrameParams fp0;
    std::memset(&fp0, 0, sizeof(fp0));
    UPCHK(c->fp.upload(std::vector<FrameParams>{fp0}));
    UPCHK(hipEventCreate(&c->ev0));
    UPCHK(hipEventCreate(&c->ev1));
    UPCHK(hipHostMalloc((void**)&c->host_flag, 64, hipHostMallocMapped));
    *c->host_flag = 0;
    UPCHK(hipHostGetDevicePointer((void**)&c->host_flag_d, c->host_flag, 0));
    UPCHK(hipHostMalloc((void**)&c->host_energy, lats_vbd::kHostEnergies * sizeof(float),
                        hipHostMallocMapped | hipHostMallocCoherent));
    UPCHK(hipHostGetDevicePointer((void**)&c->host_energy_d, c->host_energy, 0));
#undef UPCHK
    c->energy_nblk = nblocks((uint64_t)V + E + B);
    {
        const int rc = scratch_setup(c->device);
        if (rc) return cleanup_fail(rc);
    }
    *out = c;
    return LATS_OK;
}

int lats_vbd_destroy(lats_vbd* c) {
    if (!c) return LATS_OK;
    (void)hipStreamSynchronize(c->stream);
    delete c;
    return LATS_OK;
}

int lats_vbd_set_strategy(lats_vbd* c, int strategy) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (strategy < LATS_STRATEGY_AUTO || strategy > LATS_STRATEGY_PATCH)
        return fail(LATS_ERR_INVALID_ARGUMENT, "unknown strategy");
    if ((strategy == LATS_STRATEGY_PERSISTENT || strategy == LATS_STRATEGY_FLOW) && c->prm.num_substep != 1)
        return fail(LATS_ERR_UNSUPPORTED, "persistent and flow strategies run one substep per frame");
    if (strategy != c->strategy) c->drop_graphs();   // captured frames hold the strategy's kernels
    c->strategy = strategy;
    return LATS_OK;
}

int lats_vbd_set_patches(lats_vbd* c, uint32_t patches, uint32_t colours_per_launch, uint32_t threads) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (patches % kXcds) return fail(LATS_ERR_INVALID_ARGUMENT, "patches must be a multiple of 8 (0: derived)");
    if (patches > c->V) return fail(LATS_ERR_INVALID_ARGUMENT, "more patches than vertices");
    if (threads != 256 && threads != 512 && threads != 1024)
        return fail(LATS_ERR_INVALID_ARGUMENT, "patch workgroups must have 256, 512 or 1024 threads");
    if (patches != c->patch_target || colours_per_launch != c->patch_colours || threads != c->patch_threads)
        c->drop_graphs();
    c->patch_target = patches;
    c->patch_colours = colours_per_launch;
    c->patch_threads = threads;
    return LATS_OK;
}

int lats_vbd_get_patches(lats_vbd* c, lats_patch_info* info) {
    if (!c || !info) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    int rc = patch_prepare(c);
    if (rc) return rc;
    const PatchStats& st = c->patch_stats;
    info->patches = st.patches;
    info->colours_per_launch = st.colours_per_launch;
    info->launches_per_sweep = st.launches_per_sweep;
    info->threads = patch_block(c);
    info->lanes = patch_lanes(c);
    info->max_region = st.max_region;
    info->updates_per_sweep = st.updates_per_sweep;
    info->region_slots_per_sweep = st.region_slots;
    info->ell_entries_per_sweep = st.ell_entries;
    info->build_ms = st.build_ms;
    return LATS_OK;
}

int lats_vbd_set_solver(lats_vbd* c, int solver) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (solver != LATS_SOLVER_VBD && solver != LATS_SOLVER_XPBD) return fail(LATS_ERR_INVALID_ARGUMENT, "unknown solver");
    if (solver == LATS_SOLVER_XPBD) {
        if (c->has_obstacle) return fail(LATS_ERR_UNSUPPORTED, "the XPBD solver has no obstacle contact");
        const int rc = build_xpbd(c);
        if (rc) return rc;
    }
    if (solver != c->solver) {
        HIPCHK(hipStreamSynchronize(c->stream));
        c->solver = solver;
        c->drop_graphs();
    }
    return LATS_OK;
}

int lats_vbd_get_solver(lats_vbd* c, int* solver) {
    if (!c || !solver) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *solver = c->solver;
    return LATS_OK;
}

int lats_vbd_xpbd_clusters(lats_vbd* c, int which, uint32_t* csr, uint64_t cap, uint64_t* size) {
    if (!c || !size) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (which != 0 && which != 1) return fail(LATS_ERR_INVALID_ARGUMENT, "which must be 0 (springs) or 1 (bending)");
    const int rc = build_xpbd(c);
    if (rc) return rc;
    const std::vector<uint32_t>& v = which ? c->xb_csr : c->xs_csr;
    *size = v.size();
    if (csr) std::memcpy(csr, v.data(), std::min<uint64_t>(cap, v.size()) * sizeof(uint32_t));
    return LATS_OK;
}

int lats_vbd_set_obstacle(lats_vbd* c, const float* x, uint32_t nv, const int32_t* faces, uint32_t nf,
                          float thickness, float stiffness, float radius, float cell) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (nf && c->solver == LATS_SOLVER_XPBD) return fail(LATS_ERR_UNSUPPORTED, "the XPBD solver has no obstacle contact");
    if (nf == 0) {
        c->has_obstacle = false;
    } else {
        if (!x || !faces) return fail(LATS_ERR_INVALID_ARGUMENT, "null obstacle arrays");
        if (!(thickness > 0.f) || !(stiffness > 0.f) || !(radius > 0.f))
            return fail(LATS_ERR_INVALID_ARGUMENT, "thickness, stiffness and radius must be > 0");
        ObstacleHash h;
        std::string err;
        if (!build_obstacle_hash(x, nv, faces, nf, radius, cell, h, err)) return fail(LATS_ERR_INVALID_ARGUMENT, err);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(c->ob_x.upload(std::vector<float>(x, x + 3 * (size_t)nv)));
        HIPCHK(c->ob_f.upload(std::vector<int32_t>(faces, faces + 3 * (size_t)nf)));
        HIPCHK(c->ob_n.upload(face_normals(x, faces, nf)));
        HIPCHK(c->hb_belongs.upload(h.belongs));
        HIPCHK(c->hb_start.upload(h.start));
        HIPCHK(c->hb_count.upload(h.count));
        HIPCHK(c->hb_faces.upload(h.faces.empty() ? std::vector<uint32_t>{0u} : h.faces));
        HIPCHK(c->ct_count.upload(std::vector<uint32_t>(c->V, 0u)));
        HIPCHK(c->ct_fid.alloc((size_t)kMaxContacts * c->V));
        HIPCHK(c->ct_xs.alloc((size_t)kMaxContacts * c->V));
        HIPCHK(c->ct_n.alloc((size_t)kMaxContacts * c->V));
        c->ob_hash = std::move(h);
        c->thickness = thickness;
        c->ob_stiffness = stiffness;
        c->ob_radius = radius;
        c->has_obstacle = true;
    }
    c->drop_graphs();
    return LATS_OK;
}

int lats_vbd_get_contacts(lats_vbd* c, uint32_t* counts, int32_t* fids) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    std::vector<uint32_t> hc(c->V, 0u), hf((size_t)kMaxContacts * c->V, 0u);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->has_obstacle) {
        HIPCHK(hipMemcpy(hc.data(), c->ct_count.p, hc.size() * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hf.data(), c->ct_fid.p, hf.size() * 4, hipMemcpyDeviceToHost));
    }
    for (uint32_t s = 0; s < c->V; s++) {
        const uint32_t vid = c->perm[s];
        if (counts) counts[vid] = hc[s];
        if (fids)
            for (int k = 0; k < kMaxContacts; k++)
                fids[(size_t)kMaxContacts * vid + k] = (uint32_t)k < hc[s] ? (int32_t)hf[(size_t)k * c->V + s] : -1;
    }
    return LATS_OK;
}

int lats_vbd_set_lanes_per_vertex(lats_vbd* c, uint32_t lanes) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (lanes != 1 && lanes != 4 && lanes != 8 && lanes != 16)
        return fail(LATS_ERR_INVALID_ARGUMENT, "lanes per vertex must be 1, 4, 8 or 16");
    if (lanes != c->lanes) {
        c->lanes = lanes;
        c->drop_graphs();
    }
    return LATS_OK;
}

int lats_vbd_set_regions(lats_vbd* c, uint32_t regions) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (regions < 1 || regions > kMaxRegions)
        return fail(LATS_ERR_INVALID_ARGUMENT, "regions must be between 1 and 8");
    if (regions != c->regions) {
        split_regions(c, regions);
        c->drop_graphs();
    }
    return LATS_OK;
}

int lats_vbd_set_colour_block(lats_vbd* c, uint32_t threads) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (threads != 64 && threads != 128 && threads != 256)
        return fail(LATS_ERR_INVALID_ARGUMENT, "colour-pass blocks must have 64, 128 or 256 threads");
    if (threads != c->colour_block) {
        c->colour_block = threads;
        c->drop_graphs();
    }
    return LATS_OK;
}

int lats_vbd_get_colour_block(lats_vbd* c, uint32_t* threads) {
    if (!c || !threads) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *threads = c->colour_block;
    return LATS_OK;
}

int lats_vbd_get_regions(lats_vbd* c, uint32_t* regions) {
    if (!c || !regions) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *regions = c->regions;
    return LATS_OK;
}

int lats_vbd_get_lanes_per_vertex(lats_vbd* c, uint32_t* lanes) {
    if (!c || !lanes) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *lanes = c->lanes;
    return LATS_OK;
}

int lats_vbd_set_persistent_blocks(lats_vbd* c, uint32_t blocks) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    c->persist_blocks = blocks;
    return LATS_OK;
}

int lats_vbd_get_strategy(lats_vbd* c, int* strategy, uint32_t* blocks) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (strategy) *strategy = c->strategy;
    if (blocks && c->strategy == LATS_STRATEGY_PATCH) {
        int rc = patch_prepare(c);
        if (rc) return rc;
        *blocks = c->patch_built_p;
    } else if (blocks) {
        uint32_t g = 0;
        int rc = persistent_grid(c, &g);
        if (rc) return rc;
        *blocks = g;
    }
    return LATS_OK;
}

// Times `frames` frames with the graph strategy and with the persistent one
// at several grid sizes, keeps the fastest, restores the state.
int lats_vbd_autotune(lats_vbd* c, uint32_t frames, int* chosen, uint32_t* chosen_blocks) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (c->solver != LATS_SOLVER_VBD) return fail(LATS_ERR_UNSUPPORTED, "the autotuner tunes the VBD colour passes");
    if (!frames) frames = 1;
    const size_t nf = 3 * (size_t)c->V;
    DBuf<float> sx, sx0, sv;
    HIPCHK(sx.alloc(nf));
    HIPCHK(sx0.alloc(nf));
    HIPCHK(sv.alloc(nf));
    FrameParams fps;
    // Device-to-device copies go on the context's stream: a synchronous
    // hipMemcpy between device buffers may return before the copy is done
    // and is not ordered with the (non-blocking) stream's later kernels.
    HIPCHK(hipMemcpyAsync(sx.p, c->x.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(sx0.p, c->x0.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(sv.p, c->v.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(&fps, c->fp.p, sizeof(fps), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    auto restore = [&]() -> int {
        HIPCHK(hipMemcpyAsync(c->x.p, sx.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->x0.p, sx0.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->v.p, sv.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->fp.p, &fps, sizeof(fps), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return LATS_OK;
    };
    struct Cand { int strategy; uint32_t blocks; uint32_t lanes; uint32_t regions = 1; uint32_t threads = 256; };
    std::vector<Cand> cands;
    const uint32_t keep_lanes = c->lanes, keep_regions = c->regions, keep_threads = c->colour_block;
    for (uint32_t lanes : {1u, 4u, 8u, 16u}) {
        for (uint32_t threads : {64u, 128u, 256u}) {
            cands.push_back({LATS_STRATEGY_GRAPH, 0, lanes, 1, threads});
            cands.push_back({LATS_STRATEGY_GRAPH, 0, lanes, kMaxRegions, threads});
        }
        if (c->prm.num_substep != 1) continue;
        uint32_t g0 = 0;
        const uint32_t keep = c->persist_blocks;
        c->persist_blocks = 0;
        c->lanes = lanes;
        int rc = persistent_grid(c, &g0);
        c->persist_blocks = keep;
        if (rc) { c->lanes = keep_lanes; return rc; }
        for (uint32_t g : {g0 / 2, g0, 2 * g0, c->persist_max_blocks})
            if (g >= 1 && g <= c->persist_max_blocks) cands.push_back({LATS_STRATEGY_PERSISTENT, g, lanes});
        if (lanes == 1 || c->has_obstacle) continue;
        uint32_t fcap = 0;
        c->lanes = lanes;
        rc = flow_capacity(c, &fcap);
        if (rc) { c->lanes = keep_lanes; return rc; }
        for (uint32_t tiles : {128u, 256u, 512u, 1024u})
            if (tiles <= fcap && tiles * 32u <= c->V) cands.push_back({LATS_STRATEGY_FLOW, tiles, lanes});
    }
    c->lanes = keep_lanes;
    // patch strategy (several colour passes per launch): measured to pay on
    // cloths of up to ~10^4-10^5 vertices, where a colour pass is mostly
    // launch overhead; candidates {patches, colours per launch (0 = all),
    // threads}, 8 lanes per vertex
    const uint32_t keep_pt = c->patch_target, keep_pf = c->patch_colours, keep_pb = c->patch_threads;
    if (c->V >= 64 && c->V <= (1u << 18)) {
        auto round8 = [](uint32_t n) { return std::max(8u, std::min(4096u, n / 8u * 8u)); };
        const uint32_t pa = round8(c->V / 256u), pb = round8(c->V / 64u);
        // (and one patch per CU, 256, for cloths too small for that at 64
        // vertices per patch: more halo, every CU busy)
        const uint32_t pc = round8(std::min(256u, c->V / 16u));
        for (const Cand& k : {Cand{LATS_STRATEGY_PATCH, pa, 8, 0, 1024}, Cand{LATS_STRATEGY_PATCH, pa, 8, 0, 512},
                              Cand{LATS_STRATEGY_PATCH, pb, 8, 0, 256}, Cand{LATS_STRATEGY_PATCH, pb, 8, 0, 512},
                              Cand{LATS_STRATEGY_PATCH, pa, 8, 4, 1024}, Cand{LATS_STRATEGY_PATCH, pa, 8, 4, 512}})
            if (k.blocks <= c->V) cands.push_back(k);
        if (pc != pa && pc != pb && pc <= c->V) cands.push_back(Cand{LATS_STRATEGY_PATCH, pc, 8, 0, 512});
    }
    float best = 3.4e38f;
    Cand win = cands[0];
    const int keep_strategy = c->strategy;
    const uint32_t keep_blocks = c->persist_blocks;
    auto use = [&](uint32_t lanes, uint32_t regions, uint32_t threads) {
        if (c->lanes != lanes || c->regions != regions || c->colour_block != threads) {
            c->lanes = lanes;
            c->colour_block = threads;
            if (c->regions != regions) split_regions(c, regions);
            c->drop_graphs();
        }
    };
    // a candidate's settings (patch candidates carry patches / colours per
    // launch / threads in blocks / regions / threads)
    auto apply = [&](const Cand& k) {
        if (c->strategy != k.strategy) c->drop_graphs();
        c->strategy = k.strategy;
        if (k.strategy == LATS_STRATEGY_PATCH) {
            if (c->lanes != k.lanes || c->patch_target != k.blocks || c->patch_colours != k.regions ||
                c->patch_threads != k.threads)
                c->drop_graphs();
            c->lanes = k.lanes;
            c->patch_target = k.blocks;
            c->patch_colours = k.regions;
            c->patch_threads = k.threads;
        } else {
            c->persist_blocks = k.blocks;
            use(k.lanes, k.regions, k.threads);
        }
    };
    auto restore_settings = [&]() {
        if (c->strategy != keep_strategy) c->drop_graphs();
        c->strategy = keep_strategy;
        c->persist_blocks = keep_blocks;
        c->patch_target = keep_pt;
        c->patch_colours = keep_pf;
        c->patch_threads = keep_pb;
        use(keep_lanes, keep_regions, keep_threads);
    };
    // Two passes over the candidates, each candidate's best time counts:
    // the first timings after a context is created can be several per cent
    // off (clocks and caches settling), enough to swap candidates that close.
    std::vector<float> best_ms(cands.size(), 3.4e38f);
    for (int pass = 0; pass < 2; pass++) {
        for (size_t i = 0; i < cands.size(); i++) {
            const Cand& k = cands[i];
            apply(k);
            c->cooperative_refused = false;
            int rc = lats_vbd_run_frames(c, 1, nullptr);   // warm-up / capture
            if (!rc) rc = lats_vbd_run_frames(c, frames, nullptr);
            if (rc && k.strategy == LATS_STRATEGY_PATCH && rc == LATS_ERR_UNSUPPORTED) {
                // (a layout that does not fit LDS is no candidate)
                rc = restore();
                if (rc) return rc;
                continue;
            }
            if (rc == LATS_ERR_HIP && (k.strategy == LATS_STRATEGY_PERSISTENT || k.strategy == LATS_STRATEGY_FLOW) &&
                c->cooperative_refused) {
                // the runtime refused the cooperative launch itself (nothing
                // was enqueued; e.g. the device is shared with another
                // process): no candidate here
                c->cooperative_refused = false;
                rc = restore();
                if (rc) return rc;
                continue;
            }
            if (rc) {
                restore_settings();
                restore();
                return rc;
            }
            best_ms[i] = std::min(best_ms[i], c->last_ms);
            rc = restore();
            if (rc) return rc;
        }
    }
    c->tune_results.clear();
    for (size_t i = 0; i < cands.size(); i++) {
        if (best_ms[i] < best) {
            best = best_ms[i];
            win = cands[i];
        }
        const Cand& k = cands[i];
        for (float v : {(float)k.strategy, (float)k.blocks, (float)k.lanes, (float)k.regions, best_ms[i],
                        (float)k.threads})
            c->tune_results.push_back(v);
    }
    restore_settings();
    apply(win);
    if (chosen) *chosen = win.strategy;
    if (chosen_blocks) *chosen_blocks = win.blocks;
    return LATS_OK;
}

int lats_vbd_autotune_results(lats_vbd* c, float* out, uint32_t cap, uint32_t* count) {
    if (!c || !count) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    const uint32_t n = (uint32_t)(c->tune_results.size() / 6);
    *count = n;
    if (out) std::memcpy(out, c->tune_results.data(), std::min(cap, n) * 6 * sizeof(float));
    return LATS_OK;
}

int lats_vbd_set_energy_bending(lats_vbd* c, int on) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if ((on != 0) != (c->energy_bending != 0)) {
        c->energy_bending = on != 0;
        c->drop_graphs();
    }
    return LATS_OK;
}

int lats_vbd_set_state(lats_vbd* c, const float* x, const float* v) {
    if (!c || !x || !v) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    const size_t n = 3 * (size_t)c->V;
    std::vector<float> hx(n), hv(n);
    for (uint32_t s = 0; s < c->V; s++)
        for (int k = 0; k < 3; k++) {
            hx[3 * s + k] = x[3 * (size_t)c->perm[s] + k];
            hv[3 * s + k] = v[3 * (size_t)c->perm[s] + k];
        }
    // stream-ordered copies, complete before hx / hv go out of scope
    HIPCHK(hipMemcpyAsync(c->x.p, hx.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->x0.p, hx.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->v.p, hv.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LATS_OK;
}

int lats_vbd_get_state(lats_vbd* c, float* x, float* v) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    const size_t n = 3 * (size_t)c->V;
    std::vector<float> hx(n), hv(n);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(hx.data(), c->x.p, n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hv.data(), c->v.p, n * 4, hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < c->V; s++)
        for (int k = 0; k < 3; k++) {
            if (x) x[3 * (size_t)c->perm[s] + k] = hx[3 * s + k];
            if (v) v[3 * (size_t)c->perm[s] + k] = hv[3 * s + k];
        }
    return LATS_OK;
}

int lats_vbd_set_external_force(lats_vbd* c, const float f[3]) {
    if (!c || !f) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    return set_frame_params(c, nullptr, 0, f);
}

int lats_vbd_physics_step(lats_vbd* c, int record) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    int rc = run_frame(c, record != 0);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    return flow_check(c);
}

int lats_vbd_get_energy(lats_vbd* c, float* out, uint32_t cap, uint32_t* count) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t n = std::min(cap, c->energy_rows);
    if (count) *count = c->energy_rows;
    if (n && out) HIPCHK(hipMemcpy(out, c->energy.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return LATS_OK;
}

int lats_vbd_run_frames(lats_vbd* c, uint32_t nframes, const float* forces) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (forces) {
        if (c->forces.n < 3 * (size_t)nframes) HIPCHK(c->forces.alloc(3 * (size_t)nframes));
        HIPCHK(hipMemcpyAsync(c->forces.p, forces, 3 * (size_t)nframes * sizeof(float), hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int rc = set_frame_params(c, c->forces.p, 0, nullptr);
        if (rc) return rc;
    }
    int rc;
    const bool vbd = c->solver == LATS_SOLVER_VBD;
    if (vbd && c->strategy == LATS_STRATEGY_PERSISTENT) {
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        rc = run_persistent(c, nframes);
        if (rc) return rc;
    } else if (vbd && c->strategy == LATS_STRATEGY_FLOW) {
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        rc = run_flow(c, nframes);
        if (rc) return rc;
    } else {
        hipGraphExec_t g;
        rc = get_graph(c, false, &g);
        if (rc) return rc;
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        for (uint32_t f = 0; f < nframes; f++) HIPCHK(hipGraphLaunch(g, c->stream));
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    rc = flow_check(c);
    if (rc) return rc;
    if (forces) {
        rc = set_frame_params(c, nullptr, 0, forces + 3 * (size_t)(nframes ? nframes - 1 : 0));
        if (rc) return rc;
    }
    return LATS_OK;
}

int lats_vbd_last_elapsed_ms(lats_vbd* c, float* ms) {
    if (!c || !ms) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *ms = c->last_ms;
    return LATS_OK;
}

int lats_vbd_synchronize(lats_vbd* c) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return LATS_OK;
}

// ---- scheduler slot ---------------------------------------------------------
static int ensure_async(lats_vbd* c, uint32_t idx) {
    const uint32_t k = idx % LATS_NUM_ASYNC_BUFFERS;
    if (!c->iterbuf[k].p) {
        HIPCHK(c->iterbuf[k].alloc(3 * (size_t)c->V));
        HIPCHK(c->beginbuf[k].alloc(3 * (size_t)c->V));
    }
    return LATS_OK;
}
// A task the asynchronous pass left without a buffer (-1u) iterates sa_x.
static float* iter_buffer(lats_vbd* c, uint32_t idx) {
    return (idx == LATS_BUFFER_DEFAULT || idx == LATS_BUFFER_NONE) ? c->x.p
                                                                   : c->iterbuf[idx % LATS_NUM_ASYNC_BUFFERS].p;
}
static float* begin_buffer(lats_vbd* c, uint32_t idx) { return c->beginbuf[idx % LATS_NUM_ASYNC_BUFFERS].p; }

static int dispatch_energy(lats_vbd* c, const float* xb) {
    int rc = ensure_energy_rows(c, c->energy_rows + 1);
    if (rc) return rc;
    enqueue_energy(c, xb, c->energy_rows);
    hipLaunchKernelGGL(k_energy_finalize, dim3(1), dim3(kBlock), 0, c->stream, c->partial.p, c->energy_nblk,
                       c->energy_rows, c->energy.p);
    c->energy_rows++;
    HIPCHK(hipGetLastError());
    return LATS_OK;
}

int lats_vbd_set_record_energy(lats_vbd* c, int on) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    c->record_dispatch = on != 0;
    if (on) c->energy_rows = 0;
    return LATS_OK;
}

int lats_vbd_dispatch(lats_vbd* c, const lats_launch_param* p) {
    if (!c || !p) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    const uint32_t nf = 3 * c->V;
    auto need = [&](uint32_t idx) -> int {
        return (idx == LATS_BUFFER_DEFAULT || idx == LATS_BUFFER_NONE) ? LATS_OK : ensure_async(c, idx);
    };
    auto copy_to_start_and_iter = [&](const float* src, uint32_t out_idx) -> int {
        int rc = need(out_idx);
        if (rc) return rc;
        hipLaunchKernelGGL(k_copy_to_2, dim3(nblocks(nf)), dim3(kBlock), 0, c->stream, src,
                           out_idx == LATS_BUFFER_DEFAULT ? nullptr : begin_buffer(c, out_idx),
                           iter_buffer(c, out_idx), nf);
        return LATS_OK;
    };
    // GpuSolver::fn_dispatch prev hook (example3_hybrid_cpu_gpu.cpp:1769-1825)
    auto prev = [&]() -> int {
        int rc = need(p->buffer_idx);
        if (rc) return rc;
        if (p->num_input_buffers && p->left_buffer_idx != LATS_BUFFER_NONE) {
            rc = need(p->left_buffer_idx);
            if (rc) return rc;
            for (uint32_t k = 0; k < p->num_input_buffers; k++) {
                const uint32_t in = p->input_buffer_idxs[k];
                rc = need(in);
                if (rc) return rc;
                float* bb = p->is_allocated_to_main_device ? begin_buffer(c, in) : begin_buffer(c, p->left_buffer_idx);
                hipLaunchKernelGGL(k_read_and_solve_conflict, dim3(nblocks(c->V)), dim3(kBlock), 0, c->stream, bb,
                                   iter_buffer(c, in), iter_buffer(c, p->left_buffer_idx), 0.5f, c->V);
            }
        } else if (p->num_input_buffers) {
            const uint32_t in = p->input_buffer_idxs[p->num_input_buffers - 1];
            rc = need(in);
            if (rc) return rc;
            rc = copy_to_start_and_iter(iter_buffer(c, in), p->buffer_idx);
            if (rc) return rc;
        } else if (p->left_buffer_idx == LATS_BUFFER_DEFAULT) {
            rc = copy_to_start_and_iter(c->x.p, p->buffer_idx);
            if (rc) return rc;
        }
        if (c->record_dispatch && p->iter_idx == 0 && p->cluster_idx == 0) return dispatch_energy(c, iter_buffer(c, p->buffer_idx));
        return LATS_OK;
    };
    // post hook (:1827-1855)
    auto post = [&]() -> int {
        if (p->right_buffer_idx != LATS_BUFFER_NONE) {
            int rc = copy_to_start_and_iter(iter_buffer(c, p->buffer_idx), p->right_buffer_idx);
            if (rc) return rc;
        }
        if (c->record_dispatch &&
            ((p->function_id == LATS_FN_VBD_ALL_IN_ONE && p->cluster_idx + 1 == c->C) ||
             p->function_id == LATS_FN_CONSTRAINT_LAST_NODE))
            return dispatch_energy(c, iter_buffer(c, p->buffer_idx));
        return LATS_OK;
    };
    int rc = LATS_OK;
    auto xpbd_task = [&](bool stretch) -> int {
        // XPBD tasks iterate sa_x (the reference's XPBD path has no asynchronous buffers)
        if (p->buffer_idx != LATS_BUFFER_DEFAULT && p->buffer_idx != LATS_BUFFER_NONE)
            return fail(LATS_ERR_UNSUPPORTED, "XPBD tasks iterate sa_x only");
        int r = build_xpbd(c);
        if (r) return r;
        const size_t nclusters = (stretch ? c->xs_bound.size() : c->xb_bound.size()) - 1;
        if (p->cluster_idx >= nclusters) return fail(LATS_ERR_INVALID_ARGUMENT, "cluster_idx out of range");
        if (stretch) enqueue_xpbd_stretch(c, p->cluster_idx);
        else enqueue_xpbd_bending(c, p->cluster_idx);
        return LATS_OK;
    };
    switch (p->function_id) {
        case LATS_FN_ADDITIONAL_ROOT:
        case LATS_FN_ADDITIONAL_TERMINAL:
            break;
        case LATS_FN_RESET_CONSTRAINS:
            rc = build_xpbd(c);
            if (rc) return rc;
            enqueue_xpbd_reset(c);
            break;
        case LATS_FN_XPBD_STRETCH:
            rc = xpbd_task(true);
            break;
        case LATS_FN_XPBD_BENDING:
            rc = xpbd_task(false);
            break;
        case LATS_FN_PREDICT_POSITION:
            enqueue_predict(c);
            enqueue_detect(c);     // collision_detection() follows predict (example2_async_iter.cpp:1104-1106)
            break;
        case LATS_FN_UPDATE_VELOCITY:
            enqueue_update(c, 1);
            break;
        case LATS_FN_VBD_ALL_IN_ONE: {
            if (p->cluster_idx >= c->C) return fail(LATS_ERR_INVALID_ARGUMENT, "cluster_idx out of range");
            rc = prev();
            if (rc) return rc;
            enqueue_colour(c, iter_buffer(c, p->buffer_idx), p->cluster_idx);
            rc = post();
            break;
        }
        case LATS_FN_CONSTRAINT_LAST_NODE: {
            rc = prev();
            if (rc) return rc;
            rc = post();
            if (rc) return rc;
            float* src = iter_buffer(c, p->buffer_idx);
            if (src != c->x.p)
                hipLaunchKernelGGL(k_copy_to_2, dim3(nblocks(nf)), dim3(kBlock), 0, c->stream, src, c->x.p,
                                   (float*)nullptr, nf);
            break;
        }
        default:
            return fail(LATS_ERR_INVALID_ARGUMENT, "unknown function id");
    }
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return LATS_OK;
}

// ---- host <-> device signalling ------------------------------------------
int lats_vbd_stream_wait_host(lats_vbd* c, uint32_t value) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    HIPCHK(hipStreamWaitValue32(c->stream, c->host_flag_d, value, hipStreamWaitValueGte, 0xffffffffu));
    return LATS_OK;
}
int lats_vbd_host_signal(lats_vbd* c, uint32_t value) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    __atomic_store_n(c->host_flag, value, __ATOMIC_SEQ_CST);
    return LATS_OK;
}
int lats_vbd_record_marker(lats_vbd* c, uint32_t marker) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    auto it = c->markers.find(marker);
    if (it == c->markers.end()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        it = c->markers.emplace(marker, e).first;
    }
    HIPCHK(hipEventRecord(it->second, c->stream));
    return LATS_OK;
}
int lats_vbd_wait_marker(lats_vbd* c, uint32_t marker) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    auto it = c->markers.find(marker);
    if (it == c->markers.end()) return fail(LATS_ERR_INVALID_ARGUMENT, "marker never recorded");
    HIPCHK(hipEventSynchronize(it->second));
    return LATS_OK;
}
int lats_vbd_query_marker(lats_vbd* c, uint32_t marker, int* done) {
    if (!c || !done) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    auto it = c->markers.find(marker);
    if (it == c->markers.end()) return fail(LATS_ERR_INVALID_ARGUMENT, "marker never recorded");
    const hipError_t e = hipEventQuery(it->second);
    if (e == hipErrorNotReady) {
        *done = 0;
        return LATS_OK;
    }
    HIPCHK(e);
    *done = 1;
    return LATS_OK;
}
int lats_vbd_marker_elapsed_ms(lats_vbd* c, uint32_t m0, uint32_t m1, float* ms) {
    if (!c || !ms) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    auto a = c->markers.find(m0), b = c->markers.find(m1);
    if (a == c->markers.end() || b == c->markers.end()) return fail(LATS_ERR_INVALID_ARGUMENT, "marker never recorded");
    HIPCHK(hipEventElapsedTime(ms, a->second, b->second));
    return LATS_OK;
}

// Position buffers of the asynchronous frame as the host sees them (vertex
// order); copies run on the context's stream through scratch3.
static int buffer_of(lats_vbd* c, int which, uint32_t idx, float** out) {
    switch (which) {
        case LATS_BUF_ITER:
        case LATS_BUF_BEGIN: {
            if (idx == LATS_BUFFER_DEFAULT || idx == LATS_BUFFER_NONE) {
                if (which == LATS_BUF_BEGIN) return fail(LATS_ERR_INVALID_ARGUMENT, "the default buffer has no begin copy");
                *out = c->x.p;
                return LATS_OK;
            }
            const int rc = ensure_async(c, idx);
            if (rc) return rc;
            *out = which == LATS_BUF_ITER ? iter_buffer(c, idx) : begin_buffer(c, idx);
            return LATS_OK;
        }
        case LATS_BUF_X: *out = c->x.p; return LATS_OK;
        case LATS_BUF_X_START: *out = c->x0.p; return LATS_OK;
        case LATS_BUF_V: *out = c->v.p; return LATS_OK;
        default: return fail(LATS_ERR_INVALID_ARGUMENT, "unknown buffer kind");
    }
}

int lats_vbd_buffer_download(lats_vbd* c, int which, uint32_t idx, float* host) {
    if (!c || !host) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    float* src = nullptr;
    int rc = buffer_of(c, which, idx, &src);
    if (rc) return rc;
    // vertex v sits in slot inv[v]
    hipLaunchKernelGGL(k_gather_rows, dim3(nblocks(c->V)), dim3(kBlock), 0, c->stream, src, c->inv_d.p, c->scratch3.p,
                       c->V);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, c->scratch3.p, 3 * (size_t)c->V * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return LATS_OK;
}

int lats_vbd_buffer_upload(lats_vbd* c, int which, uint32_t idx, const float* host) {
    if (!c || !host) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    float* dst = nullptr;
    int rc = buffer_of(c, which, idx, &dst);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->scratch3.p, host, 3 * (size_t)c->V * sizeof(float), hipMemcpyHostToDevice, c->stream));
    // slot s holds vertex perm[s]
    hipLaunchKernelGGL(k_gather_rows, dim3(nblocks(c->V)), dim3(kBlock), 0, c->stream, c->scratch3.p, c->perm_d.p, dst,
                       c->V);
    HIPCHK(hipGetLastError());
    return LATS_OK;
}

int lats_host_alloc(size_t bytes, void** out) {
    if (!out) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    HIPCHK(hipHostMalloc(out, bytes, hipHostMallocDefault));
    return LATS_OK;
}

int lats_host_free(void* p) {
    if (p) HIPCHK(hipHostFree(p));
    return LATS_OK;
}

int lats_vbd_device_positions(lats_vbd* c, void** d) {
    if (!c || !d) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *d = c->x.p;
    return LATS_OK;
}
int lats_vbd_get_stream(lats_vbd* c, void** stream) {
    if (!c || !stream) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *stream = static_cast<void*>(c->stream);
    return LATS_OK;
}
int lats_vbd_copy_permutation(lats_vbd* c, uint32_t* perm) {
    if (!c || !perm) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    std::copy(c->perm.begin(), c->perm.end(), perm);
    return LATS_OK;
}
// ---- frame pieces (the reference's physics_step_vbd split in its stages)
int lats_vbd_frame_begin(lats_vbd* c) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    enqueue_predict(c);
    enqueue_detect(c);
    HIPCHK(hipGetLastError());
    return LATS_OK;
}

int lats_vbd_iterate(lats_vbd* c, uint32_t sweeps) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    if (c->solver != LATS_SOLVER_VBD) return fail(LATS_ERR_UNSUPPORTED, "colour sweeps belong to the VBD solver");
    if (sweeps == 0) return LATS_OK;
    if (uses_patches(c)) {
        int rc = patch_prepare(c);
        if (rc) return rc;
    }
    // one graph per distinct sweep count (a handful in practice)
    auto it = c->sweep_graphs.find(sweeps);
    if (it == c->sweep_graphs.end()) {
        if (c->sweep_graphs.size() >= 16) c->drop_graphs();
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        if (uses_patches(c))
            enqueue_patch_sweeps(c, sweeps, false);
        else
            for (uint32_t k = 0; k < sweeps; k++)
                for (uint32_t col = 0; col < c->C; col++) enqueue_colour(c, c->x.p, col);
        HIPCHK(hipStreamEndCapture(c->stream, &graph));
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(LATS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        it = c->sweep_graphs.emplace(sweeps, exec).first;
    }
    HIPCHK(hipGraphLaunch(it->second, c->stream));
    return LATS_OK;
}

int lats_vbd_frame_end(lats_vbd* c) {
    if (!c) return fail(LATS_ERR_INVALID_ARGUMENT, "null context");
    enqueue_update(c, 1);
    HIPCHK(hipGetLastError());
    return LATS_OK;
}

int lats_vbd_energy_now(lats_vbd* c, float* e) {
    if (!c || !e) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    const float* xs[1] = {c->x.p};
    return lats::energies_of(c, xs, 1, e);
}

// lats_vbd_energy_now's kernels with their block partials in `scratch` and the
// result left in device memory: no allocation, no host wait.
int lats_vbd_energy_noalloc(lats_vbd* c, const float* x_dev, float* energy_dev) {
    if (!c || !energy_dev) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (c->energy_nblk > kScratchPartials)
        return fail(LATS_ERR_INVALID_ARGUMENT,
                    "cloth too large for the no-allocation energy: vertices + springs + bending pairs = " +
                        std::to_string((uint64_t)c->V + c->E + c->B) + " > " +
                        std::to_string((uint64_t)kScratchPartials * kBlock));
    double* part = scratch_address(c->device);
    if (!part) return fail(LATS_ERR_HIP, "scratch address of the context's device was not looked up");
    enqueue_energy_partials(c, x_dev ? x_dev : c->x.p, part, c->stream);
    hipLaunchKernelGGL(k_energy_finalize, dim3(1), dim3(kBlock), 0, c->stream, part, c->energy_nblk, 0u, energy_dev);
    HIPCHK(hipGetLastError());
    return LATS_OK;
}

int lats_vbd_time_colour_kernels(lats_vbd* c, uint32_t iters, float* mean_ms, uint32_t* launches,
                                 uint64_t* vertex_updates) {
    if (!c || !mean_ms) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (iters == 0) return fail(LATS_ERR_INVALID_ARGUMENT, "iterations must be positive");
    const size_t nf = 3 * (size_t)c->V;
    const uint32_t nl = iters * c->C;
    uint64_t vu = 0;
    for (uint32_t col = 0; col < c->C; col++) vu += (uint64_t)c->ccount[col] * iters;
    // the sweeps as one graph: launches run back to back as in a frame graph
    hipGraph_t graph;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t it = 0; it < iters; it++)
        for (uint32_t col = 0; col < c->C; col++) enqueue_colour(c, c->x.p, col);
    HIPCHK(hipStreamEndCapture(c->stream, &graph));
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(LATS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    float ms = 0.f;
    HIPCHK(hipMemcpyAsync(c->scratch3.p, c->x.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
    e = hipGraphLaunch(exec, c->stream);   // warm-up replay
    if (e == hipSuccess) e = hipMemcpyAsync(c->x.p, c->scratch3.p, nf * 4, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    if (e == hipSuccess) e = hipGraphLaunch(exec, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    (void)hipStreamSynchronize(c->stream);   // also after a failed step: nothing may still run the graph
    (void)hipGraphExecDestroy(exec);
    if (e != hipSuccess) return fail(LATS_ERR_HIP, hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(c->x.p, c->scratch3.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *mean_ms = ms / (float)nl;
    if (launches) *launches = nl;
    if (vertex_updates) *vertex_updates = vu;
    return LATS_OK;
}

int lats_vbd_time_xpbd_kernels(lats_vbd* c, uint32_t iters, float* stretch_ms, float* bending_ms,
                               uint32_t* stretch_launches, uint32_t* bending_launches) {
    if (!c || !stretch_ms || !bending_ms) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (iters == 0) return fail(LATS_ERR_INVALID_ARGUMENT, "iterations must be positive");
    if (c->solver != LATS_SOLVER_XPBD || !c->xpbd_built)
        return fail(LATS_ERR_UNSUPPORTED, "select the XPBD solver first (lats_vbd_set_solver)");
    const size_t nf = 3 * (size_t)c->V;
    const uint32_t ns = (uint32_t)c->xs_bound.size() - 1, nb = (uint32_t)c->xb_bound.size() - 1;
    float* out[2] = {stretch_ms, bending_ms};
    uint32_t* count[2] = {stretch_launches, bending_launches};
    HIPCHK(hipMemcpyAsync(c->scratch3.p, c->x.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
    for (int which = 0; which < 2; which++) {
        const uint32_t n = which ? nb : ns;
        uint32_t nl = 0;
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t sz = which ? c->xb_bound[k + 1] - c->xb_bound[k] : c->xs_bound[k + 1] - c->xs_bound[k];
            nl += sz ? iters : 0u;
        }
        *out[which] = 0.f;
        if (count[which]) *count[which] = nl;
        if (!nl) continue;
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (uint32_t it = 0; it < iters; it++)
            for (uint32_t k = 0; k < n; k++) {
                if (which) enqueue_xpbd_bending(c, k);
                else enqueue_xpbd_stretch(c, k);
            }
        HIPCHK(hipStreamEndCapture(c->stream, &graph));
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(LATS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        float ms = 0.f;
        e = hipGraphLaunch(exec, c->stream);   // warm-up replay
        if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
        if (e == hipSuccess) e = hipGraphLaunch(exec, c->stream);
        if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
        (void)hipStreamSynchronize(c->stream);
        (void)hipGraphExecDestroy(exec);
        if (e != hipSuccess) return fail(LATS_ERR_HIP, hipGetErrorString(e));
        *out[which] = ms / (float)nl;
    }
    // positions back; the multipliers restart at zero with the next frame
    HIPCHK(hipMemcpyAsync(c->x.p, c->scratch3.p, nf * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LATS_OK;
}

int lats_vbd_device_bytes(lats_vbd* c, uint64_t* bytes) {
    if (!c || !bytes) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t b = 0;
    b += (c->x.n + c->x0.n + c->v.n + c->mass.n + c->rest.n + c->scratch3.n + c->forces.n + c->energy.n) * 4;
    b += (c->meta.n + c->perm_d.n + c->inv_d.n) * 4;
    b += (c->ell_e.n + c->e_slots.n) * 8 + (c->ell_bq.n + c->Qd.n + c->b_slots.n) * 16 + c->ell_bi.n * 12;
    b += c->partial.n * 8;
    for (int k = 0; k < LATS_NUM_ASYNC_BUFFERS; k++) b += (c->iterbuf[k].n + c->beginbuf[k].n) * 4;
    *bytes = b;
    return LATS_OK;
}

}  // extern "C"
