// vbd_patch.hip -- the patch strategy of the VBD colour loop (see
// vbd_patch.h): F colour passes per launch, each workgroup on one patch of
// the cloth held in LDS, halo updates recomputed locally.
//
// The per-vertex arithmetic is vbd_vertex_group's (group_load_ell /
// group_load_own / group_finish), so every update is bit-identical to the
// graph strategy's, which follows the reference's FP32 order
// (xpbd_constraints.h:4027-4205, example2_async_iter.cpp:939-1014).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/errors.h"
#include "vbd_context.h"
#include "vbd_kernels.h"
#include "vbd_patch.h"
#include "vbd_patch_build.h"

using namespace lats;

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

// Uniform tables (heads, steps) are read through the constant address space:
// scalar loads, which no workgroup barrier orders behind the vector loads in
// flight (a barrier's fences make the compiler treat later global loads as
// clobbered and issue them as vector loads, which wait on every load before).
template <class T>
using Uniform = const __attribute__((address_space(4))) T*;
__device__ __forceinline__ PatchHead head_at(Uniform<PatchHead> p, uint32_t j) {
    return PatchHead{p[j].roff, p[j].rsize, p[j].owned, p[j].nstate, p[j].nsteps, p[j].step_off, 0u, 0u};
}
__device__ __forceinline__ PatchStep step_at(Uniform<PatchStep> p, uint32_t j) {
    return PatchStep{p[j].off, p[j].n, p[j].eoff, p[j].boff, p[j].ew, p[j].bw};
}

// Orders one wave's LDS writes and reads of the staged terms: the lanes of
// a wave run in lockstep and a wave's LDS accesses complete in issue order,
// so a wavefront-scope fence (which keeps the compiler from moving the
// accesses across it) is all the exchange needs.
__device__ __forceinline__ void staged_exchange() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The pair's four points (pair order) of a local pair entry.
__device__ __forceinline__ void patch_pair_points(const PatchMem& mem, const PatchPair& e, F3& p0, F3& p1, F3& p2,
                                                  F3& p3) {
    p0 = mem.ldx(nullptr, e.i01 & 0xffffu);
    p1 = mem.ldx(nullptr, e.i01 >> 16);
    p2 = mem.ldx(nullptr, e.i23 & 0xffffu);
    p3 = mem.ldx(nullptr, e.i23 >> 16);
}

// The arithmetic of one patch update on its group of 8 lanes (lane = 0..7):
// vbd_vertex_group's terms -- VBD::compute_inertia, compute_stretch_mass_spring,
// compute_bending_quadratic (xpbd_constraints.h:4027-4205) -- summed as the
// reference sums them (example2_async_iter.cpp:939-992), then the Newton
// step of vbd_step (:993-1014) on lane 0.  Lane j computes the terms of
// incidences j, j + 8, ...; every lane writes its terms to the group's LDS
// words `scr`, then lane v adds value v of lanes 0, 1, ..., 7 (round after
// round) in that order -- the additions lane 0 makes in vbd_vertex_group, in
// the same order, so every sum is bit-identical -- and lane 0 collects the
// finished sums with DPP moves.  Spring values: lane v < 6 sums S's entry v
// (xx, xy, xz, yy, yz, zz), lanes 6 / 7 fs.x / fs.y, and lane 0 fs.z with
// DPP adds; pair values: lanes 0, 3, 5 sum hb (the diagonal they hold), 6 / 7
// fb.x / fb.y, lane 1 fb.z.  Lanes past a degree hold +0 terms, which leave
// a sum started at +0 unchanged.  A pair entry names the pair's four points
// (the vertex's own among them: the same LDS word as xk) and carries
// Q[o][o], so no lane selects points by the vertex's position in the pair.
// valid == false: the group has no vertex; it still takes part in the DPP
// moves and stores nothing.
__device__ __forceinline__ void patch_update(const VbdArgs& a, const GroupLoads<8>& L, F3 xk, uint32_t s,
                                             uint32_t sx, uint32_t i, uint32_t n, uint32_t eoff, uint32_t boff,
                                             F3 ext, uint32_t lane, bool valid, const PatchMem& mem, float* scr) {
    constexpr uint32_t G = kPatchLanes, PE = GroupLoads<8>::PE, PB = GroupLoads<8>::PB;
    static_assert(PE == 1 && PB == 2, "one spring round and two pair rounds are staged");
    const uint2* pe = L.pe;
    const float4* pq = L.pq;
    const PatchPair* pp = reinterpret_cast<const PatchPair*>(L.pi);
    const uint32_t meta = valid ? L.meta : 0u;
    const uint32_t edeg = meta & 0xffu, bdeg = (meta >> 8) & 0xffu;
    float4* const s4 = reinterpret_cast<float4*>(scr);
    const PatchPair* const ell_pair = reinterpret_cast<const PatchPair*>(a.ell_bi);

    const F3 pxo = mem.ldx(nullptr, pe[0].x);
    F3 pg[PB][4];
#pragma unroll
    for (uint32_t k = 0; k < PB; k++) patch_pair_points(mem, pp[k], pg[k][0], pg[k][1], pg[k][2], pg[k][3]);
    __builtin_amdgcn_sched_barrier(0);   // every gather issued before the arithmetic

    Sym3 Hin;
    F3 fin;
    inertia_term(xk, L.x0, L.v0, L.m, (meta >> 16) & 1u, a.g, ext, a.h, a.h2inv, Hin, fin);

    // the terms of the prefetched rounds, staged
    float fsz = 0.f;
    {
        Sym3 t;
        F3 tf;
        spring_term_masked(a, xk, lane < edeg, pe[0], pxo, t, tf);
        s4[2 * lane] = float4{t.xx, t.xy, t.xz, t.yy};
        s4[2 * lane + 1] = float4{t.yz, t.zz, tf.x, tf.y};
#pragma unroll
        for (int q = 0; q < (int)G; q++) fsz = fsz + lane_plus(tf.z, q);
    }
#pragma unroll
    for (uint32_t k = 0; k < PB; k++) {
        float th = 0.f;
        F3 tf = F3{0.f, 0.f, 0.f};
        if (k * G + lane < bdeg) {
            tf = bending_force(pq[k], pg[k][0], pg[k][1], pg[k][2], pg[k][3], a.kb);
            th = a.kb * __uint_as_float(pp[k].qoo);
        }
        s4[16 + 8 * k + lane] = float4{th, tf.x, tf.y, tf.z};
    }
    staged_exchange();
    // spring chain of value `lane` over the prefetched round
    float acc = 0.f;
#pragma unroll
    for (uint32_t q = 0; q < G; q++) acc = acc + scr[8 * q + lane];
    // spring rounds past the prefetched one (degrees above 8): the same
    // staging, one round at a time, after the reads above
    for (uint32_t r = PE * G; r < edeg; r += G) {
        const uint32_t j = r + lane;
        const bool on = j < edeg;
        const uint2 e = on ? a.ell_e[eoff + j * n + i] : uint2{0u, 0u};
        Sym3 t;
        F3 tf;
        spring_term_of(a, xk, on, e, on ? mem.ldx(nullptr, e.x) : F3{0.f, 0.f, 0.f}, t, tf);
        staged_exchange();
        s4[2 * lane] = float4{t.xx, t.xy, t.xz, t.yy};
        s4[2 * lane + 1] = float4{t.yz, t.zz, tf.x, tf.y};
#pragma unroll
        for (int q = 0; q < (int)G; q++) fsz = fsz + lane_plus(tf.z, q);
        staged_exchange();
#pragma unroll
        for (uint32_t q = 0; q < G; q++) acc = acc + scr[8 * q + lane];
    }
    // pair chains over the prefetched rounds
    const uint32_t bv = lane == 6u ? 1u : (lane == 7u ? 2u : (lane == 1u ? 3u : 0u));
    float accb = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < PB; k++)
#pragma unroll
        for (uint32_t q = 0; q < G; q++) accb = accb + scr[64 + 32 * k + 4 * q + bv];
    // pair rounds past the prefetched ones (degrees above 16)
    for (uint32_t r = PB * G; r < bdeg; r += G) {
        const uint32_t j = r + lane;
        float th = 0.f;
        F3 tf = F3{0.f, 0.f, 0.f};
        if (j < bdeg) {
            const uint32_t k = boff + j * n + i;
            const PatchPair e = ell_pair[k];
            F3 p0, p1, p2, p3;
            patch_pair_points(mem, e, p0, p1, p2, p3);
            tf = bending_force(a.ell_bq[k], p0, p1, p2, p3, a.kb);
            th = a.kb * __uint_as_float(e.qoo);
        }
        staged_exchange();
        s4[16 + lane] = float4{th, tf.x, tf.y, tf.z};
        staged_exchange();
#pragma unroll
        for (uint32_t q = 0; q < G; q++) accb = accb + scr[64 + 4 * q + bv];
    }
    staged_exchange();   // this group's reads before its next writes

    // lane v's value: H = inertia + S (then + hb on the diagonal), f = fin +
    // fs (then + fb)
    const bool diag = lane == 0u || lane == 3u || lane == 5u;
    const float base = lane < 6u ? (diag ? Hin.xx : 0.f) : (lane == 6u ? fin.x : fin.y);
    float val = base + acc;
    if (diag || lane >= 6u) val = val + accb;
    float fz = fin.z + fsz;
    fz = fz + lane_plus(accb, 1);
    const Sym3 H = Sym3{val, lane_plus(val, 1), lane_plus(val, 2), lane_plus(val, 3), lane_plus(val, 4),
                        lane_plus(val, 5)};
    const F3 f = F3{lane_plus(val, 6), lane_plus(val, 7), fz};
    if (!valid || lane != 0) return;
    vertex_solve(a, H, f, xk, s, sx, mem);
}

// One launch: workgroup b runs patch (b % 8) * per_xcd + b / 8 through the
// group's colour steps.  B threads, 8 lanes per vertex update, so B / 8
// updates per round; a step takes ceil(n / (B / 8)) rounds.
// LDS (patch_lds): the staged sums of every group of 8 threads, positions
// of the region (rsize float4), the frame state of the first nstate slots
// (x_start | mass, v | meta bits), the global slots of the first `owned`
// (the write-back reads them there, not from the region list in memory
// again).  The only global loads after the region are a round's update
// entries and ELL entries, which no step writes: those of the next round
// are issued before the current round's gathers and arithmetic, so their
// latency overlaps them.
template <int B>
__global__ __launch_bounds__(B) void k_vbd_patch(PatchArgs k, VbdArgs a) {
    constexpr uint32_t G = kPatchLanes;
    extern __shared__ float4 s_lds[];
    const uint32_t b = blockIdx.x;
    const uint32_t p = (b % kXcds) * k.per_xcd + b / kXcds;
    // the head's sizes arrive while the region and the first round load:
    // their addresses follow from p and the launch's strides alone
    const PatchHead hd = head_at((Uniform<PatchHead>)k.heads, p);
    const Uniform<PatchStep> steps = (Uniform<PatchStep>)k.steps + p * k.sstride;
    const uint32_t* const region = k.region + p * k.rstride;
    const float4* const state = k.state + 2u * p * k.rstride;
    const uint32_t nsteps = k.nsteps;
    const PatchLds lay = patch_lds(B, hd.rsize, hd.nstate, hd.owned);
    char* const s_base = reinterpret_cast<char*>(s_lds);
    float* const s_scr = reinterpret_cast<float*>(s_base);
    float4* const s_pos = reinterpret_cast<float4*>(s_base + lay.pos);
    float4* const s_x0m = reinterpret_cast<float4*>(s_base + lay.x0m);
    float4* const s_vm = reinterpret_cast<float4*>(s_base + lay.vm);
    uint32_t* const s_slot = reinterpret_cast<uint32_t*>(s_base + lay.slot);
    const PatchMem mem{s_pos};
    constexpr uint32_t S = B / G;   // updates per round
    const uint32_t lane = threadIdx.x % G, gi = threadIdx.x / G;
    const uint32_t wave_first = (threadIdx.x / 64) * (64 / G);   // first update index of this wave in a round
    auto next_step = [&](uint32_t j) {
        while (j < nsteps && steps[j].n == 0) j++;
        return j;
    };
    // The global loads of one round; lanes past the step's end load entry 0.
    auto load = [&](const PatchStep& st, uint32_t base, GroupLoads<G>& L, uint2& ent) {
        const uint32_t i = base + gi;
        const uint32_t ii = i < st.n ? i : 0u;
        ent = k.entries[st.off + ii];
        group_load_ell<G>(a, ii, st.n, st.eoff, st.boff, st.ew, st.bw, lane, L);
    };

    // the first round's loads, then the region (they overlap).  Every round
    // load is unconditional, from valid addresses (a step's entry `off` is
    // valid even when the step is empty): a load on one path only would make
    // the compiler wait for every load in flight where the paths join.
    uint32_t j = next_step(0), base = 0;
    PatchStep st = step_at(steps, j < nsteps ? j : 0u);
    // two register sets for the rounds' loads, used alternately (a copy from
    // one to the other would wait for the loads in flight)
    GroupLoads<G> LA, LB;
    uint2 eA, eB;
    load(st, 0u, LA, eA);
    {
        // RU slots per thread per pass, every load of a pass issued before
        // its stores (addresses clamped into the padded region, stores masked
        // by the patch's own sizes)
        constexpr uint32_t RU = B >= 1024 ? 2u : (B >= 512 ? 4u : 8u);
        // (the frame state comes in region order, so its loads need no
        // slot index: only the positions wait for one)
        const uint32_t rs = k.rstride, ns = k.nsmax;
        for (uint32_t r0 = 0; r0 < rs; r0 += RU * B) {
            uint32_t sl[RU];
            float4 xm[RU], vm[RU];
#pragma unroll
            for (uint32_t u = 0; u < RU; u++) {
                const uint32_t r = r0 + u * B + threadIdx.x;
                sl[u] = region[r < rs ? r : rs - 1u];
                const uint32_t rn = r < ns ? r : ns - 1u;
                xm[u] = state[2u * rn];
                vm[u] = state[2u * rn + 1u];
            }
            F3 xp[RU];
#pragma unroll
            for (uint32_t u = 0; u < RU; u++) xp[u] = ld3(k.xin, sl[u]);
            // (every value used here, so no load sinks into a masked store
            // below and waits there)
#pragma unroll
            for (uint32_t u = 0; u < RU; u++)
                asm volatile("" ::"v"(xp[u].x), "v"(xp[u].y), "v"(xp[u].z), "v"(xm[u].x), "v"(xm[u].y), "v"(xm[u].z),
                             "v"(xm[u].w), "v"(vm[u].x), "v"(vm[u].y), "v"(vm[u].z), "v"(vm[u].w));
#pragma unroll
            for (uint32_t u = 0; u < RU; u++) {
                const uint32_t r = r0 + u * B + threadIdx.x;
                if (r < hd.rsize) s_pos[r] = float4{xp[u].x, xp[u].y, xp[u].z, 0.f};
                if (r < hd.owned) s_slot[r] = sl[u];
                if (r < hd.nstate) {
                    s_x0m[r] = xm[u];
                    s_vm[r] = vm[u];
                }
            }
        }
    }
    const F3 ext = F3{a.fp->ext_cur[0], a.fp->ext_cur[1], a.fp->ext_cur[2]};
    __syncthreads();   // region in LDS
    // One round: issue the next round's loads into (nxt, enxt), update the
    // vertices of the current one from (cur, ecur); false after the last.
    auto round = [&](GroupLoads<G>& cur, uint2& ecur, GroupLoads<G>& nxt, uint2& enxt) {
        uint32_t j2 = j, base2 = base + S;
        PatchStep st2 = st;
        if (base2 >= st.n) {
            base2 = 0;
            j2 = next_step(j + 1);
            if (j2 < nsteps) st2 = step_at(steps, j2);
        }
        load(st2, base2, nxt, enxt);   // (after the last round: this round's first, unused)
        if (base + wave_first < st.n) {   // (the whole wave is past the step's end otherwise)
            const uint32_t i = base + gi;
            const bool valid = i < st.n;
            const float4 xm = s_x0m[ecur.y], vm = s_vm[ecur.y];
            cur.x0 = F3{xm.x, xm.y, xm.z};
            cur.m = xm.w;
            cur.v0 = F3{vm.x, vm.y, vm.z};
            cur.meta = __float_as_uint(vm.w);
            const F3 xk = mem.ldx(nullptr, ecur.y);
            patch_update(a, cur, xk, ecur.x, ecur.y, valid ? i : 0u, st.n, st.eoff, st.boff, ext, lane, valid, mem,
                         s_scr + gi * kStagedWords);
        }
        if (j2 != j) __syncthreads();   // the step's LDS writes before the next step's reads
        st = st2;
        j = j2;
        base = base2;
        return j < nsteps;
    };
    while (j < nsteps) {
        if (!round(LA, eA, LB, eB)) break;
        if (!round(LB, eB, LA, eA)) break;
    }
    for (uint32_t r = threadIdx.x; r < hd.owned; r += B) {
        const float4 v = s_pos[r];
        st3(k.xout, s_slot[r], F3{v.x, v.y, v.z});
    }
}

// The frame state of region entry e (x_start | mass, v | meta bits) into
// st[2e], st[2e + 1], for the n entries of every group's regions.
__global__ __launch_bounds__(256) void k_patch_stage(const uint32_t* region, uint32_t n, const float* x0,
                                                    const float* v, const float* mass, const uint32_t* meta,
                                                    float4* st) {
    const uint32_t e = blockIdx.x * 256u + threadIdx.x;
    if (e >= n) return;
    const uint32_t s = region[e];
    const F3 a = ld3(x0, s), w = ld3(v, s);
    st[2u * e] = float4{a.x, a.y, a.z, mass[s]};
    st[2u * e + 1u] = float4{w.x, w.y, w.z, __uint_as_float(meta[s])};
}

const void* patch_kernel(uint32_t threads) {
    switch (threads) {
        case 256: return (const void*)k_vbd_patch<256>;
        case 512: return (const void*)k_vbd_patch<512>;
        default: return (const void*)k_vbd_patch<1024>;
    }
}

}  // namespace

namespace lats {

// Patch updates run on 8 lanes whatever the context's lanes per vertex (which
// the other strategies use).
uint32_t patch_lanes(const lats_vbd*) { return kPatchLanes; }
// Threads per workgroup of the launches (256, 512 or 1024).
uint32_t patch_block(const lats_vbd* c) {
    return (c->patch_threads == 256 || c->patch_threads == 512) ? c->patch_threads : 1024u;
}

// Builds the patch layout of context c for P patches and F colours per
// launch from the context's device arrays.
int patch_build(lats_vbd* c, uint32_t P, uint32_t F) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t V = c->V;
    std::vector<uint32_t> meta(V);
    std::vector<float> x(3 * (size_t)V);
    std::vector<uint2> ell_e(c->ell_e.n);
    std::vector<float4> ell_bq(c->ell_bq.n);
    std::vector<PairIds> ell_bi(c->ell_bi.n);
    HIPCHK(hipMemcpyAsync(meta.data(), c->meta.p, 4 * (size_t)V, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(x.data(), c->x.p, 12 * (size_t)V, hipMemcpyDeviceToHost, c->stream));
    if (!ell_e.empty())
        HIPCHK(hipMemcpyAsync(ell_e.data(), c->ell_e.p, ell_e.size() * sizeof(uint2), hipMemcpyDeviceToHost, c->stream));
    if (!ell_bq.empty())
        HIPCHK(hipMemcpyAsync(ell_bq.data(), c->ell_bq.p, ell_bq.size() * sizeof(float4), hipMemcpyDeviceToHost,
                              c->stream));
    if (!ell_bi.empty())
        HIPCHK(hipMemcpyAsync(ell_bi.data(), c->ell_bi.p, ell_bi.size() * sizeof(PairIds), hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int lds_limit = 0;
    HIPCHK(hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    PatchSource src;
    src.V = V;
    src.C = c->C;
    src.colours = c->colours.data();
    src.meta = meta.data();
    src.x = x.data();
    src.ell_e = ell_e.data();
    src.ell_bq = ell_bq.data();
    src.ell_bi = ell_bi.data();
    PatchHost h;
    std::string err;
    // (room for the staged sums of the largest workgroup)
    const uint32_t staged = patch_staged_bytes(1024);
    if (!patch_build_host(src, P, F, (uint32_t)lds_limit > staged ? (uint32_t)lds_limit - staged : 0u, h, err))
        return fail(err.rfind("patch", 0) == 0 ? LATS_ERR_UNSUPPORTED : LATS_ERR_INVALID_ARGUMENT, err);
    HIPCHK(c->pt_heads.upload(h.heads));
    HIPCHK(c->pt_steps.upload(h.steps));
    HIPCHK(c->pt_region.upload(h.region));
    if (c->pt_state.n != 2 * h.region.size()) HIPCHK(c->pt_state.alloc(2 * h.region.size()));
    HIPCHK(c->pt_entries.upload(h.entries));
    HIPCHK(c->pt_ell_e.upload(h.ell_e));
    HIPCHK(c->pt_ell_bq.upload(h.ell_bq));
    HIPCHK(c->pt_ell_bi.upload(h.ell_bi));
    if (c->xalt.n != 3 * (size_t)V) HIPCHK(c->xalt.alloc(3 * (size_t)V));
    c->pt_group_lds = h.group_lds;
    c->pt_group_roff = h.group_roff;
    c->pt_group_rstride = h.group_rstride;
    c->pt_group_c0 = h.group_c0;
    c->pt_group_nsmax = h.group_nsmax;
    h.stats.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->patch_stats = h.stats;
    c->patch_built_p = h.stats.patches;
    c->patch_built_f = h.stats.colours_per_launch;
    return LATS_OK;
}

// The patch count used when none is set: about 256 slots per patch, at
// least 64 and at most 1024 patches, a multiple of 8.
uint32_t patch_default_count(const lats_vbd* c) {
    uint32_t P = std::max(64u, std::min(1024u, c->V / 256u));
    P = std::max(8u, P / 8u * 8u);
    while (P > 8 && P > c->V) P -= 8;
    return P;
}

int patch_prepare(lats_vbd* c) {
    if (c->V < 8) return fail(LATS_ERR_UNSUPPORTED, "patch strategy needs at least 8 vertices");
    const uint32_t P = c->patch_target ? c->patch_target : patch_default_count(c);
    uint32_t F = c->patch_colours ? std::min(c->patch_colours, c->C) : c->C;
    F = std::min(F, kMaxPatchSteps);
    if (c->patch_built_p == P && c->patch_built_f == F) return LATS_OK;
    c->drop_graphs();
    c->patch_built_p = c->patch_built_f = 0;
    int rc = patch_build(c, P, F);
    if (rc) return rc;
    // dynamic LDS above the default limit must be allowed per kernel
    const uint32_t most = *std::max_element(c->pt_group_lds.begin(), c->pt_group_lds.end());
    for (uint32_t threads : {256u, 512u, 1024u})
        HIPCHK(hipFuncSetAttribute(patch_kernel(threads), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(most + patch_staged_bytes(threads))));
    return LATS_OK;
}

uint32_t patch_launches_per_sweep(const lats_vbd* c) { return (uint32_t)c->pt_group_lds.size(); }

// Launch of group g reading positions from xin and writing them to xout.
void enqueue_patch_launch(lats_vbd* c, uint32_t g, const float* xin, float* xout) {
    const uint32_t P = c->patch_built_p;
    PatchArgs k;
    k.heads = c->pt_heads.p + (size_t)g * P;
    k.steps = c->pt_steps.p + (size_t)g * P * c->patch_built_f;
    k.region = c->pt_region.p + c->pt_group_roff[g];
    k.entries = c->pt_entries.p;
    k.state = c->pt_state.p + 2 * (size_t)c->pt_group_roff[g];
    k.xin = xin;
    k.xout = xout;
    k.per_xcd = P / kXcds;
    k.rstride = c->pt_group_rstride[g];
    k.nsmax = c->pt_group_nsmax[g];
    k.nsteps = c->pt_group_c0[g + 1] - c->pt_group_c0[g];
    k.sstride = c->patch_built_f;
    VbdArgs a = c->args(nullptr);
    a.ell_e = c->pt_ell_e.p;
    a.ell_bq = c->pt_ell_bq.p;
    // (local pair entries: PatchPair records in the global entries' place)
    a.ell_bi = reinterpret_cast<const PairIds*>(c->pt_ell_bi.p);
    const uint32_t threads = patch_block(c);
    const size_t shm = c->pt_group_lds[g] + patch_staged_bytes(threads);
    const dim3 grid(P), block(threads);
    switch (threads) {
        case 256: hipLaunchKernelGGL(k_vbd_patch<256>, grid, block, shm, c->stream, k, a); break;
        case 512: hipLaunchKernelGGL(k_vbd_patch<512>, grid, block, shm, c->stream, k, a); break;
        default: hipLaunchKernelGGL(k_vbd_patch<1024>, grid, block, shm, c->stream, k, a); break;
    }
}

// The frame state of every region entry, for the patch launches that follow
// (x_start and v stay fixed while a frame's sweeps run).
void enqueue_patch_stage(lats_vbd* c) {
    const uint32_t n = (uint32_t)c->pt_region.n;
    hipLaunchKernelGGL(k_patch_stage, dim3((n + 255u) / 256u), dim3(256), 0, c->stream, c->pt_region.p, n, c->x0.p,
                       c->v.p, c->mass.p, c->meta.p, c->pt_state.p);
}

}  // namespace lats
