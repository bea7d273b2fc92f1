// vbd_kernels.h -- device-side pieces shared by the single-cloth and the
// batched VBD kernels: frame parameters, the argument block of one cloth and
// the fused per-vertex update.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "vbd_collide.h"
#include "vbd_common.h"

namespace lats {

// One colour of one cloth: slots [start, start + count), ELL bases and
// ELL widths (largest spring / bending-pair degree in the colour).
struct ColourInfo {
    uint32_t start, count, eoff, boff, ew, bw;
};

// One incidence of a vertex in a bending pair: the pair's other three
// slots in pair order, and in the top two bits of `a` the vertex's own
// position o in the pair (the vertex's row of Q comes with it, in ell_bq).
struct PairIds {
    uint32_t a, b, c;
};
// The pair's four points in pair order: the vertex itself at position o,
// the others (g0, g1, g2) around it.
__device__ __forceinline__ void pair_points(uint32_t o, F3 xk, F3 g0, F3 g1, F3 g2, F3& p0, F3& p1, F3& p2, F3& p3) {
    p0 = (o == 0) ? xk : g0;
    p1 = (o == 1) ? xk : (o == 0) ? g0 : g1;
    p2 = (o == 2) ? xk : (o <= 1) ? g1 : g2;
    p3 = (o == 3) ? xk : g2;
}

// Most bands a colour pass cuts a colour into (one per XCD).
constexpr uint32_t kMaxRegions = 8;

struct FrameParams {
    const float* forces;   // per-frame external force (3 floats / frame) or null
    uint32_t frame;        // frame counter, advanced by the last kernel of a frame
    uint32_t pad;
    float ext[4];          // constant external force when forces == null
    float ext_cur[4];      // force of the running frame (written by predict)
};

struct VbdArgs {
    float* x;                  // positions being iterated (slot order)
    const float* x0;           // x_start
    const float* v;            // frame-start velocity
    const float* mass;
    const uint32_t* meta;
    const uint2* ell_e;
    const float4* ell_bq;
    const PairIds* ell_bi;
    const FrameParams* fp;
    float h, h2inv, ks, kb;
    float g[3];
    // obstacle contacts of the running frame (ccount == null: none)
    const uint32_t* ccount;    // contacts per slot
    const float4* ct_xs;       // (closest point, stiffness), entry k of slot s at k * nslots + s
    const float4* ct_n;        // face normal
    uint32_t nslots;
    float thickness;
};

// 12-byte position records, read and written with one dwordx3 access.
struct __attribute__((packed, aligned(4))) P3 {
    float x, y, z;
};
__device__ __forceinline__ F3 ld3(const float* p, uint32_t s) {
    const P3 v = *reinterpret_cast<const P3*>(p + 3 * (size_t)s);
    return F3{v.x, v.y, v.z};
}
__device__ __forceinline__ void st3(float* p, uint32_t s, F3 a) {
    *reinterpret_cast<P3*>(p + 3 * (size_t)s) = P3{a.x, a.y, a.z};
}

// Value of lane (this lane + q) of the same row of 16 lanes (DPP row shift,
// a VALU move; q is a constant after unrolling).  Lane 0 of an aligned group
// of G <= 16 lanes reads lane q of its own group.
__device__ __forceinline__ float lane_plus(float v, int q) {
#define LATS_ROW_SHL(n) __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x100 + (n), 0xF, 0xF, true))
    switch (q) {
        case 1: return LATS_ROW_SHL(1);
        case 2: return LATS_ROW_SHL(2);
        case 3: return LATS_ROW_SHL(3);
        case 4: return LATS_ROW_SHL(4);
        case 5: return LATS_ROW_SHL(5);
        case 6: return LATS_ROW_SHL(6);
        case 7: return LATS_ROW_SHL(7);
        case 8: return LATS_ROW_SHL(8);
        case 9: return LATS_ROW_SHL(9);
        case 10: return LATS_ROW_SHL(10);
        case 11: return LATS_ROW_SHL(11);
        case 12: return LATS_ROW_SHL(12);
        case 13: return LATS_ROW_SHL(13);
        case 14: return LATS_ROW_SHL(14);
        case 15: return LATS_ROW_SHL(15);
        default: return v;
    }
#undef LATS_ROW_SHL
}

// How a vertex update reads and writes the per-vertex state (x, x_start, v).
// PlainMem: ordinary cached accesses (kernel boundaries order the colour
// passes).  CoherentMem: device-scope relaxed atomic accesses, coherent
// across the L2s of all XCDs without cache maintenance, for kernels in
// which blocks hand positions to each other inside one launch.
//
// ld / st serve the vertex's own state (x_start, v); ldx / stx serve the
// positions, own and gathered, so a policy can hold positions elsewhere.
// Policies are passed by value (PlainMem and ScopedMem carry no state).
// kStaged: the positions live outside a.x while a frame runs and the frame's
// last step writes them back.
struct PlainMem {
    static constexpr bool kStaged = false;
    static __device__ __forceinline__ F3 ld(const float* p, uint32_t s) { return ld3(p, s); }
    static __device__ __forceinline__ void st(float* p, uint32_t s, F3 a) { st3(p, s, a); }
    static __device__ __forceinline__ F3 ldx(const float* p, uint32_t s) { return ld3(p, s); }
    static __device__ __forceinline__ void stx(float* p, uint32_t s, F3 a) { st3(p, s, a); }
};
template <int Scope>
struct ScopedMem {
    static __device__ __forceinline__ float ld1(const float* p) {
        return __hip_atomic_load(const_cast<float*>(p), __ATOMIC_RELAXED, Scope);
    }
    static __device__ __forceinline__ void st1(float* p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, Scope); }
    static __device__ __forceinline__ F3 ld(const float* p, uint32_t s) {
        return F3{ld1(p + 3 * s), ld1(p + 3 * s + 1), ld1(p + 3 * s + 2)};
    }
    static __device__ __forceinline__ void st(float* p, uint32_t s, F3 a) {
        st1(p + 3 * s, a.x);
        st1(p + 3 * s + 1, a.y);
        st1(p + 3 * s + 2, a.z);
    }
    static constexpr bool kStaged = false;
    static __device__ __forceinline__ F3 ldx(const float* p, uint32_t s) { return ld(p, s); }
    static __device__ __forceinline__ void stx(float* p, uint32_t s, F3 a) { st(p, s, a); }
};
using CoherentMem = ScopedMem<__HIP_MEMORY_SCOPE_AGENT>;
// LdsMem: the positions of one cloth, slots [vbase, vbase + last], held in
// LDS (slot s at lds[3 * (s - vbase)]) by the block that owns the cloth for
// a whole frame; the own state stays in HBM.  A gathered slot outside the
// cloth (an ELL padding entry, whose term is masked) is clamped into it, so
// every LDS access stays inside the cloth's 12 * (last + 1) bytes.
struct LdsMem {
    static constexpr bool kStaged = true;
    float* lds;
    uint32_t vbase, last;
    __device__ __forceinline__ uint32_t at(uint32_t s) const {
        const uint32_t i = s - vbase;
        return 3u * (i < last ? i : last);
    }
    static __device__ __forceinline__ F3 ld(const float* p, uint32_t s) { return ld3(p, s); }
    static __device__ __forceinline__ void st(float* p, uint32_t s, F3 a) { st3(p, s, a); }
    __device__ __forceinline__ F3 ldx(const float*, uint32_t s) const {
        const uint32_t k = at(s);
        return F3{lds[k], lds[k + 1], lds[k + 2]};
    }
    __device__ __forceinline__ void stx(float*, uint32_t s, F3 a) const {
        const uint32_t k = at(s);
        lds[k] = a.x;
        lds[k + 1] = a.y;
        lds[k + 2] = a.z;
    }
};

// One incident spring / bending pair term, zero for lanes past the degree.
// The terms are the values the reference adds, taken on their own; every sum
// they enter starts at +0, and such a sum (never -0 in round-to-nearest) is
// unchanged by a +0 or a -0 term, so the sums are the reference's.
__device__ __forceinline__ void spring_term_of(const VbdArgs& a, F3 xk, bool on, uint2 e, F3 xo, Sym3& t, F3& tf) {
    t = sym_zero();
    tf = F3{0.f, 0.f, 0.f};
    if (on) stretch_parts(xk, xo, __uint_as_float(e.y), a.ks, t, tf);
}
// The same, computed on every lane and masked afterwards: a lane past the
// degree works on a valid placeholder entry (finite: l >= epsilon).  Used
// where the neighbour was gathered ahead: computing the term unconditionally
// keeps that gather out of a branch, so it is issued with the others.
__device__ __forceinline__ void spring_term_masked(const VbdArgs& a, F3 xk, bool on, uint2 e, F3 xo, Sym3& t,
                                                   F3& tf) {
    Sym3 u;
    F3 uf;
    stretch_parts(xk, xo, __uint_as_float(e.y), a.ks, u, uf);
    t = Sym3{on ? u.xx : 0.f, on ? u.xy : 0.f, on ? u.xz : 0.f, on ? u.yy : 0.f, on ? u.yz : 0.f, on ? u.zz : 0.f};
    tf = F3{on ? uf.x : 0.f, on ? uf.y : 0.f, on ? uf.z : 0.f};
}

// One VBD vertex update on one lane: VBD::compute_inertia +
// compute_stretch_mass_spring + compute_bending_quadratic
// (xpbd_constraints.h:4027-4205) summed as Hf = inertia; Hf += sum(stretch);
// Hf += sum(bending) (example2_async_iter.cpp:939-992), then the guarded 3x3
// Newton step of vbd_step (:993-1014).  i = index inside the colour, n =
// colour size, ew / bw = the colour's ELL widths.
// The incidences are loaded ahead of their use, in chunks: the first chunk
// of kOneChunk spring entries and of kOneChunk bending entries together with
// the vertex's own state, every later chunk in one go, each from rows
// clamped to the colour's width (padding entries are valid slots), so a
// vertex waits for about (edeg + bdeg) / kOneChunk ELL round trips rather
// than one per incidence.  The sums run in incidence order as in the
// reference; a spring past the degree adds a +0 term (a sum that starts at
// +0 never becomes -0, so +0 changes nothing), a pair past the degree adds
// nothing.
constexpr uint32_t kOneChunk = 4;
template <class Mem = PlainMem>
__device__ __forceinline__ void vbd_vertex_ext(const VbdArgs& a, uint32_t s, uint32_t i, uint32_t n, uint32_t eoff,
                                               uint32_t boff, uint32_t ew, uint32_t bw, F3 ext, Mem mem = Mem{}) {
    const uint32_t e_on = ew != 0u, b_on = bw != 0u;
    uint2 pe[kOneChunk];
    float4 pq[kOneChunk];
    PairIds pi[kOneChunk];
    auto load_springs = [&](uint32_t r0) {
#pragma unroll
        for (uint32_t k = 0; k < kOneChunk; k++)
            pe[k] = a.ell_e[(eoff + (r0 + k < ew ? r0 + k : ew - 1) * n + i) * e_on];
    };
    auto load_pairs = [&](uint32_t r0) {
#pragma unroll
        for (uint32_t k = 0; k < kOneChunk; k++) {
            const uint32_t at = (boff + (r0 + k < bw ? r0 + k : bw - 1) * n + i) * b_on;
            pq[k] = a.ell_bq[at];
            pi[k] = a.ell_bi[at];
        }
    };
    load_springs(0);
    load_pairs(0);
    __builtin_amdgcn_sched_barrier(0);   // the ELL loads go out first
    const F3 xk = mem.ldx(a.x, s);
    const F3 x0 = mem.ld(a.x0, s);
    const F3 v0 = mem.ld(a.v, s);
    const float m = a.mass[s];
    const uint32_t meta = a.meta[s];
    const uint32_t edeg = meta & 0xffu, bdeg = (meta >> 8) & 0xffu;

    // inertia
    Sym3 H;
    F3 f;
    inertia_term(xk, x0, v0, m, (meta >> 16) & 1u, a.g, ext, a.h, a.h2inv, H, f);

    // stretch: local sum from zero, then added to Hf
    Sym3 S = sym_zero();
    F3 fs = F3{0.f, 0.f, 0.f};
    for (uint32_t r0 = 0; r0 < edeg; r0 += kOneChunk) {
        if (r0) load_springs(r0);
        F3 xo[kOneChunk];
#pragma unroll
        for (uint32_t k = 0; k < kOneChunk; k++) xo[k] = mem.ldx(a.x, pe[k].x);
#pragma unroll
        for (uint32_t k = 0; k < kOneChunk; k++) {
            Sym3 t;
            F3 tf;
            spring_term_masked(a, xk, r0 + k < edeg, pe[k], xo[k], t, tf);
            sym_add(S, t);
            fs = add3(fs, tf);
        }
    }
    sym_add(H, S);
    f = add3(f, fs);

    // bending: diagonal Hessian only (k * Q[o][o] * I)
    float hb = 0.f;
    F3 fb = F3{0.f, 0.f, 0.f};
    for (uint32_t r0 = 0; r0 < bdeg; r0 += kOneChunk) {
        if (r0) load_pairs(r0);
        F3 g[kOneChunk][3];
#pragma unroll
        for (uint32_t k = 0; k < kOneChunk; k++) {
            g[k][0] = mem.ldx(a.x, pi[k].a & 0x3fffffffu);
            g[k][1] = mem.ldx(a.x, pi[k].b);
            g[k][2] = mem.ldx(a.x, pi[k].c);
        }
#pragma unroll
        for (uint32_t k = 0; k < kOneChunk; k++) {
            if (r0 + k < bdeg) {
                const uint32_t o = pi[k].a >> 30;
                F3 p0, p1, p2, p3;
                pair_points(o, xk, g[k][0], g[k][1], g[k][2], p0, p1, p2, p3);
                bending_term(pq[k], o, p0, p1, p2, p3, a.kb, hb, fb);
            }
        }
    }
    H.xx = H.xx + hb;
    H.yy = H.yy + hb;
    H.zz = H.zz + hb;
    f = add3(f, fb);

    // obstacle contacts: local sum from zero, then added (the contact
    // Hessian is not exactly symmetric, so the full 3x3 is solved)
    if (a.ccount) {
        const uint32_t nc = a.ccount[s];
        if (nc) {
            float Hc[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
            F3 fc = F3{0.f, 0.f, 0.f};
            for (uint32_t k = 0; k < nc; k++) {
                const float4 q = a.ct_xs[k * a.nslots + s];
                const float4 nn = a.ct_n[k * a.nslots + s];
                contact_term(xk, F3{q.x, q.y, q.z}, F3{nn.x, nn.y, nn.z}, q.w, a.thickness, Hc, fc);
            }
            const float M[3][3] = {{H.xx + Hc[0][0], H.xy + Hc[0][1], H.xz + Hc[0][2]},
                                   {H.xy + Hc[1][0], H.yy + Hc[1][1], H.yz + Hc[1][2]},
                                   {H.xz + Hc[2][0], H.yz + Hc[2][1], H.zz + Hc[2][2]}};
            f = add3(f, fc);
            F3 dx;
            if (newton_step_full(M, f, dx)) mem.stx(a.x, s, add3(xk, dx));
            return;
        }
    }

    F3 dx;
    if (newton_step(H, f, dx)) mem.stx(a.x, s, add3(xk, dx));
}

// The same update with G lanes per vertex (lane = 0..G-1 inside an aligned
// group of G lanes).  Lane j computes the terms of incident springs / pairs
// j, j+G, ...; lane 0 then adds them in incidence order, reading the other
// lanes' terms with DPP row shifts, so every sum is formed in the reference's
// order and the result is bit-identical to vbd_vertex_ext.  All G lanes of a group must call this
// together (valid == false: the group has no vertex; it still takes part in
// the shuffles and stores nothing).
//
// ew / bw are the colour's ELL widths (its largest spring / pair degree).
// The ELL entries of the first PE spring rounds and PB pair rounds are loaded
// before the vertex's degrees arrive (padding entries are valid slots), so
// the neighbour gathers follow after one memory round trip instead of two.
// The loads of one vertex update with G lanes per vertex that come from
// memory the update does not write: the ELL entries of the first PE spring
// rounds and PB pair rounds, and the vertex's own frame state.  Split from
// the arithmetic (group_finish) so a caller can issue them one update ahead.
template <int G>
struct GroupLoads {
    static constexpr uint32_t PE = (8 + G - 1) / G;    // spring rounds prefetched (degrees <= 8)
    static constexpr uint32_t PB = (16 + G - 1) / G;   // pair rounds prefetched (degrees <= 16)
    uint2 pe[PE];
    float4 pq[PB];
    PairIds pi[PB];
    F3 x0, v0;
    float m;
    uint32_t meta;
};

// Every load below is unconditional, from an address that is always
// valid: ELL rows past the colour's width are clamped to its last row
// (a colour without springs / pairs reads entry 0), and padding entries
// hold slot 0.  Values a lane does not own are masked where they are
// used.  A conditional load (cond ? load : 0) makes the compiler wait for
// it before the next one, which serialises the prefetch rounds.
template <int G>
__device__ __forceinline__ void group_load_ell(const VbdArgs& a, uint32_t i, uint32_t n, uint32_t eoff, uint32_t boff,
                                               uint32_t ew, uint32_t bw, uint32_t lane, GroupLoads<G>& L) {
    // (index arithmetic without branches: a zero width multiplies the index
    // by zero, so no load waits behind a branch on the width)
    const uint32_t e_on = ew != 0u, b_on = bw != 0u;
#pragma unroll
    for (uint32_t k = 0; k < GroupLoads<G>::PE; k++) {
        const uint32_t j = k * G + lane;
        const uint32_t row = j < ew ? j : ew - 1;
        L.pe[k] = a.ell_e[(eoff + row * n + i) * e_on];
    }
#pragma unroll
    for (uint32_t k = 0; k < GroupLoads<G>::PB; k++) {
        const uint32_t j = k * G + lane;
        const uint32_t row = j < bw ? j : bw - 1;
        const uint32_t at = (boff + row * n + i) * b_on;
        L.pq[k] = a.ell_bq[at];
        L.pi[k] = a.ell_bi[at];
    }
}
template <int G, class Mem>
__device__ __forceinline__ void group_load_own(const VbdArgs& a, uint32_t s, GroupLoads<G>& L, Mem mem) {
    L.x0 = mem.ld(a.x0, s);
    L.v0 = mem.ld(a.v, s);
    L.m = a.mass[s];
    L.meta = a.meta[s];
}

template <int G, class Mem>
__device__ __forceinline__ void group_finish(const VbdArgs& a, const GroupLoads<G>& L, F3 xk, uint32_t s, uint32_t sx,
                                             uint32_t i, uint32_t n, uint32_t eoff, uint32_t boff, F3 ext,
                                             uint32_t lane, bool valid, Mem mem);

// The end of a vertex update once H and f hold the inertia, stretch and
// bending sums: obstacle contacts of slot s (their local sum from zero, then
// added; the contact Hessian is not exactly symmetric, so the full 3x3 is
// solved), then the guarded Newton step; the new position goes to sx.
template <class Mem>
__device__ __forceinline__ void vertex_solve(const VbdArgs& a, const Sym3& H, F3 f, F3 xk, uint32_t s, uint32_t sx,
                                             Mem mem) {
    if (a.ccount) {
        const uint32_t nc = a.ccount[s];
        if (nc) {
            float Hc[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
            F3 fc = F3{0.f, 0.f, 0.f};
            for (uint32_t k = 0; k < nc; k++) {
                const float4 q = a.ct_xs[k * a.nslots + s];
                const float4 nn = a.ct_n[k * a.nslots + s];
                contact_term(xk, F3{q.x, q.y, q.z}, F3{nn.x, nn.y, nn.z}, q.w, a.thickness, Hc, fc);
            }
            const float M[3][3] = {{H.xx + Hc[0][0], H.xy + Hc[0][1], H.xz + Hc[0][2]},
                                   {H.xy + Hc[1][0], H.yy + Hc[1][1], H.yz + Hc[1][2]},
                                   {H.xz + Hc[2][0], H.yz + Hc[2][1], H.zz + Hc[2][2]}};
            f = add3(f, fc);
            F3 dx;
            if (newton_step_full(M, f, dx)) mem.stx(a.x, sx, add3(xk, dx));
            return;
        }
    }
    F3 dx;
    if (newton_step(H, f, dx)) mem.stx(a.x, sx, add3(xk, dx));
}

template <int G, class Mem = PlainMem>
__device__ __forceinline__ void vbd_vertex_group(const VbdArgs& a, uint32_t s, uint32_t i, uint32_t n, uint32_t eoff,
                                                 uint32_t boff, uint32_t ew, uint32_t bw, F3 ext, uint32_t lane,
                                                 bool valid, Mem mem = Mem{}) {
    GroupLoads<G> L;
    group_load_ell<G>(a, i, n, eoff, boff, ew, bw, lane, L);
    // the ELL loads go out first: their addresses need no argument-block read
    // when the caller has the ELL bases and colour geometry in registers
    __builtin_amdgcn_sched_barrier(0);
    const F3 xk = mem.ldx(a.x, s);
    group_load_own<G>(a, s, L, mem);
    // every load above is issued before any gather below (the scheduler
    // would otherwise issue later rounds' ELL loads after the first wait)
    __builtin_amdgcn_sched_barrier(0);
    group_finish<G>(a, L, xk, s, s, i, n, eoff, boff, ext, lane, valid, mem);
}

// The gathers and the arithmetic of the update: s = the vertex's slot (own
// state, contacts), sx = where its position is read and written through mem.
template <int G, class Mem>
__device__ __forceinline__ void group_finish(const VbdArgs& a, const GroupLoads<G>& L, F3 xk, uint32_t s, uint32_t sx,
                                             uint32_t i, uint32_t n, uint32_t eoff, uint32_t boff, F3 ext,
                                             uint32_t lane, bool valid, Mem mem) {
    constexpr uint32_t PE = GroupLoads<G>::PE, PB = GroupLoads<G>::PB;
    const uint2* pe = L.pe;
    const float4* pq = L.pq;
    const PairIds* pi = L.pi;
    const F3 x0 = L.x0, v0 = L.v0;
    const float m = L.m;
    const uint32_t meta = valid ? L.meta : 0u;
    const uint32_t edeg = meta & 0xffu, bdeg = (meta >> 8) & 0xffu;

    // neighbour gathers of the prefetched rounds (lanes past the degree load
    // a valid slot and mask the term)
    F3 pxo[PE];
#pragma unroll
    for (uint32_t k = 0; k < PE; k++) pxo[k] = mem.ldx(a.x, pe[k].x);
    F3 pg[PB][3];
#pragma unroll
    for (uint32_t k = 0; k < PB; k++) {
        pg[k][0] = mem.ldx(a.x, pi[k].a & 0x3fffffffu);
        pg[k][1] = mem.ldx(a.x, pi[k].b);
        pg[k][2] = mem.ldx(a.x, pi[k].c);
    }
    __builtin_amdgcn_sched_barrier(0);   // every gather issued before the arithmetic

    Sym3 H;
    F3 f;
    inertia_term(xk, x0, v0, m, (meta >> 16) & 1u, a.g, ext, a.h, a.h2inv, H, f);

    Sym3 S = sym_zero();
    F3 fs = F3{0.f, 0.f, 0.f};
    // only lane 0 of the group uses the sums: it adds the terms of lanes
    // 0, 1, ..., G-1 in that order.  Lanes past the degree hold +0 terms, and
    // adding +0 leaves a sum that started at +0 unchanged (in round-to-nearest
    // such a sum is never -0), so the sums equal the reference's sums over the
    // incident springs alone.  Unconditional adds let every DPP move fold into
    // its add (one v_add_f32_dpp per term).
    auto add_spring_round = [&](const Sym3& t, const F3& tf) {
#pragma unroll
        for (int q = 0; q < G; q++) {
            S.xx = S.xx + lane_plus(t.xx, q);
            S.xy = S.xy + lane_plus(t.xy, q);
            S.xz = S.xz + lane_plus(t.xz, q);
            S.yy = S.yy + lane_plus(t.yy, q);
            S.yz = S.yz + lane_plus(t.yz, q);
            S.zz = S.zz + lane_plus(t.zz, q);
            fs.x = fs.x + lane_plus(tf.x, q);
            fs.y = fs.y + lane_plus(tf.y, q);
            fs.z = fs.z + lane_plus(tf.z, q);
        }
    };
    // every prefetched round runs (no early exit that the gathers could be
    // moved behind); a round with no lane on adds +0 terms, a no-op
#pragma unroll
    for (uint32_t k = 0; k < PE; k++) {
        const uint32_t r = k * G;
        Sym3 t;
        F3 tf;
        spring_term_masked(a, xk, r + lane < edeg, pe[k], pxo[k], t, tf);
        add_spring_round(t, tf);
    }
    for (uint32_t r = PE * G; r < edeg; r += G) {
        const uint32_t j = r + lane;
        const bool on = j < edeg;
        const uint2 e = on ? a.ell_e[eoff + j * n + i] : uint2{0u, 0u};
        Sym3 t;
        F3 tf;
        spring_term_of(a, xk, on, e, on ? mem.ldx(a.x, e.x) : F3{0.f, 0.f, 0.f}, t, tf);
        add_spring_round(t, tf);
    }
    sym_add(H, S);
    f = add3(f, fs);

    float hb = 0.f;
    F3 fb = F3{0.f, 0.f, 0.f};
    auto add_pair_round = [&](float th, const F3& tf) {   // +0 past the degree, as above
#pragma unroll
        for (int q = 0; q < G; q++) {
            hb = hb + lane_plus(th, q);
            fb.x = fb.x + lane_plus(tf.x, q);
            fb.y = fb.y + lane_plus(tf.y, q);
            fb.z = fb.z + lane_plus(tf.z, q);
        }
    };
#pragma unroll
    for (uint32_t k = 0; k < PB; k++) {
        const uint32_t r = k * G;
        if (r >= bdeg) break;
        float th = 0.f;
        F3 tf = F3{0.f, 0.f, 0.f};
        if (r + lane < bdeg) {
            const uint32_t o = pi[k].a >> 30;
            F3 p0, p1, p2, p3;
            pair_points(o, xk, pg[k][0], pg[k][1], pg[k][2], p0, p1, p2, p3);
            bending_parts(pq[k], o, p0, p1, p2, p3, a.kb, th, tf);
        }
        add_pair_round(th, tf);
    }
    for (uint32_t r = PB * G; r < bdeg; r += G) {
        const uint32_t j = r + lane;
        float th = 0.f;
        F3 tf = F3{0.f, 0.f, 0.f};
        if (j < bdeg) {
            const uint32_t k = boff + j * n + i;
            const float4 q = a.ell_bq[k];
            const PairIds id = a.ell_bi[k];
            const uint32_t o = id.a >> 30;
            F3 p0, p1, p2, p3;
            pair_points(o, xk, mem.ldx(a.x, id.a & 0x3fffffffu), mem.ldx(a.x, id.b), mem.ldx(a.x, id.c), p0, p1, p2,
                        p3);
            bending_parts(q, o, p0, p1, p2, p3, a.kb, th, tf);
        }
        add_pair_round(th, tf);
    }
    H.xx = H.xx + hb;
    H.yy = H.yy + hb;
    H.zz = H.zz + hb;
    f = add3(f, fb);
    if (!valid || lane != 0) return;
    vertex_solve(a, H, f, xk, s, sx, mem);
}

// Words of LDS one group of 8 lanes uses in group_finish_staged8 (the
// colour pass's 8-lane update; the patch strategy has its own): the
// spring terms of a round (8 lanes x 8 values) and the bending terms of the
// two prefetched rounds (2 x 8 lanes x 4 values), plus 8 words that keep
// the groups of a wave on distinct banks.
constexpr uint32_t kGroupStagedWords = 136;
// Orders one wave's LDS writes and reads of the staged terms: the lanes of
// a wave run in lockstep and a wave's LDS accesses complete in issue order,
// so a wavefront-scope fence (which keeps the compiler from moving the
// accesses across it) is all the exchange needs.
__device__ __forceinline__ void group_staged_exchange() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// group_finish for 8 lanes per vertex with the sums formed one value per
// lane: every lane writes its terms to the group's LDS words `scr`, then
// lane v adds value v of lanes 0, 1, ..., 7 (round after round) in that
// order -- the same additions, in the same order, as lane 0 forms in
// group_finish, so every sum is bit-identical -- and lane 0 collects the
// finished sums with DPP moves.  Eight chains run side by side instead of
// one lane adding every value (8 adds per value instead of 8 per value and
// lane-0 only).  Spring values: lane v < 6 sums S's entry v (xx, xy, xz, yy,
// yz, zz), lanes 6 / 7 fs.x / fs.y, and fs.z is summed in lane 0 with DPP
// adds.  Bending values: lanes 0, 3, 5 sum hb (the diagonal they hold), 6 /
// 7 fb.x / fb.y, lane 1 fb.z.  scr: kGroupStagedWords words, 16-byte aligned,
// used by this group only.
template <class Mem>
__device__ __forceinline__ void group_finish_staged8(const VbdArgs& a, const GroupLoads<8>& L, F3 xk, uint32_t s,
                                                     uint32_t sx, uint32_t i, uint32_t n, uint32_t eoff,
                                                     uint32_t boff, F3 ext, uint32_t lane, bool valid, Mem mem,
                                                     float* scr) {
    constexpr uint32_t G = 8, PE = GroupLoads<8>::PE, PB = GroupLoads<8>::PB;
    static_assert(PE == 1 && PB == 2, "one spring round and two pair rounds are staged");
    const uint2* pe = L.pe;
    const float4* pq = L.pq;
    const PairIds* pi = L.pi;
    const F3 x0 = L.x0, v0 = L.v0;
    const float m = L.m;
    const uint32_t meta = valid ? L.meta : 0u;
    const uint32_t edeg = meta & 0xffu, bdeg = (meta >> 8) & 0xffu;
    float4* const s4 = reinterpret_cast<float4*>(scr);

    const F3 pxo = mem.ldx(a.x, pe[0].x);
    F3 pg[PB][3];
#pragma unroll
    for (uint32_t k = 0; k < PB; k++) {
        pg[k][0] = mem.ldx(a.x, pi[k].a & 0x3fffffffu);
        pg[k][1] = mem.ldx(a.x, pi[k].b);
        pg[k][2] = mem.ldx(a.x, pi[k].c);
    }
    __builtin_amdgcn_sched_barrier(0);   // every gather issued before the arithmetic

    Sym3 Hin;
    F3 fin;
    inertia_term(xk, x0, v0, m, (meta >> 16) & 1u, a.g, ext, a.h, a.h2inv, Hin, fin);

    // the terms of the prefetched rounds (+0 past the degree), staged
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
            const uint32_t o = pi[k].a >> 30;
            F3 p0, p1, p2, p3;
            pair_points(o, xk, pg[k][0], pg[k][1], pg[k][2], p0, p1, p2, p3);
            bending_parts(pq[k], o, p0, p1, p2, p3, a.kb, th, tf);
        }
        s4[16 + 8 * k + lane] = float4{th, tf.x, tf.y, tf.z};
    }
    group_staged_exchange();
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
        spring_term_of(a, xk, on, e, on ? mem.ldx(a.x, e.x) : F3{0.f, 0.f, 0.f}, t, tf);
        group_staged_exchange();
        s4[2 * lane] = float4{t.xx, t.xy, t.xz, t.yy};
        s4[2 * lane + 1] = float4{t.yz, t.zz, tf.x, tf.y};
#pragma unroll
        for (int q = 0; q < (int)G; q++) fsz = fsz + lane_plus(tf.z, q);
        group_staged_exchange();
#pragma unroll
        for (uint32_t q = 0; q < G; q++) acc = acc + scr[8 * q + lane];
    }
    // bending chains over the prefetched rounds
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
            const float4 q = a.ell_bq[k];
            const PairIds id = a.ell_bi[k];
            const uint32_t o = id.a >> 30;
            F3 p0, p1, p2, p3;
            pair_points(o, xk, mem.ldx(a.x, id.a & 0x3fffffffu), mem.ldx(a.x, id.b), mem.ldx(a.x, id.c), p0, p1, p2,
                        p3);
            bending_parts(q, o, p0, p1, p2, p3, a.kb, th, tf);
        }
        group_staged_exchange();
        s4[16 + lane] = float4{th, tf.x, tf.y, tf.z};
        group_staged_exchange();
#pragma unroll
        for (uint32_t q = 0; q < G; q++) accb = accb + scr[64 + 4 * q + bv];
    }
    group_staged_exchange();   // this group's reads before its next writes

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

// vbd_vertex_group<8> with the sums formed through the group's LDS words
// `scr` (kGroupStagedWords of them): bit-identical results.
template <class Mem = PlainMem>
__device__ __forceinline__ void vbd_vertex_group_staged8(const VbdArgs& a, uint32_t s, uint32_t i, uint32_t n,
                                                         uint32_t eoff, uint32_t boff, uint32_t ew, uint32_t bw, F3 ext,
                                                         uint32_t lane, bool valid, float* scr, Mem mem = Mem{}) {
    GroupLoads<8> L;
    group_load_ell<8>(a, i, n, eoff, boff, ew, bw, lane, L);
    __builtin_amdgcn_sched_barrier(0);
    const F3 xk = mem.ldx(a.x, s);
    group_load_own<8>(a, s, L, mem);
    __builtin_amdgcn_sched_barrier(0);
    group_finish_staged8(a, L, xk, s, s, i, n, eoff, boff, ext, lane, valid, mem, scr);
}

}  // namespace lats
