// vbd_patch.h -- the patch strategy: several colour passes in one launch.
//
// The cloth's slots are split into P patches (tiles of the rest shape).  A
// launch runs a group of F consecutive colours; workgroup p owns patch p.  It
// copies the positions of its region (its patch, the halo it must update
// and the vertices those updates read) from the input buffer into LDS, runs
// the F colour steps on LDS with a workgroup barrier between steps, and
// writes the positions of the slots it owns to the output buffer.
//
// A colour step of patch p updates every vertex of that colour that a later
// step of the group reads (the halo) as well as its own: in Gauss-Seidel
// order a vertex's update reads its neighbours' latest values, and a
// neighbour of an earlier colour of the group is updated in this launch, so
// the patch computes that update itself, from the same inputs as the owner
// does, with the same arithmetic: the values are identical.  Reads come from
// the input buffer only and writes go to the output buffer only, so no
// workgroup ever sees another's writes of the same launch; the next launch
// swaps the buffers.  Every slot is owned by exactly one patch, so the output
// buffer holds every position after the launch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "vbd_kernels.h"

struct lats_vbd;

namespace lats {

// Most colour steps one launch runs.
constexpr uint32_t kMaxPatchSteps = 16;
// Workgroups go to the 8 XCDs round-robin; patches are numbered in rows of
// the tile grid, so workgroup b runs patch (b % 8) * (P / 8) + b / 8 and each
// XCD works on one band of neighbouring patches (halo reads of neighbours
// then hit the XCD's own L2).  P is a multiple of 8.
constexpr uint32_t kXcds = 8;

// One colour step of one patch: its updates are entries [off, off + n) of
// the entry list ({slot, LDS index}), with a local ELL at bases eoff / boff
// and widths ew / bw.  Local ELL entries hold LDS indices: a spring entry
// {LDS index of the other end, rest length}; a pair entry the vertex's row
// of Q (as in the global layout) and a PatchPair.
struct PatchPair {
    uint32_t i01;   // LDS indices of the pair's points 0 | 1 << 16 (the vertex itself among them)
    uint32_t i23;   // points 2 | 3 << 16
    uint32_t qoo;   // bits of Q[o][o], o = the vertex's position in the pair
};
static_assert(sizeof(PatchPair) == sizeof(PairIds), "local pair entries take the global entries' place");
struct PatchStep {
    uint32_t off, n, eoff, boff, ew, bw;
};
// One patch of one group: its region is region[roff, roff + rsize) (global
// slots; LDS index = position in that list), of which the first `owned` are
// the slots the patch owns and the first `nstate` (owned, then the halo it
// updates) have their frame state staged in LDS; its steps are
// steps[step_off, step_off + nsteps).  Regions and steps lie at fixed strides
// per group (region[roff + rsize, ...) up to the stride repeats slot
// region[roff]), so a launch needs no head to address them; roff and
// step_off restate those addresses for host code.
struct PatchHead {
    uint32_t roff, rsize, owned, nstate, nsteps, step_off, pad0, pad1;
};
// LDS bytes of one workgroup: positions of the region, frame state of the
// first nstate slots (x_start and mass, v and the meta bits), global slots of
// the first `owned` (where the launch writes its results).
inline uint32_t patch_lds_bytes(uint32_t rsize, uint32_t nstate, uint32_t owned) {
    return 16u * (rsize + 2u * nstate) + 4u * owned;
}
// A patch update runs on 8 lanes (one group); its sums are formed one value
// per lane through LDS words of its group: the spring terms of a round (8
// lanes x 8 values) and the pair terms of the two prefetched rounds (2 x 8
// lanes x 4 values), plus 8 words that keep the groups of a wave on distinct
// banks.
constexpr uint32_t kPatchLanes = 8;
constexpr uint32_t kStagedWords = 136;
// LDS bytes a launch of `threads` threads adds for the staged sums.
__host__ __device__ inline uint32_t patch_staged_bytes(uint32_t threads) {
    return threads / kPatchLanes * kStagedWords * 4u;
}
// Byte offsets of one workgroup's LDS sections: staged sums at 0, then the
// region's positions, the frame state (x_start | mass, v | meta) of the
// first nstate slots, the global slots of the first `owned`; `end` = bytes
// used.  The kernel places its sections with it and the host sizes its
// launches with it.
struct PatchLds {
    uint32_t pos, x0m, vm, slot, end;
};
__host__ __device__ inline PatchLds patch_lds(uint32_t threads, uint32_t rsize, uint32_t nstate, uint32_t owned) {
    PatchLds l;
    l.pos = patch_staged_bytes(threads);
    l.x0m = l.pos + 16u * rsize;
    l.vm = l.x0m + 16u * nstate;
    l.slot = l.vm + 16u * nstate;
    l.end = l.slot + 4u * owned;
    return l;
}

struct PatchArgs {
    const PatchHead* heads;   // the group's P heads
    const PatchStep* steps;   // the group's steps: patch p's at p * sstride
    const uint32_t* region;   // the group's regions: patch p's at p * rstride
    const uint2* entries;
    const float4* state;      // the group's frame state in region order (two per entry)
    const float* xin;
    float* xout;
    uint32_t per_xcd;         // P / 8
    uint32_t rstride;         // region stride (the group's largest region)
    uint32_t nsmax;           // most slots with frame state in one region of the group
    uint32_t nsteps;          // colour steps of the group
    uint32_t sstride;         // step records per patch
};

// Positions held in LDS as float4 (16-byte aligned single accesses).
struct PatchMem {
    static constexpr bool kStaged = true;
    float4* lds;
    static __device__ __forceinline__ F3 ld(const float* p, uint32_t s) { return ld3(p, s); }
    static __device__ __forceinline__ void st(float* p, uint32_t s, F3 a) { st3(p, s, a); }
    __device__ __forceinline__ F3 ldx(const float*, uint32_t r) const {
        const float4 v = lds[r];
        return F3{v.x, v.y, v.z};
    }
    __device__ __forceinline__ void stx(float*, uint32_t r, F3 a) const { lds[r] = float4{a.x, a.y, a.z, 0.f}; }
};

// Host-side summary of a built layout.
struct PatchStats {
    uint32_t patches = 0, colours_per_launch = 0, launches_per_sweep = 0;
    uint32_t max_region = 0;          // largest region (LDS slots)
    uint64_t updates_per_sweep = 0;   // vertex updates of one sweep, halos included
    uint64_t region_slots = 0;        // sum of region sizes over a sweep's launches
    uint64_t ell_entries = 0;         // local ELL entries (springs + pairs) of a sweep
    double build_ms = 0.0;
};

}  // namespace lats
