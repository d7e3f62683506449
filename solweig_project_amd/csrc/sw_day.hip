// Persistent batched day loop: one launch runs steps [t0, t1) of every item
// of a batch (tiles of uneven size, edit scenarios), phase after phase, with
// grid-wide barriers between the phases instead of kernel boundaries.
//
// Phases of a day step: the sun march with the wall tail writes every tile's
// walk sources, a barrier, the flux step reads them (its ground-view walk reads
// the neighbouring tiles' sources), a barrier. A night step is one flux phase
// and a barrier. Within a phase, the blocks of each XCD group draw the tiles
// of that group's contiguous share from a counter, one tile at a time (a
// static share per block left the phase waiting for the blocks with the long
// sun-march rays). A tile's thermal-delay state can therefore move between
// blocks and XCDs from one step to the next; the barriers' device-scope
// release and acquire make every phase's stores visible to the next.
//
// The barrier waits back off exponentially (up to 16 sleeps of 8128 cycles):
// thousands of waves polling one counter with short sleeps slowed the blocks
// still working 3-18x (their loads queue behind the polls).
//
// This translation unit is compiled with machine LICM disabled: with the
// day-step bodies inside the phase loops, LICM hoists the bodies' constants
// and lane-derived indices out of the loops and keeps them live across every
// phase (the flux body then needs 190-256 VGPRs instead of 64). The bodies read
// the lane index through sw_tid() and their step records through constant
// address-space pointers for the same reason, and the loops pass every
// loop-invariant argument through an empty asm before each tile.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "sw_day_body.h"

namespace {

// Sync workspace words (uint32): [0] blocks past the barrier, summed over the
// XCD groups, [1..8] per-group arrivals, [9] error flag (1: a barrier wait timed
// out), [10] phases completed by block 0, [11] the grid's block count,
// [64 + 16 x] tile counter of group x (64 bytes apart).
constexpr int kSyncGlobal = 0, kSyncGroup = 1, kSyncErr = SW_DAY_SYNC_ERR, kSyncPhases = 10, kSyncGrid = 11,
              kSyncTiles = 64;
static_assert(kSyncTiles + 8 * 16 <= SW_DAY_SYNC_WORDS, "sync workspace");

// Grid barrier over nb co-resident blocks in 8 groups (group x = block & 7,
// nx blocks each). Counters only grow: barrier number `gen` completes when
// every group's counter reaches gen * nx and the global counter gen * 8.
// The waits are bounded by `limit` wall-clock ticks; on a time-out the error
// flag is set and every block leaves at its next barrier. Returns false when
// the flag is set.
__device__ __forceinline__ bool grid_barrier(uint32_t *sync, uint32_t gen, uint32_t nx, uint64_t limit) {
    // release by every wave: its stores of the phase complete and the XCD's L2 written back, so
    // that blocks on other XCDs see them after the barrier
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t x = blockIdx.x & 7u;
        const uint32_t old = atomicAdd(sync + kSyncGroup + x, 1u);
        if (old + 1u == gen * nx) atomicAdd(sync + kSyncGlobal, 1u);
        const uint64_t t_end = wall_clock64() + limit;
        int backoff = 1;
        while (__hip_atomic_load(sync + kSyncGlobal, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < gen * 8u) {
            if (__hip_atomic_load(sync + kSyncErr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
            if (wall_clock64() > t_end) {
                atomicOr(sync + kSyncErr, 1u);
                break;
            }
            for (int k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(127);
            if (backoff < 16) backoff <<= 1;
        }
    }
    __syncthreads();
    // acquire by every wave: no cache of this CU or XCD keeps a line older than the other blocks' stores
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return __hip_atomic_load(sync + kSyncErr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
}

// The next tile position of this block's group: every block of the group draws
// positions from the group's counter until one is past the group's share. Each
// block draws once past the end, so phase `phase` of a group with `cn` tiles
// and `per` blocks starts at counter value phase * (cn + per).
__device__ __forceinline__ int next_tile(uint32_t *ctr, uint32_t base) {
    __shared__ int s_pos;
    __syncthreads();   // every wave is done with the previous position
    if (threadIdx.x == 0) s_pos = (int)(atomicAdd(ctr, 1u) - base);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(s_pos);
}

typedef const SW_CONST BatchItemDev *ItemC;

template <int FIELD>
__device__ __forceinline__ int item_of(ItemC items, int nitems, int g) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int b0 = FIELD == 0 ? items[mid].sun_block0 : items[mid].flux_block0;
        if (b0 <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// The loop-invariant arguments of a phase body, passed through an empty asm
// before each tile (see the file comment).
struct Args {
    uint64_t items, model;
    int32_t nitems, step, t0, out_mask, halo, tile;
};

__device__ __forceinline__ void opaque(Args &a) {
    asm volatile("" : "+s"(a.items), "+s"(a.model), "+s"(a.nitems), "+s"(a.step), "+s"(a.t0), "+s"(a.out_mask),
                 "+s"(a.halo), "+s"(a.tile)::"memory");
}

__device__ __forceinline__ void sun_tile(Args a) {
    opaque(a);
    ItemC its = (ItemC)a.items;
    ItemC it = its + item_of<0>(its, a.nitems, a.tile);
    sun_step_body<true>((const float4 *)it->tile.hpack, (const float *)it->tile.walls,
                        (const float *)it->tile.aspect_rad, (const uint8_t *)it->tile.cls,
                        (const float *)it->tile.buildings, (const sw_step_t *)it->steps,
                        (const sw_model_t *)a.model, (const int4 *)it->sun_tab, (const uint32_t *)it->tile.stats,
                        (uint2 *)it->tile.src, a.step, it->tile.rows, it->tile.cols, a.tile - it->sun_block0);
}

template <bool WALK>
__device__ __forceinline__ void flux_tile(Args a) {
    opaque(a);
    ItemC its = (ItemC)a.items;
    ItemC it = its + item_of<1>(its, a.nitems, a.tile);
    const long n = (long)it->tile.rows * it->tile.cols;
    SW_GLOBAL float *out = it->out + (long)(a.step - a.t0) * __builtin_popcount((unsigned)a.out_mask) * n;
    flux_body<WALK>((const uint32_t *)it->tile.masks, (const uint2 *)it->tile.src, (const float *)it->tile.svfbuveg,
                    (const float *)it->tile.tan_asvf, (const float *)it->tile.tan_svfalfa,
                    (const uint32_t *)it->tile.walkface, (const float *)it->tile.buildings,
                    (const uint8_t *)it->tile.cls, (const sw_step_t *)it->steps, (const sw_model_t *)a.model,
                    (float *)it->tile.state, (float *)out, a.out_mask, a.step, WALK ? a.halo : -1, it->tile.rows,
                    it->tile.cols, a.tile - it->flux_block0);
    __syncthreads();   // the next tile reuses the walk tile in LDS
}

}  // namespace

// Steps [s0, s1) (s1 - s0 <= 64); day_bits bit (i - s0): step i is a day step;
// step i writes output slot i - out_t0 of each item. Grid: a multiple of 8
// blocks, all co-resident (cooperative launch). The tiles of a phase are split
// into 8 contiguous shares, one per XCD group (block b is in group b & 7, the
// XCD-aware order of the per-step kernels); a group's blocks draw its tiles
// from its counter.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
k_day_persistent(const BatchItemDev *__restrict__ items, int nitems, int total, const sw_model_t *__restrict__ model,
                 uint64_t day_bits, int s0, int s1, int out_t0, int out_mask, int halo, uint32_t *__restrict__ sync,
                 uint64_t limit) {
    const int per = (int)(gridDim.x >> 3), x = (int)(blockIdx.x & 7u);
    const uint32_t nx = (uint32_t)per;
    const int share = (total + 7) >> 3;
    const int c0 = min(x * share, total), cn = min((x + 1) * share, total) - c0;
    uint32_t *ctr = sync + kSyncTiles + 16 * x;
    uint32_t gen = 0, phase = 0;
    Args a;
    a.items = (uint64_t)items;
    a.model = (uint64_t)model;
    a.nitems = nitems;
    a.t0 = out_t0;
    a.out_mask = out_mask;
    a.halo = halo;
#define SW_TILE_LOOP(BODY)                                                                   \
    {                                                                                        \
        const uint32_t base = phase++ * (uint32_t)(cn + per);                                \
        for (int p = next_tile(ctr, base); p < cn; p = next_tile(ctr, base)) {               \
            a.tile = c0 + p;                                                                 \
            BODY;                                                                            \
        }                                                                                    \
    }
    for (int i = s0; i < s1; ++i) {
        a.step = i;
        if ((day_bits >> (i - s0)) & 1ull) {
            SW_TILE_LOOP(sun_tile(a))
            if (!grid_barrier(sync, ++gen, nx, limit)) return;
            SW_TILE_LOOP(flux_tile<true>(a))
        } else {
            SW_TILE_LOOP(flux_tile<false>(a))
        }
        if (!grid_barrier(sync, ++gen, nx, limit)) return;
        if (blockIdx.x == 0 && threadIdx.x == 0) sync[kSyncPhases] = gen;
    }
#undef SW_TILE_LOOP
    if (blockIdx.x == 0 && threadIdx.x == 0) sync[kSyncGrid] = gridDim.x;
}

// zero the barrier counters (and, with `all`, the error flag) before a launch;
// a kernel rather than hipMemsetAsync, so that a captured graph replays it
__global__ void k_sync_clear(uint32_t *sync, int all) {
    for (int i = (int)threadIdx.x; i < SW_DAY_SYNC_WORDS; i += (int)blockDim.x)
        if (all || i != kSyncErr) sync[i] = 0u;
}

extern "C" {

int sw_solve_batch_persistent(const sw_batch_item_t *items_dev, const sw_batch_item_t *items_host, int nitems,
                              const sw_model_t *model, const int32_t *is_day, int t0, int t1, int out_mask,
                              int walk_halo, int reset_state, uint32_t *sync, double timeout_s, void *stream) {
    const char *who = "sw_solve_batch_persistent";
    if (!items_dev || !items_host || nitems <= 0 || !model || !is_day || !sync)
        return swi_fmt_err(SW_E_ARG, "%s: NULL pointer or empty batch", who);
    if (t0 < 0 || t1 < t0) return swi_fmt_err(SW_E_ARG, "%s: invalid step range", who);
    if (out_mask <= 0 || out_mask > 255) return swi_fmt_err(SW_E_ARG, "%s: out_mask selects nothing", who);
    if (walk_halo < 0 || walk_halo > 48) return swi_fmt_err(SW_E_ARG, "%s: walk halo out of range", who);
    if (!(timeout_s > 0.0 && timeout_s <= 60.0)) return swi_fmt_err(SW_E_ARG, "%s: timeout outside (0, 60] s", who);
    long sun_total = 0, flux_total = 0;
    int rc = swi_check_batch(items_host, nitems, &sun_total, &flux_total, who);
    if (rc) return rc;
    if (sun_total != flux_total) return swi_fmt_err(SW_E_ARG, "%s: sun and flux grids differ", who);
    if (reset_state) {
        for (int k = 0; k < nitems; ++k) {
            const sw_tile_t &t = items_host[k].tile;
            rc = sw_fill_f32(t.state, 0.f, 6 * (int64_t)t.rows * t.cols, stream);
            if (rc) return rc;
        }
    }
    if (t1 == t0) return SW_OK;
    int dev = 0, cus = 0, per_cu = 0, khz = 0;
    rc = swi_hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (!rc) rc = swi_hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    if (!rc) rc = swi_hip_check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
    const size_t lds = flux_lds_bytes(walk_halo);
    if (!rc && lds > 65536)
        rc = swi_hip_check(hipFuncSetAttribute((const void *)k_day_persistent,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                           "hipFuncSetAttribute");
    if (!rc)
        rc = swi_hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_day_persistent, 256, lds),
                           "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (rc) return rc;
    long grid = (long)per_cu * cus;
    const long need = ((flux_total + 7) / 8) * 8;
    if (grid > need) grid = need;
    grid = (grid / 8) * 8;
    if (grid < 8) return swi_fmt_err(SW_E_ARG, "%s: fewer than 8 co-resident blocks", who);
    const uint64_t limit = (uint64_t)(timeout_s * (double)khz * 1e3);
    // steps in launches of at most 64 (the day/night pattern is one 64-bit argument); the first
    // launch clears the error flag, later ones keep it so that a time-out stops the rest
    for (int s0 = t0; s0 < t1; s0 += 64) {
        const int s1 = (t1 - s0 > 64) ? s0 + 64 : t1;
        uint64_t bits = 0;
        for (int i = s0; i < s1; ++i)
            if (is_day[i - t0]) bits |= 1ull << (i - s0);
        hipLaunchKernelGGL(k_sync_clear, dim3(1), dim3(64), 0, swi_stream(stream), sync, s0 == t0 ? 1 : 0);
        rc = swi_hip_check(hipGetLastError(), "k_sync_clear");
        if (rc) return rc;
        const BatchItemDev *it = (const BatchItemDev *)items_dev;
        int ni = nitems, tot = (int)flux_total, a0 = s0, a1 = s1, ot = t0, om = out_mask, hl = walk_halo;
        void *args[] = {(void *)&it, (void *)&ni, (void *)&tot, (void *)&model, (void *)&bits, (void *)&a0,
                        (void *)&a1, (void *)&ot, (void *)&om, (void *)&hl, (void *)&sync, (void *)&limit};
        // under stream capture: a plain launch of the same grid (a graph node), resident at once on a
        // device that runs nothing else; otherwise a cooperative launch, which checks that it is
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        rc = swi_hip_check(hipStreamIsCapturing(swi_stream(stream), &cap), "hipStreamIsCapturing");
        if (rc) return rc;
        if (cap == hipStreamCaptureStatusActive)
            rc = swi_hip_check(hipLaunchKernel((const void *)k_day_persistent, dim3((unsigned)grid), dim3(256), args,
                                               lds, swi_stream(stream)),
                               "hipLaunchKernel(k_day_persistent)");
        else
            rc = swi_hip_check(hipLaunchCooperativeKernel((const void *)k_day_persistent, dim3((unsigned)grid),
                                                          dim3(256), args, (unsigned)lds, swi_stream(stream)),
                               "hipLaunchCooperativeKernel(k_day_persistent)");
        if (rc) return rc;
    }
    return SW_OK;
}

}  // extern "C"
