// st_internal.h - device-side data structures and kernel launchers shared by
// the C-ABI implementation (st_api.hip) and the kernels (st_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "st_common.h"

namespace st {

// Rays of one trace job (a single st_trace call, one batched sun step or one
// served request). The trace kernels consume a list of jobs split into
// fixed-size chunks; every chunk belongs to exactly one job.
struct JobDev {
    const float4* prims;     // [n_prims][4] FP32 primitive records, leaf order
    const uint32_t* meta;    // [n_prims] element index | receiver | kind
    const float4* optics;    // [n_elements] (reflectivity, slope, spec, 0)
    const float4* nodes;     // [n_nodes][4]
    const SunFrame* sun;
    uint64_t ray_begin;      // first Philox ray index of this job's range
    uint64_t ray_count;
    uint64_t record_base;    // records are indexed by ray - record_base
    uint32_t seed_lo, seed_hi, job_id;
    int32_t max_depth;
    uint32_t flags;
    int32_t flux_slot;       // primitive slot of the flux element (-1 none)
    int32_t agg_slot;        // receiver whose hit count is kept in registers (-1 none)
    int32_t flux_nx, flux_ny;
    uint32_t order_head;     // trace order (set_trace_order): slots before the first whole block
    uint32_t order_blocks;   // whole blocks traced cell by cell (0: index order)
    uint32_t order_kind;     // kOrderHalton or kOrderStratified (with order_blocks > 0)
    uint32_t flux_bins;      // flux_nx * flux_ny
    uint32_t flux_copy_mask; // flux holds flux_copy_mask + 1 copies of the map; wave w adds to copy w & mask
    int64_t* flux;
    uint64_t* hits;          // [n_elements]
    uint64_t* stats;         // [ST_STAT_COUNT]
    float4* records;         // [rays][max_depth]
    uint32_t* rec_elem;      // [rays][max_depth]
    const uint32_t* grid_start;  // [kGridRes^2 + 1] sun-plane cell lists, or null (BVH for sun rays)
    const int32_t* grid_items;   // primitive slots of every cell
    const uint2* bounce_range;   // [n_elements] (begin, end) list of a first reflection off the element;
                                 // begin > end: none (BVH). Null: BVH for every reflected ray
    const int32_t* bounce_items; // primitive slots
    const uint2* pair_range;     // [bounce items] list of a second reflection whose first mirror's
                                 // list entry was that item (begin > end: none). Null: BVH
    const int32_t* pair_items;
};

// Halton sun positions: indices congruent modulo kHaltonCells = 2^8 * 3^5 share
// the first 8 binary digits of a and the first 5 ternary digits of b, so their
// sun-plane points lie in one grid column and two adjacent rows. A job's range
// is traced block by block of kHaltonBlock aligned indices, 64 rays of one such
// class per wave, when its indices (the sequence's 32-bit index) stay below 2^32.
constexpr uint32_t kHaltonCells = 256u * 243u, kHaltonBlock = 64u * kHaltonCells;
// Halton positions take the 32-bit ray index (the reference's unsigned launch
// index, cuda/sun.cu:94-134): jobs whose ray range passes 2^32 are rejected
constexpr uint64_t kHaltonMaxRays = 1ull << 32;
// Stratified positions: a sweep of kStratSweep rays gives each of the 65,536
// sun-plane cells one group of 64 rays, cells in a keyed Feistel order (so that
// every prefix of a sweep samples the plane evenly). A job's whole sweeps are
// traced in Hilbert-curve order of the cells instead, so that the groups a wave
// traces one after the other lie side by side and their reflected rays walk
// similar lists and nodes (jobs whose range starts on a group boundary).
constexpr uint64_t kStratSweep = 64ull << 16;
constexpr uint32_t kOrderHalton = 1, kOrderStratified = 2;
inline void set_trace_order(JobDev* J, int32_t position_sampling) {
    J->order_head = 0;
    J->order_blocks = 0;
    J->order_kind = 0;
    const uint64_t b = J->ray_begin, e = J->ray_begin + J->ray_count;
    uint64_t blk = 0;
    if (position_sampling == ST_POS_HALTON && e <= (1ull << 32)) blk = kHaltonBlock;
    if (position_sampling == ST_POS_STRATIFIED && b % 64 == 0) blk = kStratSweep;
    if (!blk) return;
    const uint64_t b1 = (b + blk - 1) / blk * blk, e1 = e / blk * blk;
    if (e1 <= b1 || (e1 - b1) / blk >= (1ull << 32)) return;
    J->order_head = (uint32_t)(b1 - b);
    J->order_blocks = (uint32_t)((e1 - b1) / blk);
    J->order_kind = blk == kHaltonBlock ? kOrderHalton : kOrderStratified;
}

// a grid is used unless its lists grow long: a wave whose rays share a cell
// tests its list in lockstep at ~40 vector instructions per entry, while a
// BVH walk costs each lane ~40 per node over ~20 nodes of the 10k field, so
// even lists of a few hundred entries beat the walk; the average bounds the
// grid's memory (kGridMaxAverage * kGridCells items)
constexpr uint32_t kGridMaxAverage = 16, kGridMaxLongest = 512;
// scenes of fewer elements trace sun rays through the BVH (a few nodes): the
// grid's passes over its 65,536 cells cost more than they save
constexpr int32_t kGridMinElements = 16;
__host__ __device__ inline bool grid_lists_pay(uint32_t total, uint32_t longest) {
    return total > 0 && total <= kGridMaxAverage * (uint32_t)kGridCells && longest <= kGridMaxLongest;
}

struct LaunchShape {
    int block = 256;     // threads per block (multiple of 64)
    int grid = 0;        // persistent blocks (0 -> 8 per CU)
    int chunk = 2048;    // rays per chunk (multiple of 64)
    int kernel = ST_KERNEL_MEGA;
    int stack = 32;      // traversal stack entries per lane in LDS (16 or 32, >= BVH depth; 0: kStackDepth in private memory)
    int waves = 4;       // waves per SIMD the register allocation is built for
    bool queued = false; // continuing rays parked in LDS, traced in full waves
    bool lds_hits = false;  // hit counts of elements below kLdsHitElements kept per wave in
                            // LDS (launches with a scene that small; queued, block 256, stack 8 or
                            // 16 at 5 waves/SIMD, or stack 0 at 6 or 7 for listed launches)
    bool group_bvh = false; // queued form tracing list and BVH segments in separate steps
                            // (use_grouped_bvh; compiled for stack 16 at 5 waves/SIMD and stack 0 at 6, 7 or 8)
};
// Scenes this small count element hits in LDS: their few counters would
// otherwise take a device-scope atomic from nearly every wave of the launch.
constexpr int kLdsHitElements = 128;
// a compiled trace variant: form 0 whole rays, 1 queued, 2 queued with list and
// BVH segments in separate steps
bool trace_variant_compiled(int block, int stack, int waves, int form, bool lds_hits);
bool trace_variant_exists(int block, int stack, int waves, bool queued);
// traversal stack entries of a shape's stack setting
inline int stack_entries(int stack) { return stack > 0 ? stack : kStackDepth; }
bool grouped_variant_exists(int block, int stack, int waves);
// A launch of jobs whose reflected rays partly have candidate lists and partly
// walk the BVH (any launch with first-reflection lists: segments after a
// second mirror, and those without a second-reflection list, walk it) traces
// the two kinds of segment in separate steps when its shape allows it.
inline void use_grouped_bvh(LaunchShape& sh, bool mixed) {
    sh.group_bvh = mixed && sh.queued && trace_variant_compiled(sh.block, sh.stack, sh.waves, 2, sh.lds_hits);
}
// waves of this trace variant the device holds at once (0 if unknown)
int trace_resident_waves(const LaunchShape& shape);
constexpr bool kDefaultQueued = true;  // form used when the caller and the tuner leave it open

// ---- launchers (all asynchronous on `stream`) -----------------------------
// frames + packing for one scene: elements in element order -> prims in slot order
hipError_t launch_frames(hipStream_t s, const st_element_f64* elems, int32_t n, const int32_t* elem_slot,
                         const double* aim_override /* n*3 or null */, const double* zrot_override /* n or null */,
                         double* frames, double* ref_aabb, float* bounds_slot, float4* prims, uint32_t* meta,
                         float4* optics);
// FP64 bisector aim of all mirrors (in place on the device element array)
hipError_t launch_aim(hipStream_t s, st_element_f64* elems, int32_t n, const double* sun_unit, const double* target);
// statistics (ST_STAT_COUNT u64) and sun frame of each of `count` jobs, gathered
// into stats_out[count][ST_STAT_COUNT] and frames_out[count] (one launch per 32
// jobs instead of two copies per job)
hipError_t launch_round_results(hipStream_t s, const uint64_t* const* stats, const SunFrame* const* suns, int32_t count,
                                uint64_t* stats_out, SunFrame* frames_out);
// the same into a copy: dst = src with every mirror re-aimed (one pass instead
// of a device copy and an in-place aim)
hipError_t launch_aim_copy(hipStream_t s, const st_element_f64* src, st_element_f64* dst, int32_t n,
                           const double* sun_unit, const double* target);
// batched: per step k re-aim mirrors along the bisector of sun[k] and target
hipError_t launch_frames_batch(hipStream_t s, const st_element_f64* elems, int32_t n, const int32_t* elem_slot,
                               int32_t steps, const double* sun_units /* steps*3 */, const double* target,
                               double* ref_aabb /* steps*n*6 */, float* bounds_slot /* steps*n*6 */,
                               float4* prims /* steps*n*4 */, double* frames /* steps*n*9 or null */);
// refit inner nodes level by level (deepest first); nodes per step are
// contiguous blocks of n_nodes*4 float4, bounds per step n*6 floats
hipError_t launch_refit(hipStream_t s, const int32_t* child, const int32_t* level_nodes,
                        const int32_t* level_offsets_host, int32_t levels, int32_t n_nodes, int32_t n_prims,
                        int32_t steps, const float* bounds_slot, float4* nodes);
hipError_t launch_sun_plane(hipStream_t s, const double* ref_aabb, int32_t n, int32_t steps,
                            const double* sun_units /* steps*3 */, const st_sun_desc* sun, SunFrame* out);
// jobs == nullptr traces `single` (passed by value); work_counter: 2 x u32,
// zero before the first launch, left zero by every launch
hipError_t launch_trace(hipStream_t s, const JobDev* jobs, const uint64_t* chunk_prefix, int32_t n_jobs,
                        uint64_t total_chunks, uint32_t* work_counter, const LaunchShape& shape,
                        const JobDev& single);
// sun-plane candidate grids of `steps` scene states (bounds per step n*6
// floats, one SunFrame per step): counts (zeroed, steps * kGridCells u32) ->
// start (steps * (kGridCells + 1)), stats[2k] = total items, stats[2k+1] =
// longest cell list. With jobs, job k's grid pointers are set (items at
// items + k * item_cap) when grid_lists_pay and the total fits item_cap, else
// cleared.
// prims: the steps' primitive records (n * 4 float4 per step), meta shared.
hipError_t launch_grid_count(hipStream_t s, const float* bounds, const float4* prims, const uint32_t* meta, int32_t n,
                             int32_t steps, const SunFrame* sun, double tan_max, uint32_t* counts, uint32_t* start,
                             uint32_t* stats, JobDev* jobs, const int32_t* items, uint32_t item_cap);
// cursor zeroed (steps * kGridCells u32); steps whose total exceeds item_cap are skipped
hipError_t launch_grid_fill(hipStream_t s, const float* bounds, const float4* prims, const uint32_t* meta, int32_t n,
                            int32_t steps, const SunFrame* sun, double tan_max, const uint32_t* start, uint32_t* cursor,
                            int32_t* items, uint32_t item_cap);
// candidate lists of reflected rays (k_cone_lists). Task t belongs to scene
// state k = t / tasks_per_step (frames n*9, bounds n*6, nodes n_nodes*4 and one
// SunFrame per state) and is, with i = t % tasks_per_step: first reflections
// off element i (pair_items null), or second reflections of rays that left
// element pair_owner[t] and hit primitive pair_items[t]. The list goes to
// scratch[t * kMaxBounceList ...]; counts[t] gets its length (kNoBounceList:
// none) and, when range is given, range[t] = (t * kMaxBounceList, + length) or
// (1, 0) for none, so the scratch serves as the item array directly. First-
// reflection tasks run in BVH leaf order (their outputs stay at t).
// sup_range / sup_items: candidate lists (per element) holding every primitive
// an element's first reflections reach, whose boxes first-reflection tasks
// test instead of walking the BVH.
// launch_list_compact packs counted lists at items[range[t].x ...].
constexpr uint32_t kNoBounceList = 0xFFFFFFFFu;
constexpr uint32_t kMaxBounceList = 64;
hipError_t launch_cone_lists(hipStream_t s, const st_element_f64* elems, int32_t tasks, int32_t tasks_per_step,
                             int32_t n, int32_t n_nodes, const int32_t* elem_slot, const uint32_t* meta,
                             const double* frames, const float* bounds_slot, const float4* nodes, const SunFrame* sun,
                             double theta_sun, const int32_t* pair_items, const int32_t* pair_owner, uint32_t* counts,
                             uint2* range, int32_t* scratch, const uint2* sup_range = nullptr,
                             const int32_t* sup_items = nullptr);
// first-reflection candidates of every mirror over all bisector re-aims at
// target (k_reaim_lists): range[e] / scratch[e * kMaxBounceList ...] as
// k_cone_lists writes them, by element; spheres: n double4 of scratch
hipError_t launch_reaim_lists(hipStream_t s, const st_element_f64* elems, int32_t n, const int32_t* elem_slot,
                              const uint32_t* meta, const double* target, double theta_sun, double4* spheres,
                              uint2* range, int32_t* scratch);
hipError_t launch_list_compact(hipStream_t s, int32_t tasks, const uint2* range, const int32_t* scratch,
                               int32_t* items);
hipError_t launch_fill_u64(hipStream_t s, void* p, uint64_t n_u64, uint64_t value);
// A small flux map's bins are few and its hot spot fewer: atomics of every
// wave on them would queue at the same addresses. A map of fewer than
// kFluxDirectBins bins is accumulated in flux_copies(bins) copies (a power of
// two, so that the copies span at least kFluxSpreadBins bins, at most
// kFluxMaxCopies of them), which launch_flux_gather adds into the output
// [jobs][bins] (copies [jobs][copies][bins]) and clears for the next launch.
// Larger maps spread their atomics by themselves (copies measured no gain), and
// launches of fewer than kFluxCopyMinRays rays queue too few atomics to pay for
// the gather.
constexpr uint64_t kFluxDirectBins = 1u << 15, kFluxSpreadBins = 1u << 17, kFluxCopyMinRays = 1u << 20;
constexpr uint32_t kFluxMaxCopies = 64;
inline uint32_t flux_copies(uint64_t bins, uint64_t rays) {
    uint32_t c = 1;
    if (rays < kFluxCopyMinRays) return c;
    while (bins && bins < kFluxDirectBins && c < kFluxMaxCopies && c * bins < kFluxSpreadBins) c <<= 1;
    return c;
}
hipError_t launch_flux_gather(hipStream_t s, int64_t* copies, uint32_t n_copies, uint64_t bins, int32_t jobs,
                              int64_t* out);
hipError_t launch_flux_watts(hipStream_t s, const int64_t* flux, int32_t nbins, const SunFrame* sun,
                             double num_rays, double bin_area, double* out, double* power_per_ray_dev);

int device_cu_count();
// one launch of the VALU calibration kernel over `blocks` blocks of 256 lanes
// (out: blocks * 256 floats); *lane_ops = FMAs it issues
hipError_t launch_valu_calibrate(hipStream_t s, float* out, int blocks, double* lane_ops);

}  // namespace st
