// st_api_internal.h - handle structs behind the opaque C-ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <vector>

#include "st_bvh.h"
#include "st_internal.h"

namespace st {

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
int dev_alloc(void** p, size_t bytes, const char* what);
bool launch_injected();

// owning device allocation
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset();
    int ensure(size_t n, const char* what);  // (re)allocate to at least n bytes
};

// launch shape for `rays` rays; block/grid/waves > 0 override, else the scene's
// tuned shape when there is one, else defaults: the queued form with the
// traversal stack in private memory when the BVH fits it, at 7 waves per SIMD
// (72 registers; the LDS-stack variants hold 5, the 8-wave variants' 64
// registers spill inside the loops), on as many blocks as the device holds at
// once. bvh_depth: the deepest scene in the launch; min_elements: the smallest
// scene in the launch (when it has at most kLdsHitElements elements, the
// launch counts element hits below that index per wave in LDS: a small scene's
// few counters would otherwise take every wave's device atomics). An explicit
// waves-per-SIMD budget takes the smallest compiled stack that holds the BVH;
// the caller checks that the variant exists (trace_variant_exists).
LaunchShape pick_shape(uint64_t rays, int32_t block, int32_t grid, int32_t kernel, int32_t bvh_depth,
                       const LaunchShape* tuned, int32_t min_elements, int32_t waves = 0);

// First-reflection candidates of a scene's mirrors over every bisector re-aim
// at a target (k_reaim_lists): while every mirror is aimed by the bisector rule
// at that target for the sun being traced, each mirror's first reflections
// point at the target, and the scene state's first-reflection lists test these
// candidates instead of walking the BVH. Kept for one (target, sun cone angle,
// element geometry) key.
struct ReaimCache {
    DevBuf spheres, range, items, target_dev;
    bool valid = false;
    double target[3] = {0, 0, 0};
    double theta = -1.0;
    uint64_t epoch = 0;
};
// the candidates of scene s for `target` and sun cone angle theta (rebuilt
// when the key changed), asynchronous on stream st
int ensure_reaim(ReaimCache& c, st_scene s, const double target[3], double theta, hipStream_t st);

struct BatchCache {
    bool valid = false;
    int32_t steps = 0, max_depth = 0, flux_element = 0, flux_nx = 0, flux_ny = 0, block = 0, grid = 0;
    uint64_t seed = 0, nbins = 0, total_chunks = 0;
    uint32_t flags = 0;
    void* stream = nullptr;
    st_sun_desc sun{};
    double target[3] = {0, 0, 0};
    std::vector<double> suns;
    std::vector<uint64_t> budgets;
    int64_t* flux_out = nullptr;
    uint64_t* hits_out = nullptr;
    uint64_t* stats_out = nullptr;
    LaunchShape shape;
    DevBuf d_sun_units, d_target, d_ref_aabb, d_bounds, d_prims, d_nodes, d_sun, d_jobs, d_prefix, d_counter;
    // per-step sun-plane grids (items of step k at d_grid_items + k * grid_cap)
    // and first-reflection lists (range [steps][n], items in the per-task
    // scratch), rebuilt inside every run of the batch
    bool grids = false, lists = false;
    uint32_t grid_cap = 0;
    DevBuf d_frames, d_grid_counts, d_grid_start, d_grid_cursor, d_grid_stats, d_grid_items, d_bounce_range,
        d_bounce_scratch;
    // first-reflection candidates over every re-aim at the batch's target (the
    // reflection axis of a re-aimed mirror points at the target whatever the
    // sun), rebuilt inside every run; each step's lists test these instead of
    // walking the BVH
    DevBuf d_sup_spheres, d_sup_range, d_sup_scratch;
    // per-step copies of small flux maps (route_flux_copies)
    DevBuf d_flux_copies;
    uint32_t flux_copies = 1;
};
// device bytes a batch may spend on its per-step grids and reflection lists
constexpr uint64_t kBatchListBytes = 3ull << 30;
constexpr uint64_t kFluxCopyBytes = 256ull << 20;  // bound on the flux-map copies of one launch

// Routes the flux of jobs[0, n) (maps of `bins` bins; J.flux = the output, or
// null; `rays` in the launch) to flux_copies(bins, rays) copies per job in buf
// (grown and cleared as needed; fewer copies when they would exceed
// kFluxCopyBytes). *copies = the number used (1: none, jobs unchanged);
// launch_flux_gather(buf, *copies, bins, n, outputs) after the trace adds them
// into the outputs.
int route_flux_copies(DevBuf& buf, JobDev* jobs, int32_t n, uint64_t bins, uint64_t rays, hipStream_t st,
                      uint32_t* copies);

}  // namespace st

// A context stays alive while scenes or servers created from it exist:
// st_context_destroy drops the caller's reference, the last owner frees it.
struct st_context_s {
    int device = 0;
    hipStream_t stream = nullptr;  // scene set-up work (sun planes, grids, lists, refits)
    hipEvent_t setup_done = nullptr;  // orders a caller's trace stream after queued set-up work
    std::mutex order_mu;
    std::atomic<int> refs{1};
};

namespace st {
constexpr uint64_t kListMinRays = 1ull << 26;      // traced rays before the first reflection-list build
constexpr double kListSavingPerRay = 2.0e-11;      // seconds a list saves per traced ray (10k field)
void context_retain(st_context ctx);
void context_release(st_context ctx);
// A call that traces on a caller's stream (not the context's) first makes that
// stream wait for the set-up work still queued on the context stream (sun
// plane, grid, lists of the scene it reads): an event, no host wait; skipped
// while the caller's stream is being captured into a graph (the caller orders
// such work). The other direction is the caller's: set-up calls on a scene
// (sun, aim, elements, lists) must not overlap a trace of it still running on
// the caller's stream.
int order_before_trace(st_context ctx, hipStream_t user);
}  // namespace st

namespace st {
// context reference held by a scene or server; declared first so it is
// released after every device buffer of its owner has been freed
struct ContextRef {
    st_context ctx = nullptr;
    ContextRef() = default;
    ContextRef(const ContextRef&) = delete;
    ContextRef& operator=(const ContextRef&) = delete;
    void hold(st_context c) {
        context_retain(c);
        ctx = c;
    }
    ~ContextRef() {
        if (ctx) context_release(ctx);
    }
};
}  // namespace st

struct st_scene_s {
    st::ContextRef ctx_ref;
    st_context ctx = nullptr;
    int32_t n = 0, n_nodes = 0;
    std::vector<st_element_f64> host_elems;
    st::BvhTopology topo;
    std::vector<int32_t> level_offsets;  // host: nodes of level l are level_nodes[off[l], off[l+1])
    st::DevBuf d_elems, d_elem_slot, d_frames, d_ref_aabb, d_bounds, d_prims, d_meta, d_optics, d_nodes, d_child,
        d_level_nodes, d_sun, d_sun_unit, d_aim, d_zrot, d_counter;
    // sun-plane candidate grid (rebuilt with the sun plane; unused when its
    // lists would be long, e.g. at grazing sun elevations)
    st::DevBuf d_grid_counts, d_grid_start, d_grid_cursor, d_grid_items, d_grid_stats;
    st::DevBuf d_flux_copies;  // copies of a small flux map (route_flux_copies)
    // every mirror aimed by the bisector rule at aim_target for the sun aim_sun
    // (st_scene_set_sun_and_aim, st_scene_aim_heliostats); geometry changes
    // (element updates) bump geom_epoch
    bool aim_bisector = false;
    double aim_target[3] = {0, 0, 0}, aim_sun[3] = {0, 0, 0};
    uint64_t geom_epoch = 0;
    st::ReaimCache reaim;
    bool grid_valid = false;
    uint32_t grid_items = 0, grid_longest = 0;
    // candidate lists of first reflections (per element) and second
    // reflections (per first-reflection list entry), rebuilt with the sun plane
    st::DevBuf d_bounce_counts, d_bounce_range, d_bounce_items, d_pair_counts, d_pair_range, d_pair_items,
        d_pair_owner;
    std::vector<uint32_t> h_counts;
    std::vector<uint2> h_bounce_range, h_pair_range;
    std::vector<int32_t> h_pair_owner;
    bool bounce_valid = false, pair_valid = false;
    double pair_missing = 1.0;  // share of first-reflection list entries without a second-reflection list
    uint64_t list_builds = 0, tuned_list_build = ~0ull;  // the autotuner's choice of grouping holds for
                                                         // the lists (scene state) it was tuned on
    bool lists_stale = true;         // scene state changed since the lists were built
    uint64_t rays_since_change = 0;  // rays traced on the current state without lists
    uint64_t list_min_rays = st::kListMinRays;  // traced rays that repay a rebuild
    st::DevBuf d_list_scratch;
    uint64_t bounce_items = 0, pair_items = 0;
    bool sun_set = false;
    st_sun_desc sun{};
    double sun_unit[3] = {0, 0, 1};
    st_sun_plane plane{};
    st::BatchCache batch;
    hipGraphExec_t batch_graph_exec = nullptr;
    bool tuned_valid = false;
    st::LaunchShape tuned;
    // the scene's last st_trace: its stream and an event after it (traces on
    // another stream wait for it: they share d_counter and d_flux_copies), and
    // the variant it launched (st_scene_last_launch)
    hipEvent_t trace_done = nullptr;
    hipStream_t trace_stream = nullptr;
    st_launch_info last_launch{};
    ~st_scene_s();
};

namespace st {
// Largest angle between a sampled sun ray and the sun axis. Pillbox: the half
// angle. Gaussian: sigma * sqrt(-2 ln u) with u >= 2^-24 (the smallest
// uniform) = 5.768 sigma. Widened for FP32 direction rounding.
inline double sun_cone_angle(const st_sun_desc& sun) {
    const double th = sun.shape == ST_SUN_GAUSSIAN ? 5.7686 * sun.param : sun.param;
    return th * 1.001 + 1e-6;
}
int scene_refresh(st_scene s, bool use_overrides);
int scene_sun_plane(st_scene s);  // also rebuilds the sun-plane candidate grid
int scene_grid(st_scene s);
int scene_bounce_lists(st_scene s);
int scene_ensure_lists(st_scene s, uint64_t rays, bool force);
int resolve_flux_element(st_scene s, int32_t requested, int32_t nx, int32_t ny, int32_t* slot);
int fill_job(st_scene s, const st_trace_desc* d, const st_outputs* out, const void* prims, const void* nodes,
             const void* sun, JobDev* J);
int validate_trace(st_scene s, const st_trace_desc* d, const st_outputs* out);
int order_scene_traces(st_scene s, hipStream_t st);
int mark_scene_trace(st_scene s, hipStream_t st);
st_launch_info launch_info(const LaunchShape& sh, uint64_t chunks, uint32_t copies);
int trace_shape(st_scene s, const st_trace_desc* d, const JobDev& J, LaunchShape* out, uint64_t* chunks);
}  // namespace st
