/*
 * soltrace_hip.h - C ABI of the MI355X-native SolTrace-style ray tracer.
 *
 * Plain C, plain pointers and sizes. Every entry point returns an int status
 * (ST_OK == 0, negative on error) and records a message retrievable with
 * st_last_error(). No exceptions cross this boundary; the C++ drop-in layer
 * (SolTraceSystem & co.) rethrows std::runtime_error like the reference's
 * CUDA_CHECK / OPTIX_CHECK macros (lib/util_check.hpp:40-63, 93-117).
 *
 * What each group replaces in the reference (uwsbel/optixSoltraceDemos):
 *   st_context_*      OptiX device-context creation, SolTraceSystem ctor
 *                     (lib/soltrace_system.cpp:49-70) and clean_up (:267-297).
 *   st_scene_*        GeometryManager::collect_geometry_info / create_geometries
 *                     / update_geometry_info (lib/geometry_manager.cpp:10-94,
 *                     150-220, 223-265), dataManager::allocate/update
 *                     GeometryDataArray (lib/data_manager.cpp:44-68), the FP64
 *                     element frames (lib/element.cpp:77-164, lib/math_util.h:21-59)
 *                     and the sun-plane kernels (lib/sun_utils.cu:53-229,
 *                     lib/geometry_manager.cpp:97-148).
 *   st_trace          optixLaunch of __raygen__sun_source + intersection /
 *                     closest-hit programs (lib/soltrace_system.cpp:145-170;
 *                     cuda/sun.cu:94-134, cuda/intersection.cu, cuda/materials.cu),
 *                     plus the offline flux binning of
 *                     scripts/post_processing/flux_map_calculator.py:95-179.
 *   st_trace_batch    the per-frame run()/update() loop
 *                     (lib/soltrace_system.cpp:172-184, demos/demo_dynamic_scene.cpp:82-97)
 *                     folded into one device-side sweep.
 *   st_server_*       no reference counterpart (the reference is one blocking
 *                     launch at a time); continuous-batching trace service.
 */
#ifndef SOLTRACE_HIP_H
#define SOLTRACE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ST_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define ST_OK 0
#define ST_ERR_INVALID -1     /* bad argument (null, out of range, bad enum) */
#define ST_ERR_DEVICE -2      /* HIP runtime error / no device               */
#define ST_ERR_ALLOC -3       /* device or host allocation failed            */
#define ST_ERR_STATE -4       /* call out of order (e.g. trace before sun)   */
#define ST_ERR_REJECTED -5    /* server admission rejected (HBM footprint)   */
#define ST_ERR_CANCELLED -6   /* server job cancelled                        */
#define ST_ERR_INJECTED -7    /* fault injected through st_debug_inject_fault */

/* ---- enums (values follow lib/soltrace_type.h order) -------------------- */
enum st_surface_type {
    ST_SURFACE_FLAT = 0,      /* SurfaceType::FLAT      */
    ST_SURFACE_PARABOLIC = 1, /* SurfaceType::PARABOLIC */
    ST_SURFACE_CYLINDER = 3   /* SurfaceType::CYLINDER (capped, axis = local y) */
};
enum st_aperture_type {
    ST_APERTURE_RECTANGLE = 0, /* ApertureType::RECTANGLE */
    ST_APERTURE_CIRCLE = 1     /* ApertureType::CIRCLE: zero AABB in the reference,
                                  rejected here (no intersection program exists) */
};
enum st_role {
    ST_ROLE_MIRROR = 0,   /* reflects (cuda/materials.cu:26-84, 170-231)        */
    ST_ROLE_RECEIVER = 1  /* records & absorbs (cuda/materials.cu:86-163)       */
};
enum st_sun_shape {
    ST_SUN_PILLBOX = 0,   /* uniform over a cone of half-angle `param` rad (cuda/sun.cu:72-91) */
    ST_SUN_GAUSSIAN = 1   /* isotropic 2-D Gaussian, per-axis sigma `param` rad */
};
enum st_position_sampling {
    ST_POS_PHILOX = 0,    /* sun-plane point from Philox uniforms (default)   */
    ST_POS_HALTON = 1,    /* Halton(2,3) of the ray index, as cuda/sun.cu:16-38 */
    ST_POS_STRATIFIED = 2 /* 256 x 256 cells of the sun plane; every 64 consecutive rays
                             (one wave) share a cell, each sweep of 65,536 groups visits
                             every cell once in a seeded Feistel order; uniform within
                             the cell from the ray's own Philox draw */
};

/* trace flags */
#define ST_TRACE_RECORDS 0x1u          /* write reference hit records (float4 per slot) */
#define ST_TRACE_FLUX_DIRECT_SUN 0x2u  /* bin receiver hits that came straight from the sun
                                          (reference CSV keeps them; SolTrace drops them,
                                          scripts/post_processing/clean_direct_sun.py:6-16) */
#define ST_TRACE_ZERO_OUTPUTS 0x4u     /* zero flux / element_hits / stats before tracing */
#define ST_TRACE_BVH_ONLY 0x8u         /* find every hit by BVH traversal instead of the
                                          candidate lists (identical results) */
#define ST_TRACE_COUNT_WORK 0x10u      /* also count node visits and primitive tests
                                          (stats 11-14; slows the trace down) */
#define ST_JOB_KEEP_AIM 0x100u         /* server job: trace the scene's stored aims instead of
                                          re-aiming every mirror at the job's sun (bisector,
                                          target = centre of the scene's last receiver) */

/* trace kernels */
#define ST_KERNEL_AUTO 0
#define ST_KERNEL_MEGA 1       /* one lane per ray, all bounces in registers        */
#define ST_KERNEL_QUEUED 2     /* sun rays in full waves; continuing rays parked in  
                                  LDS and traced in full waves of 64                */
#define ST_KERNEL_QUEUED_GROUPED 3  /* the queued form tracing segments with a
                                  candidate list and segments that walk the BVH in
                                  separate steps (AUTO picks it where reflected
                                  rays mix the two)                                 */

/* statistics counters (uint64 each) */
enum st_stat {
    ST_STAT_RAYS = 0,          /* sun rays launched                                  */
    ST_STAT_SEGMENTS = 1,      /* closest-hit queries (ray segments traced)          */
    ST_STAT_MIRROR_HITS = 2,   /* recorded mirror interactions (tag 1)               */
    ST_STAT_RECEIVER_HITS = 3, /* recorded receiver interactions (tag 2), all        */
    ST_STAT_DIRECT_SUN = 4,    /* ... of which the receiver was the first thing hit  */
    ST_STAT_ABSORBED = 5,      /* absorbed on a mirror by the reflectivity draw      */
    ST_STAT_ESCAPED = 6,       /* left the scene (miss program)                      */
    ST_STAT_BACK_FACE = 7,     /* stopped on a flat receiver's back face             */
    ST_STAT_TRUNCATED = 8,     /* hit something at the depth cap (not recorded)      */
    ST_STAT_CAP_HITS = 9,      /* receiver hits on a cylinder cap                    */
    ST_STAT_FLUX_BINNED = 10,  /* hits added to the flux map                         */
    /* work counters, filled only with ST_TRACE_COUNT_WORK                             */
    ST_STAT_NODE_VISITS = 11,  /* BVH inner nodes visited (two child-box tests each)  */
    ST_STAT_FLAT_TESTS = 12,   /* flat-primitive intersection tests                  */
    ST_STAT_PARABOLIC_TESTS = 13, /* parabolic-primitive intersection tests          */
    ST_STAT_CYLINDER_TESTS = 14,  /* cylinder intersection tests                      */
    ST_STAT_COUNT = 16
};

/* ---- POD inputs --------------------------------------------------------- */

/* One optical element in FP64, as the reference's Element
 * (lib/element.h:41-102) + Surface (lib/Surface.h) + Aperture (lib/Aperture.h),
 * plus optics the reference parses but drops (lib/soltrace_system.cpp:576-590). */
typedef struct st_element_f64 {
    double origin[3];
    double aim_point[3];
    double zrot_deg;        /* Element::m_zrot, degrees                          */
    double dim_x, dim_y;    /* aperture width / height (cylinder: 2r, length)    */
    double curv_x, curv_y;  /* SurfaceParabolic c1, c2 (1/m)                     */
    double reflectivity;    /* 1.0 = every hit reflects                          */
    double slope_error;     /* rad, per-axis Gaussian sigma of the normal        */
    double spec_error;      /* rad, per-axis Gaussian sigma of the reflected ray */
    int32_t surface;        /* st_surface_type                                   */
    int32_t aperture;       /* st_aperture_type                                  */
    int32_t role;           /* st_role                                           */
    int32_t reserved;
} st_element_f64;

typedef struct st_sun_desc {
    double vector[3];       /* points TO the sun; normalised inside              */
    double param;           /* pillbox half-angle or gaussian sigma, rad          */
    double dni;             /* W/m^2 for power normalisation (0 -> 1000)          */
    int32_t shape;          /* st_sun_shape                                       */
    int32_t position_sampling; /* st_position_sampling                            */
} st_sun_desc;

/* Sun plane (parallelogram rays are launched from), FP64, as LaunchParams
 * sun_v0..v3 (cuda/Soltrace.h:43-47). */
typedef struct st_sun_plane {
    double v0[3], v1[3], v2[3], v3[3];
    double distance;        /* plane distance along the sun vector               */
    double area;            /* |v1-v0| * |v3-v0|, m^2                            */
    double box_half_angle;  /* angle used to pad the box (pillbox a, gaussian 4s)*/
} st_sun_plane;

typedef struct st_trace_desc {
    uint64_t num_rays;      /* rays in the whole job (sets power per ray)        */
    uint64_t ray_begin;     /* first ray index traced by this call               */
    uint64_t ray_count;     /* rays traced by this call (0 -> num_rays-ray_begin) */
    uint64_t seed;          /* Philox key                                         */
    uint32_t job_id;        /* Philox counter word 3                              */
    int32_t max_depth;      /* reference MAX_TRACE_DEPTH (cuda/Soltrace.h:11) = 5 */
    uint32_t flags;         /* ST_TRACE_*                                         */
    int32_t kernel;         /* ST_KERNEL_*                                        */
    int32_t flux_element;   /* element binned into the flux map; -1 = last receiver */
    int32_t flux_nx, flux_ny; /* flux grid; 0 -> no flux map                      */
    int32_t block_size;     /* 0 -> tuned default                                 */
    int32_t grid_blocks;    /* 0 -> tuned default (persistent kernel; untuned: the
                               blocks the device holds at once)                   */
    int32_t waves_per_simd; /* 0 -> tuned default (untuned queued forms: 7 with the
                               traversal stack in private memory); else the register
                               budget of the kernel variant to run (queued forms: 3,
                               4 or 5 with the stack in LDS, 6, 7 or 8 with it in
                               private memory; whole-ray form: 4, 5, 6 or 8); a block
                               size and budget without a compiled variant is an error */
    void* stream;           /* hipStream_t; NULL = the scene's own stream          */
} st_trace_desc;

/* Device output buffers, caller-owned (any of them may be NULL).
 * Accumulated with atomics unless ST_TRACE_ZERO_OUTPUTS is set. */
typedef struct st_outputs {
    int64_t* flux;            /* [flux_ny][flux_nx] hit counts. Flat receiver: x runs
                                 along -x_b and y along y_b of the element frame (the
                                 parallelogram v1 = -W x_b, v2 = H y_b of
                                 lib/element.cpp:136-139 that the flux scripts bin),
                                 ix = floor((W/2 - x) nx / W). Cylinder: x is
                                 theta = atan2(-z_local, x_local) mod 2 pi, y along
                                 the axis (scripts/post_processing/
                                 plot_flux_map_gpu.py:185-206)                     */
    uint64_t* element_hits;   /* [n_elements] recorded interactions per element   */
    uint64_t* stats;          /* [ST_STAT_COUNT]                                   */
    float* records;           /* [ray_count][max_depth][4] = (tag,x,y,z), tag 0 sun,
                                 1 mirror, 2 receiver, all-zero = empty slot
                                 (hit_point_buffer layout, lib/soltrace_system.cpp:118-125) */
    uint32_t* record_elements;/* [ray_count][max_depth] element index, 0xFFFFFFFF empty,
                                 0xFFFFFFFE sun origin                               */
} st_outputs;

typedef struct st_context_s* st_context;
typedef struct st_scene_s* st_scene;
typedef struct st_server_s* st_server;

/* ---- library / context ------------------------------------------------- */
int st_abi_version(void);
const char* st_last_error(void);            /* thread-local message of the last failure */
int st_device_count(int* count);
/* Free and total device memory of a GPU (either pointer may be NULL); the
 * drop-in's memory report around run() takes free memory before and after,
 * as the reference does with cudaMemGetInfo (lib/soltrace_system.cpp:77, 151). */
int st_device_memory(int device, size_t* free_bytes, size_t* total_bytes);
int st_context_create(int device, st_context* out);
int st_context_destroy(st_context ctx);

/* Device buffers for hosts without HIP headers (the C++ drop-in layer owns
 * hit_point_buffer this way, lib/soltrace_system.cpp:118-125). */
int st_buffer_alloc(st_context ctx, size_t bytes, void** dev_ptr);
int st_buffer_free(st_context ctx, void* dev_ptr);
int st_buffer_upload(st_context ctx, void* dev_dst, const void* host_src, size_t bytes);
int st_buffer_download(st_context ctx, void* host_dst, const void* dev_src, size_t bytes);
int st_buffer_zero(st_context ctx, void* dev_ptr, size_t bytes);

/* ---- scenes ------------------------------------------------------------ */
/* Upload FP64 elements, compute FP64 frames on the GPU, pack FP32 primitives,
 * build the BVH (host binned SAH over the FP64 bounds) and refit it on the GPU. */
int st_scene_create(st_context ctx, const st_element_f64* elements, int32_t n, st_scene* out);
int st_scene_destroy(st_scene scene);
int st_scene_num_elements(st_scene scene, int32_t* n);
/* Replace aim points / zrot of all elements (FP64, host arrays of n*3 and n),
 * recompute frames on the GPU and refit the BVH. NULL zrot keeps the old ones. */
int st_scene_update_aim(st_scene scene, const double* aim_points, const double* zrot_deg);
/* Replace the element set (same count and types) and refit. */
int st_scene_update_elements(st_scene scene, const st_element_f64* elements, int32_t n);
/* New element poses (same count and types; NULL keeps them) and a new sun
 * (NULL keeps it) in one set-up: frames, refit, then one sun plane and grid.
 * The per-frame SolTraceSystem::update() (lib/soltrace_system.cpp:172-184). */
int st_scene_update(st_scene scene, const st_element_f64* elements, int32_t n, const st_sun_desc* sun);
/* Re-aim every mirror at `target` along the bisector of the sun vector and the
 * direction to target (FP64, on the GPU), then refit. */
int st_scene_aim_heliostats(st_scene scene, const double sun_vector[3], const double target[3]);
/* Set the sun and recompute the sun plane on the GPU (FP64). */
int st_scene_set_sun(st_scene scene, const st_sun_desc* sun);
/* st_scene_set_sun + st_scene_aim_heliostats(sun->vector, target) with one
 * sun-plane and grid build for the new state instead of two: the per-frame
 * update of lib/soltrace_system.cpp:172-184 when the sun moves and the field
 * tracks it (demos/demo_dynamic_scene.cpp:82-97). */
int st_scene_set_sun_and_aim(st_scene scene, const st_sun_desc* sun, const double target[3]);
int st_scene_get_sun_plane(st_scene scene, st_sun_plane* out);
/* Host copies of derived FP64 data for inspection and tests. */
int st_scene_get_frames(st_scene scene, double* l2g_rows /* n*9, row-major L2G */);
int st_scene_get_aabbs(st_scene scene, float* lo_hi /* n*6 reference AABBs */);
int st_scene_bvh_info(st_scene scene, int32_t* n_nodes, int32_t* depth, int32_t* max_leaf);
int st_scene_synchronize(st_scene scene);

/* Candidate lists (new). Sun rays test the primitives binned under their
 * sun-plane cell; first and second reflections off mirrors test per-mirror
 * and per-mirror-pair lists of everything their bounded reflection cone can
 * reach. Hits are identical to BVH traversal. The sun-plane grid is rebuilt
 * with the sun plane; the reflection lists lazily, once enough rays have been
 * traced on the current scene state to repay their build (2^26 at first, then
 * scaled by the measured build time; earlier traces use the BVH). The
 * autotuner builds them.
 * st_scene_build_lists builds them now; st_scene_list_info reports their sizes
 * (grid_items / reflection_items / pair_items, 0 when unused). */
int st_scene_build_lists(st_scene scene);
int st_scene_list_info(st_scene scene, uint64_t* grid_items, uint64_t* reflection_items, uint64_t* pair_items);

/* Launch-shape autotuner (new): times every trace-kernel variant (block size,
 * traversal stack depth, waves per SIMD) at several persistent grid sizes on
 * sample_rays rays of this scene with its current sun, keeps the fastest as
 * the scene's default (used when st_trace_desc.block_size and grid_blocks are
 * 0) and reports it. */
typedef struct st_tune_result {
    int32_t block, grid, stack, waves;  /* stack: LDS stack entries per lane; 0 = the
                                           stack in private memory (frees LDS for 8
                                           waves per SIMD) */
    int32_t trials;
    int32_t queued;               /* 1: ST_KERNEL_QUEUED form chosen; 2: the same
                                     with list and BVH segments traced in separate
                                     steps                                        */
    double best_ms, default_ms;   /* kernel time of the sample, tuned and untuned;
                                     best_ms can exceed default_ms by up to 1 %:
                                     queued = 2 is kept when within 1 % of the
                                     fastest shape (it holds up at lower sun)     */
} st_tune_result;
int st_scene_autotune(st_scene scene, uint64_t sample_rays, int32_t max_depth, st_tune_result* out);

/* ---- tracing ----------------------------------------------------------- */
/* Trace on desc->stream (or the scene stream). Synchronous unless stream != NULL.
 * Replaces optixLaunch in SolTraceSystem::run() (lib/soltrace_system.cpp:145-170)
 * and the offline binning of scripts/post_processing/flux_map_calculator.py:95-179.
 * desc->kernel: an explicit form runs as asked (ST_KERNEL_QUEUED_GROUPED fails
 * with ST_ERR_INVALID when no grouped variant fits the shape); ST_KERNEL_AUTO
 * takes the tuned form or the default and picks the grouped form where
 * reflected rays mix listed and BVH segments. Position sampling ST_POS_HALTON
 * needs ray_begin + count <= 2^32 (the radical inverse takes a 32-bit index,
 * as the reference's unsigned launch index, cuda/sun.cu:94-134).
 * Traces of one scene on different streams are ordered (each waits for the
 * scene's previous trace): they share the scene's chunk counter and flux copies. */
int st_trace(st_scene scene, const st_trace_desc* desc, const st_outputs* out);

/* st_trace without device allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory - code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc. Same arguments, results
 * and errors as st_trace, but it calls no hipMalloc/hipFree: candidate lists are
 * not built on demand (those the scene already holds are used; results are the
 * same either way), and the chunk counter and the copies of a small flux map
 * live in a static device scratch of 2 MiB + 256 B per GPU (held until the
 * process ends; every flux-copy set st_trace would use fits, larger ones are
 * rejected with ST_ERR_INVALID). The scratch is cleared by kernels at each call
 * and keeps the last call's contents; calls that overlap on one GPU (several
 * streams or threads) share it and race - serialise them. */
int st_trace_noalloc(st_scene scene, const st_trace_desc* desc, const st_outputs* out);

/* The trace-kernel variant the scene's last st_trace launched (new; tests and
 * profiling). */
typedef struct st_launch_info {
    int32_t kernel;          /* ST_KERNEL_MEGA / _QUEUED / _QUEUED_GROUPED launched */
    int32_t block_size, grid_blocks, stack, waves_per_simd, chunk_rays;  /* stack 0: private */
    int32_t lds_hits;        /* element hits counted per wave in LDS              */
    int32_t flux_copies;     /* copies the flux map was accumulated in (1 = none) */
    uint64_t chunks;
} st_launch_info;
int st_scene_last_launch(st_scene scene, st_launch_info* info);
/* FP64 normalisation: flux_wm2[i] = flux[i] * dni * area / num_rays / bin_area. */
int st_flux_to_watts(st_scene scene, const st_trace_desc* desc, const int64_t* flux_dev,
                     double* flux_wm2_dev, double* power_per_ray);

/* Batched sweep: `steps` sun positions over one field. For step k the mirrors
 * are re-aimed at `target` (bisector), the BVH is refit, the sun plane is
 * recomputed and budgets[k] rays (job_id = k) are traced; all of it captured in
 * one hipGraph. Outputs are per step: flux[k][ny][nx], element_hits[k][n],
 * stats[k][ST_STAT_COUNT] (device, caller-owned, zeroed by the call). */
typedef struct st_batch_desc {
    int32_t steps;
    int32_t max_depth;
    const double* sun_vectors;   /* host [steps][3] */
    const uint64_t* budgets;     /* host [steps]    */
    double target[3];            /* aim target for every mirror */
    st_sun_desc sun;             /* shape/param/dni/sampling (vector ignored) */
    uint64_t seed;
    int32_t flux_element, flux_nx, flux_ny;
    uint32_t flags;
    int32_t use_graph;           /* 1 = capture and replay in a hipGraph */
    int32_t block_size, grid_blocks;
    void* stream;
} st_batch_desc;
int st_trace_batch(st_scene scene, const st_batch_desc* desc, int64_t* flux, uint64_t* element_hits,
                   uint64_t* stats);

/* ---- serving ------------------------------------------------------------- */
typedef struct st_server_config {
    uint64_t hbm_budget_bytes;  /* admission limit over resident job footprints */
    int32_t max_active_jobs;    /* slots in the continuous batch                */
    uint64_t chunk_rays;        /* most rays of one job per scheduling round    */
    int32_t block_size, grid_blocks;
    uint64_t round_rays;        /* most rays of one round over all jobs, shortest
                                   remaining job first (0: 4 x chunk_rays)      */
} st_server_config;
typedef struct st_job_desc {
    int32_t scene_slot;          /* index returned by st_server_add_scene        */
    st_sun_desc sun;
    uint64_t num_rays;
    uint64_t seed;
    uint32_t job_id;
    int32_t max_depth;
    int32_t flux_nx, flux_ny;
    uint32_t flags;
} st_job_desc;
typedef struct st_job_result {
    int32_t status;              /* ST_OK, ST_ERR_REJECTED, ST_ERR_CANCELLED, ... */
    int32_t done;
    double submit_s, start_s, finish_s; /* host steady-clock seconds          */
    double power_per_ray;
    uint64_t stats[ST_STAT_COUNT];
} st_job_result;
int st_server_create(st_context ctx, const st_server_config* cfg, st_server* out);
int st_server_destroy(st_server srv);
/* Register a device-resident scene (the server takes a reference; the scene must
 * outlive the server). Returns the slot index. */
int st_server_add_scene(st_server srv, st_scene scene, int32_t* slot);
int st_server_submit(st_server srv, const st_job_desc* job, int64_t* job_handle);
int st_server_poll(st_server srv, int64_t job_handle, st_job_result* result);
int st_server_wait(st_server srv, int64_t job_handle, st_job_result* result);
/* Copy a finished job's outputs to host memory (NULL skips a buffer). */
int st_server_fetch(st_server srv, int64_t job_handle, int64_t* flux_host, uint64_t* element_hits_host);
int st_server_cancel(st_server srv, int64_t job_handle);
int st_server_release(st_server srv, int64_t job_handle);
typedef struct st_server_stats {
    uint64_t submitted, admitted, rejected, completed, cancelled;
    uint64_t rays_traced, launches;
    uint64_t resident_bytes, peak_resident_bytes;
    double busy_s;
} st_server_stats;
int st_server_get_stats(st_server srv, st_server_stats* stats);

/* ---- measurement --------------------------------------------------------- */
/* No counterpart in the reference. Plain (non-packed) FP32 vector-instruction
 * issue rate of the context's device, from independent FMA chains: the roofline
 * the trace kernel is measured against (it runs in plain FP32; the published
 * FP32 peak counts packed two-wide FMAs). Lane operations per second. */
int st_calibrate_valu(st_context ctx, double* lane_ops_per_s);

/* ---- testing hooks ------------------------------------------------------ */
/* Make the next `count` device allocations (or launches, what=1) fail with
 * ST_ERR_INJECTED; host-side only, no kernel ever faults. */
int st_debug_inject_fault(int32_t what, int32_t count);

#ifdef __cplusplus
}
#endif
#endif /* SOLTRACE_HIP_H */
