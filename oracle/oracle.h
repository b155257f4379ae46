/*
 * oracle.h - FP64 scalar CPU oracle for the SolTrace-style hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the HIP path is compared
 * against; it is loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py. The product library never links or calls it.
 *
 * It restates, in double precision and one ray at a time, the reference's
 *   sun sampling        cuda/sun.cu:16-38, 72-134
 *   sun plane           lib/sun_utils.cu:53-179, lib/geometry_manager.cpp:97-148
 *   element frames      lib/math_util.h:21-59, lib/element.cpp:77-164
 *   reference AABBs     lib/element.cpp:172-275
 *   intersections       cuda/intersection.cu:47-85 (flat), 177-296 (capped
 *                       cylinder), 318-448 (parabolic)
 *   shading             cuda/materials.cu:26-253 (mirror, receivers, miss)
 *   flux binning        scripts/post_processing/flux_map_calculator.py:95-179
 * with the extensions the reference lacks (Philox4x32-10 draws, Gaussian sun,
 * slope / specular error, reflectivity), defined in oracle.c.
 *
 * The element / sun / trace structs have the same layout as the ones in
 * include/soltrace_hip.h so the same bytes can be handed to both.
 */
#ifndef ST_ORACLE_H
#define ST_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_element {
    double origin[3];
    double aim_point[3];
    double zrot_deg;
    double dim_x, dim_y;
    double curv_x, curv_y;
    double reflectivity;
    double slope_error;
    double spec_error;
    int32_t surface;  /* 0 flat, 1 parabolic, 3 cylinder */
    int32_t aperture; /* 0 rectangle */
    int32_t role;     /* 0 mirror, 1 receiver */
    int32_t reserved;
} or_element;

typedef struct or_sun {
    double vector[3];
    double param;
    double dni;
    int32_t shape;             /* 0 pillbox, 1 gaussian */
    int32_t position_sampling; /* 0 philox, 1 halton, 2 stratified */
} or_sun;

typedef struct or_sun_plane {
    double v0[3], v1[3], v2[3], v3[3];
    double distance;
    double area;
    double box_half_angle;
} or_sun_plane;

typedef struct or_trace_desc {
    uint64_t num_rays;
    uint64_t ray_begin;
    uint64_t ray_count;
    uint64_t seed;
    uint32_t job_id;
    int32_t max_depth;
    uint32_t flags; /* bit0 records, bit1 flux includes direct-sun hits */
    int32_t kernel;
    int32_t flux_element;
    int32_t flux_nx, flux_ny;
    int32_t block_size;
    int32_t grid_blocks;
    int32_t reserved;
    void* stream;
} or_trace_desc;

/* host output buffers (any may be NULL) */
typedef struct or_outputs {
    int64_t* flux;             /* [ny][nx] */
    uint64_t* element_hits;    /* [n]      */
    uint64_t* stats;           /* [16]     */
    double* records;           /* [count][max_depth][4] */
    uint32_t* record_elements; /* [count][max_depth]    */
} or_outputs;

typedef struct or_hit {
    double t;
    double p[3];      /* world hit point            */
    double n[3];      /* world normal (unit)        */
    double local[3];  /* element-frame coordinates  */
    int32_t elem;
    int32_t cap;      /* cylinder cap hit           */
} or_hit;

typedef struct or_scene or_scene;

/* Philox4x32-10 (Salmon et al., SC'11) */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_uniform(uint32_t x);           /* ((x>>9)+0.5) * 2^-23, in (0,1) */
double or_halton(uint32_t index, uint32_t base);
uint32_t or_stratified_cell(uint64_t ray, uint64_t seed, uint32_t job); /* cx | cy << 8 */

/* element frame: L2G rows (x,y,z basis vectors are the columns) */
void or_element_frame(const or_element* e, double l2g[9]);
/* reference AABB (lib/element.cpp:172-275) */
void or_element_ref_aabb(const or_element* e, double lo[3], double hi[3]);

or_scene* or_scene_create(const or_element* el, int32_t n);
void or_scene_destroy(or_scene* s);
int or_scene_set_sun(or_scene* s, const or_sun* sun);
int or_scene_get_sun_plane(const or_scene* s, or_sun_plane* out);
int or_scene_get_frames(const or_scene* s, double* l2g);

/* single-ray helpers for known-answer tests */
int or_closest_hit(const or_scene* s, const double o[3], const double d[3], double tmin, double tmax,
                   int32_t brute_force, or_hit* hit);
int or_intersect_element(const or_scene* s, int32_t elem, const double o[3], const double d[3],
                         double tmin, double tmax, or_hit* hit);
void or_sample_sun(const or_scene* s, uint64_t ray, uint64_t seed, uint32_t job, double pos[3],
                   double dir[3]);
void or_perturb(const double a[3], double sigma, double xi_r, double xi_phi, double out[3]);
void or_reflect(const double d[3], const double n[3], double out[3]);
/* mirror interaction: faceforward, slope error on the normal, reflect, specular
 * error on the reflected ray (the composition the trace uses) */
void or_mirror_direction(const double d[3], const double n[3], double sigma_slope, double sigma_spec,
                         double xs_r, double xs_p, double xp_r, double xp_p, double out[3]);
/* tests only: switch on planted mistake `code` (0 = none), returns the previous
 * code. 1 cylinder theta mirrored, 2 cylinder theta + pi, 3 flat map axes
 * swapped, 4 flat map x along +x_b, 5 slope error on the reflected ray,
 * 6 specular error on the normal. */
int or_debug_plant(int code);

/* the full trace: rays [ray_begin, ray_begin+ray_count) with `threads` workers */
int or_trace(const or_scene* s, const or_trace_desc* desc, const or_outputs* out, int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
