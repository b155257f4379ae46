/*
 * oracle.c - FP64 scalar CPU oracle (test infrastructure, see oracle.h).
 *
 * Every function cites the reference file:line it restates. Where the
 * reference has no definition (Philox draws, Gaussian sun, slope/specular
 * error, reflectivity, flux binning on the device) the published method is
 * named instead. Double precision throughout; one ray at a time.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979323846
#define OR_MAX_STACK 128
#define OR_STATS 16

/* Planted mistakes (tests only): the pinning tests switch one on to show that
 * their check fails on it, then switch it off. 0 = none; see or_debug_plant(). */
static int g_plant = 0;
int or_debug_plant(int code) {
    int prev = g_plant;
    g_plant = code;
    return prev;
}

/* ------------------------------------------------------------------------ */
/* small vector helpers                                                       */
static inline void v3set(double* r, double x, double y, double z) { r[0] = x; r[1] = y; r[2] = z; }
static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double* a, const double* b, double* r) {
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    v3set(r, x, y, z);
}
static inline void sub3(const double* a, const double* b, double* r) { v3set(r, a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
static inline void axpy3(double s, const double* x, const double* y, double* r) {
    v3set(r, s * x[0] + y[0], s * x[1] + y[1], s * x[2] + y[2]);
}
static inline void normalize3(double* a) {
    double n = sqrt(dot3(a, a));
    if (n > 0) { a[0] /= n; a[1] /= n; a[2] /= n; }
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11. Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
 * 0x9E3779B9 / 0xBB67AE85, ten rounds. Replaces curand XORWOW of
 * cuda/sun.cu:74-75 (NVIDIA closed library, not reproduced). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 23-bit uniform ((x >> 9) + 0.5) / 2^23 in (0,1): it and 1 - u are exactly
 * representable in FP32 and FP64, so the GPU (FP32) and this oracle (FP64)
 * start every draw from the same number. */
double or_uniform(uint32_t x) { return ((double)(x >> 9) + 0.5) * (1.0 / 8388608.0); }

static void draw4(uint64_t ray, uint32_t block, uint32_t job, uint64_t seed, double u[4]) {
    uint32_t ctr[4] = {(uint32_t)ray, (uint32_t)(ray >> 32), block, job};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    or_philox4x32_10(ctr, key, r);
    for (int i = 0; i < 4; ++i) u[i] = or_uniform(r[i]);
}

/* Radical inverse, cuda/sun.cu:16-24 (`halton(index, base)`), evaluated
 * exactly: digits accumulate in integers, one division at the end. */
double or_halton(uint32_t index, uint32_t base) {
    uint64_t num = 0, den = 1;
    uint32_t i = index;
    while (i > 0) {
        num = num * base + (i % base);
        den *= base;
        i /= base;
    }
    return den > 1 ? (double)num / (double)den : 0.0;
}

/* ------------------------------------------------------------------------ */
/* Element frames: lib/math_util.h:21-32 (normal_to_euler), :41-59
 * (get_rotation_matrix_G2L, Spencer & Murty eq. 2), lib/element.cpp:77-100
 * (update_euler_angles, get_rotation_matrix = G2L^T). Output row-major L2G:
 * its columns are the element x, y, z axes in world coordinates. */
void or_element_frame(const or_element* e, double l2g[9]) {
    double n[3];
    sub3(e->aim_point, e->origin, n);
    normalize3(n);
    double yaw = atan2(n[0], n[2]);
    double pitch = asin(n[1]);
    double roll = e->zrot_deg * OR_PI / 180.0;
    double ca = cos(yaw), sa = sin(yaw), cb = cos(pitch), sb = sin(pitch), cg = cos(roll), sg = sin(roll);
    double g2l[9] = {ca * cg + sa * sb * sg, -cb * sg, -sa * cg + ca * sb * sg,
                     ca * sg - sa * sb * cg, cb * cg,  -sa * sg - ca * sb * cg,
                     sa * cb,                sb,       ca * cb};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) l2g[r * 3 + c] = g2l[c * 3 + r];
}

static void frame_axes(const double l2g[9], double x[3], double y[3], double z[3]) {
    v3set(x, l2g[0], l2g[3], l2g[6]); /* Matrix33d::get_x_basis, lib/vector3d.h:151-153 */
    v3set(y, l2g[1], l2g[4], l2g[7]);
    v3set(z, l2g[2], l2g[5], l2g[8]);
}

static void l2g_point(const double l2g[9], const double* o, const double* p, double* r) {
    v3set(r, l2g[0] * p[0] + l2g[1] * p[1] + l2g[2] * p[2] + o[0],
             l2g[3] * p[0] + l2g[4] * p[1] + l2g[5] * p[2] + o[1],
             l2g[6] * p[0] + l2g[7] * p[1] + l2g[8] * p[2] + o[2]);
}

/* Reference AABB, lib/element.cpp:172-275: rectangle (non-cylinder) = 4 rotated
 * corners of the flat aperture (parabolic sag ignored, :181-206); cylinder =
 * 8 corners of (+-w/2, +-h/2, +-w/2) (:209-267). Used for the sun plane only. */
void or_element_ref_aabb(const or_element* e, double lo[3], double hi[3]) {
    double l2g[9];
    or_element_frame(e, l2g);
    double w = e->dim_x, h = e->dim_y;
    int ncorner = e->surface == 3 ? 8 : 4;
    for (int k = 0; k < 3; ++k) { lo[k] = 1e300; hi[k] = -1e300; }
    for (int c = 0; c < ncorner; ++c) {
        double p[3] = {(c & 1) ? w / 2 : -w / 2, (c & 2) ? h / 2 : -h / 2, 0.0};
        if (ncorner == 8) p[2] = (c & 4) ? w / 2 : -w / 2;
        double g[3];
        l2g_point(l2g, e->origin, p, g);
        for (int k = 0; k < 3; ++k) { lo[k] = fmin(lo[k], g[k]); hi[k] = fmax(hi[k], g[k]); }
    }
}

/* Conservative bounds for the oracle's own BVH (not a reference quantity):
 * the reference box plus the paraboloid sag z = cx/2 x^2 + cy/2 y^2. */
static void element_bounds(const or_element* e, double lo[3], double hi[3]) {
    double l2g[9];
    or_element_frame(e, l2g);
    double w = e->dim_x, h = e->dim_y;
    double zlo = 0, zhi = 0;
    if (e->surface == 1) {
        double sx = 0.5 * e->curv_x * (w / 2) * (w / 2), sy = 0.5 * e->curv_y * (h / 2) * (h / 2);
        zlo = fmin(0, sx) + fmin(0, sy);
        zhi = fmax(0, sx) + fmax(0, sy);
    } else if (e->surface == 3) {
        zlo = -w / 2;
        zhi = w / 2;
    }
    for (int k = 0; k < 3; ++k) { lo[k] = 1e300; hi[k] = -1e300; }
    for (int c = 0; c < 8; ++c) {
        double p[3] = {(c & 1) ? w / 2 : -w / 2, (c & 2) ? h / 2 : -h / 2, (c & 4) ? zhi : zlo};
        double g[3];
        l2g_point(l2g, e->origin, p, g);
        for (int k = 0; k < 3; ++k) { lo[k] = fmin(lo[k], g[k]); hi[k] = fmax(hi[k], g[k]); }
    }
    for (int k = 0; k < 3; ++k) {
        double pad = 1e-9 * (1.0 + fabs(lo[k]) + fabs(hi[k]));
        lo[k] -= pad;
        hi[k] += pad;
    }
}

/* ------------------------------------------------------------------------ */
/* scene                                                                      */
typedef struct or_prim {
    int32_t surface, role;
    double c[3];        /* center (flat, cylinder) / rectangle center (parabolic) */
    double ax[3], ay[3], az[3]; /* flat: x, y, x*y; parabolic: e1, e2, n; cylinder: base_x, axis, base_z */
    double hw, hh;      /* flat: w/2,h/2; parabolic: L1/2, L2/2; cylinder: r, h/2 */
    double cx, cy;
    double rho, slope, spec;
} or_prim;

typedef struct or_node {
    double lo[3], hi[3];
    int32_t left, right; /* children (inner) */
    int32_t start, count; /* leaf range into prim_index, count > 0 for leaves */
} or_node;

struct or_scene {
    int32_t n;
    or_element* el;
    or_prim* prim;
    double* frames; /* n*9 */
    double (*blo)[3];
    double (*bhi)[3];
    or_node* nodes;
    int32_t n_nodes;
    int32_t* prim_index;
    int32_t sun_set;
    or_sun sun;
    or_sun_plane plane;
    double sun_w[3], sun_u[3], sun_v[3]; /* sampling basis, cuda/sun.cu:77-80 */
};

/* Per-element geometry, lib/element.cpp:114-164 (toDeviceGeometryData) and the
 * GeometryDataST constructors (cuda/GeometryDataST.h:41-103), in FP64.
 * Parabolic: v1 = -w*x_b, v2 = h*y_b, e1 = v1/|v1| = -x_b, e2 = y_b,
 * n = normalize(e2 x e1) (cuda/intersection.cu:339-347), center = origin. */
static void make_prim(const or_element* e, const double l2g[9], or_prim* p) {
    double x[3], y[3], z[3];
    frame_axes(l2g, x, y, z);
    memset(p, 0, sizeof(*p));
    p->surface = e->surface;
    p->role = e->role;
    memcpy(p->c, e->origin, sizeof(p->c));
    p->rho = e->reflectivity;
    p->slope = e->slope_error;
    p->spec = e->spec_error;
    if (e->surface == 0) {
        memcpy(p->ax, x, sizeof(x));
        memcpy(p->ay, y, sizeof(y));
        cross3(x, y, p->az);
        normalize3(p->az);
        p->hw = e->dim_x / 2;
        p->hh = e->dim_y / 2;
    } else if (e->surface == 1) {
        v3set(p->ax, -x[0], -x[1], -x[2]);
        memcpy(p->ay, y, sizeof(y));
        cross3(p->ay, p->ax, p->az);
        normalize3(p->az);
        p->hw = e->dim_x / 2;
        p->hh = e->dim_y / 2;
        p->cx = e->curv_x;
        p->cy = e->curv_y;
    } else {
        /* cylinder: base_x = x_b, base_z = z_b, axis = base_z x base_x = y_b
         * (cuda/intersection.cu:192-195), r = w/2, half height = h/2 */
        memcpy(p->ax, x, sizeof(x));
        cross3(z, x, p->ay);
        memcpy(p->az, z, sizeof(z));
        p->hw = e->dim_x / 2;
        p->hh = e->dim_y / 2;
    }
}

static int build_node(or_scene* s, int32_t* idx, int32_t count, int32_t start) {
    int32_t me = s->n_nodes++;
    or_node* nd = &s->nodes[me];
    for (int k = 0; k < 3; ++k) { nd->lo[k] = 1e300; nd->hi[k] = -1e300; }
    for (int i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            nd->lo[k] = fmin(nd->lo[k], s->blo[idx[i]][k]);
            nd->hi[k] = fmax(nd->hi[k], s->bhi[idx[i]][k]);
        }
    if (count <= 2) {
        nd->start = start;
        nd->count = count;
        nd->left = nd->right = -1;
        return me;
    }
    /* median split along the widest centroid axis (simple, exact) */
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            double c = 0.5 * (s->blo[idx[i]][k] + s->bhi[idx[i]][k]);
            clo[k] = fmin(clo[k], c);
            chi[k] = fmax(chi[k], c);
        }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
    /* insertion-free selection: sort by centroid (count is small overall) */
    for (int i = 1; i < count; ++i) {
        int32_t v = idx[i];
        double cv = s->blo[v][ax] + s->bhi[v][ax];
        int j = i - 1;
        while (j >= 0 && s->blo[idx[j]][ax] + s->bhi[idx[j]][ax] > cv) { idx[j + 1] = idx[j]; --j; }
        idx[j + 1] = v;
    }
    int32_t half = count / 2;
    int32_t l = build_node(s, idx, half, start);
    int32_t r = build_node(s, idx + half, count - half, start + half);
    nd = &s->nodes[me];
    nd->left = l;
    nd->right = r;
    nd->count = 0;
    return me;
}

static int cmp_dbl_idx_axis;
static const or_scene* cmp_scene;
static int cmp_centroid(const void* a, const void* b) {
    int32_t i = *(const int32_t*)a, j = *(const int32_t*)b;
    double ci = cmp_scene->blo[i][cmp_dbl_idx_axis] + cmp_scene->bhi[i][cmp_dbl_idx_axis];
    double cj = cmp_scene->blo[j][cmp_dbl_idx_axis] + cmp_scene->bhi[j][cmp_dbl_idx_axis];
    return ci < cj ? -1 : (ci > cj ? 1 : (i < j ? -1 : (i > j)));
}

/* top levels sorted with qsort so 10k-element fields build fast */
static int build_rec(or_scene* s, int32_t* idx, int32_t count, int32_t start) {
    if (count > 64) {
        int32_t me = s->n_nodes++;
        or_node* nd = &s->nodes[me];
        double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
        for (int k = 0; k < 3; ++k) { nd->lo[k] = 1e300; nd->hi[k] = -1e300; }
        for (int i = 0; i < count; ++i)
            for (int k = 0; k < 3; ++k) {
                nd->lo[k] = fmin(nd->lo[k], s->blo[idx[i]][k]);
                nd->hi[k] = fmax(nd->hi[k], s->bhi[idx[i]][k]);
                double c = 0.5 * (s->blo[idx[i]][k] + s->bhi[idx[i]][k]);
                clo[k] = fmin(clo[k], c);
                chi[k] = fmax(chi[k], c);
            }
        int ax = 0;
        for (int k = 1; k < 3; ++k)
            if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
        cmp_dbl_idx_axis = ax;
        cmp_scene = s;
        qsort(idx, (size_t)count, sizeof(int32_t), cmp_centroid);
        int32_t half = count / 2;
        int32_t l = build_rec(s, idx, half, start);
        int32_t r = build_rec(s, idx + half, count - half, start + half);
        nd = &s->nodes[me];
        nd->left = l;
        nd->right = r;
        nd->count = 0;
        return me;
    }
    return build_node(s, idx, count, start);
}

or_scene* or_scene_create(const or_element* el, int32_t n) {
    if (!el || n <= 0) return NULL;
    for (int i = 0; i < n; ++i) {
        if (el[i].aperture != 0) return NULL;
        if (el[i].surface != 0 && el[i].surface != 1 && el[i].surface != 3) return NULL;
    }
    or_scene* s = (or_scene*)calloc(1, sizeof(or_scene));
    s->n = n;
    s->el = (or_element*)malloc(sizeof(or_element) * (size_t)n);
    memcpy(s->el, el, sizeof(or_element) * (size_t)n);
    s->prim = (or_prim*)malloc(sizeof(or_prim) * (size_t)n);
    s->frames = (double*)malloc(sizeof(double) * 9 * (size_t)n);
    s->blo = malloc(sizeof(double[3]) * (size_t)n);
    s->bhi = malloc(sizeof(double[3]) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        or_element_frame(&el[i], &s->frames[9 * i]);
        make_prim(&el[i], &s->frames[9 * i], &s->prim[i]);
        element_bounds(&el[i], s->blo[i], s->bhi[i]);
    }
    s->nodes = (or_node*)calloc((size_t)(2 * n + 1), sizeof(or_node));
    s->prim_index = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int i = 0; i < n; ++i) s->prim_index[i] = i;
    build_rec(s, s->prim_index, n, 0);
    return s;
}

void or_scene_destroy(or_scene* s) {
    if (!s) return;
    free(s->el);
    free(s->prim);
    free(s->frames);
    free(s->blo);
    free(s->bhi);
    free(s->nodes);
    free(s->prim_index);
    free(s);
}

int or_scene_get_frames(const or_scene* s, double* l2g) {
    if (!s || !l2g) return -1;
    memcpy(l2g, s->frames, sizeof(double) * 9 * (size_t)s->n);
    return 0;
}

/* Sampling basis of cuda/sun.cu:77-80: w = normalize(dir),
 * u = normalize(cross(|w.x| > 0.99 ? y : x, w)), v = cross(w, u). */
static void cone_basis(const double* w, double* u, double* v) {
    double a[3];
    if (fabs(w[0]) > 0.99) v3set(a, 0, 1, 0); else v3set(a, 1, 0, 0);
    cross3(a, w, u);
    normalize3(u);
    cross3(w, u, v);
}

/* Sun plane, lib/geometry_manager.cpp:97-148 with the two reductions of
 * lib/sun_utils.cu:53-104 (d = max |corner . s|) and :106-179 (uv bounds of the
 * corners projected onto the plane, padded by |corner . s| * tan(angle)). */
int or_scene_set_sun(or_scene* s, const or_sun* sun) {
    if (!s || !sun) return -1;
    s->sun = *sun;
    if (s->sun.dni <= 0) s->sun.dni = 1000.0;
    double sv[3] = {sun->vector[0], sun->vector[1], sun->vector[2]};
    normalize3(sv);
    double angle = sun->shape == 1 ? 4.0 * sun->param : sun->param;
    double tana = tan(angle);
    double d = 0.0;
    double (*lo)[3] = malloc(sizeof(double[3]) * (size_t)s->n);
    double (*hi)[3] = malloc(sizeof(double[3]) * (size_t)s->n);
    for (int i = 0; i < s->n; ++i) {
        or_element_ref_aabb(&s->el[i], lo[i], hi[i]);
        for (int c = 0; c < 8; ++c) {
            double p[3] = {(c & 1) ? hi[i][0] : lo[i][0], (c & 2) ? hi[i][1] : lo[i][1], (c & 4) ? hi[i][2] : lo[i][2]};
            d = fmax(d, fabs(dot3(p, sv)));
        }
    }
    double axis[3], u[3], v[3];
    if (fabs(sv[0]) < 0.9) v3set(axis, 1, 0, 0); else v3set(axis, 0, 1, 0);
    cross3(axis, sv, u);
    normalize3(u);
    cross3(sv, u, v);
    normalize3(v);
    double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
    for (int i = 0; i < s->n; ++i)
        for (int c = 0; c < 8; ++c) {
            double p[3] = {(c & 1) ? hi[i][0] : lo[i][0], (c & 2) ? hi[i][1] : lo[i][1], (c & 4) ? hi[i][2] : lo[i][2]};
            double buf = fabs(dot3(p, sv)) * tana;
            double pc[3] = {p[0] - d * sv[0], p[1] - d * sv[1], p[2] - d * sv[2]};
            double k = dot3(pc, sv);
            double proj[3] = {p[0] - k * sv[0], p[1] - k * sv[1], p[2] - k * sv[2]};
            double pu = dot3(proj, u), pv = dot3(proj, v);
            umin = fmin(umin, pu - buf);
            umax = fmax(umax, pu + buf);
            vmin = fmin(vmin, pv - buf);
            vmax = fmax(vmax, pv + buf);
        }
    free(lo);
    free(hi);
    or_sun_plane* pl = &s->plane;
    for (int k = 0; k < 3; ++k) {
        pl->v0[k] = umin * u[k] + vmin * v[k] + d * sv[k];
        pl->v1[k] = umax * u[k] + vmin * v[k] + d * sv[k];
        pl->v2[k] = umax * u[k] + vmax * v[k] + d * sv[k];
        pl->v3[k] = umin * u[k] + vmax * v[k] + d * sv[k];
    }
    pl->distance = d;
    double e1[3], e2[3];
    sub3(pl->v1, pl->v0, e1);
    sub3(pl->v3, pl->v0, e2);
    pl->area = sqrt(dot3(e1, e1)) * sqrt(dot3(e2, e2));
    pl->box_half_angle = angle;
    v3set(s->sun_w, -sv[0], -sv[1], -sv[2]);
    cone_basis(s->sun_w, s->sun_u, s->sun_v);
    s->sun_set = 1;
    return 0;
}

int or_scene_get_sun_plane(const or_scene* s, or_sun_plane* out) {
    if (!s || !out || !s->sun_set) return -1;
    *out = s->plane;
    return 0;
}

/* Isotropic 2-D Gaussian deviation of unit vector a (per-axis sigma): polar
 * Box-Muller, theta = sigma*sqrt(-2 ln xi_r) (Rayleigh), phi = 2 pi xi_phi, in
 * the basis of cone_basis(). Used for the Gaussian sun, the slope error (on the
 * normal) and the specular error (on the reflected ray), SolTrace's error model. */
void or_perturb(const double a[3], double sigma, double xi_r, double xi_phi, double out[3]) {
    double u[3], v[3];
    cone_basis(a, u, v);
    double th = sigma * sqrt(-2.0 * log(xi_r));
    double ph = 2.0 * OR_PI * xi_phi;
    double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    for (int k = 0; k < 3; ++k) out[k] = ct * a[k] + st * (cp * u[k] + sp * v[k]);
    normalize3(out);
}

/* reflect(d, n) = d - 2 (n.d) n (sutil vec_math, used at cuda/materials.cu:51) */
void or_reflect(const double d[3], const double n[3], double out[3]) {
    double k = 2.0 * dot3(n, d);
    for (int i = 0; i < 3; ++i) out[i] = d[i] - k * n[i];
}

/* Outgoing direction of a mirror interaction (cuda/materials.cu:26-84, with
 * SolTrace's error model added): the surface normal n is turned to face the
 * incoming ray d (faceforward, :38), tilted by the slope error (sigma_slope,
 * draws xs_r, xs_p), d is reflected about it (:51), and the reflected ray is
 * tilted by the specular error (sigma_spec, draws xp_r, xp_p). */
void or_mirror_direction(const double d[3], const double n[3], double sigma_slope, double sigma_spec,
                         double xs_r, double xs_p, double xp_r, double xp_p, double out[3]) {
    double nf[3] = {n[0], n[1], n[2]};
    if (dot3(d, nf) > 0) { nf[0] = -nf[0]; nf[1] = -nf[1]; nf[2] = -nf[2]; }
    double rd[3], tmp[3];
    if (sigma_slope > 0) {
        if (g_plant == 5) {
            /* planted: slope error applied to the reflected ray */
            or_reflect(d, nf, rd);
            normalize3(rd);
            or_perturb(rd, sigma_slope, xs_r, xs_p, tmp);
            memcpy(rd, tmp, sizeof(tmp));
        } else {
            or_perturb(nf, sigma_slope, xs_r, xs_p, tmp);
            memcpy(nf, tmp, sizeof(tmp));
            or_reflect(d, nf, rd);
            normalize3(rd);
        }
    } else {
        or_reflect(d, nf, rd);
        normalize3(rd);
    }
    if (sigma_spec > 0) {
        if (g_plant == 6) {
            /* planted: specular error applied to the normal before reflection */
            double n2[3];
            or_perturb(nf, sigma_spec, xp_r, xp_p, n2);
            or_reflect(d, n2, rd);
            normalize3(rd);
        } else {
            or_perturb(rd, sigma_spec, xp_r, xp_p, tmp);
            memcpy(rd, tmp, sizeof(tmp));
        }
    }
    memcpy(out, rd, sizeof(rd));
}

/* Sun ray: __raygen__sun_source, cuda/sun.cu:94-134. Position on the sun
 * parallelogram v0 + a (v1-v0) + b (v3-v0) (:28-38) with (a,b) = Halton(2,3)
 * of the ray index or Philox uniforms; direction uniform over the cap of
 * half-angle alpha around -s (:72-91): z = cos(alpha) + (1-cos(alpha)) xi2,
 * phi = 2 pi xi1, evaluated as 1-z = vers(alpha) (1-xi2) for accuracy. */
/* Stratified sun-plane cell of a ray (position sampling mode 2, new): groups of
 * 64 consecutive rays share one cell of a 256 x 256 grid; within a sweep of
 * 65,536 groups the cells follow a 4-round Feistel permutation of the group
 * index (8-bit halves, round function ((R ^ k) * 0x9E3779B1) >> 24) keyed by a
 * Philox draw of (sweep, seed, job). Every cell is visited once per sweep. */
uint32_t or_stratified_cell(uint64_t ray, uint64_t seed, uint32_t job) {
    const uint64_t group = ray >> 6;
    const uint64_t sweep = group >> 16;
    uint32_t q = (uint32_t)(group & 0xFFFFu);
    uint32_t ctr[4] = {(uint32_t)sweep, (uint32_t)(sweep >> 32), 0x53545254u, job};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t k[4];
    or_philox4x32_10(ctr, key, k);
    uint32_t L = q >> 8, R = q & 0xFFu;
    for (int r = 0; r < 4; ++r) {
        uint32_t F = ((R ^ k[r]) * 0x9E3779B1u) >> 24;
        uint32_t nl = R;
        R = L ^ F;
        L = nl;
    }
    return (L << 8) | R;
}

void or_sample_sun(const or_scene* s, uint64_t ray, uint64_t seed, uint32_t job, double pos[3], double dir[3]) {
    double u[4];
    draw4(ray, 0, job, seed, u);
    double a, b;
    if (s->sun.position_sampling == 1) {
        a = or_halton((uint32_t)ray, 2);
        b = or_halton((uint32_t)ray, 3);
    } else if (s->sun.position_sampling == 2) {
        uint32_t cell = or_stratified_cell(ray, seed, job);
        a = ((double)(cell & 255u) + u[0]) / 256.0;
        b = ((double)(cell >> 8) + u[1]) / 256.0;
    } else {
        a = u[0];
        b = u[1];
    }
    const or_sun_plane* pl = &s->plane;
    for (int k = 0; k < 3; ++k) pos[k] = pl->v0[k] + a * (pl->v1[k] - pl->v0[k]) + b * (pl->v3[k] - pl->v0[k]);
    if (s->sun.shape == 1) {
        or_perturb(s->sun_w, s->sun.param, u[3], u[2], dir);
        return;
    }
    double half = 0.5 * s->sun.param;
    double vers = 2.0 * sin(half) * sin(half);
    double sv = vers * (1.0 - u[3]);
    double z = 1.0 - sv;
    double r = sqrt(sv * (2.0 - sv));
    double ph = 2.0 * OR_PI * u[2];
    double cp = cos(ph), sp = sin(ph);
    for (int k = 0; k < 3; ++k) dir[k] = r * (cp * s->sun_u[k] + sp * s->sun_v[k]) + z * s->sun_w[k];
    normalize3(dir);
}

/* ------------------------------------------------------------------------ */
/* intersections (t in (tmin, tmax), closest wins)                            */

/* __intersection__rectangle_flat, cuda/intersection.cu:47-85: plane through the
 * center with normal x*y, then local (x,y) = (x.v, y.v), inclusive bounds. */
static int isect_flat(const or_prim* p, const double* o, const double* d, double tmin, double tmax, or_hit* h) {
    double dt = dot3(d, p->az);
    double oc[3];
    sub3(p->c, o, oc);
    double t = dot3(p->az, oc) / dt;
    if (!(t > tmin && t < tmax)) return 0;
    double q[3], v[3];
    axpy3(t, d, o, q);
    sub3(q, p->c, v);
    double x = dot3(p->ax, v), y = dot3(p->ay, v);
    if (!(x >= -p->hw && x <= p->hw && y >= -p->hh && y <= p->hh)) return 0;
    h->t = t;
    memcpy(h->p, q, sizeof(q));
    memcpy(h->n, p->az, sizeof(h->n));
    v3set(h->local, x, y, 0);
    h->cap = 0;
    return 1;
}

/* __intersection__rectangle_parabolic, cuda/intersection.cu:318-448: ray in the
 * rectangle frame (e1, e2, n) centred on the aperture; A t^2 + B t + C = 0 with
 * A = cx/2 dx^2 + cy/2 dy^2, B = cx ox dx + cy oy dy - dz,
 * C = cx/2 ox^2 + cy/2 oy^2 - oz. |A| < 1e-12: linear, t = -C/B, t > 0.
 * Otherwise t1 = (-B - sqrt)/(2A), t2 = (-B + sqrt)/(2A); take t1 if
 * 0 < t1 < t2, else t2 if t2 > 0 (no retry when the chosen root falls outside
 * the aperture, :387-394). Then tmin <= t <= tmax, |x| <= L1/2, |y| <= L2/2,
 * normal = normalize(-cx x, -cy y, 1) in the frame. Roots are evaluated with the
 * cancellation-free pair q/A, C/q (same values, stable arithmetic). */
static int isect_parabolic(const or_prim* p, const double* o, const double* d, double tmin, double tmax, or_hit* h) {
    double r[3];
    sub3(o, p->c, r);
    double ox = dot3(r, p->ax), oy = dot3(r, p->ay), oz = dot3(r, p->az);
    double dx = dot3(d, p->ax), dy = dot3(d, p->ay), dz = dot3(d, p->az);
    double A = 0.5 * p->cx * dx * dx + 0.5 * p->cy * dy * dy;
    double B = p->cx * ox * dx + p->cy * oy * dy - dz;
    double C = 0.5 * p->cx * ox * ox + 0.5 * p->cy * oy * oy - oz;
    double t = 0;
    int valid = 0;
    if (fabs(A) < 1e-12) {
        t = -C / B;
        valid = t > 0;
    } else {
        double disc = B * B - 4 * A * C;
        if (disc >= 0) {
            double sq = sqrt(disc);
            double q = B >= 0 ? -0.5 * (B + sq) : -0.5 * (B - sq);
            double ra = q / A, rb = q != 0 ? C / q : ra;
            double t1 = B >= 0 ? ra : rb; /* (-B - sq) / 2A */
            double t2 = B >= 0 ? rb : ra; /* (-B + sq) / 2A */
            if (t1 > 0 && t1 < t2) { t = t1; valid = 1; }
            else if (t2 > 0) { t = t2; valid = 1; }
        }
    }
    if (!valid || t < tmin || t > tmax) return 0;
    double x = ox + t * dx, y = oy + t * dy;
    if (x < -p->hw || x > p->hw || y < -p->hh || y > p->hh) return 0;
    double nl[3] = {-p->cx * x, -p->cy * y, 1.0};
    normalize3(nl);
    for (int k = 0; k < 3; ++k) h->n[k] = nl[0] * p->ax[k] + nl[1] * p->ay[k] + nl[2] * p->az[k];
    normalize3(h->n);
    h->t = t;
    axpy3(t, d, o, h->p);
    v3set(h->local, x, y, oz + t * dz);
    h->cap = 0;
    return 1;
}

/* __intersection__cylinder_y_capped, cuda/intersection.cu:177-296: local frame
 * (base_x, base_z x base_x, base_z); wall: quadratic in (x,z), first t1 then t2
 * with tmin < t < tmax and |y| <= h/2; caps y = +-h/2 when |dy| > 1e-6 with a
 * disk test; t = min(wall, caps); normal radial or +-y. The wall roots are
 * evaluated about the point of closest approach (same roots, no cancellation). */
static int isect_cylinder(const or_prim* p, const double* o, const double* d, double tmin, double tmax, or_hit* h) {
    double r[3];
    sub3(o, p->c, r);
    double ox = dot3(r, p->ax), oy = dot3(r, p->ay), oz = dot3(r, p->az);
    double dx = dot3(d, p->ax), dy = dot3(d, p->ay), dz = dot3(d, p->az);
    double rad = p->hw, hh = p->hh;
    double tbig = tmax + 1.0;
    double t_wall = tbig;
    double A = dx * dx + dz * dz;
    if (A > 0) {
        double tc = -(ox * dx + oz * dz) / A;
        double px = ox + tc * dx, pz = oz + tc * dz;
        double h2 = rad * rad - (px * px + pz * pz);
        if (h2 >= 0) {
            double dt = sqrt(h2 / A);
            double t1 = tc - dt, t2 = tc + dt;
            if (t1 > tmin && t1 < tmax && fabs(oy + t1 * dy) <= hh) t_wall = t1;
            else if (t2 > tmin && t2 < tmax && fabs(oy + t2 * dy) <= hh) t_wall = t2;
        }
    }
    double t_cap = tbig;
    if (fabs(dy) > 1e-6) {
        for (int sgn = -1; sgn <= 1; sgn += 2) {
            double t = (sgn * hh - oy) / dy;
            double hx = ox + t * dx, hz = oz + t * dz;
            if (t > tmin && t < tmax && hx * hx + hz * hz <= rad * rad) t_cap = fmin(t_cap, t);
        }
    }
    double t = fmin(t_wall, t_cap);
    if (t >= tmax || t <= tmin) return 0;
    double lx = ox + t * dx, ly = oy + t * dy, lz = oz + t * dz;
    double nl[3];
    int cap = !(t == t_wall);
    if (!cap) { v3set(nl, lx, 0, lz); normalize3(nl); }
    else v3set(nl, 0, signbit(ly) ? -1.0 : 1.0, 0);
    for (int k = 0; k < 3; ++k) h->n[k] = nl[0] * p->ax[k] + nl[1] * p->ay[k] + nl[2] * p->az[k];
    h->t = t;
    axpy3(t, d, o, h->p);
    v3set(h->local, lx, ly, lz);
    h->cap = cap;
    return 1;
}

static int isect_prim(const or_prim* p, const double* o, const double* d, double tmin, double tmax, or_hit* h) {
    if (p->surface == 0) return isect_flat(p, o, d, tmin, tmax, h);
    if (p->surface == 1) return isect_parabolic(p, o, d, tmin, tmax, h);
    return isect_cylinder(p, o, d, tmin, tmax, h);
}

int or_intersect_element(const or_scene* s, int32_t elem, const double o[3], const double d[3], double tmin,
                         double tmax, or_hit* hit) {
    if (!s || elem < 0 || elem >= s->n) return -1;
    int r = isect_prim(&s->prim[elem], o, d, tmin, tmax, hit);
    if (r) hit->elem = elem;
    return r;
}

/* closest hit: smallest t; equal t -> lower element index (deterministic
 * stand-in for OptiX's unspecified order among equal hits) */
static inline int better(const or_hit* c, double t, int32_t e) { return t < c->t || (t == c->t && e < c->elem); }

static int slab(const or_node* nd, const double* o, const double* inv, double tmin, double tmax) {
    double t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        double a = (nd->lo[k] - o[k]) * inv[k], b = (nd->hi[k] - o[k]) * inv[k];
        if (isnan(a) || isnan(b)) { /* origin on the slab plane with d_k == 0 */
            if (o[k] < nd->lo[k] || o[k] > nd->hi[k]) return 0;
            continue;
        }
        if (a > b) { double tmp = a; a = b; b = tmp; }
        t0 = fmax(t0, a);
        t1 = fmin(t1, b);
    }
    return t0 <= t1;
}

static int closest(const or_scene* s, const double* o, const double* d, double tmin, double tmax, int brute, or_hit* best) {
    best->t = tmax;
    best->elem = -1;
    or_hit h;
    if (brute) {
        for (int i = 0; i < s->n; ++i) {
            double lim = best->elem >= 0 ? nextafter(best->t, 1e300) : best->t;
            if (isect_prim(&s->prim[i], o, d, tmin, lim, &h) && better(best, h.t, i)) {
                *best = h;
                best->elem = i;
            }
        }
        return best->elem >= 0;
    }
    double inv[3];
    for (int k = 0; k < 3; ++k) inv[k] = 1.0 / d[k];
    int32_t stack[OR_MAX_STACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const or_node* nd = &s->nodes[stack[--sp]];
        if (!slab(nd, o, inv, tmin, best->t)) continue;
        if (nd->count > 0) {
            for (int j = 0; j < nd->count; ++j) {
                int32_t i = s->prim_index[nd->start + j];
                /* tmax slightly above the current best so an equal-t hit with a
                 * lower index is still seen */
                double lim = best->elem >= 0 ? nextafter(best->t, 1e300) : best->t;
                if (isect_prim(&s->prim[i], o, d, tmin, lim, &h) && better(best, h.t, i)) {
                    *best = h;
                    best->elem = i;
                }
            }
        } else {
            if (sp + 2 > OR_MAX_STACK) return -1;
            stack[sp++] = nd->left;
            stack[sp++] = nd->right;
        }
    }
    return best->elem >= 0;
}

int or_closest_hit(const or_scene* s, const double o[3], const double d[3], double tmin, double tmax,
                   int32_t brute_force, or_hit* hit) {
    if (!s) return -1;
    return closest(s, o, d, tmin, tmax, brute_force, hit);
}

/* ------------------------------------------------------------------------ */
/* the trace                                                                  */
typedef struct or_worker {
    const or_scene* s;
    const or_trace_desc* desc;
    const or_outputs* out;
    uint64_t begin, end; /* local ray offsets within the call */
    int64_t* flux;
    uint64_t* hits;
    uint64_t stats[OR_STATS];
    int32_t flux_elem;
    int err;
} or_worker;

/* Flux bin of a receiver hit, scripts/post_processing/flux_map_calculator.py:95-179
 * and plot_flux_map_gpu.py:150-206 (indices clipped to the grid, map [ny][nx] as
 * flux_2D[raybins_y, raybins_x] at plot_flux_map_gpu.py:263).
 *  - Flat: the scripts describe the receiver as the parallelogram (v1, v2,
 *    anchor) that lib/element.cpp:136-139 makes of an element: v1 = -W x_b,
 *    v2 = H y_b, so the map's x runs along -x_b:
 *    ix = floor((W/2 - x_local) nx / W), iy = floor((y_local + H/2) ny / H).
 *    (plot_flux_map_gpu.py:63-66 hard-codes exactly this parallelogram for the
 *    three-heliostat receiver at (0,0,10) aimed at (0,5,0).)
 *  - Cylinder: theta = atan2(-z_local, x_local) mod 2 pi, which equals the
 *    scripts' global atan2(y - c_y, x - c_x) for their tower frame base_x =
 *    (1,0,0), base_z = (0,-1,0) (plot_flux_map_gpu.py:125-130, 185, 201);
 *    z along base_z x base_x; ix = floor(theta nx / 2 pi), iy = floor((z + H/2) ny / H). */
static int flux_bin(const or_prim* p, const or_hit* h, int nx, int ny) {
    double fx, fy;
    if (p->surface == 3) {
        double th = g_plant == 1 ? atan2(h->local[2], h->local[0]) : atan2(-h->local[2], h->local[0]);
        if (g_plant == 2) th += OR_PI;
        if (th < 0) th += 2 * OR_PI;
        if (th >= 2 * OR_PI) th -= 2 * OR_PI;
        fx = th * nx / (2 * OR_PI);
        fy = (h->local[1] + p->hh) * ny / (2 * p->hh);
    } else if (g_plant == 3) {
        fx = (p->hh - h->local[1]) * nx / (2 * p->hh);
        fy = (h->local[0] + p->hw) * ny / (2 * p->hw);
    } else if (g_plant == 4) {
        fx = (h->local[0] + p->hw) * nx / (2 * p->hw);
        fy = (h->local[1] + p->hh) * ny / (2 * p->hh);
    } else {
        fx = (p->hw - h->local[0]) * nx / (2 * p->hw);
        fy = (h->local[1] + p->hh) * ny / (2 * p->hh);
    }
    long ix = (long)floor(fx), iy = (long)floor(fy);
    if (ix < 0) ix = 0;
    if (ix > nx - 1) ix = nx - 1;
    if (iy < 0) iy = 0;
    if (iy > ny - 1) iy = ny - 1;
    return (int)(iy * nx + ix);
}

static void trace_ray(or_worker* w, uint64_t local) {
    const or_scene* s = w->s;
    const or_trace_desc* dsc = w->desc;
    uint64_t ray = dsc->ray_begin + local;
    int maxd = dsc->max_depth;
    double* rec = w->out->records ? w->out->records + local * (uint64_t)maxd * 4 : NULL;
    uint32_t* rid = w->out->record_elements ? w->out->record_elements + local * (uint64_t)maxd : NULL;
    if (rec) memset(rec, 0, sizeof(double) * 4 * (size_t)maxd);
    if (rid) for (int k = 0; k < maxd; ++k) rid[k] = 0xFFFFFFFFu;

    double o[3], d[3];
    or_sample_sun(s, ray, dsc->seed, dsc->job_id, o, d);
    w->stats[0]++;
    if (rec) { rec[0] = 0.0; rec[1] = o[0]; rec[2] = o[1]; rec[3] = o[2]; }
    if (rid) rid[0] = 0xFFFFFFFEu;
    int depth = 0;
    double tmin = 1e-3; /* cuda/sun.cu:123 */
    for (;;) {
        or_hit h;
        w->stats[1]++;
        int r = closest(s, o, d, tmin, 1e16, 0, &h);
        if (r < 0) { w->err = 1; return; }
        if (r == 0) { w->stats[6]++; return; } /* __miss__ms, cuda/materials.cu:236-253 */
        int nd = depth + 1;
        const or_prim* p = &s->prim[h.elem];
        if (p->role == 1) {
            /* __closesthit__receiver (:86-121): flat records the front face only
             * (d . (x*y) < 0), back face stops silently;
             * __closesthit__receiver__cylinder__y (:123-163): any face. */
            if (p->surface == 0 && !(dot3(d, p->az) < 0)) { w->stats[7]++; return; }
            if (nd < maxd) {
                if (rec) { double* q = rec + 4 * nd; q[0] = 2.0; q[1] = h.p[0]; q[2] = h.p[1]; q[3] = h.p[2]; }
                if (rid) rid[nd] = (uint32_t)h.elem;
                if (w->hits) w->hits[h.elem]++;
                w->stats[3]++;
                int direct = depth == 0;
                if (direct) w->stats[4]++;
                if (h.cap) w->stats[9]++;
                if (w->flux && h.elem == w->flux_elem && !h.cap && (!direct || (dsc->flags & 2u))) {
                    w->flux[flux_bin(p, &h, dsc->flux_nx, dsc->flux_ny)]++;
                    w->stats[10]++;
                }
            } else {
                w->stats[8]++;
            }
            return;
        }
        /* __closesthit__mirror / __closesthit__mirror__parabolic (:26-84, :170-231) */
        if (nd >= maxd) { w->stats[8]++; return; }
        if (rec) { double* q = rec + 4 * nd; q[0] = 1.0; q[1] = h.p[0]; q[2] = h.p[1]; q[3] = h.p[2]; }
        if (rid) rid[nd] = (uint32_t)h.elem;
        if (w->hits) w->hits[h.elem]++;
        w->stats[2]++;
        /* Draws of bounce nd: block 2nd-1 = (reflectivity, slope xr, slope xp,
         * specular xr), block 2nd = (specular xp, -, -, -), each drawn only when
         * one of its optics terms is active. Absorbed when u >= rho; rho is
         * rounded to FP32 as the device stores it, so both sides compare the
         * same numbers. */
        const double rho = (double)(float)p->rho;
        double ue[4] = {0, 0, 0, 0};
        if (rho < 1.0 || p->slope > 0 || p->spec > 0) draw4(ray, (uint32_t)(2 * nd - 1), dsc->job_id, dsc->seed, ue);
        if (rho < 1.0 && ue[0] >= rho) { w->stats[5]++; return; }
        double us[4] = {0, 0, 0, 0};
        if (p->spec > 0) draw4(ray, (uint32_t)(2 * nd), dsc->job_id, dsc->seed, us);
        double rd[3];
        or_mirror_direction(d, h.n, p->slope, p->spec, ue[1], ue[2], ue[3], us[0], rd);
        memcpy(o, h.p, sizeof(o));
        memcpy(d, rd, sizeof(d));
        tmin = 1e-2; /* cuda/materials.cu:70 */
        depth = nd;
    }
}

static void* worker_main(void* arg) {
    or_worker* w = (or_worker*)arg;
    for (uint64_t i = w->begin; i < w->end && !w->err; ++i) trace_ray(w, i);
    return NULL;
}

int or_trace(const or_scene* s, const or_trace_desc* desc, const or_outputs* out, int32_t threads) {
    if (!s || !desc || !out || !s->sun_set) return -1;
    if (desc->max_depth < 1 || desc->max_depth > 32) return -1;
    uint64_t count = desc->ray_count ? desc->ray_count : desc->num_rays - desc->ray_begin;
    /* Halton positions take the 32-bit launch index of cuda/sun.cu:94-134 */
    if (s->sun.position_sampling == 1 && desc->ray_begin + count > (1ull << 32)) return -3;
    int32_t fe = desc->flux_element;
    if (fe < 0) {
        for (int i = s->n - 1; i >= 0; --i)
            if (s->prim[i].role == 1) { fe = i; break; }
    }
    int use_flux = out->flux && desc->flux_nx > 0 && desc->flux_ny > 0 && fe >= 0 && fe < s->n;
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > count) threads = count > 0 ? (int32_t)count : 1;
    or_worker* ws = (or_worker*)calloc((size_t)threads, sizeof(or_worker));
    pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    size_t nbins = use_flux ? (size_t)desc->flux_nx * (size_t)desc->flux_ny : 0;
    for (int t = 0; t < threads; ++t) {
        ws[t].s = s;
        ws[t].desc = desc;
        ws[t].out = out;
        ws[t].begin = count * (uint64_t)t / (uint64_t)threads;
        ws[t].end = count * (uint64_t)(t + 1) / (uint64_t)threads;
        ws[t].flux = use_flux ? (int64_t*)calloc(nbins, sizeof(int64_t)) : NULL;
        ws[t].hits = out->element_hits ? (uint64_t*)calloc((size_t)s->n, sizeof(uint64_t)) : NULL;
        ws[t].flux_elem = fe;
    }
    for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, worker_main, &ws[t]);
    worker_main(&ws[0]);
    for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
    int err = 0;
    for (int t = 0; t < threads; ++t) {
        err |= ws[t].err;
        if (use_flux) for (size_t b = 0; b < nbins; ++b) out->flux[b] += ws[t].flux[b];
        if (out->element_hits) for (int i = 0; i < s->n; ++i) out->element_hits[i] += ws[t].hits[i];
        if (out->stats) for (int k = 0; k < OR_STATS; ++k) out->stats[k] += ws[t].stats[k];
        free(ws[t].flux);
        free(ws[t].hits);
    }
    free(ws);
    free(th);
    return err ? -2 : 0;
}
