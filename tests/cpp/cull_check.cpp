// Brute-force checks of the conservative tests that build candidate lists
// (st_common.h): cone_meets_box (the reflection-list BVH walk: a cone from a
// mirror's box), rect_sweep_misses (reflection lists of flat mirrors) and
// grid_range / cell_meets (sun-plane grid cells of flat primitives). Each may
// only drop candidates that no ray can reach; this program samples rays that do
// reach and counts every one a test would have dropped.
//
//   cull_check cone <seed> <phi_scale>    cone_meets_box with the cone angle
//                                          scaled by phi_scale
//   cull_check sweep <seed> <phi_scale>   rect_sweep_misses with the cone angle
//                                          scaled by phi_scale when building it
//   cull_check sweeprect <seed> <phi_scale> rect_sweep_misses_rect (flat candidates'
//                                          rectangles), the same scaling
//   cull_check grid <seed> <tan_scale>    grid_range / cell_meets with the cone's
//                                          tangent scaled by tan_scale
//   cull_check gridbox <seed> <tan_scale> the same for a curved primitive (box
//                                          footprint), rays ending in its box
//   cull_check reaim <seed> <phi_scale>   sphere_cone_reaches as k_reaim_lists
//                                          uses it (bisector-aimed mirror, axis
//                                          to the target), phi scaled
// A scale of 1 is the shipped test (violations must be 0); a scale below 1
// plants a too-narrow cone, which the sampling must catch.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "st_common.h"

using namespace st;

namespace {

std::mt19937_64 rng;
double uni() { return std::uniform_real_distribution<double>(0.0, 1.0)(rng); }
double uni(double lo, double hi) { return lo + (hi - lo) * uni(); }

void normalize(double v[3]) {
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) v[k] /= l;
}
void cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
double dot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

void random_unit(double v[3]) {
    const double z = uni(-1.0, 1.0), p = uni(0.0, 2.0 * M_PI), r = std::sqrt(1.0 - z * z);
    v[0] = r * std::cos(p);
    v[1] = r * std::sin(p);
    v[2] = z;
}

// f: 3x3 row-major with columns x, y, z (an orthonormal right-handed frame)
void random_frame(double f[9]) {
    double z[3], t[3], x[3], y[3];
    random_unit(z);
    random_unit(t);
    cross(t, z, x);
    normalize(x);
    cross(z, x, y);
    for (int k = 0; k < 3; ++k) {
        f[3 * k] = x[k];
        f[3 * k + 1] = y[k];
        f[3 * k + 2] = z[k];
    }
}

// a direction within angle phi of the unit axis a, uniform on the cap
void cone_direction(const double a[3], double phi, double u[3]) {
    double t[3], p[3], q[3];
    random_unit(t);
    cross(a, t, p);
    normalize(p);
    cross(a, p, q);
    const double cz = 1.0 - uni() * (1.0 - std::cos(phi)), sz = std::sqrt(std::fmax(0.0, 1.0 - cz * cz));
    const double ang = uni(0.0, 2.0 * M_PI);
    for (int k = 0; k < 3; ++k) u[k] = cz * a[k] + sz * (std::cos(ang) * p[k] + std::sin(ang) * q[k]);
    normalize(u);
}

// a float box holding every point of [c - h, c + h] (rounded outwards)
void box_around(const double c[3], const double h[3], float b[6]) {
    for (int k = 0; k < 3; ++k) {
        b[k] = std::nextafter((float)(c[k] - h[k]), -INFINITY);
        b[3 + k] = std::nextafter((float)(c[k] + h[k]), INFINITY);
    }
}

bool ray_hits_box(const double o[3], const double d[3], const float b[6]) {
    double t0 = 0.0, t1 = 1e300;
    for (int k = 0; k < 3; ++k) {
        if (std::fabs(d[k]) < 1e-300) {
            if (o[k] < b[k] || o[k] > b[3 + k]) return false;
            continue;
        }
        double ta = (b[k] - o[k]) / d[k], tb = (b[3 + k] - o[k]) / d[k];
        if (ta > tb) std::swap(ta, tb);
        t0 = std::fmax(t0, ta);
        t1 = std::fmin(t1, tb);
    }
    return t0 <= t1;
}

// Reflection sweep: rays leave a random rectangle R in directions within phi
// of a. (i) Boxes placed around points those rays pass through must never be
// dropped; (ii) every box that is dropped is shot at with many such rays, none
// of which may enter it.
int check_sweep(double phi_scale) {
    long reach_boxes = 0, reach_dropped = 0, dropped = 0, dropped_hit = 0, boxes = 0;
    for (int cfg = 0; cfg < 150; ++cfg) {
        double f[9];
        random_frame(f);
        const double pc[3] = {uni(-800, 800), uni(-800, 800), uni(-20, 40)};
        const double hw = uni(0.3, 6.0), hh = uni(0.3, 6.0);
        const double phi = std::exp(uni(std::log(2e-3), std::log(0.3)));
        double a[3];
        random_unit(a);
        RectSweep S;
        rect_sweep_init(S, pc, hw, hh, f, a, phi * phi_scale);
        auto ray = [&](double o[3], double u[3]) {
            const double x = uni(-hw, hw), y = uni(-hh, hh);
            for (int k = 0; k < 3; ++k) o[k] = pc[k] + x * f[3 * k] + y * f[3 * k + 1];
            cone_direction(a, phi, u);
        };
        for (int i = 0; i < 1000; ++i) {  // (i) reachable boxes
            double o[3], u[3];
            ray(o, u);
            const double t = std::exp(uni(std::log(1e-2), std::log(400.0)));
            const double q[3] = {o[0] + t * u[0], o[1] + t * u[1], o[2] + t * u[2]};
            const double h[3] = {uni(0.0, 2.0), uni(0.0, 2.0), uni(0.0, 2.0)};
            double c[3];
            for (int k = 0; k < 3; ++k) c[k] = q[k] + uni(-h[k], h[k]);  // q inside, not centred
            float b[6];
            box_around(c, h, b);
            ++reach_boxes;
            if (rect_sweep_misses(S, b)) ++reach_dropped;
        }
        for (int i = 0; i < 100; ++i) {  // (ii) random boxes near R
            const double c[3] = {pc[0] + uni(-40, 40), pc[1] + uni(-40, 40), pc[2] + uni(-40, 40)};
            const double h[3] = {uni(0.1, 6.0), uni(0.1, 6.0), uni(0.1, 6.0)};
            float b[6];
            box_around(c, h, b);
            ++boxes;
            if (!rect_sweep_misses(S, b)) continue;
            ++dropped;
            for (int r = 0; r < 1000; ++r) {
                double o[3], u[3];
                ray(o, u);
                if (ray_hits_box(o, u, b)) {
                    ++dropped_hit;
                    break;
                }
            }
        }
    }
    std::printf("sweep reachable_boxes %ld reachable_dropped %ld boxes %ld dropped %ld dropped_but_hit %ld\n",
                reach_boxes, reach_dropped, boxes, dropped, dropped_hit);
    return 0;
}

// exact ray-rectangle test (rectangle centre c, frame f columns x, y, normal;
// half sizes hw, hh), forward hits only
bool ray_hits_rect(const double o[3], const double d[3], const double c[3], const double f[9], double hw, double hh) {
    const double n[3] = {f[2], f[5], f[8]}, X[3] = {f[0], f[3], f[6]}, Y[3] = {f[1], f[4], f[7]};
    const double dn = dot(d, n);
    if (std::fabs(dn) < 1e-300) return false;
    const double oc[3] = {c[0] - o[0], c[1] - o[1], c[2] - o[2]};
    const double t = dot(oc, n) / dn;
    if (!(t > 0.0)) return false;
    const double w[3] = {o[0] + t * d[0] - c[0], o[1] + t * d[1] - c[1], o[2] + t * d[2] - c[2]};
    return std::fabs(dot(w, X)) <= hw && std::fabs(dot(w, Y)) <= hh;
}

// The same sweep against flat candidates' rectangles (rect_sweep_misses_rect):
// (i) rectangles through points those rays pass are never dropped; (ii) every
// dropped rectangle is missed by many such rays.
int check_sweep_rect(double phi_scale) {
    long reach = 0, reach_dropped = 0, dropped = 0, dropped_hit = 0, rects = 0, box_kept = 0;
    for (int cfg = 0; cfg < 150; ++cfg) {
        double f[9];
        random_frame(f);
        const double pc[3] = {uni(-800, 800), uni(-800, 800), uni(-20, 40)};
        const double hw = uni(0.3, 6.0), hh = uni(0.3, 6.0);
        const double phi = std::exp(uni(std::log(2e-3), std::log(0.3)));
        double a[3];
        random_unit(a);
        RectSweep S;
        rect_sweep_init(S, pc, hw, hh, f, a, phi * phi_scale);
        auto ray = [&](double o[3], double u[3]) {
            const double x = uni(-hw, hw), y = uni(-hh, hh);
            for (int k = 0; k < 3; ++k) o[k] = pc[k] + x * f[3 * k] + y * f[3 * k + 1];
            cone_direction(a, phi, u);
        };
        auto cand = [&](const double c[3], const double g[9], double chw, double chh, RectCand& R) {
            rect_cand_init(R, c, chw, chh, g);
        };
        for (int i = 0; i < 1000; ++i) {  // (i) reachable rectangles
            double o[3], u[3];
            ray(o, u);
            const double t = std::exp(uni(std::log(1e-2), std::log(400.0)));
            const double q[3] = {o[0] + t * u[0], o[1] + t * u[1], o[2] + t * u[2]};
            double g[9];
            random_frame(g);
            const double chw = uni(0.2, 6.0), chh = uni(0.2, 6.0), x = uni(-chw, chw), y = uni(-chh, chh);
            double c[3];
            for (int k = 0; k < 3; ++k) c[k] = q[k] - x * g[3 * k] - y * g[3 * k + 1];  // q on it, anywhere
            RectCand R;
            cand(c, g, chw, chh, R);
            ++reach;
            if (rect_sweep_misses_rect(S, R)) ++reach_dropped;
        }
        for (int i = 0; i < 100; ++i) {  // (ii) random rectangles near R
            const double c[3] = {pc[0] + uni(-40, 40), pc[1] + uni(-40, 40), pc[2] + uni(-40, 40)};
            double g[9];
            random_frame(g);
            const double chw = uni(0.2, 6.0), chh = uni(0.2, 6.0);
            RectCand R;
            cand(c, g, chw, chh, R);
            ++rects;
            if (!rect_sweep_misses_rect(S, R)) continue;
            ++dropped;
            // how many of the dropped ones the box test kept: the gain of the rectangle test
            double h[3];
            for (int k = 0; k < 3; ++k)
                h[k] = std::fabs(g[3 * k]) * chw + std::fabs(g[3 * k + 1]) * chh;
            float b[6];
            box_around(c, h, b);
            if (!rect_sweep_misses(S, b)) ++box_kept;
            for (int r = 0; r < 1000; ++r) {
                double o[3], u[3];
                ray(o, u);
                if (ray_hits_rect(o, u, c, g, chw, chh)) {
                    ++dropped_hit;
                    break;
                }
            }
        }
    }
    std::printf("sweeprect reachable %ld reachable_dropped %ld rects %ld dropped %ld dropped_box_kept %ld "
                "dropped_but_hit %ld\n",
                reach, reach_dropped, rects, dropped, box_kept, dropped_hit);
    return 0;
}

// Reflection-list walk: rays leave a random box hb (a mirror's traversal box)
// in directions within phi of a; a candidate box B is kept when the cone from
// the origin meets B - hb (the walk's Minkowski difference, pieces bounded as
// in a scene 2 km across). (i) Boxes around points such rays pass through must
// be kept; (ii) boxes that are dropped are shot at with such rays, which must
// all miss.
int check_cone(double phi_scale) {
    long reach_boxes = 0, reach_dropped = 0, dropped = 0, dropped_hit = 0, boxes = 0;
    for (int cfg = 0; cfg < 150; ++cfg) {
        const double c0[3] = {uni(-800, 800), uni(-800, 800), uni(-20, 40)};
        const double h0[3] = {uni(0.2, 6.0), uni(0.2, 6.0), uni(0.2, 6.0)};
        float hb[6];
        box_around(c0, h0, hb);
        const double phi = std::exp(uni(std::log(2e-3), std::log(0.3)));
        double a[3];
        random_unit(a);
        Cone C;
        for (int k = 0; k < 3; ++k) C.a[k] = a[k];
        const double ph = phi * phi_scale;
        C.tp = std::tan(ph);
        C.cp = std::cos(ph);
        C.sp = std::sin(ph);
        const double reach = 2.0 * std::sqrt(3.0) * 2000.0 + 2.0 * kConeFirst;
        C.ratio = std::pow(reach / kConeFirst, 1.0 / (kConePieces - 2));
        auto meets = [&](const float b[6]) {
            const double dlo[3] = {(double)b[0] - hb[3], (double)b[1] - hb[4], (double)b[2] - hb[5]};
            const double dhi[3] = {(double)b[3] - hb[0], (double)b[4] - hb[1], (double)b[5] - hb[2]};
            return cone_meets_box(C, dlo, dhi);
        };
        auto ray = [&](double o[3], double u[3]) {
            for (int k = 0; k < 3; ++k) o[k] = uni((double)hb[k], (double)hb[3 + k]);
            cone_direction(a, phi, u);
        };
        for (int i = 0; i < 1000; ++i) {  // (i) reachable boxes
            double o[3], u[3];
            ray(o, u);
            const double t = std::exp(uni(std::log(1e-2), std::log(1500.0)));
            const double h[3] = {uni(0.0, 2.0), uni(0.0, 2.0), uni(0.0, 2.0)};
            double c[3];
            for (int k = 0; k < 3; ++k) c[k] = o[k] + t * u[k] + uni(-h[k], h[k]);
            float b[6];
            box_around(c, h, b);
            ++reach_boxes;
            if (!meets(b)) ++reach_dropped;
        }
        for (int i = 0; i < 100; ++i) {  // (ii) random boxes around the apex
            const double c[3] = {c0[0] + uni(-60, 60), c0[1] + uni(-60, 60), c0[2] + uni(-60, 60)};
            const double h[3] = {uni(0.1, 6.0), uni(0.1, 6.0), uni(0.1, 6.0)};
            float b[6];
            box_around(c, h, b);
            ++boxes;
            if (meets(b)) continue;
            ++dropped;
            for (int r = 0; r < 1000; ++r) {
                double o[3], u[3];
                ray(o, u);
                if (ray_hits_box(o, u, b)) {
                    ++dropped_hit;
                    break;
                }
            }
        }
    }
    std::printf("cone reachable_boxes %ld reachable_dropped %ld boxes %ld dropped %ld dropped_but_hit %ld\n",
                reach_boxes, reach_dropped, boxes, dropped, dropped_hit);
    return 0;
}

// Sun-plane grid: a sun plane for a random sun, a random flat rectangle below
// it; rays that hit the rectangle, traced back along directions within the sun
// cone to their sun-plane origin, must start in a listed cell. With flat =
// false the primitive is listed as a curved one (the box footprint only) and
// the sampled rays end anywhere in its traversal box.
int check_grid(double tan_scale, bool flat) {
    long hits = 0, unlisted = 0, listed_cells = 0, range_cells = 0;
    for (int cfg = 0; cfg < 1000; ++cfg) {
        const double el = uni(5.0, 89.0) * M_PI / 180.0, az = uni(0.0, 2.0 * M_PI);
        double s[3] = {std::sin(az) * std::cos(el), std::cos(az) * std::cos(el), std::sin(el)};
        double e1[3], e2[3];
        const double up[3] = {0.0, 0.0, 1.0}, east[3] = {1.0, 0.0, 0.0};
        cross(std::fabs(s[2]) < 0.9 ? up : east, s, e1);
        normalize(e1);
        cross(s, e1, e2);
        const double D = uni(200.0, 1500.0), L1 = uni(100.0, 2000.0), L2 = uni(100.0, 2000.0);
        SunFrame F;
        std::memset(&F, 0, sizeof(F));
        for (int k = 0; k < 3; ++k) {
            F.sun[k] = s[k];
            F.v0[k] = D * s[k] - 0.5 * L1 * e1[k] - 0.5 * L2 * e2[k];
            F.v1[k] = F.v0[k] + L1 * e1[k];
            F.v3[k] = F.v0[k] + L2 * e2[k];
            F.v2[k] = F.v1[k] + L2 * e2[k];
        }
        F.distance = D;
        F.grid_rows = (cfg & 1) ? kHaltonRows : kGridRes;
        const double alpha = std::exp(uni(std::log(1e-3), std::log(2e-2)));
        // a flat rectangle whose centre projects inside the plane, below it
        double f[9];
        random_frame(f);
        const double hw = uni(0.5, 8.0), hh = uni(0.5, 8.0);
        // axial distance below the plane from 10 m to beyond the plane's own distance
        const double depth = uni(10.0, D + 300.0), pa = uni(0.05, 0.95), pb = uni(0.05, 0.95);
        double cen[3];
        for (int k = 0; k < 3; ++k) cen[k] = F.v0[k] + pa * L1 * e1[k] + pb * L2 * e2[k] - depth * s[k];
        float4 prim[4];
        prim[0] = {(float)cen[0], (float)cen[1], (float)cen[2], (float)hw};
        prim[1] = {(float)f[0], (float)f[3], (float)f[6], (float)hh};
        prim[2] = {(float)f[1], (float)f[4], (float)f[7], 0.f};
        prim[3] = {(float)f[2], (float)f[5], (float)f[8], 0.f};
        const uint32_t meta = flat ? kKindFlat : kKindCylinder;
        // traversal box: the rectangle's corners, padded as element boxes are
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 3; ++k) {
                const double v = cen[k] + ((c & 1) ? hw : -hw) * f[3 * k] + ((c & 2) ? hh : -hh) * f[3 * k + 1];
                lo[k] = std::fmin(lo[k], v);
                hi[k] = std::fmax(hi[k], v);
            }
        float bounds[6];
        for (int k = 0; k < 3; ++k) {
            const double pad = 1e-3 + 2e-6 * std::fmax(std::fabs(lo[k]), std::fabs(hi[k]));
            bounds[k] = (float)(lo[k] - pad);
            bounds[3 + k] = (float)(hi[k] + pad);
        }
        Footprint fp;
        const bool any = grid_range(bounds, prim, &meta, 0, F, std::tan(alpha) * tan_scale, fp);
        if (any)
            for (int iy = fp.r[2]; iy <= fp.r[3]; ++iy)
                for (int ix = fp.r[0]; ix <= fp.r[1]; ++ix) {
                    ++range_cells;
                    listed_cells += cell_meets(fp, ix, iy);
                }
        const double w[3] = {-s[0], -s[1], -s[2]};
        for (int i = 0; i < 4000; ++i) {
            double q[3], u[3];
            const double x = uni(-hw, hw), y = uni(-hh, hh);
            for (int k = 0; k < 3; ++k)
                q[k] = flat ? cen[k] + x * f[3 * k] + y * f[3 * k + 1] : uni((double)bounds[k], (double)bounds[3 + k]);
            cone_direction(w, alpha, u);
            const double t = (D - dot(q, s)) / -dot(u, s);  // back along u to the plane x . s = D
            double o[3], r[3];
            for (int k = 0; k < 3; ++k) {
                o[k] = q[k] - t * u[k];
                r[k] = o[k] - F.v0[k];
            }
            const double a = dot(r, e1) / L1, b = dot(r, e2) / L2;
            if (!(a >= 0.0 && a < 1.0 && b >= 0.0 && b < 1.0)) continue;
            const int ix = (int)(a * kGridRes), iy = (int)(b * F.grid_rows);
            ++hits;
            const bool ok = any && ix >= fp.r[0] && ix <= fp.r[1] && iy >= fp.r[2] && iy <= fp.r[3] && cell_meets(fp, ix, iy);
            if (!ok) ++unlisted;
        }
    }
    std::printf("%s hits %ld unlisted %ld range_cells %ld listed_cells %ld\n", flat ? "grid" : "gridbox", hits,
                unlisted, range_cells, listed_cells);
    return 0;
}

}  // namespace

// Re-aim candidates (k_reaim_lists): a mirror A (rectangle w x h) aimed by the
// bisector rule at a target for a random sun reflects sun rays (directions
// within phi of reflect(-sun, n)) from its rectangle. (i) Spheres around points
// those rays pass through must pass sphere_cone_reaches with the kernel's
// arguments (A's circumscribed sphere + 1 mm, axis from A's centre to the
// target, phi + 1e-6); (ii) spheres that fail are shot at with many such rays,
// none of which may enter them.
int check_reaim(double phi_scale) {
    long reach = 0, reach_dropped = 0, spheres = 0, dropped = 0, dropped_hit = 0;
    for (int cfg = 0; cfg < 200; ++cfg) {
        const double pc[3] = {uni(-800, 800), uni(-800, 800), uni(-3, 3)};
        const double target[3] = {uni(-50, 50), uni(-50, 50), uni(20, 250)};
        const double hw = uni(0.3, 6.0), hh = uni(0.3, 6.0);
        const double phi = std::exp(uni(std::log(2e-3), std::log(0.2)));
        double sun[3];
        do random_unit(sun);
        while (sun[2] < 0.05);
        double aim[3];
        bisector_aim_point(pc, sun, target, aim);
        double n[3] = {aim[0] - pc[0], aim[1] - pc[1], aim[2] - pc[2]};
        normalize(n);
        double x[3], y[3], tmp[3];
        random_unit(tmp);
        cross(tmp, n, x);
        normalize(x);
        cross(n, x, y);
        // the traced reflection axis, reflect(-sun, n)
        double a[3];
        const double dn = -dot(sun, n);
        for (int k = 0; k < 3; ++k) a[k] = -sun[k] - 2.0 * dn * n[k];
        normalize(a);
        // the kernel's arguments
        double t[3] = {target[0] - pc[0], target[1] - pc[1], target[2] - pc[2]};
        normalize(t);
        const double ra = std::sqrt(hw * hw + hh * hh) + 1e-3, pk = phi * phi_scale + 1e-6;
        const double cp = std::cos(pk), sp = std::sin(pk);
        auto ray = [&](double o[3], double u[3]) {
            const double px = uni(-hw, hw), py = uni(-hh, hh);
            for (int k = 0; k < 3; ++k) o[k] = pc[k] + px * x[k] + py * y[k];
            cone_direction(a, phi, u);
        };
        for (int i = 0; i < 1000; ++i) {  // (i) reachable spheres
            double o[3], u[3];
            ray(o, u);
            const double s = std::exp(uni(std::log(1e-2), std::log(1500.0)));
            const double rb = uni(0.05, 8.0);
            double v[3];
            random_unit(v);
            const double off = uni(0.0, rb);
            const double c[3] = {o[0] + s * u[0] + off * v[0], o[1] + s * u[1] + off * v[1], o[2] + s * u[2] + off * v[2]};
            ++reach;
            if (!sphere_cone_reaches(pc, ra, t, cp, sp, c, rb + 1e-3)) ++reach_dropped;
        }
        for (int i = 0; i < 200; ++i) {  // (ii) random spheres near A
            const double c[3] = {pc[0] + uni(-60, 60), pc[1] + uni(-60, 60), pc[2] + uni(-20, 60)};
            const double rb = uni(0.1, 8.0);
            ++spheres;
            if (sphere_cone_reaches(pc, ra, t, cp, sp, c, rb + 1e-3)) continue;
            ++dropped;
            for (int r = 0; r < 1000; ++r) {
                double o[3], u[3];
                ray(o, u);
                // closest approach of the ray to the sphere's centre
                const double w[3] = {c[0] - o[0], c[1] - o[1], c[2] - o[2]};
                const double s = std::fmax(0.0, dot(w, u));
                const double q[3] = {o[0] + s * u[0] - c[0], o[1] + s * u[1] - c[1], o[2] + s * u[2] - c[2]};
                if (dot(q, q) <= rb * rb) {
                    ++dropped_hit;
                    break;
                }
            }
        }
    }
    std::printf("reaim reachable %ld reachable_dropped %ld spheres %ld dropped %ld dropped_but_hit %ld\n", reach,
                reach_dropped, spheres, dropped, dropped_hit);
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: cull_check cone|sweep|sweeprect|grid|gridbox <seed> <scale>\n");
        return 2;
    }
    rng.seed(std::strtoull(argv[2], nullptr, 10));
    const double scale = std::atof(argv[3]);
    if (!std::strcmp(argv[1], "cone")) return check_cone(scale);
    if (!std::strcmp(argv[1], "sweep")) return check_sweep(scale);
    if (!std::strcmp(argv[1], "sweeprect")) return check_sweep_rect(scale);
    if (!std::strcmp(argv[1], "grid")) return check_grid(scale, true);
    if (!std::strcmp(argv[1], "gridbox")) return check_grid(scale, false);
    if (!std::strcmp(argv[1], "reaim")) return check_reaim(scale);
    return 2;
}
