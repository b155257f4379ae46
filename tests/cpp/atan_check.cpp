// Host check of st::atan2_poly (st_common.h), the atan2 of the cylindrical
// flux bin: largest error in radians against the libm atan2 of the same FP32
// inputs over a dense sweep of angles and radii, the axes and signed zeros.
//   atan_check          the shipped function
//   atan_check planted  the same polynomial without its last coefficient
//                       (a mistake the check must catch)
#include <cmath>
#include <cstdio>
#include <cstring>

#include "st_common.h"

static float planted_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mx > 0.f ? mn / mx : 0.f;
    const float s = a * a;
    float r = -0.0161657367f;
    r = fmaf(s, r, 0.0429096138f);
    r = fmaf(s, r, -0.0752896400f);
    r = fmaf(s, r, 0.1065626393f);
    r = fmaf(s, r, -0.1420889944f);
    r = fmaf(s, r, 0.1999355085f);
    r = fmaf(s, r, -0.3333314528f);
    r = fmaf(s, r, 1.0f);
    r *= a;
    if (ay > ax) r = 1.5707963267948966f - r;
    if (x < 0.f) r = 3.141592653589793f - r;
    return copysignf(r, y);
}

int main(int argc, char** argv) {
    const bool planted = argc > 1 && std::strcmp(argv[1], "planted") == 0;
    auto f = [&](float y, float x) { return planted ? planted_atan2(y, x) : st::atan2_poly(y, x); };
    double worst = 0, worst_th = 0;
    long n = 0;
    const int kAngles = 2000003;
    const float radii[] = {1e-3f, 0.37f, 1.0f, 8.0f, 123.4f, 1e4f};
    for (float r : radii)
        for (int i = 0; i < kAngles; ++i) {
            const double th = -M_PI + 2.0 * M_PI * (i + 0.5) / kAngles;
            const float y = (float)(r * std::sin(th)), x = (float)(r * std::cos(th));
            double d = std::fabs((double)f(y, x) - std::atan2((double)y, (double)x));
            if (d > M_PI) d = 2 * M_PI - d;  // +pi and -pi are the same direction
            if (d > worst) {
                worst = d;
                worst_th = th;
            }
            ++n;
        }
    // axes and signed zeros: exact
    const float cases[][3] = {{0.f, 1.f, 0.f}, {1.f, 0.f, (float)(M_PI / 2)}, {-1.f, 0.f, (float)(-M_PI / 2)},
                              {0.f, -1.f, (float)M_PI}, {-0.f, -1.f, (float)-M_PI}, {0.f, 0.f, 0.f},
                              {1.f, 1.f, (float)(M_PI / 4)}, {-1.f, -1.f, (float)(-3 * M_PI / 4)}};
    int bad = 0;
    for (const auto& c : cases)
        if (std::fabs((double)f(c[0], c[1]) - (double)c[2]) > 2.5e-7) ++bad;
    std::printf("samples %ld max_error_rad %.3e at_theta %.6f axis_case_failures %d\n", n, worst, worst_th, bad);
    return 0;
}
