// Host check of the renderer's facing test cross_z_sign (swegl_amd/csrc/core.h)
// against the reference's expression: normalize(cross(l, r)).z compared with
// 0 (points.cpp:32-37, points.hpp:80-90; renderer.cpp:72-75, 258).  Random
// edge vectors over many magnitudes, exact zeros, tiny and huge values,
// infinities and NaNs.  Exits 1 on any mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>

#include "../../swegl_amd/csrc/core.h"

static float ref_nz(sw::f3 l, sw::f3 r)
{
    float x = l.y * r.z - l.z * r.y, y = l.z * r.x - l.x * r.z, z = l.x * r.y - l.y * r.x;
    float len = std::sqrt(x * x + y * y + z * z);
    if (len != 0)
        z /= len;
    return z;
}

int main()
{
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    std::uniform_int_distribution<int> e(-140, 130);
    const float specials[] = {0.0f, -0.0f, 1e-45f, -1e-45f, 1e-38f, 3e38f, -3e38f,
                              std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                              std::numeric_limits<float>::quiet_NaN(), 1.0f, -1.0f, 0x1p-60f, 0x1p60f};
    long n = 0, bad = 0;
    auto rnd = [&]() -> float {
        int k = (int)(rng() % 16);
        if (k == 0)
            return specials[rng() % (sizeof specials / sizeof specials[0])];
        return std::ldexp(u(rng), k < 8 ? (int)(rng() % 24) - 12 : e(rng));
    };
    for (long i = 0; i < 5000000; i++) {
        sw::f3 l = {rnd(), rnd(), rnd()}, r = {rnd(), rnd(), rnd()};
        if (i % 5 == 0) {   // screen-space-like: x, y of hundreds, z of depth
            l = {u(rng) * 500, u(rng) * 500, u(rng)};
            r = {u(rng) * 500, u(rng) * 500, u(rng)};
        }
        const float z = ref_nz(l, r);
        for (int dir : {1, -1}) {
            const bool want = dir > 0 ? z > 0.0f : z < 0.0f;
            n++;
            if (sw::cross_z_sign(l, r, dir) != want && bad++ < 10)
                std::printf("mismatch dir %d l=(%a,%a,%a) r=(%a,%a,%a)\n", dir, l.x, l.y, l.z, r.x, r.y, r.z);
        }
    }
    std::printf("facing: %ld comparisons, %ld mismatches\n", n, bad);
    return bad ? 1 : 0;
}
