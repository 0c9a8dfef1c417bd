// Host check of the renderer's pow32 (swegl_amd/csrc/shade.h) against
// (float)pow((double)s, 32.0) from the C library, the expression of
// pixel_shaders.cpp:193.  pow32 is plain IEEE double arithmetic, which the
// host and gfx950 evaluate identically.  Prints the mismatch count and how
// often the double-double fallback ran; exits 1 on any mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../swegl_amd/csrc/shade.h"

static bool fast_path(float s)
{
    double p = (double)s;
    for (int i = 0; i < 5; i++)
        p *= p;
    return (float)(p * (1.0 - 0x1p-46)) == (float)(p * (1.0 + 0x1p-46));
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 2000000L;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<float> uni(0.0f, 1.2f);
    std::uniform_real_distribution<double> expo(-150.0, 1.0);
    long bad = 0, slow = 0, total = 0;
    auto check = [&](float s) {
        float want = (float)std::pow((double)s, 32.0);
        float got = sw::pow32(s);
        total++;
        if (!fast_path(s))
            slow++;
        if (std::memcmp(&want, &got, 4) != 0) {
            if (bad < 10)
                std::printf("mismatch s=%a want=%a got=%a\n", s, want, got);
            bad++;
        }
    };
    for (long i = 0; i < n; i++) {
        check(uni(rng));
        check((float)std::exp2(expo(rng)));
    }
    // every float in [1/16, 1.2], where s^32 is a normal float (this
    // includes the few values whose s^32 lies next to a float rounding
    // boundary, e.g. 0x1.555c82p-2, which take the double-double path),
    // and a sample of the denormal inputs
    for (uint32_t b = 0x3D800000u; b < 0x3F99999Au; b++) {
        float s;
        std::memcpy(&s, &b, 4);
        check(s);
    }
    for (uint32_t b = 0; b < 0x00100000u; b += 97) {
        float s;
        std::memcpy(&s, &b, 4);
        check(s);
    }
    std::printf("pow32: %ld values, %ld mismatches, double-double fallback on %ld\n", total, bad, slow);
    return bad ? 1 : 0;
}
