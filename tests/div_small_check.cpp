// CPU check of csrc/div_small.h against the IEEE division (built and run by
// tests/test_host.py with g++ -O2 -fopenmp -ffp-contract=off).  Prints the number of
// mismatches and exits 1 when there are any.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../flub_amd/csrc/div_small.h"

static float from_bits(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static uint32_t bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

int main()
{
    unsigned long long bad = 0, checked = 0;
    for (int d = 1; d <= 6; ++d) {
        const float df = (float)d, y = 1.0f / df;
        // every x with 2^-127 <= |x| < 2^-120: subnormal and normal quotients
        // around the FMA path's bound 2^-123 (the function is odd in x)
        unsigned long long band_bad = 0;
#pragma omp parallel for reduction(+ : band_bad) schedule(static)
        for (long long u = 0x00400000ll; u < 0x03800000ll; ++u) {
            const float x = from_bits((uint32_t)u);
            if (bits(div_small(x, df, y)) != bits(x / df)) ++band_bad;
        }
        bad += band_bad;
        checked += 0x03800000ull - 0x00400000ull;
        // a deterministic sample of all finite x, both signs, and the specials
        uint64_t s = 0x9E3779B97F4A7C15ull * (uint64_t)d;
        for (int t = 0; t < 2000000; ++t) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            float x = from_bits((uint32_t)s);
            if (x != x) continue;
            const float q = div_small(x, df, y), e = x / df;
            if (bits(q) != bits(e)) ++bad;
            ++checked;
        }
        const float sp[] = {0.0f, -0.0f, 1.0f / 0.0f, -1.0f / 0.0f, 0x1p-149f, -0x1p-149f, 0x1.fffffep127f};
        for (float x : sp) {
            if (bits(div_small(x, df, y)) != bits(x / df)) ++bad;
            ++checked;
        }
    }
    printf("checked %llu, mismatches %llu\n", checked, bad);
    return bad ? 1 : 0;
}
