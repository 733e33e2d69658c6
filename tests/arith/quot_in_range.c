/* The tracer's quotInRange (artful_amd/csrc/hip/tracer.h) against IEEE division, on the CPU: the GPU's
 * steps (the hardware reciprocal r0, then fma's, then the sign of x on a zero) written with C's fmaf,
 * with r0 taken as 1/y correctly rounded and one ulp either way (the hardware's is within one ulp).
 * Over the range it claims: 2^-20 <= y <= 2^20 with 2^-100 <= |x| <= 2^60, x = +-0, or y = 1 with any
 * finite x, denormals included; random and structured cases. Prints the failures and their count;
 * exits 1 if there are any. Also checks what the drag's use of it rests on: 1 - 2x (an exact doubling
 * fused) plus a^2 x x is 1 for every |x| < 2^-26. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint32_t bitsOf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static float floatOf(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static float quotInRange(float x, float y, float r0) {
    const float r = fmaf(fmaf(-y, r0, 1.f), r0, r0);
    const float q0 = x * r;
    const float q1 = fmaf(fmaf(-y, q0, x), r, q0);
    return copysignf(fmaf(fmaf(-y, q1, x), r, q1), x);
}

static uint64_t state = 0x9E3779B97F4A7C15ull;
static uint32_t next(void) {
    state ^= state << 13, state ^= state >> 7, state ^= state << 17;
    return (uint32_t)(state >> 32);
}

/* a float of either sign with its unbiased exponent in [lo, hi] and a random significand */
static float randomIn(int lo, int hi) {
    const uint32_t e = (uint32_t)(lo + (int)(next() % (uint32_t)(hi - lo + 1)) + 127);
    return floatOf((next() & 0x80000000u) | (e << 23) | (next() & 0x7FFFFFu));
}

static long failures = 0;

static void check(float x, float y) {
    const float want = x / y, r = 1.f / y;
    const float r0s[3] = {nextafterf(r, 0.f), r, nextafterf(r, INFINITY)};
    for (int k = 0; k < 3; ++k) {
        const float got = quotInRange(x, y, r0s[k]);
        if (bitsOf(got) != bitsOf(want)) {
            if (failures < 10) printf("x %a y %a r0 %a: %a, IEEE %a\n", x, y, r0s[k], got, want);
            ++failures;
        }
    }
}

int main(void) {
    long cases = 0;
    for (int i = 0; i < 3000000; ++i, ++cases) check(randomIn(-100, 60), fabsf(randomIn(-20, 19)));
    /* divisors near 1 and the drag's: (1e-3, 2) */
    for (int i = 0; i < 1000000; ++i, ++cases) check(randomIn(-100, 60), fabsf(randomIn(-10, 0)));
    /* exact and near-exact quotients: x = q y rounded, and its neighbours */
    for (int i = 0; i < 1000000; ++i) {
        const float q = randomIn(-60, 30), y = fabsf(randomIn(-20, 19)), x = q * y;
        if (!(fabsf(x) >= 0x1p-100f && fabsf(x) <= 0x1p60f)) continue;
        check(x, y), check(nextafterf(x, 0.f), y), check(nextafterf(x, INFINITY), y), cases += 3;
    }
    /* powers of two and small integers as divisors */
    for (int e = -20; e <= 19; ++e)
        for (int i = 0; i < 2000; ++i, ++cases) check(randomIn(-100, 60), ldexpf(1.f + (float)(i % 7), e));
    /* zeros of both signs */
    for (int i = 0; i < 100000; ++i, cases += 2) {
        const float y = fabsf(randomIn(-20, 19));
        check(0.f, y), check(-0.f, y);
    }
    /* y = 1: every finite x, denormals included, one of each exponent and sign many times over */
    for (uint32_t e = 0; e < 255; ++e)
        for (int i = 0; i < 2000; ++i, ++cases) check(floatOf((next() & 0x80000000u) | (e << 23) | (next() & 0x7FFFFFu)), 1.f);
    /* the drag's divisor is 1 for |x| < 2^-26: max(fma(-2, x, 1) + (a^2 x) x, 1e-3) with a^2 <= 1 */
    long ones = 0;
    for (int i = 0; i < 2000000; ++i) {
        const float x = floatOf((next() & 0x80000000u) | (next() % bitsOf(0x1p-26f)));
        const float a2 = floatOf(next() % (bitsOf(1.f) + 1u));
        float d = fmaf(-2.f, x, 1.f) + a2 * x * x;
        d = d > 1e-3f ? d : 1e-3f;
        if (d != 1.f) {
            if (ones < 10) printf("divisor %a at x %a, a^2 %a\n", d, x, a2);
            ++ones;
        }
    }
    printf("%ld quotients, %ld wrong; %ld divisors not 1\n", cases, failures, ones);
    return failures || ones ? 1 : 0;
}
