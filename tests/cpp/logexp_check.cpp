// Sweeps the pyramid transform of float images (enblend-enfuse_amd/csrc/
// logexp.h) against the C library's log1p and expm1: every float exponent
// with random mantissas for log1p, and the pyramid values of those floats
// plus a uniform sweep of [0, 90] for expm1.  Prints the largest error in
// units in the last place of the double result and how many float results
// of the inverse differ from the C library's.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../enblend-enfuse_amd/csrc/logexp.h"

static double ulp(double x)
{
    x = fabs(x);
    if (x == 0.0) return 0x1p-1074;
    int e;
    frexp(x, &e);
    return ldexp(1.0, e - 53);
}

static uint64_t state = 0x9e3779b97f4a7c15ull;
static uint32_t next()
{
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return (uint32_t)(state >> 11);
}

int main()
{
    double log_err = 0.0, exp_err = 0.0;
    long n = 0, float_diff = 0, nexp = 0;
    for (int e = 0; e < 255; ++e)
        for (int i = 0; i < 40000; ++i) {
            const uint32_t bits = ((uint32_t)e << 23) | (next() & 0x7fffffu);
            float f;
            memcpy(&f, &bits, 4);
            const double v = f;
            const double want = log1p(v);
            const double got = enbl::log1p_nonneg(v);
            const double err = fabs(got - want) / ulp(want);
            if (err > log_err) log_err = err;
            ++n;
            // the inverse on the pyramid value of this sample
            const double c = 1.0 + want;
            const double we = expm1(c - 1.0), ge = enbl::expm1_nonneg(c - 1.0);
            if (isfinite(we)) {
                const double ee = fabs(ge - we) / ulp(we);
                if (ee > exp_err) exp_err = ee;
            }
            float_diff += (float)we != (float)ge;
            ++nexp;
        }
    for (long i = 0; i <= 9000000; ++i) {
        const double x = 90.0 * (double)i / 9000000.0;
        const double we = expm1(x), ge = enbl::expm1_nonneg(x);
        const double ee = fabs(ge - we) / ulp(we);
        if (ee > exp_err) exp_err = ee;
        float_diff += (float)we != (float)ge;
        ++nexp;
    }
    // special values
    int special_ok = enbl::log1p_nonneg(0.0) == 0.0 && enbl::log1p_nonneg((double)INFINITY) == (double)INFINITY &&
                     enbl::expm1_nonneg(0.0) == 0.0 && enbl::expm1_nonneg(1000.0) == (double)INFINITY &&
                     enbl::log1p_nonneg(0x1p-149) == 0x1p-149;
    printf("log1p samples %ld max_ulp %.3f\n", n, log_err);
    printf("expm1 samples %ld max_ulp %.3f float_differences %ld\n", nexp, exp_err, float_diff);
    printf("special %d\n", special_ok);
    return 0;
}
