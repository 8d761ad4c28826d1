/* Exhaustive check of the division by a constant C (12 or 60) that flux3's
 * kernels use (hicar_amd/csrc/kernels_std.hip, div_const): for every 32-bit
 * pattern x, the three-operation quotient (q = x*R, e = fma(-q, C, x),
 * copysign(fma(e, R, q), x), R = 1/C rounded) taken where the kernel takes
 * it (|x| >= 2^-124 or x = +-0, finite) must equal the IEEE quotient x / C
 * bit for bit; elsewhere the kernel divides.  Prints the number of
 * mismatches and exits non-zero if there is one. */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float C, R;

static float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float div_const(float x)
{
    const uint32_t a = f_bits(x) & 0x7fffffffu;
    if (a - 1u < 0x017fffffu || a >= 0x7f800000u) return x / C;
    const float q = x * R;
    const float e = fmaf(-q, C, x);
    return copysignf(fmaf(e, R, q), x);
}

typedef struct { uint64_t lo, hi, bad, fast; } Job;

static void *run(void *p)
{
    Job *j = p;
    for (uint64_t u = j->lo; u < j->hi; u++) {
        const float x = bits_f((uint32_t)u);
        const uint32_t a = (uint32_t)u & 0x7fffffffu;
        if (!(a - 1u < 0x017fffffu || a >= 0x7f800000u)) j->fast++;
        volatile float ref = x / C;
        const float got = div_const(x);
        if (isnan(ref) ? !isnan(got) : f_bits(got) != f_bits(ref)) j->bad++;
    }
    return NULL;
}

int main(int argc, char **argv)
{
    C = (float)atof(argv[1]);
    R = 1.0f / C;
    enum { T = 8 };
    pthread_t th[T];
    Job jb[T];
    const uint64_t n = 1ull << 32;
    for (int t = 0; t < T; t++) {
        jb[t] = (Job){n * t / T, n * (t + 1) / T, 0, 0};
        pthread_create(&th[t], NULL, run, &jb[t]);
    }
    uint64_t bad = 0, fast = 0;
    for (int t = 0; t < T; t++) {
        pthread_join(th[t], NULL);
        bad += jb[t].bad;
        fast += jb[t].fast;
    }
    printf("C=%g R=%08x fast=%llu mismatches=%llu\n", C, f_bits(R), (unsigned long long)fast,
           (unsigned long long)bad);
    return bad != 0;
}
