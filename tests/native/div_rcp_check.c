/* Host restatement of the device's division by a precomputed reciprocal
 * (div_rcp and div_rcp_acc in nemolite2d-cuda_amd/csrc/nemo2d_core.h):
 * q0 = RN(x*y), two FMA remainder corrections, the quotient given x's sign,
 * and the IEEE division for |x| outside [2^-960, 2^960) other than zero; the
 * unguarded form's quotient wherever its guard holds.  Checked bit for bit
 * against x / d on random operands over the whole exponent range, on
 * quotients within 2^-64 relative of a rounding midpoint (the hard cases of
 * correct rounding), and on zeros, subnormals, infinities and NaN.  Also the
 * one-correction form (div_rcp1_acc), which the library uses only when every
 * divisor d of the grid has |d*RN(1/d) - 1| <= 2^-54 (rcp_one_corr): then
 * RN(x*y) is already a faithful quotient and one correction rounds it
 * correctly (Markstein); it is checked on such divisors.  (On divisors that
 * do not qualify, RN(x*y) can be 1-2 ulps off for quotients whose
 * significand is near 2, and one correction then rounds wrongly only for
 * quotients within ~2^-52 ulp of a rounding midpoint, which random
 * operands do not reach; those mismatches are counted, not expected.)
 * Usage: div_rcp_check <n> <seed>; prints tried, mismatches, guarded,
 * one-correction tried, and one-correction mismatches on divisors that do
 * not qualify. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double div_rcp(double x, double d, double y) {
    double q = x * y;
    double r = fma(-d, q, x);
    q = fma(r, y, q);
    r = fma(-d, q, x);
    q = copysign(fma(r, y, q), x);
    uint64_t b;
    memcpy(&b, &x, 8);
    const unsigned e = (unsigned)(b >> 52) & 0x7ffu;
    if (e - 63u >= 1920u && x != 0.0) q = x / d;
    return q;
}

/* The unguarded form (div_rcp_acc): corrections written q - (d*q - x)*y,
 * exact when the numerator is zero or normal with frexp exponent -959..960;
 * returns 1 and the quotient when that holds, else 0. */
static int div_rcp_acc(double x, double d, double y, double* out) {
    double q = x * y;
    double r = fma(d, q, -x);
    q = fma(-r, y, q);
    r = fma(d, q, -x);
    q = fma(-r, y, q);
    *out = q;
    if (isnan(x) || isinf(x) || (x != 0.0 && fabs(x) < 0x1p-1022)) return 0;
    int e;
    frexp(x, &e);
    return e >= -959 && e <= 960;
}

/* The one-correction form: q0 = RN(x*y), then q0 - (d*q0 - x)*y. */
static int div_rcp1_acc(double x, double d, double y, double* out) {
    double q = x * y;
    const double r = fma(d, q, -x);
    q = fma(-r, y, q);
    *out = q;
    if (isnan(x) || isinf(x) || (x != 0.0 && fabs(x) < 0x1p-1022)) return 0;
    int e;
    frexp(x, &e);
    return e >= -959 && e <= 960;
}

/* The host's test of a divisor for the one-correction form (the library's
 * rcp_one_corr): |d*RN(1/d) - 1| <= 2^-54, with a margin for the rounding
 * of the FMA that evaluates it. */
static int one_corr_ok(double d) {
    const double y = 1.0 / d;
    return fabs(fma(d, y, -1.0)) <= 0x1p-54 * (1.0 - 0x1p-40);
}

static uint64_t s_state;
static uint64_t rnd(void) { /* splitmix64 */
    uint64_t z = (s_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static double mk(int exp, uint64_t mant, int neg) {
    uint64_t b = ((uint64_t)(neg ? 1 : 0) << 63) | ((uint64_t)(exp + 1023) << 52) |
                 (mant & 0xfffffffffffffull);
    double v;
    memcpy(&v, &b, 8);
    return v;
}
static long bad = 0, tried = 0, guarded = 0, tried1 = 0, fails1_unqualified = 0;
static void check(double x, double d) {
    const double y = 1.0 / d;
    const double a = div_rcp(x, d, y), b = x / d;
    ++tried;
    double c1;
    if (div_rcp1_acc(x, d, y, &c1)) {
        if (one_corr_ok(d)) {
            ++tried1;
            if (memcmp(&c1, &b, 8) != 0) {
                if (bad < 10) fprintf(stderr, "one-correction mismatch x=%a d=%a: %a vs %a\n", x, d, c1, b);
                ++bad;
            }
        } else if (memcmp(&c1, &b, 8) != 0) {
            ++fails1_unqualified;
        }
    }
    if (memcmp(&a, &b, 8) != 0 && !(isnan(a) && isnan(b))) {
        if (bad < 10) fprintf(stderr, "mismatch x=%a d=%a: %a vs %a\n", x, d, a, b);
        ++bad;
    }
    double c;
    if (div_rcp_acc(x, d, y, &c)) {
        ++guarded;
        if (memcmp(&c, &b, 8) != 0) {
            if (bad < 10) fprintf(stderr, "unguarded mismatch x=%a d=%a: %a vs %a\n", x, d, c, b);
            ++bad;
        }
    }
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    s_state = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
    /* divisors of the model's uniform grids (cell areas, spacings, doubled
     * spacings, depths, the friction factor) and random ones in range */
    const double fixed[] = {1.0e6, 1.0e3, 2.0e3, 100.0, 100.0 + 0.0, 1.0 + 0.0015 * 20.0,
                            1.0 + 0.00015 * 20.0, 4.0e6, 2.5e5, 1.0 + 0.0015 * 12.0, 50.0, 7.0,
                            150.0, 75.0, 3.0};
    const int nf = (int)(sizeof fixed / sizeof fixed[0]);
    for (long i = 0; i < n; ++i) {
        double d;
        if (i % 4 != 3) d = fixed[rnd() % nf];
        else d = mk((int)(rnd() % 121) - 60, rnd(), 0);
        /* random x over the full finite range */
        check(mk((int)(rnd() % 2046) - 1022, rnd(), (int)(rnd() & 1)), d);
        /* x inside the fast range */
        check(mk((int)(rnd() % 1920) - 960, rnd(), (int)(rnd() & 1)), d);
        /* quotient with significand just below 2 (where RN(x*y) is farthest
         * from x / d in ulps) */
        {
            const double q = mk((int)(rnd() % 1800) - 900, 0xfffffffffffffull - (rnd() & 0xfffffu),
                                (int)(rnd() & 1));
            const double x = (double)((long double)q * (long double)d);
            check(x, d);
            check(nextafter(x, -INFINITY), d);
        }
        /* near-midpoint quotient: x = RN(d * (q + ulp(q)/2)) in long double */
        {
            const double q = mk((int)(rnd() % 1800) - 900, rnd(), (int)(rnd() & 1));
            const long double mid = (long double)q + (long double)(nextafter(q, INFINITY) - q) / 2;
            const double x = (double)(mid * (long double)d);
            check(x, d);
            check(nextafter(x, INFINITY), d);
            check(nextafter(x, -INFINITY), d);
        }
    }
    /* specials and the edges of the fast range */
    const double sp[] = {0.0, -0.0, 0x1p-1074, -0x1p-1074, 0x1p-1022, 0x1p-960, -0x1p-960,
                         0x1.fffffffffffffp-961, 0x1p960, 0x1.fffffffffffffp959, -0x1p960,
                         INFINITY, -INFINITY, NAN, 1.0, -1.0, 3.0, 0x1.fffffffffffffp1023};
    for (int k = 0; k < (int)(sizeof sp / sizeof sp[0]); ++k)
        for (int j = 0; j < nf; ++j) check(sp[k], fixed[j]);
    printf("%ld %ld %ld %ld %ld\n", tried, bad, guarded, tried1, fails1_unqualified);
    return 0;
}
