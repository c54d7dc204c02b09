/* CPU check of the quotient sequence the point-cost kernels use in place of
 * an IEEE division by the cost scale s (one multiply by y = RN(1/s) and two
 * Markstein corrections).  Built and run by tests/test_quotient.py; the
 * device code in pdot_common.h (exact_quotient) performs the same operations.
 * Returns the number of a in the sample whose quotient differs from a / s in
 * any bit. */
#include <math.h>
#include <stdint.h>

static double quotient(double a, double s, double y) {
    const int s_ok = s >= 0x1p-1000 && s <= 0x1p+1000;
    const double lo = fmax(0x1p-960, 0x1p-900 * s), hi = 0x1p+900 * s;
    if (!(s_ok && a >= lo && a <= hi)) return a / s;
    const double q0 = a * y;
    const double r0 = fma(-q0, s, a);
    const double q1 = fma(r0, y, q0);
    const double r1 = fma(-q1, s, a);
    return fma(r1, y, q1);
}

static int differs(double x, double z) {
    union { double d; uint64_t u; } p = {x}, q = {z};
    return p.u != q.u;
}

long quotient_mismatches(const double* a, long n, double s) {
    const double y = 1.0 / s;
    long bad = 0;
    for (long k = 0; k < n; ++k) bad += differs(quotient(a[k], s, y), a[k] / s);
    return bad;
}

/* xorshift sample: count values a = s * u with u uniform in [0, 1) (53 random
 * bits), plus each of their neighbours one ulp away. */
long quotient_mismatches_random(double s, long n, uint64_t seed) {
    const double y = 1.0 / s;
    uint64_t x = seed ? seed : 88172645463325252ULL;
    long bad = 0;
    for (long k = 0; k < n; ++k) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        const double u = (double)(x >> 11) * 0x1p-53;
        const double a = s * u;
        const double cand[3] = {a, nextafter(a, 0.0), nextafter(a, INFINITY)};
        for (int c = 0; c < 3; ++c) bad += differs(quotient(cand[c], s, y), cand[c] / s);
    }
    return bad;
}

/* Tiny numerators: a = m * 2^e with m uniform in [1, 2) (53 random bits) and e
 * uniform in [emin, emax], including subnormal a (e < -1022 rounds m * 2^e to
 * the subnormal grid), plus each a's neighbours one ulp away. */
long quotient_mismatches_tiny(double s, long n, int emin, int emax, uint64_t seed) {
    const double y = 1.0 / s;
    uint64_t x = seed ? seed : 88172645463325252ULL;
    long bad = 0;
    for (long k = 0; k < n; ++k) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        const double mant = 1.0 + (double)(x >> 11) * 0x1p-53;
        const int e = emin + (int)((x & 0x7ff) % (uint64_t)(emax - emin + 1));
        const double a = ldexp(mant, e);
        const double cand[3] = {a, nextafter(a, 0.0), nextafter(a, INFINITY)};
        for (int c = 0; c < 3; ++c) bad += differs(quotient(cand[c], s, y), cand[c] / s);
    }
    return bad;
}
