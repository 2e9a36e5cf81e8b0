/*
 * Blackframe XPU/HIP — deterministic scalar math shared by host and device code.
 *
 * Transport decisions (a sampled direction that grazes an edge, a roulette threshold) fork a
 * whole path when one ulp differs, so every transcendental the wavefront kernels evaluate comes
 * from this header rather than from a platform libm: the same C source compiled by gcc for a CPU
 * and by hipcc for gfx950 performs the same IEEE binary64 operations (+, *, fma, sqrt, division
 * are correctly rounded on both) and therefore returns the same float.
 *
 * Methods (all public textbook methods, constants derived here):
 *   sin/cos : Cody–Waite reduction by pi/2 with a two-term split of pi/2 (the high term keeps
 *             33 significant bits so k*hi is exact for |k| < 2^20), then Taylor series of degree
 *             19 (sin) / 18 (cos) on |r| <= pi/4 evaluated by Horner's rule with fused
 *             multiply-add in binary64; truncation error < 1e-19, far below binary32 spacing.
 *   hypot   : sqrt(x*x + y*y) in binary64; the products are exact for binary32 inputs.
 *   atan2   : reduction to a ratio t in [0, 1] by octant, the half-angle identity
 *             atan(t) = 2 atan(t / (1 + sqrt(1 + t^2))) applied twice (|t| <= tan(pi/16)), then
 *             the Taylor series of atan through t^31 (truncation error < 1e-24).
 *   exp/expm1: Cody–Waite reduction by ln 2 (two-term split, the high term exact in k*hi), then
 *             the Taylor series of exp(r) - 1 through r^17 on |r| <= ln2/2, scaled by 2^k.
 *   log     : m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m - 1) / (m + 1)) by its
 *             series through s^25, plus e ln 2 with the same split constant.
 * Exactly rounded primitives (nextafter, frexp, ldexp) are implemented on the bit patterns.
 */
#ifndef BLACKFRAME_MATH_H
#define BLACKFRAME_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define BF_HD __host__ __device__ static inline
#else
#define BF_HD static inline
#endif

BF_HD uint32_t bf_float_bits(float value) {
    uint32_t bits;
    memcpy(&bits, &value, sizeof bits);
    return bits;
}
BF_HD float bf_bits_float(uint32_t bits) {
    float value;
    memcpy(&value, &bits, sizeof value);
    return value;
}
BF_HD uint64_t bf_double_bits(double value) {
    uint64_t bits;
    memcpy(&bits, &value, sizeof bits);
    return bits;
}
BF_HD double bf_bits_double(uint64_t bits) {
    double value;
    memcpy(&value, &bits, sizeof value);
    return value;
}

BF_HD int bf_isnan(float v) { return (bf_float_bits(v) & 0x7FFFFFFFu) > 0x7F800000u; }
BF_HD int bf_isfinite(float v) { return (bf_float_bits(v) & 0x7F800000u) != 0x7F800000u; }
BF_HD int bf_isnand(double v) { return (bf_double_bits(v) & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull; }
BF_HD int bf_isfinited(double v) { return (bf_double_bits(v) & 0x7FF0000000000000ull) != 0x7FF0000000000000ull; }
BF_HD int bf_signbit(float v) { return (bf_float_bits(v) >> 31) != 0u; }

#define BF_INFINITY_F (bf_bits_float(0x7F800000u))
#define BF_NEG_INFINITY_F (bf_bits_float(0xFF800000u))

/* nextafter toward +inf / -inf (C99 nextafter semantics for those directions). */
BF_HD float bf_next_up_f(float x) {
    if (bf_isnan(x) || bf_float_bits(x) == 0x7F800000u) return x;
    if (x == 0.0f) return bf_bits_float(1u);
    uint32_t b = bf_float_bits(x);
    return bf_bits_float(x > 0.0f ? b + 1u : b - 1u);
}
BF_HD float bf_next_down_f(float x) {
    if (bf_isnan(x) || bf_float_bits(x) == 0xFF800000u) return x;
    if (x == 0.0f) return bf_bits_float(0x80000001u);
    uint32_t b = bf_float_bits(x);
    return bf_bits_float(x > 0.0f ? b - 1u : b + 1u);
}
/* nextafterf(x, 0): one step toward zero (x finite, non-zero) */
BF_HD float bf_next_toward_zero_f(float x) {
    if (bf_isnan(x) || x == 0.0f) return x;
    return bf_bits_float(bf_float_bits(x) - 1u);
}
BF_HD double bf_next_up(double x) {
    if (bf_isnand(x) || bf_double_bits(x) == 0x7FF0000000000000ull) return x;
    if (x == 0.0) return bf_bits_double(1ull);
    uint64_t b = bf_double_bits(x);
    return bf_bits_double(x > 0.0 ? b + 1ull : b - 1ull);
}
BF_HD double bf_next_down(double x) {
    if (bf_isnand(x) || bf_double_bits(x) == 0xFFF0000000000000ull) return x;
    if (x == 0.0) return bf_bits_double(0x8000000000000001ull);
    uint64_t b = bf_double_bits(x);
    return bf_bits_double(x > 0.0 ? b - 1ull : b + 1ull);
}

/* frexpf for finite non-zero x: x = m * 2^e with 0.5 <= |m| < 1 (subnormals normalised). */
BF_HD float bf_frexpf(float x, int* exponent) {
    uint32_t b = bf_float_bits(x);
    uint32_t mag = b & 0x7FFFFFFFu;
    if (mag == 0u || mag >= 0x7F800000u) {
        *exponent = 0;
        return x;
    }
    int bias = 0;
    if (mag < 0x00800000u) { /* subnormal: scale by 2^25 exactly */
        x = x * 33554432.0f;
        b = bf_float_bits(x);
        mag = b & 0x7FFFFFFFu;
        bias = -25;
    }
    int e = (int)(mag >> 23) - 126;
    *exponent = e + bias;
    return bf_bits_float((b & 0x807FFFFFu) | 0x3F000000u);
}

/* ldexpf with a single correct rounding (round-to-nearest-even) for subnormal results. */
BF_HD float bf_ldexpf(float x, int n) {
    if (x == 0.0f || !bf_isfinite(x)) return x;
    /* exact in binary64 for every float and every |n| the callers use; one rounding to float */
    if (n > 300) n = 300;
    if (n < -300) n = -300;
    double scaled = (double)x;
    while (n > 0) {
        int step = n > 1000 ? 1000 : n;
        scaled *= bf_bits_double((uint64_t)(1023 + step) << 52);
        n -= step;
    }
    while (n < 0) {
        int step = n < -1000 ? -1000 : n;
        scaled *= bf_bits_double((uint64_t)(1023 + step) << 52);
        n -= step;
    }
    return (float)scaled;
}

BF_HD float bf_hypotf(float x, float y) {
    double dx = (double)x;
    double dy = (double)y;
    return (float)sqrt(fma(dx, dx, dy * dy));
}

/* sin and cos of a binary32 argument, rounded once from binary64. */
BF_HD void bf_sincosf(float angle, float* sine, float* cosine) {
    const double two_over_pi = 0.63661977236758134308;
    const double pio2_hi = 1.57079632673412561417e+00; /* 33 significant bits */
    const double pio2_lo = 6.07710050650619224932e-11;
    double x = (double)angle;
    double k = rint(x * two_over_pi);
    double r = fma(-k, pio2_hi, x);
    r = fma(-k, pio2_lo, r);
    double r2 = r * r;
    /* sin(r) = r * (1 - r^2/3! + r^4/5! - ...) */
    double s = -1.0 / 121645100408832000.0; /* -1/19! */
    s = fma(s, r2, 1.0 / 355687428096000.0);   /* 1/17! */
    s = fma(s, r2, -1.0 / 1307674368000.0);    /* -1/15! */
    s = fma(s, r2, 1.0 / 6227020800.0);        /* 1/13! */
    s = fma(s, r2, -1.0 / 39916800.0);         /* -1/11! */
    s = fma(s, r2, 1.0 / 362880.0);            /* 1/9! */
    s = fma(s, r2, -1.0 / 5040.0);             /* -1/7! */
    s = fma(s, r2, 1.0 / 120.0);               /* 1/5! */
    s = fma(s, r2, -1.0 / 6.0);                /* -1/3! */
    double sin_r = fma(s * r2, r, r);
    double c = 1.0 / 6402373705728000.0;       /* 1/18! */
    c = fma(c, r2, -1.0 / 20922789888000.0);   /* -1/16! */
    c = fma(c, r2, 1.0 / 87178291200.0);       /* 1/14! */
    c = fma(c, r2, -1.0 / 479001600.0);        /* -1/12! */
    c = fma(c, r2, 1.0 / 3628800.0);           /* 1/10! */
    c = fma(c, r2, -1.0 / 40320.0);            /* -1/8! */
    c = fma(c, r2, 1.0 / 720.0);               /* 1/6! */
    c = fma(c, r2, -1.0 / 24.0);               /* -1/4! */
    c = fma(c, r2, 0.5);                       /* +1/2 (subtracted below) */
    double cos_r = fma(-c, r2, 1.0);
    long long quadrant = ((long long)k) & 3;
    double sv, cv;
    if (quadrant == 0) {
        sv = sin_r;
        cv = cos_r;
    } else if (quadrant == 1) {
        sv = cos_r;
        cv = -sin_r;
    } else if (quadrant == 2) {
        sv = -sin_r;
        cv = -cos_r;
    } else {
        sv = -cos_r;
        cv = sin_r;
    }
    *sine = (float)sv;
    *cosine = (float)cv;
}

BF_HD float bf_sinf(float a) {
    float s, c;
    bf_sincosf(a, &s, &c);
    return s;
}
BF_HD float bf_cosf(float a) {
    float s, c;
    bf_sincosf(a, &s, &c);
    return c;
}

/* atan(t) for 0 <= t <= 1 in binary64 via two half-angle reductions and a Taylor series. */
BF_HD double bf_atan_unit(double t) {
    /* atan(t) = 2 atan(t / (1 + sqrt(1 + t^2))) */
    double t1 = t / (1.0 + sqrt(fma(t, t, 1.0)));
    double t2 = t1 / (1.0 + sqrt(fma(t1, t1, 1.0)));
    double u = t2 * t2;
    /* |t2| <= tan(pi/16) ~ 0.199: terms through t^31 give error < 1e-23 */
    double p = -1.0 / 31.0;
    p = fma(p, u, 1.0 / 29.0);
    p = fma(p, u, -1.0 / 27.0);
    p = fma(p, u, 1.0 / 25.0);
    p = fma(p, u, -1.0 / 23.0);
    p = fma(p, u, 1.0 / 21.0);
    p = fma(p, u, -1.0 / 19.0);
    p = fma(p, u, 1.0 / 17.0);
    p = fma(p, u, -1.0 / 15.0);
    p = fma(p, u, 1.0 / 13.0);
    p = fma(p, u, -1.0 / 11.0);
    p = fma(p, u, 1.0 / 9.0);
    p = fma(p, u, -1.0 / 7.0);
    p = fma(p, u, 1.0 / 5.0);
    p = fma(p, u, -1.0 / 3.0);
    double a = fma(p * u, t2, t2);
    return 4.0 * a;
}

BF_HD float bf_atan2f(float y, float x) {
    const double pi = 3.14159265358979323846;
    double ay = fabs((double)y);
    double ax = fabs((double)x);
    if (bf_isnan(x) || bf_isnan(y)) return x + y;
    if (ay == 0.0 && ax == 0.0) {
        double base = bf_signbit(x) ? pi : 0.0;
        return (float)(bf_signbit(y) ? -base : base);
    }
    double angle;
    if (ay <= ax) {
        angle = bf_atan_unit(ay / ax);
    } else {
        angle = 0.5 * pi - bf_atan_unit(ax / ay);
    }
    if (bf_signbit(x)) angle = pi - angle;
    if (bf_signbit(y)) angle = -angle;
    return (float)angle;
}

/* exp(x) - 1 in binary64 parts: x = k ln2 + r with |r| <= ln2 / 2, *em1 = exp(r) - 1 by its
 * Taylor series through r^17 / 17! (truncation below 2^-70 relative), *two_k = 2^k. */
BF_HD void bf_exp_parts(double x, double* em1, double* two_k) {
    const double ln2_hi = 6.93147180369123816490e-01; /* 32 significant bits: k * ln2_hi is exact */
    const double ln2_lo = 1.90821492927058770002e-10;
    const double k = rint(x * 1.44269504088896338700);
    const double r = fma(-k, ln2_lo, fma(-k, ln2_hi, x));
    double p = 1.0 / 355687428096000.0;     /* 1/17! */
    p = fma(p, r, 1.0 / 20922789888000.0);  /* 1/16! */
    p = fma(p, r, 1.0 / 1307674368000.0);   /* 1/15! */
    p = fma(p, r, 1.0 / 87178291200.0);     /* 1/14! */
    p = fma(p, r, 1.0 / 6227020800.0);      /* 1/13! */
    p = fma(p, r, 1.0 / 479001600.0);       /* 1/12! */
    p = fma(p, r, 1.0 / 39916800.0);        /* 1/11! */
    p = fma(p, r, 1.0 / 3628800.0);         /* 1/10! */
    p = fma(p, r, 1.0 / 362880.0);          /* 1/9! */
    p = fma(p, r, 1.0 / 40320.0);           /* 1/8! */
    p = fma(p, r, 1.0 / 5040.0);            /* 1/7! */
    p = fma(p, r, 1.0 / 720.0);             /* 1/6! */
    p = fma(p, r, 1.0 / 120.0);             /* 1/5! */
    p = fma(p, r, 1.0 / 24.0);              /* 1/4! */
    p = fma(p, r, 1.0 / 6.0);               /* 1/3! */
    p = fma(p, r, 0.5);                     /* 1/2! */
    *em1 = fma(p * r, r, r);
    *two_k = bf_bits_double((uint64_t)(1023 + (long long)k) << 52);
}

/* exp(x) - 1 of a binary32 argument, rounded once from binary64. */
BF_HD float bf_expm1f(float x) {
    if (bf_isnan(x)) return x;
    if (x > 89.0f) return bf_bits_float(0x7F800000u);
    if (x < -20.0f) return -1.0f;
    double em1, two_k;
    bf_exp_parts((double)x, &em1, &two_k);
    return (float)fma(two_k, em1, two_k - 1.0);
}

/* exp(x) of a binary32 argument, rounded once from binary64. */
BF_HD float bf_expf(float x) {
    if (bf_isnan(x)) return x;
    if (x > 89.0f) return bf_bits_float(0x7F800000u);
    if (x < -104.0f) return 0.0f;
    double em1, two_k;
    bf_exp_parts((double)x, &em1, &two_k);
    return (float)fma(two_k, em1, two_k);
}

/* log(x) of a binary32 argument, rounded once from binary64: x = m 2^e with m in
 * [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with s = (m - 1) / (m + 1), |s| <= 0.1716, summed
 * through s^25 / 25 (truncation below 2^-64 relative). */
BF_HD float bf_logf(float x) {
    if (bf_isnan(x) || x < 0.0f) return bf_bits_float(0x7FC00000u);
    if (x == 0.0f) return bf_bits_float(0xFF800000u);
    if (!bf_isfinite(x)) return x;
    const uint64_t bits = bf_double_bits((double)x);
    int e = (int)((bits >> 52) & 0x7FFu) - 1023;
    double m = bf_bits_double((bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.41421356237309504880) {
        m *= 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double u = s * s;
    double p = 1.0 / 25.0;
    p = fma(p, u, 1.0 / 23.0);
    p = fma(p, u, 1.0 / 21.0);
    p = fma(p, u, 1.0 / 19.0);
    p = fma(p, u, 1.0 / 17.0);
    p = fma(p, u, 1.0 / 15.0);
    p = fma(p, u, 1.0 / 13.0);
    p = fma(p, u, 1.0 / 11.0);
    p = fma(p, u, 1.0 / 9.0);
    p = fma(p, u, 1.0 / 7.0);
    p = fma(p, u, 1.0 / 5.0);
    p = fma(p, u, 1.0 / 3.0);
    const double log_m = 2.0 * fma(p * u, s, s);
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return (float)fma((double)e, ln2_hi, fma((double)e, ln2_lo, log_m));
}

#endif /* BLACKFRAME_MATH_H */
