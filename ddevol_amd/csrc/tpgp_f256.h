/* tpgp_f256 -- 256-bit binary floating point with directed rounding, for host and device.
 *
 * The residual bound (engine/src/residual_rev.c:8-9, 55, 96-108, 112-114) works on intervals
 * whose ends are MPFR numbers of 256-bit precision, combined by products with exact scalars and
 * by sums, each rounded toward -inf (lower ends) or +inf (upper ends).  MPFR rounds every
 * operation correctly: the result is the exact value rounded once in the stated direction.  The
 * functions here compute the same correctly rounded results, so a sequence of them gives the
 * same bits as the same sequence of mpfr_mul / mpfr_add calls at 256 bits.
 *
 * Representation: value = (-1)^neg * (m / 2^256) * 2^e with m a 256-bit integer whose top bit
 * is set (m[3] most significant), i.e. MPFR's own exponent convention (mantissa in [1/2, 1)).
 * Zero is m = 0 with a sign; e is then 0.  Exponents stay far inside int32: every operand comes
 * from binary64 values and the products and sums of a few of them.
 *
 * Sums are exact on a 384-bit window: the operand of larger exponent fills the top 256 bits, the
 * other is shifted in below it, and bits shifted past the window's end are kept as a sticky bit
 * in its lowest place.  The rounding point lies 128 bits above that place, and a difference
 * needs a normalising shift of more than one bit only when the shift was at most one bit (then
 * nothing was shifted out), so the rounded result is the exact sum's.  Products with a scalar of
 * at most 128 significant bits are exact on the 384-bit window.
 *
 * Plain C (C17) and HIP C++: the same source serves the device kernels and host-side checks.
 */
#ifndef TPGP_F256_H
#define TPGP_F256_H
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define F256_FN __host__ __device__ inline
#else
#define F256_FN static inline
#endif

typedef struct {
    uint64_t m[4]; /* significand, m[3] most significant; top bit set unless zero */
    int32_t e;     /* exponent: value = (m / 2^256) * 2^e */
    uint32_t neg;  /* sign, kept for zeros too */
} f256;

enum { F256_DOWN = 0, F256_UP = 1 }; /* toward -inf, toward +inf */

F256_FN uint64_t f256_mulhi(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

F256_FN int f256_is_zero(const f256 *x) { return x->m[3] == 0; }

F256_FN f256 f256_zero(uint32_t neg) {
    f256 r;
    r.m[0] = r.m[1] = r.m[2] = r.m[3] = 0;
    r.e = 0;
    r.neg = neg;
    return r;
}

/* The binary64 value d exactly (finite d). */
F256_FN f256 f256_from_double(double d) {
    uint64_t b;
    memcpy(&b, &d, sizeof(b));
    f256 r = f256_zero((uint32_t)(b >> 63));
    const uint64_t ex = (b >> 52) & 0x7ffu, fr = b & ((1ull << 52) - 1);
    if (ex == 0) {
        if (!fr) return r;
        const int lz = __builtin_clzll(fr);
        r.m[3] = fr << lz; /* fr * 2^-1074 = (m[3] / 2^64) * 2^(64 - lz - 1074) */
        r.e = -1010 - lz;
    } else {
        r.m[3] = ((1ull << 52) | fr) << 11;
        r.e = (int32_t)ex - 1022;
    }
    return r;
}

/* w[0..5] (w[5] most significant, top bit set) rounded to 256 bits toward dir, sign neg. */
F256_FN f256 f256_round(const uint64_t *w, int32_t e, uint32_t neg, int dir) {
    f256 r;
    r.m[0] = w[2];
    r.m[1] = w[3];
    r.m[2] = w[4];
    r.m[3] = w[5];
    r.e = e;
    r.neg = neg;
    const int inexact = (w[0] | w[1]) != 0;
    /* toward -inf moves a negative value away from zero, toward +inf a positive one */
    if (inexact && (dir == F256_UP ? !neg : neg)) {
        uint64_t carry = 1;
        for (int i = 0; i < 4; ++i) {
            r.m[i] += carry;
            carry = carry && r.m[i] == 0;
        }
        if (carry) { /* 2^256 - 1 + 1: the next binade */
            r.m[3] = 1ull << 63;
            r.e += 1;
        }
    }
    return r;
}

/* w >>= d on 384 bits; the bits shifted out are or-ed into bit 0 (sticky). */
F256_FN void f256_shr_sticky(uint64_t *w, uint64_t d) {
    if (!d) return;
    if (d >= 384) {
        const uint64_t any = w[0] | w[1] | w[2] | w[3] | w[4] | w[5];
        for (int i = 0; i < 6; ++i) w[i] = 0;
        w[0] = any != 0;
        return;
    }
    const unsigned q = (unsigned)(d >> 6), s = (unsigned)(d & 63);
    uint64_t lost = 0;
    for (unsigned i = 0; i < q; ++i) lost |= w[i];
    if (s) lost |= w[q] << (64 - s);
    uint64_t out[6];
    for (unsigned i = 0; i < 6; ++i) {
        const uint64_t lo = i + q < 6 ? w[i + q] : 0, hi = i + q + 1 < 6 ? w[i + q + 1] : 0;
        out[i] = s ? (lo >> s) | (hi << (64 - s)) : lo;
    }
    for (int i = 0; i < 6; ++i) w[i] = out[i];
    w[0] |= lost != 0;
}

/* w <<= d on 384 bits (d < 384; the bits shifted out are zero by construction). */
F256_FN void f256_shl(uint64_t *w, unsigned d) {
    if (!d) return;
    const unsigned q = d >> 6, s = d & 63;
    uint64_t out[6];
    for (int i = 5; i >= 0; --i) {
        const int a = i - (int)q, b = i - (int)q - 1;
        const uint64_t hi = a >= 0 ? w[a] : 0, lo = b >= 0 ? w[b] : 0;
        out[i] = s ? (hi << s) | (lo >> (64 - s)) : hi;
    }
    for (int i = 0; i < 6; ++i) w[i] = out[i];
}

/* Leading zeros of a nonzero 384-bit w. */
F256_FN unsigned f256_clz384(const uint64_t *w) {
    unsigned n = 0;
    for (int i = 5; i >= 0; --i) {
        if (w[i]) return n + (unsigned)__builtin_clzll(w[i]);
        n += 64;
    }
    return n;
}

/* x + y rounded toward dir (mpfr_add at 256 bits). */
F256_FN f256 f256_add(f256 x, f256 y, int dir) {
    const int xz = f256_is_zero(&x), yz = f256_is_zero(&y);
    if (xz && yz) /* (+0) + (-0) is +0, or -0 toward -inf */
        return f256_zero(x.neg == y.neg ? x.neg : (uint32_t)(dir == F256_DOWN));
    if (xz) return y;
    if (yz) return x;
    if (y.e > x.e) {
        const f256 t = x;
        x = y;
        y = t;
    }
    uint64_t a[6] = {0, 0, x.m[0], x.m[1], x.m[2], x.m[3]};
    uint64_t b[6] = {0, 0, y.m[0], y.m[1], y.m[2], y.m[3]};
    const uint64_t d = (uint64_t)((int64_t)x.e - (int64_t)y.e);
    f256_shr_sticky(b, d);
    int32_t e = x.e;
    uint32_t neg = x.neg;
    if (x.neg == y.neg) {
        uint64_t carry = 0;
        for (int i = 0; i < 6; ++i) {
            const uint64_t s = a[i] + b[i], c1 = s < a[i];
            a[i] = s + carry;
            carry = c1 | (a[i] < s);
        }
        if (carry) { /* one bit past the window: shift it back in, keep the lost bit sticky */
            const uint64_t lost = a[0] & 1;
            f256_shr_sticky(a, 1);
            a[5] |= 1ull << 63;
            a[0] |= lost;
            e += 1;
        }
    } else {
        int cmp = 1; /* |x| > |y| whenever the exponents differ */
        if (d == 0)
            for (int i = 5; i >= 0 && cmp == 1; --i)
                cmp = a[i] > b[i] ? 2 : a[i] < b[i] ? -1 : (i == 0 ? 0 : 1);
        if (cmp == 0) return f256_zero((uint32_t)(dir == F256_DOWN)); /* exact cancellation */
        if (cmp < 0) {
            for (int i = 0; i < 6; ++i) {
                const uint64_t t = a[i];
                a[i] = b[i];
                b[i] = t;
            }
            neg = y.neg;
        }
        uint64_t borrow = 0;
        for (int i = 0; i < 6; ++i) {
            const uint64_t s = a[i] - b[i], c1 = a[i] < b[i];
            a[i] = s - borrow;
            borrow = c1 | (s < borrow);
        }
        const unsigned lz = f256_clz384(a);
        f256_shl(a, lz);
        e -= (int32_t)lz;
    }
    return f256_round(a, e, neg, dir);
}

/* x * s rounded toward dir (mpfr_mul at 256 bits), for a scalar s of at most 128 significant
 * bits (s.m[0] = s.m[1] = 0). */
F256_FN f256 f256_mul(f256 x, f256 s, int dir) {
    const uint32_t neg = x.neg ^ s.neg;
    if (f256_is_zero(&x) || f256_is_zero(&s)) return f256_zero(neg);
    uint64_t p[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 2; ++j) {
        const uint64_t sj = s.m[2 + j];
        uint64_t carry = 0;
        for (int i = 0; i < 4; ++i) {
            const uint64_t lo = x.m[i] * sj, hi = f256_mulhi(x.m[i], sj);
            const uint64_t t = p[i + j] + lo, c1 = t < lo;
            const uint64_t u = t + carry, c2 = u < carry;
            p[i + j] = u;
            carry = hi + c1 + c2;
        }
        p[4 + j] += carry;
    }
    int32_t e = x.e + s.e;
    if (!(p[5] >> 63)) { /* both significands are in [1/2, 1): the product is in [1/4, 1) */
        f256_shl(p, 1);
        e -= 1;
    }
    return f256_round(p, e, neg, dir);
}

/* -x (exact). */
F256_FN f256 f256_neg(f256 x) {
    x.neg ^= 1u;
    return x;
}


/* x * y rounded toward dir (mpfr_mul or mpfr_sqr at 256 bits), both operands of full 256-bit
 * precision: the exact 512-bit product, normalised, then rounded with the low 256 bits as the
 * remainder. */
F256_FN f256 f256_mul_full(f256 x, f256 y, int dir) {
    const uint32_t neg = x.neg ^ y.neg;
    if (f256_is_zero(&x) || f256_is_zero(&y)) return f256_zero(neg);
    uint64_t p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int j = 0; j < 4; ++j) {
        const uint64_t yj = y.m[j];
        uint64_t carry = 0;
#if defined(__HIPCC__)
#pragma unroll
#endif
        for (int i = 0; i < 4; ++i) {
            const uint64_t lo = x.m[i] * yj, hi = f256_mulhi(x.m[i], yj);
            const uint64_t t = p[i + j] + lo, c1 = t < lo;
            const uint64_t u = t + carry, c2 = u < carry;
            p[i + j] = u;
            carry = hi + c1 + c2;
        }
        p[4 + j] = carry;
    }
    int32_t e = x.e + y.e;
    if (!(p[7] >> 63)) { /* the product of two significands in [1/2, 1) is in [1/4, 1) */
        for (int i = 7; i > 0; --i) p[i] = (p[i] << 1) | (p[i - 1] >> 63);
        p[0] <<= 1;
        e -= 1;
    }
    /* round on a 384-bit window: the top 256 bits, and any nonzero bit of the low 256 folded
     * into the window's lowest word (the remainder only has to be known to be zero or not) */
    const uint64_t w[6] = {(p[0] | p[1] | p[2] | p[3]) != 0, 0, p[4], p[5], p[6], p[7]};
    return f256_round(w, e, neg, dir);
}

/* The sign of x - y (-1, 0, 1) as mpfr_cmp gives it: zeros compare equal whatever their sign. */
F256_FN int f256_cmp(const f256 *x, const f256 *y) {
    const int xz = f256_is_zero(x), yz = f256_is_zero(y);
    if (xz && yz) return 0;
    if (xz) return y->neg ? 1 : -1;
    if (yz) return x->neg ? -1 : 1;
    if (x->neg != y->neg) return x->neg ? -1 : 1;
    int mag = 0; /* |x| against |y| */
    if (x->e != y->e)
        mag = x->e > y->e ? 1 : -1;
    else
        for (int i = 3; i >= 0 && !mag; --i)
            mag = x->m[i] > y->m[i] ? 1 : x->m[i] < y->m[i] ? -1 : 0;
    return x->neg ? -mag : mag;
}

/* |x| (exact). */
F256_FN f256 f256_abs(f256 x) {
    x.neg = 0;
    return x;
}

/* The sign as mpfr_sgn gives it: 0 for either zero. */
F256_FN int f256_sgn(const f256 *x) { return f256_is_zero(x) ? 0 : x->neg ? -1 : 1; }

#endif
