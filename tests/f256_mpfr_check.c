/* Checks ddevol_amd/csrc/tpgp_f256.h against MPFR at 256 bits, bit for bit.
 *
 * Build: gcc -std=c17 -O2 -I<mpfr>/include tests/f256_mpfr_check.c -lmpfr -lgmp
 * Run:   ./a.out CASES SEED   -> prints "ok <checked>" or the first mismatch, exit 1
 *
 * Every operation is compared with the MPFR call the residual bound makes
 * (engine/src/residual_rev.c:55, 105-106): mpfr_add and mpfr_mul at precision 256, toward -inf
 * and toward +inf, including the sign of zero results.  Operands are drawn to reach every path:
 * exponent gaps around the limb and window boundaries (0..3, 63..65, 127..130, 255..257,
 * 383..385, far), near-cancellations, carries through all-ones significands, sparse
 * significands, signed zeros, scalars of 1..128 significant bits, and binary64 conversions
 * (normal, subnormal, signed zero).
 */
#include <gmp.h>
#include <math.h>
#include <mpfr.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../ddevol_amd/csrc/tpgp_f256.h"

static uint64_t state;
static uint64_t next(void) { /* splitmix64 */
    uint64_t z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static void to_mpfr(mpfr_t out, const f256 *x) {
    if (f256_is_zero(x)) {
        mpfr_set_zero(out, x->neg ? -1 : 1);
        return;
    }
    mpz_t z;
    mpz_init(z);
    mpz_import(z, 4, -1, sizeof(uint64_t), 0, 0, x->m);
    mpfr_set_z_2exp(out, z, (mpfr_exp_t)x->e - 256, MPFR_RNDN); /* exact: 256 bits */
    if (x->neg) mpfr_neg(out, out, MPFR_RNDN);
    mpz_clear(z);
}

static f256 from_mpfr(const mpfr_t x) {
    f256 r = f256_zero(mpfr_signbit(x) ? 1u : 0u);
    if (mpfr_zero_p(x)) return r;
    mpz_t z;
    mpz_init(z);
    mpfr_exp_t e = mpfr_get_z_2exp(z, x);
    mpz_abs(z, z);
    const size_t bits = mpz_sizeinbase(z, 2);
    mpz_mul_2exp(z, z, 256 - bits);
    e -= (mpfr_exp_t)(256 - bits);
    size_t count = 0;
    mpz_export(r.m, &count, -1, sizeof(uint64_t), 0, 0, z);
    r.e = (int32_t)(e + 256);
    mpz_clear(z);
    return r;
}

static int same(const f256 *got, const mpfr_t want) {
    mpfr_t g;
    mpfr_init2(g, 256);
    to_mpfr(g, got);
    int ok = mpfr_equal_p(g, want) && (mpfr_signbit(g) != 0) == (mpfr_signbit(want) != 0);
    /* the representation itself must be canonical */
    if (!f256_is_zero(got) && !(got->m[3] >> 63)) ok = 0;
    if (f256_is_zero(got) && (got->m[0] | got->m[1] | got->m[2])) ok = 0;
    mpfr_clear(g);
    return ok;
}

static f256 random_value(int32_t e) {
    f256 r = f256_zero((uint32_t)(next() & 1));
    const unsigned kind = (unsigned)(next() % 6);
    for (int i = 0; i < 4; ++i) r.m[i] = next();
    if (kind == 1) /* all ones: carries through every limb */
        for (int i = 0; i < 4; ++i) r.m[i] = ~0ull;
    if (kind == 2) /* sparse: a few bits */
        for (int i = 0; i < 4; ++i) r.m[i] = (next() % 4 == 0) ? 1ull << (next() % 64) : 0;
    if (kind == 3) /* a binary64-sized significand */
        r.m[0] = r.m[1] = r.m[2] = 0, r.m[3] &= ~((1ull << 11) - 1);
    if (kind == 4) /* 128 bits */
        r.m[0] = r.m[1] = 0;
    if (kind == 5) /* trailing ones */
        r.m[0] = ~0ull, r.m[1] |= 1ull << 63;
    r.m[3] |= 1ull << 63;
    r.e = e;
    return r;
}

static const int gaps[] = {0,   1,   2,   3,   63,  64,  65,  127, 128, 129, 130,
                           191, 192, 255, 256, 257, 300, 383, 384, 385, 1000, 4000};

static int check_add(f256 x, f256 y, mpfr_t mx, mpfr_t my, mpfr_t want) {
    to_mpfr(mx, &x);
    to_mpfr(my, &y);
    for (int dir = 0; dir < 2; ++dir) {
        const f256 got = f256_add(x, y, dir);
        mpfr_add(want, mx, my, dir == F256_DOWN ? MPFR_RNDD : MPFR_RNDU);
        if (!same(&got, want)) {
            mpfr_printf("add mismatch dir=%d\n x=%.80Rb\n y=%.80Rb\n want=%.80Rb\n", dir, mx,
                        my, want);
            return 0;
        }
    }
    return 1;
}

static int check_mul(f256 x, f256 s, mpfr_t mx, mpfr_t ms, mpfr_t want) {
    to_mpfr(mx, &x);
    to_mpfr(ms, &s);
    for (int dir = 0; dir < 2; ++dir) {
        const f256 got = f256_mul(x, s, dir);
        mpfr_mul(want, mx, ms, dir == F256_DOWN ? MPFR_RNDD : MPFR_RNDU);
        if (!same(&got, want)) {
            mpfr_printf("mul mismatch dir=%d\n x=%.80Rb\n s=%.80Rb\n want=%.80Rb\n", dir, mx,
                        ms, want);
            return 0;
        }
    }
    return 1;
}

int main(int argc, char **argv) {
    const long cases = argc > 1 ? atol(argv[1]) : 100000;
    state = argc > 2 ? strtoull(argv[2], NULL, 10) : 1;
    mpfr_t mx, my, want;
    mpfr_inits2(256, mx, my, want, (mpfr_ptr)0);
    long checked = 0;
    /* binary64 conversion: normals, subnormals, signed zeros */
    for (long k = 0; k < cases / 10 + 8; ++k) {
        uint64_t b = next();
        if (k % 4 == 1) b &= 0x800fffffffffffffull;               /* subnormal */
        if (k % 4 == 2) b = (b & 0x8000000000000000ull) | (next() % 3 ? 1 : 0); /* tiny / 0 */
        double d;
        memcpy(&d, &b, sizeof(d));
        if (!isfinite(d)) continue;
        const f256 got = f256_from_double(d);
        mpfr_set_d(want, d, MPFR_RNDN);
        if (!same(&got, want)) {
            printf("from_double mismatch %a\n", d);
            return 1;
        }
        ++checked;
    }
    for (long k = 0; k < cases; ++k) {
        const int32_t e = (int32_t)(next() % 2001) - 1000;
        f256 x = random_value(e), y;
        const unsigned kind = (unsigned)(next() % 5);
        if (kind == 0) { /* exponent gap near a boundary, either operand larger */
            int g = gaps[next() % (sizeof(gaps) / sizeof(gaps[0]))];
            y = random_value(next() & 1 ? e - g : e + g);
        } else if (kind == 1) { /* near-cancellation: y = -x perturbed in low bits */
            y = x;
            y.neg ^= 1u;
            const unsigned limb = (unsigned)(next() % 4);
            y.m[limb] ^= next() >> (next() % 64);
            y.m[3] |= 1ull << 63;
            if (next() % 8 == 0) y = f256_neg(x); /* exact cancellation */
        } else if (kind == 2) { /* signed zeros */
            y = f256_zero((uint32_t)(next() & 1));
            if (next() & 1) x = f256_zero((uint32_t)(next() & 1));
        } else if (kind == 3) { /* random gap */
            y = random_value(e - (int32_t)(next() % 600) + 300);
        } else { /* equal exponents, random signs */
            y = random_value(e);
        }
        if (!check_add(x, y, mx, my, want) || !check_add(y, x, mx, my, want)) return 1;
        checked += 2;
        /* products with scalars of 1..128 significant bits */
        f256 s = random_value((int32_t)(next() % 41) - 20);
        s.m[0] = s.m[1] = 0;
        const unsigned bits = 1 + (unsigned)(next() % 128);
        if (bits <= 64) {
            s.m[2] = 0;
            s.m[3] &= bits == 64 ? ~0ull : ~((1ull << (64 - bits)) - 1);
        } else if (bits < 128) {
            s.m[2] &= ~((1ull << (128 - bits)) - 1);
        }
        s.m[3] |= 1ull << 63;
        if (next() % 16 == 0) s = f256_zero((uint32_t)(next() & 1));
        if (next() % 16 == 0) x = f256_zero((uint32_t)(next() & 1));
        if (!check_mul(x, s, mx, my, want)) return 1;
        ++checked;
    }
    /* full products and squares (the Gram terms), and comparisons */
    for (long k = 0; k < cases / 2; ++k) {
        f256 x = random_value((int32_t)(next() % 2001) - 1000), y;
        y = next() % 4 == 0 ? x : random_value((int32_t)(next() % 2001) - 1000);
        if (next() % 32 == 0) y = f256_zero((uint32_t)(next() & 1));
        if (next() % 32 == 0) x = f256_zero((uint32_t)(next() & 1));
        to_mpfr(mx, &x);
        to_mpfr(my, &y);
        for (int dir = 0; dir < 2; ++dir) {
            const f256 got = f256_mul_full(x, y, dir);
            mpfr_mul(want, mx, my, dir == F256_DOWN ? MPFR_RNDD : MPFR_RNDU);
            if (!same(&got, want)) {
                mpfr_printf("mul_full mismatch dir=%d\n x=%.80Rb\n y=%.80Rb\n", dir, mx, my);
                return 1;
            }
            const f256 sq = f256_mul_full(x, x, dir);
            mpfr_sqr(want, mx, dir == F256_DOWN ? MPFR_RNDD : MPFR_RNDU);
            if (!same(&sq, want)) {
                mpfr_printf("square mismatch dir=%d\n x=%.80Rb\n", dir, mx);
                return 1;
            }
        }
        /* comparisons, including near-equal values and signed zeros */
        f256 z = y;
        if (next() % 3 == 0) {
            z = x;
            if (next() & 1) z.m[next() % 4] ^= 1ull << (next() % 64), z.m[3] |= 1ull << 63;
        }
        to_mpfr(my, &z);
        const int c = f256_cmp(&x, &z), w = mpfr_cmp(mx, my);
        if ((c > 0) != (w > 0) || (c < 0) != (w < 0) || f256_sgn(&x) != mpfr_sgn(mx)) {
            printf("cmp mismatch at case %ld\n", k);
            return 1;
        }
        checked += 5;
    }
    /* chains as the generator runs them: z = z + x*s, both directions, from +0 */
    for (long k = 0; k < cases / 10; ++k) {
        f256 lo = f256_zero(0), hi = f256_zero(0);
        mpfr_t zlo, zhi, t;
        mpfr_inits2(256, zlo, zhi, t, (mpfr_ptr)0);
        mpfr_set_zero(zlo, 1);
        mpfr_set_zero(zhi, 1);
        const int len = 1 + (int)(next() % 12);
        for (int i = 0; i < len; ++i) {
            const double h = ldexp((double)(next() >> 11), -53 - (int)(next() % 8));
            f256 s = f256_from_double(next() & 1 ? -2 * h : 2 * h), x = random_value(
                (int32_t)(next() % 61) - 30);
            lo = f256_add(lo, f256_mul(x, s, F256_DOWN), F256_DOWN);
            hi = f256_add(hi, f256_mul(x, s, F256_UP), F256_UP);
            to_mpfr(mx, &x);
            to_mpfr(my, &s);
            mpfr_mul(t, mx, my, MPFR_RNDD);
            mpfr_add(zlo, zlo, t, MPFR_RNDD);
            mpfr_mul(t, mx, my, MPFR_RNDU);
            mpfr_add(zhi, zhi, t, MPFR_RNDU);
        }
        if (!same(&lo, zlo) || !same(&hi, zhi)) {
            printf("chain mismatch at case %ld\n", k);
            return 1;
        }
        /* the round trip through MPFR is exact */
        const f256 back = from_mpfr(zlo);
        if (memcmp(back.m, lo.m, sizeof(lo.m)) || (!f256_is_zero(&lo) && back.e != lo.e)) {
            printf("round trip mismatch at case %ld\n", k);
            return 1;
        }
        mpfr_clears(zlo, zhi, t, (mpfr_ptr)0);
        checked += 2;
    }
    mpfr_clears(mx, my, want, (mpfr_ptr)0);
    printf("ok %ld\n", checked);
    return 0;
}
