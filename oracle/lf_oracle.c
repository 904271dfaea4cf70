/*
 * lf_oracle.c — CPU restatement of the reference arithmetic (TEST INFRASTRUCTURE ONLY).
 * See lf_oracle.h for scope, pinning and the rule that only tests/bench/smoke use it.
 *
 * Citations are path:line in the reference repository (vanbaalon/LimbForge).
 */
#include "lf_oracle.h"
#include <stdlib.h>
#include <string.h>

typedef uint32_t word;
typedef uint64_t dword;
typedef unsigned __int128 qword;

/* ---------------------------------------------------------------- records */
void lfo_load(int n, const void* rec, lfo_num* x)
{
    const unsigned char* p = (const unsigned char*)rec;
    memset(x, 0, sizeof *x);
    memcpy(x->limb, p, 4 * (size_t)n);
    memcpy(&x->exponent, p + 4 * n, 4);
    memcpy(&x->sign, p + 4 * n + 4, 4);
    memcpy(&x->status, p + 4 * n + 8, 4);
}

void lfo_store(int n, const lfo_num* x, void* rec)
{
    unsigned char* p = (unsigned char*)rec;
    memcpy(p, x->limb, 4 * (size_t)n);
    memcpy(p + 4 * n, &x->exponent, 4);
    memcpy(p + 4 * n + 4, &x->sign, 4);
    memcpy(p + 4 * n + 8, &x->status, 4);
}

/* canonical zero carrying a status (core.hpp:46-49) */
lfo_num lfo_zero(uint32_t status)
{
    lfo_num r;
    memset(&r, 0, sizeof r);
    r.status = status;
    return r;
}

/* --------------------------------------------------------- bit utilities */
/* position of the highest set bit of w[0..W), or -1 (core.hpp:51-59) */
static long top_bit(const word* w, long W)
{
    for (long i = W - 1; i >= 0; --i)
        if (w[i]) return 32 * i + 31 - __builtin_clz(w[i]);
    return -1;
}

/* the 32 bits of w starting at bit p; bits outside [0, 32W) read as zero (core.hpp:60-66,
 * linalg.hpp:227-228) */
static word bits_at(const word* w, long W, long p)
{
    if (p <= -32) return 0;
    if (p < 0) return W > 0 ? w[0] << (-p) : 0;
    long i = p / 32;
    int s = (int)(p % 32);
    word x = i < W ? w[i] >> s : 0;
    if (s && i + 1 < W) x |= w[i + 1] << (32 - s);
    return x;
}

/* is any bit below position p set (core.hpp:67-72) */
static int any_below(const word* w, long W, long p)
{
    if (p <= 0) return 0;
    long whole = p / 32;
    for (long i = 0; i < W && i < whole; ++i)
        if (w[i]) return 1;
    int s = (int)(p % 32);
    return whole < W && s && (w[whole] & ((1u << s) - 1u));
}

/* w |= x << shift for shift >= 0, words beyond W dropped (core.hpp:106-109, 244-247) */
static void place(word* w, long W, const word* x, long xw, long shift)
{
    long q = shift / 32;
    int s = (int)(shift % 32);
    for (long i = 0; i < xw; ++i) {
        long j = i + q;
        if (j < W) w[j] |= x[i] << s;
        if (s && j + 1 < W) w[j + 1] |= x[i] >> (32 - s);
    }
}

static void increment(int n, lfo_num* r) /* core.hpp:73-76 */
{
    for (int i = 0; i < n; ++i)
        if (++r->limb[i]) return;
    r->limb[n - 1] = 0x80000000u;
    r->exponent++;
}

static lfo_num checked(lfo_num r) /* core.hpp:77-80 */
{
    if (r.exponent > LFO_EMAX || r.exponent < -LFO_EMAX) return lfo_zero(r.status | LFO_EXPONENT_OVERFLOW);
    return r;
}

/* RN-even of sign * int(w[0..W)) * 2^scale to n words (core.hpp:81-99; linalg.hpp:222-236 is the same
 * rule for any word count).  The range check before rounding admits emin-1, which a carry can rescue. */
lfo_num lfo_pack(int n, const word* w, int W, int64_t scale, int sign, uint32_t status)
{
    long h = top_bit(w, W);
    if (h < 0) return lfo_zero(status);
    int64_t e = scale + h;
    if (e > LFO_EMAX || e < -(int64_t)LFO_EMAX - 1) return lfo_zero(status | LFO_EXPONENT_OVERFLOW);
    lfo_num r = lfo_zero(status);
    r.sign = sign;
    r.exponent = (int32_t)e;
    long shift = h - (32L * n - 1);
    for (int i = 0; i < n; ++i) r.limb[i] = bits_at(w, W, shift + 32L * i);
    if (shift > 0) {
        int round = (int)(bits_at(w, W, shift - 1) & 1u);
        if (round && (any_below(w, W, shift - 1) || (r.limb[0] & 1u))) increment(n, &r);
    }
    return checked(r);
}

/* ----------------------------------------------------------- add / sub */
static int magnitude_compare(int n, const lfo_num* a, const lfo_num* b) /* core.hpp:100-105 */
{
    if (!a->sign) return b->sign ? -1 : 0;
    if (!b->sign) return 1;
    if (a->exponent != b->exponent) return a->exponent > b->exponent ? 1 : -1;
    for (int i = n - 1; i >= 0; --i)
        if (a->limb[i] != b->limb[i]) return a->limb[i] > b->limb[i] ? 1 : -1;
    return 0;
}

/* exact aligned sum, then one rounding (core.hpp:116-134) */
lfo_num lfo_add(int n, lfo_num a, lfo_num b)
{
    uint32_t st = a.status | b.status;
    if (st) return lfo_zero(st);
    if (!a.sign) return b;
    if (!b.sign) return a;
    if (magnitude_compare(n, &a, &b) < 0) { lfo_num t = a; a = b; b = t; }
    int64_t gap = (int64_t)a.exponent - b.exponent;
    /* |b| below a quarter ulp of the representable a: RN(a+b) = a */
    if (gap > 32L * n + 2) return a;
    word x[2 * LFO_MAX_WORDS + 2], y[2 * LFO_MAX_WORDS + 2];
    memset(x, 0, sizeof x);
    memset(y, 0, sizeof y);
    long W, a_shift, b_shift;
    int64_t scale;
    if (gap <= 32) { /* close exponents: n+2 words hold the exact sum even under cancellation */
        W = n + 2; a_shift = (long)gap; b_shift = 0;
        scale = (int64_t)b.exponent - (32L * n - 1);
    } else {         /* far exponents: the whole exact alignment in 2n+2 words */
        W = 2L * n + 2; a_shift = 32L * n + 2; b_shift = 32L * n + 2 - (long)gap;
        scale = (int64_t)a.exponent - (32L * n - 1) - (32L * n + 2);
    }
    place(x, W, a.limb, n, a_shift);
    place(y, W, b.limb, n, b_shift);
    if (a.sign == b.sign) {
        dword c = 0;
        for (long i = 0; i < W; ++i) { dword v = (dword)x[i] + y[i] + c; x[i] = (word)v; c = v >> 32; }
    } else { /* |a| >= |b|: the difference is nonnegative */
        dword borrow = 0;
        for (long i = 0; i < W; ++i) {
            dword v = (dword)y[i] + borrow;
            borrow = (dword)x[i] < v;
            x[i] = (word)((dword)x[i] - v);
        }
    }
    return lfo_pack(n, x, (int)W, scale, a.sign, LFO_OK);
}

lfo_num lfo_sub(int n, lfo_num a, lfo_num b) /* core.hpp:135 */
{
    b.sign = -b.sign;
    return lfo_add(n, a, b);
}

/* ----------------------------------------------------------------- mul */
/* schoolbook n x n -> 2n+1 words, one rounding (core.hpp:137-145) */
lfo_num lfo_mul(int n, lfo_num a, lfo_num b)
{
    uint32_t st = a.status | b.status;
    if (st || !a.sign || !b.sign) return lfo_zero(st);
    word p[2 * LFO_MAX_WORDS + 1];
    memset(p, 0, sizeof p);
    for (int i = 0; i < n; ++i) {
        dword carry = 0;
        for (int j = 0; j < n; ++j) {
            dword v = (dword)a.limb[i] * b.limb[j] + p[i + j] + carry;
            p[i + j] = (word)v;
            carry = v >> 32;
        }
        p[i + n] = (word)carry;
    }
    return lfo_pack(n, p, 2 * n + 1, (int64_t)a.exponent + b.exponent - 2 * (32L * n - 1), a.sign * b.sign, st);
}

/* square = the exact integer product rounded once, as mul(a, a) (core.hpp:146-167) */
lfo_num lfo_square(int n, lfo_num a) { return lfo_mul(n, a, a); }

/* ----------------------------------------------------------------- div */
/* Two-word by one-word division with a precomputed reciprocal (core.hpp:176-184); this is the
 * 2/1 step of Moller & Granlund, "Improved division by invariant integers" (2011), Algorithm 4.
 * Preconditions: d >= 2^31, high < d, inv = floor((2^64-1)/d) - 2^32. */
uint64_t lfo_divide_word(uint32_t high, uint32_t low, uint32_t d, uint32_t inv)
{
    dword product = (dword)high * inv;
    dword sum = (dword)(word)product + low;
    word q = (word)((product >> 32) + high + 1 + (sum >> 32));
    word r = low - q * d;
    if (r > (word)sum) { --q; r += d; }
    if (r >= d) { ++q; r -= d; }
    return (dword)q | ((dword)r << 32);
}

/* normalized base-2^32 long division (Knuth D) with remainder-based RN (core.hpp:185-215) */
lfo_num lfo_div(int n, lfo_num a, lfo_num b)
{
    uint32_t st = a.status | b.status;
    if (st) return lfo_zero(st);
    if (!b.sign) return lfo_zero(LFO_DIVISION_BY_ZERO);
    if (!a.sign) return lfo_zero(LFO_OK);
    int cmp = 0;
    for (int i = n - 1; i >= 0; --i)
        if (a.limb[i] != b.limb[i]) { cmp = a.limb[i] > b.limb[i] ? 1 : -1; break; }
    long shift = 32L * n - 1 + (cmp < 0);
    word u[2 * LFO_MAX_WORDS + 2];
    memset(u, 0, sizeof u);
    place(u, 2L * n + 1, a.limb, n, shift);
    lfo_num r = lfo_zero(LFO_OK);
    r.sign = a.sign * b.sign;
    int64_t e = (int64_t)a.exponent - b.exponent - (cmp < 0);
    if (e > LFO_EMAX || e < -(int64_t)LFO_EMAX - 1) return lfo_zero(LFO_EXPONENT_OVERFLOW);
    r.exponent = (int32_t)e;
    const dword B = (dword)1 << 32;
    word d = b.limb[n - 1];
    word inv = (word)(~(dword)0 / d - B);
    for (int j = n; j >= 0; --j) {
        dword top = ((dword)u[j + n] << 32) | u[j + n - 1];
        dword q, rem;
        if (u[j + n] >= d) { q = B - 1; rem = top - q * d; }
        else {
            dword qr = lfo_divide_word(u[j + n], u[j + n - 1], d, inv);
            q = (word)qr; rem = qr >> 32;
        }
        /* refine the estimate with the second divisor word */
        while (rem < B && q * b.limb[n - 2] > (rem << 32) + u[j + n - 2]) { --q; rem += b.limb[n - 1]; }
        dword borrow = 0;
        for (int i = 0; i < n; ++i) {
            dword p = q * b.limb[i] + borrow;
            word low = (word)p, old = u[j + i];
            u[j + i] = old - low;
            borrow = (p >> 32) + (old < low);
        }
        int under = (dword)u[j + n] < borrow;
        u[j + n] = (word)((dword)u[j + n] - borrow);
        if (under) { /* estimate one too large: add the divisor back */
            --q;
            dword carry = 0;
            for (int i = 0; i < n; ++i) { dword v = (dword)u[j + i] + b.limb[i] + carry; u[j + i] = (word)v; carry = v >> 32; }
            u[j + n] += (word)carry;
        }
        if (j < n) r.limb[j] = (word)q;
    }
    /* round: compare twice the remainder with the divisor; ties go to even */
    word twice[LFO_MAX_WORDS + 1];
    word carry = 0;
    for (int i = 0; i < n; ++i) { twice[i] = (u[i] << 1) | carry; carry = u[i] >> 31; }
    twice[n] = carry;
    int half = 0;
    if (twice[n]) half = 1;
    else
        for (int i = n - 1; i >= 0; --i)
            if (twice[i] != b.limb[i]) { half = twice[i] > b.limb[i] ? 1 : -1; break; }
    if (half > 0 || (half == 0 && (r.limb[0] & 1u))) increment(n, &r);
    return checked(r);
}

/* ---------------------------------------------------------------- sqrt */
/* restoring integer square root of Z = mantissa << (32n-1+odd), round up iff Z - Q^2 > Q
 * (core.hpp:216-235; docs/numerics.md "Square and square root") */
lfo_num lfo_sqrt(int n, lfo_num a)
{
    if (a.status || !a.sign) return lfo_zero(a.status);
    if (a.sign < 0) return lfo_zero(LFO_INVALID);
    int e = a.exponent / 2;
    if (a.exponent < 0 && a.exponent % 2) --e; /* floor division */
    int odd = a.exponent - 2 * e;
    long shift = 32L * n - 1 + odd;
    word z[2 * LFO_MAX_WORDS + 1];
    memset(z, 0, sizeof z);
    place(z, 2L * n + 1, a.limb, n, shift);
    word rem[LFO_MAX_WORDS + 1], root[LFO_MAX_WORDS + 1], trial[LFO_MAX_WORDS + 1];
    memset(rem, 0, sizeof rem);
    memset(root, 0, sizeof root);
    for (long bit = 32L * n - 1; bit >= 0; --bit) {
        word pair = (bits_at(z, 2L * n + 1, 2 * bit)) & 3u;
        word c = pair;
        for (int i = 0; i <= n; ++i) { word old = rem[i]; rem[i] = (old << 2) | c; c = old >> 30; }
        c = 0;
        for (int i = 0; i < n; ++i) { word old = root[i]; root[i] = (old << 1) | c; c = old >> 31; }
        c = 1;
        for (int i = 0; i < n; ++i) { word old = root[i]; trial[i] = (old << 1) | c; c = old >> 31; }
        trial[n] = c;
        int cmp = 0;
        for (int i = n; i >= 0; --i)
            if (rem[i] != trial[i]) { cmp = rem[i] > trial[i] ? 1 : -1; break; }
        if (cmp >= 0) {
            dword borrow = 0;
            for (int i = 0; i <= n; ++i) {
                dword v = (dword)trial[i] + borrow;
                word old = rem[i];
                rem[i] = (word)((dword)old - v);
                borrow = (dword)old < v;
            }
            root[0] |= 1u;
        }
    }
    lfo_num r = lfo_zero(LFO_OK);
    r.sign = 1;
    r.exponent = e;
    for (int i = 0; i < n; ++i) r.limb[i] = root[i];
    int gt = 0;
    if (rem[n]) gt = 1;
    else
        for (int i = n - 1; i >= 0; --i)
            if (rem[i] != root[i]) { gt = rem[i] > root[i]; break; }
    if (gt) increment(n, &r);
    return checked(r);
}

/* ----------------------------------------------------------------- fma */
/* exact value sign * int(w) * 2^scale (core.hpp:238) with a generous fixed width */
#define EXW (4 * LFO_MAX_WORDS + 16)
typedef struct { word w[EXW]; int64_t scale; int sign; } exact_t;

static int64_t exact_msb(const exact_t* x) { return x->scale + top_bit(x->w, EXW); }

/* x has the larger msb.  Its window reaches at least 32n+34 bits below its msb; a y more than one bit
 * below that window contributes only the sign of a sticky unit (core.hpp:251-270). */
static void exact_add_ordered(int n, const exact_t* x, const exact_t* y, exact_t* r)
{
    memset(r, 0, sizeof *r);
    int64_t mx = exact_msb(x), my = exact_msb(y), low = mx - (32L * n + 33);
    if (x->scale < low) low = x->scale;
    if (my < low - 1) {
        r->scale = low - 32;
        r->sign = x->sign;
        place(r->w, EXW, x->w, EXW, (long)(x->scale - r->scale));
        if (x->sign == y->sign) r->w[0] = 1;
        else
            for (int i = 0; i < EXW; ++i) { word old = r->w[i]; r->w[i] = old - 1; if (old) break; }
        return;
    }
    static __thread word t[EXW];
    memset(t, 0, sizeof t);
    r->scale = x->scale < y->scale ? x->scale : y->scale;
    place(r->w, EXW, x->w, EXW, (long)(x->scale - r->scale));
    place(t, EXW, y->w, EXW, (long)(y->scale - r->scale));
    if (x->sign == y->sign) {
        dword c = 0;
        for (int i = 0; i < EXW; ++i) { dword v = (dword)r->w[i] + t[i] + c; r->w[i] = (word)v; c = v >> 32; }
        r->sign = x->sign;
        return;
    }
    int cmp = 0;
    for (int i = EXW - 1; i >= 0; --i)
        if (r->w[i] != t[i]) { cmp = r->w[i] > t[i] ? 1 : -1; break; }
    if (!cmp) { memset(r, 0, sizeof *r); return; }
    dword borrow = 0;
    for (int i = 0; i < EXW; ++i) {
        word p = cmp > 0 ? r->w[i] : t[i], q = cmp > 0 ? t[i] : r->w[i];
        dword v = (dword)q + borrow;
        r->w[i] = (word)((dword)p - v);
        borrow = (dword)p < v;
    }
    r->sign = cmp > 0 ? x->sign : y->sign;
}

static void exact_add(int n, const exact_t* x, const exact_t* y, exact_t* r) /* core.hpp:271-275 */
{
    if (!y->sign) { *r = *x; return; }
    if (!x->sign) { *r = *y; return; }
    if (exact_msb(x) >= exact_msb(y)) exact_add_ordered(n, x, y, r);
    else exact_add_ordered(n, y, x, r);
}

/* RN(a*b + c) with one rounding (core.hpp:295-326): the exact product and the addend in one
 * two's-complement window of F = roundup4(3n+4) words when they fit, else the ordered exact sum. */
lfo_num lfo_fma(int n, lfo_num a, lfo_num b, lfo_num c)
{
    uint32_t st = a.status | b.status | c.status;
    if (st) return lfo_zero(st);
    static __thread exact_t p, q, s;
    memset(&p, 0, sizeof p);
    memset(&q, 0, sizeof q);
    p.sign = a.sign * b.sign;
    if (p.sign) {
        p.scale = (int64_t)a.exponent + b.exponent - 2 * (32L * n - 1);
        for (int i = 0; i < n; ++i) {
            dword carry = 0;
            for (int j = 0; j < n; ++j) {
                dword v = (dword)a.limb[i] * b.limb[j] + p.w[i + j] + carry;
                p.w[i + j] = (word)v;
                carry = v >> 32;
            }
            p.w[i + n] = (word)carry;
        }
    }
    q.sign = c.sign;
    if (q.sign) {
        q.scale = (int64_t)c.exponent - (32L * n - 1);
        for (int i = 0; i < n; ++i) q.w[i] = c.limb[i];
    }
    /* window fast path (core.hpp:303-315) */
    {
        const int F = (3 * n + 4 + 3) & ~3;
        int any = 0;
        int64_t lo = 0, hi = 0;
        const exact_t* terms[2] = {&p, &q};
        for (int k = 0; k < 2; ++k) {
            if (!terms[k]->sign) continue;
            int64_t l = terms[k]->scale, h = exact_msb(terms[k]);
            if (!any || l < lo) lo = l;
            if (!any || h > hi) hi = h;
            any = 1;
        }
        if (!any) return lfo_zero(LFO_OK);
        if (hi + 3 - lo < 32L * F) {
            word acc[EXW], x[EXW];
            memset(acc, 0, sizeof acc);
            for (int k = 0; k < 2; ++k) {
                const exact_t* t = terms[k];
                if (!t->sign) continue;
                memset(x, 0, sizeof x);
                place(x, F, t->w, EXW, (long)(t->scale - lo));
                if (t->sign > 0) {
                    dword cy = 0;
                    for (int i = 0; i < F; ++i) { dword v = (dword)acc[i] + x[i] + cy; acc[i] = (word)v; cy = v >> 32; }
                } else {
                    dword bw = 0;
                    for (int i = 0; i < F; ++i) {
                        dword v = (dword)x[i] + bw;
                        word old = acc[i];
                        acc[i] = (word)((dword)old - v);
                        bw = (dword)old < v;
                    }
                }
            }
            int sign = 1;
            if (acc[F - 1] >> 31) {
                sign = -1;
                dword cy = 1;
                for (int i = 0; i < F; ++i) { dword v = (dword)(word)~acc[i] + cy; acc[i] = (word)v; cy = v >> 32; }
            }
            return lfo_pack(n, acc, F, lo, sign, LFO_OK);
        }
    }
    exact_add(n, &p, &q, &s);
    if (!s.sign) return lfo_zero(LFO_OK);
    return lfo_pack(n, s.w, EXW, s.scale, s.sign, LFO_OK);
}

lfo_num lfo_fms(int n, lfo_num a, lfo_num b, lfo_num c) /* core.hpp:327-329 */
{
    c.sign = -c.sign;
    return lfo_fma(n, a, b, c);
}

/* ------------------------------------------------------------- complex */
/* composed complex operations, each real primitive rounded (core.hpp:414-419) */
void lfo_cadd(int n, const lfo_num a[2], const lfo_num b[2], lfo_num r[2])
{
    r[0] = lfo_add(n, a[0], b[0]);
    r[1] = lfo_add(n, a[1], b[1]);
}

void lfo_cmul(int n, const lfo_num a[2], const lfo_num b[2], lfo_num r[2])
{
    lfo_num re = lfo_sub(n, lfo_mul(n, a[0], b[0]), lfo_mul(n, a[1], b[1]));
    lfo_num im = lfo_add(n, lfo_mul(n, a[0], b[1]), lfo_mul(n, a[1], b[0]));
    r[0] = re;
    r[1] = im;
}

void lfo_cdiv(int n, const lfo_num a[2], const lfo_num b[2], lfo_num r[2])
{
    lfo_num d = lfo_add(n, lfo_mul(n, b[0], b[0]), lfo_mul(n, b[1], b[1]));
    lfo_num re = lfo_div(n, lfo_add(n, lfo_mul(n, a[0], b[0]), lfo_mul(n, a[1], b[1])), d);
    lfo_num im = lfo_div(n, lfo_sub(n, lfo_mul(n, a[1], b[0]), lfo_mul(n, a[0], b[1])), d);
    r[0] = re;
    r[1] = im;
}

/* RN(a*b + c*d + e) with one rounding (core.hpp:330-352, dot2_add): the exact dot of the two products with
 * e as its addend, the same exact value under the same conventions (a status anywhere gives zero with the OR
 * of the statuses, exact cancellation gives canonical zero, no nonzero product gives e itself). */
static lfo_num exact_dot_add_n(int n, const lfo_num* c, int subtract, const unsigned char* a, ptrdiff_t sa,
                               const unsigned char* b, ptrdiff_t sb, size_t K);
lfo_num lfo_dot2_add(int n, lfo_num a, lfo_num b, lfo_num c, lfo_num d, lfo_num e)
{
    unsigned char x[2 * (4 * LFO_MAX_WORDS + 12)], y[2 * (4 * LFO_MAX_WORDS + 12)];
    const size_t E = 4 * (size_t)n + 12;
    lfo_store(n, &a, x);
    lfo_store(n, &c, x + E);
    lfo_store(n, &b, y);
    lfo_store(n, &d, y + E);
    return exact_dot_add_n(n, &e, 0, x, 1, y, 1, 2);
}

/* complex a*b + c and a*b - c, each part one rounding of its exact value (core.hpp:353-358) */
void lfo_cfma(int n, const lfo_num a[2], const lfo_num b[2], const lfo_num c[2], lfo_num r[2])
{
    lfo_num m = a[1];
    m.sign = -m.sign;
    lfo_num re = lfo_dot2_add(n, a[0], b[0], m, b[1], c[0]);
    lfo_num im = lfo_dot2_add(n, a[0], b[1], a[1], b[0], c[1]);
    r[0] = re;
    r[1] = im;
}

void lfo_cfms(int n, const lfo_num a[2], const lfo_num b[2], const lfo_num c[2], lfo_num r[2])
{
    lfo_num d[2] = {c[0], c[1]};
    d[0].sign = -d[0].sign;
    d[1].sign = -d[1].sign;
    lfo_cfma(n, a, b, d, r);
}

/* ---------------------------------------------------------------- cast */
/* exact widening, RN narrowing (kernels.metal:251-258) */
lfo_num lfo_cast_words(int to_n, int from_n, const lfo_num* x)
{
    if (x->status || !x->sign) return lfo_zero(x->status);
    return lfo_pack(to_n, x->limb, from_n, (int64_t)x->exponent - (32L * from_n - 1), x->sign, LFO_OK);
}

/* --------------------------------------------------------------- batch */
static int valid_bits(int bits) { return bits >= 64 && bits <= 1024 && bits % 32 == 0; }

static size_t bidx(size_t i, size_t stride, size_t period) /* kernels.metal:7 */
{
    size_t j = stride > 1 ? i / stride : i;
    return period ? j % period : j;
}

int lfo_elementwise(int bits, int op, const void* a, const void* b, const void* c, void* out,
                    size_t count, size_t b_stride, size_t b_period, size_t c_stride, size_t c_period)
{
    if (!valid_bits(bits)) return -1;
    const int n = bits / 32;
    const size_t E = 4 * (size_t)n + 12;
    const unsigned char* pa = (const unsigned char*)a;
    const unsigned char* pb = (const unsigned char*)b;
    const unsigned char* pc = (const unsigned char*)c;
    unsigned char* po = (unsigned char*)out;
    int cplx = op == LFO_CADD || op == LFO_CMUL || op == LFO_CDIV || op == LFO_CFMA || op == LFO_CFMS;
    if (op < 0 || op > LFO_CFMS) return -1;
    for (size_t i = 0; i < count; ++i) {
        size_t jb = bidx(i, b_stride, b_period), jc = bidx(i, c_stride, c_period);
        if (cplx) {
            lfo_num x[2], y[2], z[2];
            lfo_load(n, pa + 2 * E * i, &x[0]);
            lfo_load(n, pa + 2 * E * i + E, &x[1]);
            lfo_load(n, pb + 2 * E * jb, &y[0]);
            lfo_load(n, pb + 2 * E * jb + E, &y[1]);
            if (op == LFO_CADD) lfo_cadd(n, x, y, z);
            else if (op == LFO_CMUL) lfo_cmul(n, x, y, z);
            else if (op == LFO_CDIV) lfo_cdiv(n, x, y, z);
            else {
                lfo_num w[2];
                lfo_load(n, pc + 2 * E * jc, &w[0]);
                lfo_load(n, pc + 2 * E * jc + E, &w[1]);
                if (op == LFO_CFMA) lfo_cfma(n, x, y, w, z);
                else lfo_cfms(n, x, y, w, z);
            }
            lfo_store(n, &z[0], po + 2 * E * i);
            lfo_store(n, &z[1], po + 2 * E * i + E);
            continue;
        }
        lfo_num x, y, z, r;
        lfo_load(n, pa + E * i, &x);
        if (op == LFO_SQUARE) r = lfo_square(n, x);
        else if (op == LFO_SQRT) r = lfo_sqrt(n, x);
        else {
            lfo_load(n, pb + E * jb, &y);
            switch (op) {
            case LFO_ADD: r = lfo_add(n, x, y); break;
            case LFO_SUB: r = lfo_sub(n, x, y); break;
            case LFO_MUL: r = lfo_mul(n, x, y); break;
            case LFO_DIV: r = lfo_div(n, x, y); break;
            default:
                lfo_load(n, pc + E * jc, &z);
                r = op == LFO_FMA ? lfo_fma(n, x, y, z) : lfo_fms(n, x, y, z);
            }
        }
        lfo_store(n, &r, po + E * i);
    }
    return 0;
}

int lfo_mul_chain(int bits, const void* x0, const void* f, void* out, size_t count, int steps)
{
    if (!valid_bits(bits)) return -1;
    const int n = bits / 32;
    const size_t E = 4 * (size_t)n + 12;
    for (size_t i = 0; i < count; ++i) {
        lfo_num x, y;
        lfo_load(n, (const unsigned char*)x0 + E * i, &x);
        lfo_load(n, (const unsigned char*)f + E * i, &y);
        for (int s = 0; s < steps; ++s) x = lfo_mul(n, x, y);
        lfo_store(n, &x, (unsigned char*)out + E * i);
    }
    return 0;
}

int lfo_cast(int from_bits, int to_bits, int complex_values, const void* in, void* out, size_t count)
{
    if (!valid_bits(from_bits) || !valid_bits(to_bits)) return -1;
    const int fn = from_bits / 32, tn = to_bits / 32;
    const size_t Ef = 4 * (size_t)fn + 12, Et = 4 * (size_t)tn + 12;
    const size_t parts = complex_values ? 2 : 1;
    for (size_t i = 0; i < count * parts; ++i) {
        lfo_num x, r;
        lfo_load(fn, (const unsigned char*)in + Ef * i, &x);
        r = lfo_cast_words(tn, fn, &x);
        lfo_store(tn, &r, (unsigned char*)out + Et * i);
    }
    return 0;
}

/* ------------------------------------------------------ exact dot product */
/* acc (two's complement, W words) += sign * x * 2^shift (linalg.hpp:258-266) */
static void accumulate(word* acc, long W, const word* x, long xw, long shift, int sign)
{
    long q = shift / 32;
    int s = (int)(shift % 32);
    dword carry = 0, borrow = 0;
    for (long i = 0; i + q < W; ++i) {
        word v = 0;
        if (i < xw) v = x[i] << s;
        if (s && i > 0 && i - 1 < xw) v |= x[i - 1] >> (32 - s);
        if (i > xw && !carry && !borrow) break;
        if (sign > 0) { dword t = (dword)acc[i + q] + v + carry; acc[i + q] = (word)t; carry = t >> 32; }
        else { dword t = (dword)v + borrow; borrow = (dword)acc[i + q] < t; acc[i + q] = (word)((dword)acc[i + q] - t); }
    }
}

typedef struct { int64_t scale; size_t k; } term_t;

static int term_desc(const void* u, const void* v)
{
    const term_t* a = (const term_t*)u;
    const term_t* b = (const term_t*)v;
    if (a->scale != b->scale) return a->scale > b->scale ? -1 : 1;
    return a->k < b->k ? -1 : (a->k > b->k);
}

/* RN(c -/+ sum_k a_k b_k) with one rounding (linalg.hpp:268-318): one exact positive/negative
 * accumulator pair when every term fits 254 words; otherwise exponent-sorted clusters, where only the
 * sign of the next nonzero cluster matters below the first nonzero cluster sum. */
static lfo_num exact_dot_add_n(int n, const lfo_num* c, int subtract, const unsigned char* a, ptrdiff_t sa,
                               const unsigned char* b, ptrdiff_t sb, size_t K)
{
    const size_t E = 4 * (size_t)n + 12;
    uint32_t status = c ? c->status : 0;
    lfo_num x, y;
#define AT(base, stride, k) ((base) + (ptrdiff_t)(k) * (stride) * (ptrdiff_t)E)
    for (size_t k = 0; k < K; ++k) {
        lfo_load(n, AT(a, sa, k), &x);
        lfo_load(n, AT(b, sb, k), &y);
        status |= x.status | y.status;
    }
    if (status) return lfo_zero(status);
    term_t* t = (term_t*)malloc((K + 1) * sizeof(term_t));
    size_t count = 0;
    for (size_t k = 0; k < K; ++k) {
        lfo_load(n, AT(a, sa, k), &x);
        lfo_load(n, AT(b, sb, k), &y);
        if (x.sign && y.sign) { t[count].scale = (int64_t)x.exponent + y.exponent - 2 * (32L * n - 1); t[count].k = k; ++count; }
    }
    if (!count) { free(t); return c ? *c : lfo_zero(LFO_OK); }
    if (c && c->sign) { t[count].scale = (int64_t)c->exponent - (32L * n - 1); t[count].k = K; ++count; }
    word p[2 * LFO_MAX_WORDS];
    lfo_num result;
    /* exact term k into p: 2n-word product, or the n-word addend (k == K) */
#define TERM(kk, words, sign)                                                                       \
    do {                                                                                            \
        if ((kk) == K) { for (int u = 0; u < n; ++u) p[u] = c->limb[u]; (words) = n; (sign) = c->sign; } \
        else {                                                                                      \
            lfo_load(n, AT(a, sa, kk), &x); lfo_load(n, AT(b, sb, kk), &y);                          \
            memset(p, 0, sizeof p);                                                                 \
            for (int i = 0; i < n; ++i) { dword cy = 0;                                             \
                for (int j = 0; j < n; ++j) { dword v = (dword)x.limb[i] * y.limb[j] + p[i + j] + cy; p[i + j] = (word)v; cy = v >> 32; } \
                p[i + n] = (word)cy; }                                                              \
            (words) = 2 * n; (sign) = subtract ? -x.sign * y.sign : x.sign * y.sign;                 \
        }                                                                                           \
    } while (0)
    int64_t hi = t[0].scale, lo = t[0].scale;
    for (size_t i = 1; i < count; ++i) { if (t[i].scale > hi) hi = t[i].scale; if (t[i].scale < lo) lo = t[i].scale; }
    int logc = 0;
    while (((size_t)1 << logc) < count) ++logc;
    const int64_t need = hi - lo + 64L * n + logc + 1;
    if (need <= 32L * (256 - 2)) {
        const long W = (long)(need / 32) + 2;
        word* pos = (word*)calloc((size_t)W + 1, 4);
        word* neg = (word*)calloc((size_t)W + 1, 4);
        for (size_t i = 0; i < count; ++i) {
            int words, sign;
            TERM(t[i].k, words, sign);
            accumulate(sign > 0 ? pos : neg, W, p, words, (long)(t[i].scale - lo), 1);
        }
        int cmp = 0;
        for (long i = W - 1; i >= 0; --i)
            if (pos[i] != neg[i]) { cmp = pos[i] > neg[i] ? 1 : -1; break; }
        if (!cmp) result = lfo_zero(LFO_OK);
        else {
            word* hi_acc = cmp > 0 ? pos : neg;
            const word* lo_acc = cmp > 0 ? neg : pos;
            dword borrow = 0;
            for (long i = 0; i < W; ++i) {
                dword v = (dword)lo_acc[i] + borrow;
                borrow = (dword)hi_acc[i] < v;
                hi_acc[i] = (word)((dword)hi_acc[i] - v);
            }
            result = lfo_pack(n, hi_acc, (int)W, lo, cmp, LFO_OK);
        }
        free(pos);
        free(neg);
        free(t);
        return result;
    }
    qsort(t, count, sizeof(term_t), term_desc);
    int logk = 0;
    while (((size_t)1 << logk) < K + 1) ++logk;
    const int64_t gap = 32L * n + logk + 4, width = 64L * n; /* a term is < 2^(scale + 64n) */
    word* sums[2] = {NULL, NULL};
    long sum_words[2] = {0, 0};
    int64_t sum_scale[2] = {0, 0};
    int sum_sign[2] = {0, 0}, nsums = 0;
    for (size_t i = 0; i < count && nsums < 2;) {
        size_t j = i + 1;
        int64_t low = t[i].scale;
        while (j < count && t[j].scale + width - 1 + gap >= low) { if (t[j].scale < low) low = t[j].scale; ++j; }
        long W = (long)((t[i].scale + width - low) / 32 + 3);
        word* acc = (word*)calloc((size_t)W, 4);
        for (size_t q = i; q < j; ++q) {
            int words, sign;
            TERM(t[q].k, words, sign);
            accumulate(acc, W, p, words, (long)(t[q].scale - low), sign);
        }
        int sign = 1;
        if (acc[W - 1] >> 31) {
            sign = -1;
            dword cy = 1;
            for (long u = 0; u < W; ++u) { dword z = (dword)(word)~acc[u] + cy; acc[u] = (word)z; cy = z >> 32; }
        }
        int nonzero = 0;
        for (long u = 0; u < W; ++u) nonzero |= acc[u] != 0;
        if (nonzero) { sums[nsums] = acc; sum_words[nsums] = W; sum_scale[nsums] = low; sum_sign[nsums] = sign; ++nsums; }
        else free(acc);
        i = j;
    }
    free(t);
    if (!nsums) return lfo_zero(LFO_OK);
    if (nsums == 1) {
        result = lfo_pack(n, sums[0], (int)sum_words[0], sum_scale[0], sum_sign[0], LFO_OK);
        free(sums[0]);
        return result;
    }
    /* S * 2^D -/+ 1 puts the sticky unit below the rounding position of S */
    const long D = 32L * n + 64;
    long XW = sum_words[0] + D / 32 + 1;
    word* ext = (word*)calloc((size_t)XW, 4);
    for (long i = 0; i < sum_words[0]; ++i) ext[i + D / 32] = sums[0][i];
    if (sum_sign[1] == sum_sign[0]) ext[0] = 1;
    else
        for (long i = 0; i < XW; ++i) { word old = ext[i]; ext[i] = old - 1; if (old) break; }
    result = lfo_pack(n, ext, (int)XW, sum_scale[0] - D, sum_sign[0], LFO_OK);
    free(ext);
    free(sums[0]);
    free(sums[1]);
    return result;
#undef TERM
#undef AT
}

void lfo_exact_dot_add(int bits, const void* c, int subtract, const void* a, ptrdiff_t sa,
                       const void* b, ptrdiff_t sb, size_t K, void* out)
{
    const int n = bits / 32;
    lfo_num cc, r;
    if (c) lfo_load(n, c, &cc);
    r = exact_dot_add_n(n, c ? &cc : NULL, subtract, (const unsigned char*)a, sa, (const unsigned char*)b, sb, K);
    lfo_store(n, &r, out);
}

int lfo_mixed_chain(int bits, const void* x, const void* f, const void* g, const uint32_t* len, const uint64_t* mask,
                    void* out, size_t count)
{
    if (!valid_bits(bits)) return -1;
    const int n = bits / 32;
    const size_t E = 4 * (size_t)n + 12;
    for (size_t i = 0; i < count; ++i) {
        lfo_num v, a, c;
        lfo_load(n, (const unsigned char*)x + i * E, &v);
        lfo_load(n, (const unsigned char*)f + i * E, &a);
        lfo_load(n, (const unsigned char*)g + i * E, &c);
        for (uint32_t t = 0; t < len[i]; ++t) v = ((mask[i] >> t) & 1u) ? lfo_fma(n, v, a, c) : lfo_mul(n, v, a);
        lfo_store(n, &v, (unsigned char*)out + i * E);
    }
    return 0;
}

/* --------------------------------------------------------- dense products */
int lfo_gemm(int bits, int transpose_a, const void* A, const void* B, size_t m, size_t n, size_t k,
             void* C, int subtract)
{
    if (!valid_bits(bits)) return -1;
    const int nw = bits / 32;
    const size_t E = 4 * (size_t)nw + 12;
    const unsigned char* pa = (const unsigned char*)A;
    const unsigned char* pb = (const unsigned char*)B;
    unsigned char* pc = (unsigned char*)C;
    for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < n; ++j) {
            const unsigned char* row = transpose_a ? pa + i * E : pa + i * k * E;
            ptrdiff_t sa = transpose_a ? (ptrdiff_t)m : 1;
            lfo_num old, r;
            lfo_load(nw, pc + (i * n + j) * E, &old);
            r = exact_dot_add_n(nw, subtract ? &old : NULL, subtract, row, sa, pb + j * E, (ptrdiff_t)n, k);
            lfo_store(nw, &r, pc + (i * n + j) * E);
        }
    return 0;
}

int lfo_syrk(int bits, const void* A, size_t rows, size_t cols, void* C, int lower_only, int subtract)
{
    if (!valid_bits(bits)) return -1;
    if (subtract && !lower_only) return -1;
    const int nw = bits / 32;
    const size_t E = 4 * (size_t)nw + 12;
    const unsigned char* pa = (const unsigned char*)A;
    unsigned char* pc = (unsigned char*)C;
    for (size_t i = 0; i < cols; ++i)
        for (size_t j = 0; j <= i; ++j) {
            lfo_num old, r;
            lfo_load(nw, pc + (i * cols + j) * E, &old);
            r = exact_dot_add_n(nw, subtract ? &old : NULL, subtract, pa + i * E, (ptrdiff_t)cols, pa + j * E, (ptrdiff_t)cols, rows);
            lfo_store(nw, &r, pc + (i * cols + j) * E);
        }
    if (!lower_only)
        for (size_t i = 0; i < cols; ++i)
            for (size_t j = i + 1; j < cols; ++j) memcpy(pc + (i * cols + j) * E, pc + (j * cols + i) * E, E);
    return 0;
}

/* --------------------------------------------------------- reductions */
/* Segmented dot products (kernels.metal:197-208, docs/numerics.md "Segmented dot products"): out[i] is the
 * sequential chain s <- fma(a[i K + k], b[i K + k] or b[k], s) from s = 0 (cfma for complex values). */
int lfo_segmented_dot(int bits, int complex_values, size_t segments, size_t length, int shared_right, const void* a,
                      const void* b, void* out)
{
    if (!valid_bits(bits)) return -1;
    const int n = bits / 32;
    const size_t E = 4 * (size_t)n + 12, P = complex_values ? 2 : 1;
    const unsigned char* pa = (const unsigned char*)a;
    const unsigned char* pb = (const unsigned char*)b;
    unsigned char* po = (unsigned char*)out;
    for (size_t i = 0; i < segments; ++i) {
        lfo_num s[2] = {lfo_zero(LFO_OK), lfo_zero(LFO_OK)};
        for (size_t k = 0; k < length; ++k) {
            const size_t ia = i * length + k, ib = shared_right ? k : i * length + k;
            lfo_num x[2], y[2], r[2];
            for (size_t q = 0; q < P; ++q) {
                lfo_load(n, pa + (ia * P + q) * E, &x[q]);
                lfo_load(n, pb + (ib * P + q) * E, &y[q]);
            }
            if (complex_values) {
                lfo_cfma(n, x, y, s, r);
                s[0] = r[0];
                s[1] = r[1];
            } else
                s[0] = lfo_fma(n, x[0], y[0], s[0]);
        }
        for (size_t q = 0; q < P; ++q) lfo_store(n, &s[q], po + (i * P + q) * E);
    }
    return 0;
}

/* Adjacent-pair tree sum (kernels.metal:61-112, docs/numerics.md "Pairwise reductions"): pairs (0,1), (2,3),
 * ... added with one rounding each, an unpaired last value carried unchanged, level by level; complex
 * values componentwise; an empty input gives canonical zero. */
int lfo_tree_sum(int bits, int complex_values, const void* in, size_t count, void* out)
{
    if (!valid_bits(bits)) return -1;
    const int n = bits / 32;
    const size_t E = 4 * (size_t)n + 12, P = complex_values ? 2 : 1;
    if (!count) {
        lfo_num z = lfo_zero(LFO_OK);
        for (size_t q = 0; q < P; ++q) lfo_store(n, &z, (unsigned char*)out + q * E);
        return 0;
    }
    lfo_num* v = (lfo_num*)calloc(count * P, sizeof(lfo_num));
    for (size_t i = 0; i < count * P; ++i) lfo_load(n, (const unsigned char*)in + i * E, &v[i]);
    while (count > 1) {
        const size_t next = (count + 1) / 2;
        for (size_t j = 0; j < next; ++j)
            for (size_t q = 0; q < P; ++q)
                v[j * P + q] = 2 * j + 1 < count ? lfo_add(n, v[2 * j * P + q], v[(2 * j + 1) * P + q]) : v[2 * j * P + q];
        count = next;
    }
    for (size_t q = 0; q < P; ++q) lfo_store(n, &v[q], (unsigned char*)out + q * E);
    free(v);
    return 0;
}

/* ----------------------------------------------------- Cholesky / solves */
/* Host threads of the factorization and solves: the panel rows below a diagonal block, the rows of a
 * trailing update and the right-hand sides are independent given what precedes them (the reference's host
 * path runs them on its worker pool, linalg.mm:418-441), so they run in parallel; every entry is the same
 * rounding of the same exact value whatever the thread count.  1 (the default) runs serially. */
static int lfo_threads = 1;
void lfo_set_threads(int threads) { lfo_threads = threads > 0 ? threads : 1; }

/* D(c; x, y; K) = RN(c - sum x_k y_k), one rounding (linalg.mm:409-411) */
static lfo_num D(int n, const lfo_num* c, const unsigned char* x, ptrdiff_t sx, const unsigned char* y, ptrdiff_t sy, size_t K)
{
    return exact_dot_add_n(n, c, 1, x, sx, y, sy, K);
}

/* The reference's host order (linalg.mm:411-456 with FactorOptions::gpu = false): panel [k0, k1) row by
 * row (diagonal block, then the rows below), then the trailing updates by that block. */
size_t lfo_cholesky(int bits, const void* A, size_t n, void* L, size_t block, int* pivot_sign, uint32_t* pivot_status)
{
    const int nw = bits / 32;
    const size_t E = 4 * (size_t)nw + 12;
    const unsigned char* pa = (const unsigned char*)A;
    unsigned char* W = (unsigned char*)malloc(n * n * E + 1);
    size_t pivot = n;
    if (pivot_sign) *pivot_sign = 0;
    if (pivot_status) *pivot_status = 0;
#define EL(i, j) (W + ((i) * n + (j)) * E)
    lfo_num zero = lfo_zero(0), v, s, dgn;
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            if (j > i) lfo_store(nw, &zero, EL(i, j));
            else memcpy(EL(i, j), pa + (i * n + j) * E, E);
        }
    const size_t nb = block ? (block < n ? block : n) : n;
    size_t k0 = 0, k1 = nb < n ? nb : n, p = 0;
    /* columns [k0, ke) of row i */
#define PANEL(i, kk0, ke)                                                                      \
    for (size_t jj = (kk0); jj < (ke); ++jj) {                                                 \
        lfo_load(nw, EL(i, jj), &v);                                                           \
        v = D(nw, &v, EL(i, kk0), 1, EL(jj, kk0), 1, jj - (kk0));                              \
        lfo_load(nw, EL(jj, jj), &dgn);                                                        \
        v = lfo_div(nw, v, dgn);                                                               \
        lfo_store(nw, &v, EL(i, jj));                                                          \
    }
#define FACTOR_PANEL(kk0, kk1, out)                                                            \
    do {                                                                                       \
        size_t pp = (kk1);                                                                     \
        for (size_t i = (kk0); i < (kk1) && pp == (kk1); ++i) {                                \
            PANEL(i, kk0, i);                                                                  \
            lfo_load(nw, EL(i, i), &v);                                                        \
            s = D(nw, &v, EL(i, kk0), 1, EL(i, kk0), 1, i - (kk0));                             \
            if (s.status || s.sign <= 0) {                                                     \
                pp = i; pivot = i;                                                             \
                if (pivot_sign) *pivot_sign = s.status ? 0 : s.sign;                           \
                if (pivot_status) *pivot_status = s.status;                                    \
            } else { s = lfo_sqrt(nw, s); lfo_store(nw, &s, EL(i, i)); }                       \
        }                                                                                      \
        size_t first = pp < (kk1) ? pp + 1 : (kk1);                                            \
        _Pragma("omp parallel for schedule(dynamic, 2) private(v, dgn) num_threads(lfo_threads)")  \
        for (size_t r = first; r < n; ++r) { PANEL(r, kk0, pp); }                              \
        (out) = pp;                                                                            \
    } while (0)
#define UPDATE(kk0, kk1, r0, c1)                                                               \
    _Pragma("omp parallel for schedule(dynamic, 2) private(v) num_threads(lfo_threads)")     \
    for (size_t i = (r0); i < n; ++i)                                                          \
        for (size_t j = (r0); j <= i && j < (c1); ++j) {                                       \
            lfo_load(nw, EL(i, j), &v);                                                        \
            v = D(nw, &v, EL(i, kk0), 1, EL(j, kk0), 1, (kk1) - (kk0));                         \
            lfo_store(nw, &v, EL(i, j));                                                       \
        }
    if (n) FACTOR_PANEL(0, k1, p);
    while (p == k1 && k1 < n) {
        size_t k2 = k1 + nb < n ? k1 + nb : n;
        if (k2 < n) { UPDATE(k0, k1, k1, k2); UPDATE(k0, k1, k2, n); }
        else { UPDATE(k0, k1, k1, n); }
        FACTOR_PANEL(k1, k2, p);
        k0 = k1;
        k1 = k2;
    }
    if (p < n) {
        lfo_num bad = lfo_zero(LFO_INVALID);
        for (size_t i = p; i < n; ++i)
            for (size_t j = p; j <= i; ++j) lfo_store(nw, &bad, EL(i, j));
    }
    memcpy(L, W, n * n * E);
    free(W);
    return n ? pivot : 0;
#undef PANEL
#undef FACTOR_PANEL
#undef UPDATE
#undef EL
}

/* Triangular solves in the reference's host order (linalg.mm:459-486, upper = false). */
int lfo_solve(int bits, const void* L, size_t n, const void* B, size_t nrhs, void* X, size_t block, int passes)
{
    const int nw = bits / 32;
    const size_t E = 4 * (size_t)nw + 12;
    const unsigned char* F = (const unsigned char*)L;
    unsigned char* W = (unsigned char*)X;
    if (!n || !nrhs) return 0;
    if (W != (unsigned char*)B) memmove(W, B, n * nrhs * E);
    const size_t nb = block ? (block < n ? block : n) : n;
    lfo_num v, dgn;
#define FL(i, j) (F + ((i) * n + (j)) * E)
#define XW(i, c) (W + ((i) * nrhs + (c)) * E)
    /* out rows (M x nrhs) = D(out; line i of the factor view, X rows [r0, r0+K)) */
#define SOLVE_UPDATE(base, line_stride, k_stride, M, r0, K, out)                               \
    _Pragma("omp parallel for schedule(dynamic, 2) private(v) num_threads(lfo_threads)")     \
    for (size_t i = 0; i < (M); ++i)                                                           \
        for (size_t c = 0; c < nrhs; ++c) {                                                    \
            unsigned char* o = (out) + (i * nrhs + c) * E;                                     \
            lfo_load(nw, o, &v);                                                               \
            v = D(nw, &v, (base) + i * (line_stride) * E, (ptrdiff_t)(k_stride), XW(r0, c), (ptrdiff_t)nrhs, (K)); \
            lfo_store(nw, &v, o);                                                              \
        }
    if (passes & 1)
        for (size_t r0 = 0; r0 < n; r0 += nb) {
            size_t r1 = r0 + nb < n ? r0 + nb : n;
#pragma omp parallel for schedule(dynamic, 1) private(v, dgn) num_threads(lfo_threads)
            for (size_t c = 0; c < nrhs; ++c)
                for (size_t i = r0; i < r1; ++i) {
                    lfo_load(nw, XW(i, c), &v);
                    v = D(nw, &v, FL(i, r0), 1, XW(r0, c), (ptrdiff_t)nrhs, i - r0);
                    lfo_load(nw, FL(i, i), &dgn);
                    v = lfo_div(nw, v, dgn);
                    lfo_store(nw, &v, XW(i, c));
                }
            if (r1 < n) { SOLVE_UPDATE(FL(r1, r0), n, 1, n - r1, r0, r1 - r0, XW(r1, 0)); }
        }
    if (passes & 2)
        for (size_t r0 = (n - 1) / nb * nb;; r0 -= nb) {
            size_t r1 = r0 + nb < n ? r0 + nb : n;
#pragma omp parallel for schedule(dynamic, 1) private(v, dgn) num_threads(lfo_threads)
            for (size_t c = 0; c < nrhs; ++c)
                for (size_t i = r1; i-- > r0;) {
                    lfo_load(nw, XW(i, c), &v);
                    if (i + 1 < r1) v = D(nw, &v, FL(i + 1, i), (ptrdiff_t)n, XW(i + 1, c), (ptrdiff_t)nrhs, r1 - i - 1);
                    lfo_load(nw, FL(i, i), &dgn);
                    v = lfo_div(nw, v, dgn);
                    lfo_store(nw, &v, XW(i, c));
                }
            if (r0) { SOLVE_UPDATE(FL(r0, 0), 1, n, r0, r0, r1 - r0, XW(0, 0)); }
            if (!r0) break;
        }
    return 0;
#undef SOLVE_UPDATE
#undef FL
#undef XW
}
