// exact_dot.cpp — the host exact dot product behind limbforge::exact_dot_add / exact_dot
// (include/limbforge/linalg.hpp; the reference's contract at include/limbforge/linalg.hpp:268-321):
// RN(c - sum_k a_k b_k) (subtract) or RN(c + sum ...), one rounding, on the CPU.
//
// Every nonzero product a_k b_k is an exact 2N-word integer times 2^scale; the addend c is one more
// (N-word) term.  Terms are sorted by scale and split into clusters: a new cluster starts where every
// remaining term lies more than 32N + ceil(log2(K+1)) + 4 bits below the current cluster's lowest bit.
// Each cluster is summed exactly in a two's-complement accumulator.  With S the first nonzero cluster sum
// and F the sum of everything after it, |F| < 2^(lsb(S) - 32N - 2), so RN(S + F) = RN(S + sign(F) * eps)
// for a tiny eps below S's last bit, and sign(F) is the sign of the next nonzero cluster sum.
// Conventions (the reference's): any status among c and the a_k, b_k gives zero with the OR of those
// statuses; with no nonzero product the result is c itself; exact cancellation gives canonical zero.
#include "limbforge/linalg.hpp"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace limbforge {
namespace detail {
namespace {
using u32 = std::uint32_t;
using u64 = std::uint64_t;
using i64 = std::int64_t;

constexpr i64 EMAX = 1000000000;

struct View {
    const u32* limb;
    int exponent, sign;
    u32 status;
};

View view(const void* p, int n)
{
    const u32* w = static_cast<const u32*>(p);
    return {w, int(w[n]), int(w[n + 1]), w[n + 2]};
}

void write_zero(u32* out, int n, u32 status)
{
    std::memset(out, 0, sizeof(u32) * (n + 3));
    out[n + 2] = status;
}

// Bit p..p+31 of the magnitude w (bits outside w are zero).
u32 bits_at(const std::vector<u32>& w, i64 p)
{
    if (p <= -32) return 0;
    if (p < 0) return w[0] << (-p);
    const u64 i = u64(p) / 32;
    const int s = int(p % 32);
    u32 x = i < w.size() ? w[i] >> s : 0;
    if (s && i + 1 < w.size()) x |= w[i + 1] << (32 - s);
    return x;
}

// out = RN(sign * integer(w) * 2^scale) to n words, ties to even; exponent range [-1e9, 1e9] (a value at
// -1e9 - 1 before rounding may round up into the range).
void round_to(const std::vector<u32>& w, i64 scale, int sign, int n, u32* out)
{
    i64 top = i64(w.size());
    while (top && !w[top - 1]) --top;
    if (!top) {
        write_zero(out, n, 0);
        return;
    }
    const i64 h = 32 * (top - 1) + 31 - __builtin_clz(w[top - 1]);
    i64 e = scale + h;
    if (e > EMAX || e < -EMAX - 1) {
        write_zero(out, n, 2);
        return;
    }
    const i64 shift = h - (32 * i64(n) - 1);
    for (int i = 0; i < n; ++i) out[i] = bits_at(w, shift + 32 * i);
    if (shift > 0) {
        const i64 rb = shift - 1;
        const bool round = (w[rb / 32] >> (rb % 32)) & 1;
        bool below = (rb % 32) && (w[rb / 32] & ((u32(1) << (rb % 32)) - 1));
        for (i64 i = 0; i < rb / 32 && !below; ++i) below = w[i] != 0;
        if (round && (below || (out[0] & 1))) {
            int i = 0;
            while (i < n && ++out[i] == 0) ++i;
            if (i == n) {
                out[n - 1] = 0x80000000u;
                ++e;
            }
        }
    }
    if (e > EMAX || e < -EMAX) {
        write_zero(out, n, 2);
        return;
    }
    out[n] = u32(int(e));
    out[n + 1] = u32(sign);
    out[n + 2] = 0;
}

// p (2n words) = x * y for n-word magnitudes, schoolbook with 64-bit partial sums.
void product(const u32* x, const u32* y, int n, u32* p)
{
    std::fill(p, p + 2 * n, 0u);
    for (int i = 0; i < n; ++i) {
        u64 carry = 0;
        for (int j = 0; j < n; ++j) {
            const u64 t = u64(x[i]) * y[j] + p[i + j] + carry;
            p[i + j] = u32(t);
            carry = t >> 32;
        }
        p[i + n] = u32(carry);
    }
}

// acc (two's complement) += sign * x * 2^shift; acc is wide enough for the cluster's exact sum.
void accumulate(std::vector<u32>& acc, const u32* x, int words, u64 shift, int sign)
{
    const u64 w0 = shift / 32;
    const int s = int(shift % 32);
    u64 carry = sign > 0 ? 0 : 1; // a - v = a + ~v + 1
    for (u64 i = w0; i < acc.size(); ++i) {
        const i64 k = i64(i - w0);
        u32 v = k < words ? x[k] << s : 0;
        if (s && k >= 1 && k - 1 < words) v |= x[k - 1] >> (32 - s);
        if (sign < 0) v = ~v;
        else if (k > words && !carry) break;
        const u64 t = u64(acc[i]) + v + carry;
        acc[i] = u32(t);
        carry = t >> 32;
    }
}

struct Term {
    i64 scale;
    std::size_t k; // K = the addend
};

struct Sum {
    int sign;
    i64 scale;
    std::vector<u32> mag;
};
} // namespace

void exact_dot_host(int bits, const void* c, bool subtract, const void* a, std::ptrdiff_t sa, const void* b, std::ptrdiff_t sb,
                    std::size_t K, void* out_record)
{
    if (bits < 64 || bits > 1024 || bits % 32) throw std::invalid_argument("exact_dot_add: bits must be a multiple of 32 in [64,1024]");
    const int n = bits / 32;
    const std::size_t rec = std::size_t(n) + 3;
    u32* out = static_cast<u32*>(out_record);
    const u32* A = static_cast<const u32*>(a);
    const u32* B = static_cast<const u32*>(b);
    auto av = [&](std::size_t k) { return view(A + std::ptrdiff_t(k) * sa * std::ptrdiff_t(rec), n); };
    auto bv = [&](std::size_t k) { return view(B + std::ptrdiff_t(k) * sb * std::ptrdiff_t(rec), n); };
    u32 status = c ? view(c, n).status : 0;
    for (std::size_t k = 0; k < K; ++k) status |= av(k).status | bv(k).status;
    if (status) {
        write_zero(out, n, status);
        return;
    }
    std::vector<Term> t;
    t.reserve(K + 1);
    for (std::size_t k = 0; k < K; ++k) {
        const View x = av(k), y = bv(k);
        if (x.sign && y.sign) t.push_back({i64(x.exponent) + y.exponent - 2 * (32 * i64(n) - 1), k});
    }
    if (t.empty()) {
        if (c) std::memcpy(out, c, rec * 4);
        else write_zero(out, n, 0);
        return;
    }
    const View cv = c ? view(c, n) : View{nullptr, 0, 0, 0};
    if (cv.sign) t.push_back({i64(cv.exponent) - (32 * i64(n) - 1), K});
    std::vector<u32> p(2 * std::size_t(n));
    // exact term q into p: magnitude words and sign
    auto term = [&](const Term& q, int& words) -> int {
        if (q.k == K) {
            std::copy(cv.limb, cv.limb + n, p.begin());
            words = n;
            return cv.sign;
        }
        const View x = av(q.k), y = bv(q.k);
        product(x.limb, y.limb, n, p.data());
        words = 2 * n;
        return (subtract ? -1 : 1) * x.sign * y.sign;
    };
    std::sort(t.begin(), t.end(), [](const Term& u, const Term& v) { return u.scale > v.scale; });
    int logk = 0;
    while ((std::size_t(1) << logk) < K + 1) ++logk;
    const i64 gap = 32 * i64(n) + logk + 4, width = 64 * i64(n); // every term is below 2^(scale + 64n)
    std::vector<Sum> sums;
    for (std::size_t i = 0; i < t.size() && sums.size() < 2;) {
        std::size_t j = i + 1;
        i64 low = t[i].scale;
        while (j < t.size() && t[j].scale + width - 1 + gap >= low) {
            low = std::min(low, t[j].scale);
            ++j;
        }
        // the cluster's terms lie below 2^(t[i].scale + width); two spare words hold carries and the sign
        std::vector<u32> acc(std::size_t((t[i].scale + width - low) / 32 + 3), 0u);
        for (std::size_t q = i; q < j; ++q) {
            int words = 0;
            const int sign = term(t[q], words);
            accumulate(acc, p.data(), words, u64(t[q].scale - low), sign);
        }
        int sign = 1;
        if (acc.back() >> 31) { // negative: negate into a magnitude
            sign = -1;
            u64 carry = 1;
            for (auto& v : acc) {
                const u64 z = u64(u32(~v)) + carry;
                v = u32(z);
                carry = z >> 32;
            }
        }
        if (std::any_of(acc.begin(), acc.end(), [](u32 v) { return v != 0; })) sums.push_back({sign, low, std::move(acc)});
        i = j;
    }
    if (sums.empty()) {
        write_zero(out, n, 0);
        return;
    }
    const Sum& s = sums[0];
    if (sums.size() == 1) {
        round_to(s.mag, s.scale, s.sign, n, out);
        return;
    }
    // S * 2^D +/- 1: D places eps below the rounding position even when S has fewer than n words
    const int D = 32 * n + 64;
    std::vector<u32> ext(s.mag.size() + D / 32 + 1, 0u);
    std::copy(s.mag.begin(), s.mag.end(), ext.begin() + D / 32);
    if (sums[1].sign == s.sign) ext[0] = 1;
    else
        for (auto& v : ext) {
            const u32 old = v;
            v = old - 1;
            if (old) break;
        }
    round_to(ext, s.scale - D, s.sign, n, out);
}
} // namespace detail
} // namespace limbforge
