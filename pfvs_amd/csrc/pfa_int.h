// pfa_int.h -- exact integer building blocks shared by the host and the
// gfx950 device code (checked 128-bit arithmetic, binary gcd).
//
// The leaf decisions of the coniZpM search are exact; these helpers carry
// the checks that the reference performs in pfvs/fp_kernel/fpk_common.h:38-178
// (checked int128 add/mul, Stein's gcd with an early exit below a required
// value).  Written for AMDGPU first: 64-bit ctz via __builtin_ctzll lowers to
// s_ff1/v_ffbl, and every loop is bounded.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PFA_HD __host__ __device__
#define PFA_DEV __device__
#else
#define PFA_HD
#define PFA_DEV
#endif

typedef __int128 pfa_i128;
typedef unsigned __int128 pfa_u128;

#define PFA_I64MAX ((pfa_i128)INT64_MAX)

// |x| <= INT64_MAX (INT64_MIN excluded, so negation never overflows)
PFA_HD static inline int pfa_fits64(pfa_i128 x) { return x <= PFA_I64MAX && x >= -PFA_I64MAX; }
PFA_HD static inline pfa_u128 pfa_abs128(pfa_i128 x) { return x < 0 ? (pfa_u128)0 - (pfa_u128)x : (pfa_u128)x; }

// r = a + b, 1 on signed overflow (two's complement sign test)
PFA_HD static inline int pfa_add_ovf(pfa_i128 a, pfa_i128 b, pfa_i128 *r)
{
    pfa_i128 s = (pfa_i128)((pfa_u128)a + (pfa_u128)b);
    *r = s;
    return ((a >= 0) == (b >= 0)) && ((s >= 0) != (a >= 0));
}

PFA_HD static inline int pfa_sub_ovf(pfa_i128 a, pfa_i128 b, pfa_i128 *r)
{
    pfa_i128 s = (pfa_i128)((pfa_u128)a - (pfa_u128)b);
    *r = s;
    return ((a >= 0) != (b >= 0)) && ((s >= 0) != (a >= 0));
}

// r = a * b, 1 if |a b| >= 2^127.  Magnitudes below 2^63 multiply directly;
// otherwise the 128x128 product is formed from 64-bit halves and checked.
PFA_HD static inline int pfa_mul_ovf(pfa_i128 a, pfa_i128 b, pfa_i128 *r)
{
    pfa_u128 ua = pfa_abs128(a), ub = pfa_abs128(b);
    const pfa_u128 lim = (pfa_u128)1 << 63;
    if (ua < lim && ub < lim) { *r = a * b; return 0; }
    uint64_t ah = (uint64_t)(ua >> 64), bh = (uint64_t)(ub >> 64);
    if (ah && bh) return 1;
    // one operand < 2^64: (hi * 2^64 + lo) * small
    pfa_u128 big = ah ? ua : ub, sm = ah ? ub : ua;
    if ((sm >> 64) != 0) return 1;
    pfa_u128 lo = (pfa_u128)(uint64_t)big * (uint64_t)sm;
    pfa_u128 hi = (pfa_u128)(uint64_t)(big >> 64) * (uint64_t)sm;
    if ((hi >> 63) != 0) return 1;
    pfa_u128 m = lo + (hi << 64);
    if (m < lo || (m >> 127) != 0) return 1;
    *r = ((a < 0) != (b < 0)) ? -(pfa_i128)m : (pfa_i128)m;
    return 0;
}

// r += a * b (int64 factors), checked
PFA_HD static inline int pfa_mac64(pfa_i128 *r, int64_t a, int64_t b)
{
    return pfa_add_ovf(*r, (pfa_i128)a * (pfa_i128)b, r);
}

PFA_HD static inline int pfa_ctz64(uint64_t x) { return __builtin_ctzll(x); }

PFA_HD static inline int pfa_ctz128(pfa_u128 x)
{
    uint64_t lo = (uint64_t)x;
    return lo ? pfa_ctz64(lo) : 64 + pfa_ctz64((uint64_t)(x >> 64));
}

// Stein's binary gcd on 32 bits
PFA_HD static inline uint32_t pfa_gcd32(uint32_t u, uint32_t v)
{
    if (u == 0) return v;
    if (v == 0) return u;
    const int sh = __builtin_ctz(u | v);
    u >>= __builtin_ctz(u);
    do {
        v >>= __builtin_ctz(v);
        if (u > v) { uint32_t t = v; v = u; u = t; }
        v -= u;
    } while (v);
    return u << sh;
}

// Stein's binary gcd (the 32-bit loop when both operands fit)
PFA_HD static inline uint64_t pfa_gcd64(uint64_t u, uint64_t v)
{
    if (((u | v) >> 32) == 0) return pfa_gcd32((uint32_t)u, (uint32_t)v);
    if (u == 0) return v;
    if (v == 0) return u;
    int sh = pfa_ctz64(u | v);
    u >>= pfa_ctz64(u);
    do {
        v >>= pfa_ctz64(v);
        if (u > v) { uint64_t t = v; v = u; u = t; }
        v -= u;
    } while (v);
    return u << sh;
}

// gcd(u, v) if it is >= need, else 0 (u, v not both 0).  Once the odd part
// of the smaller operand drops below ceil(need / 2^shift) the gcd cannot
// reach need (the same early exit as fpk_gcd64_ge, fpk_common.h:113-128).
template <bool CNT>
PFA_HD static inline uint64_t pfa_gcd64_ge_t(uint64_t u, uint64_t v, uint64_t need,
                                             unsigned long long *iters)
{
    if (u == 0) return v >= need ? v : 0;
    if (v == 0) return u >= need ? u : 0;
    int sh = pfa_ctz64(u | v);
    uint64_t ns = (need >> sh) + ((need & ((1ULL << sh) - 1)) != 0);
    u >>= pfa_ctz64(u);
    do {
        if (CNT) ++*iters;
        v >>= pfa_ctz64(v);
        if (u > v) { uint64_t t = v; v = u; u = t; }
        if (u < ns) return 0;
        v -= u;
    } while (v);
    return u << sh;
}

PFA_HD static inline uint64_t pfa_gcd64_ge(uint64_t u, uint64_t v, uint64_t need)
{
    return pfa_gcd64_ge_t<false>(u, v, need, nullptr);
}

PFA_HD static inline pfa_u128 pfa_gcd128(pfa_u128 u, pfa_u128 v)
{
    if ((u >> 64) == 0 && (v >> 64) == 0) return pfa_gcd64((uint64_t)u, (uint64_t)v);
    if (u == 0) return v;
    if (v == 0) return u;
    int sh = pfa_ctz128(u | v);
    u >>= pfa_ctz128(u);
    do {
        v >>= pfa_ctz128(v);
        if (u > v) { pfa_u128 t = v; v = u; u = t; }
        v -= u;
        if ((u >> 64) == 0 && (v >> 64) == 0)
            return (pfa_u128)pfa_gcd64((uint64_t)u, (uint64_t)v) << sh;
    } while (v);
    return u << sh;
}

// floor division on int64 (as Python's //)
PFA_HD static inline int64_t pfa_fdiv64(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}

// a / b and a % b on int128 (C semantics: the quotient truncated), through
// the int64 operations when both operands fit (the GPU has no integer
// divider: an int128 division is a long shift-subtract loop, an int64 one a
// short inline sequence); INT64_MIN / -1 stays on the int128 path
PFA_HD static inline bool pfa_div_in64(pfa_i128 a, pfa_i128 b)
{
    const pfa_i128 lo = -((pfa_i128)1 << 63), hi = ((pfa_i128)1 << 63) - 1;
    return a >= lo && a <= hi && b >= lo && b <= hi && !(a == lo && b == -1);
}

PFA_HD static inline pfa_i128 pfa_tdiv128(pfa_i128 a, pfa_i128 b)
{
    if (pfa_div_in64(a, b)) return (pfa_i128)((int64_t)a / (int64_t)b);
    return a / b;
}

// 256-bit two's complement integers (products of two int128), for exact
// determinants whose entries fit int128 but whose products do not.
struct pfa_w256 {
    pfa_u128 lo, hi;
};

PFA_HD static inline pfa_w256 pfa_umul256(pfa_u128 a, pfa_u128 b)
{
    const uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64), b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
    const pfa_u128 p00 = (pfa_u128)a0 * b0, p01 = (pfa_u128)a0 * b1, p10 = (pfa_u128)a1 * b0,
                   p11 = (pfa_u128)a1 * b1;
    const pfa_u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;      // < 3 2^64
    pfa_w256 r;
    r.lo = (mid << 64) | (uint64_t)p00;
    r.hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
    return r;
}

PFA_HD static inline pfa_w256 pfa_neg256(pfa_w256 x)
{
    pfa_w256 r;
    r.lo = ~x.lo + 1;
    r.hi = ~x.hi + (r.lo == 0 ? 1 : 0);
    return r;
}

PFA_HD static inline pfa_w256 pfa_smul256(pfa_i128 a, pfa_i128 b)
{
    const pfa_w256 r = pfa_umul256(pfa_abs128(a), pfa_abs128(b));
    return (a < 0) != (b < 0) ? pfa_neg256(r) : r;
}

PFA_HD static inline pfa_w256 pfa_sub256(pfa_w256 x, pfa_w256 y)
{
    pfa_w256 r;
    r.lo = x.lo - y.lo;
    r.hi = x.hi - y.hi - (x.lo < y.lo ? 1 : 0);
    return r;
}

// x >> s, arithmetic, 0 <= s < 256
PFA_HD static inline pfa_w256 pfa_sar256(pfa_w256 x, int s)
{
    if (s == 0) return x;
    pfa_w256 r;
    if (s < 128) {
        r.lo = (x.lo >> s) | (x.hi << (128 - s));
        r.hi = (pfa_u128)((pfa_i128)x.hi >> s);
    } else {
        r.lo = (pfa_u128)((pfa_i128)x.hi >> (s - 128));
        r.hi = (pfa_i128)x.hi < 0 ? ~(pfa_u128)0 : (pfa_u128)0;
    }
    return r;
}

// x < y, both nonnegative
PFA_HD static inline int pfa_ult256(pfa_w256 x, pfa_w256 y)
{
    return x.hi < y.hi || (x.hi == y.hi && x.lo < y.lo);
}

// A divisor b != 0 prepared for exact divisions: b = 2^s o with o odd, and
// o's inverse modulo 2^128 (Newton's iteration x <- x (2 - o x) doubles
// the correct low bits, from 3 since o o = 1 mod 8).
struct pfa_exdiv {
    pfa_i128 b;
    pfa_u128 inv;
    int s;
};

PFA_HD static inline pfa_exdiv pfa_exdiv_prep(pfa_i128 b)
{
    pfa_exdiv D;
    D.b = b;
    D.s = pfa_ctz128((pfa_u128)b);
    const pfa_u128 o = (pfa_u128)(b >> D.s);
    pfa_u128 x = o;
    for (int k = 0; k < 6; ++k) x *= (pfa_u128)2 - o * x;
    D.inv = x;
    return D;
}

// The int128 quotient N / b when b divides N and the quotient fits int128
// (returns 1), else 0: (N >> s) times o's inverse modulo 2^128 is the
// quotient modulo 2^128, and q b == N proves it (if the true quotient does
// not fit int128, no int128 q satisfies q b == N).
PFA_HD static inline int pfa_exdiv256(pfa_w256 N, const pfa_exdiv &D, pfa_i128 *q)
{
    const pfa_w256 t = pfa_sar256(N, D.s);
    const pfa_i128 r = (pfa_i128)(t.lo * D.inv);
    const pfa_w256 back = pfa_smul256(r, D.b);
    *q = r;
    return back.lo == N.lo && back.hi == N.hi;
}

PFA_HD static inline pfa_i128 pfa_tmod128(pfa_i128 a, pfa_i128 b)
{
    if (pfa_div_in64(a, b)) return (pfa_i128)((int64_t)a % (int64_t)b);
    return a % b;
}

PFA_HD static inline pfa_i128 pfa_fdiv128(pfa_i128 a, pfa_i128 b)
{
    pfa_i128 q = pfa_tdiv128(a, b);
    if ((pfa_tmod128(a, b) != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}
