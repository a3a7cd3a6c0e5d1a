// Randomised equality checks of the post-processing arithmetic in
// pfvs_amd/csrc/pfa_post.h, built with the host compiler by
// tests/test_post_math.py: each fast form against its reference form.
// Prints "ok <cases>" or the first mismatch, exit status 0 / 1.
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cmath>
#include <random>

#include "../pfvs_amd/csrc/pfa_post.h"

static int fail(const char *what, long long a, long long b, long long c, long long d)
{
    std::printf("MISMATCH %s %lld %lld %lld %lld\n", what, a, b, c, d);
    return 1;
}

int main()
{
    std::mt19937_64 rng(12345);
    auto uni = [&](int64_t lo, int64_t hi) {
        return lo + (int64_t)(rng() % (uint64_t)(hi - lo + 1));
    };
    long cases = 0;
    // floor division against the int128 form
    for (int it = 0; it < 100000; ++it) {
        const int sh = (int)uni(1, 52);
        int64_t a = uni(-((int64_t)1 << sh) + 1, ((int64_t)1 << sh) - 1);
        int64_t b = uni(1, ((int64_t)1 << uni(1, 52)) - 1);
        if (pfa_fdiv_small(a, b) != (int64_t)pfa_fdiv128(a, b)) return fail("fdiv", a, b, 0, 0);
        ++cases;
    }
    // expansion counts: typical magnitudes, boundaries, and values that
    // force the int128 form
    for (int it = 0; it < 40000; ++it) {
        const int regime = it % 4;
        int64_t Q = uni(20, 500), M0 = uni(1, 3000), q = uni(1, 20000), Kn0 = uni(-5000, 5000);
        uint64_t kg = (uint64_t)uni(0, 12);
        if (regime == 1) { kg = (uint64_t)uni(0, 3); q = uni(1, 40); Q = uni(20, 100); }   // many g
        if (regime == 2) { M0 = uni(1, ((int64_t)1 << 31) - 1); Kn0 = uni(-((int64_t)1 << 31) + 1, ((int64_t)1 << 31) - 1); }
        if (regime == 3) {                  // large K_gcd and Q, few g (gmax <= 64)
            kg = (uint64_t)uni(1, ((int64_t)1 << 40));
            Q = uni(1, ((int64_t)1 << 20));
            const int64_t QK = (int64_t)((pfa_i128)Q * (pfa_i128)kg);
            q = uni(QK / 64 + 1, QK);
        }
        // n_exp is only asked for points whose q is consistent with the
        // divisibility (q + M0 Kn0 divisible by K_gcd); mimic that half the time
        if ((it & 1) && regime != 3) {
            const int64_t k = kg == 0 ? 1 : (int64_t)kg;
            const int64_t r = (int64_t)(((pfa_i128)q + (pfa_i128)M0 * Kn0) % k);
            q -= r;
            if (q <= 0) q += k;
        }
        const int64_t f = pfa_count_expansions_fast(M0, Kn0, kg, q, Q);
        const int64_t s = pfa_count_expansions(M0, Kn0, kg, q, Q);
        if (f != s) return fail("expansions", M0, Kn0, (long long)kg, q);
        ++cases;
    }
    // bounded dot products against int128
    for (int it = 0; it < 50000; ++it) {
        const int d = (int)uni(1, 40);
        int64_t row[40];
        int32_t c[40];
        const int rb = (int)uni(1, 62), cb = (int)uni(1, 30);
        int64_t rmax = 0, cmax = 0;
        for (int t = 0; t < d; ++t) {
            row[t] = uni(-((int64_t)1 << rb) + 1, ((int64_t)1 << rb) - 1);
            c[t] = (int32_t)uni(-((int64_t)1 << cb) + 1, ((int64_t)1 << cb) - 1);
            rmax = std::max<int64_t>(rmax, row[t] < 0 ? -row[t] : row[t]);
            cmax = std::max<int64_t>(cmax, c[t] < 0 ? -(int64_t)c[t] : c[t]);
        }
        pfa_i128 ref = 0;
        for (int t = 0; t < d; ++t) ref += (pfa_i128)row[t] * c[t];
        int ok = 1;
        const int64_t v = pfa_dot_c(row, c, d, rmax, cmax, &ok);
        if (ok != pfa_fits64(ref) || (ok && v != (int64_t)ref)) return fail("dot", d, rb, cb, ok);
        ++cases;
    }
    // Montgomery products and signed reduction against plain modular arithmetic
    const uint32_t primes[3] = {2147483647u, 2147483629u, 1000000007u};
    for (uint32_t p : primes) {
        const PfaModP P{p, pfa_mont_pinv(p), (uint64_t)(((pfa_u128)1 << 64) / p)};
        const uint64_t Rinv_num = ((pfa_u128)1 << 32) % p;           // 2^32 mod p
        for (int it = 0; it < 50000; ++it) {
            const uint32_t a = (uint32_t)uni(0, p - 1), b = (uint32_t)uni(0, p - 1);
            const uint32_t m = pfa_mont(a, b, p, P.pinv);
            // m * 2^32 == a b (mod p)
            if ((uint64_t)(((pfa_u128)m * Rinv_num) % p) != (uint64_t)(((pfa_u128)a * b) % p))
                return fail("mont", a, b, p, m);
            const int64_t x = (int64_t)rng();
            int64_t r = (int64_t)(((pfa_i128)x) % (pfa_i128)p);
            if (r < 0) r += p;
            if ((int64_t)pfa_mod_fast(x, P) != r) return fail("mod", x, p, 0, 0);
            cases += 2;
        }
    }
    // int128 division through int64 when both operands fit (pfa_tdiv128,
    // pfa_tmod128, pfa_fdiv128) against the int128 operators: random
    // magnitudes up to 2^100, both signs, and the boundary values
    {
        auto big = [&]() {
            const int bits = (int)uni(0, 100);
            pfa_i128 v = (pfa_i128)(rng() >> 1);
            if (bits > 63) v = (v << (bits - 63)) ^ (pfa_i128)rng();
            else v >>= (63 - bits);
            return (rng() & 1) ? -v : v;
        };
        const pfa_i128 m63 = -((pfa_i128)1 << 63), p63 = ((pfa_i128)1 << 63) - 1;
        const pfa_i128 edge[] = {0, 1, -1, 2, -2, 7, -7, m63, m63 + 1, p63, p63 - 1, p63 + 1, m63 - 1,
                                 (pfa_i128)1 << 64, -((pfa_i128)1 << 64)};
        auto check = [&](pfa_i128 a, pfa_i128 b) -> bool {
            if (b == 0) return true;
            const pfa_i128 q = a / b, r = a % b;
            pfa_i128 f = q;
            if (r != 0 && ((a < 0) != (b < 0))) f--;
            ++cases;
            return pfa_tdiv128(a, b) == q && pfa_tmod128(a, b) == r && pfa_fdiv128(a, b) == f;
        };
        for (pfa_i128 a : edge)
            for (pfa_i128 b : edge)
                if (!check(a, b)) return fail("div128 edge", (long long)a, (long long)(a >> 64), (long long)b, (long long)(b >> 64));
        for (int it = 0; it < 200000; ++it) {
            const pfa_i128 a = big(), b = big();
            if (!check(a, b)) return fail("div128", (long long)a, (long long)(a >> 64), (long long)b, (long long)(b >> 64));
        }
    }
    std::printf("ok %ld\n", cases);
    return 0;
}
