// Prints, for seeded vectors v of many sizes and magnitudes (zeros and a
// leading zero included), the unimodular U of pfa_unimodular (plain Euclid
// chain) and of pfa_unimodular_reduced from pfvs_amd/csrc/pfa_prep.h, built
// with the host compiler by tests/test_unimodular.py, which checks them
// exactly.  One line per case:
//   n | v | rc_plain | rc_reduced | U_reduced (n*n) | U_plain row 0 . v
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>

#include "../pfvs_amd/csrc/pfa_prep.h"

static void put128(pfa_i128 x)
{
    char buf[64];
    int n = 0;
    bool neg = x < 0;
    unsigned __int128 u = neg ? (unsigned __int128)(-x) : (unsigned __int128)x;
    do { buf[n++] = (char)('0' + (int)(u % 10)); u /= 10; } while (u);
    if (neg) std::putchar('-');
    while (n) std::putchar(buf[--n]);
}

int main()
{
    std::mt19937_64 rng(2024);
    auto uni = [&](int64_t lo, int64_t hi) { return lo + (int64_t)(rng() % (uint64_t)(hi - lo + 1)); };
    const int sizes[] = {2, 3, 5, 8, 13, 20, 25, 30, 35, 40, 41};
    PfaTeam T{0, 1};
    int flag = 0;
    for (int n : sizes)
        for (int rep = 0; rep < 12; ++rep) {
            const int regime = rep % 4;          // 0 small, 1 ~2^31, 2 ~2^40..2^52, 3 sparse large
            std::vector<int64_t> v(n);
            for (int i = 0; i < n; ++i) {
                int bits = regime == 0 ? 20 : regime == 1 ? 32 : (int)uni(40, 52);
                int64_t m = ((int64_t)1 << bits) - 1;
                v[i] = uni(-m, m);
                if (regime == 3 && uni(0, 2) == 0) v[i] = 0;
            }
            if (rep == 5) v[0] = 0;
            std::vector<int64_t> U(n * n), U0(n * n);
            std::vector<pfa_i128> x(n), r(n);
            std::vector<int> ord(n);
            int rc0 = pfa_unimodular(T, v.data(), n, U0.data(), &flag);
            int rc1 = pfa_unimodular_reduced(T, v.data(), n, U.data(), x.data(), r.data(), ord.data(), &flag);
            pfa_i128 g0 = 0;
            for (int c = 0; c < n; ++c) g0 += (pfa_i128)U0[c] * v[c];
            std::printf("%d |", n);
            for (int i = 0; i < n; ++i) std::printf(" %lld", (long long)v[i]);
            std::printf(" | %d | %d |", rc0, rc1);
            for (int e = 0; e < n * n; ++e) std::printf(" %lld", (long long)U[e]);
            std::printf(" | ");
            put128(g0);
            std::printf("\n");
        }
    return 0;
}
