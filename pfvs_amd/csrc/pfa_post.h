// pfa_post.h -- the exact per-point filter of coniZpM's post-processing
// (pfvs/coni.py:275-412), evaluated where the points are produced.
//
// For a lattice point c of p-vector p: M = Binter c, Kn = Z Binter c
// (int128, must fit int64: fp_kernel.pyx:591-609), q = c^T mat c.
//   K_gcd = gcd(Kn[1:]) (0 -> 1); the point is kept only if q < Q K_gcd
//   (coni.py:313-320).  Each g = 1 .. (Q K_gcd - 1) // q admits the K0 with
//   M0 K0 = g rawQperp - Q, rawQperp = (q + M0 Kn0) / K_gcd, capped by
//   K' > 0: K0 <= ceil(Kn0 g / K_gcd) - 1 (coni.py:327-365).  `n_exp`
//   counts those (g, K0); 0 means the K-integrality filter drops the point.
//   det((kappa M)[1:, 1:]) != 0 (coni.py:397-407): decided exactly by
//   determinants modulo 31-bit primes until their product exceeds the
//   Hadamard bound (then every residue 0 proves det = 0).
#pragma once
#include <math.h>
#include "pfa_int.h"

// ceil / floor of a / b for b > 0 (int128)
PFA_HD static inline pfa_i128 pfa_cdiv_pos(pfa_i128 a, pfa_i128 b)
{
    return -pfa_fdiv128(-a, b);
}

// Number of (g, K0) pairs of a point with M0 = M[0], Kn0 = Kn[0],
// K_gcd = gcd(Kn[1:]) (kg; 0 means 1), form value q and Q; -1 if some
// value leaves int128 (never on valid input).
PFA_HD static inline int64_t pfa_count_expansions(int64_t M0v, int64_t Kn0, uint64_t kg, int64_t q,
                                                 int64_t Q)
{
    if (kg == 0) kg = 1;
    if ((kg >> 62) != 0) return -1;
    pfa_i128 Kgcd = (pfa_i128)kg;
    pfa_i128 QK = (pfa_i128)Q * Kgcd;
    if (!((pfa_i128)q < QK)) return 0;
    if (q <= 0) return -1;                    // never: M0 >= linmin > 0 forces q > 0
    pfa_i128 M0 = M0v;
    if (M0 <= 0) return -1;
    pfa_i128 num = (pfa_i128)q + M0 * (pfa_i128)Kn0;
    if (num % Kgcd != 0) return -1;          // exact by construction
    pfa_i128 raw = num / Kgcd;
    pfa_i128 gmax = (QK - 1) / q;
    if (gmax > 1000000) return -1;
    int64_t cnt = 0;
    for (pfa_i128 g = 1; g <= gmax; ++g) {
        pfa_i128 Qp = g * raw;
        pfa_i128 lo = pfa_cdiv_pos(Qp - Q, M0);
        pfa_i128 up = pfa_fdiv128(Qp - Q, M0);
        pfa_i128 cap = pfa_cdiv_pos((pfa_i128)Kn0 * g, Kgcd) - 1;
        if (cap < up) up = cap;
        if (up >= lo) cnt += (int64_t)(up - lo + 1);
    }
    return cnt;
}

// floor(a / b) for b > 0 and |a|, b < 2^53: the double quotient is within 1
// of a / b, and one remainder test corrects it.
PFA_HD static inline int64_t pfa_fdiv_small(int64_t a, int64_t b)
{
    int64_t q = (int64_t)floor((double)a / (double)b);
    const int64_t r = a - q * b;
    if (r < 0) q--;
    else if (r >= b) q++;
    return q;
}

// pfa_count_expansions in int64 when every value stays below 2^52 (the
// common case); otherwise the int128 form.  Same result.
PFA_HD static inline int64_t pfa_count_expansions_fast(int64_t M0, int64_t Kn0, uint64_t kg,
                                                      int64_t q, int64_t Q)
{
    const int64_t B31 = (int64_t)1 << 31, B52 = (int64_t)1 << 52;
    const int64_t k = kg == 0 ? 1 : (int64_t)(kg < (uint64_t)B31 ? kg : (uint64_t)B31);
    if (!(M0 > 0 && M0 < B31 && q > 0 && q < B31 && Q > 0 && Q < B31 && Kn0 > -B31 && Kn0 < B31 &&
          k < B31))
        return pfa_count_expansions(M0, Kn0, kg, q, Q);
    const int64_t QK = Q * k;                       // < 2^62
    if (!(q < QK)) return 0;
    if (QK >= B52) return pfa_count_expansions(M0, Kn0, kg, q, Q);
    const int64_t num = q + M0 * Kn0;               // |.| < 2^63
    if (k != 1 && num % k != 0) return -1;          // exact by construction
    const int64_t raw = k == 1 ? num : num / k;
    const int64_t gmax = pfa_fdiv_small(QK - 1, q);
    if (gmax > 1000000) return -1;
    const double gd = (double)gmax;
    if (!((raw < 0 ? -(double)raw : (double)raw) * gd < 4.0e15) ||
        !((Kn0 < 0 ? -(double)Kn0 : (double)Kn0) * gd < 4.0e15))
        return pfa_count_expansions(M0, Kn0, kg, q, Q);
    int64_t cnt = 0;
    for (int64_t g = 1; g <= gmax; ++g) {
        const int64_t a = g * raw - Q;
        const int64_t lo = -pfa_fdiv_small(-a, M0);
        int64_t up = pfa_fdiv_small(a, M0);
        const int64_t cap = -pfa_fdiv_small(-(Kn0 * g), k) - 1;
        if (cap < up) up = cap;
        if (up >= lo) cnt += up - lo + 1;
    }
    return cnt;
}

// sum_t row[t] c[t] for a row bounded by |row[t]| <= rmax and |c| <= cmax:
// 32-bit products when rmax < 2^31, 64-bit when the bound keeps every
// partial sum below 2^62, else int128 with *ok = 0 if the sum leaves int64.
PFA_HD static inline int64_t pfa_dot_c(const int64_t *row, const int32_t *c, int d, int64_t rmax,
                                       int64_t cmax, int *ok)
{
    const double bound = (double)rmax * (double)cmax * (double)d;
    if (bound < 4.0e18) {
        int64_t s = 0;
        if (rmax < ((int64_t)1 << 31)) {
            for (int t = 0; t < d; ++t) s += (int64_t)(int32_t)row[t] * (int64_t)c[t];
        } else {
            for (int t = 0; t < d; ++t) s += row[t] * (int64_t)c[t];
        }
        return s;
    }
    pfa_i128 s = 0;
    for (int t = 0; t < d; ++t) s += (pfa_i128)row[t] * c[t];
    if (!pfa_fits64(s)) *ok = 0;
    return (int64_t)s;
}

// The same for rows M, Kn of length h.
PFA_HD static inline int64_t pfa_num_expansions(const int64_t *M, const int64_t *Kn, int h,
                                               int64_t q, int64_t Q)
{
    pfa_u128 kg = 0;
    for (int i = 1; i < h; ++i) kg = pfa_gcd128(kg, pfa_abs128((pfa_i128)Kn[i]));
    if ((kg >> 62) != 0) return -1;
    return pfa_count_expansions(M[0], Kn[0], (uint64_t)kg, q, Q);
}

// Arithmetic modulo an odd prime p < 2^31: Barrett reduction with
// m = floor(2^64 / p), and Montgomery products with pinv = -p^-1 mod 2^32.
struct PfaModP {
    uint32_t p;
    uint32_t pinv;
    uint64_t m;
};

// -p^-1 mod 2^32 for odd p (Newton iteration on the 2-adic inverse)
PFA_HD static inline uint32_t pfa_mont_pinv(uint32_t p)
{
    uint32_t x = p;                        // correct to 3 bits: p * p = 1 mod 8
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;
    return 0u - x;
}

// a b 2^-32 mod p for a, b < p (result < p)
PFA_HD static inline uint32_t pfa_mont(uint32_t a, uint32_t b, uint32_t p, uint32_t pinv)
{
    const uint64_t t = (uint64_t)a * b;                    // < 2^62
    const uint32_t m = (uint32_t)t * pinv;
    const uint64_t u = (t + (uint64_t)m * p) >> 32;       // < 2 p
    return (uint32_t)(u >= p ? u - p : u);
}

PFA_HD static inline uint64_t pfa_umulhi64(uint64_t a, uint64_t b)
{
    return (uint64_t)(((pfa_u128)a * b) >> 64);
}

// x mod p for any x < 2^64 (q is at most one short of floor(x / p))
PFA_HD static inline uint32_t pfa_reduce(uint64_t x, const PfaModP &P)
{
    uint64_t q = pfa_umulhi64(x, P.m);
    uint64_t r = x - q * P.p;
    while (r >= P.p) r -= P.p;                 // at most twice
    return (uint32_t)r;
}

PFA_HD static inline uint32_t pfa_mulmod(uint32_t a, uint32_t b, const PfaModP &P)
{
    return pfa_reduce((uint64_t)a * b, P);
}

// inverse of a (0 < a < p) by the extended Euclidean algorithm
PFA_HD static inline uint32_t pfa_invmod(uint32_t a, uint32_t p)
{
    int64_t t = 0, nt = 1, r = p, nr = a;
    while (nr) {
        int64_t q = r / nr, x;
        x = t - q * nt; t = nt; nt = x;
        x = r - q * nr; r = nr; nr = x;
    }
    return (uint32_t)(t < 0 ? t + p : t);
}

PFA_HD static inline uint32_t pfa_mod(int64_t x, uint32_t p)
{
    int64_t r = x % (int64_t)p;
    return (uint32_t)(r < 0 ? r + p : r);
}

// x mod p in [0, p) without a 64-bit division
PFA_HD static inline uint32_t pfa_mod_fast(int64_t x, const PfaModP &P)
{
    const uint64_t u = x < 0 ? (uint64_t)(-(x + 1)) + 1u : (uint64_t)x;     // |x|
    const uint32_t r = u < P.p ? (uint32_t)u : pfa_reduce(u, P);
    return (x < 0 && r) ? P.p - r : r;
}

// det(A) mod p for the n x n int64 matrix A; element (i, j) of A and of the
// work array w lives at index (i * n + j) * stride (stride > 1 interleaves
// the matrices of many lanes so their accesses coalesce).
PFA_HD static inline uint32_t pfa_det_mod(const int64_t *A, int n, const PfaModP &P, uint32_t *w,
                                          int64_t stride)
{
    const uint32_t p = P.p;
    for (int e = 0; e < n * n; ++e) w[e * stride] = pfa_mod(A[e * stride], p);
    uint32_t det = 1;
    for (int k = 0; k < n; ++k) {
        int piv = -1;
        for (int i = k; i < n; ++i) if (w[(i * n + k) * stride]) { piv = i; break; }
        if (piv < 0) return 0;
        if (piv != k) {
            for (int j = k; j < n; ++j) {
                uint32_t t = w[(k * n + j) * stride];
                w[(k * n + j) * stride] = w[(piv * n + j) * stride];
                w[(piv * n + j) * stride] = t;
            }
            det = det ? p - det : 0;
        }
        uint32_t pv = w[(k * n + k) * stride];
        det = pfa_mulmod(det, pv, P);
        uint32_t inv = pfa_invmod(pv, p);
        for (int i = k + 1; i < n; ++i) {
            uint32_t f = w[(i * n + k) * stride];
            if (!f) continue;
            f = pfa_mulmod(f, inv, P);
            for (int j = k + 1; j < n; ++j) {
                uint32_t sub = pfa_mulmod(f, w[(k * n + j) * stride], P);
                uint32_t x = w[(i * n + j) * stride];
                w[(i * n + j) * stride] = x >= sub ? x - sub : x + p - sub;
            }
        }
    }
    return det;
}

// log2 of the Hadamard bound prod_i |row_i|_2 of the n x n matrix, rounded
// up generously (each row norm computed in double with 1e-9 bits of slack).
PFA_HD static inline double pfa_hadamard_log2(const int64_t *A, int n, int64_t stride)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        double r = 0.0;
        for (int j = 0; j < n; ++j) {
            double x = (double)A[(i * n + j) * stride];
            r += x * x;
        }
        if (r == 0.0) return -1e300;            // a zero row: det = 0
        s += 0.5 * log2(r) + 1e-9;
    }
    return s + 1.0;
}

// Exact fraction-free (Bareiss) elimination in checked int128 (same
// element layout as pfa_det_mod): 0 singular, 1 nonsingular, 2 an
// intermediate left int128.  On rank-deficient matrices it stops as soon as
// a column has no pivot, after r eliminations of minors of size <= r + 1.
PFA_HD static inline int pfa_bareiss_decide(const int64_t *A, int n, pfa_i128 *w, int64_t stride)
{
    for (int e = 0; e < n * n; ++e) w[e * stride] = A[e * stride];
    pfa_i128 prev = 1;
    for (int k = 0; k < n; ++k) {
        int piv = -1;
        for (int i = k; i < n; ++i) if (w[(i * n + k) * stride] != 0) { piv = i; break; }
        if (piv < 0) return 0;
        if (piv != k)
            for (int j = k; j < n; ++j) {
                pfa_i128 t = w[(k * n + j) * stride];
                w[(k * n + j) * stride] = w[(piv * n + j) * stride];
                w[(piv * n + j) * stride] = t;
            }
        const pfa_i128 pk = w[(k * n + k) * stride];
        for (int i = k + 1; i < n; ++i) {
            const pfa_i128 ik = w[(i * n + k) * stride];
            for (int j = k + 1; j < n; ++j) {
                pfa_i128 a, b, x;
                if (pfa_mul_ovf(w[(i * n + j) * stride], pk, &a) ||
                    pfa_mul_ovf(ik, w[(k * n + j) * stride], &b) || pfa_sub_ovf(a, b, &x))
                    return 2;
                w[(i * n + j) * stride] = x / prev;        // exact (Sylvester's identity)
            }
        }
        prev = pk;
    }
    return 1;
}

// 1 nonsingular, 0 singular, 2 undecided.  One modular determinant settles
// almost every nonsingular matrix; exact Bareiss settles low-rank ones;
// otherwise determinants modulo further primes until their product
// exceeds the Hadamard bound.
PFA_HD static inline int pfa_nonsingular(const int64_t *A, int n, const PfaModP *primes, int np,
                                         uint32_t *w, pfa_i128 *bw, int64_t stride)
{
    double need = pfa_hadamard_log2(A, n, stride);
    if (need < -1e299) return 0;
    if (np < 1) return 2;
    if (pfa_det_mod(A, n, primes[0], w, stride) != 0) return 1;
    int r = pfa_bareiss_decide(A, n, bw, stride);
    if (r != 2) return r;
    double have = log2((double)primes[0].p) - 1e-9;
    if (have > need) return 0;
    for (int k = 1; k < np; ++k) {
        if (pfa_det_mod(A, n, primes[k], w, stride) != 0) return 1;
        have += log2((double)primes[k].p) - 1e-9;
        if (have > need) return 0;
    }
    return 2;
}
