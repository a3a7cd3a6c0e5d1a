// pfa_prep.h -- per-p preparation of the coniZpM search, written once for a
// "team" of lanes: on the device a team is one wavefront (64 lanes, one
// workgroup) with the working matrices in LDS; on the host a team is one
// thread.  Scalar control flow is evaluated identically by every lane from
// LDS data; loops over columns / matrix entries are split across lanes, and
// every reduction that feeds a floating-point decision is evaluated by each
// lane in the same order as the host, so both produce the same basis.
//
// Steps (pfvs/fp_kernel/pfv_lattice.h:604-680, restated):
//   Z = kappa.p; v = (Z p)^T Mbasis; O = basis of v^perp (Bezout
//   elimination, 151-188); LLL(O) with an exact int128 Gram matrix
//   (393-521); Binter = Mbasis O^T with the Binter[0] == 0 columns first;
//   ZB = Z Binter; mat = -Binter^T ZB; cut-aware basis T = [K | w]
//   (554-602) so that Binter[0] = (0, ..., 0, g); H = row HNF of ZB[1:]
//   (250-313).  Every intermediate is checked: overflow returns -1 (the
//   problem then goes to the exact host path).
// Then the FP64 Cholesky factor with its proven backward-error margin phi
// (replacing the reference's exact Bareiss factor, fpk_common.h:232-309).
#pragma once
#include <math.h>
#include <string.h>
#include "pfa_int.h"
#include "pfa_search.h"

#if defined(__HIP_DEVICE_COMPILE__)
#define PFA_SYNC() __syncthreads()
#else
#define PFA_SYNC() ((void)0)
#endif

struct PfaTeam {
    int lane, n;
};

// Algorithmic work of the preparation, counted on the host path only (the
// device build compiles it away): one lane-op per arithmetic operation on
// one element (a product, sum, quotient or rounding of any width; an exact
// multiply-add counts 2), in the steps' textbook counts.  Moves, swaps and
// overflow checks are not counted.
#if defined(__HIP_DEVICE_COMPILE__)
#define PFA_OPS(x) ((void)0)
#define PFA_OPS_NOW() 0.0
#else
inline thread_local double pfa_prep_ops = 0.0;
#define PFA_OPS(x) (pfa_prep_ops += (double)(x))
#define PFA_OPS_NOW() pfa_prep_ops
#endif

// Working storage of one preparation.  The part the LLL and HNF loops
// touch lives in LDS on the device ("hot"), sized so that 16 wavefronts
// of h11 <= 20 share a CU's LDS (10 KB each):
//   GG  the exact Gram matrix (int128) of the rows being reduced, lower
//       triangle ((h+1)(h+2)/2 entries, GG[pfa_gi(i, j)]);
//   W1  h*h int64: the basis rows being reduced;
//       H, the HNF (d*d int128), spans GG and W1 once the last LLL is done;
//   mu  the Gram-Schmidt coefficients, strict lower triangle (mu[pfa_mi(k,
//       j)], j < k); outside the LLL the same region holds the scratch hv,
//       hw (h int128), hpiv (h int) of the HNF and the unimodular
//       completion, and lll_form's Gy (h int128) before its LLL starts;
//   bb, mw, vec (h), flag;
//   sp  (optional) per Gram-Schmidt row kk, the running value of bb[kk]
//       before each of its kk terms (strict lower triangle, as mu): a row
//       exchange then moves a row's values instead of recomputing them.
//       Without it (a launch that needs the LDS for occupancy) the row is
//       recomputed; both give the same bits.
// The part only O(h^3) products touch ("cold") is per-workgroup global
// memory on the device: Z, W2, W3 (h*h int64), Binter, ZB (h*(h-1)), mat
// ((h-1)^2).  On the host both parts are one buffer.
struct PfaWS {
    int h, d;
    int64_t *Z, *Bi, *ZB, *mat, *W1, *W2, *W3, *vec;
    pfa_i128 *GG, *hv, *hw;
    double *mu, *bb, *mw, *sp;
    int *flag, *hpiv;
};

#define PFA_WS_ROUND(x) (((size_t)(x) + 15) & ~(size_t)15)

#if defined(__HIP_DEVICE_COMPILE__)
// lane l's double on every lane, l wave-uniform (two readlanes: no LDS
// round trip, the same bits as a shuffle from lane l)
__device__ static inline double pfa_readlane_f64(double x, int l)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
#endif

// index of (i, j) in a symmetric matrix stored as its lower triangle
PFA_HD static inline int pfa_gi(int i, int j)
{
    return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// index of (k, j), j < k, in a strict lower triangle
PFA_HD static inline int pfa_mi(int k, int j) { return k * (k - 1) / 2 + j; }

PFA_HD static inline size_t pfa_ws_gg_bytes(int h) { return PFA_WS_ROUND((size_t)(h + 1) * (h + 2) / 2 * 16); }

// the mu region: the strict triangle, or the HNF scratch, or Gy
PFA_HD static inline size_t pfa_ws_mu_bytes(int h)
{
    size_t w = (size_t)h * (h - 1) / 2 * 8;
    const size_t scr = PFA_WS_ROUND((size_t)h * 16) * 2 + PFA_WS_ROUND((size_t)h * 4);
    if (scr > w) w = scr;
    if ((size_t)h * 16 > w) w = (size_t)h * 16;
    return PFA_WS_ROUND(w);
}

// Bytes of the hot (LDS) and cold (global) parts for h, 16-byte aligned.
// (GG and W1 together hold H: (h+1)(h+2) 8 + 8 h^2 >= 16 (h-1)^2.)
PFA_HD static inline size_t pfa_ws_lds_bytes(int h, bool sp = true)
{
    size_t w = 0;
    w += pfa_ws_gg_bytes(h);                             // GG (H with W1)
    w += PFA_WS_ROUND((size_t)h * h * 8);                // W1
    w += pfa_ws_mu_bytes(h);                             // mu / hv, hw, hpiv / Gy
    w += PFA_WS_ROUND((size_t)h * 8) * 3;                // bb, mw, vec
    w += 16;                                             // flag
    if (sp) w += PFA_WS_ROUND((size_t)h * (h - 1) / 2 * 8);   // sp
    return w;
}

PFA_HD static inline size_t pfa_ws_cold_bytes(int h)
{
    size_t d = h - 1, w = 0;
    w += PFA_WS_ROUND((size_t)h * h * 8) * 3;            // Z, W2, W3
    w += PFA_WS_ROUND((size_t)h * d * 8) * 2;            // Bi, ZB
    w += PFA_WS_ROUND(d * d * 8);                        // mat
    return w;
}

PFA_HD static inline size_t pfa_ws_bytes(int h) { return pfa_ws_lds_bytes(h) + pfa_ws_cold_bytes(h); }

PFA_HD static inline void pfa_ws_carve(PfaWS &W, char *hot, char *cold, int h, bool sp = true)
{
    size_t d = h - 1;
    auto take = [](char *&p, size_t bytes) { char *r = p; p += PFA_WS_ROUND(bytes); return r; };
    W.h = h;
    W.d = (int)d;
    W.GG = (pfa_i128 *)take(hot, pfa_ws_gg_bytes(h));
    W.W1 = (int64_t *)take(hot, (size_t)h * h * 8);
    char *mr = take(hot, pfa_ws_mu_bytes(h));
    W.mu = (double *)mr;
    W.hv = (pfa_i128 *)mr;
    W.hw = (pfa_i128 *)(mr + PFA_WS_ROUND((size_t)h * 16));
    W.hpiv = (int *)(mr + 2 * PFA_WS_ROUND((size_t)h * 16));
    W.bb = (double *)take(hot, (size_t)h * 8);
    W.mw = (double *)take(hot, (size_t)h * 8);
    W.vec = (int64_t *)take(hot, (size_t)h * 8);
    W.flag = (int *)take(hot, 16);
    W.sp = sp ? (double *)take(hot, (size_t)h * (h - 1) / 2 * 8) : nullptr;
    W.Z = (int64_t *)take(cold, (size_t)h * h * 8);
    W.W2 = (int64_t *)take(cold, (size_t)h * h * 8);
    W.W3 = (int64_t *)take(cold, (size_t)h * h * 8);
    W.Bi = (int64_t *)take(cold, (size_t)h * d * 8);
    W.ZB = (int64_t *)take(cold, (size_t)h * d * 8);
    W.mat = (int64_t *)take(cold, d * d * 8);
}

// team-wide "any lane failed": lanes set *flag, then all read it
PFA_HD static inline int pfa_team_any(const PfaTeam &T, int *flag, int mine)
{
    if (mine) *flag = 1;
    PFA_SYNC();
    int r = *flag;
    PFA_SYNC();
    return r;
}

PFA_HD static inline void pfa_team_clear(const PfaTeam &T, int *flag)
{
    if (T.lane == 0) *flag = 0;
    PFA_SYNC();
}

// Extended Euclid with floor division (signs as the reference's Bezout
// steps, pfv_lattice.h:116-130).
PFA_HD static inline int64_t pfa_xgcd(int64_t a, int64_t b, int64_t *s, int64_t *t)
{
    int64_t r0 = a, r1 = b, s0 = 1, s1 = 0, t0 = 0, t1 = 1;
    while (r1 != 0) {
        int64_t q = pfa_fdiv64(r0, r1), x;
        x = r0 - q * r1; r0 = r1; r1 = x;
        x = s0 - q * s1; s0 = s1; s1 = x;
        x = t0 - q * t1; t0 = t1; t1 = x;
    }
    *s = s0;
    *t = t0;
    return r0;
}

// U (n x n, stride n) unimodular with U v = (gcd, 0, ..., 0); v is read
// from shared memory and not modified.  Returns 0, -1 overflow, -2 degenerate.
PFA_HD static inline int pfa_unimodular(const PfaTeam &T, const int64_t *v, int n, int64_t *U, int *flag)
{
    for (int e = T.lane; e < n * n; e += T.n) U[e] = (e / n) == (e % n);
    pfa_team_clear(T, flag);
    int64_t w0 = v[0];
    int bad = 0;
    for (int k = 1; k < n; ++k) {
        int64_t wk = v[k];
        if (wk == 0) continue;
        int64_t s, t, g = pfa_xgcd(w0, wk, &s, &t);
        if (g == 0) return -2;
        PFA_OPS(6 * n + 4);
        int64_t m10 = -pfa_fdiv64(wk, g), m11 = pfa_fdiv64(w0, g);
        w0 = g;
        for (int r = T.lane; r < n; r += T.n) {      // each lane owns columns r
            pfa_i128 a = (pfa_i128)s * U[r] + (pfa_i128)t * U[k * n + r];
            pfa_i128 b = (pfa_i128)m10 * U[r] + (pfa_i128)m11 * U[k * n + r];
            if (!pfa_fits64(a) || !pfa_fits64(b)) bad = 1;
            U[r] = (int64_t)a;
            U[k * n + r] = (int64_t)b;
        }
    }
    return pfa_team_any(T, flag, bad) ? -1 : 0;
}

// nearest integer to a / b (b != 0), ties up; |a| < 2^126
PFA_HD static inline pfa_i128 pfa_round_div128(pfa_i128 a, pfa_i128 b)
{
    if (b < 0) { a = -a; b = -b; }
    return pfa_fdiv128(2 * a + b, 2 * b);
}

PFA_HD static inline int pfa_mag_below(pfa_i128 a, int bits)
{
    const pfa_i128 lim = (pfa_i128)1 << bits;
    return a < lim && a > -lim;
}

// r -= q * u over n entries with |r| kept below 2^125 (0 on overflow)
PFA_HD static inline int pfa_axpy128(pfa_i128 *r, pfa_i128 q, const int64_t *u, int n)
{
    if (q == 0) return 1;
    const double qd = q < 0 ? -(double)q : (double)q;
    for (int c = 0; c < n; ++c) {
        if (u[c] == 0) continue;
        const double ud = u[c] < 0 ? -(double)u[c] : (double)u[c];
        if (qd * ud >= 0x1p124) return 0;
        r[c] -= q * (pfa_i128)u[c];
        if (!pfa_mag_below(r[c], 125)) return 0;
    }
    return 1;
}

// The same U as pfa_unimodular (U v = (g, 0, ..., 0), |det U| = 1) for
// vectors whose plain Euclid chain overflows int64: there the Bezout row
// (entries up to max|v|) times v[k] becomes kernel row k.  Here the chain
// starts at the first nonzero entry i0 of v, and every new kernel row and
// the Bezout row are formed in int128, then reduced against the earlier
// kernel rows in reverse order of processing (row m is nonzero only at the
// indices processed before it and at its own index, where its entry is
// nonzero), which keeps their entries near max|v|.  One lane computes (a
// rare path); x, r (int128, n) and ord (int, n) are scratch.  The chain
// meets the same gcd steps as pfa_unimodular's, so g is the same.  Returns
// 0, or -1 when an entry would still leave its range.
PFA_HD static inline int pfa_unimodular_reduced(const PfaTeam &T, const int64_t *v, int n,
                                                int64_t *U, pfa_i128 *x, pfa_i128 *r, int *ord,
                                                int *flag)
{
    pfa_team_clear(T, flag);
    int bad = 0;
    if (T.lane == 0) {
        int i0 = -1;
        for (int i = 0; i < n && i0 < 0; ++i)
            if (v[i] != 0) i0 = i;
        int no = 0;
        if (i0 < 0) {                           // v = 0: the identity, as the plain chain
            for (int e = 0; e < n * n; ++e) U[e] = (e / n) == (e % n);
        } else {
            ord[no++] = i0;
            for (int i = 0; i < n; ++i)
                if (i != i0) ord[no++] = i;
            for (int e = 0; e < n * n; ++e) U[e] = 0;
            for (int c = 0; c < n; ++c) x[c] = c == i0;
        }
        int64_t w = i0 < 0 ? 0 : v[i0];
        for (int m = 1; m < n && i0 >= 0 && !bad; ++m) {
            const int k = ord[m];
            const int64_t vk = v[k];
            if (vk == 0) {
                for (int c = 0; c < n; ++c) r[c] = c == k;
            } else {
                int64_t s, t, g = pfa_xgcd(w, vk, &s, &t);
                if (g == 0) { bad = 1; break; }
                const int64_t m10 = -pfa_fdiv64(vk, g), m11 = pfa_fdiv64(w, g);
                if (!pfa_mag_below(m10, 62) || !pfa_mag_below(s, 62)) { bad = 1; break; }
                for (int c = 0; c < n; ++c) {
                    r[c] = (pfa_i128)m10 * x[c] + (c == k ? (pfa_i128)m11 : 0);
                    x[c] = (pfa_i128)s * x[c] + (c == k ? (pfa_i128)t : 0);
                }
                w = g;
            }
            // reduce r against kernel rows m-1 .. 1, then store it as row m
            for (int j = m - 1; j >= 1 && !bad; --j) {
                const int64_t *uj = U + (size_t)j * n;
                if (!pfa_axpy128(r, pfa_round_div128(r[ord[j]], uj[ord[j]]), uj, n)) bad = 1;
            }
            for (int c = 0; c < n && !bad; ++c) {
                if (!pfa_fits64(r[c])) bad = 1;
                else U[(size_t)m * n + c] = (int64_t)r[c];
            }
            // reduce the Bezout row against rows m .. 1; it must stay small
            // for the next step's products
            for (int j = m; j >= 1 && !bad; --j) {
                const int64_t *uj = U + (size_t)j * n;
                if (!pfa_axpy128(x, pfa_round_div128(x[ord[j]], uj[ord[j]]), uj, n)) bad = 1;
            }
            for (int c = 0; c < n && !bad; ++c)
                if (!pfa_mag_below(x[c], 62)) bad = 1;
        }
        if (!bad && i0 >= 0)
            for (int c = 0; c < n; ++c) U[c] = (int64_t)x[c];
    }
    return pfa_team_any(T, flag, bad) ? -1 : 0;
}

// Gram-Schmidt row kk (mu[kk][0:kk], bb[kk]) from the exact Gram matrix:
//   mu[kk][j] = (G[kk][j] - sum_{l<j} mu[j][l] mu[kk][l] bb[l]) / bb[j],
//   bb[kk]    =  G[kk][kk] - sum_{l<kk} mu[kk][l] mu[kk][l] bb[l].
// On the device the lanes run it right-looking: lane j < kk keeps the
// running numerator of column j, lane kk that of bb[kk]; at step l lane l's
// value is final and is broadcast.  Each lane subtracts its terms in the
// host's order (l ascending) with the host's association, so every value is
// bit-identical to the sequential host loop.
PFA_HD static inline void pfa_gs_row(const PfaTeam &T, const PfaWS &W, int kk, const pfa_i128 *GG)
{
    double *mu = W.mu, *bb = W.bb, *sp = W.sp;
    const int rk = pfa_mi(kk, 0);
    if (T.lane == 0) PFA_OPS(1.5 * kk * (kk - 1) + 5 * kk + 1);
#if defined(__HIP_DEVICE_COMPILE__)
    if (T.n > kk) {
        const int j = T.lane;
        double r = j <= kk ? (double)GG[pfa_gi(kk, j)] : 0.0;
        const int rj = pfa_mi(j, 0);
        for (int l = 0; l < kk; ++l) {
            const double m = pfa_readlane_f64(r, l) / bb[l];
            if (j == l) mu[rk + l] = m;
            else if (j > l && j < kk) r -= mu[rj + l] * m * bb[l];
            else if (j == kk) {
                if (sp) sp[rk + l] = r;
                r -= m * m * bb[l];
            }
        }
        if (j == kk) bb[kk] = r;
        PFA_SYNC();
        return;
    }
#endif
    for (int j = 0; j < kk; ++j) {
        double dd = (double)GG[pfa_gi(kk, j)];
        for (int l = 0; l < j; ++l) dd -= mu[pfa_mi(j, l)] * mu[rk + l] * bb[l];
        mu[rk + j] = dd / bb[j];
    }
    double dd = (double)GG[pfa_gi(kk, kk)];
    for (int l = 0; l < kk; ++l) {
        if (sp && T.lane == 0) sp[rk + l] = dd;
        dd -= mu[rk + l] * mu[rk + l] * bb[l];
    }
    if (T.lane == 0) bb[kk] = dd;
    PFA_SYNC();
}

// Coefficients mw[0:m] of the row `last` (row m of GG) on the reduced
// basis: mw[j] = (G[m][j] - sum_{l<j} mu[j][l] mw[l] bb[l]) / bb[j], right-
// looking over the lanes on the device as pfa_gs_row (bit-identical).
PFA_HD static inline void pfa_gs_last(const PfaTeam &T, const PfaWS &W, int m, const pfa_i128 *GG)
{
    double *mu = W.mu, *bb = W.bb, *mw = W.mw;
    if (T.lane == 0) PFA_OPS(1.5 * m * (m - 1) + 2 * m);
#if defined(__HIP_DEVICE_COMPILE__)
    if (T.n >= m) {
        const int j = T.lane;
        double r = j < m ? (double)GG[pfa_gi(m, j)] : 0.0;
        const int rj = pfa_mi(j, 0);
        for (int l = 0; l < m; ++l) {
            const double w = pfa_readlane_f64(r, l) / bb[l];
            if (j == l) mw[l] = w;
            else if (j > l && j < m) r -= mu[rj + l] * w * bb[l];
        }
        PFA_SYNC();
        return;
    }
#endif
    for (int j = 0; j < m; ++j) {
        double dd = (double)GG[pfa_gi(m, j)];
        for (int l = 0; l < j; ++l) dd -= mu[pfa_mi(j, l)] * mw[l] * bb[l];
        if (T.lane == 0) mw[j] = dd / bb[j];
        PFA_SYNC();
    }
}

// b_k -= q b_j on the rows (columns split over lanes) and on GG (exact).
// Row index m means `last`.
PFA_HD static inline int pfa_lll_sub(const PfaTeam &T, const PfaWS &W, int64_t *B, int64_t *last,
                                     int m, int n, pfa_i128 *GG, int M, int kk, int jj, int64_t q)
{
    int64_t *bk = kk == m ? last : &B[kk * n];
    const int64_t *bj = jj == m ? last : &B[jj * n];
    int bad = 0;
    if (T.lane == 0) PFA_OPS(2 * n + 2 * M + 3);
    for (int c = T.lane; c < n; c += T.n) {
        pfa_i128 x = (pfa_i128)bk[c] - (pfa_i128)q * bj[c];
        if (!pfa_fits64(x)) bad = 1;
        bk[c] = (int64_t)x;
    }
    // new G_kk = G_kk - 2 q G_kj + q^2 G_jj, computed before row kk changes
    pfa_i128 kj = GG[pfa_gi(kk, jj)], jv = GG[pfa_gi(jj, jj)], t1, t2;
    int badk = pfa_mul_ovf(jv, (pfa_i128)q, &t1) || pfa_sub_ovf(t1, 2 * kj, &t1) ||
               pfa_mul_ovf(t1, (pfa_i128)q, &t2) || pfa_add_ovf(GG[pfa_gi(kk, kk)], t2, &t2);
    PFA_SYNC();
    for (int l = T.lane; l < M; l += T.n) {
        if (l == kk) continue;
        pfa_i128 t3;
        if (pfa_mul_ovf(GG[pfa_gi(jj, l)], (pfa_i128)q, &t3) ||
            pfa_sub_ovf(GG[pfa_gi(kk, l)], t3, &t3))
            bad = 1;
        GG[pfa_gi(kk, l)] = t3;
    }
    if (T.lane == 0) GG[pfa_gi(kk, kk)] = t2;
    return pfa_team_any(T, W.flag, bad || badk);
}

// LLL (delta 0.99) of the m rows of B (length n, stride n) given their exact
// Gram matrix GG (lower triangle); `last` (row m of GG) is size-reduced
// afterwards.  0, -1 overflow, -2 rank loss, -3 no convergence.
PFA_HD static inline int pfa_lll_core(const PfaTeam &T, const PfaWS &W, int64_t *B, int m, int n,
                                      pfa_i128 *GG, int64_t *last)
{
    const int M = m + (last != nullptr);
    double *mu = W.mu, *bb = W.bb, *mw = W.mw;
    int iters = 0;
    pfa_team_clear(T, W.flag);
    if (m >= 1) {
        pfa_gs_row(T, W, 0, GG);
        if (!(bb[0] > 0)) return -2;
    }
    if (m >= 2) pfa_gs_row(T, W, 1, GG);
    PFA_SYNC();
    int k = 1;
    while (k < m) {
        if (++iters > 100000) return -3;
        for (int pass = 0; pass < 4; ++pass) {
            int changed = 0;
            for (int j = k - 1; j >= 0; --j) {
                double q = nearbyint(mu[pfa_mi(k, j)]);
                if (T.lane == 0) PFA_OPS(1);
                if (q == 0.0) continue;
                if (T.lane == 0) PFA_OPS(2 * j + 1);
                if (!(fabs(q) <= 9.0e18)) return -1;
                if (pfa_lll_sub(T, W, B, last, m, n, GG, M, k, j, (int64_t)q)) return -1;
                for (int l = T.lane; l < j; l += T.n) mu[pfa_mi(k, l)] -= q * mu[pfa_mi(j, l)];
                if (T.lane == 0) mu[pfa_mi(k, j)] -= q;
                changed = 1;
                PFA_SYNC();
            }
            if (!changed) break;
            pfa_gs_row(T, W, k, GG);
            PFA_SYNC();
        }
        if (!(bb[k] > 0)) return -2;
        const double mk1 = mu[pfa_mi(k, k - 1)];
        if (T.lane == 0) PFA_OPS(4);
        if (bb[k] < (0.99 - mk1 * mk1) * bb[k - 1]) {
            for (int c = T.lane; c < n; c += T.n) {
                int64_t t = B[k * n + c];
                B[k * n + c] = B[(k - 1) * n + c];
                B[(k - 1) * n + c] = t;
            }
            // exchange indices k-1 and k of the symmetric GG: entries (k, l)
            // and (k-1, l) for l outside {k-1, k}, and the two diagonals
            // ((k, k-1) stays)
            for (int l = T.lane; l < M; l += T.n) {
                if (l == k || l == k - 1) {
                    if (l == k) {
                        pfa_i128 t = GG[pfa_gi(k, k)];
                        GG[pfa_gi(k, k)] = GG[pfa_gi(k - 1, k - 1)];
                        GG[pfa_gi(k - 1, k - 1)] = t;
                    }
                    continue;
                }
                pfa_i128 t = GG[pfa_gi(k, l)];
                GG[pfa_gi(k, l)] = GG[pfa_gi(k - 1, l)];
                GG[pfa_gi(k - 1, l)] = t;
            }
            PFA_SYNC();
            if (k >= 2 && !W.sp) {
                // rows below k-1 are unchanged (what the reference recomputes
                // of them comes out the same): only the new row k-1
                k--;
                pfa_gs_row(T, W, k, GG);
            } else if (k >= 2) {
                // rows below k-1 are unchanged, and the new row k-1 (the old
                // row k's vector) has the old row k's first k-1 coefficients
                // and running values, bit for bit (the same operations on
                // the same operands); its b* is the old row's running value
                // before its term k-1.  What the reference recomputes here
                // is moved.
                double *sp = W.sp;
                const int ra = pfa_mi(k - 1, 0), rb = pfa_mi(k, 0);
                for (int l = T.lane; l < k - 1; l += T.n) {
                    mu[ra + l] = mu[rb + l];
                    sp[ra + l] = sp[rb + l];
                }
                if (T.lane == 0) bb[k - 1] = sp[rb + k - 1];
                k--;
            } else {
                pfa_gs_row(T, W, 0, GG);
                pfa_gs_row(T, W, 1, GG);
            }
            PFA_SYNC();
        } else {
            k++;
            if (k < m) pfa_gs_row(T, W, k, GG);
            PFA_SYNC();
        }
    }
    if (last && m > 0) {
        for (int pass = 0; pass < 4; ++pass) {
            pfa_gs_last(T, W, m, GG);
            int changed = 0;
            for (int j = m - 1; j >= 0; --j) {
                double q = nearbyint(mw[j]);
                if (T.lane == 0) PFA_OPS(1);
                if (q == 0.0) continue;
                if (T.lane == 0) PFA_OPS(2 * j + 1);
                if (!(fabs(q) <= 9.0e18)) return -1;
                if (pfa_lll_sub(T, W, B, last, m, n, GG, M, m, j, (int64_t)q)) return -1;
                for (int l = T.lane; l < j; l += T.n) mw[l] -= q * mu[pfa_mi(j, l)];
                if (T.lane == 0) mw[j] -= q;
                changed = 1;
                PFA_SYNC();
            }
            if (!changed) break;
        }
    }
    PFA_SYNC();
    return 0;
}

// Euclidean LLL of the m rows of B
PFA_HD static inline int pfa_lll_euclid(const PfaTeam &T, const PfaWS &W, int64_t *B, int m, int n)
{
    if (m < 2) return 0;
    int bad = 0;
    if (T.lane == 0) PFA_OPS((double)m * (m + 1) * n);
    for (int e = T.lane; e < m * m; e += T.n) {
        int i = e / m, j = e % m;
        if (j > i) continue;
        pfa_i128 s = 0;
        for (int c = 0; c < n; ++c)
            if (pfa_add_ovf(s, (pfa_i128)B[i * n + c] * B[j * n + c], &s)) bad = 1;
        W.GG[pfa_gi(i, j)] = s;
    }
    pfa_team_clear(T, W.flag);
    if (pfa_team_any(T, W.flag, bad)) return -1;
    return pfa_lll_core(T, W, B, m, n, W.GG, nullptr);
}

// LLL of the m rows of B w.r.t. the form G (n x n); `last` size-reduced.
// Gy (n entries) uses W.vec as scratch.
PFA_HD static inline int pfa_lll_form(const PfaTeam &T, const PfaWS &W, int64_t *B, int m, int n,
                                      const int64_t *G, int64_t *last)
{
    const int M = m + (last != nullptr);
    pfa_i128 *Gy = (pfa_i128 *)W.mu;        // mu is free until pfa_lll_core
    int bad = 0;
    if (T.lane == 0) PFA_OPS(2.0 * M * n * n + (double)M * (M + 1) * n);
    pfa_team_clear(T, W.flag);
    for (int j = 0; j < M; ++j) {
        const int64_t *y = j == m ? last : &B[j * n];
        for (int i = T.lane; i < n; i += T.n) {
            pfa_i128 r = 0;
            for (int c = 0; c < n; ++c)
                if (pfa_add_ovf(r, (pfa_i128)G[i * n + c] * y[c], &r)) bad = 1;
            Gy[i] = r;
        }
        PFA_SYNC();
        for (int i = T.lane; i <= j; i += T.n) {
            const int64_t *x = i == m ? last : &B[i * n];
            pfa_i128 s = 0, t;
            for (int c = 0; c < n; ++c) {
                if (!x[c]) continue;
                if (pfa_mul_ovf(Gy[c], (pfa_i128)x[c], &t) || pfa_add_ovf(s, t, &s)) bad = 1;
            }
            W.GG[pfa_gi(i, j)] = s;
        }
        PFA_SYNC();
    }
    if (pfa_team_any(T, W.flag, bad)) return -1;
    return pfa_lll_core(T, W, B, m, n, W.GG, last);
}

// max |a_e| over n entries, the same value on every lane of the team
PFA_HD static inline uint64_t pfa_team_absmax(const PfaTeam &T, const int64_t *a, int n)
{
    uint64_t m = 0;
    for (int e = T.lane; e < n; e += T.n) {
        const uint64_t x = a[e] < 0 ? 0 - (uint64_t)a[e] : (uint64_t)a[e];
        m = x > m ? x : m;
    }
#if defined(__HIP_DEVICE_COMPILE__)
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t y = __shfl_xor(m, o, 64);
        m = y > m ? y : m;
    }
#endif
    return m;
}

// Sums of k products of entries bounded by amax and bmax stay below 2^62,
// so they are exact in int64 and need no overflow checks.
PFA_HD static inline bool pfa_int64_safe(uint64_t amax, uint64_t bmax, int k)
{
    return (double)amax * (double)bmax * (double)k < 0x1p62;
}

// C = A B (int64, exact checks); entries split over lanes.  C must not alias.
PFA_HD static inline int pfa_matmul(const PfaTeam &T, const PfaWS &W, const int64_t *A,
                                    const int64_t *B, int64_t *C, int m, int k, int n)
{
    int bad = 0;
    if (T.lane == 0) PFA_OPS(2.0 * m * k * n);
    if (pfa_int64_safe(pfa_team_absmax(T, A, m * k), pfa_team_absmax(T, B, k * n), k)) {
        for (int e = T.lane; e < m * n; e += T.n) {
            const int i = e / n, j = e % n;
            int64_t s = 0;
            for (int l = 0; l < k; ++l) s += A[i * k + l] * B[l * n + j];
            C[e] = s;
        }
        PFA_SYNC();
        return 0;
    }
    for (int e = T.lane; e < m * n; e += T.n) {
        int i = e / n, j = e % n;
        pfa_i128 s = 0;
        for (int l = 0; l < k; ++l)
            if (pfa_add_ovf(s, (pfa_i128)A[i * k + l] * B[l * n + j], &s)) bad = 1;
        if (!pfa_fits64(s)) bad = 1;
        C[e] = (int64_t)s;
    }
    pfa_team_clear(T, W.flag);
    return pfa_team_any(T, W.flag, bad) ? -1 : 0;
}

PFA_HD static inline void pfa_copy(const PfaTeam &T, int64_t *dst, const int64_t *src, int n)
{
    PFA_SYNC();
    for (int e = T.lane; e < n; e += T.n) dst[e] = src[e];
    PFA_SYNC();
}

// ---------------------------------------------------------------- HNF
// Row HNF (upper echelon, positive pivots, entries above a pivot reduced
// into [0, pivot)) of the r x n int64 matrix A, rows inserted one at a time
// into an already reduced HNF (pfv_lattice.h:205-313).  Entries are kept
// below 2^125; -1 on overflow.  H: r x n int128 (zero rows last).
// Row ops are split over lanes; pivots/quotients are read by every lane.
#define PFA_HLIM ((pfa_i128)1 << 125)

PFA_HD static inline int pfa_hnf_muladd(pfa_i128 a, pfa_i128 b, pfa_i128 c, pfa_i128 *r)
{
    pfa_i128 x;
    if (pfa_mul_ovf(a, b, &x) || pfa_add_ovf(x, c, &x) || !(x < PFA_HLIM && x > -PFA_HLIM)) return 1;
    *r = x;
    return 0;
}

// row[c0:] -= q prow[c0:] (team); returns the failure flag of this lane
PFA_HD static inline int pfa_hnf_rowsub(const PfaTeam &T, pfa_i128 *row, const pfa_i128 *prow,
                                        pfa_i128 q, int c0, int n)
{
    int bad = 0;
    if (q == 0) return 0;
    if (T.lane == 0) PFA_OPS(2 * (n - c0));
    for (int c = c0 + T.lane; c < n; c += T.n)
        if (pfa_hnf_muladd(-q, prow[c], row[c], &row[c])) bad = 1;
    return bad;
}

// reduce row by rows t0..m-1 of H at their pivots (in order)
PFA_HD static inline int pfa_hnf_reduce(const PfaTeam &T, int *flag, pfa_i128 *row,
                                        const pfa_i128 *H, const int *piv, int t0, int m, int n)
{
    int bad = 0;
    for (int l = t0; l < m; ++l) {
        int j = piv[l];
        pfa_i128 q = pfa_fdiv128(row[j], H[l * n + j]);
        if (T.lane == 0) PFA_OPS(1);
        PFA_SYNC();
        bad |= pfa_hnf_rowsub(T, row, &H[l * n], q, j, n);
        PFA_SYNC();
    }
    return pfa_team_any(T, flag, bad);
}

PFA_HD static inline int pfa_hnf(const PfaTeam &T, const PfaWS &W, const int64_t *A, int r, int n,
                                 pfa_i128 *H, pfa_i128 *v, pfa_i128 *w, int *piv)
{
    int m = 0;
    pfa_team_clear(T, W.flag);
    for (int k = 0; k < r; ++k) {
        for (int c = T.lane; c < n; c += T.n) v[c] = A[k * n + c];
        PFA_SYNC();
        int t = 0;
        for (;;) {
            int j = 0;
            while (j < n && v[j] == 0) ++j;
            if (j == n) break;
            while (t < m && piv[t] < j) ++t;
            if (t == m || piv[t] > j) {
                // new pivot j: insert v as row t
                PFA_SYNC();
                for (int l = m; l > t; --l)
                    for (int c = T.lane; c < n; c += T.n) H[l * n + c] = H[(l - 1) * n + c];
                PFA_SYNC();
                if (T.lane == 0)
                    for (int l = m; l > t; --l) piv[l] = piv[l - 1];
                int neg = v[j] < 0;
                for (int c = T.lane; c < n; c += T.n) H[t * n + c] = neg ? -v[c] : v[c];
                if (T.lane == 0) piv[t] = j;
                PFA_SYNC();
                m++;
                if (pfa_hnf_reduce(T, W.flag, &H[t * n], H, piv, t + 1, m, n)) return -1;
                break;
            }
            pfa_i128 *hr = &H[t * n];
            pfa_i128 a = hr[j], b = v[j];
            int bad = 0;
            if (pfa_tmod128(b, a) == 0) {
                pfa_i128 q = pfa_tdiv128(b, a);
                if (T.lane == 0) PFA_OPS(2);
                PFA_SYNC();
                bad = pfa_hnf_rowsub(T, v, hr, q, j, n);
                if (pfa_team_any(T, W.flag, bad)) return -1;
            } else {
                pfa_i128 r0 = a, r1 = b, s0 = 1, s1 = 0, x0 = 0, x1 = 1;
                if (T.lane == 0) PFA_OPS(8 * (n - j) + 2);
                while (r1) {
                    pfa_i128 q = pfa_tdiv128(r0, r1), tt;
                    if (T.lane == 0) PFA_OPS(7);
                    tt = r0 - q * r1; r0 = r1; r1 = tt;
                    tt = s0 - q * s1; s0 = s1; s1 = tt;
                    tt = x0 - q * x1; x0 = x1; x1 = tt;
                }
                if (r0 < 0) { r0 = -r0; s0 = -s0; x0 = -x0; }
                pfa_i128 ag = pfa_tdiv128(a, r0), bg = pfa_tdiv128(b, r0);
                PFA_SYNC();
                for (int c = j + T.lane; c < n; c += T.n) {
                    pfa_i128 hn = 0, vn = 0, t1 = 0;
                    if (pfa_mul_ovf(s0, hr[c], &t1) || pfa_hnf_muladd(x0, v[c], t1, &hn)) bad = 1;
                    if (pfa_mul_ovf(-bg, hr[c], &t1) || pfa_hnf_muladd(ag, v[c], t1, &vn)) bad = 1;
                    w[c] = hn;
                    v[c] = vn;
                }
                PFA_SYNC();
                for (int c = j + T.lane; c < n; c += T.n) hr[c] = w[c];
                if (pfa_team_any(T, W.flag, bad)) return -1;
                if (pfa_hnf_reduce(T, W.flag, hr, H, piv, t + 1, m, n)) return -1;
            }
            if (pfa_hnf_reduce(T, W.flag, v, H, piv, t + 1, m, n)) return -1;
            for (int l = 0; l < t; ++l)
                if (pfa_hnf_reduce(T, W.flag, &H[l * n], H, piv, t, m, n)) return -1;
            t++;
        }
        for (int l = 0; l < t && l < m; ++l)
            if (pfa_hnf_reduce(T, W.flag, &H[l * n], H, piv, t < m ? t : m, m, n)) return -1;
    }
    for (int e = m * n + T.lane; e < r * n; e += T.n) H[e] = 0;
    PFA_SYNC();
    return 0;
}

// --------------------------------------------------- lattice build
// ---------------------------------------------------------------------
// A certificate that a direction's tree is empty, before the lattice is
// built.  Every point has M0 >= linmin and q(M) = -M^T Z M <= qmax with M
// in the real subspace V = {M : w.M = 0}, w = Z p, where q is positive
// definite; so M0^2 <= r q(M) <= r qmax with r = max over V of M0^2 / q.
// With the bordered matrix of -Z and w ordered (j, border, the other
// coordinates, 0) for a j != 0 with w_j != 0 and Z_jj != 0, its leading
// principal minors P_k of orders k >= 3 are the bordered minors of q with
// one constraint, so q is positive definite on V iff every such P_k is
// negative; then 1 / r = P_{h+1} / P_h (the Schur complement of the
// coordinate M0, ordered last: the minimum of q over V with M0 = 1).  The
// minors are the pivots of a fraction-free (Bareiss) elimination, exact in
// int128; a value beyond it gives no certificate.
//
// The tree is declared empty when 5 linmin^2 |P_{h+1}| > 9 (qmax + 1) |P_h|,
// i.e. linmin^2 > 1.8 r (qmax + 1).  In the cut-aware basis M0 = g c_{d-1},
// the root's coordinate, and the search's root interval is |U_{d-1,d-1} c|
// <= R with R^2 = (Qe + slack + Kerr Qe)(1 + Kerr)^2, Qe = qmax / (1 - phi)
// (1 + 8u), slack <= 1e-4 + 1e-9 (Qe + 1); since q <= (1 + phi) |U c|^2,
// g^2 / U_{d-1,d-1}^2 <= (1 + phi) r, so the interval reaches at most
// M0^2 <= r (1 + phi) / (1 - phi) (qmax (1 + 1e-8) + 2e-4) < 1.7 r (qmax + 1)
// for phi < 0.25 (a larger phi sends the problem to the exact host path),
// and the interval's 1e-9 rounding allowance stays inside the margin.  The
// search would enter the root, find no candidate and stop: one node, as
// the record's root-empty flag counts it.  Lanes own rows; the scratch is
// the cold workspace (before the lattice is built: Z, then W2 .. mat as
// (h+1)^2 int128, which fits for h >= 3) and hv.  Returns 1 (empty) or 0.
PFA_HD static inline int pfa_empty_certificate(const PfaTeam &T, const PfaWS &W, const int64_t *kappa,
                                               const int64_t *p, int64_t qmax, int64_t linmin)
{
    const int h = W.h, n = h + 1;
    if (linmin <= 0 || qmax < 0 || h < 3) return 0;
    int64_t *Zt = W.Z;                                   // (h, h)
    pfa_i128 *w = W.hv;                                  // (h)
    pfa_i128 *Mx = (pfa_i128 *)W.W2;                     // (n, n) over W2 .. mat
    int bad = 0;
    if (T.lane == 0) PFA_OPS(2.0 * h * h * (h - 1) + 2.0 * h * (h - 1));
    pfa_team_clear(T, W.flag);
    for (int e = T.lane; e < h * h; e += T.n) {
        pfa_i128 s = 0;
        for (int c = 1; c < h; ++c)
            if (pfa_add_ovf(s, (pfa_i128)kappa[(size_t)e * h + c] * p[c], &s)) bad = 1;
        if (!pfa_fits64(s)) bad = 1;
        Zt[e] = (int64_t)s;
    }
    if (pfa_team_any(T, W.flag, bad)) return 0;
    for (int a = T.lane; a < h; a += T.n) {
        pfa_i128 s = 0;
        for (int b = 1; b < h; ++b)
            if (pfa_add_ovf(s, (pfa_i128)Zt[a * h + b] * p[b], &s)) bad = 1;
        w[a] = s;
    }
    if (pfa_team_any(T, W.flag, bad)) return 0;
    // j: the smallest |w_j| != 0 with j != 0 and Z_jj != 0 (the same on every lane)
    int j = -1;
    for (int i = 1; i < h; ++i)
        if (w[i] != 0 && Zt[i * h + i] != 0 && (j < 0 || pfa_abs128(w[i]) < pfa_abs128(w[j]))) j = i;
    if (j < 0) return 0;
    // position r of the ordering -> coordinate (h: the border)
    auto ord = [&](int r) { return r == 0 ? j : r == 1 ? h : r == h ? 0 : (r - 1 < j ? r - 1 : r); };
    for (int r = T.lane; r < n; r += T.n) {
        const int a = ord(r);
        for (int c = 0; c < n; ++c) {
            const int b = ord(c);
            Mx[r * n + c] = a == h ? (b == h ? (pfa_i128)0 : w[b]) : (b == h ? w[a] : -(pfa_i128)Zt[a * h + b]);
        }
    }
    PFA_SYNC();
    pfa_i128 prev = 1, Ph = 0;
    for (int k = 0; k < n - 1; ++k) {
        const pfa_i128 piv = Mx[k * n + k];              // the leading minor of order k + 1
        if (piv == 0 || (k >= 2 && piv > 0)) return 0;   // not (provably) positive definite on V
        if (k + 1 == h) Ph = piv;
        // entries (r, c) > k: (piv M_rc - M_rk M_kc) / prev, exact; the
        // products need 256 bits, the quotients (minors) must fit int128
        const pfa_exdiv D = pfa_exdiv_prep(prev);
        const int m = n - 1 - k;
        if (T.lane == 0) PFA_OPS(4.0 * m * m);
        for (int e = T.lane; e < m * m; e += T.n) {
            const int r = k + 1 + e / m, c = k + 1 + e % m;
            const pfa_w256 N = pfa_sub256(pfa_smul256(piv, Mx[r * n + c]), pfa_smul256(Mx[r * n + k], Mx[k * n + c]));
            pfa_i128 q;
            if (!pfa_exdiv256(N, D, &q)) bad = 1;
            Mx[r * n + c] = q;
        }
        if (pfa_team_any(T, W.flag, bad)) return 0;
        prev = piv;
    }
    const pfa_i128 Pn = Mx[(n - 1) * n + (n - 1)];
    if (!(Pn < 0 && Ph < 0)) return 0;
    // 5 linmin^2 |Pn| > 9 (qmax + 1) |Ph|
    pfa_i128 a, b;
    if (pfa_mul_ovf((pfa_i128)5 * linmin, (pfa_i128)linmin, &a) || pfa_mul_ovf(9, (pfa_i128)qmax + 1, &b)) return 0;
    return pfa_ult256(pfa_smul256(b, -Ph), pfa_smul256(a, -Pn));
}

// Builds Z, Binter (W.Bi), ZB, mat for p (length h, p[0] = 0).  Returns 0
// or 10 * rc - stage (stage 1 orthogonal, 2 LLL, 4 cut-aware basis, 0 a
// product overflow), as pfl_build's status (pfv_lattice.h:604-680).
PFA_HD static inline int pfa_build_lattice(const PfaTeam &T, const PfaWS &W, const int64_t *kappa,
                                           const int64_t *Mbasis, const int64_t *p)
{
    const int h = W.h, d = W.d;
    int bad = 0, rc;
    pfa_team_clear(T, W.flag);
    if (T.lane == 0) PFA_OPS(2.0 * h * h * h + 4.0 * h * h);
    if (pfa_int64_safe(pfa_team_absmax(T, kappa, h * h * h), pfa_team_absmax(T, p, h), h)) {
        for (int e = T.lane; e < h * h; e += T.n) {
            int64_t s = 0;
            for (int k = 0; k < h; ++k) s += kappa[(size_t)e * h + k] * p[k];
            W.Z[e] = s;
        }
    } else {
        for (int e = T.lane; e < h * h; e += T.n) {
            pfa_i128 s = 0;
            for (int k = 0; k < h; ++k)
                if (pfa_add_ovf(s, (pfa_i128)kappa[(size_t)e * h + k] * p[k], &s)) bad = 1;
            if (!pfa_fits64(s)) bad = 1;
            W.Z[e] = (int64_t)s;
        }
    }
    if (pfa_team_any(T, W.flag, bad)) return -1;
    // T = Z p into W.W2[0:h]; v = T^T Mbasis into W.vec
    int64_t *Tv = W.W2;
    for (int i = T.lane; i < h; i += T.n) {
        pfa_i128 s = 0;
        for (int k = 0; k < h; ++k)
            if (pfa_add_ovf(s, (pfa_i128)W.Z[i * h + k] * p[k], &s)) bad = 1;
        if (!pfa_fits64(s)) bad = 1;
        Tv[i] = (int64_t)s;
    }
    if (pfa_team_any(T, W.flag, bad)) return -1;
    for (int j = T.lane; j < h; j += T.n) {
        pfa_i128 s = 0;
        for (int l = 0; l < h; ++l)
            if (pfa_add_ovf(s, (pfa_i128)Tv[l] * Mbasis[l * h + j], &s)) bad = 1;
        if (!pfa_fits64(s)) bad = 1;
        W.vec[j] = (int64_t)s;
    }
    if (pfa_team_any(T, W.flag, bad)) return -1;
    // orthogonal lattice: rows 1.. of the unimodular U (W1); the reduced
    // chain when the plain one overflows (hv, hw, hpiv are free until the HNF)
    rc = pfa_unimodular(T, W.vec, h, W.W1, W.flag);
    if (rc == -1) rc = pfa_unimodular_reduced(T, W.vec, h, W.W1, W.hv, W.hw, W.hpiv, W.flag);
    if (rc) return rc * 10 - 1;
    int64_t *O = W.W1 + h;                            // (d, h) rows
    PFA_SYNC();
    if ((rc = pfa_lll_euclid(T, W, O, d, h))) return rc * 10 - 2;
    // BT[a][i] = sum_l Mbasis[i][l] O[a][l]  -> W2 (d, h)
    int64_t *BT = W.W2;
    if (T.lane == 0) PFA_OPS(2.0 * d * h * h + 2.0 * d * d * h);
    if (pfa_int64_safe(pfa_team_absmax(T, Mbasis, h * h), pfa_team_absmax(T, O, d * h), h)) {
        for (int e = T.lane; e < d * h; e += T.n) {
            const int a = e / h, i = e % h;
            int64_t s = 0;
            for (int l = 0; l < h; ++l) s += Mbasis[i * h + l] * O[a * h + l];
            BT[e] = s;
        }
    } else {
        for (int e = T.lane; e < d * h; e += T.n) {
            int a = e / h, i = e % h;
            pfa_i128 s = 0;
            for (int l = 0; l < h; ++l)
                if (pfa_add_ovf(s, (pfa_i128)Mbasis[i * h + l] * O[a * h + l], &s)) bad = 1;
            if (!pfa_fits64(s)) bad = 1;
            BT[e] = (int64_t)s;
        }
    }
    if (pfa_team_any(T, W.flag, bad)) return -1;
    // columns with Binter[0] == 0 first (stable); order in W3
    int64_t *order = W.W3;
    if (T.lane == 0) {
        int no = 0;
        for (int a = 0; a < d; ++a) if (BT[a * h] == 0) order[no++] = a;
        for (int a = 0; a < d; ++a) if (BT[a * h] != 0) order[no++] = a;
    }
    PFA_SYNC();
    for (int e = T.lane; e < h * d; e += T.n) {
        int i = e / d, a = e % d;
        W.Bi[e] = BT[order[a] * h + i];
    }
    PFA_SYNC();
    if (pfa_matmul(T, W, W.Z, W.Bi, W.ZB, h, h, d)) return -1;
    if (pfa_int64_safe(pfa_team_absmax(T, W.Bi, h * d), pfa_team_absmax(T, W.ZB, h * d), h)) {
        for (int e = T.lane; e < d * d; e += T.n) {
            const int a = e / d, b = e % d;
            int64_t s = 0;
            for (int i = 0; i < h; ++i) s += W.Bi[i * d + a] * W.ZB[i * d + b];
            W.mat[e] = -s;
        }
    } else {
        for (int e = T.lane; e < d * d; e += T.n) {
            int a = e / d, b = e % d;
            pfa_i128 s = 0;
            for (int i = 0; i < h; ++i)
                if (pfa_add_ovf(s, (pfa_i128)W.Bi[i * d + a] * W.ZB[i * d + b], &s)) bad = 1;
            if (!pfa_fits64(s)) bad = 1;
            W.mat[e] = (int64_t)(-s);
        }
    }
    if (pfa_team_any(T, W.flag, bad)) return -1;

    // cut-aware basis: V l = (g, 0, ...) for l = Binter[0, :]
    int nz = 0;
    for (int a = 0; a < d; ++a) nz += W.Bi[a] != 0;
    if (nz == 0) return 0;
    int64_t *l = W.vec;
    PFA_SYNC();
    for (int a = T.lane; a < d; a += T.n) l[a] = W.Bi[a];
    PFA_SYNC();
    int64_t *V = W.W1;                                 // (d, d); row 0 = w, rows 1.. = K
    rc = pfa_unimodular(T, l, d, V, W.flag);
    if (rc == -1) rc = pfa_unimodular_reduced(T, l, d, V, W.hv, W.hw, W.hpiv, W.flag);
    if (rc) return rc * 10 - 4;
    PFA_SYNC();
    if ((rc = pfa_lll_form(T, W, V + d, d - 1, d, W.mat, V))) return rc * 10 - 4;
    int64_t *Tm = W.W2;                                // T[i][a]: column a = K_a, last = w
    for (int e = T.lane; e < d * d; e += T.n) {
        int i = e / d, a = e % d;
        Tm[e] = a < d - 1 ? V[d + a * d + i] : V[i];
    }
    PFA_SYNC();
    if (pfa_matmul(T, W, W.Bi, Tm, W.W3, h, d, d)) return -14;
    pfa_copy(T, W.Bi, W.W3, h * d);
    if (pfa_matmul(T, W, W.ZB, Tm, W.W3, h, d, d)) return -14;
    pfa_copy(T, W.ZB, W.W3, h * d);
    if (pfa_matmul(T, W, W.mat, Tm, W.W3, d, d, d)) return -14;   // MT = mat T
    if (pfa_int64_safe(pfa_team_absmax(T, Tm, d * d), pfa_team_absmax(T, W.W3, d * d), d)) {
        for (int e = T.lane; e < d * d; e += T.n) {
            const int a = e / d, b = e % d;
            int64_t s = 0;
            for (int i = 0; i < d; ++i) s += Tm[i * d + a] * W.W3[i * d + b];
            W.W1[e] = s;
        }
    } else {
        for (int e = T.lane; e < d * d; e += T.n) {
            int a = e / d, b = e % d;
            pfa_i128 s = 0;
            for (int i = 0; i < d; ++i)
                if (pfa_add_ovf(s, (pfa_i128)Tm[i * d + a] * W.W3[i * d + b], &s)) bad = 1;
            if (!pfa_fits64(s)) bad = 1;
            W.W1[e] = (int64_t)s;
        }
    }
    if (pfa_team_any(T, W.flag, bad)) return -14;
    pfa_copy(T, W.mat, W.W1, d * d);
    return 0;
}

// -------------------------------------------------- FP64 Cholesky + phi
// U (d x d, upper) with U^T U = mat + E in FP64.  By Higham, Accuracy and
// Stability of Numerical Algorithms, Thm 10.3, |E| <= gamma_{d+1} |U|^T |U|
// for any summation order.  Then for integer c,
//   |c^T E c| <= gamma_{d+1} | |U| |c| |^2 <= gamma_{d+1} ||U||_F^2 ||U^-1||_2^2 |U c|^2,
// and ||U^-1||_2 <= ||X||_F / (1 - rho) for X ~ U^-1 with
// rho >= ||I - U X||_F (Neumann series), rho from the computed residual
// plus its own rounding bound gamma_{d+1} || |U| |X| ||_F.  So
//   (1 - phi) |U c|^2 <= q(c) <= (1 + phi) |U c|^2,
//   phi = gamma_{d+1} ||U||_F^2 ||X||_F^2 / (1 - rho)^2   (with 1% slack).
// Norm sums run on one lane in a fixed order (all lanes compute them), so
// the host and device get the same phi.  Returns 0, PFA_NOT_PD if the
// factorization broke down (the exact host path then decides), or
// PFA_HOST if phi is too large to be useful or mat is not exact in FP64.
PFA_HD static inline int pfa_cholesky(const PfaTeam &T, const int64_t *mat, int d, double *U,
                                      double *X, double *phi_out, int *flag)
{
    int bad = 0;
    // the factor and U^-1 (d^3 / 6 multiply-adds each), the residual of U X
    // (d^3 / 6 terms of 5) and the norms
    if (T.lane == 0) PFA_OPS(1.5 * d * d * d + 2.0 * d * d);
    for (int e = T.lane; e < d * d; e += T.n) {
        int64_t m = mat[e];
        if (m > (INT64_C(1) << 53) || m < -(INT64_C(1) << 53)) bad = 1;
        U[e] = 0.0;
    }
    pfa_team_clear(T, flag);
    if (pfa_team_any(T, flag, bad)) return PFA_HOST;
    for (int k = 0; k < d; ++k) {
        double s = (double)mat[k * d + k];
        for (int l = 0; l < k; ++l) s -= U[l * d + k] * U[l * d + k];
        if (!(s > 0.0)) return PFA_NOT_PD;
        double ukk = sqrt(s);
        PFA_SYNC();
        if (T.lane == 0) U[k * d + k] = ukk;
        for (int j = k + 1 + T.lane; j < d; j += T.n) {
            double t = (double)mat[k * d + j];
            for (int l = 0; l < k; ++l) t -= U[l * d + k] * U[l * d + j];
            U[k * d + j] = t / ukk;
        }
        PFA_SYNC();
    }
    // X = U^-1 (upper), column j by back substitution
    for (int j = T.lane; j < d; j += T.n) {
        for (int i = d - 1; i >= 0; --i) {
            if (i > j) { X[i * d + j] = 0.0; continue; }
            double t = i == j ? 1.0 : 0.0;
            for (int l = i + 1; l <= j; ++l) t -= U[i * d + l] * X[l * d + j];
            X[i * d + j] = t / U[i * d + i];
        }
    }
    PFA_SYNC();
    const double u = 0x1p-53;
    const double g = (d + 1) * u / (1.0 - (d + 1) * u);
    double nU = 0.0, nX = 0.0, nR = 0.0, nA = 0.0;
    for (int i = 0; i < d; ++i)
        for (int j = i; j < d; ++j) {
            nU += U[i * d + j] * U[i * d + j];
            nX += X[i * d + j] * X[i * d + j];
            double r = (i == j) ? 1.0 : 0.0, a = 0.0;
            for (int l = i; l <= j; ++l) {
                r -= U[i * d + l] * X[l * d + j];
                a += fabs(U[i * d + l] * X[l * d + j]);
            }
            nR += r * r;
            nA += a * a;
        }
    // strictly-lower entries of U X are exactly 0 (both upper triangular)
    double rho = (sqrt(nR) + g * sqrt(nA)) * 1.01 + 1e-300;
    if (!(rho < 0.5)) return PFA_HOST;
    double phi = g * nU * nX / ((1.0 - rho) * (1.0 - rho)) * 1.01;
    if (!(phi < 0.25)) return PFA_HOST;
    *phi_out = phi;
    return 0;
}
