// pfa_bound.h -- the lattice minimum behind coniZp's dilation bound, on
// host threads (pfvs/dilation.py:1-30 and 152-231; the reference's C form
// is pfvs/fp_kernel/coni_bound.h:255-399).
//
// For a direction p_hat = (0, p_r) (p_r has m = h11 - 1 entries):
//   A = kappa . p_hat,   S = -A_rr^-1,   s = a00 - a_r0^T A_rr^-1 a_r0,
//   Lambda = {K in Z^m : p_r . K = 0}   (rank n = m - 1),
//   mu0 = min over nonzero K in Lambda of K^T S K.
// A coni PFV of this direction with tadpole Q and dilation delta has
// delta < Q / mu0 whenever s <= 0 and S is positive definite on Lambda.
//
// Exact throughout: a basis O of Lambda (rows of the unimodular completion
// of p_r); a fraction-free elimination of [A_rr | O^T | a_r0], giving
// d = det and Y = d A_rr^-1 [O^T | a_r0] in integers; the sign of -s; the
// integer Gram matrix G of S on Lambda with S|Lambda = G / L; positive
// definiteness by Sylvester's criterion (the leading minors, fraction free);
// an LLL reduction of G (floating point chooses the steps, the transformed
// Gram matrix and the transform stay exact); then every y with y^T Gr y at
// most the smallest diagonal entry, by the library's Fincke-Pohst search
// (the caller runs it), and the smallest nonzero value.
//
// Status: 0 decided; 1 no bound (A_rr singular, s > 0, S not positive
// definite on Lambda, or rank 0); < 0 not decided here (a value beyond
// int128 / int64, no LLL convergence): the caller's exact path decides.
#pragma once
#include <math.h>
#include <stdint.h>
#include "pfa_int.h"
#include "pfa_prep.h"

#define PFA_BOUND_MAXM PFA_DMAX          // m = h11 - 1

struct PfaBoundForm {
    int m, n;
    int64_t O[PFA_BOUND_MAXM * PFA_BOUND_MAXM];    // (n, m): rows span Lambda
    int64_t G[PFA_BOUND_MAXM * PFA_BOUND_MAXM];    // (n, n): S on Lambda = G / L
    int64_t V[PFA_BOUND_MAXM * PFA_BOUND_MAXM];    // (n, n): rows = reduced vectors
    int64_t Gr[PFA_BOUND_MAXM * PFA_BOUND_MAXM];   // (n, n): V G V^T (LLL-reduced)
    int64_t L;
    pfa_i128 sn, sd;                                // -s = sn / sd, sd > 0, in lowest terms
};

// a * b + c, 1 on overflow
static inline int pfa_bnd_muladd(pfa_i128 a, pfa_i128 b, pfa_i128 c, pfa_i128 *r)
{
    pfa_i128 t;
    return pfa_mul_ovf(a, b, &t) || pfa_add_ovf(t, c, r);
}

// Fraction-free elimination (Bareiss) of the m x w system M (row-major,
// int128), with row swaps; then back substitution of each column c >= m:
// Y[:, c - m] = det * M_A^-1 M[:, c].  *det = the product of the pivots'
// last leading minor (+- det A, consistent with Y).  Returns 0, 1 singular,
// -1 overflow.
static inline int pfa_bnd_solve(pfa_i128 *M, int m, int w, pfa_i128 *det, pfa_i128 *Y)
{
    pfa_i128 prev = 1;
    for (int k = 0; k < m; ++k) {
        int piv = k;
        while (piv < m && M[piv * w + k] == 0) ++piv;
        if (piv == m) return 1;
        if (piv != k)
            for (int j = 0; j < w; ++j) {
                const pfa_i128 t = M[k * w + j];
                M[k * w + j] = M[piv * w + j];
                M[piv * w + j] = t;
            }
        for (int i = k + 1; i < m; ++i) {
            for (int j = k + 1; j < w; ++j) {
                pfa_i128 t1, t2;
                if (pfa_mul_ovf(M[k * w + k], M[i * w + j], &t1) ||
                    pfa_mul_ovf(M[i * w + k], M[k * w + j], &t2) || pfa_sub_ovf(t1, t2, &t1))
                    return -1;
                M[i * w + j] = pfa_tdiv128(t1, prev);       // exact
            }
            M[i * w + k] = 0;
        }
        prev = M[k * w + k];
    }
    const pfa_i128 d = M[(m - 1) * w + (m - 1)];
    *det = d;
    const int nc = w - m;
    for (int c = 0; c < nc; ++c)
        for (int i = m - 1; i >= 0; --i) {
            pfa_i128 s;
            if (pfa_mul_ovf(d, M[i * w + m + c], &s)) return -1;
            for (int j = i + 1; j < m; ++j) {
                pfa_i128 t;
                if (pfa_mul_ovf(M[i * w + j], Y[j * nc + c], &t) || pfa_sub_ovf(s, t, &s)) return -1;
            }
            const pfa_i128 q = pfa_tdiv128(s, M[i * w + i]);
            if (q * M[i * w + i] != s) return -1;                // cannot happen (Cramer)
            Y[i * nc + c] = q;
        }
    return 0;
}

// 1 if the symmetric n x n int64 matrix G is positive definite (every
// leading minor > 0, by fraction-free elimination), 0 if not, -1 overflow
static inline int pfa_bnd_pd(const int64_t *G, int n)
{
    pfa_i128 M[PFA_BOUND_MAXM * PFA_BOUND_MAXM];
    for (int e = 0; e < n * n; ++e) M[e] = G[e];
    pfa_i128 prev = 1;
    for (int k = 0; k < n; ++k) {
        if (M[k * n + k] <= 0) return 0;                          // the k-th leading minor
        for (int i = k + 1; i < n; ++i) {
            for (int j = k + 1; j < n; ++j) {
                pfa_i128 t1, t2;
                if (pfa_mul_ovf(M[k * n + k], M[i * n + j], &t1) ||
                    pfa_mul_ovf(M[i * n + k], M[k * n + j], &t2) || pfa_sub_ovf(t1, t2, &t1))
                    return -1;
                M[i * n + j] = pfa_tdiv128(t1, prev);
            }
        }
        prev = M[k * n + k];
    }
    return 1;
}

// LLL (delta 0.99) of the rows of V (identity on entry) for the positive
// definite form G: Gr = V G V^T updated exactly by every row operation;
// the Gram-Schmidt data, recomputed in double from Gr, only choose the
// operations.  Returns 0, -1 overflow, -2 no convergence.
static inline int pfa_bnd_lll(const int64_t *G, int n, int64_t *V, int64_t *Gr_out)
{
    pfa_i128 Gr[PFA_BOUND_MAXM * PFA_BOUND_MAXM];
    double mu[PFA_BOUND_MAXM][PFA_BOUND_MAXM], bb[PFA_BOUND_MAXM];
    for (int e = 0; e < n * n; ++e) {
        V[e] = (e / n) == (e % n);
        Gr[e] = G[e];
    }
    auto gs_row = [&](int k) {
        for (int j = 0; j < k; ++j) {
            double t = (double)Gr[k * n + j];
            for (int l = 0; l < j; ++l) t -= mu[j][l] * mu[k][l] * bb[l];
            mu[k][j] = t / bb[j];
        }
        double t = (double)Gr[k * n + k];
        for (int l = 0; l < k; ++l) t -= mu[k][l] * mu[k][l] * bb[l];
        bb[k] = t;
    };
    gs_row(0);
    if (!(bb[0] > 0)) return -2;
    if (n > 1) gs_row(1);
    int k = 1;
    long iters = 0;
    while (k < n) {
        if (++iters > 200000) return -2;
        for (int j = k - 1; j >= 0; --j) {                        // size reduction of row k
            const double qd = nearbyint(mu[k][j]);
            if (qd == 0.0) continue;
            if (!(fabs(qd) < 1e15)) return -1;
            const int64_t q = (int64_t)qd;
            for (int c = 0; c < n; ++c) {
                pfa_i128 t = (pfa_i128)V[k * n + c] - (pfa_i128)q * V[j * n + c];
                if (!pfa_fits64(t)) return -1;
                V[k * n + c] = (int64_t)t;
            }
            // Gr: row and column k of (b_k - q b_j)
            pfa_i128 gkj = Gr[k * n + j], gjj = Gr[j * n + j], nk;
            pfa_i128 t1, t2;
            if (pfa_mul_ovf((pfa_i128)q, gjj, &t1) || pfa_mul_ovf((pfa_i128)q, t1, &t1) ||
                pfa_mul_ovf((pfa_i128)(2 * q), gkj, &t2) || pfa_sub_ovf(Gr[k * n + k], t2, &nk) ||
                pfa_add_ovf(nk, t1, &nk))
                return -1;
            for (int l = 0; l < n; ++l) {
                if (l == k) continue;
                pfa_i128 t;
                if (pfa_mul_ovf((pfa_i128)q, Gr[j * n + l], &t) || pfa_sub_ovf(Gr[k * n + l], t, &t)) return -1;
                Gr[k * n + l] = Gr[l * n + k] = t;
            }
            Gr[k * n + k] = nk;
            for (int l = 0; l < j; ++l) mu[k][l] -= qd * mu[j][l];
            mu[k][j] -= qd;
        }
        gs_row(k);
        if (!(bb[k] > 0)) return -2;
        if (bb[k] < (0.99 - mu[k][k - 1] * mu[k][k - 1]) * bb[k - 1]) {
            for (int c = 0; c < n; ++c) {                          // swap b_k, b_{k-1}
                const int64_t t = V[k * n + c];
                V[k * n + c] = V[(k - 1) * n + c];
                V[(k - 1) * n + c] = t;
            }
            for (int l = 0; l < n; ++l) {
                const pfa_i128 t = Gr[k * n + l];
                Gr[k * n + l] = Gr[(k - 1) * n + l];
                Gr[(k - 1) * n + l] = t;
            }
            for (int l = 0; l < n; ++l) {
                const pfa_i128 t = Gr[l * n + k];
                Gr[l * n + k] = Gr[l * n + k - 1];
                Gr[l * n + k - 1] = t;
            }
            k = k > 1 ? k - 1 : 1;
            if (k == 1) gs_row(0);
            gs_row(k - 1);
            gs_row(k);
        } else if (++k < n) {
            gs_row(k);
        }
    }
    for (int e = 0; e < n * n; ++e) {
        if (!pfa_fits64(Gr[e])) return -1;
        Gr_out[e] = (int64_t)Gr[e];
    }
    return 0;
}

// The form of direction p_r (m entries) for kappa (h = m + 1): F.O, F.G,
// F.L, then LLL into F.V, F.Gr.  Status as above.
static inline int pfa_bound_form(int h, const int64_t *kappa, const int64_t *p_r, PfaBoundForm &F)
{
    const int m = h - 1, n = m - 1;
    F.m = m;
    F.n = n;
    if (h < 3 || m > PFA_BOUND_MAXM) return n < 1 ? 1 : -3;
    // A = kappa . p_hat (p_hat_0 = 0)
    int64_t A[(PFA_BOUND_MAXM + 1) * (PFA_BOUND_MAXM + 1)];
    for (int a = 0; a < h; ++a)
        for (int b = 0; b < h; ++b) {
            pfa_i128 s = 0;
            for (int c = 1; c < h; ++c)
                if (pfa_bnd_muladd((pfa_i128)kappa[((size_t)a * h + b) * h + c], (pfa_i128)p_r[c - 1], s, &s))
                    return -1;
            if (!pfa_fits64(s)) return -1;
            A[a * h + b] = (int64_t)s;
        }
    // Lambda: rows 1 .. m-1 of U with U p_r = (g, 0, ..., 0), U unimodular
    {
        int64_t U[PFA_BOUND_MAXM * PFA_BOUND_MAXM];
        int flag = 0;
        const PfaTeam T{0, 1};
        int rc = pfa_unimodular(T, p_r, m, U, &flag);
        if (rc == -1) {
            pfa_i128 x[PFA_BOUND_MAXM], r[PFA_BOUND_MAXM];
            int ord[PFA_BOUND_MAXM];
            rc = pfa_unimodular_reduced(T, p_r, m, U, x, r, ord, &flag);
        }
        if (rc) return rc == -2 ? 1 : -1;
        for (int a = 0; a < n; ++a)
            for (int i = 0; i < m; ++i) F.O[a * m + i] = U[(a + 1) * m + i];
    }
    // [A_rr | O^T | a_r0] -> det, Y = det A_rr^-1 [O^T | a_r0]
    const int w = m + n + 1;
    pfa_i128 M[PFA_BOUND_MAXM * (3 * PFA_BOUND_MAXM)], Y[PFA_BOUND_MAXM * (2 * PFA_BOUND_MAXM)];
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) M[i * w + j] = A[(i + 1) * h + j + 1];
        for (int a = 0; a < n; ++a) M[i * w + m + a] = F.O[a * m + i];
        M[i * w + m + n] = A[(i + 1) * h];
    }
    pfa_i128 d;
    int rc = pfa_bnd_solve(M, m, w, &d, Y);
    if (rc) return rc;                                          // 1 singular, -1 overflow
    const int nc = n + 1;
    // -s = (a_r0^T y - d a00) / d must be >= 0
    pfa_i128 sn;
    if (pfa_mul_ovf(-d, (pfa_i128)A[0], &sn)) return -1;
    for (int i = 0; i < m; ++i)
        if (pfa_bnd_muladd((pfa_i128)A[(i + 1) * h], Y[i * nc + n], sn, &sn)) return -1;
    if (sn != 0 && ((sn < 0) != (d < 0))) return 1;            // s > 0
    {
        pfa_i128 a = d < 0 ? -sn : sn, b = d < 0 ? -d : d;
        const pfa_u128 gs = pfa_gcd128(pfa_abs128(a), (pfa_u128)b);
        F.sn = a / (pfa_i128)gs;
        F.sd = b / (pfa_i128)gs;
    }
    // S on Lambda = Gn / d, Gn = -O Y[:, :n]; reduced by g = gcd(Gn, d)
    pfa_i128 Gn[PFA_BOUND_MAXM * PFA_BOUND_MAXM];
    pfa_u128 g = pfa_abs128(d);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            pfa_i128 s = 0;
            for (int i = 0; i < m; ++i) {
                if (!F.O[a * m + i]) continue;
                pfa_i128 t;
                if (pfa_mul_ovf((pfa_i128)F.O[a * m + i], Y[i * nc + b], &t) || pfa_sub_ovf(s, t, &s)) return -1;
            }
            Gn[a * n + b] = s;
            g = pfa_gcd128(g, pfa_abs128(s));
        }
    for (int e = 0; e < n * n; ++e) {
        const pfa_i128 v = (Gn[e] / (pfa_i128)g) * (d < 0 ? -1 : 1);
        if (!pfa_fits64(v)) return -1;
        F.G[e] = (int64_t)v;
    }
    const pfa_i128 L = pfa_abs128(d) / g;
    if (!pfa_fits64(L)) return -1;
    F.L = (int64_t)L;
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < a; ++b)
            if (F.G[a * n + b] != F.G[b * n + a]) return -4;    // cannot happen (S symmetric)
    const int pd = pfa_bnd_pd(F.G, n);
    if (pd == 0) return 1;
    rc = pfa_bnd_lll(F.G, n, F.V, F.Gr);
    if (pd > 0) return rc ? (rc == -2 ? -2 : -1) : 0;
    // the minors of the unreduced form left int128: decide on the reduced
    // form Gr = V G V^T instead (V unimodular: the same definiteness); a
    // reduction that met a nonpositive Gram-Schmidt value decides nothing.
    // When the reduced form's minors leave int128 too (L large), the
    // callers' FP64 factor of Gr decides: its proven margin phi < 1 gives
    // c^T Gr c >= (1 - phi) |U c|^2 > 0 for c != 0, and a factor that fails
    // leaves the direction undecided (never "no bound").
    if (rc) return -1;
    const int pd2 = pfa_bnd_pd(F.Gr, n);
    return pd2 == 0 ? 1 : 0;
}

// K = O^T V^T y for a vector y in the reduced coordinates; 0 or -1 overflow
static inline int pfa_bound_K(const PfaBoundForm &F, const int32_t *y, int64_t *K)
{
    const int m = F.m, n = F.n;
    pfa_i128 z[PFA_BOUND_MAXM];
    for (int a = 0; a < n; ++a) {
        pfa_i128 s = 0;
        for (int b = 0; b < n; ++b)
            if (pfa_bnd_muladd((pfa_i128)F.V[b * n + a], (pfa_i128)y[b], s, &s)) return -1;
        z[a] = s;
    }
    for (int i = 0; i < m; ++i) {
        pfa_i128 s = 0;
        for (int a = 0; a < n; ++a)
            if (F.O[a * m + i] && pfa_bnd_muladd((pfa_i128)F.O[a * m + i], z[a], s, &s)) return -1;
        if (!pfa_fits64(s)) return -1;
        K[i] = (int64_t)s;
    }
    return 0;
}
