/*
 * pfvs_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain-C, single-threaded restatement of the reference's CPU coniZpM
 * path, used by tests/ to check the HIP product path and by bench.py as the
 * timed CPU baseline.  Nothing in the product links or calls this file.
 *
 * It follows, step by step:
 *   - per-p setup  pfvs/fp_kernel/pfv_lattice.h:604-680 (Z = kappa.p, the
 *     orthogonal lattice 151-188, LLL 393-550, cut-aware basis 554-602,
 *     row-insertion HNF 250-313);
 *   - exact factor pfvs/fp_kernel/fpk_common.h:232-309 (int128 Bareiss
 *     LDL^T rounded to U, M0-bound data m0N, m0V, m0Va);
 *   - the search   pfvs/fp_kernel/fpk_search_impl.h:217-455 in its double
 *     instantiation (fpk_common.h:345-357), without the sparse-candidate
 *     lists (which skip candidates but never change the nodes entered);
 *   - the batch    pfvs/fp_kernel/fp_kernel.pyx:480-622 (M = Binter c,
 *     Kn = ZB c in int128, status 1 on overflow).
 * H rows must fit int64 here (the reference's fast path,
 * fp_kernel.h:510-519); otherwise the p-vector gets status 1.
 */
#include <gmp.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

#define OMAX 64

/* ---------------------------------------------------------------- ints */
static int oadd(i128 a, i128 b, i128 *r) { return __builtin_add_overflow(a, b, r); }
static int osub(i128 a, i128 b, i128 *r) { return __builtin_sub_overflow(a, b, r); }
static u128 oabs(i128 x) { return x < 0 ? (u128)0 - (u128)x : (u128)x; }
/* checked int128 product (fpk_common.h:69-78) */
static int omul(i128 a, i128 b, i128 *r)
{
    u128 ua = oabs(a), ub = oabs(b);
    const u128 lim = (u128)1 << 63;
    if (ua < lim && ub < lim) { *r = a * b; return 0; }
    if (ua && ub > ((((u128)1) << 127) - 1) / ua) return 1;
    u128 m = ua * ub;
    *r = ((a < 0) != (b < 0)) ? -(i128)m : (i128)m;
    return 0;
}
static int fits64(i128 x) { return x <= (i128)INT64_MAX && x > (i128)INT64_MIN; }
static int ctz64(uint64_t x) { return __builtin_ctzll(x); }
static int ctz128(u128 x) { uint64_t lo = (uint64_t)x; return lo ? ctz64(lo) : 64 + ctz64((uint64_t)(x >> 64)); }

static uint64_t gcd64(uint64_t u, uint64_t v)        /* fpk_common.h:96-109 */
{
    if (!u) return v;
    if (!v) return u;
    int sh = ctz64(u | v);
    u >>= ctz64(u);
    do {
        v >>= ctz64(v);
        if (u > v) { uint64_t t = u; u = v; v = t; }
        v -= u;
    } while (v);
    return u << sh;
}
/* gcd if >= need else 0 (fpk_common.h:113-128) */
static uint64_t gcd64_ge(uint64_t u, uint64_t v, uint64_t need)
{
    if (!u) return v >= need ? v : 0;
    if (!v) return u >= need ? u : 0;
    int sh = ctz64(u | v);
    uint64_t ns = (need >> sh) + ((need & ((1ULL << sh) - 1)) != 0);
    u >>= ctz64(u);
    do {
        v >>= ctz64(v);
        if (u > v) { uint64_t t = u; u = v; v = t; }
        if (u < ns) return 0;
        v -= u;
    } while (v);
    return u << sh;
}
static u128 gcd128(u128 u, u128 v)                   /* fpk_common.h:131-147 */
{
    if ((u >> 64) == 0 && (v >> 64) == 0) return gcd64((uint64_t)u, (uint64_t)v);
    if (!u) return v;
    if (!v) return u;
    int sh = ctz128(u | v);
    u >>= ctz128(u);
    do {
        v >>= ctz128(v);
        if (u > v) { u128 t = u; u = v; v = t; }
        v -= u;
        if ((u >> 64) == 0 && (v >> 64) == 0) return (u128)gcd64((uint64_t)u, (uint64_t)v) << sh;
    } while (v);
    return u << sh;
}

/* ------------------------------------------------------------- lattice */
static int64_t fdiv64(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}
static int64_t xgcd64(int64_t a, int64_t b, int64_t *s, int64_t *t)
{   /* floor-division extended Euclid (pfv_lattice.h:116-130) */
    int64_t r0 = a, r1 = b, s0 = 1, s1 = 0, t0 = 0, t1 = 1;
    while (r1) {
        int64_t q = fdiv64(r0, r1), x;
        x = r0 - q * r1; r0 = r1; r1 = x;
        x = s0 - q * s1; s0 = s1; s1 = x;
        x = t0 - q * t1; t0 = t1; t1 = x;
    }
    *s = s0; *t = t0;
    return r0;
}
/* U (n x n) unimodular with U v = (g, 0, ..) (pfv_lattice.h:162-188) */
static int unimod(const int64_t *v, int n, int64_t *U)
{
    int64_t w[OMAX];
    for (int i = 0; i < n; ++i) {
        w[i] = v[i];
        for (int j = 0; j < n; ++j) U[i * n + j] = (i == j);
    }
    for (int k = 1; k < n; ++k) {
        if (!w[k]) continue;
        int64_t s, t, g = xgcd64(w[0], w[k], &s, &t);
        if (!g) return -2;
        int64_t m10 = -fdiv64(w[k], g), m11 = fdiv64(w[0], g);
        w[0] = g; w[k] = 0;
        for (int r = 0; r < n; ++r) {
            i128 a = (i128)s * U[r] + (i128)t * U[k * n + r];
            i128 b = (i128)m10 * U[r] + (i128)m11 * U[k * n + r];
            if (!fits64(a) || !fits64(b)) return -1;
            U[r] = (int64_t)a; U[k * n + r] = (int64_t)b;
        }
    }
    return 0;
}

/* LLL (delta 0.99) of the m rows of B (length n) with exact Gram GG (stride
 * m+1); optional row `last` (index m) size-reduced afterwards.
 * pfv_lattice.h:393-505. */
static int lll_core(int64_t *B, int m, int n, i128 *GG, int64_t *last)
{
    const int L = m + 1, M = m + (last != NULL);
    static __thread double mu[OMAX][OMAX], bb[OMAX], mw[OMAX];
    int iters = 0;
#define ROW(kk) ((kk) == m ? last : &B[(kk) * n])
#define SUBROW(kk, jj, qi) do {                                              \
        int64_t *bk = ROW(kk); const int64_t *bj = ROW(jj);                  \
        for (int c = 0; c < n; ++c) {                                        \
            i128 x = (i128)bk[c] - (i128)(qi) * bj[c];                       \
            if (!fits64(x)) return -1;                                       \
            bk[c] = (int64_t)x;                                              \
        }                                                                    \
        i128 kj = GG[(kk) * L + (jj)], jjv = GG[(jj) * L + (jj)], t1, t2;    \
        if (omul(jjv, (i128)(qi), &t1) || osub(t1, 2 * kj, &t1) ||           \
            omul(t1, (i128)(qi), &t2) || oadd(GG[(kk) * L + (kk)], t2, &t2)) \
            return -1;                                                       \
        for (int l = 0; l < M; ++l) {                                        \
            if (l == (kk)) continue;                                         \
            i128 t3;                                                         \
            if (omul(GG[(jj) * L + l], (i128)(qi), &t3) ||                   \
                osub(GG[(kk) * L + l], t3, &t3)) return -1;                  \
            GG[(kk) * L + l] = t3; GG[l * L + (kk)] = t3;                    \
        }                                                                    \
        GG[(kk) * L + (kk)] = t2;                                            \
    } while (0)
#define GS(kk) do {                                                          \
        for (int j = 0; j < (kk); ++j) {                                     \
            double d = (double)GG[(kk) * L + j];                             \
            for (int l = 0; l < j; ++l) d -= mu[j][l] * mu[(kk)][l] * bb[l]; \
            mu[(kk)][j] = d / bb[j];                                         \
        }                                                                    \
        double d = (double)GG[(kk) * L + (kk)];                              \
        for (int l = 0; l < (kk); ++l) d -= mu[(kk)][l] * mu[(kk)][l] * bb[l]; \
        bb[(kk)] = d;                                                        \
    } while (0)
    if (m >= 1) { GS(0); if (!(bb[0] > 0)) return -2; }
    if (m >= 2) GS(1);
    int k = 1;
    while (k < m) {
        if (++iters > 100000) return -3;
        for (int pass = 0; pass < 4; ++pass) {
            int changed = 0;
            for (int j = k - 1; j >= 0; --j) {
                double q = nearbyint(mu[k][j]);
                if (q == 0.0) continue;
                if (!(fabs(q) <= 9.0e18)) return -1;
                int64_t qi = (int64_t)q;
                SUBROW(k, j, qi);
                for (int l = 0; l < j; ++l) mu[k][l] -= q * mu[j][l];
                mu[k][j] -= q;
                changed = 1;
            }
            if (!changed) break;
            GS(k);
        }
        if (!(bb[k] > 0)) return -2;
        if (bb[k] < (0.99 - mu[k][k - 1] * mu[k][k - 1]) * bb[k - 1]) {
            for (int c = 0; c < n; ++c) {
                int64_t t = B[k * n + c]; B[k * n + c] = B[(k - 1) * n + c]; B[(k - 1) * n + c] = t;
            }
            for (int l = 0; l < M; ++l) { i128 t = GG[k * L + l]; GG[k * L + l] = GG[(k - 1) * L + l]; GG[(k - 1) * L + l] = t; }
            for (int l = 0; l < M; ++l) { i128 t = GG[l * L + k]; GG[l * L + k] = GG[l * L + k - 1]; GG[l * L + k - 1] = t; }
            k = k > 1 ? k - 1 : 1;
            if (k == 1) GS(0);
            GS(k - 1);
            GS(k);
        } else {
            k++;
            if (k < m) GS(k);
        }
    }
    if (last && m > 0) {
        for (int pass = 0; pass < 4; ++pass) {
            for (int j = 0; j < m; ++j) {
                double d = (double)GG[m * L + j];
                for (int l = 0; l < j; ++l) d -= mu[j][l] * mw[l] * bb[l];
                mw[j] = d / bb[j];
            }
            int changed = 0;
            for (int j = m - 1; j >= 0; --j) {
                double q = nearbyint(mw[j]);
                if (q == 0.0) continue;
                if (!(fabs(q) <= 9.0e18)) return -1;
                int64_t qi = (int64_t)q;
                int rl = m;
                SUBROW(rl, j, qi);
                for (int l = 0; l < j; ++l) mw[l] -= q * mu[j][l];
                mw[j] -= q;
                changed = 1;
            }
            if (!changed) break;
        }
    }
#undef ROW
#undef SUBROW
#undef GS
    return 0;
}

static int lll_euclid(int64_t *B, int m, int n)       /* pfv_lattice.h:508-521 */
{
    if (m < 2) return 0;
    static __thread i128 GG[(OMAX + 1) * (OMAX + 1)];
    const int L = m + 1;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) {
            i128 s = 0;
            for (int c = 0; c < n; ++c)
                if (oadd(s, (i128)B[i * n + c] * B[j * n + c], &s)) return -1;
            GG[i * L + j] = GG[j * L + i] = s;
        }
    return lll_core(B, m, n, GG, NULL);
}

static int lll_form(int64_t *B, int m, int n, const int64_t *G, int64_t *last)
{                                                     /* pfv_lattice.h:525-550 */
    static __thread i128 GG[(OMAX + 1) * (OMAX + 1)];
    i128 Gy[OMAX];
    const int L = m + 1, M = m + (last != NULL);
    for (int j = 0; j < M; ++j) {
        const int64_t *y = j == m ? last : &B[j * n];
        for (int i = 0; i < n; ++i) {
            i128 r = 0;
            for (int c = 0; c < n; ++c)
                if (oadd(r, (i128)G[i * n + c] * y[c], &r)) return -1;
            Gy[i] = r;
        }
        for (int i = 0; i <= j; ++i) {
            const int64_t *x = i == m ? last : &B[i * n];
            i128 s = 0, t;
            for (int c = 0; c < n; ++c) {
                if (!x[c]) continue;
                if (omul(Gy[c], (i128)x[c], &t) || oadd(s, t, &s)) return -1;
            }
            GG[i * L + j] = GG[j * L + i] = s;
        }
    }
    return lll_core(B, m, n, GG, last);
}

static int matmul64(const int64_t *A, const int64_t *B, int64_t *C, int m, int k, int n)
{
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            i128 s = 0;
            for (int l = 0; l < k; ++l)
                if (oadd(s, (i128)A[i * k + l] * B[l * n + j], &s)) return -1;
            if (!fits64(s)) return -1;
            C[i * n + j] = (int64_t)s;
        }
    return 0;
}

/* -------- HNF: rows inserted one at a time (pfv_lattice.h:205-313) ----- */
#define HLIM ((i128)1 << 125)
static int fit125(i128 x) { return x < HLIM && x > -HLIM; }
static i128 fdiv128(i128 a, i128 b)
{
    i128 q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}
static int muladd(i128 a, i128 b, i128 c, i128 *r)
{
    i128 x;
    if (omul(a, b, &x) || oadd(x, c, &x) || !fit125(x)) return -1;
    *r = x;
    return 0;
}
static int rowsub(i128 *row, const i128 *prow, i128 q, int c0, int n)
{
    if (!q) return 0;
    for (int c = c0; c < n; ++c) if (muladd(-q, prow[c], row[c], &row[c])) return -1;
    return 0;
}
static int reduce_by(i128 *row, const i128 *H, const int *piv, int t0, int m, int n)
{
    for (int l = t0; l < m; ++l) {
        int j = piv[l];
        if (rowsub(row, &H[l * n], fdiv128(row[j], H[l * n + j]), j, n)) return -1;
    }
    return 0;
}
static int hnf_insert(const int64_t *A, int r, int n, i128 *H)
{
    int piv[OMAX];
    i128 v[OMAX], w[OMAX];
    int m = 0;
    for (int k = 0; k < r; ++k) {
        for (int c = 0; c < n; ++c) v[c] = A[k * n + c];
        int t = 0;
        for (;;) {
            int j = 0;
            while (j < n && v[j] == 0) ++j;
            if (j == n) break;
            while (t < m && piv[t] < j) ++t;
            if (t == m || piv[t] > j) {
                for (int l = m; l > t; --l) {
                    piv[l] = piv[l - 1];
                    for (int c = 0; c < n; ++c) H[l * n + c] = H[(l - 1) * n + c];
                }
                m++;
                if (v[j] < 0) for (int c = j; c < n; ++c) v[c] = -v[c];
                for (int c = 0; c < n; ++c) H[t * n + c] = v[c];
                piv[t] = j;
                if (reduce_by(&H[t * n], H, piv, t + 1, m, n)) return -1;
                break;
            }
            i128 *hr = &H[t * n];
            i128 a = hr[j], b = v[j];
            if (b % a == 0) {
                if (rowsub(v, hr, b / a, j, n)) return -1;
            } else {
                i128 r0 = a, r1 = b, s0 = 1, s1 = 0, x0 = 0, x1 = 1;
                while (r1) {
                    i128 q = r0 / r1, tt;
                    tt = r0 - q * r1; r0 = r1; r1 = tt;
                    tt = s0 - q * s1; s0 = s1; s1 = tt;
                    tt = x0 - q * x1; x0 = x1; x1 = tt;
                }
                if (r0 < 0) { r0 = -r0; s0 = -s0; x0 = -x0; }
                i128 ag = a / r0, bg = b / r0;
                for (int c = j; c < n; ++c) {
                    i128 hn, vn, t1;
                    if (omul(s0, hr[c], &t1) || muladd(x0, v[c], t1, &hn)) return -1;
                    if (omul(-bg, hr[c], &t1) || muladd(ag, v[c], t1, &vn)) return -1;
                    w[c] = hn; v[c] = vn;
                }
                for (int c = j; c < n; ++c) hr[c] = w[c];
                if (reduce_by(hr, H, piv, t + 1, m, n)) return -1;
            }
            if (reduce_by(v, H, piv, t + 1, m, n)) return -1;
            for (int l = 0; l < t; ++l)
                if (reduce_by(&H[l * n], H, piv, t, m, n)) return -1;
            t++;
        }
        for (int l = 0; l < t && l < m; ++l)
            if (reduce_by(&H[l * n], H, piv, t < m ? t : m, m, n)) return -1;
    }
    for (int l = m; l < r; ++l) for (int c = 0; c < n; ++c) H[l * n + c] = 0;
    return 0;
}

/* per-p build (pfv_lattice.h:604-680), coni, extra LLL, cut-aware basis */
typedef struct {
    int64_t Z[OMAX * OMAX], Binter[OMAX * OMAX], ZB[OMAX * OMAX], mat[OMAX * OMAX];
    i128 H[OMAX * OMAX];
} or_lat;

static int cut_basis(or_lat *R, int h, int d)          /* pfv_lattice.h:554-602 */
{
    int64_t l[OMAX], w[OMAX];
    static __thread int64_t V[OMAX * OMAX], T[OMAX * OMAX], tmp[OMAX * OMAX], MT[OMAX * OMAX];
    int rc, nz = 0;
    for (int a = 0; a < d; ++a) { l[a] = R->Binter[a]; nz += l[a] != 0; }
    if (!nz) return 0;
    if ((rc = unimod(l, d, V))) return rc;
    for (int a = 0; a < d; ++a) w[a] = V[a];
    if ((rc = lll_form(&V[d], d - 1, d, R->mat, w))) return rc;
    for (int a = 0; a < d - 1; ++a) for (int i = 0; i < d; ++i) T[i * d + a] = V[d + a * d + i];
    for (int i = 0; i < d; ++i) T[i * d + d - 1] = w[i];
    if ((rc = matmul64(R->Binter, T, tmp, h, d, d))) return rc;
    memcpy(R->Binter, tmp, sizeof(int64_t) * h * d);
    if ((rc = matmul64(R->ZB, T, tmp, h, d, d))) return rc;
    memcpy(R->ZB, tmp, sizeof(int64_t) * h * d);
    if ((rc = matmul64(R->mat, T, MT, d, d, d))) return rc;
    for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) {
            i128 s = 0;
            for (int i = 0; i < d; ++i) if (oadd(s, (i128)T[i * d + a] * MT[i * d + b], &s)) return -1;
            if (!fits64(s)) return -1;
            tmp[a * d + b] = (int64_t)s;
        }
    memcpy(R->mat, tmp, sizeof(int64_t) * d * d);
    return 0;
}

int or_build(int h, const int64_t *kappa, const int64_t *Mbasis, const int64_t *p, or_lat *R)
{
    const int d = h - 1;
    if (h < 2 || h > OMAX) return -2;
    int64_t T[OMAX], v[OMAX];
    static __thread int64_t O[OMAX * OMAX], BT[OMAX * OMAX];
    int rc;
    for (int i = 0; i < h * h; ++i) {
        i128 s = 0;
        for (int k = 0; k < h; ++k)
            if (oadd(s, (i128)kappa[i * h + k] * p[k], &s)) return -1;
        if (!fits64(s)) return -1;
        R->Z[i] = (int64_t)s;
    }
    if ((rc = matmul64(R->Z, p, T, h, h, 1))) return rc;
    if ((rc = matmul64(T, Mbasis, v, 1, h, h))) return rc;
    {
        static __thread int64_t Uf[OMAX * OMAX];
        if ((rc = unimod(v, h, Uf))) return rc * 10 - 1;
        memcpy(O, &Uf[h], sizeof(int64_t) * d * h);
    }
    if ((rc = lll_euclid(O, d, h))) return rc * 10 - 2;
    for (int a = 0; a < d; ++a)
        for (int i = 0; i < h; ++i) {
            i128 s = 0;
            for (int l = 0; l < h; ++l)
                if (oadd(s, (i128)Mbasis[i * h + l] * O[a * h + l], &s)) return -1;
            if (!fits64(s)) return -1;
            BT[a * h + i] = (int64_t)s;
        }
    int order[OMAX], no = 0;
    for (int a = 0; a < d; ++a) if (BT[a * h] == 0) order[no++] = a;
    for (int a = 0; a < d; ++a) if (BT[a * h] != 0) order[no++] = a;
    for (int i = 0; i < h; ++i) for (int a = 0; a < d; ++a) R->Binter[i * d + a] = BT[order[a] * h + i];
    if ((rc = matmul64(R->Z, R->Binter, R->ZB, h, h, d))) return rc;
    for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) {
            i128 s = 0;
            for (int i = 0; i < h; ++i)
                if (oadd(s, (i128)R->Binter[i * d + a] * R->ZB[i * d + b], &s)) return -1;
            if (!fits64(s) || s == (i128)INT64_MIN) return -1;
            R->mat[a * d + b] = (int64_t)(-s);
        }
    if ((rc = cut_basis(R, h, d))) return rc * 10 - 4;
    rc = hnf_insert(&R->ZB[d], h - 1, d, R->H);
    return rc ? rc * 10 - 5 : 0;
}

/* ------------------------------------------------------ exact factor */
static u128 inv2k(u128 o) { u128 x = o; for (int i = 0; i < 6; ++i) x *= (u128)2 - o * x; return x; }
static int bdiv(i128 a, i128 b, i128 c, i128 dd, i128 e, i128 *q)
{   /* (a b - c d)/e, exact (fpk_common.h:240-257) */
    double ab = (double)a * (double)b, cd = (double)c * (double)dd, ed = (double)e;
    double qf = (ab - cd) / ed;
    double qerr = 1e-15 * (fabs(ab) + fabs(cd)) / ed + 2.0;
    int s = ctz128((u128)e);
    if (s > 100 || fabs(qf) + qerr >= ldexp(1.0, 125 - s)) return -1;
    u128 N = (u128)a * (u128)b - (u128)c * (u128)dd;
    u128 Q = (N >> s) * inv2k((u128)e >> s);
    if (s > 0) {
        u128 mask = ((u128)1 << (128 - s)) - 1;
        Q &= mask;
        if ((Q >> (127 - s)) & 1) Q |= ~mask;
    }
    *q = (i128)Q;
    return 0;
}
/* U (upper, mat = U^T U), m0N, m0V, m0Va; -9 not PD, -1 needs wider ints */
int or_factor(const int64_t *mat, int dim, const int64_t *linvec,
              double *U, double *m0N, double *m0V, double *m0Va)
{
    enum { MD = 32 };
    if (dim > MD) return -1;
    i128 A[MD][MD + 1];
    int na = dim + (linvec != NULL);
    for (int i = 0; i < dim; ++i) {
        for (int j = 0; j < dim; ++j) A[i][j] = mat[i * dim + j];
        if (linvec) A[i][dim] = linvec[i];
    }
    i128 prev = 1;
    double N = 0.0, V[MD], Va[MD];
    for (int j = 0; j < dim; ++j) V[j] = Va[j] = 0.0;
    for (int k = 0; k < dim; ++k) {
        i128 pk = A[k][k];
        if (pk <= 0) return -9;
        double pkd = (double)pk, prevd = (double)prev;
        double ukk = sqrt(pkd / prevd);
        U[k * dim + k] = ukk;
        for (int j = 0; j < k; ++j) U[k * dim + j] = 0.0;
        for (int j = k + 1; j < dim; ++j) U[k * dim + j] = ukk * ((double)A[k][j] / pkd);
        if (linvec) {
            double z = (double)A[k][dim] / prevd;
            N += z * z * (prevd / pkd);
            m0N[k] = N;
            for (int j = k + 1; j < dim; ++j) {
                double t = ((double)A[k][j] / pkd) * z;
                V[j] += t; Va[j] += fabs(t);
                m0V[k * dim + j] = V[j]; m0Va[k * dim + j] = Va[j];
            }
        }
        for (int i = k + 1; i < dim; ++i)
            for (int j = i; j < na; ++j) {
                i128 q;
                if (bdiv(pk, A[i][j], A[i][k], A[k][j], prev, &q)) return -1;
                A[i][j] = q;
                if (j < dim && j != i) A[j][i] = q;
            }
        prev = pk;
    }
    return 0;
}

/* The same factor in exact GMP arithmetic, for when int128 is too narrow
 * (fp_kernel.h:193-306): fraction-free elimination, U_kk = sqrt(p_k/p_{k-1}),
 * U_kj = U_kk A_kj / p_k; M0 data from z = R^{-T} linvec; m0Va = |m0V|. */
static int or_factor_gmp(const int64_t *mat, int dim, const int64_t *linvec,
                         double *U, double *m0N, double *m0V, double *m0Va)
{
    const int n2 = dim * dim;
    mpz_t *A = malloc(sizeof(mpz_t) * n2);
    mpz_t prev, t;
    mpq_t r;
    for (int k = 0; k < n2; ++k) mpz_init_set_si(A[k], mat[k]);
    mpz_init_set_ui(prev, 1);
    mpz_init(t);
    mpq_init(r);
    int st = 0;
    for (int k = 0; k < dim && !st; ++k) {
        mpz_t *pk = &A[k * dim + k];
        if (mpz_sgn(*pk) <= 0) { st = -9; break; }
        mpq_set_num(r, *pk); mpq_set_den(r, prev); mpq_canonicalize(r);
        double ukk = sqrt(mpq_get_d(r));
        U[k * dim + k] = ukk;
        for (int j = 0; j < k; ++j) U[k * dim + j] = 0.0;
        for (int j = k + 1; j < dim; ++j) {
            mpq_set_num(r, A[k * dim + j]); mpq_set_den(r, *pk); mpq_canonicalize(r);
            U[k * dim + j] = ukk * mpq_get_d(r);
        }
        for (int i = k + 1; i < dim; ++i)
            for (int j = i; j < dim; ++j) {
                mpz_mul(t, *pk, A[i * dim + j]);
                mpz_submul(t, A[i * dim + k], A[k * dim + j]);
                mpz_divexact(A[i * dim + j], t, prev);
                if (j != i) mpz_set(A[j * dim + i], A[i * dim + j]);
            }
        mpz_set(prev, *pk);
    }
    if (!st && linvec) {
        mpq_t *z = malloc(sizeof(mpq_t) * dim), *V = malloc(sizeof(mpq_t) * dim);
        mpq_t R, acc, N;
        mpq_inits(R, acc, N, NULL);
        for (int k = 0; k < dim; ++k) { mpq_init(z[k]); mpq_init(V[k]); }
        for (int k = 0; k < dim; ++k) {
            mpq_set_si(z[k], linvec[k], 1);
            for (int m = 0; m < k; ++m) {
                mpq_set_num(R, A[m * dim + k]); mpq_set_den(R, A[m * dim + m]); mpq_canonicalize(R);
                mpq_mul(acc, R, z[m]);
                mpq_sub(z[k], z[k], acc);
            }
        }
        mpz_set_ui(prev, 1);
        for (int i = 0; i < dim; ++i) {
            mpq_mul(acc, z[i], z[i]);
            mpq_set_num(R, prev); mpq_set_den(R, A[i * dim + i]); mpq_canonicalize(R);
            mpq_mul(acc, acc, R);
            mpq_add(N, N, acc);
            m0N[i] = mpq_get_d(N);
            for (int j = i + 1; j < dim; ++j) {
                mpq_set_num(R, A[i * dim + j]); mpq_set_den(R, A[i * dim + i]); mpq_canonicalize(R);
                mpq_mul(acc, R, z[i]);
                mpq_add(V[j], V[j], acc);
                m0V[i * dim + j] = mpq_get_d(V[j]);
                m0Va[i * dim + j] = fabs(m0V[i * dim + j]);
            }
            mpz_set(prev, A[i * dim + i]);
        }
        for (int k = 0; k < dim; ++k) { mpq_clear(z[k]); mpq_clear(V[k]); }
        mpq_clears(R, acc, N, NULL);
        free(z); free(V);
    }
    for (int k = 0; k < n2; ++k) mpz_clear(A[k]);
    free(A);
    mpz_clear(prev); mpz_clear(t); mpq_clear(r);
    return st;
}

/* ------------------------------------------------------------ search */
typedef struct {
    int dim, strict, nrows, m0_level;
    int64_t Q, qmax, linmin;
    const int64_t *mat, *linvec, *H;       /* H: (nrows, dim) int64 */
    long max_out;
    double eps;
    const double *fac;    /* optional exact factor: U, m0N, m0V, m0Va */
    /* optional: rows (nleaf, dim) int64 whose gcd at a leaf is the exact g
     * (used when some rows of H do not fit int64: H then prunes with the
     * rows that fit, and the leaf decides with the complete gcd) */
    const int64_t *leafrows;
    int nleaf;
} or_problem;

typedef struct {
    int32_t *pts; int64_t *qs; long n, cap;
    long long n_nodes, n_cand, n_leaf;
} or_out;

/* smallest gcd passing the cut when q >= q_lb (fpk_search_impl.h:43-51) */
static uint64_t fneed(double q_lb, int64_t Q, int strict)
{
    if (!(q_lb > 0)) return 0;
    double t = q_lb / (double)Q;
    t -= 1e-9 * (1.0 + t);
    if (t < 0) t = 0;
    if (t >= 1.8e19) return UINT64_MAX;
    return strict ? (uint64_t)floor(t) + 1 : (uint64_t)ceil(t);
}
static int gfails(u128 g, int64_t Q, double q_lb, int strict)  /* 56-62 */
{
    if (g == 0 || (g >> 64) != 0) return 0;
    double Qg = (double)Q * (double)(uint64_t)g;
    return strict ? (Qg <= q_lb) : (Qg < q_lb);
}
static int fbounds(double R, double off, double uinv, int64_t *lo, int64_t *hi)  /* 67-78 */
{
    double a = (-R - off) * uinv, b = (R - off) * uinv;
    a = ceil(a - 1e-9 * (1.0 + fabs(a)));
    b = floor(b + 1e-9 * (1.0 + fabs(b)));
    if (!(a >= -2147483647.0 && b <= 2147483647.0)) return -8;
    *lo = (int64_t)a; *hi = (int64_t)b;
    return 0;
}
static int exact_q(const int64_t *mat, const int32_t *c, int dim, i128 *q_out)   /* 158-178 */
{
    i128 q = 0;
    for (int i = 0; i < dim; ++i) {
        if (!c[i]) continue;
        i128 row = 0, t;
        for (int j = i + 1; j < dim; ++j) row += (i128)mat[i * dim + j] * c[j];
        if (omul(row, 2, &t) || oadd(t, (i128)mat[i * dim + i] * c[i], &t) ||
            omul(t, c[i], &t) || oadd(q, t, &q)) return 1;
    }
    *q_out = q;
    return 0;
}
static int push(or_out *o, const int32_t *c, int dim, int64_t q, long max_out)
{
    if (o->n >= max_out) return -2;
    if (o->n == o->cap) {
        long cap = o->cap ? 2 * o->cap : 64;
        o->pts = realloc(o->pts, sizeof(int32_t) * cap * dim);
        o->qs = realloc(o->qs, sizeof(int64_t) * cap);
        if (!o->pts || !o->qs) return -7;
        o->cap = cap;
    }
    memcpy(&o->pts[o->n * dim], c, sizeof(int32_t) * dim);
    o->qs[o->n++] = q;
    return 0;
}

/* exploration aid: stop a search after this many nodes (0: unlimited) */
long long or_node_budget = 0;
void or_set_node_budget(long long b) { or_node_budget = b; }

/* Fincke-Pohst search in double with the reference's pruning
 * (fpk_search_impl.h:217-455 with FPK_SEARCH_SPARSE 0).  Returns 0, -2,
 * -8, -9 or 1 (a leaf needs more than int128). */
int or_search(const or_problem *P, or_out *out)
{
    const int dim = P->dim;
    if (dim < 1 || dim > OMAX) return -6;
    out->n_nodes = out->n_cand = out->n_leaf = 0;
    if (P->qmax < 0) return 0;
    double U[OMAX * OMAX], m0N[OMAX], m0V[OMAX * OMAX], m0Va[OMAX * OMAX], Uinv[OMAX];
    memset(m0V, 0, sizeof m0V); memset(m0Va, 0, sizeof m0Va);
    if (P->fac) {
        const int d2 = dim * dim;
        memcpy(U, P->fac, sizeof(double) * d2);
        memcpy(m0N, P->fac + d2, sizeof(double) * dim);
        memcpy(m0V, P->fac + d2 + dim, sizeof(double) * d2);
        memcpy(m0Va, P->fac + 2 * d2 + dim, sizeof(double) * d2);
    } else {
        int st = or_factor(P->mat, dim, P->linvec, U, m0N, m0V, m0Va);
        if (st == -1) {
            memset(m0V, 0, sizeof m0V); memset(m0Va, 0, sizeof m0Va);
            st = or_factor_gmp(P->mat, dim, P->linvec, U, m0N, m0V, m0Va);
        }
        if (st) return st;
    }
    for (int i = 0; i < dim; ++i) Uinv[i] = 1.0 / U[i * dim + i];
    int m0_level = -1;
    if (P->linvec) {
        for (int j = 0; j < dim; ++j) if (P->linvec[j]) { m0_level = j; break; }
        if (m0_level < 0 && 0 < P->linmin) return 0;
    }
    const int use_gcd = P->nrows > 0;
    int lstart[OMAX + 1], order[OMAX * 2], cnt[OMAX + 1];
    memset(cnt, 0, sizeof cnt);
    int rlev[OMAX * 2];
    for (int r = 0; r < P->nrows; ++r) {
        int lv = -1;
        for (int j = 0; j < dim; ++j) if (P->H[r * dim + j]) { lv = j; break; }
        rlev[r] = lv;
        if (lv >= 0) cnt[lv]++;
    }
    lstart[0] = 0;
    for (int l = 0; l < dim; ++l) lstart[l + 1] = lstart[l] + cnt[l];
    {
        int fill[OMAX + 1];
        memcpy(fill, lstart, sizeof(int) * (dim + 1));
        for (int r = 0; r < P->nrows; ++r) if (rlev[r] >= 0) order[fill[rlev[r]]++] = r;
    }
    const double qmax_d = (double)P->qmax;
    const double slack = fmax(P->eps, 1e-9 * (qmax_d + 1.0));
    const double Kerr = 2.0 * (double)(dim + 8) * 0x1p-53;
    const int64_t *mat = P->mat, *linvec = P->linvec, *Hs = P->H;
    int32_t c[OMAX];
    int64_t cur[OMAX], hi[OMAX];
    double rem[OMAX], off[OMAX], err[OMAX], aoff[OMAX];
    i128 m0p[OMAX];
    u128 g[OMAX + 1];
    i128 pre[OMAX * 2];
    for (int j = 0; j < dim; ++j) c[j] = 0;

#define ENTER(i) do {                                                             \
        double o_ = 0.0, a_ = 0.0;                                                \
        for (int j_ = (i) + 1; j_ < dim; ++j_) {                                  \
            double t_ = U[(i) * dim + j_] * c[j_];                                \
            o_ += t_; a_ += fabs(t_);                                             \
        }                                                                         \
        off[i] = o_; aoff[i] = a_;                                                \
        if (use_gcd)                                                              \
            for (int k_ = lstart[i]; k_ < lstart[(i) + 1]; ++k_) {                \
                int r_ = order[k_]; i128 s_ = 0;                                  \
                for (int j_ = (i) + 1; j_ < dim; ++j_) s_ += (i128)Hs[r_ * dim + j_] * c[j_]; \
                pre[r_] = s_;                                                     \
            }                                                                     \
        double R_ = sqrt(fmax(rem[i] + slack + err[i], 0.0));                    \
        R_ += Kerr * (R_ + 2.0 * a_) + 1e-300;                                    \
        int empty_ = 0;                                                           \
        if (use_gcd && g[(i) + 1] != 0 && (g[(i) + 1] >> 64) == 0) {             \
            double am_ = R_ + 2.0 * a_;                                           \
            double Rg2_ = (double)P->Q * (double)(uint64_t)g[(i) + 1]             \
                - (qmax_d - rem[i]) + slack + err[i] + Kerr * (am_ * am_ + rem[i]); \
            if (Rg2_ < 0.0) empty_ = 1;                                           \
            else {                                                                \
                double Rg_ = sqrt(Rg2_);                                          \
                Rg_ += Kerr * (Rg_ + 2.0 * a_) + 1e-300;                          \
                if (Rg_ < R_) R_ = Rg_;                                           \
            }                                                                     \
        }                                                                         \
        if (!empty_ && linvec && (i) >= m0_level && m0_level >= 0) {             \
            double vc_ = 0.0, va_ = 0.0;                                          \
            for (int j_ = (i) + 1; j_ < dim; ++j_) {                              \
                vc_ += m0V[(i) * dim + j_] * c[j_];                               \
                va_ += m0Va[(i) * dim + j_] * fabs((double)c[j_]);                \
            }                                                                     \
            double sq_ = sqrt(fmax(m0N[i] * (rem[i] + slack + err[i]), 0.0));     \
            double ub_ = (double)m0p[i] - vc_ + sq_                               \
                + Kerr * (va_ + sq_ + fabs((double)m0p[i])) + 1e-9;               \
            if (ub_ < (double)P->linmin) empty_ = 1;                              \
        }                                                                         \
        int64_t lo_ = 1, hi_ = 0;                                                 \
        if (!empty_ && fbounds(R_, off[i], Uinv[i], &lo_, &hi_)) return -8;       \
        if (!empty_ && (i) == m0_level) {                                         \
            i128 need_ = (i128)P->linmin - m0p[i], l_ = linvec[i];                \
            if (l_ > 0) {                                                         \
                i128 b_ = need_ >= 0 ? (need_ + l_ - 1) / l_ : -((-need_) / l_);  \
                if (b_ > lo_) lo_ = b_ > hi_ ? hi_ + 1 : (int64_t)b_;             \
            } else {                                                              \
                i128 nl_ = -l_;                                                   \
                i128 b_ = need_ <= 0 ? (-need_) / nl_ : -((need_ + nl_ - 1) / nl_); \
                if (b_ < hi_) hi_ = b_ < lo_ ? lo_ - 1 : (int64_t)b_;             \
            }                                                                     \
        }                                                                         \
        cur[i] = lo_; hi[i] = hi_;                                                \
        if (++out->n_nodes > or_node_budget && or_node_budget > 0) return -12;    \
    } while (0)

    int i = dim - 1;
    rem[i] = qmax_d; err[i] = Kerr * qmax_d; m0p[i] = 0; g[dim] = 0;
    ENTER(i);
    while (i < dim) {
        if (cur[i] > hi[i]) { c[i] = 0; i++; continue; }
        int32_t v = (int32_t)cur[i]++;
        c[i] = v;
        out->n_cand++;
        double x = U[i * dim + i] * v + off[i];
        double nr = rem[i] - x * x;
        double ai = fabs(U[i * dim + i] * v) + aoff[i];
        double en = err[i] + Kerr * (ai * ai + rem[i]);
        if (nr < -(slack + en)) continue;
        i128 m0 = m0p[i] + (linvec ? (i128)linvec[i] * v : 0);
        if (i == m0_level && m0 < P->linmin) continue;
        if (use_gcd) {
            u128 gs = g[i + 1];
            double q_lb = (qmax_d - nr) - (slack + en);
            uint64_t need = fneed(q_lb, P->Q, P->strict);
            int pruned = 0;
            for (int k = lstart[i]; k < lstart[i + 1]; ++k) {
                if (gfails(gs, P->Q, q_lb, P->strict)) { pruned = 1; break; }
                int r = order[k];
                u128 a = oabs(pre[r] + (i128)Hs[r * dim + i] * v);
                if ((gs >> 64) == 0 && (a >> 64) == 0 && (gs | a) != 0) {
                    uint64_t gg = gcd64_ge((uint64_t)gs, (uint64_t)a, need);
                    if (!gg) { pruned = 1; break; }
                    gs = gg;
                } else {
                    gs = gcd128(gs, a);
                }
            }
            if (pruned || gfails(gs, P->Q, q_lb, P->strict)) continue;
            g[i] = gs;
        }
        if (i > 0) {
            m0p[i - 1] = m0; rem[i - 1] = nr; err[i - 1] = en;
            i--;
            ENTER(i);
            continue;
        }
        out->n_leaf++;
        i128 q;
        if (exact_q(mat, c, dim, &q)) return 1;
        if (q > P->qmax) continue;
        u128 gl = use_gcd ? g[0] : 0;
        if (P->leafrows) {
            /* gcd(H c) = gcd(L c) for L with the row lattice of H (|L c| < 2^100) */
            gl = 0;
            for (int r = 0; r < P->nleaf; ++r) {
                i128 s_ = 0;
                for (int j = 0; j < dim; ++j) s_ += (i128)P->leafrows[r * dim + j] * c[j];
                gl = gcd128(gl, oabs(s_));
            }
        }
        if ((use_gcd || P->leafrows) && gl != 0 && (gl >> 64) == 0) {
            i128 Qg = (i128)P->Q * (i128)(uint64_t)gl;
            if (P->strict ? (Qg <= q) : (Qg < q)) continue;
        }
        int rc = push(out, c, dim, (int64_t)q, P->max_out);
        if (rc) return rc;
    }
#undef ENTER
    return 0;
}

void or_out_free(or_out *o) { free(o->pts); free(o->qs); o->pts = NULL; o->qs = NULL; o->n = o->cap = 0; }

/* ----------------------------------------------------------- batch API */
/* One single-problem search, for the kernel fixtures.  pts/qs are malloc'd. */
int or_kernel(int dim, const int64_t *mat, int64_t Q, int64_t qmax, int nrows,
              const int64_t *H, int strict, const int64_t *linvec, int64_t linmin,
              long max_out, double eps, const double *fac, int32_t **pts, int64_t **qs,
              long *n, long long *counts)
{
    or_problem P = {dim, strict, nrows, -1, Q, qmax, linmin, mat, linvec, H, max_out, eps, fac,
                    NULL, 0};
    or_out o;
    memset(&o, 0, sizeof o);
    int st = or_search(&P, &o);
    *pts = o.pts; *qs = o.qs; *n = o.n;
    counts[0] = o.n_nodes; counts[1] = o.n_cand; counts[2] = o.n_leaf;
    return st;
}

/* The batched coni pipeline (fp_kernel.pyx:480-622) for one geometry:
 * every p row (length h, leading 0) is built and searched; points are
 * appended as M (h), Kn (h), q, pidx.  status[ip]: 0 done, 1 exact path,
 * < 0 kernel status.  Returns the number of points (arrays malloc'd). */
long or_coni_batch(int h, const int64_t *kappa, const int64_t *Mbasis, const int64_t *ps,
                   long n_p, int64_t Q, int64_t qmax, int64_t linmin, long max_out,
                   int64_t **Mo, int64_t **Ko, int64_t **qo, int64_t **po, int32_t *status,
                   long long *counts)
{
    const int d = h - 1;
    or_lat *R = malloc(sizeof(or_lat));
    int64_t *H64 = malloc(sizeof(int64_t) * OMAX * OMAX);
    long cnt = 0, cap = 0;
    int64_t *Mb = NULL, *Kb = NULL, *qb = NULL, *pb = NULL;
    counts[0] = counts[1] = counts[2] = 0;
    for (long ip = 0; ip < n_p; ++ip) {
        status[ip] = 0;
        int st = or_build(h, kappa, Mbasis, &ps[ip * h], R);
        if (st) { status[ip] = 1; continue; }
        /* HNF rows beyond int64 leave the pruning; the leaf then takes the
         * exact gcd of Kn[1:] = ZB[1:] c, whose rows span the same lattice */
        int bad = 0, wide = 0;
        for (int r = 0; r < d; ++r) {
            int fit = 1;
            for (int k = 0; k < d; ++k) fit &= fits64(R->H[r * d + k]);
            for (int k = 0; k < d; ++k) H64[r * d + k] = fit ? (int64_t)R->H[r * d + k] : 0;
            wide |= !fit;
        }
        or_problem P = {d, 1, d, -1, Q, qmax, linmin, R->mat, &R->Binter[0], H64, max_out, 1e-4,
                        NULL, wide ? &R->ZB[d] : NULL, wide ? d : 0};
        or_out o;
        memset(&o, 0, sizeof o);
        st = or_search(&P, &o);
        counts[0] += o.n_nodes; counts[1] += o.n_cand; counts[2] += o.n_leaf;
        if (st) { status[ip] = st; or_out_free(&o); continue; }
        if (cnt + o.n > cap) {
            cap = 2 * cap > cnt + o.n ? 2 * cap : cnt + o.n + 64;
            Mb = realloc(Mb, sizeof(int64_t) * cap * h);
            Kb = realloc(Kb, sizeof(int64_t) * cap * h);
            qb = realloc(qb, sizeof(int64_t) * cap);
            pb = realloc(pb, sizeof(int64_t) * cap);
        }
        long start = cnt;
        for (long t = 0; t < o.n; ++t) {
            for (int i = 0; i < h; ++i) {
                i128 am = 0, ak = 0;
                for (int a = 0; a < d; ++a) {
                    am += (i128)R->Binter[i * d + a] * o.pts[t * d + a];
                    ak += (i128)R->ZB[i * d + a] * o.pts[t * d + a];
                }
                if (!fits64(am) || !fits64(ak)) bad = 1;
                Mb[cnt * h + i] = (int64_t)am;
                Kb[cnt * h + i] = (int64_t)ak;
            }
            qb[cnt] = o.qs[t];
            pb[cnt] = ip;
            cnt++;
        }
        or_out_free(&o);
        if (bad) { cnt = start; status[ip] = 1; }
    }
    free(R); free(H64);
    *Mo = Mb; *Ko = Kb; *qo = qb; *po = pb;
    return cnt;
}

/* Build only, exposing the lattice data (for lattice-level checks). */
int or_build_export(int h, const int64_t *kappa, const int64_t *Mbasis, const int64_t *p,
                    int64_t *Z, int64_t *Binter, int64_t *ZB, int64_t *mat, int64_t *H_lohi)
{
    or_lat *R = malloc(sizeof(or_lat));
    int st = or_build(h, kappa, Mbasis, p, R);
    const int d = h - 1;
    if (st == 0 || st == -15) {
        memcpy(Z, R->Z, sizeof(int64_t) * h * h);
        memcpy(Binter, R->Binter, sizeof(int64_t) * h * d);
        memcpy(ZB, R->ZB, sizeof(int64_t) * h * d);
        memcpy(mat, R->mat, sizeof(int64_t) * d * d);
        if (st == 0)
            for (int k = 0; k < d * d; ++k) {
                u128 u = (u128)R->H[k];
                H_lohi[2 * k] = (int64_t)(uint64_t)u;
                H_lohi[2 * k + 1] = (int64_t)(uint64_t)(u >> 64);
            }
    }
    free(R);
    return st;
}

void or_free(void *p) { free(p); }
