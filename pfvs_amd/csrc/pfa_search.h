// pfa_search.h -- the Fincke-Pohst node and candidate steps of the coniZpM
// search, shared (same source, same floating-point operation order) by the
// plain one-thread-per-problem kernel, the batched wave-cooperative kernel
// and the host path, so all three visit exactly the same nodes.
//
// Contract (pfvs/fp_kernel/fp_kernel.h:25-63): return exactly the c in Z^d
// with q(c) = c^T mat c <= qmax, linvec.c >= linmin and g == 0 or Q g > q(c)
// (strict), g = gcd(H c).  Coordinates are set from c[d-1] down to c[0]
// (fpk_search_impl.h:217-455); floating point only prunes, every decision
// at a leaf is exact.
//
// Factor.  U is an FP64 Cholesky factor of mat computed on the device, not
// an exactly rounded one.  Its backward error is carried as phi: for every
// integer c, (1 - phi) |U c|^2 <= q(c) <= (1 + phi) |U c|^2 (proof in the
// preparation code).  The reference's pruning (radii, error recursion with
// Kerr = 2 (d + 8) 2^-53 for evaluating U c in floating point) is kept, with
// the ellipsoid budget Qe >= qmax / (1 - phi), lower bounds on q scaled by
// (1 - phi) and GCD budgets Q g scaled by QGmul >= 1 / (1 - phi).
// The M0 upper-bound test of the reference (fpk_search_impl.h:276-286) is
// not used: M0 is cut exactly at its level, which keeps the output exact.
#pragma once
#include <math.h>
#include "pfa_int.h"

#define PFA_DMAX 40                  // largest d = h11 - 1 supported (h11 <= 41)

// Status codes (per problem)
enum {
    PFA_OK = 0,
    PFA_HOST = 1,          // hand the problem to the exact host path
    PFA_TOO_MANY = -2,     // more points than allowed
    PFA_I32 = -8,          // a coordinate range exceeds int32
    PFA_NOT_PD = -9,       // mat is not positive definite (decided exactly)
};

// A problem as the search sees it (pointers into the device/host record).
struct PfaProb {
    int d, m0_level, strict, leafg;   // leafg: 1 = recompute g at leaves
    int root_empty;                   // 1: proven before the factor that the root has no candidate
    int64_t qmax, Q, linmin;
    double Qe, QGmul, phi, slack, Kerr;
    uint64_t hmask;                   // bit i: a GCD row becomes known at level i
    const double *U;                  // (d, d) row-major, upper triangular
    const double *Uinv;               // (d,) 1 / U_ii
    const int64_t *mat;               // (d, d)
    const int64_t *linvec;            // (d,)
    const int64_t *H;                 // (d, d): row of level i at H + i d
    const int64_t *L;                 // leaf gcd rows (nL, d) as int128 pairs, when leafg
    int nL;
    int64_t matmax, hmax;             // max |mat|, max |H| over the pruning rows
};

// One node: level i with c[i+1:] fixed; c[i] runs over [cur, hi].
struct PfaNode {
    int32_t i, cur, hi;
    int32_t cmax;    // max_{j>i} |c_j|
    double rem, err, off, aoff;
    pfa_i128 m0p;    // sum_{j>i} linvec_j c_j (exact)
    pfa_i128 pre;    // sum_{j>i} H_ij c_j for the row of level i
    pfa_i128 qp;     // c[i+1:]^T mat c[i+1:] (exact)
    pfa_i128 mrow;   // sum_{j>i} mat_ij c_j (exact)
    pfa_u128 g;      // gcd of the rows determined above level i
};

struct PfaCounts {
    unsigned long long nodes, cand, leaf;
};

// Terms of the algorithmic work of a search (host path only), with
// k = d - 1 - i the coordinates set above a node at level i, r = 1 if an H
// row leads at column i, s = iterations of the binary gcd:
//   sk = sum_nodes k, skm0 = sum_nodes k [i >= m0_level], skr = sum_nodes k r,
//   sr = sum_cand r, ss = sum_cand s, sleaf = sum_leaf d (d + 1) / 2.
struct PfaWork {
    unsigned long long sk, skm0, skr, sr, ss, sleaf;
};

// fpk_search_impl.h:43-51: smallest gcd passing the cut when q >= q_lb,
// rounded down so it never demands more than the exact cut.
PFA_HD static inline uint64_t pfa_need(double q_lb, int64_t Q, int strict)
{
    if (!(q_lb > 0)) return 0;
    double t = q_lb / (double)Q;
    t -= 1e-9 * (1.0 + t);
    if (t < 0) t = 0;
    if (t >= 1.8e19) return UINT64_MAX;
    return strict ? (uint64_t)floor(t) + 1 : (uint64_t)ceil(t);
}

// 1 if a vector with q >= q_lb must fail the GCD cut (g == 0 never prunes)
PFA_HD static inline int pfa_gcd_fails(pfa_u128 g, int64_t Q, double q_lb, int strict)
{
    if (g == 0 || (g >> 64) != 0) return 0;
    double Qg = (double)Q * (double)(uint64_t)g;
    return strict ? (Qg <= q_lb) : (Qg < q_lb);
}

// all integers v with |U_ii v + off| <= R, widened outward (-8 beyond int32)
PFA_HD static inline int pfa_bounds(double R, double off, double uinv, int32_t *lo, int32_t *hi)
{
    double a = (-R - off) * uinv, b = (R - off) * uinv;
    a = ceil(a - 1e-9 * (1.0 + fabs(a)));
    b = floor(b + 1e-9 * (1.0 + fabs(b)));
    if (!(a >= -2147483647.0 && b <= 2147483647.0)) return PFA_I32;
    *lo = (int32_t)a;
    *hi = (int32_t)b;
    return 0;
}

// Enter level i, second half: with the center nd.off and its magnitude
// nd.aoff set, the radius (with the GCD narrowing) and the integer interval
// [cur, hi] of c[i], cut exactly by the M0 bound at its level.  Returns 0
// or PFA_I32.
PFA_HD static inline int pfa_enter_finish(const PfaProb &P, PfaNode &nd)
{
    const int i = nd.i;
    const double o = nd.off, a = nd.aoff;
    const double slack = P.slack, Kerr = P.Kerr;
    double R = sqrt(fmax(nd.rem + slack + nd.err, 0.0));
    R += Kerr * (R + 2.0 * a) + 1e-300;
    int empty = 0;
    if (P.hmask && nd.g != 0 && (nd.g >> 64) == 0) {
        // GCD narrowing: a child survives only if Q g >= (1 - phi) * its
        // partial norm, i.e. t_i^2 <= Q g / (1 - phi) - (Qe - rem) + margins
        double am = R + 2.0 * a;
        double Rg2 = (double)P.Q * (double)(uint64_t)nd.g * P.QGmul - (P.Qe - nd.rem) + slack +
                     nd.err + Kerr * (am * am + nd.rem);
        if (Rg2 < 0.0) {
            empty = 1;
        } else {
            double Rg = sqrt(Rg2);
            Rg += Kerr * (Rg + 2.0 * a) + 1e-300;
            if (Rg < R) R = Rg;
        }
    }
    int32_t lo = 1, hi = 0;
    if (!empty && pfa_bounds(R, o, P.Uinv[i], &lo, &hi)) return PFA_I32;
    if (!empty && i == P.m0_level) {
        // M0 = m0p + l_i v >= linmin: an exact linear bound on v
        pfa_i128 need = (pfa_i128)P.linmin - nd.m0p;
        pfa_i128 l = P.linvec[i];
        if (l > 0) {
            pfa_i128 b = need >= 0 ? pfa_tdiv128(need + l - 1, l) : -pfa_tdiv128(-need, l);
            if (b > lo) lo = b > hi ? hi + 1 : (int32_t)b;
        } else {
            pfa_i128 nl = -l;
            pfa_i128 b = need <= 0 ? pfa_tdiv128(-need, nl) : -pfa_tdiv128(need + nl - 1, nl);
            if (b < hi) hi = b < lo ? lo - 1 : (int32_t)b;
        }
    }
    nd.cur = lo;
    nd.hi = hi;
    return 0;
}

// 1 when the exact prefixes of level i fit the int64 form: |mat|, |H| <
// 2^31 and every partial sum bounded by max(|mat|, |H|) max|c_j| (d - 1 - i)
// < 2^62
PFA_HD static inline int pfa_prefix_small(const PfaProb &P, const PfaNode &nd)
{
    const int64_t B31 = (int64_t)1 << 31, bm = P.matmax > P.hmax ? P.matmax : P.hmax;
    return bm < B31 && (double)bm * (double)nd.cmax * (double)(P.d - 1 - nd.i) < 4.0e18;
}

// Enter level i (nd.i = i): nd.rem, err, m0p, g (of level i+1) and qp are
// set by the caller; computes the center off = sum_{j > i} U_ij c_j (j
// ascending), aoff, the exact prefixes mrow and pre, then the interval.
// Returns 0 or PFA_I32.
PFA_HD static inline int pfa_enter(const PfaProb &P, const int32_t *c, PfaNode &nd)
{
    const int d = P.d, i = nd.i;
    const double *Ui = P.U + (size_t)i * d;
    const int64_t *Mi = P.mat + (size_t)i * d;
    const int64_t *Hi = P.H + (size_t)i * d;
    const int hrow = (int)((P.hmask >> i) & 1);
    double o = 0.0, a = 0.0;
    if (pfa_prefix_small(P, nd)) {
        int64_t mr = 0, sp = 0;
        for (int j = i + 1; j < d; ++j) {
            const int32_t cj = c[j];
            double t = Ui[j] * (double)cj;
            o += t;
            a += fabs(t);
            mr += (int64_t)(int32_t)Mi[j] * cj;
            sp += (int64_t)(int32_t)Hi[j] * cj;
        }
        nd.mrow = mr;
        nd.pre = hrow ? sp : 0;
    } else {
        for (int j = i + 1; j < d; ++j) {
            double t = Ui[j] * (double)c[j];
            o += t;
            a += fabs(t);
        }
        pfa_i128 mr = 0;
        for (int j = i + 1; j < d; ++j) mr += (pfa_i128)Mi[j] * c[j];
        nd.mrow = mr;
        nd.pre = 0;
        if (hrow) {
            pfa_i128 s = 0;
            for (int j = i + 1; j < d; ++j) s += (pfa_i128)Hi[j] * c[j];
            nd.pre = s;
        }
    }
    nd.off = o;
    nd.aoff = a;
    return pfa_enter_finish(P, nd);
}

// Test candidate c[i] = v of node nd; on success fill the child's inherited
// state (level i-1, or the leaf when i == 0) and return 1; 0 if cut;
// PFA_CAND_OVF if exact int128 arithmetic overflowed.
#define PFA_CAND_OVF 2
template <bool CNT = false>
PFA_HD static inline int pfa_cand(const PfaProb &P, const PfaNode &nd, int32_t v, PfaNode &ch,
                                  unsigned long long *gcd_iters = nullptr)
{
    const int i = nd.i, d = P.d;
    const double slack = P.slack, Kerr = P.Kerr;
    const double uii = P.U[(size_t)i * d + i];
    double x = uii * (double)v + nd.off;
    double nr = nd.rem - x * x;
    double ai = fabs(uii * (double)v) + nd.aoff;
    double en = nd.err + Kerr * (ai * ai + nd.rem);
    if (nr < -(slack + en)) return 0;
    pfa_i128 m0 = nd.m0p + (pfa_i128)P.linvec[i] * v;
    if (i == P.m0_level && m0 < (pfa_i128)P.linmin) return 0;
    pfa_u128 gn = nd.g;
    if (P.hmask) {
        // partial norm >= plb, and q(c) >= (1 - phi) partial
        double plb = (P.Qe - nr) - (slack + en);
        double q_lb = plb - P.phi * (fabs(plb) + 1.0);
        if ((P.hmask >> i) & 1) {
            if (pfa_gcd_fails(gn, P.Q, q_lb, P.strict)) return 0;
            pfa_u128 a = pfa_abs128(nd.pre + (pfa_i128)P.H[(size_t)i * d + i] * v);
            if ((gn >> 64) == 0 && (a >> 64) == 0 && (gn | a) != 0) {
                uint64_t need = pfa_need(q_lb, P.Q, P.strict);
                uint64_t gg = pfa_gcd64_ge_t<CNT>((uint64_t)gn, (uint64_t)a, need, gcd_iters);
                if (!gg) return 0;
                gn = gg;
            } else {
                gn = pfa_gcd128(gn, a);
            }
        }
        if (pfa_gcd_fails(gn, P.Q, q_lb, P.strict)) return 0;
    }
    // exact partial form: qp + v (mat_ii v + 2 mrow)
    pfa_i128 t, q;
    if (pfa_mul_ovf(nd.mrow, 2, &t) ||
        pfa_add_ovf(t, (pfa_i128)P.mat[(size_t)i * d + i] * v, &t) ||
        pfa_mul_ovf(t, (pfa_i128)v, &t) || pfa_add_ovf(nd.qp, t, &q))
        return PFA_CAND_OVF;
    ch.i = i - 1;
    ch.rem = nr;
    ch.err = en;
    ch.m0p = m0;
    ch.g = gn;
    ch.qp = q;
    const int32_t av = v < 0 ? -v : v;
    ch.cmax = av > nd.cmax ? av : nd.cmax;
    return 1;
}

// The exact gcd of the leaf rows at c.  leafg 1: the rows L (nL x d, int64,
// stored as (value, 0) pairs) are the complete gcd rows, g from scratch.
// leafg 2: g0 (the exact gcd of the pruning rows) combined with rows L
// whose entries need int128 (pairs lo, hi); their values are reduced mod
// g0 first.  *ok = 0 if undecidable in int128.
PFA_HD static inline pfa_u128 pfa_leaf_gcd(const PfaProb &P, const int32_t *c, pfa_u128 g0, int *ok)
{
    *ok = 1;
    const int d = P.d;
    if (P.leafg == 1) {
        pfa_u128 g = 0;
        for (int r = 0; r < P.nL; ++r) {
            const int64_t *Lr = P.L + (size_t)r * d * 2;
            pfa_i128 s = 0;
            for (int j = 0; j < d; ++j) s += (pfa_i128)Lr[2 * j] * c[j];     // |.| < 2^100
            g = pfa_gcd128(g, pfa_abs128(s));
        }
        return g;
    }
    pfa_u128 g = g0;
    for (int r = 0; r < P.nL; ++r) {
        if (g == 0) {
            // nothing to reduce by: the row's value itself, checked in int128
            const int64_t *Lr = P.L + (size_t)r * d * 2;
            pfa_i128 s = 0, t;
            for (int j = 0; j < d; ++j) {
                pfa_i128 e = (pfa_i128)(((pfa_u128)(uint64_t)Lr[2 * j + 1] << 64) | (uint64_t)Lr[2 * j]);
                if (pfa_mul_ovf(e, (pfa_i128)c[j], &t) || pfa_add_ovf(s, t, &s)) { *ok = 0; return 0; }
            }
            g = pfa_abs128(s);
            continue;
        }
        if ((g >> 62) != 0) { *ok = 0; return 0; }
        const pfa_i128 m = (pfa_i128)g;
        const int64_t *Lr = P.L + (size_t)r * d * 2;
        pfa_i128 s = 0;
        for (int j = 0; j < d; ++j) {
            pfa_i128 e = (pfa_i128)(((pfa_u128)(uint64_t)Lr[2 * j + 1] << 64) | (uint64_t)Lr[2 * j]);
            pfa_i128 em = e % m;                                  // |em| < 2^62
            s += em * c[j];                                       // |.| < 2^100
        }
        g = pfa_gcd128(g, pfa_abs128(s % m));
    }
    return g;
}

// g of a kept leaf when it does not fit 64 bits
#define PFA_G_WIDE UINT64_MAX

// Leaf decision for the vector c whose last step produced ch (ch.i == -1):
// 1 keep (q in *q_out), 0 drop, PFA_CAND_OVF undecidable here.  *g_out:
// the exact gcd of the pruning and leaf rows at c -- the rows of the HNF of
// ZB[1:] (or ZB[1:] itself), so gcd(Kn[1:]) with Kn = ZB c, 0 when every
// row vanishes -- or PFA_G_WIDE beyond 64 bits.
PFA_HD static inline int pfa_leaf(const PfaProb &P, const PfaNode &ch, const int32_t *c,
                                  int64_t *q_out, uint64_t *g_out)
{
    pfa_i128 q = ch.qp;
    if (q > (pfa_i128)P.qmax) return 0;
    pfa_u128 g = ch.g;
    if (P.leafg) {
        int ok;
        g = pfa_leaf_gcd(P, c, ch.g, &ok);
        if (!ok) return PFA_CAND_OVF;
    }
    if (P.hmask || P.leafg) {
        if (g != 0 && (g >> 64) == 0) {
            pfa_i128 Qg = (pfa_i128)P.Q * (pfa_i128)(uint64_t)g;
            if (P.strict ? (Qg <= q) : (Qg < q)) return 0;
        }
    }
    *q_out = (int64_t)q;
    *g_out = (g >> 64) == 0 ? (uint64_t)g : PFA_G_WIDE;
    return 1;
}

// Root node of a problem (level d-1), before pfa_enter.
PFA_HD static inline void pfa_root(const PfaProb &P, PfaNode &nd)
{
    nd.i = P.d - 1;
    nd.rem = P.Qe;
    nd.err = P.Kerr * P.Qe;
    nd.m0p = 0;
    nd.g = 0;
    nd.qp = 0;
    nd.pre = 0;
    nd.mrow = 0;
    nd.cur = 1;
    nd.hi = 0;
    nd.cmax = 0;
}

// The root node, entered: its interval (or, for a problem whose root was
// proven empty at preparation, whose factor and rows are not built, the
// empty interval).  Returns 0 or PFA_I32; c[] must be all zero.
PFA_HD static inline int pfa_root_enter(const PfaProb &P, const int32_t *c, PfaNode &nd)
{
    pfa_root(P, nd);
    if (P.root_empty) return 0;
    return pfa_enter(P, c, nd);
}

// 1 if the problem has no node at all (M0 identically 0 below linmin, or
// qmax < 0), matching fp_kernel.h:391 and 435.
PFA_HD static inline int pfa_trivially_empty(const PfaProb &P)
{
    return P.qmax < 0 || (P.m0_level < 0 && P.linmin > 0);
}

// The whole search of one problem, depth first, in the reference's order:
// the plain form (one thread per problem) and the host path.  emit(c, q, g)
// (g as pfa_leaf's) returns nonzero to stop (-2).  Returns a status.  With
// CNT the terms of the algorithmic work are added to *wk.
template <bool CNT, class Emit>
PFA_HD static inline int pfa_dfs_t(const PfaProb &P, PfaNode *st, int32_t *c, PfaCounts &cnt,
                                   Emit &emit, PfaWork *wk)
{
    const int d = P.d;
    if (pfa_trivially_empty(P)) return 0;
    for (int j = 0; j < d; ++j) c[j] = 0;
    auto entered = [&](int lev) {
        cnt.nodes++;
        if (CNT) {
            const unsigned long long k = (unsigned long long)(d - 1 - lev);
            wk->sk += k;
            if (P.m0_level >= 0 && lev >= P.m0_level) wk->skm0 += k;
            if ((P.hmask >> lev) & 1) wk->skr += k;
        }
    };
    int i = d - 1;
    if (pfa_root_enter(P, c, st[i])) return PFA_I32;
    entered(i);
    while (i < d) {
        PfaNode &nd = st[i];
        if (nd.cur > nd.hi) { c[i] = 0; i++; continue; }
        int32_t v = nd.cur++;
        c[i] = v;
        cnt.cand++;
        PfaNode ch;
        int r;
        if (CNT) {
            wk->sr += (P.hmask >> i) & 1;
            r = pfa_cand<true>(P, nd, v, ch, &wk->ss);
        } else {
            r = pfa_cand(P, nd, v, ch);
        }
        if (r == 0) continue;
        if (r == PFA_CAND_OVF) return PFA_HOST;
        if (i > 0) {
            st[i - 1].i = i - 1;
            st[i - 1].rem = ch.rem;
            st[i - 1].err = ch.err;
            st[i - 1].m0p = ch.m0p;
            st[i - 1].g = ch.g;
            st[i - 1].qp = ch.qp;
            st[i - 1].cmax = ch.cmax;
            i--;
            if (pfa_enter(P, c, st[i])) return PFA_I32;
            entered(i);
            continue;
        }
        cnt.leaf++;
        if (CNT) wk->sleaf += (unsigned long long)d * (d + 1) / 2;
        int64_t q;
        uint64_t g;
        int lr = pfa_leaf(P, ch, c, &q, &g);
        if (lr == PFA_CAND_OVF) return PFA_HOST;
        if (lr == 1 && emit(c, q, g)) return PFA_TOO_MANY;
    }
    return 0;
}

template <class Emit>
PFA_HD static inline int pfa_dfs(const PfaProb &P, PfaNode *st, int32_t *c, PfaCounts &cnt,
                                 Emit &emit)
{
    return pfa_dfs_t<false>(P, st, c, cnt, emit, (PfaWork *)nullptr);
}
