// pfa_box.h -- the p-vector enumeration on the input side of coniZpM: all
// integer p with |p|_inf <= B and H p >= rhs (row-wise), optionally
// primitive (gcd of the entries 1), in lexicographic order (coordinate 0
// outermost, ascending).  The contract of latticepts.box_enum as the
// reference calls it (pfvs/pvectors.py:58-89, pfvs/distributed.py:121-139,
// tests/test_paper_examples.py:119-121; latticepts >= 0.1.2 is a
// dependency the reference does not vendor).
//
// The box is split into units: every assignment of the first s coordinates
// (a prefix).  A unit is enumerated depth first by one team (one host
// thread; the code is written for a wavefront team as well).  The team's lanes own the rows of H
// (row r: lane r mod T.n), so a node's child interval is a team reduction:
// with part_r = sum_{k<j} H_rk p_k set and rest_r(j+1) = B sum_{k>j} |H_rk|
// the most the unset coordinates can add, p_j = v survives row r iff
//   H_rj v >= rhs_r - part_r - rest_r(j+1),
// an interval in v.  Exactly the points of the box satisfying every row
// reach the last level, where the team tests up to T.n values at once
// (primitivity: gcd(g, |v|) == 1 with g the gcd of the prefix).  Counting
// and writing runs share the code, so a unit's points land at offset
// (points of earlier units) in order.
#pragma once
#include <stdint.h>
#include "pfa_int.h"

#define PFA_BOX_MAXM 256             // rows of H
#define PFA_BOX_MAXN 64              // coordinates
#define PFA_BOX_RPL_DEV 4            // rows per lane on the device (64 lanes)

#if defined(__HIP_DEVICE_COMPILE__)
#define PFA_SYNC_BOX() __syncthreads()
#else
#define PFA_SYNC_BOX() ((void)0)
#endif

struct PfaBox {
    int m, n, B, primitive, s;
    const int32_t *Ht;               // (n, m): column j of H at Ht + j m
    const int64_t *rest;             // (n + 1, m): B sum_{k >= j} |H_rk| at rest + j m
    const int64_t *rhs;              // (m,)
};

struct PfaBoxTeam {
    int lane, n;
};

PFA_HD static inline int64_t pfa_box_max(int64_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t y = __shfl_xor(x, o, 64);
        x = y > x ? y : x;
    }
#endif
    return x;
}

PFA_HD static inline int64_t pfa_box_min(int64_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t y = __shfl_xor(x, o, 64);
        x = y < x ? y : x;
    }
#endif
    return x;
}

PFA_HD static inline int64_t pfa_box_fdiv(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}

PFA_HD static inline uint64_t pfa_box_gcd(uint64_t a, uint64_t b)
{
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// The values v of coordinate j that keep every row reachable, given the
// partial sums `part` of the coordinates before j: [*lo, *hi] (empty when
// *lo > *hi).  Same value on every lane.
template <int RPL>
PFA_HD static inline void pfa_box_interval(const PfaBox &X, const PfaBoxTeam &T, const int64_t *part,
                                           int j, int64_t *lo_out, int64_t *hi_out)
{
    int64_t lo = -X.B, hi = X.B;
    const int32_t *a = X.Ht + (size_t)j * X.m;
    const int64_t *rs = X.rest + (size_t)(j + 1) * X.m;
    for (int t = 0; t < RPL; ++t) {
        const int r = T.lane + T.n * t;
        if (r >= X.m) break;
        const int64_t need = X.rhs[r] - part[t] - rs[r];       // a v >= need
        const int64_t ar = a[r];
        if (ar > 0) {
            const int64_t b = -pfa_box_fdiv(-need, ar);        // ceil(need / ar)
            lo = b > lo ? b : lo;
        } else if (ar < 0) {
            const int64_t b = pfa_box_fdiv(need, ar);          // floor(need / ar)
            hi = b < hi ? b : hi;
        } else if (need > 0) {
            hi = lo - 1;
        }
    }
    lo = pfa_box_max(lo);
    hi = pfa_box_min(hi);
    *lo_out = lo;
    *hi_out = hi;
}

template <int RPL>
PFA_HD static inline void pfa_box_apply(const PfaBox &X, const PfaBoxTeam &T, int64_t *part, int j,
                                        int64_t v)
{
    const int32_t *a = X.Ht + (size_t)j * X.m;
    for (int t = 0; t < RPL; ++t) {
        const int r = T.lane + T.n * t;
        if (r >= X.m) break;
        part[t] += (int64_t)a[r] * v;
    }
}

// The last coordinate: the values v in [lo, hi] (every one satisfies the
// rows) that keep p primitive (gcd(g, |v|) == 1) when asked; counted, and
// written after the prefix p[0:n-1] at out + (base + k) n for the k-th one
// when out != nullptr (rows at or beyond cap are not written).
PFA_HD static inline int64_t pfa_box_last(const PfaBox &X, const PfaBoxTeam &T, const int32_t *p,
                                          uint64_t g, int64_t lo, int64_t hi, int64_t *out,
                                          int64_t base, int64_t cap)
{
    int64_t cnt = 0;
    for (int64_t v0 = lo; v0 <= hi; v0 += T.n) {
        const int64_t v = v0 + T.lane;
        int ok = v <= hi;
        if (ok && X.primitive) ok = pfa_box_gcd(g, (uint64_t)(v < 0 ? -v : v)) == 1;
#if defined(__HIP_DEVICE_COMPILE__)
        const unsigned long long bal = __ballot(ok);
        const int before = __popcll(bal & ((1ULL << T.lane) - 1));
        const int k = __popcll(bal);
#else
        const int before = 0, k = ok;
#endif
        if (ok && out) {
            const int64_t row = base + cnt + before;
            if (row < cap) {
                int64_t *o = out + (size_t)row * X.n;
                for (int c = 0; c < X.n - 1; ++c) o[c] = p[c];
                o[X.n - 1] = v;
            }
        }
        cnt += k;
    }
    return cnt;
}

// Points of unit u (prefix digits of u in base 2B+1, coordinate 0 most
// significant); written from row `base` when out != nullptr.  p, cur, hi
// (n int32) and g (n uint64) are team-shared scratch (LDS on the device).
template <int RPL>
PFA_HD static inline int64_t pfa_box_unit(const PfaBox &X, const PfaBoxTeam &T, int64_t u, int32_t *p,
                                          int32_t *cur, int32_t *hi, uint64_t *g, int64_t *out,
                                          int64_t base, int64_t cap)
{
    const int n = X.n, s = X.s, W = 2 * X.B + 1;
    int64_t part[RPL];
    for (int t = 0; t < RPL; ++t) part[t] = 0;
    PFA_SYNC_BOX();                     // the previous unit's reads of p are done
    // the prefix, most significant digit first
    int64_t rem = u;
    for (int j = s - 1; j >= 0; --j) {
        const int64_t dgt = rem % W;
        rem /= W;
        if (T.lane == 0) p[j] = (int32_t)(dgt - X.B);
    }
    PFA_SYNC_BOX();
    uint64_t gp = 0;
    for (int j = 0; j < s; ++j) {
        int64_t lo, hh;
        pfa_box_interval<RPL>(X, T, part, j, &lo, &hh);
        const int64_t v = p[j];
        if (v < lo || v > hh) return 0;
        if (j == n - 1) return pfa_box_last(X, T, p, gp, v, v, out, base, cap);
        pfa_box_apply<RPL>(X, T, part, j, v);
        gp = pfa_box_gcd(gp, (uint64_t)(v < 0 ? -v : v));
    }
    int64_t cnt = 0;
    int j = s;
    {
        int64_t lo, hh;
        pfa_box_interval<RPL>(X, T, part, j, &lo, &hh);
        if (j == n - 1) return pfa_box_last(X, T, p, gp, lo, hh, out, base, cap);
        if (lo > hh) return 0;
        if (T.lane == 0) {
            cur[j] = (int32_t)lo;
            hi[j] = (int32_t)hh;
            p[j] = (int32_t)lo;
            g[j] = pfa_box_gcd(gp, (uint64_t)(lo < 0 ? -lo : lo));
        }
        PFA_SYNC_BOX();
        pfa_box_apply<RPL>(X, T, part, j, lo);
    }
    for (;;) {
        // level j holds cur[j]: the child interval, then the leaves or a descent
        int64_t lo, hh;
        pfa_box_interval<RPL>(X, T, part, j + 1, &lo, &hh);
        if (j + 1 == n - 1) {
            cnt += pfa_box_last(X, T, p, g[j], lo, hh, out, base + cnt, cap);
        } else if (lo <= hh) {
            ++j;
            PFA_SYNC_BOX();
            if (T.lane == 0) {
                cur[j] = (int32_t)lo;
                hi[j] = (int32_t)hh;
                p[j] = (int32_t)lo;
                g[j] = pfa_box_gcd(g[j - 1], (uint64_t)(lo < 0 ? -lo : lo));
            }
            PFA_SYNC_BOX();
            pfa_box_apply<RPL>(X, T, part, j, lo);
            continue;
        }
        // next value at level j, backing up through exhausted levels
        for (;;) {
            const int64_t old = cur[j];
            pfa_box_apply<RPL>(X, T, part, j, -old);
            if (old < hi[j]) {
                const int64_t v = old + 1;
                PFA_SYNC_BOX();
                if (T.lane == 0) {
                    cur[j] = (int32_t)v;
                    p[j] = (int32_t)v;
                    g[j] = pfa_box_gcd(j > s ? g[j - 1] : gp, (uint64_t)(v < 0 ? -v : v));
                }
                PFA_SYNC_BOX();
                pfa_box_apply<RPL>(X, T, part, j, v);
                break;
            }
            if (--j < s) return cnt;
        }
    }
}
