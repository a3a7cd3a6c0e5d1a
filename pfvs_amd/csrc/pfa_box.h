// pfa_box.h -- the p-vector enumeration on the input side of coniZpM: all
// integer p with |p|_inf <= B and H p >= rhs (row-wise), optionally
// primitive (gcd of the entries 1), in lexicographic order (coordinate 0
// outermost, ascending).  The contract of latticepts.box_enum as the
// reference calls it (pfvs/pvectors.py:58-89, pfvs/distributed.py:121-139,
// tests/test_paper_examples.py:119-121; latticepts >= 0.1.2 is a
// dependency the reference does not vendor).
//
// The box is split into units: every assignment of the first s coordinates
// (a prefix).  A unit is enumerated depth first by one thread, a GPU thread
// on the device and a host thread on the host, with the same code
// (pfa_box_unit_thread): no barriers, no cross-lane traffic, every loop
// bounded by the box.  With part_r = sum_{k<j} H_rk p_k set and rest_r(j+1)
// = B sum_{k>j} |H_rk| the most the unset coordinates can add, p_j = v
// survives row r iff
//   H_rj v >= rhs_r - part_r - rest_r(j+1),
// an interval in v.  Exactly the points of the box satisfying every row
// reach the last level, where each value is tested for primitivity
// (gcd(g, |v|) == 1 with g the gcd of the prefix).  Counting and writing
// runs share the code, so a unit's points land at offset (points of
// earlier units) in order.
#pragma once
#include <stdint.h>
#include "pfa_int.h"

#define PFA_BOX_MAXM 256             // rows of H
#define PFA_BOX_MAXN 64              // coordinates

struct PfaBox {
    int m, n, B, primitive, s;
    const int32_t *Ht;               // (n, m): column j of H at Ht + j m
    const int64_t *rest;             // (n + 1, m): B sum_{k >= j} |H_rk| at rest + j m
    const int64_t *rhs;              // (m,)
};

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

// ------------------------------------------------------------------------
// One unit per thread.  The partial sums, current values, upper bounds and
// prefix gcds of thread tid live in a scratch array strided by the T
// threads of the launch (element k of thread tid at k * T + tid), so
// neighbouring GPU threads touch neighbouring words; a host thread uses
// T = 1.
struct PfaBoxScr {
    int64_t *part;                   // (m, T)
    int32_t *cur, *hi;               // (n, T)
    uint64_t *g;                     // (n, T)
    int tid, T;
};

PFA_HD static inline void pfa_box_t_interval(const PfaBox &X, const PfaBoxScr &S, int j, int64_t *lo_out,
                                             int64_t *hi_out)
{
    int64_t lo = -X.B, hi = X.B;
    const int32_t *a = X.Ht + (size_t)j * X.m;
    const int64_t *rs = X.rest + (size_t)(j + 1) * X.m;
    for (int r = 0; r < X.m; ++r) {
        const int64_t need = X.rhs[r] - S.part[(size_t)r * S.T + S.tid] - rs[r];   // a v >= need
        const int64_t ar = a[r];
        if (ar > 0) {
            const int64_t b = -pfa_box_fdiv(-need, ar);
            lo = b > lo ? b : lo;
        } else if (ar < 0) {
            const int64_t b = pfa_box_fdiv(need, ar);
            hi = b < hi ? b : hi;
        } else if (need > 0) {
            hi = lo - 1;
        }
    }
    *lo_out = lo;
    *hi_out = hi;
}

PFA_HD static inline void pfa_box_t_apply(const PfaBox &X, const PfaBoxScr &S, int j, int64_t v)
{
    const int32_t *a = X.Ht + (size_t)j * X.m;
    for (int r = 0; r < X.m; ++r) S.part[(size_t)r * S.T + S.tid] += (int64_t)a[r] * v;
}

// the values v in [lo, hi] of the last coordinate (primitive: gcd(g, |v|)
// == 1), counted, and written after the prefix cur[0 : n-1] when out
PFA_HD static inline int64_t pfa_box_t_last(const PfaBox &X, const PfaBoxScr &S, uint64_t g, int64_t lo,
                                            int64_t hi, int64_t *out, int64_t base, int64_t cap)
{
    int64_t cnt = 0;
    for (int64_t v = lo; v <= hi; ++v) {
        if (X.primitive && pfa_box_gcd(g, (uint64_t)(v < 0 ? -v : v)) != 1) continue;
        if (out && base + cnt < cap) {
            int64_t *o = out + (size_t)(base + cnt) * X.n;
            for (int c = 0; c < X.n - 1; ++c) o[c] = S.cur[(size_t)c * S.T + S.tid];
            o[X.n - 1] = v;
        }
        cnt++;
    }
    return cnt;
}

// points of unit u in order, written from row `base` when out != nullptr
PFA_HD static inline int64_t pfa_box_unit_thread(const PfaBox &X, const PfaBoxScr &S, int64_t u,
                                                 int64_t *out, int64_t base, int64_t cap)
{
    const int n = X.n, s = X.s;
    const int64_t W = 2 * (int64_t)X.B + 1;
#define PFA_BT(arr, k) (S.arr[(size_t)(k) * S.T + S.tid])
    for (int r = 0; r < X.m; ++r) PFA_BT(part, r) = 0;
    int64_t rem = u;
    for (int j = s - 1; j >= 0; --j) {
        PFA_BT(cur, j) = (int32_t)(rem % W - X.B);
        rem /= W;
    }
    uint64_t gp = 0;
    for (int j = 0; j < s; ++j) {
        int64_t lo, hh;
        pfa_box_t_interval(X, S, j, &lo, &hh);
        const int64_t v = PFA_BT(cur, j);
        if (v < lo || v > hh) return 0;
        if (j == n - 1) return pfa_box_t_last(X, S, gp, v, v, out, base, cap);
        pfa_box_t_apply(X, S, j, v);
        gp = pfa_box_gcd(gp, (uint64_t)(v < 0 ? -v : v));
    }
    int64_t cnt = 0;
    int j = s;
    {
        int64_t lo, hh;
        pfa_box_t_interval(X, S, j, &lo, &hh);
        if (j == n - 1) return pfa_box_t_last(X, S, gp, lo, hh, out, base, cap);
        if (lo > hh) return 0;
        PFA_BT(cur, j) = (int32_t)lo;
        PFA_BT(hi, j) = (int32_t)hh;
        PFA_BT(g, j) = pfa_box_gcd(gp, (uint64_t)(lo < 0 ? -lo : lo));
        pfa_box_t_apply(X, S, j, lo);
    }
    for (;;) {
        int64_t lo, hh;
        pfa_box_t_interval(X, S, j + 1, &lo, &hh);
        if (j + 1 == n - 1) {
            cnt += pfa_box_t_last(X, S, PFA_BT(g, j), lo, hh, out, base + cnt, cap);
        } else if (lo <= hh) {
            ++j;
            PFA_BT(cur, j) = (int32_t)lo;
            PFA_BT(hi, j) = (int32_t)hh;
            PFA_BT(g, j) = pfa_box_gcd(PFA_BT(g, j - 1), (uint64_t)(lo < 0 ? -lo : lo));
            pfa_box_t_apply(X, S, j, lo);
            continue;
        }
        for (;;) {
            const int64_t old = PFA_BT(cur, j);
            pfa_box_t_apply(X, S, j, -old);
            if (old < PFA_BT(hi, j)) {
                const int64_t v = old + 1;
                PFA_BT(cur, j) = (int32_t)v;
                PFA_BT(g, j) = pfa_box_gcd(j > s ? PFA_BT(g, j - 1) : gp, (uint64_t)(v < 0 ? -v : v));
                pfa_box_t_apply(X, S, j, v);
                break;
            }
            if (--j < s) return cnt;
        }
    }
#undef PFA_BT
}
