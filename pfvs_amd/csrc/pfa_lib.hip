// pfa_lib.hip -- libpfvs_amd.so: the coniZpM search on MI355X (gfx950).
//
// Pipeline per call (one HIP stream, inputs uploaded once):
//   k_prep      one wavefront (64 lanes; the LLL/HNF working set in LDS,
//               8.9 KB at h11 = 20, the matrix products' in global memory)
//               per p-vector, one launch per h11 size class: lattice, HNF,
//               FP64 Cholesky with its error margin, record.
//   k_roots     one lane per p-vector: top-level node -> frontier buffer.
//   k_expand    batched form: breadth-first expansion of the frontier, one
//               launch per level, until it holds enough nodes to start
//               every persistent wavefront with a deep deque.
//   search      PLAIN: k_search<true>, one launch per batch, one
//               wavefront per p-vector (the same wave-cooperative step,
//               no work sharing between p-vectors).  LANE: k_lane, one
//               thread per p-vector runs the whole depth-first search.
//               BATCHED: k_search<false>, persistent wavefronts each holding a
//               LIFO deque of nodes (its top 128 entries cached in LDS);
//               every step expands up to 64 sibling candidates of the top
//               nodes at once (prefix sum over their ranges), pushes
//               surviving children with a ballot and writes leaves into
//               chunks of point slots; deep deques donate their bottom to
//               a ticketed frontier ring that idle wavefronts take from;
//               an outstanding-work counter ends the kernel.
//   k_post      one lane per point slot: Kn = ZB c and M0 (exact double
//               FMAs, rows through the scalar cache), the K-integrality
//               count, the list of kept points.
//   k_filter    det((kappa M)[1:,1:]) != 0 for the kept points, one group
//               of h11 - 1 lanes per point (left-looking Bareiss);
//               k_det_mod decides the points whose values left 2^53.
// The host path runs the same preparation and search code on host threads;
// the finish (canonical order, copy-out, PFV assembly) runs on host threads
// after the device is released.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <type_traits>
#include <numeric>
#include <thread>
#include <string>
#include <vector>

#include "pfa_box.h"
#include "pfa_bound.h"
#include "pfa_int.h"
#include "pfa_post.h"
#include "pfa_prep.h"
#include "pfa_search.h"
#include "../../include/pfvs_amd.h"

#define PFA_MAX_H11 (PFA_DMAX + 1)

// ------------------------------------------------------------ record
// Per-problem record in 8-byte words; offsets depend on the batch's max d.
#define PFA_HDR 20                 // header words
struct RecL {
    int D;
    int U, Uinv, mat, lin, H, L, Bi, ZB, words;
};

static RecL rec_layout(int D)
{
    RecL r;
    r.D = D;
    int o = PFA_HDR;
    r.U = o; o += D * D;
    r.Uinv = o; o += D;
    r.mat = o; o += D * D;
    r.lin = o; o += D;
    r.H = o; o += D * D;
    r.L = o; o += 2 * (D + 1) * D;
    r.Bi = o; o += (D + 1) * D;
    r.ZB = o; o += (D + 1) * D;
    r.words = (o + 1) & ~1;
    return r;
}

// header (int32 view): [0] status [1] d [20] m0_level [21] leafg [22] nL
// [23] strict [24] h [25] code [26] geo [27] root proven empty (no factor,
// rows or Binter written); (int64 view) [1] qmax [2] Q
// [3] linmin [4..8] Qe QGmul phi slack Kerr (double) [9] hmask
// [14] max |Binter| [15] max |ZB| [16] max |mat| [17] max |H| of the
// pruning rows (int64 bounds for the search's exact prefix sums)
// A record's header fields and row pointers.  H64 / H32: the types the
// header is read through (plain, or the constant address space for a
// wave-uniform record on the device: the header does not change during a
// search, only the status word, which is not read here).
template <class H64, class H32>
PFA_HD static inline PfaProb rec_view_as(const int64_t *r, const RecL &L)
{
    const H64 *r64 = (const H64 *)r;
    const H32 *ri = (const H32 *)r;
    PfaProb P;
    P.d = ri[1];
    P.m0_level = ri[20];
    P.leafg = ri[21];
    P.nL = ri[22];
    P.strict = ri[23];
    P.root_empty = ri[27];
    P.qmax = r64[1];
    P.Q = r64[2];
    P.linmin = r64[3];
    P.Qe = __builtin_bit_cast(double, (int64_t)r64[4]);
    P.QGmul = __builtin_bit_cast(double, (int64_t)r64[5]);
    P.phi = __builtin_bit_cast(double, (int64_t)r64[6]);
    P.slack = __builtin_bit_cast(double, (int64_t)r64[7]);
    P.Kerr = __builtin_bit_cast(double, (int64_t)r64[8]);
    P.hmask = (uint64_t)r64[9];
    P.matmax = r64[16];
    P.hmax = r64[17];
    P.U = (const double *)(r + L.U);
    P.Uinv = (const double *)(r + L.Uinv);
    P.mat = r + L.mat;
    P.linvec = r + L.lin;
    P.H = r + L.H;
    P.L = r + L.L;
    return P;
}

PFA_HD static inline PfaProb rec_view(const int64_t *r, const RecL &L)
{
    return rec_view_as<int64_t, int32_t>(r, L);
}

// through the scalar cache, for a record every lane of the wavefront reads
typedef __attribute__((address_space(4))) const int64_t pfa_chdr64;
typedef __attribute__((address_space(4))) const int32_t pfa_chdr32;
__device__ static inline PfaProb rec_view_uni(const int64_t *r, const RecL &L)
{
    return rec_view_as<pfa_chdr64, pfa_chdr32>(r, L);
}

PFA_HD static inline int32_t &rec_status(int64_t *r) { return ((int32_t *)r)[0]; }

// A search that may meet several events in one problem (the batched form,
// whose wavefronts reach them in no fixed order) records them with a fixed
// precedence instead of last-writer-wins: PFA_HOST (the exact host path
// searches the problem again and decides it) over a kernel status (PFA_I32)
// over PFA_OK.  The result then does not depend on timing.
PFA_HD static inline int status_rank(int32_t s) { return s == PFA_OK ? 0 : (s == PFA_HOST ? 2 : 1); }

__device__ static inline void rec_raise_status(int64_t *r, int32_t s)
{
    int32_t *p = (int32_t *)r;
    int32_t old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (status_rank(s) > status_rank(old)) {
        const int32_t prev = atomicCAS(p, old, s);
        if (prev == old) return;
        old = prev;
    }
}

// scalar constants of the search for a problem of dimension d with factor
// margin phi (slack and Kerr as fpk_search_consts, fpk_search_impl.h:30-39)
PFA_HD static inline void rec_set_consts(int64_t *r, int d, int64_t qmax, double phi)
{
    double *rd = (double *)r;
    const double u = 0x1p-53;
    double inv = 1.0 / (1.0 - phi);
    double Qe = (double)qmax * inv * (1.0 + 8 * u);
    rd[4] = Qe;
    rd[5] = inv * (1.0 + 8 * u);
    rd[6] = phi;
    rd[7] = fmax(1e-4, 1e-9 * (Qe + 1.0));
    rd[8] = 2.0 * (double)(d + 8) * u;
}

struct GeoArgs {
    const int32_t *h;
    const int64_t *Q, *qmax, *linmin, *kappa, *Mbasis, *kofs, *mofs;
};

// --------------------------------------------------- preparation (team)
PFA_HD static inline void prep_one(const PfaTeam &T, char *ws_hot, char *ws_cold, const GeoArgs &G,
                                   const int64_t *ps, int MH, const int32_t *pgeo, int64_t ip,
                                   int64_t *rec, RecL L, bool sp = true)
{
    const int geo = pgeo[ip];
    const int h = G.h[geo], d = h - 1;
    int32_t *ri = (int32_t *)rec;
    const int64_t *p = ps + ip * MH;
    if (T.lane == 0) {
        for (int k = 0; k < PFA_HDR; ++k) rec[k] = 0;
        ri[0] = PFA_HOST;
        ri[1] = d;
        rec[1] = G.qmax[geo];
        rec[2] = G.Q[geo];
        rec[3] = G.linmin[geo];
        ri[23] = 1;
        ri[24] = h;
        ri[26] = geo;
    }
    PfaWS W;
    pfa_ws_carve(W, ws_hot, ws_cold, h, sp);
    if (pfa_empty_certificate(T, W, G.kappa + G.kofs[geo], p, G.qmax[geo], G.linmin[geo])) {
        // the search would enter the root and find no candidate: one node
        PFA_SYNC();
        if (T.lane == 0) {
            ri[20] = d - 1;
            ri[25] = 1;
            ri[27] = 1;
            rec_set_consts(rec, d, rec[1], 0.0);
            ri[0] = PFA_OK;
        }
        PFA_SYNC();
        return;
    }
    int st = pfa_build_lattice(T, W, G.kappa + G.kofs[geo], G.Mbasis + G.mofs[geo], p);
    if (st) {
        PFA_SYNC();
        if (T.lane == 0) ri[25] = st;
        return;
    }
    // H = HNF(ZB[1:]) over the GG and W1 regions (free after the last
    // LLL; the scratch hv, hw, hpiv shares the mu region)
    pfa_i128 *H = W.GG;
    int hst = pfa_hnf(T, W, W.ZB + d, d, d, H, W.hv, W.hw, W.hpiv);
    PFA_SYNC();
    // row levels of H (pivot columns) and whether each row fits int64;
    // H's rows go to the record before the factor reuses W1
    int *lvl = (int *)W.W3;             // d levels + d fit flags
    if (T.lane == 0) {
        for (int r = 0; r < d; ++r) {
            int lv = -1, fit = 1;
            if (!hst)
                for (int j = 0; j < d; ++j) {
                    pfa_i128 x = H[r * d + j];
                    if (lv < 0 && x != 0) lv = j;
                    if (!pfa_fits64(x)) fit = 0;
                }
            lvl[r] = lv;
            lvl[d + r] = fit;
        }
    }
    int64_t *rU = rec + L.U, *rM = rec + L.mat, *rH = rec + L.H, *rL = rec + L.L;
    for (int e = T.lane; e < d * d; e += T.n) rH[e] = 0;
    PFA_SYNC();
    for (int e = T.lane; e < d * d; e += T.n) {
        int r = e / d, j = e % d;
        if (lvl[r] >= 0 && lvl[d + r]) rH[lvl[r] * d + j] = (int64_t)H[r * d + j];
    }
    PFA_SYNC();
    double *U = (double *)W.W1, *X = (double *)W.W2, phi = 0.0;
    int cst = pfa_cholesky(T, W.mat, d, U, X, &phi, W.flag);
    if (cst) {
        PFA_SYNC();
        if (T.lane == 0) ri[25] = cst == PFA_NOT_PD ? -60 : -70;
        return;
    }
    for (int e = T.lane; e < d * d; e += T.n) {
        ((double *)rU)[e] = U[e];
        rM[e] = W.mat[e];
    }
    for (int e = T.lane; e < d; e += T.n) {
        ((double *)(rec + L.Uinv))[e] = 1.0 / U[e * d + e];
        rec[L.lin + e] = W.Bi[e];
    }
    for (int e = T.lane; e < h * d; e += T.n) {
        rec[L.Bi + e] = W.Bi[e];
        rec[L.ZB + e] = W.ZB[e];
    }
    if (T.lane == 0) {
        // entry bounds of Binter and ZB for the point transform
        int64_t mb = 0, mz = 0;
        for (int e = 0; e < h * d; ++e) {
            const int64_t b = W.Bi[e] < 0 ? -W.Bi[e] : W.Bi[e], z = W.ZB[e] < 0 ? -W.ZB[e] : W.ZB[e];
            mb = b > mb ? b : mb;
            mz = z > mz ? z : mz;
        }
        rec[14] = mb;
        rec[15] = mz;
    }
    PFA_SYNC();
    uint64_t hmask = 0;
    int leafg = hst != 0;
    for (int r = 0; r < d; ++r) {
        if (lvl[r] < 0) continue;
        if (lvl[d + r]) hmask |= 1ULL << lvl[r];
        else leafg = 1;
    }
    if (leafg) {
        // exact leaf gcd from Kn[1:] = ZB[1:] c (int64 rows)
        for (int e = T.lane; e < d * d; e += T.n) {
            rL[2 * e] = W.ZB[d + e];
            rL[2 * e + 1] = 0;
        }
    }
    int m0 = -1;
    for (int j = 0; j < d; ++j)
        if (W.Bi[j] != 0) { m0 = j; break; }
    PFA_SYNC();
    if (T.lane == 0) {
        rec[9] = (int64_t)hmask;
        // bounds of the search's exact prefix sums
        int64_t mm = 0, hm2 = 0;
        for (int e = 0; e < d * d; ++e) {
            const int64_t x = rM[e] < 0 ? -rM[e] : rM[e], y = rH[e] < 0 ? -rH[e] : rH[e];
            mm = x > mm ? x : mm;
            hm2 = y > hm2 ? y : hm2;
        }
        rec[16] = mm;
        rec[17] = hm2;
        ri[20] = m0;
        ri[21] = leafg;
        ri[22] = leafg ? d : 0;
        ri[25] = hst ? -15 : 0;
        rec_set_consts(rec, d, rec[1], phi);
        ri[0] = PFA_OK;
    }
    PFA_SYNC();
}

// order (optional): the p-vectors this launch prepares, by index (a size
// class of a mixed batch, its LDS sized for the class's largest h11).
// sp: the workspace holds the Gram-Schmidt running values (pfa_ws_carve)
__global__ __launch_bounds__(64) void k_prep(GeoArgs G, const int64_t *ps, int MH,
                                             const int32_t *pgeo, int64_t n, const int32_t *order,
                                             int64_t *recs, RecL L, char *cold, size_t cold_bytes,
                                             int sp)
{
    extern __shared__ __align__(16) char smem[];
    PfaTeam T{(int)threadIdx.x, 64};
    char *mine = cold + (size_t)blockIdx.x * cold_bytes;
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int64_t ip = order ? (int64_t)order[i] : i;
        prep_one(T, smem, mine, G, ps, MH, pgeo, ip, recs + ip * L.words, L, sp != 0);
        __syncthreads();
    }
}

// Factor-only preparation for the single-problem kernel API: the record
// already holds mat, lin, H, L and the header; computes U, Uinv, phi.
PFA_HD static inline void factor_one(const PfaTeam &T, char *wsmem, int64_t *rec, RecL L)
{
    int32_t *ri = (int32_t *)rec;
    const int d = ri[1];
    double *U = (double *)wsmem, *X = U + d * d;
    int *flag = (int *)(X + d * d);
    double phi = 0.0;
    int cst = pfa_cholesky(T, rec + L.mat, d, U, X, &phi, flag);
    if (cst) {
        PFA_SYNC();
        if (T.lane == 0) { ri[0] = PFA_HOST; ri[25] = cst == PFA_NOT_PD ? -60 : -70; }
        return;
    }
    for (int e = T.lane; e < d * d; e += T.n) ((double *)(rec + L.U))[e] = U[e];
    for (int e = T.lane; e < d; e += T.n) ((double *)(rec + L.Uinv))[e] = 1.0 / U[e * d + e];
    PFA_SYNC();
    if (T.lane == 0) {
        rec_set_consts(rec, d, rec[1], phi);
        ri[0] = PFA_OK;
    }
    PFA_SYNC();
}

__global__ __launch_bounds__(64) void k_factor(int64_t n, int64_t *recs, RecL L)
{
    extern __shared__ __align__(16) char smem[];
    PfaTeam T{(int)threadIdx.x, 64};
    for (int64_t ip = blockIdx.x; ip < n; ip += gridDim.x) {
        factor_one(T, smem, recs + ip * L.words, L);
        __syncthreads();
    }
}

// ------------------------------------------------------- frontier entries
// Entry: PfaNode (128 B) | pidx | pad | c[D] (int32)
struct EntL {
    int D, bytes;
};
static EntL ent_layout(int D) { return EntL{D, ((144 + 4 * D) + 15) & ~15}; }

// Frontier ring shared by all wavefronts: ticket t lives in slot t % cap;
// seq[slot] == t + 1 once the entry of ticket t is published.
struct QueueArgs {
    char *ring;
    unsigned long long *seq;
    unsigned long long cap;
    // words 16 k apart (one 128-byte line each): k = 0 head (consumer
    // tickets), 1 tail (producer tickets), 2 outstanding work units,
    // 3 waiting wavefronts, 4 consumed, 5 error (1 ring full, 2 deadline),
    // 6 point count, 8 point-buffer overflow
    unsigned long long *ctl;
};

// Breadth-first split of the search before the persistent launch: the
// roots go to frontier buffer 0, and expansion t turns the nodes of buffer
// t % 2 (all children of expansion t - 1) into their children in the other
// buffer, until a frontier holds `target` nodes or its children could
// outgrow the deques (`room`) or a buffer (`cap`).  The persistent search
// then starts every wavefront with a contiguous slice of the last frontier.
// st[t] nodes and st[64 + t] their candidates in the frontier of expansion
// t; st[128] stop flag, st[129] the expansion that stopped (the last
// frontier), st[130] its node count, st[131] its buffer.
#define PFA_SPLIT_MAX 48
struct SplitArgs {
    char *F0, *F1;
    unsigned long long cap;       // entries per buffer
    unsigned long long *st;
    unsigned long long target;    // stop once a frontier holds this many nodes
    unsigned long long room;      // most nodes the deques take at the start
};

struct PointArgs {
    int32_t *pidx;
    int64_t *q;
    uint64_t *g;                  // the leaf's gcd(Kn[1:]) (pfa_leaf), or null
    int32_t *c;                   // cap x D
    unsigned long long cap;
    unsigned long long *count;
    int *overflow;
};

struct CountArgs {
    unsigned long long *c;        // nodes, cand, leaf
};

// A node entry's coordinates to the lane's c (those at or below the node's
// level i are 0 there), and the lane's c to a child's entry (its entries
// below its level are 0 in c), four at a time with every read of a block
// issued before its writes: an element at a time, each write waited on its
// read.  The last block reads at most index D - 1.
__device__ static inline void coords_in(int32_t *c, const int32_t *ec, int D, int i)
{
    for (int k0 = 0; k0 < D; k0 += 4) {
        int32_t x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = ec[k0 + q < D ? k0 + q : D - 1];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (k0 + q < D) c[k0 + q] = k0 + q > i ? x[q] : 0;
    }
}

__device__ static inline void coords_out(int32_t *ec, const int32_t *c, int D)
{
    for (int k0 = 0; k0 < D; k0 += 4) {
        int32_t x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = c[k0 + q < D ? k0 + q : D - 1];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (k0 + q < D) ec[k0 + q] = x[q];
    }
}

__device__ static inline void entry_copy(char *dst, const char *src, int bytes)
{
    const int4 *s = (const int4 *)src;
    int4 *t = (int4 *)dst;
    for (int k = 0; k < bytes / 16; ++k) t[k] = s[k];
}

// one lane per p-vector: its root node (entered, counted as a node)
// appended to frontier buffer 0 with its candidate count
__global__ void k_roots(int64_t n, int64_t *recs, RecL L, SplitArgs S, EntL E, CountArgs C)
{
    int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    int64_t *rec = recs + ip * L.words;
    if (rec_status(rec) != PFA_OK) return;
    PfaProb P = rec_view(rec, L);
    if (pfa_trivially_empty(P)) return;
    PfaNode nd;
    int32_t c[PFA_DMAX];
    for (int j = 0; j < PFA_DMAX; ++j) c[j] = 0;
    if (pfa_root_enter(P, c, nd)) { rec_raise_status(rec, PFA_I32); return; }
    atomicAdd(&C.c[0], 1ULL);
    if (nd.hi < nd.cur) return;                          // no candidate: a node without children
    unsigned long long t = atomicAdd(&S.st[0], 1ULL);
    atomicAdd(&S.st[64], (unsigned long long)((long long)nd.hi - nd.cur + 1));
    char *e = S.F0 + t * E.bytes;
    *(PfaNode *)e = nd;
    ((int32_t *)(e + 128))[0] = (int32_t)ip;
    int32_t *ec = (int32_t *)(e + 144);
    for (int j = 0; j < E.D; ++j) ec[j] = 0;
}

// ------------------------------------------------ one lane per problem
// The reference-shaped form: a thread runs a whole problem's depth-first
// search (PFA_MODE_LANE).
__global__ void k_lane(int64_t n, int64_t *recs, RecL L, PointArgs Pt, int D, CountArgs C)
{
    int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    int64_t *rec = recs + ip * L.words;
    if (rec_status(rec) != PFA_OK) return;
    PfaProb P = rec_view(rec, L);
    PfaNode st[PFA_DMAX];
    int32_t c[PFA_DMAX];
    PfaCounts cnt{0, 0, 0};
    auto emit = [&](const int32_t *cc, int64_t q, uint64_t g) -> int {
        unsigned long long k = atomicAdd(Pt.count, 1ULL);
        if (k >= Pt.cap) { *Pt.overflow = 1; return 0; }
        Pt.pidx[k] = (int32_t)ip;
        Pt.q[k] = q;
        if (Pt.g) Pt.g[k] = g;
        for (int j = 0; j < D; ++j) Pt.c[k * D + j] = j < P.d ? cc[j] : 0;
        return 0;
    };
    int r = pfa_dfs(P, st, c, cnt, emit);
    if (r) rec_status(rec) = r;
    atomicAdd(&C.c[0], cnt.nodes);
    atomicAdd(&C.c[1], cnt.cand);
    atomicAdd(&C.c[2], cnt.leaf);
}

// --------------------------------------------- batched and plain forms
// k_search<false> (PFA_MODE_BATCHED): one persistent launch.  Each
// wavefront owns a deque of nodes (stack top for depth-first expansion,
// bottom for donation).  A step expands up to 64 candidates taken from the
// top nodes; children are pushed with a ballot; leaves are written with one
// atomic per wavefront.  Wavefronts whose deque runs dry take a ticket of
// the frontier ring and wait for its entry; busy wavefronts that see
// waiting ones donate their bottom (shallowest, largest) nodes.  `work`
// counts queued + in-process units, so waiting wavefronts leave exactly
// when the whole search is done; a wall-clock deadline bounds every wait
// and the launch itself.
//
// k_search<true> (PFA_MODE_PLAIN): the one-launch-per-batch form, one
// wavefront per p-vector.  Workgroup b searches p-vectors b, b + grid, ...
// from their roots with the same step (64 candidates of its top nodes per
// step, ballot pushes, chunked point slots), but no frontier ring, no
// donation and no stealing: a p-vector's tree stays on the wavefront that
// started it.  A deque that outgrows stack_cap stops the launch with error
// 3, and the host reruns it with twice the capacity.
struct BatchArgs {
    int64_t *recs;
    RecL L;
    EntL E;
    QueueArgs Q;
    char *stacks;
    int stack_cap, donate_max, donate_min, pt_chunk;
    PointArgs Pt;
    CountArgs C;
    unsigned long long deadline;      // wall-clock ticks (0: none)
    int64_t n_p;                      // plain form: the p-vectors
    SplitArgs S;                      // batched form: the split's last frontier
};

#define PFA_CSTRIDE (PFA_DMAX + 1)        // odd stride: conflict-free LDS rows

// Inclusive scans over the wavefront's 64 lanes with DPP moves (row_shr
// 1, 2, 4, 8 within rows of 16, then row_bcast 15 and 31 across rows): no
// LDS round trip per step, as a shuffle has.  A lane whose source is out
// of its row, or whose row is masked off, takes the identity (0, or -1 for
// the max of values >= -1).  Checked against the shuffle scans on 20,000
// waves of random values, partly active waves included.
template <int CTRL, int RM>
__device__ static inline int dpp_mov(int v, int id)
{
    return __builtin_amdgcn_update_dpp(id, v, CTRL, RM, 0xf, false);
}

__device__ static inline int wave_scan_incl(int v)
{
    v += dpp_mov<0x111, 0xf>(v, 0);
    v += dpp_mov<0x112, 0xf>(v, 0);
    v += dpp_mov<0x114, 0xf>(v, 0);
    v += dpp_mov<0x118, 0xf>(v, 0);
    v += dpp_mov<0x142, 0xa>(v, 0);
    v += dpp_mov<0x143, 0xc>(v, 0);
    return v;
}

__device__ static inline int wave_scan_max(int v)
{
    v = max(v, dpp_mov<0x111, 0xf>(v, -1));
    v = max(v, dpp_mov<0x112, 0xf>(v, -1));
    v = max(v, dpp_mov<0x114, 0xf>(v, -1));
    v = max(v, dpp_mov<0x118, 0xf>(v, -1));
    v = max(v, dpp_mov<0x142, 0xa>(v, -1));
    v = max(v, dpp_mov<0x143, 0xc>(v, -1));
    return v;
}

template <int CTRL, int RM>
__device__ static inline long long dpp_mov64(long long v)
{
    const int lo = dpp_mov<CTRL, RM>((int)(uint32_t)(uint64_t)v, 0);
    const int hi = dpp_mov<CTRL, RM>((int)(uint32_t)((uint64_t)v >> 32), 0);
    return (long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ static inline unsigned long long ld_agent(unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// publish stack entries [from, from + k) at tickets t .. t + k - 1 (lane l: entry l)
__device__ static inline void ring_publish(const BatchArgs &a, const char *stk, int from, int k,
                                           unsigned long long t, int lane)
{
    const int ES = a.E.bytes;
    for (int l = lane; l < k; l += 64) {
        unsigned long long tk = t + l, slot = tk % a.Q.cap;
        entry_copy(a.Q.ring + slot * ES, stk + (size_t)(from + l) * ES, ES);
        __threadfence();
        atomicExch(&a.Q.seq[slot], tk + 1);
    }
}

// hand k bottom entries to the ring (lane 0 reserves); returns k or 0
__device__ static inline int donate(const BatchArgs &a, char *stk, int bot, int k, int forced,
                                    int lane, unsigned long long *s_u64)
{
    if (lane == 0) {
        unsigned long long tail = ld_agent(&a.Q.ctl[16]), used = ld_agent(&a.Q.ctl[64]);
        unsigned long long room = a.Q.cap - (tail - used);
        unsigned long long t = ~0ULL;
        // keep room for every other wavefront's largest donation
        if (room > (unsigned long long)k + 64ULL * 4096ULL) {
            atomicAdd(&a.Q.ctl[32], (unsigned long long)k);          // units before tickets
            t = atomicAdd(&a.Q.ctl[16], (unsigned long long)k);
        } else if (forced) {
            atomicExch(&a.Q.ctl[80], 1ULL);
        }
        s_u64[1] = t;
    }
    __syncthreads();
    unsigned long long t = s_u64[1];
    __syncthreads();
    if (t == ~0ULL) return 0;
    ring_publish(a, stk, bot, k, t, lane);
    __syncthreads();
    return k;
}

// The top of each wavefront's deque is cached in LDS: entries with index
// >= cb live in a circular buffer of PFA_SL entries (slot = index mod
// PFA_SL) and are authoritative there; entries below cb are in the deque's
// global array at their own index.  Every step's node reads and child
// pushes touch the top only, so they stay in LDS; entries move to and from
// global memory only when the cached part outgrows the buffer, when the
// top 64 are not all cached, and for donation and compaction.
#define PFA_SL 128

// copy deque entries [from, to) between the LDS cache and the global array
// (to_global: LDS -> global), wave-cooperatively in 16-byte words
__device__ static inline void cache_move(char *cache, char *stk, int from, int to, int ES,
                                         int lane, bool to_global)
{
    __syncthreads();                       // earlier LDS writes of any lane first
    const int w = ES / 16, total = (to - from) * w;
    for (int t = lane; t < total; t += 64) {
        const int idx = from + t / w, k = t % w;
        int4 *g = (int4 *)(stk + (size_t)idx * ES) + k;
        int4 *l = (int4 *)(cache + (size_t)(idx & (PFA_SL - 1)) * ES) + k;
        if (to_global) *g = *l;
        else *l = *g;
    }
    __syncthreads();
}

template <bool PLAIN>
__global__ __launch_bounds__(64) void k_search(BatchArgs a)
{
    const int lane = threadIdx.x;
    const int ES = a.E.bytes, D = a.E.D;
    char *stk = a.stacks + (size_t)blockIdx.x * a.stack_cap * ES;
    extern __shared__ __align__(16) char s_cache[];     // PFA_SL entries
    __shared__ int s_incl[64];
    __shared__ int s_mark[64];
    __shared__ unsigned long long s_u64[2];
    __shared__ int32_t s_c[64 * PFA_CSTRIDE];
    int32_t *c = s_c + lane * PFA_CSTRIDE;
    auto ent = [&](int idx) -> char * { return s_cache + (size_t)(idx & (PFA_SL - 1)) * ES; };
    int cb = 0;                                         // cache base
    const unsigned long long t0 = wall_clock64();
    const unsigned long long c0 = clock64();
    unsigned long long n_nodes = 0, n_cand = 0, n_leaf = 0, n_wait = 0, n_fetch = 0, n_don = 0;
    int bot = 0, top = 0, have_unit = 0, steps = 0;
    unsigned long long waiting = 0;
    unsigned long long pt_next = 0, pt_end = 0;        // this wavefront's point slots
    int64_t ip_next = blockIdx.x;                       // plain form: next p-vector
    if (!PLAIN) {
        // this wavefront's contiguous slice of the split's last frontier
        // becomes its deque (in the global array; nothing cached yet)
        const unsigned long long nf = a.S.st[130];
        const char *Ff = a.S.st[131] ? a.S.F1 : a.S.F0;
        const unsigned long long lo = nf * blockIdx.x / gridDim.x;
        const unsigned long long hi = nf * (blockIdx.x + 1) / gridDim.x;
        if (hi > lo) {
            const int n = (int)(hi - lo);
            const int4 *src = (const int4 *)(Ff + lo * ES);
            int4 *dst = (int4 *)stk;
            for (int k = lane; k < n * (ES / 16); k += 64) dst[k] = src[k];
            top = n;
            cb = n;
            have_unit = 1;
        }
    }
    for (;;) {
        if (PLAIN && top == bot) {
            // the next p-vector of this wavefront whose root opens: its root
            // node to cache slot 0 (counted as a node, as k_roots does)
            int got = 0;
            while (!got && ip_next < a.n_p) {
                const int64_t ip = ip_next;
                ip_next += gridDim.x;
                if (lane == 0) {
                    int64_t *rec = a.recs + ip * a.L.words;
                    int ok = 0;
                    if (rec_status(rec) == PFA_OK) {
                        PfaProb P = rec_view(rec, a.L);
                        if (!pfa_trivially_empty(P)) {
                            PfaNode nd;
                            for (int j = 0; j < PFA_DMAX; ++j) c[j] = 0;
                            if (pfa_root_enter(P, c, nd)) {
                                rec_raise_status(rec, PFA_I32);
                            } else {
                                char *e = ent(0);
                                *(PfaNode *)e = nd;
                                ((int32_t *)(e + 128))[0] = (int32_t)ip;
                                int32_t *ec = (int32_t *)(e + 144);
                                for (int j = 0; j < D; ++j) ec[j] = 0;
                                n_nodes++;
                                ok = 1;
                            }
                        }
                    }
                    s_u64[0] = ok;
                }
                __syncthreads();
                got = (int)s_u64[0];
                __syncthreads();
            }
            if (!got) break;
            bot = 0;
            top = 1;
            cb = 0;
        }
        if (!PLAIN && top == bot) {
            if (lane == 0) {
                if (have_unit) atomicAdd(&a.Q.ctl[32], ~0ULL);          // unit done
                unsigned long long h = atomicAdd(&a.Q.ctl[0], 1ULL), slot = h % a.Q.cap;
                atomicAdd(&a.Q.ctl[48], 1ULL);
                int got = 0;
                const unsigned long long w0 = clock64();
                for (unsigned it = 0;; ++it) {
                    if (ld_agent(&a.Q.seq[slot]) == h + 1) { got = 1; break; }
                    if ((it & 15) == 15) {
                        if (ld_agent(&a.Q.ctl[32]) == 0) break;
                        if (ld_agent(&a.Q.ctl[80])) break;
                        if (a.deadline && wall_clock64() - t0 > a.deadline) {
                            atomicExch(&a.Q.ctl[80], 2ULL);
                            break;
                        }
                    }
                    __builtin_amdgcn_s_sleep(8);
                }
                n_wait += clock64() - w0;
                n_fetch++;
                atomicAdd(&a.Q.ctl[48], ~0ULL);
                s_u64[0] = got ? h : ~0ULL;
            }
            have_unit = 0;
            __syncthreads();
            const unsigned long long h = s_u64[0];
            __syncthreads();
            if (h == ~0ULL) break;
            __threadfence();                               // acquire the entry
            const int4 *src = (const int4 *)(a.Q.ring + (h % a.Q.cap) * ES);
            int4 *dst = (int4 *)ent(0);
            for (int k = lane; k < ES / 16; k += 64) dst[k] = src[k];
            __syncthreads();
            if (lane == 0) atomicAdd(&a.Q.ctl[64], 1ULL);
            bot = 0;
            top = 1;
            cb = 0;
            have_unit = 1;
        }
        if ((++steps & 255) == 0 && a.deadline) {
            // stop at the deadline (error 2) or once another wavefront has
            // recorded an error (kept: the first error stands)
            if (lane == 0) {
                const int late = wall_clock64() - t0 > a.deadline;
                if (late) atomicCAS(&a.Q.ctl[80], 0ULL, 2ULL);
                s_u64[0] = (late || ld_agent(&a.Q.ctl[80])) ? 1 : 0;
            }
            __syncthreads();
            const int stop = (int)s_u64[0];
            __syncthreads();
            if (stop) break;
        }
        // gather up to 64 candidates from the top nodes (lane l: node top-1-l),
        // all of them cached
        const int live = top - bot;
        const int ntop = live < 64 ? live : 64;
        if (cb > top - ntop) {
            cache_move(s_cache, stk, top - ntop, cb, ES, lane, false);
            cb = top - ntop;
        }
        int cur = 0, hi = -1;
        if (lane < ntop) {
            const PfaNode *nd = (const PfaNode *)ent(top - 1 - lane);
            cur = nd->cur;
            hi = nd->hi;
        }
        long long cl = lane < ntop ? (long long)hi - cur + 1 : 0;
        int cnt = cl <= 0 ? 0 : (cl > 64 ? 65 : (int)cl);
        const int incl = wave_scan_incl(cnt);
        s_incl[lane] = incl;
        s_mark[lane] = -1;
        const int total = __builtin_amdgcn_readlane(incl, 63);
        const int take = total < 64 ? total : 64;
        const unsigned long long full = __ballot(lane < ntop && incl <= 64);
        const int F = __popcll(full);
        // lane -> node: each node with candidates marks its first lane (they
        // are distinct), and a running max over the lanes carries the node
        // to the lanes after it -- the smallest node whose scan exceeds the
        // lane, as a binary search over s_incl finds it
        __syncthreads();
        if (cnt > 0 && incl - cnt < 64) s_mark[incl - cnt] = lane;
        __syncthreads();
        int j = wave_scan_max(s_mark[lane]);
        if (lane >= take) j = -1;
        PfaNode nd;
        int32_t pidx = 0, v = 0;
        if (j >= 0) {
            const char *e = ent(top - 1 - j);
            nd = *(const PfaNode *)e;
            pidx = ((const int32_t *)(e + 128))[0];
            // (one loop of the wave-uniform length D: no divergent trip counts)
            coords_in(c, (const int32_t *)(e + 144), D, nd.i);
            v = nd.cur + (lane - (j > 0 ? s_incl[j - 1] : 0));
        }
        int part_new_cur = 0;
        if (F < ntop && lane == 0) {
            const PfaNode *pn = (const PfaNode *)ent(top - 1 - F);
            part_new_cur = pn->cur + (64 - (F > 0 ? s_incl[F - 1] : 0));
        }
        __syncthreads();                     // every parent read before any write
        int newtop = top - F;
        if (F < ntop && lane == 0)
            ((PfaNode *)ent(top - 1 - F))->cur = part_new_cur;
        int has_child = 0, emit = 0;
        int64_t q = 0;
        uint64_t lg = 0;
        PfaNode ch;
        // one p-vector for every candidate of the step (the usual case): its
        // header through the scalar cache
        const int32_t p0 = __builtin_amdgcn_readfirstlane(j >= 0 ? pidx : -1);
        const bool uni = p0 >= 0 && !__ballot(j >= 0 && pidx != p0);
        if (j >= 0) {
            const int64_t *rec = a.recs + (int64_t)pidx * a.L.words;
            PfaProb P = uni ? rec_view_uni(a.recs + (int64_t)p0 * a.L.words, a.L) : rec_view(rec, a.L);
            n_cand++;
            int r = pfa_cand(P, nd, v, ch);
            if (r == PFA_CAND_OVF) {
                rec_raise_status((int64_t *)rec, PFA_HOST);
            } else if (r == 1) {
                c[nd.i] = v;
                if (nd.i > 0) {
                    n_nodes++;
                    if (pfa_enter(P, c, ch)) rec_raise_status((int64_t *)rec, PFA_I32);
                    else has_child = ch.hi >= ch.cur;
                } else {
                    n_leaf++;
                    int lr = pfa_leaf(P, ch, c, &q, &lg);
                    if (lr == PFA_CAND_OVF) rec_raise_status((int64_t *)rec, PFA_HOST);
                    else emit = lr == 1;
                }
            }
        }
        const unsigned long long em = __ballot(emit);
        if (em) {
            const int ne = __popcll(em);
            if (pt_next + ne > pt_end) {
                // a new chunk of point slots; the rest of the old one is
                // marked empty (pidx -1)
                if (lane == 0) s_u64[0] = atomicAdd(a.Pt.count, (unsigned long long)a.pt_chunk);
                __syncthreads();
                for (unsigned long long k = pt_next + lane; k < pt_end; k += 64)
                    if (k < a.Pt.cap) a.Pt.pidx[k] = -1;
                pt_next = s_u64[0];
                pt_end = pt_next + a.pt_chunk;
                __syncthreads();
            }
            if (emit) {
                unsigned long long k = pt_next + __popcll(em & ((1ULL << lane) - 1));
                if (k < a.Pt.cap) {
                    a.Pt.pidx[k] = pidx;
                    a.Pt.q[k] = q;
                    if (a.Pt.g) a.Pt.g[k] = lg;
                    for (int t = 0; t < D; ++t) a.Pt.c[k * D + t] = c[t];
                } else {
                    *a.Pt.overflow = 1;
                }
            }
            pt_next += ne;
        }
        const unsigned long long cm = __ballot(has_child);
        const int nch = __popcll(cm);
        if (newtop + nch > a.stack_cap) {
            __syncthreads();
            // the whole deque to the global array, then compact or donate there
            cache_move(s_cache, stk, cb, newtop, ES, lane, true);
            if (bot > 0) {
                // compact [bot, newtop) to 0 in chunks of bot entries
                for (int base = bot; base < newtop; base += bot) {
                    int n = newtop - base < bot ? newtop - base : bot;
                    for (int k = lane; k < n; k += 64)
                        entry_copy(stk + (size_t)(base - bot + k) * ES, stk + (size_t)(base + k) * ES, ES);
                    __syncthreads();
                }
                newtop -= bot;
                bot = 0;
            }
            if (PLAIN && newtop + nch > a.stack_cap) {
                // no donation in the plain form: the host reruns the launch
                // with a larger deque
                if (lane == 0) atomicCAS(&a.Q.ctl[80], 0ULL, 3ULL);
                break;
            }
            if (newtop + nch > a.stack_cap) {
                int k = newtop / 2;
                if (!donate(a, stk, 0, k, 1, lane, s_u64)) break;     // ring full: error set
                for (int base = k; base < newtop; base += k) {
                    int n = newtop - base < k ? newtop - base : k;
                    for (int t = lane; t < n; t += 64)
                        entry_copy(stk + (size_t)(base - k + t) * ES, stk + (size_t)(base + t) * ES, ES);
                    __syncthreads();
                }
                newtop -= k;
            }
            cb = newtop;                     // nothing cached
        }
        if (newtop + nch - cb > PFA_SL) {    // room for the children: oldest cached out
            const int nb = newtop + nch - PFA_SL;
            cache_move(s_cache, stk, cb, nb, ES, lane, true);
            cb = nb;
        }
        if (has_child) {
            int slot = newtop + __popcll(cm & ((1ULL << lane) - 1));
            char *e = ent(slot);
            *(PfaNode *)e = ch;
            ((int32_t *)(e + 128))[0] = pidx;
            coords_out((int32_t *)(e + 144), c, D);   // (c is 0 at and below ch.i)
        }
        top = newtop + nch;
        __syncthreads();                     // writes visible to the next step
        // feed waiting wavefronts from the bottom of the deque (the shallowest
        // nodes: the largest subtrees), only from deques deep enough to keep
        // this wavefront's own steps full; the waiting count is refreshed
        // every 16 steps
        if (!PLAIN && (steps & 15) == 0) {
            if (lane == 0) s_u64[0] = ld_agent(&a.Q.ctl[48]);
            __syncthreads();
            waiting = s_u64[0];
            __syncthreads();
        }
        if (!PLAIN && waiting > 0 && top - bot >= a.donate_min) {
            {
                int k = (top - bot) / 4;
                if ((unsigned long long)k > waiting) k = (int)waiting;
                if (k > a.donate_max) k = a.donate_max;
                waiting -= k;
                if (cb < bot + k) {          // donated entries come from the global array
                    cache_move(s_cache, stk, cb, bot + k, ES, lane, true);
                    cb = bot + k;
                }
                int got = donate(a, stk, bot, k, 0, lane, s_u64);
                bot += got;
                n_don += got;
            }
        }
    }
    for (unsigned long long k = pt_next + lane; k < pt_end; k += 64)
        if (k < a.Pt.cap) a.Pt.pidx[k] = -1;
    if (n_nodes) atomicAdd(&a.C.c[0], n_nodes);
    if (n_cand) atomicAdd(&a.C.c[1], n_cand);
    if (n_leaf) atomicAdd(&a.C.c[2], n_leaf);
    if (lane == 0) {
        atomicAdd(&a.C.c[3], (unsigned long long)steps);
        atomicAdd(&a.C.c[4], n_wait);
        atomicAdd(&a.C.c[5], clock64() - c0);
        atomicAdd(&a.C.c[6], n_don);
        atomicAdd(&a.C.c[7], n_fetch);
    }
}

// Expansion t of the breadth-first split (SplitArgs): every node of the
// frontier in buffer t % 2 has each of its candidates tested (pfa_cand);
// survivors above level 0 are entered (pfa_enter) and, when they have
// candidates, appended to the other buffer; level-0 survivors are leaves
// (pfa_leaf) written to the point buffer.  A workgroup (one wavefront)
// takes 64 nodes at a time, stages them in LDS and runs their candidates
// 64 per round (lane -> candidate by a prefix sum over the nodes' ranges).
// Every workgroup reads the same frontier sizes and so takes the same
// decision to expand or to stop.  The same nodes, candidates and leaves are
// counted as in the depth-first forms.
__device__ static inline long long wave_scan_incl64(long long v)
{
    v += dpp_mov64<0x111, 0xf>(v);
    v += dpp_mov64<0x112, 0xf>(v);
    v += dpp_mov64<0x114, 0xf>(v);
    v += dpp_mov64<0x118, 0xf>(v);
    v += dpp_mov64<0x142, 0xa>(v);
    v += dpp_mov64<0x143, 0xc>(v);
    return v;
}

__global__ __launch_bounds__(64) void k_expand(int t, int64_t *recs, RecL L, EntL E, SplitArgs S,
                                               PointArgs Pt, CountArgs C, int pt_chunk)
{
    const int lane = threadIdx.x;
    unsigned long long *st = S.st;
    if (st[128]) return;
    const unsigned long long n_in = st[t], cs = st[64 + t];
    if (n_in == 0 || n_in >= S.target || cs > S.room || cs > S.cap || t + 1 >= 64) {
        if (blockIdx.x == 0 && lane == 0) {
            st[129] = t;
            st[128] = 1;
        }
        return;
    }
    const int ES = E.bytes, D = E.D;
    const char *Fin = (t & 1) ? S.F1 : S.F0;
    char *Fout = (t & 1) ? S.F0 : S.F1;
    extern __shared__ __align__(16) char s_nodes[];      // 64 entries
    __shared__ long long s_incl[64];
    __shared__ int s_mark[64];
    __shared__ unsigned long long s_u64[2];
    __shared__ int32_t s_c[64 * PFA_CSTRIDE];
    int32_t *c = s_c + lane * PFA_CSTRIDE;
    unsigned long long n_nodes = 0, n_cand = 0, n_leaf = 0;
    unsigned long long n_cc = 0;                         // the children's candidates (wave total)
    unsigned long long pt_next = 0, pt_end = 0;
    for (unsigned long long base = (unsigned long long)blockIdx.x * 64ULL; base < n_in;
         base += (unsigned long long)gridDim.x * 64ULL) {
        const int nb = n_in - base < 64 ? (int)(n_in - base) : 64;
        __syncthreads();                                  // the previous block's reads done
        {
            const int4 *src = (const int4 *)(Fin + base * ES);
            int4 *dst = (int4 *)s_nodes;
            for (int k = lane; k < nb * (ES / 16); k += 64) dst[k] = src[k];
        }
        __syncthreads();
        long long cnt = 0;
        if (lane < nb) {
            const PfaNode *nd = (const PfaNode *)(s_nodes + (size_t)lane * ES);
            cnt = (long long)nd->hi - nd->cur + 1;
            if (cnt < 0) cnt = 0;
        }
        const long long incl = wave_scan_incl64(cnt);
        s_incl[lane] = incl;
        __syncthreads();
        const long long total = s_incl[63];
        for (long long r0 = 0; r0 < total; r0 += 64) {
            const long long g = r0 + lane;
            // lane -> node, as in the search step: each node whose
            // candidates meet the round's window marks its first lane in it
            // (at most one node covers the window's first lane), then a
            // running max over the lanes
            s_mark[lane] = -1;
            __syncthreads();
            const long long ex = incl - cnt;
            if (cnt > 0 && incl > r0 && ex < r0 + 64) s_mark[ex > r0 ? (int)(ex - r0) : 0] = lane;
            __syncthreads();
            int j = wave_scan_max(s_mark[lane]);
            if (g >= total) j = -1;
            __syncthreads();                          // the marks read before the next round's
            PfaNode nd;
            int32_t pidx = 0, v = 0;
            if (j >= 0) {
                const char *e = s_nodes + (size_t)j * ES;
                nd = *(const PfaNode *)e;
                pidx = ((const int32_t *)(e + 128))[0];
                coords_in(c, (const int32_t *)(e + 144), D, nd.i);
                v = nd.cur + (int32_t)(g - (j > 0 ? s_incl[j - 1] : 0));
            }
            int has_child = 0, emit = 0;
            int64_t q = 0;
            uint64_t lg = 0;
            PfaNode ch;
            const int32_t p0 = __builtin_amdgcn_readfirstlane(j >= 0 ? pidx : -1);
            const bool uni = p0 >= 0 && !__ballot(j >= 0 && pidx != p0);
            if (j >= 0) {
                const int64_t *rec = recs + (int64_t)pidx * L.words;
                PfaProb P = uni ? rec_view_uni(recs + (int64_t)p0 * L.words, L) : rec_view(rec, L);
                n_cand++;
                int r = pfa_cand(P, nd, v, ch);
                if (r == PFA_CAND_OVF) {
                    rec_raise_status((int64_t *)rec, PFA_HOST);
                } else if (r == 1) {
                    c[nd.i] = v;
                    if (nd.i > 0) {
                        n_nodes++;
                        if (pfa_enter(P, c, ch)) rec_raise_status((int64_t *)rec, PFA_I32);
                        else has_child = ch.hi >= ch.cur;
                    } else {
                        n_leaf++;
                        int lr = pfa_leaf(P, ch, c, &q, &lg);
                        if (lr == PFA_CAND_OVF) rec_raise_status((int64_t *)rec, PFA_HOST);
                        else emit = lr == 1;
                    }
                }
            }
            const unsigned long long em = __ballot(emit);
            if (em) {
                const int ne = __popcll(em);
                if (pt_next + ne > pt_end) {
                    if (lane == 0) s_u64[0] = atomicAdd(Pt.count, (unsigned long long)pt_chunk);
                    __syncthreads();
                    for (unsigned long long k = pt_next + lane; k < pt_end; k += 64)
                        if (k < Pt.cap) Pt.pidx[k] = -1;
                    pt_next = s_u64[0];
                    pt_end = pt_next + pt_chunk;
                    __syncthreads();
                }
                if (emit) {
                    unsigned long long k = pt_next + __popcll(em & ((1ULL << lane) - 1));
                    if (k < Pt.cap) {
                        Pt.pidx[k] = pidx;
                        Pt.q[k] = q;
                        if (Pt.g) Pt.g[k] = lg;
                        for (int tt = 0; tt < D; ++tt) Pt.c[k * D + tt] = c[tt];
                    } else {
                        *Pt.overflow = 1;
                    }
                }
                pt_next += ne;
            }
            const unsigned long long cm = __ballot(has_child);
            if (cm) {
                const int nch = __popcll(cm);
                // the children's candidates, summed over the wavefront
                long long cc = has_child ? (long long)ch.hi - ch.cur + 1 : 0;
                for (int o = 32; o > 0; o >>= 1) cc += __shfl_xor(cc, o, 64);
                if (lane == 0) s_u64[0] = atomicAdd(&st[t + 1], (unsigned long long)nch);
                n_cc += (unsigned long long)cc;
                __syncthreads();
                const unsigned long long ob = s_u64[0];
                __syncthreads();
                if (has_child) {
                    // (cap >= the candidates of this frontier >= its children)
                    char *e = Fout + (ob + __popcll(cm & ((1ULL << lane) - 1))) * (unsigned long long)ES;
                    *(PfaNode *)e = ch;
                    ((int32_t *)(e + 128))[0] = pidx;
                    coords_out((int32_t *)(e + 144), c, D);   // (c is 0 at and below ch.i)
                }
            }
        }
    }
    for (unsigned long long k = pt_next + lane; k < pt_end; k += 64)
        if (k < Pt.cap) Pt.pidx[k] = -1;
    for (int o = 32; o > 0; o >>= 1) {
        n_nodes += __shfl_xor(n_nodes, o, 64);
        n_cand += __shfl_xor(n_cand, o, 64);
        n_leaf += __shfl_xor(n_leaf, o, 64);
    }
    if (lane == 0) {
        if (n_nodes) atomicAdd(&C.c[0], n_nodes);
        if (n_cand) atomicAdd(&C.c[1], n_cand);
        if (n_leaf) atomicAdd(&C.c[2], n_leaf);
        if (n_cc) atomicAdd(&st[64 + t + 1], n_cc);
    }
}

// After the split: the last frontier (node count, buffer) for the
// persistent search, and one outstanding work unit per wavefront that gets
// a nonempty slice of it (min(n, W) of the W wavefronts).
__global__ void k_split_final(SplitArgs S, unsigned long long *ctl, int W, int n_exp)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long *st = S.st;
    const int t = st[128] ? (int)st[129] : n_exp;
    const unsigned long long n = st[t];
    st[130] = n;
    st[131] = (unsigned long long)(t & 1);
    ctl[32] = n < (unsigned long long)W ? n : (unsigned long long)W;
}

// n bytes at p set to v (the buffer clears of a captured segment: a graph
// gets kernel nodes, not memset nodes); 16-byte stores where aligned
__global__ void k_fill(uint8_t *p, size_t n, uint8_t v)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t head = (16 - ((uintptr_t)p & 15)) & 15, h = head < n ? head : n;
    const size_t n16 = (n - h) / 16;
    const uint32_t w = 0x01010101u * v;
    uint4 *q = (uint4 *)(p + h);
    for (size_t i = t0; i < n16; i += stride) q[i] = make_uint4(w, w, w, w);
    for (size_t i = t0; i < h; i += stride) p[i] = v;
    for (size_t i = h + 16 * n16 + t0; i < n; i += stride) p[i] = v;
}

// Status words of the records: op 0 saves them to buf, op 1 restores them
// from buf (a search changes only the status word of a record), op 2
// gathers (status, stage code) pairs for the copy back.
__global__ void k_status_io(int64_t n, int64_t *recs, int words, int32_t *buf, int op)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int32_t *ri = (int32_t *)(recs + i * words);
        if (op == 0) buf[i] = ri[0];
        else if (op == 1) ri[0] = buf[i];
        else { buf[2 * i] = ri[0]; buf[2 * i + 1] = ri[25]; }
    }
}

// ------------------------------------------------------------ post
struct PostArgs {
    int64_t n;
    const int32_t *pidx;
    const int64_t *q;
    const int32_t *c;
    int D, MH;
    const int64_t *recs;
    RecL L;
    int64_t *M, *Kn;
    int64_t *nexp;
    int8_t *bad;                  // per problem: M/Kn overflow
    int filter;
};

PFA_HD static inline void post_one(const PostArgs &a, int64_t k)
{
    const int32_t ip = a.pidx[k];
    if (ip < 0) {                                  // an unused point slot
        if (a.filter) a.nexp[k] = -1;
        return;
    }
    const int64_t *rec = a.recs + (int64_t)ip * a.L.words;
    const int32_t *ri = (const int32_t *)rec;
    const int d = ri[1], h = d + 1;
    const int32_t *c = a.c + k * a.D;
    const int64_t *Bi = rec + a.L.Bi, *ZB = rec + a.L.ZB;
    int64_t *M = a.M + k * a.MH, *Kn = a.Kn + k * a.MH;
    int bad = 0;
    for (int i = 0; i < a.MH; ++i) {
        if (i >= h) { M[i] = 0; Kn[i] = 0; continue; }
        pfa_i128 sm = 0, sk = 0;
        for (int t = 0; t < d; ++t) {
            sm += (pfa_i128)Bi[i * d + t] * c[t];
            sk += (pfa_i128)ZB[i * d + t] * c[t];
        }
        if (!pfa_fits64(sm) || !pfa_fits64(sk)) bad = 1;
        M[i] = (int64_t)sm;
        Kn[i] = (int64_t)sk;
    }
    if (bad) a.bad[ip] = 1;
    if (a.filter) a.nexp[k] = bad ? -1 : pfa_num_expansions(M, Kn, h, a.q[k], rec[2]);
}

// N = (kappa M)[1:, 1:] for a point (element e at N[e * stride]); 0, or 1 if
// an entry leaves int64.  kappa is symmetric: N is too.
PFA_HD static inline int kappa_M(const int64_t *kappa, int h, const int64_t *M, int64_t *N,
                                 int64_t stride)
{
    const int n = h - 1;
    for (int i = 1; i < h; ++i)
        for (int j = i; j < h; ++j) {
            pfa_i128 s = 0;
            const int64_t *kij = kappa + ((size_t)i * h + j) * h;
            for (int k = 0; k < h; ++k)
                if (kij[k]) s += (pfa_i128)kij[k] * M[k];
            if (!pfa_fits64(s)) return 1;
            N[((i - 1) * n + (j - 1)) * stride] = (int64_t)s;
            N[((j - 1) * n + (i - 1)) * stride] = (int64_t)s;
        }
    return 0;
}

// ------------------------------------------------------------ k_post
// One lane per point slot, single-wavefront blocks (64 consecutive slots):
// Kn = ZB c and M = Binter c (only M[0] unless the rows are kept) in
// int128, checked into int64 (a failure marks the p-vector for the exact
// host path); K_gcd = gcd(Kn[1:]) and the number n_exp of (g, K0)
// expansions (0: the K-integrality filter drops the point).  Without kept
// rows, K_gcd is the search leaf's exact gcd of the HNF rows of ZB[1:] at c
// (the same integer) and only Kn[0] is formed, when a bound proves every
// entry of Kn within int64.  Kept rows are
// staged in LDS and stored as one contiguous block per wavefront; points
// with n_exp > 0 are appended to the list of the det test (one atomic per
// wavefront).
//
// ------------------------------------------------------------ k_filter
// det((kappa M)[1:, 1:]) != 0, exactly, for the listed points.  A
// wavefront holds R = 64 / hl points (hl: the largest h of the launch);
// lane i of a group computes M_i and owns row i of N (n = h - 1 rows; the
// row's entries of the pivot columns so far in registers).  Per point:
//   1. the columns of N from kappa as int32 (|kappa| < 2^31, else step 3),
//      staged in LDS for the geometry of the block's contiguous span of
//      points: with 24-bit integer multiply-adds when sum|M| < 2^23 and
//      max|kappa| sum|M| < 2^31 (every partial sum then fits int32), else
//      in double when max|kappa| sum|M| < 2^52 (every product and partial
//      sum is then an exact integer), else step 3;
//   2. fraction-free elimination (Bareiss) in double, left-looking: column
//      k is built at step k and updated by the steps before it, every
//      product and difference checked below 2^53 so each value is the
//      exact integer minor: a column without a pivot among the unused rows
//      proves det = 0 (a rank-r matrix builds r + 1 columns), n pivots
//      prove det != 0.  The exact division by the previous pivot is a
//      product with its rounded reciprocal, rounded to the nearest integer:
//      y = x fl(1/prev) is within |q| 2^-52 (1 + 2^-52) of the integer
//      quotient q, so rint(y) = q whenever |y| < 2^50 (checked with the
//      products);
//   3. points whose values reach 2^53 go on a second list (with their M)
//      for k_det_mod: N in int128 from the int64 kappa (an entry outside
//      int64 leaves the point undecided, 2), then det modulo 31-bit primes
//      with Montgomery products (rows are only scaled by units mod p, which
//      keeps det != 0 invariant) until a residue is nonzero (det != 0) or
//      the primes' product exceeds the columns' Hadamard bound (det = 0).
// Pivot rows are marked used instead of swapped (only the sign of det
// changes).  Totals (points, kept, expansions, nonsingular, undecided) go
// to tot[0..4].
struct FiltArgs {
    int64_t n;                    // point slots
    const int32_t *pidx;          // < 0: unused slot
    const int64_t *q;
    const uint64_t *g;            // per slot gcd(Kn[1:]) from the search's leaf (PFA_G_WIDE: not
                                  // known), or null
    const int32_t *c;             // D int32 per point
    const int64_t *Min;           // instead of c: M rows given (stride MH), all tested
    int D, MH, hl;                // MH = row stride of M, Kn (= hl)
    const int64_t *recs;
    RecL L;
    int filter;                   // count expansions and test det
    int64_t *M, *Kn, *nexp;       // stored when non-null (points kept)
    int8_t *nonsing;              // stored when non-null
    int8_t *bad;                  // per p-vector: M or Kn outside int64
    unsigned long long *tot;      // [0] points [1] kept [2] expansions [3] nonsingular [4] undecided
    GeoArgs G;                    // int64 kappa: step 3
    const int32_t *kT;            // per distinct kappa: kT[(k n + i) n + j] = kappa[i+1][j+1][k]
    const int64_t *kTofs;         // per geometry (identical kappas share one copy)
    const double *kTmax;          // per geometry max |kappa| (inf when >= 2^31)
    int lds_floats;               // LDS floats per block (0: no staging)
    // sparse form of kappa per distinct kappa (geometries where every
    // |kappa| < 2^23): for column j and row s of N, the nonzero terms
    // kappa[s+1][j+1][k] packed (k << 24 | value & 0xffffff) in
    // spKV[spKVofs + spR[spRofs + j n + s] .. + spR[spRofs + j n + s + 1])
    const int32_t *spR, *spKV;
    const int64_t *spRofs;        // per geometry, -1: no sparse form
    const int64_t *spKVofs;
    const int32_t *spInts;        // per geometry: LDS ints of its sparse form (n^2 + 1 + nnz)
    const PfaModP *primes;
    int np;
    int64_t modcap;               // list capacity; beyond it a point stays undecided
    int64_t *modkp, *modnexp, *modM;   // list: slot, n_exp, M row (stride MH)
    unsigned long long *modcount;
    int64_t *keepkp, *keepnexp;   // points with n_exp > 0 (k_post -> k_filter)
    int64_t *rawkp, *rawnexp, *rawmeta;   // k_post's per-block runs of them (capacity n)
    unsigned long long *pcount;   // per p-vector: lattice points (k_post; null: not counted)
    int64_t *keepmeta;            // their p-vector << 32 | geometry << 8 | h
    int n_geo;
    unsigned long long *keepcount;
};

#define PFA_TWO52 4503599627370496.0
#define PFA_TWO53 9007199254740992.0

struct DetLanes {
    int lane, nl, R, r, j, src0;
    bool inlane;
    uint64_t gm;                  // ballot bits of this lane's group
};

__device__ static inline DetLanes det_lanes(int nl)
{
    DetLanes d;
    d.lane = threadIdx.x & 63;
    d.nl = nl;
    d.R = 64 / nl;
    d.r = d.lane / nl;
    d.j = d.lane - d.r * nl;
    d.inlane = d.r < d.R;
    d.gm = d.inlane ? (((nl >= 64) ? ~0ull : ((1ull << nl) - 1)) << (d.r * nl)) : 0ull;
    d.src0 = d.inlane ? d.r * nl : d.lane;
    return d;
}

__device__ static inline uint64_t shfl_u64(uint64_t v, int src)
{
    const int lo = __shfl((int)(uint32_t)v, src), hi = __shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

__device__ static inline uint64_t wave_sum_u64(uint64_t v)
{
    for (int o = 32; o >= 1; o >>= 1) v += shfl_u64(v, (threadIdx.x & 63) ^ o);
    return v;
}

// NMAX >= D (the launch's largest dimension): the coordinates sit in
// registers as doubles, and a row product runs as unrolled FMAs when
// max|row| sum|c| < 2^53 and max|row| < 2^31 (every product and partial sum
// is then an exact integer; the low word of each int64 entry holds it),
// else as pfa_dot_c (int64 under its bound, int128 beyond).  The fast form
// reads all NMAX entries, unconditionally, so the loads issue together:
// cd[t] = 0 for t >= d adds nothing, and the records buffer extends past
// the last row.
template <int NMAX>
__device__ __forceinline__ int64_t post_dot(const int64_t *row, const double *cd, const int32_t *cc,
                                            int d, bool fast, int64_t rmax, int64_t cmax, int *ok)
{
    if (fast) {
        // every low word read into registers first (one load each, then
        // one wait), then the FMAs in order
        const int2 *r2 = (const int2 *)row;
        int32_t rv[NMAX];
#pragma unroll
        for (int t = 0; t < NMAX; ++t) rv[t] = r2[t].x;
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < NMAX; ++t) s = fma((double)rv[t], cd[t], s);
        return (int64_t)s;
    }
    return pfa_dot_c(row, cc, d, rmax, cmax, ok);
}

// The same product when the whole wavefront shares one p-vector: the row is
// read through the scalar cache (constant address space, wave-uniform
// address), the coordinates are the lane's own.
typedef __attribute__((address_space(4))) const int64_t pfa_cint64;
// (all NMAX entries are read, so the loads merge: cd[t] = 0 for t >= d
// adds nothing, and the records buffer extends past the last row)
template <int NMAX>
__device__ __forceinline__ int64_t post_dot_uni(const int64_t *row, const double *cd)
{
    const pfa_cint64 *r64 = (const pfa_cint64 *)row;
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < NMAX; ++t) s = fma((double)(int32_t)r64[t], cd[t], s);
    return (int64_t)s;
}

// Each block (one wavefront) takes a contiguous run of 64-slot groups, so
// the per-p-vector point counts accumulate in a register while the run stays
// on one p-vector (one atomic per change), and the kept points go to the
// block's own stretch of the raw list, then to the shared list with one
// atomic per block: thousands of wavefronts adding to a few counters
// serialize in the memory-side atomic units.
template <int NMAX>
__global__ __launch_bounds__(64) void k_post(FiltArgs a)
{
    extern __shared__ int64_t stage[];             // rows 64 x MH (when kept), then c 64 x D
    const int lane = threadIdx.x;
    const bool store = a.M != nullptr;
    int32_t *cst = (int32_t *)(stage + (store ? 64 * a.MH : 0));
    unsigned long long t_pts = 0, t_kept = 0;
    const int64_t groups = (a.n + 63) / 64;
    const int64_t g0 = groups * blockIdx.x / gridDim.x, g1 = groups * (blockIdx.x + 1) / gridDim.x;
    const int64_t r0 = g0 * 64;                    // this block's stretch of the raw list
    unsigned long long kc = 0;                     // entries in it
    int32_t cp = -1;                               // p-vector of the running point count
    unsigned long long cpn = 0;
    for (int64_t grp = g0; grp < g1; ++grp) {
        const int64_t base = grp * 64;
        const int64_t kp = base + lane;
        const int32_t pid = kp < a.n ? a.pidx[kp] : -1;
        const bool valid = pid >= 0;
        const int64_t *rec = a.recs + (valid ? (int64_t)pid * a.L.words : 0);
        const int32_t *ri = (const int32_t *)rec;
        const int d = valid ? ri[1] : 0, h = d + 1;
        // one p-vector for the whole wavefront (the usual case: slots are
        // handed out in chunks per p-vector): its rows come through the
        // scalar cache
        const int32_t p0 = __builtin_amdgcn_readfirstlane(pid);
        const bool uni = p0 >= 0 && !__ballot(valid && pid != p0);
        const int64_t *urec = a.recs + (uni ? (int64_t)p0 * a.L.words : 0);
        const int du = __builtin_amdgcn_readfirstlane(d);
        // the wavefront's 64 c vectors are one contiguous block: stage it
        const int64_t nc = (a.n - base < 64 ? a.n - base : 64) * a.D;
        __syncthreads();
        for (int64_t e = lane; e < nc; e += 64) cst[e] = a.c[base * a.D + e];
        __syncthreads();
        const int32_t *cc = cst + lane * a.D;
        int ok = 1;
        int64_t Kn0 = 0, M0 = 0, cmax = 0;
        uint64_t kg = 0;
        double cd[NMAX], csum = 0.0;
#pragma unroll
        for (int t = 0; t < NMAX; ++t) {
            const int32_t x = t < d ? cc[t] : 0;
            cd[t] = (double)x;
            csum += fabs(cd[t]);
            const int64_t ax = x < 0 ? -(int64_t)x : x;
            cmax = ax > cmax ? ax : cmax;
        }
        const int64_t zmax = valid ? rec[15] : 0, bmax = valid ? rec[14] : 0;
        const bool fz = zmax < ((int64_t)1 << 31) && (double)zmax * csum < PFA_TWO53;
        const bool fb = bmax < ((int64_t)1 << 31) && (double)bmax * csum < PFA_TWO53;
        // Kn (staged when kept).  Unless the rows are kept, gcd(Kn[1:]) is
        // the search leaf's g when every Kn entry provably fits int64
        // (max|ZB| sum|c| < 2^62): then only Kn[0] is computed here.
        const bool gk = !store && a.g && valid && (double)zmax * csum < 4.0e18 && a.g[kp] != PFA_G_WIDE;
        if (gk) kg = a.g[kp];
        for (int i = 0; i < (gk ? 1 : a.MH); ++i) {
            int64_t v = 0;
            if (valid && i < h) {
                v = uni && fz ? post_dot_uni<NMAX>(urec + a.L.ZB + (int64_t)i * du, cd)
                              : post_dot<NMAX>(rec + a.L.ZB + (int64_t)i * d, cd, cc, d, fz, zmax, cmax, &ok);
                if (i == 0) Kn0 = v;
                else if (kg != 1) kg = pfa_gcd64(kg, (uint64_t)(v < 0 ? -v : v));
            }
            if (store) stage[lane * a.MH + i] = v;
        }
        const int64_t rows = (a.n - base < 64 ? a.n - base : 64) * a.MH;
        if (store) {
            __syncthreads();
            for (int64_t e = lane; e < rows; e += 64) a.Kn[base * a.MH + e] = stage[e];
            __syncthreads();
        }
        // M (all entries when kept, else M[0])
        for (int i = 0; i < (store ? a.MH : 1); ++i) {
            int64_t v = 0;
            if (valid && i < h) {
                v = uni && fb ? post_dot_uni<NMAX>(urec + a.L.Bi + (int64_t)i * du, cd)
                              : post_dot<NMAX>(rec + a.L.Bi + (int64_t)i * d, cd, cc, d, fb, bmax, cmax, &ok);
                if (i == 0) M0 = v;
            }
            if (store) stage[lane * a.MH + i] = v;
        }
        if (store) {
            __syncthreads();
            for (int64_t e = lane; e < rows; e += 64) a.M[base * a.MH + e] = stage[e];
            __syncthreads();
        }
        if (valid && !ok) a.bad[a.pidx[kp]] = 1;
        int64_t ne = 0;
        if (a.filter && valid) ne = ok ? pfa_count_expansions_fast(M0, Kn0, kg, a.q[kp], rec[2]) : -1;
        if (a.nexp && valid) a.nexp[kp] = ne;
        if (a.nonsing && valid) a.nonsing[kp] = 0;
        if (a.pcount) {
            // lattice points per p-vector: a running count, flushed with
            // one atomic when the p-vector changes; a group's slots are
            // taken one p-vector at a time (the first remaining lane's)
            unsigned long long rest = __ballot(valid);
            while (rest) {
                const int32_t pq = __shfl(pid, (int)__builtin_ctzll(rest));
                const unsigned long long same = __ballot(valid && pid == pq);
                if (pq != cp) {
                    if (lane == 0 && cpn) atomicAdd(&a.pcount[cp], cpn);
                    cp = pq;
                    cpn = 0;
                }
                cpn += __popcll(same);
                rest &= ~same;
            }
        }
        // n_exp > 0: to the block's stretch of the raw list (entry k <=
        // slot - r0, so it stays inside the stretch)
        const unsigned long long km = __ballot(ne > 0);
        const int cnt = __popcll(km);
        if (ne > 0) {
            const int64_t at = r0 + (int64_t)kc + __popcll(km & ((1ull << lane) - 1));
            a.rawkp[at] = kp;
            a.rawnexp[at] = ne;
            a.rawmeta[at] = ((int64_t)a.pidx[kp] << 32) | ((int64_t)ri[26] << 8) | (int64_t)h;
        }
        kc += cnt;
        const unsigned long long vm = __ballot(valid);
        if (lane == 0) {
            t_pts += __popcll(vm);
            t_kept += cnt;
        }
    }
    if (a.pcount && lane == 0 && cpn) atomicAdd(&a.pcount[cp], cpn);
    if (kc) {
        // the stretch to the shared list: one atomic for the block
        unsigned long long b0 = 0;
        if (lane == 0) b0 = atomicAdd(a.keepcount, kc);
        b0 = shfl_u64(b0, 0);
        __syncthreads();                           // the stretch's stores, visible to every lane
        for (unsigned long long k = lane; k < kc; k += 64) {
            a.keepkp[b0 + k] = a.rawkp[r0 + k];
            a.keepnexp[b0 + k] = a.rawnexp[r0 + k];
            a.keepmeta[b0 + k] = a.rawmeta[r0 + k];
        }
    }
    if (a.tot && lane == 0) {
        atomicAdd(&a.tot[0], t_pts);
        atomicAdd(&a.tot[1], t_kept);
    }
}

// LDS writes of any lane of the wavefront before reads of any other
// (wavefront-scope release, wave barrier, acquire)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Entry N[s][c] = sum_{k <= n} kappa[s+1][c+1][k] M[k] of column c at the
// lane's row s (kT[(k n + s) n + c]), in four forms, each exact under the
// condition k_filter checks for the group:
//  - col_i24_lds: the zero-padded LDS layout lds[(c NMAX + s)(NMAX + 1) + k]
//    (rowp = the row's slice of column c), M as int32 in registers, 24-bit
//    multiply-adds (|kappa|, |M| < 2^23 and every partial sum below 2^31);
//  - col_sparse_lds: the sparse form staged in LDS (column c, row s: the
//    packed terms of R[c n + s] .. R[c n + s + 1]), the same int32 sums;
//  - col_i24: from kT in global memory, the same int32 sums;
//  - col_f64: from kT in global memory in double (every product and
//    partial sum an integer below 2^53).
// gM / gMd: the group's M (M[k] at gM[k]) in LDS.
template <int NMAX>
__device__ __forceinline__ int32_t col_i24_lds(const int32_t *rowp, const int32_t *mk)
{
    int32_t acc = 0;
#pragma unroll
    for (int k = 0; k <= NMAX; ++k) acc += __mul24(rowp[k], mk[k]);
    return acc;
}

__device__ __forceinline__ int32_t col_sparse_lds(const int32_t *sp, const int32_t *gM, int n, int s,
                                                  int c)
{
    const int32_t *R = sp + c * n + s, *KV = sp + n * n + 1;
    int32_t acc = 0;
    for (int t = R[0]; t < R[1]; ++t) {
        const int32_t x = KV[t];
        acc += (x << 8 >> 8) * gM[(x >> 24) & 63];
    }
    return acc;
}

__device__ __forceinline__ int32_t col_i24(const int32_t *kt, const int32_t *gM, int n, int s, int c)
{
    int32_t acc = 0;
    for (int k = 0; k <= n; ++k) acc += __mul24(kt[((int64_t)k * n + s) * n + c], gM[k]);
    return acc;
}

__device__ __forceinline__ double col_f64(const int32_t *kt, const double *gMd, int n, int s, int c)
{
    double w = 0.0;
    for (int k = 0; k <= n; ++k) w = fma((double)kt[((int64_t)k * n + s) * n + c], gMd[k], w);
    return w;
}

// f(T), f(T + 1), ..., f(K - 1) with constant arguments
template <int T, int K, class F>
__device__ __forceinline__ void static_for(F &f)
{
    if constexpr (T < K) {
        f(T);
        static_for<T + 1, K>(f);
    }
}

// f(integral_constant<K>), f(<K + 1>), ... while f returns true, below N;
// true if every call returned true
template <int K, int N, class F>
__device__ __forceinline__ bool static_steps(F &f)
{
    if constexpr (K < N) {
        if (!f(std::integral_constant<int, K>{})) return false;
        return static_steps<K + 1, N>(f);
    } else {
        return true;
    }
}

// SP: the launch has geometries with a sparse form of kappa (a separate
// instantiation, so the dense-only kernel keeps its registers)

template <int NMAX, bool SP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_filter(FiltArgs a)
{
    // dynamic LDS: kappa (lds_floats ints, rounded up to 4), then per
    // wavefront the groups' step tables: (pivot, 1 / previous pivot) as
    // double2 and the pivot row, R NMAX entries each
    extern __shared__ int32_t lds_kT[];
    __shared__ int64_t lds_ofs;
    __shared__ int lds_cnt;
    __shared__ int lds_sparse;
    __shared__ int32_t s_Mi[256];                  // the groups' M: int32 (small) ...
    __shared__ double s_Md[256];                   // ... and double (exact)
    __shared__ int64_t s_kofs[128];                // kappa offset and bound of geometries < 128
    __shared__ double s_kmax[128];
    // per wavefront: Binter of one p-vector as int32 (NMAX <= 24), the
    // p-vector staged (bp, wave-uniform) and its max|Binter|
    constexpr int BSZ = NMAX <= 24 ? (NMAX + 1) * NMAX : 1;
    __shared__ __align__(16) int32_t s_B[4 * BSZ];   // rows at stride NMAX (a multiple of 4)
    // per wavefront: each group's c at a stride of hl rounded up to 4, read
    // four entries at a time (at most 64 / hl groups: 128 + NMAX entries);
    // later each group's M, zero-padded at a stride of NMAX + 4 (a launch of
    // class NMAX has hl > NMAX - 3 unless NMAX = 8: at most 3 x 24, 4 x 20,
    // 6 x 16, 2 x 28 entries, and 32 x 12 for NMAX = 8)
    constexpr int CSZ = NMAX == 8 ? 384 : 128 + NMAX;
    __shared__ __align__(16) int32_t s_cg[4 * CSZ];
    const DetLanes L = det_lanes(a.hl);
    const int R = L.R, i = L.j, nlmax = a.hl - 1;
    const int wv = threadIdx.x >> 6;
    const int SW = R * NMAX;
    double2 *s_e = (double2 *)(lds_kT + ((a.lds_floats + 3) & ~3)) + wv * SW;
    int32_t *s_p = (int32_t *)((double2 *)(lds_kT + ((a.lds_floats + 3) & ~3)) + 4 * SW) + wv * SW;
    {
        const int rr = L.inlane ? L.r : 0;
        s_e += rr * NMAX;
        s_p += rr * NMAX;
    }
    // the points: the list k_post made, or every slot when M is given
    const int64_t cnt = a.keepkp ? (int64_t)*a.keepcount : a.n;
    auto slot = [&](int64_t t) -> int64_t { return a.keepkp ? a.keepkp[t] : t; };
    // each block takes a contiguous span of rounds (4 wavefronts x R points)
    const int64_t rounds = (cnt + 4 * R - 1) / (4 * R);
    const int64_t span = (rounds + gridDim.x - 1) / gridDim.x;
    const int64_t rd0 = (int64_t)blockIdx.x * span;
    const int64_t rd1 = rd0 + span < rounds ? rd0 + span : rounds;
    if (rd0 >= rd1) return;
    // stage kappa of the geometry of the span's first point
    int64_t cofs = -1;
    int sparse = 0;
    if (a.lds_floats > 0) {
        // the kappa of the span's first point: its sparse form when it has
        // one that fits, else the padded dense layout when that fits
        if (threadIdx.x == 0) {
            int64_t o = -1;
            int c = 0, spg = 0;
            const int64_t t0 = rd0 * 4 * R;
            if (t0 < cnt) {
                const int64_t kp = slot(t0);
                const int32_t *ri = (const int32_t *)(a.recs + (int64_t)a.pidx[kp] * a.L.words);
                const int h = ri[24], g = ri[26];
                if (SP && h > 1 && h - 1 <= NMAX && a.spRofs[g] >= 0 && a.spInts[g] <= a.lds_floats) {
                    o = a.kTofs[g];
                    c = h - 1;
                    spg = g + 1;
                } else if (h > 1 && h - 1 <= NMAX && (NMAX + 1) * NMAX * NMAX <= a.lds_floats &&
                           a.kTmax[g] < PFA_TWO52) {
                    o = a.kTofs[g];
                    c = h - 1;
                }
            }
            lds_ofs = o;                       // identifies the staged kappa
            lds_cnt = c;
            lds_sparse = spg;                  // geometry + 1 of the sparse form, or 0
        }
        __syncthreads();
        cofs = lds_ofs;
        sparse = SP && lds_sparse > 0;
        if (SP && sparse) {
            // ranges, then the packed terms
            const int g = lds_sparse - 1, nr = lds_cnt * lds_cnt + 1;
            const int32_t *Rg = a.spR + a.spRofs[g];
            const int32_t *KVg = a.spKV + a.spKVofs[g];
            for (int e = threadIdx.x; e < a.spInts[g]; e += blockDim.x)
                lds_kT[e] = e < nr ? Rg[e] : KVg[e - nr];
            __syncthreads();
        }
        // zero-padded copy by column and row: lds[(c NMAX + s)(NMAX + 1) + k]
        // = kT[(k n + s) n + c]
        const int nn = lds_cnt, cnt = (nn + 1) * nn * nn;
        if (cofs >= 0 && !sparse) {
            for (int e = threadIdx.x; e < (NMAX + 1) * NMAX * NMAX; e += blockDim.x) lds_kT[e] = 0;
            __syncthreads();
            for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
                const int k = e / (nn * nn), r = e - k * nn * nn, s2 = r / nn, c2 = r - s2 * nn;
                lds_kT[(c2 * NMAX + s2) * (NMAX + 1) + k] = a.kT[cofs + e];
            }
        }
        __syncthreads();
    }
    const int ng = a.n_geo < 128 ? a.n_geo : 128;
    for (int g = threadIdx.x; g < ng; g += blockDim.x) {
        s_kofs[g] = a.kTofs[g];
        s_kmax[g] = a.kTmax[g];
    }
    __syncthreads();
    auto kofs_of = [&](int g) -> int64_t { return g < 128 ? s_kofs[g] : a.kTofs[g]; };
    auto kmax_of = [&](int g) -> double { return g < 128 ? s_kmax[g] : a.kTmax[g]; };
    int32_t *sB = s_B + wv * BSZ;
    int32_t bp = -1;                               // p-vector whose Binter sB holds
    int64_t bB = 0;                                // its max|Binter|
    unsigned long long t_exp = 0, t_ns = 0, t_und = 0;
    for (int64_t rd = rd0; rd < rd1; ++rd) {
        const int64_t tl = (rd * 4 + wv) * R + L.r;
        const bool need = L.inlane && tl < cnt;
        int64_t kp = 0, ne = 0, pi = 0;
        int h = 0, geo = 0;
        if (need) {
            if (a.keepkp) {
                // one level of loads: slot, n_exp and the point's metadata
                kp = a.keepkp[tl];
                ne = a.keepnexp[tl];
                const int64_t m = a.keepmeta[tl];
                pi = m >> 32;
                geo = (int)((m >> 8) & 0xffffff);
                h = (int)(m & 0xff);
            } else {
                kp = tl;
                ne = 1;
                pi = a.pidx[kp];
                const int32_t *r0 = (const int32_t *)(a.recs + pi * a.L.words);
                h = r0[24];
                geo = r0[26];
            }
        }
        const int64_t *rec = a.recs + pi * a.L.words;
        const int n = h > 0 ? h - 1 : 0;
        const bool rowl = need && i < n;          // the lane's row s = i of N
        int res = (need && n == 0) ? 1 : 0;

        // M_i (checked into int64 by k_post already)
        int64_t Mi = 0;
        if (a.Min) {
            if (need && i < h) Mi = a.Min[kp * a.MH + i];
        } else {
            // lane t < n holds c[t]; every lane gets all of c by shuffles.
            // M_i = Binter[i] . c in double, valid when max|Binter| sum|c| <
            // 2^53 (every product and partial sum is then an exact integer)
            // and |Binter| < 2^31 (the low word of each entry holds it); else
            // again in int128 (the whole wavefront runs those shuffles)
            const int32_t cv = (need && i < n) ? a.c[kp * a.D + i] : 0;
            const int64_t *bi = rec + a.L.Bi + (int64_t)i * n;
            const int32_t *b32 = (const int32_t *)bi;
            const bool row = need && i < h;
            // the wavefront's points usually share one p-vector: its Binter
            // (|entries| < 2^31) is copied once into the wavefront's LDS
            // slice and the rows are read from there
            bool lb = false;
            if constexpr (NMAX <= 24) {
                const int32_t p0 = __builtin_amdgcn_readfirstlane(need ? (int32_t)pi : -1);
                if (p0 >= 0 && p0 != bp) {                 // wave-uniform
                    const int64_t *r0 = a.recs + (int64_t)p0 * a.L.words;
                    const int h0 = __builtin_amdgcn_readfirstlane(h);   // lane 0 holds p0's point
                    const int64_t b0 = r0[14];
                    wave_lds_sync();                       // the old copy's reads
                    if (b0 < ((int64_t)1 << 31) && h0 >= 2 && h0 - 1 <= NMAX) {
                        for (int e = L.lane; e < h0 * (h0 - 1); e += 64) {
                            const int rw = e / (h0 - 1);
                            sB[rw * NMAX + (e - rw * (h0 - 1))] = (int32_t)r0[a.L.Bi + e];
                        }
                        bp = p0;
                        bB = b0;
                    } else {
                        bp = -1;
                    }
                    wave_lds_sync();
                }
                lb = need && bp >= 0 && pi == (int64_t)bp;
            }
            const int64_t bmax = lb ? bB : rec[14];
            const bool r32 = row && bmax < ((int64_t)1 << 31);
            // the group's c through LDS, four entries a read (entries at or
            // beyond n are multiplied by 0)
            const int hlp = (a.hl + 3) & ~3;
            int32_t *cg = s_cg + wv * CSZ + (L.inlane ? L.r : 0) * hlp;
            wave_lds_sync();
            if (L.inlane && i < hlp) cg[i] = cv;
            wave_lds_sync();
            double acc = 0.0, csum = 0.0;
            if (!__ballot(row && !lb)) {
                // every row from LDS, all NMAX entries read unconditionally
                // where the slice exists (NMAX <= 24: inside it; cd = 0 adds
                // nothing, and a lane without a row does not use its sum)
                const int32_t *br = sB + (i < h ? i : 0) * NMAX;
#pragma unroll 1
                for (int t0 = 0; t0 < NMAX; t0 += 4) {
                    const int4 c4 = *(const int4 *)(cg + t0);
                    int4 b4 = make_int4(0, 0, 0, 0);
                    if constexpr (NMAX <= 24) b4 = *(const int4 *)(br + t0);
                    const int32_t cc[4] = {c4.x, c4.y, c4.z, c4.w}, bb[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int t = t0 + u;
                        const double cd = t < n ? (double)cc[u] : 0.0;
                        csum += fabs(cd);
                        if constexpr (NMAX <= 24) acc = fma((double)bb[u], cd, acc);
                        else if (row && t < n) acc = fma((double)br[t], cd, acc);   // (no LDS slice: lb never holds)
                    }
                }
            } else {
                // every entry read, unconditionally, so the loads issue
                // together (cd = 0 adds nothing; a lane without a row reads
                // its record's first words, and its sum is not used)
                const int32_t *bq = row ? b32 : (const int32_t *)rec;
#pragma unroll 1
                for (int t0 = 0; t0 < NMAX; t0 += 4) {
                    const int4 c4 = *(const int4 *)(cg + t0);
                    const int32_t cc[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int t = t0 + u;
                        const double cd = t < n ? (double)cc[u] : 0.0;
                        csum += fabs(cd);
                        acc = fma((double)bq[2 * t], cd, acc);
                    }
                }
            }
            const bool fast = r32 && (double)bmax * csum < PFA_TWO53;
            Mi = fast ? (int64_t)acc : 0;
            if (__ballot(row && !fast)) {
                pfa_i128 sm = 0;
                for (int t = 0; t < nlmax; ++t) {
                    const int32_t ct = __shfl(cv, L.src0 + t);
                    if (row && !fast && t < n) sm += (pfa_i128)bi[t] * ct;
                }
                if (row && !fast) Mi = (int64_t)sm;
            }
        }

        // 1. the group bound sum|M| (a shuffle tree) decides the arithmetic
        // of the entries of N: 24-bit integer or double (the whole wavefront
        // takes one form); terms capped at 2^57 so the sum cannot wrap (a
        // capped sum fails every test below)
        uint64_t am = (need && i < h) ? (uint64_t)(Mi < 0 ? -Mi : Mi) : 0;
        if (am > ((uint64_t)1 << 57)) am = (uint64_t)1 << 57;
        for (int off = 1; off < a.hl; off <<= 1) {
            const uint64_t o = shfl_u64(am, (L.inlane && i + off < a.hl) ? L.lane + off : L.lane);
            am += o;
        }
        const double msum = (double)shfl_u64(am, L.src0);
        const double kmax = need ? kmax_of(geo) : 0.0;
        const bool exact = !rowl || kmax * msum < PFA_TWO52;
        const bool small = !rowl || (msum < 8388608.0 && kmax < 8388608.0 && kmax * msum < 2147483648.0);
        const bool exact_g = !(__ballot(rowl && !exact) & L.gm);
        const bool all_small = !__ballot(!small);
        // the group's M through LDS (lane i of the group holds M_i)
        wave_lds_sync();
        s_Mi[threadIdx.x] = (int32_t)Mi;
        s_Md[threadIdx.x] = (double)Mi;
        wave_lds_sync();
        const int32_t *gM = s_Mi + (wv << 6) + L.src0;
        const double *gMd = s_Md + (wv << 6) + L.src0;
        // the form of the column builds: 0 dense LDS, 1 sparse LDS, 2 global
        // int32, 3 global double
        const bool lds = rowl && kofs_of(geo) == cofs;
        const bool all_lds = !__ballot(rowl && !lds);
        int form = all_small ? 2 : 3;
        if (SP && sparse && all_lds && all_small) form = 1;
        if (NMAX <= 24 && !sparse && cofs >= 0 && all_lds && all_small) form = 0;
        constexpr int MKN = NMAX <= 24 ? NMAX + 1 : 1;
        int32_t mk[MKN];
#pragma unroll
        for (int k = 0; k < MKN; ++k) mk[k] = 0;
        if constexpr (NMAX <= 24) {
            // the group's M zero-padded to NMAX + 4 entries (lanes at or
            // beyond h hold M = 0), read four entries at a time
            if (form == 0) {
                constexpr int MS = NMAX + 4;
                int32_t *gz = s_cg + wv * CSZ + (L.inlane ? L.r : 0) * MS;
                if (L.inlane) {
                    gz[i] = (int32_t)Mi;
                    for (int e = i; e < MS - a.hl; e += a.hl) gz[a.hl + e] = 0;
                }
                wave_lds_sync();
#pragma unroll
                for (int k0 = 0; k0 <= NMAX; k0 += 4) {
                    const int4 x4 = *(const int4 *)(gz + k0);
                    const int32_t xs[4] = {x4.x, x4.y, x4.z, x4.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (k0 + u <= NMAX) mk[k0 + u] = xs[u];
                }
            }
        }
        const int32_t *rowp = lds_kT + (i < NMAX ? i : NMAX - 1) * (NMAX + 1);
        const int32_t *kt = a.kT + (rowl ? kofs_of(geo) : 0);
        auto build = [&](int c) -> double {
            if constexpr (NMAX <= 24) {
                if (form == 0) return (double)col_i24_lds<NMAX>(rowp + c * NMAX * (NMAX + 1), mk);
            }
            if (!rowl || c >= n) return 0.0;
            if (form == 1) return (double)col_sparse_lds(lds_kT, gM, n, i, c);
            if (form == 2) return (double)col_i24(kt, gM, n, i, c);
            return col_f64(kt, gMd, n, i, c);
        };

        // 2. Bareiss elimination in double, left-looking: column k is built
        // at step k and takes the updates of steps t < k,
        //   v <- (v pk_t - pc_t v[p_t]) / pk_{t-1}     (pk_{-1} = 1)
        // on the rows that are no pivot yet (pc_t: the row's entry of the
        // pivot column of step t), which gives every value the
        // right-looking order gives, in particular the same integer minors.
        // Every product and difference is checked below 2^53; the exact
        // division is a product with the rounded reciprocal, rounded to the
        // nearest integer: y = x fl(1/pk_{t-1}) lies within |q| 2^-52
        // (1 + 2^-52) of the integer quotient q, so rint(y) = q whenever
        // |y| < 2^50 (checked with the products).  The pivot of step k is
        // the lowest row that is no pivot yet with a nonzero entry; a column
        // without one proves det = 0 (only the columns up to the rank are
        // built), n pivots prove det != 0.
        bool run = need && n > 0 && exact_g, trip = false;
        int ustep = 255;                          // the step this row was the pivot of
        double prev = 1.0;
        double pc[NMAX];
        // one step: column k built, the steps t < k applied, the pivot
        // chosen; false once no group of the wavefront runs.  The first steps
        // (all of them up to NMAX 20, the first 6 beyond) are unrolled with k
        // a constant: pc[k] a plain register, the updates a fixed sequence
        // (rank-r points stop after r + 1 steps).  The rest of the loop stays
        // rolled, its loops over t running to NMAX with wave-uniform
        // predicates so every index of pc is a constant.
        auto step = [&](auto KC) -> bool {
            constexpr bool CK = !std::is_same<decltype(KC), int>::value;
            int k;
            if constexpr (CK) k = decltype(KC)::value;
            else k = KC;
            if (k >= nlmax) return false;
            if (run && k >= n) { run = false; res = 1; }
            if (!__ballot(run)) return false;
            double v = build(k);
            int okb = 1;
            auto upd = [&](int t) {
                const int pt = s_p[t];
                const double2 e = s_e[t];
                const double cp = __shfl(v, L.inlane ? L.src0 + pt : L.lane);
                if (run && i < n && ustep > t) {
                    const double x1 = v * e.x, x2 = pc[t] * cp, x = x1 - x2, y = x * e.y;
                    // |x1|, |x2|, |x| < 2^53 and |y| < 2^50 (8|y| is exact)
                    const double m = fmax(fmax(fabs(x1), fabs(x2)), fmax(fabs(x), 8.0 * fabs(y)));
                    okb &= (int)(m < PFA_TWO53);
                    v = rint(y);
                }
            };
            if constexpr (CK) {
                static_for<0, decltype(KC)::value>(upd);
            } else {
#pragma unroll
                for (int t = 0; t < NMAX; ++t)
                    if (t < k) upd(t);            // wave-uniform
            }
            if (run && (__ballot(!okb) & L.gm)) { run = false; trip = true; }
            const uint64_t cm = __ballot(run && i < n && ustep > k && v != 0.0) & L.gm;
            const int pl = cm ? (int)__builtin_ctzll(cm) : L.lane;
            const double pk = __shfl(v, pl);
            if (run && !cm) { run = false; res = 0; }
            if constexpr (CK) {
                pc[decltype(KC)::value] = v;
            } else {
#pragma unroll
                for (int t = 0; t < NMAX; ++t)
                    if (t == k) pc[t] = v;
            }
            if (run) {
                if (L.lane == pl) ustep = k;
                if (i == 0) {
                    s_p[k] = pl - L.src0;
                    s_e[k] = make_double2(pk, 1.0 / prev);
                }
                prev = pk;
            }
            wave_lds_sync();
            return true;
        };
        constexpr int NU = NMAX <= 20 ? NMAX : 6;
        if (static_steps<0, NU>(step)) {
            if constexpr (NU < NMAX) {
                for (int k = NU; k < nlmax; ++k)
                    if (!step(k)) break;
            }
        }
        if (run) res = 1;

        // 3. the modular list
        const bool modn = need && n > 0 && (!exact_g || trip);
        int64_t t = -1;
        if (modn && i == 0) {
            t = (int64_t)atomicAdd(a.modcount, 1ull);
            if (t >= a.modcap) t = -1;
        }
        t = (int64_t)shfl_u64((uint64_t)t, L.src0);
        if (modn && t >= 0) {
            if (i < a.MH) a.modM[t * a.MH + i] = Mi;
            if (i == 0) {
                a.modkp[t] = kp;
                a.modnexp[t] = ne;
            }
        }
        if (modn && t < 0) res = 2;             // list full: left to the host
        if (need && i == 0) {
            if (a.nonsing) a.nonsing[kp] = (int8_t)(modn && t >= 0 ? 2 : res);
            if (!(modn && t >= 0)) {
                if (res == 1) { t_ns++; t_exp += (unsigned long long)ne; }
                if (res == 2) t_und++;
            }
        }
    }
    if (a.tot) {
        const unsigned long long v2 = wave_sum_u64(t_exp), v3 = wave_sum_u64(t_ns),
                                 v4 = wave_sum_u64(t_und);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&a.tot[2], v2);
            atomicAdd(&a.tot[3], v3);
            atomicAdd(&a.tot[4], v4);
        }
    }
}

// Step 3 for the points k_filter listed (M rows from the list).
template <int NMAX>
__global__ __launch_bounds__(256) void k_det_mod(FiltArgs a)
{
    const DetLanes L = det_lanes(a.hl);
    const int R = L.R, j = L.j, nlmax = a.hl - 1;
    const unsigned long long cntu = *a.modcount;
    const int64_t cnt = (int64_t)cntu < a.modcap ? (int64_t)cntu : a.modcap;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long t_exp = 0, t_ns = 0, t_und = 0;
    for (int64_t base = wave * R; base < cnt; base += nwaves * R) {
        const int64_t t = base + L.r;
        const bool need = L.inlane && t < cnt;
        const int64_t kp = need ? a.modkp[t] : 0;
        const int64_t *rec = a.recs + (need ? (int64_t)a.pidx[kp] * a.L.words : 0);
        const int32_t *ri = (const int32_t *)rec;
        const int h = need ? ri[24] : 0, geo = need ? ri[26] : 0;
        const int n = h > 0 ? h - 1 : 0;
        const bool col = need && j < n;
        const int64_t *Mrow = a.modM + (need ? t : 0) * a.MH;
        int res = 2;
        bool modn = need && n > 0;
        int64_t Ni[NMAX];
#pragma unroll
        for (int s = 0; s < NMAX; ++s) Ni[s] = 0;
        bool fits = true;
        double cn = 0.0;
        if (col) {
            const int64_t *kap = a.G.kappa + a.G.kofs[geo];
#pragma unroll
            for (int s = 0; s < NMAX; ++s) {
                if (s >= n) continue;
                const int64_t *kij = kap + ((int64_t)(s + 1) * h + (j + 1)) * h;
                pfa_i128 acc = 0;
                for (int k = 0; k < h; ++k) {
                    pfa_i128 u;
                    if (pfa_add_ovf(acc, (pfa_i128)kij[k] * Mrow[k], &u)) fits = false;
                    acc = u;
                }
                if (!pfa_fits64(acc)) fits = false;
                Ni[s] = (int64_t)acc;
                cn += (double)Ni[s] * (double)Ni[s];
            }
        }
        if (modn && (__ballot(modn && !fits) & L.gm)) { modn = false; res = 2; }
        if (modn && (__ballot(modn && col && cn == 0.0) & L.gm)) { modn = false; res = 0; }
        const double tb = (modn && col) ? 0.5 * log2(cn) + 1e-9 : 0.0;
        double need_bits = 1.0;
        for (int q = 0; q < a.hl; ++q) need_bits += __shfl(tb, L.inlane ? L.src0 + q : L.lane);
        double have = 0.0;
        for (int pi = 0; pi < a.np; ++pi) {
            if (!__ballot(modn)) break;
            const PfaModP P = a.primes[pi];
            uint32_t wm[NMAX];
#pragma unroll
            for (int s = 0; s < NMAX; ++s)
                wm[s] = (modn && col && s < n) ? pfa_mod_fast(Ni[s], P) : 0u;
            bool runm = modn, zero = false;
            uint64_t usedm = 0;
            for (int k = 0; k < nlmax; ++k) {
                if (runm && k >= n) runm = false;
                if (!__ballot(runm)) break;
                int cand = -1;
#pragma unroll
                for (int s = NMAX - 1; s >= 0; --s)
                    if (s < n && !((usedm >> s) & 1) && wm[s] != 0u) cand = s;
                const int srck = L.inlane ? L.src0 + k : L.lane;
                const int p = __shfl(cand, srck);
                bool act = runm;
                if (act && p < 0) { runm = false; act = false; zero = true; }
                uint32_t pj = 0;
#pragma unroll
                for (int s = 0; s < NMAX; ++s)
                    if (s == p) pj = wm[s];
                const uint32_t pk = (uint32_t)__shfl((int)pj, srck);
#pragma unroll
                for (int s = 0; s < NMAX; ++s) {
                    const uint32_t ik = (uint32_t)__shfl((int)wm[s], srck);
                    if (act && col && s < n && s != p && !((usedm >> s) & 1)) {
                        const uint32_t x1 = pfa_mont(pk, wm[s], P.p, P.pinv);
                        const uint32_t x2 = pfa_mont(ik, pj, P.p, P.pinv);
                        wm[s] = x1 >= x2 ? x1 - x2 : x1 + (P.p - x2);
                    }
                }
                if (act) usedm |= 1ull << p;
            }
            if (modn && !zero) { res = 1; modn = false; }
            if (modn) {
                have += log2((double)P.p) - 1e-9;
                if (have > need_bits) { res = 0; modn = false; }
            }
        }
        if (need && j == 0) {
            if (a.nonsing) a.nonsing[kp] = (int8_t)res;
            if (res == 1) { t_ns++; t_exp += (unsigned long long)a.modnexp[t]; }
            if (res == 2) t_und++;
        }
    }
    if (a.tot) {
        const unsigned long long v2 = wave_sum_u64(t_exp), v3 = wave_sum_u64(t_ns),
                                 v4 = wave_sum_u64(t_und);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&a.tot[2], v2);
            atomicAdd(&a.tot[3], v3);
            atomicAdd(&a.tot[4], v4);
        }
    }
}

// ---------------------------------------------------------- k_compact
// The points that can yield PFVs -- n_exp > 0 (the filter's list) and a
// verdict other than singular (1 nonsingular, 2 undecided) -- gathered into
// dense arrays for the copy back: p-vector, q, n_exp, verdict, and the rows
// M = Binter c and Kn = ZB c recomputed in int128 (a value beyond int64 has
// already marked its p-vector for the exact host path in k_post; such a
// row is stored truncated and dropped with its p-vector).  One atomic per
// wavefront reserves the output slots; slots beyond `cap` are counted, not
// written, and the host reruns with room for all.
struct CompactArgs {
    const int64_t *keepkp, *keepnexp;
    const unsigned long long *keepcount;
    const int8_t *nonsing;
    const int32_t *pidx;
    const int64_t *q;
    const int32_t *c;
    int D, HS;
    const int64_t *recs;
    RecL L;
    int32_t *o_pidx;
    int64_t *o_q, *o_M, *o_Kn, *o_nexp;
    int8_t *o_ns;
    unsigned long long *o_count;
    unsigned long long cap;
};

__global__ __launch_bounds__(64) void k_compact(CompactArgs a)
{
    const int lane = threadIdx.x;
    const int64_t nk = (int64_t)*a.keepcount;
    for (int64_t base = (int64_t)blockIdx.x * 64; base < nk; base += (int64_t)gridDim.x * 64) {
        const int64_t t = base + lane;
        int64_t kp = -1;
        int8_t ns = 0;
        if (t < nk) {
            kp = a.keepkp[t];
            ns = a.nonsing[kp];
        }
        const bool take = kp >= 0 && ns != 0;
        const unsigned long long m = __ballot(take);
        if (!m) continue;
        unsigned long long o0 = 0;
        if (lane == 0) o0 = atomicAdd(a.o_count, (unsigned long long)__popcll(m));
        o0 = shfl_u64(o0, 0);
        if (!take) continue;
        const unsigned long long o = o0 + __popcll(m & ((1ull << lane) - 1));
        if (o >= a.cap) continue;
        const int32_t ip = a.pidx[kp];
        const int64_t *rec = a.recs + (int64_t)ip * a.L.words;
        const int d = ((const int32_t *)rec)[1];
        const int32_t *cc = a.c + kp * a.D;
        const int64_t *Bi = rec + a.L.Bi, *ZB = rec + a.L.ZB;
        for (int i = 0; i < a.HS; ++i) {
            pfa_i128 sm = 0, sk = 0;
            if (i <= d)
                for (int j = 0; j < d; ++j) {
                    sm += (pfa_i128)Bi[(int64_t)i * d + j] * cc[j];
                    sk += (pfa_i128)ZB[(int64_t)i * d + j] * cc[j];
                }
            a.o_M[o * a.HS + i] = (int64_t)sm;
            a.o_Kn[o * a.HS + i] = (int64_t)sk;
        }
        a.o_pidx[o] = ip;
        a.o_q[o] = a.q[kp];
        a.o_nexp[o] = a.keepnexp[t];
        a.o_ns[o] = ns;
    }
}

// ============================================================ host side
#define HIPCHK(x)                                                                    \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            snprintf(errbuf, 256, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                     __LINE__);                                                        \
            throw 1;                                                                   \
        }                                                                              \
    } while (0)

// grow-only device buffers kept per device until pfg_release
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

// Pinned host staging of one call's results.  Two sets per device: a call
// copies its results into one while holding the device, then releases the
// device and assembles from the set, so the next call's kernels overlap it.
struct Stage {
    std::mutex mu;
    std::vector<DevBuf> bufs;
};

struct DevCtx {
    std::mutex mu;
    hipStream_t stream = nullptr;
    std::vector<DevBuf> bufs;
    Stage stage[2];
    int next_stage = 0;                    // (under mu)
    int cus = 0;
    size_t lds_max = 0;                    // per workgroup
    size_t lds_cu = 0;                     // per CU
    unsigned long long pcap_hint = 0;      // largest point count seen
    unsigned long long pts_reserved = 0;   // point slots pfa_reserve set aside (0: none)
    double ppp_hint = 0;                   // largest points per p-vector seen (n_p >= 64)
    double ms_alloc = 0;                   // this run's buffer growth (host ms)
    unsigned long long sv_hint = 0;        // largest survivor count seen (keep_points 2)
    hipEvent_t ev[8] = {};                 // stage timings (created once)
    int wclk_khz = 0;                      // wall-clock counter rate
    // instantiated graphs of the pipeline's segments, by launch key (the
    // oldest is dropped beyond PFA_GRAPHS_MAX)
    std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
    unsigned long long graph_hits = 0, graph_captures = 0;
    DevBuf up_host[2];                     // pinned staging of packed uploads (inputs, filter forms)
};

#define PFA_GRAPHS_MAX 48

static std::mutex g_ctx_mu;
static std::vector<DevCtx *> g_ctx;

static DevCtx *ctx_for(int device)
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if ((int)g_ctx.size() <= device) g_ctx.resize(device + 1, nullptr);
    if (!g_ctx[device]) g_ctx[device] = new DevCtx();
    return g_ctx[device];
}

// stream and device facts, once per device (call with C->mu held)
static void ctx_init(DevCtx *C, int device, char *errbuf)
{
    HIPCHK(hipSetDevice(device));
    if (C->stream) return;
    HIPCHK(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    C->cus = prop.multiProcessorCount;
    C->lds_max = prop.sharedMemPerBlock;
    C->lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                      : prop.sharedMemPerBlock;
    for (auto &e : C->ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipDeviceGetAttribute(&C->wclk_khz, hipDeviceAttributeWallClockRate, device));
}

enum {
    B_GEO_H, B_GEO_Q, B_GEO_QMAX, B_GEO_LIN, B_KAPPA, B_MBASIS, B_KOFS, B_MOFS, B_PS, B_PGEO,
    B_RECS, B_RING, B_SEQ, B_STACKS, B_CTL, B_PT_PIDX, B_PT_Q, B_PT_C, B_CNT, B_M, B_KN, B_NEXP, B_BAD,
    B_PRIMES, B_KT, B_KTOFS, B_KTMAX, B_MODCNT, B_MODLIST, B_MODNEXP, B_MODM, B_NONSING, B_PREPCOLD,
    B_TOT, B_KEEPCNT, B_KEEPKP, B_KEEPNEXP, B_KEEPMETA, B_PREPORDER,
    B_SPR, B_SPKV, B_SPROFS, B_SPKVOFS, B_SPINTS,
    B_SV_PIDX, B_SV_Q, B_SV_M, B_SV_KN, B_SV_NEXP, B_SV_NS, B_SV_CNT, B_STATSAVE, B_STATOUT,
    B_PCOUNT, B_F0, B_F1, B_SPLITST, B_BOX, B_BOXOUT, B_PT_G, B_KEEPRAW, B_UPIN, B_UPFILT,
    B_NBUF
};

static void *dev_get(DevCtx *C, int slot, size_t bytes, char *errbuf)
{
    // k_post reads whole NMAX-wide rows of the records (up to PFA_DMAX words
    // past a p-vector's last row): the records buffer keeps 1 KB beyond
    if (slot == B_RECS) bytes += 1024;
    if ((int)C->bufs.size() <= slot) C->bufs.resize(slot + 1);
    DevBuf &b = C->bufs[slot];
    if (b.n < bytes) {
        auto t0 = std::chrono::steady_clock::now();
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
        size_t want = bytes + bytes / 4 + 4096;
        HIPCHK(hipMalloc(&b.p, want));
        b.n = want;
        C->ms_alloc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return b.p;
}

// Grow-only pinned host buffers, per device, for the results' copy back
// (no zero fill, no page faults after the first use, full-rate DMA).
enum { H_PIDX, H_Q, H_M, H_KN, H_NEXP, H_NS, H_STAT, H_NBUF };

static void *host_get(DevCtx *C, Stage &S, int slot, size_t bytes, char *errbuf)
{
    if ((int)S.bufs.size() <= slot) S.bufs.resize(slot + 1);
    DevBuf &b = S.bufs[slot];
    if (b.n < bytes) {
        auto t0 = std::chrono::steady_clock::now();
        if (b.p) (void)hipHostFree(b.p);
        b.p = nullptr;
        b.n = 0;
        size_t want = bytes + bytes / 4 + 4096;
        HIPCHK(hipHostMalloc(&b.p, want, hipHostMallocDefault));
        b.n = want;
        C->ms_alloc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return b.p;
}

#define CTL_BYTES (16 * 8 * 16)

// A call's small inputs sent with one copy: packed (256-byte aligned) into
// a grow-only pinned host buffer of the device, then one host-to-device
// copy into one device buffer; each pageable copy would be a staged
// transfer of its own.  The staging buffer is rewritten by the next call
// only, after this call's stream has finished.
struct Packed {
    std::vector<const void *> src;
    std::vector<size_t> off, len;
    size_t total = 0;
    size_t add(const void *p, size_t bytes)
    {
        src.push_back(p);
        off.push_back(total);
        len.push_back(bytes);
        const size_t at = total;
        total += (std::max<size_t>(bytes, 16) + 255) & ~(size_t)255;
        return at;
    }
};

static void pinned_reserve(DevCtx *C, int which, size_t bytes, char *errbuf)
{
    DevBuf &h = C->up_host[which];
    if (h.n >= bytes) return;
    auto t0 = std::chrono::steady_clock::now();
    if (h.p) (void)hipHostFree(h.p);
    h.p = nullptr;
    h.n = 0;
    const size_t want = bytes + bytes / 4 + 4096;
    HIPCHK(hipHostMalloc(&h.p, want, hipHostMallocDefault));
    h.n = want;
    C->ms_alloc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static char *packed_upload(DevCtx *C, int which, int dslot, const Packed &P, hipStream_t s, char *errbuf)
{
    pinned_reserve(C, which, std::max<size_t>(P.total, 16), errbuf);
    char *h = (char *)C->up_host[which].p;
    for (size_t k = 0; k < P.src.size(); ++k)
        if (P.len[k]) memcpy(h + P.off[k], P.src[k], P.len[k]);
    char *d = (char *)dev_get(C, dslot, std::max<size_t>(P.total, 16), errbuf);
    if (P.total) HIPCHK(hipMemcpyAsync(d, h, P.total, hipMemcpyHostToDevice, s));
    return d;
}


// the k largest primes below 2^31 with their Barrett constants
static std::vector<PfaModP> primes31(int k)
{
    std::vector<PfaModP> out;
    for (uint32_t x = 2147483647u; (int)out.size() < k; x -= 2) {
        bool pr = true;
        for (uint32_t f = 3; (uint64_t)f * f <= x; f += 2)
            if (x % f == 0) { pr = false; break; }
        if (pr) out.push_back(PfaModP{x, pfa_mont_pinv(x), (uint64_t)(((pfa_u128)1 << 64) / x)});
    }
    return out;
}

static double ms_between(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms;
}

struct HostPts {
    std::vector<int32_t> pidx, c;
    std::vector<int64_t> q;
};

static int batch_dmax(const pfa_run_input *in)
{
    int D = 1;
    for (int g = 0; g < in->n_geo; ++g) D = std::max(D, (int)in->h[g] - 1);
    return D;
}

// Size classes of the preparation: p-vectors grouped by the h11 of their
// geometry into classes of largest h11 at most 12, 20, 28 and the maximum;
// within a class by descending h11 (long problems first).  One class
// without an index list when every p-vector has the same h11.
struct PrepClass {
    int h;                        // largest h11 in the class
    int64_t n;
    std::vector<int32_t> idx;
};

static std::vector<PrepClass> prep_classes(const pfa_run_input *in)
{
    int hlo = PFA_MAX_H11, hhi = 0;
    for (int g = 0; g < in->n_geo; ++g) {
        hlo = std::min<int>(hlo, in->h[g]);
        hhi = std::max<int>(hhi, in->h[g]);
    }
    if (hlo == hhi || in->n_p < 2) return {PrepClass{hhi, in->n_p, {}}};
    static const int edges[] = {12, 20, 28, PFA_MAX_H11};
    std::vector<std::vector<int32_t>> by_h(PFA_MAX_H11 + 1);
    for (int64_t i = 0; i < in->n_p; ++i) by_h[in->h[in->pgeo[i]]].push_back((int32_t)i);
    std::vector<PrepClass> out;
    int lo = 0;
    for (int e : edges) {
        PrepClass c{0, 0, {}};
        for (int h = e; h > lo; --h)
            if (!by_h[h].empty()) {
                if (!c.h) c.h = h;
                c.idx.insert(c.idx.end(), by_h[h].begin(), by_h[h].end());
            }
        c.n = (int64_t)c.idx.size();
        if (c.n) out.push_back(std::move(c));
        lo = e;
    }
    std::reverse(out.begin(), out.end());    // largest class first
    return out;
}

// Host threads for the final assembly: the caller's count, else the
// hardware's, at most 16.
static int finish_threads(const pfa_run_input *in)
{
    int n = in->host_threads > 0 ? in->host_threads : (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 16));
}

// fn(lo, hi) over [0, n) in chunks of `chunk`, on up to nth threads.
template <class F>
static void parallel_chunks(int64_t n, int64_t chunk, int nth, F fn)
{
    const int64_t nchunks = (n + chunk - 1) / chunk;
    nth = (int)std::max<int64_t>(1, std::min<int64_t>(nth, nchunks));
    if (nth == 1) {
        if (n > 0) fn((int64_t)0, n);
        return;
    }
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (int64_t c; (c = next.fetch_add(1)) < nchunks;)
            fn(c * chunk, std::min(n, (c + 1) * chunk));
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
}

// canonical order: by p-vector, then (M, Kn) lexicographically.  A stable
// counting sort by p-vector, then every p-vector's points sorted on their
// own (small, cache-resident groups) on nth threads.
static void canonical_order(std::vector<int64_t> &idx, const int64_t *pidx, const int64_t *M,
                            const int64_t *Kn, int MH, int64_t n_p, int nth)
{
    std::vector<int64_t> start(n_p + 1, 0);
    for (int64_t k : idx) start[pidx[k] + 1]++;
    for (int64_t i = 0; i < n_p; ++i) start[i + 1] += start[i];
    std::vector<int64_t> by_p(idx.size()), pos(start.begin(), start.end() - 1);
    for (int64_t k : idx) by_p[pos[pidx[k]]++] = k;
    auto less = [&](int64_t a, int64_t b) {
        const int64_t *ma = M + a * MH, *mb = M + b * MH;
        for (int i = 0; i < MH; ++i) if (ma[i] != mb[i]) return ma[i] < mb[i];
        const int64_t *ka = Kn + a * MH, *kb = Kn + b * MH;
        for (int i = 0; i < MH; ++i) if (ka[i] != kb[i]) return ka[i] < kb[i];
        return false;
    };
    // threads by the points to sort (starting them costs tens of us each)
    nth = (int)std::max<int64_t>(1, std::min<int64_t>(nth, (int64_t)idx.size() / 4096));
    parallel_chunks(n_p, 64, nth, [&](int64_t lo, int64_t hi) {
        for (int64_t ip = lo; ip < hi; ++ip)
            std::sort(by_p.begin() + start[ip], by_p.begin() + start[ip + 1], less);
    });
    idx.swap(by_p);
}

static void fill_geo_offsets(const pfa_run_input *in, std::vector<int64_t> &kofs,
                             std::vector<int64_t> &mofs, int64_t &ktot, int64_t &mtot)
{
    kofs.resize(in->n_geo);
    mofs.resize(in->n_geo);
    ktot = mtot = 0;
    for (int g = 0; g < in->n_geo; ++g) {
        int64_t h = in->h[g];
        kofs[g] = ktot;
        mofs[g] = mtot;
        ktot += h * h * h;
        mtot += h * h;
    }
}

// kappa[i+1][j+1][k] as int32 at kT[ofs + (k n + i) n + j] (n = h - 1; the
// lanes of a point group read consecutive j), one copy per distinct kappa
// (kTofs per geometry), and the largest |kappa| of each geometry
// (infinity when one reaches 2^31).
// The sparse form of one kappa (see FiltArgs): ranges spR (n^2 + 1, by
// column then row) and packed terms spKV; empty when an entry needs 24 bits
// or more, or when more than a quarter of the terms are nonzero (the dense
// build's fixed, predicate-free layout is faster there).
static void build_sparse(const int64_t *kap, int64_t h, std::vector<int32_t> &R,
                         std::vector<int32_t> &KV)
{
    const int64_t n = h - 1;
    R.assign(n * n + 1, 0);
    KV.clear();
    for (int64_t j = 0; j < n; ++j)
        for (int64_t sr = 0; sr < n; ++sr) {
            for (int64_t k = 0; k < h; ++k) {
                const int64_t v = kap[((sr + 1) * h + (j + 1)) * h + k];
                if (!v) continue;
                if (v >= (1 << 23) || v <= -(1 << 23)) {
                    R.clear();
                    KV.clear();
                    return;
                }
                KV.push_back((int32_t)(((uint32_t)k << 24) | ((uint32_t)v & 0xffffffu)));
            }
            R[j * n + sr + 1] = (int32_t)KV.size();
        }
    if (4 * (int64_t)KV.size() > n * n * h) {
        R.clear();
        KV.clear();
    }
}

static void build_kT(const pfa_run_input *in, const std::vector<int64_t> &kofs,
                     std::vector<int32_t> &kT, std::vector<int64_t> &kTofs,
                     std::vector<double> &kTmax, std::vector<int32_t> &spR,
                     std::vector<int32_t> &spKV, std::vector<int64_t> &spRofs,
                     std::vector<int64_t> &spKVofs, std::vector<int32_t> &spInts)
{
    spRofs.assign(in->n_geo, -1);
    spKVofs.assign(in->n_geo, 0);
    spInts.assign(in->n_geo, 0);
    spR.clear();
    spKV.clear();
    kTofs.assign(in->n_geo, 0);
    kTmax.assign(in->n_geo, 0.0);
    kT.clear();
    std::vector<std::pair<uint64_t, int>> seen;    // (hash, first geometry)
    for (int g = 0; g < in->n_geo; ++g) {
        const int64_t h = in->h[g], n = h - 1;
        const int64_t *kap = in->kappa + kofs[g];
        uint64_t hs = 1469598103934665603ull ^ (uint64_t)h;
        for (int64_t e = 0; e < h * h * h; ++e) hs = (hs ^ (uint64_t)kap[e]) * 1099511628211ull;
        int same = -1;
        for (auto &pr : seen)
            if (pr.first == hs && in->h[pr.second] == h &&
                !memcmp(in->kappa + kofs[pr.second], kap, 8 * h * h * h)) {
                same = pr.second;
                break;
            }
        if (same >= 0) {
            kTofs[g] = kTofs[same];
            kTmax[g] = kTmax[same];
            spRofs[g] = spRofs[same];
            spKVofs[g] = spKVofs[same];
            spInts[g] = spInts[same];
            continue;
        }
        {
            std::vector<int32_t> R, KV;
            build_sparse(kap, h, R, KV);
            if (!R.empty()) {
                spRofs[g] = (int64_t)spR.size();
                spKVofs[g] = (int64_t)spKV.size();
                spInts[g] = (int32_t)(R.size() + KV.size());
                spR.insert(spR.end(), R.begin(), R.end());
                spKV.insert(spKV.end(), KV.begin(), KV.end());
            }
        }
        seen.push_back({hs, g});
        kTofs[g] = (int64_t)kT.size();
        kT.resize(kT.size() + h * n * n);
        double mx = 0.0;
        for (int64_t k = 0; k < h; ++k)
            for (int64_t i = 0; i < n; ++i)
                for (int64_t jj = 0; jj < n; ++jj) {
                    const int64_t v = kap[((i + 1) * h + (jj + 1)) * h + k];
                    kT[kTofs[g] + (k * n + i) * n + jj] = (int32_t)v;      // exact when mx < 2^31
                    mx = std::max(mx, std::fabs((double)v));
                }
        kTmax[g] = mx < 2147483648.0 ? mx : INFINITY;
    }
    if (kT.empty()) kT.push_back(0);
    if (spR.empty()) spR.push_back(0);
    if (spKV.empty()) spKV.push_back(0);
}

// ---------------------------------------------------- captured segments
// A segment is a fixed run of launches on the library's stream between two
// host decisions (prep; the search's set-up and the search; the
// post-processing).  run_seg first describes the body's launches (kernel,
// grid, block, LDS bytes, argument bytes) into a key; with graphs on, a key
// seen before replays its instantiated graph and a new key is captured
// (hipStreamBeginCapture .. EndCapture), instantiated, kept and launched.
// Buffer clears in a body are k_fill launches.  With graphs off
// (PFVS_GRAPH=0) the body launches directly.
struct Seg {
    hipStream_t s;
    char *errbuf;
    int describe;                          // 1: append to key only
    std::string key;
    void put(const void *p, size_t n) { key.append((const char *)p, n); }
    template <class P, class A> void put_arg(const A &x)
    {
        P y = (P)x;
        put(&y, sizeof y);
    }
    template <class... P, class... A>
    void k(void (*fn)(P...), dim3 g, dim3 b, size_t lds, A... args)
    {
        static_assert(sizeof...(P) == sizeof...(A), "argument count");
        if (describe) {
            put(&fn, sizeof fn);
            put(&g, sizeof g);
            put(&b, sizeof b);
            put(&lds, sizeof lds);
            (put_arg<P>(args), ...);
            return;
        }
        hipLaunchKernelGGL(fn, g, b, lds, s, ((P)args)...);
        HIPCHK(hipGetLastError());
    }
    void zero(void *p, size_t n)
    {
        if (!n) return;
        const int grid = (int)std::min<size_t>((n / 16 + 255) / 256 + 1, 2048);
        k(k_fill, dim3(grid), dim3(256), 0, (uint8_t *)p, n, (uint8_t)0);
    }
};

static bool graphs_on()
{
    const char *e = getenv("PFVS_GRAPH");
    return !(e && e[0] == '0');
}

// graph = false: launched directly.  A segment that starts on an idle
// device (right after the host's uploads, or after it waited for the
// search) runs its first kernel sooner that way: a graph launch costs the
// host about 0.2 ms before the device sees any of it.
template <class F>
static void run_seg(DevCtx *C, hipStream_t s, F body, char *errbuf, bool graph = true)
{
    Seg sg{s, errbuf, 0, std::string()};
    if (!graph || !graphs_on()) {
        body(sg);
        return;
    }
    sg.describe = 1;
    body(sg);
    if (sg.key.empty()) return;
    hipGraphExec_t ex = nullptr;
    for (auto &g : C->graphs)
        if (g.first == sg.key) { ex = g.second; break; }
    if (ex) {
        C->graph_hits++;
    } else {
        Seg cap{s, errbuf, 0, std::string()};
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        body(cap);
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamEndCapture(s, &gr));
        HIPCHK(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
        if ((int)C->graphs.size() >= PFA_GRAPHS_MAX) {
            (void)hipGraphExecDestroy(C->graphs.front().second);
            C->graphs.erase(C->graphs.begin());
        }
        C->graphs.push_back({sg.key, ex});
        C->graph_captures++;
    }
    HIPCHK(hipGraphLaunch(ex, s));
}

static void *upload(DevCtx *C, int slot, const void *src, size_t bytes, hipStream_t s, char *errbuf)
{
    void *d = dev_get(C, slot, std::max<size_t>(bytes, 16), errbuf);
    if (bytes) HIPCHK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, s));
    return d;
}

// Post-processing over n point slots: k_post (unless M rows are given),
// k_filter, then k_det_mod on the points it listed.  Outputs go to the
// non-null pointers of `a` (filled in by the caller: slots, inputs,
// outputs, bad, tot).  filter_setup uploads kappa in the filter's forms
// and the primes and sizes the lists and the launches; filter_body
// launches (a segment body: no host work, no uploads).
struct FiltLaunch {
    FiltArgs a;
    int nmax = 8, grid = 0, pgrid = 0;
    size_t lds = 0, plds = 0;
    bool any_sparse = false;
    unsigned long long *d_mc = nullptr, *d_kc = nullptr;
};

static FiltLaunch filter_setup(DevCtx *C, hipStream_t s, const pfa_run_input *gin,
                               const std::vector<int64_t> &kofs, FiltArgs a, char *errbuf)
{
    static const std::vector<PfaModP> pr = primes31(96);
    std::vector<int32_t> kT, spR, spKV, spInts;
    std::vector<double> kTmax;
    std::vector<int64_t> kTofs, spRofs, spKVofs;
    build_kT(gin, kofs, kT, kTofs, kTmax, spR, spKV, spRofs, spKVofs, spInts);
    FiltLaunch F;
    // LDS for kappa: the padded layout of the launch's NMAX if it fits, and
    // room for the largest sparse form up to 48 KB
    const int nl = a.hl - 1;
    const int nmax = nl <= 8 ? 8 : nl <= 12 ? 12 : nl <= 16 ? 16 : nl <= 20 ? 20 : nl <= 24 ? 24 : nl <= 32 ? 32 : 40;
    int lds_floats = (nmax + 1) * nmax * nmax;
    if (lds_floats > 64 * 1024 / 4) lds_floats = 0;
    if (nmax > 24)
        for (int g = 0; g < gin->n_geo; ++g)
            if (spRofs[g] >= 0 && spInts[g] <= 12 * 1024) lds_floats = std::max(lds_floats, (int)spInts[g]);
    F.d_mc = (unsigned long long *)dev_get(C, B_MODCNT, 8, errbuf);
    Packed P;
    const size_t o_kT = P.add(kT.data(), 4 * kT.size()), o_kTo = P.add(kTofs.data(), 8 * kTofs.size()),
                 o_kTm = P.add(kTmax.data(), 8 * kTmax.size()), o_spR = P.add(spR.data(), 4 * spR.size()),
                 o_spKV = P.add(spKV.data(), 4 * spKV.size()),
                 o_spRo = P.add(spRofs.data(), 8 * spRofs.size()),
                 o_spKVo = P.add(spKVofs.data(), 8 * spKVofs.size()),
                 o_spI = P.add(spInts.data(), 4 * spInts.size()),
                 o_pr = P.add(pr.data(), sizeof(PfaModP) * pr.size());
    char *d_f = packed_upload(C, 1, B_UPFILT, P, s, errbuf);
    a.kT = (const int32_t *)(d_f + o_kT);
    a.kTofs = (const int64_t *)(d_f + o_kTo);
    a.kTmax = (const double *)(d_f + o_kTm);
    a.spR = (const int32_t *)(d_f + o_spR);
    a.spKV = (const int32_t *)(d_f + o_spKV);
    a.spRofs = (const int64_t *)(d_f + o_spRo);
    a.spKVofs = (const int64_t *)(d_f + o_spKVo);
    a.spInts = (const int32_t *)(d_f + o_spI);
    a.n_geo = gin->n_geo;
    a.lds_floats = a.filter ? lds_floats : 0;
    a.primes = (const PfaModP *)(d_f + o_pr);
    a.np = (int)pr.size();
    a.modcap = std::max<int64_t>(1, std::min<int64_t>(a.n, 1 << 20));
    a.modkp = (int64_t *)dev_get(C, B_MODLIST, 8 * a.modcap, errbuf);
    a.modnexp = (int64_t *)dev_get(C, B_MODNEXP, 8 * a.modcap, errbuf);
    a.modM = (int64_t *)dev_get(C, B_MODM, 8 * a.modcap * a.MH, errbuf);
    a.modcount = F.d_mc;
    if (a.n > 0 && !a.Min) {
        F.d_kc = (unsigned long long *)dev_get(C, B_KEEPCNT, 8, errbuf);
        a.keepkp = (int64_t *)dev_get(C, B_KEEPKP, 8 * a.n, errbuf);
        a.keepnexp = (int64_t *)dev_get(C, B_KEEPNEXP, 8 * a.n, errbuf);
        a.keepmeta = (int64_t *)dev_get(C, B_KEEPMETA, 8 * a.n, errbuf);
        a.rawkp = (int64_t *)dev_get(C, B_KEEPRAW, 24 * a.n, errbuf);
        a.rawnexp = a.rawkp + a.n;
        a.rawmeta = a.rawnexp + a.n;
        a.keepcount = F.d_kc;
        // contiguous runs of 64-slot groups, 16 wavefronts per CU
        F.pgrid = (int)std::max<int64_t>(1, std::min<int64_t>((a.n + 63) / 64, (int64_t)std::max(C->cus, 1) * 16));
        F.plds = (a.M ? (size_t)64 * a.MH * 8 : 0) + (size_t)64 * a.D * 4;
    }
    const int R = 64 / a.hl;
    // list length unknown on the host: size the grid by the slots
    const int64_t rounds = (a.n + 4 * R - 1) / (4 * R);
    F.grid = (int)std::min<int64_t>(rounds, 16384);
    // kappa, then 4 wavefronts x R groups x NMAX steps of (double2, int32)
    F.lds = (size_t)((a.lds_floats + 3) & ~3) * 4 + (size_t)4 * R * nmax * 20;
    // the sparse form where the padded dense layout cannot be staged (NMAX
    // > 24); below that the unrolled dense build is faster even at 20 %
    // density
    for (int g = 0; g < gin->n_geo; ++g)
        F.any_sparse |= nmax > 24 && spRofs[g] >= 0 && spInts[g] <= a.lds_floats;
    F.nmax = nmax;
    F.a = a;
#define PFA_FILTER_ATTR(NM)                                                                    \
    HIPCHK(hipFuncSetAttribute((const void *)(F.any_sparse ? k_filter<NM, true> : k_filter<NM, false>), \
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)F.lds))
    switch (nmax) {
    case 8: PFA_FILTER_ATTR(8); break;
    case 12: PFA_FILTER_ATTR(12); break;
    case 16: PFA_FILTER_ATTR(16); break;
    case 20: PFA_FILTER_ATTR(20); break;
    case 24: PFA_FILTER_ATTR(24); break;
    case 32: PFA_FILTER_ATTR(32); break;
    default: PFA_FILTER_ATTR(40); break;
    }
#undef PFA_FILTER_ATTR
    return F;
}

static void filter_body(Seg &sg, const FiltLaunch &F)
{
    const FiltArgs &a = F.a;
    sg.zero(F.d_mc, 8);
    if (a.n <= 0) return;
    if (!a.Min) {
        sg.zero(F.d_kc, 8);
        auto kpost = a.D <= 8 ? k_post<8> : a.D <= 12 ? k_post<12> : a.D <= 16 ? k_post<16> :
                     a.D <= 20 ? k_post<20> : a.D <= 24 ? k_post<24> : a.D <= 32 ? k_post<32> :
                     k_post<40>;
        sg.k(kpost, dim3(F.pgrid), dim3(64), F.plds, a);
        if (!a.filter) return;
    }
#define PFA_FILTER_LAUNCH(NM)                                                                  \
    do {                                                                                        \
        sg.k(F.any_sparse ? k_filter<NM, true> : k_filter<NM, false>, dim3(F.grid), dim3(256),  \
             F.lds, a);                                                                         \
        if (a.filter) sg.k(k_det_mod<NM>, dim3(512), dim3(256), 0, a);                         \
    } while (0)
    switch (F.nmax) {
    case 8: PFA_FILTER_LAUNCH(8); break;
    case 12: PFA_FILTER_LAUNCH(12); break;
    case 16: PFA_FILTER_LAUNCH(16); break;
    case 20: PFA_FILTER_LAUNCH(20); break;
    case 24: PFA_FILTER_LAUNCH(24); break;
    case 32: PFA_FILTER_LAUNCH(32); break;
    default: PFA_FILTER_LAUNCH(40); break;
    }
#undef PFA_FILTER_LAUNCH
}

// coniZpM's post-processing (pfvs/coni.py:313-412) over points grouped by
// key (ascending) and in canonical (M, Kn) order within a key: for the kept
// (n_exp > 0, or all when nexp is null) nonsingular points of a key, with
// Kgcd = gcd(Kn[1:]) (0 -> 1), raw = (q + M0 Kn0) / Kgcd and
// g = 1 .. (Q Kgcd - 1) / q, every K0 from ceil((g raw - Q) / M0) to
// min(floor((g raw - Q) / M0), ceil(Kn0 g / Kgcd) - 1) gives
// K = (K0, g Kn[1:] / Kgcd) with the point's M; rows by g, then the points
// in their order, then K0.  The tadpole -K.M <= Q is checked (an error, as
// coni.py:374-386 raises).  A key with an undecided point (verdict 2), a
// value beyond int64, g beyond 10^6, or skip[key] set is left to the caller
// (done[key] = 0).  When nexp is given, each point's row count must equal
// it (the device's count checked against this expansion).
struct AsmIn {
    int64_t n_keys, n;
    const int64_t *key, *M, *Kn, *q, *nexp;
    const int8_t *nonsing;
    int stride;                       // of M, Kn and of the output rows
    const int32_t *kh;                // per key: h11
    const int64_t *kQ;                // per key: Q
    const int8_t *skip;               // per key (null: none)
};

struct AsmOut {
    int64_t n_rows = 0;
    int64_t *Ks = nullptr, *Ms = nullptr, *keys = nullptr;
    int8_t *done = nullptr;
};

// Returns 0, or 1 with a message (tadpole violation, count mismatch).
static int assemble_core(const AsmIn &A, AsmOut &O, int nth, char *errbuf)
{
    const int S = A.stride;
    const int64_t nk = A.n_keys, n = A.n;
    // threads by the points to expand (a key without points costs nothing;
    // starting 16 threads twice costs about a millisecond)
    nth = (int)std::max<int64_t>(1, std::min<int64_t>(nth, n / 4096));
    O.done = (int8_t *)calloc((size_t)std::max<int64_t>(nk, 1), 1);
    std::vector<int64_t> pb(nk + 1, 0), rows(nk, -1), base(nk, 0);
    for (int64_t t = 0; t < n; ++t) pb[A.key[t] + 1]++;
    for (int64_t i = 0; i < nk; ++i) pb[i + 1] += pb[i];
    std::mutex emu;
    std::string emsg;
    // rows of key ip (-1: left to the caller); writes them from row `at`
    // when write is set
    auto one = [&](int64_t ip, bool write, int64_t at) -> int64_t {
        if (A.skip && A.skip[ip]) return -1;
        const int h = A.kh[ip];
        const pfa_i128 Q = A.kQ[ip];
        struct Pt { int64_t t; pfa_i128 kg, raw; int64_t gmax; };
        std::vector<Pt> pts;
        int64_t G = 0;
        for (int64_t t = pb[ip]; t < pb[ip + 1]; ++t) {
            if (A.nexp && A.nexp[t] <= 0) continue;
            if (A.nonsing[t] == 2) return -1;
            if (A.nonsing[t] != 1) continue;
            const int64_t *Kn = A.Kn + t * S, *M = A.M + t * S;
            uint64_t kg = 0;
            for (int i = 1; i < h; ++i)
                kg = pfa_box_gcd(kg, (uint64_t)(Kn[i] < 0 ? -(pfa_i128)Kn[i] : (pfa_i128)Kn[i]));
            if (kg == 0) kg = 1;
            const pfa_i128 q = A.q[t], QK = Q * (pfa_i128)kg, M0 = M[0];
            if (!(q < QK)) {                             // no g: the point gives no row
                if (A.nexp) return -1;                   // (the filter counted some)
                continue;
            }
            if (q <= 0 || M0 <= 0) return -1;
            const pfa_i128 num = q + M0 * (pfa_i128)Kn[0];
            if (num % (pfa_i128)kg != 0) return -1;
            const pfa_i128 gm = (QK - 1) / q;
            if (gm > 1000000) return -1;
            pts.push_back(Pt{t, (pfa_i128)kg, num / (pfa_i128)kg, (int64_t)gm});
            G = std::max<int64_t>(G, (int64_t)gm);
        }
        int64_t cnt = 0;
        std::vector<int64_t> per(pts.size(), 0);
        for (int64_t g = 1; g <= G; ++g)
            for (size_t k = 0; k < pts.size(); ++k) {
                const Pt &P = pts[k];
                if (g > P.gmax) continue;
                const int64_t *Kn = A.Kn + P.t * S, *M = A.M + P.t * S;
                const pfa_i128 M0 = M[0], a = (pfa_i128)g * P.raw - Q;
                const pfa_i128 lo = pfa_cdiv_pos(a, M0);
                pfa_i128 up = pfa_fdiv128(a, M0);
                const pfa_i128 cap = pfa_cdiv_pos((pfa_i128)Kn[0] * g, P.kg) - 1;
                if (cap < up) up = cap;
                if (up < lo) continue;
                if (!pfa_fits64(lo) || !pfa_fits64(up)) return -1;
                for (pfa_i128 K0 = lo; K0 <= up; ++K0) {
                    pfa_i128 tad = -(K0 * (pfa_i128)M[0]);
                    int64_t *Ko = write ? O.Ks + (at + cnt) * S : nullptr;
                    if (Ko) Ko[0] = (int64_t)K0;
                    for (int i = 1; i < h; ++i) {
                        const pfa_i128 Ki = (pfa_i128)g * ((pfa_i128)Kn[i] / P.kg);
                        if (!pfa_fits64(Ki)) return -1;
                        tad -= Ki * (pfa_i128)M[i];
                        if (Ko) Ko[i] = (int64_t)Ki;
                    }
                    if (tad > Q) {
                        std::lock_guard<std::mutex> lk(emu);
                        if (emsg.empty()) {
                            char b[256];
                            snprintf(b, sizeof b, "Tadpole violation: -dot(K,M)=%lld > Q=%lld",
                                     (long long)tad, (long long)Q);
                            emsg = b;
                        }
                        return -1;
                    }
                    if (write) {
                        int64_t *Mo = O.Ms + (at + cnt) * S;
                        for (int i = 0; i < S; ++i) Mo[i] = i < h ? M[i] : 0;
                        for (int i = h; i < S; ++i) O.Ks[(at + cnt) * S + i] = 0;
                        O.keys[at + cnt] = ip;
                    }
                    ++cnt;
                    ++per[k];
                }
            }
        if (A.nexp)
            for (size_t k = 0; k < pts.size(); ++k)
                if (per[k] != A.nexp[pts[k].t]) {
                    std::lock_guard<std::mutex> lk(emu);
                    if (emsg.empty()) emsg = "expansion count differs from the filter's";
                    return -1;
                }
        return cnt;
    };
    parallel_chunks(nk, 64, nth, [&](int64_t lo, int64_t hi) {
        for (int64_t ip = lo; ip < hi; ++ip) rows[ip] = one(ip, false, 0);
    });
    if (!emsg.empty()) {
        snprintf(errbuf, 256, "%s", emsg.c_str());
        return 1;
    }
    int64_t tot = 0;
    for (int64_t ip = 0; ip < nk; ++ip) {
        base[ip] = tot;
        if (rows[ip] >= 0) tot += rows[ip];
    }
    const size_t nb = (size_t)std::max<int64_t>(tot, 1);
    O.n_rows = tot;
    O.Ks = (int64_t *)malloc(8 * nb * S);
    O.Ms = (int64_t *)malloc(8 * nb * S);
    O.keys = (int64_t *)malloc(8 * nb);
    parallel_chunks(nk, 64, nth, [&](int64_t lo, int64_t hi) {
        for (int64_t ip = lo; ip < hi; ++ip)
            if (rows[ip] >= 0) {
                one(ip, true, base[ip]);
                O.done[ip] = 1;
            }
    });
    return 0;
}

// The finish of a run: the core over the run's sorted output, keys = the
// p-vectors, skipping those with a status.
static int assemble_pfvs(const pfa_run_input *in, pfa_run_output *out, int nth, char *errbuf)
{
    const int64_t n_p = in->n_p;
    std::vector<int32_t> kh(std::max<int64_t>(n_p, 1));
    std::vector<int64_t> kQ(std::max<int64_t>(n_p, 1));
    std::vector<int8_t> skip(std::max<int64_t>(n_p, 1));
    for (int64_t ip = 0; ip < n_p; ++ip) {
        kh[ip] = in->h[in->pgeo[ip]];
        kQ[ip] = in->Q[in->pgeo[ip]];
        skip[ip] = out->pstat[ip] != 0;
    }
    AsmIn A{n_p, out->n, out->pidx, out->M, out->Kn, out->q, out->nexp, out->nonsing,
            PFA_MAX_H11, kh.data(), kQ.data(), skip.data()};
    AsmOut O;
    int rc = assemble_core(A, O, nth, errbuf);
    out->pdone = O.done;
    out->n_pfv = O.n_rows;
    out->Ks = O.Ks;
    out->Ms = O.Ms;
    out->pfv_pidx = O.keys;
    return rc;
}

// survivors_only: the arrays hold every point (the host path) and only
// those the filter keeps for the PFVs (n_exp > 0, not singular) are
// returned.  pcount: lattice points per p-vector.
static int finish_run(const pfa_run_input *in, pfa_run_output *out, int64_t npts, int HS,
                      const int32_t *pidx32, const int64_t *q, const int64_t *M, const int64_t *Kn,
                      const int64_t *nexp, const int8_t *nonsing, const int32_t *status,
                      const int32_t *code, const int8_t *bad, const int64_t *pcount,
                      bool survivors_only, char *errbuf)
{
    const int MH = PFA_MAX_H11;
    const int64_t n_p = in->n_p;
    const int nth = finish_threads(in);
    auto t0 = std::chrono::steady_clock::now();
    out->pstat = (int8_t *)malloc(std::max<int64_t>(n_p, 1));
    out->pcode = (int32_t *)malloc(sizeof(int32_t) * std::max<int64_t>(n_p, 1));
    for (int64_t i = 0; i < n_p; ++i) {
        int s = status[i];
        if (s == PFA_OK && bad[i]) s = PFA_HOST;
        out->pstat[i] = (int8_t)(s == PFA_OK ? 0 : (s == PFA_HOST ? 1 : s));
        out->pcode[i] = code[i];
    }
    std::vector<int64_t> idx;
    idx.reserve(npts);
    for (int64_t k = 0; k < npts; ++k)
        if (pidx32[k] >= 0 && out->pstat[pidx32[k]] == 0 &&
            !(survivors_only && nexp && nonsing && (nexp[k] <= 0 || nonsing[k] == 0)))
            idx.push_back(k);
    out->pcount = (int64_t *)malloc(sizeof(int64_t) * std::max<int64_t>(n_p, 1));
    for (int64_t i = 0; i < n_p; ++i) out->pcount[i] = pcount ? pcount[i] : 0;
    std::vector<int64_t> pidx64(npts);
    for (int64_t k = 0; k < npts; ++k) pidx64[k] = pidx32[k];
    int64_t n_valid = 0;
    for (int64_t k = 0; k < npts; ++k) n_valid += pidx32[k] >= 0;
    if (in->sort) canonical_order(idx, pidx64.data(), M, Kn, HS, n_p, nth);
    const int64_t n = (int64_t)idx.size();
    out->n = n;
    size_t nb = std::max<int64_t>(n, 1);
    out->M = (int64_t *)malloc(sizeof(int64_t) * nb * MH);
    out->Kn = (int64_t *)malloc(sizeof(int64_t) * nb * MH);
    out->q = (int64_t *)malloc(sizeof(int64_t) * nb);
    out->pidx = (int64_t *)malloc(sizeof(int64_t) * nb);
    out->nexp = (int64_t *)malloc(sizeof(int64_t) * nb);
    out->nonsing = (int8_t *)malloc(nb);
    parallel_chunks(n, 1 << 14, nth, [&](int64_t lo, int64_t hi) {
        for (int64_t t = lo; t < hi; ++t) {
            int64_t k = idx[t];
            for (int i = 0; i < MH; ++i) {
                out->M[t * MH + i] = i < HS ? M[k * HS + i] : 0;
                out->Kn[t * MH + i] = i < HS ? Kn[k * HS + i] : 0;
            }
            out->q[t] = q[k];
            out->pidx[t] = pidx64[k];
            out->nexp[t] = nexp ? nexp[k] : -1;
            out->nonsing[t] = nonsing ? nonsing[k] : 2;
        }
    });
    // filter totals over every point of a finished p-vector (the returned
    // ones, or all of them when only the survivors are returned)
    int64_t kept = 0, nex = 0, nns = 0, und = 0, nhost = 0;
    if (nexp)
        for (int64_t k = 0; k < npts; ++k) {
            if (pidx32[k] < 0 || out->pstat[pidx32[k]] != 0) continue;
            if (nexp[k] > 0) {
                kept++;
                if (nonsing && nonsing[k] == 1) { nex += nexp[k]; nns++; }
                if (nonsing && nonsing[k] == 2) und++;
            }
        }
    for (int64_t i = 0; i < n_p; ++i) nhost += out->pstat[i] == 1;
    out->stats[PFA_ST_POINTS] = (double)n_valid;
    out->stats[PFA_ST_KEPT] = (double)kept;
    out->stats[PFA_ST_EXPANSIONS] = (double)nex;
    out->stats[PFA_ST_NONSING] = (double)nns;
    out->stats[PFA_ST_UNDECIDED] = (double)und;
    out->stats[PFA_ST_HOST_P] = (double)nhost;
    int rc = 0;
    if (in->assemble && in->sort && in->filter && nexp && nonsing)
        rc = assemble_pfvs(in, out, nth, errbuf);
    out->stats[PFA_ST_MS_FINISH] =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

// --------------------------------------------------------- device run
// The search over the prepared records (statuses PFA_OK where searchable)
// in the chosen form, growing and rerunning on overflow: the point buffer
// (all forms) and the plain form's deques.  A rerun restores the record
// statuses the previous attempt changed (saved on the device first).
// nodes per wavefront the split aims at, and the deque capacity that
// takes such a slice with room for the depth-first search below it
#define PFA_SPLIT_DEFAULT 2048
static int default_stack_cap(int D, int split_per_wave)
{
    return std::max(64 * (D + 2), split_per_wave > 0 ? 2 * std::min(split_per_wave, 1 << 16) + 128 : 0);
}

struct SearchTune {
    int mode, waves_per_cu, budget, stack_cap, split;
};

struct SearchRun {
    unsigned long long npts = 0, queue_peak = 0;
    double ms_roots = 0, ms_search = 0, wasted_ms = 0;
    int passes = 0, launches = 0, attempts = 0, nwaves = 0, stack_cap = 0;
    int32_t *d_pidx = nullptr;
    int64_t *d_q = nullptr;
    uint64_t *d_g = nullptr;
    int32_t *d_c = nullptr;
    unsigned long long *d_cnt = nullptr;     // nodes, cand, leaf, steps, wait, cycles, donated, fetches
};

static void search_device(DevCtx *C, hipStream_t s, int64_t *d_recs, const RecL &L, int64_t n_p,
                          int D, const SearchTune &tn, unsigned long long pcap, SearchRun &R,
                          char *errbuf)
{
    const EntL E = ent_layout(D);
    // the batched and plain forms keep PFA_SL deque entries per wavefront in
    // LDS: no more wavefronts per CU than LDS holds resident
    const size_t bat_lds = (size_t)PFA_SL * E.bytes;
    const size_t bat_lds_all = bat_lds + 64 * 4 + 16 + 64 * PFA_CSTRIDE * 4;
    int waves_per_cu = tn.waves_per_cu > 0 ? tn.waves_per_cu : 4;
    waves_per_cu = std::max(1, std::min<int>(waves_per_cu, (int)(C->lds_cu / bat_lds_all)));
    const int donate_max = tn.budget > 0 ? tn.budget : 8;
    const int donate_min = 64;               // deque depth before a wavefront donates
    const int pt_chunk = 256;                // point slots reserved per atomic
    const bool plain = tn.mode == PFA_MODE_PLAIN, lane = tn.mode == PFA_MODE_LANE;
    const int split_per_wave = tn.split == 0 ? PFA_SPLIT_DEFAULT : tn.split;
    int stack_cap = tn.stack_cap > 0 ? std::max(tn.stack_cap, 128)
                                     : default_stack_cap(D, (plain || lane) ? -1 : split_per_wave);
    // plain: one workgroup per p-vector up to 8192, each p-vector's tree on
    // the wavefront that starts it
    const int nwaves = plain ? (int)std::max<int64_t>(1, std::min<int64_t>(n_p, 8192))
                             : std::max(1, C->cus * waves_per_cu);
    // batched: the split's last frontier (at most the roots, or `room`
    // nodes) starts the deques, so every slice fits with room to grow
    if (!plain && !lane)
        stack_cap = std::max<int64_t>(stack_cap, (n_p + nwaves - 1) / nwaves + 128);
    const int n_exp = (plain || lane || tn.split < 0) ? 0 : std::min(D, PFA_SPLIT_MAX - 1);
    const unsigned long long room = (unsigned long long)nwaves * (unsigned long long)(stack_cap / 2);
    const unsigned long long fcap = std::max<unsigned long long>(room, (unsigned long long)n_p) + 64;
    // room for every root plus the donation reserve kept by donate()
    const unsigned long long ring_cap = (unsigned long long)std::max<int64_t>(2 * n_p + 64, 1 << 20) +
                                        2 * 64ULL * 4096ULL;
    const char *dl_env = getenv("PFVS_SEARCH_DEADLINE_S");
    const double deadline_s = dl_env ? atof(dl_env) : 120.0;
    const unsigned long long deadline =
        (unsigned long long)(deadline_s * 1000.0 * (double)std::max(C->wclk_khz, 1));
    hipEvent_t *ev = C->ev;
    unsigned long long *d_ctl = (unsigned long long *)dev_get(C, B_CTL, CTL_BYTES, errbuf);
    int *d_ovf = (int *)(d_ctl + 16 * 8);
    R.d_cnt = (unsigned long long *)dev_get(C, B_CNT, 64, errbuf);
    int32_t *d_save = (int32_t *)dev_get(C, B_STATSAVE, 4 * (size_t)std::max<int64_t>(n_p, 1), errbuf);
    const int sgrid = (int)std::min<int64_t>((n_p + 255) / 256 + 1, 4096);
    R.wasted_ms = 0;
    for (int attempt = 0;; ++attempt) {
        R.attempts = attempt + 1;
        R.d_pidx = (int32_t *)dev_get(C, B_PT_PIDX, 4 * pcap, errbuf);
        R.d_q = (int64_t *)dev_get(C, B_PT_Q, 8 * pcap, errbuf);
        R.d_c = (int32_t *)dev_get(C, B_PT_C, 4 * pcap * D, errbuf);
        R.d_g = (uint64_t *)dev_get(C, B_PT_G, 8 * pcap, errbuf);
        PointArgs Pt{R.d_pidx, R.d_q, R.d_g, R.d_c, pcap, d_ctl + 16 * 6, d_ovf};
        CountArgs Cn{R.d_cnt};
        R.passes = 1;
        R.launches = (n_p > 0 || !(plain || lane)) ? 1 : 0;
        unsigned long long err = 0;
        char *ring = nullptr, *stacks = nullptr;
        unsigned long long *seq = nullptr;
        if (!lane) {
            if (!plain) {
                ring = (char *)dev_get(C, B_RING, (size_t)ring_cap * E.bytes, errbuf);
                seq = (unsigned long long *)dev_get(C, B_SEQ, 8 * ring_cap, errbuf);
            }
            stacks = (char *)dev_get(C, B_STACKS, (size_t)nwaves * stack_cap * E.bytes, errbuf);
        }
        QueueArgs Qa{ring, seq, ring_cap, d_ctl};
        SplitArgs Sp{nullptr, nullptr, fcap, nullptr,
                     (unsigned long long)nwaves * (unsigned long long)std::max(split_per_wave, 1), room};
        if (!lane && !plain) {
            Sp.F0 = (char *)dev_get(C, B_F0, (size_t)fcap * E.bytes, errbuf);
            Sp.F1 = (char *)dev_get(C, B_F1, (size_t)fcap * E.bytes, errbuf);
            Sp.st = (unsigned long long *)dev_get(C, B_SPLITST, 256 * 8, errbuf);
        }
        BatchArgs A{d_recs, L, E, Qa, stacks, stack_cap, donate_max, donate_min, pt_chunk,
                    Pt, Cn, deadline, n_p, Sp};
        auto kern = plain ? k_search<true> : k_search<false>;
        if (!lane)
            HIPCHK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)bat_lds));
        // set-up: clears, the statuses saved (first attempt) or restored,
        // the roots of the batched form; then the search
        HIPCHK(hipEventRecord(ev[2], s));
        run_seg(C, s, [&](Seg &sg) {
            sg.zero(d_ctl, CTL_BYTES);
            sg.zero(R.d_cnt, 64);
            if (n_p > 0)
                sg.k(k_status_io, dim3(sgrid), dim3(256), 0, n_p, d_recs, L.words, d_save,
                     attempt == 0 ? 0 : 1);
            if (!lane && !plain) {
                sg.zero(seq, 8 * ring_cap);
                sg.zero(Sp.st, 256 * 8);
                if (n_p > 0) sg.k(k_roots, dim3((n_p + 255) / 256), dim3(256), 0, n_p, d_recs, L, Sp, E, Cn);
            }
        }, errbuf);
        HIPCHK(hipEventRecord(ev[3], s));
        run_seg(C, s, [&](Seg &sg) {
            if (lane) {
                if (n_p > 0) sg.k(k_lane, dim3((n_p + 63) / 64), dim3(64), 0, n_p, d_recs, L, Pt, D, Cn);
            } else if (plain) {
                if (n_p > 0) sg.k(kern, dim3(nwaves), dim3(64), bat_lds, A);
            } else {
                // the breadth-first split, then the persistent search from
                // its last frontier
                for (int t = 0; t < n_exp; ++t)
                    sg.k(k_expand, dim3(nwaves), dim3(64), (size_t)64 * E.bytes, t, d_recs, L, E, Sp, Pt,
                         Cn, pt_chunk);
                sg.k(k_split_final, dim3(1), dim3(64), 0, Sp, d_ctl, nwaves, n_exp);
                sg.k(kern, dim3(nwaves), dim3(64), bat_lds, A);
            }
        }, errbuf);
        HIPCHK(hipEventRecord(ev[4], s));
        // control words up to the overflow flag (word 16 * 8)
        unsigned long long hctl[16 * 9];
        HIPCHK(hipMemcpyAsync(hctl, d_ctl, sizeof hctl, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (!lane) {
            R.queue_peak = hctl[16];             // tickets issued: roots + donations
            err = hctl[80];
            if (err == 1) {
                snprintf(errbuf, 256, "frontier ring full (capacity %llu)", ring_cap);
                throw 1;
            }
            if (err == 2) {
                snprintf(errbuf, 256, "batched search exceeded its %.0f s deadline", deadline_s);
                throw 1;
            }
            if (!plain && hctl[32] != 0) {
                snprintf(errbuf, 256, "batched search ended with %lld units open", (long long)hctl[32]);
                throw 1;
            }
        }
        R.npts = hctl[16 * 6];
        const int ovf = (int)(((const int *)(hctl + 16 * 8))[0]);
        R.ms_roots = ms_between(ev[2], ev[3]);
        R.ms_search = ms_between(ev[3], ev[4]);
        C->pcap_hint = std::max(C->pcap_hint, R.npts);
        if (!ovf && err != 3) {
            R.npts = std::min<unsigned long long>(R.npts, pcap);   // trailing reserved slots
            break;
        }
        R.wasted_ms += R.ms_roots + R.ms_search;
        if (ovf) pcap = std::max<unsigned long long>(2 * pcap, R.npts + R.npts / 4 + 1024);
        if (err == 3) stack_cap *= 2;
        if (attempt > 8) {
            snprintf(errbuf, 256, ovf ? "point buffer overflow" : "plain search deque overflow");
            throw 1;
        }
    }
    R.nwaves = nwaves;
    R.stack_cap = stack_cap;
}

static int run_device(const pfa_run_input *in, pfa_run_output *out, char *errbuf)
{
    auto t_wall = std::chrono::steady_clock::now();
    DevCtx *C = ctx_for(in->device);
    std::unique_lock<std::mutex> lk(C->mu);
    ctx_init(C, in->device, errbuf);
    C->ms_alloc = 0;
    hipStream_t s = C->stream;
    const int MH = PFA_MAX_H11, D = batch_dmax(in);
    const int64_t n_p = in->n_p;
    const RecL L = rec_layout(D);
    std::vector<int64_t> kofs, mofs;
    int64_t ktot, mtot;
    fill_geo_offsets(in, kofs, mofs, ktot, mtot);
    int hmax = D + 1;
    size_t lds = pfa_ws_lds_bytes(hmax);
    if (lds > C->lds_max) {
        snprintf(errbuf, 256, "h11 = %d needs %zu B of LDS (device has %zu)", hmax, lds, C->lds_max);
        return 1;
    }
    auto up = [&](int slot, const void *src, size_t bytes) -> void * {
        void *d = dev_get(C, slot, std::max<size_t>(bytes, 16), errbuf);
        if (bytes) HIPCHK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, s));
        return d;
    };
    auto t_up = std::chrono::steady_clock::now();
    Packed P;
    const size_t o_h = P.add(in->h, 4 * in->n_geo), o_Q = P.add(in->Q, 8 * in->n_geo),
                 o_qm = P.add(in->qmax, 8 * in->n_geo), o_lm = P.add(in->linmin, 8 * in->n_geo),
                 o_k = P.add(in->kappa, 8 * ktot), o_mb = P.add(in->Mbasis, 8 * mtot),
                 o_ko = P.add(kofs.data(), 8 * in->n_geo), o_mo = P.add(mofs.data(), 8 * in->n_geo),
                 o_ps = P.add(in->ps, 8 * n_p * MH), o_pg = P.add(in->pgeo, 4 * n_p);
    char *d_in = packed_upload(C, 0, B_UPIN, P, s, errbuf);
    GeoArgs G;
    G.h = (const int32_t *)(d_in + o_h);
    G.Q = (const int64_t *)(d_in + o_Q);
    G.qmax = (const int64_t *)(d_in + o_qm);
    G.linmin = (const int64_t *)(d_in + o_lm);
    G.kappa = (const int64_t *)(d_in + o_k);
    G.Mbasis = (const int64_t *)(d_in + o_mb);
    G.kofs = (const int64_t *)(d_in + o_ko);
    G.mofs = (const int64_t *)(d_in + o_mo);
    const int64_t *d_ps = (const int64_t *)(d_in + o_ps);
    const int32_t *d_pgeo = (const int32_t *)(d_in + o_pg);
    int64_t *d_recs = (int64_t *)dev_get(C, B_RECS, 8 * (size_t)L.words * std::max<int64_t>(n_p, 1), errbuf);
    const double ms_upload =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_up).count();

    hipEvent_t *ev = C->ev;
    HIPCHK(hipEventRecord(ev[0], s));
    if (n_p > 0) {
        // one workgroup per p-vector up to 16384 (each with its own cold
        // workspace in global memory), looping beyond.  A batch of mixed
        // h11 runs one launch per size class, each with the LDS of its
        // largest h11 (a uniform h11 = 40 workspace would cut the
        // occupancy of small-h11 p-vectors threefold), largest first.
        std::vector<PrepClass> cls = prep_classes(in);
        int32_t *d_order = nullptr;
        size_t cold_max = 0;
        for (const PrepClass &c : cls)
            cold_max = std::max(cold_max, pfa_ws_cold_bytes(c.h) * (size_t)std::min<int64_t>(c.n, 16384));
        char *d_cold = (char *)dev_get(C, B_PREPCOLD, cold_max, errbuf);
        if (cls.size() > 1) {
            std::vector<int32_t> ord;
            ord.reserve(n_p);
            for (const PrepClass &c : cls) ord.insert(ord.end(), c.idx.begin(), c.idx.end());
            d_order = (int32_t *)up(B_PREPORDER, ord.data(), 4 * ord.size());
        }
        HIPCHK(hipFuncSetAttribute((const void *)k_prep, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        run_seg(C, s, [&](Seg &sg) {
            int64_t off = 0;
            for (const PrepClass &c : cls) {
                const int grid = (int)std::min<int64_t>(c.n, 16384);
                // the Gram-Schmidt running values in LDS unless they would
                // keep a launch of large h11 from holding all its p-vectors
                // at once (h11 <= 24: at least 9 per CU either way)
                const int64_t per_cu = (c.n + std::max(C->cus, 1) - 1) / std::max(C->cus, 1);
                const bool sp = c.h <= 24 || (size_t)per_cu * pfa_ws_lds_bytes(c.h, true) <= C->lds_cu;
                sg.k(k_prep, dim3(grid), dim3(64), pfa_ws_lds_bytes(c.h, sp), G, d_ps, MH, d_pgeo, c.n,
                     d_order ? d_order + off : nullptr, d_recs, L, d_cold, pfa_ws_cold_bytes(c.h), (int)sp);
                off += c.n;
            }
        }, errbuf, false);
    }
    HIPCHK(hipEventRecord(ev[1], s));

    // points buffer: grow and rerun the search on overflow
    unsigned long long pcap = std::max<unsigned long long>(
        std::max<int64_t>(1 << 16, 16 * n_p), C->pcap_hint + C->pcap_hint / 8);
    unsigned long long est =
        (unsigned long long)std::min(1.25 * C->ppp_hint * (double)n_p, (double)(1 << 26));
    // after pfa_reserve the per-p-vector estimate stays within the
    // reservation (a search that overflows grows the buffer and reruns)
    if (C->pts_reserved > 0) est = std::min(est, std::max(C->pts_reserved, pcap));
    pcap = std::max(pcap, est);
    SearchRun R;
    SearchTune tn{in->mode, in->waves_per_cu, in->budget, in->stack_cap, in->split};
    search_device(C, s, d_recs, L, n_p, D, tn, pcap, R, errbuf);
    // post: M, Kn, the expansion count and the det filter (k_filter).
    // keep_points 1: every point's rows are stored and copied back;
    // 2 (with the filter): only the points that can yield PFVs, gathered
    // by k_compact; 0: counters only
    const int64_t n = (int64_t)R.npts;
    const int HS = D + 1;                    // row stride of M, Kn on the device
    const int keep = in->keep_points == 2 && !in->filter ? 1 : in->keep_points;
    const bool rows = keep == 1;
    int64_t *d_M = rows ? (int64_t *)dev_get(C, B_M, 8 * std::max<int64_t>(n, 1) * HS, errbuf) : nullptr;
    int64_t *d_Kn = rows ? (int64_t *)dev_get(C, B_KN, 8 * std::max<int64_t>(n, 1) * HS, errbuf) : nullptr;
    int64_t *d_nexp = rows && in->filter ? (int64_t *)dev_get(C, B_NEXP, 8 * std::max<int64_t>(n, 1), errbuf) : nullptr;
    int8_t *d_ns = keep && in->filter ? (int8_t *)dev_get(C, B_NONSING, std::max<int64_t>(n, 1), errbuf) : nullptr;
    int8_t *d_bad = (int8_t *)dev_get(C, B_BAD, std::max<int64_t>(n_p, 1), errbuf);
    unsigned long long *d_tot = (unsigned long long *)dev_get(C, B_TOT, 64, errbuf);
    unsigned long long *d_pcount =
        (unsigned long long *)dev_get(C, B_PCOUNT, 8 * (size_t)std::max<int64_t>(n_p, 1), errbuf);
    FiltArgs Fa{};
    Fa.n = n;
    Fa.pidx = R.d_pidx;
    Fa.q = R.d_q;
    Fa.g = R.d_g;
    Fa.c = R.d_c;
    Fa.Min = nullptr;
    Fa.D = D;
    Fa.MH = HS;
    Fa.hl = HS;
    Fa.recs = d_recs;
    Fa.L = L;
    Fa.filter = in->filter;
    Fa.M = d_M;
    Fa.Kn = d_Kn;
    Fa.nexp = d_nexp;
    Fa.nonsing = d_ns;
    Fa.bad = d_bad;
    Fa.tot = d_tot;
    Fa.pcount = d_pcount;
    Fa.G = G;
    const FiltLaunch F = filter_setup(C, s, in, kofs, Fa, errbuf);
    // statuses and stage codes of the p-vectors, gathered for one copy
    int32_t *d_stat = (int32_t *)dev_get(C, B_STATOUT, 8 * (size_t)std::max<int64_t>(n_p, 1), errbuf);
    HIPCHK(hipEventRecord(ev[5], s));
    run_seg(C, s, [&](Seg &sg) {
        sg.zero(d_pcount, 8 * (size_t)std::max<int64_t>(n_p, 1));
        sg.zero(d_bad, std::max<int64_t>(n_p, 1));
        sg.zero(d_tot, 64);
        filter_body(sg, F);
        if (n_p > 0)
            sg.k(k_status_io, dim3((int)std::min<int64_t>((n_p + 255) / 256, 4096)), dim3(256), 0, n_p,
                 d_recs, L.words, d_stat, 2);
    }, errbuf, false);
    // the survivors (keep 2): compacted; their count read back, with a
    // second pass if they outgrew the buffer
    int64_t nsv = 0;
    int32_t *d_sv_pidx = nullptr;
    int64_t *d_sv_q = nullptr, *d_sv_M = nullptr, *d_sv_Kn = nullptr, *d_sv_nexp = nullptr;
    int8_t *d_sv_ns = nullptr;
    if (keep == 2 && n > 0) {
        unsigned long long *d_svc = (unsigned long long *)dev_get(C, B_SV_CNT, 64, errbuf);
        unsigned long long cap = std::max<unsigned long long>(1 << 16, C->sv_hint + C->sv_hint / 4);
        for (int pass = 0; pass < 2; ++pass) {
            CompactArgs Ca{};
            // the filter's list of points with n_exp > 0 (filled by k_post)
            Ca.keepkp = (const int64_t *)dev_get(C, B_KEEPKP, 8 * n, errbuf);
            Ca.keepnexp = (const int64_t *)dev_get(C, B_KEEPNEXP, 8 * n, errbuf);
            Ca.keepcount = (const unsigned long long *)dev_get(C, B_KEEPCNT, 8, errbuf);
            Ca.nonsing = d_ns;
            Ca.pidx = R.d_pidx;
            Ca.q = R.d_q;
            Ca.c = R.d_c;
            Ca.D = D;
            Ca.HS = HS;
            Ca.recs = d_recs;
            Ca.L = L;
            Ca.o_pidx = d_sv_pidx = (int32_t *)dev_get(C, B_SV_PIDX, 4 * cap, errbuf);
            Ca.o_q = d_sv_q = (int64_t *)dev_get(C, B_SV_Q, 8 * cap, errbuf);
            Ca.o_M = d_sv_M = (int64_t *)dev_get(C, B_SV_M, 8 * cap * HS, errbuf);
            Ca.o_Kn = d_sv_Kn = (int64_t *)dev_get(C, B_SV_KN, 8 * cap * HS, errbuf);
            Ca.o_nexp = d_sv_nexp = (int64_t *)dev_get(C, B_SV_NEXP, 8 * cap, errbuf);
            Ca.o_ns = d_sv_ns = (int8_t *)dev_get(C, B_SV_NS, cap, errbuf);
            Ca.o_count = d_svc;
            Ca.cap = cap;
            const int cgrid = (int)std::min<int64_t>((n + 63) / 64, 4096);
            run_seg(C, s, [&](Seg &sg) {
                sg.zero(d_svc, 8);
                sg.k(k_compact, dim3(cgrid), dim3(64), 0, Ca);
            }, errbuf);
            unsigned long long cnt = 0;
            HIPCHK(hipMemcpyAsync(&cnt, d_svc, 8, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            nsv = (int64_t)cnt;
            C->sv_hint = std::max(C->sv_hint, cnt);
            if (cnt <= cap) break;
            cap = cnt + cnt / 4 + 1024;
        }
    }
    HIPCHK(hipEventRecord(ev[6], s));
    // results back, into the pinned staging buffers
    HIPCHK(hipEventSynchronize(ev[6]));
    auto t_d2h = std::chrono::steady_clock::now();
    const int64_t nh = keep == 1 ? n : (keep == 2 ? nsv : 0);   // points copied back
    const size_t nb1 = (size_t)std::max<int64_t>(nh, 1);
    Stage &S = C->stage[C->next_stage];
    C->next_stage ^= 1;
    std::unique_lock<std::mutex> sl(S.mu);    // waits for the call two back to finish with it
    int32_t *h_pidx = (int32_t *)host_get(C, S, H_PIDX, 4 * nb1, errbuf);
    int64_t *h_q = (int64_t *)host_get(C, S, H_Q, 8 * nb1, errbuf);
    int64_t *h_M = (int64_t *)host_get(C, S, H_M, 8 * nb1 * HS, errbuf);
    int64_t *h_Kn = (int64_t *)host_get(C, S, H_KN, 8 * nb1 * HS, errbuf);
    int64_t *h_nexp = (int64_t *)host_get(C, S, H_NEXP, 8 * nb1, errbuf);
    int8_t *h_ns = (int8_t *)host_get(C, S, H_NS, nb1, errbuf);
    int32_t *h_stat = (int32_t *)host_get(C, S, H_STAT, 8 * (size_t)std::max<int64_t>(n_p, 1), errbuf);
    std::vector<int32_t> h_status(n_p), h_code(n_p);
    std::vector<int8_t> h_bad(n_p);
    std::vector<int64_t> h_pcount(std::max<int64_t>(n_p, 1), 0);
    unsigned long long htot[8] = {0};
    HIPCHK(hipMemcpyAsync(htot, d_tot, 64, hipMemcpyDeviceToHost, s));
    if (nh > 0 && keep == 1) {
        HIPCHK(hipMemcpyAsync(h_pidx, R.d_pidx, 4 * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_q, R.d_q, 8 * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_M, d_M, 8 * n * HS, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_Kn, d_Kn, 8 * n * HS, hipMemcpyDeviceToHost, s));
        if (in->filter) {
            HIPCHK(hipMemcpyAsync(h_nexp, d_nexp, 8 * n, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(h_ns, d_ns, n, hipMemcpyDeviceToHost, s));
        }
    } else if (nh > 0) {
        HIPCHK(hipMemcpyAsync(h_pidx, d_sv_pidx, 4 * nh, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_q, d_sv_q, 8 * nh, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_M, d_sv_M, 8 * nh * HS, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_Kn, d_sv_Kn, 8 * nh * HS, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_nexp, d_sv_nexp, 8 * nh, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_ns, d_sv_ns, nh, hipMemcpyDeviceToHost, s));
    }
    if (n_p > 0) {
        HIPCHK(hipMemcpyAsync(h_stat, d_stat, 8 * n_p, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_bad.data(), d_bad, n_p, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h_pcount.data(), d_pcount, 8 * n_p, hipMemcpyDeviceToHost, s));
    }
    unsigned long long hc[8] = {0};
    HIPCHK(hipMemcpyAsync(hc, R.d_cnt, 64, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int64_t ip = 0; ip < n_p; ++ip) {
        h_status[ip] = h_stat[2 * ip];
        h_code[ip] = h_stat[2 * ip + 1];
    }
    out->stats[PFA_ST_NODES] = (double)hc[0];
    out->stats[PFA_ST_CAND] = (double)hc[1];
    out->stats[PFA_ST_LEAF] = (double)hc[2];
    out->stats[PFA_ST_PASSES] = R.passes;
    out->stats[PFA_ST_SEARCH_LAUNCHES] = R.launches;
    out->stats[PFA_ST_MS_PREP] = ms_between(ev[0], ev[1]);
    out->stats[PFA_ST_MS_ROOTS] = R.ms_roots;
    out->stats[PFA_ST_MS_SEARCH] = R.ms_search;
    out->stats[PFA_ST_MS_POST] = ms_between(ev[5], ev[6]);
    // (a search rerun after the point buffer overflowed counts too)
    out->stats[PFA_ST_MS_TOTAL] = out->stats[PFA_ST_MS_PREP] + R.ms_roots + R.ms_search +
                                  out->stats[PFA_ST_MS_POST] + R.wasted_ms;
    out->stats[PFA_ST_QUEUE_PEAK] = (double)R.queue_peak;
    out->stats[PFA_ST_MS_D2H] =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_d2h).count();
    out->stats[PFA_ST_ATTEMPTS] = (double)R.attempts;
    out->stats[PFA_ST_MS_ALLOC] = C->ms_alloc;
    out->stats[PFA_ST_MS_UPLOAD] = ms_upload;
    out->stats[PFA_ST_STEPS] = (double)hc[3];
    out->stats[PFA_ST_WAIT_CYC] = (double)hc[4];
    out->stats[PFA_ST_WAVE_CYC] = (double)hc[5];
    out->stats[PFA_ST_DONATED] = (double)hc[6];
    out->stats[PFA_ST_FETCHES] = (double)hc[7];
    out->stats[PFA_ST_WAVES] = R.nwaves;
    out->stats[PFA_ST_GRAPHS] = graphs_on() ? (double)C->graphs.size() : 0.0;
    out->stats[PFA_ST_GRAPH_HITS] = (double)C->graph_hits;
    out->stats[PFA_ST_GRAPH_CAPTURES] = (double)C->graph_captures;
    out->stats[PFA_ST_SURVIVORS] = (double)nsv;
    if (n_p >= 64) C->ppp_hint = std::max(C->ppp_hint, (double)R.npts / (double)n_p);
    lk.unlock();                              // the device is free; assemble from the stage
    // counts-only totals from the device (htot: points, kept, expansions,
    // nonsingular, undecided)
    double st_keep[5] = {(double)htot[1], (double)htot[2], (double)htot[3], (double)htot[4],
                         (double)htot[0]};
    int rc = finish_run(in, out, nh, HS, h_pidx, h_q, h_M, h_Kn,
                        in->filter ? h_nexp : nullptr, in->filter ? h_ns : nullptr,
                        h_status.data(), h_code.data(), h_bad.data(), h_pcount.data(), false,
                        errbuf);
    if (keep != 1) {
        out->stats[PFA_ST_POINTS] = st_keep[4];
        out->stats[PFA_ST_KEPT] = st_keep[0];
        out->stats[PFA_ST_EXPANSIONS] = st_keep[1];
        out->stats[PFA_ST_NONSING] = st_keep[2];
        out->stats[PFA_ST_UNDECIDED] = st_keep[3];
    }
    out->stats[PFA_ST_MS_WALL] =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
    return rc;
}

// ------------------------------------------------------------ host run
static int run_host(const pfa_run_input *in, pfa_run_output *out, char *errbuf)
{
    auto t0 = std::chrono::steady_clock::now();
    const int MH = PFA_MAX_H11, D = batch_dmax(in);
    const int64_t n_p = in->n_p;
    const RecL L = rec_layout(D);
    std::vector<int64_t> kofs, mofs;
    int64_t ktot, mtot;
    fill_geo_offsets(in, kofs, mofs, ktot, mtot);
    GeoArgs G{in->h, in->Q, in->qmax, in->linmin, in->kappa, in->Mbasis, kofs.data(), mofs.data()};
    std::vector<int64_t> recs((size_t)L.words * std::max<int64_t>(n_p, 1));
    int nth = in->host_threads > 0 ? in->host_threads : (int)std::thread::hardware_concurrency();
    nth = std::max(1, std::min<int>(nth, (int)std::max<int64_t>(n_p, 1)));
    std::vector<HostPts> tp(nth);
    std::vector<PfaCounts> tc(nth, PfaCounts{0, 0, 0});
    std::vector<PfaWork> tw(nth, PfaWork{0, 0, 0, 0, 0, 0});
    std::vector<int64_t> tn(nth, 0);                 // points found per thread
    std::vector<double> tpw(nth, 0.0), tpw_max(nth, 0.0);   // preparation lane-ops
    std::vector<int64_t> pcount(std::max<int64_t>(n_p, 1), 0);   // points per p-vector
    const bool store = in->keep_points || in->filter; // else counts only
    std::atomic<int64_t> next{0};
    auto worker = [&](int t) {
        std::vector<char> ws(pfa_ws_bytes(D + 1) + 64);
        std::vector<PfaNode> st(PFA_DMAX);
        int32_t c[PFA_DMAX];
        PfaTeam T{0, 1};
        for (;;) {
            int64_t ip = next.fetch_add(1);
            if (ip >= n_p) break;
            int64_t *rec = recs.data() + ip * L.words;
            const double ops0 = PFA_OPS_NOW();
            prep_one(T, ws.data(), ws.data() + pfa_ws_lds_bytes(D + 1), G, in->ps, MH, in->pgeo, ip,
                     rec, L);
            const double ops = PFA_OPS_NOW() - ops0;
            tpw[t] += ops;
            tpw_max[t] = std::max(tpw_max[t], ops);
            if (rec_status(rec) != PFA_OK) continue;
            PfaProb P = rec_view(rec, L);
            HostPts &hp = tp[t];
            size_t mark = hp.q.size();
            int64_t found = 0;
            auto emit = [&](const int32_t *cc, int64_t q, uint64_t) -> int {
                found++;
                if (!store) return 0;
                hp.pidx.push_back((int32_t)ip);
                hp.q.push_back(q);
                for (int j = 0; j < D; ++j) hp.c.push_back(j < P.d ? cc[j] : 0);
                return 0;
            };
            int r = pfa_dfs_t<true>(P, st.data(), c, tc[t], emit, &tw[t]);
            if (r) {
                rec_status(rec) = r;
                hp.pidx.resize(mark);
                hp.q.resize(mark);
                hp.c.resize(mark * D);
            } else {
                tn[t] += found;
                pcount[ip] = found;
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t) th.emplace_back(worker, t);
    for (auto &x : th) x.join();
    auto t1 = std::chrono::steady_clock::now();
    // gather
    HostPts all;
    for (auto &hp : tp) {
        all.pidx.insert(all.pidx.end(), hp.pidx.begin(), hp.pidx.end());
        all.q.insert(all.q.end(), hp.q.begin(), hp.q.end());
        all.c.insert(all.c.end(), hp.c.begin(), hp.c.end());
    }
    const int64_t n = (int64_t)all.q.size();
    const int HS = D + 1;
    std::vector<int64_t> M((size_t)n * HS), Kn((size_t)n * HS), nexp(n);
    std::vector<int8_t> ns(n, 2), bad(std::max<int64_t>(n_p, 1), 0);
    PostArgs Pa{n, all.pidx.data(), all.q.data(), all.c.data(), D, HS, recs.data(), L, M.data(),
                Kn.data(), nexp.data(), bad.data(), in->filter};
    static const std::vector<PfaModP> pr = primes31(96);
    // M, Kn, the integrality count and the det filter, in contiguous chunks
    // over the host threads (a point's bad flag is a plain store of 1)
    auto post_worker = [&](int t) {
        const int64_t k0 = n * t / nth, k1 = n * (t + 1) / nth;
        std::vector<int64_t> N((HS - 1) * (HS - 1));
        std::vector<uint32_t> w((HS - 1) * (HS - 1));
        std::vector<pfa_i128> bw((HS - 1) * (HS - 1));
        for (int64_t k = k0; k < k1; ++k) {
            post_one(Pa, k);
            if (!in->filter) continue;
            if (nexp[k] <= 0) { ns[k] = 0; continue; }
            const int64_t *rec = recs.data() + (int64_t)all.pidx[k] * L.words;
            const int32_t *ri = (const int32_t *)rec;
            const int h = ri[24], geo = ri[26];
            if (kappa_M(in->kappa + kofs[geo], h, M.data() + k * HS, N.data(), 1)) { ns[k] = 2; continue; }
            ns[k] = (int8_t)pfa_nonsingular(N.data(), h - 1, pr.data(), (int)pr.size(), w.data(),
                                            bw.data(), 1);
        }
    };
    {
        std::vector<std::thread> pth;
        for (int t = 0; t < nth; ++t) pth.emplace_back(post_worker, t);
        for (auto &x : pth) x.join();
    }
    std::vector<int32_t> status(n_p), code(n_p);
    for (int64_t ip = 0; ip < n_p; ++ip) {
        const int32_t *ri = (const int32_t *)(recs.data() + ip * L.words);
        status[ip] = ri[0];
        code[ip] = ri[25];
    }
    PfaCounts tot{0, 0, 0};
    for (auto &x : tc) { tot.nodes += x.nodes; tot.cand += x.cand; tot.leaf += x.leaf; }
    PfaWork wk{0, 0, 0, 0, 0, 0};
    for (auto &x : tw) {
        wk.sk += x.sk; wk.skm0 += x.skm0; wk.skr += x.skr;
        wk.sr += x.sr; wk.ss += x.ss; wk.sleaf += x.sleaf;
    }
    out->stats[PFA_ST_SUM_K] = (double)wk.sk;
    out->stats[PFA_ST_SUM_K_M0] = (double)wk.skm0;
    out->stats[PFA_ST_SUM_K_R] = (double)wk.skr;
    out->stats[PFA_ST_SUM_R] = (double)wk.sr;
    out->stats[PFA_ST_SUM_S] = (double)wk.ss;
    out->stats[PFA_ST_SUM_LEAF] = (double)wk.sleaf;
    // lane-ops: 4 per coordinate of off/aoff, M0 bound and H prefix updates
    // plus 24 per node; 12 per candidate plus 4 + 5 s where an H row leads;
    // 4 per entry of the leaf's exact form plus 8
    out->stats[PFA_ST_WORK] = 4.0 * (double)(wk.sk + wk.skm0 + wk.skr) + 24.0 * (double)tot.nodes +
                              12.0 * (double)tot.cand + 4.0 * (double)wk.sr + 5.0 * (double)wk.ss +
                              4.0 * (double)wk.sleaf + 8.0 * (double)tot.leaf;
    out->stats[PFA_ST_NODES] = (double)tot.nodes;
    out->stats[PFA_ST_CAND] = (double)tot.cand;
    out->stats[PFA_ST_LEAF] = (double)tot.leaf;
    out->stats[PFA_ST_PASSES] = 1;
    for (int t = 0; t < nth; ++t) {
        out->stats[PFA_ST_PREP_WORK] += tpw[t];
        out->stats[PFA_ST_PREP_WORK_MAX] = std::max(out->stats[PFA_ST_PREP_WORK_MAX], tpw_max[t]);
    }
    out->stats[PFA_ST_MS_SEARCH] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    int rc = finish_run(in, out, n, HS, all.pidx.data(), all.q.data(), M.data(), Kn.data(),
                        in->filter ? nexp.data() : nullptr, in->filter ? ns.data() : nullptr,
                        status.data(), code.data(), bad.data(), pcount.data(),
                        in->keep_points == 2 && in->filter, errbuf);
    if (!store) {
        int64_t np = 0;
        for (int64_t x : tn) np += x;
        out->stats[PFA_ST_POINTS] = (double)np;
    }
    out->stats[PFA_ST_MS_WALL] =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    out->stats[PFA_ST_MS_TOTAL] = out->stats[PFA_ST_MS_WALL];
    return rc;
}

static int validate_run(const pfa_run_input *in, char *errbuf)
{
    if (in->n_geo < 0 || in->n_p < 0) { snprintf(errbuf, 256, "negative sizes"); return 1; }
    if (in->n_geo >= (1 << 24)) { snprintf(errbuf, 256, "more than 2^24 geometries"); return 1; }
    for (int g = 0; g < in->n_geo; ++g)
        if (in->h[g] < 3 || in->h[g] > PFA_MAX_H11) {
            snprintf(errbuf, 256, "h11 = %d outside 3..%d", in->h[g], PFA_MAX_H11);
            return 1;
        }
    for (int64_t i = 0; i < in->n_p; ++i)
        if (in->pgeo[i] < 0 || in->pgeo[i] >= in->n_geo) {
            snprintf(errbuf, 256, "pgeo: geometry index out of range");
            return 1;
        }
    return 0;
}

extern "C" int pfa_coni_run(const pfa_run_input *in, pfa_run_output *out)
{
    memset(out, 0, sizeof(*out));
    char *errbuf = out->err;
    if (validate_run(in, errbuf)) return 1;
    try {
        if (in->mode == PFA_MODE_HOST) return run_host(in, out, errbuf);
        return run_device(in, out, errbuf);
    } catch (int) {
        return 1;
    }
}

extern "C" void pfa_run_output_free(pfa_run_output *out)
{
    free(out->M); free(out->Kn); free(out->q); free(out->pidx);
    free(out->pstat); free(out->pcode); free(out->nexp); free(out->nonsing);
    free(out->Ks); free(out->Ms); free(out->pfv_pidx); free(out->pdone); free(out->pcount);
    out->M = out->Kn = out->q = out->pidx = out->nexp = out->pcount = nullptr;
    out->Ks = out->Ms = out->pfv_pidx = nullptr;
    out->pstat = nullptr; out->nonsing = nullptr; out->pcode = nullptr; out->pdone = nullptr;
    out->n = 0;
    out->n_pfv = 0;
}

extern "C" int pfa_assemble(int h, int64_t Q, int64_t n_keys, int64_t n, const int64_t *key,
                            const int64_t *M, const int64_t *Kn, const int64_t *q,
                            const int8_t *nonsing, int threads, pfa_assemble_output *out)
{
    memset(out, 0, sizeof(*out));
    char *errbuf = out->err;
    if (h < 2 || h > PFA_MAX_H11 || n < 0 || n_keys < 0) {
        snprintf(errbuf, 256, "pfa_assemble: bad sizes (h %d, n %lld, keys %lld)", h, (long long)n,
                 (long long)n_keys);
        return 1;
    }
    for (int64_t t = 0; t < n; ++t)
        if (key[t] < 0 || key[t] >= n_keys) {
            snprintf(errbuf, 256, "pfa_assemble: key out of range");
            return 1;
        }
    const int nth = threads > 0 ? threads : std::max(1, (int)std::thread::hardware_concurrency());
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    canonical_order(idx, key, M, Kn, h, n_keys, nth);
    std::vector<int64_t> sk(n), sM((size_t)n * h), sK((size_t)n * h), sq(n);
    std::vector<int8_t> sns(n);
    for (int64_t t = 0; t < n; ++t) {
        const int64_t k = idx[t];
        sk[t] = key[k];
        sq[t] = q[k];
        sns[t] = nonsing[k];
        memcpy(&sM[t * h], M + k * h, 8 * h);
        memcpy(&sK[t * h], Kn + k * h, 8 * h);
    }
    std::vector<int32_t> kh(std::max<int64_t>(n_keys, 1), h);
    std::vector<int64_t> kQ(std::max<int64_t>(n_keys, 1), Q);
    AsmIn A{n_keys, n, sk.data(), sM.data(), sK.data(), sq.data(), nullptr, sns.data(), h,
            kh.data(), kQ.data(), nullptr};
    AsmOut O;
    int rc = assemble_core(A, O, nth, errbuf);
    out->n_rows = O.n_rows;
    out->Ks = O.Ks;
    out->Ms = O.Ms;
    out->keys = O.keys;
    out->done = O.done;
    return rc;
}

extern "C" void pfa_assemble_output_free(pfa_assemble_output *out)
{
    free(out->Ks); free(out->Ms); free(out->keys); free(out->done);
    out->Ks = out->Ms = out->keys = nullptr;
    out->done = nullptr;
    out->n_rows = 0;
}

// ----------------------------------------------- reference-compatible ABI
extern "C" int pfg_coni_batch(const pfg_input *in, pfg_output *out)
{
    memset(out, 0, sizeof(*out));
    pfa_run_input r;
    memset(&r, 0, sizeof r);
    r.n_geo = in->n_geo; r.h = in->h; r.Q = in->Q; r.qmax = in->qmax; r.linmin = in->linmin;
    r.kappa = in->kappa; r.Mbasis = in->Mbasis; r.n_p = in->n_p; r.ps = in->ps; r.pgeo = in->pgeo;
    r.device = in->device; r.mode = PFA_MODE_BATCHED; r.filter = 0; r.sort = 1; r.keep_points = 1;
    r.verbose = in->verbose;
    pfa_run_output o;
    int rc = pfa_coni_run(&r, &o);
    if (rc) {
        memcpy(out->err, o.err, 256);
        pfa_run_output_free(&o);
        return rc;
    }
    out->n = o.n;
    out->M = o.M; out->Kn = o.Kn; out->q = o.q; out->pidx = o.pidx;
    // pstat 0/1 only: a kernel status (< 0) is redone by the host path too
    out->pstat = o.pstat;
    for (int64_t i = 0; i < in->n_p; ++i) if (out->pstat[i] != 0) out->pstat[i] = 1;
    out->seconds_gpu = o.stats[PFA_ST_MS_TOTAL] * 1e-3;
    free(o.pcode); free(o.nexp); free(o.nonsing); free(o.pcount);
    return 0;
}

extern "C" void pfg_output_free(pfg_output *out)
{
    free(out->M); free(out->Kn); free(out->q); free(out->pidx); free(out->pstat);
    out->M = out->Kn = out->q = out->pidx = nullptr;
    out->pstat = nullptr;
    out->n = 0;
}

extern "C" int pfg_max_h11(void) { return PFA_MAX_H11; }
extern "C" const char *pfg_backend(void) { return "hip"; }

extern "C" int pfg_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" void pfg_release(int device)
{
    if (device < 0) return;
    DevCtx *C;
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        if (device >= (int)g_ctx.size() || !g_ctx[device]) return;
        C = g_ctx[device];
    }
    std::lock_guard<std::mutex> lk(C->mu);
    if (hipSetDevice(device) != hipSuccess) return;
    if (C->stream) (void)hipStreamSynchronize(C->stream);
    // graphs refer to the buffers: drop them first
    for (auto &g : C->graphs) (void)hipGraphExecDestroy(g.second);
    C->graphs.clear();
    for (auto &b : C->bufs) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
    }
    for (Stage &S : C->stage) {
        std::lock_guard<std::mutex> sl(S.mu);
        for (auto &b : S.bufs) {
            if (b.p) (void)hipHostFree(b.p);
            b.p = nullptr;
            b.n = 0;
        }
    }
    for (DevBuf &b : C->up_host) {
        if (b.p) (void)hipHostFree(b.p);
        b.p = nullptr;
        b.n = 0;
    }
}

// --------------------------------------------------- p-vector enumeration
// prefix length: the fewest leading coordinates giving at least `target`
// units, at most n and at most 2^24 units
static int box_prefix(int n, int B, int64_t target)
{
    const int64_t W = 2 * (int64_t)B + 1;
    int s = 0;
    int64_t units = 1;
    while (s < n && units < target && units * W <= (1 << 24)) {
        units *= W;
        ++s;
    }
    return s;
}

// device form: one thread per unit, grid-stride; counts (write 0) or rows
// at their offsets (write 1)
__global__ __launch_bounds__(256) void k_box(PfaBox X, PfaBoxScr S0, int64_t n_units, int write,
                                             int64_t *counts, const int64_t *offs, int64_t *out,
                                             int64_t cap)
{
    PfaBoxScr S = S0;
    S.tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t u = S.tid; u < n_units; u += S.T) {
        if (write) pfa_box_unit_thread(X, S, u, out, offs[u], cap);
        else counts[u] = pfa_box_unit_thread(X, S, u, nullptr, 0, 0);
    }
}

extern "C" int pfa_box_enum(const int32_t *H, int m, int n, const int64_t *rhs, int B, int primitive,
                            int device, int64_t cap, int64_t *out, int64_t *n_out, int64_t *n_total,
                            double *ms, char *err)
{
    char *errbuf = err;
    *n_out = 0;
    *n_total = 0;
    if (ms) *ms = 0;
    if (m < 0 || m > PFA_BOX_MAXM || n < 1 || n > PFA_BOX_MAXN || B < 0 || B > (1 << 20) || cap < 0) {
        snprintf(errbuf, 256, "pfa_box_enum: unsupported sizes (m %d <= %d, 1 <= n %d <= %d, B %d)",
                 m, PFA_BOX_MAXM, n, PFA_BOX_MAXN, B);
        return 2;
    }
    // H^T and the reachable maxima rest[j][r] = B sum_{k >= j} |H_rk|
    const int mm = std::max(m, 1);
    std::vector<int32_t> Ht((size_t)n * mm, 0);
    std::vector<int64_t> rest((size_t)(n + 1) * mm, 0);
    for (int r = 0; r < m; ++r)
        for (int j = 0; j < n; ++j) Ht[(size_t)j * m + r] = H[(size_t)r * n + j];
    for (int j = n - 1; j >= 0; --j)
        for (int r = 0; r < m; ++r) {
            const int64_t a = H[(size_t)r * n + j];
            rest[(size_t)j * m + r] = rest[(size_t)(j + 1) * m + r] + (int64_t)B * (a < 0 ? -a : a);
        }
    std::vector<int64_t> rhs_v(mm, 0);
    for (int r = 0; r < m; ++r) rhs_v[r] = rhs[r];
    auto t0 = std::chrono::steady_clock::now();
    const int nth = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    const int s = box_prefix(n, B, device >= 0 ? 65536 : 64 * nth);
    int64_t n_units = 1;
    for (int k = 0; k < s; ++k) n_units *= 2 * (int64_t)B + 1;
    std::vector<int64_t> counts(n_units), offs(n_units);
    int64_t tot = 0;
    auto prefix = [&]() {
        tot = 0;
        for (int64_t u = 0; u < n_units; ++u) { offs[u] = tot; tot += counts[u]; }
    };
    try {
        if (device < 0) {
            // host threads: the device's per-unit walk, each thread with its
            // own scratch (T = 1)
            PfaBox X{m, n, B, primitive ? 1 : 0, s, Ht.data(), rest.data(), rhs_v.data()};
            auto run = [&](bool write) {
                parallel_chunks(n_units, 16, nth, [&](int64_t lo, int64_t hi) {
                    int64_t part[PFA_BOX_MAXM];
                    int32_t cur[PFA_BOX_MAXN], hh[PFA_BOX_MAXN];
                    uint64_t g[PFA_BOX_MAXN];
                    const PfaBoxScr S{part, cur, hh, g, 0, 1};
                    for (int64_t u = lo; u < hi; ++u) {
                        if (write) pfa_box_unit_thread(X, S, u, out, offs[u], cap);
                        else counts[u] = pfa_box_unit_thread(X, S, u, nullptr, 0, 0);
                    }
                });
            };
            run(false);
            prefix();
            run(true);
        } else {
            DevCtx *C = ctx_for(device);
            std::lock_guard<std::mutex> lk(C->mu);
            ctx_init(C, device, errbuf);
            hipStream_t st = C->stream;
            const int T = (int)std::min<int64_t>(std::max<int64_t>(n_units, 1), (int64_t)C->cus * 256);
            const int grid = (T + 255) / 256, Tg = grid * 256;
            // inputs and the per-thread scratch (element k of thread t at k Tg + t)
            size_t o = 0;
            auto take = [&](size_t bytes) { size_t at = o; o += (bytes + 255) & ~(size_t)255; return at; };
            const size_t oHt = take(4 * Ht.size()), oRest = take(8 * rest.size()), oRhs = take(8 * rhs_v.size());
            const size_t oPart = take(8 * (size_t)mm * Tg), oCur = take(4 * (size_t)n * Tg),
                         oHi = take(4 * (size_t)n * Tg), oG = take(8 * (size_t)n * Tg);
            const size_t oCnt = take(8 * (size_t)n_units), oOffs = take(8 * (size_t)n_units);
            char *d = (char *)dev_get(C, B_BOX, o, errbuf);
            HIPCHK(hipMemcpyAsync(d + oHt, Ht.data(), 4 * Ht.size(), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(d + oRest, rest.data(), 8 * rest.size(), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(d + oRhs, rhs_v.data(), 8 * rhs_v.size(), hipMemcpyHostToDevice, st));
            PfaBox X{m, n, B, primitive ? 1 : 0, s, (const int32_t *)(d + oHt), (const int64_t *)(d + oRest),
                     (const int64_t *)(d + oRhs)};
            PfaBoxScr S{(int64_t *)(d + oPart), (int32_t *)(d + oCur), (int32_t *)(d + oHi),
                        (uint64_t *)(d + oG), 0, Tg};
            int64_t *d_cnt = (int64_t *)(d + oCnt), *d_offs = (int64_t *)(d + oOffs);
            hipLaunchKernelGGL(k_box, dim3(grid), dim3(256), 0, st, X, S, n_units, 0, d_cnt,
                               (const int64_t *)d_offs, (int64_t *)nullptr, (int64_t)0);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(counts.data(), d_cnt, 8 * (size_t)n_units, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            prefix();
            const int64_t rows = std::min(tot, cap);
            if (rows > 0) {
                int64_t *d_out = (int64_t *)dev_get(C, B_BOXOUT, 8 * (size_t)rows * n, errbuf);
                HIPCHK(hipMemcpyAsync(d_offs, offs.data(), 8 * (size_t)n_units, hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(k_box, dim3(grid), dim3(256), 0, st, X, S, n_units, 1, d_cnt,
                                   (const int64_t *)d_offs, d_out, rows);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(out, d_out, 8 * (size_t)rows * n, hipMemcpyDeviceToHost, st));
            }
            HIPCHK(hipStreamSynchronize(st));
        }
        *n_total = tot;
        *n_out = std::min(tot, cap);
    } catch (int) {
        return 2;
    }
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return *n_total > cap ? 1 : 0;
}

// ------------------------------------------ coniZp's dilation bounds
// mu0 for every direction (pfa_bound.h), on host threads: the exact form,
// then every y with y^T Gr y at most Gr's smallest diagonal entry by the
// library's Fincke-Pohst search (pfa_dfs on a record with no cuts), and the
// smallest nonzero value.
static int coni_mu0_one(int h, const int64_t *kappa, const int64_t *p_r, int64_t *num, int64_t *den,
                        int64_t *K)
{
    std::unique_ptr<PfaBoundForm> Fp(new PfaBoundForm);
    PfaBoundForm &F = *Fp;
    int st = pfa_bound_form(h, kappa, p_r, F);
    if (st) return st;
    const int n = F.n;
    int64_t best = INT64_MAX;
    for (int k = 0; k < n; ++k) best = std::min(best, F.Gr[k * n + k]);
    if (best <= 0) return 1;
    const RecL L = rec_layout(n);
    std::vector<int64_t> rec((size_t)L.words, 0);
    int64_t *r = rec.data();
    int32_t *ri = (int32_t *)r;
    ri[0] = PFA_HOST;
    ri[1] = n;
    r[1] = best;
    r[2] = 1;
    r[3] = INT64_MIN;
    ri[20] = -1;
    ri[23] = 1;
    ri[24] = n + 1;
    int64_t mm = 0;
    for (int e = 0; e < n * n; ++e) {
        r[L.mat + e] = F.Gr[e];
        mm = std::max<int64_t>(mm, F.Gr[e] < 0 ? -F.Gr[e] : F.Gr[e]);
    }
    r[16] = mm;
    std::vector<char> ws((size_t)2 * n * n * 8 + 64);
    const PfaTeam T{0, 1};
    factor_one(T, ws.data(), r, L);
    if (rec_status(r) != PFA_OK) return -5;                     // ill-conditioned for FP64
    PfaProb P = rec_view(r, L);
    std::vector<PfaNode> stk(PFA_DMAX);
    int32_t c[PFA_DMAX], y[PFA_DMAX];
    PfaCounts cnt{0, 0, 0};
    int64_t qmin = INT64_MAX;
    auto emit = [&](const int32_t *cc, int64_t q, uint64_t) -> int {
        if (q > 0 && q < qmin) {
            qmin = q;
            for (int j = 0; j < n; ++j) y[j] = cc[j];
        }
        return 0;
    };
    const int rr = pfa_dfs(P, stk.data(), c, cnt, emit);
    if (rr) return -6;
    if (qmin == INT64_MAX) return -7;                           // cannot happen: the best basis vector
    if (pfa_bound_K(F, y, K)) return -1;
    const int64_t g = (int64_t)pfa_gcd64((uint64_t)qmin, (uint64_t)F.L);
    *num = qmin / g;
    *den = F.L / g;
    return 0;
}

extern "C" int pfa_coni_mu0(int h, const int64_t *kappa, int64_t n_p, const int64_t *ps, int threads,
                            int32_t *status, int64_t *num, int64_t *den, int64_t *K, char *err)
{
    if (h < 2 || h > PFA_MAX_H11 || n_p < 0) {
        snprintf(err, 256, "pfa_coni_mu0: h11 outside 2..%d or a negative count", PFA_MAX_H11);
        return 1;
    }
    const int m = h - 1;
    const int nth = threads > 0 ? threads
                                : std::max(1, std::min(64, (int)std::thread::hardware_concurrency()));
    parallel_chunks(n_p, 64, nth, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            num[i] = den[i] = 0;
            for (int j = 0; j < m; ++j) K[i * m + j] = 0;
            status[i] = coni_mu0_one(h, kappa, ps + i * m, num + i, den + i, K + i * m);
        }
    });
    return 0;
}

// ------------------------------------------------------------ coniZpK
// The lattice points behind the coni PFVs of a direction p_hat with
// dilation delta > D0 (the reference's coniZpK, pfvs/coni.py:554-652, with
// its C kernel pfvs/fp_kernel/coni_zpk.h).  Notation as pfa_bound.h:
// A = kappa . p_hat, S = -A_rr^-1, -s = sn / sd, Lambda = {K : p_r . K = 0}
// of rank n = m - 1.  A PFV with p = p_hat / delta has K_r in Lambda and
// tadpole Q = M0 eps + (M0^2 / delta)(-s) + delta K_r^T S K_r with eps > 0,
// so for -s > 0 the dilation delta > D0 forces K_r^T S K_r < Q / D0: K_r is
// one of finitely many short vectors of Lambda (the library's search over
// the reduced form Gr = L S in reduced coordinates x: x^T Gr x D0 < Q L).
// The lattice points c of the preparation (M = Binter c, Kn = ZB c) satisfy
// W c = delta K_r with W = ZB[1:].  For K_r = g K' (K' primitive) and o the
// order of K' in Lambda / Gamma, Gamma = W Z^m, the c with W c parallel to
// K_r are c = j c1 + t k (W c1 = o K', W k = 0) with delta = j o / g:
//   delta > D0                <=>  j o > g D0,
//   the tadpole with eps > 0  <=>  g L sn M0^2 < sd j o (Q L - g j o kn),
// kn = L K'^T S K', and M0 = j (l . c1) + t (l . k) >= M0min (l = Binter[0]),
// an arithmetic progression of M0 for each j.  Gamma and k come from a
// column echelon form of W in Lambda coordinates, Yc E = [Hc | 0] (E
// unimodular, Hc lower triangular with a positive diagonal): k = E e_{m-1};
// with Hadj = Hdet Hc^-1 and y' the Lambda coordinates of K',
// o = Hdet / gcd(Hdet, Hadj y') and c1 = E [Hadj y' / gcd(...); 0].  All
// exact (int128, checked); a value beyond it leaves the direction to the
// caller.  Status per direction: 0 done; 1 no short-vector bound (s >= 0,
// A_rr singular, S not positive definite on Lambda); < 0 undecided here
// (the caller searches that direction to its dilation bound instead).
struct ZpkPts {
    std::vector<int64_t> M, Kn, q, pidx;
};


// floor(a / b), b > 0
static inline pfa_i128 zpk_fdiv(pfa_i128 a, pfa_i128 b)
{
    pfa_i128 q = a / b;
    if ((a % b != 0) && (a < 0)) --q;
    return q;
}

// s a + t b = g = gcd(a, b) >= 0
static void zpk_xgcd(pfa_i128 a, pfa_i128 b, pfa_i128 *s, pfa_i128 *t, pfa_i128 *g)
{
    pfa_i128 r0 = a, r1 = b, s0 = 1, s1 = 0, t0 = 0, t1 = 1;
    while (r1 != 0) {
        const pfa_i128 q = r0 / r1;
        pfa_i128 x = r0 - q * r1; r0 = r1; r1 = x;
        x = s0 - q * s1; s0 = s1; s1 = x;
        x = t0 - q * t1; t0 = t1; t1 = x;
    }
    if (r0 < 0) { r0 = -r0; s0 = -s0; t0 = -t0; }
    *s = s0; *t = t0; *g = r0;
}

static int zpk_one(int h, const int64_t *kappa, const int64_t *Mbasis, const int64_t *p, int64_t Q,
                   int64_t M0min, int64_t D0, int64_t pidx, int64_t room, ZpkPts &out)
{
    const int m = h - 1, n = m - 1;
    if (h < 3 || m > PFA_DMAX) return -20;
    if (D0 < 1 || M0min < 1 || Q < 1) return -21;
    std::unique_ptr<PfaBoundForm> Fp(new PfaBoundForm);
    PfaBoundForm &F = *Fp;
    int st = pfa_bound_form(h, kappa, p + 1, F);
    if (st == 1) return 1;
    if (st) return -30 + st;
    if (F.sn <= 0) return 1;                                    // -s <= 0: no bound
    // the preparation's lattice: Binter (h x m), ZB (h x m), mat (m x m)
    std::vector<char> ws(pfa_ws_bytes(h) + 64);
    PfaWS Wk;
    pfa_ws_carve(Wk, ws.data(), ws.data() + pfa_ws_lds_bytes(h), h);
    const PfaTeam T{0, 1};
    if (pfa_build_lattice(T, Wk, kappa, Mbasis, p)) return -40;
    const std::vector<int64_t> Bi(Wk.Bi, Wk.Bi + h * m), ZB(Wk.ZB, Wk.ZB + h * m),
        mat(Wk.mat, Wk.mat + m * m);
    // Lambda coordinates of W's columns: U unimodular with U p_r = (g, 0, ...)
    // (the rows 1.. are F.O); z = U^-T v has z_0 = 0 for v in Lambda
    std::vector<int64_t> U(m * m);
    {
        int flag = 0, rc = pfa_unimodular(T, p + 1, m, U.data(), &flag);
        if (rc == -1) {
            pfa_i128 x[PFA_DMAX], r[PFA_DMAX];
            int ord[PFA_DMAX];
            rc = pfa_unimodular_reduced(T, p + 1, m, U.data(), x, r, ord, &flag);
        }
        if (rc) return -41;
        for (int a = 0; a < n; ++a)
            for (int i = 0; i < m; ++i)
                if (F.O[a * m + i] != U[(a + 1) * m + i]) return -42;   // cannot happen
    }
    const int w = 2 * m;
    std::vector<pfa_i128> Mx((size_t)m * w), Y((size_t)m * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            Mx[i * w + j] = U[j * m + i];
            Mx[i * w + m + j] = ZB[(i + 1) * m + j];
        }
    pfa_i128 det;
    if (pfa_bnd_solve(Mx.data(), m, w, &det, Y.data()) || (det != 1 && det != -1)) return -43;
    std::vector<pfa_i128> A((size_t)n * m), E((size_t)m * m, 0);
    for (int c = 0; c < m; ++c) {
        if (Y[c] != 0) return -44;                              // (W c lies in Lambda)
        for (int a = 0; a < n; ++a) A[a * m + c] = det * Y[(a + 1) * m + c];
    }
    for (int i = 0; i < m; ++i) E[i * m + i] = 1;
    // column operations [col c1, col c2] <- [col c1, col c2] [[a11, a12], [a21, a22]]
    auto colop = [&](int c1, int c2, pfa_i128 a11, pfa_i128 a12, pfa_i128 a21, pfa_i128 a22) -> int {
        auto one = [&](std::vector<pfa_i128> &X, int rows) -> int {
            for (int r = 0; r < rows; ++r) {
                const pfa_i128 x = X[r * m + c1], y = X[r * m + c2];
                pfa_i128 u, v, s1, s2;
                if (pfa_mul_ovf(a11, x, &u) || pfa_mul_ovf(a21, y, &v) || pfa_add_ovf(u, v, &s1)) return 1;
                if (pfa_mul_ovf(a12, x, &u) || pfa_mul_ovf(a22, y, &v) || pfa_add_ovf(u, v, &s2)) return 1;
                X[r * m + c1] = s1;
                X[r * m + c2] = s2;
            }
            return 0;
        };
        return one(A, n) || one(E, m);
    };
    for (int r = 0; r < n; ++r) {
        for (int c = r + 1; c < m; ++c) {
            if (A[r * m + c] == 0) continue;
            pfa_i128 s, t, g;
            zpk_xgcd(A[r * m + r], A[r * m + c], &s, &t, &g);
            if (colop(r, c, s, -A[r * m + c] / g, t, A[r * m + r] / g)) return -45;
        }
        if (A[r * m + r] == 0) return -46;                     // W of rank below n
        if (A[r * m + r] < 0 && colop(r, r == 0 ? 1 : 0, -1, 0, 0, 1)) return -45;
        for (int c = 0; c < r; ++c) {                           // left of the pivot into [0, pivot)
            const pfa_i128 q = zpk_fdiv(A[r * m + c], A[r * m + r]);
            if (q != 0 && colop(c, r, 1, 0, -q, 1)) return -45;
        }
    }
    std::vector<int64_t> k(m);
    for (int r = 0; r < n; ++r)
        if (A[r * m + m - 1] != 0) return -47;                  // cannot happen
    for (int i = 0; i < m; ++i) {
        if (!pfa_fits64(E[i * m + m - 1])) return -45;
        k[i] = (int64_t)E[i * m + m - 1];
    }
    for (int i = 0; i < m; ++i) {                               // W k = 0, exactly
        pfa_i128 s = 0;
        for (int c = 0; c < m; ++c)
            if (pfa_mac64(&s, ZB[(i + 1) * m + c], k[c])) return -45;
        if (s != 0) return -47;
    }
    // Hdet and Hadj = Hdet Hc^-1 (lower triangular), by forward substitution
    pfa_i128 Hdet = 1;
    for (int r = 0; r < n; ++r)
        if (pfa_mul_ovf(Hdet, A[r * m + r], &Hdet)) return -48;
    std::vector<pfa_i128> Hadj((size_t)n * n, 0);
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) {
            pfa_i128 s = i == j ? Hdet : 0;
            for (int c = j; c < i; ++c) {
                pfa_i128 t;
                if (pfa_mul_ovf(A[i * m + c], Hadj[c * n + j], &t) || pfa_sub_ovf(s, t, &s)) return -48;
            }
            if (s % A[i * m + i] != 0) return -48;              // cannot happen (Hdet Hc^-1 is integral)
            Hadj[i * n + j] = s / A[i * m + i];
        }
    pfa_i128 lk = 0;
    for (int c = 0; c < m; ++c)
        if (pfa_mac64(&lk, Bi[c], k[c])) return -49;
    if (lk == 0) return -49;                                    // M0 would not depend on t
    const pfa_i128 alk = lk < 0 ? -lk : lk;
    // the short vectors: x^T Gr x <= (Q L - 1) / D0, through the search
    pfa_i128 QL, coefL;
    if (pfa_mul_ovf((pfa_i128)Q, (pfa_i128)F.L, &QL)) return -50;
    const pfa_i128 qs = (QL - 1) / D0;
    if (qs < 1) return 0;
    if (!pfa_fits64(qs)) return -50;
    if (pfa_mul_ovf((pfa_i128)F.L, F.sn, &coefL)) return -50;      // L sn (times g below)
    const RecL L = rec_layout(n);
    std::vector<int64_t> rec((size_t)L.words, 0);
    int64_t *r = rec.data();
    int32_t *ri = (int32_t *)r;
    ri[0] = PFA_HOST;
    ri[1] = n;
    r[1] = (int64_t)qs;
    r[2] = 1;
    r[3] = INT64_MIN;
    ri[20] = -1;
    ri[23] = 1;
    ri[24] = n + 1;
    int64_t mm = 0;
    for (int e = 0; e < n * n; ++e) {
        r[L.mat + e] = F.Gr[e];
        mm = std::max<int64_t>(mm, F.Gr[e] < 0 ? -F.Gr[e] : F.Gr[e]);
    }
    r[16] = mm;
    std::vector<char> fws((size_t)2 * n * n * 8 + 64);
    factor_one(T, fws.data(), r, L);
    if (rec_status(r) != PFA_OK) return -51;                   // ill-conditioned for FP64
    PfaProb P = rec_view(r, L);
    std::vector<PfaNode> stk(PFA_DMAX);
    int32_t cc[PFA_DMAX];
    PfaCounts cnt{0, 0, 0};
    const size_t n0 = out.q.size();
    int bad = 0;
    // the points of one short vector x (q = x^T Gr x)
    auto per_K = [&](const int32_t *x, int64_t val) -> int {
        pfa_i128 z[PFA_DMAX], u[PFA_DMAX], c1[PFA_DMAX];
        pfa_u128 gu = 0;
        for (int a = 0; a < n; ++a) {                           // Lambda coordinates V^T x
            pfa_i128 s = 0;
            for (int b = 0; b < n; ++b)
                if (pfa_mac64(&s, F.V[b * n + a], x[b])) return -60;
            z[a] = s;
            gu = pfa_gcd128(gu, pfa_abs128(s));
        }
        const pfa_i128 g = (pfa_i128)gu;
        pfa_i128 g2;
        if (g <= 0 || pfa_mul_ovf(g, g, &g2) || val % g2 != 0) return -61;
        const pfa_i128 kn = val / g2;
        for (int a = 0; a < n; ++a) {                           // y': K' in Lambda coordinates
            z[a] /= g;
            if (!pfa_fits64(z[a])) return -62;
        }
        pfa_u128 og = (pfa_u128)Hdet;
        for (int i = 0; i < n; ++i) {
            pfa_i128 s = 0;
            for (int c = 0; c <= i; ++c) {
                pfa_i128 t;
                if (pfa_mul_ovf(Hadj[i * n + c], z[c], &t) || pfa_add_ovf(s, t, &s)) return -62;
            }
            u[i] = s;
            og = pfa_gcd128(og, pfa_abs128(s));
        }
        const pfa_i128 o = Hdet / (pfa_i128)og;
        for (int i = 0; i < m; ++i) {                           // c1 = E [u / og; 0]
            pfa_i128 s = 0;
            for (int a = 0; a < n; ++a) {
                pfa_i128 t;
                if (pfa_mul_ovf(E[i * m + a], u[a] / (pfa_i128)og, &t) || pfa_add_ovf(s, t, &s)) return -62;
            }
            c1[i] = s;
        }
        for (int i = 0; i < m; ++i) {                           // K' = O^T y' and W c1 = o K'
            pfa_i128 s = 0, t = 0;
            for (int a = 0; a < n; ++a)
                if (pfa_mac64(&s, F.O[a * m + i], (int64_t)z[a])) return -62;
            for (int c = 0; c < m; ++c) {
                pfa_i128 e;
                if (pfa_mul_ovf((pfa_i128)ZB[(i + 1) * m + c], c1[c], &e) || pfa_add_ovf(t, e, &t)) return -62;
            }
            pfa_i128 oK;
            if (pfa_mul_ovf(o, s, &oK) || oK != t) return -63;      // cannot happen
        }
        pfa_i128 lc1 = 0;
        for (int c = 0; c < m; ++c) {
            pfa_i128 t;
            if (pfa_mul_ovf((pfa_i128)Bi[c], c1[c], &t) || pfa_add_ovf(lc1, t, &lc1)) return -62;
        }
        // j: j o > g D0 and g j o kn < Q L
        pfa_i128 go, gokn, gD0, coef, sdo;
        if (pfa_mul_ovf(g, o, &go) || pfa_mul_ovf(go, kn, &gokn) || pfa_mul_ovf(g, (pfa_i128)D0, &gD0) ||
            pfa_mul_ovf(g, coefL, &coef) || pfa_mul_ovf(F.sd, o, &sdo))
            return -64;
        const pfa_i128 jlo = zpk_fdiv(gD0, o) + 1, jhi = (QL - 1) / gokn;
        if (jlo > jhi) return 0;
        if (jhi - jlo > ((pfa_i128)1 << 26)) return -65;       // too many dilations to walk here
        const double coef_d = (double)coef, sdo_d = (double)sdo;
        for (pfa_i128 j = jlo; j <= jhi; ++j) {
            pfa_i128 diff, R, jl;
            if (pfa_mul_ovf(gokn, j, &diff) || pfa_sub_ovf(QL, diff, &diff)) return -64;
            if (pfa_mul_ovf(j, lc1, &jl)) return -64;
            // the smallest M0 >= M0min with M0 = j lc1 (mod |lk|)
            pfa_i128 off = (jl - (pfa_i128)M0min) % alk;
            if (off < 0) off += alk;
            pfa_i128 M0 = (pfa_i128)M0min + off;
            const double M0d = (double)M0;
            if (M0d * M0d * coef_d > sdo_d * (double)j * (double)diff * (1 + 1e-9) + 1.0) continue;
            if (pfa_mul_ovf(sdo, j, &R) || pfa_mul_ovf(R, diff, &R)) return -64;
            for (;; M0 += alk) {
                pfa_i128 lhs;
                if (pfa_mul_ovf(M0, M0, &lhs) || pfa_mul_ovf(lhs, coef, &lhs) || lhs >= R) break;
                const pfa_i128 tt = (M0 - jl) / lk;             // exact: M0 = j lc1 (mod lk)
                if ((int64_t)(out.q.size() - n0) >= room) return -66;
                int64_t c[PFA_DMAX];
                for (int i = 0; i < m; ++i) {
                    pfa_i128 a, b;
                    if (pfa_mul_ovf(j, c1[i], &a) || pfa_mul_ovf(tt, (pfa_i128)k[i], &b) || pfa_add_ovf(a, b, &a) ||
                        !pfa_fits64(a))
                        return -67;
                    c[i] = (int64_t)a;
                }
                int64_t Mr[PFA_DMAX + 1], Kr[PFA_DMAX + 1];
                for (int i = 0; i < h; ++i) {
                    pfa_i128 a = 0, b = 0;
                    for (int e = 0; e < m; ++e)
                        if (pfa_mac64(&a, Bi[i * m + e], c[e]) || pfa_mac64(&b, ZB[i * m + e], c[e])) return -67;
                    if (!pfa_fits64(a) || !pfa_fits64(b)) return -67;
                    Mr[i] = (int64_t)a;
                    Kr[i] = (int64_t)b;
                }
                if (Mr[0] != (int64_t)M0) return -68;           // cannot happen
                pfa_i128 q = 0;
                for (int i = 0; i < m; ++i) {
                    pfa_i128 s = 0, t;
                    for (int e = 0; e < m; ++e)
                        if (pfa_mac64(&s, mat[i * m + e], c[e])) return -67;
                    if (pfa_mul_ovf(s, (pfa_i128)c[i], &t) || pfa_add_ovf(q, t, &q)) return -67;
                }
                if (!pfa_fits64(q)) return -67;
                out.M.insert(out.M.end(), Mr, Mr + h);
                out.Kn.insert(out.Kn.end(), Kr, Kr + h);
                out.q.push_back((int64_t)q);
                out.pidx.push_back(pidx);
            }
        }
        return 0;
    };
    auto emit = [&](const int32_t *x, int64_t q, uint64_t) -> int {
        if (q <= 0) return 0;                                   // x = 0
        const int rc = per_K(x, q);
        if (rc) { bad = rc; return 1; }
        return 0;
    };
    const int rr = pfa_dfs(P, stk.data(), cc, cnt, emit);
    if (bad) return bad;
    if (rr) return -52;
    return 0;
}

extern "C" int pfa_coni_zpk(int h, const int64_t *kappa, const int64_t *Mbasis, int64_t n_p,
                            const int64_t *ps, int64_t Q, int64_t M0min, int64_t D0, int64_t max_out,
                            int threads, int32_t *status, pfa_zpk_output *out, char *err)
{
    memset(out, 0, sizeof(*out));
    if (h < 3 || h > PFA_MAX_H11 || n_p < 0 || D0 < 1 || M0min < 1 || Q < 1 || max_out < 0) {
        snprintf(err, 256, "pfa_coni_zpk: bad arguments (h11 %d of 3..%d, %lld p-vectors, D0 %lld, "
                 "M0min %lld, Q %lld)", h, PFA_MAX_H11, (long long)n_p, (long long)D0,
                 (long long)M0min, (long long)Q);
        return 1;
    }
    const int nth = threads > 0 ? threads
                                : std::max(1, std::min(64, (int)std::thread::hardware_concurrency()));
    std::vector<ZpkPts> part((size_t)n_p);
    parallel_chunks(n_p, 4, nth, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            ZpkPts &P = part[i];
            const int st = zpk_one(h, kappa, Mbasis, ps + i * h, Q, M0min, D0, i, max_out, P);
            if (st) P = ZpkPts();
            status[i] = st;
        }
    });
    int64_t n = 0;
    for (const ZpkPts &P : part) n += (int64_t)P.q.size();
    out->n = n;
    const size_t rows = (size_t)std::max<int64_t>(n, 1);
    out->M = (int64_t *)malloc(8 * rows * h);
    out->Kn = (int64_t *)malloc(8 * rows * h);
    out->q = (int64_t *)malloc(8 * rows);
    out->pidx = (int64_t *)malloc(8 * rows);
    if (!out->M || !out->Kn || !out->q || !out->pidx) {
        pfa_zpk_output_free(out);
        snprintf(err, 256, "pfa_coni_zpk: out of host memory for %lld points", (long long)n);
        return 1;
    }
    int64_t k = 0;
    for (const ZpkPts &P : part) {
        const int64_t c = (int64_t)P.q.size();
        if (!c) continue;
        memcpy(out->M + k * h, P.M.data(), 8 * (size_t)c * h);
        memcpy(out->Kn + k * h, P.Kn.data(), 8 * (size_t)c * h);
        memcpy(out->q + k, P.q.data(), 8 * (size_t)c);
        memcpy(out->pidx + k, P.pidx.data(), 8 * (size_t)c);
        k += c;
    }
    return 0;
}

extern "C" void pfa_zpk_output_free(pfa_zpk_output *out)
{
    if (!out) return;
    free(out->M);
    free(out->Kn);
    free(out->q);
    free(out->pidx);
    memset(out, 0, sizeof(*out));
}

extern "C" int pfa_reserve(int device, int64_t n_p, int64_t n_points, int h_max, char *err)
{
    char *errbuf = err;
    try {
        if (device < 0 || n_p < 0 || n_points < 0 || h_max < 3 || h_max > PFA_MAX_H11) {
            snprintf(errbuf, 256, "pfa_reserve: bad arguments (n_p %lld, n_points %lld, h_max %d)",
                     (long long)n_p, (long long)n_points, h_max);
            return 1;
        }
        DevCtx *C = ctx_for(device);
        std::lock_guard<std::mutex> lk(C->mu);
        ctx_init(C, device, errbuf);
        const int D = h_max - 1, HS = D + 1;
        const RecL L = rec_layout(D);
        const EntL E = ent_layout(D);
        const size_t np1 = (size_t)std::max<int64_t>(n_p, 1), n1 = (size_t)std::max<int64_t>(n_points, 1);
        dev_get(C, B_RECS, 8 * (size_t)L.words * np1, errbuf);
        dev_get(C, B_PREPCOLD, pfa_ws_cold_bytes(h_max) * std::min<size_t>(np1, 16384), errbuf);
        // a search sizes its point buffer at 9/8 of the largest count seen
        const size_t pc = n1 + n1 / 8 + 1;
        C->pts_reserved = std::max<unsigned long long>(C->pts_reserved, pc);
        dev_get(C, B_PT_PIDX, 4 * pc, errbuf);
        dev_get(C, B_PT_Q, 8 * pc, errbuf);
        dev_get(C, B_PT_G, 8 * pc, errbuf);
        dev_get(C, B_PT_C, 4 * pc * D, errbuf);
        const unsigned long long ring_cap =
            (unsigned long long)std::max<int64_t>(2 * n_p + 64, 1 << 20) + 2 * 64ULL * 4096ULL;
        dev_get(C, B_RING, (size_t)ring_cap * E.bytes, errbuf);
        dev_get(C, B_SEQ, 8 * ring_cap, errbuf);
        // deques and the split's frontier buffers at the default settings,
        // for up to 4 wavefronts per CU
        const int sc = default_stack_cap(D, PFA_SPLIT_DEFAULT);
        dev_get(C, B_STACKS, (size_t)std::max(1, C->cus * 4) * sc * E.bytes, errbuf);
        const unsigned long long room = (unsigned long long)std::max(1, C->cus * 4) * (unsigned long long)(sc / 2);
        const unsigned long long fcap = std::max<unsigned long long>(room, (unsigned long long)np1) + 64;
        dev_get(C, B_F0, (size_t)fcap * E.bytes, errbuf);
        dev_get(C, B_F1, (size_t)fcap * E.bytes, errbuf);
        dev_get(C, B_SPLITST, 256 * 8, errbuf);
        dev_get(C, B_M, 8 * n1 * HS, errbuf);
        dev_get(C, B_KN, 8 * n1 * HS, errbuf);
        dev_get(C, B_NEXP, 8 * n1, errbuf);
        dev_get(C, B_NONSING, n1, errbuf);
        dev_get(C, B_BAD, np1, errbuf);
        dev_get(C, B_KEEPKP, 8 * n1, errbuf);
        dev_get(C, B_KEEPNEXP, 8 * n1, errbuf);
        dev_get(C, B_KEEPMETA, 8 * n1, errbuf);
        dev_get(C, B_KEEPRAW, 24 * n1, errbuf);
        // packed uploads: the p-vectors and a few MB of geometry tables
        const size_t up_in = np1 * (8 * (size_t)h_max + 4) + ((size_t)4 << 20);
        pinned_reserve(C, 0, up_in, errbuf);
        dev_get(C, B_UPIN, up_in, errbuf);
        pinned_reserve(C, 1, (size_t)4 << 20, errbuf);
        dev_get(C, B_UPFILT, (size_t)4 << 20, errbuf);
        // the filter's list of points for the modular determinant (up to 2^20)
        const size_t mc = std::min<size_t>(n1, (size_t)1 << 20);
        dev_get(C, B_MODLIST, 8 * mc, errbuf);
        dev_get(C, B_MODNEXP, 8 * mc, errbuf);
        dev_get(C, B_MODM, 8 * mc * HS, errbuf);
        for (Stage &S : C->stage) {
            std::lock_guard<std::mutex> sl(S.mu);
            host_get(C, S, H_PIDX, 4 * n1, errbuf);
            host_get(C, S, H_Q, 8 * n1, errbuf);
            host_get(C, S, H_M, 8 * n1 * HS, errbuf);
            host_get(C, S, H_KN, 8 * n1 * HS, errbuf);
            host_get(C, S, H_NEXP, 8 * n1, errbuf);
            host_get(C, S, H_NS, n1, errbuf);
        }
        C->pcap_hint = std::max<unsigned long long>(C->pcap_hint, (unsigned long long)n_points);
        return 0;
    } catch (int) {
        return 1;
    }
}

extern "C" int pfa_device_cus(int device)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
    return prop.multiProcessorCount;
}

// A plain copy of 16-byte words: the reference point for HBM counters.
__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int pfa_copy_probe(int device, int64_t bytes, int reps, double *ms_per_copy)
{
    char errbuf[256];
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 1;
    try {
        const int64_t n16 = bytes / 16;
        if (n16 <= 0 || reps <= 0) return 1;
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipMalloc(&a, 16 * n16));
        HIPCHK(hipMalloc(&b, 16 * n16));
        HIPCHK(hipMemset(a, 1, 16 * n16));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const int grid = (int)std::min<int64_t>((n16 + 255) / 256, 8192);
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, (const uint4 *)a, (uint4 *)b, n16);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, (const uint4 *)a, (uint4 *)b, n16);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e1, 0));
        HIPCHK(hipEventSynchronize(e1));
        *ms_per_copy = ms_between(e0, e1) / reps;
        rc = 0;
    } catch (int) {
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

extern "C" int pfa_device_clock_khz(int device)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
    return prop.clockRate;
}

extern "C" const char *pfa_arch(void) { return "gfx950"; }

// ------------------------------------------------------- exact det filter
extern "C" int pfa_det_filter(int n_geo, const int32_t *h, const int64_t *kappa, int64_t count,
                              const int32_t *geo, const int64_t *M, int device, int8_t *out,
                              char *err)
{
    char errbuf[256] = {0};
    try {
        int hmax = 1;
        for (int g = 0; g < n_geo; ++g) {
            if (h[g] < 1 || h[g] > PFA_MAX_H11) {
                snprintf(err, 256, "h11 outside 1..%d", PFA_MAX_H11);
                return 1;
            }
            hmax = std::max(hmax, (int)h[g]);
        }
        for (int64_t k = 0; k < count; ++k)
            if (geo[k] < 0 || geo[k] >= n_geo) {
                snprintf(err, 256, "geometry index out of range");
                return 1;
            }
        pfa_run_input gin;
        memset(&gin, 0, sizeof gin);
        gin.n_geo = n_geo;
        gin.h = h;
        gin.kappa = kappa;
        std::vector<int64_t> kofs, mofs;
        int64_t ktot = 0, mtot = 0;
        fill_geo_offsets(&gin, kofs, mofs, ktot, mtot);
        const int MH = hmax;
        if (device < 0) {
            static std::vector<PfaModP> pr = primes31(96);
            std::vector<int64_t> N((size_t)MH * MH);
            std::vector<uint32_t> w((size_t)MH * MH);
            std::vector<pfa_i128> bw((size_t)MH * MH);
            for (int64_t k = 0; k < count; ++k) {
                const int hh = h[geo[k]];
                if (kappa_M(kappa + kofs[geo[k]], hh, M + k * MH, N.data(), 1)) { out[k] = 2; continue; }
                out[k] = (int8_t)pfa_nonsingular(N.data(), hh - 1, pr.data(), (int)pr.size(), w.data(),
                                                 bw.data(), 1);
            }
            return 0;
        }
        DevCtx *C = ctx_for(device);
        std::lock_guard<std::mutex> lk(C->mu);
        ctx_init(C, device, errbuf);
        hipStream_t s = C->stream;
        auto up = [&](int slot, const void *src, size_t bytes) -> void * {
            void *d = dev_get(C, slot, std::max<size_t>(bytes, 16), errbuf);
            if (bytes) HIPCHK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, s));
            return d;
        };
        const int D = std::max(hmax - 1, 1);
        const RecL L = rec_layout(D);
        std::vector<int64_t> recs((size_t)L.words * std::max(n_geo, 1), 0);
        for (int g = 0; g < n_geo; ++g) {
            int32_t *ri = (int32_t *)(recs.data() + (size_t)g * L.words);
            ri[24] = h[g];
            ri[26] = g;
        }
        GeoArgs G{};
        G.kappa = (const int64_t *)up(B_KAPPA, kappa, 8 * ktot);
        G.kofs = (const int64_t *)up(B_KOFS, kofs.data(), 8 * kofs.size());
        const int32_t *d_geo = (const int32_t *)up(B_PT_PIDX, geo, 4 * count);
        const int64_t *d_M = (const int64_t *)up(B_M, M, 8 * count * MH);
        const int64_t *d_recs = (const int64_t *)up(B_RECS, recs.data(), 8 * recs.size());
        int8_t *d_ns = (int8_t *)dev_get(C, B_NONSING, std::max<int64_t>(count, 1), errbuf);
        int8_t *d_bad = (int8_t *)dev_get(C, B_BAD, std::max(n_geo, 1), errbuf);
        FiltArgs Fa{};
        Fa.n = count;
        Fa.pidx = d_geo;
        Fa.Min = d_M;
        Fa.D = D;
        Fa.MH = MH;
        Fa.hl = MH;
        Fa.recs = d_recs;
        Fa.L = L;
        Fa.filter = 1;
        Fa.nonsing = d_ns;
        Fa.bad = d_bad;
        Fa.G = G;
        const FiltLaunch F = filter_setup(C, s, &gin, kofs, Fa, errbuf);
        run_seg(C, s, [&](Seg &sg) { filter_body(sg, F); }, errbuf);
        if (count > 0) HIPCHK(hipMemcpyAsync(out, d_ns, count, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    } catch (...) {
        snprintf(err, 256, "%s", errbuf);
        return 1;
    }
}

// ------------------------------------------- allocation-free det filter
// pfa_det_filter_noalloc: the same kernels (k_filter, k_det_mod) on
// caller-owned device arrays, with every other buffer carved from one
// file-scope device array (`scratch`).  Nothing is allocated per call: the
// per-geometry tables (record headers, transposed kappa, its magnitude)
// are built by k_na_setup from the caller's kappa, and the geometry sizes
// and the prime table travel as a kernel argument.  The array's device
// address is looked up once per device (hipGetSymbolAddress) and only
// that address is handed to kernels.  Calls that overlap on one device
// share the array and race.
#define PFA_NA_MAX_GEO 16
#define PFA_NA_MAX_POINTS 65536
#define PFA_NA_NPRIMES 96
#define PFA_NA_RECW 32                         // record words: the header only
namespace na {
constexpr size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t RECS = 0;
constexpr size_t KOFS = RECS + up256((size_t)PFA_NA_MAX_GEO * PFA_NA_RECW * 8);
constexpr size_t TOFS = KOFS + up256(PFA_NA_MAX_GEO * 8);
constexpr size_t TMAX = TOFS + up256(PFA_NA_MAX_GEO * 8);
constexpr size_t BAD = TMAX + up256(PFA_NA_MAX_GEO * 8);
constexpr size_t MODCNT = BAD + up256(PFA_NA_MAX_GEO);
constexpr size_t PRIMES = MODCNT + up256(8);
constexpr size_t KT = PRIMES + up256(PFA_NA_NPRIMES * sizeof(PfaModP));
constexpr size_t KT_PER_GEO = (size_t)PFA_MAX_H11 * PFA_DMAX * PFA_DMAX;     // int32
constexpr size_t MODKP = KT + up256(PFA_NA_MAX_GEO * KT_PER_GEO * 4);
constexpr size_t MODNEXP = MODKP + up256((size_t)PFA_NA_MAX_POINTS * 8);
constexpr size_t MODM = MODNEXP + up256((size_t)PFA_NA_MAX_POINTS * 8);
constexpr size_t BYTES = MODM + up256((size_t)PFA_NA_MAX_POINTS * PFA_MAX_H11 * 8);
}  // namespace na

// about 27 MB per device, held until the process ends
static __device__ __align__(256) char scratch[na::BYTES];

struct NaTables {
    int n_geo;
    int32_t h[PFA_NA_MAX_GEO];
    int64_t kofs[PFA_NA_MAX_GEO];             // into the caller's kappa
    int64_t tofs[PFA_NA_MAX_GEO];             // into the scratch kT
    PfaModP primes[PFA_NA_NPRIMES];
};

// Block g < n_geo: geometry g's record header (h, geometry), offsets, and
// kT[(k n + i) n + j] = kappa[i+1][j+1][k] with max |kappa| (infinity from
// 2^31 on: the int32 copy is then not exact).  Block n_geo: the prime
// table and the counters that start at zero.
__global__ __launch_bounds__(256) void k_na_setup(NaTables T, const int64_t *kappa, char *scr)
{
    __shared__ double s_mx[256];
    const int g = blockIdx.x;
    if (g == T.n_geo) {
        PfaModP *pr = (PfaModP *)(scr + na::PRIMES);
        for (int i = threadIdx.x; i < PFA_NA_NPRIMES; i += blockDim.x) pr[i] = T.primes[i];
        if (threadIdx.x == 0) *(unsigned long long *)(scr + na::MODCNT) = 0;
        for (int i = threadIdx.x; i < PFA_NA_MAX_GEO; i += blockDim.x) ((int8_t *)(scr + na::BAD))[i] = 0;
        return;
    }
    const int h = T.h[g], n = h - 1;
    int64_t *rec = (int64_t *)(scr + na::RECS) + (size_t)g * PFA_NA_RECW;
    for (int w = threadIdx.x; w < PFA_NA_RECW; w += blockDim.x) rec[w] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        ((int32_t *)rec)[24] = h;
        ((int32_t *)rec)[26] = g;
        ((int64_t *)(scr + na::KOFS))[g] = T.kofs[g];
        ((int64_t *)(scr + na::TOFS))[g] = T.tofs[g];
    }
    const int64_t *kap = kappa + T.kofs[g];
    int32_t *kt = (int32_t *)(scr + na::KT) + T.tofs[g];
    double mx = 0.0;
    for (int e = threadIdx.x; e < h * n * n; e += blockDim.x) {
        const int k = e / (n * n), r = e - k * n * n, i = r / n, j = r - i * n;
        const int64_t v = kap[((int64_t)(i + 1) * h + (j + 1)) * h + k];
        kt[e] = (int32_t)v;
        mx = fmax(mx, fabs((double)v));
    }
    s_mx[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s_mx[threadIdx.x] = fmax(s_mx[threadIdx.x], s_mx[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        ((double *)(scr + na::TMAX))[g] = s_mx[0] < 2147483648.0 ? s_mx[0] : INFINITY;
}

static std::mutex g_na_mu;
static std::vector<char *> g_na_ptr;         // per device: the scratch array's address

// the device address of `scratch` on `device` (looked up once; call with
// the device current)
static char *na_scratch(int device, char *errbuf)
{
    std::lock_guard<std::mutex> lk(g_na_mu);
    if ((int)g_na_ptr.size() <= device) g_na_ptr.resize(device + 1, nullptr);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_na_ptr[device] = (char *)p;
    }
    return g_na_ptr[device];
}

extern "C" int pfa_det_filter_noalloc(int device, int n_geo, const int32_t *h,
                                      const int64_t *kappa, int64_t count, const int32_t *geo,
                                      const int64_t *M, int hstride, int8_t *out, void *stream,
                                      char *err)
{
    char errbuf[256] = {0};
    try {
        if (n_geo < 1 || n_geo > PFA_NA_MAX_GEO) {
            snprintf(err, 256, "pfa_det_filter_noalloc: n_geo = %d outside 1..%d", n_geo,
                     PFA_NA_MAX_GEO);
            return 1;
        }
        if (count < 0 || count > PFA_NA_MAX_POINTS) {
            snprintf(err, 256, "pfa_det_filter_noalloc: count = %lld outside 0..%d",
                     (long long)count, PFA_NA_MAX_POINTS);
            return 1;
        }
        NaTables T;
        memset(&T, 0, sizeof T);
        T.n_geo = n_geo;
        int hmax = 2;
        int64_t ko = 0, to = 0;
        for (int g = 0; g < n_geo; ++g) {
            if (h[g] < 2 || h[g] > PFA_MAX_H11) {
                snprintf(err, 256, "pfa_det_filter_noalloc: h11 = %d outside 2..%d", h[g], PFA_MAX_H11);
                return 1;
            }
            T.h[g] = h[g];
            T.kofs[g] = ko;
            T.tofs[g] = to;
            ko += (int64_t)h[g] * h[g] * h[g];
            to += (int64_t)h[g] * (h[g] - 1) * (h[g] - 1);
            hmax = std::max(hmax, (int)h[g]);
        }
        if (hstride < hmax || hstride > PFA_MAX_H11) {
            snprintf(err, 256, "pfa_det_filter_noalloc: row stride %d outside %d..%d", hstride, hmax,
                     PFA_MAX_H11);
            return 1;
        }
        static const std::vector<PfaModP> pr = primes31(PFA_NA_NPRIMES);
        for (int i = 0; i < PFA_NA_NPRIMES; ++i) T.primes[i] = pr[i];
        HIPCHK(hipSetDevice(device));
        char *scr = na_scratch(device, errbuf);
        hipStream_t s = (hipStream_t)stream;
        hipLaunchKernelGGL(k_na_setup, dim3(n_geo + 1), dim3(256), 0, s, T, kappa, scr);
        HIPCHK(hipGetLastError());
        if (count == 0) return 0;
        const int MH = hstride, D = std::max(hmax - 1, 1), nl = MH - 1;
        RecL L{};
        L.D = D;
        L.words = PFA_NA_RECW;
        FiltArgs a{};
        a.n = count;
        a.pidx = geo;
        a.Min = M;
        a.D = D;
        a.MH = MH;
        a.hl = MH;
        a.recs = (const int64_t *)(scr + na::RECS);
        a.L = L;
        a.filter = 1;
        a.nonsing = out;
        a.bad = (int8_t *)(scr + na::BAD);
        a.G.kappa = kappa;
        a.G.kofs = (const int64_t *)(scr + na::KOFS);
        a.kT = (const int32_t *)(scr + na::KT);
        a.kTofs = (const int64_t *)(scr + na::TOFS);
        a.kTmax = (const double *)(scr + na::TMAX);
        a.n_geo = n_geo;
        a.primes = (const PfaModP *)(scr + na::PRIMES);
        a.np = PFA_NA_NPRIMES;
        a.modcap = count;
        a.modkp = (int64_t *)(scr + na::MODKP);
        a.modnexp = (int64_t *)(scr + na::MODNEXP);
        a.modM = (int64_t *)(scr + na::MODM);
        a.modcount = (unsigned long long *)(scr + na::MODCNT);
        const int nmax = nl <= 8 ? 8 : nl <= 12 ? 12 : nl <= 16 ? 16 : nl <= 20 ? 20 : nl <= 24 ? 24 :
                         nl <= 32 ? 32 : 40;
        int lds_floats = (nmax + 1) * nmax * nmax;
        if (lds_floats > 64 * 1024 / 4) lds_floats = 0;     // kappa read from the global table
        a.lds_floats = lds_floats;
        const int R = 64 / a.hl;
        const int grid = (int)std::min<int64_t>((count + 4 * R - 1) / (4 * R), 16384);
        const size_t lds = (size_t)((lds_floats + 3) & ~3) * 4 + (size_t)4 * R * nmax * 20;
#define PFA_NA_LAUNCH(NM)                                                                      \
    do {                                                                                        \
        HIPCHK(hipFuncSetAttribute((const void *)k_filter<NM, false>,                          \
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));      \
        hipLaunchKernelGGL((k_filter<NM, false>), dim3(grid), dim3(256), lds, s, a);           \
        HIPCHK(hipGetLastError());                                                              \
        hipLaunchKernelGGL(k_det_mod<NM>, dim3(512), dim3(256), 0, s, a);                       \
        HIPCHK(hipGetLastError());                                                              \
    } while (0)
        if (nmax == 8) PFA_NA_LAUNCH(8);
        else if (nmax == 12) PFA_NA_LAUNCH(12);
        else if (nmax == 16) PFA_NA_LAUNCH(16);
        else if (nmax == 20) PFA_NA_LAUNCH(20);
        else if (nmax == 24) PFA_NA_LAUNCH(24);
        else if (nmax == 32) PFA_NA_LAUNCH(32);
        else PFA_NA_LAUNCH(40);
#undef PFA_NA_LAUNCH
        return 0;
    } catch (...) {
        snprintf(err, 256, "%s", errbuf);
        return 1;
    }
}

extern "C" int pfa_kernel_batch(const pfa_kernel_input *in, pfa_kernel_output *out)
{
    memset(out, 0, sizeof(*out));
    char *errbuf = out->err;
    try {
        int D = 1;
        for (int k = 0; k < in->n_prob; ++k) {
            if (in->dims[k] < 1 || in->dims[k] > PFA_DMAX) {
                snprintf(errbuf, 256, "problem dimension outside 1..%d", PFA_DMAX);
                return 1;
            }
            if (in->nl[k] > in->dims[k] + 1) {
                snprintf(errbuf, 256, "too many leaf rows");
                return 1;
            }
            D = std::max(D, (int)in->dims[k]);
        }
        const RecL L = rec_layout(D);
        const int64_t n = in->n_prob;
        std::vector<int64_t> recs((size_t)L.words * std::max<int64_t>(n, 1), 0);
        int64_t om = 0, oh = 0, ol = 0, olin = 0;
        for (int64_t k = 0; k < n; ++k) {
            int64_t *r = recs.data() + k * L.words;
            int32_t *ri = (int32_t *)r;
            const int d = in->dims[k];
            ri[0] = PFA_HOST;
            ri[1] = d;
            r[1] = in->qmax[k];
            r[2] = in->Q[k];
            r[3] = in->has_lin[k] ? in->linmin[k] : INT64_MIN;
            ri[23] = in->strict[k];
            ri[24] = d + 1;
            for (int e = 0; e < d * d; ++e) r[L.mat + e] = in->mat[om + e];
            om += (int64_t)d * d;
            int m0 = -1;
            for (int j = 0; j < d; ++j) {
                r[L.lin + j] = in->has_lin[k] ? in->linvec[olin + j] : 0;
                if (m0 < 0 && r[L.lin + j] != 0) m0 = j;
            }
            if (in->has_lin[k]) olin += d;
            ri[20] = m0;
            uint64_t hm = 0;
            for (int rr = 0; rr < in->nh[k]; ++rr) {
                const int64_t *hr = in->H + oh + (int64_t)rr * d;
                int lv = -1;
                for (int j = 0; j < d; ++j) if (hr[j]) { lv = j; break; }
                if (lv < 0) continue;
                if ((hm >> lv) & 1) { snprintf(errbuf, 256, "H is not in echelon form"); return 1; }
                hm |= 1ULL << lv;
                for (int j = 0; j < d; ++j) r[L.H + (int64_t)lv * d + j] = hr[j];
            }
            oh += (int64_t)in->nh[k] * d;
            r[9] = (int64_t)hm;
            int64_t mm = 0, hm2 = 0;
            for (int e = 0; e < d * d; ++e) {
                const int64_t x = r[L.mat + e] < 0 ? -r[L.mat + e] : r[L.mat + e];
                const int64_t y = r[L.H + e] < 0 ? -r[L.H + e] : r[L.H + e];
                mm = x > mm ? x : mm;
                hm2 = y > hm2 ? y : hm2;
            }
            r[16] = mm;
            r[17] = hm2;
            ri[21] = in->nl[k] ? 2 : 0;
            ri[22] = in->nl[k];
            for (int e = 0; e < in->nl[k] * d * 2; ++e) r[L.L + e] = in->L[ol + e];
            ol += (int64_t)in->nl[k] * d * 2;
        }
        // factor and search
        std::vector<int32_t> pidx;
        std::vector<int64_t> qs;
        std::vector<int32_t> cs;
        std::vector<int64_t> nodes(n, 0);
        if (in->mode == PFA_MODE_HOST) {
            std::vector<char> ws((size_t)2 * D * D * 8 + 64);
            PfaTeam T{0, 1};
            std::vector<PfaNode> st(PFA_DMAX);
            int32_t c[PFA_DMAX];
            for (int64_t k = 0; k < n; ++k) {
                int64_t *r = recs.data() + k * L.words;
                factor_one(T, ws.data(), r, L);
                if (rec_status(r) != PFA_OK) continue;
                PfaProb P = rec_view(r, L);
                PfaCounts cnt{0, 0, 0};
                size_t mark = qs.size();
                auto emit = [&](const int32_t *cc, int64_t q, uint64_t) -> int {
                    pidx.push_back((int32_t)k);
                    qs.push_back(q);
                    for (int j = 0; j < D; ++j) cs.push_back(j < P.d ? cc[j] : 0);
                    return 0;
                };
                int rr = pfa_dfs(P, st.data(), c, cnt, emit);
                nodes[k] = (int64_t)cnt.nodes;
                if (rr) {
                    rec_status(r) = rr;
                    pidx.resize(mark); qs.resize(mark); cs.resize(mark * D);
                }
            }
        } else {
            DevCtx *C = ctx_for(in->device);
            std::lock_guard<std::mutex> lk(C->mu);
            ctx_init(C, in->device, errbuf);
            hipStream_t s = C->stream;
            int64_t *d_recs = (int64_t *)dev_get(C, B_RECS, 8 * recs.size(), errbuf);
            HIPCHK(hipMemcpyAsync(d_recs, recs.data(), 8 * recs.size(), hipMemcpyHostToDevice, s));
            if (n > 0)
                hipLaunchKernelGGL(k_factor, dim3((unsigned)std::min<int64_t>(n, 65536)), dim3(64),
                                   (size_t)2 * D * D * 8 + 64, s, n, d_recs, L);
            unsigned long long pcap = std::max<int64_t>(1 << 14, 64 * n);
            // the same search forms as the pipeline (lane, plain, batched),
            // with its overflow reruns
            SearchRun R;
            SearchTune tn{in->mode, 0, 0, 0, 0};
            search_device(C, s, d_recs, L, n, D, tn, pcap, R, errbuf);
            const unsigned long long npts = R.npts;
            HIPCHK(hipMemcpyAsync(recs.data(), d_recs, 8 * recs.size(), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            int32_t *d_pidx = R.d_pidx;
            int64_t *d_q = R.d_q;
            int32_t *d_c = R.d_c;
            pidx.resize(npts); qs.resize(npts); cs.resize(npts * D);
            if (npts) {
                HIPCHK(hipMemcpy(pidx.data(), d_pidx, 4 * npts, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(qs.data(), d_q, 8 * npts, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(cs.data(), d_c, 4 * npts * D, hipMemcpyDeviceToHost));
            }
        }
        // the chunked forms leave unused point slots (pidx -1): drop them
        {
            int64_t w = 0;
            for (int64_t t = 0; t < (int64_t)qs.size(); ++t) {
                if (pidx[t] < 0) continue;
                pidx[w] = pidx[t];
                qs[w] = qs[t];
                memmove(&cs[w * D], &cs[t * D], 4 * D);
                w++;
            }
            pidx.resize(w); qs.resize(w); cs.resize(w * D);
        }
        // per-problem depth-first order: lexicographic in (c[d-1], ..., c[0])
        const int64_t np = (int64_t)qs.size();
        std::vector<int64_t> idx(np);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
            if (pidx[a] != pidx[b]) return pidx[a] < pidx[b];
            for (int j = D - 1; j >= 0; --j)
                if (cs[a * D + j] != cs[b * D + j]) return cs[a * D + j] < cs[b * D + j];
            return false;
        });
        out->dmax = D;
        out->status = (int32_t *)malloc(4 * std::max<int64_t>(n, 1));
        out->nodes = (int64_t *)malloc(8 * std::max<int64_t>(n, 1));
        for (int64_t k = 0; k < n; ++k) {
            out->status[k] = ((int32_t *)(recs.data() + k * L.words))[0];
            out->nodes[k] = nodes[k];
        }
        int64_t keep = 0;
        for (int64_t t = 0; t < np; ++t) keep += out->status[pidx[idx[t]]] == PFA_OK;
        out->n = keep;
        out->pts = (int32_t *)malloc(4 * std::max<int64_t>(keep, 1) * D);
        out->qs = (int64_t *)malloc(8 * std::max<int64_t>(keep, 1));
        out->pidx = (int64_t *)malloc(8 * std::max<int64_t>(keep, 1));
        int64_t o = 0;
        for (int64_t t = 0; t < np; ++t) {
            int64_t k = idx[t];
            if (out->status[pidx[k]] != PFA_OK) continue;
            memcpy(out->pts + o * D, &cs[k * D], 4 * D);
            out->qs[o] = qs[k];
            out->pidx[o] = pidx[k];
            o++;
        }
        return 0;
    } catch (int) {
        return 1;
    }
}

extern "C" void pfa_kernel_output_free(pfa_kernel_output *out)
{
    free(out->pts); free(out->qs); free(out->pidx); free(out->status); free(out->nodes);
    out->pts = nullptr; out->qs = nullptr; out->pidx = nullptr; out->status = nullptr;
    out->nodes = nullptr;
    out->n = 0;
}
