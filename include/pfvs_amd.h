/*
 * pfvs_amd.h -- C ABI of libpfvs_amd.so, the MI355X (gfx950) coniZpM search.
 *
 * Plain C types only (pointers and sizes); no torch or HIP types.  All
 * arrays are host memory owned by the caller, except the output arrays,
 * which the library allocates with malloc and the caller releases with the
 * matching *_free call.
 *
 * Part 1 replaces the reference's GPU library interface one for one (the
 * ctypes binding in pfvs/gpu.py:37-82 loads these names):
 *   pfg_coni_batch   <- pfvs/fp_kernel/cuda/pfvs_gpu.cu:64-97, 489-751
 *   pfg_output_free  <- pfvs_gpu.cu:64-97
 *   pfg_max_h11      <- pfvs_gpu.cu:46-48, 519-522
 *   pfg_backend      <- pfvs_gpu.cu:64-97 (returns "hip")
 *   pfg_device_count <- pfvs_gpu.cu:64-97
 *   pfg_release      <- pfvs_gpu.cu:753-758
 * Part 2 adds entry points of this library: the run call with a choice of
 * search form and counters, the single-problem kernel batch, and the
 * serving session.
 */
#ifndef PFVS_AMD_H
#define PFVS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ part 1 */

/* Same field order and meaning as the reference's pfg_input (gpu.py:37-43):
 * per geometry h11, Q, qmax = floor(dilation Q + 1e-7), linmin =
 * ceil(M0min - 1e-9), kappa (h^3, concatenated) and Mbasis (h^2, columns
 * are basis vectors); p-vectors padded to pfg_max_h11() entries with the
 * leading coni 0; pgeo the geometry of each p-vector. */
typedef struct {
    int n_geo;
    const int32_t *h;
    const int64_t *Q, *qmax, *linmin;
    const int64_t *kappa;
    const int64_t *Mbasis;
    int64_t n_p;
    const int64_t *ps;
    const int32_t *pgeo;
    int device;
    int64_t batch;
    int verbose;
} pfg_input;

/* Points grouped by p-vector (ascending), sorted by (M, Kn) within one;
 * rows padded to pfg_max_h11() with zeros; pstat[i] 0 done, 1 redo on the
 * exact host path (none of its points are returned). */
typedef struct {
    int64_t n;
    int64_t *M, *Kn;
    int64_t *q, *pidx;
    int8_t *pstat;
    double seconds_gpu;
    char err[256];
} pfg_output;

/* 0 on success; nonzero with a message in out->err. */
int pfg_coni_batch(const pfg_input *in, pfg_output *out);
void pfg_output_free(pfg_output *out);
int pfg_max_h11(void);
const char *pfg_backend(void);
int pfg_device_count(void);
void pfg_release(int device);

/* ------------------------------------------------------------ part 2 */

enum { PFA_MODE_PLAIN = 0, PFA_MODE_BATCHED = 1, PFA_MODE_HOST = 2, PFA_MODE_LANE = 3 };

/* Geometry/p-vector layout as pfg_input.  mode: PFA_MODE_PLAIN (one
 * launch per p-vector batch, one wavefront per p-vector: wave-cooperative
 * candidate expansion of that p-vector's tree, no work sharing),
 * PFA_MODE_BATCHED (persistent wavefronts, LIFO frontier deques, a global
 * frontier ring with work stealing across p-vectors), PFA_MODE_LANE (one
 * thread per p-vector, the reference kernel's shape), PFA_MODE_HOST (the
 * same search on host threads).  Tunables <= 0 take defaults.  filter != 0
 * also evaluates the exact per-point filter (K-integrality count and
 * det != 0).  keep_points: 0 counters only; 1 every lattice point; 2 (with
 * filter) only the points that can yield PFVs (n_exp > 0, det != 0 or
 * undecided), compacted on the device -- what coniZpM's post-processing
 * reads. */
typedef struct {
    int n_geo;
    const int32_t *h;
    const int64_t *Q, *qmax, *linmin;
    const int64_t *kappa;
    const int64_t *Mbasis;
    int64_t n_p;
    const int64_t *ps;
    const int32_t *pgeo;
    int device;
    int mode;
    int waves_per_cu;     /* batched: persistent wavefronts per CU (default 4) */
    int budget;           /* batched: most nodes one wavefront donates at once (default 8) */
    int stack_cap;        /* batched: entries per wavefront deque (default max(64 (d + 2),
                             2 split + 128)) */
    int host_threads;     /* host mode */
    int filter;
    int sort;             /* 1: group by p and sort by (M, Kn) */
    int keep_points;      /* 0 counters only, 1 all points, 2 the filter's survivors */
    int verbose;
    int assemble;         /* with filter and sort: also the finished PFVs (K, M) of every
                             p-vector the library can finish (see pfa_run_output.pdone) */
    int split;            /* batched: nodes per wavefront the breadth-first split expands the
                             roots to before the persistent search (default 2048; < 0 none) */
} pfa_run_input;

/* stats[]: see PFA_ST_* */
enum {
    PFA_ST_NODES = 0, PFA_ST_CAND, PFA_ST_LEAF, PFA_ST_POINTS, PFA_ST_PASSES,
    PFA_ST_MS_PREP, PFA_ST_MS_ROOTS, PFA_ST_MS_SEARCH, PFA_ST_MS_POST, PFA_ST_MS_TOTAL,
    PFA_ST_SEARCH_LAUNCHES, PFA_ST_HOST_P, PFA_ST_QUEUE_PEAK, PFA_ST_MS_WALL,
    PFA_ST_KEPT, PFA_ST_EXPANSIONS, PFA_ST_NONSING, PFA_ST_UNDECIDED,
    PFA_ST_STEPS, PFA_ST_WAIT_CYC, PFA_ST_WAVE_CYC, PFA_ST_DONATED, PFA_ST_FETCHES,
    /* host mode only: terms of the algorithmic work (lane-ops) of the search */
    PFA_ST_SUM_K, PFA_ST_SUM_K_M0, PFA_ST_SUM_K_R, PFA_ST_SUM_R, PFA_ST_SUM_S, PFA_ST_SUM_LEAF,
    PFA_ST_WORK,
    PFA_ST_MS_FINISH,     /* host ms of the final assembly: statuses, canonical sort, copy-out */
    PFA_ST_MS_ALLOC,      /* device: host ms spent growing device and pinned buffers */
    PFA_ST_MS_D2H,        /* device: host ms of copying the results back */
    PFA_ST_ATTEMPTS,      /* device: searches run (more than 1 after a point-buffer overflow) */
    PFA_ST_MS_UPLOAD,     /* device: host ms of the uploads (pageable inputs) */
    PFA_ST_WAVES,         /* device: wavefronts of the search launch */
    PFA_ST_SURVIVORS,     /* device, keep_points 2: points copied back */
    PFA_ST_GRAPHS,        /* device: instantiated segment graphs held (0: graphs off) */
    PFA_ST_GRAPH_HITS,    /* device: segment launches replayed from a held graph (cumulative) */
    PFA_ST_GRAPH_CAPTURES,/* device: segments captured and instantiated (cumulative) */
    PFA_ST_PREP_WORK,     /* host mode: algorithmic lane-ops of the preparation, all p-vectors */
    PFA_ST_PREP_WORK_MAX, /* host mode: the same for the p-vector with the most */
    PFA_ST_N
};

typedef struct {
    int64_t n;            /* points */
    int64_t *M, *Kn;      /* n x pfg_max_h11() */
    int64_t *q, *pidx;
    int8_t *pstat;        /* n_p: 0 done, 1 exact host path, < 0 kernel status */
    int32_t *pcode;       /* n_p: detailed code (preparation stage codes; 1: the search
                             tree was proven empty before the lattice, one node) */
    int64_t *nexp;        /* n: (g, K0) pairs of the point (filter), -1 unknown */
    int8_t *nonsing;      /* n: 1 det != 0, 0 det == 0, 2 undecided */
    double stats[PFA_ST_N];
    char err[256];
    /* assemble: the PFVs of coniZpM's post-processing (pfvs/coni.py:313-412)
     * for every p-vector with pdone[i] = 1 -- rows (n_pfv x pfg_max_h11(),
     * zero-padded) in the reference's order: by p-vector, then g, then the
     * points in (M, Kn) order, then K0.  pdone[i] = 0: the caller finishes
     * p-vector i (a point left undecided, pstat != 0, or a value beyond
     * int64). */
    int64_t n_pfv;
    int64_t *Ks, *Ms, *pfv_pidx;
    int8_t *pdone;        /* n_p */
    int64_t *pcount;      /* n_p: lattice points of each p-vector (every keep_points mode) */
} pfa_run_output;

int pfa_coni_run(const pfa_run_input *in, pfa_run_output *out);
void pfa_run_output_free(pfa_run_output *out);

/* coniZpM's post-processing of given lattice points (pfvs/coni.py:275-412,
 * restated natively; the same code finishes pfa_coni_run's assemble) for
 * one geometry (h11 = h, tadpole bound Q): points with rows M and Kn (h
 * int64 each), q = c^T mat c, key = the p-vector index (0 <= key <
 * n_keys) and the verdict of det((kappa M)[1:, 1:]) (1 nonsingular, 0
 * singular, 2 undecided), in any order.  The library orders them by
 * (key, M, Kn), expands each point into its (g, K0) rows (g = 1 ..
 * (Q Kgcd - 1) / q, K0 on the tadpole with K' > 0) and keeps the rows of
 * nonsingular points: by key, then g, then point, then K0.  Rows (n_rows x
 * h) and their keys go to malloc'd arrays; done[k] = 0 leaves key k to the
 * caller (an undecided point, or a value beyond int64).  threads <= 0:
 * all cores.  Returns 0, or 1 with a message in out->err (a tadpole
 * violation, bad sizes).  Replaces the numpy post-processing of
 * pfvs/coni.py:275-412 in this library's coniZpM. */
typedef struct {
    int64_t n_rows;
    int64_t *Ks, *Ms, *keys;
    int8_t *done;         /* n_keys */
    char err[256];
} pfa_assemble_output;

int pfa_assemble(int h, int64_t Q, int64_t n_keys, int64_t n, const int64_t *key,
                 const int64_t *M, const int64_t *Kn, const int64_t *q, const int8_t *nonsing,
                 int threads, pfa_assemble_output *out);
void pfa_assemble_output_free(pfa_assemble_output *out);

/* coniZp's dilation bound (pfvs/dilation.py:152-231; the reference's C
 * form pfvs/fp_kernel/coni_bound.h:255-399, bound through
 * fp_kernel.pyx's _coni_mu0_batch): for each direction p_r (row i of ps,
 * h - 1 entries; p_hat = (0, p_r)), A = kappa . p_hat, S = -A_rr^-1 and
 * Lambda = {K : p_r . K = 0}; mu0 = min over nonzero K in Lambda of
 * K^T S K.  status[i]: 0 mu0 = num[i] / den[i] (reduced) and K (row i,
 * h - 1 entries) a minimizer; 1 no bound (A_rr singular, s > 0, or S not
 * positive definite on Lambda); < 0 not decided here (a value beyond
 * int64/int128, a form too ill-conditioned for the FP64 search): the
 * caller's exact path decides.  Host threads (threads <= 0: all cores).
 * Returns 0, or 1 with a message for bad sizes. */
int pfa_coni_mu0(int h, const int64_t *kappa, int64_t n_p, const int64_t *ps, int threads,
                 int32_t *status, int64_t *num, int64_t *den, int64_t *K, char *err);

/* coniZpK's lattice points (pfvs/coni.py:554-652, whose C kernel is
 * pfvs/fp_kernel/coni_zpk.h, bound through fp_kernel.pyx's _coni_zpk_batch):
 * for each direction (row i of ps, h entries, ps[i][0] = 0) the lattice
 * points c of the coni PFVs with dilation delta > D0 -- c = j c1 + t k for
 * every short K_r of Lambda (K_r^T S K_r < Q / D0) with j o > g D0 and
 * g L sn M0^2 < sd j o (Q L - g j o kn), M0 >= M0min -- as rows M =
 * Binter c, Kn = ZB c (h entries each) with q = c^T mat c and pidx = i, in
 * direction order; a point can repeat and some of its PFVs can have
 * delta <= D0 (the caller filters and deduplicates after the PFV
 * assembly).  status[i]: 0 done; 1 no short-vector bound (s >= 0, A_rr
 * singular, S not positive definite on Lambda); < 0 not decided here (a
 * value beyond int128, a form too ill-conditioned for the FP64 search,
 * more than max_out points for the direction): the caller searches that
 * direction to its dilation bound instead, and none of its points are
 * returned.  Host threads (threads <= 0: all cores).  The library
 * mallocs the arrays; pfa_zpk_output_free frees them.  Returns 0, or 1
 * with a message for bad arguments. */
typedef struct {
    int64_t n;                /* points */
    int64_t *M, *Kn;          /* n x h */
    int64_t *q, *pidx;        /* n */
} pfa_zpk_output;

int pfa_coni_zpk(int h, const int64_t *kappa, const int64_t *Mbasis, int64_t n_p, const int64_t *ps,
                 int64_t Q, int64_t M0min, int64_t D0, int64_t max_out, int threads, int32_t *status,
                 pfa_zpk_output *out, char *err);
void pfa_zpk_output_free(pfa_zpk_output *out);

/* Single-problem kernel batch (the contract of conipfv_kernel,
 * pfvs/fp_kernel/fp_kernel.pyx:258-318 / fp_kernel.h:25-63): for problem
 * k of dimension dims[k], mat (d*d), H (nh[k] rows of d, echelon form,
 * int64), leaf rows L (nl[k] rows, int64; used when some H rows were
 * dropped for not fitting int64), linvec (d, or has_lin[k] = 0), Q, qmax,
 * linmin, strict.  Arrays of all problems are concatenated.  Outputs the
 * points in depth-first order per problem (pts int32 (n, dmax), qs, pidx),
 * per-problem status, and per-problem node counts (host mode; the device
 * kernels count nodes only in total, and report 0 here). */
typedef struct {
    int n_prob;
    const int32_t *dims, *nh, *nl, *has_lin, *strict;
    const int64_t *mat, *H, *L, *linvec, *Q, *qmax, *linmin;
    int device;
    int mode;
} pfa_kernel_input;

typedef struct {
    int64_t n;
    int32_t *pts;         /* n x dmax */
    int64_t *qs, *pidx;
    int32_t *status;      /* n_prob */
    int64_t *nodes;       /* n_prob */
    int dmax;
    char err[256];
} pfa_kernel_output;

int pfa_kernel_batch(const pfa_kernel_input *in, pfa_kernel_output *out);
void pfa_kernel_output_free(pfa_kernel_output *out);

/* The exact filter det((kappa M)[1:, 1:]) != 0 of coniZpM's
 * post-processing (pfvs/coni.py:397-407) on its own, for `count` vectors M
 * (row stride max(h)) of geometries geo[k] (kappa of geometry g: h[g]^3
 * int64, concatenated): out[k] = 1 nonsingular, 0 singular, 2 undecided
 * (an entry of kappa M outside int64).  device >= 0 runs the HIP kernel,
 * device < 0 the host code.  Returns 0, or 1 with a message in err (256
 * bytes).  No counterpart in the reference, which evaluates the filter
 * with flint inside coniZpM; exported for testing the filter alone. */
int pfa_det_filter(int n_geo, const int32_t *h, const int64_t *kappa, int64_t count,
                   const int32_t *geo, const int64_t *M, int device, int8_t *out, char *err);

/* pfa_det_filter on device-resident arrays, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  kappa (sum of h[g]^3 int64), geo
 * (count int32, each in 0..n_geo-1: not checked, the array is on the
 * device), M (count rows of stride hstride >= max h, int64) and out (count
 * int8: 1 nonsingular, 0 singular, 2 undecided) are the caller's device
 * memory; h (n_geo int32) is host memory.  Work is queued on `stream` (a
 * hipStream_t of the HIP runtime the library is linked against, or null);
 * the call returns without waiting.  Every other buffer lives in one
 * fixed device array of about 27 MB per device, held until the process
 * ends; calls that overlap on one device share it and race (serialize
 * them, e.g. on one stream).  Limits: n_geo <= 16, 2 <= h <= 41,
 * count <= 65536.  Returns 0, or 1 with a message in err (256 bytes) for
 * sizes beyond the limits. */
int pfa_det_filter_noalloc(int device, int n_geo, const int32_t *h, const int64_t *kappa,
                           int64_t count, const int32_t *geo, const int64_t *M, int hstride,
                           int8_t *out, void *stream, char *err);

/* Copies `bytes` (rounded down to 16) between two fresh device buffers
 * with a plain copy kernel, `reps` timed times after one untimed;
 * *ms_per_copy gets the mean.  The reference point for HBM traffic
 * counters.  No counterpart in the reference.  Returns 0 or 1. */
int pfa_copy_probe(int device, int64_t bytes, int reps, double *ms_per_copy);

/* All integer p with |p_j| <= B (n coordinates) and H p >= rhs row-wise
 * (H: m x n int32, row-major), primitive ones only when `primitive`, in
 * lexicographic order (p_0 outermost, ascending).  Replaces the
 * enumeration the reference takes from latticepts (box_enum(B, H, rhs, cap,
 * primitive) at pfvs/distributed.py:121-139 and
 * tests/test_paper_examples.py:119-121; enum_lattice_points at
 * pfvs/pvectors.py:58-89).  Runs on host threads (device < 0); device >= 0
 * returns 2 (the device form is not enabled in this build).  The first
 * min(total, cap) points go to out (cap x n
 * int64); *n_out gets that number, *n_total the count of all, *ms (if not
 * null) the time.  Returns 0; 1 when total > cap (out holds the first cap
 * points); 2 on an error (message in err, 256 bytes).  m <= 256, n <= 64. */
int pfa_box_enum(const int32_t *H, int m, int n, const int64_t *rhs, int B, int primitive,
                 int device, int64_t cap, int64_t *out, int64_t *n_out, int64_t *n_total,
                 double *ms, char *err);

/* Grows the per-device buffers (device memory and both pinned result
 * staging sets) to hold a batch of n_p p-vectors and n_points points at
 * h11 <= h_max, so that later calls of that size allocate nothing; also
 * sizes the first point buffer of later searches.  For long-running
 * callers (a serving loop) to pay allocation once, at start.  No
 * counterpart in the reference.  Returns 0, or 1 with a message in err
 * (256 bytes). */
int pfa_reserve(int device, int64_t n_p, int64_t n_points, int h_max, char *err);

/* Library facts: number of compute units and peak engine clock (kHz) of
 * `device`, and the build's architecture string. */
int pfa_device_cus(int device);
int pfa_device_clock_khz(int device);
const char *pfa_arch(void);

#ifdef __cplusplus
}
#endif
#endif
