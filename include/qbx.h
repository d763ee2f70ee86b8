/*
 * qbx.h -- C ABI of libqbx.so, the gfx950 (MI355X) exhaustive QUBO solver.
 *
 * Every entry point returns 0 on success and a negative QBX_E* code on failure; the
 * message of the last failure on the calling thread is returned by qbx_last_error().
 * No exception crosses this boundary.  Input pointers are host memory owned by the
 * caller and read only during the call.  Output structs are caller-allocated.  Device
 * memory belongs to the context.  One context is used by one host thread at a time.
 *
 * Conventions shared with the reference (src/qubo_lib/solvers/qbf.py):
 *   - Q is n x n, row-major, any (possibly non-symmetric) matrix; E(x) = x^T Q x
 *     (qbf.py:264, qbf.py:499).
 *   - The state integer has x[0] as its most significant bit (qbf.py:373-393).
 *   - Ties are broken exactly as the reference QBF breaks them with suffix width
 *     B = tie_B: the answer is the minimum of (E, rank) with
 *     rank(s) = (invgray(s >> B) << B) | (s & (2^B - 1))  (qbf.py:212-230, 395-412),
 *     and tie_B = -1 selects the reference CPU default B = min(n, 20) (qbf.py:443-446).
 */
#ifndef QBX_H
#define QBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QBX_OK 0
#define QBX_EINVAL (-1)     /* bad argument (ValueError on the Python side) */
#define QBX_ERANGE (-2)     /* energies do not fit the selected arithmetic */
#define QBX_EDEVICE (-3)    /* HIP runtime error */
#define QBX_ENOMEM (-4)     /* allocation failure */
#define QBX_EBUSY (-5)      /* server queue full / budget exhausted */

#define QBX_KIND_I32 0      /* int32 matrices, exact int32 energies */
#define QBX_KIND_F64 1      /* double matrices, quantized int32 filter + FP64 re-walk */
#define QBX_MODE_SINGLE 0   /* one problem spread over the whole GPU */
#define QBX_MODE_BATCH 1    /* many problems sharing the GPU (batches, serving) */

typedef struct qbx_ctx qbx_ctx_t;

typedef struct {
    int32_t tie_B;        /* reference suffix width B for the tie rule; -1 = min(n, 20) */
    int32_t walk_bits;    /* bits walked by one lane (L); -1 = tuned default */
    int32_t inner_bits;   /* unrolled inner block bits (M: 1, 4, 6 or 7); -1 = 7 */
    int32_t flags;        /* QBX_FLAG_* bits, 0 = defaults */
} qbx_opts_t;

/* opts.flags, float problems (same answers either way): by default the packed 16-bit filter
 * runs where it can for problems of n >= 22 in batches (QBX_MODE_BATCH, and serving) and
 * n >= 28 solved alone, the int32 filter below (measured: the packed filter's set-up does
 * not pay for smaller problems); QBX_FLAG_FLOAT_INT32_FILTER takes the int32 filter always
 * (result path QBX_PATH_FLOAT), QBX_FLAG_FLOAT_PACKED_FILTER the packed one wherever it can
 * run (QBX_PATH_FLOAT_PAIR16).  At most one of the two. */
#define QBX_FLAG_FLOAT_INT32_FILTER 1
#define QBX_FLAG_FLOAT_PACKED_FILTER 2

typedef struct {
    int64_t E_int;        /* exact minimum energy (integer solves) */
    double E;             /* minimum energy (all solves; exact FP64 sum for float ones) */
    uint64_t state;       /* argmin state integer, x[0] = most significant bit */
    uint64_t rank;        /* rank of `state` under the tie rule */
    uint64_t states;      /* states enumerated (2^n) */
    uint64_t tiles;       /* work tiles (64 lanes each) the solve was cut into */
    double kernel_ms;     /* device time of the solve's kernels (HIP events); batch: whole launch */
    int32_t n;
    int32_t tie_B;
    int32_t walk_bits;
    int32_t inner_bits;
    int32_t lane_bits;
    int32_t status;       /* 0 = ok, QBX_ERANGE = energy bound outside the int32 path */
    int32_t path;         /* walk that ran: QBX_PATH_* */
    int32_t reserved;
} qbx_result_t;

/* qbx_result_t.path.  Integer problems whose energy bound (sum |Q_ii| + sum_{i<j} |Q_ij + Q_ji|)
 * is at most QBX_PAIR_BOUND walk two states per 32-bit operation (16-bit energies:
 * QBX_PATH_PAIR16); other integer problems whose energy differences inside a walk block
 * are expected to fit 16 bits do the same relative to each block's base, with a per-group
 * guard and int32 groups where it fails (QBX_PATH_PAIR16_REBASED); the rest walk in int32.
 * All give the same answers.  Float problems take a filter over quantized tables with FP64
 * re-walks of the blocks that pass it: in int32 (QBX_PATH_FLOAT), or, where a coarse
 * scale of the quantized block-relative energies fits 16 bits, on the packed block-relative
 * walk (QBX_PATH_FLOAT_PAIR16).  Both give the same answers. */
#define QBX_PATH_INT32 0
#define QBX_PATH_PAIR16 1
#define QBX_PATH_FLOAT 2
#define QBX_PATH_PAIR16_REBASED 3
#define QBX_PATH_FLOAT_PAIR16 4
#define QBX_PAIR_BOUND 15359

typedef struct {
    char name[64];        /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t compute_units;
    int32_t clock_khz;    /* peak engine clock */
    int32_t simd_per_cu;  /* 4 on CDNA */
    int32_t wave_size;    /* 64 */
    int64_t total_mem;    /* bytes */
    double valu_peak_ops; /* CUs * SIMDs * 16 lanes * clock: int32 VALU lane-ops per second */
} qbx_devinfo_t;

/* ------------------------------------------------------------------ context */
/* Opens HIP device `device` and a stream of default priority on it. */
int qbx_open(int device, qbx_ctx_t **ctx);
int qbx_close(qbx_ctx_t *ctx);
const char *qbx_last_error(void);
int qbx_device_info(qbx_ctx_t *ctx, qbx_devinfo_t *info);

/* ------------------------------------------------------------ single solves
 * Replaces the reference's hot loop QBF.iterate_prefix_interval (qbf.py:188-230) and the
 * QB_F1 kernel launch (qbf1.py:69-77, kernel qbf1.py:179-249), together with the caches
 * they need (qbf.py:246-289; qbf1.py:105-118), for one problem.
 *
 * qbx_solve_i32: exact int32 arithmetic.  Requires
 *   bound = sum_i |Q_ii| + sum_{i<j} |Q_ij + Q_ji| <= 2^31 - 2, else QBX_ERANGE.
 * qbx_solve_f64: Q given as doubles (float32 and integer inputs convert exactly); the int32
 *   walk runs over tables quantized to 2^s units as a filter with a proven error bound
 *   (n(n+1)/4 units per state), and the blocks it passes are walked exactly in FP64.  The
 *   answer is the FP64 (E, rank) minimum.  Requires finite entries with bound <= 1e300,
 *   else QBX_ERANGE.  1 <= n <= 56 for both. */
int qbx_solve_i32(qbx_ctx_t *ctx, const int32_t *Q, int n, const qbx_opts_t *opts,
                  qbx_result_t *out);
int qbx_solve_f64(qbx_ctx_t *ctx, const double *Q, int n, const qbx_opts_t *opts,
                  qbx_result_t *out);

/* qbx_solve_i32 for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * kernels, same answer and result fields as qbx_solve_i32; it calls no hipMalloc, hipFree,
 * hipHostMalloc nor any other allocating routine: the solve's device memory (tables,
 * reduction state, per-chunk slots) is a fixed array of QBX_NOALLOC_SCRATCH_BYTES that every
 * GPU holds from module load to the end of the process, and Q is staged from the stack.
 * A solve needing more scratch than that (n >= 44 at the default tiling: 2^20 and more
 * per-chunk slots; more walk bits shrink them) is rejected with QBX_EINVAL.  All calls on
 * one GPU share that GPU's array: calls that overlap on one GPU (from several threads or
 * contexts) race.  Each call resets what must start at zero (reduction state, tickets). */
#define QBX_NOALLOC_SCRATCH_BYTES (16 << 20)
int qbx_solve_i32_noalloc(qbx_ctx_t *ctx, const int32_t *Q, int n, const qbx_opts_t *opts,
                          qbx_result_t *out);

/* A solve with the reference's per-prefix metrics beside its answer (`out` is exactly
 * qbx_solve_*'s).  The reference walks the 2^A prefixes P (the top A = n - B bits of the
 * state, B = tie_B) in Gray order, step i visiting P = i ^ (i >> 1), and records each step's
 * best energy (energy_arr, qbf.py:342-351, 409-410) and duration (time_arr, qbf.py:362-365).
 * Here, for i = 0 .. 2^A - 1:
 *   minima[i]  = the least energy over the 2^B states of prefix i ^ (i >> 1)
 *                (int64 for qbx_solve_minima_i32, double -- FP64-exact -- for _f64);
 *   wave_ms[i] = (nullable) device time the walk spent on that prefix: the wall time of the
 *                waves that walked its states, each wave's share prorated by its lanes, in ms
 *                (the waves run concurrently, so the sum exceeds the solve's duration).
 * The walk then cuts lanes no wider than B bits (walk_bits is capped at B, inner_bits at B);
 * float problems walk without the shared bound across lanes (every lane finds its own
 * minimum), so they take longer than qbx_solve_f64.  Needs 1 <= B and A <= 28. */
#define QBX_MAX_PREFIX_BITS 28
int qbx_solve_minima_i32(qbx_ctx_t *ctx, const int32_t *Q, int n, const qbx_opts_t *opts, int64_t *minima,
                         double *wave_ms, qbx_result_t *out);
int qbx_solve_minima_f64(qbx_ctx_t *ctx, const double *Q, int n, const qbx_opts_t *opts, double *minima,
                         double *wave_ms, qbx_result_t *out);

/* One part of a solve: the states in part `part` of `parts` equal, contiguous ranges of the
 * solve's chunks (chunks are runs of tiles, tiles runs of the state's leading bits; needs
 * parts <= chunks).  out holds the (E, rank) minimum over that part (out.states, out.tiles:
 * the part's share); the lexicographic minimum of (E, rank) over all parts is exactly
 * qbx_solve_*'s answer.  This is one GPU's share of a solve spread over `parts` GPUs, whose
 * merge is a minimum over `parts` 16-byte keys.  Counterpart of the reference's bit-fixing
 * split into independent tasks (parallel.py:145-157), which also splits by leading bits. */
int qbx_solve_part_i32(qbx_ctx_t *ctx, const int32_t *Q, int n, const qbx_opts_t *opts, int part, int parts,
                       qbx_result_t *out);
int qbx_solve_part_f64(qbx_ctx_t *ctx, const double *Q, int n, const qbx_opts_t *opts, int part, int parts,
                       qbx_result_t *out);

/* ------------------------------------------------------------------ batches
 * `count` independent problems solved in one persistent launch (plus one table-building
 * launch).  Q_i is stored row-major at Qpacked + offsets[i] with n = ns[i].  Replaces a
 * loop of QBF().solve calls, i.e. the per-task do_solve of the reference's Parallel
 * bruteforce path (src/qubo_lib/solvers/parallel.py:145-157, 381-385), for problems
 * that each fit one GPU.  out[i].kernel_ms is the time of the whole launch. */
int qbx_batch_i32(qbx_ctx_t *ctx, int count, const int32_t *Qpacked, const int64_t *offsets,
                  const int32_t *ns, const qbx_opts_t *opts, qbx_result_t *out);
int qbx_batch_f64(qbx_ctx_t *ctx, int count, const double *Qpacked, const int64_t *offsets,
                  const int32_t *ns, const qbx_opts_t *opts, qbx_result_t *out);

/* Batch plans: the launches of one batch shape (count, kind, each n, opts) captured once in
 * a HIP graph with their own device arena; every run copies new matrices of that shape in
 * and replays the graph (copy in, table build, persistent walks, results out).  Results
 * are identical to qbx_batch_*; out[i].kernel_ms is the whole replay, copies included.
 * Same reference counterpart as qbx_batch_* (parallel.py:145-157, 381-385). */
typedef struct qbx_batch_plan qbx_batch_plan_t;
int qbx_batch_plan_create(qbx_ctx_t *ctx, int kind, int count, const int32_t *ns, const qbx_opts_t *opts,
                          qbx_batch_plan_t **plan);
int qbx_batch_plan_run_i32(qbx_batch_plan_t *plan, const int32_t *Qpacked, const int64_t *offsets,
                           qbx_result_t *out);
int qbx_batch_plan_run_f64(qbx_batch_plan_t *plan, const double *Qpacked, const int64_t *offsets,
                           qbx_result_t *out);
int qbx_batch_plan_destroy(qbx_batch_plan_t *plan);

/* ------------------------------------------------------- device-built RandomQUBO
 * Batches of RandomQUBO instances whose matrices are drawn on the device, bit for bit as
 * the reference draws them on the host: Generator(PCG64(seed)).integers(low, high, (N, N),
 * dtype) for integer dtypes (high exclusive), low + random((N, N), dtype) * (high - low) for
 * float dtypes (random_qubo.py:19-33; the reference's use_gpu=True path draws on the host
 * and copies).  The caller seeds PCG64 (numpy's SeedSequence) and passes each instance's
 * 128-bit state and increment.  Integer dtypes solve on the int32 path (QBX_ERANGE per
 * problem when the energy bound does not fit), float dtypes on the FP64 path.
 * qbx_random_qubo_read returns the generated matrices (packed n*n each, int32 for integer
 * dtypes, double for float dtypes) instead of solving them. */
#define QBX_GEN_INT16 16
#define QBX_GEN_INT32 32
#define QBX_GEN_FLOAT32 (-32)
#define QBX_GEN_FLOAT64 (-64)
typedef struct {
    uint64_t state_hi, state_lo, inc_hi, inc_lo;
} qbx_pcg64_t;
/* The generator states numpy's Generator(PCG64(seed)) starts from, for integer seeds
 * 0 <= seed < 2^64 (SeedSequence hash, then PCG64 seeding; random_qubo.py:21 and tsp.py:55
 * seed their generators so).  Host only: needs no context and no GPU. */
int qbx_pcg64_from_seeds(int count, const uint64_t *seeds, qbx_pcg64_t *out);
int qbx_batch_random_qubo(qbx_ctx_t *ctx, int count, const qbx_pcg64_t *rng, const int32_t *ns,
                          int64_t low, int64_t high, int dtype, const qbx_opts_t *opts, qbx_result_t *out);
int qbx_random_qubo_read(qbx_ctx_t *ctx, int count, const qbx_pcg64_t *rng, const int32_t *ns,
                         int64_t low, int64_t high, int dtype, void *Qout);
/* Batches of TSP.init_random_instance(seed, n_cities, low, high, dtype) instances (1..7
 * cities, dtype QBX_GEN_FLOAT32 or QBX_GEN_FLOAT64 for the city coordinates) whose QUBOs
 * are built on the device, bit for bit as the reference builds them (tsp.py:50-75,
 * 272-321: uniform cities, distances in the cities' type, tour cost plus A times the
 * constraints, folded to the upper triangle, cast to dtype); solved on the FP64 path.
 * The cities and their distances are computed in the calling process, with the C
 * library's pow and sqrt (the functions numpy's scalar `**` and np.sqrt call: pow(x, 2)
 * is not always the correctly rounded x*x), and staged as 21 doubles per problem; the
 * (n_cities^2)^2 QUBO is built on the device.  The reference builds it with Python loops
 * on the host.  qbx_random_tsp_read returns the matrices ((n_cities^2)^2 doubles each,
 * packed) instead of solving them. */
int qbx_batch_random_tsp(qbx_ctx_t *ctx, int count, const qbx_pcg64_t *rng, const int32_t *n_cities,
                         double low, double high, int dtype, const qbx_opts_t *opts, qbx_result_t *out);
int qbx_random_tsp_read(qbx_ctx_t *ctx, int count, const qbx_pcg64_t *rng, const int32_t *n_cities,
                        double low, double high, int dtype, double *Qout);

/* ------------------------------------------------------------------- tiling
 * Walk bits (L: states per lane = 2^L, lanes per problem = 2^(n-L)) used when
 * opts->walk_bits = -1.  qbx_set_walk_table installs a per-n table measured by the
 * autotuner (qubo_lib_amd.tune) for one kind and mode: walk_bits[n], n < count <= 57;
 * entries < 0, or invalid for the kernel variant n gets, keep the built-in rule.
 * qbx_plan_info reports the tiling a solve of size n would get, without running it.
 * The reference sizes its tiling by memory instead (QBF.define_chunk_size and
 * adjust_AB_to_fit_memory, qbf.py:430-466; QB_F1's fixed B = 18, qbf1.py:94-95). */
int qbx_set_walk_table(qbx_ctx_t *ctx, int kind, int mode, const int32_t *walk_bits, int count);
int qbx_plan_info(qbx_ctx_t *ctx, int kind, int mode, int n, const qbx_opts_t *opts, qbx_result_t *out);

/* ------------------------------------------------------------------ serving
 * A continuous-batching server: a native worker thread owns a stream pair and a pool of
 * device slots.  Requests submitted at any time are admitted when a slot is free and the
 * in-flight work (sum of 2^n) stays within the budget; every scheduling step is one finite
 * persistent launch over chunks of all admitted requests (oldest first), so requests that
 * arrive while others run share the GPU with them.  Results are identical to qbx_solve_*.
 * No counterpart in the reference, which solves one problem per call (qbf.py:48-70) and
 * spreads several over processes only through Parallel (parallel.py:124-158). */
typedef struct qbx_server qbx_server_t;

typedef struct {
    int32_t max_n;            /* largest n admitted (<= 56); sizes each slot; default 40 */
    int32_t max_inflight;     /* device slots (admitted requests); default 512 */
    double work_budget;       /* max sum of 2^n over admitted requests; default 2^40 */
    double step_states;       /* states scheduled per step launch; default 2^34 */
    int32_t walk_bits;        /* -1: per-n default */
    int32_t flags;            /* reserved, 0 */
} qbx_server_opts_t;

typedef struct {
    uint64_t id;              /* from qbx_server_submit_* */
    int32_t status;           /* QBX_OK or a QBX_E* code */
    int32_t n;
    int64_t E_int;
    double E;
    uint64_t state;
    uint64_t rank;
    double t_submit_ms;       /* server clock (ms since qbx_server_start) */
    double t_admit_ms;
    double t_done_ms;         /* when the worker saw the result on the host */
} qbx_request_result_t;

typedef struct {
    uint64_t submitted, completed, steps, chunks, admissions;
    double busy_ms;           /* device time covered by steps: the union of their intervals (HIP events) */
} qbx_server_stats_t;

int qbx_server_start(qbx_ctx_t *ctx, const qbx_server_opts_t *opts, qbx_server_t **out);
/* Copies Q; tie_B = -1 selects min(n, 20).  Never blocks on the GPU. */
int qbx_server_submit_i32(qbx_server_t *s, const int32_t *Q, int n, int tie_B, uint64_t *id);
int qbx_server_submit_f64(qbx_server_t *s, const double *Q, int n, int tie_B, uint64_t *id);
/* A burst of requests in one call (packed as for qbx_batch_*): ids[i] for problem i, in
 * order.  Every request is checked before any is queued, so a bad one queues none (the
 * error names it); they are then copied and queued 256 at a time, and the worker starts on
 * the first of them while the rest are still being copied. */
int qbx_server_submit_many_i32(qbx_server_t *s, int count, const int32_t *Qpacked, const int64_t *offsets,
                               const int32_t *ns, int tie_B, uint64_t *ids);
int qbx_server_submit_many_f64(qbx_server_t *s, int count, const double *Qpacked, const int64_t *offsets,
                               const int32_t *ns, int tie_B, uint64_t *ids);
/* Moves up to `max` completed results into out[]; waits up to wait_ms for the first one. */
int qbx_server_poll(qbx_server_t *s, qbx_request_result_t *out, int max, int wait_ms, int *count);
int qbx_server_stats(qbx_server_t *s, qbx_server_stats_t *out);
/* Finishes every submitted request, joins the worker, frees the server. */
int qbx_server_stop(qbx_server_t *s);

#ifdef __cplusplus
}
#endif

#endif /* QBX_H */
