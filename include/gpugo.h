/*
 * gpugo.h -- C ABI of libgpugo.so, the MI355X (gfx950) interval
 * branch-and-bound global optimizer.
 *
 * Plain C types only: pointers, sizes, scalars and opaque handles.  Every
 * function returns an int status (GPUGO_OK == 0, negative on failure) unless
 * stated otherwise; no exception or abort crosses this boundary.  The text of
 * the last failure on the calling thread is returned by gpugo_last_error().
 *
 * Host pointers passed in are borrowed for the duration of the call only
 * (their contents are copied).  Device memory belongs to the handle that
 * allocated it until the matching *_destroy call.  Result structs are
 * allocated by the caller.
 *
 * Reference interfaces replaced (paths relative to the GPUGO repository,
 * GPUGO_1.1/ omitted):
 *   - the interval-arithmetic module `ia.*` (Updated Interval Arithmetic
 *     Operations/IA.py:57-570, and the per-folder IA<Fn>.py modules, e.g.
 *     1-Ackley/IAAckley.py:147-467)                        -> gpugo_ia_eval
 *   - the per-objective code lists <fn>Value_IA / <fn>Sample_IA /
 *     <fn>Derivative_IA that each driver defines (1-Ackley/ackley50.py:116-479)
 *     and Numba's JIT of them                               -> gpugo_module_*
 *   - the kernels sampling / ruleOutGPU / constructFlagArr
 *     (1-Ackley/ackley50.py:602-810) together with the host driver loop
 *     (1-Ackley/ackley50.py:950-1395)                       -> gpugo_problem_*,
 *                                                              gpugo_step,
 *                                                              gpugo_solve
 *   - completeRegion (1-Ackley/ackley50.py:820-948), the host rebuild of a
 *     subregion from (parent, pos, cycleIdx)               -> gpugo_boxes_copy
 */
#ifndef GPUGO_H
#define GPUGO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GPUGO_OK            0
#define GPUGO_E_INVALID    -1  /* bad argument (null pointer, size, enum)      */
#define GPUGO_E_DOMAIN     -2  /* NaN or inverted interval in the input        */
#define GPUGO_E_NOMEM      -3  /* device allocation failed                     */
#define GPUGO_E_COMPILE    -4  /* hipRTC failure; log in gpugo_last_error()    */
#define GPUGO_E_HIP        -5  /* HIP runtime error                            */
#define GPUGO_E_NODEV      -6  /* no usable GPU                                */
#define GPUGO_E_CAPACITY   -7  /* box queue or parent arena full: the solve
                                  stopped without a certificate                */

/* ---- interval operations (gpugo_ia_eval) ---------------------------------
 * Binary ops read (a, b); scalar ops read (a, s); unary ops read a only.    */
enum gpugo_op {
  GPUGO_OP_ADD   = 0,   /* a + b                         IA.py:57-85 add     */
  GPUGO_OP_SUB   = 1,   /* a - b                         IA.py:87-115 minus  */
  GPUGO_OP_MUL   = 2,   /* a * b                         IA.py:117-152       */
  GPUGO_OP_DIV   = 3,   /* a / b, extended division      IA.py:154-207       */
  GPUGO_OP_NEG   = 4,   /* -a                                                */
  GPUGO_OP_SCALE = 5,   /* a * s, any sign of s          IA.py:306-333 times */
  GPUGO_OP_ADDC  = 6,   /* a + s                         IA.py:335-361 plus  */
  GPUGO_OP_RSUBC = 7,   /* s - a                                             */
  GPUGO_OP_DIVC  = 8,   /* a / s, s != 0                                     */
  GPUGO_OP_SQR   = 9,   /* a^2                           IA.py:209-304 pow 2 */
  GPUGO_OP_POWN  = 10,  /* a^n, n = (int)s >= 0          IA.py:209-304       */
  GPUGO_OP_SQRT  = 11,  /* sqrt(a), domain a >= 0        power(., 1/2)       */
  GPUGO_OP_EXP   = 12,  /* exp(a)                        IA.py:532-570       */
  GPUGO_OP_SIN   = 13,  /* sin(a)                        IA.py:363-446       */
  GPUGO_OP_COS   = 14,  /* cos(a)                        IA.py:448-530       */
  GPUGO_OP_COUNT = 15
};

enum gpugo_prec {
  GPUGO_PREC_F64 = 0,    /* FP64 intervals                                  */
  GPUGO_PREC_F32 = 1     /* FP32 intervals (inputs rounded outward to FP32) */
};

enum gpugo_round {
  GPUGO_ROUND_EXACT = 0, /* +,-,*,/,sqrt give the true round-down/round-up
                            results (bit-identical to IEEE directed rounding) */
  GPUGO_ROUND_FAST  = 1, /* round-to-nearest then one-ulp outward step:
                            rigorous, at most one ulp wider per operation     */
  GPUGO_ROUND_DIGITS15 = 2 /* FP64 only: the IA of the paper's results, each
                            result widened to the enclosing decimal of 15
                            significant digits (outRound, n = 15,
                            1-Ackley/IAAckley.py:106-141); never narrower
                            than EXACT.  gpugo_ia_eval and single-problem
                            modules, unrolled or windowed (no split,
                            persistent or FP32 form)                          */
};

/* ---- library ----------------------------------------------------------- */
int         gpugo_init(int device);          /* select and initialise a GPU */
const char *gpugo_last_error(void);          /* thread-local, never NULL    */
int         gpugo_version(void);             /* major*10000+minor*100+patch */
int         gpugo_device_count(void);

/* Batched interval op on host arrays of n intervals.  Replaces the device
 * functions of IA.py (see enum gpugo_op for line references).  blo/bhi may
 * be NULL for unary and scalar ops. */
int gpugo_ia_eval(int op, int prec, int rounding, int64_t n,
                  const double *alo, const double *ahi,
                  const double *blo, const double *bhi, double s,
                  double *olo, double *ohi);

/* ---- objective modules (hipRTC) -----------------------------------------
 * A module is one objective's generated HIP source (value, gradient/KKT and
 * sample code lists plus the fused kernels), compiled for gfx950.
 * Replaces Numba's @cuda.jit of the per-driver code lists
 * (1-Ackley/ackley50.py:116-479). */
typedef struct gpugo_module_s *gpugo_module_t;

int  gpugo_module_compile(const char *source, const char *name,
                          const char *extra_options, gpugo_module_t *out);
/* Compile only (no GPU needed, e.g. to fill an on-disk cache ahead of time).
 * The code object is kept per thread; gpugo_rtc_code_copy copies it out. */
int  gpugo_rtc_compile(const char *source, const char *name,
                       const char *extra_options, size_t *code_size);
int  gpugo_rtc_code_copy(void *buf, size_t size);
int  gpugo_module_load(const void *code_object, size_t size,
                       gpugo_module_t *out);
/* Copy the compiled code object out (for an on-disk cache).  With buf NULL
 * only *size is set. */
int  gpugo_module_code(gpugo_module_t m, void *buf, size_t *size);
void gpugo_module_destroy(gpugo_module_t m);

/* Evaluate the module's objective on n explicit boxes (host arrays, row-major
 * n x d).  Any output pointer may be NULL to skip it.
 *   flo/fhi      value enclosure                 (<fn>Value_IA)
 *   glo/ghi      gradient enclosure, n x d       (<fn>Derivative_IA)
 *   kkt          1 if the monotonicity test prunes the box
 *   ub           min over n_samples diagonal points of the upper bound of f
 *                                                (<fn>Sample_IA)
 * dom_lo/dom_hi is the initial box used by the monotonicity test's boundary
 * rule (1-Ackley/ackley50.py:458-477).  rounding is enum gpugo_round. */
int gpugo_eval_boxes(gpugo_module_t m, int32_t d,
                     const double *params, int64_t n_params,
                     int64_t n, const double *lo, const double *hi,
                     const double *dom_lo, const double *dom_hi,
                     int32_t n_samples, int32_t rounding,
                     double *flo, double *fhi, double *glo, double *ghi,
                     int32_t *kkt, double *ub);

/* FP32 value enclosure of n explicit boxes, for modules emitted with the
 * FP32 code list (precision-F32 problems): the parity hook of the pruning
 * pass.  Coefficients are rounded outward to FP32 intervals. */
int gpugo_eval_boxes_f32(gpugo_module_t m, int32_t d, const double *params, int64_t n_params,
                         int64_t n, const double *lo, const double *hi, int32_t rounding,
                         double *flo, double *fhi);

/* Windowed (large-d) modules: built from an objective's loop templates, they
 * keep each template sum per parent record outside the parent's split window
 * (win + t) mod d, t < c, so a child costs the terms the window touches
 * rather than all d (the reference evaluates all d coordinates per child,
 * 6-Levy/levy10000.py:160-232); the rest sums are also kept at the parent's
 * n_samples diagonal points for the incumbent samples.  Problems on such a
 * module take no flags; precision is FP64, or GPUGO_PREC_F32 when the module
 * carries the FP32 windowed value list (gpugo_eval_boxes_f32 then evaluates
 * it on explicit boxes).  This is the test hook of the window rule
 * shared by the device code: for each row of args (win, c, d, klo, khi, nt:
 * terms j in [0, nt) using coordinates j + k, k in [klo, khi]) it writes
 * (n, lo0, hi0, lo1, hi1), the n <= 2 ranges [lo, hi) of terms whose
 * coordinates meet the window.  Runs on the host; needs no GPU. */
int gpugo_window_ranges(int64_t n, const int32_t *args, int32_t *out);

/* gpugo_problem_desc.flags */
#define GPUGO_FLAG_MEAN_VALUE 1  /* also bound each child by the mean-value form
                                    f(c) + sum_j G_j (X_j - c_j), c the midpoint,
                                    G the gradient enclosure (not in the
                                    reference; tightens lower bounds)          */
#define GPUGO_FLAG_SPLIT      4  /* term-parallel split+bound: the lanes of a group
                                    (the module's G, gpugo_amd.emit split=G) share
                                    one child and split its sums; for small steps
                                    that leave the GPU partly idle.  Needs a module
                                    built from loop templates; not with
                                    GPUGO_PREC_F32.  Certified enclosures are as
                                    rigorous; bits differ from the unsplit form
                                    (sums associate differently).               */
#define GPUGO_FLAG_PERSIST    8  /* with GPUGO_FLAG_SPLIT: run the steps of a
                                    gpugo_step/gpugo_solve call in one cooperative
                                    launch (select, split+bound and compaction
                                    separated by grid barriers on resident
                                    workgroups; no launch gaps, no host polling).
                                    Same results as GPUGO_FLAG_SPLIT alone.     */
#define GPUGO_FLAG_EAGER_COMPACT 16 /* k_select 1: compact the box queue after
                                    every step.  By default a one-box-per-step
                                    problem compacts every M steps (M <= 256,
                                    M * wid^cycle <= max_queue / 8); the steps
                                    between leave consumed and ruled-out
                                    entries in place and find the next box
                                    from per-chunk minimum keys.  Same boxes,
                                    bounds and results either way; queue_peak
                                    then counts the stale entries too, and the
                                    queue can fill up to M * wid^cycle entries
                                    sooner.                                    */
#define GPUGO_FLAG_NO_TABLES  32 /* windowed (large-d) modules: evaluate every
                                    window term per child instead of reading
                                    the step's term tables (each term over a
                                    child depends only on the digits of its
                                    window coordinates, so by default it is
                                    evaluated once per digit combination and
                                    looked up).  Same bits either way; for
                                    measurement and parity tests.              */
#define GPUGO_FLAG_NO_BAND    64 /* k_select > 1: the select passes over the
                                    whole queue every step.  By default each
                                    compaction also copies the entries below
                                    a threshold (about 2048 of them, set by
                                    the last full pass) into a side list of
                                    at most 16384, and the select takes the K
                                    smallest from it while it holds at least
                                    K; every other entry lies at or above the
                                    threshold.  Same selections and results
                                    either way; for measurement and parity
                                    tests.                                     */

/* ---- problems -------------------------------------------------------------
 * One box-constrained minimisation.  Replaces the module-level constants and
 * the while-loop of each driver (1-Ackley/ackley50.py:958-1015, :1033-1380). */
typedef struct {
  int32_t       d;          /* dimension                                     */
  int32_t       wid;        /* subintervals per split coordinate (wid, :527) */
  int32_t       cycle;      /* coordinates split per step (cycleSize, :530)  */
  int32_t       k_select;   /* boxes selected per step (reference: 1)        */
  int32_t       n_samples;  /* diagonal samples per child (numSamples :1015) */
  int32_t       precision;  /* enum gpugo_prec for the pruning pass          */
  int32_t       rounding;   /* enum gpugo_round                              */
  int32_t       flags;      /* GPUGO_FLAG_* bits                             */
  const double *lo;         /* initial box, d values each (initialRegion)    */
  const double *hi;
  const double *params;     /* objective coefficients                        */
  int64_t       n_params;
  double        tol_f;      /* certify ub - lb <= tol_f; <= 0 disables       */
  double        tol_x;      /* reference x-width output rule; <= 0 disables  */
  double        init_ub;    /* initial upper bound (lowestUpperVal, :994)    */
  int64_t       max_queue;  /* box-queue capacity (entries)                  */
  int64_t       max_parents;/* parent-arena capacity (boxes)                 */
  int64_t       max_iters;  /* iteration cap for gpugo_solve                 */
} gpugo_problem_desc;

enum gpugo_status {
  GPUGO_ST_RUNNING     = 0,
  GPUGO_ST_CERTIFIED   = 1,  /* ub - lb <= tol_f                             */
  GPUGO_ST_X_CONVERGED = 2,  /* reference rule: queue drained into outputs   */
  GPUGO_ST_MAX_ITERS   = 3,
  GPUGO_ST_EXHAUSTED   = 4,  /* queue empty                                  */
  GPUGO_ST_OVERFLOW    = 5   /* capacity exceeded, no certificate            */
};

typedef struct {
  double  lb;               /* certified lower bound of the global minimum   */
  double  ub;               /* best upper bound found (incumbent)            */
  double  width;            /* ub - lb, rounded up                           */
  int32_t status;           /* enum gpugo_status                             */
  int32_t overflow;         /* with GPUGO_ST_OVERFLOW: 1 box queue full,
                               2 parent arena full                          */
  int64_t iterations;       /* steps taken                                   */
  int64_t boxes_bounded;    /* children that received a value enclosure      */
  int64_t pruned_value;     /* children pruned by lb > ub                    */
  int64_t pruned_kkt;       /* children pruned by the monotonicity test      */
  int64_t queue_size;       /* live boxes after the last step                */
  int64_t queue_peak;
  int64_t n_out;            /* boxes in the output list (x-width rule)       */
  double  ms;               /* wall time inside gpugo_step/gpugo_solve       */
  double  bound_ms;         /* summed split+bound kernel time (HIP events),
                               0 unless gpugo_problem_timing enabled it      */
  int64_t bound_launches;   /* split+bound launches that did work            */
} gpugo_result;

typedef struct gpugo_problem_s *gpugo_problem_t;

int  gpugo_problem_create(gpugo_module_t m, const gpugo_problem_desc *desc,
                          gpugo_problem_t *out);
/* Return the problem to its initial state (root box, init_ub) keeping all
 * device allocations; the next step starts a new solve. */
int  gpugo_problem_reset(gpugo_problem_t p);
/* enable != 0: time every split+bound launch with HIP events on the
 * problem's stream (reported in gpugo_result.bound_ms). */
int  gpugo_problem_timing(gpugo_problem_t p, int enable);
/* Run up to `iters` steps (fewer if the problem terminates). */
int  gpugo_step(gpugo_problem_t p, int64_t iters, gpugo_result *res);
/* Run to termination or desc.max_iters. */
int  gpugo_solve(gpugo_problem_t p, gpugo_result *res);
/* One solve that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  The problem's
 * device state (queues, parent arena, counters, incumbent) lives in a static
 * arena of the library, GPUGO_STATIC_ARENA_BYTES per GPU, held from the
 * library's load until the process ends; the caller passes the description
 * (host arrays, read during the call) and the result struct.  Capacities are
 * desc.max_queue / desc.max_parents (0: 2^20 queue entries, 2^15 parent
 * records); a problem whose buffers exceed the arena is refused with
 * GPUGO_E_CAPACITY before any GPU work.  Work runs on `stream` (a
 * hipStream_t; NULL: the default stream) and the call returns after it has
 * finished.  Every call on one GPU uses that GPU's one arena: calls that
 * overlap on a GPU (other host threads) race.  No persistent-loop flag. */
#define GPUGO_STATIC_ARENA_BYTES (128ull << 20)
int  gpugo_solve_static(gpugo_module_t m, const gpugo_problem_desc *desc, void *stream,
                        gpugo_result *res);
/* Copy live boxes (or, when the x-width rule finished, the output boxes) to
 * the host: lo/hi are max x d row-major, lb has max entries.  *n receives the
 * number of boxes available (which may exceed max). */
int  gpugo_boxes_copy(gpugo_problem_t p, int64_t max, double *lo, double *hi,
                      double *lb, int64_t *n);
void gpugo_problem_destroy(gpugo_problem_t p);

/* ---- batches: many independent problems resident on one GPU -------------
 * No counterpart in the reference, which solves one problem per script run
 * (1-Ackley/ackley50.py:950-1395); each slot runs exactly the per-problem
 * loop above (same select, split+bound and compaction semantics), so a
 * problem's result in a batch equals its gpugo_solve result with the same
 * schedule, rounding and flags.
 *
 * A batch module (gpugo_amd.emit.batch_module_source) holds several
 * objective families in variable-dimension form (runtime d <= dmax).  The
 * batch owns `slots` device-resident problem slots: queues, parent arenas,
 * incumbents and states stay in HBM across the whole run.  Problems are
 * admitted into free slots at any time (continuous batching); every launch
 * advances all running slots by `steps_per_launch` steps (one hipGraph when
 * use_graph is set); finished slots are reported and freed.
 * Threading: a batch handle is used by one host thread at a time. */
typedef struct {
  int32_t slots;             /* concurrent problems                          */
  int32_t steps_per_launch;  /* B&B steps per launch (>= 1)                  */
  int32_t rounding;          /* enum gpugo_round (module must hold it)       */
  int32_t use_graph;         /* 1: capture the steps of a launch in a graph  */
  int64_t max_queue;         /* live boxes per slot (0: 65536)               */
  int64_t max_parents;       /* parent records per slot (0: 16384)           */
  int32_t bound_blocks;      /* persistent split+bound grid (0: 8 per CU)    */
  int32_t grow;              /* growth tiers, 0..3.  0: a slot whose problem
                                outgrows its queue or parent arena stops with
                                GPUGO_ST_OVERFLOW.  n > 0: the batch also owns
                                n larger tiers; tier k has 8^k times the
                                per-slot max_queue and max_parents and
                                slots >> (4 k) slots (at least 1).  A problem
                                whose next step might not fit its slot stops
                                before that step, moves (queue, parent
                                records, incumbent, counters) into a free slot
                                of the next tier and resumes there, with the
                                same result as in a slot large enough from the
                                start; only the last tier reports overflow.
                                gpugo_batch_bytes counts all tiers.          */
  int32_t promote_age;       /* with grow > 0: a tier-0 problem that has run
                                this many steps moves to tier 1 the same way
                                (0: never), so tier 0's launches stay short
                                for the problems admitted since (served
                                form).  Same results.                        */
  int32_t pad0;
} gpugo_batch_desc;

typedef struct gpugo_batch_s *gpugo_batch_t;

/* Device bytes a batch with this description allocates (admission by memory:
 * compare with gpugo_mem_info before creating). */
int  gpugo_batch_bytes(gpugo_module_t m, const gpugo_batch_desc *desc, int64_t *bytes);
int  gpugo_mem_info(int64_t *free_bytes, int64_t *total_bytes);
int  gpugo_batch_create(gpugo_module_t m, const gpugo_batch_desc *desc, gpugo_batch_t *out);
/* Admit one problem of objective family `family` into a free slot.  desc
 * uses the per-problem fields (d <= dmax, lo, hi, params in the family's
 * dmax layout, wid, cycle, k_select, n_samples, tol_f, init_ub, max_iters,
 * flags); tol_x and the capacities are not used (per-slot capacities come
 * from the batch).  Returns GPUGO_E_CAPACITY when no slot is free.  The
 * problem starts with the next launch. */
int  gpugo_batch_admit(gpugo_batch_t b, int32_t family, const gpugo_problem_desc *desc, int64_t tag,
                       int32_t *slot);
/* Enqueue `launches` launches.  The host keeps one launch in flight and
 * collects the slots that finished (see gpugo_batch_results).  Returns early
 * when no slot is running. */
int  gpugo_batch_run(gpugo_batch_t b, int64_t launches);
/* Continuous batching without waiting (no counterpart in the reference,
 * which has no served form: for a request loop over several batches, so a
 * batch busy with long steps does not hold up the others).  Collects the
 * launches whose results have landed, then enqueues one more launch of every
 * tier that has running slots and fewer than two launches in flight;
 * *launched (or null) receives the launches enqueued.  Never blocks. */
int  gpugo_batch_advance(gpugo_batch_t b, int64_t *launched);
/* Wait for the launch in flight and collect what finished. */
int  gpugo_batch_sync(gpugo_batch_t b);
/* Move up to `max` finished results (tags[i], res[i]) out of the batch;
 * res.ms is the time from admission to the host seeing the result. */
int  gpugo_batch_results(gpugo_batch_t b, int64_t max, int64_t *tags, gpugo_result *res, int64_t *n);
/* running: slots admitted and not yet reported; free_slots: slots that can
 * take a problem now; pending: finished results not yet taken. */
int  gpugo_batch_counts(gpugo_batch_t b, int64_t *running, int64_t *free_slots, int64_t *pending);
void gpugo_batch_destroy(gpugo_batch_t b);

#ifdef __cplusplus
}
#endif

#endif /* GPUGO_H */
