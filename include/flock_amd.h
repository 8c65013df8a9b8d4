/*
 * flock_amd.h -- C ABI of the MI355X transportation-simplex solver.
 *
 * The solver replaces the hot path of researchgroup-zx93920/flock's
 * `uvModel_parallel` (src/parallel_simplex_method/parallel_uv_method/uv_model_parallel.h:19-88):
 * Vogel initial basis, UV duals, reduced-cost sweep + argmin, stepping-stone pivot.
 * All arithmetic is exact integer: int32 costs/supplies/demands/flows/duals, int64
 * objective.  Everything here is plain C: pointers, sizes and status codes.
 *
 * Thread-safety: every entry point may be called from any host thread; calls share
 * one HIP device (the current device at first use) and serialise on an internal lock
 * where they touch shared state.  Server handles are independent of each other.
 */
#ifndef FLOCK_AMD_H
#define FLOCK_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference exit()s on CUDA errors, parallel_structs.h:96-104,
 *      and silently stops at MAX_ITERATIONS, uv_model_parallel.cu:301) ---- */
enum flock_status {
  FLOCK_OK = 0,
  FLOCK_E_ITERATION_LIMIT = 1, /* pivot budget exhausted before optimality           */
  FLOCK_E_FOREST = 2,          /* basis is not a spanning tree (PIVOT_uv.cu:372-377)  */
  FLOCK_E_UNBALANCED = 3,      /* sum(supply) != sum(demand)                          */
  FLOCK_E_INVALID = 4,         /* bad sizes / negative supply or demand / null ptr    */
  FLOCK_E_RANGE = 5,           /* costs or totals too large for exact int32 duals     */
  FLOCK_E_TOO_LARGE = 6,       /* m+n beyond what the entry point supports            */
  FLOCK_E_DEVICE = 7,          /* HIP runtime error (message in flock_last_error)     */
  FLOCK_E_NONINTEGRAL = 8,     /* float costs have no exact decimal scale <= 10^6     */
  FLOCK_E_BUDGET = 9,          /* request larger than the server's memory budget      */
  FLOCK_E_BUSY = 10            /* server queue full; retry after polling             */
};

/* Initial basic feasible solution (BFS_METHOD, parallel_structs.h:45). */
enum flock_initial { FLOCK_INIT_VOGEL = 0, FLOCK_INIT_NWC = 1 };

/* Entering-arc (pricing) rule.
 *   DANTZIG: full reduced-cost sweep, most negative r, ties -> lowest row-major index
 *            (thrust::min_element, uv_model_parallel.cu:169; host scan :184-194).  With
 *            this rule the pivot sequence equals the sequential reference's.
 *   CANDIDATES: one full sweep keeps the best cell of each of a few interleaved slots per
 *            sweep lane; minor iterations re-price that list with the current duals and
 *            pivot on its most negative entry until none is negative, then sweep again.
 *   MULTI:   stepping-stone multi-pivot: the most negative candidates' cycles are found
 *            in parallel, tree arcs are claimed atomically, and every candidate that
 *            won all its arcs pivots (the reference's parallel_ss design,
 *            PIVOT_ss.cu:294-580, with exact integer conflict resolution).
 *   SWITCH:  the reference's switch_hybrid: MULTI rounds (one full sweep, then claim
 *            rounds on its candidates) until a round makes fewer than 10 pivots, then
 *            DANTZIG to optimality (switch_model_parallel.cu:166-191, 292-307). */
enum flock_pricing { FLOCK_PRICE_DANTZIG = 0, FLOCK_PRICE_CANDIDATES = 1, FLOCK_PRICE_MULTI = 2,
                     FLOCK_PRICE_AUTO = 3, FLOCK_PRICE_SWITCH = 4 };

typedef struct flock_options {
  int32_t initial;          /* flock_initial                                        */
  int32_t pricing;          /* flock_pricing                                        */
  int64_t max_iterations;   /* <= 0: 50*(m+n) + 100000                              */
  int32_t threads;          /* workgroup size for the per-instance engine; 0 = auto.
                             * 256/512 need the tree and pivot snapshot in one workgroup's
                             * LDS (about 44 bytes per vertex); larger instances get 1024 */
  int32_t team;             /* with threads == 64 (a wavefront per instance): instances
                             * per workgroup, 1, 2 or 4; 0 = the tuned policy's          */
  int32_t reserved[8];
} flock_options;

/* One balanced instance, host or device memory depending on the entry point. */
typedef struct flock_instance {
  int32_t m, n;             /* sources (rows) x sinks (columns)                      */
  const int32_t* cost;      /* m*n, row-major i*n+j                                  */
  const int32_t* supply;    /* m                                                     */
  const int32_t* demand;    /* n                                                     */
} flock_instance;

typedef struct flock_flow {  /* flowInformation (structs.h:45-52), exact integer qty */
  int32_t source, destination, qty;
} flock_flow;

/* Result of one solve.  Optional arrays are caller-owned and may be NULL:
 *   flows        capacity >= m+n-1: basic arcs with positive flow, in basis-slot order
 *                (the reference's active_flows, retrieve_solution_on_current_tree)
 *   basis_row/col/flow  capacity m+n-1: every basic arc incl. zero-flow ones, by slot
 *   u, v         capacity m and n: duals with u[0] = 0                              */
typedef struct flock_result {
  int64_t objective;        /* sum(flow * cost), exact                               */
  int64_t iterations;       /* pivots performed                                      */
  int32_t status;           /* flock_status                                          */
  int32_t n_flows;          /* number written to flows (active_flows)                */
  double  solve_ms;         /* device time of the solve (HIP events)                 */
  flock_flow* flows;
  int32_t* basis_row;
  int32_t* basis_col;
  int32_t* basis_flow;
  int32_t* u;
  int32_t* v;
  int64_t sweeps;           /* full reduced-cost sweeps performed                    */
  /* device time per phase in microseconds, measured inside the solving workgroup:
   * [0] transpose + Vogel, [1] tree build + duals, [2] full reduced-cost sweeps,
   * [3] candidate re-pricing, [4] pivot cycle discovery + leaving arc,
   * [5] pivot flow update + subtree re-rooting (incremental duals) */
  double phase_us[6];
  /* batches: device-clock microseconds from the start of the batch launch until this
   * instance's solve started / finished (its latency inside the batch) */
  double start_us, done_us;
} flock_result;

/* Last error message of the calling thread (never NULL). */
const char* flock_last_error(void);

/* Library / device set-up; optional (every entry point initialises lazily). */
int flock_init(int device);

void flock_default_options(flock_options* opt);

/* Plain solver: one instance per call, host arrays in and out, synchronous.
 * Replaces uvModel_parallel(problem, flows) + execute() + create_flows()
 * (uv_model_parallel.cu:20-65, 227-400).  m+n up to 16383: above 8000 the basis tree
 * lives in global memory (the other entry points keep it in LDS: m+n up to 8000). */
int flock_solve(const flock_instance* inst, const flock_options* opt, flock_result* res);

/* Plain solver that allocates nothing: the same solve as flock_solve (same engine and
 * pivot rules; host arrays in and out, synchronous), with every device buffer taken
 * from a fixed 48 MiB device array that the library holds on each device until the
 * process ends -- about 8 bytes per cost cell plus O(m+n), enough for 2000 x 2000;
 * larger instances return FLOCK_E_TOO_LARGE.  It never calls hipMalloc/hipFree.
 * Set-up (the array's device address, two timing events) happens in flock_init, or
 * on the first call on a device.  Calls on one device share the array and run one at
 * a time (an internal lock).  Instances of any size take the per-instance engine
 * (never the graph path of large flock_solve calls).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid allocation). */
int flock_solve_noalloc(const flock_instance* inst, const flock_options* opt, flock_result* res);

/* Batched solver: `count` independent host instances solved in one persistent
 * work-queue launch (workgroup per instance, largest first).  res[k] receives instance k.
 * Returns FLOCK_OK when the batch ran; per-instance status is in res[k].status. */
int flock_batch_solve(int32_t count, const flock_instance* inst, const flock_options* opt,
                      flock_result* res);

/* ---- device-resident batches (inputs already in HBM; used for timing) ----------- */
typedef struct flock_batch flock_batch;
/* Upload `count` host instances once; they stay resident until flock_batch_free. */
int flock_batch_create(int32_t count, const flock_instance* inst, flock_batch** out);
/* Solve every resident instance (one launch on the library stream), no host copies. */
int flock_batch_run(flock_batch* b, const flock_options* opt);
/* Device time of the last run in ms (HIP events around the launch). */
double flock_batch_last_ms(const flock_batch* b);
/* Copy results of the last run to the host (res[k] for instance k). */
int flock_batch_results(flock_batch* b, flock_result* res);
void flock_batch_free(flock_batch* b);

/* ---- continuous-batching server ---------------------------------------------------
 * Requests are admitted against an HBM budget into device slots as they free up.  The
 * device works in time slices: each slice is one launch that advances every admitted
 * request (oldest first) until the slice deadline, so new arrivals join at the next
 * slice while others are still on the GPU; admission copies overlap the running slice.
 * Costs, basis trees and duals stay resident on the device between slices. */
typedef struct flock_server flock_server;
typedef struct flock_server_config {
  int64_t hbm_budget_bytes; /* device bytes the server may hold for admitted requests */
  int32_t max_slots;        /* instances in flight at once                           */
  int32_t step_pivots;      /* optional cap on pivots per request per slice (0: none)  */
  int32_t slice_us;         /* slice length in microseconds (0: 4000)                */
  int32_t minor_limit;      /* candidate pricing: minor iterations per full sweep for
                               this server's requests (0: the library's tuning)       */
  int32_t reserved[6];
} flock_server_config;

int flock_server_create(const flock_server_config* cfg, const flock_options* opt,
                        flock_server** out);
/* Copy the instance in and queue it; returns a ticket >= 0 in *ticket.
 * FLOCK_E_BUDGET if the instance alone exceeds the budget. */
int flock_server_submit(flock_server* s, const flock_instance* inst, int64_t* ticket);
/* Advance the server: while a slice is running, admit queued requests (their uploads
 * overlap the slice); once it has ended, collect its finished requests, admit, and
 * start the next slice.  Repeats until at least one request has finished or
 * `max_wait_ms` passes (0: never waits for a running slice), then reports up to `cap`
 * finished requests: tickets[k] with res[k] filled (res[k]'s optional arrays are used
 * when non-NULL).  *n_done receives the count.  The next slice keeps running after
 * poll returns. */
int flock_server_poll(flock_server* s, double max_wait_ms, int32_t cap, int64_t* tickets,
                      flock_result* res, int32_t* n_done);
/* Requests submitted but not yet reported. */
int32_t flock_server_pending(const flock_server* s);
void flock_server_destroy(flock_server* s);

/* ---- size-class policy (which engine configuration solves an instance) ----------
 * Instances with m+n <= wave_max_v get one 64-lane wavefront each, wave_pack (1, 2 or
 * 4) per workgroup; larger ones get a 256-thread workgroup up to wg256_max_v, else a
 * 1024-thread workgroup.  flock_autotune() measures candidates on a sample and installs
 * the fastest; flock_tuning_set/get read and write the policy directly. */
typedef struct flock_tuning {
  int32_t wave_max_v;
  int32_t wave_pack;
  int32_t wg256_max_v;
  int32_t minor_limit;      /* candidate pricing: minor iterations per full sweep     */
  /* launches of at most one instance per compute unit (the plain solver, small batches)
   * favour per-instance latency: 256 threads up to few_256_max_v, 512 up to
   * few_512_max_v, else 1024 */
  int32_t few_256_max_v;
  int32_t few_512_max_v;
  /* must be 0 (team sweeps, where idle workgroups helped sweep one large matrix, are
   * not part of this library; large matrices take the graph path below) */
  int32_t team_min_cells;
  /* plain solves of matrices with at least this many cells run as a graph of sweep-grid
   * and pivot-round kernels (0: never) */
  int32_t graph_min_cells;
  /* per mille (0..1000) of each launch's instances, largest first, whose workgroups run at
   * raised wave priority, so that co-resident workgroups do not slow the instances that
   * set the launch's makespan */
  int32_t prio_permille;
  /* large-instance path (graph of sweep grid and pivot rounds): the sweep tile's shape, in
   * 16-byte cost vectors per thread of a 256-thread tile (1..64; default 4) */
  int32_t sweep_vectors;
  /* large-instance path: threads of the set-up and pivot-round workgroup, 512 or 1024
   * (512 applies to instances with m+n <= 2048 whose basis tree fits LDS) */
  int32_t large_threads;
  /* large-instance path: candidate pricing's minor iterations per full sweep (0: minor_limit) */
  int32_t large_minor_limit;
} flock_tuning;
int flock_tuning_get(flock_tuning* t);
int flock_tuning_set(const flock_tuning* t);
/* Autotuner: uploads the `count` sample instances once, then measures candidate
 * policies on them (one warm-up and one timed batch launch each, coordinate descent
 * over the minor-iteration limit, the 256- vs 1024-thread threshold, instances per
 * workgroup (a wavefront per instance, 1, 2 or 4 per workgroup, for the smallest
 * instances) and the raised-priority share), checks that every candidate returns the
 * same objectives, installs the fastest and returns it in *out with its batch time in
 * *best_ms. */
int flock_autotune(int32_t count, const flock_instance* inst, const flock_options* opt, flock_tuning* out,
                   double* best_ms);

/* Autotuner of the large-instance path: solves `inst` (which must take the graph path:
 * at least graph_min_cells cells, or m+n > 8000) with each sweep tile shape (1, 2, 4, 8
 * vectors per thread), then at the best shape with the other set-up / round workgroup
 * size (512 or 1024 threads), then (candidate pricing) with the path's minor limit at
 * 32, 64 and 128, two solves each, checks that the objectives agree, installs the
 * fastest and returns it in *out with its solve time (device ms) in *best_ms. */
int flock_autotune_large(const flock_instance* inst, const flock_options* opt, flock_tuning* out, double* best_ms);

/* ---- measurement ----------------------------------------------------------------
 * The large-instance path's reduced-cost sweep (the grid of sweep-tile workgroups over
 * the padded int32 matrix) on `inst`, launched `reps` times back to back on the
 * library stream with all duals zero, timed with HIP events around the launches:
 * *sweep_ms receives the average time per sweep including the gaps between launches
 * (Sigma time / reps).  *empty_ms receives the same for an empty kernel with the same
 * grid: the launch-to-completion floor of one launch.  *grid (optional) receives the
 * number of workgroups.  Not part of the reference's interface: it exists to measure
 * the sweep against the HBM roofline (4*m*n + 4*(m+n) bytes per sweep). */
int flock_measure_sweep(const flock_instance* inst, int32_t reps, double* sweep_ms, double* empty_ms,
                        int32_t* grid);

/* ---- float-cost adapter of the reference's ProblemInstance (structs.h:21-40) ------
 * Finds the smallest k in 0..6 such that every float cost equals float(N/10^k) for an
 * integer N (the decimal the float was parsed from) and writes N; never rounds.
 * Returns FLOCK_E_NONINTEGRAL if no such k exists.  *scale receives k. */
int flock_scale_float_costs(int64_t count, const float* costs, int32_t* out, int32_t* scale);

#ifdef __cplusplus
}
#endif
#endif /* FLOCK_AMD_H */
