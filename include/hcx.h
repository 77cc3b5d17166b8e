/*
 * hcx.h — C-ABI of the MI355X homotopy-continuation path tracker (libhcx.so).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * an hcx_status (message via hcx_last_error(), thread-local).  Citations name
 * the reference interface (C-H-Chien/Homotopy-Continuation-Tracker-on-GPU,
 * paths relative to its root) that each entry point replaces.
 *
 * Conventions
 *  - Complex numbers are hcx_complex {re, im} in FP64 (layout of complex128).
 *  - "d_" pointers are device (HBM) pointers owned by the caller; "h_" pointers
 *    are host pointers.  The library owns only what a *_create call returns.
 *  - Calls that take a stream are asynchronous on that stream (a hipStream_t
 *    passed as void*; NULL = the default stream) unless stated otherwise.
 *  - Per-path outputs are indexed path = instance * num_tracks + root.
 */
#ifndef HCX_H
#define HCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HCX_ABI_VERSION 1

typedef struct { double re, im; } hcx_complex;

typedef enum {
  HCX_OK = 0,
  HCX_ERR_INVALID_ARG = 1,
  HCX_ERR_HIP = 2,
  HCX_ERR_OUT_OF_MEMORY = 3,
  HCX_ERR_UNSUPPORTED = 4,
  HCX_ERR_IO = 5,
  HCX_ERR_NOT_READY = 6,
  HCX_ERR_BUDGET = 7,
  HCX_ERR_NO_DEVICE = 8
} hcx_status;

/* Message of the last failing call on this thread ("" if none). */
const char* hcx_last_error(void);
int hcx_abi_version(void);
/* Number of visible HIP devices (0 when no GPU or no HIP runtime). */
int hcx_device_count(void);

/* ------------------------------------------------------------------------ */
/* Tracker options.  Replaces the compile-time switches of                  */
/* GPU-HC/magmaHC/definitions.hpp:5-51 and the GPU-HC settings keys of       */
/* GPU-HC/problems/<p>/gpuhc_settings.yaml (parsed GPU_HC_Solver.cpp:44-68). */
/* ------------------------------------------------------------------------ */
typedef enum {
  HCX_PREC_FP64 = 0,   /* everything complex FP64 */
  HCX_PREC_MIXED = 1   /* FP32 RK4 predictor (Jacobian, dH/dt, LU); FP64 Newton corrector (FP32 before the
                          end zone when fp32_corrector = 1); FP64 refinement; a path that ends without a
                          refined root is re-tracked from t = 0 in FP64 after the launch (escalation) */
} hcx_precision;

typedef struct {
  int max_steps;          /* GPUHC_Max_Steps: predictor steps incl. rejected (loop runs max_steps+1 times) */
  int max_corr;           /* GPUHC_Max_Correction_Steps (>= 1) */
  int incr_steps;         /* GPUHC_Num_Of_Steps_to_Increase_Delta_t */
  double dt0;             /* initial step, 0.05 (kernel_HC_Solver_generalized_3views_4pts.cu:78) */
  double end_zone;        /* end zone width, 0.05 -> t >= 0.95 (:121-131) */
  double t_tol;           /* stop when 1 - t <= t_tol, 1e-7 (:116) */
  double dt_min;          /* fail the path when dt < dt_min (0 = off, reference behaviour) */
  double corr_tol;        /* Newton success: |dx|^2 < corr_tol * |x|^2, 1e-6 (:214) */
  double inf_tol;         /* infinity failure: |x|^2 > inf_tol, 1e14 (:215) */
  int refine_iters;       /* FP64 Newton iterations on H(x,1) after tracking (0 = off) */
  double refine_tol;      /* refinement success: |dx| <= refine_tol * max(1,|x|) (1e-10; Newton converges
                             quadratically, so the accepted root is far more accurate than the last step) */
  double real_tol;        /* root is real if all |Im x_i| <= real_tol (definitions.hpp:25-26) */
  double unique_tol;      /* duplicate if all |dRe|,|dIm| < unique_tol (definitions.hpp:24) */
  int precision;          /* hcx_precision */
  int lanes_per_path;     /* lanes of a wave64 per path (G); 0 = autotuned / n */
  int waves_per_cu;       /* persistent grid size per CU; 0 = from occupancy; < 0: occupancy less that
                             many waves (room for other solves in flight on other streams) */
  int specialize;         /* 1: size-specialised kernel instance when one exists (n = G); 0: generic */
  int retrack_duplicates; /* mixed precision: re-track in FP64 every path whose refined root duplicates
                             another root of its instance (path jumping), so roots and counts are the
                             FP64 tracker's (1, default) */
  int fp32_corrector;     /* mixed precision: 1 = Newton corrector in FP32 before the end zone (FP64 in it);
                             0 (default) = FP64 corrector at every t, so step acceptance is decided in FP64 */
  double real_ref_tol;    /* HCX_FLAG_REAL_REF: the reference's real test on the tracked root,
                             ZERO_IMAG_PART_TOL_FOR_SP = 1e-4 (definitions.hpp:25, Evaluations.cpp:96-104) */
  int replicas;           /* FP64, one-segment launches on sizes with a latency instance: R copies of each
                             path's lane group, each summing every R-th term of the evaluation (a power of
                             two, R * lanes <= 64); 0 = automatic (launches of at most one wave per SIMD),
                             1 = off */
} hcx_track_opts;

void hcx_track_opts_default(hcx_track_opts* o);

/* ------------------------------------------------------------------------ */
/* Problem = one polynomial-system type with its device-resident tables and */
/* start solutions.  Replaces GPU_HC_Solver::Allocate_Arrays /               */
/* Read_Problem_Data / Data_Transfer_From_Host_To_Device                     */
/* (GPU-HC/magmaHC/GPU_HC_Solver.cpp:104-205).                               */
/* ------------------------------------------------------------------------ */
typedef struct {
  int num_vars;           /* n  (Num_Of_Vars), 1..32 */
  int num_coeffs;         /* C  (Num_Of_Coeffs_From_Params); index C is the constant 1 */
  int max_order_t;        /* M  (Max_Order_Of_T), >= 1 */
  int hx_terms, hx_parts; /* dHdx_Max_Terms / dHdx_Max_Parts */
  int ht_terms, ht_parts; /* dHdt_Max_Terms / dHdt_Max_Parts */
  const int* h_hx_idx;    /* [n eq][n var][hx_terms][hx_parts]  dHdx_indx.txt layout */
  const int* h_ht_idx;    /* [n eq][ht_terms][ht_parts]         dHdt_indx.txt layout */
  int num_tracks;         /* K start solutions (Num_Of_Tracks) */
  const hcx_complex* h_start_sols; /* [K][n] (start_sols.txt, without the homogenising 1) */
  /* Optional parameter-affine coefficient map c_j(p) = A[j][0] + sum_k A[j][1+k] p_k,
   * used by the on-device P2C of hcx_track when a batch gives target parameters. */
  int num_params;                  /* k (0 = no map) */
  const hcx_complex* h_coeff_affine; /* [C][1+k] or NULL */
  const hcx_complex* h_start_params; /* [k] start parameters p0 or NULL */
  /* Optional parameter-polynomial coefficient map, the alternative to the
   * affine map for coefficients that are polynomials in the parameters (the
   * reference's P2C problems, whose generated p2c-<p>.h headers play this role,
   * Data_Reader.cpp:59-89; and the PH type, coefficient = p_a * p_b):
   *   c_j(p) = sum_{u < cmap_terms} cmap_scalar[j][u] * prod_{d < cmap_deg} p[cmap_params[j][u][d]]
   * with p[k] = 1 as the padding factor.  On the straight-line parameter path
   * p(t) = p0 + t (p1 - p0) (createUnivPoly4Coeffs.m) each c_j is a polynomial
   * in t of degree <= cmap_deg <= max_order_t, built on the device per instance.
   * Needs num_params = k and h_start_params; h_coeff_affine must then be NULL. */
  int cmap_terms, cmap_deg;
  const hcx_complex* h_cmap_scalar;  /* [C][cmap_terms] (0 = padding term) */
  const int* h_cmap_params;          /* [C][cmap_terms][cmap_deg], entries in 0..k */
} hcx_problem_desc;

typedef struct hcx_problem hcx_problem;

typedef struct {
  int num_vars, num_coeffs, max_order_t, num_tracks, num_params;
  int hx_terms, ht_terms, hx_vars, ht_vars; /* padded terms; max variable factors per term */
  /* Algorithmic FLOPs (LAPACK Working Note 41 convention: complex mul 6, add 2)
   * of the reference algorithm per unit of work. */
  double flops_coef;      /* coefficient polynomials at one t */
  double flops_hx;        /* one dH/dx evaluation (all n rows, padding excluded) */
  double flops_ht;        /* one dH/dt evaluation */
  double flops_h;         /* one H evaluation */
  double flops_lu;        /* one complex LU + solve, 1 RHS */
  double flops_stage;     /* one RK4 stage */
  double flops_newton;    /* one Newton iteration */
} hcx_problem_info;

hcx_status hcx_problem_create(const hcx_problem_desc* d, hcx_problem** out);
hcx_status hcx_problem_destroy(hcx_problem* p);
hcx_status hcx_problem_get_info(const hcx_problem* p, hcx_problem_info* info);

/* ------------------------------------------------------------------------ */
/* Batched tracking.  Generic, size-as-data replacement of the per-problem   */
/* launchers kernel_HC_Solver_<p>(queue, max_steps, max_corr, incr,          */
/*   d_startSols_array, d_Track_array, d_Hx_idx, d_Ht_idx, d_phc_coeffs_Hx,  */
/*   d_phc_coeffs_Ht, d_converge, d_infinity, d_debug)                       */
/* (GPU-HC/magmaHC/gpu-kernels/magmaHC-kernels.hpp:23-364; implementation    */
/*  e.g. kernel_HC_Solver_generalized_3views_4pts.cu:267-377), extended from */
/* one instance to a batch of instances tracked in one persistent launch.    */
/* ------------------------------------------------------------------------ */
typedef struct {
  int num_instances;                 /* I; 0 is a valid empty batch (no buffers needed) */
  /* Either per-instance coefficient-polynomial-in-t tables (the P2C output,
   * Data_Reader.cpp:59-89):  Hx[I][C+1][M+1] and Ht[I][C+1][M] ...           */
  const hcx_complex* d_coeffs_hx;
  const hcx_complex* d_coeffs_ht;
  /* ... or per-instance target parameters p1 [I][k] (needs the affine or the
   * parameter-polynomial coefficient map); p(t) = (1-t) p0 + t p1.
   * d_start_params [I][k] may be NULL (problem's p0 for every instance).  When
   * d_coeffs_hx is non-NULL these are ignored.  d_coeff_work must then hold
   * I*(C+1)*(2M+1) hcx_complex (M = max_order_t; 3 per coefficient at M = 1). */
  const hcx_complex* d_target_params;
  const hcx_complex* d_start_params;
  hcx_complex* d_coeff_work;
} hcx_batch_desc;

#include "hcx_flags.h"

typedef struct {
  hcx_complex* d_roots;      /* [P][n] final x (refined when refine_iters > 0) */
  uint32_t* d_flags;         /* [P] HCX_FLAG_* */
  int32_t* d_steps;          /* [P] predictor steps incl. rejected (may be NULL) */
  int32_t* d_newton;         /* [P] Newton iterations while tracking (may be NULL) */
  int32_t* d_refine;         /* [P] refinement iterations (may be NULL) */
  double* d_t_final;         /* [P] t at stop (reference debug slot, :262-264) (may be NULL) */
  double* d_dt_final;        /* [P] dt at stop (may be NULL) */
  double* d_residual;        /* [P] max_i |H_i(x,1)| after refinement (may be NULL) */
  uint64_t* d_inst_done;     /* [I] device wall clock (100 MHz) when the instance's
                                last path finished (may be NULL) */
  uint64_t* d_batch_start;   /* [1] device wall clock when the launch started (may be NULL) */
} hcx_result_desc;

/* Tracks every (instance, start root) path of a batch on `stream`.
 * Asynchronous.  The launch workspace (path-queue counters, output
 * descriptors, per-instance counters, the FP64 re-track list) belongs to the
 * stream: launches on one stream share it in stream order, launches on
 * different streams never share it, so one problem may be tracked on several
 * streams at once.  A stream's workspace grows, after waiting for the
 * stream, when a launch needs more room; during graph capture it cannot grow
 * (HCX_ERR_INVALID_ARG: run the same launch once uncaptured first), and a
 * workspace used under capture is never freed or resized afterwards. */
hcx_status hcx_track(hcx_problem* p, const hcx_batch_desc* b, const hcx_track_opts* o,
                     const hcx_result_desc* r, void* stream);

/* hcx_track that allocates nothing.  No counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Inputs and outputs
 * are the caller's as for hcx_track (a batch of target parameters still
 * brings its d_coeff_work); the launch workspace is a static device array of
 * about 4.3 MiB that each GPU holds until the process ends, with room for
 * HCX_NOALLOC_MAX_INSTANCES instance counters (used when d_inst_done is set)
 * and HCX_NOALLOC_MAX_PATHS re-track entries (used in mixed precision); a
 * larger batch is rejected with HCX_ERR_INVALID_ARG.  Every call on one GPU
 * shares that array and it keeps the last call's contents (the launch
 * clears what it needs): calls that overlap on one GPU (several streams or
 * threads) race. */
#define HCX_NOALLOC_MAX_INSTANCES 65536
#define HCX_NOALLOC_MAX_PATHS 1048576
hcx_status hcx_track_noalloc(hcx_problem* p, const hcx_batch_desc* b, const hcx_track_opts* o,
                             const hcx_result_desc* r, void* stream);

/* Frees the launch workspace of `stream` after waiting for the stream (call
 * before destroying a stream that tracked; a workspace that a captured graph
 * refers to stays alive).  No counterpart in the reference, whose launcher
 * uses no workspace (kernel_HC_Solver_generalized_3views_4pts.cu:299-300). */
hcx_status hcx_stream_release(void* stream);

/* Same work with the reference's launch shape: each lane group tracks exactly
 * one path and exits (grid = paths / groups per wave, no refill, no
 * persistence) — one launch per problem instance is the plain per-problem
 * baseline (kernel_HC_Solver_generalized_3views_4pts.cu:299-300,361-371).
 * In mixed precision the FP64 re-track pass (escalations, duplicate roots)
 * follows as after hcx_track. */
hcx_status hcx_track_static(hcx_problem* p, const hcx_batch_desc* b, const hcx_track_opts* o,
                            const hcx_result_desc* r, void* stream);

/* Several problem types in ONE persistent launch (uneven batch).  Segment s
 * tracks problems[s] with batches[s] into results[s]. */
hcx_status hcx_track_multi(int num_segments, hcx_problem* const* problems,
                           const hcx_batch_desc* batches, const hcx_result_desc* results,
                           const hcx_track_opts* o, void* stream);

/* Reference-shaped synchronous launcher for ONE instance: tables and flags as
 * in kernel_HC_Solver_<p> (magmaHC-kernels.hpp:23-364).  d_start_sols and
 * d_track are [K][n+1] (last entry 1, Data_Reader.cpp:110-113), d_converge /
 * d_infinity are bool [K], d_debug is (t, dt) per path.  Returns the elapsed
 * seconds of the launch like the reference (<0 on error).  Allocates nothing:
 * its counters live with the problem (hcx_problem_create), and calls on one
 * problem are serialised. */
double hcx_kernel_HC_Solver(void* stream, const hcx_problem* p, int max_steps, int max_corr,
                            int incr_steps, const hcx_complex* d_start_sols, hcx_complex* d_track,
                            const hcx_complex* d_coeffs_hx, const hcx_complex* d_coeffs_ht,
                            uint8_t* d_converge, uint8_t* d_infinity, hcx_complex* d_debug);

hcx_status hcx_sync(void* stream);

/* ------------------------------------------------------------------------ */
/* Launch-shape autotuner (no counterpart in the reference, whose launch      */
/* shape is fixed per problem: kernel_HC_Solver_<p>.cu:299-300).  Times       */
/* lanes-per-path (paths per wavefront = 64 / G), kernel variant and grid     */
/* density on a sample batch and stores the winner in the problem; later     */
/* launches use it when hcx_track_opts.lanes_per_path == 0.                  */
/* ------------------------------------------------------------------------ */
typedef struct {
  int lanes_per_path;     /* G: paths per wavefront = floor(64 / G) */
  int specialize;         /* size-specialised kernel instance (1) or generic (0) */
  int waves_per_cu;       /* persistent grid density (0 = occupancy) */
  double best_ms;         /* sample launch time of the chosen shape */
  int candidates;         /* shapes timed */
} hcx_tune_result;

hcx_status hcx_autotune(hcx_problem* p, const hcx_batch_desc* sample, const hcx_track_opts* o,
                        hcx_tune_result* out, void* stream);
/* Sets the stored launch shape directly (lanes_per_path = 0 clears it). */
hcx_status hcx_problem_set_launch_shape(hcx_problem* p, int lanes_per_path, int specialize, int waves_per_cu);

/* ------------------------------------------------------------------------ */
/* HIP graphs: one uneven multi-type batch (hcx_track_multi) captured once   */
/* and replayed with hcx_graph_launch (same device buffers every replay).    */
/* ------------------------------------------------------------------------ */
typedef struct hcx_graph hcx_graph;
hcx_status hcx_graph_capture_multi(int num_segments, hcx_problem* const* problems, const hcx_batch_desc* batches,
                                   const hcx_result_desc* results, const hcx_track_opts* o, hcx_graph** out);
hcx_status hcx_graph_launch(hcx_graph* g, void* stream);
hcx_status hcx_graph_destroy(hcx_graph* g);

/* ------------------------------------------------------------------------ */
/* Serving loop: continuous batching of solve requests (one parameter        */
/* instance of a registered problem type each).  Replaces the reference's    */
/* one-instance-per-synchronous-launch flow (GPU_HC_Solver::Solve_by_GPU_HC,  */
/* GPU_HC_Solver.cpp:207-384, driven per RANSAC target by                    */
/* cmd/magmaHC-main.cpp:93-112).  Requests are queued per type and launched  */
/* as micro-batches on up to four streams while an HBM budget admits them;   */
/* tables and start roots stay resident; results are returned per request.   */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t hbm_budget_bytes;   /* device memory for in-flight requests (params, P2C tables, results) */
  int max_batch_requests;     /* requests per micro-batch launch */
  int streams;                /* concurrent launches, 1..4 */
  int max_queue;              /* queued requests before submit rejects */
} hcx_server_opts;

typedef struct {
  int64_t submitted, completed, rejected;
  int64_t inflight_requests, inflight_paths;
  int64_t launches, paths_done, converged_paths;
  int64_t hbm_in_use, peak_inflight_paths;
  double p50_ms, p99_ms, mean_ms, max_ms;   /* submit -> result latency */
} hcx_server_stats;

typedef struct {
  int type_id, num_paths, refined, real;
  double latency_ms;   /* submit -> result available */
  double queue_ms;     /* submit -> launch */
} hcx_request_result;

typedef struct hcx_server hcx_server;

void hcx_server_opts_default(hcx_server_opts* o);
hcx_status hcx_server_create(const hcx_server_opts* so, const hcx_track_opts* to, hcx_server** out);
/* Registers a resident problem type (needs its parameter-affine map); all
 * types are added before the first submit. */
hcx_status hcx_server_add_type(hcx_server* s, hcx_problem* p, int* type_id);
/* Queues one request: h_target_params holds the type's k parameters.
 * Thread-safe; returns immediately. */
hcx_status hcx_server_submit(hcx_server* s, int type_id, const hcx_complex* h_target_params, uint64_t* request_id);
/* *done = 1 once the request finished; then fills *out and (if non-NULL)
 * h_roots [K][n] and h_flags [K], and forgets the request.  With out, h_roots
 * and h_flags all NULL it only reports *done and keeps the request. */
hcx_status hcx_server_poll(hcx_server* s, uint64_t request_id, int* done, hcx_request_result* out,
                           hcx_complex* h_roots, uint32_t* h_flags);
/* Waits until every submitted request has finished (or the timeout). */
hcx_status hcx_server_drain(hcx_server* s, double timeout_s);
hcx_status hcx_server_get_stats(hcx_server* s, hcx_server_stats* out);
hcx_status hcx_server_reset_stats(hcx_server* s);
hcx_status hcx_server_destroy(hcx_server* s);

/* ------------------------------------------------------------------------ */
/* Solution collection (GPU-HC/magmaHC/Evaluations.cpp:66-176).  Sets        */
/* HCX_FLAG_UNIQUE and HCX_FLAG_UNIQUE_REF per instance on device and        */
/* returns host counts.                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t tracked;    /* Num_Of_Coverged_Sols (reference semantics: t reached 1) */
  int64_t infinity;   /* Num_Of_Inf_Sols */
  int64_t refined;    /* root-level converged: tracked and FP64-refined */
  int64_t real;       /* refined and real (|Im x_i| <= real_tol on the refined root) */
  int64_t unique;     /* refined, first occurrence of its root within the instance */
  /* The reference's own integer outputs, per instance: */
  int64_t unique_ref; /* Num_Of_Unique_Sols by Find_Unique_Sols (Evaluations.cpp:127-176): a tracked
                         path counts when no later path of its instance (any outcome) has the same
                         root (|dRe|, |dIm| < unique_tol in every coordinate), skipping paths that
                         duplicate the last path found to have a later duplicate (the Skip set is
                         replaced, not accumulated); a duplicate pair therefore adds 0.  Computed on
                         the output roots: the tracked roots when refine_iters = 0. */
  int64_t real_ref;   /* Num_Of_Real_Sols (Evaluations.cpp:88-110): paths with HCX_FLAG_REAL_REF */
} hcx_counts;

hcx_status hcx_collect(int n, int num_instances, int num_tracks, const hcx_complex* d_roots,
                       uint32_t* d_flags, double unique_tol, hcx_counts* out, void* stream);

/* Host-side collection on host arrays (same semantics as hcx_collect). */
hcx_status hcx_collect_host(int n, int num_instances, int num_tracks, const hcx_complex* h_roots,
                            uint32_t* h_flags, double unique_tol, hcx_counts* out);

/* ------------------------------------------------------------------------ */
/* Parametric polynomial systems: synthetic generator, table compiler, P2C.  */
/* Replaces the offline generator auto-data-gen-tools/autogen_gpuhc.m:503-699 */
/* (dHdx/dHdt index tables) and the generated P2C headers                    */
/* GPU-HC/magmaHC/PHC_Coeffs/p2c-<p>.h for parameter-affine coefficients.    */
/* ------------------------------------------------------------------------ */
typedef struct hcx_system hcx_system;

typedef struct {
  int num_vars;          /* n */
  int num_params;        /* k */
  int num_quadratic;     /* equations 0..num_quadratic-1 have degree 2, the rest degree 3 */
  int monomials;         /* monomials per equation (>= n+1); <= 0: dense (all of degree <= d) */
  uint64_t seed;
} hcx_synth_spec;

typedef struct {
  int num_vars, num_params, num_coeffs;
  int max_degree, max_terms;     /* max monomial degree, max monomials per equation */
  int hx_terms, hx_parts, ht_terms, ht_parts; /* compiled reference-format table sizes */
} hcx_system_info;

hcx_status hcx_system_create_synthetic(const hcx_synth_spec* spec, hcx_system** out);
/* exps [C][n] exponents of monomial j, eq [C] equation of monomial j
 * (monomial j carries coefficient j), affine [C][1+k] (may be NULL). */
hcx_status hcx_system_create(int n, int num_params, int num_coeffs, const int* exps, const int* eq,
                             const hcx_complex* affine, hcx_system** out);
hcx_status hcx_system_destroy(hcx_system* s);
hcx_status hcx_system_get_info(const hcx_system* s, hcx_system_info* info);
hcx_status hcx_system_get_monomials(const hcx_system* s, int* exps, int* eq);
hcx_status hcx_system_get_affine(const hcx_system* s, hcx_complex* affine);
/* Reference-format tables (dHdx_indx.txt / dHdt_indx.txt layouts). */
hcx_status hcx_system_compile_tables(const hcx_system* s, int* hx_idx, int* ht_idx);
/* c_j(p) for one parameter vector (host P2C). */
hcx_status hcx_system_eval_coeffs(const hcx_system* s, const hcx_complex* params, hcx_complex* coeffs);
/* Coefficient tables of the parameter homotopy p(t) = (1-t) p0 + t p1:
 * Hx[C+1][2] = {c(p0), c(p1)-c(p0)}, Ht[C+1][1] = {c(p1)-c(p0)}; entry C = {1,0}. */
hcx_status hcx_system_param_homotopy_coeffs(const hcx_system* s, const hcx_complex* p0,
                                            const hcx_complex* p1, hcx_complex* hx, hcx_complex* ht);
/* Problem descriptor pieces of a system: fills *d with sizes, table pointers
 * (owned by s) and the affine map; the caller adds start solutions. */
hcx_status hcx_system_problem_desc(const hcx_system* s, hcx_problem_desc* d);

/* Total-degree homotopy H = (1-t) gamma G + t F with G_i = x_i^{d_i} - 1
 * (reference B.3; the gamma trick of definitions.hpp:18,35).  Builds the
 * merged system (F's monomials plus x_i^{d_i} and 1) as a new hcx_system
 * *out_sys with C' coefficients, and its instance tables hx [C'+1][2],
 * ht [C'+1][1] for target coefficients c (length C of s).  out_sys is owned
 * by the caller (hcx_system_destroy). Sizes: query with hx = ht = NULL. */
hcx_status hcx_total_degree_homotopy(const hcx_system* s, const hcx_complex* target_coeffs,
                                     hcx_complex gamma, hcx_system** out_sys, int* out_num_coeffs,
                                     hcx_complex* hx, hcx_complex* ht);
/* Start roots of G: root r in mixed radix over the degrees, x_i = exp(2 pi i m_i / d_i). */
hcx_status hcx_total_degree_start_roots(const hcx_system* s, int count, const int64_t* root_index,
                                        hcx_complex* roots);
/* Seeded subset of start-root indices: the first `count` of a seeded
 * permutation of [0, prod d_i) (all of them in order if count >= prod d_i). */
hcx_status hcx_total_degree_subset(const hcx_system* s, int count, uint64_t seed, int64_t* root_index);

/* Deterministic CN(0,1) draws (SplitMix64 + Box-Muller): z = (a + i b)/sqrt(2). */
hcx_status hcx_random_cn(uint64_t seed, int64_t count, hcx_complex* out);
/* gamma = exp(2 pi i u), u ~ U[0,1) from the same generator. */
hcx_complex hcx_random_gamma(uint64_t seed);

/* ------------------------------------------------------------------------ */
/* Reference on-disk problem folders (GPU-HC/problems/<p>/):              */
/* gpuhc_settings.yaml, start_params.txt,                                    */
/* target_params.txt, start_sols.txt, dHdx_indx.txt, dHdt_indx.txt.          */
/* Replaces Data_Reader::Read_* (GPU-HC/magmaHC/Data_Reader.cpp:91-225) and   */
/* the YAML parsing of GPU_HC_Solver.cpp:44-68.                              */
/* ------------------------------------------------------------------------ */
typedef struct {
  char problem_name[128];
  char gpuhc_type[16];          /* "P2C" or "PH" */
  int max_steps, max_corr, incr_steps;
  int num_vars, num_params, num_tracks;
  int hx_terms, hx_parts, ht_terms, ht_parts, max_order_t, num_coeffs;
} hcx_ref_settings;

hcx_status hcx_ref_read_settings(const char* problem_dir, hcx_ref_settings* s);
/* Each reader fills a caller buffer sized from the settings; *count returns
 * the number of values read (complex or int). */
hcx_status hcx_ref_read_params(const char* path, hcx_complex* out, int cap, int* count);
hcx_status hcx_ref_read_start_sols(const char* path, int n, hcx_complex* out, int cap_tracks, int* tracks);
hcx_status hcx_ref_read_indices(const char* path, int* out, int cap, int* count);
/* Writes converged roots like Evaluations::Write_Converged_Sols (Evaluations.cpp:66-86). */
hcx_status hcx_ref_write_converged(const char* path, int n, int num_instances, int num_tracks,
                                   const hcx_complex* h_roots, const uint32_t* h_flags, uint32_t mask);

/* Parameter-polynomial coefficient map file (written by the problem compiler,
 * hcx/compiler.py; it takes the place of the generated p2c-<p>.h of a P2C
 * problem, GPU-HC/magmaHC/PHC_Coeffs/, called from Data_Reader.cpp:59-89):
 *   line 1: "C T D k"; then C*T lines "re im i_1 .. i_D", coefficient-major,
 * meaning c_j(p) = sum_u scalar[j][u] prod_d p[i_d] with p[k] = 1.
 * With scalar == NULL only the sizes are returned; otherwise scalar holds
 * cap_terms >= C*T entries and params cap_terms*D. */
hcx_status hcx_ref_read_cmap(const char* path, int* C, int* T, int* D, int* k, hcx_complex* scalar, int* params,
                             int cap_terms);

/* PH-type problem (GPUHC_Type: PH, trifocal_2op1p_30x30) -> P2C form.
 * PH tables are transposed, equation index fastest:
 *   ph_ht [ht_terms][ht_parts][n], ph_hx [n var][hx_terms][hx_parts][n eq],
 * parts = {scalar, p_a, p_b, x ...}, parameter index P = the constant 1
 * (dev-eval-indxing-trifocal_2op1p_30x30.cuh:44-168,
 *  kernel_HC_Solver_trifocal_2op1p_30x30.cu:124-139).  Each distinct unordered
 * pair {a, b} other than {P, P} becomes one coefficient p_a * p_b, numbered in
 * order of first use (dH/dt table first, then dH/dx); {P, P} and padding take
 * the constant index C.  Outputs: *C, ht [n][ht_terms][ht_parts-1],
 * hx [n][n][hx_terms][hx_parts-1] in the P2C layouts, pairs [C][2]
 * (cap_pairs >= C).  The coefficient map is then scalar 1, params pairs[j]
 * with k = P (degree 2, max_order_t = 2). */
hcx_status hcx_ref_ph_to_p2c(int n, int P, int ht_terms, int ht_parts, const int* ph_ht, int hx_terms, int hx_parts,
                             const int* ph_hx, int* C, int* ht, int* hx, int* pairs, int cap_pairs);

#ifdef __cplusplus
}
#endif
#endif /* HCX_H */
