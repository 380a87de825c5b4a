/*
 * usbs_hip.h -- C ABI of the MI355X (gfx950) backend for the USBS hot path.
 *
 * The reference (rangell/usbs) is pure Python on JAX and has no FFI; its
 * replaceable boundary is a set of Python call signatures.  Each entry point
 * below names the reference function whose arithmetic it replaces.  The
 * Python mirror in usbs_amd/ binds these through ctypes behind the
 * reference's own signatures (a trailing backend="hip" argument selects it).
 *
 * Conventions
 *  - All pointers are HOST pointers to plain arrays; the library copies inputs
 *    into device memory it owns (a context) and copies results back.
 *  - Dense (n x p) arrays are row-major, as numpy's default (C order).
 *  - Every function returns USBS_OK (0) or a negative status; the message of
 *    the last failure is available from usbs_last_error().
 *  - A context is used by one host thread at a time and owns one HIP stream.
 */
#ifndef USBS_HIP_H
#define USBS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define USBS_OK 0
#define USBS_ERR_INVALID (-1)   /* bad argument or unsupported size          */
#define USBS_ERR_HIP (-2)       /* a HIP runtime call failed                  */
#define USBS_ERR_OOM (-3)       /* device allocation failed                   */
#define USBS_ERR_NONFINITE (-4) /* non-finite input detected                  */
#define USBS_ERR_STATE (-5)     /* call order (e.g. no problem uploaded)      */
#define USBS_ERR_TIMEOUT (-6)   /* a bounded device-side wait expired         */

#define USBS_PREC_FP64 0  /* plain FP64 everywhere                                  */
#define USBS_PREC_MIXED 1 /* FP32 storage for Lanczos / sketch, FP64 accumulation,   */
                          /* FP64 Ritz solve, IPM, duals and convergence test       */
#define USBS_PREC_BF16 2  /* usbs_qap_kron_matvec only: BF16 operands, FP32 accum.  */

#define USBS_EXEC_PEROP 0      /* one kernel per phase, host reads control per restart;
                                  usbs_solve and usbs_cgal replay each restart's
                                  launches from a HIP graph                          */
#define USBS_EXEC_PERSISTENT 1 /* one cooperative launch per Lanczos call            */
#define USBS_EXEC_AUTO 2       /* persistent where every row tile fits in LDS for a
                                  restart (light rows, <= 256-row tiles) and there
                                  are no more tiles than compute units, else per-op */
#define USBS_EXEC_TUNE 3       /* usbs_solve: time the first Lanczos call in both
                                  forms (same results) and keep the faster; the
                                  choice is kept per context until the next problem;
                                  elsewhere as USBS_EXEC_AUTO                       */

typedef struct usbs_ctx usbs_ctx;

/* Create a context on HIP device `device` (its own stream and memory). */
int usbs_create(usbs_ctx** out, int device);
/* Free the context and every device buffer it owns. */
void usbs_destroy(usbs_ctx* ctx);
/* Message of the last failure on this context ("" if none). */
const char* usbs_last_error(const usbs_ctx* ctx);
/* Library version string. */
const char* usbs_version(void);
/* Free and total bytes of the context's device (for admission by footprint;
 * no counterpart in the reference). */
int usbs_device_memory(usbs_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);

/*
 * Upload an SDP in the reference's SDPState layout (utils/common.py:11-26):
 *   C: n x n sparse cost in COO form (both triangles, duplicates summed), the
 *      BCOO (indices, data) pair of SDPState.C;
 *   A_indices: nA x 3 int32 row-major (constraint, row, col), A_data: nA;
 *   b, b_ineq_mask, SCALE_A: m.
 * Replaces the device placement XLA does for SDPState (no reference line).
 */
int usbs_set_problem(usbs_ctx* ctx, int64_t n, int64_t m, int64_t nnzC, const int32_t* C_rows,
                     const int32_t* C_cols, const double* C_vals, int64_t nA,
                     const int32_t* A_indices, const double* A_data, const double* b,
                     const double* b_ineq_mask, const double* SCALE_A);

/*
 * Smallest eigenpairs of C_coef*C + A*(adjoint_left_vec) by thick-restart
 * Lanczos.  Replaces eigsh_smallest (solver/lanczos.py:239-284); the
 * reference's callers pass C_arg = -C, i.e. C_coef = -1 (solver/usbs.py:590).
 * eigvals: num_desired; eigvecs: n x num_desired row-major.
 * info (may be NULL): [restarts, num_converged, lanczos_steps, breakdowns].
 */
int usbs_lanczos_smallest(usbs_ctx* ctx, double C_coef, const double* adjoint_left_vec,
                          const double* q0, int num_desired, int inner_iterations,
                          int max_restarts, double tolerance, int precision, int exec_mode,
                          double* eigvals, double* eigvecs, int* info);


/*
 * Batched form of usbs_lanczos_smallest: ninst independent instances, one per
 * context (each with its own uploaded problem, sizes may differ; all on one
 * device), advanced together in one persistent launch per restart (Ritz and
 * rotation kernels cover every instance).  Each instance's results are
 * bitwise the same as its own usbs_lanczos_smallest call with
 * USBS_EXEC_PERSISTENT.  Arrays are per instance; info holds 4 ints per
 * instance.  No counterpart in the reference, which solves one problem per
 * call (solver/lanczos.py:239-284 is the per-instance arithmetic).
 */
int usbs_lanczos_smallest_batch(usbs_ctx* const* ctxs, int ninst, double C_coef,
                                const double* const* adjoint_left_vecs, const double* const* q0s,
                                int num_desired, int inner_iterations, int max_restarts,
                                double tolerance, int precision, double* const* eigvals,
                                double* const* eigvecs, int* info);

/*
 * The lifted QAP's cost operator on a batch of vectors: U_b = W Y_b D^T for
 * b < nvec, i.e. C v for C = W (x) D padded by an empty leading row and
 * column (utils/qap_helpers.py:157-172, the cost whose matvec the reference
 * runs inside the operator at solver/lanczos.py:86), with Y_b[c][j] =
 * v_b[1 + l c + j] and (C v_b)[1 + l a + i] = U_b[a][i], (C v_b)[0] = 0.
 * W, D: l x l row-major; Y, U: nvec x l x l.  precision USBS_PREC_BF16: BF16
 * operands on the matrix cores with FP32 accumulation (l <= 64);
 * USBS_PREC_FP64: FP64 FMA (l <= 48).  The product runs reps times on
 * device-resident inputs; kernel_ms (may be NULL) receives the mean time of
 * one launch.
 */
int usbs_qap_kron_matvec(usbs_ctx* ctx, int32_t l, const double* W, const double* D,
                         int64_t nvec, const double* Y, int32_t precision, int32_t reps,
                         double* U, double* kernel_ms);

/*
 * The constraint operator A(.) on the columns of U, summed over them
 * (apply_A_operator_batched, solver/utils.py:27-31; apply_A_operator_slim,
 * utils.py:11-16, is p = 1):
 *   out[c] = sum_t [A_indices[t][0] == c] A_data[t] sum_q U[i_t][q] U[j_t][q],
 * with (c, i_t, j_t) = A_indices[t] (nA x 3 int32 row-major), U n x p row-major,
 * out m.  Needs no uploaded problem.
 */
int usbs_apply_A_operator(usbs_ctx* ctx, int64_t n, int64_t m, int64_t nA,
                          const int32_t* A_indices, const double* A_data, int32_t p,
                          const double* U, double* out);

/*
 * The adjoint A*(z) applied to the columns of U (apply_A_adjoint_batched,
 * solver/utils.py:34-37; apply_A_adjoint_slim, utils.py:19-24, is p = 1):
 *   out[i][q] = sum_t [i_t == i] A_data[t] z[c_t] U[j_t][q];
 * U and out n x p row-major, z m.  Needs no uploaded problem.
 */
int usbs_apply_A_adjoint(usbs_ctx* ctx, int64_t n, int64_t m, int64_t nA,
                         const int32_t* A_indices, const double* A_data, const double* z,
                         int32_t p, const double* U, double* out);

/* Parameters of usbs() (solver/usbs.py:436-459), plus the execution choices. */
typedef struct usbs_params {
  int32_t k_curr, k_past;
  int32_t max_iters;
  int32_t lanczos_inner_iterations, lanczos_max_restarts;
  int32_t subprob_max_iters;
  int32_t sketch_dim;  /* r > 0: Nystrom sketch P (n x r); 0: diag-only P (n);       */
                       /* -1: dense X (n x n, FP64; P0 / P_out carry X0 / X)        */
  int32_t rounding;    /* 1: max-cut rounding at every condition check (needs r > 0); */
                       /* 2: QAP rounding (needs usbs_set_qap_rounding, n = l^2 + 1) */
  int32_t precision;   /* USBS_PREC_*                                               */
  int32_t exec_mode;   /* USBS_EXEC_* for the Lanczos calls                          */
  int32_t warm;        /* 1: start from y, z and P the context kept from its last    */
                       /*    usbs_solve (device-resident; y0, z0, P0 are ignored);    */
                       /*    z and P are taken from that solve's SCALE_X to this one's */
                       /*    (SCALE_A must be the same), y carries over               */
  int32_t reserved;
  double trace_ub, trace_factor, rho, beta;
  double max_time, obj_gap_eps, infeas_gap_eps, max_infeas_eps;
  double subprob_eps, cond_exp_base;
  double SCALE_C, SCALE_X;
  double cut_scale;    /* rounding reports sgn^T (C / cut_scale) sgn for the uploaded */
                       /* (scaled) C; the reference's callers pass C / SCALE_C      */
                       /* (scripts/warm_start_maxcut_usbs.py:121), i.e. SCALE_C;    */
                       /* 0 means SCALE_C                                           */
  double lanczos_rtol; /* Ritz pairs count as converged below                       */
                       /* max(subprob_eps, lanczos_rtol * max|theta|); 0: the       */
                       /* reference's absolute test (solver/lanczos.py:205-208)     */
} usbs_params;

/* One condition-check record: the fields of the reference's log line
 * (solver/usbs.py:514-530) plus solver statistics. */
typedef struct usbs_record {
  double t, time, pen_dual_obj, cand_pen_dual_obj, lb_spec_est, primal_obj, dual_obj, obj_ub;
  double obj_gap, infeas_gap, max_infeas, callback_val;
  double ipm_iters, lb_iters, lanczos_restarts, descent;
} usbs_record;

/*
 * The whole USBS loop on the device (solver/usbs.py:436-789): cold or warm
 * start from (y0, z0, tr_X0, primal_obj0, P0), or, with params->warm, from the
 * y, z and P that this context kept on the device at the end of its previous
 * solve (same n, m and sketch size; the problem's values may have changed, as
 * in utils/maxcut_helpers.py:95-161 where y, z, P carry over and only the
 * scalars tr_X0, primal_obj0 are recomputed by the caller).  Omega the sketch
 * test matrix
 * (n x r row-major, NULL for diag-only), q0 the Lanczos start vector
 * (solver/usbs.py:691-692).  Outputs y_out (m), z_out (m), P_out (n x r or n),
 * scal_out = [tr_X, primal_obj, iterations, obj_gap, infeas_gap, max_infeas].
 * Every condition check appends a record to log (up to log_cap).
 */
int usbs_solve(usbs_ctx* ctx, const usbs_params* params, const double* y0, const double* z0,
               double tr_X0, double primal_obj0, const double* Omega, const double* P0,
               const double* q0, double* y_out, double* z_out, double* P_out, double* scal_out,
               usbs_record* log, int log_cap, int* log_len);

/* One solve of a usbs_solve_batch call: its context (with its problem
 * uploaded), its parameters, and the arguments and outputs of usbs_solve. */
typedef struct usbs_batch_item {
  usbs_ctx* ctx;
  const usbs_params* params;
  const double *y0, *z0;
  double tr_X0, primal_obj0;
  const double *Omega, *P0, *q0;
  double *y_out, *z_out, *P_out, *scal_out;
  usbs_record* log;
  int32_t log_cap;
  int32_t* log_len;
} usbs_batch_item;

/*
 * count usbs_solve calls on one device in lockstep, for throughput when many
 * problems are solved at once (the reference solves one problem per call,
 * solver/usbs.py:436-789; each item here gives that call's results).  Each
 * solve runs on its own context and stream; per iteration one persistent
 * (cooperative) Lanczos launch per restart serves every solve still running,
 * and each solve's outputs and records are bitwise those of usbs_solve with
 * exec_mode USBS_EXEC_PERSISTENT (record times aside).  The items must use
 * distinct contexts on one device and share precision, k_curr, k_past,
 * lanczos_inner_iterations and lanczos_max_restarts; exec_mode must be
 * USBS_EXEC_AUTO or USBS_EXEC_PERSISTENT.  A solve leaves the batch when its
 * stop test holds.  On an error every item's context holds the message,
 * which names the failing item.
 */
int usbs_solve_batch(const usbs_batch_item* items, int32_t count);

/* Parameters of cgal() (solver/cgal.py:21-39), plus the execution choices. */
typedef struct usbs_cgal_params {
  int32_t max_iters;
  int32_t lanczos_inner_iterations, lanczos_max_restarts;
  int32_t sketch_dim;  /* r > 0: Nystrom sketch P (n x r) of the primal             */
  int32_t rounding;    /* 1: max-cut rounding at every condition check              */
  int32_t precision;   /* USBS_PREC_*                                               */
  int32_t exec_mode;   /* USBS_EXEC_* for the Lanczos calls                          */
  int32_t reserved;
  double trace_ub, beta0;
  double max_time, obj_gap_eps, infeas_gap_eps, max_infeas_eps;
  double subprob_eps, cond_exp_base;
  double SCALE_C, SCALE_X;
  double cut_scale;    /* as usbs_params.cut_scale                                  */
} usbs_cgal_params;

/*
 * The whole CGAL loop on the device (solver/cgal.py:21-243): rank-one
 * Frank-Wolfe steps from the smallest eigenvector of -C + A*(y + beta (z -
 * proj_b)), sketch P = (1 - eta) P + eta trace_ub u (u^T Omega), dual step
 * clipped to [0, beta0].  Same arguments and outputs as usbs_solve; records
 * carry t, time, primal_obj, obj_gap, infeas_gap, max_infeas, callback_val
 * (the fields of the reference's log line, cgal.py:77-88) and NaN elsewhere.
 * The sketch is required (the reference's dense-X mode is not provided).
 */
int usbs_cgal_solve(usbs_ctx* ctx, const usbs_cgal_params* params, const double* y0,
                    const double* z0, double tr_X0, double primal_obj0, const double* Omega,
                    const double* P0, const double* q0, double* y_out, double* z_out,
                    double* P_out, double* scal_out, usbs_record* log, int log_cap, int* log_len);

/*
 * Nystrom fixed-rank reconstruction from the sketch (utils/common.py:99-116):
 * E (n x r row-major), Lambda (r), singular values in descending order.  A
 * core that is not positive definite (e.g. P = 0) yields NaN, as XLA does.
 */
int usbs_reconstruct_from_sketch(usbs_ctx* ctx, int64_t n, int32_t r, const double* Omega,
                                 const double* P, double* E, double* Lambda);

/*
 * Max-cut rounding callback (utils/maxcut_helpers.py:292-302) against the
 * uploaded C: values[j] = sgn_j^T C sgn_j with sgn = 2 (E > 0) - 1 from the
 * Nystrom factor; computed as an exact integer sum when every 4*C_ij is an
 * integer.  signs (n x r row-major, may be NULL) returns the sign vectors.
 */
int usbs_maxcut_round(usbs_ctx* ctx, int32_t r, const double* P, const double* Omega,
                      double* values, int8_t* signs);

/*
 * Sketch update (solver/usbs.py:580, 636; diag-only form usbs.py:574, 630
 * when r == 0): P = eta Pbar + F (F^T Omega), Pbar' = eta Pbar + F_c (F_c^T
 * Omega) with F_c the first kc columns.  F: n x k, Omega, Pbar: n x r
 * row-major.  precision selects FP64 or FP32 storage.
 */
int usbs_sketch_update(usbs_ctx* ctx, int64_t n, int32_t k, int32_t kc, int32_t r,
                       const double* F, const double* Omega, const double* Pbar, double eta,
                       int32_t precision, double* P_out, double* Pbar_out);

/*
 * The sketch update of usbs_sketch_update on caller-owned DEVICE arrays,
 * allocating nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops that forbid allocation).  F (n x k), Omega, Pbar, P_out, Pbar_out
 * (n x r; one column when r == 0, the diag-only form) are column-major with
 * leading dimension ld >= n, doubles for USBS_PREC_FP64 and floats for
 * USBS_PREC_MIXED; Pbar is left unchanged.  Intermediates live in a fixed
 * 2 MiB device array per GPU: sizes whose row-tile partials of F^T Omega do
 * not fit (ntiles * k * r + 1024 doubles) are refused with USBS_ERR_INVALID.
 * Runs on the context's stream and returns when done; calls that overlap on
 * one device share that array and race.
 */
int usbs_sketch_update_noalloc(usbs_ctx* ctx, int64_t n, int32_t k, int32_t kc, int32_t r,
                               int64_t ld, const void* F, const void* Omega, const void* Pbar,
                               double eta, int32_t precision, void* P_out, void* Pbar_out);

/*
 * Primal objective, constraint map, dual candidate and infeasibility sums of
 * one USBS iteration (solver/usbs.py:571-585, 637-652) for the uploaded
 * problem: z = eta zbar + A(F F^T), zbar' = eta zbar + A(F_c F_c^T),
 * ycand = y - (b - z - upsilon) / rho;  out6 = [tr(F^T C F), tr(F_c^T C F_c),
 * b.ycand, zbar.ycand, |infeas/SCALE_A|^2, max |infeas|/SCALE_A].
 */
int usbs_objective_infeas(usbs_ctx* ctx, int32_t k, int32_t kc, const double* F, double eta,
                          const double* zbar, const double* y, const double* upsilon, double rho,
                          double* z_out, double* zbar_next_out, double* ycand_out, double* out6);

/*
 * Minimum-cost assignment of an n x n cost matrix (row-major) by the
 * Hungarian algorithm with the reference's step order, tie-breaking and
 * n^2-step bound (munkres, utils/munkres.py:14-120): mask (n x n row-major)
 * receives the reference's result matrix, 1 at the assigned entries (a
 * mask left incomplete by the step bound is returned as the reference
 * returns it).  Host-only: needs no context and no GPU.
 */
int usbs_munkres(int32_t n, const double* cost, int32_t* mask);

/*
 * QAP rounding (qap_round, utils/qap_helpers.py:625-653) of a sketch of the
 * lifted QAP's X (N = l^2 + 1 rows, r columns, row-major): the Nystrom factor
 * E is formed on the device; for each column i < l (clamped to r - 1) and its
 * negative, the assignment minimising max(c) - c for c = E[1:, i] as an l x l
 * matrix is priced by trace(W Pi D Pi^T).  values (may be NULL): 2 l
 * candidate objectives (positive columns, then negative); perm (may be NULL):
 * the best assignment (-1 for a row the assignment left open); best: the
 * smaller minimum.
 */
int usbs_qap_round(usbs_ctx* ctx, int32_t l, int32_t r, const double* P, const double* Omega,
                   const double* W, const double* D, double* values, int32_t* perm,
                   double* best);

/* W (flow) and D (distance), l x l row-major, for usbs_solve's QAP rounding
 * (usbs_params.rounding = 2): the callback_nonstatic_args {"W", "D"} of the
 * reference's QAP scripts (scripts/warm_start_qap_usbs.py). */
int usbs_set_qap_rounding(usbs_ctx* ctx, int32_t l, const double* W, const double* D);

/*
 * Live kernel timing (measurement support; no counterpart in the reference).
 * While enabled, usbs_solve brackets the Lanczos step kernels of every
 * restart with HIP events on the context's stream -- the one cooperative
 * launch in the persistent form, the restart's step launches (gaps included)
 * in the per-op form -- and accumulates the duration of the restarts that ran
 * at least one step, with their step counts, so that a caller can divide
 * algorithmic bytes by measured kernel time.
 */
typedef struct usbs_stats {
  double lz_kernel_ms;  /* summed duration of those launches                        */
  int64_t lz_launches;  /* number of launches that ran steps                         */
  int64_t lz_steps;     /* Lanczos steps they ran                                    */
  int64_t lz_sum_cols;  /* sum over those steps of the valid basis column count i+1  */
  int64_t n, nnz;       /* operator size of the uploaded problem (merged pattern)     */
  int32_t value_bytes;  /* bytes per stored basis / operator value of the last solve */
  int32_t lz_form;      /* USBS_EXEC_PEROP or _PERSISTENT: the form the timed calls ran */
} usbs_stats;

/* Enable (1) or disable (0) the timing; enabling resets the counters. */
int usbs_set_timing(usbs_ctx* ctx, int enable);
/* Copy the counters accumulated since timing was enabled. */
int usbs_get_stats(const usbs_ctx* ctx, usbs_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* USBS_HIP_H */
