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

#define USBS_EXEC_PEROP 0      /* one kernel per phase, host reads control per restart */
#define USBS_EXEC_PERSISTENT 1 /* one cooperative launch per Lanczos call            */

typedef struct usbs_ctx usbs_ctx;

/* Create a context on HIP device `device` (its own stream and memory). */
int usbs_create(usbs_ctx** out, int device);
/* Free the context and every device buffer it owns. */
void usbs_destroy(usbs_ctx* ctx);
/* Message of the last failure on this context ("" if none). */
const char* usbs_last_error(const usbs_ctx* ctx);
/* Library version string. */
const char* usbs_version(void);

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

#ifdef __cplusplus
}
#endif
#endif /* USBS_HIP_H */
