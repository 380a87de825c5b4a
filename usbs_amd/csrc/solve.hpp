// USBS outer loop on the device: the scalar state block and the launchers the
// loop driver (solve.hip) calls.
#pragma once
#include "engine.hpp"

namespace usbs {

// Offsets into the per-solve FP64 scalar block (device memory).  Every
// scalar of the reference's StateStruct (solver/usbs.py:467-500) that is
// not a vector lives here, so the loop runs without host round trips.
enum Sc {
  SC_ETA = 0,       // eta from the prox subproblem
  SC_TRS,           // trace(S)
  SC_SUMLAM,        // sum of the first k_curr clipped eigenvalues of S
  SC_ZZ,            // zbar.zbar      (dots3 writes 4 consecutive: ZZ, ZY, ZBU, pad)
  SC_ZY,            // zbar.y
  SC_ZBU,           // zbar.(b - ups)
  SC_PAD0,
  SC_BYC,           // b.ycand        (constraint update writes 4: BYC, ZYC, INF2, INFMAX)
  SC_ZYC,           // zbar.ycand
  SC_INF2,          // |infeas / SCALE_A|^2
  SC_INFMAX,        // max |infeas| / SCALE_A
  SC_OBJF,          // tr(F^T C F)    (objective writes 2: OBJF, OBJFC)
  SC_OBJFC,         // tr(Fc^T C Fc)
  SC_LB,            // lb_spec_est
  SC_TRX,           // tr_X
  SC_TRXB,          // tr_X_bar
  SC_P,             // primal_obj
  SC_PB,            // bar_primal_obj
  SC_PEN,           // pen_dual_obj
  SC_CANDPEN,       // cand_pen_dual_obj
  SC_OBJUB,         // obj_ub
  SC_GAP,           // obj_gap
  SC_INFGAP,        // infeas_gap
  SC_MAXINF,        // max_infeas
  SC_DESCENT,       // 1 if the last step was a descent step
  SC_T,             // iteration counter t
  SC_BY,            // b.y (dual_obj of the log line)
  SC_CB,            // callback value (max-cut rounding)
  SC_IPM_IT,        // IPM iterations of the last quad subproblem
  SC_LB_IT,         // IPM iterations of the last lower-bound subproblem
  SC_UPSGAP,        // upsilon gap of the alternating loop
  SC_NYS_RHO,       // rho of the last Nystrom reconstruction
  SC_LAM0,          // -(smallest eigenvalue of -C + A*(ycand))
  SC_BNORM,         // |b / SCALE_A / SCALE_X|
  SC_NSUBIT,        // alternating-loop iterations
  SC_COUNT = 48
};

int launch_assemble(const Problem& P, double coefC, const double* z1, double a1, const double* z2,
                    double a2, const double* z3, double a3, double* out64, float* out32,
                    hipStream_t s);
// heavy != 0: a wavefront per row (operators whose rows average 32+ entries)
template <typename TV, typename TX>
int launch_spmm(int n, const int* row_ptr, const int* col, const TV* val, const TX* X,
                long long ldx, int p, double* Y, long long ldy, int heavy, hipStream_t s);
int launch_reduce(const double* partial, int ntiles, int count, double scale, double* out,
                  hipStream_t s);
template <typename TA, typename TB>
int launch_tsgram(int n, int tile, int ntiles, const TA* A, long long lda, int p, const TB* B,
                  long long ldb, int q, double* partial, double* out, hipStream_t s);
template <typename TA, typename TC>
int launch_tsmm(int n, const TA* A, long long lda, int p, const double* B, int q, TC* C,
                long long ldc, double beta, const TC* C0, long long ldc0, hipStream_t s);
// Constraint tiles of the Q_base Gram (d = k(k+1)/2): 256 constraints,
// doubled while the d x d partials of more than 256 tiles would exceed 2^20
// doubles.
int qbase_tile_rows(int m, int d);
template <typename TV>
int launch_qbase(const Problem& P, int k, const TV* V, long long ldv, double* partial, double* out,
                 hipStream_t s);
template <typename TF>
int launch_constraint_update(const Problem& P, int k, int kc, const TF* F, long long ldf,
                             const double* eta, const double* zbar, const double* y,
                             const double* ups, double rho, double* z, double* zbn, double* ycand,
                             double* partial, double* out4, hipStream_t s);
template <typename T>
int launch_sketch_update(int n, int k, int kc, int r, const T* F, long long ldf, const double* FO,
                         const double* eta, T* Pbar, T* P, long long ldp, hipStream_t s);
template <typename T>
int launch_objective(const Problem& P, int k, int kc, const T* F, long long ldf, double* partial,
                     double* out2, hipStream_t s);
int launch_dots3(const Problem& P, const double* z, const double* y, const double* ups,
                 double* partial, double* out4, hipStream_t s);
int launch_select(int m, const double* flag, double* y, const double* ycand, hipStream_t s);
template <typename TF>
int launch_upsilon(const Problem& P, int k, const TF* F, long long ldf, const double* eta,
                   const double* zbar, const double* y, double rho, double* ups, double* partial,
                   double* out, hipStream_t s);
// Weights of the cut values sgn^T (C / s) sgn: integer 4 C_ij / s when every
// one is an integer (exact int64 sums), else the FP64 values of C scaled by
// inv_scale = 1 / s at the end.
struct CutWeights {
  const int* cw4 = nullptr;
  int use_int = 0;
  double inv_scale = 1.0;
};
// The uploaded C itself (s = 1).
inline CutWeights cut_weights_of(const Problem& P) {
  CutWeights w;
  w.cw4 = P.cw4;
  w.use_int = P.cut_integer ? 1 : 0;
  return w;
}
// s != 1: integer weights round(4 C_ij / s) built from the host copy of C into
// `pool` when each is within 1e-9 relative of an integer (C / s is the
// unscaled cost of a state whose C was multiplied by SCALE_C = 1 / s), else
// the FP64 path with inv_scale.  Returns 0 or a USBS_ERR_* code.
int make_cut_weights(Ctx* c, std::vector<DevBuf>& pool, double s, CutWeights* out);
// api_ops.hip: E = (P + rho Omega) G (row-major n x r); the QAP rounding on
// the host from E (utils/qap_helpers.py:625-653).
template <typename T>
int launch_nystrom_E(int n, int r, const T* P, const T* Om, long long ld, const double* G,
                     const double* rho, double* E, hipStream_t s);
int qap_round_host(int l, int r, const double* E, const double* W, const double* D, double* values,
                   int* perm, double* best);
template <typename T>
int launch_round(const Problem& P, const CutWeights& cw, int r, const T* Pm, const T* Om, long long ld, const double* G,
                 const double* rho, signed char* sg, long long* ipart, double* dpart, double* out,
                 hipStream_t s);

// dense_kernels.hip
struct IpmArgs {
  int k, kc, max_iters, first;
  double tu, rho, eps;
  const double* Gq;   // d x d  Q_base (unscaled)
  const double* Gz;   // k x k  V^T A*(zbar) V
  const double* Gh;   // k x k  V^T (A*(y - (b - ups)/rho) - C) V
  double* sc;         // scalar block
  double* eta_io;     // warm start eta (in/out), unscaled as the reference's eta
  double* S_io;       // k x k warm start S (in/out)
  double* GF;         // k x k: E_S diag(sqrt(lam)) (F = V GF)
  double* Epast;      // k x (k - kc): E_S[:, kc:]
};
int launch_ipm_quad(const IpmArgs& a, hipStream_t s);
int launch_ipm_lb(int k, int max_iters, double tu, double eps, const double* Glb, double* sc,
                  hipStream_t s);
int launch_nystrom_core(int n, int r, const double* PtPO, const double* OtO, double* G,
                        double* lam, double* sc, hipStream_t s);
struct ScalarArgs {
  double tu, tf, beta, SC, SX;
  const double* lanczos_alpha;  // smallest Ritz value of -C + A*(ycand) at [0]
};
int launch_scalars(const ScalarArgs& a, double* sc, hipStream_t s);

}  // namespace usbs
