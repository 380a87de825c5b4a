// The USBS outer loop (solver/usbs.py:436-789) driven from the host with all
// state resident on the device.  One body = the kernel sequence below; the
// host reads the scalar block back only at condition checks (every body, or
// on the geometric schedule of utils/loop.py:85-98).
//
//   dots3 / assemble+SpMM+Gram (h_1, q_12) / Q_base Gram   IPM inputs
//   IPM (one workgroup) + eigh(S) -> eta, S, F = V E sqrt(lam)
//   [upsilon alternation for inequality rows]
//   constraint update: z, zbar_next, ycand, b.ycand, zbar.ycand, infeasibility
//   sketch F^T Omega + P/Pbar update;  objective tr(F^T C F)
//   assemble -C + A*(ycand) -> Lanczos (per-op or persistent)
//   SpMM+Gram (g_1) -> lower-bound IPM -> scalar bookkeeping -> y select
//   V_next = orth([V E_S[:, kc:], cand_vecs])
//   condition: Nystrom rounding (max-cut) + gaps
#include <algorithm>
#include <chrono>
#include <memory>
#include <cmath>
#include <vector>

#include "../../include/usbs_hip.h"
#include "solve.hpp"
#include "solve_util.hpp"

namespace usbs {

// Penalised dual objective of the initial point (solver/usbs.py:704-706).
__global__ void k_init_pen(double* sc, const double* by4, const double* alpha, double tu) {
  if (threadIdx.x != 0) return;
  const double lam = -alpha[0];
  const double pen = by4[1] + tu * (lam < 0.0 ? 0.0 : lam);
  sc[SC_PEN] = pen;
  sc[SC_CANDPEN] = pen;
  sc[SC_BY] = by4[1];
}

__global__ void k_set_by(double* sc, const double* by4) {
  if (threadIdx.x == 0) sc[SC_BY] = by4[1];
}

// the early rounding's two scalars (its own block) into the scalar block
__global__ void k_copy_round(double* sc, const double* rsc) {
  if (threadIdx.x == 0) {
    sc[SC_NYS_RHO] = rsc[SC_NYS_RHO];
    sc[SC_CB] = rsc[SC_CB];
  }
}

__global__ void k_copy1(double* dst, const double* src) {
  if (threadIdx.x == 0) *dst = *src;
}

// Dense-X storage (solver/usbs.py:575-576, 631-632):
// X[i, j] = eta Xbar[i, j] + sum_{a < k} A[i, a] B[j, a]  (X, Xbar: n x n
// column-major with leading dimension ldx, FP64; A, B: n x k column-major
// with leading dimension ld).  X may be Xbar (each entry is read and then
// written by one thread).  A block covers 256 rows and 32 columns; the 32
// rows of B it needs sit in LDS.
template <typename T>
__global__ void __launch_bounds__(256)
k_dense_rankk(int n, int k, const double* __restrict__ eta_ptr, const double* Xbar, double* X,
              long long ldx, const T* __restrict__ A, const T* __restrict__ B, long long ld) {
  __shared__ double Bs[16][32];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j0 = blockIdx.y * 32;
  for (int idx = threadIdx.x; idx < 16 * 32; idx += 256) {
    const int a = idx / 32, jj = idx % 32;
    Bs[a][jj] = (a < k && j0 + jj < n) ? (double)B[(j0 + jj) + (long long)a * ld] : 0.0;
  }
  __syncthreads();
  if (i >= n) return;
  double ai[16];
#pragma unroll
  for (int a = 0; a < 16; ++a) ai[a] = a < k ? (double)A[i + (long long)a * ld] : 0.0;
  const double eta = *eta_ptr;
  const int jn = min(32, n - j0);
  for (int jj = 0; jj < jn; ++jj) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 16; ++a) s += ai[a] * Bs[a][jj];
    const long long e = i + (long long)(j0 + jj) * ldx;
    X[e] = eta * Xbar[e] + s;
  }
}

template <typename T>
struct Solver {
  Ctx* c = nullptr;
  const Problem* P = nullptr;
  usbs_params p{};
  hipStream_t s = nullptr;
  int n = 0, m = 0, k = 0, kc = 0, kp = 0, r = 0, d = 0, mL = 0;
  long long ld = 0;
  std::vector<DevBuf> pool;
  double *y = nullptr, *z = nullptr, *zbar = nullptr, *zbn = nullptr, *ycand = nullptr,
         *ups = nullptr, *tmp4 = nullptr;
  T *Qa = nullptr, *Qb = nullptr, *F = nullptr, *PO = nullptr, *Pbar = nullptr;
  double* W = nullptr;
  double* op64 = nullptr;
  float* op32 = nullptr;
  double *partial = nullptr, *dpart = nullptr;
  long long* ipart = nullptr;
  double *Gq = nullptr, *Gz = nullptr, *Gh = nullptr, *Glb = nullptr, *FO = nullptr, *GF = nullptr,
         *Epast = nullptr, *S_io = nullptr, *eta_io = nullptr, *PtPO = nullptr, *OtO = nullptr,
         *Gn = nullptr, *lam = nullptr, *cut = nullptr, *sc = nullptr, *q0 = nullptr;
  signed char* sg = nullptr;
  CutWeights cutw{};       // weights of the rounding's cut values (C / cut_scale)
  bool dense = false;      // dense-X storage (sketch_dim = -1): X, Xbar n x n FP64
  double *X = nullptr, *Xbar = nullptr;
  T* VS = nullptr;         // V S (n x k), the factor of the X update
  double* dE = nullptr;    // QAP rounding: the Nystrom factor (n x r row-major)
  std::vector<double> hE;  // ... its host copy
  double qap_cb = NAN;     // ... the last rounded QAP objective
  LzInst<T> lz{}, qr{};
  LzInst<T>* d_lz_k = nullptr;   // device copies for the persistent form (nd = k, nd = kc)
  LzInst<T>* d_lz_kc = nullptr;
  int* d_map = nullptr;
  int nb = 1;
  // the bundle QR as one cooperative launch (qr_mode 1: persistent single
  // solves whose tiles are resident row-in-registers tiles): descriptors of
  // qr over each bundle buffer (d_qr[0] over qr_buf0, d_qr[1] over the
  // other); -1 not planned yet, 0 per-op launches.  Batch members keep the
  // per-op launches (their streams run side by side).
  LzInst<T>* d_qr = nullptr;
  T* qr_buf0 = nullptr;
  int qr_mode = -1;
  bool batch_member = false;
  size_t cache_bytes = 0;            // LDS per persistent workgroup when tiles are cached
  int host_ctrl[LZ_NCTRL] = {0};
  int lz_restarts = 0;
  LzTiming ptm{};          // per-op step timing (usbs_set_timing)
  bool use_ptm = false;
  bool tuning = false;     // USBS_EXEC_TUNE: the next Lanczos call picks the form
  LzGraphs graphs[2];      // per-op restarts replayed as HIP graphs (nd = k, nd = kc)
  int lz_hint = 2;        // restarts the last persistent call needed: its first early-stop check
  int tev_launched = 0;   // restarts launched by the last timed persistent call
  hipStream_t s2 = nullptr;          // side stream: the lower-bound IPM overlaps the Lanczos call
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // recorded on the main stream after the candidate operator -C + A*(ycand)
  // is assembled into op64; the side stream's lower-bound projection reads
  // op64 and waits for it
  hipEvent_t ev_asm = nullptr;
  std::vector<hipEvent_t> tev;       // timing: 2 events per persistent launch
  bool tev_pending = false;
  std::vector<int> lstat_h;
  // the scalar block and the Lanczos control words of a record, copied into
  // the context's pinned host buffers behind ev_rec (record_launch /
  // record_finish)
  double* rec_h = nullptr;
  int* ctrl_h = nullptr;
  hipEvent_t ev_rec = nullptr;
  // max-cut rounding launched early: when the driver sets round_early
  // before body(), body_pre queues the rounding of the coming condition
  // check on the side stream right after the sketch update (its inputs, the
  // sketch and Omega, do not change again in the iteration), with its own
  // partial-sum buffer; it overlaps the candidate Lanczos call, and
  // cond_launch waits for ev_round instead of running it.
  // Its scalars go to rsc (a block of SC_COUNT doubles), copied into sc when
  // the check runs, so the two streams never write the same block.
  bool round_early = false, round_done = false;
  double* partial_r = nullptr;
  double* rsc = nullptr;
  hipEvent_t ev_round = nullptr;

  ~Solver() {
    if (s2) (void)hipStreamSynchronize(s2);  // the context's side stream: kept
    if (ev_rec) { (void)hipEventSynchronize(ev_rec); (void)hipEventDestroy(ev_rec); }
    if (ev_round) (void)hipEventDestroy(ev_round);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_asm) (void)hipEventDestroy(ev_asm);
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    if (ptm.e0) (void)hipEventDestroy(ptm.e0);
    for (auto& g : graphs) lanczos_graphs_clear(&g);
    if (ptm.e1) (void)hipEventDestroy(ptm.e1);
    ctx_free_pool(pool);
  }

  // Fold the timed launches of the last persistent Lanczos call into the
  // context's counters (waits for that call to finish).
  int collect_timing() {
    if (!tev_pending) return USBS_OK;
    tev_pending = false;
    const int R = tev_launched;
    if (R <= 0) return USBS_OK;
    if (hipEventSynchronize(tev[2 * R - 1]) != hipSuccess ||
        hipMemcpy(lstat_h.data(), lz.lstat, sizeof(int) * 2 * R, hipMemcpyDeviceToHost) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "reading Lanczos timing");
    for (int q = 0; q < R; ++q) {
      if (lstat_h[2 * q] <= 0) continue;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, tev[2 * q], tev[2 * q + 1]) != hipSuccess)
        return set_err(c, USBS_ERR_HIP, "event timing");
      c->stats.lz_kernel_ms += ms;
      c->stats.lz_launches += 1;
      c->stats.lz_steps += lstat_h[2 * q];
      c->stats.lz_sum_cols += lstat_h[2 * q + 1];
    }
    return USBS_OK;
  }

  template <typename X>
  X* alloc(size_t count) {
    return (X*)ctx_alloc(c, pool, sizeof(X) * (count ? count : 1));
  }
  T* V() const { return Qa + ld; }      // current bundle basis = Qa[:, 1..k]
  T* P_() const { return PO; }
  T* Om() const { return PO + (long long)r * ld; }

  int setup(Ctx* ctx, const usbs_params& prm) {
    c = ctx;
    P = &ctx->P;
    p = prm;
    s = ctx->stream;
    n = P->n;
    m = P->m;
    kc = p.k_curr;
    kp = p.k_past;
    k = kc + kp;
    dense = p.sketch_dim == -1;
    r = dense ? 0 : p.sketch_dim;
    d = k * (k + 1) / 2;
    mL = p.lanczos_inner_iterations;
    ld = ((long long)n + 63) / 64 * 64;
    const int ntl = P->ntiles > P->mntiles ? P->ntiles : P->mntiles;
    if (!(y = alloc<double>(m)) || !(z = alloc<double>(m)) || !(zbar = alloc<double>(m)) ||
        !(zbn = alloc<double>(m)) || !(ycand = alloc<double>(m)) || !(ups = alloc<double>(m)) ||
        !(tmp4 = alloc<double>(8)))
      return USBS_ERR_OOM;
    if (!(Qa = alloc<T>(ld * (k + 1))) || !(Qb = alloc<T>(ld * (k + 1))) || !(F = alloc<T>(ld * k)) ||
        !(W = alloc<double>(ld * k)))
      return USBS_ERR_OOM;
    const int rr = r > 0 ? r : 1;
    if (!(PO = alloc<T>(ld * 2 * rr)) || !(Pbar = alloc<T>(ld * rr))) return USBS_ERR_OOM;
    if (!(op64 = alloc<double>(P->nnz))) return USBS_ERR_OOM;
    if (dense && (!(X = alloc<double>((size_t)ld * n)) || !(Xbar = alloc<double>((size_t)ld * n)) ||
                  !(VS = alloc<T>((size_t)ld * k))))
      return USBS_ERR_OOM;
    if (sizeof(T) == 4 && !(op32 = alloc<float>(P->nnz))) return USBS_ERR_OOM;
    const int qt = qbase_tile_rows(m, d);
    const size_t pw = std::max(
        std::max((size_t)P->ntiles * (size_t)std::max(std::max(k * k, 2 * rr * rr), std::max(k * rr, 16)),
                 (size_t)P->mntiles * 4),
        std::max((size_t)((m + qt - 1) / qt) * d * d, (size_t)objective_ntiles(*P) * 16));
    if (!(partial = alloc<double>(pw)) || !(partial_r = alloc<double>(pw)) || !(rsc = alloc<double>(SC_COUNT)) ||
        !(dpart = alloc<double>((size_t)ntl * rr)) ||
        !(ipart = alloc<long long>((size_t)ntl * rr)))
      return USBS_ERR_OOM;
    if (!(Gq = alloc<double>(d * d)) || !(Gz = alloc<double>(k * k)) || !(Gh = alloc<double>(k * k)) ||
        !(Glb = alloc<double>(k * k)) || !(FO = alloc<double>(k * rr)) || !(GF = alloc<double>(k * k)) ||
        !(Epast = alloc<double>(k * (kp > 0 ? kp : 1))) || !(S_io = alloc<double>(k * k)) ||
        !(eta_io = alloc<double>(1)) || !(PtPO = alloc<double>(2 * rr * rr)) ||
        !(OtO = alloc<double>(rr * rr)) || !(Gn = alloc<double>(rr * rr)) || !(lam = alloc<double>(rr)) ||
        !(cut = alloc<double>(rr + 1)) || !(sc = alloc<double>(SC_COUNT)) || !(q0 = alloc<double>(n)) ||
        !(sg = alloc<signed char>((size_t)n * rr)))
      return USBS_ERR_OOM;
    if (p.rounding == 1 && r > 0) {
      const int rcw = make_cut_weights(c, pool, p.cut_scale > 0.0 ? p.cut_scale : p.SCALE_C, &cutw);
      if (rcw) return rcw;
    }
    if (p.rounding == 2) {
      if (c->qap_l < 1 || (long long)c->qap_l * c->qap_l + 1 != n || r < 1)
        return set_err(c, USBS_ERR_STATE, "QAP rounding needs usbs_set_qap_rounding with l^2 + 1 = n "
                                          "and a sketch");
      if (!(dE = alloc<double>((size_t)n * r))) return USBS_ERR_OOM;
      hE.resize((size_t)n * r);
    }
    int rc = make_lanczos<T>(c, pool, *P, mL, k, p.lanczos_max_restarts, p.subprob_eps,
                             sizeof(T) == 4 ? (const T*)op32 : (const T*)op64, lz);
    if (rc) return rc;
    lz.rtol = p.lanczos_rtol;
    rc = make_lanczos<T>(c, pool, *P, k, k, 1, 0.0, (const T*)nullptr, qr);
    qr.exact_norm = 1;
    if (rc) return rc;
    if (lz.ldq != ld || qr.ldq != ld) return set_err(c, USBS_ERR_INVALID, "leading dimension mismatch");
    if ((!c->side && hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) ||
        hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ev_asm, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ev_rec, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ev_round, hipEventDisableTiming) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "stream/event creation failed");
    s2 = c->side;
    if ((!c->rec_h && hipHostMalloc((void**)&c->rec_h, sizeof(double) * SC_COUNT) != hipSuccess) ||
        (!c->ctrl_h && hipHostMalloc((void**)&c->ctrl_h, sizeof(int) * (LZ_NCTRL + 1)) != hipSuccess))
      return set_err(c, USBS_ERR_OOM, "pinned host buffers for the records");
    rec_h = c->rec_h;
    ctrl_h = c->ctrl_h;
    tuning = p.exec_mode == USBS_EXEC_TUNE;
    if (c->timing) {
      if (p.exec_mode == USBS_EXEC_PERSISTENT || tuning) {
        tev.assign(2 * lz.max_restarts, nullptr);
        lstat_h.assign(2 * lz.max_restarts, 0);
        for (auto& e : tev)
          if (hipEventCreate(&e) != hipSuccess) return set_err(c, USBS_ERR_HIP, "event creation failed");
      }
      if (p.exec_mode == USBS_EXEC_PEROP || tuning) {
        if (hipEventCreate(&ptm.e0) != hipSuccess || hipEventCreate(&ptm.e1) != hipSuccess)
          return set_err(c, USBS_ERR_HIP, "event creation failed");
        use_ptm = !tuning;
      }
      c->stats.n = n;
      c->stats.nnz = P->nnz;
      c->stats.value_bytes = (int)sizeof(T);
      c->stats.lz_form = p.exec_mode;
    }
    if (p.exec_mode == USBS_EXEC_PERSISTENT || tuning) {
      if (lanczos_plan_persistent<T>(lz, c->device, c->persist_wgs, &cache_bytes))
        return set_err(c, USBS_ERR_HIP, "no persistent Lanczos launch fits the device");
      nb = lz.nblocks;
      if (!(d_lz_k = alloc<LzInst<T>>(1)) || !(d_lz_kc = alloc<LzInst<T>>(1)) || !(d_map = alloc<int>(2 * (nb + lz.ntiles))))
        return USBS_ERR_OOM;
      std::vector<int2> map(nb + lz.ntiles);
      for (int i = 0; i < nb; ++i) map[i] = make_int2(0, i);
      for (int i = 0; i < lz.ntiles; ++i) map[nb + i] = make_int2(0, i);
      LzInst<T> a = lz, b = lz;
      a.nd = k;
      b.nd = kc;
      if (hipMemcpy(d_lz_k, &a, sizeof(a), hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(d_lz_kc, &b, sizeof(b), hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(d_map, map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice) != hipSuccess)
        return set_err(c, USBS_ERR_HIP, "upload of Lanczos descriptors failed");
    }
    return USBS_OK;
  }

#define CK(expr)                                                                    \
  do {                                                                              \
    int _rc = (expr);                                                               \
    if (_rc) return set_err(c, USBS_ERR_HIP, std::string(#expr) + " -> " +         \
                                                 hipGetErrorString((hipError_t)_rc)); \
  } while (0)

  int fill(double* ptr, long long count, double v) {
    k_fill<double><<<(unsigned)((count + 255) / 256), 256, 0, s>>>(ptr, count, v);
    return (int)hipGetLastError();
  }

  // USBS_EXEC_TUNE: run the first Lanczos call in both forms once untimed
  // (the per-op run records its graphs), then again in both forms timed with
  // events on the stream (same inputs, bitwise the same results), and keep
  // the faster form for the rest of the solve (and for later solves of this
  // problem on this context).
  int tune_lanczos(int nd) {
    LzInst<T> L = lz;
    L.nd = nd;
    LzGraphs* gr = &graphs[nd == k ? 0 : 1];
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) return set_err(c, USBS_ERR_HIP, "event creation failed");
    int rc = USBS_OK;
    float ms_perop = 0.f, ms_pers = 0.f;
    int* ctrl_ptr = lz.ctrl;
    const LzStop stop{&ctrl_ptr, 2, nullptr};
    auto perop = [&]() {
      k_lz_start<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(L, q0);
      return hipGetLastError() != hipSuccess || lanczos_perop<T>(L, s, host_ctrl, nullptr, gr);
    };
    auto persistent = [&]() {
      k_lz_start<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(L, q0);
      return hipGetLastError() != hipSuccess ||
             lanczos_persistent<T>(nd == k ? d_lz_k : d_lz_kc, 1, nb, mL, lz.max_restarts, s, d_map,
                                   d_map + 2 * nb, lz.ntiles, LZ_SPIN_LIMIT, cache_bytes, nullptr,
                                   lanczos_tile_form(lz), &stop);
    };
    do {
      if (perop() || persistent()) { rc = USBS_ERR_HIP; break; }
      if (hipEventRecord(ev[0], s) != hipSuccess || perop()) { rc = USBS_ERR_HIP; break; }
      if (hipEventRecord(ev[1], s) != hipSuccess || persistent()) { rc = USBS_ERR_HIP; break; }
      if (hipEventRecord(ev[2], s) != hipSuccess || hipEventSynchronize(ev[2]) != hipSuccess ||
          hipEventElapsedTime(&ms_perop, ev[0], ev[1]) != hipSuccess ||
          hipEventElapsedTime(&ms_pers, ev[1], ev[2]) != hipSuccess) rc = USBS_ERR_HIP;
    } while (false);
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return set_err(c, rc, "Lanczos form tuning failed");
    const int mode = ms_perop < ms_pers ? USBS_EXEC_PEROP : USBS_EXEC_PERSISTENT;
    p.exec_mode = mode;
    tuning = false;
    c->tuned_mode = mode;
    c->tuned_m = mL;
    c->tuned_prec = p.precision;
    c->tune_ms[0] = ms_perop;
    c->tune_ms[1] = ms_pers;
    if (c->timing) {
      c->stats.lz_form = mode;
      use_ptm = mode == USBS_EXEC_PEROP;
    }
    lz_restarts = host_ctrl[LZ_RESTART];
    return USBS_OK;
  }

  // the start of a Lanczos call: column 0 = q0, control words reset
  int lanczos_start(int nd) {
    LzInst<T> L = lz;
    L.nd = nd;
    k_lz_start<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(L, q0);
    CK((int)hipGetLastError());
    return USBS_OK;
  }

  // Lanczos on the assembled operator (op64/op32 already hold -C + A*(z)).
  int lanczos(int nd) {
    if (tuning) return tune_lanczos(nd);
    LzInst<T> L = lz;
    L.nd = nd;
    if (lanczos_start(nd)) return USBS_ERR_HIP;
    if (p.exec_mode == USBS_EXEC_PEROP) {
      if (use_ptm) {
        ptm.ms = 0.0;
        ptm.launches = ptm.steps = ptm.sum_cols = 0;
      }
      CK(lanczos_perop<T>(L, s, host_ctrl, use_ptm ? &ptm : nullptr, &graphs[nd == k ? 0 : 1]));
      lz_restarts = host_ctrl[LZ_RESTART];
      if (use_ptm) {
        c->stats.lz_kernel_ms += ptm.ms;
        c->stats.lz_launches += ptm.launches;
        c->stats.lz_steps += ptm.steps;
        c->stats.lz_sum_cols += ptm.sum_cols;
      }
    } else {
      if (tev_pending) {
        const int rc = collect_timing();
        if (rc) return rc;
      }
      int* ctrl_ptr = lz.ctrl;
      int launched = 0;
      const LzStop stop{&ctrl_ptr, lz_hint, &launched};
      CK(lanczos_persistent<T>(nd == k ? d_lz_k : d_lz_kc, 1, nb, mL, lz.max_restarts, s, d_map,
                               d_map + 2 * nb, lz.ntiles, LZ_SPIN_LIMIT, cache_bytes,
                               tev.empty() ? nullptr : tev.data(), lanczos_tile_form(lz), &stop));
      tev_launched = launched;
      tev_pending = !tev.empty();
    }
    return USBS_OK;
  }

  int assemble_op(double coefC, const double* z1, double a1, const double* z2, double a2,
                  const double* z3, double a3, bool want32) {
    return launch_assemble(*P, coefC, z1, a1, z2, a2, z3, a3, op64, want32 ? op32 : nullptr, s);
  }

  // out (k x k) = V^T M V for the operator currently in op64.
  int project(double* out, hipStream_t st) {
    CK((launch_spmm<double, T>(n, P->row_ptr, P->col, op64, V(), ld, k, W, ld, P->lz_heavy, st)));
    CK((launch_tsgram<T, double>(n, P->tile, P->ntiles, V(), ld, k, W, ld, k, partial, out, st)));
    return USBS_OK;
  }
  int project(double* out) { return project(out, s); }

  int rounding() {
    if (round_done) {  // launched early on the side stream
      round_done = false;
      CK((int)hipStreamWaitEvent(s, ev_round, 0));
      k_copy_round<<<1, 64, 0, s>>>(sc, rsc);
      CK((int)hipGetLastError());
      return USBS_OK;
    }
    return rounding_on(s, partial, sc);
  }
  bool rounds_early() const { return r > 0 && p.rounding == 1; }
  // the rounding on stream st with partial-sum buffer part, its scalars
  // (SC_NYS_RHO, SC_CB) written into the block sco
  int rounding_on(hipStream_t st, double* part, double* sco) {
    if (r <= 0 || !p.rounding) return USBS_OK;
    CK((launch_tsgram<T, T>(n, P->tile, P->ntiles, P_(), ld, r, PO, ld, 2 * r, part, PtPO, st)));
    CK(launch_nystrom_core(n, r, PtPO, OtO, Gn, lam, sco, st));
    if (p.rounding == 2) {
      // QAP: the factor on the device, the assignments on the host
      CK(launch_nystrom_E<T>(n, r, P_(), Om(), ld, Gn, sco + SC_NYS_RHO, dE, st));
      if (hipMemcpyAsync(hE.data(), dE, sizeof(double) * hE.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return set_err(c, USBS_ERR_HIP, "reading the Nystrom factor");
      const int rc = qap_round_host(c->qap_l, r, hE.data(), c->qap_W.data(), c->qap_D.data(), nullptr,
                                    nullptr, &qap_cb);
      if (rc) return set_err(c, rc, "QAP rounding failed");
      return USBS_OK;
    }
    CK(launch_round<T>(*P, cutw, r, P_(), Om(), ld, Gn, sco + SC_NYS_RHO, sg, ipart, dpart, cut, st));
    k_copy1<<<1, 64, 0, st>>>(sco + SC_CB, cut + r);
    CK((int)hipGetLastError());
    return USBS_OK;
  }

  // one USBS iteration up to the candidate Lanczos call (its operator
  // -C + A*(ycand) assembled; the lower-bound IPM forked onto the side stream)
  int body_pre() {
    const double tu = p.trace_ub, rho = p.rho;
    // --- prox subproblem inputs (solver/usbs.py:53-61)
    CK(launch_dots3(*P, zbar, y, P->has_ineq ? ups : nullptr, partial, sc + SC_ZZ, s));
    CK(assemble_op(-1.0, y, 1.0, P->b, -1.0 / rho, P->has_ineq ? ups : nullptr, 1.0 / rho, false));
    if (project(Gh)) return USBS_ERR_HIP;
    CK(assemble_op(0.0, zbar, 1.0, nullptr, 0.0, nullptr, 0.0, false));
    if (project(Gz)) return USBS_ERR_HIP;
    CK(launch_qbase<T>(*P, k, V(), ld, partial, Gq, s));
    IpmArgs ia{};
    ia.k = k; ia.kc = kc; ia.max_iters = p.subprob_max_iters; ia.first = 1;
    ia.tu = tu; ia.rho = rho; ia.eps = p.subprob_eps;
    ia.Gq = Gq; ia.Gz = Gz; ia.Gh = Gh; ia.sc = sc; ia.eta_io = eta_io; ia.S_io = S_io;
    ia.GF = GF; ia.Epast = Epast;
    CK(launch_ipm_quad(ia, s));
    CK((launch_tsmm<T, T>(n, V(), ld, k, GF, k, F, ld, 0.0, nullptr, 0, s)));
    if (P->has_ineq) {
      // upsilon alternation (solver/usbs.py:384-431)
      double gap = 1.0;
      int it = 0;
      while (true) {
        CK(launch_upsilon<T>(*P, k, F, ld, sc + SC_ETA, zbar, y, rho, ups, partial, sc + SC_UPSGAP, s));
        ++it;
        if (hipMemcpyAsync(&gap, sc + SC_UPSGAP, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
          return set_err(c, USBS_ERR_HIP, "reading upsilon gap");
        if (!(gap > p.subprob_eps) || it >= p.subprob_max_iters) break;
        CK(launch_dots3(*P, zbar, y, ups, partial, sc + SC_ZZ, s));
        CK(assemble_op(-1.0, y, 1.0, P->b, -1.0 / rho, ups, 1.0 / rho, false));
        if (project(Gh)) return USBS_ERR_HIP;
        ia.first = 0;
        CK(launch_ipm_quad(ia, s));
        CK((launch_tsmm<T, T>(n, V(), ld, k, GF, k, F, ld, 0.0, nullptr, 0, s)));
      }
    }
    // --- primal / sketch / dual candidate (solver/usbs.py:568-585)
    if (dense) {
      // X = eta Xbar + V S V^T, then Xbar = eta Xbar + Fc Fc^T (usbs.py:576, 632)
      const dim3 g((unsigned)((n + 255) / 256), (unsigned)((n + 31) / 32));
      CK((launch_tsmm<T, T>(n, V(), ld, k, S_io, k, VS, ld, 0.0, nullptr, 0, s)));
      k_dense_rankk<T><<<g, 256, 0, s>>>(n, k, sc + SC_ETA, Xbar, X, ld, VS, V(), ld);
      CK((int)hipGetLastError());
      k_dense_rankk<T><<<g, 256, 0, s>>>(n, kc, sc + SC_ETA, Xbar, Xbar, ld, F, F, ld);
      CK((int)hipGetLastError());
    }
    CK(launch_constraint_update<T>(*P, k, kc, F, ld, sc + SC_ETA, zbar, y,
                                   P->has_ineq ? ups : nullptr, rho, z, zbn, ycand, partial,
                                   sc + SC_BYC, s));
    // The candidate Lanczos call needs only ycand.  The sketch update, the
    // objective, a due check's rounding and the spectral lower bound go to
    // the side stream and overlap it; body_post waits for them (ev_join).
    // Nothing on the main stream touches F, the sketch, `partial` or the
    // scalar block until then.
    CK((int)hipEventRecord(ev_fork, s));
    CK((int)hipStreamWaitEvent(s2, ev_fork, 0));
    if (r > 0) CK((launch_tsgram<T, T>(n, P->tile, P->ntiles, F, ld, k, Om(), ld, r, partial, FO, s2)));
    CK(launch_sketch_update<T>(n, k, kc, r, F, ld, FO, sc + SC_ETA, Pbar, P_(), ld, s2));
    CK(launch_objective<T>(*P, k, kc, F, ld, partial, sc + SC_OBJF, s2));
    const bool early = round_early && rounds_early();
    round_early = false;
    if (early) {
      if (rounding_on(s2, partial_r, rsc)) return USBS_ERR_HIP;
      CK((int)hipEventRecord(ev_round, s2));
      round_done = true;
    }
    // --- candidate eigenvectors (solver/usbs.py:588-599)
    CK(assemble_op(-1.0, ycand, 1.0, nullptr, 0.0, nullptr, 0.0, sizeof(T) == 4));
    CK((int)hipEventRecord(ev_asm, s));
    // --- spectral lower bound with the old V (solver/usbs.py:606-621): the
    // projection reads the candidate operator just assembled on the main
    // stream
    CK((int)hipStreamWaitEvent(s2, ev_asm, 0));
    if (project(Glb, s2)) return USBS_ERR_HIP;
    CK(launch_ipm_lb(k, p.subprob_max_iters, tu, p.subprob_eps, Glb, sc, s2));
    CK((int)hipEventRecord(ev_join, s2));
    return USBS_OK;
  }

  // the candidate Lanczos call happens between body_pre and body_post
  int body() {
    int rc = body_pre();
    if (rc) return rc;
    if (lanczos(kc)) return USBS_ERR_HIP;
    return body_post();
  }

  int body_post() {
    const double tu = p.trace_ub;
    CK((int)hipStreamWaitEvent(s, ev_join, 0));
    ScalarArgs sa{};
    sa.tu = tu; sa.tf = p.trace_factor; sa.beta = p.beta; sa.SC = p.SCALE_C; sa.SX = p.SCALE_X;
    sa.lanczos_alpha = lz.alpha;
    CK(launch_scalars(sa, sc, s));
    CK(launch_select(m, sc + SC_DESCENT, y, ycand, s));
    std::swap(zbar, zbn);
    // --- V_next = qr([V E_S[:, kc:], cand_vecs]) (solver/usbs.py:640-642)
    if (kp > 0) CK((launch_tsmm<T, T>(n, V(), ld, k, Epast, kp, Qb + ld, ld, 0.0, nullptr, 0, s)));
    if (hipMemcpyAsync(Qb + ld * (1 + kp), lz.Q, sizeof(T) * ld * kc, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "copy of candidate eigenvectors");
    LzInst<T> q = qr;
    q.Q = Qb;
    if (qr_mode < 0) {
      const int rc = plan_qr();
      if (rc) return rc;
    }
    // resets the control words and barrier counters (column 0 is set
    // below); the persistent QR's barrier error stays set for the records to
    // report
    k_lz_start<T><<<(unsigned)((LZ_BAR_STRIDE * q.nbar + 255) / 256), 256, 0, s>>>(q, q0, qr_mode);
    CK(fill_T(Qb, ld));
    if (qr_mode == 1)
      CK(orth_persistent<T>(d_qr + (Qb == qr_buf0 ? 0 : 1), lz.ntiles, k, s, d_map));
    else
      CK(orth_perop<T>(q, s));
    std::swap(Qa, Qb);
    return USBS_OK;
  }

  int plan_qr() {
    qr_mode = 0;
    if (batch_member || tuning || p.exec_mode != USBS_EXEC_PERSISTENT || cache_bytes == 0 ||
        nb != lz.ntiles || !orth_persistent_fits<T>(qr, c->device))
      return USBS_OK;
    if (!(d_qr = alloc<LzInst<T>>(2))) return USBS_ERR_OOM;
    LzInst<T> h[2] = {qr, qr};
    h[0].Q = Qa;
    h[1].Q = Qb;
    h[0].nblocks = h[1].nblocks = lz.ntiles;
    // the error word starts clear (a reused buffer holds old contents); the
    // per-call resets leave it alone from here on
    if (hipMemcpy(d_qr, h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemsetAsync(qr.ctrl, 0, sizeof(int) * LZ_NCTRL, s) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "upload of the QR descriptors failed");
    qr_buf0 = Qa;
    qr_mode = 1;
    return USBS_OK;
  }

  int fill_T(T* ptr, long long count) {
    k_fill<T><<<(unsigned)((count + 255) / 256), 256, 0, s>>>(ptr, count, (T)0);
    return (int)hipGetLastError();
  }

  // A record in two parts: record_launch queues the copies of the scalar
  // block (and, persistent, the Lanczos control words) behind the stream's
  // work; record_finish waits for them and fills rec.
  int record_launch() {
    if (hipMemcpyAsync(rec_h, sc, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, s) != hipSuccess ||
        (p.exec_mode == USBS_EXEC_PERSISTENT &&
         hipMemcpyAsync(ctrl_h, lz.ctrl, sizeof(int) * LZ_NCTRL, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        (qr_mode == 1 &&
         hipMemcpyAsync(ctrl_h + LZ_NCTRL, qr.ctrl + LZ_ERROR, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipEventRecord(ev_rec, s) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "reading the scalar block");
    return USBS_OK;
  }
  int record_finish(usbs_record& rec, double t0) {
    if (hipEventSynchronize(ev_rec) != hipSuccess) return set_err(c, USBS_ERR_HIP, "reading the scalar block");
    const double* h = rec_h;
    if (tev_pending) {
      const int rc = collect_timing();
      if (rc) return rc;
    }
    if (p.exec_mode == USBS_EXEC_PERSISTENT) {
      const int* ctrl = ctrl_h;
      if (ctrl[LZ_ERROR]) return set_err(c, USBS_ERR_TIMEOUT, "persistent Lanczos barrier wait expired");
      if (qr_mode == 1 && ctrl[LZ_NCTRL]) return set_err(c, USBS_ERR_TIMEOUT, "persistent QR barrier wait expired");
      lz_restarts = ctrl[LZ_RESTART];
      lz_hint = lz_restarts > 1 ? lz_restarts : 1;
    }
    const double scx = p.SCALE_C * p.SCALE_X;
    rec.t = h[SC_T];
    rec.time = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t0;
    rec.pen_dual_obj = h[SC_PEN];
    rec.cand_pen_dual_obj = h[SC_CANDPEN];
    rec.lb_spec_est = h[SC_LB];
    rec.primal_obj = h[SC_P] / scx;
    rec.dual_obj = h[SC_BY];
    rec.obj_ub = h[SC_OBJUB];
    rec.obj_gap = h[SC_GAP];
    rec.infeas_gap = h[SC_INFGAP];
    rec.max_infeas = h[SC_MAXINF];
    rec.callback_val = (r > 0 && p.rounding == 1) ? h[SC_CB] : (r > 0 && p.rounding == 2) ? qap_cb : NAN;
    rec.ipm_iters = h[SC_IPM_IT];
    rec.lb_iters = h[SC_LB_IT];
    rec.lanczos_restarts = lz_restarts;
    rec.descent = h[SC_DESCENT];
    return USBS_OK;
  }
#undef CK
};

// One usbs_solve, in the stages the batched driver (usbs_solve_batch) also
// runs: prepare (buffers, start point, scalar block), the initial bundle
// (operator assembly, a Lanczos call the caller issues, bundle and
// penalised dual), condition checks, and the outputs.  The statements are
// those of the single-solve loop, in its order.
template <typename T>
struct SolveRun {
  Ctx* c = nullptr;
  usbs_params p{};
  Solver<T> S;
  std::vector<DevBuf> tmp;  // per-call uploads and the output staging
  double t_origin = 0.0;
  int nlog = 0;
  long long iters = 0;
  bool go = true;
  usbs_record* log = nullptr;
  int log_cap = 0;

  ~SolveRun() { ctx_free_pool(tmp); }

  double* up(const double* h, size_t cnt) {
    double* dptr = (double*)ctx_alloc(c, tmp, sizeof(double) * cnt);
    if (dptr && hipMemcpy(dptr, h, sizeof(double) * cnt, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return dptr;
  }

  int prepare(Ctx* cc, const usbs_params& pp, const double* y0, const double* z0, double trX0,
              double pobj0, const double* Omega, const double* P0, const double* q0h,
              usbs_record* lg, int cap) {
    c = cc;
    p = pp;
    log = lg;
    log_cap = cap;
    int rc = S.setup(c, p);
    if (rc) return rc;
    hipStream_t s = c->stream;
    const int n = S.n, m = S.m, r = S.r;
    const long long ld = S.ld;
    const int rk = r > 0 ? r : 1;  // columns of P (the diag-only P is one column)
    const bool warm = p.warm != 0;
    if (warm) {
      if (c->keep_n != n || c->keep_m != m || c->keep_r != rk || c->keep_ld != ld)
        return set_err(c, USBS_ERR_STATE, "warm start needs a previous solve of the same size on this context");
      // y carries over; z and P are taken from the kept solve's SCALE_X to this one's
      if (hipMemcpyAsync(S.y, c->keep_y, sizeof(double) * m, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return set_err(c, USBS_ERR_HIP, "copy of the kept start point");
      k_rescale<double><<<(unsigned)((m + 255) / 256), 256, 0, s>>>(m, c->keep_z, S.z, c->keep_sx, p.SCALE_X);
      k_rescale<double><<<(unsigned)((m + 255) / 256), 256, 0, s>>>(m, c->keep_z, S.zbar, c->keep_sx, p.SCALE_X);
    } else if (hipMemcpy(S.y, y0, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(S.z, z0, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(S.zbar, z0, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess) {
      return set_err(c, USBS_ERR_HIP, "upload of the start point");
    }
    if (hipMemcpy(S.q0, q0h, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "upload of q0");
    S.fill(S.ups, m, 0.0);
    S.fill_T(S.Qa, ld * (S.k + 1));
    S.fill_T(S.Qb, ld * (S.k + 1));
    S.fill_T(S.PO, ld * 2 * (r > 0 ? r : 1));
    if (warm) {
      const long long cnt = ld * rk;
      k_rescale<T><<<(unsigned)((cnt + 255) / 256), 256, 0, s>>>(cnt, c->keep_P, S.P_(), c->keep_sx, p.SCALE_X);
      k_rescale<T><<<(unsigned)((cnt + 255) / 256), 256, 0, s>>>(cnt, c->keep_P, S.Pbar, c->keep_sx, p.SCALE_X);
    }
    if (r > 0) {
      double* d_om = up(Omega, (size_t)n * r);
      if (!d_om) return USBS_ERR_OOM;
      k_load_colmajor<T><<<(unsigned)(((long long)n * r + 255) / 256), 256, 0, s>>>(n, r, d_om, S.Om(), ld);
      if (warm) {
      } else if (P0) {
        double* d_p = up(P0, (size_t)n * r);
        if (!d_p) return USBS_ERR_OOM;
        k_load_colmajor<T><<<(unsigned)(((long long)n * r + 255) / 256), 256, 0, s>>>(n, r, d_p, S.P_(), ld);
        k_load_colmajor<T><<<(unsigned)(((long long)n * r + 255) / 256), 256, 0, s>>>(n, r, d_p, S.Pbar, ld);
      } else {
        S.fill_T(S.Pbar, ld * r);
      }
      if (launch_tsgram<T, T>(n, c->P.tile, c->P.ntiles, S.Om(), ld, r, S.Om(), ld, r, S.partial, S.OtO, s))
        return set_err(c, USBS_ERR_HIP, "Omega Gram");
    } else if (S.dense) {
      // dense X: P0 carries X0 (n x n row-major); X and Xbar start from it
      S.fill_T(S.Pbar, ld);
      const long long nn = (long long)n * n;
      if (P0) {
        double* d_x = up(P0, (size_t)nn);
        if (!d_x) return USBS_ERR_OOM;
        k_load_colmajor<double><<<(unsigned)((nn + 255) / 256), 256, 0, s>>>(n, n, d_x, S.X, ld);
        k_load_colmajor<double><<<(unsigned)((nn + 255) / 256), 256, 0, s>>>(n, n, d_x, S.Xbar, ld);
      } else {
        S.fill(S.X, ld * n, 0.0);
        S.fill(S.Xbar, ld * n, 0.0);
      }
    } else {
      if (warm) {
      } else if (P0) {
        double* d_p = up(P0, (size_t)n);
        if (!d_p) return USBS_ERR_OOM;
        k_load_colmajor<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, 1, d_p, S.P_(), ld);
        k_load_colmajor<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, 1, d_p, S.Pbar, ld);
      } else {
        S.fill_T(S.Pbar, ld);
      }
    }
    // scalar block
    std::vector<double> h(SC_COUNT, 0.0);
    h[SC_TRX] = trX0;
    h[SC_TRXB] = trX0;
    h[SC_P] = pobj0;
    h[SC_PB] = pobj0;
    h[SC_OBJUB] = INFINITY;
    h[SC_GAP] = 1.1 * p.obj_gap_eps;
    h[SC_INFGAP] = 1.1 * p.infeas_gap_eps;
    h[SC_MAXINF] = 1.1 * p.max_infeas_eps;
    h[SC_CB] = NAN;
    {
      std::vector<double> bh(m), sah(m);
      if (hipMemcpy(bh.data(), c->P.b, sizeof(double) * m, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(sah.data(), c->P.scale_a, sizeof(double) * m, hipMemcpyDeviceToHost) != hipSuccess)
        return set_err(c, USBS_ERR_HIP, "reading b");
      double nb2 = 0.0;
      for (int i = 0; i < m; ++i) {
        const double v = (bh[i] / sah[i]) / p.SCALE_X;
        nb2 += v * v;
      }
      h[SC_BNORM] = std::sqrt(nb2);
    }
    if (hipMemcpy(S.sc, h.data(), sizeof(double) * SC_COUNT, hipMemcpyHostToDevice) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "upload of the scalar block");

    t_origin =
        std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    return USBS_OK;
  }

  // the operator of the initial Lanczos call, -C + A*(y0)
  int assemble_initial() {
    hipStream_t s = c->stream;
    // initial bundle: k smallest eigenpairs of -C + A*(y0) (solver/usbs.py:693-706)
    if (launch_assemble(c->P, -1.0, S.y, 1.0, nullptr, 0.0, nullptr, 0.0, S.op64,
                        sizeof(T) == 4 ? S.op32 : nullptr, s))
      return set_err(c, USBS_ERR_HIP, "assemble");
    return USBS_OK;
  }

  // after the initial Lanczos call: the bundle and the penalised dual
  int after_initial() {
    hipStream_t s = c->stream;
    const long long ld = S.ld;
    if (hipMemcpyAsync(S.Qa + ld, S.lz.Q, sizeof(T) * ld * S.k, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "copy of the initial bundle");
    if (launch_dots3(c->P, S.y, c->P.b, nullptr, S.partial, S.tmp4, s)) return set_err(c, USBS_ERR_HIP, "dots");
    k_init_pen<<<1, 64, 0, s>>>(S.sc, S.tmp4, S.lz.alpha, p.trace_ub);

    return USBS_OK;
  }

  // condition check (solver/usbs.py:503-540): rounding, record, stop test
  // a condition check: cond_launch queues its kernels and the record's
  // copies, cond_finish waits for them (the batched driver launches every
  // due check before it finishes any)
  int cond() {
    const int rc = cond_launch();
    return rc ? rc : cond_finish();
  }
  int cond_launch() {
    hipStream_t s = c->stream;
    int rc2 = S.rounding();
    if (rc2) return rc2;
    if (launch_dots3(c->P, S.y, c->P.b, nullptr, S.partial, S.tmp4, s)) return set_err(c, USBS_ERR_HIP, "dots");
    k_set_by<<<1, 64, 0, s>>>(S.sc, S.tmp4);
    return S.record_launch();
  }
  int cond_finish() {
    usbs_record rec{};
    const int rc2 = S.record_finish(rec, t_origin);
    if (rc2) return rc2;
    if (log && nlog < log_cap) log[nlog] = rec;
    ++nlog;
    go = (rec.t == 0.0) ||
         (rec.time < p.max_time && (rec.obj_gap > p.obj_gap_eps || rec.infeas_gap > p.infeas_gap_eps ||
                                    rec.max_infeas > p.max_infeas_eps));
    return USBS_OK;
  }

  int finish(double* y_out, double* z_out, double* P_out, double* scal_out, int* log_len) {
    hipStream_t s = c->stream;
    const int n = S.n, m = S.m, r = S.r;
    const long long ld = S.ld;
    const int rk = r > 0 ? r : 1;
    if (log_len) *log_len = nlog;
    // outputs
    // P (n x r, or n for diag-only), or X (n x n) in the dense mode
    const int rr = S.dense ? n : (r > 0 ? r : 1);
    double* d_out = (double*)ctx_alloc(c, tmp, sizeof(double) * (size_t)n * rr);
    if (!d_out) return USBS_ERR_OOM;
    if (S.dense)
      k_store_rowmajor<double><<<(unsigned)(((long long)n * rr + 255) / 256), 256, 0, s>>>(n, rr, S.X, ld, d_out);
    else
      k_store_rowmajor<T><<<(unsigned)(((long long)n * rr + 255) / 256), 256, 0, s>>>(n, rr, S.P_(), ld, d_out);
    std::vector<double> hs(SC_COUNT);
    if ((y_out && hipMemcpyAsync(y_out, S.y, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        (z_out && hipMemcpyAsync(z_out, S.z, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        (P_out && hipMemcpyAsync(P_out, d_out, sizeof(double) * (size_t)n * rr, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipMemcpyAsync(hs.data(), S.sc, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "reading results");
    // keep y, z, P on the device for a warm start of the next solve (not
    // in the dense mode, which has no sketch to keep)
    if (S.dense) {
      ctx_free_pool(c->keep_bufs);
      c->keep_y = c->keep_z = c->keep_P = nullptr;
      c->keep_n = c->keep_m = 0;
      c->keep_r = -1;
    } else if (c->keep_n != n || c->keep_m != m || c->keep_r != rk || c->keep_ld != ld) {
      ctx_free_pool(c->keep_bufs);
      c->keep_y = (double*)ctx_alloc(c, c->keep_bufs, sizeof(double) * m);
      c->keep_z = (double*)ctx_alloc(c, c->keep_bufs, sizeof(double) * m);
      c->keep_P = (double*)ctx_alloc(c, c->keep_bufs, sizeof(double) * ld * rk);
      if (!c->keep_y || !c->keep_z || !c->keep_P) {
        ctx_free_pool(c->keep_bufs);
        c->keep_n = c->keep_m = 0;
        c->keep_r = -1;
        return USBS_ERR_OOM;
      }
      c->keep_n = n;
      c->keep_m = m;
      c->keep_r = rk;
      c->keep_ld = ld;
    }
    c->keep_sx = p.SCALE_X;
    if (!S.dense) {
      const long long cnt = ld * rk;
      k_cast<T, double><<<(unsigned)((cnt + 255) / 256), 256, 0, s>>>(cnt, S.P_(), c->keep_P);
      if (hipMemcpyAsync(c->keep_y, S.y, sizeof(double) * m, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(c->keep_z, S.z, sizeof(double) * m, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return set_err(c, USBS_ERR_HIP, "keeping the final state");
    }
    if (scal_out) {
      scal_out[0] = hs[SC_TRX];
      scal_out[1] = hs[SC_P];
      scal_out[2] = (double)iters;
      scal_out[3] = hs[SC_GAP];
      scal_out[4] = hs[SC_INFGAP];
      scal_out[5] = hs[SC_MAXINF];
    }
    return USBS_OK;
  }
};

template <typename T>
static int run_solve(Ctx* c, const usbs_params& p, const double* y0, const double* z0, double trX0,
                     double pobj0, const double* Omega, const double* P0, const double* q0h,
                     double* y_out, double* z_out, double* P_out, double* scal_out,
                     usbs_record* log, int log_cap, int* log_len) {
  SolveRun<T> R;
  int rc = R.prepare(c, p, y0, z0, trX0, pobj0, Omega, P0, q0h, log, log_cap);
  if (rc) return rc;
  if ((rc = R.assemble_initial())) return rc;
  if ((rc = R.S.lanczos(R.S.k))) return rc;
  if ((rc = R.after_initial())) return rc;
  if ((rc = R.cond())) return rc;
  Solver<T>& S = R.S;
  if (p.cond_exp_base > 1.0) {
    for (long long steps : geometric_schedule(p.max_iters, p.cond_exp_base)) {
      if (!R.go) break;
      for (long long i = 0; i < steps; ++i) {
        S.round_early = i == steps - 1;  // the check follows this iteration
        rc = S.body();
        if (rc) return rc;
        ++R.iters;
      }
      rc = R.cond();
      if (rc) return rc;
    }
  } else {
    for (long long it = 0; it < p.max_iters && R.go; ++it) {
      S.round_early = true;  // a check follows every iteration
      rc = S.body();
      if (rc) return rc;
      ++R.iters;
      rc = R.cond();
      if (rc) return rc;
    }
  }
  return R.finish(y_out, z_out, P_out, scal_out, log_len);
}

// The condition-check schedule of one solve: entry i is the number of
// iterations before the i-th check after the initial one -- 1 each without a
// geometric base, else utils/loop.py:85-98's steps.
struct CondSchedule {
  std::vector<long long> steps;
  long long count = 0;
  bool unit = true;
  explicit CondSchedule(const usbs_params& p) {
    unit = !(p.cond_exp_base > 1.0);
    if (unit) {
      count = p.max_iters;
    } else {
      steps = geometric_schedule(p.max_iters, p.cond_exp_base);
      count = (long long)steps.size();
    }
  }
  long long at(long long i) const { return unit ? 1 : steps[(size_t)i]; }
};

struct StreamEvent {
  hipEvent_t e = nullptr;
  ~StreamEvent() { if (e) (void)hipEventDestroy(e); }
};

// B solves in lockstep on one device.  Each runs the statements of run_solve
// on its own context and stream; per iteration every active solve enqueues its
// work up to the candidate Lanczos call, then one persistent cooperative
// launch per restart advances all their Lanczos calls together on a batch
// stream (the same step, Ritz and rotation kernels as a single persistent
// call, whose results do not depend on the workgroup plan), and each solve
// continues on its own stream.  A solve whose stop test holds, or whose
// schedule ends, leaves the batch; the launch plan is rebuilt for the rest.
template <typename T>
static int run_batch(const std::vector<Ctx*>& cs, const std::vector<usbs_params>& ps,
                     const usbs_batch_item* items, int& failed) {
  const int B = (int)cs.size();
  Ctx* c0 = cs[0];
  std::vector<std::unique_ptr<SolveRun<T>>> R(B);
  int rc = USBS_OK;
  for (int q = 0; q < B; ++q) {
    failed = q;
    const usbs_batch_item& it = items[q];
    R[q].reset(new SolveRun<T>());
    rc = R[q]->prepare(cs[q], ps[q], it.y0, it.z0, it.tr_X0, it.primal_obj0, it.Omega, it.P0,
                       it.q0, it.log, it.log_cap);
    if (rc) return rc;
    R[q]->S.batch_member = true;
    if ((rc = R[q]->assemble_initial())) return rc;
  }
  failed = -1;
  // launch plans (initial, loop): declared before the batch stream's guard so
  // that the stream drains before their descriptors are released
  LzBatch<T> lb_init, lb;
  hipStream_t bs = nullptr;
  if (hipStreamCreateWithFlags(&bs, hipStreamNonBlocking) != hipSuccess)
    return set_err(c0, USBS_ERR_HIP, "batch stream creation failed");
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  } bs_guard{bs};
  std::vector<StreamEvent> ready(B);
  StreamEvent lz_done;
  for (auto& e : ready)
    if (hipEventCreateWithFlags(&e.e, hipEventDisableTiming) != hipSuccess)
      return set_err(c0, USBS_ERR_HIP, "event creation failed");
  if (hipEventCreateWithFlags(&lz_done.e, hipEventDisableTiming) != hipSuccess)
    return set_err(c0, USBS_ERR_HIP, "event creation failed");
  const int m = ps[0].lanczos_inner_iterations, max_restarts = ps[0].lanczos_max_restarts;

  // one batched Lanczos call over the solves in `act` (descriptors in lb)
  auto batched_lanczos = [&](const std::vector<int>& act, LzBatch<T>& lb, int nd) -> int {
    int hint = 1;
    for (int q : act) {
      Solver<T>& S = R[q]->S;
      failed = q;
      if (S.lanczos_start(nd)) return USBS_ERR_HIP;
      if (hipEventRecord(ready[q].e, S.s) != hipSuccess || hipStreamWaitEvent(bs, ready[q].e, 0) != hipSuccess)
        return set_err(cs[q], USBS_ERR_HIP, "batch fork");
      hint = std::max(hint, S.lz_hint);
    }
    failed = 0;
    const int e = lb.run(bs, max_restarts, hint);
    if (e) return set_err(c0, USBS_ERR_HIP, std::string("batched Lanczos: ") + hipGetErrorString((hipError_t)e));
    if (hipEventRecord(lz_done.e, bs) != hipSuccess) return set_err(c0, USBS_ERR_HIP, "batch join");
    for (int q : act)
      if (hipStreamWaitEvent(R[q]->S.s, lz_done.e, 0) != hipSuccess) {
        failed = q;
        return set_err(cs[q], USBS_ERR_HIP, "batch join");
      }
    failed = -1;
    return USBS_OK;
  };
  auto plan = [&](const std::vector<int>& act, LzBatch<T>& lb, int nd) -> int {
    if (hipStreamSynchronize(bs) != hipSuccess) return set_err(c0, USBS_ERR_HIP, "synchronise");
    std::vector<LzInst<T>> L;
    for (int q : act) {
      L.push_back(R[q]->S.lz);
      L.back().nd = nd;
    }
    failed = 0;
    const int e = lb.build(c0, L, m);
    if (!e) failed = -1;
    return e;
  };

  // initial bundles (nd = k), then the first condition check
  std::vector<int> act(B);
  for (int q = 0; q < B; ++q) act[q] = q;
  {
    if ((rc = plan(act, lb_init, R[0]->S.k))) return rc;
    if ((rc = batched_lanczos(act, lb_init, R[0]->S.k))) return rc;
    for (int q = 0; q < B; ++q) {
      failed = q;
      if ((rc = R[q]->after_initial()) || (rc = R[q]->cond())) return rc;
    }
    if (hipStreamSynchronize(bs) != hipSuccess) return set_err(c0, USBS_ERR_HIP, "synchronise");
  }
  // per solve: schedule position, iterations left before its next check
  std::vector<CondSchedule> sched;
  std::vector<long long> pos(B, 0), left(B, 0);
  std::vector<char> on(B, 0);
  auto checks = [&](int q) -> int {  // run the checks that are due; leave on[q] set while the solve goes on
    while (on[q] && left[q] == 0) {
      failed = q;
      int e = R[q]->cond();
      if (e) return e;
      ++pos[q];
      on[q] = R[q]->go && pos[q] < sched[q].count;
      if (on[q]) left[q] = sched[q].at(pos[q]);
    }
    return USBS_OK;
  };
  for (int q = 0; q < B; ++q) {
    sched.emplace_back(ps[q]);
    on[q] = R[q]->go && sched[q].count > 0;
    if (on[q]) left[q] = sched[q].at(0);
    if ((rc = checks(q))) return rc;
  }
  std::vector<int> planned;
  while (true) {
    act.clear();
    for (int q = 0; q < B; ++q)
      if (on[q]) act.push_back(q);
    if (act.empty()) break;
    if (act != planned) {
      if ((rc = plan(act, lb, R[act[0]]->S.kc))) return rc;
      planned = act;
    }
    for (int q : act) {
      failed = q;
      R[q]->S.round_early = left[q] == 1;  // its check is due after this iteration
      if ((rc = R[q]->S.body_pre())) return rc;
    }
    if ((rc = batched_lanczos(act, lb, R[act[0]]->S.kc))) return rc;
    // every solve's rest of the iteration, then every due condition check,
    // is enqueued before the first check waits on its stream, so the solves
    // run them together
    for (int q : act) {
      failed = q;
      if ((rc = R[q]->S.body_post())) return rc;
      ++R[q]->iters;
      --left[q];
    }
    for (int q : act) {
      if (!(on[q] && left[q] == 0)) continue;
      failed = q;
      if ((rc = R[q]->cond_launch())) return rc;
    }
    for (int q : act) {
      if (!(on[q] && left[q] == 0)) continue;
      failed = q;
      if ((rc = R[q]->cond_finish())) return rc;
      ++pos[q];
      on[q] = R[q]->go && pos[q] < sched[q].count;
      if (on[q]) left[q] = sched[q].at(pos[q]);
      if ((rc = checks(q))) return rc;  // a further check due at once
    }
  }
  for (int q = 0; q < B; ++q) {
    failed = q;
    const usbs_batch_item& it = items[q];
    if ((rc = R[q]->finish(it.y_out, it.z_out, it.P_out, it.scal_out, it.log_len))) return rc;
  }
  failed = -1;
  return USBS_OK;
}

}  // namespace usbs

using namespace usbs;

extern "C" int usbs_set_qap_rounding(usbs_ctx* h, int32_t l, const double* W, const double* D) {
  if (!h) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (l < 1 || l > 64 || !W || !D) return set_err(c, USBS_ERR_INVALID, "need 1 <= l <= 64 and W, D");
  for (int e = 0; e < l * l; ++e)
    if (!std::isfinite(W[e]) || !std::isfinite(D[e])) return set_err(c, USBS_ERR_NONFINITE, "non-finite W / D");
  c->qap_l = l;
  c->qap_W.assign(W, W + (size_t)l * l);
  c->qap_D.assign(D, D + (size_t)l * l);
  return USBS_OK;
}

extern "C" int usbs_set_timing(usbs_ctx* h, int enable) {
  if (!h) return USBS_ERR_INVALID;
  h->c.timing = enable != 0;
  if (enable) h->c.stats = usbs_stats{};
  return USBS_OK;
}

extern "C" int usbs_get_stats(const usbs_ctx* h, usbs_stats* out) {
  if (!h || !out) return USBS_ERR_INVALID;
  *out = h->c.stats;
  return USBS_OK;
}

// The argument checks of usbs_solve (and of each usbs_solve_batch item).
static int check_solve_args(Ctx* c, const usbs_params& p, const double* y0, const double* z0,
                            const double* Omega, const double* q0) {
  if (!c->has_problem) return set_err(c, USBS_ERR_STATE, "no problem uploaded");
  if (!q0 || (!p.warm && (!y0 || !z0))) return set_err(c, USBS_ERR_INVALID, "null y0 / z0 / q0");
  const int k = p.k_curr + p.k_past;
  if (p.k_curr < 1 || p.k_past < 0 || k > 11)
    return set_err(c, USBS_ERR_INVALID, "need k_curr >= 1, k_past >= 0, k_curr + k_past <= 11");
  if (p.lanczos_inner_iterations < k || p.lanczos_inner_iterations > 32 ||
      p.lanczos_inner_iterations > c->P.n)
    return set_err(c, USBS_ERR_INVALID, "lanczos_inner_iterations must be in [k, min(n, 32)]");
  if (!(p.lanczos_rtol >= 0.0) || !(p.lanczos_rtol < 1.0))
    return set_err(c, USBS_ERR_INVALID, "lanczos_rtol must be in [0, 1)");
  if (p.lanczos_max_restarts < 1 || p.subprob_max_iters < 1 || p.max_iters < 0)
    return set_err(c, USBS_ERR_INVALID, "iteration limits must be positive");
  if (p.sketch_dim < -1 || p.sketch_dim > 48)
    return set_err(c, USBS_ERR_INVALID, "sketch_dim in [-1, 48] (-1: dense X)");
  if (p.sketch_dim == -1 && (Omega || p.rounding || p.warm))
    return set_err(c, USBS_ERR_INVALID, "the dense-X mode takes no Omega, rounding or warm start");
  if (p.sketch_dim == -1 && (long long)c->P.n * c->P.n > (1ll << 31))
    return set_err(c, USBS_ERR_INVALID, "the dense-X mode holds n x n doubles: n too large");
  if (p.sketch_dim > 0 && !Omega) return set_err(c, USBS_ERR_INVALID, "Omega required when sketch_dim > 0");
  if (p.rounding && p.sketch_dim == 0) return set_err(c, USBS_ERR_INVALID, "rounding needs a sketch");
  if (p.rounding < 0 || p.rounding > 2) return set_err(c, USBS_ERR_INVALID, "rounding must be 0, 1 or 2");
  if (!(p.cond_exp_base >= 1.0 && p.cond_exp_base <= 2.0))
    return set_err(c, USBS_ERR_INVALID, "cond_exp_base must be in [1, 2]");
  if (!(p.rho > 0.0) || !(p.trace_ub > 0.0) || !(p.trace_factor > 0.0))
    return set_err(c, USBS_ERR_INVALID, "rho, trace_ub and trace_factor must be positive");
  if (p.precision != USBS_PREC_FP64 && p.precision != USBS_PREC_MIXED)
    return set_err(c, USBS_ERR_INVALID, "unknown precision");
  if (p.exec_mode != USBS_EXEC_PEROP && p.exec_mode != USBS_EXEC_PERSISTENT &&
      p.exec_mode != USBS_EXEC_TUNE)
    return set_err(c, USBS_ERR_INVALID, "unknown exec_mode");
  if (!p.warm)
    for (int i = 0; i < c->P.m; ++i)
      if (!std::isfinite(y0[i]) || !std::isfinite(z0[i])) return set_err(c, USBS_ERR_NONFINITE, "non-finite y0/z0");
  return USBS_OK;
}

extern "C" int usbs_solve(usbs_ctx* h, const usbs_params* prm, const double* y0, const double* z0,
                          double tr_X0, double primal_obj0, const double* Omega, const double* P0,
                          const double* q0, double* y_out, double* z_out, double* P_out,
                          double* scal_out, usbs_record* log, int log_cap, int* log_len) {
  if (!h || !prm) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (!c->has_problem) return set_err(c, USBS_ERR_STATE, "no problem uploaded");
  usbs_params p = *prm;
  if (p.exec_mode == USBS_EXEC_AUTO)
    p.exec_mode = resolve_exec_mode(c->P, p.lanczos_inner_iterations, p.precision, c->num_cus);
  if (p.exec_mode == USBS_EXEC_TUNE && c->tuned_mode >= 0 &&
      c->tuned_m == p.lanczos_inner_iterations && c->tuned_prec == p.precision)
    p.exec_mode = c->tuned_mode;
  int rc = check_solve_args(c, p, y0, z0, Omega, q0);
  if (rc) return rc;
  const bool coop_form = p.exec_mode == USBS_EXEC_PERSISTENT || p.exec_mode == USBS_EXEC_TUNE;
  CoopGuard coop(coop_form ? (const void*)c : nullptr);
  if (coop_form && !coop.ok()) return set_err(c, USBS_ERR_STATE, COOP_BUSY_MSG);
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  if (p.precision == USBS_PREC_FP64)
    return run_solve<double>(c, p, y0, z0, tr_X0, primal_obj0, Omega, P0, q0, y_out, z_out, P_out,
                             scal_out, log, log_cap, log_len);
  return run_solve<float>(c, p, y0, z0, tr_X0, primal_obj0, Omega, P0, q0, y_out, z_out, P_out,
                          scal_out, log, log_cap, log_len);
}

extern "C" int usbs_solve_batch(const usbs_batch_item* items, int32_t count) {
  if (!items || count < 1) return USBS_ERR_INVALID;
  std::vector<Ctx*> cs(count);
  std::vector<usbs_params> ps(count);
  for (int q = 0; q < count; ++q) {
    if (!items[q].ctx || !items[q].params) return USBS_ERR_INVALID;
    cs[q] = &items[q].ctx->c;
  }
  Ctx* c0 = cs[0];
  auto fail_all = [&](int rc, const std::string& msg) {
    for (Ctx* c : cs) set_err(c, rc, msg);
    return rc;
  };
  for (int q = 0; q < count; ++q) {
    Ctx* c = cs[q];
    for (int j = 0; j < q; ++j)
      if (cs[j] == c) return fail_all(USBS_ERR_INVALID, "a context appears twice in the batch");
    if (c->device != c0->device) return fail_all(USBS_ERR_INVALID, "batch items on different devices");
    usbs_params p = *items[q].params;
    if (p.exec_mode != USBS_EXEC_AUTO && p.exec_mode != USBS_EXEC_PERSISTENT)
      return fail_all(USBS_ERR_INVALID, "a batch runs the persistent form: exec_mode must be auto or persistent");
    p.exec_mode = USBS_EXEC_PERSISTENT;
    const usbs_params& p0 = *items[0].params;
    if (p.precision != p0.precision || p.k_curr != p0.k_curr || p.k_past != p0.k_past ||
        p.lanczos_inner_iterations != p0.lanczos_inner_iterations ||
        p.lanczos_max_restarts != p0.lanczos_max_restarts)
      return fail_all(USBS_ERR_INVALID, "batch items must share precision, k_curr, k_past, "
                                        "lanczos_inner_iterations and lanczos_max_restarts");
    const usbs_batch_item& it = items[q];
    const int rc = check_solve_args(c, p, it.y0, it.z0, it.Omega, it.q0);
    if (rc) return fail_all(rc, "batch item " + std::to_string(q) + ": " + c->err);
    ps[q] = p;
  }
  CoopGuard coop(c0);
  if (!coop.ok()) return fail_all(USBS_ERR_STATE, COOP_BUSY_MSG);
  if (hipSetDevice(c0->device) != hipSuccess) return fail_all(USBS_ERR_HIP, "hipSetDevice");
  int failed = -1;
  const int rc = ps[0].precision == USBS_PREC_FP64 ? run_batch<double>(cs, ps, items, failed)
                                                   : run_batch<float>(cs, ps, items, failed);
  if (rc) {
    const Ctx* src = failed >= 0 ? cs[failed] : c0;
    return fail_all(rc, "batch item " + std::to_string(failed >= 0 ? failed : 0) + ": " + src->err);
  }
  return USBS_OK;
}
