// The CGAL loop (solver/cgal.py:21-243) with all vectors resident on the
// device.  One body:
//
//   constraint pass 1: proj_b, v = y + beta (z - proj_b) and the dual /
//                      infeasibility sums of the current point
//   assemble -C + A*(v) -> Lanczos, one eigenpair (per-op or persistent)
//   u^T C u (objective kernel), |u|^2, u^T Omega, zu = A(u u^T), v.zu
//   [host: gaps, eta, step scalars]
//   sketch P = (1 - eta) P + (eta trace_ub) u (u^T Omega)
//   constraint pass 2: z' = (1 - eta) z + eta trace_ub zu, proj_b', |z' - proj_b'|^2
//   [host: dual step size]  y += step (z' - proj_b')
//
// Reductions over constraints use per-tile partials summed in tile order on
// the host, so results do not depend on the launch shape.
#include <chrono>
#include <cmath>
#include <vector>

#include "../../include/usbs_hip.h"
#include "solve.hpp"
#include "solve_util.hpp"

namespace usbs {

#define CG_NP 8  // partial sums per constraint tile

// Pass 1 (cgal.py:100-105, 124-128): proj_b = (1-mask) b + mask min(z + y/beta, b),
// v = y + beta (z - proj_b); partials [v.z, y.(z - proj_b), |z - proj_b|^2,
// |(z - pb2)/SA/SX|^2, |pb2/SA/SX|^2, max |(z - pb2)/SA|] with pb2 the
// projection of z alone.
__global__ void __launch_bounds__(USBS_NT)
k_cgal_pass1(int m, int mtile, const double* __restrict__ y, const double* __restrict__ z,
             const double* __restrict__ b, const double* __restrict__ mask,
             const double* __restrict__ sa, double beta, double SX, double* __restrict__ v,
             double* __restrict__ partial) {
  __shared__ double red[8];
  const int tl = blockIdx.x;
  const int c0 = tl * mtile, c1 = min(m, c0 + mtile);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, mx = 0.0;
  for (int i = c0 + (int)threadIdx.x; i < c1; i += USBS_NT) {
    const double mk = mask[i], bi = b[i], zi = z[i], yi = y[i];
    const double pb = (1.0 - mk) * bi + mk * fmin(zi + (1.0 / beta) * yi, bi);
    const double d = zi - pb;
    const double vi = yi + beta * d;
    v[i] = vi;
    s0 += vi * zi;
    s1 += yi * d;
    s2 += d * d;
    const double pb2 = (1.0 - mk) * bi + mk * fmin(zi, bi);
    const double e = (zi - pb2) / sa[i];
    const double es = e / SX;
    s3 += es * es;
    const double ps = (pb2 / sa[i]) / SX;
    s4 += ps * ps;
    mx = fmax(mx, fabs(e));
  }
  s0 = block_sum1(s0, red);
  s1 = block_sum1(s1, red);
  s2 = block_sum1(s2, red);
  s3 = block_sum1(s3, red);
  s4 = block_sum1(s4, red);
  mx = wave_max(mx);
  __shared__ double rmx[USBS_NT / 64];
  if ((threadIdx.x & 63) == 0) rmx[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int w = 0; w < USBS_NT / 64; ++w) a = fmax(a, rmx[w]);
    double* o = partial + (size_t)tl * CG_NP;
    o[0] = s0; o[1] = s1; o[2] = s2; o[3] = s3; o[4] = s4; o[5] = a; o[6] = 0.0; o[7] = 0.0;
  }
}

// zu = A(u u^T) (solver/utils.py:11-16) per constraint in triple order, and
// per-tile partials of v.zu (= u^T A*(v) u).
template <typename T>
__global__ void __launch_bounds__(USBS_NT)
k_cgal_Au(int m, int mtile, const int* __restrict__ con_ptr, const int* __restrict__ ct_i,
          const int* __restrict__ ct_j, const double* __restrict__ ct_d,
          const T* __restrict__ u, const double* __restrict__ v, double* __restrict__ zu,
          double* __restrict__ partial) {
  __shared__ double red[8];
  const int tl = blockIdx.x;
  const int c0 = tl * mtile, c1 = min(m, c0 + mtile);
  double s = 0.0;
  for (int c = c0 + (int)threadIdx.x; c < c1; c += USBS_NT) {
    double a = 0.0;
    for (int p = con_ptr[c]; p < con_ptr[c + 1]; ++p)
      a += ct_d[p] * (double)u[ct_i[p]] * (double)u[ct_j[p]];
    zu[c] = a;
    s += v[c] * a;
  }
  s = block_sum1(s, red);
  if (threadIdx.x == 0) partial[(size_t)tl * CG_NP + 6] = s;
}

// Pass 2 (cgal.py:160-171): z' = (1 - eta) z + (eta tu us) zu; with the new
// beta, d = z' - proj_b(z' + y/beta); d kept in zu; partial |d|^2.
__global__ void __launch_bounds__(USBS_NT)
k_cgal_pass2(int m, int mtile, double* __restrict__ z, double* __restrict__ zu,
             const double* __restrict__ y, const double* __restrict__ b,
             const double* __restrict__ mask, double a, double cz, double beta,
             double* __restrict__ partial) {
  __shared__ double red[8];
  const int tl = blockIdx.x;
  const int c0 = tl * mtile, c1 = min(m, c0 + mtile);
  double s = 0.0;
  for (int i = c0 + (int)threadIdx.x; i < c1; i += USBS_NT) {
    const double zn = a * z[i] + cz * zu[i];
    z[i] = zn;
    const double mk = mask[i], bi = b[i];
    const double pb = (1.0 - mk) * bi + mk * fmin(zn + (1.0 / beta) * y[i], bi);
    const double d = zn - pb;
    zu[i] = d;
    s += d * d;
  }
  s = block_sum1(s, red);
  if (threadIdx.x == 0) partial[(size_t)tl * CG_NP + 7] = s;
}

__global__ void k_cgal_dual(int m, double* __restrict__ y, const double* __restrict__ d, double step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) y[i] = y[i] + step * d[i];
}

// P = a P + u (cu FO) (cgal.py:150-155), FO = u^T Omega (1 x r).
template <typename T>
__global__ void k_cgal_sketch(int n, int r, T* __restrict__ P, long long ld, const T* __restrict__ u,
                              const double* __restrict__ FO, double a, double cu) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * r) return;
  const int j = (int)(idx / n), i = (int)(idx - (long long)j * n);
  T* pj = P + (long long)j * ld;
  pj[i] = (T)(a * (double)pj[i] + (double)u[i] * (cu * FO[j]));
}

template <typename T>
struct CgalSolver {
  Ctx* c = nullptr;
  const Problem* P = nullptr;
  usbs_cgal_params p{};
  hipStream_t s = nullptr;
  int n = 0, m = 0, r = 0, mL = 0;
  long long ld = 0;
  std::vector<DevBuf> pool;
  double *y = nullptr, *z = nullptr, *v = nullptr, *zu = nullptr, *mpart = nullptr;
  double *op64 = nullptr, *partial = nullptr, *dpart = nullptr, *obj2 = nullptr, *uu = nullptr,
         *FO = nullptr, *PtPO = nullptr, *OtO = nullptr, *Gn = nullptr, *lam = nullptr,
         *cut = nullptr, *sc = nullptr, *q0 = nullptr;
  long long* ipart = nullptr;
  float* op32 = nullptr;
  T* PO = nullptr;  // P | Omega, n x 2r column-major
  signed char* sg = nullptr;
  CutWeights cutw{};  // weights of the rounding's cut values (C / cut_scale)
  LzInst<T> lz{};
  LzInst<T>* d_lz = nullptr;
  int* d_map = nullptr;
  int nb = 1;
  size_t cache_bytes = 0;
  int host_ctrl[LZ_NCTRL] = {0};
  std::vector<double> hpart;
  LzGraphs graphs;  // per-op restarts replayed as HIP graphs

  ~CgalSolver() {
    lanczos_graphs_clear(&graphs);
    ctx_free_pool(pool);
  }

  template <typename X>
  X* alloc(size_t count) {
    return (X*)ctx_alloc(c, pool, sizeof(X) * (count ? count : 1));
  }
  T* P_() const { return PO; }
  T* Om() const { return PO + (long long)r * ld; }

#define CK(expr)                                                                    \
  do {                                                                              \
    int _rc = (expr);                                                               \
    if (_rc) return set_err(c, USBS_ERR_HIP, std::string(#expr) + " -> " +         \
                                                 hipGetErrorString((hipError_t)_rc)); \
  } while (0)

  int setup(Ctx* ctx, const usbs_cgal_params& prm) {
    c = ctx;
    P = &ctx->P;
    p = prm;
    s = ctx->stream;
    n = P->n;
    m = P->m;
    r = p.sketch_dim;
    mL = p.lanczos_inner_iterations;
    ld = ((long long)n + 63) / 64 * 64;
    const int ntl = P->ntiles > P->mntiles ? P->ntiles : P->mntiles;
    if (!(y = alloc<double>(m)) || !(z = alloc<double>(m)) || !(v = alloc<double>(m)) ||
        !(zu = alloc<double>(m)) || !(mpart = alloc<double>((size_t)P->mntiles * CG_NP)))
      return USBS_ERR_OOM;
    if (!(PO = alloc<T>(ld * 2 * r)) || !(op64 = alloc<double>(P->nnz))) return USBS_ERR_OOM;
    if (sizeof(T) == 4 && !(op32 = alloc<float>(P->nnz))) return USBS_ERR_OOM;
    if (!(partial = alloc<double>(std::max((size_t)P->ntiles * std::max(2 * r * r, 64),
                                           (size_t)objective_ntiles(*P) * 16))) ||
        !(dpart = alloc<double>((size_t)ntl * r)) || !(ipart = alloc<long long>((size_t)ntl * r)) ||
        !(obj2 = alloc<double>(2)) || !(uu = alloc<double>(1)) || !(FO = alloc<double>(r)) ||
        !(PtPO = alloc<double>(2 * r * r)) || !(OtO = alloc<double>(r * r)) ||
        !(Gn = alloc<double>(r * r)) || !(lam = alloc<double>(r)) || !(cut = alloc<double>(r + 1)) ||
        !(sc = alloc<double>(SC_COUNT)) || !(q0 = alloc<double>(n)) ||
        !(sg = alloc<signed char>((size_t)n * r)))
      return USBS_ERR_OOM;
    if (p.rounding) {
      const int rcw = make_cut_weights(c, pool, p.cut_scale > 0.0 ? p.cut_scale : p.SCALE_C, &cutw);
      if (rcw) return rcw;
    }
    int rc = make_lanczos<T>(c, pool, *P, mL, 1, p.lanczos_max_restarts, p.subprob_eps,
                             sizeof(T) == 4 ? (const T*)op32 : (const T*)op64, lz);
    if (rc) return rc;
    if (lz.ldq != ld) return set_err(c, USBS_ERR_INVALID, "leading dimension mismatch");
    if (p.exec_mode == USBS_EXEC_PERSISTENT) {
      if (lanczos_plan_persistent<T>(lz, c->device, c->persist_wgs, &cache_bytes))
        return set_err(c, USBS_ERR_HIP, "no persistent Lanczos launch fits the device");
      nb = lz.nblocks;
      if (!(d_lz = alloc<LzInst<T>>(1)) || !(d_map = alloc<int>(2 * (nb + lz.ntiles))))
        return USBS_ERR_OOM;
      std::vector<int2> map(nb + lz.ntiles);
      for (int i = 0; i < nb; ++i) map[i] = make_int2(0, i);
      for (int i = 0; i < lz.ntiles; ++i) map[nb + i] = make_int2(0, i);
      if (hipMemcpy(d_lz, &lz, sizeof(lz), hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(d_map, map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice) != hipSuccess)
        return set_err(c, USBS_ERR_HIP, "upload of Lanczos descriptors failed");
    }
    hpart.assign((size_t)P->mntiles * CG_NP, 0.0);
    return USBS_OK;
  }

  int lanczos() {
    k_lz_start<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(lz, q0);
    CK((int)hipGetLastError());
    if (p.exec_mode == USBS_EXEC_PEROP) {
      CK(lanczos_perop<T>(lz, s, host_ctrl, nullptr, &graphs));
    } else {
      int* ctrl_ptr = lz.ctrl;
      const LzStop stop{&ctrl_ptr, 2, nullptr};
      CK(lanczos_persistent<T>(d_lz, 1, nb, mL, lz.max_restarts, s, d_map, d_map + 2 * nb,
                               lz.ntiles, LZ_SPIN_LIMIT, cache_bytes, nullptr,
                               lanczos_tile_form(lz), &stop));
    }
    return USBS_OK;
  }

  int rounding() {
    if (r <= 0 || !p.rounding) return USBS_OK;
    CK((launch_tsgram<T, T>(n, P->tile, P->ntiles, P_(), ld, r, PO, ld, 2 * r, partial, PtPO, s)));
    CK(launch_nystrom_core(n, r, PtPO, OtO, Gn, lam, sc, s));
    CK(launch_round<T>(*P, cutw, r, P_(), Om(), ld, Gn, sc + SC_NYS_RHO, sg, ipart, dpart, cut, s));
    return USBS_OK;
  }

  // Sum of column e of the constraint-tile partials, in tile order.
  double msum(int e) const {
    double a = 0.0;
    for (int t = 0; t < P->mntiles; ++t) a += hpart[(size_t)t * CG_NP + e];
    return a;
  }
  double mmax(int e) const {
    double a = 0.0;
    for (int t = 0; t < P->mntiles; ++t) a = std::fmax(a, hpart[(size_t)t * CG_NP + e]);
    return a;
  }
  int read_partials() {
    if (hipMemcpyAsync(hpart.data(), mpart, sizeof(double) * hpart.size(), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "reading partial sums");
    return USBS_OK;
  }

  // One iteration; gaps of the current point go to g[0..2].
  int body(long long t, double& trX, double& pobj, double g[3]) {
    const double tu = p.trace_ub, SCX = p.SCALE_C * p.SCALE_X;
    const double beta = p.beta0 * std::sqrt((double)t + 1.0);
    k_cgal_pass1<<<P->mntiles, USBS_NT, 0, s>>>(m, P->mtile, y, z, P->b, P->mask, P->scale_a, beta,
                                                 p.SCALE_X, v, mpart);
    CK((int)hipGetLastError());
    CK(launch_assemble(*P, -1.0, v, 1.0, nullptr, 0.0, nullptr, 0.0, op64,
                       sizeof(T) == 4 ? op32 : nullptr, s));
    int rc = lanczos();
    if (rc) return rc;
    const T* u = lz.Q;  // the eigenvector (column 0)
    CK(launch_objective<T>(*P, 1, 1, u, ld, partial, obj2, s));
    CK((launch_tsgram<T, T>(n, P->tile, P->ntiles, u, ld, 1, u, ld, 1, partial, uu, s)));
    if (r > 0) CK((launch_tsgram<T, T>(n, P->tile, P->ntiles, u, ld, 1, Om(), ld, r, partial, FO, s)));
    k_cgal_Au<T><<<P->mntiles, USBS_NT, 0, s>>>(m, P->mtile, P->con_ptr, P->ct_i, P->ct_j,
                                                 P->ct_d, u, v, zu, mpart);
    CK((int)hipGetLastError());
    double h[4];  // lambda, u^T C u, |u|^2
    if (hipMemcpyAsync(h, lz.alpha, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(h + 1, obj2, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(h + 2, uu, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "reading eigenpair scalars");
    rc = read_partials();
    if (rc) return rc;
    // cgal.py:115-135 (u = 0 when the smallest eigenvalue is >= 0)
    const double us = (h[0] >= 0.0) ? 0.0 : 1.0;
    const double uCu = us * h[1], uAvu = us * msum(6), unorm2 = us * h[2];
    double sgap = tu * uCu - pobj + msum(0);
    sgap -= tu * uAvu;
    double og = sgap - msum(1);
    og -= 0.5 * beta * msum(2);
    og = og / SCX;
    og /= 1.0 + std::fabs(pobj) / SCX;
    g[0] = og;
    g[1] = std::sqrt(msum(3)) / (1.0 + std::sqrt(msum(4)));
    g[2] = mmax(5) / p.SCALE_X;
    const double eta = 2.0 / ((double)t + 2.0);
    if (r > 0) {
      const long long tot = (long long)n * r;
      k_cgal_sketch<T><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, r, P_(), ld, u, FO, 1.0 - eta,
                                                                     eta * tu * us);
      CK((int)hipGetLastError());
    }
    trX = (1.0 - eta) * trX + eta * tu * unorm2;
    const double beta2 = p.beta0 * std::sqrt((double)t + 2.0);
    k_cgal_pass2<<<P->mntiles, USBS_NT, 0, s>>>(m, P->mtile, z, zu, y, P->b, P->mask, 1.0 - eta,
                                                 eta * tu * us, beta2, mpart);
    CK((int)hipGetLastError());
    rc = read_partials();
    if (rc) return rc;
    double step = 4.0 * beta2 * eta * eta * tu * tu / msum(7);
    step = step < 0.0 ? 0.0 : (step > p.beta0 ? p.beta0 : step);
    k_cgal_dual<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(m, y, zu, step);
    CK((int)hipGetLastError());
    pobj = (1.0 - eta) * pobj + eta * tu * uCu;
    return USBS_OK;
  }
#undef CK
};

template <typename T>
static int run_cgal(Ctx* c, const usbs_cgal_params& p, const double* y0, const double* z0,
                    double trX0, double pobj0, const double* Omega, const double* P0,
                    const double* q0h, double* y_out, double* z_out, double* P_out,
                    double* scal_out, usbs_record* log, int log_cap, int* log_len) {
  CgalSolver<T> S;
  int rc = S.setup(c, p);
  if (rc) return rc;
  hipStream_t s = c->stream;
  const int n = S.n, m = S.m, r = S.r;
  const long long ld = S.ld;
  std::vector<DevBuf> tmp;
  struct G { std::vector<DevBuf>& t; ~G() { ctx_free_pool(t); } } guard{tmp};
  auto up = [&](const double* h, size_t cnt) -> double* {
    double* d = (double*)ctx_alloc(c, tmp, sizeof(double) * cnt);
    if (d && hipMemcpy(d, h, sizeof(double) * cnt, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
  };
  if (hipMemcpy(S.y, y0, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(S.z, z0, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(S.q0, q0h, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "upload of the start point");
  double* d_om = up(Omega, (size_t)n * r);
  if (!d_om) return USBS_ERR_OOM;
  const unsigned gr = (unsigned)(((long long)n * r + 255) / 256);
  k_load_colmajor<T><<<gr, 256, 0, s>>>(n, r, d_om, S.Om(), ld);
  if (P0) {
    double* d_p = up(P0, (size_t)n * r);
    if (!d_p) return USBS_ERR_OOM;
    k_load_colmajor<T><<<gr, 256, 0, s>>>(n, r, d_p, S.P_(), ld);
  } else {
    k_fill<T><<<(unsigned)((ld * r + 255) / 256), 256, 0, s>>>(S.P_(), ld * r, (T)0);
  }
  if (launch_tsgram<T, T>(n, c->P.tile, c->P.ntiles, S.Om(), ld, r, S.Om(), ld, r, S.partial, S.OtO, s))
    return set_err(c, USBS_ERR_HIP, "Omega Gram");
  double trX = trX0, pobj = pobj0;
  double gaps[3] = {1.1 * p.obj_gap_eps, 1.1 * p.infeas_gap_eps, 1.1 * p.max_infeas_eps};
  const double t_origin =
      std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  long long t = 0;
  int nlog = 0;
  auto cond = [&](bool& go) -> int {
    int rc2 = S.rounding();
    if (rc2) return rc2;
    double cb = NAN;
    if (r > 0 && p.rounding) {
      if (hipMemcpyAsync(&cb, S.cut + r, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
        return set_err(c, USBS_ERR_HIP, "reading the cut value");
    }
    if (hipStreamSynchronize(s) != hipSuccess) return set_err(c, USBS_ERR_HIP, "synchronise");
    usbs_record rec;
    const double nan = NAN;
    rec.t = (double)t;
    rec.time = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() -
               t_origin;
    rec.pen_dual_obj = rec.cand_pen_dual_obj = rec.lb_spec_est = rec.dual_obj = rec.obj_ub = nan;
    rec.primal_obj = pobj / (p.SCALE_C * p.SCALE_X);
    rec.obj_gap = gaps[0];
    rec.infeas_gap = gaps[1];
    rec.max_infeas = gaps[2];
    rec.callback_val = cb;
    rec.ipm_iters = rec.lb_iters = rec.descent = nan;
    rec.lanczos_restarts = (p.exec_mode == USBS_EXEC_PEROP) ? S.host_ctrl[LZ_RESTART] : nan;
    if (log && nlog < log_cap) log[nlog] = rec;
    ++nlog;
    go = (t == 0) || (rec.time < p.max_time &&
                      (gaps[0] > p.obj_gap_eps || gaps[1] > p.infeas_gap_eps || gaps[2] > p.max_infeas_eps));
    return USBS_OK;
  };
  bool go = true;
  rc = cond(go);
  if (rc) return rc;
  auto step = [&]() -> int {
    const int rc2 = S.body(t, trX, pobj, gaps);
    ++t;
    return rc2;
  };
  if (p.cond_exp_base > 1.0) {
    for (long long steps : geometric_schedule(p.max_iters, p.cond_exp_base)) {
      if (!go) break;
      for (long long i = 0; i < steps; ++i)
        if ((rc = step())) return rc;
      if ((rc = cond(go))) return rc;
    }
  } else {
    for (long long it = 0; it < p.max_iters && go; ++it) {
      if ((rc = step())) return rc;
      if ((rc = cond(go))) return rc;
    }
  }
  if (log_len) *log_len = nlog;
  double* d_out = (double*)ctx_alloc(c, tmp, sizeof(double) * (size_t)n * r);
  if (!d_out) return USBS_ERR_OOM;
  k_store_rowmajor<T><<<gr, 256, 0, s>>>(n, r, S.P_(), ld, d_out);
  if ((y_out && hipMemcpyAsync(y_out, S.y, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      (z_out && hipMemcpyAsync(z_out, S.z, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      (P_out && hipMemcpyAsync(P_out, d_out, sizeof(double) * (size_t)n * r, hipMemcpyDeviceToHost, s) !=
                    hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "reading results");
  if (scal_out) {
    scal_out[0] = trX;
    scal_out[1] = pobj;
    scal_out[2] = (double)t;
    scal_out[3] = gaps[0];
    scal_out[4] = gaps[1];
    scal_out[5] = gaps[2];
  }
  return USBS_OK;
}

}  // namespace usbs

using namespace usbs;

extern "C" int usbs_cgal_solve(usbs_ctx* h, const usbs_cgal_params* prm, const double* y0,
                               const double* z0, double tr_X0, double primal_obj0,
                               const double* Omega, const double* P0, const double* q0,
                               double* y_out, double* z_out, double* P_out, double* scal_out,
                               usbs_record* log, int log_cap, int* log_len) {
  if (!h || !prm) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (!c->has_problem) return set_err(c, USBS_ERR_STATE, "no problem uploaded");
  usbs_cgal_params p = *prm;
  if (p.exec_mode == USBS_EXEC_AUTO || p.exec_mode == USBS_EXEC_TUNE)
    p.exec_mode = resolve_exec_mode(c->P, p.lanczos_inner_iterations, p.precision, c->num_cus);
  if (!y0 || !z0 || !q0) return set_err(c, USBS_ERR_INVALID, "null y0 / z0 / q0");
  if (p.lanczos_inner_iterations < 1 || p.lanczos_inner_iterations > 32 ||
      p.lanczos_inner_iterations > c->P.n)
    return set_err(c, USBS_ERR_INVALID, "lanczos_inner_iterations must be in [1, min(n, 32)]");
  if (p.lanczos_max_restarts < 1 || p.max_iters < 0)
    return set_err(c, USBS_ERR_INVALID, "iteration limits must be positive");
  if (p.sketch_dim < 1 || p.sketch_dim > 48 || !Omega)
    return set_err(c, USBS_ERR_INVALID, "CGAL needs a sketch: sketch_dim in [1, 48] and Omega");
  if (!(p.cond_exp_base >= 1.0 && p.cond_exp_base <= 2.0))
    return set_err(c, USBS_ERR_INVALID, "cond_exp_base must be in [1, 2]");
  if (!(p.beta0 > 0.0) || !(p.trace_ub > 0.0))
    return set_err(c, USBS_ERR_INVALID, "beta0 and trace_ub must be positive");
  if (p.precision != USBS_PREC_FP64 && p.precision != USBS_PREC_MIXED)
    return set_err(c, USBS_ERR_INVALID, "unknown precision");
  if (p.exec_mode != USBS_EXEC_PEROP && p.exec_mode != USBS_EXEC_PERSISTENT)
    return set_err(c, USBS_ERR_INVALID, "unknown exec_mode");
  for (int i = 0; i < c->P.m; ++i)
    if (!std::isfinite(y0[i]) || !std::isfinite(z0[i])) return set_err(c, USBS_ERR_NONFINITE, "non-finite y0/z0");
  CoopGuard coop(p.exec_mode == USBS_EXEC_PERSISTENT ? (const void*)c : nullptr);
  if (p.exec_mode == USBS_EXEC_PERSISTENT && !coop.ok())
    return set_err(c, USBS_ERR_STATE, COOP_BUSY_MSG);
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  if (p.precision == USBS_PREC_FP64)
    return run_cgal<double>(c, p, y0, z0, tr_X0, primal_obj0, Omega, P0, q0, y_out, z_out, P_out,
                            scal_out, log, log_cap, log_len);
  return run_cgal<float>(c, p, y0, z0, tr_X0, primal_obj0, Omega, P0, q0, y_out, z_out, P_out,
                         scal_out, log, log_cap, log_len);
}
