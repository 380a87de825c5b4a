// Standalone C-ABI entry points for the individual hot-path operations, so a
// caller (or a test) can run each one on host arrays: the sketch update, the
// constraint map with the dual candidate and infeasibility sums, the primal
// objective, the Nystrom reconstruction and the max-cut rounding.  They use
// the same kernels as the usbs_solve loop.
#include <cmath>
#include <vector>

#include "../../include/usbs_hip.h"
#include "munkres.hpp"
#include "solve.hpp"

namespace usbs {

struct Scratch {
  Ctx* c;
  std::vector<DevBuf> pool;
  explicit Scratch(Ctx* cc) : c(cc) {}
  ~Scratch() { ctx_free_pool(pool); }
  template <typename X>
  X* alloc(size_t n) { return (X*)ctx_alloc(c, pool, sizeof(X) * (n ? n : 1)); }
  template <typename X>
  X* up(const X* h, size_t n) {
    X* d = alloc<X>(n);
    if (d && n && hipMemcpy(d, h, sizeof(X) * n, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
  }
};

template <typename T>
__global__ void k_rm_to_cm(long long n, int r, const double* src, T* dst, long long ld) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * r) return;
  dst[(i / r) + (i % r) * ld] = (T)src[i];
}

template <typename T>
__global__ void k_cm_to_rm(long long n, int r, const T* src, long long ld, double* dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * r) return;
  dst[i] = (double)src[(i / r) + (i % r) * ld];
}

// E = (P + rho Omega) G, row-major n x r output (utils/common.py:110-112).
template <typename T>
__global__ void k_nystrom_E(int n, int r, const T* P, const T* Om, long long ld,
                            const double* G, const double* rho_ptr, double* E) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double rho = *rho_ptr;
  for (int j = 0; j < r; ++j) {
    double s = 0.0;
    for (int x = 0; x < r; ++x) s += ((double)P[i + x * ld] + rho * (double)Om[i + x * ld]) * G[x * r + j];
    E[(long long)i * r + j] = s;
  }
}

template <typename T>
static T* upload_cm(Scratch& S, const double* h, long long n, int r, long long ld, hipStream_t s) {
  double* tmp = S.up(h, (size_t)n * r);
  T* d = S.alloc<T>((size_t)ld * r);
  if (!tmp || !d) return nullptr;
  k_rm_to_cm<T><<<(unsigned)((n * r + 255) / 256), 256, 0, s>>>(n, r, tmp, d, ld);
  return d;
}

// Nystrom core + rounding shared by the two entry points below.  PO holds
// [P | Omega] column-major with leading dimension ld.
static int nystrom_prepare(Ctx* c, Scratch& S, int n, int r, const double* P, const double* Om,
                           double** d_PO, long long& ld, double** d_G, double** d_lam,
                           double** d_sc) {
  hipStream_t s = c->stream;
  ld = ((long long)n + 63) / 64 * 64;
  double* PO = S.alloc<double>((size_t)ld * 2 * r);
  if (!PO) return USBS_ERR_OOM;
  double* tp = S.up(P, (size_t)n * r);
  double* to = S.up(Om, (size_t)n * r);
  if (!tp || !to) return USBS_ERR_OOM;
  k_rm_to_cm<double><<<(unsigned)(((long long)n * r + 255) / 256), 256, 0, s>>>(n, r, tp, PO, ld);
  k_rm_to_cm<double><<<(unsigned)(((long long)n * r + 255) / 256), 256, 0, s>>>(n, r, to, PO + r * ld, ld);
  const int tile = tile_rows_for(n), ntiles = (n + tile - 1) / tile;
  double* part = S.alloc<double>((size_t)ntiles * 2 * r * r);
  double* PtPO = S.alloc<double>(2 * r * r);
  double* OtO = S.alloc<double>(r * r);
  double* G = S.alloc<double>(r * r);
  double* lam = S.alloc<double>(r);
  double* sc = S.alloc<double>(SC_COUNT);
  if (!part || !PtPO || !OtO || !G || !lam || !sc) return USBS_ERR_OOM;
  if (launch_tsgram<double, double>(n, tile, ntiles, PO, ld, r, PO, ld, 2 * r, part, PtPO, s) ||
      launch_tsgram<double, double>(n, tile, ntiles, PO + r * ld, ld, r, PO + r * ld, ld, r, part, OtO, s) ||
      launch_nystrom_core(n, r, PtPO, OtO, G, lam, sc, s))
    return set_err(c, USBS_ERR_HIP, "Nystrom core launch failed");
  *d_PO = PO;
  *d_G = G;
  *d_lam = lam;
  *d_sc = sc;
  return USBS_OK;
}

template <typename T>
int launch_nystrom_E(int n, int r, const T* P, const T* Om, long long ld, const double* G,
                     const double* rho, double* E, hipStream_t s) {
  k_nystrom_E<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, r, P, Om, ld, G, rho, E);
  return (int)hipGetLastError();
}
template int launch_nystrom_E<double>(int, int, const double*, const double*, long long,
                                      const double*, const double*, double*, hipStream_t);
template int launch_nystrom_E<float>(int, int, const float*, const float*, long long,
                                     const double*, const double*, double*, hipStream_t);

// QAP rounding on the host from the Nystrom factor E (N x r row-major,
// N = l^2 + 1), utils/qap_helpers.py:625-653: for columns i = 0..l-1 (the
// reference gathers them with clamped indices, so i >= r reads column r-1),
// the assignment minimising max(c) - c for c = E[1:, i] and for c = -E[1:, i]
// (an l x l matrix, c[a][j] = E[1 + l a + j]), each priced by
// trace(W Pi D Pi^T); the smaller of the two minima (positive unless the
// negative one is strictly smaller).  values: 2 l (positive candidates, then
// negative ones); perm: the best assignment.
int qap_round_host(int l, int r, const double* E, const double* W, const double* D,
                   double* values, int* perm, double* best) {
  std::vector<double> cost((size_t)l * l);
  std::vector<int> mask, asg(l), best_pos(l), best_neg(l);
  double min_pos = 0.0, min_neg = 0.0;
  for (int sign = 0; sign < 2; ++sign) {
    for (int i = 0; i < l; ++i) {
      const int col = i < r ? i : r - 1;
      double mx = 0.0;
      for (int e = 0; e < l * l; ++e) {
        const double v = E[(size_t)(1 + e) * r + col];
        cost[e] = sign ? -v : v;
        mx = (e == 0 || cost[e] > mx) ? cost[e] : mx;
      }
      for (int e = 0; e < l * l; ++e) cost[e] = mx - cost[e];
      munkres_mask(l, cost.data(), mask);
      mask_to_assign(l, mask, asg.data());
      const double v = qap_objective_mask(l, W, D, mask);
      if (values) values[sign * l + i] = v;
      double& mn = sign ? min_neg : min_pos;
      std::vector<int>& bp = sign ? best_neg : best_pos;
      if (i == 0 || v < mn) { mn = v; bp = asg; }
    }
  }
  const bool pos = min_pos < min_neg;
  *best = pos ? min_pos : min_neg;
  if (perm)
    for (int a = 0; a < l; ++a) perm[a] = pos ? best_pos[a] : best_neg[a];
  return USBS_OK;
}

}  // namespace usbs

using namespace usbs;

extern "C" {

int usbs_munkres(int32_t n, const double* cost, int32_t* mask_out) {
  if (n < 1 || n > 4096 || !cost || !mask_out) return USBS_ERR_INVALID;
  for (long long e = 0; e < (long long)n * n; ++e)
    if (!std::isfinite(cost[e])) return USBS_ERR_NONFINITE;
  std::vector<int> mask;
  munkres_mask(n, cost, mask);
  for (size_t e = 0; e < mask.size(); ++e) mask_out[e] = mask[e];
  return USBS_OK;
}

int usbs_qap_round(usbs_ctx* h, int32_t l, int32_t r, const double* P, const double* Omega,
                   const double* W, const double* D, double* values, int32_t* perm,
                   double* best) {
  if (!h || !P || !Omega || !W || !D || !best) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (l < 1 || l > 64 || r < 1 || r > 48) return set_err(c, USBS_ERR_INVALID, "need 1 <= l <= 64, 1 <= r <= 48");
  const int n = l * l + 1;
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  Scratch S(c);
  double *PO, *G, *lam, *sc;
  long long ld;
  int rc = nystrom_prepare(c, S, n, r, P, Omega, &PO, ld, &G, &lam, &sc);
  if (rc) return rc;
  double* dE = S.alloc<double>((size_t)n * r);
  if (!dE) return USBS_ERR_OOM;
  hipStream_t s = c->stream;
  if (launch_nystrom_E(n, r, PO, PO + r * ld, ld, G, sc + SC_NYS_RHO, dE, s))
    return set_err(c, USBS_ERR_HIP, "Nystrom factor");
  std::vector<double> E((size_t)n * r);
  if (hipMemcpyAsync(E.data(), dE, sizeof(double) * E.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "reading the Nystrom factor");
  std::vector<int> pm(l);
  rc = qap_round_host(l, r, E.data(), W, D, values, pm.data(), best);
  if (rc) return set_err(c, rc, "QAP rounding failed");
  if (perm)
    for (int a = 0; a < l; ++a) perm[a] = pm[a];
  return USBS_OK;
}

int usbs_reconstruct_from_sketch(usbs_ctx* h, int64_t n, int32_t r, const double* Omega,
                                 const double* P, double* E, double* Lambda) {
  if (!h || !Omega || !P || !E || !Lambda) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (n < 1 || r < 1 || r > 48 || n > INT32_MAX / 2) return set_err(c, USBS_ERR_INVALID, "bad n or r");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  Scratch S(c);
  double *PO, *G, *lam, *sc;
  long long ld;
  int rc = nystrom_prepare(c, S, (int)n, r, P, Omega, &PO, ld, &G, &lam, &sc);
  if (rc) return rc;
  double* dE = S.alloc<double>((size_t)n * r);
  if (!dE) return USBS_ERR_OOM;
  hipStream_t s = c->stream;
  k_nystrom_E<double><<<(unsigned)((n + 255) / 256), 256, 0, s>>>((int)n, r, PO, PO + r * ld, ld, G,
                                                          sc + SC_NYS_RHO, dE);
  if (hipMemcpyAsync(E, dE, sizeof(double) * n * r, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(Lambda, lam, sizeof(double) * r, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "reconstruct_from_sketch");
  return USBS_OK;
}

int usbs_maxcut_round(usbs_ctx* h, int32_t r, const double* P, const double* Omega, double* values,
                      int8_t* signs) {
  if (!h || !P || !Omega || !values) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (!c->has_problem) return set_err(c, USBS_ERR_STATE, "no problem uploaded");
  if (r < 1 || r > 48) return set_err(c, USBS_ERR_INVALID, "r in [1, 48]");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  const Problem& Pb = c->P;
  const int n = Pb.n;
  Scratch S(c);
  double *PO, *G, *lam, *sc;
  long long ld;
  int rc = nystrom_prepare(c, S, n, r, P, Omega, &PO, ld, &G, &lam, &sc);
  if (rc) return rc;
  signed char* sg = S.alloc<signed char>((size_t)n * r);
  long long* ip = S.alloc<long long>((size_t)Pb.ntiles * r);
  double* dp = S.alloc<double>((size_t)Pb.ntiles * r);
  double* out = S.alloc<double>(r + 1);
  if (!sg || !ip || !dp || !out) return USBS_ERR_OOM;
  hipStream_t s = c->stream;
  if (launch_round<double>(Pb, cut_weights_of(Pb), r, PO, PO + r * ld, ld, G, sc + SC_NYS_RHO, sg, ip, dp, out, s))
    return set_err(c, USBS_ERR_HIP, "rounding launch failed");
  std::vector<double> hv(r + 1);
  std::vector<signed char> hs;
  if (signs) hs.resize((size_t)n * r);
  if (hipMemcpyAsync(hv.data(), out, sizeof(double) * (r + 1), hipMemcpyDeviceToHost, s) != hipSuccess ||
      (signs && hipMemcpyAsync(hs.data(), sg, (size_t)n * r, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "reading rounding results");
  for (int j = 0; j < r; ++j) values[j] = hv[j];
  if (signs)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < r; ++j) signs[(size_t)i * r + j] = hs[(size_t)j * n + i];
  return USBS_OK;
}

int usbs_sketch_update(usbs_ctx* h, int64_t n, int32_t k, int32_t kc, int32_t r, const double* F,
                       const double* Omega, const double* Pbar, double eta, int32_t precision,
                       double* P_out, double* Pbar_out) {
  if (!h || !F || !Pbar || !P_out || !Pbar_out) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (n < 1 || k < 1 || k > 16 || kc < 0 || kc > k || r < 0 || r > 48 || (r > 0 && !Omega))
    return set_err(c, USBS_ERR_INVALID, "bad sketch sizes");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  Scratch S(c);
  hipStream_t s = c->stream;
  const long long ld = ((long long)n + 63) / 64 * 64;
  const int tile = tile_rows_for((int)n), ntiles = (int)((n + tile - 1) / tile);
  const int rr = r > 0 ? r : 1;
  double* d_eta = S.up(&eta, 1);
  double* FO = S.alloc<double>(k * rr);
  double* part = S.alloc<double>((size_t)ntiles * k * rr);
  double* out = S.alloc<double>((size_t)n * 2 * rr);
  if (!d_eta || !FO || !part || !out) return USBS_ERR_OOM;
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    T* dF = upload_cm<T>(S, F, n, k, ld, s);
    T* dPb = upload_cm<T>(S, Pbar, n, rr, ld, s);
    T* dP = S.alloc<T>((size_t)ld * rr);
    if (!dF || !dPb || !dP) return (int)USBS_ERR_OOM;
    if (r > 0) {
      T* dO = upload_cm<T>(S, Omega, n, r, ld, s);
      if (!dO) return (int)USBS_ERR_OOM;
      if (launch_tsgram<T, T>((int)n, tile, ntiles, dF, ld, k, dO, ld, r, part, FO, s))
        return set_err(c, USBS_ERR_HIP, "F^T Omega");
    }
    if (launch_sketch_update<T>((int)n, k, kc, r, dF, ld, FO, d_eta, dPb, dP, ld, s))
      return set_err(c, USBS_ERR_HIP, "sketch update");
    k_cm_to_rm<T><<<(unsigned)((n * rr + 255) / 256), 256, 0, s>>>(n, rr, dP, ld, out);
    k_cm_to_rm<T><<<(unsigned)((n * rr + 255) / 256), 256, 0, s>>>(n, rr, dPb, ld, out + n * rr);
    return (int)USBS_OK;
  };
  int rc = (precision == USBS_PREC_FP64) ? run(double(0)) : run(float(0));
  if (rc) return rc;
  if (hipMemcpyAsync(P_out, out, sizeof(double) * n * rr, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(Pbar_out, out + n * rr, sizeof(double) * n * rr, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "reading sketch");
  return USBS_OK;
}

int usbs_objective_infeas(usbs_ctx* h, int32_t k, int32_t kc, const double* F, double eta,
                          const double* zbar, const double* y, const double* upsilon, double rho,
                          double* z_out, double* zbar_next_out, double* ycand_out, double* out6) {
  if (!h || !F || !zbar || !y || !out6) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (!c->has_problem) return set_err(c, USBS_ERR_STATE, "no problem uploaded");
  if (k < 1 || k > 16 || kc < 0 || kc > k) return set_err(c, USBS_ERR_INVALID, "bad k");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  const Problem& Pb = c->P;
  const int n = Pb.n, m = Pb.m;
  Scratch S(c);
  hipStream_t s = c->stream;
  const long long ld = ((long long)n + 63) / 64 * 64;
  const int ntl = std::max(std::max(Pb.ntiles, Pb.mntiles), objective_ntiles(Pb));
  double* dF = upload_cm<double>(S, F, n, k, ld, s);
  double* d_eta = S.up(&eta, 1);
  double* dzb = S.up(zbar, m);
  double* dy = S.up(y, m);
  double* du = upsilon ? S.up(upsilon, m) : nullptr;
  double *dz = S.alloc<double>(m), *dzn = S.alloc<double>(m), *dyc = S.alloc<double>(m);
  double* part = S.alloc<double>((size_t)ntl * 16);
  double* o6 = S.alloc<double>(6);
  if (!dF || !d_eta || !dzb || !dy || (upsilon && !du) || !dz || !dzn || !dyc || !part || !o6)
    return USBS_ERR_OOM;
  if (launch_objective<double>(Pb, k, kc, dF, ld, part, o6, s) ||
      launch_constraint_update<double>(Pb, k, kc, dF, ld, d_eta, dzb, dy, du, rho, dz, dzn, dyc, part,
                                       o6 + 2, s))
    return set_err(c, USBS_ERR_HIP, "objective / constraint update");
  if (hipMemcpyAsync(out6, o6, sizeof(double) * 6, hipMemcpyDeviceToHost, s) != hipSuccess ||
      (z_out && hipMemcpyAsync(z_out, dz, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      (zbar_next_out && hipMemcpyAsync(zbar_next_out, dzn, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      (ycand_out && hipMemcpyAsync(ycand_out, dyc, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "reading objective");
  return USBS_OK;
}

}  // extern "C"
