// The constraint operator and its adjoint as standalone C-ABI entry points
// (solver/utils.py:11-37): z = A(U U^T) summed over the columns of U, and
// A*(z) U column by column.  The solve loop never calls these (its kernels
// fuse the same sums into the value assembly and the constraint update); they
// are the reference's own operator functions for callers that use them
// directly.
//
// The triples are grouped by output index on the host (a counting sort that
// keeps input order within a group), so each output is one thread's ordered
// sum: no atomics, the same result on every run.
#include <cmath>
#include <vector>

#include "../../include/usbs_hip.h"
#include "engine.hpp"

namespace usbs {

// out[c] = sum_{t in group c} a[t] * sum_q U[i_t, q] U[j_t, q]  (U row-major n x p).
__global__ void __launch_bounds__(256)
k_apply_A_operator(long long m, int p, const int* __restrict__ gptr, const int* __restrict__ gi,
                   const int* __restrict__ gj, const double* __restrict__ ga,
                   const double* __restrict__ U, double* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  double s = 0.0;
  for (int t = gptr[c]; t < gptr[c + 1]; ++t) {
    const double* ui = U + (long long)gi[t] * p;
    const double* uj = U + (long long)gj[t] * p;
    double d = 0.0;
    for (int q = 0; q < p; ++q) d += ui[q] * uj[q];
    s += ga[t] * d;
  }
  out[c] = s;
}

// out[i, q] = sum_{t in group i} a[t] z[c_t] U[j_t, q]  (out, U row-major n x p).
__global__ void __launch_bounds__(256)
k_apply_A_adjoint(long long n, int p, const int* __restrict__ gptr, const int* __restrict__ gc,
                  const int* __restrict__ gj, const double* __restrict__ ga,
                  const double* __restrict__ z, const double* __restrict__ U,
                  double* __restrict__ out) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * p) return;
  const long long i = idx / p;
  const int q = (int)(idx - i * p);
  double s = 0.0;
  for (int t = gptr[i]; t < gptr[i + 1]; ++t) s += ga[t] * z[gc[t]] * U[(long long)gj[t] * p + q];
  out[idx] = s;
}

namespace {

struct Grouped {
  std::vector<int> ptr, a, b;  // group pointers; the two other indices of each triple
  std::vector<double> data;
};

// Group the triples by column `key` of A_indices (0: constraint, 1: row),
// input order within a group; a and b receive columns oa and ob.
Grouped group_triples(int64_t ngroups, int64_t nA, const int32_t* Ai, const double* Ad, int key,
                      int oa, int ob) {
  Grouped g;
  g.ptr.assign((size_t)ngroups + 1, 0);
  for (int64_t t = 0; t < nA; ++t) g.ptr[(size_t)Ai[3 * t + key] + 1]++;
  for (int64_t x = 0; x < ngroups; ++x) g.ptr[x + 1] += g.ptr[x];
  g.a.resize(nA > 0 ? nA : 1);
  g.b.resize(nA > 0 ? nA : 1);
  g.data.resize(nA > 0 ? nA : 1);
  std::vector<int> fill(g.ptr.begin(), g.ptr.end() - 1);
  for (int64_t t = 0; t < nA; ++t) {
    const int e = fill[Ai[3 * t + key]]++;
    g.a[e] = Ai[3 * t + oa];
    g.b[e] = Ai[3 * t + ob];
    g.data[e] = Ad[t];
  }
  return g;
}

int check_triples(Ctx* c, int64_t n, int64_t m, int64_t nA, const int32_t* Ai, const double* Ad) {
  if (n < 1 || m < 1 || nA < 0 || n > INT32_MAX || m > INT32_MAX || nA > INT32_MAX)
    return set_err(c, USBS_ERR_INVALID, "invalid sizes");
  if (nA > 0 && (!Ai || !Ad)) return set_err(c, USBS_ERR_INVALID, "null triple arrays");
  for (int64_t t = 0; t < nA; ++t) {
    const int32_t cc = Ai[3 * t], ii = Ai[3 * t + 1], jj = Ai[3 * t + 2];
    if (cc < 0 || cc >= m || ii < 0 || ii >= n || jj < 0 || jj >= n)
      return set_err(c, USBS_ERR_INVALID, "A_indices out of range at triple " + std::to_string(t));
    if (!std::isfinite(Ad[t])) return set_err(c, USBS_ERR_NONFINITE, "non-finite A_data");
  }
  return USBS_OK;
}

struct Pool {
  Ctx* c;
  std::vector<DevBuf> bufs;
  explicit Pool(Ctx* cc) : c(cc) {}
  ~Pool() { ctx_free_pool(bufs); }
  template <typename X>
  X* up(const X* h, size_t count) {
    X* d = (X*)ctx_alloc(c, bufs, sizeof(X) * (count ? count : 1));
    if (d && count && hipMemcpy(d, h, sizeof(X) * count, hipMemcpyHostToDevice) != hipSuccess)
      return nullptr;
    return d;
  }
  template <typename X>
  X* alloc(size_t count) { return (X*)ctx_alloc(c, bufs, sizeof(X) * (count ? count : 1)); }
};

}  // namespace
}  // namespace usbs

using namespace usbs;

extern "C" int usbs_apply_A_operator(usbs_ctx* h, int64_t n, int64_t m, int64_t nA,
                                     const int32_t* A_indices, const double* A_data, int32_t p,
                                     const double* U, double* out) {
  if (!h) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (!U || !out || p < 1) return set_err(c, USBS_ERR_INVALID, "null U / out or p < 1");
  int rc = check_triples(c, n, m, nA, A_indices, A_data);
  if (rc) return rc;
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  const Grouped g = group_triples(m, nA, A_indices, A_data, 0, 1, 2);
  Pool S(c);
  const int* dptr = S.up(g.ptr.data(), g.ptr.size());
  const int* di = S.up(g.a.data(), g.a.size());
  const int* dj = S.up(g.b.data(), g.b.size());
  const double* da = S.up(g.data.data(), g.data.size());
  const double* dU = S.up(U, (size_t)n * p);
  double* dout = S.alloc<double>((size_t)m);
  if (!dptr || !di || !dj || !da || !dU || !dout) return USBS_ERR_OOM;
  hipStream_t s = c->stream;
  k_apply_A_operator<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(m, p, dptr, di, dj, da, dU, dout);
  USBS_HIP(c, hipGetLastError());
  USBS_HIP(c, hipMemcpyAsync(out, dout, sizeof(double) * m, hipMemcpyDeviceToHost, s));
  USBS_HIP(c, hipStreamSynchronize(s));
  return USBS_OK;
}

extern "C" int usbs_apply_A_adjoint(usbs_ctx* h, int64_t n, int64_t m, int64_t nA,
                                    const int32_t* A_indices, const double* A_data,
                                    const double* z, int32_t p, const double* U, double* out) {
  if (!h) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (!z || !U || !out || p < 1) return set_err(c, USBS_ERR_INVALID, "null z / U / out or p < 1");
  int rc = check_triples(c, n, m, nA, A_indices, A_data);
  if (rc) return rc;
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  const Grouped g = group_triples(n, nA, A_indices, A_data, 1, 0, 2);
  Pool S(c);
  const int* dptr = S.up(g.ptr.data(), g.ptr.size());
  const int* dc = S.up(g.a.data(), g.a.size());
  const int* dj = S.up(g.b.data(), g.b.size());
  const double* da = S.up(g.data.data(), g.data.size());
  const double* dz = S.up(z, (size_t)m);
  const double* dU = S.up(U, (size_t)n * p);
  double* dout = S.alloc<double>((size_t)n * p);
  if (!dptr || !dc || !dj || !da || !dz || !dU || !dout) return USBS_ERR_OOM;
  hipStream_t s = c->stream;
  const long long tot = n * (long long)p;
  k_apply_A_adjoint<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, p, dptr, dc, dj, da, dz, dU, dout);
  USBS_HIP(c, hipGetLastError());
  USBS_HIP(c, hipMemcpyAsync(out, dout, sizeof(double) * tot, hipMemcpyDeviceToHost, s));
  USBS_HIP(c, hipStreamSynchronize(s));
  return USBS_OK;
}
