// Small kernels and host helpers shared by the USBS and CGAL drivers.
#pragma once
#include <cmath>
#include <vector>

#include "lanczos.hpp"

namespace usbs {

template <typename T>
__global__ void k_load_colmajor(long long n, int r, const double* src, T* dst, long long ld) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * r) return;
  const long long row = i / r, c = i % r;
  dst[row + c * ld] = (T)src[i];
}

template <typename T>
__global__ void k_store_rowmajor(long long n, int r, const T* src, long long ld, double* dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * r) return;
  const long long row = i / r, c = i % r;
  dst[i] = (double)src[row + c * ld];
}

template <typename T>
__global__ void k_fill(T* p, long long count, T v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// dst = (T)((src / sx_old) * sx_new): a state kept in the units of SCALE_X =
// sx_old taken to sx_new, with the operations of unscaling then scaling it
// (utils/common.py:29-96 with SCALE_A = 1), so that the result equals that of
// the same transform on the host bit for bit.
template <typename TD>
__global__ void k_rescale(long long count, const double* __restrict__ src, TD* __restrict__ dst,
                          double sx_old, double sx_new) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (TD)((src[i] / sx_old) * sx_new);
}

template <typename TS, typename TD>
__global__ void k_cast(long long count, const TS* __restrict__ src, TD* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (TD)src[i];
}

// Lanczos per-call reset: Q[:, 0] = q0, alpha = beta = 0, control and barrier 0
// (keep_error: the barrier error word stays set across calls).
template <typename T>
__global__ void k_lz_start(LzInst<T> L, const double* q0, int keep_error = 0) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < L.n) L.Q[i] = (T)q0[i];
  if (i < L.m) { L.alpha[i] = 0.0; L.beta[i] = 0.0; }
  if (i < LZ_NCTRL && !(keep_error && i == LZ_ERROR)) L.ctrl[i] = 0;
  for (long long x = i; x < (long long)LZ_BAR_STRIDE * L.nbar; x += (long long)gridDim.x * blockDim.x) L.bar[x] = 0u;
  for (long long x = i; x < 2 * L.nbar; x += (long long)gridDim.x * blockDim.x) L.lstat[x] = 0;
}

inline std::vector<long long> geometric_schedule(long long maxiter, double base) {
  // utils/loop.py:85-98
  const double num = std::ceil(std::log((base - 1) * maxiter - 1) / std::log(base)) + 1;
  std::vector<long long> steps;
  long long total = 0;
  for (long long i = 0; i < (long long)num; ++i) {
    const long long st = (long long)std::pow(base, (double)i);
    steps.push_back(st);
  }
  std::vector<long long> kept;
  long long cum = 0;
  for (long long st : steps) {
    cum += st;
    if (cum < maxiter) { kept.push_back(st); total += st; }
  }
  kept.push_back(maxiter - total);
  return kept;
}


}  // namespace usbs
