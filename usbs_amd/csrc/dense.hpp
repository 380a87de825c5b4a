// Small dense FP64 linear algebra for one workgroup, operands in LDS.
//
// Used by the Ritz solve (m <= 64), the IPM subproblems (k <= 12,
// svec dimension d = k(k+1)/2 <= 78), the bundle eigendecomposition and the
// Nystrom core.  All matrices are row-major with leading dimension = their
// column count.  Every routine is called by all threads of the block and
// ends with the block synchronised.
#pragma once
#include "dev_common.hpp"

namespace usbs {

// ---------------------------------------------------------------------------
// Symmetric eigendecomposition by parallel cyclic Jacobi (round-robin
// ordering of disjoint rotation pairs, schedule precomputed in LDS).  A
// rotation is skipped when |a_pq| <= 1e-16 |A|_F (backward error at the level
// of a LAPACK eigensolver) and the sweeps stop after a sweep that rotates
// nothing.  A (m x m, m <= 64) is destroyed.  On exit w[0..m) holds the
// eigenvalues in ascending order and column j of V the eigenvector of w[j].
// tmp: m*m doubles; coef: 2*m doubles; part: m*m + 2*m + 4 ints; red: >= 16
// doubles.  Returns the number of sweeps.
__device__ int dense_eigh(double* A, int m, double* w, double* V, double* tmp, double* coef,
                          int* part, double* red) {
  const int t = threadIdx.x, nt = blockDim.x;
  const int sh = (m <= 32) ? 5 : 6;          // threads per row of the 2-D loops
  const int tx = t & ((1 << sh) - 1), ty = t >> sh, ny = nt >> sh;
  const int M = m + (m & 1);
  const int half = M / 2;
  int* flag = part + m;                      // any rotation in the current sweep
  int* sched = part + m + 2;                 // (M-1) rounds x half pairs x (p, q)
  double fro = 0.0;
  for (int i = ty; i < m; i += ny)
    for (int j = tx; j < m; j += (1 << sh)) {
      V[i * m + j] = (i == j) ? 1.0 : 0.0;
      fro += A[i * m + j] * A[i * m + j];
    }
  for (int idx = t; idx < (M - 1) * half; idx += nt) {
    const int r = idx / half, kp = idx - (idx / half) * half;
    const int pa = (kp == 0) ? 0 : 1 + ((kp - 1 + r) % (M - 1));
    const int pb = 1 + ((M - 2 - kp + r) % (M - 1));
    sched[2 * idx] = pa < pb ? pa : pb;
    sched[2 * idx + 1] = pa < pb ? pb : pa;
  }
  fro = block_sum1(fro, red);
  const double thr = 1e-16 * sqrt(fro);
  double* Ac = A;
  double* An = tmp;
  int sweep = 0;
  for (; sweep < 40 && m > 1; ++sweep) {
    if (t == 0) flag[0] = 0;
    for (int r = 0; r < M - 1; ++r) {
      const int* pr = sched + 2 * r * half;
      for (int kp = t; kp < half; kp += nt) {
        const int p = pr[2 * kp], q = pr[2 * kp + 1];
        if (q >= m) {
          part[p] = p; coef[2 * p] = 1.0; coef[2 * p + 1] = 0.0;
          continue;
        }
        const double apq = Ac[p * m + q], app = Ac[p * m + p], aqq = Ac[q * m + q];
        double c = 1.0, s = 0.0;
        if (fabs(apq) > thr) {
          const double theta = (aqq - app) / (2.0 * apq);
          const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          c = 1.0 / sqrt(tt * tt + 1.0);
          s = tt * c;
          flag[0] = 1;
        }
        part[p] = q; part[q] = p;
        coef[2 * p] = c; coef[2 * p + 1] = -s;
        coef[2 * q] = c; coef[2 * q + 1] = s;
      }
      __syncthreads();
      for (int i = ty; i < m; i += ny) {
        const int pi = part[i];
        const double ai = coef[2 * i], bi = coef[2 * i + 1];
        const double* Ai = Ac + i * m;
        const double* Api = Ac + pi * m;
        for (int j = tx; j < m; j += (1 << sh)) {
          const int pj = part[j];
          const double aj = coef[2 * j], bj = coef[2 * j + 1];
          double v = ai * (aj * Ai[j] + bj * Ai[pj]) + bi * (aj * Api[j] + bj * Api[pj]);
          if (pi == j && i != j) v = 0.0;
          An[i * m + j] = v;
        }
      }
      // V <- V J, in place per (row, pair)
      for (int idx = t; idx < m * half; idx += nt) {
        const int i = idx / half, kp = idx - (idx / half) * half;
        const int p = pr[2 * kp], q = pr[2 * kp + 1];
        if (q >= m) continue;
        const double vp = V[i * m + p], vq = V[i * m + q];
        V[i * m + p] = coef[2 * p] * vp + coef[2 * p + 1] * vq;
        V[i * m + q] = coef[2 * q] * vq + coef[2 * q + 1] * vp;
      }
      __syncthreads();
      double* sw = Ac; Ac = An; An = sw;
    }
    const int done = (flag[0] == 0);
    __syncthreads();
    if (done) break;
  }
  // sort ascending (stable in index), permute eigenvector columns via An
  for (int i = t; i < m; i += nt) {
    const double di = Ac[i * m + i];
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double dj = Ac[j * m + j];
      rank += (dj < di) || (dj == di && j < i);
    }
    part[i] = rank;
  }
  __syncthreads();
  for (int i = t; i < m; i += nt) w[part[i]] = Ac[i * m + i];
  for (int i = ty; i < m; i += ny)
    for (int j = tx; j < m; j += (1 << sh)) An[i * m + part[j]] = V[i * m + j];
  __syncthreads();
  for (int i = ty; i < m; i += ny)
    for (int j = tx; j < m; j += (1 << sh)) V[i * m + j] = An[i * m + j];
  __syncthreads();
  return sweep;
}

// ---------------------------------------------------------------------------
// Solve A x = b by Gaussian elimination with partial pivoting (first maximal
// |pivot|, as LAPACK's idamax; same pivots and elimination as getrf/getrs).
// A (d x d) and b are destroyed; b returns x.  Column k is eliminated with
// the pivot row in place, then wave 0 swaps it into position k and searches
// the next pivot, so each column costs two block barriers.  A zero pivot
// yields inf/NaN, as getrs does.
__device__ void dense_lu_solve1(double* A, int d, double* b, int* ired) {
  const int t = threadIdx.x, nt = blockDim.x;
  auto search = [&](int k) {  // wave 0 only
    double best = -1.0;
    int bi = k;
    for (int i = k + t; i < d; i += 64) {
      const double v = fabs(A[i * d + k]);
      if (v > best) { best = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if (t == 0) ired[0] = bi;
  };
  if (t < 64) search(0);
  __syncthreads();
  for (int k = 0; k < d; ++k) {
    const int piv = ired[0];
    const double* Ap = A + piv * d;
    const double akk = Ap[k];
    const double bp = b[piv];
    // eliminate column k from every remaining row except the pivot row
    // rows k..d-1 over groups of 32 threads; columns k+1..d (column d = rhs)
    for (int i = k + (t >> 5); i < d; i += (nt >> 5)) {
      if (i == piv) continue;
      const double l = A[i * d + k] / akk;
      for (int j = k + 1 + (t & 31); j <= d; j += 32) {
        if (j < d) A[i * d + j] -= l * Ap[j];
        else b[i] -= l * bp;
      }
    }
    __syncthreads();
    if (t < 64) {
      // zero the eliminated column entries and swap rows k <-> piv
      for (int i = k + t; i < d; i += 64)
        if (i != piv) A[i * d + k] = 0.0;
      if (piv != k) {
        for (int j = k + t; j < d; j += 64) {
          const double a = A[k * d + j]; A[k * d + j] = A[piv * d + j]; A[piv * d + j] = a;
        }
        if (t == 0) { const double x = b[k]; b[k] = b[piv]; b[piv] = x; }
      }
      if (k + 1 < d) search(k + 1);
    }
    __syncthreads();
  }
  // back substitution by wave 0
  if (t < 64) {
    for (int k = d - 1; k >= 0; --k) {
      const double xk = b[k] / A[k * d + k];
      for (int i = t; i < k; i += 64) b[i] -= A[i * d + k] * xk;
      if (t == 0) b[k] = xk;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// In-place LDL^T of a symmetric positive definite d x d matrix (lower
// triangle used, no pivoting): one block barrier per column.  On exit
// A[i][k] (i > k) holds the unscaled column (L[i][k] * D[k]) and A[k][k] = D[k].
__device__ void dense_ldlt(double* A, int d) {
  const int t = threadIdx.x, nt = blockDim.x;
  const int tr = t >> 4, nr = nt >> 4, tc = t & 15;
  for (int k = 0; k < d - 1; ++k) {
    const double dk = A[k * d + k];
    // trailing update A[i][j] -= A[i][k] A[j][k] / dk for k < j <= i
    for (int i = k + 1 + tr; i < d; i += nr) {
      const double lik = A[i * d + k] / dk;
      for (int j = k + 1 + tc; j <= i; j += 16) A[i * d + j] -= lik * A[j * d + k];
    }
    __syncthreads();
  }
}

// Solve (L D L^T) X = B for nrhs right-hand sides with the dense_ldlt factor
// (B row-major d x nrhs, overwritten).  One wavefront per right-hand side.
__device__ void dense_ldlt_solve(const double* A, int d, double* B, int nrhs) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (w < nrhs) {
    // forward: y_i = b_i - sum_{k<i} (A[i][k] / D[k]) y_k  (column-oriented)
    for (int k = 0; k < d; ++k) {
      const double yk = B[k * nrhs + w];
      const double dk = A[k * d + k];
      for (int i = k + 1 + lane; i < d; i += 64) B[i * nrhs + w] -= (A[i * d + k] / dk) * yk;
    }
    // diagonal and backward: x_k = y_k / D[k] - sum_{i>k} (A[i][k] / D[k]) x_i
    for (int k = d - 1; k >= 0; --k) {
      const double dk = A[k * d + k];
      double s = 0.0;
      for (int i = k + 1 + lane; i < d; i += 64) s += (A[i * d + k] / dk) * B[i * nrhs + w];
      s = wave_sum(s);
      if (lane == 0) B[k * nrhs + w] = B[k * nrhs + w] / dk - s;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Cholesky A = L L^T of a symmetric m x m matrix in place (lower triangle
// holds L, strict upper set to zero).  Returns false (for every thread) if a
// pivot is not positive or not finite.
__device__ bool dense_cholesky(double* A, int m, int* flag) {
  const int t = threadIdx.x, nt = blockDim.x;
  if (t == 0) flag[0] = 1;
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    if (t == 0) {
      const double d = A[k * m + k];
      if (!(d > 0.0) || !(d < INFINITY)) flag[0] = 0;
      A[k * m + k] = sqrt(d);
    }
    __syncthreads();
    if (!flag[0]) return false;
    const double lkk = A[k * m + k];
    for (int i = k + 1 + t; i < m; i += nt) A[i * m + k] /= lkk;
    __syncthreads();
    const int r = m - k - 1;
    for (int idx = t; idx < r * r; idx += nt) {
      const int i = k + 1 + idx / r, j = k + 1 + idx % r;
      if (j <= i) A[i * m + j] -= A[i * m + k] * A[j * m + k];
    }
    __syncthreads();
  }
  for (int idx = t; idx < m * m; idx += nt)
    if (idx % m > idx / m) A[idx] = 0.0;
  __syncthreads();
  return true;
}

// Per-thread positive-definiteness test of M + s*dM (k x k, row-major, in
// LDS) by an unpivoted Cholesky in the thread's own scratch (k*k doubles).
__device__ __forceinline__ bool thread_is_pd(const double* M, const double* dM, double s, int k,
                                             double* L) {
  for (int i = 0; i < k; ++i)
    for (int j = 0; j <= i; ++j) L[i * k + j] = M[i * k + j] + s * dM[i * k + j];
  for (int j = 0; j < k; ++j) {
    double d = L[j * k + j];
    for (int p = 0; p < j; ++p) d -= L[j * k + p] * L[j * k + p];
    if (!(d > 0.0)) return false;
    d = sqrt(d);
    L[j * k + j] = d;
    for (int i = j + 1; i < k; ++i) {
      double v = L[i * k + j];
      for (int p = 0; p < j; ++p) v -= L[i * k + p] * L[j * k + p];
      L[i * k + j] = v / d;
    }
  }
  return true;
}

// svec index a <-> (b, c), b >= c, in numpy.tril_indices (row-major) order.
__device__ __forceinline__ void svec_pair(int a, int& b, int& c) {
  b = (int)((sqrt(8.0 * a + 1.0) - 1.0) * 0.5);
  while ((b + 1) * (b + 2) / 2 <= a) ++b;
  while (b * (b + 1) / 2 > a) --b;
  c = a - b * (b + 1) / 2;
}

// v = svec(M) for symmetric-ish k x k M (off-diagonal: (M_bc + M_cb)/sqrt 2).
__device__ __forceinline__ void dense_svec(const double* M, int k, double* v) {
  const int d = k * (k + 1) / 2;
  const double u = 0.70710678118654752440;
  for (int a = threadIdx.x; a < d; a += blockDim.x) {
    int b, c;
    svec_pair(a, b, c);
    v[a] = (b == c) ? M[b * k + c] : (M[b * k + c] * u + M[c * k + b] * u);
  }
}

// M = svec^{-1}(v).
__device__ __forceinline__ void dense_svec_inv(const double* v, int k, double* M) {
  const int d = k * (k + 1) / 2;
  const double u = 0.70710678118654752440;
  for (int a = threadIdx.x; a < d; a += blockDim.x) {
    int b, c;
    svec_pair(a, b, c);
    if (b == c) M[b * k + c] = v[a];
    else { M[b * k + c] = v[a] * u; M[c * k + b] = v[a] * u; }
  }
}

}  // namespace usbs
