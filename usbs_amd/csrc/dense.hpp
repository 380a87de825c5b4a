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
// The same parallel cyclic Jacobi for m <= 64 with the data laid out for a
// whole block: each round, threads t < M/2 compute the rotations of the
// round-robin pairing, then every thread applies the two-sided rotation to
// whole 2x2 blocks (pair a rows x pair b columns) and the one-sided rotation
// to (row, pair) entries of V, with one barrier after each half.  Rows are
// padded (lda = M + 1) so that threads reading one column of consecutive rows
// hit different banks.
//   A : M x lda in LDS, M = m rounded up to even; when m is odd row/column m
//       must be zero.  Destroyed.
//   Vw: M x lda LDS scratch.  cs: M doubles.  sched: (M-1) * M ints.
//   w : m eigenvalues ascending;  Y: m x m row-major, column j = eigenvector
//       of w[j] (stable order for equal eigenvalues).
// Threshold and stopping rule as dense_eigh (|a_pq| <= 1e-16 |A|_F is set to
// zero; stop after a sweep without a rotation).  Called by all threads of the
// block (blockDim a multiple of 64, at least M/2); returns the sweeps.
__device__ int block_eigh(double* A, int lda, int m, double* w, double* Y, double* Vw, double* cs,
                          int* sched, double* red) {
  const int t = threadIdx.x, nt = blockDim.x;
  const int M = m + (m & 1), half = M / 2;
  if (m == 1) {
    if (t == 0) { w[0] = A[0]; Y[0] = 1.0; }
    __syncthreads();
    return 0;
  }
  for (int idx = t; idx < (M - 1) * half; idx += nt) {
    const int r = idx / half, kp = idx - (idx / half) * half;
    const int pa = (kp == 0) ? 0 : 1 + ((kp - 1 + r) % (M - 1));
    const int pb = 1 + ((M - 2 - kp + r) % (M - 1));
    sched[2 * idx] = pa < pb ? pa : pb;
    sched[2 * idx + 1] = pa < pb ? pb : pa;
  }
  // thread decomposition: column pair bl, row pairs al0, al0 + astep, ...
  const int astep = nt / half;
  const int bl = t % half, al0 = t / half;
  const bool act = t < astep * half;
  double fro = 0.0;
  if (act)
    for (int i = al0; i < M; i += astep)
      for (int j = bl; j < M; j += half) {
        Vw[i * lda + j] = (i == j) ? 1.0 : 0.0;
        fro += A[i * lda + j] * A[i * lda + j];
      }
  fro = block_sum1(fro, red);
  const double thr = 1e-16 * sqrt(fro);
  __syncthreads();
  int sweep = 0;
  for (; sweep < 40; ++sweep) {
    int any = 0;
    for (int r = 0; r < M - 1; ++r) {
      const int* pr = sched + 2 * r * half;
      int rot = 0;
      if (t < half) {
        const int p = pr[2 * t], q = pr[2 * t + 1];
        double c = 1.0, s = 0.0;
        if (q < m) {
          const double apq = A[p * lda + q];
          if (fabs(apq) > thr) {
            const double app = A[p * lda + p], aqq = A[q * lda + q];
            const double theta = (aqq - app) / (2.0 * apq);
            const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(tt * tt + 1.0);
            s = tt * c;
            rot = 1;
          }
        }
        cs[2 * t] = c;
        cs[2 * t + 1] = s;
      }
      any |= __syncthreads_or(rot);
      if (act) {
        const int pb = pr[2 * bl], qb = pr[2 * bl + 1];
        const double cb = cs[2 * bl], sb = cs[2 * bl + 1];
        for (int a = al0; a < half; a += astep) {
          const int pa = pr[2 * a], qa = pr[2 * a + 1];
          const double ca = cs[2 * a], sa = cs[2 * a + 1];
          double* rp = A + pa * lda;
          double* rq = A + qa * lda;
          const double x00 = rp[pb], x01 = rp[qb], x10 = rq[pb], x11 = rq[qb];
          const double y00 = ca * x00 - sa * x10, y01 = ca * x01 - sa * x11;
          const double y10 = sa * x00 + ca * x10, y11 = sa * x01 + ca * x11;
          double z00 = cb * y00 - sb * y01, z01 = sb * y00 + cb * y01;
          double z10 = cb * y10 - sb * y11, z11 = sb * y10 + cb * y11;
          if (a == bl && qa < m) { z01 = 0.0; z10 = 0.0; }
          rp[pb] = z00; rp[qb] = z01; rq[pb] = z10; rq[qb] = z11;
        }
        if (qb < m)
          for (int i = al0; i < m; i += astep) {
            double* vi = Vw + i * lda;
            const double vp = vi[pb], vq = vi[qb];
            vi[pb] = cb * vp - sb * vq;
            vi[qb] = cb * vq + sb * vp;
          }
      }
      __syncthreads();
    }
    if (!any) { ++sweep; break; }
  }
  // ascending order, ties by index; column permutation of V into Y
  for (int i = t; i < m; i += nt) {
    const double di = A[i * lda + i];
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double dj = A[j * lda + j];
      rank += (dj < di) || (dj == di && j < i);
    }
    w[rank] = di;
    for (int x = 0; x < m; ++x) Y[x * m + rank] = Vw[x * lda + i];
  }
  __syncthreads();
  return sweep;
}

// ---------------------------------------------------------------------------
// LDL^T of a symmetric d x d matrix without pivoting (the IPM's SPD Newton
// block), in place: on return the strict lower triangle holds the unit factor
// L and rD[k] = 1 / D[k].  Right-looking: step k updates the trailing lower
// triangle with column k (16 threads per row).  A zero pivot yields inf/NaN,
// as an LU solve of a singular matrix does.
__device__ void dense_ldlt(double* A, int d, double* rD) {
  const int t = threadIdx.x, nt = blockDim.x;
  const int tr = t >> 4, nr = nt >> 4, tc = t & 15;
  for (int k = 0; k < d - 1; ++k) {
    const double rdk = 1.0 / A[k * d + k];
    // trailing update A[i][j] -= (A[i][k] / D_k) A[j][k] for k < j <= i
    for (int i = k + 1 + tr; i < d; i += nr) {
      const double lik = A[i * d + k] * rdk;
      for (int j = k + 1 + tc; j <= i; j += 16) A[i * d + j] -= lik * A[j * d + k];
    }
    __syncthreads();
  }
  for (int k = t; k < d; k += nt) rD[k] = 1.0 / A[k * d + k];
  __syncthreads();
  for (int i = tr; i < d; i += nr)
    for (int j = tc; j < i; j += 16) A[i * d + j] *= rD[j];
  __syncthreads();
}

// Broadcast lane `src` (uniform) of a double to the whole wavefront.
__device__ __forceinline__ double readlane_d(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Solve (L D L^T) X = B with the dense_ldlt factor for nrhs right-hand sides
// (B row-major d x ldb, columns 0..nrhs-1 overwritten), d <= 128.  One
// wavefront per right-hand side holds it in registers (rows lane and
// lane + 64); forward and backward substitution are column-oriented, so each
// step is one broadcast and one FMA per row.
__device__ void dense_ldlt_solve(const double* A, int d, const double* rD, double* B, int ldb,
                                 int nrhs) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (w < nrhs) {
    const int i0 = lane, i1 = lane + 64;
    double y0 = (i0 < d) ? B[i0 * ldb + w] : 0.0;
    double y1 = (i1 < d) ? B[i1 * ldb + w] : 0.0;
    for (int k = 0; k < d; ++k) {
      const double yk = (k < 64) ? readlane_d(y0, k) : readlane_d(y1, k - 64);
      if (i0 > k && i0 < d) y0 -= A[i0 * d + k] * yk;
      if (i1 > k && i1 < d) y1 -= A[i1 * d + k] * yk;
    }
    if (i0 < d) y0 *= rD[i0];
    if (i1 < d) y1 *= rD[i1];
    for (int k = d - 1; k > 0; --k) {
      const double xk = (k < 64) ? readlane_d(y0, k) : readlane_d(y1, k - 64);
      if (i0 < k) y0 -= A[k * d + i0] * xk;
      if (i1 < k) y1 -= A[k * d + i1] * xk;
    }
    if (i0 < d) B[i0 * ldb + w] = y0;
    if (i1 < d) B[i1 * ldb + w] = y1;
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
