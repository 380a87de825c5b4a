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
// Broadcast lane `src` (uniform) of a double to the whole wavefront.
__device__ __forceinline__ double readlane_d(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
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

// ---------------------------------------------------------------------------
// Symmetric eigendecomposition for m <= 32 by Householder tridiagonalisation
// (the reflectors accumulated into Q as they are formed), multisection on
// Sturm counts for every eigenvalue, eigenvectors of the tridiagonal from
// twisted factorisations, Gram-Schmidt (two passes) inside clusters
// (consecutive eigenvalues closer than 1e-3 |T|), and Y = Q Z.  Eigenvalues
// come out in ascending order by construction; each is bracketed to
// max(2 eps |lambda|, eps |T|).  The code is written for latency: small
// reductions are done redundantly by every thread from LDS broadcasts
// rather than with cross-lane shuffles, and the Sturm and twisted-
// factorisation chains run on a register copy of the tridiagonal with a
// Newton-refined reciprocal.
//   A  : m x lda in LDS (lda >= m), symmetric; destroyed.
//   Qw : m x lda LDS scratch (the accumulated orthogonal factor).
//   w  : m eigenvalues ascending;  Y: m x m row-major, column j = eigenvector
//        of w[j].
//   scr: tridiag_scratch_doubles(m, blockDim.x) doubles of LDS.
// Called by all threads of the block (blockDim 256).  Returns 1 (in every
// thread) when a cluster is degenerate -- a vector loses more than half its
// norm to the Gram-Schmidt projection, as for an exactly repeated eigenvalue
// -- and the caller must use another eigensolver; 0 otherwise.
#define TD_MAXM 32
__host__ __device__ constexpr int tridiag_scratch_doubles(int m, int nt) {
  return 9 * m + 64 + (nt + 1) / 2 + 3 * m * m;
}

__device__ __forceinline__ double td_guard(double q, double pivmin) {
  return fabs(q) < pivmin ? -pivmin : q;
}

// 1 / q: hardware reciprocal refined by one Newton step.
__device__ __forceinline__ double td_rcp(double q) {
  double r = __builtin_amdgcn_rcp(q);
  const double e = fma(-q, r, 1.0);
  return fma(r, e, r);
}

// Number of eigenvalues of the tridiagonal (register copy dr, e2r; entries
// at and past m ignored) below x.  Straight-line code: steps past m keep q.
__device__ __forceinline__ int td_sturm(const double (&dr)[TD_MAXM], const double (&e2r)[TD_MAXM],
                                        int m, double x, double pivmin) {
  double q = td_guard(dr[0] - x, pivmin);
  int c = q < 0.0 ? 1 : 0;
#pragma unroll
  for (int i = 1; i < TD_MAXM; ++i) {
    const double qn = td_guard((dr[i] - x) - e2r[i - 1] * td_rcp(q), pivmin);
    const bool in = i < m;
    c += (in && qn < 0.0) ? 1 : 0;
    q = in ? qn : q;
  }
  return c;
}

__device__ int tridiag_eigh(double* A, int lda, int m, double* w, double* Y, double* Qw,
                            double* scr) {
  constexpr int MT = TD_MAXM;
  const int t = threadIdx.x, nt = blockDim.x, lane = t & 63;
  double* d = scr;           // diagonal
  double* e = d + m;         // e[i] couples i and i + 1
  double* e2 = e + m;
  double* hv = e2 + m;       // current reflector, by absolute row (hv[j+1] = 1)
  double* p = hv + m;        // A22 v by absolute row, then Gram-Schmidt coefficients
  double* lo = p + m;
  double* hi = lo + m;
  double* misc = hi + m;     // 2m + 64 doubles: [0] tnorm, [1] pivmin, [2] tau, [8..8+m) Q v,
                             // [8+m..8+2m) p*v
  int* cnt = (int*)(misc + 2 * m + 64);                  // nt ints
  double* dp = misc + 2 * m + 64 + (nt + 1) / 2;         // m x m: D+ of eigenvector j (row j)
  double* dm = dp + m * m;                               // m x m: D-
  double* Z = dm + m * m;                                // m x m: eigenvectors of the tridiagonal
  double* qv = misc + 8;
  double* pvp = qv + m;     // p_r v_r of the current step
  for (int idx = t; idx < m * m; idx += nt) {
    const int r = idx / m, c = idx - (idx / m) * m;
    Qw[r * lda + c] = r == c ? 1.0 : 0.0;
  }
  // 1. Householder: step j maps A[j+1.., j] to (|x|, 0, ...), Q <- Q H_j
  for (int j = 0; j + 2 < m; ++j) {
    const int n1 = m - j - 1;
    __syncthreads();
    if (t < 64) {
      const double xi = lane < n1 ? A[(j + 1 + lane) * lda + j] : 0.0;
      const double sg = wave_sum((lane >= 1 && lane < n1) ? xi * xi : 0.0);
      const double x0 = __shfl(xi, 0, 64);
      double v0 = 1.0, mu = x0, tj = 0.0;
      if (sg != 0.0) {
        mu = sqrt(x0 * x0 + sg);
        v0 = x0 <= 0.0 ? x0 - mu : -sg / (x0 + mu);
        tj = 2.0 * v0 * v0 / (sg + v0 * v0);
      }
      if (lane < n1) hv[j + 1 + lane] = lane == 0 ? 1.0 : (sg != 0.0 ? xi / v0 : 0.0);
      if (lane == 0) {
        misc[2] = tj;
        e[j] = mu;
        d[j] = A[j * lda + j];
      }
    }
    __syncthreads();
    const double tj = misc[2];
    if (tj == 0.0) continue;  // uniform
    {
      // p = tau A22 v: 8 threads per row
      const int g = t >> 3, sub = t & 7;
      const bool isp = g < n1;
      double s = 0.0;
      if (isp)
        for (int c = sub; c < n1; c += 8) s = fma(A[(j + 1 + g) * lda + j + 1 + c], hv[j + 1 + c], s);
      s += __shfl_xor(s, 4, 8);
      s += __shfl_xor(s, 2, 8);
      s += __shfl_xor(s, 1, 8);
      if (isp && sub == 0) {
        p[j + 1 + g] = tj * s;
        pvp[g] = tj * s * hv[j + 1 + g];
      }
    }
    {
      // (Q v)_i: 8 threads per row of Q
      const int g = t >> 3, sub = t & 7;
      double s = 0.0;
      if (g < m)
        for (int c = sub; c < n1; c += 8) s = fma(Qw[g * lda + j + 1 + c], hv[j + 1 + c], s);
      s += __shfl_xor(s, 4, 8);
      s += __shfl_xor(s, 2, 8);
      s += __shfl_xor(s, 1, 8);
      if (g < m && sub == 0) qv[g] = s;
    }
    __syncthreads();
    double K0 = 0.0, K1 = 0.0, K2 = 0.0, K3 = 0.0;
#pragma unroll
    for (int c = 0; c < MT; c += 4) {
      K0 += c < n1 ? pvp[c < n1 ? c : 0] : 0.0;
      K1 += c + 1 < n1 ? pvp[c + 1 < n1 ? c + 1 : 0] : 0.0;
      K2 += c + 2 < n1 ? pvp[c + 2 < n1 ? c + 2 : 0] : 0.0;
      K3 += c + 3 < n1 ? pvp[c + 3 < n1 ? c + 3 : 0] : 0.0;
    }
    const double K = 0.5 * tj * ((K0 + K1) + (K2 + K3));
    for (int idx = t; idx < n1 * n1; idx += nt) {
      const int r = j + 1 + idx / n1, c = j + 1 + idx - (idx / n1) * n1;
      const double wr = p[r] - K * hv[r], wc = p[c] - K * hv[c];
      A[r * lda + c] -= hv[r] * wc + wr * hv[c];
    }
    for (int idx = t; idx < m * n1; idx += nt) {
      const int i = idx / n1, c = j + 1 + idx - (idx / n1) * n1;
      Qw[i * lda + c] -= tj * qv[i] * hv[c];
    }
  }
  __syncthreads();
  if (t == 0) {
    if (m >= 2) {
      d[m - 2] = A[(m - 2) * lda + m - 2];
      e[m - 2] = A[(m - 1) * lda + m - 2];
    }
    d[m - 1] = A[(m - 1) * lda + m - 1];
    e[m - 1] = 0.0;
    double gl = 0.0, gu = 0.0, emax = 0.0;
    for (int i = 0; i < m; ++i) {
      const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < m ? fabs(e[i]) : 0.0);
      gl = (i == 0 || d[i] - r < gl) ? d[i] - r : gl;
      gu = (i == 0 || d[i] + r > gu) ? d[i] + r : gu;
      e2[i] = e[i] * e[i];
      emax = e2[i] > emax ? e2[i] : emax;
    }
    const double tnorm = fmax(fabs(gl), fabs(gu));
    const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax);
    const double fudge = 2.1 * 2.220446049250313e-16 * tnorm * m + 4.0 * pivmin;
    misc[0] = tnorm;
    misc[1] = pivmin;
    for (int i = 0; i < m; ++i) {
      lo[i] = gl - fudge;
      hi[i] = gu + fudge;
    }
  }
  __syncthreads();
  const double tnorm = misc[0], pivmin = misc[1];
  const double eps = 2.220446049250313e-16;
  double dr[MT], e2r[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) {
    dr[i] = i < m ? d[i] : 0.0;
    e2r[i] = i < m ? e2[i] : 0.0;
  }
  // 2. every eigenvalue by multisection: P points per eigenvalue per round
  {
    int P = nt / m;
    if (P > 32) P = 32;
    const int jj = t / P, pp = t - (t / P) * P;
    for (int it = 0; it < 80; ++it) {
      if (jj < m) {
        const double l0 = lo[jj], h0 = hi[jj];
        const double h = (h0 - l0) / (P + 1);
        cnt[t] = td_sturm(dr, e2r, m, l0 + (pp + 1) * h, pivmin);
      }
      __syncthreads();
      int open = 0;
      if (t < m) {
        const double l0 = lo[t], h0 = hi[t];
        const double h = (h0 - l0) / (P + 1);
        int ps = P;
        for (int q = 0; q < P; ++q)
          if (cnt[t * P + q] > t) { ps = q; break; }
        const double nl = ps == 0 ? l0 : l0 + ps * h;
        const double nh = ps == P ? h0 : l0 + (ps + 1) * h;
        lo[t] = nl;
        hi[t] = nh;
        const double tol = fmax(2.0 * eps * fmax(fabs(nl), fabs(nh)), eps * tnorm);
        open = (nh - nl > tol && nh - nl < h0 - l0) ? 1 : 0;
      }
      if (!__syncthreads_or(open)) break;
    }
    if (t < m) w[t] = 0.5 * (lo[t] + hi[t]);
  }
  __syncthreads();
  // 3. eigenvectors of the tridiagonal: D+ (threads < m) and D- (64..64+m)
  const double pv = fmax(pivmin, eps * (tnorm > 0.0 ? tnorm : 1.0));
  if (t < m) {
    const double lam = w[t];
    double* q = dp + t * m;
    double qq = td_guard(dr[0] - lam, pv);
    q[0] = qq;
#pragma unroll
    for (int i = 1; i < MT; ++i) {
      const double qn = td_guard((dr[i] - lam) - e2r[i - 1] * td_rcp(qq), pv);
      if (i < m) q[i] = qn;
      qq = i < m ? qn : qq;
    }
  } else if (t >= 64 && t < 64 + m) {
    // backward chain on the reversed index k = m - 1 - i, read from LDS
    const int j = t - 64;
    const double lam = w[j];
    double* q = dm + j * m;
    double qq = td_guard(d[m - 1] - lam, pv);
    q[m - 1] = qq;
    for (int i = m - 2; i >= 0; --i) {
      qq = td_guard((d[i] - lam) - e2[i] * td_rcp(qq), pv);
      q[i] = qq;
    }
  }
  __syncthreads();
  if (t < m) {
    const int j = t;
    const double lam = w[j];
    const double* qp = dp + j * m;
    const double* qm = dm + j * m;
    int r = 0;
    double best = 0.0;
    for (int i = 0; i < m; ++i) {
      const double g = fabs(qp[i] + qm[i] - (d[i] - lam));
      if (i == 0 || g < best) { best = g; r = i; }
    }
    Z[r * m + j] = 1.0;
    double s = 1.0, z = 1.0;
    for (int i = r - 1; i >= 0; --i) {
      z *= -e[i] * td_rcp(qp[i]);
      Z[i * m + j] = z;
      s = fma(z, z, s);
    }
    z = 1.0;
    for (int i = r + 1; i < m; ++i) {
      z *= -e[i - 1] * td_rcp(qm[i]);
      Z[i * m + j] = z;
      s = fma(z, z, s);
    }
    const double inv = 1.0 / sqrt(s);
    for (int i = 0; i < m; ++i) Z[i * m + j] *= inv;
  }
  __syncthreads();
  // 4. two Gram-Schmidt passes inside clusters, in ascending order
  const double ortol = fmax(1e-3 * tnorm, 2.2250738585072014e-308);
  int degenerate = 0;
  for (int j = 1; j < m; ++j) {
    int c0 = j;
    while (c0 > 0 && w[c0] - w[c0 - 1] < ortol) --c0;
    if (c0 == j) continue;
    const int nb = j - c0;
    for (int pass = 0; pass < 2; ++pass) {
      {
        const int i = t >> 3, sub = t & 7;
        double s = 0.0;
        if (i < nb)
          for (int row = sub; row < m; row += 8) s = fma(Z[row * m + c0 + i], Z[row * m + j], s);
        s += __shfl_xor(s, 4, 8);
        s += __shfl_xor(s, 2, 8);
        s += __shfl_xor(s, 1, 8);
        if (i < nb && sub == 0) p[i] = s;
      }
      __syncthreads();
      for (int row = t; row < m; row += nt) {
        double s = 0.0;
        for (int i = 0; i < nb; ++i) s = fma(p[i], Z[row * m + c0 + i], s);
        Z[row * m + j] -= s;
      }
      __syncthreads();
    }
    if (t < 64) {
      const double z = lane < m ? Z[lane * m + j] : 0.0;
      const double s = wave_sum(z * z);
      if (lane < m) Z[lane * m + j] = z / sqrt(s);
      if (lane == 0) misc[3] = (s >= 0.25) ? 0.0 : 1.0;
    }
    __syncthreads();
    if (misc[3] != 0.0) degenerate = 1;
    __syncthreads();
  }
  if (degenerate) return 1;
  // 5. Y = Q Z
  for (int idx = t; idx < m * m; idx += nt) {
    const int r = idx / m, c = idx - (idx / m) * m;
    const double* qr = Qw + r * lda;
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int i = 0; i < MT; i += 2) {
      s0 = fma(i < m ? qr[i] : 0.0, i < m ? Z[(i < m ? i : 0) * m + c] : 0.0, s0);
      s1 = fma(i + 1 < m ? qr[i + 1] : 0.0, i + 1 < m ? Z[(i + 1 < m ? i + 1 : 0) * m + c] : 0.0, s1);
    }
    Y[idx] = s0 + s1;
  }
  __syncthreads();
  return 0;
}

}  // namespace usbs
