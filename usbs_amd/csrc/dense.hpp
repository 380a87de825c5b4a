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
// Eigendecomposition of the thick-restart arrowhead T of a Lanczos restart
// (m <= 32): diagonal alpha, couplings beta[a] between rows k and a for
// a < k (the arrow) and between a and a + 1 for a >= k (the Krylov tail);
// k >= m means T = diag(alpha).
//
// T is taken in the row order P = (m-1, m-2, ..., k, 0, 1, ..., k-1): the
// tail, reversed, comes first and is already tridiagonal, and the arrow is
// the trailing (k+1) x (k+1) block B = [alpha_k b^T; b diag(alpha_0..k-1)].
// Only B is tridiagonalised, by Householder reflectors that map column s of
// the trailing block to (mu, 0, ...) and are accumulated into the k x k
// factor Qk.  Each step is computed by all four wavefronts at once: the
// reflector, p = tau B22 v, K and w redundantly in every wavefront (identical
// arithmetic, so identical values), then each wavefront updates its own rows
// (row mod 4) of B, written into the other of two LDS copies, and of Qk in
// place -- one workgroup barrier per step.
//
// Eigenvalues of the tridiagonal: multisection on Sturm counts, 8 points per
// eigenvalue per round, eight eigenvalues per wavefront, each wavefront
// iterating on its own (ballots, no workgroup barrier) until its intervals
// are below max(2 eps |lambda|, eps |T|).  Eigenvectors: twisted
// factorisations (both Sturm chains of an eigenvalue in one thread, in
// registers), Gram-Schmidt (two passes) inside clusters of eigenvalues closer
// than 1e-3 |T|, then Y = P^T blockdiag(I, Qk) Z.
//   w  : m eigenvalues, ascending (LDS);  Y: m x m row-major (LDS), column j
//        = eigenvector of w[j] in T's own row order.
//   alpha, beta: T's coefficients (read only).
//   scr: arrow_scratch_doubles(m) doubles of LDS.
// Called by all 256 threads of the block.  Returns 1 (in every thread) when
// a cluster is degenerate -- a vector loses more than half its norm to the
// Gram-Schmidt projection, as for an exactly repeated eigenvalue -- and the
// caller must use another eigensolver; 0 otherwise.
#define TD_MAXM 32
#define AB_LD (TD_MAXM + 1)
// Each B copy is AB_ROWS x AB_BLD: the unrolled reads of a step run up to 31
// rows and columns past the trailing block's corner, into zeros.
#define AB_ROWS (2 * TD_MAXM)
#define AB_BLD (2 * TD_MAXM)
__host__ __device__ constexpr int arrow_scratch_doubles(int m) {
  return 2 * AB_ROWS * AB_BLD + (TD_MAXM + 1) * AB_LD + 4 * 3 * TD_MAXM + m * m + 3 * m + 64;
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
// A guarded pivot is never zero, so its sign bit counts it.
__device__ __forceinline__ int td_negbit(double q) {
  return (int)((unsigned)(__double_as_longlong(q) >> 32) >> 31);
}
__device__ __forceinline__ int td_sturm(const double (&dr)[TD_MAXM], const double (&e2r)[TD_MAXM],
                                        int m, double x, double pivmin) {
  double q = td_guard(dr[0] - x, pivmin);
  int c = td_negbit(q);
  if (m == TD_MAXM) {  // every entry in use: no masking
#pragma unroll
    for (int i = 1; i < TD_MAXM; ++i) {
      q = td_guard((dr[i] - x) - e2r[i - 1] * td_rcp(q), pivmin);
      c += td_negbit(q);
    }
    return c;
  }
#pragma unroll
  for (int i = 1; i < TD_MAXM; ++i) {
    const double qn = td_guard((dr[i] - x) - e2r[i - 1] * td_rcp(q), pivmin);
    const bool in = i < m;
    c += in ? td_negbit(qn) : 0;
    q = in ? qn : q;
  }
  return c;
}

// Sum / min / max over the wavefront, the same value in every lane: a DPP
// butterfly inside each row of 16 lanes (each level combines the same two
// numbers in both lanes of a pair), then the four row results read from
// lanes 0, 16, 32, 48 and combined in a fixed order.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += lane_xor<1>(v);
  v += lane_xor<2>(v);
  v += lane_xor<4>(v);
  v += lane_xor<8>(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
__device__ __forceinline__ double wave_min_dpp(double v) {
  v = fmin(v, lane_xor<1>(v));
  v = fmin(v, lane_xor<2>(v));
  v = fmin(v, lane_xor<4>(v));
  v = fmin(v, lane_xor<8>(v));
  return fmin(fmin(readlane_d(v, 0), readlane_d(v, 16)), fmin(readlane_d(v, 32), readlane_d(v, 48)));
}
__device__ __forceinline__ double wave_max_dpp(double v) {
  v = fmax(v, lane_xor<1>(v));
  v = fmax(v, lane_xor<2>(v));
  v = fmax(v, lane_xor<4>(v));
  v = fmax(v, lane_xor<8>(v));
  return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}

__device__ __forceinline__ int arrow_eigh(int m, int k, const double* __restrict__ alpha,
                                          const double* __restrict__ beta, double* w, double* Y,
                                          double* scr) {
  constexpr int MT = TD_MAXM;
  constexpr int P = 8;  // multisection points per eigenvalue and round
  const int t = threadIdx.x, nt = blockDim.x, lane = t & 63, wv = t >> 6;
  const int kp = k < m ? k : m - 1;  // arrow size of the reordered T
  const int j0 = m - kp - 1;         // reordered row of T's row k (first row of B)
  double* Bb = scr;                          // two copies of B, AB_ROWS x AB_BLD each
  double* Qk = scr + 2 * AB_ROWS * AB_BLD;   // kp x kp, row stride AB_LD, zero pad row
  double* bc = Qk + (TD_MAXM + 1) * AB_LD;   // per wavefront: hv, w, Qk v (32 each)
  double* Z = bc + 4 * 3 * TD_MAXM;          // m x m: eigenvectors of the tridiagonal
  double* d = Z + m * m;             // diagonal
  double* e = d + m;                 // e[p] couples p and p + 1
  double* misc = e + m;              // [0] degenerate flag

  // B (local index 0 = T's row k, 1 + a = T's row a) in copy 0, zeros
  // elsewhere in both copies; Qk = I with zeros around it
  for (int i2 = t; i2 < AB_ROWS * AB_BLD; i2 += nt) {  // pairs of doubles
    const int idx = 2 * i2;
    const int r = (idx / AB_BLD) & (AB_ROWS - 1), c = idx & (AB_BLD - 1);
    const bool first = idx < AB_ROWS * AB_BLD;
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int cu = c + u;
      if (first && r <= kp && cu <= kp) {
        if (r == cu) v[u] = alpha[r == 0 ? kp : r - 1];
        else if (k < m && (r == 0 || cu == 0)) v[u] = beta[(r == 0 ? cu : r) - 1];
      }
    }
    *reinterpret_cast<double2*>(Bb + idx) = make_double2(v[0], v[1]);
  }
  for (int idx = t; idx < (TD_MAXM + 1) * AB_LD; idx += nt) {
    const int r = idx / AB_LD, c = idx - (idx / AB_LD) * AB_LD;
    Qk[idx] = (r == c && r < kp) ? 1.0 : 0.0;
  }
  for (int p = t; p < j0; p += nt) {
    d[p] = alpha[m - 1 - p];
    e[p] = beta[m - 2 - p];
  }
  __syncthreads();

  // 1. Householder on B: step s maps B[s+1.., s] to (mu, 0, ...).  Values a
  // wavefront shares between its lanes (hv, w, Qk v) go through its own LDS
  // slots: written, then read back as broadcasts by the same wavefront.
  double* hvs = bc + wv * 3 * TD_MAXM;
  double* wvs = hvs + TD_MAXM;
  double* qvs = wvs + TD_MAXM;
  int cur = 0;
  for (int s = 0; s + 2 <= kp; ++s) {
    const double* Bc = Bb + (cur ? AB_ROWS * AB_BLD : 0);
    double* Bn = Bb + (cur ? 0 : AB_ROWS * AB_BLD);
    const int n1 = kp - s;  // trailing block rows s+1 .. kp
    // the reflector, in every lane from broadcast reads of row s (= column
    // s; entries past kp are zero)
    const double* xrow = Bc + s * AB_BLD + s + 1;
    double xs[TD_MAXM];
#pragma unroll
    for (int c = 0; c < TD_MAXM; ++c) xs[c] = xrow[c];
    asm volatile("" ::: "memory");
    double sg0 = 0.0, sg1 = 0.0, sg2 = 0.0, sg3 = 0.0;
#pragma unroll
    for (int c = 0; c < TD_MAXM; c += 4) {
      if (c > 0) sg0 = fma(xs[c], xs[c], sg0);
      sg1 = fma(xs[c + 1], xs[c + 1], sg1);
      sg2 = fma(xs[c + 2], xs[c + 2], sg2);
      sg3 = fma(xs[c + 3], xs[c + 3], sg3);
    }
    const double sg = (sg0 + sg1) + (sg2 + sg3);
    const double x0 = xs[0];
    double v0 = 1.0, mu = x0, tj = 0.0;
    if (sg != 0.0) {
      mu = sqrt(x0 * x0 + sg);
      v0 = x0 <= 0.0 ? x0 - mu : -sg / (x0 + mu);
      tj = 2.0 * v0 * v0 / (sg + v0 * v0);
    }
    if (t == 0) {
      d[j0 + s] = Bc[s * AB_BLD + s];
      e[j0 + s] = mu;
    }
    if (tj == 0.0) continue;  // uniform: B and Qk unchanged
    const double rv = 1.0 / v0;
    double hb[TD_MAXM];
    hb[0] = 1.0;
#pragma unroll
    for (int c = 1; c < TD_MAXM; ++c) hb[c] = xs[c] * rv;  // 0 past n1
    const double hv = lane < n1 ? xrow[lane] * rv : 0.0;
    const double hvl = lane == 0 ? 1.0 : hv;
    if (lane < TD_MAXM) hvs[lane] = hvl;
    // p = tau B22 v (lane = column; B symmetric, so column l = row l; rows
    // past kp are zero)
    const double* bcol = Bc + (s + 1) * AB_BLD + s + 1 + (lane < n1 ? lane : 0);
    double bv[TD_MAXM];
#pragma unroll
    for (int c = 0; c < TD_MAXM; ++c) bv[c] = bcol[c * AB_BLD];
    asm volatile("" ::: "memory");  // every load issued before the first FMA waits
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
#pragma unroll
    for (int c = 0; c < TD_MAXM; c += 4) {
      acc0 = fma(bv[c], hb[c], acc0);
      acc1 = fma(bv[c + 1], hb[c + 1], acc1);
      acc2 = fma(bv[c + 2], hb[c + 2], acc2);
      acc3 = fma(bv[c + 3], hb[c + 3], acc3);
    }
    const double p = lane < n1 ? tj * ((acc0 + acc1) + (acc2 + acc3)) : 0.0;
    const double K = 0.5 * tj * wave_sum_dpp(p * hvl);
    const double wl = p - K * hvl;
    // this wavefront's rows of Qk: qv = Qk v (lane l < 8: row wv + 4 l); reads
    // past column kp - 1 meet zeros of hb
    const int qr = wv + 4 * lane;
    if (lane < 8) {
      const double* qrow = Qk + (qr < kp ? qr : 0) * AB_LD + s;
      double qa[TD_MAXM];
#pragma unroll
      for (int c = 0; c < TD_MAXM; ++c) qa[c] = qrow[c];
      asm volatile("" ::: "memory");
      double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
#pragma unroll
      for (int c = 0; c < TD_MAXM; c += 4) {
        q0 = fma(qa[c], hb[c], q0);
        q1 = fma(qa[c + 1], hb[c + 1], q1);
        q2 = fma(qa[c + 2], hb[c + 2], q2);
        q3 = fma(qa[c + 3], hb[c + 3], q3);
      }
      qvs[lane] = (q0 + q1) + (q2 + q3);
    }
    if (lane < TD_MAXM) wvs[lane] = wl;
    __builtin_amdgcn_wave_barrier();
    // rank-2 update of this wavefront's rows into Bn, Qk -= tau (Qk v) v^T:
    // two rows per pass, lane & 31 = column
    const int col = lane & 31, half = lane >> 5;
    const double hvc = hvs[col], wc = wvs[col];
    double hr[4], wr[4], qq[4], bo[4], qo[4];
#pragma unroll
    for (int ps = 0; ps < 4; ++ps) {
      const int ri = 2 * ps + half;  // this wavefront's ri-th row
      const int r = wv + 4 * ri;     // trailing-block row
      hr[ps] = hvs[r];
      wr[ps] = wvs[r];
      qq[ps] = qvs[ri];
      bo[ps] = Bc[(s + 1 + r) * AB_BLD + s + 1 + col];  // zero rows / columns past kp
      qo[ps] = Qk[(r < kp ? r : 0) * AB_LD + s + col];
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int ps = 0; ps < 4; ++ps) {
      const int r = wv + 4 * (2 * ps + half);
      if (r < n1 && col < n1) Bn[(s + 1 + r) * AB_BLD + s + 1 + col] = bo[ps] - (hr[ps] * wc + wr[ps] * hvc);
      if (r < kp && col < n1) Qk[r * AB_LD + s + col] = qo[ps] - tj * qq[ps] * hvc;
    }
    cur ^= 1;
    __syncthreads();
  }
  {
    const double* Bf = Bb + (cur ? AB_ROWS * AB_BLD : 0);
    if (t == 0) {
      if (kp >= 1) {
        d[m - 2] = Bf[(kp - 1) * AB_BLD + kp - 1];
        e[m - 2] = Bf[kp * AB_BLD + kp - 1];
      }
      d[m - 1] = Bf[kp * AB_BLD + kp];
      e[m - 1] = 0.0;
    }
  }
  __syncthreads();

  // 2. Gershgorin bounds (every wavefront), register copy of the tridiagonal
  const double eps = 2.220446049250313e-16;
  double gl, gu, emax;
  {
    const bool in = lane < m;
    const double di = in ? d[lane] : 0.0, ei = in ? e[lane] : 0.0;
    const double em = (in && lane > 0) ? e[lane - 1] : 0.0;
    const double rr = fabs(em) + fabs(ei);
    gl = wave_min_dpp(in ? di - rr : INFINITY);
    gu = wave_max_dpp(in ? di + rr : -INFINITY);
    emax = wave_max_dpp(ei * ei);
  }
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax);
  const double fudge = 2.1 * eps * tnorm * m + 4.0 * pivmin;
  double dr[MT], e2r[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) {
    dr[i] = i < m ? d[i] : 0.0;
    const double ev = i < m ? e[i] : 0.0;
    e2r[i] = ev * ev;
  }

  // 3. eigenvalue jj = 8 wv + (lane >> 3), point lane & 7
  {
    const int g = lane >> 3, pp = lane & 7, jj = 8 * wv + g;
    double lo = gl - fudge, hi = gu + fudge;
    for (int it = 0; it < 80; ++it) {
      const double h = (hi - lo) / (P + 1);
      const int cnt = jj < m ? td_sturm(dr, e2r, m, lo + (pp + 1) * h, pivmin) : 0;
      const unsigned long long bal = __ballot(cnt > jj);
      const unsigned seg = (unsigned)(bal >> (8 * g)) & 0xFFu;
      const int ps = seg ? __builtin_ctz(seg) : P;
      const double nl = ps == 0 ? lo : lo + ps * h;
      const double nh = ps == P ? hi : lo + (ps + 1) * h;
      const double tol = fmax(2.0 * eps * fmax(fabs(nl), fabs(nh)), eps * tnorm);
      const bool open = jj < m && nh - nl > tol && nh - nl < hi - lo;
      lo = nl;
      hi = nh;
      if (!__any(open)) break;
    }
    if (jj < m && pp == 0) w[jj] = 0.5 * (lo + hi);
  }
  __syncthreads();

  // 4. eigenvectors of the tridiagonal: thread j < m, both chains in registers
  const double pv = fmax(pivmin, eps * (tnorm > 0.0 ? tnorm : 1.0));
  if (t < m) {
    const double lam = w[t];
    double qp[MT], qm[MT];
    {
      double qq = td_guard(dr[0] - lam, pv);
      qp[0] = qq;
#pragma unroll
      for (int i = 1; i < MT; ++i) {
        const double qn = td_guard((dr[i] - lam) - e2r[i - 1] * td_rcp(qq), pv);
        qp[i] = qn;
        qq = i < m ? qn : qq;
      }
      double qb = 0.0;
#pragma unroll
      for (int i = MT - 1; i >= 0; --i) {
        const double nb2 = (i == m - 1) ? td_guard(dr[i] - lam, pv)
                                        : td_guard((dr[i] - lam) - e2r[i] * td_rcp(qb), pv);
        qb = i <= m - 1 ? nb2 : qb;
        qm[i] = qb;
      }
    }
    int r = 0;
    double best = 0.0;
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      const double gg = fabs(qp[i] + qm[i] - (dr[i] - lam));
      if (i < m && (i == 0 || gg < best)) {
        best = gg;
        r = i;
      }
    }
    // z_r = 1; downwards z_i = -e_i / qp_i z_{i+1}; upwards z_i = -e_{i-1} / qm_i z_{i-1}
    double zz = 1.0;
#pragma unroll
    for (int i = MT - 1; i >= 0; --i) {
      if (i < r) zz *= -e[i] * td_rcp(qp[i]);
      qp[i] = zz;  // z_i for i <= r
    }
    zz = 1.0;
#pragma unroll
    for (int i = 1; i < MT; ++i) {
      if (i > r) zz *= -e[i - 1] * td_rcp(qm[i]);
      qm[i] = zz;  // z_i for i > r
    }
    double ss = 0.0;
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      const double z = i <= r ? qp[i] : qm[i];
      if (i < m) ss = fma(z, z, ss);
    }
    const double inv = 1.0 / sqrt(ss);
#pragma unroll
    for (int i = 0; i < MT; ++i)
      if (i < m) Z[i * m + t] = (i <= r ? qp[i] : qm[i]) * inv;
  }
  __syncthreads();

  // 5. two Gram-Schmidt passes inside clusters, in ascending order
  const double ortol = fmax(1e-3 * tnorm, 2.2250738585072014e-308);
  unsigned long long starts;
  {
    const bool in = lane < m;
    const double wl2 = in ? w[lane] : 0.0, wp = (in && lane > 0) ? w[lane - 1] : 0.0;
    starts = __ballot(in && (lane == 0 || !(wl2 - wp < ortol)));
  }
  int degenerate = 0;
  for (int j = 1; j < m; ++j) {
    const int c0 = 63 - __builtin_clzll(starts & ((2ull << j) - 1ull));
    if (c0 == j) continue;  // uniform
    const int nbk = j - c0;
    for (int pass = 0; pass < 2; ++pass) {
      {
        const int i = t >> 3, sub = t & 7;
        double sacc = 0.0;
        if (i < nbk)
          for (int row = sub; row < m; row += 8) sacc = fma(Z[row * m + c0 + i], Z[row * m + j], sacc);
        sacc += __shfl_xor(sacc, 4, 8);
        sacc += __shfl_xor(sacc, 2, 8);
        sacc += __shfl_xor(sacc, 1, 8);
        if (i < nbk && sub == 0) misc[8 + i] = sacc;
      }
      __syncthreads();
      for (int row = t; row < m; row += nt) {
        double sacc = 0.0;
        for (int i = 0; i < nbk; ++i) sacc = fma(misc[8 + i], Z[row * m + c0 + i], sacc);
        Z[row * m + j] -= sacc;
      }
      __syncthreads();
    }
    if (t < 64) {
      const double z = lane < m ? Z[lane * m + j] : 0.0;
      const double sn = wave_sum(z * z);
      if (lane < m) Z[lane * m + j] = z / sqrt(sn);
      if (lane == 0) misc[0] = (sn >= 0.25) ? 0.0 : 1.0;
    }
    __syncthreads();
    if (misc[0] != 0.0) degenerate = 1;
    __syncthreads();
  }
  if (degenerate) return 1;

  // 6. Y = P^T blockdiag(I_{j0+1}, Qk) Z: reordered row p is T's row
  //    m-1-p (p < m-kp) or p-(m-kp)
  {
    const int j = t & 31;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int p = (t >> 5) + 8 * it;
      if (p < m && j < m) {
        double y;
        if (p <= j0) {
          y = Z[p * m + j];
        } else {
          const double* qrow = Qk + (p - j0 - 1) * AB_LD;
          const double* zc = Z + (j0 + 1) * m + j;
          double y0 = 0.0, y1 = 0.0;
          int c = 0;
          for (; c + 2 <= kp; c += 2) {
            y0 = fma(qrow[c], zc[c * m], y0);
            y1 = fma(qrow[c + 1], zc[(c + 1) * m], y1);
          }
          if (c < kp) y0 = fma(qrow[c], zc[c * m], y0);
          y = y0 + y1;
        }
        const int x = p < m - kp ? m - 1 - p : p - (m - kp);
        Y[x * m + j] = y;
      }
    }
  }
  __syncthreads();
  return 0;
}

}  // namespace usbs
