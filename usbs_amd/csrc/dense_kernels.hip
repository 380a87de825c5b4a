// Single-workgroup FP64 kernels of the USBS loop: the two prox-bundle IPM
// subproblems (solver/usbs.py:25-356) with the bundle eigendecomposition
// epilogue, the Nystrom core of the max-cut rounding (utils/common.py:99-116),
// and the scalar bookkeeping / convergence test of one iteration
// (solver/usbs.py:581-652).  One wavefront runs each; all operands are in
// LDS, so the workgroup barriers are cheap.
#include "dense.hpp"
#include "solve.hpp"

namespace usbs {

#define IPM_NT 256
#define IPM_MAXK 11

__device__ __forceinline__ double dotd(const double* x, const double* y, int d, double* red) {
  double s = 0.0;
  for (int a = threadIdx.x; a < d; a += blockDim.x) s += x[a] * y[a];
  return block_sum1(s, red);
}

// Three dot products of length d in one block reduction: each component is
// reduced exactly as dotd reduces it (same terms, same tree), with one set
// of barriers for all three.
__device__ __forceinline__ void dotd3(const double* x1, const double* y1, const double* x2,
                                      const double* y2, const double* x3, const double* y3, int d,
                                      double* red, double& r1, double& r2, double& r3) {
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int a = threadIdx.x; a < d; a += blockDim.x) {
    s1 += x1[a] * y1[a];
    s2 += x2[a] * y2[a];
    s3 += x3[a] * y3[a];
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  s3 = wave_sum(s3);
  const int nw = blockDim.x >> 6, wid = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[wid] = s1;
    red[nw + wid] = s2;
    red[2 * nw + wid] = s3;
  }
  __syncthreads();
  r1 = 0.0;
  r2 = 0.0;
  r3 = 0.0;
  for (int w = 0; w < nw; ++w) {
    r1 += red[w];
    r2 += red[nw + w];
    r3 += red[2 * nw + w];
  }
  __syncthreads();
}

// 0.5 * U (T (x) Sinv + Sinv (x) T) U^T, entry (a, b)  (solver/usbs.py:92-93)
__device__ __forceinline__ double symkron(const double* T, const double* Si, int k, int a, int b) {
  int b1, c1, b2, c2;
  svec_pair(a, b1, c1);
  svec_pair(b, b2, c2);
  const double u = 0.70710678118654752440;
  const int na = (b1 == c1) ? 1 : 2, nb = (b2 == c2) ? 1 : 2;
  const double ua = (na == 1) ? 1.0 : u, ub = (nb == 1) ? 1.0 : u;
  double s = 0.0;
  for (int x = 0; x < na; ++x) {
    const int p = x == 0 ? b1 : c1, q = x == 0 ? c1 : b1;
    for (int y = 0; y < nb; ++y) {
      const int pp = y == 0 ? b2 : c2, qq = y == 0 ? c2 : b2;
      s += ua * ub * (T[p * k + pp] * Si[q * k + qq] + Si[p * k + pp] * T[q * k + qq]);
    }
  }
  return 0.5 * s;
}

// ---------------------------------------------------------------------------
// Per-lane K x K Cholesky in registers (K is a compile-time constant so the
// packed factor stays in VGPRs).  Returns false on a non-positive pivot.  A
// lane that meets one carries on with pivot 1 (its factor is discarded)
// instead of returning, so no load sits behind a branch and the next
// column's operands load while the square root and division run.
template <int K>
__device__ __forceinline__ bool lane_chol(const double* M, const double* dM, double s, double* L,
                                          double* rdiag) {
  bool ok = true;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    double d = M[c * K + c] + s * dM[c * K + c];
#pragma unroll
    for (int p = 0; p < c; ++p) d -= L[c * (c + 1) / 2 + p] * L[c * (c + 1) / 2 + p];
    if (!(d > 0.0)) {
      ok = false;
      d = 1.0;
    }
    d = sqrt(d);
    const double rd = 1.0 / d;
    L[c * (c + 1) / 2 + c] = d;
    rdiag[c] = rd;
#pragma unroll
    for (int i = c + 1; i < K; ++i) {
      double v = M[i * K + c] + s * dM[i * K + c];
#pragma unroll
      for (int p = 0; p < c; ++p) v -= L[i * (i + 1) / 2 + p] * L[c * (c + 1) / 2 + p];
      L[i * (i + 1) / 2 + c] = v * rd;
    }
  }
  return ok;
}

// S^{-1} for SPD S: lane j < K solves L L^T x = e_j for column j.
template <int K>
__device__ __forceinline__ void lane_spd_inverse(const double* S, double* Si) {
  const int j = threadIdx.x;
  if (j < K) {
    double L[K * (K + 1) / 2], rd[K];
    lane_chol<K>(S, S, 0.0, L, rd);
    double y[K], x[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
      for (int p = 0; p < i; ++p) v -= L[i * (i + 1) / 2 + p] * y[p];
      y[i] = v * rd[i];
    }
#pragma unroll
    for (int i = K - 1; i >= 0; --i) {
      double v = y[i];
#pragma unroll
      for (int p = i + 1; p < K; ++p) v -= L[p * (p + 1) / 2 + i] * x[p];
      x[i] = v * rd[i];
    }
#pragma unroll
    for (int i = 0; i < K; ++i) Si[i * K + j] = x[i];
  }
  __syncthreads();
}

// Largest 0.9^j (j < max_iters) with M + 0.9^j dM positive definite, or 0:
// the vmapped eigvalsh test of solver/usbs.py:135-144 with a Cholesky test
// per candidate (each thread tests one candidate of a batch of blockDim).
// pw0 = 0.9^threadIdx.x, the first batch's candidate of this thread.
template <int K>
__device__ __forceinline__ double pd_step(const double* M, const double* dM, int max_iters, double pw0,
                          double* red) {
  const int t = threadIdx.x;
  for (int j0 = 0; j0 < max_iters; j0 += blockDim.x) {
    double v = 0.0;
    if (j0 + t < max_iters) {
      const double s = (j0 == 0) ? pw0 : pow(0.9, (double)(j0 + t));
      double L[K * (K + 1) / 2], rd[K];
      if (lane_chol<K>(M, dM, s, L, rd)) v = s;
    }
    v = wave_max(v);
    __syncthreads();
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double b = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) b = red[w] > b ? red[w] : b;
    __syncthreads();
    if (b > 0.0) return b;
  }
  return 0.0;
}

// pd_step for M = S and M = T together when max_iters <= 128: threads
// 0..127 test S's candidates, threads 128..255 T's (pwh = 0.9^(threadIdx.x
// mod 128)), so the two searches share one Cholesky latency.  Same candidates
// and the same maxima as two pd_step calls.
template <int K>
__device__ __forceinline__ void pd_step2(const double* S, const double* dS, const double* T,
                                         const double* dT, int max_iters, double pwh, double* red,
                                         double& cS, double& cT) {
  const int t = threadIdx.x, h = t >> 7;
  double v = 0.0;
  if ((t & 127) < max_iters) {
    double L[K * (K + 1) / 2], rd[K];
    if (lane_chol<K>(h ? T : S, h ? dT : dS, pwh, L, rd)) v = pwh;
  }
  v = wave_max(v);
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  cS = red[1] > red[0] ? red[1] : red[0];
  cT = red[3] > red[2] ? red[3] : red[2];
  __syncthreads();
}

__device__ __forceinline__ void mat_from_svec(const double* v, int k, double* M) {
  dense_svec_inv(v, k, M);
  __syncthreads();
}

struct IpmLds {
  double *coeff, *q12, *h1, *sI, *sS, *sT, *F1, *ordv, *dsS, *dsT, *tmpv;
  double *S, *T, *Si, *dS, *dT, *tk, *red;
  double* R;     // d x 3 right-hand sides of the Woodbury solve
  double* rD;    // d reciprocal pivots of the LDL^T factor
  double* lp;    // d x 8 scaled panel columns of the LDL^T
  int* ired;
  int *pb, *pc;  // svec index -> (b, c)
};

__device__ IpmLds ipm_layout(double* lds, int k) {
  const int d = k * (k + 1) / 2, kk = k * k;
  IpmLds L;
  double* p = lds;
  L.coeff = p; p += (d * d > 5 * kk + 4 * k + 64 ? d * d : 5 * kk + 4 * k + 64);
  L.q12 = p; p += d; L.h1 = p; p += d; L.sI = p; p += d; L.sS = p; p += d; L.sT = p; p += d;
  L.F1 = p; p += d; L.ordv = p; p += d; L.dsS = p; p += d; L.dsT = p; p += d; L.tmpv = p; p += d;
  L.S = p; p += kk; L.T = p; p += kk; L.Si = p; p += kk; L.dS = p; p += kk; L.dT = p; p += kk;
  L.tk = p; p += kk;
  L.R = p; p += 3 * d;
  L.rD = p; p += d;
  L.lp = p; p += 8 * d;
  L.red = p; p += 16;
  L.ired = (int*)p; p += 16;
  L.pb = (int*)p; p += (d + 1) / 2 + 1;
  L.pc = (int*)p;
  return L;
}

__host__ __device__ constexpr int ipm_lds_doubles(int k) {
  return (k * (k + 1) / 2) * (k * (k + 1) / 2) > 5 * k * k + 4 * k + 64
             ? (k * (k + 1) / 2) * (k * (k + 1) / 2) + 23 * (k * (k + 1) / 2) + 6 * k * k + 36
             : 5 * k * k + 4 * k + 64 + 23 * (k * (k + 1) / 2) + 6 * k * k + 36;
}

static size_t ipm_lds_bytes(int k) { return sizeof(double) * ipm_lds_doubles(k); }

__device__ void ipm_tables(const IpmLds& L, int k) {
  const int d = k * (k + 1) / 2;
  for (int a = threadIdx.x; a < d; a += blockDim.x) {
    int b, c;
    svec_pair(a, b, c);
    L.pb[a] = b;
    L.pc[a] = c;
    L.sI[a] = (b == c) ? 1.0 : 0.0;
  }
  __syncthreads();
}

// 0.5 * U (T (x) Sinv + Sinv (x) T) U^T, entry (a, b)  (solver/usbs.py:92-93).
// All four (x, y) terms are evaluated, with the operands loaded
// unconditionally (no branch per load); a term outside na x nb gets weight 0,
// and fma(0, X, s) = s for finite X, so only the na x nb terms contribute, in
// the same order.
__device__ __forceinline__ double symkron(const IpmLds& L, int k, int a, int b) {
  const int b1 = L.pb[a], c1 = L.pc[a], b2 = L.pb[b], c2 = L.pc[b];
  const double* T = L.T;
  const double* Si = L.Si;
  const double u = 0.70710678118654752440;
  const int na = (b1 == c1) ? 1 : 2, nb = (b2 == c2) ? 1 : 2;
  const double ua = (na == 1) ? 1.0 : u, ub = (nb == 1) ? 1.0 : u;
  const double uab = ua * ub;
  double s = 0.0;
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    const int p = x == 0 ? b1 : c1, q = x == 0 ? c1 : b1;
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int pp = y == 0 ? b2 : c2, qq = y == 0 ? c2 : b2;
      const double X = T[p * k + pp] * Si[q * k + qq] + Si[p * k + pp] * T[q * k + qq];
      s = fma((x < na && y < nb) ? uab : 0.0, X, s);
    }
  }
  return 0.5 * s;
}


// LDL^T (no pivoting) of the SPD Newton block A (D x D row-major in LDS,
// D <= 66), in place: on return the strict lower triangle holds the unit
// factor L and rD[k] = 1 / D_k.  The rows 0..63 are held in registers by
// column slots: thread (row i = t mod 64, wavefront w) holds A[i][j] for the
// columns j = 4 s + w < 64, as a window that drops its first slot every
// four steps.  For D > 64 every thread also holds A[64][i], A[65][i] and
// the 2 x 2 corner (the same values in each wavefront).  Step k: the owner
// wavefront of column k has written it, and 1 / D_k, into an LDS slot (two
// slots, alternating); after one workgroup barrier every thread reads them,
// stores l_i = A[i][k] / D_k as L[i][k] (owner wavefront), updates column
// k + 1 first and, if it owns it, writes that column and its pivot's
// reciprocal into the other slot, then applies A[i][j] -= l_i A[j][k] to its
// other columns.  Each entry thus sees the right-looking sequence, k
// ascending, with one barrier per column.  Entries above the diagonal or
// already final are updated too and never read.  A zero pivot yields
// inf/NaN, as an LU solve of a singular matrix does.  Lp: at least
// 2 (D + 1) doubles of LDS.
template <int D>
__device__ __forceinline__ void ipm_ldlt(double* A, double* rD, double* Lp) {
  static_assert(D >= 1 && D <= 66, "ipm_ldlt: D <= 66");
  constexpr bool XR = D > 64;              // rows 64, 65 (D = 66)
  constexpr int RM = D < 64 ? D : 64;      // rows and columns of the slot layout
  constexpr int NS = (RM + 3) / 4;         // column slots per thread
  constexpr int SL = D + 1;                // LDS slot: column k (D entries), then 1 / D_k
  const int t = threadIdx.x, i = t & 63, w = t >> 6;
  const int ic = i < RM ? i : RM - 1;
  double a[NS];  // a[u] = A[i][4 (q + u) + w] in step group q
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int j = 4 * u + w, jc = j < RM ? j : RM - 1;
    a[u] = A[ic * D + jc];
  }
  double e0 = 0.0, e1 = 0.0, g00 = 0.0, g10 = 0.0, g11 = 0.0;
  if constexpr (XR) {
    e0 = A[64 * D + ic];
    e1 = A[65 * D + ic];
    g00 = A[64 * D + 64];
    g10 = A[65 * D + 64];
    g11 = A[65 * D + 65];
  }
  if (w == 0) {  // column 0 and 1 / D_0 into slot 0
    if (i < RM) Lp[i] = a[0];
    if (i == 0) {
      Lp[D] = 1.0 / a[0];
      if (D == 1) rD[0] = Lp[D];
      if (XR) {
        Lp[64] = e0;  // lane 0: A[64][0], A[65][0]
        Lp[65] = e1;
      }
    }
  }
  for (int q = 0; 4 * q < D - 1; ++q) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = 4 * q + r;
      if (k < D - 1) {
        const double* cb = Lp + (r & 1) * SL;
        double* cn = Lp + ((r + 1) & 1) * SL;
        __syncthreads();
        const double rdk = cb[D];
        const double ck = cb[ic];  // A[i][k]
        const double li = ck * rdk;
        double l64 = 0.0, l65 = 0.0, c64 = 0.0, c65 = 0.0;
        if constexpr (XR) {
          c64 = cb[64];
          c65 = cb[65];
          l64 = c64 * rdk;
          l65 = c65 * rdk;
        }
        if (t == 0) {
          rD[k] = rdk;
          if (XR) {
            if (k < 64) A[64 * D + k] = l64;
            A[65 * D + k] = l65;
          }
        }
        if (k < RM && w == r && i > k && i < RM) A[i * D + k] = li;
        if constexpr (XR) {
          e0 = fma(-l64, ck, e0);  // A[64][i] -= l_64 A[i][k]
          e1 = fma(-l65, ck, e1);
          g00 = fma(-l64, c64, g00);
          g10 = fma(-l65, c64, g10);
          g11 = fma(-l65, c65, g11);
        }
        // column k + 1 first, by its owner wavefront, into the next slot
        const int u1 = (r + 1) / 4;  // 0 or 1 (static once r is unrolled)
        const bool own = w == ((r + 1) & 3);
        const bool ahead = k + 1 < RM && own;
        if (ahead) {
          const double v1 = fma(-li, cb[k + 1], a[u1]);
          a[u1] = v1;
          if (k + 2 <= D - 1) {
            if (i < RM) cn[i] = v1;
            if (i == k + 1) cn[D] = 1.0 / v1;
            if constexpr (XR) {
              const double n64 = readlane_d(e0, k + 1), n65 = readlane_d(e1, k + 1);
              if (i == 0) {
                cn[64] = n64;
                cn[65] = n65;
              }
            }
          } else if (i == k + 1) {
            rD[k + 1] = 1.0 / v1;  // the last pivot
          }
        } else if (XR && k + 1 == 64 && own && i == 0) {  // the corner's first column
          cn[64] = g00;
          cn[65] = g10;
          cn[D] = 1.0 / g00;
        }
#pragma unroll
        for (int u = 0; u < NS; ++u) {
          const int j = 4 * (q + u) + w;
          const double cj = cb[j < RM ? j : RM - 1];
          if (!(u == u1 && ahead)) a[u] = fma(-li, cj, a[u]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u + 1 < NS; ++u) a[u] = a[u + 1];
  }
  if constexpr (XR) {
    if (t == 0) rD[65] = 1.0 / g11;
  }
  __syncthreads();
}

// Solve (L D L^T) X = R for nrhs <= 4 right-hand sides (R row-major D x 3)
// with the ipm_ldlt factor.  One wavefront per right-hand side keeps it in
// registers (rows lane, lane + 64); both substitutions are column-oriented:
// one broadcast (readlane) and one FMA per row and step.  The factor entries
// of eight steps are loaded together, unconditionally from clamped in-bounds
// addresses, and masked by a select, so the loads do not wait on the
// substitution chain.  The steps run in the same order as one at a time.
template <int D>
__device__ __forceinline__ void ipm_ldlt_solve(const double* Lm, const double* rD, double* R, int nrhs) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (w < nrhs) {
    const int i0 = lane, i1 = lane + 64;
    const int i0c = i0 < D ? i0 : D - 1, i1c = i1 < D ? i1 : D - 1;
    double y0 = (i0 < D) ? R[i0 * 3 + w] : 0.0;
    double y1 = (D > 64 && i1 < D) ? R[i1 * 3 + w] : 0.0;
    constexpr int G = 8;  // factor entries loaded per group, ahead of the group's steps
#pragma unroll
    for (int k0 = 0; k0 < D - 1; k0 += G) {
      double la[G], lb[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int k = k0 + g < D - 1 ? k0 + g : D - 2;
        la[g] = Lm[i0c * D + k];
        lb[g] = (D > 64) ? Lm[i1c * D + k] : 0.0;
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int k = k0 + g;
        if (k < D - 1) {
          const double yk = (k < 64) ? readlane_d(y0, k) : readlane_d(y1, k - 64);
          y0 -= ((i0 > k && i0 < D) ? la[g] : 0.0) * yk;
          if (D > 64) y1 -= ((i1 > k && i1 < D) ? lb[g] : 0.0) * yk;
        }
      }
    }
    y0 *= rD[i0c];
    if (D > 64) y1 *= rD[i1c];
#pragma unroll
    for (int k0 = D - 1; k0 > 0; k0 -= G) {
      double la[G], lb[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int k = k0 - g > 0 ? k0 - g : 1;
        la[g] = Lm[k * D + i0c];
        lb[g] = (D > 64) ? Lm[k * D + i1c] : 0.0;
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int k = k0 - g;
        if (k > 0) {
          const double xk = (k < 64) ? readlane_d(y0, k) : readlane_d(y1, k - 64);
          y0 -= ((i0 < k) ? la[g] : 0.0) * xk;
          if (D > 64) y1 -= ((i1 < k) ? lb[g] : 0.0) * xk;
        }
      }
    }
    if (i0 < D) R[i0 * 3 + w] = y0;
    if (D > 64 && i1 < D) R[i1 * 3 + w] = y1;
  }
  __syncthreads();
}

// Solve (B - U W^T) x = r where B (D x D) is SPD: U = [u1, u2], W = [w1, w2]
// (nu = 1 or 2 columns), by the Sherman-Morrison-Woodbury identity
// x = z0 + Z (I - W^T Z)^{-1} W^T z0 with z0 = B^{-1} r, Z = B^{-1} U.  Same
// solution as a dense solve of the full matrix, up to rounding.  r in/out;
// R: D x 3 LDS scratch; B is destroyed.
template <int D>
__device__ __forceinline__ void woodbury_solve(double* B, double* r, const double* u1, const double* u2,
                               const double* w1, const double* w2, int nu, double* R, double* rD,
                               double* lp, double* red) {
  const int t = threadIdx.x;
  for (int a = t; a < D; a += blockDim.x) {
    R[a * 3 + 0] = r[a];
    R[a * 3 + 1] = u1[a];
    R[a * 3 + 2] = (nu > 1) ? u2[a] : 0.0;
  }
  ipm_ldlt<D>(B, rD, lp);
  ipm_ldlt_solve<D>(B, rD, R, nu + 1);
  if (t < 64) {
    double g0 = 0.0, g1 = 0.0, m00 = 0.0, m01 = 0.0, m10 = 0.0, m11 = 0.0;
    for (int a = t; a < D; a += 64) {
      const double z0 = R[a * 3], z1 = R[a * 3 + 1], z2 = R[a * 3 + 2];
      const double wa = w1[a], wb = (nu > 1) ? w2[a] : 0.0;
      g0 += wa * z0; g1 += wb * z0;
      m00 += wa * z1; m01 += wa * z2; m10 += wb * z1; m11 += wb * z2;
    }
    g0 = wave_sum(g0); g1 = wave_sum(g1);
    m00 = wave_sum(m00); m01 = wave_sum(m01);
    m10 = wave_sum(m10); m11 = wave_sum(m11);
    double c0, c1 = 0.0;
    if (nu > 1) {
      const double a00 = 1.0 - m00, a01 = -m01, a10 = -m10, a11 = 1.0 - m11;
      const double det = a00 * a11 - a01 * a10;
      c0 = (a11 * g0 - a01 * g1) / det;
      c1 = (a00 * g1 - a10 * g0) / det;
    } else {
      c0 = g0 / (1.0 - m00);
    }
    if (t == 0) { red[0] = c0; red[1] = c1; }
  }
  __syncthreads();
  const double c0 = red[0], c1 = red[1];
  for (int a = t; a < D; a += blockDim.x) r[a] = R[a * 3] + R[a * 3 + 1] * c0 + R[a * 3 + 2] * c1;
  __syncthreads();
}

// mu of the current point (solver/usbs.py:75-77, 163-165)
__device__ __forceinline__ double ipm_mu(const IpmLds& L, int k, int d, double eta, double zeta, double omega) {
  dense_svec(L.S, k, L.sS);
  dense_svec(L.T, k, L.sT);
  __syncthreads();
  double st, trS, unused;
  dotd3(L.sS, L.sT, L.sI, L.sS, L.sI, L.sS, d, L.red, st, trS, unused);
  return (st + eta * zeta + omega * (1 - trS - eta)) / (2 * k + 4.0);
}

// Initial T, zeta, omega from T0 (in L.T) and zeta0 (usbs.py:67-74 / 228-235).
__device__ __forceinline__ void ipm_center(const IpmLds& L, int k, double& zeta, double& omega) {
  double mx = zeta, mn = zeta;
  for (int i = 0; i < k; ++i) {
    const double v = L.T[i * k + i];
    mx = v > mx ? v : mx;
    mn = v < mn ? v : mn;
  }
  double width = mx - mn;
  width = width < 100.0 ? 100.0 : width;
  omega = -1.01 * mn;
  omega = omega < width ? width : omega;
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x) L.T[i * k + i] += omega;
  zeta += omega;
  __syncthreads();
}

// Shared step: sizes, line search, update; returns the new mu and step.
template <int K>
__device__ __forceinline__ void ipm_step(const IpmLds& L, int max_iters, double pw0, double pwh, double& eta, double& zeta,
                         double& omega, double& mu, double& step, double d_eta, double d_zeta,
                         double d_omega) {
  constexpr int D = K * (K + 1) / 2;
  mat_from_svec(L.dsS, K, L.dS);
  mat_from_svec(L.dsT, K, L.dT);
  const double r1 = d_eta / eta, r2 = d_zeta / zeta, r3 = d_omega / omega;
  double rmin = r1;
  if (r2 < rmin) rmin = r2;
  if (r3 < rmin) rmin = r3;
  if (r1 != r1 || r2 != r2 || r3 != r3) rmin = NAN;  // jnp.min propagates NaN
  double st = -1.0 / rmin;
  st = st > 1.0 ? 1.0 : st;
  st = 0.99 * ((st <= 0.0) ? 1.0 : st);
  double cS, cT;
  if (max_iters <= IPM_NT / 2) {
    pd_step2<K>(L.S, L.dS, L.T, L.dT, max_iters, pwh, L.red, cS, cT);
  } else {
    cS = pd_step<K>(L.S, L.dS, max_iters, pw0, L.red);
    cT = pd_step<K>(L.T, L.dT, max_iters, pw0, L.red);
  }
  const double cand = cS < cT ? cS : cT;
  st = st < cand ? st : cand;
  __syncthreads();
  for (int idx = threadIdx.x; idx < K * K; idx += blockDim.x) {
    L.S[idx] += st * L.dS[idx];
    L.T[idx] += st * L.dT[idx];
  }
  __syncthreads();
  eta += st * d_eta;
  zeta += st * d_zeta;
  omega += st * d_omega;
  double mun = ipm_mu(L, K, D, eta, zeta, omega);
  mun *= (st > 0.2) ? (0.5 - 0.4 * st * st) : 1.0;
  mu = mun > mu ? mu : mun;
  step = st;
}

// out = (cq Gq) v for the D x D matrix Gq (an LDS copy): 4 threads per row,
// each summing every 4th column, combined in a fixed order.  The row's
// operands are loaded unconditionally from clamped in-bounds addresses, all
// in flight together; columns past D add nothing.
template <int D>
__device__ __forceinline__ void qmatvec(const double* Gq, double cq, const double* v, double* out) {
  const int t = threadIdx.x, q = t & 3;
  for (int a0 = 0; a0 < D; a0 += IPM_NT / 4) {
    const int a = a0 + (t >> 2);
    const double* g = Gq + (a < D ? a : D - 1) * D;
    double gv[(D + 3) / 4], vv[(D + 3) / 4];
#pragma unroll
    for (int i = 0; i < (D + 3) / 4; ++i) {
      const int b = q + 4 * i, bc = b < D ? b : D - 1;
      gv[i] = g[bc];
      vv[i] = v[bc];
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < (D + 3) / 4; ++i)
      if (q + 4 * i < D) s += (cq * gv[i]) * vv[i];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (a < D && q == 0) out[a] = s;
  }
  __syncthreads();
}

template <int K>
__global__ void __launch_bounds__(IPM_NT) k_ipm_quad(IpmArgs A) {
  constexpr int D = K * (K + 1) / 2;
  const int k = K, d = D;
  extern __shared__ double lds[];
  const int t = threadIdx.x;
  IpmLds L = ipm_layout(lds, k);
  // LDS copy of the D x D block Gq (after the shared layout; the launch adds
  // its bytes): every IPM iteration reads it three times.
  double* gq = lds + ipm_lds_doubles(K);
  {
    constexpr int NL = (D * D + IPM_NT - 1) / IPM_NT;
    double g[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int idx = t + IPM_NT * i;
      g[i] = A.Gq[idx < D * D ? idx : D * D - 1];
    }
#pragma unroll
    for (int i = 0; i < NL; ++i)
      if (t + IPM_NT * i < D * D) gq[t + IPM_NT * i] = g[i];
  }
  ipm_tables(L, k);
  const double pw0 = pow(0.9, (double)t), pwh = pow(0.9, (double)(t & 127));
  double* sc = A.sc;
  const double tu = A.tu, rho = A.rho;
  const double trXb = sc[SC_TRXB] > 0.0 ? sc[SC_TRXB] : tu;
  const double pb = sc[SC_PB];
  const double cq = tu * tu / rho;
  const double q22 = (tu * tu / (rho * trXb * trXb)) * sc[SC_ZZ];
  const double h2 = (tu / trXb) * (sc[SC_ZY] - pb - (sc[SC_ZBU] / rho));
  dense_svec(A.Gz, k, L.q12);
  dense_svec(A.Gh, k, L.h1);
  __syncthreads();
  for (int a = t; a < d; a += IPM_NT) {
    L.q12[a] *= tu * tu / (rho * trXb);
    L.h1[a] *= tu;
  }
  // warm start (solve_step_subprob init, usbs.py:422-427, then usbs.py:64-65)
  double eta_in;
  if (A.first) {
    eta_in = (tu / trXb) * ((pb == 0.0) ? 0.00001 : 0.99 * (1.0 / (k + 1.0)));
    for (int idx = t; idx < k * k; idx += IPM_NT)
      L.S[idx] = (idx / k == idx % k) ? tu * (0.99 * (1.0 / (k + 1.0))) : 0.0;
  } else {
    eta_in = A.eta_io[0];
    for (int idx = t; idx < k * k; idx += IPM_NT) L.S[idx] = A.S_io[idx];
  }
  __syncthreads();
  double eta = (trXb / tu) * eta_in;
  for (int idx = t; idx < k * k; idx += IPM_NT) L.S[idx] = L.S[idx] / tu;
  __syncthreads();
  dense_svec(L.S, k, L.sS);
  __syncthreads();
  qmatvec<D>(gq, cq, L.sS, L.tmpv);
  for (int a = t; a < d; a += IPM_NT) L.tmpv[a] = L.tmpv[a] + eta * L.q12[a] + L.h1[a];
  __syncthreads();
  mat_from_svec(L.tmpv, k, L.T);
  double zeta = dotd(L.q12, L.sS, d, L.red) + eta * q22 + h2;
  double omega;
  ipm_center(L, k, zeta, omega);
  double mu = ipm_mu(L, k, d, eta, zeta, omega);
  double step = 1.0;
  int it = 0;
  const double eps = A.eps;
  while (it < A.max_iters && mu > eps && step > eps * eps) {
    // sS, sT are current (ipm_mu recomputes them)
    double trS, q12sS, unused;
    dotd3(L.sI, L.sS, L.q12, L.sS, L.q12, L.sS, d, L.red, trS, q12sS, unused);
    const double k1 = (1 - trS - eta) / omega;
    const double k2 = (zeta / eta) + q22;
    qmatvec<D>(gq, cq, L.sS, L.F1);
    for (int a = t; a < d; a += IPM_NT)
      L.F1[a] = (L.F1[a] + eta * L.q12[a] + L.h1[a]) + (-L.sT[a] + omega * L.sI[a]);
    __syncthreads();
    double F2 = q12sS + eta * q22 + h2;
    F2 = F2 + (-zeta + omega);
    lane_spd_inverse<K>(L.S, L.Si);
    const double den = k1 * k2 + 1;
    // coeff = B - U W^T, B = Q11 + sym-kron(T, S^-1) (SPD),
    // U = [q12, svec(I)], W = [k1 q12 + svec(I), q12 - k2 svec(I)] / den
    for (int idx = t; idx < d * d; idx += IPM_NT) {
      const int a = idx / D, b = idx - (idx / D) * D;
      L.coeff[idx] = cq * gq[idx] + symkron(L, k, a, b);
    }
    dense_svec(L.Si, k, L.tmpv);
    for (int a = t; a < d; a += IPM_NT) {
      L.dsT[a] = (k1 * L.q12[a] + L.sI[a]) / den;
    }
    __syncthreads();
    const double c_a = (mu / omega + trS + eta - 1 + k1 * (F2 - mu / eta + zeta)) / den;
    const double c_b = (F2 - mu / eta + zeta - k2 * (mu / omega + trS + eta - 1)) / den;
    for (int a = t; a < d; a += IPM_NT) {
      double v = -L.F1[a] + mu * L.tmpv[a] - L.sT[a];
      v += c_a * L.q12[a];
      v += c_b * L.sI[a];
      L.dsS[a] = v;
      L.ordv[a] = (L.q12[a] - k2 * L.sI[a]) / den;
    }
    __syncthreads();
    woodbury_solve<D>(L.coeff, L.dsS, L.q12, L.sI, L.dsT, L.ordv, 2, L.R, L.rD, L.lp, L.red);
    // back-substitution for the other directions (usbs.py:111-122)
    // p1, p2 and q12 . dsS (for d_zeta) in one reduction; tmpv and ordv
    // hold k1 q12 + svec(I) and q12 - k2 svec(I)
    for (int a = t; a < d; a += IPM_NT) {
      L.tmpv[a] = k1 * L.q12[a] + L.sI[a];
      L.ordv[a] = L.q12[a] - k2 * L.sI[a];
    }
    __syncthreads();
    double p1, p2, q12dsS;
    dotd3(L.tmpv, L.dsS, L.ordv, L.dsS, L.q12, L.dsS, d, L.red, p1, p2, q12dsS);
    const double d_eta = (-p1 - (mu / omega) - trS - eta + 1 + k1 * (-F2 + mu / eta - zeta)) / den;
    const double d_omega = (-F2 + mu / eta - zeta + k2 * (mu / omega + trS + eta - 1) - p2) / den;
    qmatvec<D>(gq, cq, L.dsS, L.dsT);
    for (int a = t; a < d; a += IPM_NT)
      L.dsT[a] = L.dsT[a] + d_eta * L.q12[a] + d_omega * L.sI[a] + L.F1[a];
    __syncthreads();
    const double d_zeta = q12dsS + d_eta * q22 + d_omega + F2;
    ipm_step<K>(L, A.max_iters, pw0, pwh, eta, zeta, omega, mu, step, d_eta, d_zeta, d_omega);
    ++it;
  }
  // outputs (usbs.py:191-192) and the bundle factor (usbs.py:568-570)
  const double eta_out = (tu / trXb) * eta;
  for (int idx = t; idx < k * k; idx += IPM_NT) A.S_io[idx] = tu * L.S[idx];
  if (t == 0) {
    A.eta_io[0] = eta_out;
    sc[SC_ETA] = eta_out;
    sc[SC_IPM_IT] = it;
  }
  __syncthreads();
  double* Ae = L.coeff;
  double* Ve = Ae + k * k;
  double* tmp = Ve + k * k;
  double* w = tmp + k * k;
  double* coef = w + k;
  int* part = (int*)(coef + 2 * k + 2);
  for (int idx = t; idx < k * k; idx += IPM_NT) Ae[idx] = tu * L.S[idx];
  __syncthreads();
  double trs = 0.0;
  for (int i = 0; i < k; ++i) trs += tu * L.S[i * k + i];
  dense_eigh(Ae, k, w, Ve, tmp, coef, part, L.red);
  for (int idx = t; idx < k * k; idx += IPM_NT) {
    const int i = idx / k, x = idx % k;
    const double lam = w[x] < 0.0 ? 0.0 : w[x];
    A.GF[idx] = Ve[idx] * sqrt(lam);
    if (x >= A.kc) A.Epast[i * (k - A.kc) + (x - A.kc)] = Ve[idx];
  }
  if (t == 0) {
    double s = 0.0;
    for (int x = 0; x < A.kc; ++x) s += (w[x] < 0.0 ? 0.0 : w[x]);
    sc[SC_TRS] = trs;
    sc[SC_SUMLAM] = s;
  }
}

// Spectral lower bound (solver/usbs.py:195-356); writes sc[SC_LB].
template <int K>
__global__ void __launch_bounds__(IPM_NT)
k_ipm_lb(int max_iters, double tu, double eps, const double* Glb, double* sc) {
  constexpr int D = K * (K + 1) / 2;
  const int k = K, d = D;
  extern __shared__ double lds[];
  const int t = threadIdx.x;
  IpmLds L = ipm_layout(lds, k);
  ipm_tables(L, k);
  const double pw0 = pow(0.9, (double)t), pwh = pow(0.9, (double)(t & 127));
  const double trXb = sc[SC_TRXB] > 0.0 ? sc[SC_TRXB] : tu;
  const double pb = sc[SC_PB];
  double* g1 = L.h1;
  dense_svec(Glb, k, g1);
  __syncthreads();
  for (int a = t; a < d; a += IPM_NT) g1[a] *= tu;
  const double g2 = (tu / trXb) * (sc[SC_ZYC] - pb);
  for (int idx = t; idx < k * k; idx += IPM_NT)
    L.S[idx] = (idx / k == idx % k) ? 0.99 * (1.0 / (k + 1.0)) : 0.0;
  __syncthreads();
  double eta = (pb == 0.0) ? 0.00001 : 0.99 * (1.0 / (k + 1.0));
  mat_from_svec(g1, k, L.T);
  double zeta = g2, omega;
  ipm_center(L, k, zeta, omega);
  double mu = ipm_mu(L, k, d, eta, zeta, omega);
  double step = 1.0;
  int it = 0;
  while (it < max_iters && mu > eps) {
    const double trS = dotd(L.sI, L.sS, d, L.red);
    const double k1 = (1 - trS - eta) / omega;
    lane_spd_inverse<K>(L.S, L.Si);
    const double ze = zeta / eta;
    const double cI = ze / (k1 * ze + 1);
    // coeff = sym-kron(T, S^-1) + cI svec(I) svec(I)^T  (B - U W^T with W = -cI svec(I))
    for (int idx = t; idx < d * d; idx += IPM_NT) {
      const int a = idx / D, b = idx - (idx / D) * D;
      L.coeff[idx] = symkron(L, k, a, b);
    }
    dense_svec(L.Si, k, L.tmpv);
    for (int a = t; a < d; a += IPM_NT) L.ordv[a] = -cI * L.sI[a];
    __syncthreads();
    const double cf = ze * (-k1 * (mu / eta - g2 - omega) + mu / omega + trS + eta - 1) / (-k1 * ze - 1);
    for (int a = t; a < d; a += IPM_NT) {
      double v = cf * L.sI[a];
      v = v + ((g2 - mu / eta) * L.sI[a] - g1[a]);
      v = v + mu * L.tmpv[a];
      L.dsS[a] = v;
    }
    __syncthreads();
    woodbury_solve<D>(L.coeff, L.dsS, L.sI, L.sI, L.ordv, L.ordv, 1, L.R, L.rD, L.lp, L.red);
    const double p1 = dotd(L.sI, L.dsS, d, L.red);
    const double d_eta = (p1 - k1 * (mu / eta - g2 - omega) + mu / omega + trS + eta - 1) / (-k1 * ze - 1);
    double d_omega = -ze * d_eta + mu / eta;
    d_omega = d_omega - g2 - omega;
    for (int a = t; a < d; a += IPM_NT)
      L.dsT[a] = d_omega * L.sI[a] - L.sT[a] + g1[a] + omega * L.sI[a];
    __syncthreads();
    const double d_zeta = d_omega - zeta + g2 + omega;
    ipm_step<K>(L, max_iters, pw0, pwh, eta, zeta, omega, mu, step, d_eta, d_zeta, d_omega);
    ++it;
  }
  dense_svec(L.S, k, L.sS);
  __syncthreads();
  const double gs = dotd(g1, L.sS, d, L.red);
  if (t == 0) {
    sc[SC_LB] = sc[SC_BYC] - gs - eta * g2;
    sc[SC_LB_IT] = it;
  }
}

#define USBS_K_SWITCH(k, CALL) \
  switch (k) {                 \
    case 1: CALL(1); break;    \
    case 2: CALL(2); break;    \
    case 3: CALL(3); break;    \
    case 4: CALL(4); break;    \
    case 5: CALL(5); break;    \
    case 6: CALL(6); break;    \
    case 7: CALL(7); break;    \
    case 8: CALL(8); break;    \
    case 9: CALL(9); break;    \
    case 10: CALL(10); break;  \
    case 11: CALL(11); break;  \
    default: return (int)hipErrorInvalidValue; \
  }

int launch_ipm_quad(const IpmArgs& a, hipStream_t s) {
#define Q_CALL(KK) \
  k_ipm_quad<KK><<<1, IPM_NT, ipm_lds_bytes(KK) + sizeof(double) * (KK * (KK + 1) / 2) * (KK * (KK + 1) / 2), s>>>(a)
  USBS_K_SWITCH(a.k, Q_CALL)
#undef Q_CALL
  return (int)hipGetLastError();
}

int launch_ipm_lb(int k, int max_iters, double tu, double eps, const double* Glb, double* sc,
                  hipStream_t s) {
#define L_CALL(KK) k_ipm_lb<KK><<<1, IPM_NT, ipm_lds_bytes(KK), s>>>(max_iters, tu, eps, Glb, sc)
  USBS_K_SWITCH(k, L_CALL)
#undef L_CALL
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Nystrom core (utils/common.py:99-116) from the r x r Grams P^T P, Omega^T P,
// Omega^T Omega:  rho = sqrt(n) 1e-6 |P|_2,  B = sym(Omega^T (P + rho Omega)),
// L = chol(B),  M = (P + rho Omega) L^{-T} = E Sigma W^T with
// M^T M = L^{-1} P_rho^T P_rho L^{-T}.  Outputs G = L^{-T} W Sigma^{-1}
// (so E = P_rho G), lam = max(sigma^2 - rho, 0) in descending sigma order,
// sc[SC_NYS_RHO] = rho.  A Cholesky failure gives NaN G and lam.
__global__ void __launch_bounds__(IPM_NT)
k_nystrom_core(int n, int r, const double* PtPO, const double* OtO, double* G, double* lam,
               double* sc) {
  extern __shared__ double lds[];
  const int t = threadIdx.x, rr = r * r;
  double* A = lds;
  double* V = A + rr;
  double* tmp = V + rr;
  double* w = tmp + rr;
  double* coef = w + r;
  double* red = coef + 2 * r;
  double* B = red + 16;
  double* M = B + rr;
  double* X = M + rr;
  int* part = (int*)(X + rr);
  int* flag = part + r * r + 2 * r + 8;
  // PtPO = P^T [P | Omega] (r x 2r row-major): P^T P and P^T Omega
#define PtP(i, j) PtPO[(i) * 2 * r + (j)]
#define OtP(i, j) PtPO[(j) * 2 * r + r + (i)]
  for (int idx = t; idx < rr; idx += IPM_NT) A[idx] = PtP(idx / r, idx % r);
  __syncthreads();
  dense_eigh(A, r, w, V, tmp, coef, part, red);
  const double lmax = w[r - 1];
  const double normP = sqrt(lmax > 0.0 ? lmax : 0.0);
  const double rho = sqrt((double)n) * 1e-6 * normP;
  for (int idx = t; idx < rr; idx += IPM_NT) {
    const int i = idx / r, j = idx % r;
    const double bij = OtP(i, j) + rho * OtO[i * r + j];
    const double bji = OtP(j, i) + rho * OtO[j * r + i];
    B[idx] = 0.5 * (bij + bji);
    // P_rho^T P_rho = P^T P + rho (P^T Omega + Omega^T P) + rho^2 Omega^T Omega
    M[idx] = PtP(i, j) + rho * (OtP(j, i) + OtP(i, j)) + rho * rho * OtO[idx];
  }
  __syncthreads();
  const bool ok = dense_cholesky(B, r, flag);
  if (!ok) {
    for (int idx = t; idx < rr; idx += IPM_NT) G[idx] = NAN;
    for (int j = t; j < r; j += IPM_NT) lam[j] = NAN;
    if (t == 0) sc[SC_NYS_RHO] = rho;
    return;
  }
  // X = L^{-1} M L^{-T}: forward solves on the columns of M, then of X^T.
  if (t < r) {
    for (int i = 0; i < r; ++i) {
      double v = M[i * r + t];
      for (int p = 0; p < i; ++p) v -= B[i * r + p] * tmp[p * r + t];
      tmp[i * r + t] = v / B[i * r + i];
    }
  }
  __syncthreads();
  if (t < r) {
    for (int i = 0; i < r; ++i) {
      double v = tmp[t * r + i];
      for (int p = 0; p < i; ++p) v -= B[i * r + p] * X[p * r + t];
      X[i * r + t] = v / B[i * r + i];
    }
  }
  __syncthreads();
  for (int idx = t; idx < rr; idx += IPM_NT) {
    const int i = idx / r, j = idx % r;
    A[idx] = 0.5 * (X[i * r + j] + X[j * r + i]);
  }
  __syncthreads();
  dense_eigh(A, r, w, V, tmp, coef, part, red);
  // G = L^{-T} W (descending order) Sigma^{-1}: back substitution with L^T.
  if (t < r) {
    const int src = r - 1 - t;
    const double sig = sqrt(w[src] > 0.0 ? w[src] : 0.0);
    for (int i = r - 1; i >= 0; --i) {
      double v = V[i * r + src];
      for (int p = i + 1; p < r; ++p) v -= B[p * r + i] * X[p * r + t];
      X[i * r + t] = v / B[i * r + i];
    }
    for (int i = 0; i < r; ++i) G[i * r + t] = X[i * r + t] / sig;
    const double l = w[src] - rho;
    lam[t] = l > 0.0 ? l : 0.0;
  }
  if (t == 0) sc[SC_NYS_RHO] = rho;
#undef PtP
#undef OtP
}

int launch_nystrom_core(int n, int r, const double* PtPO, const double* OtO, double* G,
                        double* lam, double* sc, hipStream_t s) {
  if (r < 1 || r > 48) return (int)hipErrorInvalidValue;
  const size_t lds = sizeof(double) * (8 * r * r + 6 * r + 64);
  k_nystrom_core<<<1, IPM_NT, lds, s>>>(n, r, PtPO, OtO, G, lam, sc);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Scalar bookkeeping of one USBS iteration (solver/usbs.py:581-652).
__global__ void k_scalars(ScalarArgs a, double* sc) {
  if (threadIdx.x != 0) return;
  const double eta = sc[SC_ETA];
  const double primal = eta * sc[SC_PB] + sc[SC_OBJF];
  const double trX = eta * sc[SC_TRXB] + sc[SC_TRS];
  double lam0 = -a.lanczos_alpha[0];
  sc[SC_LAM0] = lam0;
  const double lp = lam0 < 0.0 ? 0.0 : lam0;
  const double cand_pen = sc[SC_BYC] + a.tu * lp;
  double ub = sc[SC_BYC] + (a.tu / a.tf) * lp;
  ub = ub > sc[SC_OBJUB] ? sc[SC_OBJUB] : ub;
  const double pen = sc[SC_PEN];
  const bool descent = a.beta * (pen - sc[SC_LB]) <= pen - cand_pen;
  sc[SC_DESCENT] = descent ? 1.0 : 0.0;
  sc[SC_PEN] = descent ? cand_pen : pen;
  sc[SC_CANDPEN] = cand_pen;
  sc[SC_OBJUB] = ub;
  sc[SC_P] = primal;
  sc[SC_TRX] = trX;
  sc[SC_TRXB] = eta * sc[SC_TRXB] + sc[SC_SUMLAM];
  sc[SC_PB] = eta * sc[SC_PB] + sc[SC_OBJFC];
  const double scx = a.SC * a.SX;
  const double ap = fabs(primal), au = fabs(ub);
  double gap = fabs(ub - primal) / scx;
  gap /= 1.0 + (ap < au ? au : ap) / scx;
  sc[SC_GAP] = gap;
  sc[SC_INFGAP] = (sqrt(sc[SC_INF2]) / a.SX) / (1.0 + sc[SC_BNORM]);
  sc[SC_MAXINF] = sc[SC_INFMAX] / a.SX;
  sc[SC_T] += 1.0;
}

int launch_scalars(const ScalarArgs& a, double* sc, hipStream_t s) {
  k_scalars<<<1, 64, 0, s>>>(a, sc);
  return (int)hipGetLastError();
}

}  // namespace usbs
