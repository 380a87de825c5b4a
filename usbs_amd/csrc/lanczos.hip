// Thick-restart Lanczos for the smallest eigenpairs of the assembled operator
// (solver/lanczos.py:59-284), two execution forms over the same tile code:
//
//  * per-op: one kernel per phase of a Lanczos step, the host reads the
//    control words once per restart (lanczos_perop);
//  * persistent: per restart, one cooperative launch runs every Lanczos step
//    of one or many instances with device-wide barriers between phases, then
//    the Ritz and rotation kernels run (lanczos_persistent).
//
// Step i (columns 0..i of Q valid, the reference's Q_valid mask at
// lanczos.py:102) is three phases over row tiles:
//   A  [i > first step: beta = |w_{i-1}|, q_i = w_{i-1} / beta]
//      y = Op q_i, part1 = Q^T y              (SpMV fused with CGS pass 1 dot)
//   B  h1 = sum part1; y -= Q h1; part2 = Q^T y
//   C  h2 = sum part2; y -= Q h2; part3 = |y|^2; w_i = y; alpha_i = h1_i + h2_i
// Normalising the previous residual is fused into the next step's SpMV: the
// operator gathers (T)(w_{i-1}[c] / beta), the same value the basis column
// holds, so a step needs three device-wide reductions.  Light-row (RR)
// tiles need two: phase B also sums |y|^2 and phase C takes
// beta^2 = |y|^2 - |h2|^2 (y - Q h2 with h2 = Q^T y removes exactly |h2|^2,
// and h2 is a rounding-level correction, so nothing cancels), except on a
// step whose dot products fill every butterfly slot, which keeps part3.  A breakdown
// (beta < 1e-15, lanczos.py:107-117) is handled before phase A by
// orthogonalising a counter-based random vector in place of w_{i-1}.  After
// the last step the residual is normalised into Q[:, m].
//
// Reductions over rows: per tile, column-wise sums in a fixed thread order,
// then tile partials summed in a fixed order (sum_tile_partials), so both
// forms give bitwise the same numbers.  The persistent form keeps its tile's
// rows of Q, the work vector and the CSR slice in LDS for the whole restart
// when they fit.
#include <cstring>
#include <type_traits>
#include <tuple>
#include <vector>

#include "lanczos.hpp"
#include "dense.hpp"

namespace usbs {

// Row tiles: 256 rows up to n = 65536 (at most 256 tiles); above, the
// smallest power of two >= 512 rows that gives at most 512 tiles, so that a
// compute unit runs two workgroups of the row-wise kernels (every block sums
// all tile partials, which keeps the count from growing further).
int tile_rows_for(int n) {
  if (n <= 256 * 256) return 256;
  int t = 512;
  while ((n + t - 1) / t > 512) t *= 2;
  return t;
}

int lanczos_heavy_tile_rows(int n) {
  int t = 16;
  while ((n + t - 1) / t > 256) t *= 2;
  return t;
}

// Heavy-row SpMV: rows of a wavefront taken together, entries per lane and row per round.
#ifndef LZ_HEAVY_RG
#define LZ_HEAVY_RG 2
#endif
#ifndef LZ_HEAVY_UN
#define LZ_HEAVY_UN 8
#endif

#define LZ_RED (USBS_NT + 64)  // doubles of reduction scratch in every tile kernel
// Tiles of more than 256 rows reduce row-wise in registers and transpose
// through LDS (block_reduce_cols): that scratch is larger.
#define LZ_BIG (USBS_RCH * USBS_NT + USBS_RCH * USBS_NPART + 64)

// Ritz eigensolve: arrowhead tridiagonalisation + multisection for m <= 32
// (LZ_TRIDIAG_MAX, arrow_eigh), parallel Jacobi above or on a degenerate
// cluster.  LDS: Y, w, cs, red, the small ints, then either Jacobi's T, V and
// pairing schedule or arrow_eigh's scratch (one after the other in time).
#define LZ_TRIDIAG_MAX 32
static inline size_t ritz_lds_bytes(int m) {
  const int M = m + (m & 1);
  const size_t head = (size_t)(m * m + 2 * M + 16 + (8 + M) / 2);
  const size_t jac = (size_t)(2 * M * (M + 1) + ((M - 1) * M + 1) / 2);
  const size_t arr = m <= LZ_TRIDIAG_MAX ? (size_t)arrow_scratch_doubles(m) : 0;
  return sizeof(double) * (head + (jac > arr ? jac : arr));
}
static inline size_t tile_lds_bytes(int maxc) { return sizeof(double) * (size_t)(LZ_BIG + maxc + 8); }

// Row-wise reduction (tiles of more than 256 rows, where the Q block streams
// from HBM): each thread owns whole rows, so every Q value is read once per
// phase, coalesced, and the per-thread sums are transposed through LDS.
template <typename T>
__device__ __forceinline__ bool lz_rowwise(const LzInst<T>& L) { return L.tile > 256; }
template <typename T>
__device__ __forceinline__ double* lz_hbuf(const LzInst<T>& L, double* red) {
  return red + (lz_rowwise(L) ? LZ_BIG : LZ_RED);
}
static inline size_t rot_lds_bytes(int m) { return sizeof(double) * (size_t)(m * m + 8); }

// ----------------------------------------------------------------------------
// tile primitives.  Every routine is called by all USBS_NT threads of a block
// and ends synchronised.  (rl, j) of a tile's Q block is qb[j * qld + rl].

// out[j] = sum_rl Q(rl, j) y(rl) for j < c: thread t sums rows g, g+G, ... of
// column t / G (g = t % G, G = NT / c), then the G sums of a column are added
// in g order.
template <typename T, typename M = MemPlain>
__device__ void col_dots(const T* qb, long long qld, const T* yb, int R, int c, double* red,
                         double* out) {
  const int t = threadIdx.x;
  int G = USBS_NT / c;
  if (G < 1) G = 1;
  const int j = t / G, g = t - (t / G) * G;
  if (j < c) {
    const T* qj = qb + (long long)j * qld;
    double s = 0.0;
    for (int rl = g; rl < R; rl += G) s += (double)qj[rl] * (double)yb[rl];
    red[t] = s;
  }
  __syncthreads();
  if (t < c) {
    double s = 0.0;
    for (int q = 0; q < G; ++q) s += red[t * G + q];
    M::st(out + t, s);
  }
  __syncthreads();
}

// out[0] = sum_rl y(rl)^2 (thread-strided partial sums, then in thread order).
template <typename T, typename M = MemPlain>
__device__ void self_dot(const T* yb, int R, double* red, double* out) {
  const int t = threadIdx.x;
  double s = 0.0;
  for (int rl = t; rl < R; rl += USBS_NT) s += (double)yb[rl] * (double)yb[rl];
  red[t] = s;
  __syncthreads();
  if (t < 64) {
    double v = 0.0;
    for (int q = t * (USBS_NT / 64); q < (t + 1) * (USBS_NT / 64); ++q) v += red[q];
    // 64 sums of 4, combined in lane order through LDS
    red[USBS_NT + t] = v;
  }
  __syncthreads();
  if (t == 0) {
    double v = 0.0;
    for (int q = 0; q < 64; ++q) v += red[USBS_NT + q];
    M::st(out, v);
  }
  __syncthreads();
}

// y(rl) = (T)(y(rl) - sum_{j<c} Q(rl, j) h[j]), j ascending.
template <typename T>
__device__ void row_update(const T* qb, long long qld, T* yb, int R, int c, const double* h) {
  for (int rl = threadIdx.x; rl < R; rl += USBS_NT) {
    double s = 0.0;
    for (int j = 0; j < c; ++j) s += (double)qb[(long long)j * qld + rl] * h[j];
    yb[rl] = (T)((double)yb[rl] - s);
  }
  __syncthreads();
}

// beta = sqrt(sum of the part3 tile partials), identical in every block.
template <typename T, typename M = MemPlain>
__device__ double lz_beta(const LzInst<T>& L, double* red) {
  double* s = red + USBS_NT;
  sum_tile_partials<M>(L.part3, L.ntiles, L.pstride, 1, s, red);
  return sqrt(s[0]);
}

// Rotation Q[:, :k_next] = Q[:, :m] Yr; Q[:, k_next] = Q[:, m] (lanczos.py:217-221).
// Eight output columns at a time, eight independent FMA chains, each summed
// over x in order.  Yr: m x m in LDS followed by 8 readable doubles (reads
// past the last column feed only outputs that are not stored).
template <typename T, int MAXC>
__device__ void tile_rotate(const LzInst<T>& L, int tl, int knext, const double* Yr) {
  const int m = L.m;
  const int kk = knext < m ? knext : m;
  const int r0 = tl * L.tile;
  const int r1 = min(L.n, r0 + L.tile);
  for (int r = r0 + (int)threadIdx.x; r < r1; r += USBS_NT) {
    double q[MAXC];
#pragma unroll
    for (int x = 0; x < MAXC; ++x) q[x] = (x < m) ? (double)L.Q[r + (long long)x * L.ldq] : 0.0;
    const T qm = L.Q[r + (long long)m * L.ldq];
    for (int j0 = 0; j0 < kk; j0 += 8) {
      double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int x = 0; x < MAXC; ++x)
        if (x < m) {
          const double* y = Yr + x * m + j0;
#pragma unroll
          for (int u = 0; u < 8; ++u) s[u] += q[x] * y[u];
        }
      T* out = L.Q + r + (long long)j0 * L.ldq;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (j0 + u < kk) out[u * L.ldq] = (T)s[u];
    }
    if (knext < m) L.Q[r + (long long)knext * L.ldq] = qm;
  }
}


// The arrowhead T of restart k (lanczos.py:128-146): diagonal alpha, arrow
// beta[a] at (a, k) for a < k, tridiagonal beta[a] at (a, a+1) for a >= k.
__device__ void lz_build_T(int m, int k, const double* alpha, const double* beta, double* Tw,
                           int M, int lda) {
  for (int idx = threadIdx.x; idx < M * lda; idx += blockDim.x) {
    const int i = idx / lda, j = idx - (idx / lda) * lda;
    double v = 0.0;
    if (i < m && j < m) {
      if (i == j) {
        v = alpha[i];
      } else {
        const int a = i < j ? i : j, b = i < j ? j : i;
        if (b == k && a < k) v = beta[a];
        else if (b == a + 1 && a >= k) v = beta[a];
      }
    }
    Tw[idx] = v;
  }
}

// ----------------------------------------------------------------------------
// Ritz step (one block): arrowhead T, eigh, convergence count, saved set,
// reorder, new alpha/beta and the rotation Yr (lanczos.py:128-168, 202-221).
// The convergence count, the gamma-ratio choice and the saved-set order are
// computed by wavefront 0, a lane per Ritz value (m <= 48 < 64), with the
// loops' tie-breaking: first non-converged value, first NaN ratio, else the
// first index of the largest ratio.
__device__ __forceinline__ void lz_ritz(int m, int nd, double tol, double rtol, int max_restarts, double* alpha, double* beta,
                        double* Yr, int* ctrl, double* lds) {
  const int t = threadIdx.x, nt = blockDim.x, lane = t & 63;
  const int M = m + (m & 1), lda = M + 1;
  double* Y = lds;                  // m x m sorted eigenvectors
  double* w = Y + m * m;            // m eigenvalues
  double* cs = w + M;               // M rotation coefficients
  double* red = cs + M;             // 16
  int* sh = (int*)(red + 16);       // small int scratch
  int* part = sh + 8;               // m ints: saved-set order
  double* U = (double*)(part + M);  // Jacobi's Tw, Vw, sched, or arrow_eigh's scratch
  double* Tw = U;                   // M x lda working copy of T (destroyed)
  double* Vw = Tw + M * lda;        // M x lda eigenvector accumulation
  int* sched = (int*)(Vw + M * lda);  // (M-1) * M ints: Jacobi pairing
  const int k = ctrl[LZ_K];
  bool jacobi = m > LZ_TRIDIAG_MAX;
  if (!jacobi && arrow_eigh(m, k, alpha, beta, w, Y, U)) jacobi = true;  // degenerate cluster
  if (jacobi) {
    lz_build_T(m, k, alpha, beta, Tw, M, lda);
    __syncthreads();
    block_eigh(Tw, lda, m, w, Y, Vw, cs, sched, red);
  }
  const double bm = beta[m - 1];
  if (t < 64) {
    const bool in = lane < m;
    const double wj = in ? w[lane] : 0.0;
    // convergence test (lanczos.py:205-208); with rtol > 0 the threshold has
    // a floor relative to the spectrum's scale (the FP32 basis policy)
    const double scale = fmax(fabs(readlane_d(wj, 0)), fabs(readlane_d(wj, m - 1)));
    const double thr = (rtol > 0.0 && rtol * scale > tol) ? rtol * scale : tol;
    const double res = bm * fabs(in ? Y[(m - 1) * m + lane] : 0.0);
    const unsigned long long bad = __ballot(in && !(res < thr));
    const int nc = bad ? __builtin_ctzll(bad) : m;
    int delta = m - nd;
    const int d2 = 2 * (m - nc) / 5;
    if (d2 < delta) delta = d2;
    const double lt = readlane_d(wj, nc < m ? nc : m - 1);
    const int kr = lane + delta;
    const double wkr = __shfl(wj, kr < m ? kr : m - 1, 64);
    double g = 0.0;
    if (kr <= m) g = (wj - lt) / (wkr - lt);
    const unsigned long long nanm = __ballot(in && g != g);
    int kl;
    if (nanm) {
      kl = __builtin_ctzll(nanm);
    } else {
      const double gmax = wave_max_dpp(in ? g : -INFINITY);
      kl = __builtin_ctzll(__ballot(in && g == gmax));
    }
    const int kright = kl + delta;
    const bool kept = in && (kl >= lane || kright < lane);
    const unsigned long long km = __ballot(kept);
    const unsigned long long inm = __ballot(in);
    const unsigned long long below = (1ull << lane) - 1ull;
    const int knext = __popcll(km);
    if (in) part[kept ? __popcll(km & below) : knext + __popcll(~km & inm & below)] = lane;
    if (t == 0) {
      sh[0] = knext;
      sh[1] = nc;
      const int restart = ctrl[LZ_RESTART] + 1;
      ctrl[LZ_K] = knext;
      ctrl[LZ_RESTART] = restart;
      ctrl[LZ_NCONV] = nc;
      ctrl[LZ_DONE] = (nc >= nd || restart >= max_restarts) ? 1 : 0;
    }
  }
  __syncthreads();
  const int knext = sh[0];
  for (int j = t; j < m; j += nt) {
    const int o = part[j];
    alpha[j] = (j < knext) ? w[o] : 0.0;
    beta[j] = (j < knext) ? bm * Y[(m - 1) * m + o] : 0.0;
  }
  for (int idx = t; idx < m * m; idx += nt) {
    const int x = idx / m, j = idx % m;
    Yr[idx] = Y[x * m + part[j]];
  }
  __syncthreads();
}


// ----------------------------------------------------------------------------
// phases of step i for tile tl.  Cached (C) tiles read Q, y and the CSR slice
// from LDS: qc (ldc x (m+1) column-major), yc (ldc), rp/ci/cv.

template <typename T, bool C>
struct TileCtx {
  int r0, R;
  T* qc;        // C: LDS Q block
  int ldc;
  T* yc;        // C: LDS work vector
  const int* rp;  // C: local row pointers (R + 1)
  const int* ci;  // C: column indices of the slice
  const T* cv;    // C: values of the slice
};

// Phase A (mode 0: SpMV; mode 2: y = Q[:, i+1] for the column-by-column QR;
// mode 3: only normalise w into column i).  first: q_i is already in Q.
// RW: the kernel may meet tiles of more than 256 rows (row-wise reductions);
// kernels instantiated with RW = false carry no row-wise code.
template <typename T, bool C, int MAXC, bool RW, typename M = MemPlain>
__device__ void phaseA(const LzInst<T>& L, const TileCtx<T, C>& X, int tl, int i, bool first,
                       int mode, double* red, double beta_known = -1.0) {
  const int t = threadIdx.x;
  const T* gb = L.w + (long long)((i + 1) & 1) * L.ldq;  // w_{i-1}
  T* wc = L.w + (long long)(i & 1) * L.ldq;               // this step's work vector (per-op)
  double beta = 1.0;
  if (!first) beta = beta_known >= 0.0 ? beta_known : lz_beta<T, M>(L, red);
  T* Qi = L.Q + (long long)i * L.ldq + X.r0;
  if (!first) {
    for (int rl = t; rl < X.R; rl += USBS_NT) {
      const T q = (T)((double)M::ld(gb + X.r0 + rl) / beta);
      M::st(Qi + rl, q);
      if constexpr (C) X.qc[(long long)i * X.ldc + rl] = q;
    }
  }
  if (mode == 3) { __syncthreads(); return; }
  T* yb;
  if constexpr (C) yb = X.yc; else yb = wc + X.r0;
  const T* qb;
  long long qld;
  if constexpr (C) { qb = X.qc; qld = X.ldc; } else { qb = L.Q + X.r0; qld = L.ldq; }
  const bool rowwise = RW && lz_rowwise(L);
  double acc[MAXC];
#pragma unroll
  for (int j = 0; j < MAXC; ++j) acc[j] = 0.0;
  const T* xq = L.Q + (long long)i * L.ldq;
  if (mode == 0 && L.heavy) {
    // a wavefront per row, LZ_HEAVY_RG rows of a wavefront at a time: lane l
    // takes entries l, l+64, ... of each row, LZ_HEAVY_UN per row per round,
    // all loads of a round in flight together; each row's lane sums add in
    // entry order, then a butterfly over the wavefront, and lane 0 writes.
    // Heavy tiles have at most 256 rows (column-wise dots).
    constexpr int RG = LZ_HEAVY_RG, UN = LZ_HEAVY_UN;
    const int lane = t & 63, wv = t >> 6;
    const int* ci;
    const T* cv;
    const int* rp;
    if constexpr (C) { ci = X.ci; cv = X.cv; rp = X.rp; }
    else { ci = L.col; cv = L.val; rp = L.row_ptr + X.r0; }
    for (int q0 = wv; q0 < X.R; q0 += RG * (USBS_NT / 64)) {
      int pb[RG], pe[RG];
      double s[RG];
      int rounds = 0;
#pragma unroll
      for (int g = 0; g < RG; ++g) {
        const int rl = q0 + g * (USBS_NT / 64);
        pb[g] = rl < X.R ? rp[rl] : 0;
        pe[g] = rl < X.R ? rp[rl + 1] : 0;
        s[g] = 0.0;
        const int nr = (pe[g] - pb[g] + 64 * UN - 1) / (64 * UN);
        rounds = nr > rounds ? nr : rounds;
      }
      for (int it = 0; it < rounds; ++it) {
        int c[RG][UN];
        T v[RG][UN];
#pragma unroll
        for (int g = 0; g < RG; ++g)
#pragma unroll
          for (int u = 0; u < UN; ++u) {
            // past the row's end: load entry 0 (in bounds) and mask it
            const int e = pb[g] + (it * UN + u) * 64 + lane;
            const bool ok = e < pe[g];
            const int cl = ci[ok ? e : 0];
            const T vl = cv[ok ? e : 0];
            c[g][u] = ok ? cl : 0;
            v[g][u] = ok ? vl : (T)0;
          }
#pragma unroll
        for (int g = 0; g < RG; ++g)
#pragma unroll
          for (int u = 0; u < UN; ++u) {
            const T x = first ? M::ld(xq + c[g][u]) : (T)((double)M::ld(gb + c[g][u]) / beta);
            s[g] += (double)v[g][u] * (double)x;
          }
      }
#pragma unroll
      for (int g = 0; g < RG; ++g) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s[g] += __shfl_xor(s[g], off, 64);
        const int rl = q0 + g * (USBS_NT / 64);
        if (lane == 0 && rl < X.R) yb[rl] = (T)s[g];
      }
    }
    __syncthreads();
    col_dots<T, M>(qb, qld, yb, X.R, i + 1, red, L.part1 + (long long)tl * L.pstride);
    return;
  }
  for (int rl = t; rl < X.R; rl += USBS_NT) {
    double s = 0.0;
    if (mode == 0) {
      if constexpr (C) {
        const int p1 = X.rp[rl + 1];
        for (int p = X.rp[rl]; p < p1; ++p) {
          const int c = X.ci[p];
          const T x = first ? M::ld(xq + c) : (T)((double)M::ld(gb + c) / beta);
          s += (double)X.cv[p] * (double)x;
        }
      } else {
        const int r = X.r0 + rl;
        const int p1 = L.row_ptr[r + 1];
        for (int p = L.row_ptr[r]; p < p1; ++p) {
          const int c = L.col[p];
          const T x = first ? xq[c] : (T)((double)gb[c] / beta);
          s += (double)L.val[p] * (double)x;
        }
      }
    } else {
      s = (double)L.Q[(long long)(i + 1) * L.ldq + X.r0 + rl];
    }
    const T yv = (T)s;
    yb[rl] = yv;
    if (rowwise) {
      // loads of columns past i read column i (in bounds) and are masked, so
      // every load of the row is issued without a branch
      const double yd = (double)yv;
#pragma unroll
      for (int j = 0; j < MAXC; ++j) {
        const double qv = (double)qb[(long long)(j <= i ? j : i) * qld + rl];
        acc[j] += (j <= i) ? qv * yd : 0.0;
      }
    }
  }
  if (rowwise) {
    block_reduce_cols<MAXC>(acc, i + 1, red, red + USBS_RCH * USBS_NT,
                            L.part1 + (long long)tl * L.pstride);
    return;
  }
  __syncthreads();
  col_dots<T, M>(qb, qld, yb, X.R, i + 1, red, L.part1 + (long long)tl * L.pstride);
}

// Phases B (final = 0) and C (final = 1).
template <typename T, bool C, int MAXC, bool RW, typename M = MemPlain>
__device__ void phaseBC(const LzInst<T>& L, const TileCtx<T, C>& X, int tl, int i, int final_pass,
                        double* red) {
  const int t = threadIdx.x;
  double* h = lz_hbuf(L, red);
  const double* src = final_pass ? L.part2 : L.part1;
  sum_tile_partials<M>(src, L.ntiles, L.pstride, i + 1, h, red);
  if (tl == 0)
    for (int j = t; j <= i; j += USBS_NT) M::st((final_pass ? L.h2 : L.h1) + j, h[j]);
  T* yb;
  const T* qb;
  long long qld;
  T* wc = L.w + (long long)(i & 1) * L.ldq + X.r0;
  if constexpr (C) { yb = X.yc; qb = X.qc; qld = X.ldc; } else { yb = wc; qb = L.Q + X.r0; qld = L.ldq; }
  double* part_out = (final_pass ? L.part3 : L.part2) + (long long)tl * L.pstride;
  if (RW && lz_rowwise(L)) {
    // one pass over the tile's Q rows: update, then the next dots
    double acc[MAXC];
#pragma unroll
    for (int j = 0; j < MAXC; ++j) acc[j] = 0.0;
    for (int rl = t; rl < X.R; rl += USBS_NT) {
      double q[MAXC];
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < MAXC; ++j) {
        const double qv = (double)qb[(long long)(j <= i ? j : i) * qld + rl];
        q[j] = (j <= i) ? qv : 0.0;
        s += q[j] * ((j <= i) ? h[j] : 0.0);
      }
      const T yv = (T)((double)yb[rl] - s);
      yb[rl] = yv;
      const double yd = (double)yv;
      if (final_pass) {
        acc[0] += yd * yd;
      } else {
#pragma unroll
        for (int j = 0; j < MAXC; ++j)
          if (j <= i) acc[j] += q[j] * yd;
      }
    }
    block_reduce_cols<MAXC>(acc, final_pass ? 1 : i + 1, red, red + USBS_RCH * USBS_NT, part_out);
  } else {
    row_update(qb, qld, yb, X.R, i + 1, h);
    if (!final_pass) {
      col_dots<T, M>(qb, qld, yb, X.R, i + 1, red, part_out);
    } else {
      if constexpr (C)
        for (int rl = t; rl < X.R; rl += USBS_NT) M::st(wc + rl, yb[rl]);
      self_dot<T, M>(yb, X.R, red, part_out);
    }
  }
  if (final_pass && tl == 0 && t == 0) {
    L.alpha[i] = M::ld(L.h1 + i) + h[i];
    L.ctrl[LZ_STEPS] += 1;
  }
}

// counter_normal kept out of line: the breakdown path is cold, and inlined its
// transcendental code would raise the register count of the step kernels.
__device__ __attribute__((noinline)) double counter_normal_cold(uint64_t key, uint64_t idx) {
  return counter_normal(key, idx);
}

// Breakdown restart vector for the residual w_{i-1} (lanczos.py:107-117):
// counter normals keyed by step i, orthogonalised twice against columns
// 0..i-1 and left in w_{i-1} with |.|^2 partials in part3.  sub = 0, 1, 2 are
// the three reduction phases.
template <typename T>
__device__ void breakdown_sub(const LzInst<T>& L, int tl, int i, int sub, double* red) {
  const int t = threadIdx.x;
  const int r0 = tl * L.tile;
  const int R = min(L.n, r0 + L.tile) - r0;
  T* gb = L.w + (long long)((i + 1) & 1) * L.ldq + r0;
  const T* qb = L.Q + r0;
  if (sub == 0) {
    for (int rl = t; rl < R; rl += USBS_NT) gb[rl] = (T)counter_normal_cold((uint64_t)i, (uint64_t)(r0 + rl));
    __syncthreads();
    col_dots(qb, L.ldq, (const T*)gb, R, i, red, L.part1 + (long long)tl * L.pstride);
  } else {
    double* h = lz_hbuf(L, red);
    sum_tile_partials(sub == 1 ? L.part1 : L.part2, L.ntiles, L.pstride, i, h, red);
    row_update(qb, L.ldq, gb, R, i, h);
    if (sub == 1) col_dots(qb, L.ldq, (const T*)gb, R, i, red, L.part2 + (long long)tl * L.pstride);
    else self_dot((const T*)gb, R, red, L.part3 + (long long)tl * L.pstride);
  }
}

// The same restart vector for a tile cached in LDS (the coherent persistent
// kernel): built in the LDS work vector with the arithmetic of
// breakdown_sub, and published to w_{i-1} in global memory in the third pass.
template <typename T, bool C>
__device__ void breakdown_sub_lds(const LzInst<T>& L, const TileCtx<T, C>& X, int tl, int i, int sub,
                                  double* red) {
  using M = MemCoherent;
  const int t = threadIdx.x;
  T* yc = X.yc;
  if (sub == 0) {
    for (int rl = t; rl < X.R; rl += USBS_NT) yc[rl] = (T)counter_normal_cold((uint64_t)i, (uint64_t)(X.r0 + rl));
    __syncthreads();
    col_dots<T, M>(X.qc, X.ldc, yc, X.R, i, red, L.part1 + (long long)tl * L.pstride);
  } else {
    double* h = lz_hbuf(L, red);
    sum_tile_partials<M>(sub == 1 ? L.part1 : L.part2, L.ntiles, L.pstride, i, h, red);
    row_update(X.qc, X.ldc, yc, X.R, i, h);
    if (sub == 1) {
      col_dots<T, M>(X.qc, X.ldc, yc, X.R, i, red, L.part2 + (long long)tl * L.pstride);
    } else {
      T* gb = L.w + (long long)((i + 1) & 1) * L.ldq + X.r0;
      for (int rl = t; rl < X.R; rl += USBS_NT) M::st(gb + rl, yc[rl]);
      self_dot<T, M>(yc, X.R, red, L.part3 + (long long)tl * L.pstride);
    }
  }
}

// beta_{i-1} from part3: recorded in beta[i-1]; returns whether it broke down.
template <typename T, typename M = MemPlain>
__device__ bool lz_record_beta(const LzInst<T>& L, int i, bool writer, double* red,
                               double* beta_out = nullptr) {
  const double b = lz_beta<T, M>(L, red);
  if (beta_out) *beta_out = b;
  const bool brk = b < 1e-15;
  if (writer && threadIdx.x == 0) {
    L.beta[i - 1] = b;
    L.ctrl[LZ_BREAKDOWN] = brk ? 1 : 0;
    if (brk) L.ctrl[LZ_NBREAK] += 1;
  }
  __syncthreads();
  return brk;
}


// ----------------------------------------------------------------------------
// per-op kernels (one block per tile; Q and the work vector in global memory)

template <typename T>
__device__ TileCtx<T, false> plain_tile(const LzInst<T>& L, int tl) {
  TileCtx<T, false> X{};
  X.r0 = tl * L.tile;
  X.R = min(L.n, X.r0 + L.tile) - X.r0;
  return X;
}

// Phase A of step i.  record (restart steps without a k_lz_pre launch before
// them): every block computes beta_{i-1} from the part3 partials, block 0
// records it, and on a breakdown the step is not taken: block 0 marks it in
// ctrl[LZ_PENDING] and the host resumes there (lanczos_perop).
template <typename T, int MAXC, bool RW>
__global__ void __launch_bounds__(USBS_NT) k_lz_A(LzInst<T> L, int i, int first, int mode,
                                                  int record) {
  extern __shared__ double lds[];
  if (L.ctrl[LZ_PENDING]) return;
  double beta = -1.0;
  if (record && !first) {
    beta = lz_beta(L, lds);
    __syncthreads();
    const bool brk = beta < 1e-15;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      if (brk) {
        L.ctrl[LZ_PENDING] = i;
      } else {
        L.beta[i - 1] = beta;
        L.ctrl[LZ_BREAKDOWN] = 0;
      }
    }
    if (brk) return;
  }
  phaseA<T, false, MAXC, RW>(L, plain_tile(L, blockIdx.x), blockIdx.x, i, first != 0, mode, lds,
                             beta);
}

template <typename T, int MAXC, bool RW>
__global__ void __launch_bounds__(USBS_NT) k_lz_BC(LzInst<T> L, int i, int final_pass) {
  extern __shared__ double lds[];
  if (L.ctrl[LZ_PENDING]) return;
  phaseBC<T, false, MAXC, RW>(L, plain_tile(L, blockIdx.x), blockIdx.x, i, final_pass, lds);
}

// beta_{i-1} and, on breakdown, the restart vector: one block walking every
// tile in order (same tile arithmetic as the multi-block phases).
template <typename T>
__global__ void __launch_bounds__(USBS_NT) k_lz_pre(LzInst<T> L, int i) {
  extern __shared__ double lds[];
  if (threadIdx.x == 0) L.ctrl[LZ_PENDING] = 0;
  if (!lz_record_beta(L, i, true, lds)) return;
  for (int sub = 0; sub < 3; ++sub) {
    for (int tl = 0; tl < L.ntiles; ++tl) breakdown_sub(L, tl, i, sub, lds);
    __syncthreads();
  }
}

#define LZ_RITZ_NT 256
template <typename T>
__global__ void __launch_bounds__(LZ_RITZ_NT) k_lz_ritz(LzInst<T> L) {
  extern __shared__ double lds[];
  if (L.ctrl[LZ_PENDING]) return;
  lz_ritz(L.m, L.nd, L.tol, L.rtol, L.max_restarts, L.alpha, L.beta, L.Yr, L.ctrl, lds);
}

// Ritz step of restart r for every instance still running (block = instance).
template <typename T>
__global__ void __launch_bounds__(LZ_RITZ_NT) k_lz_ritz_multi(const LzInst<T>* __restrict__ insts,
                                                              int r) {
  extern __shared__ double lds[];
  const LzInst<T>& L = insts[blockIdx.x];
  if (L.ctrl[LZ_DONE] || L.ctrl[LZ_RESTART] != r || L.ctrl[LZ_ERROR]) return;
  lz_ritz(L.m, L.nd, L.tol, L.rtol, L.max_restarts, L.alpha, L.beta, L.Yr, L.ctrl, lds);
}

template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_rotate(LzInst<T> L) {
  extern __shared__ double lds[];
  if (L.ctrl[LZ_PENDING]) return;
  const int knext = L.ctrl[LZ_K];
  const int mm = L.m * L.m;
  for (int idx = threadIdx.x; idx < mm; idx += USBS_NT) lds[idx] = L.Yr[idx];
  if (threadIdx.x < 8) lds[mm + threadIdx.x] = 0.0;
  __syncthreads();
  tile_rotate<T, MAXC>(L, blockIdx.x, knext, lds);
}

// Rotation after the Ritz step of restart r (block = (instance, tile)).
template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_rotate_multi(const LzInst<T>* __restrict__ insts,
                                                             const int2* __restrict__ tilemap, int r) {
  extern __shared__ double lds[];
  const int2 mp = tilemap[blockIdx.x];
  const LzInst<T>& L = insts[mp.x];
  if (L.ctrl[LZ_RESTART] != r + 1 || L.ctrl[LZ_ERROR]) return;
  const int knext = L.ctrl[LZ_K];
  const int mm = L.m * L.m;
  for (int idx = threadIdx.x; idx < mm; idx += USBS_NT) lds[idx] = L.Yr[idx];
  if (threadIdx.x < 8) lds[mm + threadIdx.x] = 0.0;
  __syncthreads();
  tile_rotate<T, MAXC>(L, mp.y, knext, lds);
}

// ----------------------------------------------------------------------------
// Row-in-registers (RR) tiles: light operator rows (a thread per row in the
// SpMV) in tiles of at most USBS_NT rows.  Thread t owns row r0 + t of its
// tile and holds that row of Q in registers (q[MAXC]) and its element of the
// work vector (y), so the CGS updates y -= Q h are register dot products and
// only the column sums cross threads: a reduce-scatter butterfly over each
// wavefront's lanes, then the four wavefronts added in a fixed order.  The
// per-op kernels load q from Q at the start of each phase; the persistent
// kernel keeps q and y in registers for the whole restart.  Both use these
// routines, so they agree bitwise.

// Slots of the butterfly: a power of two >= the most columns a dot product
// has (m = MAXC - 1).
template <int MAXC>
struct RRSlots {
  static constexpr int C = (MAXC - 1 <= 16) ? 16 : (MAXC - 1 <= 32) ? 32 : 64;
};

// Reduce-scatter over the 64 lanes: stage s exchanges half of the live slots
// with lane ^ (1 << s) (lane_xor: the large early stages are DPP moves); a
// lane keeps the upper half when lane bit s is set.  After log2(C) stages
// lane l holds the wavefront's sum of slot rr_slot<C>(l), the bit reversal
// of its low log2(C) bits; the remaining stages add the lanes that share a
// slot.  Each sum is formed by one fixed lane, so the result is deterministic.
template <int C>
__device__ __forceinline__ int rr_slot(int lane) {
  constexpr int LG = (C >= 64) ? 6 : (C >= 32) ? 5 : (C >= 16) ? 4 : (C >= 8) ? 3 : (C >= 4) ? 2 : (C >= 2) ? 1 : 0;
  int s = 0;
  for (int b = 0; b < LG; ++b) s |= ((lane >> b) & 1) << (LG - 1 - b);
  return s;
}
template <int C>
__device__ __forceinline__ void rr_butterfly(double (&v)[C]) {
  const int lane = threadIdx.x & 63;
  static_for<6>([&](auto ST) {
    constexpr int mask = 1 << ST;
    constexpr int half = C >> (ST + 1);
    if constexpr (half >= 1) {
      const bool up = (lane & mask) != 0;
      static_for<half>([&](auto S) {
        const double a = v[S], b = v[S + half];
        const double send = up ? a : b;
        const double keep = up ? b : a;
        v[S] = keep + lane_xor<mask>(send);
      });
    } else {
      v[0] += lane_xor<mask>(v[0]);
    }
  });
}

// out[j] (j < c <= C) = sum over the block's rows of v[j]: the butterfly,
// then wavefronts 0..3 added as (w0 + w1) + (w2 + w3).  red: 4C doubles of LDS.
template <int C, typename M>
__device__ __forceinline__ void rr_reduce_cols(double (&v)[C], int c, double* red, double* out) {
  rr_butterfly<C>(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane < C) red[wv * C + rr_slot<C>(lane)] = v[0];
  __syncthreads();
  if ((int)threadIdx.x < c) {
    const int j = threadIdx.x;
    M::st(out + j, (red[j] + red[C + j]) + (red[2 * C + j] + red[3 * C + j]));
  }
  __syncthreads();
}

// h[j] (LDS) = sum over tiles of part[tile * stride + j], j < c <= 64: P = 8
// lanes per column (4 above 32 columns); lane q of a column adds tiles q,
// q + P, ... in order, then the P lanes are added by DPP (xor 1, 2, 4).  Every
// lane of the block takes part; the order depends only on (c, ntiles).
template <typename M>
__device__ __forceinline__ void rr_sum_tiles(const double* __restrict__ part, int ntiles, int stride,
                                             int c, double* out) {
  const int t = threadIdx.x;
  const int P = c <= 32 ? 8 : 4;
  const int j = t / P, q = t & (P - 1);
  const int jj = j < c ? j : 0;
  double s = 0.0;
  for (int tl0 = q; tl0 < ntiles; tl0 += 16 * P) {
    double v[16];
    static_for<16>([&](auto U) {
      const int tl = tl0 + U * P;
      v[U] = M::ld(part + (size_t)(tl < ntiles ? tl : 0) * stride + jj);
    });
    static_for<16>([&](auto U) {
      if (tl0 + U * P < ntiles) s += v[U];
    });
  }
  s += lane_xor<1>(s);
  s += lane_xor<2>(s);
  if (P == 8) s += lane_xor<4>(s);
  if (q == 0 && j < c) out[j] = s;
  __syncthreads();
}

// beta = sqrt(sum of the part3 tile partials), identical in every block.
template <typename T, typename M>
__device__ __forceinline__ double rr_beta(const LzInst<T>& L, double* red) {
  double* s = red + USBS_NT;
  rr_sum_tiles<M>(L.part3, L.ntiles, L.pstride, 1, s);
  const double b = sqrt(s[0]);
  __syncthreads();
  return b;
}

// Whether step i's phase B sums |y|^2 with its dot products (one more
// column) and phase C takes beta from it: when the butterfly has a slot left
// and the instance does not ask for exact norms.
template <int MAXC, typename T>
__device__ __forceinline__ bool rr_pyth(const LzInst<T>& L, int i) {
  return !L.exact_norm && i + 2 <= RRSlots<MAXC>::C;
}

// |w_{i-1}| at the start of step i: phase C of step i-1 left it in h2[m]
// (rr_pyth; a breakdown restart vector's norm is stored there too), else the
// part3 partials.  known >= 0: the caller holds that value already.
template <typename T, int MAXC, typename M>
__device__ __forceinline__ double rr_wnorm(const LzInst<T>& L, int i, double* red, double known = -1.0) {
  if (rr_pyth<MAXC>(L, i - 1)) return known >= 0.0 ? known : M::ld(L.h2 + L.m);
  return rr_beta<T, M>(L, red);
}

// beta_{i-1} recorded in beta[i-1] (writer); returns whether it broke down.
template <typename T, int MAXC, typename M>
__device__ __forceinline__ bool rr_record_beta(const LzInst<T>& L, int i, bool writer, double* red,
                                               double* beta_out = nullptr, double known = -1.0) {
  const double b = rr_wnorm<T, MAXC, M>(L, i, red, known);
  if (beta_out) *beta_out = b;
  const bool brk = b < 1e-15;
  if (writer && threadIdx.x == 0) {
    L.beta[i - 1] = b;
    L.ctrl[LZ_BREAKDOWN] = brk ? 1 : 0;
    if (brk) L.ctrl[LZ_NBREAK] += 1;
  }
  __syncthreads();
  return brk;
}

// The thread's row of Q[:, 0..c-1] (columns past c-1 and rows past the tile
// read column 0 / the tile's first row, in bounds, and are zeroed).
template <typename T, int MAXC, typename M>
__device__ __forceinline__ void rr_load_q(const LzInst<T>& L, int r, bool valid, int c, T (&q)[MAXC]) {
  static_for<MAXC>([&](auto J) {
    constexpr int j = J;
    const T v = M::ld(L.Q + (long long)(j < c ? j : 0) * L.ldq + r);
    q[j] = (valid && j < c) ? v : (T)0;
  });
}

// v[j] = q[j] y for j <= i (and j < MAXC), v[i+1] = y^2 when sq, else 0.
template <typename T, int MAXC, int C>
__device__ __forceinline__ void rr_products(const T (&q)[MAXC], double yd, int i, double (&v)[C],
                                            bool sq = false) {
  const double y2 = sq ? yd * yd : 0.0;
  static_for<C>([&](auto J) {
    constexpr int j = J;
    if constexpr (j < MAXC) v[j] = (j <= i) ? (double)q[j] * yd : (j == i + 1 ? y2 : 0.0);
    else v[j] = (j == i + 1) ? y2 : 0.0;
  });
}

// sum_{j < c} q[j] h[j], j ascending.  h: at least MAXC readable doubles
// (the terms past c are loaded with the others and not added).  hh (when
// given): sum_{j < c} h[j]^2, over j = 4u + v as four chains (v = 0..3)
// added (0 + 1) + (2 + 3).
template <typename T, int MAXC>
__device__ __forceinline__ double rr_qdot(const T (&q)[MAXC], const double* h, int c,
                                          double* hh = nullptr) {
  double hv[MAXC];
  static_for<MAXC>([&](auto J) { hv[J] = h[J]; });
  double s = 0.0;
  static_for<MAXC>([&](auto J) {
    if (J < c) s += (double)q[J] * hv[J];
  });
  if (hh) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    static_for<MAXC>([&](auto J) {
      if (J < c) a[J & 3] = fma(hv[J], hv[J], a[J & 3]);
    });
    *hh = (a[0] + a[1]) + (a[2] + a[3]);
  }
  return s;
}

// Loads of phase A that do not need beta, issued before it is summed so that
// all of them (and the beta partials) are in flight together: the thread's
// element of w_{i-1} and the operand values of its row's first RR_PF
// operator entries (w_{i-1} raw, or q_i on the first step).  Entries past
// the row read a valid entry of the row (or column 0 of an empty row).
#define RR_PF 8
template <typename T>
struct RRPre {
  T w;
  T x[RR_PF];
};
template <typename T, typename M>
__device__ __forceinline__ void rr_prefetch(const LzInst<T>& L, int tl, int i, bool first, int mode,
                                            const int* rp, const int* ci, RRPre<T>& pf) {
  const int t = threadIdx.x;
  const int r0 = tl * L.tile, R = min(L.n, r0 + L.tile) - r0;
  const bool valid = t < R;
  const int r = r0 + (valid ? t : 0);
  const T* gb = L.w + (long long)((i + 1) & 1) * L.ldq;
  pf.w = first ? (T)0 : M::ld(gb + r);
  if (mode == 0) {
    const T* src = first ? L.Q + (long long)i * L.ldq : gb;
    const int p0 = valid ? rp[t] : 0, p1 = valid ? rp[t + 1] : 0;
    static_for<RR_PF>([&](auto U) {
      const int p = p0 + U;
      const int c = p < p1 ? ci[p] : (p1 > p0 ? ci[p0] : 0);
      pf.x[U] = M::ld(src + c);
    });
  }
}

// Phase A of step i for the thread's row: mode 0 y = Op q_i (CSR rows rp[t]..
// rp[t+1] of ci/cv), mode 2 y = Q[:, i+1] (column-by-column QR), mode 3 only
// normalise w_{i-1} into Q[:, i].  first: q_i is already in Q (and q).
// part1 = Q^T y.  beta: |w_{i-1}| (unused when first).  pf: rr_prefetch of
// the same step.  The operator sum runs over the row's entries in order.
template <typename T, int MAXC, typename M>
__device__ __forceinline__ void rr_phaseA(const LzInst<T>& L, int tl, int i, bool first, int mode, double beta,
                          const RRPre<T>& pf, T (&q)[MAXC], T& y, const int* rp, const int* ci,
                          const T* cv, double* red) {
  constexpr int C = RRSlots<MAXC>::C;
  const int t = threadIdx.x;
  const int r0 = tl * L.tile, R = min(L.n, r0 + L.tile) - r0;
  const bool valid = t < R;
  const int r = r0 + (valid ? t : 0);
  const T* gb = L.w + (long long)((i + 1) & 1) * L.ldq;  // w_{i-1}
  if (!first) {
    const T qi = (T)((double)pf.w / beta);
    if (valid) M::st(L.Q + (long long)i * L.ldq + r, qi);
    const T qv = valid ? qi : (T)0;
    static_for<MAXC>([&](auto J) { q[J] = (J == i) ? qv : q[J]; });
  }
  if (mode == 3) return;
  double s = 0.0;
  if (mode == 0) {
    const T* src = first ? L.Q + (long long)i * L.ldq : gb;
    if (valid) {
      const int p0 = rp[t], p1 = rp[t + 1];
      static_for<RR_PF>([&](auto U) {
        const int p = p0 + U;
        if (p < p1) {
          const T x = first ? pf.x[U] : (T)((double)pf.x[U] / beta);
          s += (double)cv[p] * (double)x;
        }
      });
      for (int pb = p0 + RR_PF; pb < p1; pb += RR_PF) {
        T xv[RR_PF];
        static_for<RR_PF>([&](auto U) {
          const int p = pb + U;
          xv[U] = M::ld(src + ci[p < p1 ? p : pb]);
        });
        static_for<RR_PF>([&](auto U) {
          const int p = pb + U;
          if (p < p1) {
            const T x = first ? xv[U] : (T)((double)xv[U] / beta);
            s += (double)cv[p] * (double)x;
          }
        });
      }
    }
  } else {
    s = (double)M::ld(L.Q + (long long)(i + 1) * L.ldq + r);
  }
  y = (T)s;
  double v[C];
  rr_products<T, MAXC, C>(q, valid ? (double)y : 0.0, i, v);
  rr_reduce_cols<C, M>(v, i + 1, red, L.part1 + (long long)tl * L.pstride);
}

// Phases B (final 0: part2 = Q^T y, and |y|^2 in column i + 1 when
// rr_pyth) and C (final 1) after y -= Q h with h the sum of the previous
// pass's tile partials.  C returns beta = |y| = sqrt(|y_B|^2 - |h2|^2) and
// tile 0 stores it in h2[m] (rr_pyth), or writes part3 = |y|^2 partials and
// returns -1.
template <typename T, int MAXC, typename M>
__device__ __forceinline__ double rr_phaseBC(const LzInst<T>& L, int tl, int i, int final_pass, T (&q)[MAXC], T& y,
                           double* red) {
  constexpr int C = RRSlots<MAXC>::C;
  const bool pyth = rr_pyth<MAXC>(L, i);
  const int t = threadIdx.x;
  const int r0 = tl * L.tile, R = min(L.n, r0 + L.tile) - r0;
  const bool valid = t < R;
  double* h = lz_hbuf(L, red);
  rr_sum_tiles<M>(final_pass ? L.part2 : L.part1, L.ntiles, L.pstride,
                  (final_pass && pyth) ? i + 2 : i + 1, h);
  if (tl == 0)
    for (int j = t; j <= i; j += USBS_NT) M::st((final_pass ? L.h2 : L.h1) + j, h[j]);
  double hh = 0.0;
  y = (T)((double)y - rr_qdot<T, MAXC>(q, h, i + 1, (final_pass && pyth) ? &hh : nullptr));
  const double yd = valid ? (double)y : 0.0;
  double beta = -1.0;
  if (!final_pass) {
    double v[C];
    rr_products<T, MAXC, C>(q, yd, i, v, pyth);
    rr_reduce_cols<C, M>(v, pyth ? i + 2 : i + 1, red, L.part2 + (long long)tl * L.pstride);
  } else {
    if (pyth) {
      const double b2 = h[i + 1] - hh;
      beta = sqrt(b2 > 0.0 ? b2 : 0.0);
      if (tl == 0 && t == 0) M::st(L.h2 + L.m, beta);
    } else {
      double v[1] = {yd * yd};
      rr_reduce_cols<1, M>(v, 1, red, L.part3 + (long long)tl * L.pstride);
    }
    if (tl == 0 && t == 0) {
      L.alpha[i] = M::ld(L.h1 + i) + h[i];
      L.ctrl[LZ_STEPS] += 1;
    }
    __syncthreads();  // h is read above and rewritten by the next sum
  }
  return beta;
}

// Breakdown restart vector for tile tl (lanczos.py:107-117): counter normals
// keyed by step i, orthogonalised twice against Q[:, 0..i-1], left in w_{i-1}
// with |.|^2 partials in part3; sub = 0, 1, 2 are the three reduction passes.
// q: the thread's row of Q[:, 0..i-1].
template <typename T, int MAXC, typename M>
__device__ __forceinline__ void rr_breakdown_sub(const LzInst<T>& L, int tl, int i, int sub, const T (&q)[MAXC],
                                 double* red) {
  constexpr int C = RRSlots<MAXC>::C;
  const int t = threadIdx.x;
  const int r0 = tl * L.tile, R = min(L.n, r0 + L.tile) - r0;
  const bool valid = t < R;
  const int r = r0 + (valid ? t : 0);
  T* gb = L.w + (long long)((i + 1) & 1) * L.ldq;
  T y;
  if (sub == 0) {
    y = (T)counter_normal_cold((uint64_t)i, (uint64_t)r);
  } else {
    double* h = lz_hbuf(L, red);
    rr_sum_tiles<M>(sub == 1 ? L.part1 : L.part2, L.ntiles, L.pstride, i, h);
    y = (T)((double)M::ld(gb + r) - rr_qdot<T, MAXC>(q, h, i));
  }
  if (valid) M::st(gb + r, y);
  const double yd = valid ? (double)y : 0.0;
  if (sub < 2) {
    double v[C];
    rr_products<T, MAXC, C>(q, yd, i - 1, v);
    rr_reduce_cols<C, M>(v, i, red, (sub == 0 ? L.part1 : L.part2) + (long long)tl * L.pstride);
  } else {
    double v[1] = {yd * yd};
    rr_reduce_cols<1, M>(v, 1, red, L.part3 + (long long)tl * L.pstride);
  }
}

// per-op RR kernels (one block per tile)
template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_A_rr(LzInst<T> L, int i, int first, int mode, int record) {
  extern __shared__ double lds[];
  if (L.ctrl[LZ_PENDING]) return;
  const int tl = blockIdx.x, t = threadIdx.x;
  RRPre<T> pf;
  rr_prefetch<T, MemPlain>(L, tl, i, first != 0, mode, L.row_ptr + tl * L.tile, L.col, pf);
  double beta = -1.0;
  if (!first) {
    beta = rr_wnorm<T, MAXC, MemPlain>(L, i, lds);
    if (record) {
      __syncthreads();
      const bool brk = beta < 1e-15;
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (brk) {
          L.ctrl[LZ_PENDING] = i;
        } else {
          L.beta[i - 1] = beta;
          L.ctrl[LZ_BREAKDOWN] = 0;
        }
      }
      if (brk) return;
    }
  }
  const int r0 = tl * L.tile, R = min(L.n, r0 + L.tile) - r0;
  const bool valid = t < R;
  const int r = r0 + (valid ? t : 0);
  T q[MAXC];
  rr_load_q<T, MAXC, MemPlain>(L, r, valid, i + 1, q);
  T y = (T)0;
  rr_phaseA<T, MAXC, MemPlain>(L, tl, i, first != 0, mode, beta, pf, q, y, L.row_ptr + r0, L.col, L.val, lds);
  if (mode != 3 && valid) L.w[(long long)(i & 1) * L.ldq + r] = y;
}

template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_BC_rr(LzInst<T> L, int i, int final_pass) {
  extern __shared__ double lds[];
  if (L.ctrl[LZ_PENDING]) return;
  const int tl = blockIdx.x, t = threadIdx.x;
  const int r0 = tl * L.tile, R = min(L.n, r0 + L.tile) - r0;
  const bool valid = t < R;
  const int r = r0 + (valid ? t : 0);
  T* wc = L.w + (long long)(i & 1) * L.ldq;
  T q[MAXC];
  rr_load_q<T, MAXC, MemPlain>(L, r, valid, i + 1, q);
  T y = wc[r];
  rr_phaseBC<T, MAXC, MemPlain>(L, tl, i, final_pass, q, y, lds);
  if (valid) wc[r] = y;
}

// beta_{i-1} and, on breakdown, the restart vector: one block walking every
// tile in order (the tile arithmetic of the multi-block phases).
template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_pre_rr(LzInst<T> L, int i) {
  extern __shared__ double lds[];
  if (threadIdx.x == 0) L.ctrl[LZ_PENDING] = 0;
  if (!rr_record_beta<T, MAXC, MemPlain>(L, i, true, lds)) return;
  for (int sub = 0; sub < 3; ++sub) {
    for (int tl = 0; tl < L.ntiles; ++tl) {
      const int t = threadIdx.x, r0 = tl * L.tile, R = min(L.n, r0 + L.tile) - r0;
      T q[MAXC];
      rr_load_q<T, MAXC, MemPlain>(L, r0 + (t < R ? t : 0), t < R, i, q);
      rr_breakdown_sub<T, MAXC, MemPlain>(L, tl, i, sub, q, lds);
    }
    __syncthreads();
  }
  if (rr_pyth<MAXC>(L, i - 1)) {  // the restart vector's norm, where phase A looks
    const double b = rr_beta<T, MemPlain>(L, lds);
    if (threadIdx.x == 0) L.h2[L.m] = b;
  }
}

// All Lanczos steps of restart r of an RR instance inside the persistent
// launch (k_lz_steps_coop), with coh_barrier between the phases and
// MemCoherent accesses for every cross-workgroup datum.  RES (every tile has
// its own workgroup): the workgroup keeps its rows of Q and the work vector
// in registers and its CSR slice in LDS for the whole restart; otherwise a
// workgroup walks tiles lw, lw + G, ... and loads them per phase.
// Basis columns a kernel instantiation holds (MAXC) for basis size m.
__host__ __device__ constexpr int lz_maxc(int m) { return m + 1 <= 17 ? 17 : m + 1 <= 33 ? 33 : 49; }
// Reduction scratch and h (rr_qdot reads MAXC entries) before the CSR slice.
__host__ __device__ constexpr int rr_head_doubles(int m) { return LZ_RED + lz_maxc(m) + 8; }
template <typename T>
static inline size_t rr_res_lds_bytes(int tile, int m, int nnz_tile) {
  size_t b = sizeof(double) * (size_t)rr_head_doubles(m);
  b += sizeof(T) * (size_t)nnz_tile + sizeof(int) * (size_t)(tile + 1 + nnz_tile);
  return (b + 255) / 256 * 256;
}

// mode0: 0 Lanczos steps (SpMV), 2 the column-by-column QR (y = Q[:, i+1]);
// the step i = m normalises the last vector (mode 3) either way.
template <typename T, int MAXC, bool RES>
__device__ __forceinline__ void rr_steps(const LzInst<T>& L, int lw, int r, unsigned long long spin_limit, double* lds,
                                         int mode0 = 0) {
  using M = MemCoherent;
  const int G = L.nblocks, t = threadIdx.x;
  int* err = L.ctrl + LZ_ERROR;
  unsigned* bar = L.bar + (size_t)LZ_BAR_STRIDE * (r < L.nbar ? r : L.nbar - 1);
  unsigned epoch = 0;
  const int k = L.ctrl[LZ_K];
  if (k >= L.m) return;  // every basis column saved: no step, Q[:, m] unchanged
  if (lw == 0 && t == 0 && r < L.nbar) {
    int sumc = 0;
    for (int i = k; i < L.m; ++i) sumc += i + 1;
    L.lstat[2 * r] = L.m - k;
    L.lstat[2 * r + 1] = sumc;
  }
  double* red = lds;
  // RES: the tile's CSR slice in LDS, values first (T), then row pointers
  // relative to the slice and columns
  T* cv = (T*)(lds + rr_head_doubles(L.m));
  int* rp = (int*)(cv + L.tile_nnz_max);
  int* ci = rp + L.tile + 1;
  T q[MAXC];
  T y = (T)0;
  const int r0 = lw * L.tile;
  const int R = RES ? min(L.n, r0 + L.tile) - r0 : 0;
  const bool valid = t < R;
  const int rr = r0 + (valid ? t : 0);
  if constexpr (RES) {
    if (mode0 == 0) {  // the QR reads no operator
      const int pbase = L.row_ptr[r0];
      const int nnzt = L.row_ptr[r0 + R] - pbase;
      for (int x = t; x <= R; x += USBS_NT) rp[x] = L.row_ptr[r0 + x] - pbase;
      for (int p = t; p < nnzt; p += USBS_NT) { ci[p] = L.col[pbase + p]; cv[p] = L.val[pbase + p]; }
    }
    rr_load_q<T, MAXC, MemPlain>(L, rr, valid, k + 1, q);
    __syncthreads();
  }
  bool ok = true;
  double bnext = -1.0;  // |w_{i-1}| from phase C of the previous step (rr_pyth)
  for (int i = k; i <= L.m && ok; ++i) {
    double beta = -1.0;
    const int mode = i == L.m ? 3 : mode0;
    RRPre<T> pf;
    if constexpr (RES) rr_prefetch<T, M>(L, lw, i, i == k, mode, rp, ci, pf);
    if (i > k) {
      if (rr_record_beta<T, MAXC, M>(L, i, lw == 0, red, &beta, bnext)) {
        for (int sub = 0; sub < 3 && ok; ++sub) {
          if constexpr (RES) {
            rr_breakdown_sub<T, MAXC, M>(L, lw, i, sub, q, red);
          } else {
            for (int tl = lw; tl < L.ntiles; tl += G) {
              const int s0 = tl * L.tile, SR = min(L.n, s0 + L.tile) - s0;
              T qt[MAXC];
              rr_load_q<T, MAXC, M>(L, s0 + (t < SR ? t : 0), t < SR, i, qt);
              rr_breakdown_sub<T, MAXC, M>(L, tl, i, sub, qt, red);
            }
          }
          ok = coh_barrier(bar, lw, G, epoch, err, spin_limit);
        }
        if (!ok) break;
        beta = rr_beta<T, M>(L, red);  // the restart vector's norm
        if constexpr (RES) rr_prefetch<T, M>(L, lw, i, false, mode, rp, ci, pf);  // w_{i-1} changed
      }
    }
    if constexpr (RES) {
      rr_phaseA<T, MAXC, M>(L, lw, i, i == k, mode, beta, pf, q, y, rp, ci, cv, red);
    } else {
      for (int tl = lw; tl < L.ntiles; tl += G) {
        const int s0 = tl * L.tile, SR = min(L.n, s0 + L.tile) - s0;
        const int sr = s0 + (t < SR ? t : 0);
        T qt[MAXC];
        rr_load_q<T, MAXC, M>(L, sr, t < SR, i + 1, qt);
        T yt = (T)0;
        rr_prefetch<T, M>(L, tl, i, i == k, mode, L.row_ptr + s0, L.col, pf);
        rr_phaseA<T, MAXC, M>(L, tl, i, i == k, mode, beta, pf, qt, yt, L.row_ptr + s0, L.col, L.val, red);
        if (mode != 3 && t < SR) M::st(L.w + (long long)(i & 1) * L.ldq + sr, yt);
      }
    }
    if (mode == 3) break;
    if (!(ok = coh_barrier(bar, lw, G, epoch, err, spin_limit))) break;
    for (int pass = 0; pass < 2 && ok; ++pass) {
      if constexpr (RES) {
        const double b = rr_phaseBC<T, MAXC, M>(L, lw, i, pass, q, y, red);
        if (pass == 1) bnext = b;
        if (pass == 1 && valid) M::st(L.w + (long long)(i & 1) * L.ldq + rr, y);
      } else {
        for (int tl = lw; tl < L.ntiles; tl += G) {
          const int s0 = tl * L.tile, SR = min(L.n, s0 + L.tile) - s0;
          const int sr = s0 + (t < SR ? t : 0);
          T* wc = L.w + (long long)(i & 1) * L.ldq;
          T qt[MAXC];
          rr_load_q<T, MAXC, M>(L, sr, t < SR, i + 1, qt);
          T yt = M::ld(wc + sr);
          const double b = rr_phaseBC<T, MAXC, M>(L, tl, i, pass, qt, yt, red);
          if (pass == 1) bnext = b;
          if (t < SR) M::st(wc + sr, yt);
        }
      }
      ok = coh_barrier(bar, lw, G, epoch, err, spin_limit);
    }
  }
}

// Destination of the per-op step launches: the stream, or (graph set) a
// chain of kernel nodes appended to a graph built node by node (no stream
// capture, so other threads' synchronous HIP calls are unaffected).
struct StepSink {
  hipStream_t s = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;
  hipError_t err = hipSuccess;

  template <typename... P, typename... A>
  void launch(void (*kern)(P...), dim3 grid, dim3 block, size_t lds, A... args) {
    if (!graph) {
      kern<<<grid, block, lds, s>>>(args...);
      return;
    }
    if (err != hipSuccess) return;
    std::tuple<P...> vals(args...);
    void* ptrs[sizeof...(P)];
    std::apply([&ptrs](auto&... v) {
      int j = 0;
      ((ptrs[j++] = (void*)&v), ...);
    }, vals);
    hipKernelNodeParams kp{};
    kp.func = (void*)kern;
    kp.gridDim = grid;
    kp.blockDim = block;
    kp.sharedMemBytes = (unsigned)lds;
    kp.kernelParams = ptrs;
    kp.extra = nullptr;
    hipGraphNode_t node = nullptr;
    err = hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &kp);
    if (err == hipSuccess) last = node;
  }
};

// Steps i = i0..m-1 of one restart (or of the QR, mode 2) as per-op launches,
// then the normalisation of Q[:, m].  Two sequences:
//  * checked: a k_lz_pre launch before every step after the first (beta and,
//    on a breakdown, the restart vector), so the steps never stop;
//  * optimistic (restarts of lanczos_perop): phase A records beta itself and
//    flags a breakdown in ctrl[LZ_PENDING], three launches a step; the host
//    resumes a flagged restart with perop_resume.
// resume (optimistic only): step i0 follows a breakdown fixed by k_lz_pre,
// so it normalises w_{i0-1} like any later step.
template <typename T, int MAXC, bool RW>
static void perop_steps_m(const LzInst<T>& L, int i0, int mode, StepSink& k, bool optimistic,
                          bool resume) {
  if (i0 >= L.m && !resume) return;  // every basis column saved: no step, Q[:, m] unchanged
  const size_t lds = tile_lds_bytes(L.m + 1);
  const dim3 g(L.ntiles), b(USBS_NT), one(1);
  const int rec = optimistic ? 1 : 0;
  // light rows in tiles of at most USBS_NT rows: the row-in-registers kernels
  const bool rr = !RW && !L.heavy && L.tile <= USBS_NT;
  auto kpre = rr ? k_lz_pre_rr<T, MAXC> : k_lz_pre<T>;
  auto kA = rr ? k_lz_A_rr<T, MAXC> : k_lz_A<T, MAXC, RW>;
  auto kBC = rr ? k_lz_BC_rr<T, MAXC> : k_lz_BC<T, MAXC, RW>;
  if (resume) k.launch(kpre, one, b, lds, L, i0);
  for (int i = i0; i < L.m; ++i) {
    if (i > i0 && !optimistic) k.launch(kpre, one, b, lds, L, i);
    // the resumed step normalises the vector k_lz_pre built (beta already recorded)
    k.launch(kA, g, b, lds, L, i, (i == i0 && !resume) ? 1 : 0, mode, (resume && i == i0) ? 0 : rec);
    k.launch(kBC, g, b, lds, L, i, 0);
    k.launch(kBC, g, b, lds, L, i, 1);
  }
  if (!optimistic) k.launch(kpre, one, b, lds, L, L.m);
  // after a resume at i0 = m the breakdown vector is already built and normalised here
  k.launch(kA, g, b, lds, L, L.m, 0, 3, (resume && i0 == L.m) ? 0 : rec);
}

template <typename T, bool RW>
static void perop_steps_rw(const LzInst<T>& L, int i0, int mode, StepSink& k, bool optimistic,
                           bool resume) {
  if (L.m + 1 <= 17) perop_steps_m<T, 17, RW>(L, i0, mode, k, optimistic, resume);
  else if (L.m + 1 <= 33) perop_steps_m<T, 33, RW>(L, i0, mode, k, optimistic, resume);
  else perop_steps_m<T, 49, RW>(L, i0, mode, k, optimistic, resume);
}

template <typename T>
static void perop_steps(const LzInst<T>& L, int i0, int mode, StepSink& k, bool optimistic = false,
                        bool resume = false) {
  if (L.tile > 256) perop_steps_rw<T, true>(L, i0, mode, k, optimistic, resume);
  else perop_steps_rw<T, false>(L, i0, mode, k, optimistic, resume);
}

template <typename T>
static void perop_steps(const LzInst<T>& L, int i0, int mode, hipStream_t s,
                        bool optimistic = false, bool resume = false) {
  StepSink k;
  k.s = s;
  perop_steps(L, i0, mode, k, optimistic, resume);
}

template <typename T, int MAXC>
static int perop_impl(const LzInst<T>& L, hipStream_t s, int* host_ctrl, LzTiming* tm,
                      LzGraphs* graphs) {
  const dim3 g(L.ntiles), b(USBS_NT);
  if (graphs) {
    static_assert(sizeof(LzInst<T>) <= sizeof(graphs->key), "LzInst key too large");
    if (graphs->key_len != (int)sizeof(L) || memcmp(graphs->key, &L, sizeof(L)) != 0) {
      lanczos_graphs_clear(graphs);
      memcpy(graphs->key, &L, sizeof(L));
      graphs->key_len = (int)sizeof(L);
    }
  }
  int k = 0;
  for (int restart = 0; restart < L.max_restarts; ++restart) {
    if (tm && hipEventRecord(tm->e0, s) != hipSuccess) return (int)hipErrorUnknown;
    if (graphs && k < L.m && k < LzGraphs::N) {
      if (!graphs->exec[k]) {
        StepSink sink;
        hipError_t e = hipGraphCreate(&sink.graph, 0);
        if (e != hipSuccess) return (int)e;
        perop_steps(L, k, 0, sink, true);
        e = sink.err;
        if (e == hipSuccess) e = hipGraphInstantiate(&graphs->exec[k], sink.graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(sink.graph);
        if (e != hipSuccess) { graphs->exec[k] = nullptr; return (int)e; }
      }
      const hipError_t e = hipGraphLaunch(graphs->exec[k], s);
      if (e != hipSuccess) return (int)e;
    } else {
      perop_steps(L, k, 0, s, true);
    }
    if (tm && hipEventRecord(tm->e1, s) != hipSuccess) return (int)hipErrorUnknown;
    k_lz_ritz<T><<<1, LZ_RITZ_NT, ritz_lds_bytes(L.m), s>>>(L);
    k_lz_rotate<T, MAXC><<<g, b, rot_lds_bytes(L.m), s>>>(L);
    hipError_t e = hipMemcpyAsync(host_ctrl, L.ctrl, sizeof(int) * LZ_NCTRL,
                                  hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return (int)e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return (int)e;
    // a breakdown stopped the restart at step p: build its vector and go on
    // from there (each pass ends at a later step, so at most m + 1 passes)
    for (int pass = 0; host_ctrl[LZ_PENDING] && pass <= L.m; ++pass) {
      perop_steps(L, host_ctrl[LZ_PENDING], 0, s, true, true);
      k_lz_ritz<T><<<1, LZ_RITZ_NT, ritz_lds_bytes(L.m), s>>>(L);
      k_lz_rotate<T, MAXC><<<g, b, rot_lds_bytes(L.m), s>>>(L);
      e = hipMemcpyAsync(host_ctrl, L.ctrl, sizeof(int) * LZ_NCTRL, hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return (int)e;
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) return (int)e;
    }
    if (host_ctrl[LZ_PENDING]) return (int)hipErrorUnknown;
    if (tm && k < L.m) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, tm->e0, tm->e1) != hipSuccess) return (int)hipErrorUnknown;
      tm->ms += ms;
      tm->launches += 1;
      tm->steps += L.m - k;
      for (int i = k; i < L.m; ++i) tm->sum_cols += i + 1;
    }
    k = host_ctrl[LZ_K];
    if (host_ctrl[LZ_DONE]) break;
  }
  return (int)hipGetLastError();
}

// Column-by-column CGS2 orthonormalisation of Q[:, 1..m] (Q[:, 0] = 0):
// step i orthonormalises column i+1 against columns 0..i with the Lanczos
// step kernels in mode 2 (a breakdown draws a counter-based random column).
template <typename T>
int orth_perop(const LzInst<T>& L, hipStream_t s) {
  if (L.m + 1 > 49) return (int)hipErrorInvalidValue;
  perop_steps(L, 0, 2, s);
  return (int)hipGetLastError();
}
template int orth_perop<double>(const LzInst<double>&, hipStream_t);
template int orth_perop<float>(const LzInst<float>&, hipStream_t);

template <typename T>
int lanczos_perop(const LzInst<T>& L, hipStream_t s, int* host_ctrl, LzTiming* tm,
                  LzGraphs* graphs) {
  if (L.m + 1 <= 17) return perop_impl<T, 17>(L, s, host_ctrl, tm, graphs);
  if (L.m + 1 <= 33) return perop_impl<T, 33>(L, s, host_ctrl, tm, graphs);
  if (L.m + 1 <= 49) return perop_impl<T, 49>(L, s, host_ctrl, tm, graphs);
  return (int)hipErrorInvalidValue;
}

// ----------------------------------------------------------------------------
// persistent form: one cooperative launch runs every Lanczos step of restart
// r for all instances; wgmap[blockIdx.x] = (instance, local workgroup index).
// C: every workgroup owns one tile and keeps it in LDS (layout lz_cache_*).

template <typename T>
static inline size_t lz_cache_bytes(int tile, int m, int nnz_tile) {
  size_t b = sizeof(double) * (size_t)(LZ_RED + m + 1 + 8);   // reduction scratch, h
  b += sizeof(T) * (size_t)tile * (size_t)(m + 2);            // Q block + work vector
  b += sizeof(T) * (size_t)nnz_tile;                          // values
  b += sizeof(int) * (size_t)(tile + 1 + nnz_tile);           // row pointers, columns
  return (b + 255) / 256 * 256;
}

// Instances with light rows in tiles of at most USBS_NT rows run rr_steps
// (C: one workgroup per tile, resident in registers).  Others: C (tiles
// cached in LDS) moves cross-workgroup data (tile partials, the work vector
// the operator gathers, Q columns) through MemCoherent accesses with
// coh_barrier, no L2 writeback or invalidate; tiles streamed from global
// memory (C false) use plain accesses and grid_barrier.  Either way, restart
// r uses counter region r of L.bar (LZ_BAR_STRIDE each).
template <typename T, bool C, int MAXC, bool RW>
__device__ __forceinline__ void tile_steps(const LzInst<T> L, int lw, int r,
                                           unsigned long long spin_limit, double* lds) {
  using M = typename std::conditional<C, MemCoherent, MemPlain>::type;
  const int G = L.nblocks, t = threadIdx.x;
  int* err = L.ctrl + LZ_ERROR;
  unsigned* bar = L.bar + (size_t)LZ_BAR_STRIDE * (r < L.nbar ? r : L.nbar - 1);
  unsigned target = 0;
  auto sync_all = [&]() {
    if constexpr (C) return coh_barrier(bar, lw, G, target, err, spin_limit);
    else return grid_barrier(bar, G, target, err, spin_limit);
  };
  const int k = L.ctrl[LZ_K];
  if (k >= L.m) return;  // every basis column saved: no step, Q[:, m] unchanged
  if (lw == 0 && t == 0 && r < L.nbar) {
    int sumc = 0;
    for (int i = k; i < L.m; ++i) sumc += i + 1;
    L.lstat[2 * r] = L.m - k;
    L.lstat[2 * r + 1] = sumc;
  }
  double* red = lds;
  TileCtx<T, C> X{};
  if constexpr (C) {
    // one tile per workgroup: load its CSR slice and Q[:, 0..k] into LDS
    X.r0 = lw * L.tile;
    X.R = min(L.n, X.r0 + L.tile) - X.r0;
    X.ldc = L.tile;
    T* qc = (T*)(lds + LZ_RED + L.m + 1 + 8);
    X.qc = qc;
    X.yc = qc + (long long)L.tile * (L.m + 1);
    T* cv = X.yc + L.tile;
    const int pbase = L.row_ptr[X.r0];
    const int nnzt = L.row_ptr[X.r0 + X.R] - pbase;
    int* rp = (int*)(cv + L.tile_nnz_max);
    int* ci = rp + L.tile + 1;
    for (int rl = t; rl <= X.R; rl += USBS_NT) rp[rl] = L.row_ptr[X.r0 + rl] - pbase;
    for (int p = t; p < nnzt; p += USBS_NT) { ci[p] = L.col[pbase + p]; cv[p] = L.val[pbase + p]; }
    for (int j = 0; j <= k; ++j)
      for (int rl = t; rl < X.R; rl += USBS_NT)
        qc[(long long)j * X.ldc + rl] = L.Q[(long long)j * L.ldq + X.r0 + rl];
    X.rp = rp; X.ci = ci; X.cv = cv;
    __syncthreads();
  }
  auto tile_of = [&](int tl) {
    if constexpr (C) return X;
    else {
      TileCtx<T, C> Y{};
      Y.r0 = tl * L.tile;
      Y.R = min(L.n, Y.r0 + L.tile) - Y.r0;
      return Y;
    }
  };
  bool ok = true;
  for (int i = k; i <= L.m && ok; ++i) {
    double beta = -1.0;  // |w_{i-1}| when known here (no breakdown)
    if (i > k) {
      // beta_{i-1} and the breakdown restart vector
      if (lz_record_beta<T, M>(L, i, lw == 0, red, &beta)) {
        beta = -1.0;
        for (int sub = 0; sub < 3 && ok; ++sub) {
          if constexpr (C) breakdown_sub_lds<T, C>(L, X, lw, i, sub, red);
          else
            for (int tl = lw; tl < L.ntiles; tl += G) breakdown_sub(L, tl, i, sub, red);
          ok = sync_all();
        }
        if (!ok) break;
      }
    }
    if (i == L.m) {
      for (int tl = lw; tl < L.ntiles; tl += G)
        phaseA<T, C, MAXC, RW, M>(L, tile_of(tl), tl, i, false, 3, red, beta);
      break;
    }
    for (int tl = lw; tl < L.ntiles; tl += G)
      phaseA<T, C, MAXC, RW, M>(L, tile_of(tl), tl, i, i == k, 0, red, beta);
    if (!(ok = sync_all())) break;
    for (int tl = lw; tl < L.ntiles; tl += G) phaseBC<T, C, MAXC, RW, M>(L, tile_of(tl), tl, i, 0, red);
    if (!(ok = sync_all())) break;
    for (int tl = lw; tl < L.ntiles; tl += G) phaseBC<T, C, MAXC, RW, M>(L, tile_of(tl), tl, i, 1, red);
    ok = sync_all();
  }
}

template <typename T, bool C, int MAXC, bool RW>
__global__ void __launch_bounds__(USBS_NT) k_lz_steps_coop(const LzInst<T>* __restrict__ insts,
                                                           const int2* __restrict__ wgmap, int r,
                                                           unsigned long long spin_limit, int mode0) {
  extern __shared__ double lds[];
  const int2 mp = wgmap[blockIdx.x];
  const LzInst<T>& L = insts[mp.x];
  if (L.ctrl[LZ_DONE] || L.ctrl[LZ_ERROR]) return;  // uniform over the instance's workgroups
  if (!RW && !L.heavy && L.tile <= USBS_NT)
    rr_steps<T, MAXC, C>(L, mp.y, r, spin_limit, lds, mode0);  // the instance read in place
  else
    tile_steps<T, C, MAXC, RW>(L, mp.y, r, spin_limit, lds);  // a copy in registers
}

// The same for launches whose instances all have light rows in tiles of at
// most 256 rows (LZ_FORM_RR): only rr_steps.  The _rr3 copy is compiled for
// three wavefronts per SIMD (LZ_FORM_RR3), so that three workgroups share a
// compute unit and a batch of solves stays co-resident with one workgroup
// per tile.
template <typename T, bool C, int MAXC>
__global__ void __launch_bounds__(USBS_NT)
k_lz_steps_coop_rr(const LzInst<T>* __restrict__ insts, const int2* __restrict__ wgmap, int r,
                   unsigned long long spin_limit, int mode0) {
  extern __shared__ double lds[];
  const int2 mp = wgmap[blockIdx.x];
  const LzInst<T>& L = insts[mp.x];
  if (L.ctrl[LZ_DONE] || L.ctrl[LZ_ERROR]) return;
  rr_steps<T, MAXC, C>(L, mp.y, r, spin_limit, lds, mode0);
}
template <typename T, bool C, int MAXC>
__global__ void __launch_bounds__(USBS_NT) __attribute__((amdgpu_waves_per_eu(3)))
k_lz_steps_coop_rr3(const LzInst<T>* __restrict__ insts, const int2* __restrict__ wgmap, int r,
                    unsigned long long spin_limit, int mode0) {
  extern __shared__ double lds[];
  const int2 mp = wgmap[blockIdx.x];
  const LzInst<T>& L = insts[mp.x];
  if (L.ctrl[LZ_DONE] || L.ctrl[LZ_ERROR]) return;
  rr_steps<T, MAXC, C>(L, mp.y, r, spin_limit, lds, mode0);
}

// The step kernel for (T, MAXC, cache, form).
template <typename T, int MAXC>
static const void* coop_kernel_m(bool cache, int form) {
  if (form == LZ_FORM_RR)
    return cache ? (const void*)k_lz_steps_coop_rr<T, true, MAXC>
                 : (const void*)k_lz_steps_coop_rr<T, false, MAXC>;
  if (form == LZ_FORM_RR3)
    return cache ? (const void*)k_lz_steps_coop_rr3<T, true, MAXC>
                 : (const void*)k_lz_steps_coop_rr3<T, false, MAXC>;
  return cache ? (const void*)k_lz_steps_coop<T, true, MAXC, false>
               : form == LZ_FORM_ROW ? (const void*)k_lz_steps_coop<T, false, MAXC, true>
                                     : (const void*)k_lz_steps_coop<T, false, MAXC, false>;
}

template <typename T, int MAXC>
static int persistent_impl(const LzInst<T>* dev_insts, int ninst, int total_wgs, int m,
                           int max_restarts, hipStream_t s, const int* dev_wgmap,
                           const int* dev_tilemap, int total_tiles, unsigned long long spin_limit,
                           size_t cache_bytes, hipEvent_t* evs, int form,
                           const LzStop* stop) {
  const bool cache = cache_bytes > 0;
  const size_t lds = cache ? cache_bytes : tile_lds_bytes(m + 1);
  const void* kern = coop_kernel_m<T, MAXC>(cache, form);
  int next_check = (stop && stop->ctrls && stop->first_check > 0) ? stop->first_check : max_restarts;
  std::vector<int> hc;
  int launched = 0;
  for (int r = 0; r < max_restarts; ++r) {
    int mode0 = 0;
    void* args[] = {(void*)&dev_insts, (void*)&dev_wgmap, (void*)&r, (void*)&spin_limit, (void*)&mode0};
    if (evs && hipEventRecord(evs[2 * r], s) != hipSuccess) return (int)hipErrorUnknown;
    hipError_t e = hipLaunchCooperativeKernel(kern, dim3(total_wgs), dim3(USBS_NT), args, lds, s);
    if (e != hipSuccess) return (int)e;
    if (evs && hipEventRecord(evs[2 * r + 1], s) != hipSuccess) return (int)hipErrorUnknown;
    k_lz_ritz_multi<T><<<ninst, LZ_RITZ_NT, ritz_lds_bytes(m), s>>>(dev_insts, r);
    k_lz_rotate_multi<T, MAXC><<<total_tiles, USBS_NT, rot_lds_bytes(m), s>>>(
        dev_insts, (const int2*)dev_tilemap, r);
    e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    launched = r + 1;
    if (launched == next_check && launched < max_restarts) {
      // later launches of finished instances return at once but still cost
      // a launch each: stop when every instance is done (or failed)
      hc.assign((size_t)LZ_NCTRL * ninst, 0);
      for (int q = 0; q < ninst; ++q) {
        e = hipMemcpyAsync(hc.data() + (size_t)q * LZ_NCTRL, stop->ctrls[q], sizeof(int) * LZ_NCTRL,
                           hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return (int)e;
      }
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) return (int)e;
      bool all = true;
      for (int q = 0; q < ninst; ++q)
        all = all && (hc[(size_t)q * LZ_NCTRL + LZ_DONE] || hc[(size_t)q * LZ_NCTRL + LZ_ERROR]);
      if (all) break;
      next_check *= 2;
    }
  }
  if (stop && stop->launched) *stop->launched = launched;
  return 0;
}

template <typename T>
size_t lanczos_cache_bytes(const LzInst<T>& L) {
  if (L.nblocks < L.ntiles) return 0;  // a workgroup would own several tiles
  if (L.tile > 256) return 0;          // row-wise tiles stream Q from memory
  // RR instances keep Q in registers: only the CSR slice is cached
  const size_t b = L.heavy ? lz_cache_bytes<T>(L.tile, L.m, L.tile_nnz_max)
                           : rr_res_lds_bytes<T>(L.tile, L.m, L.tile_nnz_max);
  return b <= LZ_CACHE_MAX_BYTES ? b : 0;
}
template size_t lanczos_cache_bytes<double>(const LzInst<double>&);
template size_t lanczos_cache_bytes<float>(const LzInst<float>&);

int lanczos_prefers_persistent(int tile, int heavy, int m, int tile_nnz_max, int value_bytes) {
  if (heavy || tile > 256) return 0;
  const size_t b = value_bytes == 4 ? rr_res_lds_bytes<float>(tile, m, tile_nnz_max)
                                    : rr_res_lds_bytes<double>(tile, m, tile_nnz_max);
  return b <= LZ_CACHE_MAX_BYTES ? 1 : 0;
}

template <typename T>
int lanczos_persistent(const LzInst<T>* dev_insts, int ninst, int total_wgs, int m,
                       int max_restarts, hipStream_t s, const int* dev_wgmap,
                       const int* dev_tilemap, int total_tiles, unsigned long long spin_limit,
                       size_t cache_bytes, hipEvent_t* evs, int form, const LzStop* stop) {
  if (m + 1 <= 17)
    return persistent_impl<T, 17>(dev_insts, ninst, total_wgs, m, max_restarts, s, dev_wgmap,
                                  dev_tilemap, total_tiles, spin_limit, cache_bytes, evs,
                                  form, stop);
  if (m + 1 <= 33)
    return persistent_impl<T, 33>(dev_insts, ninst, total_wgs, m, max_restarts, s, dev_wgmap,
                                  dev_tilemap, total_tiles, spin_limit, cache_bytes, evs,
                                  form, stop);
  if (m + 1 <= 49)
    return persistent_impl<T, 49>(dev_insts, ninst, total_wgs, m, max_restarts, s, dev_wgmap,
                                  dev_tilemap, total_tiles, spin_limit, cache_bytes, evs,
                                  form, stop);
  return (int)hipErrorInvalidValue;
}

size_t lanczos_tile_lds_bytes(int m) { return tile_lds_bytes(m + 1); }

// The step kernel persistent_impl launches for (T, m, cache, form).
template <typename T>
static const void* coop_kernel(int m, bool cache, int form) {
  if (m + 1 <= 17) return coop_kernel_m<T, 17>(cache, form);
  if (m + 1 <= 33) return coop_kernel_m<T, 33>(cache, form);
  return coop_kernel_m<T, 49>(cache, form);
}

// The column-by-column QR (orth_perop's steps) as one cooperative launch of
// the row-in-registers kernel, one workgroup per tile with its rows of Q in
// registers and only the reduction scratch in LDS: the same arithmetic as
// the per-op launches.  Restart region 0 of L.bar, zeroed by the caller.
static inline size_t orth_lds_bytes(int m) {
  return ((size_t)rr_head_doubles(m) * sizeof(double) + 255) / 256 * 256;
}
template <typename T>
int orth_persistent_fits(const LzInst<T>& L, int device) {
  if (L.m + 1 > 49 || L.tile > USBS_NT || L.heavy) return 0;
  return lanczos_max_coresident_blocks<T>(L.m, true, LZ_FORM_RR, orth_lds_bytes(L.m), device) >= L.ntiles;
}
template int orth_persistent_fits<double>(const LzInst<double>&, int);
template int orth_persistent_fits<float>(const LzInst<float>&, int);

template <typename T>
int orth_persistent(const LzInst<T>* dev_inst, int ntiles, int m, hipStream_t s, const int* dev_wgmap) {
  if (m + 1 > 49) return (int)hipErrorInvalidValue;
  const void* kern = coop_kernel<T>(m, true, LZ_FORM_RR);
  int r = 0, mode0 = 2;
  unsigned long long spin = LZ_SPIN_LIMIT;
  void* args[] = {(void*)&dev_inst, (void*)&dev_wgmap, (void*)&r, (void*)&spin, (void*)&mode0};
  const hipError_t e = hipLaunchCooperativeKernel(kern, dim3(ntiles), dim3(USBS_NT), args,
                                                  orth_lds_bytes(m), s);
  return e == hipSuccess ? 0 : (int)e;
}
template int orth_persistent<double>(const LzInst<double>*, int, int, hipStream_t, const int*);
template int orth_persistent<float>(const LzInst<float>*, int, int, hipStream_t, const int*);

template <typename T>
int lanczos_tile_form(const LzInst<T>& L) {
  if (L.tile > USBS_NT) return LZ_FORM_ROW;
  return L.heavy ? LZ_FORM_COL : LZ_FORM_RR;
}
template int lanczos_tile_form<double>(const LzInst<double>&);
template int lanczos_tile_form<float>(const LzInst<float>&);
int lanczos_launch_form(int a, int b) {
  if (a == LZ_FORM_ROW || b == LZ_FORM_ROW) return LZ_FORM_ROW;
  return (a == LZ_FORM_RR && b == LZ_FORM_RR) ? LZ_FORM_RR : LZ_FORM_COL;
}

template <typename T>
int lanczos_max_coresident_blocks(int m, bool cache, int form, size_t lds_bytes, int device) {
  int per_cu = 0;
  const void* kern = coop_kernel<T>(m, cache, form);
  if (lds_bytes > 64 * 1024 &&
      hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, USBS_NT, lds_bytes) != hipSuccess)
    return 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
  return per_cu * prop.multiProcessorCount;
}
template int lanczos_max_coresident_blocks<double>(int, bool, int, size_t, int);
template int lanczos_max_coresident_blocks<float>(int, bool, int, size_t, int);

template <typename T>
int lanczos_plan_persistent(LzInst<T>& L, int device, int want_wgs, size_t* cache_bytes) {
  int nb = L.ntiles;
  if (want_wgs > 0 && want_wgs < nb) nb = want_wgs;
  L.nblocks = nb;
  size_t cb = lanczos_cache_bytes(L);
  const int form = lanczos_tile_form(L);
  int cores = cb ? lanczos_max_coresident_blocks<T>(L.m, true, form, cb, device)
                 : lanczos_max_coresident_blocks<T>(L.m, false, form, tile_lds_bytes(L.m + 1), device);
  if (cb && cores < nb) {
    cb = 0;
    cores = lanczos_max_coresident_blocks<T>(L.m, false, form, tile_lds_bytes(L.m + 1), device);
  }
  if (cores < 1) return (int)hipErrorInvalidConfiguration;
  if (nb > cores) nb = cores;
  L.nblocks = nb;
  if (nb < L.ntiles) cb = 0;
  *cache_bytes = cb;
  return 0;
}
template int lanczos_plan_persistent<double>(LzInst<double>&, int, int, size_t*);
template int lanczos_plan_persistent<float>(LzInst<float>&, int, int, size_t*);

template int lanczos_perop<double>(const LzInst<double>&, hipStream_t, int*, LzTiming*, LzGraphs*);
template int lanczos_perop<float>(const LzInst<float>&, hipStream_t, int*, LzTiming*, LzGraphs*);

void lanczos_graphs_clear(LzGraphs* g) {
  if (!g) return;
  for (auto& e : g->exec)
    if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
  g->key_len = 0;
}
template int lanczos_persistent<double>(const LzInst<double>*, int, int, int, int, hipStream_t,
                                        const int*, const int*, int, unsigned long long, size_t,
                                        hipEvent_t*, int, const LzStop*);
template int lanczos_persistent<float>(const LzInst<float>*, int, int, int, int, hipStream_t,
                                       const int*, const int*, int, unsigned long long, size_t,
                                       hipEvent_t*, int, const LzStop*);

}  // namespace usbs
