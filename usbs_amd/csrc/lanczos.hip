// Thick-restart Lanczos for the smallest eigenpairs of the assembled operator
// (solver/lanczos.py:59-284), two execution forms over the same tile code:
//
//  * per-op: one kernel per phase of a Lanczos step, the host reads the
//    control words once per restart (lanczos_perop);
//  * persistent: one cooperative launch runs every restart of one or many
//    instances, with device-wide barriers between phases and the Ritz solve
//    done by the instance's first workgroup (lanczos_persistent).
//
// Per Lanczos step i (columns 0..i of Q are valid, the reference's Q_valid
// mask at lanczos.py:102) the tile functions perform:
//   A  w = Op q_i, part1 = Q^T w                 (SpMV fused with CGS pass 1 dot)
//   B  h1 = sum part1; w -= Q h1; part2 = Q^T w  (CGS update 1 fused with pass 2 dot)
//   C  h2 = sum part2; w -= Q h2; part3 = |w|^2  (CGS update 2 fused with the norm)
//   D  beta = sqrt(sum part3); q_{i+1} = w/beta; alpha_i = h1_i + h2_i
// with a counter-based random restart vector when beta < 1e-15
// (lanczos.py:107-117).  Q is read three times per step.
#include "lanczos.hpp"
#include "dense.hpp"

namespace usbs {

int tile_rows_for(int n) {
  int t = 256;
  while ((n + t - 1) / t > 256) t *= 2;
  return t;
}

static inline int lds_doubles_for(int m, int maxc) {
  const int red = USBS_RCH * USBS_NT + USBS_RCH * USBS_NPART + maxc + 32;
  const int ritz = 4 * m * m + 4 * m + 64;
  return red > ritz ? red : ritz;
}

// ----------------------------------------------------------------------------
// tile phases

// mode 0: w = Op q_i (SpMV); mode 1: w = counter normals (breakdown restart);
// mode 2: w = Q[:, i+1] (orthonormalise an existing column against 0..i).
// Then part1[tile] = Q[:, 0..i]^T w.
template <typename T, int MAXC>
__device__ void tile_phaseA(const LzInst<T>& L, int tl, int i, int mode, double* lds) {
  const int r0 = tl * L.tile;
  const int r1 = min(L.n, r0 + L.tile);
  const T* __restrict__ x = L.Q + (long long)i * L.ldq;
  double acc[MAXC];
#pragma unroll
  for (int j = 0; j < MAXC; ++j) acc[j] = 0.0;
  for (int r = r0 + (int)threadIdx.x; r < r1; r += USBS_NT) {
    double s = 0.0;
    if (mode == 0) {
      const int p0 = L.row_ptr[r], p1 = L.row_ptr[r + 1];
      for (int p = p0; p < p1; ++p) s += (double)L.val[p] * (double)x[L.col[p]];
    } else if (mode == 1) {
      s = counter_normal((uint64_t)(i + 1), (uint64_t)r);
    } else {
      s = (double)L.Q[r + (long long)(i + 1) * L.ldq];
    }
    const T wv = (T)s;
    L.w[r] = wv;
    const double wd = (double)wv;
#pragma unroll
    for (int j = 0; j < MAXC; ++j)
      if (j <= i) acc[j] += (double)L.Q[r + (long long)j * L.ldq] * wd;
  }
  double* red = lds;
  double* red2 = lds + USBS_RCH * USBS_NT;
  block_reduce_cols<MAXC>(acc, i + 1, red, red2, L.part1 + (long long)tl * L.pstride);
}

// h = sum of src partials; w -= Q h; then either dst = Q^T w (final = 0) or
// dst[0] = |w|^2 (final = 1).  save (if non-null, tile 0 only) keeps h.
template <typename T, int MAXC>
__device__ void tile_update(const LzInst<T>& L, int tl, int i, const double* src, double* dst,
                            double* save, int final_pass, double* lds) {
  double* red = lds;
  double* red2 = lds + USBS_RCH * USBS_NT;
  double* h = red2 + USBS_RCH * USBS_NPART;
  sum_tile_partials(src, L.ntiles, L.pstride, i + 1, h, red);
  if (save != nullptr && tl == 0)
    for (int j = threadIdx.x; j <= i; j += USBS_NT) save[j] = h[j];
  const int r0 = tl * L.tile;
  const int r1 = min(L.n, r0 + L.tile);
  double acc[MAXC];
#pragma unroll
  for (int j = 0; j < MAXC; ++j) acc[j] = 0.0;
  for (int r = r0 + (int)threadIdx.x; r < r1; r += USBS_NT) {
    double q[MAXC];
    double qh = 0.0;
#pragma unroll
    for (int j = 0; j < MAXC; ++j) {
      q[j] = (j <= i) ? (double)L.Q[r + (long long)j * L.ldq] : 0.0;
      if (j <= i) qh += q[j] * h[j];
    }
    const T wv = (T)((double)L.w[r] - qh);
    L.w[r] = wv;
    const double wd = (double)wv;
    if (final_pass) {
      acc[0] += wd * wd;
    } else {
#pragma unroll
      for (int j = 0; j < MAXC; ++j)
        if (j <= i) acc[j] += q[j] * wd;
    }
  }
  block_reduce_cols<MAXC>(acc, final_pass ? 1 : i + 1, red, red2,
                          dst + (long long)tl * L.pstride);
}

// q_{i+1} = w / beta for the tile's rows (beta from the part3 sums).
template <typename T>
__device__ void tile_normalize(const LzInst<T>& L, int tl, int i, double beta) {
  const int r0 = tl * L.tile;
  const int r1 = min(L.n, r0 + L.tile);
  T* __restrict__ q = L.Q + (long long)(i + 1) * L.ldq;
  for (int r = r0 + (int)threadIdx.x; r < r1; r += USBS_NT) q[r] = (T)((double)L.w[r] / beta);
}

template <typename T>
__device__ double lz_beta(const LzInst<T>& L, double* lds) {
  double* s = lds + USBS_RCH * USBS_NT;
  sum_tile_partials(L.part3, L.ntiles, L.pstride, 1, s, lds);
  const double b = sqrt(s[0]);
  __syncthreads();
  return b;
}

// Rotation Q[:, :k_next] = Q[:, :m] Yr; Q[:, k_next] = Q[:, m] (lanczos.py:217-221).
template <typename T, int MAXC>
__device__ void tile_rotate(const LzInst<T>& L, int tl, int knext, const double* Yr) {
  const int m = L.m;
  const int r0 = tl * L.tile;
  const int r1 = min(L.n, r0 + L.tile);
  for (int r = r0 + (int)threadIdx.x; r < r1; r += USBS_NT) {
    double q[MAXC];
#pragma unroll
    for (int x = 0; x < MAXC; ++x) q[x] = (x < m) ? (double)L.Q[r + (long long)x * L.ldq] : 0.0;
    const T qm = L.Q[r + (long long)m * L.ldq];
    for (int j = 0; j < knext && j < m; ++j) {
      double s = 0.0;
#pragma unroll
      for (int x = 0; x < MAXC; ++x)
        if (x < m) s += q[x] * Yr[x * m + j];
      L.Q[r + (long long)j * L.ldq] = (T)s;
    }
    if (knext < m) L.Q[r + (long long)knext * L.ldq] = qm;
  }
}

// ----------------------------------------------------------------------------
// Ritz step (one block): arrowhead T, eigh, convergence count, saved set,
// reorder, new alpha/beta and the rotation Yr (lanczos.py:128-168, 202-221).
__device__ void lz_ritz(int m, int nd, double tol, int max_restarts, double* alpha, double* beta,
                        double* Yr, int* ctrl, double* lds) {
  const int t = threadIdx.x, nt = blockDim.x;
  double* Tm = lds;
  double* Y = Tm + m * m;
  double* tmp = Y + m * m;
  double* w = tmp + m * m;
  double* coef = w + m;
  double* red = coef + 2 * m;
  int* sh = (int*)(red + 16);  // small int scratch
  int* part = sh + 8;          // eigh's scratch: m*m + 2m + 4 ints
  const int k = ctrl[LZ_K];
  for (int idx = t; idx < m * m; idx += nt) {
    const int i = idx / m, j = idx % m;
    double v;
    if (i == j) {
      v = alpha[i];
    } else {
      const int a = i < j ? i : j, b = i < j ? j : i;
      v = 0.0;
      if (b == k && a < k) v = beta[a];
      else if (b == a + 1 && a >= k) v = beta[a];
    }
    Tm[idx] = v;
  }
  __syncthreads();
  dense_eigh(Tm, m, w, Y, tmp, coef, part, red);
  const double bm = beta[m - 1];
  if (t == 0) {
    int nc = m;
    for (int j = 0; j < m; ++j) {
      const double res = bm * fabs(Y[(m - 1) * m + j]);
      if (!(res < tol)) { nc = j; break; }
    }
    int delta = m - nd;
    const int d2 = 2 * (m - nc) / 5;
    if (d2 < delta) delta = d2;
    const double lt = w[nc < m ? nc : m - 1];
    int kl = 0;
    bool nan_seen = false;
    double best = 0.0;
    for (int j = 0; j < m; ++j) {
      const int kr = j + delta;
      double g = 0.0;
      if (kr <= m) g = (w[j] - lt) / (w[kr < m ? kr : m - 1] - lt);
      if (g != g) {
        if (!nan_seen) { nan_seen = true; kl = j; }
      } else if (!nan_seen && (j == 0 || g > best)) {
        best = g;
        kl = j;
      }
    }
    const int kright = kl + delta;
    int cnt = 0;
    for (int j = 0; j < m; ++j)
      if (kl >= j || kright < j) part[cnt++] = j;
    const int knext = cnt;
    for (int j = 0; j < m; ++j)
      if (!(kl >= j || kright < j)) part[cnt++] = j;
    sh[0] = knext;
    sh[1] = nc;
    const int restart = ctrl[LZ_RESTART] + 1;
    ctrl[LZ_K] = knext;
    ctrl[LZ_RESTART] = restart;
    ctrl[LZ_NCONV] = nc;
    ctrl[LZ_DONE] = (nc >= nd || restart >= max_restarts) ? 1 : 0;
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
// per-op kernels (one block per tile)

template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_phaseA(LzInst<T> L, int i, int mode) {
  extern __shared__ double lds[];
  tile_phaseA<T, MAXC>(L, blockIdx.x, i, mode, lds);
}

template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_phaseB(LzInst<T> L, int i) {
  extern __shared__ double lds[];
  tile_update<T, MAXC>(L, blockIdx.x, i, L.part1, L.part2, L.h1, 0, lds);
}

template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_phaseC(LzInst<T> L, int i) {
  extern __shared__ double lds[];
  tile_update<T, MAXC>(L, blockIdx.x, i, L.part2, L.part3, L.h2, 1, lds);
}

template <typename T>
__global__ void __launch_bounds__(USBS_NT) k_lz_phaseD(LzInst<T> L, int i) {
  extern __shared__ double lds[];
  const double beta = lz_beta(L, lds);
  const bool brk = beta < 1e-15;
  if (!brk) tile_normalize(L, blockIdx.x, i, beta);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    L.alpha[i] = L.h1[i] + L.h2[i];
    L.beta[i] = beta;
    L.ctrl[LZ_BREAKDOWN] = brk ? 1 : 0;
    if (brk) L.ctrl[LZ_NBREAK] += 1;
    L.ctrl[LZ_STEPS] += 1;
  }
}

// Breakdown restart vector, one block walking every tile in order (same
// tile arithmetic as the multi-block phases, so results do not depend on
// which form ran).  No-op unless the step flagged a breakdown.
template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_breakdown(LzInst<T> L, int i) {
  extern __shared__ double lds[];
  if (L.ctrl[LZ_BREAKDOWN] == 0) return;
  for (int tl = 0; tl < L.ntiles; ++tl) tile_phaseA<T, MAXC>(L, tl, i, 1, lds);
  __syncthreads();
  for (int tl = 0; tl < L.ntiles; ++tl) tile_update<T, MAXC>(L, tl, i, L.part1, L.part2, nullptr, 0, lds);
  __syncthreads();
  for (int tl = 0; tl < L.ntiles; ++tl) tile_update<T, MAXC>(L, tl, i, L.part2, L.part3, nullptr, 1, lds);
  __syncthreads();
  const double b = lz_beta(L, lds);
  for (int tl = 0; tl < L.ntiles; ++tl) tile_normalize(L, tl, i, b);
}

#define LZ_RITZ_NT 1024
template <typename T>
__global__ void __launch_bounds__(LZ_RITZ_NT) k_lz_ritz(LzInst<T> L) {
  extern __shared__ double lds[];
  lz_ritz(L.m, L.nd, L.tol, L.max_restarts, L.alpha, L.beta, L.Yr, L.ctrl, lds);
}

// Ritz step of restart r for every instance still running (block = instance).
template <typename T>
__global__ void __launch_bounds__(LZ_RITZ_NT) k_lz_ritz_multi(const LzInst<T>* __restrict__ insts,
                                                              int r) {
  extern __shared__ double lds[];
  const LzInst<T>& L = insts[blockIdx.x];
  if (L.ctrl[LZ_DONE] || L.ctrl[LZ_RESTART] != r || L.ctrl[LZ_ERROR]) return;
  lz_ritz(L.m, L.nd, L.tol, L.max_restarts, L.alpha, L.beta, L.Yr, L.ctrl, lds);
}

template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_rotate(LzInst<T> L) {
  extern __shared__ double lds[];
  const int knext = L.ctrl[LZ_K];
  const int mm = L.m * L.m;
  for (int idx = threadIdx.x; idx < mm; idx += USBS_NT) lds[idx] = L.Yr[idx];
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
  __syncthreads();
  tile_rotate<T, MAXC>(L, mp.y, knext, lds);
}

template <typename T, int MAXC>
static int perop_impl(const LzInst<T>& L, hipStream_t s, int* host_ctrl) {
  const size_t lds = sizeof(double) * lds_doubles_for(L.m, MAXC);
  const dim3 g(L.ntiles), b(USBS_NT);
  int k = 0;
  for (int restart = 0; restart < L.max_restarts; ++restart) {
    for (int i = k; i < L.m; ++i) {
      k_lz_phaseA<T, MAXC><<<g, b, lds, s>>>(L, i, 0);
      k_lz_phaseB<T, MAXC><<<g, b, lds, s>>>(L, i);
      k_lz_phaseC<T, MAXC><<<g, b, lds, s>>>(L, i);
      k_lz_phaseD<T><<<g, b, lds, s>>>(L, i);
      k_lz_breakdown<T, MAXC><<<1, b, lds, s>>>(L, i);
    }
    k_lz_ritz<T><<<1, LZ_RITZ_NT, lds, s>>>(L);
    k_lz_rotate<T, MAXC><<<g, b, lds, s>>>(L);
    hipError_t e = hipMemcpyAsync(host_ctrl, L.ctrl, sizeof(int) * LZ_NCTRL,
                                  hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return (int)e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return (int)e;
    k = host_ctrl[LZ_K];
    if (host_ctrl[LZ_DONE]) break;
  }
  return (int)hipGetLastError();
}

// Column-by-column CGS2 orthonormalisation of Q[:, 1..m] (Q[:, 0] = 0):
// step i orthonormalises column i+1 against columns 0..i with the Lanczos
// step kernels in mode 2 (a breakdown draws a counter-based random column).
template <typename T, int MAXC>
static int orth_impl(const LzInst<T>& L, hipStream_t s) {
  const size_t lds = sizeof(double) * lds_doubles_for(L.m, MAXC);
  const dim3 g(L.ntiles), b(USBS_NT);
  for (int i = 0; i < L.m; ++i) {
    k_lz_phaseA<T, MAXC><<<g, b, lds, s>>>(L, i, 2);
    k_lz_phaseB<T, MAXC><<<g, b, lds, s>>>(L, i);
    k_lz_phaseC<T, MAXC><<<g, b, lds, s>>>(L, i);
    k_lz_phaseD<T><<<g, b, lds, s>>>(L, i);
    k_lz_breakdown<T, MAXC><<<1, b, lds, s>>>(L, i);
  }
  return (int)hipGetLastError();
}

template <typename T>
int orth_perop(const LzInst<T>& L, hipStream_t s) {
  if (L.m + 1 <= 17) return orth_impl<T, 17>(L, s);
  if (L.m + 1 <= 33) return orth_impl<T, 33>(L, s);
  return (int)hipErrorInvalidValue;
}
template int orth_perop<double>(const LzInst<double>&, hipStream_t);
template int orth_perop<float>(const LzInst<float>&, hipStream_t);

template <typename T>
int lanczos_perop(const LzInst<T>& L, hipStream_t s, int* host_ctrl) {
  if (L.m + 1 <= 17) return perop_impl<T, 17>(L, s, host_ctrl);
  if (L.m + 1 <= 33) return perop_impl<T, 33>(L, s, host_ctrl);
  if (L.m + 1 <= 49) return perop_impl<T, 49>(L, s, host_ctrl);
  return (int)hipErrorInvalidValue;
}

// ----------------------------------------------------------------------------
// persistent form: one cooperative launch runs every Lanczos step of restart
// r for all instances; wgmap[blockIdx.x] = (instance, local workgroup index).

template <typename T, int MAXC>
__global__ void __launch_bounds__(USBS_NT) k_lz_steps_coop(const LzInst<T>* __restrict__ insts,
                                                           const int2* __restrict__ wgmap, int r,
                                                           unsigned long long spin_limit) {
  extern __shared__ double lds[];
  const int2 mp = wgmap[blockIdx.x];
  const LzInst<T> L = insts[mp.x];
  if (L.ctrl[LZ_DONE] || L.ctrl[LZ_ERROR]) return;  // uniform over the instance's workgroups
  const int lw = mp.y, G = L.nblocks;
  int* err = L.ctrl + LZ_ERROR;
  unsigned* bar = L.bar + (r < L.nbar ? r : L.nbar - 1);
  unsigned target = 0;
  const int k = L.ctrl[LZ_K];
  bool ok = true;
  if (lw == 0 && threadIdx.x == 0 && r < L.nbar) {
    int sumc = 0;
    for (int i = k; i < L.m; ++i) sumc += i + 1;
    L.lstat[2 * r] = L.m - k;
    L.lstat[2 * r + 1] = sumc;
  }
  for (int i = k; i < L.m && ok; ++i) {
    for (int tl = lw; tl < L.ntiles; tl += G) tile_phaseA<T, MAXC>(L, tl, i, 0, lds);
    if (!(ok = grid_barrier(bar, G, target, err, spin_limit))) break;
    for (int tl = lw; tl < L.ntiles; tl += G)
      tile_update<T, MAXC>(L, tl, i, L.part1, L.part2, L.h1, 0, lds);
    if (!(ok = grid_barrier(bar, G, target, err, spin_limit))) break;
    for (int tl = lw; tl < L.ntiles; tl += G)
      tile_update<T, MAXC>(L, tl, i, L.part2, L.part3, L.h2, 1, lds);
    if (!(ok = grid_barrier(bar, G, target, err, spin_limit))) break;
    double beta = lz_beta(L, lds);
    const bool brk = beta < 1e-15;
    if (lw == 0 && threadIdx.x == 0) {
      L.alpha[i] = L.h1[i] + L.h2[i];
      L.beta[i] = beta;
      L.ctrl[LZ_BREAKDOWN] = brk ? 1 : 0;
      if (brk) L.ctrl[LZ_NBREAK] += 1;
      L.ctrl[LZ_STEPS] += 1;
    }
    if (brk) {
      for (int tl = lw; tl < L.ntiles; tl += G) tile_phaseA<T, MAXC>(L, tl, i, 1, lds);
      if (!(ok = grid_barrier(bar, G, target, err, spin_limit))) break;
      for (int tl = lw; tl < L.ntiles; tl += G)
        tile_update<T, MAXC>(L, tl, i, L.part1, L.part2, nullptr, 0, lds);
      if (!(ok = grid_barrier(bar, G, target, err, spin_limit))) break;
      for (int tl = lw; tl < L.ntiles; tl += G)
        tile_update<T, MAXC>(L, tl, i, L.part2, L.part3, nullptr, 1, lds);
      if (!(ok = grid_barrier(bar, G, target, err, spin_limit))) break;
      beta = lz_beta(L, lds);
    }
    for (int tl = lw; tl < L.ntiles; tl += G) tile_normalize(L, tl, i, beta);
    ok = grid_barrier(bar, G, target, err, spin_limit);
  }
}

template <typename T, int MAXC>
static int persistent_impl(const LzInst<T>* dev_insts, int ninst, int total_wgs, int m,
                           int max_restarts, hipStream_t s, const int* dev_wgmap,
                           const int* dev_tilemap, int total_tiles, unsigned long long spin_limit,
                           hipEvent_t* evs) {
  const size_t lds = sizeof(double) * lds_doubles_for(m, MAXC);
  for (int r = 0; r < max_restarts; ++r) {
    void* args[] = {(void*)&dev_insts, (void*)&dev_wgmap, (void*)&r, (void*)&spin_limit};
    if (evs && hipEventRecord(evs[2 * r], s) != hipSuccess) return (int)hipErrorUnknown;
    hipError_t e = hipLaunchCooperativeKernel((const void*)k_lz_steps_coop<T, MAXC>, dim3(total_wgs),
                                              dim3(USBS_NT), args, lds, s);
    if (e != hipSuccess) return (int)e;
    if (evs && hipEventRecord(evs[2 * r + 1], s) != hipSuccess) return (int)hipErrorUnknown;
    k_lz_ritz_multi<T><<<ninst, LZ_RITZ_NT, lds, s>>>(dev_insts, r);
    k_lz_rotate_multi<T, MAXC><<<total_tiles, USBS_NT, lds, s>>>(dev_insts, (const int2*)dev_tilemap, r);
    e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
  }
  return 0;
}

template <typename T>
int lanczos_persistent(const LzInst<T>* dev_insts, int ninst, int total_wgs, int m,
                       int max_restarts, hipStream_t s, const int* dev_wgmap,
                       const int* dev_tilemap, int total_tiles, unsigned long long spin_limit,
                       hipEvent_t* evs) {
  if (m + 1 <= 17)
    return persistent_impl<T, 17>(dev_insts, ninst, total_wgs, m, max_restarts, s, dev_wgmap,
                                  dev_tilemap, total_tiles, spin_limit, evs);
  if (m + 1 <= 33)
    return persistent_impl<T, 33>(dev_insts, ninst, total_wgs, m, max_restarts, s, dev_wgmap,
                                  dev_tilemap, total_tiles, spin_limit, evs);
  if (m + 1 <= 49)
    return persistent_impl<T, 49>(dev_insts, ninst, total_wgs, m, max_restarts, s, dev_wgmap,
                                  dev_tilemap, total_tiles, spin_limit, evs);
  return (int)hipErrorInvalidValue;
}

int lanczos_max_coresident_blocks(int m, int device) {
  int maxc = (m + 1 <= 17) ? 17 : (m + 1 <= 33 ? 33 : 49);
  const size_t lds = sizeof(double) * lds_doubles_for(m, maxc);
  int per_cu = 0;
  hipError_t e;
  if (maxc == 17)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lz_steps_coop<float, 17>, USBS_NT, lds);
  else if (maxc == 33)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lz_steps_coop<float, 33>, USBS_NT, lds);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lz_steps_coop<float, 49>, USBS_NT, lds);
  int per_cu64 = 0;
  if (maxc == 17)
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu64, k_lz_steps_coop<double, 17>, USBS_NT, lds);
  else if (maxc == 33)
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu64, k_lz_steps_coop<double, 33>, USBS_NT, lds);
  else
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu64, k_lz_steps_coop<double, 49>, USBS_NT, lds);
  if (e != hipSuccess) return 0;
  if (per_cu64 < per_cu) per_cu = per_cu64;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
  return per_cu * prop.multiProcessorCount;
}

template int lanczos_perop<double>(const LzInst<double>&, hipStream_t, int*);
template int lanczos_perop<float>(const LzInst<float>&, hipStream_t, int*);
template int lanczos_persistent<double>(const LzInst<double>*, int, int, int, int, hipStream_t,
                                        const int*, const int*, int, unsigned long long, hipEvent_t*);
template int lanczos_persistent<float>(const LzInst<float>*, int, int, int, int, hipStream_t,
                                       const int*, const int*, int, unsigned long long, hipEvent_t*);

}  // namespace usbs
