// Row- and constraint-parallel kernels around the Lanczos core: operator
// assembly, SpMM, tall-skinny products and Grams, the sketch update, the
// constraint map A(F F^T) with the dual/infeasibility reductions, the
// objective, and the max-cut rounding.
#include "solve.hpp"

#include <cmath>
#include <vector>

namespace usbs {

// vals[s] = coefC * C[s] + sum_{t in slot s} A_data[t] * (a1 z1 + a2 z2 + a3 z3)[con(t)]
// (the operator C_coef * C + A*(z) of solver/lanczos.py:86-87 in CSR form;
// A* as in solver/utils.py:19-24).
__global__ void __launch_bounds__(USBS_NT)
k_assemble(long long nnz, const double* __restrict__ cval, double coefC,
           const int* __restrict__ slot_ptr, const int* __restrict__ slot_trip,
           const double* __restrict__ tdata, const int* __restrict__ tcon,
           const double* __restrict__ z1, double a1, const double* __restrict__ z2, double a2,
           const double* __restrict__ z3, double a3, double* __restrict__ out64,
           float* __restrict__ out32) {
  const long long s = (long long)blockIdx.x * USBS_NT + threadIdx.x;
  if (s >= nnz) return;
  double v = coefC * cval[s];
  for (int p = slot_ptr[s]; p < slot_ptr[s + 1]; ++p) {
    const int t = slot_trip[p];
    const int c = tcon[t];
    double z = 0.0;
    if (z1) z += a1 * z1[c];
    if (z2) z += a2 * z2[c];
    if (z3) z += a3 * z3[c];
    v += tdata[t] * z;
  }
  if (out64) out64[s] = v;
  if (out32) out32[s] = (float)v;
}

int launch_assemble(const Problem& P, double coefC, const double* z1, double a1, const double* z2,
                    double a2, const double* z3, double a3, double* out64, float* out32,
                    hipStream_t s) {
  const long long nb = (P.nnz + USBS_NT - 1) / USBS_NT;
  k_assemble<<<dim3((unsigned)nb), dim3(USBS_NT), 0, s>>>(P.nnz, P.cval, coefC, P.slot_ptr,
                                                          P.slot_trip, P.tdata, P.tcon, z1, a1, z2,
                                                          a2, z3, a3, out64, out32);
  return (int)hipGetLastError();
}

}  // namespace usbs

namespace usbs {

// Sum over tiles of partial[tile * stride + j] by one wavefront: lane q sums
// tiles q, q + 64, ... in order, then a fixed butterfly; lane 0 holds the
// deterministic total.
__device__ __forceinline__ double wave_tile_sum(const double* __restrict__ partial, int ntiles,
                                                long long stride, long long j) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int t = lane; t < ntiles; t += 64) s += partial[(long long)t * stride + j];
  return wave_sum(s);
}

__device__ __forceinline__ double wave_tile_max(const double* __restrict__ partial, int ntiles,
                                                long long stride, long long j) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int t = lane; t < ntiles; t += 64) {
    const double v = partial[(long long)t * stride + j];
    s = v > s ? v : s;
  }
  return wave_max(s);
}

// ---------------------------------------------------------------------------
// Y (double, ldy) = M X for p <= 16 columns; M in CSR with values `val`.
template <typename TV, typename TX>
__global__ void __launch_bounds__(USBS_NT)
k_spmm(int n, const int* __restrict__ row_ptr, const int* __restrict__ col,
       const TV* __restrict__ val, const TX* __restrict__ X, long long ldx, int p,
       double* __restrict__ Y, long long ldy) {
  const int r = blockIdx.x * USBS_NT + threadIdx.x;
  if (r >= n) return;
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0;
  const int p0 = row_ptr[r], p1 = row_ptr[r + 1];
  for (int e = p0; e < p1; ++e) {
    const double v = (double)val[e];
    const long long c = col[e];
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (j < p) acc[j] += v * (double)X[c + j * ldx];
  }
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (j < p) Y[r + j * ldy] = acc[j];
}

// The same for heavy rows (32+ entries on average): a wavefront per row,
// lane l summing entries l, l+64, ..., the p sums combined by a butterfly
// over the wavefront; lane 0 writes.
template <typename TV, typename TX>
__global__ void __launch_bounds__(USBS_NT)
k_spmm_wave(int n, const int* __restrict__ row_ptr, const int* __restrict__ col,
            const TV* __restrict__ val, const TX* __restrict__ X, long long ldx, int p,
            double* __restrict__ Y, long long ldy) {
  const int r = blockIdx.x * (USBS_NT / 64) + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= n) return;  // uniform over the wavefront
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0;
  const int p1 = row_ptr[r + 1];
  for (int e = row_ptr[r] + lane; e < p1; e += 64) {
    const double v = (double)val[e];
    const long long c = col[e];
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (j < p) acc[j] += v * (double)X[c + j * ldx];
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (j < p) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
      if (lane == 0) Y[r + j * ldy] = acc[j];
    }
  }
}

template <typename TV, typename TX>
int launch_spmm(int n, const int* row_ptr, const int* col, const TV* val, const TX* X,
                long long ldx, int p, double* Y, long long ldy, int heavy, hipStream_t s) {
  if (p > 16) return (int)hipErrorInvalidValue;
  if (heavy) {
    constexpr int rows = USBS_NT / 64;
    k_spmm_wave<TV, TX><<<(n + rows - 1) / rows, USBS_NT, 0, s>>>(n, row_ptr, col, val, X, ldx, p,
                                                                  Y, ldy);
  } else {
    k_spmm<TV, TX><<<(n + USBS_NT - 1) / USBS_NT, USBS_NT, 0, s>>>(n, row_ptr, col, val, X, ldx,
                                                                   p, Y, ldy);
  }
  return (int)hipGetLastError();
}
template int launch_spmm<double, double>(int, const int*, const int*, const double*, const double*,
                                         long long, int, double*, long long, int, hipStream_t);
template int launch_spmm<double, float>(int, const int*, const int*, const double*, const float*,
                                        long long, int, double*, long long, int, hipStream_t);

// ---------------------------------------------------------------------------
// Tall-skinny Gram: partial[tile][a*q + b] = sum_{rows of tile} A[r,a] B[r,b]
// (p, q <= 80).  Rows are staged through LDS in chunks of 32.
#define GRAM_CH 32
template <typename TA, typename TB>
__global__ void __launch_bounds__(USBS_NT)
k_tsgram(int n, int tile, const TA* __restrict__ A, long long lda, int p, const TB* __restrict__ B,
         long long ldb, int q, double* __restrict__ partial) {
  extern __shared__ double lds[];
  double* As = lds;                  // GRAM_CH x p
  double* Bs = lds + GRAM_CH * p;    // GRAM_CH x q
  const int tl = blockIdx.x;
  const int r0 = tl * tile, r1 = min(n, r0 + tile);
  const int pq = p * q;
  double acc[26];
#pragma unroll
  for (int e = 0; e < 26; ++e) acc[e] = 0.0;
  for (int c0 = r0; c0 < r1; c0 += GRAM_CH) {
    const int rows = min(GRAM_CH, r1 - c0);
    for (int idx = threadIdx.x; idx < GRAM_CH * p; idx += USBS_NT) {
      const int rr = idx % GRAM_CH, a = idx / GRAM_CH;
      As[rr * p + a] = rr < rows ? (double)A[(long long)(c0 + rr) + a * lda] : 0.0;
    }
    for (int idx = threadIdx.x; idx < GRAM_CH * q; idx += USBS_NT) {
      const int rr = idx % GRAM_CH, b = idx / GRAM_CH;
      Bs[rr * q + b] = rr < rows ? (double)B[(long long)(c0 + rr) + b * ldb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 26; ++s) {
      const int e = threadIdx.x + s * USBS_NT;
      if (e < pq) {
        const int a = e / q, b = e % q;
        double v = acc[s];
        for (int rr = 0; rr < GRAM_CH; ++rr) v += As[rr * p + a] * Bs[rr * q + b];
        acc[s] = v;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < 26; ++s) {
    const int e = threadIdx.x + s * USBS_NT;
    if (e < pq) partial[(long long)tl * pq + e] = acc[s];
  }
}

// out[e] = scale * sum over tiles of partial[tile][e]; one wavefront per entry.
__global__ void __launch_bounds__(256) k_reduce_partials(const double* __restrict__ partial,
                                                          int ntiles, int count, double scale,
                                                          double* __restrict__ out) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= count) return;
  const double s = wave_tile_sum(partial, ntiles, count, e);
  if ((threadIdx.x & 63) == 0) out[e] = scale * s;
}

int launch_reduce(const double* partial, int ntiles, int count, double scale, double* out,
                  hipStream_t s) {
  k_reduce_partials<<<(count + 3) / 4, 256, 0, s>>>(partial, ntiles, count, scale, out);
  return (int)hipGetLastError();
}

template <typename TA, typename TB>
int launch_tsgram(int n, int tile, int ntiles, const TA* A, long long lda, int p, const TB* B,
                  long long ldb, int q, double* partial, double* out, hipStream_t s) {
  if (p * q > 26 * USBS_NT || p > 80 || q > 80) return (int)hipErrorInvalidValue;
  const size_t lds = sizeof(double) * GRAM_CH * (p + q);
  k_tsgram<TA, TB><<<ntiles, USBS_NT, lds, s>>>(n, tile, A, lda, p, B, ldb, q, partial);
  int rc = (int)hipGetLastError();
  if (rc) return rc;
  return launch_reduce(partial, ntiles, p * q, 1.0, out, s);
}
template int launch_tsgram<double, double>(int, int, int, const double*, long long, int,
                                           const double*, long long, int, double*, double*,
                                           hipStream_t);
template int launch_tsgram<float, double>(int, int, int, const float*, long long, int,
                                          const double*, long long, int, double*, double*,
                                          hipStream_t);
template int launch_tsgram<float, float>(int, int, int, const float*, long long, int, const float*,
                                         long long, int, double*, double*, hipStream_t);

// ---------------------------------------------------------------------------
// C[r, j] = sum_x A[r, x] B[x, j] (+ beta * C0[r, j]); B small row-major p x q
// in device memory; p, q <= 48.
template <typename TA, typename TC>
__global__ void __launch_bounds__(USBS_NT)
k_tsmm(int n, const TA* A, long long lda, int p, const double* __restrict__ B, int q,
       TC* Cout, long long ldc, double beta, const TC* C0, long long ldc0) {
  __shared__ double Bs[48 * 48];
  for (int idx = threadIdx.x; idx < p * q; idx += USBS_NT) Bs[idx] = B[idx];
  __syncthreads();
  const int r = blockIdx.x * USBS_NT + threadIdx.x;
  if (r >= n) return;
  double a[48];
#pragma unroll
  for (int x = 0; x < 48; ++x) a[x] = (x < p) ? (double)A[r + x * lda] : 0.0;
  for (int j = 0; j < q; ++j) {
    double s = 0.0;
#pragma unroll
    for (int x = 0; x < 48; ++x)
      if (x < p) s += a[x] * Bs[x * q + j];
    if (C0) s += beta * (double)C0[r + j * ldc0];
    Cout[r + j * ldc] = (TC)s;
  }
}

template <typename TA, typename TC>
int launch_tsmm(int n, const TA* A, long long lda, int p, const double* B, int q, TC* C,
                long long ldc, double beta, const TC* C0, long long ldc0, hipStream_t s) {
  if (p > 48 || q > 48) return (int)hipErrorInvalidValue;
  k_tsmm<TA, TC><<<(n + USBS_NT - 1) / USBS_NT, USBS_NT, 0, s>>>(n, A, lda, p, B, q, C, ldc, beta,
                                                                 C0, ldc0);
  return (int)hipGetLastError();
}
template int launch_tsmm<double, double>(int, const double*, long long, int, const double*, int,
                                         double*, long long, double, const double*, long long,
                                         hipStream_t);
template int launch_tsmm<float, float>(int, const float*, long long, int, const double*, int,
                                       float*, long long, double, const float*, long long,
                                       hipStream_t);
template int launch_tsmm<float, double>(int, const float*, long long, int, const double*, int,
                                        double*, long long, double, const double*, long long,
                                        hipStream_t);
template int launch_tsmm<double, float>(int, const double*, long long, int, const double*, int,
                                        float*, long long, double, const float*, long long,
                                        hipStream_t);

// ---------------------------------------------------------------------------
// Q_base Gram (solver/utils.py:53-65): G row c = svec(V^T A_c V) built in LDS
// per chunk of GRAM_CH constraints, then partial[tile] += G_chunk^T G_chunk.
// Each thread owns 4 x 4 output blocks (upper block triangle only); the
// reduce kernel mirrors the strictly lower blocks.
#define QB_BS 4
#define QB_KMAX 16  // k <= 15 (d <= 124)
template <typename TV>
__global__ void __launch_bounds__(USBS_NT)
k_qbase(int m, int mtile, int k, const TV* __restrict__ V, long long ldv,
        const int* __restrict__ con_ptr, const int* __restrict__ ct_i,
        const int* __restrict__ ct_j, const double* __restrict__ ct_d,
        double* __restrict__ partial) {
  extern __shared__ double lds[];
  const int d = k * (k + 1) / 2;
  const int dp = (d + QB_BS - 1) / QB_BS * QB_BS;  // padded width
  double* G = lds;  // GRAM_CH x dp
  double* rd = G + GRAM_CH * dp;      // chunk row: data of its first triple
  double* Vi = rd + GRAM_CH;          // chunk row: V row i of its first triple (QB_KMAX)
  double* Vj = Vi + GRAM_CH * QB_KMAX;  // and V row j
  int* pb = (int*)(Vj + GRAM_CH * QB_KMAX);
  int* pc = pb + dp;
  int* rcnt = pc + dp;                // chunk row: triple count
  int* ri = rcnt + GRAM_CH;
  int* rj = ri + GRAM_CH;
  const int nbk = dp / QB_BS;
  const int nblk = nbk * (nbk + 1) / 2;  // upper block triangle
  const int tl = blockIdx.x;
  const int c0t = tl * mtile, c1t = min(m, c0t + mtile);
  for (int a = threadIdx.x; a < dp; a += USBS_NT) {
    int bb = 0, cc = 0;
    if (a < d) {
      bb = (int)((sqrt(8.0 * a + 1.0) - 1.0) * 0.5);
      while ((bb + 1) * (bb + 2) / 2 <= a) ++bb;
      while (bb * (bb + 1) / 2 > a) --bb;
      cc = a - bb * (bb + 1) / 2;
    }
    pb[a] = bb;
    pc[a] = cc;
  }
  // this thread's blocks: bi <= bj in the upper triangle, indices s*NT + t
  constexpr int MAXB = 2;  // nblk <= 2*256 covers d <= 124
  int bi[MAXB], bj[MAXB];
#pragma unroll
  for (int s = 0; s < MAXB; ++s) {
    int id = threadIdx.x + s * USBS_NT, r = 0;
    bi[s] = -1; bj[s] = -1;
    if (id < nblk) {
      while (id >= nbk - r) { id -= nbk - r; ++r; }
      bi[s] = r;
      bj[s] = r + id;
    }
  }
  double acc[MAXB][QB_BS][QB_BS];
#pragma unroll
  for (int s = 0; s < MAXB; ++s)
#pragma unroll
    for (int x = 0; x < QB_BS; ++x)
#pragma unroll
      for (int y = 0; y < QB_BS; ++y) acc[s][x][y] = 0.0;
  const double u = 0.70710678118654752440;
  __syncthreads();
  for (int c0 = c0t; c0 < c1t; c0 += GRAM_CH) {
    const int rows = min(GRAM_CH, c1t - c0);
    // A constraint with one triple (max-cut's diagonal, most QAP rows) reads
    // its two V rows once, into LDS, with every load of the chunk in flight;
    // others loop over their triples.  The arithmetic is the same either way.
    for (int rr = threadIdx.x; rr < GRAM_CH; rr += USBS_NT) {
      int cv = 0, iv = 0, jv = 0;
      double dv = 0.0;
      if (rr < rows) {
        const int c = c0 + rr;
        const int e0 = con_ptr[c];
        cv = con_ptr[c + 1] - e0;
        if (cv >= 1) { iv = ct_i[e0]; jv = ct_j[e0]; dv = ct_d[e0]; }
      }
      rcnt[rr] = cv; ri[rr] = iv; rj[rr] = jv; rd[rr] = dv;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < GRAM_CH * k; idx += USBS_NT) {
      const int rr = idx / k, x = idx - (idx / k) * k;
      double vi = 0.0, vj = 0.0;
      if (rcnt[rr] == 1) {
        vi = (double)V[(long long)ri[rr] + x * ldv];
        vj = (double)V[(long long)rj[rr] + x * ldv];
      }
      Vi[rr * QB_KMAX + x] = vi;
      Vj[rr * QB_KMAX + x] = vj;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < GRAM_CH * dp; idx += USBS_NT) {
      const int rr = idx / dp, a = idx - (idx / dp) * dp;
      double v = 0.0;
      if (rr < rows && a < d && rcnt[rr] == 1) {
        const int bb = pb[a], cc = pc[a];
        const double td = rd[rr];
        const double vib = Vi[rr * QB_KMAX + bb], vjc = Vj[rr * QB_KMAX + cc];
        if (bb == cc) {
          v += td * (vib * vjc);
        } else {
          const double vic = Vi[rr * QB_KMAX + cc], vjb = Vj[rr * QB_KMAX + bb];
          v += td * (vib * vjc) * u + td * (vic * vjb) * u;
        }
      } else if (rr < rows && a < d) {
        const int bb = pb[a], cc = pc[a];
        const int c = c0 + rr;
        for (int e = con_ptr[c]; e < con_ptr[c + 1]; ++e) {
          const long long i = ct_i[e], j = ct_j[e];
          const double td = ct_d[e];
          const double vib = (double)V[i + bb * ldv], vjc = (double)V[j + cc * ldv];
          if (bb == cc) {
            v += td * (vib * vjc);
          } else {
            const double vic = (double)V[i + cc * ldv], vjb = (double)V[j + bb * ldv];
            v += td * (vib * vjc) * u + td * (vic * vjb) * u;
          }
        }
      }
      G[idx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < MAXB; ++s) {
      if (bi[s] >= 0) {
        const int a0 = bi[s] * QB_BS, b0 = bj[s] * QB_BS;
        for (int rr = 0; rr < GRAM_CH; ++rr) {
          double ga[QB_BS], gb[QB_BS];
#pragma unroll
          for (int x = 0; x < QB_BS; ++x) {
            ga[x] = G[rr * dp + a0 + x];
            gb[x] = G[rr * dp + b0 + x];
          }
#pragma unroll
          for (int x = 0; x < QB_BS; ++x)
#pragma unroll
            for (int y = 0; y < QB_BS; ++y) acc[s][x][y] += ga[x] * gb[y];
        }
      }
    }
    __syncthreads();
  }
  double* out = partial + (long long)tl * d * d;
#pragma unroll
  for (int s = 0; s < MAXB; ++s) {
    if (bi[s] >= 0) {
      const int a0 = bi[s] * QB_BS, b0 = bj[s] * QB_BS;
#pragma unroll
      for (int x = 0; x < QB_BS; ++x)
#pragma unroll
        for (int y = 0; y < QB_BS; ++y) {
          const int a = a0 + x, b = b0 + y;
          if (a < d && b < d && a <= b) out[a * d + b] = acc[s][x][y];
        }
    }
  }
}

__global__ void __launch_bounds__(256) k_reduce_sym(const double* __restrict__ partial, int ntiles,
                                                     int d, double* __restrict__ out) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= d * d) return;
  const int a = e / d, b = e % d;
  const int src = a <= b ? a * d + b : b * d + a;
  const double s = wave_tile_sum(partial, ntiles, (long long)d * d, src);
  if ((threadIdx.x & 63) == 0) out[e] = s;
}

int qbase_tile_rows(int m, int d) {
  int t = 256;
  while ((long long)((m + t - 1) / t) * d * d > (1 << 20) && (m + t - 1) / t > 256) t *= 2;
  return t;
}

template <typename TV>
int launch_qbase(const Problem& P, int k, const TV* V, long long ldv, double* partial, double* out,
                 hipStream_t s) {
  const int d = k * (k + 1) / 2;
  const int dp = (d + QB_BS - 1) / QB_BS * QB_BS;
  const int nbk = dp / QB_BS;
  if (nbk * (nbk + 1) / 2 > 2 * USBS_NT || k >= QB_KMAX) return (int)hipErrorInvalidValue;
  const size_t lds = sizeof(double) * (GRAM_CH * dp + GRAM_CH + 2 * GRAM_CH * QB_KMAX) +
                     sizeof(int) * (2 * dp + 3 * GRAM_CH);
  const int qt = qbase_tile_rows(P.m, d), qnt = (P.m + qt - 1) / qt;
  k_qbase<TV><<<qnt, USBS_NT, lds, s>>>(P.m, qt, k, V, ldv, P.con_ptr, P.ct_i, P.ct_j, P.ct_d,
                                        partial);
  int rc = (int)hipGetLastError();
  if (rc) return rc;
  k_reduce_sym<<<(d * d + 3) / 4, 256, 0, s>>>(partial, qnt, d, out);
  return (int)hipGetLastError();
}
template int launch_qbase<double>(const Problem&, int, const double*, long long, double*, double*,
                                  hipStream_t);
template int launch_qbase<float>(const Problem&, int, const float*, long long, double*, double*,
                                 hipStream_t);

// ---------------------------------------------------------------------------
// Per-constraint update after the bundle factor F (solver/usbs.py:571-583,
// 637-639, 648-652):
//   z     = eta z_bar + A(F F^T)            (all k columns)
//   zbn   = eta z_bar + A(F_c F_c^T)        (first kc columns)
//   ycand = y - (b - z - ups) / rho
// and per-tile partials: [b.ycand, zbar.ycand, |infeas/SA|^2, max |infeas|/SA].
template <typename TF>
__global__ void __launch_bounds__(USBS_NT)
k_constraint_update(int m, int mtile, int k, int kc, const TF* __restrict__ F, long long ldf,
                    const int* __restrict__ con_ptr, const int* __restrict__ ct_i,
                    const int* __restrict__ ct_j, const double* __restrict__ ct_d,
                    const double* __restrict__ eta_ptr,
                    const double* __restrict__ zbar,
                    const double* __restrict__ y, const double* __restrict__ b,
                    const double* __restrict__ ups, const double* __restrict__ mask,
                    const double* __restrict__ scale_a, double rho, double* __restrict__ z,
                    double* __restrict__ zbn, double* __restrict__ ycand,
                    double* __restrict__ partial) {
  __shared__ double red[8];
  const double eta = *eta_ptr;
  const int tl = blockIdx.x;
  const int c0 = tl * mtile, c1 = min(m, c0 + mtile);
  double s_by = 0.0, s_zy = 0.0, s_inf2 = 0.0, s_max = 0.0;
  for (int c = c0 + (int)threadIdx.x; c < c1; c += USBS_NT) {
    double sall = 0.0, scur = 0.0;
    for (int e = con_ptr[c]; e < con_ptr[c + 1]; ++e) {
      const long long i = ct_i[e], j = ct_j[e];
      double pa = 0.0, pc = 0.0;
      for (int x = 0; x < k; ++x) {
        const double pr = (double)F[i + x * ldf] * (double)F[j + x * ldf];
        pa += pr;
        if (x < kc) pc += pr;
      }
      sall += ct_d[e] * pa;
      scur += ct_d[e] * pc;
    }
    const double zb = zbar[c];
    const double zn = eta * zb + sall;
    z[c] = zn;
    zbn[c] = eta * zb + scur;
    const double u = ups ? ups[c] : 0.0;
    const double yc = y[c] - (1.0 / rho) * (b[c] - zn - u);
    ycand[c] = yc;
    s_by += b[c] * yc;
    s_zy += zb * yc;
    const double d1 = zn - b[c], d2 = b[c] - zn;
    const double inf = (d1 > 0.0 ? d1 : 0.0) + (1.0 - mask[c]) * (d2 > 0.0 ? d2 : 0.0);
    const double q = inf / scale_a[c];
    s_inf2 += q * q;
    const double aq = fabs(inf) / scale_a[c];
    s_max = aq > s_max ? aq : s_max;
  }
  const double a0 = block_sum1(s_by, red);
  const double a1 = block_sum1(s_zy, red);
  const double a2 = block_sum1(s_inf2, red);
  double mx = wave_max(s_max);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < USBS_NT / 64; ++w) v = red[w] > v ? red[w] : v;
    partial[tl * 4 + 0] = a0;
    partial[tl * 4 + 1] = a1;
    partial[tl * 4 + 2] = a2;
    partial[tl * 4 + 3] = v;
  }
}

// Sum the first three partial columns over tiles, max the fourth; write to
// out[0..3].  Wavefront w handles column w.
__global__ void __launch_bounds__(256) k_reduce_cu(const double* __restrict__ partial, int ntiles,
                                                    double* __restrict__ out) {
  const int w = threadIdx.x >> 6;
  const double v = (w < 3) ? wave_tile_sum(partial, ntiles, 4, w) : wave_tile_max(partial, ntiles, 4, 3);
  if ((threadIdx.x & 63) == 0) out[w] = v;
}

template <typename TF>
int launch_constraint_update(const Problem& P, int k, int kc, const TF* F, long long ldf,
                             const double* eta, const double* zbar, const double* y, const double* ups,
                             double rho, double* z, double* zbn, double* ycand, double* partial,
                             double* out4, hipStream_t s) {
  k_constraint_update<TF><<<P.mntiles, USBS_NT, 0, s>>>(
      P.m, P.mtile, k, kc, F, ldf, P.con_ptr, P.ct_i, P.ct_j, P.ct_d, eta, zbar, y, P.b,
      ups, P.mask, P.scale_a, rho, z, zbn, ycand, partial);
  int rc = (int)hipGetLastError();
  if (rc) return rc;
  k_reduce_cu<<<1, 256, 0, s>>>(partial, P.mntiles, out4);
  return (int)hipGetLastError();
}
template int launch_constraint_update<double>(const Problem&, int, int, const double*, long long,
                                              const double*, const double*, const double*, const double*,
                                              double, double*, double*, double*, double*, double*,
                                              hipStream_t);
template int launch_constraint_update<float>(const Problem&, int, int, const float*, long long,
                                             const double*, const double*, const double*, const double*,
                                             double, double*, double*, double*, double*, double*,
                                             hipStream_t);
}  // namespace usbs

namespace usbs {

// ---------------------------------------------------------------------------
// Sketch update (solver/usbs.py:580, 636):
//   P    = eta Pbar + F   (F^T Omega)
//   Pbar = eta Pbar + F_c (F_c^T Omega)       (F_c = first kc columns)
// FO = F^T Omega (k x r, row-major, FP64).  diag-only mode (r == 0) writes the
// row sums of F*F instead (usbs.py:574, 630) into the n-vectors P / Pbar.
template <typename T>
__global__ void __launch_bounds__(USBS_NT)
k_sketch_update(int n, int k, int kc, int r, const T* __restrict__ F, long long ldf,
                const double* __restrict__ FO, const double* __restrict__ eta_ptr,
                T* __restrict__ Pbar, T* __restrict__ P, long long ldp) {
  __shared__ double FOs[16 * 48];
  for (int idx = threadIdx.x; idx < k * r; idx += USBS_NT) FOs[idx] = FO[idx];
  __syncthreads();
  const int i = blockIdx.x * USBS_NT + threadIdx.x;
  if (i >= n) return;
  const double eta = *eta_ptr;
  double f[16];
#pragma unroll
  for (int x = 0; x < 16; ++x) f[x] = (x < k) ? (double)F[i + x * ldf] : 0.0;
  if (r == 0) {
    double sa = 0.0, sc = 0.0;
#pragma unroll
    for (int x = 0; x < 16; ++x) {
      sa += f[x] * f[x];
      if (x < kc) sc += f[x] * f[x];
    }
    const double pb = (double)Pbar[i];
    P[i] = (T)(eta * pb + sa);
    Pbar[i] = (T)(eta * pb + sc);
    return;
  }
  for (int j = 0; j < r; ++j) {
    double sa = 0.0, sc = 0.0;
#pragma unroll
    for (int x = 0; x < 16; ++x) {
      if (x < k) {
        const double v = f[x] * FOs[x * r + j];
        sa += v;
        if (x < kc) sc += v;
      }
    }
    const double pb = (double)Pbar[i + j * ldp];
    P[i + j * ldp] = (T)(eta * pb + sa);
    Pbar[i + j * ldp] = (T)(eta * pb + sc);
  }
}

template <typename T>
int launch_sketch_update(int n, int k, int kc, int r, const T* F, long long ldf, const double* FO,
                         const double* eta, T* Pbar, T* P, long long ldp, hipStream_t s) {
  if (k > 16 || r > 48) return (int)hipErrorInvalidValue;
  k_sketch_update<T><<<(n + USBS_NT - 1) / USBS_NT, USBS_NT, 0, s>>>(n, k, kc, r, F, ldf, FO, eta,
                                                                     Pbar, P, ldp);
  return (int)hipGetLastError();
}
template int launch_sketch_update<double>(int, int, int, int, const double*, long long,
                                          const double*, const double*, double*, double*,
                                          long long, hipStream_t);
template int launch_sketch_update<float>(int, int, int, int, const float*, long long, const double*,
                                         const double*, float*, float*, long long, hipStream_t);

// ---------------------------------------------------------------------------
// Objective (solver/usbs.py:585, 644): per column x of F, f_x^T C f_x, as
// per-tile partials (C in the merged pattern, FP64 values).
template <typename T>
__global__ void __launch_bounds__(USBS_NT)
k_objective(int n, int tile, int k, const int* __restrict__ row_ptr, const int* __restrict__ col,
            const double* __restrict__ cval, const T* __restrict__ F, long long ldf,
            double* __restrict__ partial) {
  extern __shared__ double lds[];
  const int tl = blockIdx.x;
  const int r0 = tl * tile, r1 = min(n, r0 + tile);
  double acc[16];
#pragma unroll
  for (int x = 0; x < 16; ++x) acc[x] = 0.0;
  for (int r = r0 + (int)threadIdx.x; r < r1; r += USBS_NT) {
    double cf[16];
#pragma unroll
    for (int x = 0; x < 16; ++x) cf[x] = 0.0;
    for (int e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
      const double v = cval[e];
      const long long c = col[e];
#pragma unroll
      for (int x = 0; x < 16; ++x)
        if (x < k) cf[x] += v * (double)F[c + x * ldf];
    }
#pragma unroll
    for (int x = 0; x < 16; ++x)
      if (x < k) acc[x] += (double)F[r + x * ldf] * cf[x];
  }
  block_reduce_cols<16>(acc, k, lds, lds + USBS_RCH * USBS_NT, partial + (long long)tl * 16);
}

// Heavy rows: a wavefront per row of a Lanczos row tile (lanes over the row's
// entries, butterfly), lane 0 accumulating f_x[r] (C f_x)[r].
template <typename T>
__global__ void __launch_bounds__(USBS_NT)
k_objective_wave(int n, int tile, int k, const int* __restrict__ row_ptr,
                 const int* __restrict__ col, const double* __restrict__ cval,
                 const T* __restrict__ F, long long ldf, double* __restrict__ partial) {
  extern __shared__ double lds[];
  const int tl = blockIdx.x;
  const int r0 = tl * tile, r1 = min(n, r0 + tile);
  const int lane = threadIdx.x & 63;
  double acc[16];
#pragma unroll
  for (int x = 0; x < 16; ++x) acc[x] = 0.0;
  for (int r = r0 + (int)(threadIdx.x >> 6); r < r1; r += USBS_NT / 64) {
    double cf[16];
#pragma unroll
    for (int x = 0; x < 16; ++x) cf[x] = 0.0;
    const int p1 = row_ptr[r + 1];
    for (int e = row_ptr[r] + lane; e < p1; e += 64) {
      const double v = cval[e];
      const long long c = col[e];
#pragma unroll
      for (int x = 0; x < 16; ++x)
        if (x < k) cf[x] += v * (double)F[c + x * ldf];
    }
#pragma unroll
    for (int x = 0; x < 16; ++x) {
      if (x < k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) cf[x] += __shfl_xor(cf[x], off, 64);
        if (lane == 0) acc[x] += (double)F[r + x * ldf] * cf[x];
      }
    }
  }
  block_reduce_cols<16>(acc, k, lds, lds + USBS_RCH * USBS_NT, partial + (long long)tl * 16);
}

// out[0] = sum over all k columns, out[1] = over the first kc columns.
__global__ void __launch_bounds__(1024) k_reduce_objective(const double* __restrict__ partial,
                                                           int ntiles, int k, int kc,
                                                           double* __restrict__ out) {
  __shared__ double col[16];
  const int w = threadIdx.x >> 6;
  if (w < k) {
    const double v = wave_tile_sum(partial, ntiles, 16, w);
    if ((threadIdx.x & 63) == 0) col[w] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    for (int x = 0; x < k; ++x) {
      a += col[x];
      if (x < kc) c += col[x];
    }
    out[0] = a;
    out[1] = c;
  }
}

template <typename T>
int launch_objective(const Problem& P, int k, int kc, const T* F, long long ldf, double* partial,
                     double* out2, hipStream_t s) {
  if (k > 16) return (int)hipErrorInvalidValue;
  const size_t lds = sizeof(double) * (USBS_RCH * USBS_NT + USBS_RCH * USBS_NPART);
  if (P.lz_heavy)
    k_objective_wave<T><<<P.lz_ntiles, USBS_NT, lds, s>>>(P.n, P.lz_tile, k, P.row_ptr, P.col,
                                                          P.cval, F, ldf, partial);
  else
    k_objective<T><<<P.ntiles, USBS_NT, lds, s>>>(P.n, P.tile, k, P.row_ptr, P.col, P.cval, F,
                                                  ldf, partial);
  int rc = (int)hipGetLastError();
  if (rc) return rc;
  k_reduce_objective<<<1, 1024, 0, s>>>(partial, objective_ntiles(P), k, kc, out2);
  return (int)hipGetLastError();
}
template int launch_objective<double>(const Problem&, int, int, const double*, long long, double*,
                                      double*, hipStream_t);
template int launch_objective<float>(const Problem&, int, int, const float*, long long, double*,
                                     double*, hipStream_t);

// ---------------------------------------------------------------------------
// Three dual dots over the constraints: [z.z, z.y, z.(b - ups)] with per-tile
// partials then an in-order sum (the IPM constants of usbs.py:57, 61).
__global__ void __launch_bounds__(USBS_NT)
k_dots3(int m, int mtile, const double* __restrict__ z, const double* __restrict__ y,
        const double* __restrict__ b, const double* __restrict__ ups, double* __restrict__ partial) {
  __shared__ double red[8];
  const int tl = blockIdx.x;
  const int c0 = tl * mtile, c1 = min(m, c0 + mtile);
  double a = 0.0, bb = 0.0, c = 0.0;
  for (int i = c0 + (int)threadIdx.x; i < c1; i += USBS_NT) {
    const double zi = z[i];
    a += zi * zi;
    bb += zi * y[i];
    c += zi * (b[i] - (ups ? ups[i] : 0.0));
  }
  a = block_sum1(a, red);
  bb = block_sum1(bb, red);
  c = block_sum1(c, red);
  if (threadIdx.x == 0) {
    partial[tl * 4] = a;
    partial[tl * 4 + 1] = bb;
    partial[tl * 4 + 2] = c;
    partial[tl * 4 + 3] = 0.0;
  }
}

int launch_dots3(const Problem& P, const double* z, const double* y, const double* ups,
                 double* partial, double* out4, hipStream_t s) {
  k_dots3<<<P.mntiles, USBS_NT, 0, s>>>(P.m, P.mtile, z, y, P.b, ups, partial);
  int rc = (int)hipGetLastError();
  if (rc) return rc;
  k_reduce_cu<<<1, 256, 0, s>>>(partial, P.mntiles, out4);
  return (int)hipGetLastError();
}

// y <- ycand where the device flag is set (the descent step, usbs.py:623-624).
__global__ void k_select(int m, const double* __restrict__ flag, double* __restrict__ y,
                         const double* __restrict__ ycand) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && *flag != 0.0) y[i] = ycand[i];
}

int launch_select(int m, const double* flag, double* y, const double* ycand, hipStream_t s) {
  k_select<<<(m + 255) / 256, 256, 0, s>>>(m, flag, y, ycand);
  return (int)hipGetLastError();
}

// Upsilon step for inequality rows (usbs.py:412-415):
//   z = eta zbar + A(F F^T);  ups' = mask * max(b - z - rho y, 0);
//   partial: max |ups - ups'|.
template <typename TF>
__global__ void __launch_bounds__(USBS_NT)
k_upsilon(int m, int mtile, int k, const TF* __restrict__ F, long long ldf,
          const int* __restrict__ con_ptr, const int* __restrict__ ct_i,
          const int* __restrict__ ct_j, const double* __restrict__ ct_d,
          const double* __restrict__ eta_ptr, const double* __restrict__ zbar,
          const double* __restrict__ y, const double* __restrict__ b,
          const double* __restrict__ mask, double rho, double* __restrict__ ups,
          double* __restrict__ partial) {
  __shared__ double red[8];
  const double eta = *eta_ptr;
  const int tl = blockIdx.x;
  const int c0 = tl * mtile, c1 = min(m, c0 + mtile);
  double gap = 0.0;
  for (int c = c0 + (int)threadIdx.x; c < c1; c += USBS_NT) {
    double sall = 0.0;
    for (int e = con_ptr[c]; e < con_ptr[c + 1]; ++e) {
      const long long i = ct_i[e], j = ct_j[e];
      double pa = 0.0;
      for (int x = 0; x < k; ++x) pa += (double)F[i + x * ldf] * (double)F[j + x * ldf];
      sall += ct_d[e] * pa;
    }
    const double zn = eta * zbar[c] + sall;
    double u = b[c] - zn - rho * y[c];
    u = mask[c] * (u > 0.0 ? u : 0.0);
    const double g = fabs(ups[c] - u);
    gap = g > gap ? g : gap;
    ups[c] = u;
  }
  gap = wave_max(gap);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = gap;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < USBS_NT / 64; ++w) v = red[w] > v ? red[w] : v;
    partial[tl] = v;
  }
}

__global__ void __launch_bounds__(64) k_reduce_max(const double* __restrict__ partial, int ntiles,
                                                    double* __restrict__ out) {
  const double v = wave_tile_max(partial, ntiles, 1, 0);
  if (threadIdx.x == 0) *out = v;
}

template <typename TF>
int launch_upsilon(const Problem& P, int k, const TF* F, long long ldf, const double* eta,
                   const double* zbar, const double* y, double rho, double* ups, double* partial,
                   double* out, hipStream_t s) {
  k_upsilon<TF><<<P.mntiles, USBS_NT, 0, s>>>(P.m, P.mtile, k, F, ldf, P.con_ptr, P.ct_i,
                                              P.ct_j, P.ct_d, eta, zbar, y, P.b, P.mask, rho, ups,
                                              partial);
  int rc = (int)hipGetLastError();
  if (rc) return rc;
  k_reduce_max<<<1, 64, 0, s>>>(partial, P.mntiles, out);
  return (int)hipGetLastError();
}
template int launch_upsilon<double>(const Problem&, int, const double*, long long, const double*,
                                    const double*, const double*, double, double*, double*,
                                    double*, hipStream_t);
template int launch_upsilon<float>(const Problem&, int, const float*, long long, const double*,
                                   const double*, const double*, double, double*, double*, double*,
                                   hipStream_t);

// ---------------------------------------------------------------------------
// Max-cut rounding (utils/maxcut_helpers.py:292-302).
// Signs: s[i, j] = +1 if E[i, j] > 0 else -1 where E = (P + rho Omega) G
// (G from the Nystrom core; a NaN core gives every sign -1, as XLA's NaN
// Cholesky does).  Stored as int8, column-major n x r.
template <typename T>
__global__ void __launch_bounds__(USBS_NT)
k_round_signs(int n, int r, const T* __restrict__ P, const T* __restrict__ Om, long long ld,
              const double* __restrict__ G, const double* __restrict__ rho_ptr,
              signed char* __restrict__ sg) {
  __shared__ double Gs[48 * 48];
  for (int idx = threadIdx.x; idx < r * r; idx += USBS_NT) Gs[idx] = G[idx];
  __syncthreads();
  const int i = blockIdx.x * USBS_NT + threadIdx.x;
  if (i >= n) return;
  const double rho = *rho_ptr;
  double pr[48];
#pragma unroll
  for (int x = 0; x < 48; ++x)
    pr[x] = (x < r) ? (double)P[i + x * ld] + rho * (double)Om[i + x * ld] : 0.0;
  for (int j = 0; j < r; ++j) {
    double e = 0.0;
#pragma unroll
    for (int x = 0; x < 48; ++x)
      if (x < r) e += pr[x] * Gs[x * r + j];
    sg[i + (long long)j * n] = (e > 0.0) ? 1 : -1;
  }
}

// Cut values: per tile and candidate j, the integer sum over stored entries
// of 4*C_ik s_i s_k (cut_integer problems) or the FP64 sum of C_ik s_i s_k.
// value_j = sgn_j^T C sgn_j (= sum of cut edge weights for C = L/4).
__global__ void __launch_bounds__(USBS_NT)
k_cut_values(int n, int tile, int r, const int* __restrict__ row_ptr, const int* __restrict__ col,
             const int* __restrict__ cw4, const double* __restrict__ cval, int use_int,
             const signed char* __restrict__ sg, long long* __restrict__ ipart,
             double* __restrict__ dpart) {
  __shared__ long long ired[4 * 48];
  __shared__ double dred[4 * 48];
  const int tl = blockIdx.x;
  const int r0 = tl * tile, r1 = min(n, r0 + tile);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // eight candidates per pass over the tile's CSR (each candidate's sum keeps
  // its order: rows, then entries, in storage order)
  constexpr int CG = 8;
  for (int j0 = 0; j0 < r; j0 += CG) {
    const int jn = min(CG, r - j0);
    long long ia[CG];
    double da[CG];
#pragma unroll
    for (int u = 0; u < CG; ++u) { ia[u] = 0; da[u] = 0.0; }
    for (int i = r0 + (int)threadIdx.x; i < r1; i += USBS_NT) {
      int si[CG];
#pragma unroll
      for (int u = 0; u < CG; ++u) si[u] = u < jn ? (int)sg[(long long)(j0 + u) * n + i] : 0;
      for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
        const int c = col[e];
        const int w = use_int ? cw4[e] : 0;
        const double cvv = use_int ? 0.0 : cval[e];
#pragma unroll
        for (int u = 0; u < CG; ++u) {
          if (u < jn) {
            const int sk = sg[(long long)(j0 + u) * n + c];
            if (use_int) ia[u] += (long long)(w * si[u] * sk);
            else da[u] += cvv * (double)(si[u] * sk);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < CG; ++u) {
      long long a = ia[u];
      double b = da[u];
      for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
      }
      if (lane == 0 && u < jn) { ired[wid * 48 + j0 + u] = a; dred[wid * 48 + j0 + u] = b; }
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < r; j += USBS_NT) {
    long long ia = 0;
    double da = 0.0;
    for (int w = 0; w < USBS_NT / 64; ++w) { ia += ired[w * 48 + j]; da += dred[w * 48 + j]; }
    ipart[(long long)tl * r + j] = ia;
    dpart[(long long)tl * r + j] = da;
  }
}

// out[j] = cut value of candidate j, out[r] = max_j out[j].  Lane j sums
// candidate j's tile partials in tile order (eight loads in flight); lane 0
// then takes the maximum in candidate order.  r <= 48 < blockDim (64).
__global__ void k_reduce_cut(const long long* __restrict__ ipart, const double* __restrict__ dpart,
                             int ntiles, int r, int use_int, double inv_scale,
                             double* __restrict__ out) {
  __shared__ double vals[64];
  const int j = threadIdx.x;
  if (j < r) {
    double v;
    if (use_int) {
      long long s = 0;
#pragma unroll 8
      for (int t = 0; t < ntiles; ++t) s += ipart[(long long)t * r + j];
      v = (double)s * 0.25;
    } else {
      double s = 0.0;
#pragma unroll 8
      for (int t = 0; t < ntiles; ++t) s += dpart[(long long)t * r + j];
      v = s * inv_scale;
    }
    out[j] = v;
    vals[j] = v;
  }
  __syncthreads();
  if (j == 0) {
    double best = -INFINITY;
    for (int q = 0; q < r; ++q) best = (q == 0 || vals[q] > best) ? vals[q] : best;
    out[r] = best;
  }
}

int make_cut_weights(Ctx* c, std::vector<DevBuf>& pool, double scale, CutWeights* out) {
  const Problem& P = c->P;
  *out = cut_weights_of(P);
  if (scale == 1.0) return USBS_OK;
  if (!(scale > 0.0) || !std::isfinite(scale))
    return set_err(c, USBS_ERR_INVALID, "cut scale must be positive and finite");
  std::vector<int> w(P.nnz > 0 ? P.nnz : 1, 0);
  bool integer = true;
  for (long long e = 0; e < P.nnz; ++e) {
    const double v4 = 4.0 * P.cval_host[e] / scale;
    const double rv = std::nearbyint(v4);
    if (!(std::fabs(v4) <= 1e9) || std::fabs(v4 - rv) > 1e-9 * std::fmax(1.0, std::fabs(v4))) {
      integer = false;
      break;
    }
    w[e] = (int)rv;
  }
  out->inv_scale = 1.0 / scale;
  out->use_int = integer ? 1 : 0;
  if (!integer) return USBS_OK;
  int* d = (int*)ctx_alloc(c, pool, sizeof(int) * w.size());
  if (!d) return USBS_ERR_OOM;
  if (hipMemcpy(d, w.data(), sizeof(int) * w.size(), hipMemcpyHostToDevice) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "upload of the cut weights");
  out->cw4 = d;
  return USBS_OK;
}

template <typename T>
int launch_round(const Problem& P, const CutWeights& cw, int r, const T* Pm, const T* Om,
                 long long ld, const double* G, const double* rho, signed char* sg,
                 long long* ipart, double* dpart, double* out, hipStream_t s) {
  if (r > 48) return (int)hipErrorInvalidValue;
  k_round_signs<T><<<(P.n + USBS_NT - 1) / USBS_NT, USBS_NT, 0, s>>>(P.n, r, Pm, Om, ld, G, rho, sg);
  int rc = (int)hipGetLastError();
  if (rc) return rc;
  const int use_int = cw.use_int;
  k_cut_values<<<P.ntiles, USBS_NT, 0, s>>>(P.n, P.tile, r, P.row_ptr, P.col, cw.cw4, P.cval, use_int,
                                            sg, ipart, dpart);
  rc = (int)hipGetLastError();
  if (rc) return rc;
  k_reduce_cut<<<1, 64, 0, s>>>(ipart, dpart, P.ntiles, r, use_int, cw.inv_scale, out);
  return (int)hipGetLastError();
}
template int launch_round<double>(const Problem&, const CutWeights&, int, const double*, const double*, long long,
                                  const double*, const double*, signed char*, long long*, double*,
                                  double*, hipStream_t);
template int launch_round<float>(const Problem&, const CutWeights&, int, const float*, const float*, long long,
                                 const double*, const double*, signed char*, long long*, double*,
                                 double*, hipStream_t);

}  // namespace usbs
