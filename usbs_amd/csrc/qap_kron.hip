// The lifted QAP's cost operator as batched small GEMMs on the matrix cores.
//
// utils/qap_helpers.py:157-172 builds C = W (x) D padded by an empty leading
// row and column: C[1 + l a + i, 1 + l c + j] = W[a, c] D[i, j].  Applied to
// a vector v whose Y block is Y[c, j] = v[1 + l c + j] (row-major l x l),
// (C v)[1 + l a + i] = sum_{c,j} W[a, c] Y[c, j] D[i, j] = (W Y D^T)[a, i]
// and (C v)[0] = 0.  For nvec vectors the kernels compute U_b = W Y_b D^T.
//
// BF16 form: one wavefront per vector.  W and D (bf16, zero-padded to a
// multiple of 16) sit in LDS for the whole launch; the wavefront stages its
// Y_b in LDS, forms T = Y_b D^T with v_mfma_f32_16x16x16_bf16 (FP32
// accumulation), rounds T to bf16 straight from the accumulator registers
// (the accumulator layout of a 16 x 16 tile is the B-operand layout of the
// next product, so T never leaves registers), forms U = W T the same way
// and writes U in FP32.  FP64 form: a workgroup per vector, D, Y_b and T in
// LDS, plain FP64 FMA; it is the accuracy and speed baseline.
#include <cmath>
#include <vector>

#include "../../include/usbs_hip.h"
#include "engine.hpp"

namespace usbs {
namespace {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

#define KR_NT 256
#define KR_MAXL 64      // BF16 form (LDS 6 x 64 x 68 halves)
#define KR_MAXL_F64 48  // FP64 form (LDS 3 x 48 x 49 doubles)

// float -> bf16, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__global__ void k_to_bf16(long long count, const double* __restrict__ src,
                          unsigned short* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = bf16_rne((float)src[i]);
}

__global__ void k_f32_to_f64(long long count, const float* __restrict__ src, double* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (double)src[i];
}

// 4 consecutive bf16 from LDS (8-byte aligned)
__device__ __forceinline__ bf16x4 lds4(const unsigned short* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}

// TM = padded size / 16.  LDS row stride LS = 16 TM + 4 halves.  DIRECT
// (l % 4 == 0): every 4-element chunk of a Y row is 8-byte aligned and lies
// wholly inside or outside the l columns, so the A operands load straight
// from memory and the wavefronts run without barriers; otherwise each
// wavefront stages its Y_b (zero-padded) in LDS.
template <int TM, bool DIRECT>
__global__ void __launch_bounds__(KR_NT)
k_kron_bf16(int l, long long nvec, const unsigned short* __restrict__ W,
            const unsigned short* __restrict__ D, const unsigned short* __restrict__ Y,
            float* __restrict__ U) {
  constexpr int LP = 16 * TM, LS = LP + 4;
  extern __shared__ unsigned short kl[];
  unsigned short* Ws = kl;
  unsigned short* Ds = Ws + LP * LS;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  unsigned short* Ys = Ds + LP * LS + wave * LP * LS;
  for (int idx = t; idx < LP * LP; idx += KR_NT) {
    const int r = idx / LP, c = idx - (idx / LP) * LP;
    const bool in = r < l && c < l;
    Ws[r * LS + c] = in ? W[r * l + c] : (unsigned short)0;
    Ds[r * LS + c] = in ? D[r * l + c] : (unsigned short)0;
  }
  const int row16 = lane & 15, k4 = 4 * (lane >> 4);
  const long long ll = (long long)l * l;
  if constexpr (DIRECT) __syncthreads();
  // DIRECT: each wavefront walks its own vectors.  Staged: the loop is
  // workgroup-uniform so that every wavefront passes the same barriers.
  const long long step = (long long)gridDim.x * 4;
  // DIRECT: the A operands of Y for the wavefront's next vector are loaded
  // while the current one is multiplied (two register sets).
  auto load_y = [&](long long bv, bf16x4 (&dst)[TM][TM]) {
    const unsigned short* yv = Y + (bv < nvec ? bv : 0) * ll;
#pragma unroll
    for (int mi = 0; mi < TM; ++mi)
#pragma unroll
      for (int ki = 0; ki < TM; ++ki) {
        const int r = 16 * mi + row16, c = 16 * ki + k4;
        dst[mi][ki] = (bv < nvec && r < l && c < l)
                          ? *reinterpret_cast<const bf16x4*>(yv + r * l + c)
                          : bf16x4{0, 0, 0, 0};
      }
  };
  bf16x4 ynext[TM][TM];
  if constexpr (DIRECT) load_y((long long)blockIdx.x * 4 + wave, ynext);
  for (long long b0 = (long long)blockIdx.x * 4 + (DIRECT ? wave : 0); b0 < nvec; b0 += step) {
    const long long b = DIRECT ? b0 : b0 + wave;
    const bool active = b < nvec;
    const unsigned short* yb = Y + (active ? b : 0) * ll;
    bf16x4 ya[TM][TM];  // DIRECT: A operands of Y, tile (mi, ki)
    if constexpr (DIRECT) {
#pragma unroll
      for (int mi = 0; mi < TM; ++mi)
#pragma unroll
        for (int ki = 0; ki < TM; ++ki) ya[mi][ki] = ynext[mi][ki];
      if (b0 + step < nvec) load_y(b0 + step, ynext);
    } else {
      __syncthreads();  // W, D staged; the previous vector's reads of Ys done
      for (int idx = lane; idx < LP * LP; idx += 64) {
        const int r = idx / LP, c = idx - (idx / LP) * LP;
        Ys[r * LS + c] = (active && r < l && c < l) ? yb[r * l + c] : (unsigned short)0;
      }
      __syncthreads();
    }
    // T = Y D^T: A = Y (rows c, k = j), B[k][n] = D[n][k]
    f32x4 acc[TM][TM];
#pragma unroll
    for (int mi = 0; mi < TM; ++mi)
#pragma unroll
      for (int ni = 0; ni < TM; ++ni) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ki = 0; ki < TM; ++ki) {
      bf16x4 bo[TM];
#pragma unroll
      for (int ni = 0; ni < TM; ++ni) bo[ni] = lds4(Ds + (16 * ni + row16) * LS + 16 * ki + k4);
#pragma unroll
      for (int mi = 0; mi < TM; ++mi) {
        const bf16x4 a = DIRECT ? ya[mi][ki] : lds4(Ys + (16 * mi + row16) * LS + 16 * ki + k4);
#pragma unroll
        for (int ni = 0; ni < TM; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bo[ni], acc[mi][ni], 0, 0, 0);
      }
    }
    // T tile (ci, ni) in accumulator layout = B operand of k-tile ci, n-tile ni
    bf16x4 tb[TM][TM];
#pragma unroll
    for (int ci = 0; ci < TM; ++ci)
#pragma unroll
      for (int ni = 0; ni < TM; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) tb[ci][ni][q] = (short)bf16_rne(acc[ci][ni][q]);
    // U = W T
    f32x4 u[TM][TM];
#pragma unroll
    for (int mi = 0; mi < TM; ++mi)
#pragma unroll
      for (int ni = 0; ni < TM; ++ni) u[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ki = 0; ki < TM; ++ki) {
#pragma unroll
      for (int mi = 0; mi < TM; ++mi) {
        const bf16x4 a = lds4(Ws + (16 * mi + row16) * LS + 16 * ki + k4);
#pragma unroll
        for (int ni = 0; ni < TM; ++ni)
          u[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, tb[ki][ni], u[mi][ni], 0, 0, 0);
      }
    }
    if (active) {
      float* ub = U + b * ll;
#pragma unroll
      for (int mi = 0; mi < TM; ++mi)
#pragma unroll
        for (int ni = 0; ni < TM; ++ni)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int r = 16 * mi + k4 + q, c = 16 * ni + row16;
            if (r < l && c < l) ub[r * l + c] = u[mi][ni][q];
          }
    }
  }
}

// FP64: a workgroup per vector; D, Y_b, T in LDS (row stride l + 1), W read
// through the caches.
__global__ void __launch_bounds__(KR_NT)
k_kron_f64(int l, long long nvec, const double* __restrict__ W, const double* __restrict__ D,
           const double* __restrict__ Y, double* __restrict__ U) {
  extern __shared__ double kd[];
  const int ls = l + 1;
  double* Ds = kd;
  double* Ys = Ds + l * ls;
  double* Ts = Ys + l * ls;
  const int t = threadIdx.x, ll = l * l;
  for (int idx = t; idx < ll; idx += KR_NT) {
    const int r = idx / l, c = idx - (idx / l) * l;
    Ds[r * ls + c] = D[idx];
  }
  for (long long b = blockIdx.x; b < nvec; b += gridDim.x) {
    __syncthreads();
    const double* yb = Y + b * ll;
    for (int idx = t; idx < ll; idx += KR_NT) {
      const int r = idx / l, c = idx - (idx / l) * l;
      Ys[r * ls + c] = yb[idx];
    }
    __syncthreads();
    for (int idx = t; idx < ll; idx += KR_NT) {  // T[c][i] = sum_j Y[c][j] D[i][j]
      const int c = idx / l, i = idx - (idx / l) * l;
      double s = 0.0;
      for (int j = 0; j < l; ++j) s += Ys[c * ls + j] * Ds[i * ls + j];
      Ts[c * ls + i] = s;
    }
    __syncthreads();
    double* ub = U + b * ll;
    for (int idx = t; idx < ll; idx += KR_NT) {  // U[a][i] = sum_c W[a][c] T[c][i]
      const int a = idx / l, i = idx - (idx / l) * l;
      double s = 0.0;
      for (int c = 0; c < l; ++c) s += W[a * l + c] * Ts[c * ls + i];
      ub[idx] = s;
    }
  }
}

// Grid: at most three workgroups per compute unit, each walking its
// vectors with W and D staged once (measured at l = 48, 8192 vectors:
// 2048 groups 27.8 us, 1024 24.4 us, 768 22.9 us, 512 23.1 us).
template <int TM>
hipError_t launch_bf16(int l, long long nvec, const unsigned short* W, const unsigned short* D,
                       const unsigned short* Y, float* U, int num_cus, hipStream_t s) {
  constexpr int LP = 16 * TM, LS = LP + 4;
  long long g = (nvec + 3) / 4;
  const long long cap = 3LL * (num_cus > 0 ? num_cus : 256);
  if (g > cap) g = cap;
  if (l % 4 == 0) {
    const size_t lds = sizeof(unsigned short) * (size_t)LP * LS * 2;
    k_kron_bf16<TM, true><<<(unsigned)g, KR_NT, lds, s>>>(l, nvec, W, D, Y, U);
  } else {
    const size_t lds = sizeof(unsigned short) * (size_t)LP * LS * 6;
    k_kron_bf16<TM, false><<<(unsigned)g, KR_NT, lds, s>>>(l, nvec, W, D, Y, U);
  }
  return hipGetLastError();
}

hipError_t launch_kron(int l, long long nvec, int precision, const void* W, const void* D,
                       const void* Y, void* U, int num_cus, hipStream_t s) {
  if (precision == USBS_PREC_BF16) {
    const auto* w = (const unsigned short*)W;
    const auto* d = (const unsigned short*)D;
    const auto* y = (const unsigned short*)Y;
    float* u = (float*)U;
    const int tm = (l + 15) / 16;
    if (tm == 1) return launch_bf16<1>(l, nvec, w, d, y, u, num_cus, s);
    if (tm == 2) return launch_bf16<2>(l, nvec, w, d, y, u, num_cus, s);
    if (tm == 3) return launch_bf16<3>(l, nvec, w, d, y, u, num_cus, s);
    return launch_bf16<4>(l, nvec, w, d, y, u, num_cus, s);
  }
  const size_t lds = sizeof(double) * (size_t)3 * l * (l + 1);
  long long g = nvec > 4096 ? 4096 : nvec;
  k_kron_f64<<<(unsigned)g, KR_NT, lds, s>>>(l, nvec, (const double*)W, (const double*)D,
                                             (const double*)Y, (double*)U);
  return hipGetLastError();
}

}  // namespace
}  // namespace usbs

using namespace usbs;

extern "C" int usbs_qap_kron_matvec(usbs_ctx* h, int32_t l, const double* W, const double* D,
                                    int64_t nvec, const double* Y, int32_t precision,
                                    int32_t reps, double* U, double* kernel_ms) {
  if (!h) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (!W || !D || !Y || !U) return set_err(c, USBS_ERR_INVALID, "null array argument");
  if (l < 1 || l > (precision == USBS_PREC_FP64 ? KR_MAXL_F64 : KR_MAXL))
    return set_err(c, USBS_ERR_INVALID, "l must be in [1, 48] (FP64) or [1, 64] (BF16)");
  if (nvec < 1) return set_err(c, USBS_ERR_INVALID, "nvec must be >= 1");
  if (reps < 1) return set_err(c, USBS_ERR_INVALID, "reps must be >= 1");
  if (precision != USBS_PREC_FP64 && precision != USBS_PREC_BF16)
    return set_err(c, USBS_ERR_INVALID, "precision must be USBS_PREC_FP64 or USBS_PREC_BF16");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  const long long ll = (long long)l * l, ny = nvec * ll;
  for (long long i = 0; i < ll; ++i)
    if (!std::isfinite(W[i]) || !std::isfinite(D[i])) return set_err(c, USBS_ERR_NONFINITE, "non-finite W / D");
  std::vector<DevBuf> pool;
  struct Guard {
    std::vector<DevBuf>& p;
    ~Guard() { ctx_free_pool(p); }
  } guard{pool};
  hipStream_t s = c->stream;
  const bool bf = precision == USBS_PREC_BF16;
  double* dW = (double*)ctx_alloc(c, pool, sizeof(double) * ll);
  double* dD = (double*)ctx_alloc(c, pool, sizeof(double) * ll);
  double* dY = (double*)ctx_alloc(c, pool, sizeof(double) * ny);
  if (!dW || !dD || !dY) return USBS_ERR_OOM;
  if (hipMemcpy(dW, W, sizeof(double) * ll, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dD, D, sizeof(double) * ll, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dY, Y, sizeof(double) * ny, hipMemcpyHostToDevice) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "upload failed");
  const void *kW = dW, *kD = dD, *kY = dY;
  void* kU = nullptr;
  if (bf) {
    unsigned short* bW = (unsigned short*)ctx_alloc(c, pool, 2 * ll);
    unsigned short* bD = (unsigned short*)ctx_alloc(c, pool, 2 * ll);
    unsigned short* bY = (unsigned short*)ctx_alloc(c, pool, 2 * ny);
    kU = ctx_alloc(c, pool, sizeof(float) * ny);
    if (!bW || !bD || !bY || !kU) return USBS_ERR_OOM;
    k_to_bf16<<<(unsigned)((ll + 255) / 256), 256, 0, s>>>(ll, dW, bW);
    k_to_bf16<<<(unsigned)((ll + 255) / 256), 256, 0, s>>>(ll, dD, bD);
    k_to_bf16<<<(unsigned)((ny + 255) / 256), 256, 0, s>>>(ny, dY, bY);
    kW = bW; kD = bD; kY = bY;
  } else {
    kU = ctx_alloc(c, pool, sizeof(double) * ny);
    if (!kU) return USBS_ERR_OOM;
  }
  if (launch_kron(l, nvec, precision, kW, kD, kY, kU, c->num_cus, s) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "Kronecker matvec launch failed");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "event create failed");
  int rc = USBS_OK;
  if (hipEventRecord(e0, s) != hipSuccess) rc = USBS_ERR_HIP;
  for (int r = 0; r < reps && rc == USBS_OK; ++r)
    if (launch_kron(l, nvec, precision, kW, kD, kY, kU, c->num_cus, s) != hipSuccess) rc = USBS_ERR_HIP;
  float ms = 0.f;
  if (rc == USBS_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = USBS_ERR_HIP;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return set_err(c, rc, "Kronecker matvec timing failed");
  if (kernel_ms) *kernel_ms = (double)ms / reps;
  if (bf) {
    k_f32_to_f64<<<(unsigned)((ny + 255) / 256), 256, 0, s>>>(ny, (const float*)kU, dY);
    if (hipGetLastError() != hipSuccess) return set_err(c, USBS_ERR_HIP, "conversion launch failed");
    kU = dY;
  }
  // the context's stream does not synchronise with the null stream: copy on it
  if (hipMemcpyAsync(U, kU, sizeof(double) * ny, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return set_err(c, USBS_ERR_HIP, "download failed");
  return USBS_OK;
}
