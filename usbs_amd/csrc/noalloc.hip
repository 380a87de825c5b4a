// usbs_sketch_update_noalloc: the sketch update of usbs_sketch_update on
// caller-owned device arrays, with every intermediate in a fixed device array
// of this library, so that the call allocates nothing.  It is for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops that forbid allocation in steady state).
//
// Scratch: `scratch`, a file-scope device array of SCRATCH_DOUBLES doubles
// (2 MiB).  Each GPU holds its own copy until the process ends; its address on
// a device is looked up once, when the first context on that device is
// created (noalloc_init_device, called by usbs_create with the device
// current), and that address -- never the array's name, which in host code is
// only a stand-in -- is what the kernels receive.  Every call on a device
// uses the same array: calls that overlap on one device race.  Nothing in
// the array needs to start at zero: each call writes every partial before it
// reads it.
#include <mutex>
#include <vector>

#include "../../include/usbs_hip.h"
#include "solve.hpp"

namespace usbs {

namespace {

constexpr long long SCRATCH_DOUBLES = 1ll << 18;     // 2 MiB per device
constexpr long long SCRATCH_FO = 32;                 // F^T Omega (k x r <= 16 x 48)
constexpr long long SCRATCH_PART = 1024;             // tile partials of F^T Omega
static_assert(SCRATCH_FO + 16 * 48 <= SCRATCH_PART, "scratch layout");

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

std::mutex g_mu;
std::vector<double*> g_scratch;  // per device: the address of `scratch` there

__global__ void k_set_scalar(double* dst, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = v;
}

double* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  return device >= 0 && device < (int)g_scratch.size() ? g_scratch[device] : nullptr;
}

}  // namespace

// Look up this device's copy of the scratch array (the device must be
// current).  Once per device; later calls return at once.
int noalloc_init_device(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (device < 0) return USBS_ERR_INVALID;
  if ((int)g_scratch.size() <= device) g_scratch.resize(device + 1, nullptr);
  if (g_scratch[device]) return USBS_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return USBS_ERR_HIP;
  }
  g_scratch[device] = (double*)p;
  return USBS_OK;
}

}  // namespace usbs

using namespace usbs;

extern "C" int usbs_sketch_update_noalloc(usbs_ctx* h, int64_t n, int32_t k, int32_t kc,
                                          int32_t r, int64_t ld, const void* F,
                                          const void* Omega, const void* Pbar, double eta,
                                          int32_t precision, void* P_out, void* Pbar_out) {
  if (!h) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (!F || !Pbar || !P_out || !Pbar_out || (r > 0 && !Omega))
    return set_err(c, USBS_ERR_INVALID, "null device array");
  if (n < 1 || n > INT32_MAX || k < 1 || k > 16 || kc < 0 || kc > k || r < 0 || r > 48)
    return set_err(c, USBS_ERR_INVALID, "bad sketch sizes");
  if (ld < n) return set_err(c, USBS_ERR_INVALID, "leading dimension below n");
  if (precision != USBS_PREC_FP64 && precision != USBS_PREC_MIXED)
    return set_err(c, USBS_ERR_INVALID, "unknown precision");
  if (!std::isfinite(eta)) return set_err(c, USBS_ERR_NONFINITE, "non-finite eta");
  const int tile = tile_rows_for((int)n), ntiles = (int)((n + tile - 1) / tile);
  const int rr = r > 0 ? r : 1;
  const long long need = SCRATCH_PART + (long long)ntiles * k * rr;
  if (need > SCRATCH_DOUBLES)
    return set_err(c, USBS_ERR_INVALID,
                   "sizes need " + std::to_string(need) + " scratch doubles, more than the " +
                       std::to_string(SCRATCH_DOUBLES) + " of the fixed array");
  double* base = scratch_of(c->device);
  if (!base) return set_err(c, USBS_ERR_STATE, "scratch address not set up for this device");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  hipStream_t s = c->stream;
  double* d_eta = base;
  double* FO = base + SCRATCH_FO;
  double* part = base + SCRATCH_PART;
  k_set_scalar<<<1, 64, 0, s>>>(d_eta, eta);
  USBS_HIP(c, hipGetLastError());
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    const T* dF = (const T*)F;
    const T* dO = (const T*)Omega;
    T* dP = (T*)P_out;
    T* dPb = (T*)Pbar_out;
    // Pbar_out = Pbar, then the kernel updates it in place (Pbar stays intact)
    if (hipMemcpyAsync(dPb, Pbar, sizeof(T) * (size_t)ld * rr, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return set_err(c, USBS_ERR_HIP, "copy of Pbar");
    if (r > 0 && launch_tsgram<T, T>((int)n, tile, ntiles, dF, ld, k, dO, ld, r, part, FO, s))
      return set_err(c, USBS_ERR_HIP, "F^T Omega");
    if (launch_sketch_update<T>((int)n, k, kc, r, dF, ld, FO, d_eta, dPb, dP, ld, s))
      return set_err(c, USBS_ERR_HIP, "sketch update");
    return (int)USBS_OK;
  };
  const int rc = precision == USBS_PREC_FP64 ? run(double(0)) : run(float(0));
  if (rc) return rc;
  USBS_HIP(c, hipStreamSynchronize(s));
  return USBS_OK;
}
