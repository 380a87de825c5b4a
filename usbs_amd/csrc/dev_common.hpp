// Shared device-side helpers for the USBS HIP library (gfx950 / CDNA4).
//
// Conventions used by every kernel in this library:
//  * Tall vectors and bases are column-major with an explicit leading
//    dimension; "T" is the storage type (double or float); all reductions
//    accumulate in double.
//  * Rows are split into fixed tiles of `tile` rows (a function of n only).
//    Every reduction over rows first produces one partial per tile, then sums
//    the tile partials in an order fixed by the tile count alone.  Results
//    therefore do not depend on the
//    launch shape, on how many workgroups a persistent launch has, or on which
//    other problems share the GPU: a solve is bitwise the same alone and in a
//    batch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#define USBS_NT 256          // threads per block of every row-parallel kernel
#define USBS_NPART 8         // second-level split of the LDS transpose reduction
#define USBS_DENSE_NT 64     // one wavefront for the small dense FP64 kernels

namespace usbs {

// f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>): a loop
// unrolled at template instantiation, so that arrays indexed by it have
// constant indices from the start and are kept in registers (a loop under
// #pragma unroll is unrolled only after the compiler has decided where such
// an array lives, and a lane-dependent select between two of its elements
// can then pin it to scratch memory).
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Global-memory access policies of the tile routines.  MemPlain: ordinary
// loads and stores (kernels whose cross-workgroup data meets only at kernel
// boundaries, or behind grid_barrier's cache maintenance).  MemCoherent:
// agent-scope relaxed atomic loads and stores, which bypass the XCDs'
// non-coherent L2 caches (sc1), so data written by one workgroup is seen by
// every other after coh_arrive/coh_wait without any writeback or invalidate
// of the L2.  Read-only inputs stay on plain loads under either policy.
struct MemPlain {
  template <typename X>
  static __device__ __forceinline__ X ld(const X* p) { return *p; }
  template <typename X>
  static __device__ __forceinline__ void st(X* p, X v) { *p = v; }
};
struct MemCoherent {
  template <typename X>
  static __device__ __forceinline__ X ld(const X* p) {
    return __hip_atomic_load(const_cast<X*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <typename X>
  static __device__ __forceinline__ void st(X* p, X v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
};

// The value of lane (lane ^ MASK) within the wavefront.  MASK 1, 2: DPP quad
// permutations; 4, 8: DPP row shifts (row_shl brings lane + n, row_shr lane
// - n) and a select on the lane bit; 16, 32: ds_bpermute.  DPP moves stay in
// the vector ALU, where ds_bpermute goes through the LDS crossbar.  Must be
// called with every lane of the wavefront active.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffull), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}
template <int MASK>
__device__ __forceinline__ double lane_xor(double v) {
  if constexpr (MASK == 1) {
    return dpp_f64<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  } else if constexpr (MASK == 2) {
    return dpp_f64<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  } else if constexpr (MASK == 4 || MASK == 8) {
    const double up = dpp_f64<0x100 | MASK>(v);   // row_shl: lane + MASK
    const double dn = dpp_f64<0x110 | MASK>(v);   // row_shr: lane - MASK
    return (threadIdx.x & MASK) ? dn : up;
  } else {
    return __shfl_xor(v, MASK, 64);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Deterministic block reduction of `c` per-thread accumulators, processed
// in chunks of USBS_RCH columns.  red: LDS of USBS_RCH*USBS_NT doubles,
// red2: LDS of USBS_RCH*USBS_NPART doubles.  On return out[j] (LDS or
// global) holds the block sum of acc[j] for j < c; the summation order is
// fixed (thread order within a part, then part order).
#define USBS_RCH 16
template <int MAXC>
__device__ __forceinline__ void block_reduce_cols(const double (&acc)[MAXC], int c, double* red,
                                                  double* red2, double* out) {
  const int t = threadIdx.x;
  constexpr int SEG = USBS_NT / USBS_NPART;
#pragma unroll
  for (int c0 = 0; c0 < MAXC; c0 += USBS_RCH) {
    if (c0 < c) {
      const int cc = (c - c0) < USBS_RCH ? (c - c0) : USBS_RCH;
#pragma unroll
      for (int j = 0; j < USBS_RCH; ++j)
        if (c0 + j < MAXC && j < cc) red[j * USBS_NT + t] = acc[(c0 + j < MAXC) ? c0 + j : 0];
      __syncthreads();
      for (int idx = t; idx < cc * USBS_NPART; idx += USBS_NT) {
        const int j = idx / USBS_NPART, part = idx % USBS_NPART;
        const double* p = red + j * USBS_NT + part * SEG;
        double s = 0.0;
#pragma unroll 8
        for (int q = 0; q < SEG; ++q) s += p[q];
        red2[idx] = s;
      }
      __syncthreads();
      for (int j = t; j < cc; j += USBS_NT) {
        double s = 0.0;
#pragma unroll
        for (int part = 0; part < USBS_NPART; ++part) s += red2[j * USBS_NPART + part];
        out[c0 + j] = s;
      }
      __syncthreads();
    }
  }
}

// Sum of one accumulator across the block, deterministic.
__device__ __forceinline__ double block_sum1(double v, double* lds /* >= NT/64 doubles */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += lds[w];
  __syncthreads();
  return s;
}

// Sum tile partials part[tile * stride + j] over tiles, j < c <= blockDim,
// into out (LDS).  P = blockDim / c threads share a column: thread (j, q)
// sums tiles q, q + P, ... in order, then thread j adds the P sums in order.
// The order depends only on (c, ntiles, blockDim), so it is deterministic.
// Each thread issues up to 16 loads before the first add, so the partials of
// a typical reduction arrive in one memory round trip.
// scratch: LDS of blockDim doubles (must not alias out).
template <typename M = MemPlain>
__device__ __forceinline__ void sum_tile_partials(const double* __restrict__ part, int ntiles,
                                                  int stride, int c, double* out,
                                                  double* scratch) {
  const int t = threadIdx.x, nt = blockDim.x;
  if (c <= 0) { __syncthreads(); return; }
  int P = nt / c;
  if (P < 1) P = 1;
  if (P > ntiles) P = ntiles;
  if (t < c * P) {
    const int j = t / P, q = t - (t / P) * P;
    double s = 0.0;
    for (int tl0 = q; tl0 < ntiles; tl0 += 16 * P) {
      // every load issued before the first add: tiles past the end read
      // tile q (in bounds) and are not added
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int tl = tl0 + u * P;
        v[u] = M::ld(part + (size_t)(tl < ntiles ? tl : q) * stride + j);
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (tl0 + u * P < ntiles) s += v[u];
    }
    scratch[t] = s;
  }
  __syncthreads();
  if (t < c) {
    double s = 0.0;
    for (int q = 0; q < P; ++q) s += scratch[t * P + q];
    out[t] = s;
  }
  __syncthreads();
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Standard normal number idx of stream `key`: Box-Muller on two SplitMix64
// uniforms with counter key*2^32 + 2*idx (+1).
__device__ __forceinline__ double counter_normal(uint64_t key, uint64_t idx) {
  const uint64_t base = key << 32;
  const uint64_t a = splitmix64(base + 2 * idx), b = splitmix64(base + 2 * idx + 1);
  const double u1 = (double)(a >> 11) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log1p(-u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

// Device-wide barrier among `nblocks` workgroups that share `counter`
// (monotonic, zeroed before the launch).  The caller's workgroups must all be
// resident (cooperative launch).  `target` is kept by the caller and grows by
// nblocks each call.  The L2 of an XCD is not coherent with the others, so
// thread 0 writes it back (agent-scope release fence) before arriving and
// invalidates it with the compute unit's L1 (acquire fence) after the wait.
// __syncthreads() orders only LDS operations, so two waits make those fences
// cover the whole workgroup: every wavefront waits for its own global stores
// to reach the L2 before the workgroup barrier in front of the writeback, and
// thread 0 waits for the invalidation to complete before the barrier that
// releases the other wavefronts' loads.  Without them another wavefront's
// stores could miss the writeback, or its first loads be served stale; both
// happened, rarely, as results that differed between runs.  The polling loop
// uses relaxed loads so that waiting does not flush the cache the XCD's other
// workgroups are working from.  A bounded wait: after `spin_limit` polls the
// barrier records an error in *err and every later barrier returns at once,
// so a wait that can never complete ends instead of hanging the device.
__device__ __forceinline__ bool grid_barrier(unsigned* counter, unsigned nblocks, unsigned& target,
                                             int* err, unsigned long long spin_limit) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  target += nblocks;
  __shared__ int s_ok;
  if (threadIdx.x == 0) {
    int ok = 1;
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
      ok = 0;
    } else {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      unsigned long long polls = 0;
      while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(1);
        if (++polls > spin_limit) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        if ((polls & 1023) == 0 &&
            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          ok = 0;
          break;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __builtin_amdgcn_s_waitcnt(0);
    }
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Device-wide barrier of the MemCoherent kernels, split into arrival and
// wait so that work independent of the other workgroups can run in between.
// Every cross-workgroup datum of those kernels is written and read with
// agent-scope (sc1) accesses, so the barrier needs no L2 writeback or
// invalidate: each wavefront waits for its own stores to complete, the
// workgroup synchronises, and thread 0 adds one arrival to the counter of
// its group (workgroup index mod 8; eight counters 256 bytes apart, so that
// arrivals do not queue on one address).  The wait polls the eight counters
// from eight lanes of wavefront 0 until each holds its group's arrivals for
// this epoch.  `counters` (8 x 64 unsigned) are zeroed before the launch;
// `epoch` is kept by the caller and counts completed arrivals.  Bounded: a
// wait that exceeds spin_limit polls records *err = 1, and every later
// arrival or wait returns false at once.
__device__ __forceinline__ bool coh_arrive(unsigned* counters, int lw, unsigned& epoch, int* err) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  ++epoch;
  __shared__ int s_ok_a;
  if (threadIdx.x == 0) {
    const int ok = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    if (ok) __hip_atomic_fetch_add(counters + 64 * (lw & 7), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_ok_a = ok;
  }
  __syncthreads();
  return s_ok_a != 0;
}

__device__ __forceinline__ bool coh_wait(unsigned* counters, int nblocks, unsigned epoch, int* err,
                                         unsigned long long spin_limit) {
  __shared__ int s_ok_w;
  const int t = threadIdx.x;
  if (t < 64) {
    const unsigned per = (t < 8) ? (unsigned)(nblocks / 8 + (t < nblocks % 8 ? 1 : 0)) : 0u;
    const unsigned want = per * epoch;
    unsigned long long polls = 0;
    int ok = 1;
    while (true) {
      const bool mine = (t >= 8) || want == 0u ||
                        __hip_atomic_load(counters + 64 * t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want;
      if (__all(mine)) break;
      __builtin_amdgcn_s_sleep(1);
      if (++polls > spin_limit) {
        if (t == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      if ((polls & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        ok = 0;
        break;
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    if (t == 0) s_ok_w = ok;
  }
  __syncthreads();
  return s_ok_w != 0;
}

__device__ __forceinline__ bool coh_barrier(unsigned* counters, int lw, int nblocks, unsigned& epoch,
                                            int* err, unsigned long long spin_limit) {
  if (!coh_arrive(counters, lw, epoch, err)) return false;
  return coh_wait(counters, nblocks, epoch, err, spin_limit);
}

}  // namespace usbs
