// Row- and constraint-parallel kernels around the Lanczos core: operator
// assembly, SpMM, tall-skinny products and Grams, the sketch update, the
// constraint map A(F F^T) with the dual/infeasibility reductions, the
// objective, and the max-cut rounding.
#include "engine.hpp"

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
