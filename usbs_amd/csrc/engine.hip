// Context, problem upload and the standalone Lanczos entry point.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "../../include/usbs_hip.h"


namespace usbs {

int set_err(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

void ctx_release_cache(Ctx* c) {
  for (auto& kv : c->cache.blocks) (void)hipFree(kv.second);
  c->cache.blocks.clear();
  c->cache.bytes = 0;
}

void* ctx_alloc(Ctx* c, std::vector<DevBuf>& pool, size_t bytes) {
  bytes = (bytes + 255) / 256 * 256;
  if (bytes == 0) bytes = 256;
  BufCache& bc = c->cache;
  bc.stream = c->stream;
  // the smallest cached block that fits without wasting more than half of it
  auto it = bc.blocks.lower_bound(bytes);
  if (it != bc.blocks.end() && it->first <= 2 * bytes) {
    DevBuf b{it->second, it->first, &bc};
    bc.bytes -= it->first;
    bc.blocks.erase(it);
    pool.push_back(b);
    return b.p;
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    ctx_release_cache(c);
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      c->err = "hipMalloc of " + std::to_string(bytes) + " bytes failed";
      return nullptr;
    }
  }
  pool.push_back({p, bytes, &bc});
  return p;
}

void ctx_free_pool(std::vector<DevBuf>& pool) {
  BufCache* synced = nullptr;
  for (auto& b : pool) {
    if (!b.p) continue;
    if (!b.cache) {
      (void)hipFree(b.p);
      continue;
    }
    if (b.cache != synced) {
      // the context's queued work may still use the block
      (void)hipStreamSynchronize(b.cache->stream);
      synced = b.cache;
    }
    b.cache->blocks.emplace(b.bytes, b.p);
    b.cache->bytes += b.bytes;
  }
  pool.clear();
}

template <typename X>
static X* upload(Ctx* c, std::vector<DevBuf>& pool, const X* host, size_t count) {
  X* d = (X*)ctx_alloc(c, pool, sizeof(X) * count);
  if (!d) return nullptr;
  if (count && hipMemcpy(d, host, sizeof(X) * count, hipMemcpyHostToDevice) != hipSuccess) {
    c->err = "hipMemcpy upload failed";
    return nullptr;
  }
  return d;
}

// Build the merged CSR pattern of C and of the triples' (i, j), C's values in
// it, and the slot->triple and constraint->triple maps (counting sorts, so
// each slot and each constraint lists its triples in input order).
static int build_problem(Ctx* c, int64_t n, int64_t m, int64_t nnzC, const int32_t* Cr,
                         const int32_t* Cc, const double* Cv, int64_t nA, const int32_t* Ai,
                         const double* Ad, const double* b, const double* mask, const double* sa) {
  if (n <= 0 || m <= 0 || nnzC < 0 || nA < 0 || n > (int64_t)INT32_MAX / 2 || m > INT32_MAX)
    return set_err(c, USBS_ERR_INVALID, "invalid problem sizes");
  for (int64_t e = 0; e < nnzC; ++e) {
    if (Cr[e] < 0 || Cr[e] >= n || Cc[e] < 0 || Cc[e] >= n)
      return set_err(c, USBS_ERR_INVALID, "C index out of range at entry " + std::to_string(e));
    if (!std::isfinite(Cv[e])) return set_err(c, USBS_ERR_NONFINITE, "non-finite C value");
  }
  for (int64_t t = 0; t < nA; ++t) {
    const int32_t cc = Ai[3 * t], ii = Ai[3 * t + 1], jj = Ai[3 * t + 2];
    if (cc < 0 || cc >= m || ii < 0 || ii >= n || jj < 0 || jj >= n)
      return set_err(c, USBS_ERR_INVALID, "A_indices out of range at triple " + std::to_string(t));
    if (!std::isfinite(Ad[t])) return set_err(c, USBS_ERR_NONFINITE, "non-finite A_data");
  }
  for (int64_t i = 0; i < m; ++i)
    if (!std::isfinite(b[i]) || !std::isfinite(mask[i]) || !std::isfinite(sa[i]))
      return set_err(c, USBS_ERR_NONFINITE, "non-finite b / b_ineq_mask / SCALE_A");

  // pattern keys
  std::vector<uint64_t> keys;
  keys.reserve((size_t)(nnzC + nA));
  for (int64_t e = 0; e < nnzC; ++e) keys.push_back((uint64_t)Cr[e] * (uint64_t)n + (uint64_t)Cc[e]);
  for (int64_t t = 0; t < nA; ++t)
    keys.push_back((uint64_t)Ai[3 * t + 1] * (uint64_t)n + (uint64_t)Ai[3 * t + 2]);
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const int64_t nnz = (int64_t)keys.size();
  if (nnz > INT32_MAX) return set_err(c, USBS_ERR_INVALID, "operator pattern exceeds int32");
  std::vector<int> row_ptr(n + 1, 0), col(nnz);
  for (int64_t s = 0; s < nnz; ++s) {
    row_ptr[keys[s] / n + 1]++;
    col[s] = (int)(keys[s] % n);
  }
  for (int64_t i = 0; i < n; ++i) row_ptr[i + 1] += row_ptr[i];
  auto slot_of = [&](uint64_t key) -> int64_t {
    return std::lower_bound(keys.begin(), keys.end(), key) - keys.begin();
  };
  std::vector<double> cval(nnz, 0.0);
  for (int64_t e = 0; e < nnzC; ++e) cval[slot_of((uint64_t)Cr[e] * n + Cc[e])] += Cv[e];
  bool cut_int = true;
  std::vector<int> cw4(nnz, 0);
  for (int64_t s = 0; s < nnz; ++s) {
    const double v4 = 4.0 * cval[s];
    if (std::fabs(v4) > 1e9 || v4 != std::nearbyint(v4)) cut_int = false;
    else cw4[s] = (int)std::nearbyint(v4);
  }
  std::vector<int> tslot(nA), slot_ptr(nnz + 1, 0), slot_trip(nA), con_ptr(m + 1, 0);
  std::vector<int> ct_i(nA), ct_j(nA);
  std::vector<double> ct_d(nA);
  std::vector<int> tcon(nA), ti(nA), tj(nA);
  for (int64_t t = 0; t < nA; ++t) {
    tcon[t] = Ai[3 * t];
    ti[t] = Ai[3 * t + 1];
    tj[t] = Ai[3 * t + 2];
    tslot[t] = (int)slot_of((uint64_t)ti[t] * n + tj[t]);
    slot_ptr[tslot[t] + 1]++;
    con_ptr[tcon[t] + 1]++;
  }
  for (int64_t s = 0; s < nnz; ++s) slot_ptr[s + 1] += slot_ptr[s];
  for (int64_t i = 0; i < m; ++i) con_ptr[i + 1] += con_ptr[i];
  {
    std::vector<int> fill_s(slot_ptr.begin(), slot_ptr.end() - 1), fill_c(con_ptr.begin(), con_ptr.end() - 1);
    for (int64_t t = 0; t < nA; ++t) {
      slot_trip[fill_s[tslot[t]]++] = (int)t;
      const int e = fill_c[tcon[t]]++;
      ct_i[e] = ti[t];
      ct_j[e] = tj[t];
      ct_d[e] = Ad[t];
    }
  }
  bool diag = (nA == n && m == n);
  for (int64_t t = 0; t < nA && diag; ++t)
    diag = (tcon[t] == t && ti[t] == t && tj[t] == t && Ad[t] == 1.0);
  bool ineq = false;
  for (int64_t i = 0; i < m; ++i) ineq |= (mask[i] != 0.0);

  ctx_free_pool(c->prob_bufs);
  ctx_free_pool(c->work_bufs);
  if (c->keep_n != (int)n || c->keep_m != (int)m) {
    ctx_free_pool(c->keep_bufs);
    c->keep_y = c->keep_z = c->keep_P = nullptr;
    c->keep_n = c->keep_m = 0;
    c->keep_r = -1;
  }
  c->has_problem = false;
  c->tuned_mode = -1;
  Problem& P = c->P;
  P = Problem();
  P.n = (int)n;
  P.m = (int)m;
  P.nnz = nnz;
  P.nA = nA;
  P.tile = tile_rows_for((int)n);
  P.ntiles = (int)((n + P.tile - 1) / P.tile);
  P.lz_heavy = (n <= 65536 && (int64_t)row_ptr[n] >= 32 * n) ? 1 : 0;
  P.lz_tile = P.lz_heavy ? lanczos_heavy_tile_rows((int)n) : P.tile;
  P.lz_ntiles = (int)((n + P.lz_tile - 1) / P.lz_tile);
  for (int t = 0; t < P.lz_ntiles; ++t) {
    const int64_t a = (int64_t)t * P.lz_tile, b = std::min<int64_t>(n, a + P.lz_tile);
    P.tile_nnz_max = std::max(P.tile_nnz_max, row_ptr[b] - row_ptr[a]);
  }
  P.mtile = 256;
  P.mntiles = (int)((m + P.mtile - 1) / P.mtile);
  P.diag_A = diag;
  P.has_ineq = ineq;
  P.cut_integer = cut_int;
  P.cval_host = cval;
  auto& pool = c->prob_bufs;
  std::vector<double> tdata(Ad, Ad + nA);
  if (!(P.row_ptr = upload(c, pool, row_ptr.data(), row_ptr.size()))) return USBS_ERR_OOM;
  if (!(P.col = upload(c, pool, col.data(), col.size()))) return USBS_ERR_OOM;
  if (!(P.cval = upload(c, pool, cval.data(), cval.size()))) return USBS_ERR_OOM;
  if (!(P.cw4 = upload(c, pool, cw4.data(), cw4.size()))) return USBS_ERR_OOM;
  if (!(P.slot_ptr = upload(c, pool, slot_ptr.data(), slot_ptr.size()))) return USBS_ERR_OOM;
  if (!(P.slot_trip = upload(c, pool, slot_trip.data(), slot_trip.size()))) return USBS_ERR_OOM;
  if (!(P.tcon = upload(c, pool, tcon.data(), tcon.size()))) return USBS_ERR_OOM;
  if (!(P.ti = upload(c, pool, ti.data(), ti.size()))) return USBS_ERR_OOM;
  if (!(P.tj = upload(c, pool, tj.data(), tj.size()))) return USBS_ERR_OOM;
  if (!(P.tdata = upload(c, pool, tdata.data(), tdata.size()))) return USBS_ERR_OOM;
  if (!(P.con_ptr = upload(c, pool, con_ptr.data(), con_ptr.size()))) return USBS_ERR_OOM;
  if (!(P.ct_i = upload(c, pool, ct_i.data(), ct_i.size()))) return USBS_ERR_OOM;
  if (!(P.ct_j = upload(c, pool, ct_j.data(), ct_j.size()))) return USBS_ERR_OOM;
  if (!(P.ct_d = upload(c, pool, ct_d.data(), ct_d.size()))) return USBS_ERR_OOM;
  if (!(P.b = upload(c, pool, b, (size_t)m))) return USBS_ERR_OOM;
  if (!(P.mask = upload(c, pool, mask, (size_t)m))) return USBS_ERR_OOM;
  if (!(P.scale_a = upload(c, pool, sa, (size_t)m))) return USBS_ERR_OOM;
  c->has_problem = true;
  return USBS_OK;
}

// Lanczos workspace for one instance of dimension n, basis m, storage T.
template <typename T>
int make_lanczos(Ctx* c, std::vector<DevBuf>& pool, const Problem& P, int m, int nd,
                 int max_restarts, double tol, const T* opval, LzInst<T>& L) {
  L = LzInst<T>();
  L.n = P.n;
  L.tile = P.lz_tile;
  L.ntiles = P.lz_ntiles;
  L.heavy = P.lz_heavy;
  L.m = m;
  L.nd = nd;
  L.max_restarts = max_restarts;
  L.tol = tol;
  L.pstride = ((m + 1) + 7) / 8 * 8;
  L.nblocks = 1;
  L.row_ptr = P.row_ptr;
  L.col = P.col;
  L.val = opval;
  L.ldq = ((long long)P.n + 63) / 64 * 64;
  if (!(L.Q = (T*)ctx_alloc(c, pool, sizeof(T) * L.ldq * (m + 1)))) return USBS_ERR_OOM;
  if (!(L.w = (T*)ctx_alloc(c, pool, sizeof(T) * 2 * L.ldq))) return USBS_ERR_OOM;
  L.tile_nnz_max = P.tile_nnz_max;
  const size_t ps = sizeof(double) * (size_t)L.ntiles * L.pstride;
  if (!(L.part1 = (double*)ctx_alloc(c, pool, ps))) return USBS_ERR_OOM;
  if (!(L.part2 = (double*)ctx_alloc(c, pool, ps))) return USBS_ERR_OOM;
  if (!(L.part3 = (double*)ctx_alloc(c, pool, ps))) return USBS_ERR_OOM;
  if (!(L.h1 = (double*)ctx_alloc(c, pool, sizeof(double) * (m + 1)))) return USBS_ERR_OOM;
  if (!(L.h2 = (double*)ctx_alloc(c, pool, sizeof(double) * (m + 1)))) return USBS_ERR_OOM;
  if (!(L.alpha = (double*)ctx_alloc(c, pool, sizeof(double) * m))) return USBS_ERR_OOM;
  if (!(L.beta = (double*)ctx_alloc(c, pool, sizeof(double) * m))) return USBS_ERR_OOM;
  if (!(L.Yr = (double*)ctx_alloc(c, pool, sizeof(double) * m * m))) return USBS_ERR_OOM;
  if (!(L.ctrl = (int*)ctx_alloc(c, pool, sizeof(int) * LZ_NCTRL))) return USBS_ERR_OOM;
  L.nbar = max_restarts + 1 > 64 ? max_restarts + 1 : 64;
  if (!(L.bar = (unsigned*)ctx_alloc(c, pool, sizeof(unsigned) * LZ_BAR_STRIDE * L.nbar))) return USBS_ERR_OOM;
  if (!(L.lstat = (int*)ctx_alloc(c, pool, sizeof(int) * 2 * L.nbar))) return USBS_ERR_OOM;
  return USBS_OK;
}

template int make_lanczos<double>(Ctx*, std::vector<DevBuf>&, const Problem&, int, int, int, double,
                                  const double*, LzInst<double>&);
template int make_lanczos<float>(Ctx*, std::vector<DevBuf>&, const Problem&, int, int, int, double,
                                 const float*, LzInst<float>&);

// Reset per-call Lanczos state: Q = [q0, 0...], alpha = beta = 0, ctrl = 0.
template <typename T>
__global__ void k_lz_reset(LzInst<T> L, const double* q0) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long tot = L.ldq * (L.m + 1);
  for (long long x = i; x < tot; x += (long long)gridDim.x * blockDim.x) {
    const long long r = x % L.ldq, cidx = x / L.ldq;
    L.Q[x] = (cidx == 0 && r < L.n) ? (T)q0[r] : (T)0;
  }
  if (i < L.m) { L.alpha[i] = 0.0; L.beta[i] = 0.0; }
  if (i < LZ_NCTRL) L.ctrl[i] = 0;
  for (long long x = i; x < (long long)LZ_BAR_STRIDE * L.nbar; x += (long long)gridDim.x * blockDim.x) L.bar[x] = 0u;
  for (long long x = i; x < 2 * L.nbar; x += (long long)gridDim.x * blockDim.x) L.lstat[x] = 0;
}

template <typename T>
__global__ void k_copy_cols_out(const T* Q, long long ldq, int n, int nd, double* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * nd) return;
  const long long r = i / nd, j = i % nd;
  out[i] = (double)Q[r + j * ldq];
}

template <typename T>
int reset_lanczos(const LzInst<T>& L, const double* q0, hipStream_t s) {
  k_lz_reset<T><<<1024, 256, 0, s>>>(L, q0);
  return (int)hipGetLastError();
}

template <typename T>
static int run_lanczos_standalone(Ctx* c, double C_coef, const double* y, const double* q0, int nd,
                                  int m, int max_restarts, double tol, int exec_mode,
                                  double* eigvals, double* eigvecs, int* info) {
  const Problem& P = c->P;
  std::vector<DevBuf> pool;
  struct Guard {
    std::vector<DevBuf>& p;
    ~Guard() { ctx_free_pool(p); }
  } guard{pool};
  double* d_y = upload(c, pool, y, (size_t)P.m);
  double* d_q0 = upload(c, pool, q0, (size_t)P.n);
  double* d_val64 = (double*)ctx_alloc(c, pool, sizeof(double) * P.nnz);
  float* d_val32 = nullptr;
  if (!d_y || !d_q0 || !d_val64) return USBS_ERR_OOM;
  if (sizeof(T) == 4 && !(d_val32 = (float*)ctx_alloc(c, pool, sizeof(float) * P.nnz))) return USBS_ERR_OOM;
  int rc = launch_assemble(P, C_coef, d_y, 1.0, nullptr, 0.0, nullptr, 0.0, d_val64, d_val32, c->stream);
  if (rc) return set_err(c, USBS_ERR_HIP, "assemble launch failed");
  LzInst<T> L;
  rc = make_lanczos<T>(c, pool, P, m, nd, max_restarts, tol,
                       sizeof(T) == 4 ? (const T*)d_val32 : (const T*)d_val64, L);
  if (rc) return rc;
  if (reset_lanczos(L, d_q0, c->stream)) return set_err(c, USBS_ERR_HIP, "reset launch failed");
  int host_ctrl[LZ_NCTRL] = {0};
  if (exec_mode == USBS_EXEC_PEROP) {
    rc = lanczos_perop<T>(L, c->stream, host_ctrl);
    if (rc) return set_err(c, USBS_ERR_HIP, std::string("lanczos per-op: ") + hipGetErrorString((hipError_t)rc));
  } else {
    size_t cache_bytes = 0;
    rc = lanczos_plan_persistent<T>(L, c->device, c->persist_wgs, &cache_bytes);
    if (rc) return set_err(c, USBS_ERR_HIP, "no persistent Lanczos launch fits the device");
    const int nb = L.nblocks;
    LzInst<T>* d_inst = (LzInst<T>*)ctx_alloc(c, pool, sizeof(LzInst<T>));
    int* d_map = (int*)ctx_alloc(c, pool, sizeof(int2) * (nb + L.ntiles));
    if (!d_inst || !d_map) return USBS_ERR_OOM;
    std::vector<int2> map(nb + L.ntiles);
    for (int i = 0; i < nb; ++i) map[i] = make_int2(0, i);
    for (int i = 0; i < L.ntiles; ++i) map[nb + i] = make_int2(0, i);
    USBS_HIP(c, hipMemcpyAsync(d_inst, &L, sizeof(L), hipMemcpyHostToDevice, c->stream));
    USBS_HIP(c, hipMemcpyAsync(d_map, map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice, c->stream));
    int* ctrl_ptr = L.ctrl;
    const LzStop stop{&ctrl_ptr, 2, nullptr};
    rc = lanczos_persistent<T>(d_inst, 1, nb, m, max_restarts, c->stream, d_map, d_map + 2 * nb,
                               L.ntiles, LZ_SPIN_LIMIT, cache_bytes, nullptr, lanczos_tile_form(L),
                               &stop);
    if (rc) return set_err(c, USBS_ERR_HIP, std::string("lanczos persistent: ") + hipGetErrorString((hipError_t)rc));
    USBS_HIP(c, hipMemcpyAsync(host_ctrl, L.ctrl, sizeof(host_ctrl), hipMemcpyDeviceToHost, c->stream));
    USBS_HIP(c, hipStreamSynchronize(c->stream));
    if (host_ctrl[LZ_ERROR]) return set_err(c, USBS_ERR_TIMEOUT, "persistent Lanczos barrier wait expired");
  }
  double* d_vecs = (double*)ctx_alloc(c, pool, sizeof(double) * (size_t)P.n * nd);
  if (!d_vecs) return USBS_ERR_OOM;
  const long long tot = (long long)P.n * nd;
  k_copy_cols_out<T><<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(L.Q, L.ldq, P.n, nd, d_vecs);
  USBS_HIP(c, hipGetLastError());
  USBS_HIP(c, hipMemcpyAsync(eigvals, L.alpha, sizeof(double) * nd, hipMemcpyDeviceToHost, c->stream));
  USBS_HIP(c, hipMemcpyAsync(eigvecs, d_vecs, sizeof(double) * tot, hipMemcpyDeviceToHost, c->stream));
  USBS_HIP(c, hipMemcpyAsync(host_ctrl, L.ctrl, sizeof(host_ctrl), hipMemcpyDeviceToHost, c->stream));
  USBS_HIP(c, hipStreamSynchronize(c->stream));
  if (info) {
    info[0] = host_ctrl[LZ_RESTART];
    info[1] = host_ctrl[LZ_NCONV];
    info[2] = host_ctrl[LZ_STEPS];
    info[3] = host_ctrl[LZ_NBREAK];
  }
  return USBS_OK;
}

template <typename T>
int LzBatch<T>::build(Ctx* c0, std::vector<LzInst<T>>& L, int m_) {
  ctx_free_pool(pool);
  m = m_;
  ninst = (int)L.size();
  total_tiles = 0;
  for (auto& x : L) total_tiles += x.ntiles;
  // workgroups: one per tile with the LDS cache when every instance can
  // cache and all fit together, else proportional shares without it
  size_t cb = 0;
  bool all_cache = true;
  for (int q = 0; q < ninst; ++q) {
    L[q].nblocks = L[q].ntiles;
    const size_t b = lanczos_cache_bytes(L[q]);
    if (b == 0) all_cache = false;
    cb = std::max(cb, b);
  }
  if (!all_cache) cb = 0;
  int form = lanczos_tile_form(L[0]);
  for (auto& x : L) form = lanczos_launch_form(form, lanczos_tile_form(x));
  int cap = cb ? lanczos_max_coresident_blocks<T>(m, true, form, cb, c0->device)
               : lanczos_max_coresident_blocks<T>(m, false, form, lanczos_tile_lds_bytes(m), c0->device);
  if (cb && cap < total_tiles && form == LZ_FORM_RR) {
    // the three-wavefront copy of the kernel keeps every tile resident
    const int cap3 = lanczos_max_coresident_blocks<T>(m, true, LZ_FORM_RR3, cb, c0->device);
    if (cap3 >= total_tiles) {
      form = LZ_FORM_RR3;
      cap = cap3;
    }
  }
  if (cb && cap < total_tiles) {
    cb = 0;
    cap = lanczos_max_coresident_blocks<T>(m, false, form, lanczos_tile_lds_bytes(m), c0->device);
  }
  if (cap < ninst) return set_err(c0, USBS_ERR_INVALID, "batch does not fit one co-resident launch");
  total_wgs = 0;
  for (int q = 0; q < ninst; ++q) {
    int nb = L[q].ntiles;
    if (total_tiles > cap) nb = std::max(1, (int)((long long)L[q].ntiles * cap / total_tiles));
    L[q].nblocks = nb;
    total_wgs += nb;
  }
  if (total_wgs > cap) return set_err(c0, USBS_ERR_INVALID, "batch does not fit one co-resident launch");
  cache = cb;
  std::vector<int2> wgmap, tilemap;
  for (int q = 0; q < ninst; ++q) {
    for (int i = 0; i < L[q].nblocks; ++i) wgmap.push_back(make_int2(q, i));
    for (int i = 0; i < L[q].ntiles; ++i) tilemap.push_back(make_int2(q, i));
  }
  d_insts = (LzInst<T>*)ctx_alloc(c0, pool, sizeof(LzInst<T>) * ninst);
  d_wg = (int*)ctx_alloc(c0, pool, sizeof(int2) * wgmap.size());
  d_tm = (int*)ctx_alloc(c0, pool, sizeof(int2) * tilemap.size());
  if (!d_insts || !d_wg || !d_tm) return USBS_ERR_OOM;
  USBS_HIP(c0, hipMemcpy(d_insts, L.data(), sizeof(LzInst<T>) * ninst, hipMemcpyHostToDevice));
  USBS_HIP(c0, hipMemcpy(d_wg, wgmap.data(), sizeof(int2) * wgmap.size(), hipMemcpyHostToDevice));
  USBS_HIP(c0, hipMemcpy(d_tm, tilemap.data(), sizeof(int2) * tilemap.size(), hipMemcpyHostToDevice));
  tile_form = form;
  ctrls.assign(ninst, nullptr);
  for (int q = 0; q < ninst; ++q) ctrls[q] = L[q].ctrl;
  return USBS_OK;
}

template <typename T>
int LzBatch<T>::run(hipStream_t s, int max_restarts, int first_check, int* launched) {
  const LzStop stop{ctrls.data(), first_check, launched};
  return lanczos_persistent<T>(d_insts, ninst, total_wgs, m, max_restarts, s, d_wg, d_tm,
                               total_tiles, LZ_SPIN_LIMIT, cache, nullptr, tile_form, &stop);
}

template struct LzBatch<double>;
template struct LzBatch<float>;

// Several instances (one per context, each with its own uploaded problem) in
// one persistent launch per restart.  All contexts are on one device; the
// launch runs on the first context's stream.  Workgroups: one per tile of
// every instance (tiles cached in LDS) when they fit on the device together,
// else a proportional share per instance.
template <typename T>
static int run_lanczos_batch(Ctx** cs, int ninst, double C_coef, const double* const* ys,
                             const double* const* q0s, int nd, int m, int max_restarts, double tol,
                             double* const* eigvals, double* const* eigvecs, int* info) {
  Ctx* c0 = cs[0];
  hipStream_t s = c0->stream;
  std::vector<std::vector<DevBuf>> pools(ninst);
  std::vector<DevBuf> pool0;
  struct Guard {
    std::vector<std::vector<DevBuf>>& p;
    std::vector<DevBuf>& p0;
    ~Guard() { for (auto& q : p) ctx_free_pool(q); ctx_free_pool(p0); }
  } guard{pools, pool0};
  std::vector<LzInst<T>> L(ninst);
  int total_tiles = 0;
  for (int q = 0; q < ninst; ++q) {
    Ctx* c = cs[q];
    const Problem& P = c->P;
    double* d_y = upload(c, pools[q], ys[q], (size_t)P.m);
    double* d_q0 = upload(c, pools[q], q0s[q], (size_t)P.n);
    double* d_val64 = (double*)ctx_alloc(c, pools[q], sizeof(double) * P.nnz);
    float* d_val32 = nullptr;
    if (!d_y || !d_q0 || !d_val64) return USBS_ERR_OOM;
    if (sizeof(T) == 4 && !(d_val32 = (float*)ctx_alloc(c, pools[q], sizeof(float) * P.nnz)))
      return USBS_ERR_OOM;
    if (launch_assemble(P, C_coef, d_y, 1.0, nullptr, 0.0, nullptr, 0.0, d_val64, d_val32, s))
      return set_err(c0, USBS_ERR_HIP, "assemble launch failed");
    int rc = make_lanczos<T>(c, pools[q], P, m, nd, max_restarts, tol,
                             sizeof(T) == 4 ? (const T*)d_val32 : (const T*)d_val64, L[q]);
    if (rc) return rc;
    if (reset_lanczos(L[q], d_q0, s)) return set_err(c0, USBS_ERR_HIP, "reset launch failed");
    total_tiles += L[q].ntiles;
  }
  LzBatch<T> lb;
  int rc = lb.build(c0, L, m);
  if (rc) return rc;
  if (hipStreamSynchronize(s) != hipSuccess) return set_err(c0, USBS_ERR_HIP, "synchronise");
  rc = lb.run(s, max_restarts);
  if (rc) return set_err(c0, USBS_ERR_HIP, std::string("lanczos batch: ") + hipGetErrorString((hipError_t)rc));
  for (int q = 0; q < ninst; ++q) {
    const Problem& P = cs[q]->P;
    double* d_vecs = (double*)ctx_alloc(c0, pool0, sizeof(double) * (size_t)P.n * nd);
    if (!d_vecs) return USBS_ERR_OOM;
    const long long tot = (long long)P.n * nd;
    k_copy_cols_out<T><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(L[q].Q, L[q].ldq, P.n, nd, d_vecs);
    USBS_HIP(c0, hipGetLastError());
    int hc[LZ_NCTRL];
    USBS_HIP(c0, hipMemcpyAsync(eigvals[q], L[q].alpha, sizeof(double) * nd, hipMemcpyDeviceToHost, s));
    USBS_HIP(c0, hipMemcpyAsync(eigvecs[q], d_vecs, sizeof(double) * tot, hipMemcpyDeviceToHost, s));
    USBS_HIP(c0, hipMemcpyAsync(hc, L[q].ctrl, sizeof(hc), hipMemcpyDeviceToHost, s));
    USBS_HIP(c0, hipStreamSynchronize(s));
    if (hc[LZ_ERROR]) return set_err(c0, USBS_ERR_TIMEOUT, "persistent Lanczos barrier wait expired");
    if (info) {
      info[4 * q + 0] = hc[LZ_RESTART];
      info[4 * q + 1] = hc[LZ_NCONV];
      info[4 * q + 2] = hc[LZ_STEPS];
      info[4 * q + 3] = hc[LZ_NBREAK];
    }
  }
  return USBS_OK;
}

}  // namespace usbs

using namespace usbs;

extern "C" {

const char* usbs_version(void) { return "usbs-hip 0.1 (gfx950)"; }

int usbs_create(usbs_ctx** out, int device) {
  if (!out) return USBS_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return USBS_ERR_HIP;
  if (device < 0 || device >= ndev) return USBS_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return USBS_ERR_HIP;
  if (noalloc_init_device(device) != USBS_OK) return USBS_ERR_HIP;
  usbs_ctx* h = new usbs_ctx();
  h->c.device = device;
  if (hipStreamCreateWithFlags(&h->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return USBS_ERR_HIP;
  }
  if (const char* e = getenv("USBS_PERSIST_WGS")) h->c.persist_wgs = atoi(e);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    h->c.num_cus = cus;
  *out = h;
  return USBS_OK;
}

void usbs_destroy(usbs_ctx* h) {
  if (!h) return;
  (void)hipSetDevice(h->c.device);
  (void)hipStreamSynchronize(h->c.stream);
  ctx_free_pool(h->c.work_bufs);
  ctx_free_pool(h->c.prob_bufs);
  ctx_free_pool(h->c.keep_bufs);
  ctx_release_cache(&h->c);
  if (h->c.side) {
    (void)hipStreamSynchronize(h->c.side);
    (void)hipStreamDestroy(h->c.side);
  }
  if (h->c.rec_h) (void)hipHostFree(h->c.rec_h);
  if (h->c.ctrl_h) (void)hipHostFree(h->c.ctrl_h);
  (void)hipStreamDestroy(h->c.stream);
  delete h;
}

int usbs_device_memory(usbs_ctx* h, int64_t* free_bytes, int64_t* total_bytes) {
  if (!h || !free_bytes || !total_bytes) return USBS_ERR_INVALID;
  if (hipSetDevice(h->c.device) != hipSuccess) return set_err(&h->c, USBS_ERR_HIP, "hipSetDevice");
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return set_err(&h->c, USBS_ERR_HIP, "hipMemGetInfo");
  *free_bytes = (int64_t)f;
  *total_bytes = (int64_t)t;
  return USBS_OK;
}

const char* usbs_last_error(const usbs_ctx* h) { return h ? h->c.err.c_str() : "null context"; }

int usbs_set_problem(usbs_ctx* h, int64_t n, int64_t m, int64_t nnzC, const int32_t* C_rows,
                     const int32_t* C_cols, const double* C_vals, int64_t nA,
                     const int32_t* A_indices, const double* A_data, const double* b,
                     const double* b_ineq_mask, const double* SCALE_A) {
  if (!h) return USBS_ERR_INVALID;
  if ((nnzC > 0 && (!C_rows || !C_cols || !C_vals)) || (nA > 0 && (!A_indices || !A_data)) || !b ||
      !b_ineq_mask || !SCALE_A)
    return set_err(&h->c, USBS_ERR_INVALID, "null array argument");
  if (hipSetDevice(h->c.device) != hipSuccess) return set_err(&h->c, USBS_ERR_HIP, "hipSetDevice");
  return build_problem(&h->c, n, m, nnzC, C_rows, C_cols, C_vals, nA, A_indices, A_data, b,
                       b_ineq_mask, SCALE_A);
}

int usbs_lanczos_smallest(usbs_ctx* h, double C_coef, const double* y, const double* q0,
                          int num_desired, int inner_iterations, int max_restarts, double tolerance,
                          int precision, int exec_mode, double* eigvals, double* eigvecs, int* info) {
  if (!h) return USBS_ERR_INVALID;
  Ctx* c = &h->c;
  if (!c->has_problem) return set_err(c, USBS_ERR_STATE, "no problem uploaded");
  if (!y || !q0 || !eigvals || !eigvecs) return set_err(c, USBS_ERR_INVALID, "null array argument");
  const int m = inner_iterations;
  if (m < 1 || m > 48 || m > c->P.n)
    return set_err(c, USBS_ERR_INVALID, "inner_iterations must be in [1, min(n, 48)]");
  if (num_desired < 1 || num_desired > m)
    return set_err(c, USBS_ERR_INVALID, "num_desired must be in [1, inner_iterations]");
  if (max_restarts < 1) return set_err(c, USBS_ERR_INVALID, "max_restarts must be >= 1");
  if (precision != USBS_PREC_FP64 && precision != USBS_PREC_MIXED)
    return set_err(c, USBS_ERR_INVALID, "unknown precision");
  if (exec_mode == USBS_EXEC_AUTO || exec_mode == USBS_EXEC_TUNE)
    exec_mode = resolve_exec_mode(c->P, inner_iterations, precision, c->num_cus);
  if (exec_mode != USBS_EXEC_PEROP && exec_mode != USBS_EXEC_PERSISTENT)
    return set_err(c, USBS_ERR_INVALID, "unknown exec_mode");
  for (int i = 0; i < c->P.m; ++i)
    if (!std::isfinite(y[i])) return set_err(c, USBS_ERR_NONFINITE, "non-finite adjoint_left_vec");
  for (int i = 0; i < c->P.n; ++i)
    if (!std::isfinite(q0[i])) return set_err(c, USBS_ERR_NONFINITE, "non-finite q0");
  CoopGuard coop(exec_mode == USBS_EXEC_PERSISTENT ? (const void*)c : nullptr);
  if (exec_mode == USBS_EXEC_PERSISTENT && !coop.ok()) return set_err(c, USBS_ERR_STATE, COOP_BUSY_MSG);
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, USBS_ERR_HIP, "hipSetDevice");
  if (precision == USBS_PREC_FP64)
    return run_lanczos_standalone<double>(c, C_coef, y, q0, num_desired, m, max_restarts, tolerance,
                                          exec_mode, eigvals, eigvecs, info);
  return run_lanczos_standalone<float>(c, C_coef, y, q0, num_desired, m, max_restarts, tolerance,
                                       exec_mode, eigvals, eigvecs, info);
}

int usbs_lanczos_smallest_batch(usbs_ctx* const* ctxs, int ninst, double C_coef,
                                const double* const* adjoint_left_vecs, const double* const* q0s,
                                int num_desired, int inner_iterations, int max_restarts,
                                double tolerance, int precision, double* const* eigvals,
                                double* const* eigvecs, int* info) {
  if (!ctxs || ninst < 1 || !ctxs[0] || !adjoint_left_vecs || !q0s || !eigvals || !eigvecs)
    return USBS_ERR_INVALID;
  Ctx* c0 = &ctxs[0]->c;
  std::vector<Ctx*> cs(ninst);
  for (int q = 0; q < ninst; ++q) {
    if (!ctxs[q] || !adjoint_left_vecs[q] || !q0s[q] || !eigvals[q] || !eigvecs[q])
      return set_err(c0, USBS_ERR_INVALID, "null argument in batch entry " + std::to_string(q));
    cs[q] = &ctxs[q]->c;
    if (!cs[q]->has_problem) return set_err(c0, USBS_ERR_STATE, "batch entry without a problem");
    if (cs[q]->device != c0->device) return set_err(c0, USBS_ERR_INVALID, "batch spans devices");
    const int n = cs[q]->P.n;
    if (num_desired < 1 || inner_iterations < num_desired || inner_iterations > 48 ||
        inner_iterations > n || max_restarts < 1)
      return set_err(c0, USBS_ERR_INVALID, "need 1 <= num_desired <= inner_iterations <= min(n, 48)");
    for (int i = 0; i < cs[q]->P.m; ++i)
      if (!std::isfinite(adjoint_left_vecs[q][i])) return set_err(c0, USBS_ERR_NONFINITE, "non-finite adjoint_left_vec");
  }
  if (precision != USBS_PREC_FP64 && precision != USBS_PREC_MIXED)
    return set_err(c0, USBS_ERR_INVALID, "unknown precision");
  CoopGuard coop(c0);
  if (!coop.ok()) return set_err(c0, USBS_ERR_STATE, COOP_BUSY_MSG);
  if (hipSetDevice(c0->device) != hipSuccess) return set_err(c0, USBS_ERR_HIP, "hipSetDevice");
  for (int q = 1; q < ninst; ++q)  // work queued on other contexts' streams finishes first
    if (hipStreamSynchronize(cs[q]->stream) != hipSuccess) return set_err(c0, USBS_ERR_HIP, "synchronise");
  if (precision == USBS_PREC_FP64)
    return run_lanczos_batch<double>(cs.data(), ninst, C_coef, adjoint_left_vecs, q0s, num_desired,
                                     inner_iterations, max_restarts, tolerance, eigvals, eigvecs, info);
  return run_lanczos_batch<float>(cs.data(), ninst, C_coef, adjoint_left_vecs, q0s, num_desired,
                                  inner_iterations, max_restarts, tolerance, eigvals, eigvecs, info);
}

}  // extern "C"
