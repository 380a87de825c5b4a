// Host-side engine internals: device context, resident problem data and
// workspaces.  Everything here sits behind the C ABI in include/usbs_hip.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <map>
#include <vector>

#include "../../include/usbs_hip.h"
#include "coop_guard.hpp"
#include "lanczos.hpp"

#define COOP_BUSY_MSG                                                                      \
  "a persistent (cooperative) Lanczos launch is in flight on another context of this "   \
  "process; concurrent solves must use exec_mode per-op"

namespace usbs {

// Device buffers a context has finished with, kept for its later
// allocations: a solve's ~45 workspaces then cost no hipMalloc / hipFree
// (hipFree waits for the whole device, stalling every other context's work)
// once the context has seen its sizes.  Blocks return here only after the
// context's stream has drained; work a block saw on another stream (a
// solver's side stream, a batch stream) is waited for by its owner before
// the block is released.  Freed for good by usbs_destroy, or when a
// hipMalloc fails.
struct BufCache {
  hipStream_t stream = nullptr;           // the owning context's stream
  std::multimap<size_t, void*> blocks;    // size -> block
  size_t bytes = 0;
};

// A device buffer owned by a context (cache: where it returns when freed;
// null: hipFree).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  BufCache* cache = nullptr;
};

// The operator pattern merges C's pattern with the (i, j) pattern of the
// constraint triples, so that C_coef * C + A*(z) is one CSR matrix whose
// values are assembled per call.
struct Problem {
  int n = 0, m = 0;
  long long nnz = 0, nA = 0;
  int tile = 0, ntiles = 0;     // row tiles of n
  // Lanczos row tiles: the same as tile, except for operators whose rows
  // average 32 or more entries (heavy rows, a wavefront per row in the SpMV),
  // which get tiles of 16+ rows so that enough workgroups share the work.
  int lz_tile = 0, lz_ntiles = 0;
  int lz_heavy = 0;
  int tile_nnz_max = 0;         // most operator entries in one Lanczos row tile
  int mtile = 0, mntiles = 0;   // tiles over the m constraints (256: a thread per constraint)
  bool diag_A = false;          // A = {(i, i, i) with data 1}: max-cut's diag(X) = 1
  bool has_ineq = false;        // any b_ineq_mask != 0
  bool cut_integer = false;     // every off-diagonal 4*C_ij is an integer
  int* row_ptr = nullptr;       // n+1
  int* col = nullptr;           // nnz
  double* cval = nullptr;       // nnz: C in the merged pattern (0 where C has no entry)
  int* cw4 = nullptr;           // nnz: round(4*C) when cut_integer
  std::vector<double> cval_host;  // C in the merged pattern (host copy: cut weights of C / s)
  int* slot_ptr = nullptr;      // nnz+1: triples per pattern slot
  int* slot_trip = nullptr;     // nA
  int* tcon = nullptr;          // nA: triple constraint index
  int* ti = nullptr;            // nA
  int* tj = nullptr;            // nA
  double* tdata = nullptr;      // nA
  int* con_ptr = nullptr;       // m+1: triples per constraint
  int* ct_i = nullptr;          // nA: the triples' (i, j, data) sorted by
  int* ct_j = nullptr;          //     constraint (input order within one), so
  double* ct_d = nullptr;       //     per-constraint loops read them directly
  double* b = nullptr;          // m
  double* mask = nullptr;       // m
  double* scale_a = nullptr;    // m
};

int lanczos_prefers_persistent(int tile, int heavy, int m, int tile_nnz_max, int value_bytes);
// USBS_EXEC_AUTO: the persistent form where it keeps each tile in LDS for the
// whole restart (light rows, small tiles) and the tiles fit one workgroup per
// compute unit, per-op launches otherwise (heavy rows, tiles streamed from
// HBM, or more tiles than compute units: there the device-wide barriers cost
// more than the launches of graph-replayed restarts).  m: Lanczos basis size;
// precision: USBS_PREC_*; num_cus: compute units of the device.
inline int resolve_exec_mode(const Problem& P, int m, int precision, int num_cus) {
  if (P.lz_ntiles > num_cus) return 0;
  return lanczos_prefers_persistent(P.lz_tile, P.lz_heavy, m, P.tile_nnz_max,
                                    precision == 1 ? 4 : 8);
}

// Row tiles of the objective's partials: the Lanczos tiles for heavy rows.
inline int objective_ntiles(const Problem& P) { return P.lz_heavy ? P.lz_ntiles : P.ntiles; }

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool has_problem = false;
  Problem P;
  BufCache cache;                 // freed buffers for reuse (see BufCache)
  std::vector<DevBuf> prob_bufs;  // freed on new problem / destroy
  std::vector<DevBuf> work_bufs;  // solve workspaces
  int num_cus = 256;             // compute units of the device
  int persist_wgs = 0;            // workgroups per persistent Lanczos instance (0 = one per tile)
  int tuned_mode = -1;            // USBS_EXEC_TUNE: the form measured faster for this problem
  int tuned_m = 0, tuned_prec = -1;  // ... at this Lanczos basis size and precision
  float tune_ms[2] = {0.f, 0.f};     // the measured first calls: per-op, persistent
  bool timing = false;            // bracket persistent Lanczos launches with events
  // y, z, P (column-major n x r, ld = keep_ld) of the last usbs_solve, kept
  // on the device for warm starts; survive usbs_set_problem while n, m agree
  std::vector<DevBuf> keep_bufs;
  double *keep_y = nullptr, *keep_z = nullptr, *keep_P = nullptr;
  int keep_n = 0, keep_m = 0, keep_r = -1;
  long long keep_ld = 0;
  double keep_sx = 1.0;           // SCALE_X of the solve that left them
  // QAP rounding data (usbs_set_qap_rounding): flow W and distance D, l x l
  int qap_l = 0;
  std::vector<double> qap_W, qap_D;
  usbs_stats stats{};
  // kept for the context's solves, created on first use: the side stream of
  // the lower-bound IPM, and pinned host buffers for the records (scalar
  // block, Lanczos control words)
  hipStream_t side = nullptr;
  double* rec_h = nullptr;
  int* ctrl_h = nullptr;
};

// error helpers
int set_err(Ctx* c, int code, const std::string& msg);
#define USBS_HIP(ctx, expr)                                                               \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return ::usbs::set_err(ctx, -2, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

void* ctx_alloc(Ctx* c, std::vector<DevBuf>& pool, size_t bytes);
void ctx_free_pool(std::vector<DevBuf>& pool);
// hipFree every cached block of the context.
void ctx_release_cache(Ctx* c);

template <typename T>
int make_lanczos(Ctx* c, std::vector<DevBuf>& pool, const Problem& P, int m, int nd,
                 int max_restarts, double tol, const T* opval, LzInst<T>& L);

// noalloc.hip: look up this device's copy of the fixed scratch array of
// usbs_sketch_update_noalloc (once per device, with the device current).
int noalloc_init_device(int device);

// Several Lanczos instances (possibly of different sizes, one device) advanced
// together by one persistent (cooperative) launch per restart: build() plans
// the workgroups -- one per row tile with the tile cached in LDS when all fit
// on the device at once, else proportional shares -- and uploads the
// descriptors (the instances' nblocks are set); run() issues the restarts on
// stream s, reading the instances' control words after first_check restarts
// and at each doubling to stop once all are done.
template <typename T>
struct LzBatch {
  std::vector<DevBuf> pool;
  LzInst<T>* d_insts = nullptr;
  int* d_wg = nullptr;
  int* d_tm = nullptr;
  int ninst = 0, m = 0, total_wgs = 0, total_tiles = 0, tile_form = 0;
  size_t cache = 0;
  std::vector<int*> ctrls;
  ~LzBatch() { ctx_free_pool(pool); }
  int build(Ctx* c0, std::vector<LzInst<T>>& L, int m);
  int run(hipStream_t s, int max_restarts, int first_check = 2, int* launched = nullptr);
};

// kernels in ops.hip
int launch_assemble(const Problem& P, double coefC, const double* z1, double a1, const double* z2,
                    double a2, const double* z3, double a3, double* out64, float* out32,
                    hipStream_t s);

}  // namespace usbs

// The opaque handle of the C ABI.
struct usbs_ctx {
  usbs::Ctx c;
};
