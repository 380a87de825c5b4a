// Thick-restart Lanczos on the GPU: shared declarations.
//
// One Lanczos problem ("instance") is described by LzInst: the operator in
// CSR (values assembled for C_arg + A*(v) beforehand), the basis Q
// (n x (m+1), column-major), the work vector, tile partials and the small
// FP64 state (alpha, beta, Ritz rotation, control words).
#pragma once
#include "dev_common.hpp"

namespace usbs {

enum LzCtrl {
  LZ_K = 0,          // first step of the next restart (= number of saved Ritz vectors)
  LZ_RESTART = 1,    // restarts completed
  LZ_NCONV = 2,      // num_converged of the last restart
  LZ_DONE = 3,       // 1 once num_converged >= num_desired or max_restarts reached
  LZ_BREAKDOWN = 4,  // 1 if the current step hit beta < 1e-15
  LZ_ERROR = 5,      // nonzero: a bounded wait expired (persistent launch)
  LZ_NBREAK = 6,     // number of breakdowns seen
  LZ_STEPS = 7,      // Lanczos steps taken (operator applications)
  LZ_PENDING = 8,    // per-op: step i >= 1 met a breakdown in its phase A; the
                     // restart's later kernels return at once until the host
                     // has built the breakdown vector and resumed at step i
  LZ_NCTRL = 9
};

template <typename T>
struct LzInst {
  int n, tile, ntiles, m, nd, max_restarts, pstride, nblocks;
  double tol;
  double rtol;   // Ritz residual floor relative to the largest |Ritz value|:
                 // a pair counts as converged when its residual estimate is
                 // below max(tol, rtol * max|theta|) (0: tol alone)
  const int* row_ptr;
  const int* col;
  const T* val;
  T* Q;          // n x (m+1), leading dimension ldq
  long long ldq;
  T* w;          // 2 x ldq: residual / work vectors of alternate steps
  double* part1; // ntiles x pstride, one array per reduction pass
  double* part2;
  double* part3;
  double* h1;    // m+1: first CGS coefficients of the current step
  double* h2;    // m+1: second CGS coefficients
  double* alpha; // m
  double* beta;  // m
  double* Yr;    // m x m row-major: saved Ritz vectors (columns) of T
  int* ctrl;     // LZ_NCTRL
  unsigned* bar; // barrier counters of the persistent launches: nbar regions
                 // of LZ_BAR_STRIDE, region r for restart r
  int* lstat;    // 2 x nbar: per persistent launch r, steps done and the sum
                 // over those steps of the valid column count i+1
  int nbar;      // number of counter regions in bar (>= max_restarts)
  int tile_nnz_max;  // largest number of operator entries in one row tile
  int heavy;         // 1: a wavefront per operator row in the SpMV (rows of
                     // 32+ entries on average), else a thread per row
  int exact_norm;    // 1: on light-row (RR) tiles every step takes |w| from
                     // its own reduction (the bundle QR, whose columns must
                     // be unit to storage precision); 0: |w| may come from
                     // the second CGS pass (|y|^2 - |h2|^2).  Also keeps the
                     // struct free of padding bytes (it is compared bytewise
                     // as the key of the captured per-op graphs)
};
static_assert(sizeof(LzInst<double>) == 8 * 4 + 8 * 2 + 8 * 17 + 4 * 4, "LzInst has padding");

// Tile forms of the step kernels: column-wise reductions (tiles of at most
// 256 rows with heavy rows), row-wise reductions (tiles above 256 rows), rows
// of Q in registers (light rows, tiles of at most 256 rows).  The persistent
// kernel for LZ_FORM_RR carries only that code; LZ_FORM_RR3 is the same code
// compiled for three wavefronts per SIMD (some registers spill: slower per
// step, but three workgroups share a compute unit), for batches that would
// not be co-resident otherwise.
#define LZ_FORM_COL 0
#define LZ_FORM_ROW 1
#define LZ_FORM_RR 2
#define LZ_FORM_RR3 3

// Largest LDS block of a persistent workgroup that caches its tile.
#define LZ_CACHE_MAX_BYTES (64 * 1024)
// Barrier counters of one persistent launch (restart): a region of
// LZ_BAR_STRIDE unsigned per restart, which holds grid_barrier's counter at
// offset 0 or coh_barrier's eight group counters at 0, 64, ..., 448.
#define LZ_BAR_STRIDE 512
// Polls before a device-wide barrier wait gives up and records an error
// (LZ_ERROR, reported as USBS_ERR_TIMEOUT): a few seconds, far above any
// legitimate wait inside one restart.
#define LZ_SPIN_LIMIT (1ull << 21)

int tile_rows_for(int n);
// Row tiles of the Lanczos steps for heavy-row operators (n <= 65536):
// 16 rows, doubled until there are at most 256 tiles.
int lanczos_heavy_tile_rows(int n);

// Early stop of the persistent form: after first_check restarts, and at each
// doubling of that count, the host reads the control words of every instance
// (ctrls[q]: device pointers, one per instance) and launches no further
// restart once all are done or failed.  launched (optional) receives the
// number of restarts launched.
struct LzStop {
  int* const* ctrls;
  int first_check;
  int* launched;
};

// Host launchers (lanczos.hip).  All return hipError_t as int.
// Measured time of the step kernels of per-op restarts (HIP events around
// each restart's step launches, launch gaps included).
struct LzTiming {
  hipEvent_t e0, e1;
  double ms;
  long long launches, steps, sum_cols;
};
// Per-op restarts replayed from HIP graphs: the step launches of a restart
// that starts at step k are captured once per k (for one LzInst) and then
// launched as one graph.  Owned by the caller; lanczos_graphs_clear frees it.
struct LzGraphs {
  static constexpr int N = 49;
  hipGraphExec_t exec[N] = {};
  unsigned char key[256] = {};  // bytes of the LzInst the graphs were captured for
  int key_len = 0;
};
void lanczos_graphs_clear(LzGraphs* g);
template <typename T>
int lanczos_perop(const LzInst<T>& L, hipStream_t s, int* host_ctrl, LzTiming* tm = nullptr,
                  LzGraphs* graphs = nullptr);
// Persistent form for ninst instances sharing the basis size m: per restart
// one cooperative launch of all Lanczos steps (dev_wgmap: int2 per workgroup
// = instance, local workgroup), then the Ritz kernel (one block per
// instance) and the rotation (dev_tilemap: int2 per tile = instance, tile).
// evs (optional, 2 x max_restarts events) brackets each cooperative launch;
// cache_bytes > 0 selects the form that keeps each workgroup's tile in LDS
// (from lanczos_plan_persistent).
template <typename T>
int lanczos_persistent(const LzInst<T>* dev_insts, int ninst, int total_wgs, int m,
                       int max_restarts, hipStream_t s, const int* dev_wgmap,
                       const int* dev_tilemap, int total_tiles, unsigned long long spin_limit,
                       size_t cache_bytes, hipEvent_t* evs = nullptr, int form = LZ_FORM_ROW,
                       const LzStop* stop = nullptr);
// Workgroup count (L.nblocks) and LDS cache size for a persistent launch of
// one instance: one workgroup per tile unless want_wgs > 0 asks for fewer,
// capped by what can be co-resident on the device.
template <typename T>
int lanczos_plan_persistent(LzInst<T>& L, int device, int want_wgs, size_t* cache_bytes);
template <typename T>
size_t lanczos_cache_bytes(const LzInst<T>& L);
// 1 when the persistent form would keep every tile in LDS for a restart (the
// case where it beats per-op launches): light rows, tiles of at most 256
// rows, cache within LZ_CACHE_MAX_BYTES.
int lanczos_prefers_persistent(int tile, int heavy, int m, int tile_nnz_max, int value_bytes);
// Workgroups of the persistent step kernel for (T, m, cache, form) that
// fit on the device at once with lds_bytes of LDS each.
template <typename T>
int lanczos_max_coresident_blocks(int m, bool cache, int form, size_t lds_bytes, int device);
// The tile form of an instance (LZ_FORM_*), and of a launch holding several:
// row-wise if any instance is, row-in-registers only if every one is.
template <typename T>
int lanczos_tile_form(const LzInst<T>& L);
int lanczos_launch_form(int form_a, int form_b);
// LDS of a persistent workgroup that does not cache its tile (basis size m).
size_t lanczos_tile_lds_bytes(int m);
template <typename T>
int orth_perop(const LzInst<T>& L, hipStream_t s);
// The same QR as one cooperative launch of one workgroup per tile
// (dev_inst->nblocks = ntiles, dev_wgmap the identity map over the tiles),
// for instances orth_persistent_fits accepts: light rows in tiles of at most
// 256 rows, every tile's workgroup co-resident on the device.
template <typename T>
int orth_persistent_fits(const LzInst<T>& L, int device);
template <typename T>
int orth_persistent(const LzInst<T>* dev_inst, int ntiles, int m, hipStream_t s, const int* dev_wgmap);

}  // namespace usbs
