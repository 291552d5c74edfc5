// Batched mixed-precision solve of many independent systems (ligand poses,
// titration states, serving requests) with per-system CG state in HBM.
//
// Each system keeps its own context (mesh, markers, assembled FP64 CSR, compact FP32
// operator, solver vectors). The batch holds a device table of per-slot descriptors
// (pointers into those contexts) and per-slot scalars, and a tile table: tile =
// (slot, first slice) covering `tile_slices` consecutive 64-row slices of one system.
// Every row kernel runs one block per tile over all systems at once; a system's
// phase decides which kernels touch its rows:
//
//   JOIN     admitted, first residual r = b pending
//   ACTIVE   inner FP32 CG (Jacobi-scaled operator, as in the single-system solver)
//   REPLACE  inner cycle finished: x += S xhat, FP64 residual r = b - A x pending
//   DONE     ||b - A x|| <= tol ||b|| in FP64 (or iteration limit / stagnation)
//
// One replay of the captured hipGraph is: the replacement kernels (accumulate, FP64
// CSR SpMV, residual, per-system finalize) and the cycle-start direction, then R inner
// iterations of five kernels (SpMV with (p,q), alpha, residual update with (r,r),
// beta/stop, x and p update). Dot products are written per slice (a fixed wave
// reduction) and summed per system by one block in a fixed order, so a system's
// arithmetic does not depend on the tile size, the block shape or the other systems
// in the batch: a system solved alone and in any mix gives bit-identical results as
// long as it joins at a replay boundary, which is the only place the host admits
// systems.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ngpb_internal.h"

namespace ngpb {

void build_compact(ngpb_ctx* c, bool allow_band);
void check_operator(ngpb_ctx* c);

namespace {

constexpr int kW = 64;
constexpr int kFinBS = 256;

#include "ngpb_sell.inc"

enum Phase : int { kFree = 0, kJoin = 1, kActive = 2, kReplace = 3, kDone = 4 };

struct BatchDesc {
  Sell A;
  const int32_t* rowptr;
  const int32_t* col;
  const double* val;
  const double* b;
  const double* dscale;
  double* x;
  double* t;
  float* xs;
  float* rs;
  float* ps;
  float* qs;
  double* part;  // 2 per slice
  int64_t n, nslices;
};

struct BatchScal {
  double rho, alpha, beta, rr, stop_rr, bb, target, true_rr, prev_true, ss;
  double tol, drop;
  int phase, stepped, cycle_start, iters, maxiter, repl, stagnant, status;
  int inner_at_join;  // reserved
  int pad;
};

struct Tile {
  int32_t slot;
  int32_t slice;
};

// ---------------------------------------------------------------- row kernels
// One block per tile; wave w of the block takes slices first + w, first + w + WPB, ...

// At most 72 VGPRs (7 waves per SIMD; the compiler's own choice was 101 and 4 waves):
// 128 ligands 0.565 -> 0.421 ms per SpMV over every system (5 waves: 0.492, 6: 0.435,
// 8: 0.448 ms with 24 spilled VGPRs). The few values spilled at 7 waves belong to the
// widest slice forms.
template <int WPB>
__global__ void __launch_bounds__(WPB * 64) __attribute__((amdgpu_waves_per_eu(7))) kb_spmv(const BatchDesc* __restrict__ D, const BatchScal* __restrict__ S,
                                                    const Tile* __restrict__ tiles, int tile_slices, int force) {
  const Tile tl = tiles[blockIdx.x];
  const int slot = tl.slot;
  if (!force && S[slot].phase != kActive) return;
  const BatchDesc& d = D[slot];
  const int64_t s1 = min((int64_t)tl.slice + tile_slices, d.nslices);
  const int wave = threadIdx.x >> 6;
  for (int64_t s = (int64_t)tl.slice + wave; s < s1; s += WPB) {
    const int w = d.A.width[s];
    const int64_t bp = d.A.base_ptr[s];
    double part = 0.0;
    one_slice(d.A, s, w, bp, d.ps, d.qs, &part);
    part = dev::wave_sum(part);
    if (lane_id() == 0) d.part[2 * s] = part;
  }
}

template <int WPB>
__global__ void __launch_bounds__(WPB * 64) kb_update(const BatchDesc* __restrict__ D, const BatchScal* __restrict__ S,
                                                      const Tile* __restrict__ tiles, int tile_slices) {
  const Tile tl = tiles[blockIdx.x];
  const int slot = tl.slot;
  if (S[slot].phase != kActive) return;
  const BatchDesc& d = D[slot];
  const float alpha = (float)S[slot].alpha;
  const int64_t s1 = min((int64_t)tl.slice + tile_slices, d.nslices);
  const int wave = threadIdx.x >> 6;
  for (int64_t s = (int64_t)tl.slice + wave; s < s1; s += WPB) {
    const int64_t i = s * kW + lane_id();
    double rr = 0.0;
    if (i < d.n) {
      const float ri = fmaf(-alpha, d.qs[i], d.rs[i]);
      d.rs[i] = ri;
      rr = (double)ri * (double)ri;
    }
    rr = dev::wave_sum(rr);
    if (lane_id() == 0) d.part[2 * s] = rr;
  }
}

// xhat += alpha p for a system that took a step; p = rhat + beta p if it continues
template <int WPB>
__global__ void __launch_bounds__(WPB * 64) kb_dir_x(const BatchDesc* __restrict__ D, const BatchScal* __restrict__ S,
                                                     const Tile* __restrict__ tiles, int tile_slices) {
  const Tile tl = tiles[blockIdx.x];
  const int slot = tl.slot;
  if (!S[slot].stepped) return;
  const BatchDesc& d = D[slot];
  const bool cont = S[slot].phase == kActive;
  const float alpha = (float)S[slot].alpha;
  const float beta = (float)S[slot].beta;
  const int64_t s1 = min((int64_t)tl.slice + tile_slices, d.nslices);
  const int wave = threadIdx.x >> 6;
  for (int64_t s = (int64_t)tl.slice + wave; s < s1; s += WPB) {
    const int64_t i = s * kW + lane_id();
    if (i >= d.n) continue;
    const float pi = d.ps[i];
    d.xs[i] = fmaf(alpha, pi, d.xs[i]);
    if (cont) d.ps[i] = fmaf(beta, pi, d.rs[i]);
  }
}

// first direction of a cycle: p = rhat + beta p
template <int WPB>
__global__ void __launch_bounds__(WPB * 64) kb_dir(const BatchDesc* __restrict__ D, const BatchScal* __restrict__ S,
                                                   const Tile* __restrict__ tiles, int tile_slices) {
  const Tile tl = tiles[blockIdx.x];
  const int slot = tl.slot;
  if (!S[slot].cycle_start) return;
  const BatchDesc& d = D[slot];
  const float beta = (float)S[slot].beta;
  const int64_t s1 = min((int64_t)tl.slice + tile_slices, d.nslices);
  const int wave = threadIdx.x >> 6;
  for (int64_t s = (int64_t)tl.slice + wave; s < s1; s += WPB) {
    const int64_t i = s * kW + lane_id();
    if (i < d.n) d.ps[i] = fmaf(beta, d.ps[i], d.rs[i]);
  }
}

// x += S xhat, xhat = 0 (systems whose inner cycle ended)
template <int WPB>
__global__ void __launch_bounds__(WPB * 64) kb_accum(const BatchDesc* __restrict__ D, const BatchScal* __restrict__ S,
                                                     const Tile* __restrict__ tiles, int tile_slices) {
  const Tile tl = tiles[blockIdx.x];
  const int slot = tl.slot;
  if (S[slot].phase != kReplace) return;
  const BatchDesc& d = D[slot];
  const int64_t s1 = min((int64_t)tl.slice + tile_slices, d.nslices);
  const int wave = threadIdx.x >> 6;
  for (int64_t s = (int64_t)tl.slice + wave; s < s1; s += WPB) {
    const int64_t i = s * kW + lane_id();
    if (i >= d.n) continue;
    d.x[i] += d.dscale[i] * (double)d.xs[i];
    d.xs[i] = 0.f;
  }
}

// t = A x with the FP64 CSR (row per lane, entries in column order, fma chain)
template <int WPB>
__global__ void __launch_bounds__(WPB * 64) kb_spmv64(const BatchDesc* __restrict__ D, const BatchScal* __restrict__ S,
                                                      const Tile* __restrict__ tiles, int tile_slices) {
  const Tile tl = tiles[blockIdx.x];
  const int slot = tl.slot;
  if (S[slot].phase != kReplace) return;
  const BatchDesc& d = D[slot];
  const int64_t s1 = min((int64_t)tl.slice + tile_slices, d.nslices);
  const int wave = threadIdx.x >> 6;
  for (int64_t s = (int64_t)tl.slice + wave; s < s1; s += WPB) {
    const int64_t i = s * kW + lane_id();
    if (i >= d.n) continue;
    double acc = 0.0;
    const int32_t e1 = d.rowptr[i + 1];
    for (int32_t k = d.rowptr[i]; k < e1; ++k) acc = fma(d.val[k], d.x[d.col[k]], acc);
    d.t[i] = acc;
  }
}

// r = b - t (JOIN: r = b, x = xhat = p = 0); rhat = S r; per slice (r, r) and (rhat, rhat)
template <int WPB>
__global__ void __launch_bounds__(WPB * 64) kb_resid(const BatchDesc* __restrict__ D, const BatchScal* __restrict__ S,
                                                     const Tile* __restrict__ tiles, int tile_slices) {
  const Tile tl = tiles[blockIdx.x];
  const int slot = tl.slot;
  const int ph = S[slot].phase;
  if (ph != kReplace && ph != kJoin) return;
  const bool first = ph == kJoin;
  const BatchDesc& d = D[slot];
  const int64_t s1 = min((int64_t)tl.slice + tile_slices, d.nslices);
  const int wave = threadIdx.x >> 6;
  for (int64_t s = (int64_t)tl.slice + wave; s < s1; s += WPB) {
    const int64_t i = s * kW + lane_id();
    double rr = 0.0, ss = 0.0;
    if (i < d.n) {
      double ri;
      if (first) {
        ri = d.b[i];
        d.x[i] = 0.0;
        d.xs[i] = 0.f;
        d.ps[i] = 0.f;
      } else {
        ri = d.b[i] - d.t[i];
      }
      const float si = (float)(d.dscale[i] * ri);
      d.rs[i] = si;
      rr = ri * ri;
      ss = (double)si * (double)si;
    }
    rr = dev::wave_sum(rr);
    ss = dev::wave_sum(ss);
    if (lane_id() == 0) {
      d.part[2 * s] = rr;
      d.part[2 * s + 1] = ss;
    }
  }
}

// ---------------------------------------------------------------- per-system finalize
// One block per slot; fixed-order sums of the slice partials.

__device__ __forceinline__ double slot_sum(const BatchDesc& d, int k, double* lds) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < d.nslices; i += kFinBS) s += d.part[2 * i + k];
  return dev::block_sum<kFinBS>(s, lds);
}

__global__ void __launch_bounds__(kFinBS) kb_fin_alpha(const BatchDesc* __restrict__ D, BatchScal* __restrict__ S) {
  __shared__ double lds[kFinBS / 64];
  BatchScal& sc = S[blockIdx.x];
  const int ph = sc.phase;
  if (ph == kFree) return;
  if (ph != kActive) {
    if (threadIdx.x == 0) sc.cycle_start = 0;
    return;
  }
  const double pq = slot_sum(D[blockIdx.x], 0, lds);
  if (threadIdx.x == 0) {
    sc.cycle_start = 0;
    sc.alpha = sc.rho / pq;
  }
}

__global__ void __launch_bounds__(kFinBS) kb_fin_beta(const BatchDesc* __restrict__ D, BatchScal* __restrict__ S) {
  __shared__ double lds[kFinBS / 64];
  BatchScal& sc = S[blockIdx.x];
  const int ph = sc.phase;
  if (ph != kActive) {
    if (threadIdx.x == 0 && ph != kFree) sc.stepped = 0;
    return;
  }
  const double RR = slot_sum(D[blockIdx.x], 0, lds);
  if (threadIdx.x == 0) {
    const int it = sc.iters + 1;
    sc.iters = it;
    sc.stepped = 1;
    sc.rr = RR;
    if (RR <= sc.stop_rr || it >= sc.maxiter) {
      sc.phase = kReplace;
    } else {
      sc.beta = RR / sc.rho;
      sc.rho = RR;
    }
  }
}

__global__ void __launch_bounds__(kFinBS) kb_fin_resid(const BatchDesc* __restrict__ D, BatchScal* __restrict__ S) {
  __shared__ double lds[kFinBS / 64];
  BatchScal& sc = S[blockIdx.x];
  const int ph = sc.phase;
  if (ph != kReplace && ph != kJoin) return;
  const double RR = slot_sum(D[blockIdx.x], 0, lds);
  const double SS = slot_sum(D[blockIdx.x], 1, lds);
  if (threadIdx.x != 0) return;
  sc.stepped = 0;
  if (ph == kJoin) {
    sc.bb = RR;
    sc.target = sc.tol * sc.tol * RR;
    sc.iters = 0;
    sc.repl = 0;
    sc.stagnant = 0;
    sc.prev_true = RR;
    sc.beta = 0.0;
    if (RR == 0.0) {
      sc.true_rr = 0.0;
      sc.status = 0;
      sc.phase = kDone;
      return;
    }
  } else {
    sc.repl += 1;
    sc.beta = sc.rho > 0.0 ? SS / sc.rho : 0.0;
  }
  sc.true_rr = RR;
  if (RR <= sc.target) {
    sc.status = 0;
    sc.phase = kDone;
    return;
  }
  if (sc.iters >= sc.maxiter) {
    sc.status = NGPB_E_NOCONV;
    sc.phase = kDone;
    return;
  }
  if (ph == kReplace) {
    // the FP32 inner solve no longer reduces the FP64 residual
    if (RR > 0.81 * sc.prev_true) {
      if (++sc.stagnant >= 3) {
        sc.status = NGPB_E_NOCONV;
        sc.phase = kDone;
        return;
      }
    } else {
      sc.stagnant = 0;
    }
    sc.prev_true = fmin(sc.prev_true, RR);
  }
  const double ratio = SS / RR;
  sc.stop_rr = fmax(sc.drop * sc.drop * SS, 0.25 * sc.target * ratio);
  sc.rho = SS;
  sc.rr = SS;
  sc.ss = SS;
  sc.cycle_start = 1;
  sc.phase = kActive;
}

}  // namespace
}  // namespace ngpb

// ---------------------------------------------------------------- host side

struct ngpb_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int cap = 0;
  std::string err;
  // options
  double tol = 1e-10, drop = 1e-3;
  int maxiter = 100000;
  int inner_per_replay = 8;
  // tuning
  int tile_slices = 8, wpb = 4;
  double tuned_ms = 0.0;
  // slots
  std::vector<ngpb_ctx*> sys;
  std::vector<int> reported;  // per slot: DONE seen by the host
  ngpb::DBuf<ngpb::BatchDesc> desc;
  ngpb::DBuf<ngpb::BatchScal> scal;
  ngpb::BatchScal* scal_host = nullptr;  // pinned mirror
  std::vector<ngpb::Tile> tiles_host;
  ngpb::DBuf<ngpb::Tile> tiles;
  int64_t n_tiles = 0;
  bool tiles_dirty = true;
  int tiled_systems = 0;  // systems in the current tile list
  hipGraph_t graph = nullptr;        // replacement step + inner iterations
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph_inner = nullptr;  // inner iterations only (no slot joining or replacing)
  hipGraphExec_t exec_inner = nullptr;
  int64_t replays = 0;
  double ms_total = 0.0;
};

namespace ngpb {
namespace {

template <class F>
int bguard(ngpb_batch* b, F&& f) {
  if (!b) return NGPB_E_ARG;
  try {
    NGPB_HIP(hipSetDevice(b->device));
    f();
    b->err.clear();
    return NGPB_OK;
  } catch (const Error& e) {
    b->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    b->err = "host allocation failed";
    return NGPB_E_NOMEM;
  } catch (const std::exception& e) {
    b->err = e.what();
    return NGPB_E_HIP;
  }
}

void drop_graph(ngpb_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->exec_inner) (void)hipGraphExecDestroy(b->exec_inner);
  if (b->graph_inner) (void)hipGraphDestroy(b->graph_inner);
  b->exec = nullptr;
  b->graph = nullptr;
  b->exec_inner = nullptr;
  b->graph_inner = nullptr;
}

// Tiles of the systems not yet done (with all_systems, of every system: the autotuner
// times the SpMV over the whole batch). Finished systems are skipped by every kernel
// anyway; leaving their tiles out saves launching their blocks.
void build_tiles(ngpb_batch* b, bool all_systems = false) {
  b->tiles_host.clear();
  b->tiled_systems = 0;
  for (int s = 0; s < b->cap; ++s) {
    ngpb_ctx* c = b->sys[s];
    if (!c || (!all_systems && b->scal_host[s].phase == kDone)) continue;
    ++b->tiled_systems;
    for (int64_t k = 0; k < c->n_slices; k += b->tile_slices) b->tiles_host.push_back(Tile{s, (int32_t)k});
  }
  b->n_tiles = (int64_t)b->tiles_host.size();
  b->tiles.reserve(std::max<size_t>(b->tiles_host.size(), 1));
  if (b->n_tiles)
    NGPB_HIP(hipMemcpyAsync(b->tiles.p, b->tiles_host.data(), b->tiles_host.size() * sizeof(Tile),
                            hipMemcpyHostToDevice, b->stream));
  b->tiles_dirty = false;
  drop_graph(b);
}

// One replay: the FP64 replacement step of the systems joining or ending an inner
// cycle (repl; every other system skips those kernels), then inner_per_replay FP32
// iterations. A replay without repl is the same work when no system is joining or
// replacing: the replacement kernels would all return at once.
template <int WPB>
void enqueue_replay(ngpb_batch* b, bool repl) {
  const dim3 g((unsigned)b->n_tiles), t(WPB * 64), gf((unsigned)b->cap), tf(kFinBS);
  const BatchDesc* D = b->desc.p;
  BatchScal* S = b->scal.p;
  const Tile* T = b->tiles.p;
  const int ts = b->tile_slices;
  hipStream_t st = b->stream;
  if (repl) {
    hipLaunchKernelGGL(kb_accum<WPB>, g, t, 0, st, D, S, T, ts);
    hipLaunchKernelGGL(kb_spmv64<WPB>, g, t, 0, st, D, S, T, ts);
    hipLaunchKernelGGL(kb_resid<WPB>, g, t, 0, st, D, S, T, ts);
    hipLaunchKernelGGL(kb_fin_resid, gf, tf, 0, st, D, S);
    hipLaunchKernelGGL(kb_dir<WPB>, g, t, 0, st, D, S, T, ts);
  }
  for (int k = 0; k < b->inner_per_replay; ++k) {
    hipLaunchKernelGGL(kb_spmv<WPB>, g, t, 0, st, D, S, T, ts, 0);
    hipLaunchKernelGGL(kb_fin_alpha, gf, tf, 0, st, D, S);
    hipLaunchKernelGGL(kb_update<WPB>, g, t, 0, st, D, S, T, ts);
    hipLaunchKernelGGL(kb_fin_beta, gf, tf, 0, st, D, S);
    hipLaunchKernelGGL(kb_dir_x<WPB>, g, t, 0, st, D, S, T, ts);
  }
}

void enqueue_replay_any(ngpb_batch* b, bool repl) {
  switch (b->wpb) {
    case 1: enqueue_replay<1>(b, repl); break;
    case 2: enqueue_replay<2>(b, repl); break;
    case 4: enqueue_replay<4>(b, repl); break;
    case 8: enqueue_replay<8>(b, repl); break;
    default: throw Error(NGPB_E_ARG, "waves per block must be 1, 2, 4 or 8");
  }
}

// the replay graph with (repl) or without the replacement step, captured on first use
hipGraphExec_t ensure_graph(ngpb_batch* b, bool repl) {
  if (b->tiles_dirty) build_tiles(b);
  hipGraph_t& graph = repl ? b->graph : b->graph_inner;
  hipGraphExec_t& exec = repl ? b->exec : b->exec_inner;
  if (exec || b->n_tiles == 0) return exec;
  NGPB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_replay_any(b, repl);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(b->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  NGPB_HIP(hipStreamEndCapture(b->stream, &graph));
  NGPB_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  return exec;
}

template <int WPB>
void launch_spmv_force(ngpb_batch* b) {
  hipLaunchKernelGGL(kb_spmv<WPB>, dim3((unsigned)b->n_tiles), dim3(WPB * 64), 0, b->stream, b->desc.p, b->scal.p,
                     b->tiles.p, b->tile_slices, 1);
}

void launch_spmv_force_any(ngpb_batch* b) {
  switch (b->wpb) {
    case 1: launch_spmv_force<1>(b); break;
    case 2: launch_spmv_force<2>(b); break;
    case 4: launch_spmv_force<4>(b); break;
    default: launch_spmv_force<8>(b); break;
  }
}

int active_slots(const ngpb_batch* b) {
  int k = 0;
  for (auto* c : b->sys) k += c != nullptr;
  return k;
}

}  // namespace
}  // namespace ngpb

using ngpb::Error;

extern "C" {

int ngpb_batch_create(ngpb_batch** out, int device, int max_systems) {
  if (!out) return NGPB_E_ARG;
  *out = nullptr;
  if (max_systems < 1 || max_systems > 65536) return NGPB_E_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return NGPB_E_NODEV;
  }
  if (device < 0 || device >= n) return NGPB_E_ARG;
  ngpb_batch* b = new (std::nothrow) ngpb_batch();
  if (!b) return NGPB_E_NOMEM;
  b->device = device;
  const int rc = ngpb::bguard(b, [&] {
    NGPB_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->cap = max_systems;
    b->sys.assign((size_t)max_systems, nullptr);
    b->reported.assign((size_t)max_systems, 0);
    b->desc.resize((size_t)max_systems);
    b->scal.resize((size_t)max_systems);
    NGPB_HIP(hipHostMalloc((void**)&b->scal_host, sizeof(ngpb::BatchScal) * (size_t)max_systems, hipHostMallocDefault));
    std::memset(b->scal_host, 0, sizeof(ngpb::BatchScal) * (size_t)max_systems);
    std::vector<ngpb::BatchDesc> zero((size_t)max_systems);
    std::memset(zero.data(), 0, zero.size() * sizeof(ngpb::BatchDesc));
    NGPB_HIP(hipMemcpyAsync(b->desc.p, zero.data(), zero.size() * sizeof(ngpb::BatchDesc), hipMemcpyHostToDevice,
                            b->stream));
    NGPB_HIP(hipMemcpyAsync(b->scal.p, b->scal_host, sizeof(ngpb::BatchScal) * (size_t)max_systems,
                            hipMemcpyHostToDevice, b->stream));
    NGPB_HIP(hipStreamSynchronize(b->stream));
  });
  if (rc != NGPB_OK) {
    if (b->scal_host) (void)hipHostFree(b->scal_host);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return rc;
  }
  *out = b;
  return NGPB_OK;
}

int ngpb_batch_destroy(ngpb_batch* b) {
  if (!b) return NGPB_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto* c : b->sys)
    if (c) c->in_batch = false;
  ngpb::drop_graph(b);
  if (b->scal_host) (void)hipHostFree(b->scal_host);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return NGPB_OK;
}

const char* ngpb_batch_last_error(const ngpb_batch* b) { return b ? b->err.c_str() : "null batch"; }

int ngpb_batch_set_options(ngpb_batch* b, const ngpb_solve_opts* o, int inner_per_replay) {
  return ngpb::bguard(b, [&] {
    NGPB_REQUIRE(o, NGPB_E_ARG, "null options");
    NGPB_REQUIRE(o->tol > 0.0 && o->maxiter > 0, NGPB_E_ARG, "tol and maxiter must be positive");
    NGPB_REQUIRE(o->solver == NGPB_SOLVER_CG && o->precon == NGPB_PRECON_JACOBI, NGPB_E_ARG,
                 "a batch runs mixed-precision CG with Jacobi scaling only");
    NGPB_REQUIRE(inner_per_replay >= 1 && inner_per_replay <= 256, NGPB_E_ARG, "inner_per_replay must be in [1, 256]");
    b->tol = o->tol;
    b->drop = (o->drop > 0.0 && o->drop < 1.0) ? o->drop : 1e-3;
    b->maxiter = o->maxiter;
    if (b->inner_per_replay != inner_per_replay) ngpb::drop_graph(b);
    b->inner_per_replay = inner_per_replay;
  });
}

int ngpb_batch_add(ngpb_batch* b, ngpb_ctx* c, int* slot) {
  return ngpb::bguard(b, [&] {
    NGPB_REQUIRE(c && slot, NGPB_E_ARG, "null argument");
    NGPB_REQUIRE(c->device == b->device, NGPB_E_ARG, "context and batch are on different devices");
    NGPB_REQUIRE(!c->in_batch, NGPB_E_STATE, "context is already in a batch");
    NGPB_REQUIRE(c->have_system, NGPB_E_STATE, "context has no assembled system");
    int s = -1;
    for (int k = 0; k < b->cap; ++k)
      if (!b->sys[k]) {
        s = k;
        break;
      }
    NGPB_REQUIRE(s >= 0, NGPB_E_STATE, "batch is full");
    NGPB_HIP(hipSetDevice(c->device));
    ngpb::check_operator(c);
    if (!c->have_compact || c->sell_banded) ngpb::build_compact(c, false);  // batch kernels use node order
    const int64_t n = c->n;
    c->x.resize((size_t)n);
    c->q.reserve((size_t)n);
    c->xs.reserve((size_t)n);
    c->rs.reserve((size_t)n);
    float* const pin = ngpb::spmv_input(c);  // zero guard entries either side
    c->qs.reserve((size_t)n);
    c->red.partials.reserve((size_t)std::max<int64_t>(2 * c->n_slices, 1));
    NGPB_HIP(hipStreamSynchronize(c->stream));
    ngpb::BatchDesc d;
    std::memset(&d, 0, sizeof(d));
    d.A.n = n;
    d.A.nslices = c->n_slices;
    d.A.width = c->slice_width.p;
    d.A.base_ptr = c->slice_base_ptr.p;
    d.A.wide_ptr = c->slice_wide_ptr.p;
    d.A.ov_ptr = c->slice_ov_ptr.p;
    d.A.val = c->sell_val.p;
    d.A.off = c->sell_off.p;
    d.A.base = c->sell_base.p;
    d.A.col32 = c->sell_col32.p;
    d.A.ov_cnt = c->sell_ov_cnt.p;
    d.A.ov_val = c->sell_ov_val.p;
    d.A.ov_col = c->sell_ov_col.p;
    d.A.pk_val = c->sell_pk_val.p;
    d.A.pk_zero = (uint32_t)c->sell_pk_entries;
    d.A.pk_idx = c->sell_pk_idx.p;
    d.A.pk_lp = c->sell_pk_lp.p;
    d.A.pk_dict = c->sell_pk_dict.p;
    d.A.pk_hdr = reinterpret_cast<const ngpb::PkHdr*>(c->sell_pk_hdr.p);
    d.A.cval = c->sell_cval.p;
    d.rowptr = c->rowptr.p;
    d.col = c->col.p;
    d.val = c->val.p;
    d.b = c->rhs.p;
    d.dscale = c->dscale.p;
    d.x = c->x.p;
    d.t = c->q.p;
    d.xs = c->xs.p;
    d.rs = c->rs.p;
    d.ps = pin;
    d.qs = c->qs.p;
    d.part = c->red.partials.p;
    d.n = n;
    d.nslices = c->n_slices;
    ngpb::BatchScal sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.phase = ngpb::kJoin;
    sc.tol = b->tol;
    sc.drop = b->drop;
    sc.maxiter = b->maxiter;
    b->scal_host[s] = sc;
    NGPB_HIP(hipMemcpyAsync(b->desc.p + s, &d, sizeof(d), hipMemcpyHostToDevice, b->stream));
    NGPB_HIP(hipMemcpyAsync(b->scal.p + s, b->scal_host + s, sizeof(sc), hipMemcpyHostToDevice, b->stream));
    NGPB_HIP(hipStreamSynchronize(b->stream));
    b->sys[s] = c;
    b->reported[s] = 0;
    c->in_batch = true;
    c->have_phi = false;
    b->tiles_dirty = true;
    *slot = s;
  });
}

int ngpb_batch_remove(ngpb_batch* b, int slot) {
  return ngpb::bguard(b, [&] {
    NGPB_REQUIRE(slot >= 0 && slot < b->cap && b->sys[slot], NGPB_E_ARG, "no system in this slot");
    NGPB_HIP(hipStreamSynchronize(b->stream));
    ngpb::BatchScal sc;
    std::memset(&sc, 0, sizeof(sc));
    b->scal_host[slot] = sc;
    NGPB_HIP(hipMemcpyAsync(b->scal.p + slot, b->scal_host + slot, sizeof(sc), hipMemcpyHostToDevice, b->stream));
    NGPB_HIP(hipStreamSynchronize(b->stream));
    b->sys[slot]->in_batch = false;
    b->sys[slot] = nullptr;
    b->reported[slot] = 0;
    b->tiles_dirty = true;
  });
}

// Picks the tile size (slices per block) and the waves per block by timing the
// batched compact SpMV over every system in the batch. Every candidate computes the
// same per-slice results; only the launch shape differs. Must be called while no
// system of the batch is iterating (e.g. right after adding systems).
int ngpb_batch_autotune(ngpb_batch* b, int* tile_slices, int* waves_per_block, double* ms_per_spmv) {
  return ngpb::bguard(b, [&] {
    NGPB_REQUIRE(ngpb::active_slots(b) > 0, NGPB_E_STATE, "empty batch");
    NGPB_HIP(hipStreamSynchronize(b->stream));
    for (int s = 0; s < b->cap; ++s)
      if (b->sys[s]) {
        const int ph = b->scal_host[s].phase;
        NGPB_REQUIRE(ph == ngpb::kJoin || ph == ngpb::kDone, NGPB_E_STATE,
                     "autotune needs every system of the batch to be joining or done");
      }
    const int cand_ts[] = {2, 4, 8, 16, 32};
    const int cand_w[] = {1, 2, 4, 8};
    hipEvent_t e0, e1;
    NGPB_HIP(hipEventCreate(&e0));
    NGPB_HIP(hipEventCreate(&e1));
    double best = 1e30;
    int bts = b->tile_slices, bw = b->wpb;
    for (int ts : cand_ts)
      for (int w : cand_w) {
        if (ts < w) continue;
        b->tile_slices = ts;
        b->wpb = w;
        ngpb::build_tiles(b, true);
        ngpb::launch_spmv_force_any(b);  // warm
        const int reps = 5;
        NGPB_HIP(hipEventRecord(e0, b->stream));
        for (int r = 0; r < reps; ++r) ngpb::launch_spmv_force_any(b);
        NGPB_HIP(hipEventRecord(e1, b->stream));
        NGPB_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        NGPB_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double m = (double)ms / reps;
        if (m < best) {
          best = m;
          bts = ts;
          bw = w;
        }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    NGPB_HIP(hipGetLastError());
    b->tile_slices = bts;
    b->wpb = bw;
    b->tuned_ms = best;
    b->tiles_dirty = true;
    if (tile_slices) *tile_slices = bts;
    if (waves_per_block) *waves_per_block = bw;
    if (ms_per_spmv) *ms_per_spmv = best;
  });
}

int ngpb_batch_set_tiling(ngpb_batch* b, int tile_slices, int waves_per_block) {
  return ngpb::bguard(b, [&] {
    NGPB_REQUIRE(tile_slices >= 1 && tile_slices <= 4096, NGPB_E_ARG, "tile_slices must be in [1, 4096]");
    NGPB_REQUIRE(waves_per_block == 1 || waves_per_block == 2 || waves_per_block == 4 || waves_per_block == 8,
                 NGPB_E_ARG, "waves per block must be 1, 2, 4 or 8");
    b->tile_slices = tile_slices;
    b->wpb = waves_per_block;
    b->tiles_dirty = true;
  });
}

// Runs up to max_replays graph replays (each: replacement step + inner_per_replay
// inner iterations), stopping early once every system is done. *running receives the
// number of systems not yet done.
int ngpb_batch_run(ngpb_batch* b, int max_replays, int* running) {
  return ngpb::bguard(b, [&] {
    NGPB_REQUIRE(max_replays >= 0, NGPB_E_ARG, "max_replays must be >= 0");
    int left = 0;
    for (int s = 0; s < b->cap; ++s)
      if (b->sys[s] && b->scal_host[s].phase != ngpb::kDone) ++left;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    NGPB_HIP(hipEventCreate(&e0));
    NGPB_HIP(hipEventCreate(&e1));
    for (int r = 0; r < max_replays && left > 0; ++r) {
      // the host mirror holds every slot's phase after the last replay (or join)
      bool repl = false;
      for (int s = 0; s < b->cap; ++s)
        if (b->sys[s] && (b->scal_host[s].phase == ngpb::kJoin || b->scal_host[s].phase == ngpb::kReplace))
          repl = true;
      hipGraphExec_t ex = ngpb::ensure_graph(b, repl);
      if (!ex) break;
      NGPB_HIP(hipEventRecord(e0, b->stream));
      NGPB_HIP(hipGraphLaunch(ex, b->stream));
      NGPB_HIP(hipEventRecord(e1, b->stream));
      NGPB_HIP(hipMemcpyAsync(b->scal_host, b->scal.p, sizeof(ngpb::BatchScal) * (size_t)b->cap,
                              hipMemcpyDeviceToHost, b->stream));
      NGPB_HIP(hipStreamSynchronize(b->stream));
      float ms = 0.f;
      NGPB_HIP(hipEventElapsedTime(&ms, e0, e1));
      b->ms_total += ms;
      ++b->replays;
      left = 0;
      for (int s = 0; s < b->cap; ++s) {
        ngpb_ctx* c = b->sys[s];
        if (!c) continue;
        if (b->scal_host[s].phase == ngpb::kDone) {
          if (!b->reported[s]) {
            b->reported[s] = 1;
            c->have_phi = true;
            c->phi_len = c->n;
          }
        } else {
          ++left;
        }
      }
      // once a quarter of the tiled systems are done, re-tile (and re-capture the
      // graph) without them
      if (left > 0 && 4 * (b->tiled_systems - left) >= b->tiled_systems) b->tiles_dirty = true;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (running) *running = left;
  });
}

// done = 1 once the slot's solve has finished (stats->status tells how).
int ngpb_batch_status(ngpb_batch* b, int slot, int* done, ngpb_solve_stats* st) {
  return ngpb::bguard(b, [&] {
    NGPB_REQUIRE(slot >= 0 && slot < b->cap && b->sys[slot], NGPB_E_ARG, "no system in this slot");
    const ngpb::BatchScal& sc = b->scal_host[slot];
    if (done) *done = sc.phase == ngpb::kDone ? 1 : 0;
    if (st) {
      ngpb_ctx* c = b->sys[slot];
      st->status = sc.phase == ngpb::kDone ? sc.status : NGPB_E_NOCONV;
      st->iterations = sc.iters;
      st->replacements = sc.repl;
      st->relres = sc.bb > 0.0 ? std::sqrt(sc.true_rr / sc.bb) : 0.0;
      st->ms_solve = 0.0;
      st->bytes_per_iter = ngpb::spmv_compact_bytes(c) + (double)c->n * 4.0 * 8.0;
    }
  });
}

// Replays run so far and their summed device time (HIP events around each replay).
int ngpb_batch_timing(ngpb_batch* b, int64_t* replays, double* ms_total) {
  return ngpb::bguard(b, [&] {
    if (replays) *replays = b->replays;
    if (ms_total) *ms_total = b->ms_total;
  });
}

}  // extern "C"
