// FP64 CSR SpMV and the plain Jacobi-preconditioned CG (LIS -i cg -p jacobi as
// called from lis_compute_electric_potential, src/pb_class.cpp:2243-2329).
//
// SpMV: a block of 256 threads owns 256 consecutive rows. Their column indices and
// values are read once, coalesced, and each (value, x[column]) pair is staged in
// LDS; each thread then forms its row sum in column order as a chain of fused
// multiply-adds starting from zero (the same arithmetic as the oracle's SpMV, so the
// results are bit-identical). Blocks whose rows hold more than kCap entries stage
// them in pieces of kCap, each row carrying its chain from piece to piece (reading
// such blocks' entries row by row from global memory instead made the 8.04M-row
// SpMV 10% slower: 217 against 195 us).
//
// CG iteration, three kernels: SpMV + per-block (p, q) -> x/r update + per-block
// (r, D^-1 r), (r, r) -> stop test and new direction. Each kernel first combines the
// previous kernel's per-block partial sums itself, every block in the same fixed
// order, so all blocks hold identical scalars; reductions are deterministic and no
// kernel needs a device-wide memory fence (on a multi-XCD GPU an agent-scope fence
// writes back the XCD's L2, which would serialize every block). Scalars live in
// device memory; the host polls the stop flag every `check_every` iterations and
// kernels launched after the stop exit at once.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "ngpb_internal.h"

namespace ngpb {
namespace {

constexpr int kBS = 256;
constexpr int kCap = 2048;  // staged entries per block
constexpr int kFin = 1024;  // finalize block size

// SCHED 0: block b takes chunks b, b + G, ...; SCHED 1: block b takes one contiguous
// range of chunks, with blocks renumbered so that the blocks on one XCD (block id
// mod 8) cover one contiguous eighth of the rows (L2 reuse of the x gathers); the
// grid is then the number of blocks resident at once.
template <bool DOT, int SCHED>
__global__ void __launch_bounds__(kBS) k_spmv64(int64_t n, const int32_t* __restrict__ rowptr,
                                                 const int32_t* __restrict__ col, const double* __restrict__ val,
                                                 const double* __restrict__ x, double* __restrict__ y,
                                                 const KrylovScalars* ks, double* __restrict__ partials) {
  __shared__ double s_val[kCap];
  __shared__ double s_x[kCap];
  __shared__ double s_red[kBS / 64];
  if (DOT && ks->done) return;
  double part = 0.0;
  const int64_t nchunks = (n + kBS - 1) / kBS;
  int64_t c_begin = blockIdx.x, c_end = nchunks, c_step = gridDim.x;
  if (SCHED == 1) {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const int64_t lb = (nb % 8u == 0u) ? (int64_t)((b % 8u) * (nb / 8u) + b / 8u) : (int64_t)b;
    c_begin = lb * nchunks / nb;
    c_end = (lb + 1) * nchunks / nb;
    c_step = 1;
  }
  for (int64_t chunk = c_begin; chunk < c_end; chunk += c_step) {
    const int64_t r0 = chunk * kBS;
    const int64_t r1 = min(r0 + kBS, n);
    const int64_t row = r0 + threadIdx.x;
    const int64_t e0 = rowptr[r0], e1 = rowptr[r1];
    const int64_t rb = row < n ? (int64_t)rowptr[row] : e1, re = row < n ? (int64_t)rowptr[row + 1] : e1;
    double sum = 0.0;
    // the chunk's entries staged kCap at a time; each row continues its fma chain
    // across pieces, so the order of its terms is the column order as before
    for (int64_t p0 = e0; p0 < e1; p0 += kCap) {
      const int64_t p1 = min(p0 + (int64_t)kCap, e1);
      __syncthreads();  // the previous piece's rows are done with the staging buffers
      // all loads of the piece are issued before any is used: indices past the piece's
      // last entry are clamped to it and their LDS slots are never read
      constexpr int kU = kCap / kBS;
      double v[kU];
      int c[kU];
      const int64_t kl = p1 - 1;
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t k = min(p0 + threadIdx.x + (int64_t)u * kBS, kl);
        // the matrix is read once per SpMV: non-temporal loads leave the last-level cache
        // to the CG vectors (FP64 solve of the 8M-unknown globular system 174 -> 170 ms)
        v[u] = __builtin_nontemporal_load(&val[k]);
        c[u] = __builtin_nontemporal_load(&col[k]);
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t k = p0 + threadIdx.x + (int64_t)u * kBS;
        s_val[k - p0] = v[u];
        s_x[k - p0] = x[c[u]];
      }
      __syncthreads();
      const int64_t kb = max(rb, p0), ke = min(re, p1);
      for (int64_t k = kb; k < ke; ++k) sum = fma(s_val[k - p0], s_x[k - p0], sum);
    }
    if (row < n) y[row] = sum;
    if (DOT && row < n) part += x[row] * sum;
  }
  if (DOT) {
    part = dev::block_sum<kBS>(part, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = part;
  }
}

unsigned spmv64_grid(ngpb_ctx* c) {
  const int64_t need = std::max<int64_t>(1, (c->n + kBS - 1) / kBS);
  if (c->spmv_sched == 0) return (unsigned)std::min<int64_t>(need, 2048);
  if (c->csr_resident <= 0) {
    int cus = 0, per_cu = 0;
    NGPB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    NGPB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_spmv64<true, 1>, kBS, 0));
    c->csr_resident = std::max(8, (std::max(cus, 1) * std::max(per_cu, 1)) / 8 * 8);
  }
  if (need >= c->csr_resident) return (unsigned)c->csr_resident;
  return (unsigned)(c->spmv_sched == 1 ? (need + 7) / 8 * 8 : need);
}

template <bool DOT>
void launch_spmv64(ngpb_ctx* c, unsigned grid, const double* x, double* y, double* part) {
  const KrylovScalars* ks = DOT ? c->ks.p : nullptr;
  if (c->spmv_sched == 1)
    spmv_launch(c, k_spmv64<DOT, 1>, dim3(grid), dim3(kBS), c->n, c->rowptr.p, c->col.p, c->val.p, x, y, ks, part);
  else
    spmv_launch(c, k_spmv64<DOT, 0>, dim3(grid), dim3(kBS), c->n, c->rowptr.p, c->col.p, c->val.p, x, y, ks, part);
}

// x = 0, r = b, p = D^-1 b; per block (r, D^-1 r), (b, b)
__global__ void __launch_bounds__(kBS) k_pcg_init(int64_t n, const double* __restrict__ b, const double* __restrict__ dinv,
                                                  double* __restrict__ x, double* __restrict__ r, double* __restrict__ p,
                                                  double* __restrict__ partials) {
  __shared__ double s_red[kBS / 64];
  double rz = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double bi = b[i], zi = dinv[i] * bi;
    x[i] = 0.0;
    r[i] = bi;
    p[i] = zi;
    rz += bi * zi;
    bb += bi * bi;
  }
  rz = dev::block_sum<kBS>(rz, s_red);
  bb = dev::block_sum<kBS>(bb, s_red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = rz;
    partials[2 * blockIdx.x + 1] = bb;
  }
}

__global__ void __launch_bounds__(kFin) k_fin_init(KrylovScalars* ks, const double* __restrict__ partials, int np,
                                                   double tol2, int maxiter) {
  __shared__ double s_red[kFin / 64];
  const double RZ = dev::sum_partials<kFin>(partials, np, 2, s_red);
  const double BB = dev::sum_partials<kFin>(partials + 1, np, 2, s_red);
  if (threadIdx.x == 0) {
    ks->rho = RZ;
    ks->rho2[0] = RZ;
    ks->iter2[0] = 0;
    ks->bb = BB;
    ks->rr = BB;
    ks->stop_rr = tol2 * BB;
    ks->iter = 0;
    ks->maxiter = maxiter;
    ks->status = 0;
    ks->done = (BB == 0.0) ? 1 : 0;
  }
}

// Iteration of parity par. Every block forms (p, q) from the SpMV partials in the
// same fixed order, so all blocks use the same alpha = rho / (p, q).
// x += alpha p; r -= alpha q; per block (r, D^-1 r), (r, r)
__global__ void __launch_bounds__(kBS) k_pcg_update(int64_t n, const double* __restrict__ dinv, double* __restrict__ x,
                                                    double* __restrict__ r, const double* __restrict__ p,
                                                    const double* __restrict__ q, KrylovScalars* ks,
                                                    const double* __restrict__ pq_part, int npq,
                                                    double* __restrict__ partials, int par) {
  __shared__ double s_red[kBS / 64];
  if (ks->done) return;
  const double pq = dev::sum_partials<kBS>(pq_part, npq, 1, s_red);
  const double alpha = ks->rho2[par] / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ks->pq = pq;
    ks->alpha2[par] = alpha;
  }
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    rz += ri * (dinv[i] * ri);
    rr += ri * ri;
  }
  rz = dev::block_sum<kBS>(rz, s_red);
  rr = dev::block_sum<kBS>(rr, s_red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = rz;
    partials[2 * blockIdx.x + 1] = rr;
  }
}

// Stop test on ||r||^2 (every block decides alike), then beta = (r', z') / (r, z) and
// p = D^-1 r + beta p. Block 0 records the scalars for the next iteration and the host.
__global__ void __launch_bounds__(kBS) k_pcg_dir(int64_t n, const double* __restrict__ dinv, const double* __restrict__ r,
                                                 double* __restrict__ p, KrylovScalars* ks,
                                                 const double* __restrict__ part, int np, int par) {
  __shared__ double s_red[kBS / 64];
  if (ks->done) return;
  const double RZ = dev::sum_partials<kBS>(part, np, 2, s_red);
  const double RR = dev::sum_partials<kBS>(part + 1, np, 2, s_red);
  const int it = ks->iter2[par] + 1;
  const bool conv = RR <= ks->stop_rr;
  const bool stop = conv || it >= ks->maxiter;
  const double rho = ks->rho2[par];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ks->rr = RR;
    ks->iter = it;
    ks->iter2[par ^ 1] = it;
    if (stop) {
      ks->done = 1;
      ks->status = conv ? 0 : NGPB_E_NOCONV;
      ks->stop_par = par;
    } else {
      ks->beta = RZ / rho;
      ks->rho = RZ;
      ks->rho2[par ^ 1] = RZ;
    }
  }
  if (stop) return;
  const double beta = RZ / rho;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS)
    p[i] = dinv[i] * r[i] + beta * p[i];
}

__global__ void k_diag_inv(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                           const double* __restrict__ val, double* __restrict__ dinv, int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
    const double v = val[k];
    if (!isfinite(v)) atomicOr(bad, 2);
    if (col[k] == i) d = v;
  }
  if (!(d > 0.0)) atomicOr(bad, 1);
  dinv[i] = 1.0 / d;
}

__global__ void __launch_bounds__(kBS) k_copy(int64_t n16, const uint4* __restrict__ a, uint4* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n16; i += (int64_t)gridDim.x * kBS) b[i] = a[i];
}

}  // namespace

unsigned vec_grid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 4 * kBS - 1) / (4 * kBS), 2048));
}

// opens the timing of one SpMV when profiling is on and tag is kProfEvery / 2 modulo
// kProfEvery (tag < 0: always). The solvers poll the host every check_every launches
// (32 by default, a multiple of kProfEvery), so the sample leaves out the first launch
// after each poll, which starts on an idle GPU.
void prof_begin(ngpb_ctx* c, int kind, int64_t tag) {
  c->prof_kind = -1;
  if (!c->profile || (tag >= 0 && tag % kProfEvery != kProfEvery / 2)) return;
  c->prof_kind = kind;
  c->prof_tag = tag;
  c->prof_first = true;
}

void prof_end(ngpb_ctx* c, int kind) {
  (void)kind;
  c->prof_kind = -1;
}

// the event pair for the next kernel of the SpMV being timed
hipEvent_t* prof_pair(ngpb_ctx* c) {
  if (c->prof_kind < 0) return nullptr;
  auto& p = c->kprof[c->prof_kind];
  while (p.ev.size() < p.used + 2) {
    hipEvent_t e;
    NGPB_HIP(hipEventCreate(&e));
    p.ev.push_back(e);
  }
  const size_t k = p.used / 2;
  if (p.tag.size() <= k) {
    p.tag.resize(k + 1);
    p.opens.resize(k + 1);
  }
  p.tag[k] = c->prof_tag;
  p.opens[k] = c->prof_first ? 1 : 0;
  c->prof_first = false;
  hipEvent_t* e = &p.ev[p.used];
  p.used += 2;
  return e;
}

// Call after the stream has been synchronized: adds the timed SpMVs whose tag is below
// `limit` (launches queued after the stop flag exit at once and would bias the
// average; limit < 0 keeps all) and clears the record. An SpMV's time is the sum of
// its kernels' own times.
void prof_collect(ngpb_ctx* c, int kind, int64_t limit) {
  auto& p = c->kprof[kind];
  for (size_t i = 0; i + 1 < p.used; i += 2) {
    if (limit >= 0 && p.tag[i / 2] >= limit) continue;
    float ms = 0.f;
    NGPB_HIP(hipEventElapsedTime(&ms, p.ev[i], p.ev[i + 1]));
    p.ms += ms;
    if (p.opens[i / 2]) ++p.launches;
  }
  p.used = 0;
}

// plain device copy of `bytes` (multiple of 16): reads and writes `bytes` each launch
double time_copy(ngpb_ctx* c, int64_t bytes, int reps) {
  const int64_t n16 = bytes / 16;
  DBuf<uint4> a, b;
  a.resize((size_t)n16);
  b.resize((size_t)n16);
  NGPB_HIP(hipMemsetAsync(a.p, 0, n16 * 16, c->stream));
  const unsigned g = (unsigned)std::min<int64_t>((n16 + kBS - 1) / kBS, 8192);
  hipEvent_t e0, e1;
  NGPB_HIP(hipEventCreate(&e0));
  NGPB_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_copy, dim3(g), dim3(kBS), 0, c->stream, n16, a.p, b.p);
  NGPB_HIP(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps; ++i) hipLaunchKernelGGL(k_copy, dim3(g), dim3(kBS), 0, c->stream, n16, a.p, b.p);
  NGPB_HIP(hipEventRecord(e1, c->stream));
  NGPB_HIP(hipEventSynchronize(e1));
  NGPB_HIP(hipGetLastError());
  float ms = 0.f;
  NGPB_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / std::max(reps, 1);
}

void ensure_reduce(ngpb_ctx* c, int64_t blocks) {
  const size_t need = (size_t)std::max<int64_t>(blocks, 2048) * 4;
  if (c->red.partials.n < need) c->red.partials.resize(need);
  if (!c->ks.p) c->ks.resize(1);
  if (!c->ks_host) NGPB_HIP(hipHostMalloc((void**)&c->ks_host, sizeof(KrylovScalars), hipHostMallocDefault));
}

void spmv_fp64_dot(ngpb_ctx* c, const double* x, double* y, double* partials) {
  launch_spmv64<true>(c, spmv64_grid(c), x, y, partials);
}

unsigned spmv_fp64_grid(ngpb_ctx* c) { return spmv64_grid(c); }

double spmv_fp64_bytes(const ngpb_ctx* c) {
  // nnz (value + column) + row pointers + x read + y written
  return (double)c->nnz * (8.0 + 4.0) + (double)(c->n + 1) * 4.0 + (double)c->n * 16.0;
}

void spmv_fp64(ngpb_ctx* c, const double* x, double* y) {
  launch_spmv64<false>(c, spmv64_grid(c), x, y, nullptr);
  NGPB_HIP(hipGetLastError());
}

double time_spmv_fp64(ngpb_ctx* c, int reps) {
  c->p.reserve((size_t)c->n);
  c->q.reserve((size_t)c->n);
  NGPB_HIP(hipMemsetAsync(c->p.p, 0, c->n * sizeof(double), c->stream));
  hipEvent_t e0, e1;
  NGPB_HIP(hipEventCreate(&e0));
  NGPB_HIP(hipEventCreate(&e1));
  spmv_fp64(c, c->p.p, c->q.p);  // warm-up
  NGPB_HIP(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps; ++i) spmv_fp64(c, c->p.p, c->q.p);
  NGPB_HIP(hipEventRecord(e1, c->stream));
  NGPB_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  NGPB_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / std::max(reps, 1);
}

void check_operator(ngpb_ctx* c) {
  c->dinv.resize((size_t)c->n);
  DBuf<int> bad;
  bad.resize(1);
  NGPB_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(k_diag_inv, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->rowptr.p, c->col.p,
                     c->val.p, c->dinv.p, bad.p);
  NGPB_HIP(hipGetLastError());
  int hb = 0;
  NGPB_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  NGPB_REQUIRE(!(hb & 2), NGPB_E_NUMERIC, "operator holds NaN or Inf entries");
  NGPB_REQUIRE(!(hb & 1), NGPB_E_NUMERIC, "operator has a zero or negative diagonal entry");
}

void solve_fp64(ngpb_ctx* c, const ngpb_solve_opts& o, ngpb_solve_stats& st) {
  const int64_t n = c->n;
  check_operator(c);
  c->x.resize((size_t)n);
  c->r.resize((size_t)n);
  c->p.reserve((size_t)n);
  c->q.reserve((size_t)n);
  const unsigned gs = spmv64_grid(c), gv = vec_grid(n);
  ensure_reduce(c, std::max<int64_t>(gs, 2048) + 2 * (int64_t)gv);
  double* part = c->red.partials.p;
  KrylovScalars* ks = c->ks.p;
  hipEvent_t e0, e1;
  NGPB_HIP(hipEventCreate(&e0));
  NGPB_HIP(hipEventCreate(&e1));
  NGPB_HIP(hipEventRecord(e0, c->stream));
  hipLaunchKernelGGL(k_pcg_init, dim3(gv), dim3(kBS), 0, c->stream, n, c->rhs.p, c->dinv.p, c->x.p, c->r.p, c->p.p, part);
  hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(kFin), 0, c->stream, ks, part, (int)gv, o.tol * o.tol, o.maxiter);
  NGPB_HIP(hipGetLastError());
  // partial-sum regions: SpMV (p, q) and update (r, z), (r, r)
  double* ppq = part;
  double* pup = part + std::max<int64_t>(gs, 2048);
  const int chunk = std::max(1, o.check_every);
  int64_t launch = 0;
  while (true) {
    for (int k = 0; k < chunk; ++k, ++launch) {
      const int par = (int)(launch & 1);
      prof_begin(c, 0, launch);
      launch_spmv64<true>(c, gs, c->p.p, c->q.p, ppq);
      prof_end(c, 0);
      hipLaunchKernelGGL(k_pcg_update, dim3(gv), dim3(kBS), 0, c->stream, n, c->dinv.p, c->x.p, c->r.p, c->p.p, c->q.p,
                         ks, ppq, (int)gs, pup, par);
      hipLaunchKernelGGL(k_pcg_dir, dim3(gv), dim3(kBS), 0, c->stream, n, c->dinv.p, c->r.p, c->p.p, ks, pup, (int)gv,
                         par);
    }
    NGPB_HIP(hipGetLastError());
    NGPB_HIP(hipMemcpyAsync(c->ks_host, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c->stream));
    NGPB_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c, 0, c->ks_host->iter);
    if (c->ks_host->done) break;
  }
  NGPB_HIP(hipEventRecord(e1, c->stream));
  NGPB_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  NGPB_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const KrylovScalars& h = *c->ks_host;
  st.status = h.status;
  st.iterations = h.iter;
  st.replacements = 0;
  st.relres = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
  st.ms_solve = ms;
  // SpMV + update (x, p, r, q, dinv read; x, r written) + direction (dinv, r, p read; p written)
  st.bytes_per_iter = spmv_fp64_bytes(c) + (double)n * 8.0 * (7.0 + 4.0);
  c->have_phi = true;
  c->phi_len = c->n;
}

// ------------------------------------------------------------ allocation-free entry point
// ngpb_solve_csr_noalloc runs the CG-Jacobi above on a CSR system the caller keeps in
// device memory, with its work vectors, partial sums, scalars and flags in one
// file-scope device array: the call allocates nothing (no hipMalloc / hipFree, no
// pinned host memory, no events), for callers that cannot allocate. Each GPU holds
// its own copy of the array until the process ends; its address on a device is taken
// once with hipGetSymbolAddress, and only that address is handed to kernels and
// copies. Every call on one GPU uses the same array, so calls that overlap on one GPU
// (from several host threads or streams) race; each call initializes what it reads.
constexpr int64_t kNoAllocMaxRows = int64_t(1) << 20;  // 1,048,576 rows
constexpr int64_t kNoAllocPartials = 8192;             // >= 2048 + 2 * 2048
constexpr int64_t kNoAllocDoubles = 4 * kNoAllocMaxRows + kNoAllocPartials + 64;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

__global__ void k_clear_int(int* p) { *p = 0; }

// Structure of a CSR the library did not build: row pointers from 0 to nnz, non-
// decreasing, columns in [0, n). Sets bit 4 of *bad otherwise; reads col only below nnz.
__global__ void k_csr_check(int64_t n, int64_t nnz, const int32_t* __restrict__ rowptr,
                            const int32_t* __restrict__ col, int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t a = rowptr[i], e = rowptr[i + 1];
  if ((i == 0 && a != 0) || (i == n - 1 && e != nnz) || a < 0 || e < a || e > nnz) {
    atomicOr(bad, 4);
    return;
  }
  for (int64_t k = a; k < e; ++k)
    if (col[k] < 0 || col[k] >= n) {
      atomicOr(bad, 4);
      return;
    }
}

// device address of `scratch` on the current device (taken on the first call there)
double* scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    NGPB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    NGPB_REQUIRE(ptr != nullptr, NGPB_E_HIP, "no device address for the scratch array");
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[device];
}

}  // namespace

int64_t noalloc_max_rows() { return kNoAllocMaxRows; }

void solve_csr_noalloc(int device, hipStream_t s, int64_t n, int64_t nnz, const int32_t* rowptr, const int32_t* col,
                       const double* val, const double* b, double* x, double tol, int maxiter, int check_every,
                       ngpb_solve_stats& st) {
  NGPB_REQUIRE(n > 0 && n <= kNoAllocMaxRows, NGPB_E_ARG,
               "ngpb_solve_csr_noalloc takes 1 to 1,048,576 rows (its scratch array is sized for that)");
  double* base = scratch_on(device);
  double* dinv = base;
  double* r = base + kNoAllocMaxRows;
  double* p = base + 2 * kNoAllocMaxRows;
  double* q = base + 3 * kNoAllocMaxRows;
  double* part = base + 4 * kNoAllocMaxRows;
  KrylovScalars* ks = reinterpret_cast<KrylovScalars*>(part + kNoAllocPartials);
  int* bad = reinterpret_cast<int*>(ks + 1);
  static_assert(sizeof(KrylovScalars) + sizeof(int) <= 64 * sizeof(double), "scalar area too small");
  const auto t0 = std::chrono::steady_clock::now();
  hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(1), 0, s, bad);
  hipLaunchKernelGGL(k_csr_check, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, nnz, rowptr, col, bad);
  NGPB_HIP(hipGetLastError());
  int hb = 0;
  NGPB_HIP(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  NGPB_HIP(hipStreamSynchronize(s));
  NGPB_REQUIRE(!(hb & 4), NGPB_E_ARG, "row pointers must run from 0 to nnz without decreasing, columns in [0, n)");
  hipLaunchKernelGGL(k_diag_inv, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, rowptr, col, val, dinv, bad);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  NGPB_HIP(hipStreamSynchronize(s));
  NGPB_REQUIRE(!(hb & 2), NGPB_E_NUMERIC, "operator holds NaN or Inf entries");
  NGPB_REQUIRE(!(hb & 1), NGPB_E_NUMERIC, "operator has a zero or negative diagonal entry");
  // the launch shapes of ngpb_solve under SpMV schedule 0
  const unsigned gs = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (n + kBS - 1) / kBS), 2048);
  const unsigned gv = vec_grid(n);
  double* ppq = part;
  double* pup = part + 2048;
  hipLaunchKernelGGL(k_pcg_init, dim3(gv), dim3(kBS), 0, s, n, b, dinv, x, r, p, pup);
  hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(kFin), 0, s, ks, pup, (int)gv, tol * tol, maxiter);
  NGPB_HIP(hipGetLastError());
  const int chunk = std::max(1, check_every);
  KrylovScalars h;
  std::memset(&h, 0, sizeof(h));
  for (int64_t launch = 0;;) {
    for (int k = 0; k < chunk; ++k, ++launch) {
      const int par = (int)(launch & 1);
      hipLaunchKernelGGL((k_spmv64<true, 0>), dim3(gs), dim3(kBS), 0, s, n, rowptr, col, val, p, q, ks, ppq);
      hipLaunchKernelGGL(k_pcg_update, dim3(gv), dim3(kBS), 0, s, n, dinv, x, r, p, q, ks, ppq, (int)gs, pup, par);
      hipLaunchKernelGGL(k_pcg_dir, dim3(gv), dim3(kBS), 0, s, n, dinv, r, p, ks, pup, (int)gv, par);
    }
    NGPB_HIP(hipGetLastError());
    NGPB_HIP(hipMemcpyAsync(&h, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, s));
    NGPB_HIP(hipStreamSynchronize(s));
    if (h.done) break;
  }
  st.status = h.status;
  st.iterations = h.iter;
  st.replacements = 0;
  st.relres = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
  st.ms_solve = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  // SpMV (12 nnz + 4 (n + 1) + 16 n) + update and direction vector streams, as ngpb_solve
  st.bytes_per_iter = 12.0 * (double)nnz + 4.0 * (double)(n + 1) + 16.0 * (double)n + (double)n * 8.0 * (7.0 + 4.0);
}

}  // namespace ngpb
