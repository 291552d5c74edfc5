// Device-wide exclusive prefix sum (int64) and stream compaction.
// Three passes over tiles of 2048 items: tile sums, a single-block scan of the
// tile sums, and the per-tile scan with its offset. Integer arithmetic, so the
// result does not depend on scheduling.
#include "ngpb_internal.h"

namespace ngpb {
namespace {

constexpr int kBS = 256;
constexpr int kIPT = 8;
constexpr int kTile = kBS * kIPT;

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// exclusive scan across the block of one value per thread; *total = block sum
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* lds, int64_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t inc = wave_incl_scan(v);
  if (lane == 63) lds[wid] = inc;
  __syncthreads();
  int64_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kBS / 64; ++w) {
    const int64_t t = lds[w];
    if (w < wid) before += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

// sum of tile `tile` of in[0, n)
__device__ __forceinline__ void tile_sum_body(const int64_t* in, int64_t n, int64_t tile, int64_t* out) {
  __shared__ int64_t lds[kBS / 64];
  const int64_t base = tile * kTile + (int64_t)threadIdx.x * kIPT;
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kIPT; ++k)
    if (base + k < n) s += in[base + k];
  int64_t tot;
  (void)block_excl_scan(s, lds, &tot);
  if (threadIdx.x == 0) *out = tot;
}

// in-place exclusive scan of a[0, n) by one block; *total = the sum
__device__ __forceinline__ void scan_single_body(int64_t* a, int64_t n, int64_t* total) {
  __shared__ int64_t lds[kBS / 64];
  int64_t carry = 0;
  for (int64_t t0 = 0; t0 < n; t0 += kTile) {
    const int64_t base = t0 + (int64_t)threadIdx.x * kIPT;
    int64_t v[kIPT];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
      v[k] = (base + k < n) ? a[base + k] : 0;
      s += v[k];
    }
    int64_t tot;
    int64_t ex = block_excl_scan(s, lds, &tot) + carry;
#pragma unroll
    for (int k = 0; k < kIPT; ++k)
      if (base + k < n) {
        a[base + k] = ex;
        ex += v[k];
      }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

// exclusive scan of tile `tile` of in[0, n) into out, starting from off
__device__ __forceinline__ void tile_apply_body(const int64_t* in, int64_t* out, int64_t n, int64_t tile, int64_t off) {
  __shared__ int64_t lds[kBS / 64];
  const int64_t base = tile * kTile + (int64_t)threadIdx.x * kIPT;
  int64_t v[kIPT];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kIPT; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int64_t tot;
  int64_t ex = block_excl_scan(s, lds, &tot) + off;
#pragma unroll
  for (int k = 0; k < kIPT; ++k)
    if (base + k < n) {
      out[base + k] = ex;
      ex += v[k];
    }
}

__global__ void __launch_bounds__(kBS) k_tile_sums(const int64_t* in, int64_t n, int64_t* tile_sums) {
  tile_sum_body(in, n, blockIdx.x, &tile_sums[blockIdx.x]);
}

__global__ void __launch_bounds__(kBS) k_scan_single(int64_t* a, int64_t n, int64_t* total) {
  scan_single_body(a, n, total);
}

__global__ void __launch_bounds__(kBS) k_tile_apply(const int64_t* in, int64_t* out, int64_t n,
                                                    const int64_t* tile_off) {
  tile_apply_body(in, out, n, blockIdx.x, tile_off[blockIdx.x]);
}

// Many segments in one launch each: segment i owns tiles [tile0[i], tile0[i + 1]).
__global__ void __launch_bounds__(kBS) k_tile_sums_many(const ScanSeg* __restrict__ seg,
                                                        const int64_t* __restrict__ tile0, int ns,
                                                        int64_t* __restrict__ tile_sums) {
  const int i = group_of(tile0, ns, blockIdx.x);
  tile_sum_body(seg[i].in, seg[i].n, blockIdx.x - tile0[i], &tile_sums[blockIdx.x]);
}

__global__ void __launch_bounds__(kBS) k_scan_single_many(const int64_t* __restrict__ tile0, int64_t* tile_sums,
                                                          int64_t* __restrict__ totals) {
  const int i = blockIdx.x;
  scan_single_body(tile_sums + tile0[i], tile0[i + 1] - tile0[i], &totals[i]);
}

__global__ void __launch_bounds__(kBS) k_tile_apply_many(const ScanSeg* __restrict__ seg,
                                                         const int64_t* __restrict__ tile0, int ns,
                                                         const int64_t* __restrict__ tile_off) {
  const int i = group_of(tile0, ns, blockIdx.x);
  tile_apply_body(seg[i].in, seg[i].out, seg[i].n, blockIdx.x - tile0[i], tile_off[blockIdx.x]);
}

__global__ void k_flags_to_counts(const uint8_t* f, int64_t n, int64_t* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt[i] = f[i] ? 1 : 0;
}

__global__ void k_scatter_flags(const uint8_t* f, const int64_t* pos, int64_t n, int32_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && f[i]) out[pos[i]] = (int32_t)i;
}

}  // namespace

// out may alias in. Returns the total (synchronizes the stream to read it).
int64_t exclusive_scan(ngpb_ctx* c, const int64_t* d_in, int64_t* d_out, int64_t n) {
  if (n <= 0) return 0;
  const int64_t ntiles = (n + kTile - 1) / kTile;
  c->scan_tmp.reserve((size_t)ntiles + 2);
  int64_t* tiles = c->scan_tmp.p;
  int64_t* total = c->scan_tmp.p + ntiles;
  hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)ntiles), dim3(kBS), 0, c->stream, d_in, n, tiles);
  hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBS), 0, c->stream, tiles, ntiles, total);
  hipLaunchKernelGGL(k_tile_apply, dim3((unsigned)ntiles), dim3(kBS), 0, c->stream, d_in, d_out, n, tiles);
  NGPB_HIP(hipGetLastError());
  int64_t h = 0;
  NGPB_HIP(hipMemcpyAsync(&h, total, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  return h;
}

void exclusive_scan_many(hipStream_t st, const std::vector<ScanSeg>& segs, int64_t* totals) {
  const int ns = (int)segs.size();
  if (ns == 0) return;
  std::vector<int64_t> t0((size_t)ns + 1, 0);
  for (int i = 0; i < ns; ++i) t0[i + 1] = t0[i] + (std::max<int64_t>(segs[i].n, 0) + kTile - 1) / kTile;
  const int64_t ntiles = t0[ns];
  NGPB_REQUIRE(ntiles < (int64_t)INT32_MAX, NGPB_E_ARG, "too many scan tiles");
  DBuf<ScanSeg> d_seg(st);
  DBuf<int64_t> d_t0(st), d_sums(st), d_tot(st);
  d_seg.upload(segs.data(), (size_t)ns, st);
  d_t0.upload(t0.data(), t0.size(), st);
  d_sums.resize((size_t)std::max<int64_t>(ntiles, 1));
  d_tot.resize((size_t)ns);
  if (ntiles > 0)
    hipLaunchKernelGGL(k_tile_sums_many, dim3((unsigned)ntiles), dim3(kBS), 0, st, d_seg.p, d_t0.p, ns, d_sums.p);
  hipLaunchKernelGGL(k_scan_single_many, dim3((unsigned)ns), dim3(kBS), 0, st, d_t0.p, d_sums.p, d_tot.p);
  if (ntiles > 0)
    hipLaunchKernelGGL(k_tile_apply_many, dim3((unsigned)ntiles), dim3(kBS), 0, st, d_seg.p, d_t0.p, ns, d_sums.p);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipMemcpyAsync(totals, d_tot.p, (size_t)ns * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  NGPB_HIP(hipStreamSynchronize(st));
}

int64_t compact_flags(ngpb_ctx* c, const uint8_t* d_flag, int64_t n, int32_t* d_out) {
  if (n <= 0) return 0;
  int64_t* cnt = tmp_buf<int64_t>(c, kTmpFlagCount, (size_t)n);
  hipLaunchKernelGGL(k_flags_to_counts, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, d_flag, n, cnt);
  const int64_t total = exclusive_scan(c, cnt, cnt, n);
  hipLaunchKernelGGL(k_scatter_flags, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, d_flag, cnt, n, d_out);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipStreamSynchronize(c->stream));
  return total;
}

}  // namespace ngpb
