// Atom location on the device: the owner leaf of each atom, as search_points /
// cerca_atomo / controlla_coordinate (src/pb_class.cpp:3633-3729) build lookup_table.
//
// The rule is the reference's half-open test on every axis, l - tol < x <= r - tol
// with l, r the coordinates of the leaf's corners 0 and 7; among the leaves that pass,
// the first in leaf order owns the atom (the first insertion into lookup_table wins);
// -1 when none passes.
//
// Leaves are bucketed on a uniform grid whose cell is the coarsest leaf edge, anchored
// at the domain's lower corner: an octree leaf never straddles such a cell, so each
// leaf sits in exactly one bucket (the one holding its centre). An atom tests the
// leaves of the 3 x 3 x 3 buckets around x + tol, which covers every leaf that can
// pass whatever the rounding of the bucket index, and keeps the smallest passing
// index, so the result does not depend on the order of the bucket lists.
#include <algorithm>
#include <cmath>

#include "ngpb_internal.h"

namespace ngpb {

MeshView mesh_view_of(const ngpb_ctx* c);

namespace {

struct LeafGrid {
  double g0[3];
  double cell;
  int64_t dim[3];
};

__device__ __forceinline__ int64_t bucket_of(const LeafGrid& G, const double p[3]) {
  int64_t b[3];
  for (int d = 0; d < 3; ++d) b[d] = min(max((int64_t)floor((p[d] - G.g0[d]) / G.cell), (int64_t)0), G.dim[d] - 1);
  return (b[2] * G.dim[1] + b[1]) * G.dim[0] + b[0];
}

// leaf extent (corners 0 and 7) and the largest edge, per block max
__global__ void k_leaf_extent(MeshView m, double* __restrict__ mn, double* __restrict__ mx, double* __restrict__ edge) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= m.L) return;
  const int32_t c0 = m.leaf_corner[8 * l], c7 = m.leaf_corner[8 * l + 7];
  double e = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double lo = dev::corner_coord(m, c0, d), hi = dev::corner_coord(m, c7, d);
    mn[3 * l + d] = lo;
    mx[3 * l + d] = hi;
    e = fmax(e, hi - lo);
  }
  edge[l] = e;
}

// Runs of consecutive lanes with the same bucket (leaves in Morton order mostly share
// one) take one atomic per run: the run's first lane adds the run length, and each
// lane's offset is its distance from that lane. Lanes past the end form no run.
struct LaneRun {
  int head;  // lane holding the run's first element
  int len;   // run length (valid on the head lane)
};

__device__ __forceinline__ LaneRun lane_run(bool valid, int32_t key) {
  const int lane = threadIdx.x & 63;
  const int32_t prev = __shfl(key, lane > 0 ? lane - 1 : 0, 64);
  const bool prev_valid = lane > 0;  // lanes are filled from lane 0, so a valid lane's predecessor is valid
  const bool head = valid && (!prev_valid || prev != key);
  const uint64_t heads = __ballot(head), vmask = __ballot(valid);
  // head of this lane's run: the highest head at or below the lane
  const uint64_t upto = heads & ((lane == 63) ? ~0ULL : ((2ULL << lane) - 1));
  LaneRun r;
  r.head = upto ? 63 - __clzll(upto) : 0;
  // run length on a head lane: up to the next head or the last valid lane
  const uint64_t above = heads & ~((lane == 63) ? ~0ULL : ((2ULL << lane) - 1));
  const int end = above ? __ffsll((unsigned long long)above) - 1 : (vmask ? 64 - __clzll(vmask) : 0);
  r.len = end - lane;
  return r;
}

__global__ void __launch_bounds__(256) k_leaf_bucket(MeshView m, LeafGrid G, const double* __restrict__ mn,
                                                     const double* __restrict__ mx, int32_t* __restrict__ bucket,
                                                     int32_t* __restrict__ cnt) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = l < m.L;
  int32_t b = -1;
  if (valid) {
    double ctr[3];
    for (int d = 0; d < 3; ++d) ctr[d] = 0.5 * (mn[3 * l + d] + mx[3 * l + d]);
    b = (int32_t)bucket_of(G, ctr);
    bucket[l] = b;
  }
  const LaneRun r = lane_run(valid, b);
  if (valid && r.head == (int)(threadIdx.x & 63)) atomicAdd(&cnt[b], r.len);
}

__global__ void __launch_bounds__(256) k_leaf_fill(int64_t L, const int32_t* __restrict__ bucket,
                                                   const int64_t* __restrict__ start, int32_t* __restrict__ cursor,
                                                   int32_t* __restrict__ list) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = l < L;
  const int32_t b = valid ? bucket[l] : -1;
  const int lane = threadIdx.x & 63;
  const LaneRun r = lane_run(valid, b);
  int32_t base = 0;
  if (valid && r.head == lane) base = atomicAdd(&cursor[b], r.len);
  base = __shfl(base, r.head, 64);
  if (valid) list[start[b] + base + (lane - r.head)] = (int32_t)l;
}

// one wave per atom: the lanes split each bucket's leaves, then the wave takes the
// smallest passing leaf index
constexpr int kLocateBS = 256;

__global__ void __launch_bounds__(kLocateBS) k_locate(int64_t A, const double* __restrict__ axyz, double tol,
                                                      LeafGrid G, const int64_t* __restrict__ start,
                                                      const int32_t* __restrict__ list, const double* __restrict__ mn,
                                                      const double* __restrict__ mx, int32_t* __restrict__ owner) {
  const int lane = threadIdx.x & 63;
  const int64_t a = (int64_t)blockIdx.x * (kLocateBS / 64) + (threadIdx.x >> 6);
  if (a >= A) return;  // the whole wave
  const double x[3] = {axyz[3 * a], axyz[3 * a + 1], axyz[3 * a + 2]};
  const double p[3] = {x[0] + tol, x[1] + tol, x[2] + tol};
  int64_t bc[3];
  for (int d = 0; d < 3; ++d) bc[d] = (int64_t)floor((p[d] - G.g0[d]) / G.cell);
  int32_t best = INT32_MAX;
  for (int64_t z = bc[2] - 1; z <= bc[2] + 1; ++z)
    for (int64_t y = bc[1] - 1; y <= bc[1] + 1; ++y)
      for (int64_t w = bc[0] - 1; w <= bc[0] + 1; ++w) {
        if (w < 0 || y < 0 || z < 0 || w >= G.dim[0] || y >= G.dim[1] || z >= G.dim[2]) continue;
        const int64_t b = (z * G.dim[1] + y) * G.dim[0] + w;
        for (int64_t k = start[b] + lane; k < start[b + 1]; k += 64) {
          const int32_t l = list[k];
          if (l >= best) continue;
          bool in = true;
          for (int d = 0; d < 3; ++d) in = in && (mn[3 * l + d] - tol < x[d]) && (x[d] <= mx[3 * l + d] - tol);
          if (in) best = l;
        }
      }
  for (int off = 32; off > 0; off >>= 1) best = min(best, __shfl_xor(best, off, 64));
  if (lane == 0) owner[a] = best == INT32_MAX ? -1 : best;
}

// per block g of kRedBlocks: min over mn[3 l + d], max over mx[3 l + d] (d = 0..2) and
// max over edge[l] for its contiguous range of leaves; then one block combines them
constexpr int kRedBlocks = 256;

__global__ void __launch_bounds__(256) k_extent_part(int64_t L, const double* __restrict__ mn,
                                                     const double* __restrict__ mx, const double* __restrict__ edge,
                                                     double* __restrict__ part) {
  __shared__ double s[7][256];
  double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
  const int64_t chunk = (L + kRedBlocks - 1) / kRedBlocks;
  const int64_t l0 = (int64_t)blockIdx.x * chunk, l1 = min(L, l0 + chunk);
  for (int64_t l = l0 + threadIdx.x; l < l1; l += 256) {
    for (int d = 0; d < 3; ++d) {
      v[d] = fmin(v[d], mn[3 * l + d]);
      v[3 + d] = fmax(v[3 + d], mx[3 * l + d]);
    }
    v[6] = fmax(v[6], edge[l]);
  }
  for (int k = 0; k < 7; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 7; ++k)
        s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], s[k][threadIdx.x + w])
                                  : fmax(s[k][threadIdx.x], s[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 7) part[threadIdx.x * kRedBlocks + blockIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_extent_fin(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double s[256];
  for (int k = 0; k < 7; ++k) {
    s[threadIdx.x] = part[k * kRedBlocks + threadIdx.x];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
        s[threadIdx.x] = k < 3 ? fmin(s[threadIdx.x], s[threadIdx.x + w]) : fmax(s[threadIdx.x], s[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = s[0];
    __syncthreads();
  }
}

__global__ void k_i32_to_i64(int64_t n, const int32_t* __restrict__ a, int64_t* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

}  // namespace

void run_locate_atoms(ngpb_ctx* c, int64_t A, const double* h_xyz, double tol, int32_t* h_owner) {
  const MeshView m = mesh_view_of(c);
  const int64_t L = c->L;
  if (A == 0) return;
  if (L == 0) {
    std::fill(h_owner, h_owner + A, -1);
    return;
  }
  DBuf<double> mn(c->stream), mx(c->stream), edge(c->stream), part(c->stream), ext(c->stream);
  mn.resize((size_t)3 * L);
  mx.resize((size_t)3 * L);
  edge.resize((size_t)L);
  part.resize((size_t)7 * kRedBlocks);
  ext.resize(7);
  hipLaunchKernelGGL(k_leaf_extent, dim3(blocks_for(L, 256)), dim3(256), 0, c->stream, m, mn.p, mx.p, edge.p);
  hipLaunchKernelGGL(k_extent_part, dim3(kRedBlocks), dim3(256), 0, c->stream, L, mn.p, mx.p, edge.p, part.p);
  hipLaunchKernelGGL(k_extent_fin, dim3(1), dim3(256), 0, c->stream, part.p, ext.p);
  NGPB_HIP(hipGetLastError());
  double e[7];
  NGPB_HIP(hipMemcpyAsync(e, ext.p, sizeof(e), hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  NGPB_REQUIRE(e[6] > 0.0 && std::isfinite(e[6]), NGPB_E_ARG, "leaves without extent");
  LeafGrid G;
  G.cell = e[6];
  int64_t nb = 1;
  for (int d = 0; d < 3; ++d) {
    G.g0[d] = e[d];
    G.dim[d] = std::max<int64_t>(1, (int64_t)std::floor((e[3 + d] - e[d]) / G.cell) + 1);
    nb *= G.dim[d];
  }
  NGPB_REQUIRE(nb < (int64_t)INT32_MAX, NGPB_E_ARG, "leaf bucket grid too large");
  DBuf<int32_t> bucket(c->stream), cnt(c->stream), cursor(c->stream), list(c->stream), owner(c->stream);
  DBuf<int64_t> start(c->stream);
  bucket.resize((size_t)L);
  cnt.resize((size_t)nb);
  cursor.resize((size_t)nb);
  list.resize((size_t)L);
  start.resize((size_t)nb + 1);
  owner.resize((size_t)A);
  NGPB_HIP(hipMemsetAsync(cnt.p, 0, nb * sizeof(int32_t), c->stream));
  NGPB_HIP(hipMemsetAsync(cursor.p, 0, nb * sizeof(int32_t), c->stream));
  hipLaunchKernelGGL(k_leaf_bucket, dim3(blocks_for(L, 256)), dim3(256), 0, c->stream, m, G, mn.p, mx.p, bucket.p,
                     cnt.p);
  hipLaunchKernelGGL(k_i32_to_i64, dim3(blocks_for(nb, 256)), dim3(256), 0, c->stream, nb, cnt.p, start.p);
  const int64_t total = exclusive_scan(c, start.p, start.p, nb);
  NGPB_HIP(hipMemcpyAsync(start.p + nb, &total, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_leaf_fill, dim3(blocks_for(L, 256)), dim3(256), 0, c->stream, L, bucket.p, start.p, cursor.p,
                     list.p);
  DBuf<double> axyz(c->stream);
  axyz.upload(h_xyz, (size_t)(3 * A), c->stream);
  hipLaunchKernelGGL(k_locate, dim3(blocks_for(A, kLocateBS / 64)), dim3(kLocateBS), 0, c->stream, A, axyz.p, tol, G,
                     start.p, list.p, mn.p, mx.p, owner.p);
  NGPB_HIP(hipGetLastError());
  owner.download(h_owner, (size_t)A, c->stream);
  NGPB_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace ngpb
