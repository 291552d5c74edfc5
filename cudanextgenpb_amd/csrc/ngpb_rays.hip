// Axis-aligned ray casting against a union of balls on the device: the ray table that
// ray_cache_t (src/raytracer.cpp:22-63) fills on demand by casting through NanoShaper
// (castAxisOrientedRay, src/raytracer.cpp:337-467), here for the analytic union of
// atom spheres and for every grid line through a mesh vertex at once.
//
// Per axis d (d1 < d2 the other two axes):
//   lines     the distinct (coordinate d1, coordinate d2) pairs of all vertices
//             (nodes and hanging nodes), in lexicographic order: two stable radix
//             passes over order-preserving 64-bit images of the coordinates, then
//             duplicates dropped;
//   balls     binned on a 2D grid of cell max(2 r_max, 1) in the (d1, d2) plane; a
//             line visits the 3 x 3 bins around it;
//   crossings each ball the line pierces (r^2 - du^2 - dv^2 > 0) gives the interval
//             c -+ sqrt(.); the intervals are sorted by (start, end, ball), merged
//             where they overlap or touch, and the end points of the merged intervals
//             are the crossings, with the outward normal (p - c_a) / r_a of the ball
//             owning each end point. Lines without crossings are dropped.
// The arithmetic and the order of every step are those of the host caster of the
// synthetic workloads (csrc/workload.cpp), so both give the same table bit for bit.
#include <algorithm>
#include <cmath>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "ngpb_internal.h"

namespace ngpb {

namespace {

// order-preserving image of a double (total order of the non-NaN values)
__device__ __forceinline__ uint64_t ordered_bits(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

__device__ __forceinline__ double vertex_coord(const double* node_xyz, const double* hang_xyz, int64_t N, int64_t v,
                                               int d) {
  return v < N ? node_xyz[3 * v + d] : hang_xyz[3 * (v - N) + d];
}

__global__ void k_line_keys(int64_t NV, int64_t N, const double* __restrict__ node_xyz,
                            const double* __restrict__ hang_xyz, int d1, int d2, uint64_t* __restrict__ ku,
                            uint64_t* __restrict__ kv) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= NV) return;
  ku[v] = ordered_bits(vertex_coord(node_xyz, hang_xyz, N, v, d1));
  kv[v] = ordered_bits(vertex_coord(node_xyz, hang_xyz, N, v, d2));
}

__global__ void k_line_first(int64_t n, const uint64_t* __restrict__ ku, const uint64_t* __restrict__ kv,
                             uint8_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0 || ku[i] != ku[i - 1] || kv[i] != kv[i - 1]) ? 1 : 0;
}

struct BallGrid {
  const double* xyz;
  const double* r;
  const int32_t* bin_ptr;    // gn0 * gn1 + 1
  const int32_t* bin_ball;   // balls in increasing index within each bin
  double lo0, lo1, cell;
  int gn0, gn1;
};

// line coordinates from the sorted keys (the coordinate of the sorted vertex)
__device__ __forceinline__ double key_coord(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffULL) : ~k;
  return __longlong_as_double((long long)b);
}

// visits the candidate balls of line (u, v) in the host caster's order; f(a, s) for
// each ball pierced, s = sqrt(r^2 - du^2 - dv^2)
template <class F>
__device__ __forceinline__ void for_each_pierced(const BallGrid& G, double u, double v, int axis_d1, int axis_d2,
                                                 F&& f) {
  const int b0 = (int)floor((u - G.lo0) / G.cell), b1 = (int)floor((v - G.lo1) / G.cell);
  for (int y = b1 - 1; y <= b1 + 1; ++y)
    for (int x = b0 - 1; x <= b0 + 1; ++x) {
      if (x < 0 || y < 0 || x >= G.gn0 || y >= G.gn1) continue;
      const int64_t bin = (int64_t)y * G.gn0 + x;
      for (int32_t k = G.bin_ptr[bin]; k < G.bin_ptr[bin + 1]; ++k) {
        const int32_t a = G.bin_ball[k];
        const double du = u - G.xyz[3 * a + axis_d1], dv = v - G.xyz[3 * a + axis_d2];
        const double q = G.r[a] * G.r[a] - (du * du + dv * dv);
        if (q <= 0) continue;
        f(a, sqrt(q));
      }
    }
}

__global__ void k_count_pierced(int64_t nl, const uint64_t* __restrict__ lu, const uint64_t* __restrict__ lv,
                                BallGrid G, int d1, int d2, int64_t* __restrict__ cnt) {
  const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nl) return;
  int64_t c = 0;
  for_each_pierced(G, key_coord(lu[li]), key_coord(lv[li]), d1, d2, [&](int32_t, double) { ++c; });
  cnt[li] = c;
}

struct Iv {
  double a, b;
  int32_t ia, ib;
};

__device__ __forceinline__ bool iv_less(const Iv& p, const Iv& q) {
  return p.a < q.a || (p.a == q.a && (p.b < q.b || (p.b == q.b && p.ia < q.ia)));
}

// Intervals of each line, sorted under (start, end, ball) -- a total order, so the
// result is the host caster's std::sort result -- and merged where they overlap or
// touch; mcnt = 2 x merged intervals, written to the line's slots of iv. One wave per
// line: the pierced intervals are gathered into LDS (lanes split each bin's balls,
// ballot ranks give the slots), sorted by a bitonic network (padding sorts last) and
// merged by one lane in a single pass. A line with more than kLineCap intervals is
// handled by one lane in global memory (insertion sort, then the same merge).
constexpr int kLineCap = 1024;

__global__ void __launch_bounds__(64) k_merge_pierced_wave(int64_t nl, const uint64_t* __restrict__ lu,
                                                           const uint64_t* __restrict__ lv, BallGrid G, int axis,
                                                           int d1, int d2, const int64_t* __restrict__ off,
                                                           int64_t niv, Iv* __restrict__ iv,
                                                           int64_t* __restrict__ mcnt) {
  __shared__ double sa[kLineCap], sb[kLineCap];
  __shared__ int32_t sia[kLineCap];
  const int64_t li = blockIdx.x;
  if (li >= nl) return;
  const int lane = threadIdx.x;
  const int64_t o0 = off[li], k = (li + 1 < nl ? off[li + 1] : niv) - o0;
  Iv* s = iv + o0;
  if (k > kLineCap) {
    if (lane != 0) return;
    int64_t kk = 0;
    for_each_pierced(G, key_coord(lu[li]), key_coord(lv[li]), d1, d2, [&](int32_t a, double sq) {
      const double c = G.xyz[3 * a + axis];
      s[kk++] = Iv{c - sq, c + sq, a, a};
    });
    for (int64_t i = 1; i < kk; ++i) {
      const Iv t = s[i];
      int64_t j = i - 1;
      while (j >= 0 && iv_less(t, s[j])) {
        s[j + 1] = s[j];
        --j;
      }
      s[j + 1] = t;
    }
    int64_t m = 0;
    for (int64_t i = 0; i < kk; ++i) {
      const Iv t = s[i];
      if (m > 0 && t.a <= s[m - 1].b) {
        if (t.b > s[m - 1].b) {
          s[m - 1].b = t.b;
          s[m - 1].ib = t.ib;
        }
      } else {
        s[m++] = t;
      }
    }
    mcnt[li] = 2 * m;
    return;
  }
  // gather: the same candidate bins and arithmetic as for_each_pierced
  const double u = key_coord(lu[li]), v = key_coord(lv[li]);
  const int b0 = (int)floor((u - G.lo0) / G.cell), b1 = (int)floor((v - G.lo1) / G.cell);
  const uint64_t below = (1ULL << lane) - 1;
  int n = 0;
  for (int y = b1 - 1; y <= b1 + 1; ++y)
    for (int x = b0 - 1; x <= b0 + 1; ++x) {
      if (x < 0 || y < 0 || x >= G.gn0 || y >= G.gn1) continue;
      const int64_t bin = (int64_t)y * G.gn0 + x;
      for (int32_t base = G.bin_ptr[bin]; base < G.bin_ptr[bin + 1]; base += 64) {
        const int32_t kb = base + lane;
        bool hit = false;
        double lo = 0, hi = 0;
        int32_t a = 0;
        if (kb < G.bin_ptr[bin + 1]) {
          a = G.bin_ball[kb];
          const double du = u - G.xyz[3 * a + d1], dv = v - G.xyz[3 * a + d2];
          const double q = G.r[a] * G.r[a] - (du * du + dv * dv);
          if (q > 0) {
            const double sq = sqrt(q), c = G.xyz[3 * a + axis];
            lo = c - sq;
            hi = c + sq;
            hit = true;
          }
        }
        const uint64_t mask = __ballot(hit);
        const int pos = n + __popcll(mask & below);
        if (hit && pos < kLineCap) {
          sa[pos] = lo;
          sb[pos] = hi;
          sia[pos] = a;
        }
        n += __popcll(mask);
      }
    }
  n = min(n, kLineCap);  // equal to the counted k (same test), so never clipped
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = n + lane; i < P; i += 64) {
    sa[i] = INFINITY;
    sb[i] = INFINITY;
    sia[i] = INT32_MAX;
  }
  __syncthreads();
  for (int w = 2; w <= P; w <<= 1)
    for (int j = w >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < P; i += 64) {
        const int x = i ^ j;
        if (x <= i) continue;
        const Iv p{sa[i], sb[i], sia[i], sia[i]}, q{sa[x], sb[x], sia[x], sia[x]};
        if (iv_less(q, p) == ((i & w) == 0)) {
          sa[i] = q.a;
          sb[i] = q.b;
          sia[i] = q.ia;
          sa[x] = p.a;
          sb[x] = p.b;
          sia[x] = p.ia;
        }
      }
      __syncthreads();
    }
  if (lane != 0) return;
  int64_t m = 0;
  Iv last{0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const Iv t{sa[i], sb[i], sia[i], sia[i]};
    if (m > 0 && t.a <= last.b) {
      if (t.b > last.b) {
        last.b = t.b;
        last.ib = t.ib;
      }
    } else {
      if (m > 0) s[m - 1] = last;
      last = t;
      ++m;
    }
  }
  if (m > 0) s[m - 1] = last;
  mcnt[li] = 2 * m;
}

__global__ void k_nonzero(int64_t n, const int64_t* __restrict__ c, uint8_t* __restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = c[i] > 0 ? 1 : 0;
}

// ray rr = kept line li: key, crossing count (the pointers are its scan)
__global__ void k_ray_keys(int64_t nr, const int32_t* __restrict__ kept, const uint64_t* __restrict__ lu,
                           const uint64_t* __restrict__ lv, const int64_t* __restrict__ mcnt, double* __restrict__ key,
                           int64_t* __restrict__ ptr) {
  const int64_t rr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (rr >= nr) return;
  const int64_t li = kept[rr];
  key[2 * rr] = key_coord(lu[li]);
  key[2 * rr + 1] = key_coord(lv[li]);
  ptr[rr] = mcnt[li];
}

__global__ void k_ray_fill(int64_t nr, const int32_t* __restrict__ kept, const int64_t* __restrict__ ptr,
                           const int64_t* __restrict__ off, const Iv* __restrict__ iv, const double* __restrict__ key,
                           BallGrid G, int axis, int d1, int d2, double* __restrict__ inters,
                           double* __restrict__ normals) {
  const int64_t rr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (rr >= nr) return;
  const Iv* s = iv + off[kept[rr]];
  const int64_t m = (ptr[rr + 1] - ptr[rr]) / 2;
  const double u = key[2 * rr], v = key[2 * rr + 1];
  int64_t o = ptr[rr];
  for (int64_t i = 0; i < m; ++i)
    for (int e = 0; e < 2; ++e) {
      const double tc = e == 0 ? s[i].a : s[i].b;
      const int32_t a = e == 0 ? s[i].ia : s[i].ib;
      double p[3];
      p[axis] = tc;
      p[d1] = u;
      p[d2] = v;
      inters[o] = tc;
      for (int d = 0; d < 3; ++d) normals[3 * o + d] = (p[d] - G.xyz[3 * a + d]) / G.r[a];
      ++o;
    }
}

__global__ void k_gather_u64(int64_t n, const int32_t* __restrict__ idx, const uint64_t* __restrict__ src,
                             uint64_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// keys of every line of the grid along axis 2 through (origin0 + i hf, origin1 + j hf),
// i <= n0, j <= n1, i major (the order of the distinct lines of a mesh)
__global__ void k_grid_zline_keys(int64_t n0, int64_t n1, double o0, double o1, double hf, uint64_t* __restrict__ lu,
                                  uint64_t* __restrict__ lv) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (n0 + 1) * (n1 + 1)) return;
  const int64_t i = l / (n1 + 1), j = l % (n1 + 1);
  lu[l] = ordered_bits(o0 + (double)i * hf);
  lv[l] = ordered_bits(o1 + (double)j * hf);
}

}  // namespace

struct SesDevice;
SesDevice* ses_prepare(ngpb_ctx* c, int64_t na, const double* xyz, const double* r, double rp);
void ses_release(SesDevice* D);
void ses_cast_lines(ngpb_ctx* c, const SesDevice* D, int axis, int64_t nl, const uint64_t* lu, const uint64_t* lv,
                    const double* h_xyz);

namespace {

// Rays of the nl distinct lines (lu, lv: order-preserving keys, sorted) of one axis
// into c->rays[axis]: SES rays when ses is given, sphere-union rays otherwise (balls
// bxyz / br on the device, h_xyz / h_r on the host).
void cast_axis_lines(ngpb_ctx* c, int64_t na, const double* h_xyz, const double* h_r, const SesDevice* ses,
                     const double* bxyz, const double* br, int axis, int64_t nl, const uint64_t* lu,
                     const uint64_t* lv) {
  const int d1 = axis == 0 ? 1 : 0;
  const int d2 = axis == 2 ? 1 : 2;
  auto& R = c->rays[axis];
  if (nl == 0 || na == 0) {
    R.n = R.k = 0;
    R.key.resize(2);
    R.ptr.resize(1);
    R.inters.resize(1);
    R.normals.resize(3);
    NGPB_HIP(hipMemsetAsync(R.ptr.p, 0, sizeof(int64_t), c->stream));
    NGPB_HIP(hipStreamSynchronize(c->stream));
    return;
  }
  if (ses) {
    ses_cast_lines(c, ses, axis, nl, lu, lv, h_xyz);
    return;
  }
  // balls and their bins (built on the host as the host caster does)
  double rmax = 0;
  for (int64_t a = 0; a < na; ++a) rmax = std::max(rmax, h_r[a]);
  const double cell = std::max(2.0 * rmax, 1.0);
  double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
  for (int64_t a = 0; a < na; ++a) {
    lo[0] = std::min(lo[0], h_xyz[3 * a + d1]);
    hi[0] = std::max(hi[0], h_xyz[3 * a + d1]);
    lo[1] = std::min(lo[1], h_xyz[3 * a + d2]);
    hi[1] = std::max(hi[1], h_xyz[3 * a + d2]);
  }
  const int gn0 = (int)std::floor((hi[0] - lo[0]) / cell) + 1, gn1 = (int)std::floor((hi[1] - lo[1]) / cell) + 1;
  std::vector<int32_t> bptr((size_t)gn0 * gn1 + 1, 0), bball((size_t)std::max<int64_t>(na, 1));
  std::vector<int32_t> bin_of((size_t)std::max<int64_t>(na, 1));
  for (int64_t a = 0; a < na; ++a) {
    const int b0 = (int)std::floor((h_xyz[3 * a + d1] - lo[0]) / cell);
    const int b1 = (int)std::floor((h_xyz[3 * a + d2] - lo[1]) / cell);
    bin_of[a] = b1 * gn0 + b0;
    ++bptr[(size_t)bin_of[a] + 1];
  }
  for (size_t i = 1; i < bptr.size(); ++i) bptr[i] += bptr[i - 1];
  {
    std::vector<int32_t> cur(bptr.begin(), bptr.end() - 1);
    for (int64_t a = 0; a < na; ++a) bball[(size_t)cur[(size_t)bin_of[a]]++] = (int32_t)a;
  }
  DBuf<int32_t> d_bptr(c->stream), d_bball(c->stream);
  d_bptr.upload(bptr.data(), bptr.size(), c->stream);
  d_bball.upload(bball.data(), bball.size(), c->stream);
  BallGrid G{bxyz, br, d_bptr.p, d_bball.p, lo[0], lo[1], cell, gn0, gn1};
  // pierced balls per line, their intervals sorted and merged
  DBuf<int64_t> off(c->stream), mcnt(c->stream);
  off.resize((size_t)nl);
  mcnt.resize((size_t)nl);
  hipLaunchKernelGGL(k_count_pierced, dim3(blocks_for(nl, 128)), dim3(128), 0, c->stream, nl, lu, lv, G, d1, d2,
                     off.p);
  NGPB_HIP(hipGetLastError());
  const int64_t niv = exclusive_scan(c, off.p, off.p, nl);
  DBuf<Iv> iv(c->stream);
  iv.resize((size_t)std::max<int64_t>(niv, 1));
  hipLaunchKernelGGL(k_merge_pierced_wave, dim3((unsigned)nl), dim3(64), 0, c->stream, nl, lu, lv, G, axis, d1, d2,
                     off.p, niv, iv.p, mcnt.p);
  NGPB_HIP(hipGetLastError());
  // rays = lines with crossings
  DBuf<uint8_t> keep(c->stream);
  keep.resize((size_t)nl);
  hipLaunchKernelGGL(k_nonzero, dim3(blocks_for(nl, 256)), dim3(256), 0, c->stream, nl, mcnt.p, keep.p);
  DBuf<int32_t> kept(c->stream);
  kept.resize((size_t)nl);
  const int64_t nr = compact_flags(c, keep.p, nl, kept.p);
  R.n = nr;
  R.key.resize((size_t)2 * std::max<int64_t>(nr, 1));
  R.ptr.resize((size_t)nr + 1);
  if (nr > 0) {
    hipLaunchKernelGGL(k_ray_keys, dim3(blocks_for(nr, 256)), dim3(256), 0, c->stream, nr, kept.p, lu, lv, mcnt.p,
                       R.key.p, R.ptr.p);
    R.k = exclusive_scan(c, R.ptr.p, R.ptr.p, nr);
  } else {
    R.k = 0;
  }
  NGPB_HIP(hipMemcpyAsync(R.ptr.p + nr, &R.k, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  R.inters.resize((size_t)std::max<int64_t>(R.k, 1));
  R.normals.resize((size_t)3 * std::max<int64_t>(R.k, 1));
  if (nr > 0)
    hipLaunchKernelGGL(k_ray_fill, dim3(blocks_for(nr, 128)), dim3(128), 0, c->stream, nr, kept.p, R.ptr.p, off.p,
                       iv.p, R.key.p, G, axis, d1, d2, R.inters.p, R.normals.p);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipStreamSynchronize(c->stream));
}

// solvent-excluded surface data for probe > 0 (released with the holder)
struct SesHolder {
  SesDevice* d = nullptr;
  ~SesHolder() {
    if (d) ses_release(d);
  }
};

}  // namespace

void run_cast_rays(ngpb_ctx* c, int64_t na, const double* h_xyz, const double* h_r, double probe) {
  const int64_t N = c->N, NV = c->N + c->H;
  SesHolder ses;
  if (probe > 0.0 && na > 0) ses.d = ses_prepare(c, na, h_xyz, h_r, probe);
  DBuf<double> bxyz(c->stream), br(c->stream);
  if (na > 0) {
    bxyz.upload(h_xyz, (size_t)(3 * na), c->stream);
    br.upload(h_r, (size_t)na, c->stream);
  }
  for (int axis = 0; axis < 3; ++axis) {
    const int d1 = axis == 0 ? 1 : 0;
    const int d2 = axis == 2 ? 1 : 2;
    // distinct lines in lexicographic order
    DBuf<uint64_t> ku(c->stream), kv(c->stream), su(c->stream), sv(c->stream);
    ku.resize((size_t)NV);
    kv.resize((size_t)NV);
    su.resize((size_t)NV);
    sv.resize((size_t)NV);
    hipLaunchKernelGGL(k_line_keys, dim3(blocks_for(NV, 256)), dim3(256), 0, c->stream, NV, N, c->node_xyz.p,
                       c->hang_xyz.p, d1, d2, ku.p, kv.p);
    NGPB_HIP(hipGetLastError());
    size_t tb1 = 0, tb2 = 0;
    NGPB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, kv.p, sv.p, ku.p, su.p, (int)NV, 0, 64, c->stream));
    NGPB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, su.p, ku.p, sv.p, kv.p, (int)NV, 0, 64, c->stream));
    DBuf<uint8_t> tmp(c->stream);
    tmp.resize(std::max(tb1, tb2));
    // by the second coordinate, then stably by the first
    NGPB_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb1, kv.p, sv.p, ku.p, su.p, (int)NV, 0, 64, c->stream));
    NGPB_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb2, su.p, ku.p, sv.p, kv.p, (int)NV, 0, 64, c->stream));
    DBuf<uint8_t> first(c->stream);
    first.resize((size_t)NV);
    hipLaunchKernelGGL(k_line_first, dim3(blocks_for(NV, 256)), dim3(256), 0, c->stream, NV, ku.p, kv.p, first.p);
    DBuf<int32_t> uniq(c->stream);
    uniq.resize((size_t)NV);
    const int64_t nl = compact_flags(c, first.p, NV, uniq.p);
    // line coordinates (keys) of the distinct lines
    DBuf<uint64_t> lu(c->stream), lv(c->stream);
    lu.resize((size_t)std::max<int64_t>(nl, 1));
    lv.resize((size_t)std::max<int64_t>(nl, 1));
    if (nl > 0) {
      hipLaunchKernelGGL(k_gather_u64, dim3(blocks_for(nl, 256)), dim3(256), 0, c->stream, nl, uniq.p, ku.p, lu.p);
      hipLaunchKernelGGL(k_gather_u64, dim3(blocks_for(nl, 256)), dim3(256), 0, c->stream, nl, uniq.p, kv.p, lv.p);
      NGPB_HIP(hipGetLastError());
    }
    cast_axis_lines(c, na, h_xyz, h_r, ses.d, bxyz.p, br.p, axis, nl, lu.p, lv.p);
  }
}

void cast_grid_zlines(ngpb_ctx* c, const double* origin, double hf, int64_t n0, int64_t n1, int64_t na,
                      const double* h_xyz, const double* h_r, double probe) {
  SesHolder ses;
  if (probe > 0.0 && na > 0) ses.d = ses_prepare(c, na, h_xyz, h_r, probe);
  DBuf<double> bxyz(c->stream), br(c->stream);
  if (na > 0) {
    bxyz.upload(h_xyz, (size_t)(3 * na), c->stream);
    br.upload(h_r, (size_t)na, c->stream);
  }
  const int64_t nl = (n0 + 1) * (n1 + 1);
  DBuf<uint64_t> lu(c->stream), lv(c->stream);
  lu.resize((size_t)nl);
  lv.resize((size_t)nl);
  hipLaunchKernelGGL(k_grid_zline_keys, dim3(blocks_for(nl, 256)), dim3(256), 0, c->stream, n0, n1, origin[0],
                     origin[1], hf, lu.p, lv.p);
  NGPB_HIP(hipGetLastError());
  cast_axis_lines(c, na, h_xyz, h_r, ses.d, bxyz.p, br.p, 2, nl, lu.p, lv.p);
}

}  // namespace ngpb
