// Octree mesh generation on the device: the leaf/corner/node view of a 2:1 balanced
// linear octree that tmesh_3d hands to the discretization (p(d,i), gt(i),
// is_hanging(i), num_parents(i), gparent(j,i); src/pb_class.cpp:1838-1852,2024-2036),
// built where init_tmesh / refine_surface build it on the host with p4est
// (src/pb_class.cpp:1289-1391,1467-1745).
//
// p4est and NanoShaper are not available, so the refinement rule is the one of the
// synthetic workload generator (csrc/workload.cpp, ngpbw_mesh_octree), and this file
// produces exactly its mesh:
//
//   refine   root cells of level smax tile the box; level by level every cell of
//            level s >= 1 is split when some atom sphere comes closer to the cell
//            centre than band[s-1] + half the cell diagonal
//            (|x - c_a| - r_a < thr for some atom a);
//   balance  on a dense map of the finest cells holding the level of their leaf,
//            every leaf of level s >= 2 touching (face, edge or corner) a leaf of
//            level < s - 1 is split; Jacobi sweeps until none is (at most 64);
//   order    leaves sorted by the Morton key of their anchor (unique);
//   vertices corners of leaves are regular; edge midpoints and face centres of a
//            leaf that are corners of a smaller leaf are hanging, with the edge's 2
//            or the face's 4 corners as parents; both kinds are numbered in
//            lexicographic order (x fastest) of their finest-grid coordinates;
//   boundary regular nodes on the faces of the box, in node order.
//
// Every step is a pass over cells, leaves or vertices; no step depends on the order
// in which threads run: the split tests are pure functions of the leaf, the level
// map is read in one kernel and written in the next, the Morton keys are unique,
// numbering is a prefix sum, and every writer of a hanging node's parents writes the
// same values (the parents are fixed by the point's position and the size of the
// leaves it is an edge midpoint or face centre of).
#include <algorithm>
#include <cmath>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "ngpb_internal.h"

namespace ngpb {

void install_device_mesh(ngpb_ctx* c, int64_t N, int64_t H, int64_t L, int64_t NB);

namespace {

constexpr int kMaxLevels = 10;
constexpr int kBalanceSweeps = 64;

struct OLeaf {
  int32_t x, y, z;  // anchor in finest cells
  int32_t s;        // log2 of the edge in finest cells
};

struct Geom {
  double origin[3];
  double hf;
  int64_t n[3];                // finest cells per axis
  int64_t vn[3];               // vertices per axis (n + 1)
  double band[kMaxLevels];     // band[s - 1]: split distance for a cell of level s
};

__device__ __forceinline__ double gcoord(const Geom& g, int d, int64_t i) { return g.origin[d] + (double)i * g.hf; }
__device__ __forceinline__ int64_t cell_at(const Geom& g, int64_t x, int64_t y, int64_t z) {
  return (z * g.n[1] + y) * g.n[0] + x;
}
__device__ __forceinline__ int64_t vert_at(const Geom& g, int64_t x, int64_t y, int64_t z) {
  return (z * g.vn[1] + y) * g.vn[0] + x;
}

// atoms binned on a uniform grid of edge 4 A over their bounding box, with the
// bounding sphere of the atom surfaces as a first rejection test
struct AtomBins {
  const double* xyz;
  const double* r;
  const int32_t* ptr;   // bins + 1
  const int32_t* atom;  // atoms by bin
  const double* box;    // per bin: min corner, max corner of its atom centres, largest radius
  double lo[3];
  double cell;
  int dim[3];
  double rmax;
  double bc[3];
  double brad;
  int na;
};

__device__ __forceinline__ int clamp_bin(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

// true iff min_a (|p - c_a| - r_a) < thr
__device__ bool surface_below(const AtomBins& B, const double p[3], double thr) {
  if (B.na == 0) return false;
  double dd = 0.0;
  for (int d = 0; d < 3; ++d) dd += (p[d] - B.bc[d]) * (p[d] - B.bc[d]);
  if (sqrt(dd) - B.brad >= thr) return false;
  const double rr = fmax(thr + B.rmax, 0.0);
  int c0[3], c1[3];
  for (int d = 0; d < 3; ++d) {
    c0[d] = clamp_bin((int)floor((p[d] - rr - B.lo[d]) / B.cell), B.dim[d]);
    c1[d] = clamp_bin((int)floor((p[d] + rr - B.lo[d]) / B.cell), B.dim[d]);
  }
  for (int z = c0[2]; z <= c1[2]; ++z)
    for (int y = c0[1]; y <= c1[1]; ++y)
      for (int x = c0[0]; x <= c1[0]; ++x) {
        const int64_t b = ((int64_t)z * B.dim[1] + y) * B.dim[0] + x;
        if (B.ptr[b] == B.ptr[b + 1]) continue;
        // skip a bin whose atoms are all clearly beyond thr: |p - c| >= the distance
        // from p to the box of the bin's centres; the margin keeps every skipped atom
        // one whose own test is false whatever its rounding
        const double* bx = B.box + 7 * b;
        double q = 0.0;
        for (int d = 0; d < 3; ++d) {
          const double t = fmax(fmax(bx[d] - p[d], p[d] - bx[3 + d]), 0.0);
          q += t * t;
        }
        const double lower = sqrt(q) - bx[6];
        if (lower >= thr + 1e-9 * (fabs(thr) + fabs(lower) + 1.0)) continue;
        for (int32_t k = B.ptr[b]; k < B.ptr[b + 1]; ++k) {
          const int32_t a = B.atom[k];
          const double dx = p[0] - B.xyz[3 * a], dy = p[1] - B.xyz[3 * a + 1], dz = p[2] - B.xyz[3 * a + 2];
          if (sqrt(dx * dx + dy * dy + dz * dz) - B.r[a] < thr) return true;
        }
      }
  return false;
}

constexpr int64_t kKeep = 1, kSplit = (int64_t)1 << 32;  // packed counters: low keep, high split

__global__ void k_roots(Geom g, int smax, int64_t nr0, int64_t nr1, int64_t nroot, OLeaf* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nroot) return;
  const int32_t R = 1 << smax;
  out[i] = OLeaf{(int32_t)(i % nr0) * R, (int32_t)((i / nr0) % nr1) * R, (int32_t)(i / (nr0 * nr1)) * R, smax};
}

__global__ void k_refine_flag(const OLeaf* __restrict__ work, int64_t nw, Geom g, AtomBins B,
                              uint8_t* __restrict__ flag, int64_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const OLeaf l = work[i];
  bool split = false;
  if (l.s > 0) {
    const double sz = (double)((int64_t)1 << l.s) * g.hf;
    const double c[3] = {gcoord(g, 0, l.x) + 0.5 * sz, gcoord(g, 1, l.y) + 0.5 * sz, gcoord(g, 2, l.z) + 0.5 * sz};
    const double half_diag = 0.5 * sqrt(3.0) * sz;
    split = surface_below(B, c, g.band[l.s - 1] + half_diag);
  }
  flag[i] = split ? 1 : 0;
  cnt[i] = split ? kSplit : kKeep;
}

// kept leaves to keep[keep_off + low(off)], the 8 children of split ones to
// child[8 high(off) + k] in corner order (bit 0 x, bit 1 y, bit 2 z)
__global__ void k_split_emit(const OLeaf* __restrict__ work, int64_t nw, const uint8_t* __restrict__ flag,
                       const int64_t* __restrict__ off, OLeaf* __restrict__ keep, OLeaf* __restrict__ child) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const OLeaf l = work[i];
  const int64_t o = off[i];
  if (!flag[i]) {
    keep[o & 0xffffffffLL] = l;
    return;
  }
  const int32_t h = 1 << (l.s - 1);
  OLeaf* dst = child + 8 * (o >> 32);
  for (int k = 0; k < 8; ++k)
    dst[k] = OLeaf{l.x + ((k & 1) ? h : 0), l.y + ((k & 2) ? h : 0), l.z + ((k & 4) ? h : 0), l.s - 1};
}

// level map: every finest cell of leaf l holds l.s (minus one for split leaves, which
// paints their children)
__global__ void k_paint(const OLeaf* __restrict__ lv, int64_t L, const uint8_t* __restrict__ split, Geom g,
                        int8_t* __restrict__ lev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  if (split && !split[i]) return;
  const int8_t v = (int8_t)(split ? l.s - 1 : l.s);
  const int32_t S = 1 << l.s;
  for (int32_t z = l.z; z < l.z + S; ++z)
    for (int32_t y = l.y; y < l.y + S; ++y) {
      int8_t* row = lev + cell_at(g, l.x, y, z);
      for (int32_t x = 0; x < S; ++x) row[x] = v;
    }
}

// a leaf of level s >= 2 violates 2:1 balance when a finest cell in the layer around
// it (faces, edges and corners) belongs to a leaf of level < s - 1
__global__ void k_balance_flag(const OLeaf* __restrict__ lv, int64_t L, Geom g, const int8_t* __restrict__ lev,
                               uint8_t* __restrict__ flag, int64_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  bool split = false;
  if (l.s >= 2) {
    const int32_t S = 1 << l.s;
    const int lim = l.s - 1;
    const int64_t z0 = max((int64_t)l.z - 1, (int64_t)0), z1 = min((int64_t)l.z + S, g.n[2] - 1);
    const int64_t y0 = max((int64_t)l.y - 1, (int64_t)0), y1 = min((int64_t)l.y + S, g.n[1] - 1);
    const int64_t x0 = max((int64_t)l.x - 1, (int64_t)0), x1 = min((int64_t)l.x + S, g.n[0] - 1);
    for (int64_t z = z0; z <= z1 && !split; ++z) {
      const bool zin = z >= l.z && z < l.z + S;
      for (int64_t y = y0; y <= y1 && !split; ++y) {
        const bool yin = y >= l.y && y < l.y + S;
        const int8_t* row = lev + cell_at(g, 0, y, z);
        if (zin && yin) {  // interior row: only the two cells beside the leaf
          if (l.x - 1 >= 0 && row[l.x - 1] < lim) split = true;
          if ((int64_t)l.x + S < g.n[0] && row[l.x + S] < lim) split = true;
        } else {
          for (int64_t x = x0; x <= x1; ++x)
            if (row[x] < lim) {
              split = true;
              break;
            }
        }
      }
    }
  }
  flag[i] = split ? 1 : 0;
  cnt[i] = split ? kSplit : kKeep;
}

// ---- marker-driven refinement (refine_surface, src/pb_class.cpp:1467-1745)
// In/out flag of a grid vertex as is_in_ns_surf gives it for the z direction
// (src/pb_class.cpp:769-806): the parity of the crossings at or below its z on the
// z-line through it. Z holds the rays of every grid z-line; lmap maps grid line
// i (n1 + 1) + j to its ray (-1: no crossings).
struct ZRays {
  const int32_t* lmap;
  const int64_t* ptr;
  const double* inters;
};

__device__ int zflag(const Geom& g, const ZRays& Z, int64_t x, int64_t y, int64_t z) {
  const int32_t r = Z.lmap[x * g.vn[1] + y];
  if (r < 0) return 0;
  const double zc = gcoord(g, 2, z);
  int64_t lo = Z.ptr[r], hi = Z.ptr[r + 1];
  const int64_t a = lo;
  while (lo < hi) {  // first crossing above zc
    const int64_t mid = lo + (hi - lo) / 2;
    if (Z.inters[mid] <= zc) lo = mid + 1;
    else hi = mid;
  }
  return (int)((lo - a) & 1);
}

// a vertex is regular (not hanging) when it is a corner of every leaf around it
__device__ bool vertex_regular(const Geom& g, const int8_t* __restrict__ lev, int64_t px, int64_t py, int64_t pz) {
  for (int64_t z = pz - 1; z <= pz; ++z) {
    if (z < 0 || z >= g.n[2]) continue;
    for (int64_t y = py - 1; y <= py; ++y) {
      if (y < 0 || y >= g.n[1]) continue;
      for (int64_t x = px - 1; x <= px; ++x) {
        if (x < 0 || x >= g.n[0]) continue;
        const int64_t S = (int64_t)1 << lev[cell_at(g, x, y, z)];
        if (px % S || py % S || pz % S) return false;
      }
    }
  }
  return true;
}

// flags of the regular corners of the cube (x, y, z) of edge S: bit 0 some 0, bit 1 some 1
__device__ int corner_flags(const Geom& g, const int8_t* __restrict__ lev, const ZRays& Z, int64_t x, int64_t y,
                            int64_t z, int64_t S) {
  int seen = 0;
  for (int k = 0; k < 8; ++k) {
    const int64_t px = x + ((k & 1) ? S : 0), py = y + ((k & 2) ? S : 0), pz = z + ((k & 4) ? S : 0);
    if (!vertex_regular(g, lev, px, py, pz)) continue;
    seen |= zflag(g, Z, px, py, pz) ? 2 : 1;
  }
  return seen;
}

__global__ void k_zline_map(int64_t nr, const double* __restrict__ key, Geom g, int32_t* __restrict__ lmap) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const int64_t i = llround((key[2 * r] - g.origin[0]) / g.hf), j = llround((key[2 * r + 1] - g.origin[1]) / g.hf);
  if (i >= 0 && i < g.vn[0] && j >= 0 && j < g.vn[1]) lmap[i * g.vn[1] + j] = (int32_t)r;
}

__global__ void k_fill_i32(int64_t n, int32_t v, int32_t* __restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// uniform leaves of level s
__global__ void k_uniform(Geom g, int s, int64_t m0, int64_t m1, int64_t cnt, OLeaf* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int32_t S = 1 << s;
  out[i] = OLeaf{(int32_t)(i % m0) * S, (int32_t)((i / m0) % m1) * S, (int32_t)(i / (m0 * m1)) * S, s};
}

// refinement marker: a leaf above the finest level whose regular corners are on both
// sides of the surface is split once (src/pb_class.cpp:1569-1607; the atom-box test
// there never fires, as the atom list is empty at that point)
__global__ void k_surf_split(const OLeaf* __restrict__ lv, int64_t L, Geom g, const int8_t* __restrict__ lev,
                             ZRays Z, uint8_t* __restrict__ flag, int64_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  const bool split = l.s > 0 && corner_flags(g, lev, Z, l.x, l.y, l.z, (int64_t)1 << l.s) == 3;
  flag[i] = split ? 1 : 0;
  cnt[i] = split ? kSplit : kKeep;
}

// coarsening marker (src/pb_class.cpp:1700-1741): a leaf below the coarsest level
// whose regular corners are all on one side; a family of 8 sibling leaves is merged
// when every member is marked. The first child decides for its family and writes the
// decision at its anchor cell.
__global__ void k_surf_coarsen_decide(const OLeaf* __restrict__ lv, int64_t L, Geom g, int smax,
                                      const int8_t* __restrict__ lev, ZRays Z, uint8_t* __restrict__ dmap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  if (l.s >= smax) return;
  const int64_t P = (int64_t)1 << (l.s + 1), S = (int64_t)1 << l.s;
  if (l.x % P || l.y % P || l.z % P) return;  // not a first child
  bool merge = true;
  for (int k = 0; k < 8 && merge; ++k) {
    const int64_t x = l.x + ((k & 1) ? S : 0), y = l.y + ((k & 2) ? S : 0), z = l.z + ((k & 4) ? S : 0);
    if (lev[cell_at(g, x, y, z)] != l.s) merge = false;  // sibling is not a leaf of this level
    else if (corner_flags(g, lev, Z, x, y, z, S) == 3) merge = false;
  }
  dmap[cell_at(g, l.x, l.y, l.z)] = merge ? 1 : 0;
}

// emission after the coarsening decision: a merged family becomes its parent (written
// by the first child), every other leaf is kept
__global__ void k_surf_coarsen_count(const OLeaf* __restrict__ lv, int64_t L, Geom g, int smax,
                                     const int8_t* __restrict__ lev, const uint8_t* __restrict__ dmap,
                                     int64_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  int64_t c = 1;
  if (l.s < smax) {
    const int64_t P = (int64_t)1 << (l.s + 1), S = (int64_t)1 << l.s;
    const int64_t px = l.x - l.x % P, py = l.y - l.y % P, pz = l.z - l.z % P;
    bool family = true;
    for (int k = 0; k < 8 && family; ++k)
      family = lev[cell_at(g, px + ((k & 1) ? S : 0), py + ((k & 2) ? S : 0), pz + ((k & 4) ? S : 0))] == l.s;
    if (family && dmap[cell_at(g, px, py, pz)]) c = (l.x == px && l.y == py && l.z == pz) ? 1 : 0;
  }
  cnt[i] = c;
}

__global__ void k_surf_coarsen_emit(const OLeaf* __restrict__ lv, int64_t L, Geom g, int smax,
                                    const int8_t* __restrict__ lev, const uint8_t* __restrict__ dmap,
                                    const int64_t* __restrict__ off, const int64_t* __restrict__ cnt,
                                    OLeaf* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  if (cnt[i] == 0) return;
  OLeaf l = lv[i];
  if (l.s < smax) {
    const int64_t P = (int64_t)1 << (l.s + 1), S = (int64_t)1 << l.s;
    const int64_t px = l.x - l.x % P, py = l.y - l.y % P, pz = l.z - l.z % P;
    bool family = true;
    for (int k = 0; k < 8 && family; ++k)
      family = lev[cell_at(g, px + ((k & 1) ? S : 0), py + ((k & 2) ? S : 0), pz + ((k & 4) ? S : 0))] == l.s;
    if (family && dmap[cell_at(g, px, py, pz)]) l.s += 1;  // the first child becomes the parent
  }
  out[off[i]] = l;
}

// bit interleave of a 21-bit coordinate (x to bits 0, 3, 6, ...)
__device__ __forceinline__ uint64_t spread3(uint64_t v) {
  uint64_t r = 0;
  for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ULL) << (3 * b);
  return r;
}

__global__ void k_morton(const OLeaf* __restrict__ lv, int64_t L, uint64_t* __restrict__ key,
                         int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  key[i] = spread3((uint64_t)l.x) | (spread3((uint64_t)l.y) << 1) | (spread3((uint64_t)l.z) << 2);
  idx[i] = (int32_t)i;
}

__global__ void k_gather_leaves(const OLeaf* __restrict__ in, const int32_t* __restrict__ idx, int64_t L,
                                OLeaf* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < L) out[i] = in[idx[i]];
}

__device__ __forceinline__ int64_t corner_vertex(const Geom& g, const OLeaf& l, int k) {
  const int64_t S = (int64_t)1 << l.s;
  return vert_at(g, l.x + ((k & 1) ? S : 0), l.y + ((k & 2) ? S : 0), l.z + ((k & 4) ? S : 0));
}

__global__ void k_mark_corners(const OLeaf* __restrict__ lv, int64_t L, Geom g, int32_t* __restrict__ vmap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  for (int k = 0; k < 8; ++k) vmap[corner_vertex(g, l, k)] = 1;
}

// The 18 candidate hanging points of a leaf of edge S = 2h: the 12 edge midpoints
// (edge along axis d through corner offsets (a, b) on the other two axes) and the 6
// face centres (face normal to axis d at offset 0 or S). f(v, np, parents[4]) for
// each, with the parents in increasing corner order (x fastest): the edge's lower
// and upper end, the face's corners (-,-), (+,-), (-,+), (+,+) over its two axes
// in increasing axis order.
template <class F>
__device__ __forceinline__ void for_each_hanging_candidate(const Geom& g, const OLeaf& l, F&& f) {
  const int64_t S = (int64_t)1 << l.s, h = S >> 1;
  const int64_t o[3] = {l.x, l.y, l.z};
  for (int d = 0; d < 3; ++d) {
    const int e = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;  // the other axes, e < q
    for (int ab = 0; ab < 4; ++ab) {                  // edges along d
      int64_t p[3];
      p[d] = o[d] + h;
      p[e] = o[e] + ((ab & 1) ? S : 0);
      p[q] = o[q] + ((ab & 2) ? S : 0);
      int64_t lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
      lo[d] -= h;
      hi[d] += h;
      const int64_t par[4] = {vert_at(g, lo[0], lo[1], lo[2]), vert_at(g, hi[0], hi[1], hi[2]), -1, -1};
      f(vert_at(g, p[0], p[1], p[2]), 2, par);
    }
    for (int side = 0; side < 2; ++side) {  // faces normal to d
      int64_t p[3];
      p[d] = o[d] + (side ? S : 0);
      p[e] = o[e] + h;
      p[q] = o[q] + h;
      int64_t par[4];
      for (int j = 0; j < 4; ++j) {
        int64_t c[3] = {p[0], p[1], p[2]};
        c[e] += (j & 1) ? h : -h;
        c[q] += (j & 2) ? h : -h;
        par[j] = vert_at(g, c[0], c[1], c[2]);
      }
      f(vert_at(g, p[0], p[1], p[2]), 4, par);
    }
  }
}

__global__ void k_mark_hanging(const OLeaf* __restrict__ lv, int64_t L, Geom g, int32_t* __restrict__ vmap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  if (l.s == 0) return;
  // corners were marked by the previous kernel; here values only go from 1 to 2
  for_each_hanging_candidate(g, l, [&](int64_t v, int, const int64_t*) {
    if (vmap[v] != 0) vmap[v] = 2;
  });
}

__global__ void k_vertex_counts(const int32_t* __restrict__ vmap, int64_t V, int64_t* __restrict__ cnt) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int32_t t = vmap[v];
  cnt[v] = t == 1 ? kKeep : (t == 2 ? kSplit : 0);
}

// vmap: regular vertex -> node id, hanging -> -1 - hanging id, unused -> INT32_MIN
__global__ void k_number(int32_t* __restrict__ vmap, const int64_t* __restrict__ off, int64_t V, Geom g,
                         double* __restrict__ node_xyz, uint8_t* __restrict__ on_bnd, double* __restrict__ hang_xyz) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int32_t t = vmap[v];
  const int64_t ijk[3] = {v % g.vn[0], (v / g.vn[0]) % g.vn[1], v / (g.vn[0] * g.vn[1])};
  if (t == 1) {
    const int64_t id = off[v] & 0xffffffffLL;
    bool b = false;
    for (int d = 0; d < 3; ++d) {
      node_xyz[3 * id + d] = gcoord(g, d, ijk[d]);
      b = b || ijk[d] == 0 || ijk[d] == g.n[d];
    }
    on_bnd[id] = b ? 1 : 0;
    vmap[v] = (int32_t)id;
  } else if (t == 2) {
    const int64_t k = off[v] >> 32;
    for (int d = 0; d < 3; ++d) hang_xyz[3 * k + d] = gcoord(g, d, ijk[d]);
    vmap[v] = (int32_t)(-1 - k);
  } else {
    vmap[v] = INT32_MIN;
  }
}

__global__ void k_hang_parents(const OLeaf* __restrict__ lv, int64_t L, Geom g, const int32_t* __restrict__ vmap,
                               int32_t* __restrict__ hang_np, int32_t* __restrict__ hang_par,
                               int32_t* __restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  if (l.s == 0) return;
  for_each_hanging_candidate(g, l, [&](int64_t v, int np, const int64_t* par) {
    const int32_t id = vmap[v];
    if (id >= 0 || id == INT32_MIN) return;
    const int64_t k = -1 - (int64_t)id;
    hang_np[k] = np;
    for (int j = 0; j < 4; ++j) {
      const int32_t p = j < np ? vmap[par[j]] : -1;
      if (j < np && p < 0) err[0] = 1;  // a hanging parent: the mesh is not 2:1 balanced
      hang_par[4 * k + j] = p;
    }
  });
}

__global__ void k_leaf_corners(const OLeaf* __restrict__ lv, int64_t L, Geom g, const int32_t* __restrict__ vmap,
                               int32_t* __restrict__ corner) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const OLeaf l = lv[i];
  for (int k = 0; k < 8; ++k) corner[8 * i + k] = vmap[corner_vertex(g, l, k)];
}

// one split/keep pass over `in` (flags and packed counters already computed): the
// counters are scanned, kept cells go to keep[0 ..), the children of split cells to
// `out` (resized to 8 per split cell); returns (kept, split)
std::pair<int64_t, int64_t> split_pass(ngpb_ctx* c, const OLeaf* in, int64_t n, const uint8_t* flag, int64_t* cnt,
                                       OLeaf* keep, DBuf<OLeaf>& out) {
  const int64_t tot = exclusive_scan(c, cnt, cnt, n);
  const int64_t nk = tot & 0xffffffffLL, ns = tot >> 32;
  out.resize((size_t)std::max<int64_t>(8 * ns, 1));
  hipLaunchKernelGGL(k_split_emit, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, in, n, flag, cnt, keep, out.p);
  NGPB_HIP(hipGetLastError());
  return {nk, ns};
}

// 2:1 balance on the level map of the finest cells: Jacobi sweeps of k_balance_flag
// until no leaf is split (at most kBalanceSweeps)
void balance_leaves(ngpb_ctx* c, const Geom& g, DBuf<OLeaf>& leaves, int64_t& L) {
  hipStream_t st = c->stream;
  const int64_t ncell = g.n[0] * g.n[1] * g.n[2];
  DBuf<OLeaf> next(st);
  DBuf<uint8_t> flag(st);
  DBuf<int64_t> cnt(st);
  DBuf<int8_t> lev(st);
  lev.resize((size_t)ncell);
  hipLaunchKernelGGL(k_paint, dim3(blocks_for(L, 64)), dim3(64), 0, st, leaves.p, L, nullptr, g, lev.p);
  NGPB_HIP(hipGetLastError());
  for (int sweep = 0; sweep < kBalanceSweeps; ++sweep) {
    flag.resize((size_t)L);
    cnt.resize((size_t)L);
    hipLaunchKernelGGL(k_balance_flag, dim3(blocks_for(L, 128)), dim3(128), 0, st, leaves.p, L, g, lev.p, flag.p,
                       cnt.p);
    NGPB_HIP(hipGetLastError());
    const int64_t tot = exclusive_scan(c, cnt.p, cnt.p, L);
    const int64_t nk = tot & 0xffffffffLL, ns = tot >> 32;
    if (ns == 0) break;
    NGPB_REQUIRE(nk + 8 * ns < INT32_MAX, NGPB_E_ARG, "leaf count out of range");
    hipLaunchKernelGGL(k_paint, dim3(blocks_for(L, 64)), dim3(64), 0, st, leaves.p, L, flag.p, g, lev.p);
    next.resize((size_t)(nk + 8 * ns));  // kept leaves first, children after them
    hipLaunchKernelGGL(k_split_emit, dim3(blocks_for(L, 256)), dim3(256), 0, st, leaves.p, L, flag.p, cnt.p, next.p,
                       next.p + nk);
    NGPB_HIP(hipGetLastError());
    leaves.swap(next);
    L = nk + 8 * ns;
  }
}

OctreeSizes finish_octree(ngpb_ctx* c, const Geom& g, const int64_t* n, DBuf<OLeaf>& leaves, int64_t L);

}  // namespace

OctreeSizes run_build_octree(ngpb_ctx* c, const double* origin, double hf, const int64_t* n, int smax,
                             const double* band, int64_t na, const double* h_xyz, const double* h_r) {
  Geom g{};
  for (int d = 0; d < 3; ++d) {
    g.origin[d] = origin[d];
    g.n[d] = n[d];
    g.vn[d] = n[d] + 1;
  }
  g.hf = hf;
  for (int s = 0; s < smax; ++s) g.band[s] = band[s];
  const int64_t R = (int64_t)1 << smax;
  hipStream_t st = c->stream;

  // ---- atoms binned on the host (a few arrays of the atom count), uploaded once
  AtomBins B{};
  B.na = (int)na;
  B.cell = 4.0;
  DBuf<double> bxyz(st), br(st);
  DBuf<int32_t> bptr(st), batom(st);
  DBuf<double> bbox(st);
  {
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (na) {
      for (int d = 0; d < 3; ++d) {
        lo[d] = 1e300;
        hi[d] = -1e300;
      }
      for (int64_t a = 0; a < na; ++a)
        for (int d = 0; d < 3; ++d) {
          lo[d] = std::min(lo[d], h_xyz[3 * a + d]);
          hi[d] = std::max(hi[d], h_xyz[3 * a + d]);
        }
    }
    int64_t nb = 1;
    for (int d = 0; d < 3; ++d) {
      B.lo[d] = lo[d];
      B.dim[d] = std::max(1, (int)std::floor((hi[d] - lo[d]) / B.cell) + 1);
      nb *= B.dim[d];
    }
    NGPB_REQUIRE(nb < (int64_t)INT32_MAX, NGPB_E_ARG, "atom bin grid too large");
    B.rmax = 0.0;
    for (int64_t a = 0; a < na; ++a) B.rmax = std::max(B.rmax, h_r[a]);
    B.brad = 0.0;
    if (na) {
      for (int d = 0; d < 3; ++d) B.bc[d] = 0.5 * (lo[d] + hi[d]);
      for (int64_t a = 0; a < na; ++a) {
        double dd = 0;
        for (int d = 0; d < 3; ++d) dd += (h_xyz[3 * a + d] - B.bc[d]) * (h_xyz[3 * a + d] - B.bc[d]);
        B.brad = std::max(B.brad, std::sqrt(dd) + h_r[a]);
      }
    }
    std::vector<int32_t> ptr((size_t)nb + 1, 0), bin((size_t)std::max<int64_t>(na, 1)),
        atom((size_t)std::max<int64_t>(na, 1));
    for (int64_t a = 0; a < na; ++a) {
      int64_t cidx[3];
      for (int d = 0; d < 3; ++d)
        cidx[d] = std::min(B.dim[d] - 1, std::max(0, (int)std::floor((h_xyz[3 * a + d] - lo[d]) / B.cell)));
      bin[a] = (int32_t)((cidx[2] * B.dim[1] + cidx[1]) * B.dim[0] + cidx[0]);
      ++ptr[(size_t)bin[a] + 1];
    }
    for (int64_t b = 0; b < nb; ++b) ptr[b + 1] += ptr[b];
    {
      std::vector<int32_t> cur(ptr.begin(), ptr.end() - 1);
      for (int64_t a = 0; a < na; ++a) atom[(size_t)cur[bin[a]]++] = (int32_t)a;
    }
    std::vector<double> box((size_t)7 * nb);
    for (int64_t b = 0; b < nb; ++b) {
      for (int d = 0; d < 3; ++d) {
        box[7 * b + d] = 1e300;
        box[7 * b + 3 + d] = -1e300;
      }
      box[7 * b + 6] = 0.0;
    }
    for (int64_t a = 0; a < na; ++a) {
      double* bx = box.data() + 7 * (int64_t)bin[a];
      for (int d = 0; d < 3; ++d) {
        bx[d] = std::min(bx[d], h_xyz[3 * a + d]);
        bx[3 + d] = std::max(bx[3 + d], h_xyz[3 * a + d]);
      }
      bx[6] = std::max(bx[6], h_r[a]);
    }
    bbox.upload(box.data(), box.size(), st);
    bxyz.upload(na ? h_xyz : nullptr, (size_t)(3 * na), st);
    br.upload(na ? h_r : nullptr, (size_t)na, st);
    if (!na) {
      bxyz.resize(3);
      br.resize(1);
    }
    bptr.upload(ptr.data(), ptr.size(), st);
    batom.upload(atom.data(), atom.size(), st);
    B.xyz = bxyz.p;
    B.r = br.p;
    B.ptr = bptr.p;
    B.atom = batom.p;
    B.box = bbox.p;
    NGPB_HIP(hipStreamSynchronize(st));  // the host arrays above go out of scope
  }

  // ---- top-down refinement, one level per pass
  DBuf<OLeaf> work(st), next(st), kept[kMaxLevels + 1];
  for (auto& k : kept) k.order_on(st);
  int64_t nkept[kMaxLevels + 1] = {0};
  DBuf<uint8_t> flag(st);
  DBuf<int64_t> cnt(st);
  int64_t nw = (n[0] / R) * (n[1] / R) * (n[2] / R);
  work.resize((size_t)nw);
  hipLaunchKernelGGL(k_roots, dim3(blocks_for(nw, 256)), dim3(256), 0, st, g, smax, n[0] / R, n[1] / R, nw, work.p);
  NGPB_HIP(hipGetLastError());
  int64_t L = 0;
  // Temporaries are allocated and freed in the stream's order (no device-wide wait).
  for (int pass = 0; nw > 0; ++pass) {
    flag.resize((size_t)nw);
    cnt.resize((size_t)nw);
    hipLaunchKernelGGL(k_refine_flag, dim3(blocks_for(nw, 128)), dim3(128), 0, st, work.p, nw, g, B, flag.p, cnt.p);
    NGPB_HIP(hipGetLastError());
    kept[pass].resize((size_t)nw);
    const auto ks = split_pass(c, work.p, nw, flag.p, cnt.p, kept[pass].p, next);
    nkept[pass] = ks.first;
    L += ks.first;
    work.swap(next);
    nw = 8 * ks.second;
  }
  NGPB_REQUIRE(L > 0 && L < INT32_MAX, NGPB_E_ARG, "leaf count out of range");
  DBuf<OLeaf> leaves(st);
  leaves.resize((size_t)L);
  {
    int64_t o = 0;
    for (int pass = 0; pass <= kMaxLevels; ++pass)
      if (nkept[pass]) {
        NGPB_HIP(hipMemcpyAsync(leaves.p + o, kept[pass].p, nkept[pass] * sizeof(OLeaf), hipMemcpyDeviceToDevice, st));
        o += nkept[pass];
      }
    for (auto& k : kept) k.release();
  }
  work.release();

  balance_leaves(c, g, leaves, L);
  return finish_octree(c, g, n, leaves, L);
}

OctreeSizes run_build_octree_surface(ngpb_ctx* c, const double* origin, double hf, const int64_t* n, int smax,
                                     int s_uni, int cycles, int64_t na, const double* h_xyz, const double* h_r,
                                     double probe) {
  Geom g{};
  for (int d = 0; d < 3; ++d) {
    g.origin[d] = origin[d];
    g.n[d] = n[d];
    g.vn[d] = n[d] + 1;
  }
  g.hf = hf;
  hipStream_t st = c->stream;
  const int64_t ncell = n[0] * n[1] * n[2];
  // rays of every grid z-line, taken out of the context's ray table
  cast_grid_zlines(c, origin, hf, n[0], n[1], na, h_xyz, h_r, probe);
  auto& R = c->rays[2];
  DBuf<int64_t> zptr;  // plain allocations, like the table they are taken from
  DBuf<double> zint;
  zptr.swap(R.ptr);
  zint.swap(R.inters);
  DBuf<int32_t> lmap(st);
  const int64_t nlines = g.vn[0] * g.vn[1];
  lmap.resize((size_t)nlines);
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(nlines, 256)), dim3(256), 0, st, nlines, (int32_t)-1, lmap.p);
  if (R.n > 0)
    hipLaunchKernelGGL(k_zline_map, dim3(blocks_for(R.n, 256)), dim3(256), 0, st, R.n, R.key.p, g, lmap.p);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipStreamSynchronize(st));  // R.key is released below
  // the context's table stays empty: the caller casts the mesh's rays afterwards
  R.n = R.k = 0;
  R.key.resize(2);
  R.ptr.resize(1);
  R.inters.resize(1);
  R.normals.resize(3);
  NGPB_HIP(hipMemsetAsync(R.ptr.p, 0, sizeof(int64_t), st));
  const ZRays Z{lmap.p, zptr.p, zint.p};

  // uniform start at level s_uni (init_tmesh, src/pb_class.cpp:1289-1296)
  const int64_t U = (int64_t)1 << s_uni;
  int64_t L = (n[0] / U) * (n[1] / U) * (n[2] / U);
  NGPB_REQUIRE(L > 0 && L < INT32_MAX, NGPB_E_ARG, "leaf count out of range");
  DBuf<OLeaf> leaves(st), next(st);
  leaves.resize((size_t)L);
  hipLaunchKernelGGL(k_uniform, dim3(blocks_for(L, 256)), dim3(256), 0, st, g, s_uni, n[0] / U, n[1] / U, L,
                     leaves.p);
  NGPB_HIP(hipGetLastError());
  DBuf<int8_t> lev(st);
  DBuf<uint8_t> flag(st), dmap(st);
  DBuf<int64_t> cnt(st);
  lev.resize((size_t)ncell);
  dmap.resize((size_t)ncell);
  auto paint = [&]() {
    hipLaunchKernelGGL(k_paint, dim3(blocks_for(L, 64)), dim3(64), 0, st, leaves.p, L, nullptr, g, lev.p);
    NGPB_HIP(hipGetLastError());
  };
  for (int cyc = 0; cyc < cycles; ++cyc) {
    // refinement, then 2:1 balance (tmsh.refine (0, 1))
    paint();
    flag.resize((size_t)L);
    cnt.resize((size_t)L);
    hipLaunchKernelGGL(k_surf_split, dim3(blocks_for(L, 128)), dim3(128), 0, st, leaves.p, L, g, lev.p, Z, flag.p,
                       cnt.p);
    NGPB_HIP(hipGetLastError());
    {
      const int64_t tot = exclusive_scan(c, cnt.p, cnt.p, L);
      const int64_t nk = tot & 0xffffffffLL, ns = tot >> 32;
      NGPB_REQUIRE(nk + 8 * ns < INT32_MAX, NGPB_E_ARG, "leaf count out of range");
      if (ns > 0) {
        next.resize((size_t)(nk + 8 * ns));
        hipLaunchKernelGGL(k_split_emit, dim3(blocks_for(L, 256)), dim3(256), 0, st, leaves.p, L, flag.p, cnt.p,
                           next.p, next.p + nk);
        NGPB_HIP(hipGetLastError());
        leaves.swap(next);
        L = nk + 8 * ns;
      }
    }
    balance_leaves(c, g, leaves, L);
    // coarsening, then 2:1 balance (tmsh.coarsen (0, 1))
    paint();
    hipLaunchKernelGGL(k_surf_coarsen_decide, dim3(blocks_for(L, 128)), dim3(128), 0, st, leaves.p, L, g, smax, lev.p,
                       Z, dmap.p);
    cnt.resize((size_t)L + 1);
    hipLaunchKernelGGL(k_surf_coarsen_count, dim3(blocks_for(L, 256)), dim3(256), 0, st, leaves.p, L, g, smax, lev.p,
                       dmap.p, cnt.p);
    NGPB_HIP(hipGetLastError());
    {
      DBuf<int64_t> off(st);
      off.resize((size_t)L);
      const int64_t nl = exclusive_scan(c, cnt.p, off.p, L);
      if (nl != L) {
        next.resize((size_t)nl);
        hipLaunchKernelGGL(k_surf_coarsen_emit, dim3(blocks_for(L, 256)), dim3(256), 0, st, leaves.p, L, g, smax,
                           lev.p, dmap.p, off.p, cnt.p, next.p);
        NGPB_HIP(hipGetLastError());
        leaves.swap(next);
        L = nl;
      }
    }
    balance_leaves(c, g, leaves, L);
  }
  lev.release();
  dmap.release();
  return finish_octree(c, g, n, leaves, L);
}

namespace {

OctreeSizes finish_octree(ngpb_ctx* c, const Geom& g, const int64_t* n, DBuf<OLeaf>& leaves, int64_t L) {
  hipStream_t st = c->stream;
  const int64_t V = g.vn[0] * g.vn[1] * g.vn[2];
  DBuf<OLeaf> next(st);
  // ---- Morton order
  {
    DBuf<uint64_t> key(st), key2(st);
    DBuf<int32_t> idx(st), idx2(st);
    key.resize((size_t)L);
    key2.resize((size_t)L);
    idx.resize((size_t)L);
    idx2.resize((size_t)L);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(L, 256)), dim3(256), 0, st, leaves.p, L, key.p, idx.p);
    NGPB_HIP(hipGetLastError());
    int bits = 1;
    while (((int64_t)1 << bits) < std::max({n[0], n[1], n[2]})) ++bits;
    size_t tb = 0;
    NGPB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.p, key2.p, idx.p, idx2.p, (int)L, 0, 3 * bits, st));
    DBuf<uint8_t> tmp(st);
    tmp.resize(std::max<size_t>(tb, 1));
    NGPB_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.p, key2.p, idx.p, idx2.p, (int)L, 0, 3 * bits, st));
    next.resize((size_t)L);
    hipLaunchKernelGGL(k_gather_leaves, dim3(blocks_for(L, 256)), dim3(256), 0, st, leaves.p, idx2.p, L, next.p);
    NGPB_HIP(hipGetLastError());
    leaves.swap(next);
    next.release();
  }

  // ---- vertices
  DBuf<int32_t> vmap(st);
  DBuf<int64_t> off(st);
  vmap.resize((size_t)V);
  off.resize((size_t)V);
  NGPB_HIP(hipMemsetAsync(vmap.p, 0, (size_t)V * sizeof(int32_t), st));
  hipLaunchKernelGGL(k_mark_corners, dim3(blocks_for(L, 256)), dim3(256), 0, st, leaves.p, L, g, vmap.p);
  hipLaunchKernelGGL(k_mark_hanging, dim3(blocks_for(L, 128)), dim3(128), 0, st, leaves.p, L, g, vmap.p);
  hipLaunchKernelGGL(k_vertex_counts, dim3(blocks_for(V, 256)), dim3(256), 0, st, vmap.p, V, off.p);
  NGPB_HIP(hipGetLastError());
  const int64_t vt = exclusive_scan(c, off.p, off.p, V);
  const int64_t N = vt & 0xffffffffLL, H = vt >> 32;
  NGPB_REQUIRE(N > 0 && N < INT32_MAX / 2 && H < INT32_MAX / 2, NGPB_E_ARG, "too many nodes for int32 ids");
  c->node_xyz.resize((size_t)(3 * N));
  c->is_bnd.resize((size_t)N);
  c->hang_xyz.resize((size_t)std::max<int64_t>(3 * H, 3));
  c->hang_np.resize((size_t)std::max<int64_t>(H, 1));
  c->hang_par.resize((size_t)std::max<int64_t>(4 * H, 4));
  c->leaf_corner.resize((size_t)(8 * L));
  hipLaunchKernelGGL(k_number, dim3(blocks_for(V, 256)), dim3(256), 0, st, vmap.p, off.p, V, g, c->node_xyz.p,
                     c->is_bnd.p, c->hang_xyz.p);
  off.release();
  DBuf<int32_t> err(st);
  err.resize(1);
  NGPB_HIP(hipMemsetAsync(err.p, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_hang_parents, dim3(blocks_for(L, 128)), dim3(128), 0, st, leaves.p, L, g, vmap.p,
                     c->hang_np.p, c->hang_par.p, err.p);
  hipLaunchKernelGGL(k_leaf_corners, dim3(blocks_for(L, 256)), dim3(256), 0, st, leaves.p, L, g, vmap.p,
                     c->leaf_corner.p);
  NGPB_HIP(hipGetLastError());
  int32_t h_err = 0;
  NGPB_HIP(hipMemcpyAsync(&h_err, err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  NGPB_HIP(hipStreamSynchronize(st));
  NGPB_REQUIRE(!h_err, NGPB_E_ARG, "hanging node with a hanging parent (mesh not 2:1 balanced)");

  // ---- boundary nodes (domain faces), in node order
  c->bnd_node.resize((size_t)N);
  const int64_t NB = compact_flags(c, c->is_bnd.p, N, c->bnd_node.p);
  install_device_mesh(c, N, H, L, NB);
  return OctreeSizes{N, H, L, NB};
}

}  // namespace


}  // namespace ngpb
