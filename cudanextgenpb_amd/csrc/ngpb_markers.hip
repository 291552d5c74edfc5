// In/out markers, border leaves and edge fractions on the device.
//
//   k_node_inside   is_in_ns_surf along z + the node part of create_markers
//                   (src/pb_class.cpp:768-805, 1846-1857): one thread per node.
//   k_leaf_marker   leaf marker of create_markers (src/pb_class.cpp:1910-1916).
//   k_fractions     cube_fraction_intersection / normal_intersection
//                   (src/pb_class.cpp:2422-2568): one thread per (border leaf, edge).
//
// Ray lookup is a binary search over the lexicographically sorted keys of the
// axis (exact key equality, a missing ray means no crossing).
#include <algorithm>

#include "ngpb_internal.h"

namespace ngpb {
namespace {

__constant__ int kEdgeNodes[24] = {0, 1, 1, 3, 2, 3, 0, 2, 4, 5, 5, 7, 6, 7, 4, 6, 0, 4, 1, 5, 3, 7, 2, 6};
__constant__ int kEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};

struct Rays3 {
  RayView r[3];
};

__global__ void k_node_inside(MeshView m, RayView rz, double eps_in, double eps_out, double reac_out,
                              double* __restrict__ eps, double* __restrict__ reac,
                              uint8_t* __restrict__ inside) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const double x = m.node_xyz[3 * i], y = m.node_xyz[3 * i + 1], z = m.node_xyz[3 * i + 2];
  int in = 0;
  const int64_t ray = dev::ray_find(rz, x, y);
  if (ray >= 0) {
    const int64_t b = rz.ptr[ray], e = rz.ptr[ray + 1];
    if (e > b && !(z < rz.inters[b])) {
      int64_t k = b;
      while (k < e && z >= rz.inters[k]) ++k;
      in = (int)((k - b) & 1);
    }
  }
  inside[i] = (uint8_t)in;
  eps[i] = in ? eps_in : eps_out;
  reac[i] = in ? 0.0 : reac_out;
}

__global__ void k_leaf_marker(MeshView m, const uint8_t* __restrict__ inside, uint8_t* __restrict__ marker,
                              uint8_t* __restrict__ bflag) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= m.L) return;
  int nin = 0, nh = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int32_t c = m.leaf_corner[8 * l + i];
    if (c >= 0) nin += inside[c];
    else ++nh;
  }
  uint8_t mk = 0;
  if (nin == 0) mk = 2;
  else if (nin < 8 - nh) mk = 1;
  marker[l] = mk;
  bflag[l] = (mk == 1) ? 1 : 0;
}

__global__ void k_border_flag(int64_t L, const uint8_t* __restrict__ marker, uint8_t* __restrict__ bflag) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l < L) bflag[l] = marker[l] == 1 ? 1 : 0;
}

__global__ void k_border_of(int64_t L, int32_t* __restrict__ border_of) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < L) border_of[i] = -1;
}

__global__ void k_border_of_set(int64_t nb, const int32_t* __restrict__ border, int32_t* __restrict__ border_of) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) border_of[border[b]] = (int32_t)b;
}

__global__ void k_fractions(MeshView m, Rays3 rays, int64_t nb, const int32_t* __restrict__ border,
                            double* __restrict__ frac, double* __restrict__ nrm, uint8_t* __restrict__ found) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * 12) return;
  const int64_t b = t / 12;
  const int e = (int)(t - b * 12);
  const int64_t l = border[b];
  const int dir = kEdgeAxis[e];
  const int32_t c1 = m.leaf_corner[8 * l + kEdgeNodes[2 * e]];
  const int32_t c2 = m.leaf_corner[8 * l + kEdgeNodes[2 * e + 1]];
  const double x1 = dev::corner_coord(m, c1, dir), x2 = dev::corner_coord(m, c2, dir);
  const int da = dir == 0 ? 1 : 0, db = dir == 2 ? 1 : 2;
  const double ka = dev::corner_coord(m, c1, da), kb = dev::corner_coord(m, c1, db);
  double f = 0.5, n0 = 0.0, n1 = 0.0, n2 = 0.0;
  uint8_t fd = 0;
  const RayView& r = rays.r[dir];
  const int64_t ray = dev::ray_find(r, ka, kb);
  if (ray >= 0) {
    for (int64_t k = r.ptr[ray]; k < r.ptr[ray + 1]; ++k) {
      const double x = r.inters[k];
      if (x >= x1 && x <= x2) {
        f = (x - x1) / (x2 - x1);
        n0 = r.normals[3 * k];
        n1 = r.normals[3 * k + 1];
        n2 = r.normals[3 * k + 2];
        fd = 1;
      }
    }
  }
  frac[t] = f;
  nrm[3 * t] = n0;
  nrm[3 * t + 1] = n1;
  nrm[3 * t + 2] = n2;
  found[t] = fd;
}

MeshView mesh_view(const ngpb_ctx* c) {
  MeshView m;
  m.N = c->N;
  m.H = c->H;
  m.L = c->L;
  m.node_xyz = c->node_xyz.p;
  m.hang_xyz = c->hang_xyz.p;
  m.hang_np = c->hang_np.p;
  m.hang_par = c->hang_par.p;
  m.leaf_corner = c->leaf_corner.p;
  return m;
}

RayView ray_view(const ngpb_ctx* c, int d) {
  RayView r;
  r.n = c->rays[d].n;
  r.key = c->rays[d].key.p;
  r.ptr = c->rays[d].ptr.p;
  r.inters = c->rays[d].inters.p;
  r.normals = c->rays[d].normals.p;
  return r;
}

void finish_border(ngpb_ctx* c, const uint8_t* bflag) {
  c->border.resize((size_t)std::max<int64_t>(c->L, 1));
  c->n_border = compact_flags(c, bflag, c->L, c->border.p);
  c->border_of.resize((size_t)c->L);
  hipLaunchKernelGGL(k_border_of, dim3(blocks_for(c->L, 256)), dim3(256), 0, c->stream, c->L, c->border_of.p);
  if (c->n_border > 0)
    hipLaunchKernelGGL(k_border_of_set, dim3(blocks_for(c->n_border, 256)), dim3(256), 0, c->stream,
                       c->n_border, c->border.p, c->border_of.p);
  NGPB_HIP(hipGetLastError());
}

}  // namespace

MeshView mesh_view_of(const ngpb_ctx* c) { return mesh_view(c); }

void run_markers(ngpb_ctx* c) {
  const MeshView m = mesh_view(c);
  c->eps.resize((size_t)c->N);
  c->reac.resize((size_t)c->N);
  c->leaf_marker.resize((size_t)c->L);
  uint8_t* inside = tmp_buf<uint8_t>(c, kTmpInside, (size_t)c->N);
  uint8_t* bflag = tmp_buf<uint8_t>(c, kTmpBorderFlag, (size_t)c->L);
  hipLaunchKernelGGL(k_node_inside, dim3(blocks_for(c->N, 256)), dim3(256), 0, c->stream, m, ray_view(c, 2),
                     c->pc.eps_in, c->pc.eps_out, c->pc.eps_out * c->pc.k2, c->eps.p, c->reac.p, inside);
  hipLaunchKernelGGL(k_leaf_marker, dim3(blocks_for(c->L, 256)), dim3(256), 0, c->stream, m, inside,
                     c->leaf_marker.p, bflag);
  NGPB_HIP(hipGetLastError());
  finish_border(c, bflag);
  NGPB_HIP(hipStreamSynchronize(c->stream));
}

void run_border_from_markers(ngpb_ctx* c) {
  uint8_t* bflag = tmp_buf<uint8_t>(c, kTmpBorderFlag, (size_t)c->L);
  hipLaunchKernelGGL(k_border_flag, dim3(blocks_for(c->L, 256)), dim3(256), 0, c->stream, c->L,
                     c->leaf_marker.p, bflag);
  NGPB_HIP(hipGetLastError());
  finish_border(c, bflag);
  NGPB_HIP(hipStreamSynchronize(c->stream));
}

void run_fractions(ngpb_ctx* c) {
  const int64_t nb = c->n_border;
  const size_t cap = (size_t)std::max<int64_t>(nb, 1);
  c->frac.resize(12 * cap);
  c->frac_normal.resize(36 * cap);
  c->frac_found.resize(12 * cap);
  if (nb > 0) {
    Rays3 r;
    for (int d = 0; d < 3; ++d) r.r[d] = ray_view(c, d);
    hipLaunchKernelGGL(k_fractions, dim3(blocks_for(12 * nb, 256)), dim3(256), 0, c->stream, mesh_view(c), r, nb,
                       c->border.p, c->frac.p, c->frac_normal.p, c->frac_found.p);
    NGPB_HIP(hipGetLastError());
  }
  c->have_frac = true;
}

}  // namespace ngpb
