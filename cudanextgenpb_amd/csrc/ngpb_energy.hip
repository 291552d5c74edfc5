// Polarization and direct ionic energy (poisson_boltzmann::energy,
// src/pb_class.cpp:2790-3063; energy_fast :3065-3338 is the same on meshes without
// hanging nodes), and the per-atom potentials and fields of pot_field / pot_field_fast
// (:3731-5060), which sum the same sources per atom.
//
//   k_sources     one thread per border leaf: corner eps/phi (hanging corners: mean of
//                 parents), flux edges (eps(i1) != eps(i2)) with their crossing point
//                 and flux, and, for the ionic term, the marching-cubes triangles
//                 (classifyCube + getTriangles) with per-vertex phi0, normal and area.
//   normal carry  normal_intersection only overwrites the normal when a crossing is
//                 found, so a vertex without a crossing reuses the last normal written
//                 in the reference's sequential sweep (possibly in an earlier leaf),
//                 found by walking back over the border leaves.
//   k_nbody       one thread per source, charged atoms (|q| > 1e-5) staged in LDS,
//                 FP64 sums in atom order.
//   k_sum3_*      fixed-order two-level reduction of the per-source sums; energies in
//                 FP64.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ngpb_internal.h"
#include "mc_table.inc"

namespace ngpb {

MeshView mesh_view_of(const ngpb_ctx* c);

namespace {

__constant__ int kEN[24] = {0, 1, 1, 3, 2, 3, 0, 2, 4, 5, 5, 7, 6, 7, 4, 6, 0, 4, 1, 5, 3, 7, 2, 6};
__constant__ int kEA[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
__constant__ signed char kTri[256][16];

constexpr double kPi = 3.14159265358979323846;
constexpr int kFluxSlots = 12;
constexpr int kTriSlots = 15;

// flux source: x y z flux ; vertex source: x y z nx ny nz phi_sup area
constexpr int kFS = 4;
constexpr int kVS = 8;

__device__ __forceinline__ double wha(double eps1, double eps2, double frac) {
  return 1.0 / (frac / eps1 + (1 - frac) / eps2);
}
__device__ __forceinline__ double phi0(double eps1, double eps2, double phi1, double phi2, double frac) {
  return phi1 + frac * eps2 * (phi2 - phi1) / (eps2 * frac + eps1 * (1 - frac));
}

struct EnergyArgs {
  MeshView m;
  int64_t nb;
  const int32_t* border;
  const double* eps;
  const double* phi;
  const double* frac;
  const double* nrm;
  const uint8_t* found;
  double eps_out;
  int ionic;
};

__global__ void k_sources(EnergyArgs a, double* __restrict__ fsrc, uint8_t* __restrict__ fvalid,
                          double* __restrict__ vsrc, uint8_t* __restrict__ vvalid, uint8_t* __restrict__ vneed,
                          double* __restrict__ leaf_charge, double* __restrict__ leaf_last, uint8_t* __restrict__ leaf_has,
                          int64_t* __restrict__ counts) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nb) return;
  const MeshView& m = a.m;
  const int64_t l = a.border[b];
  const int32_t c0 = m.leaf_corner[8 * l], c7 = m.leaf_corner[8 * l + 7];
  double h[3], p0[8][3];
  for (int d = 0; d < 3; ++d) h[d] = dev::corner_coord(m, c7, d) - dev::corner_coord(m, c0, d);
  const double area_h[3] = {h[1] * h[2] / h[0] * 0.25, h[0] * h[2] / h[1] * 0.25, h[0] * h[1] / h[2] * 0.25};
  double te[8], tp[8];
  for (int i = 0; i < 8; ++i) {
    const int32_t c = m.leaf_corner[8 * l + i];
    for (int d = 0; d < 3; ++d) p0[i][d] = dev::corner_coord(m, c, d);
    if (c >= 0) {
      te[i] = a.eps[c];
      tp[i] = a.phi[c];
    } else {
      const int64_t hh = -1 - (int64_t)c;
      const int np = m.hang_np[hh];
      double se = 0.0, sp = 0.0;
      for (int j = 0; j < np; ++j) {
        const int32_t pj = m.hang_par[4 * hh + j];
        se += a.eps[pj] / np;
        sp += a.phi[pj] / np;
      }
      te[i] = se;
      tp[i] = sp;
    }
  }
  // the normal most recently written in this leaf, if any
  bool has = false;
  double N[3] = {0.0, 0.0, 0.0};
  double cp = 0.0;
  int nflux = 0;
  for (int e = 0; e < 12; ++e) {
    const int64_t slot = kFluxSlots * b + e;
    const int i1 = kEN[2 * e], i2 = kEN[2 * e + 1];
    double fl;
    if (te[i1] < te[i2]) fl = 1.0;
    else if (te[i1] > te[i2]) fl = -1.0;
    else {
      fvalid[slot] = 0;
      continue;
    }
    const int ax = kEA[e];
    const double f = a.frac[12 * b + e];
    if (a.found[12 * b + e]) {
      has = true;
      for (int d = 0; d < 3; ++d) N[d] = a.nrm[36 * b + 3 * e + d];
    }
    double V[3] = {p0[i1][0], p0[i1][1], p0[i1][2]};
    V[ax] += f * h[ax];
    const double flux = -(tp[i2] - tp[i1]) * wha(te[i1], te[i2], f) * fl * area_h[ax];
    cp += flux;
    ++nflux;
    fsrc[kFS * slot + 0] = V[0];
    fsrc[kFS * slot + 1] = V[1];
    fsrc[kFS * slot + 2] = V[2];
    fsrc[kFS * slot + 3] = flux;
    fvalid[slot] = 1;
  }
  int ntri = 0;
  for (int s = 0; s < kTriSlots; ++s) vvalid[kTriSlots * b + s] = 0;
  if (a.ionic) {
    static constexpr int order[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    int ci = 0, bit = 1;
    for (int o = 0; o < 8; ++o) {
      if (te[order[o]] < a.eps_out - 1e-10) ci |= bit;
      bit *= 2;
    }
    if (ci != 0 && ci != 255) {
      for (int t = 0; t < 5 && kTri[ci][3 * t] >= 0; ++t) {
        double vt[3][3];
        int64_t vs[3];
        for (int jj = 0; jj < 3; ++jj) {
          const int edge = kTri[ci][3 * t + jj];
          const int ax = kEA[edge];
          const int i1 = kEN[2 * edge], i2 = kEN[2 * edge + 1];
          const double f = a.frac[12 * b + edge];
          const int64_t slot = kTriSlots * b + 3 * t + jj;
          vs[jj] = slot;
          uint8_t need = 0;
          if (a.found[12 * b + edge]) {
            has = true;
            for (int d = 0; d < 3; ++d) N[d] = a.nrm[36 * b + 3 * edge + d];
          } else if (!has) {
            need = 1;
          }
          double V[3] = {p0[i1][0], p0[i1][1], p0[i1][2]};
          V[ax] += f * h[ax];
          for (int d = 0; d < 3; ++d) vt[jj][d] = V[d];
          double* o = vsrc + kVS * slot;
          o[0] = V[0];
          o[1] = V[1];
          o[2] = V[2];
          o[3] = N[0];
          o[4] = N[1];
          o[5] = N[2];
          o[6] = phi0(te[i1], te[i2], tp[i1], tp[i2], f);
          vvalid[slot] = 1;
          vneed[slot] = need;
        }
        double ab[3], ac[3];
        for (int i = 0; i < 3; ++i) {
          ab[i] = vt[0][i] - vt[1][i];
          ac[i] = vt[0][i] - vt[2][i];
        }
        const double cx = ab[1] * ac[2] - ab[2] * ac[1], cy = ab[0] * ac[2] - ab[2] * ac[0],
                     cz = ab[1] * ac[0] - ab[0] * ac[1];
        const double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
        for (int jj = 0; jj < 3; ++jj) vsrc[kVS * vs[jj] + 7] = area;
        ++ntri;
      }
    }
  }
  leaf_charge[b] = cp;
  leaf_has[b] = has ? 1 : 0;
  leaf_last[3 * b] = N[0];
  leaf_last[3 * b + 1] = N[1];
  leaf_last[3 * b + 2] = N[2];
  if (nflux) atomicAdd((unsigned long long*)&counts[0], (unsigned long long)nflux);
  if (ntri) atomicAdd((unsigned long long*)&counts[1], (unsigned long long)ntri);
}

// A triangle vertex whose leaf has not written a normal before it takes the normal
// last written in an earlier leaf of the sweep (normal_intersection leaves its output
// unchanged when it finds no crossing): the search walks back to the last leaf that
// wrote one (almost always the previous leaf); zero if there is none.
__global__ void k_apply_carry(int64_t nv, const uint8_t* __restrict__ vvalid, const uint8_t* __restrict__ vneed,
                              const uint8_t* __restrict__ has, const double* __restrict__ last,
                              double* __restrict__ vsrc) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nv || !vvalid[s] || !vneed[s]) return;
  int64_t b = s / kTriSlots - 1;
  while (b >= 0 && !has[b]) --b;
  for (int d = 0; d < 3; ++d) vsrc[kVS * s + 3 + d] = b >= 0 ? last[3 * b + d] : 0.0;
}

template <int W>
__global__ void k_gather_src(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ src,
                             double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = idx[i];
  for (int k = 0; k < W; ++k) dst[W * i + k] = src[W * s + k];
}

__global__ void k_charged_flags(int64_t A, const double* __restrict__ q, uint8_t* __restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < A) f[i] = fabs(q[i]) > 1.e-5 ? 1 : 0;
}

__global__ void k_gather_atoms(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ xyz,
                               const double* __restrict__ q, double4* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t a = idx[i];
  out[i] = make_double4(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2], q[a]);
}

constexpr int kNB = 256;

// Coulomb pair energy of the charged atoms (src/pb_class.cpp:2842-2857): thread i sums
// (q_i q_j) / r_ij over j > i in increasing j with the reference's arithmetic
// (r = sqrt(dx dx + dy dy + dz dz)); the per-atom sums are then added in atom order.
__global__ void __launch_bounds__(kNB) k_coul_rows(int64_t nq, const double4* __restrict__ atoms,
                                                   double* __restrict__ rows) {
  __shared__ double4 s_at[kNB];
  const int64_t i = (int64_t)blockIdx.x * kNB + threadIdx.x;
  const bool act = i < nq;
  const double4 a = act ? atoms[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  double e = 0.0;
  for (int64_t t0 = (int64_t)blockIdx.x * kNB; t0 < nq; t0 += kNB) {  // tiles from this block's first atom
    __syncthreads();
    if (t0 + threadIdx.x < nq) s_at[threadIdx.x] = atoms[t0 + threadIdx.x];
    __syncthreads();
    const int cnt = (int)min((int64_t)kNB, nq - t0);
    for (int jj = 0; jj < cnt; ++jj) {
      const int64_t j = t0 + jj;
      if (j <= i) continue;
      const double4 b = s_at[jj];
      const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      e += a.w * b.w / r;
    }
  }
  if (act) rows[i] = e;
}

// MODE 0/1: flux sources, first += q flux / r
// MODE 2:   triangle-vertex sources, second += q (phi area / (3 4 pi)) (d . n) / r^3
// Each thread takes kNS sources, so every atom read from LDS serves kNS pairs.
// 1/r: FP32 reciprocal square root refined by one Newton step in FP64 (relative
// error ~1e-14), not rounded exactly like the reference's sqrt and divisions; the
// sums agree to ~1e-13 relative.
constexpr int kNS = 2;

__device__ __forceinline__ double inv_sqrt(double r2) {
  const double y = (double)rsqrtf((float)r2);
  return y * (1.5 - 0.5 * r2 * y * y);
}

template <int MODE>
__global__ void __launch_bounds__(kNB) k_nbody(int64_t ns, const double* __restrict__ src, int64_t nq,
                                               const double4* __restrict__ atoms, double* __restrict__ acc) {
  constexpr int VS = MODE == 2 ? kVS : kFS;
  __shared__ double4 s_at[kNB];
  const double inv_4pi = 1.0 / (4.0 * kPi);
  double v[kNS][8], wsrc[kNS], sum[kNS];
  bool act[kNS];
#pragma unroll
  for (int u = 0; u < kNS; ++u) {
    const int64_t s = ((int64_t)blockIdx.x * kNS + u) * kNB + threadIdx.x;
    act[u] = s < ns;
#pragma unroll
    for (int k = 0; k < 8; ++k) v[u][k] = (act[u] && k < VS) ? src[VS * s + k] : 0.0;
    // per-source weight: flux, or phi_sup * area / 3 / (4 pi)
    wsrc[u] = MODE == 2 ? v[u][6] * inv_4pi * v[u][7] / 3.0 : v[u][3];
    sum[u] = 0.0;
  }
  for (int64_t t0 = 0; t0 < nq; t0 += kNB) {
    __syncthreads();
    if (t0 + threadIdx.x < nq) s_at[threadIdx.x] = atoms[t0 + threadIdx.x];
    __syncthreads();
    const int cnt = (int)min((int64_t)kNB, nq - t0);
#pragma unroll 2
    for (int j = 0; j < cnt; ++j) {
      const double4 at = s_at[j];
#pragma unroll
      for (int u = 0; u < kNS; ++u) {
        const double dx = v[u][0] - at.x, dy = v[u][1] - at.y, dz = v[u][2] - at.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double ir = inv_sqrt(r2);
        if (MODE < 2) {
          sum[u] += at.w * ir;
        } else {
          const double dot = dx * v[u][3] + dy * v[u][4] + dz * v[u][5];
          sum[u] += at.w * dot * (ir * ir * ir);
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kNS; ++u) {
    const int64_t s = ((int64_t)blockIdx.x * kNS + u) * kNB + threadIdx.x;
    if (act[u]) acc[s] = sum[u] * wsrc[u];
  }
}

// Sums of three arrays in a fixed order: block g adds the g-th of kSumBlocks equal
// contiguous ranges of each array, then one block adds the block sums in block order.
constexpr int kSumBlocks = 256;

__global__ void __launch_bounds__(256) k_sum3_part(const double* __restrict__ a, int64_t na, const double* __restrict__ b,
                                                   int64_t nbv, const double* __restrict__ c, int64_t nc,
                                                   double* __restrict__ part) {
  __shared__ double s_red[256 / 64];
  const double* arr[3] = {a, b, c};
  const int64_t cnt[3] = {na, nbv, nc};
  for (int k = 0; k < 3; ++k) {
    const int64_t chunk = (cnt[k] + kSumBlocks - 1) / kSumBlocks;
    const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(cnt[k], i0 + chunk);
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) s += arr[k][i];
    s = dev::block_sum<256>(s, s_red);
    if (threadIdx.x == 0) part[k * kSumBlocks + blockIdx.x] = s;
  }
}

__global__ void __launch_bounds__(256) k_sum3_fin(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double s_red[256 / 64];
  for (int k = 0; k < 3; ++k) {
    const double s = dev::block_sum<256>(part[k * kSumBlocks + threadIdx.x], s_red);
    if (threadIdx.x == 0) out[k] = s;
  }
}

// out[0..2] = sums of a[0..na), b[0..nb), c[0..nc) (partials in scratch slot `slot`)
void sum3(ngpb_ctx* c, int slot, const double* a, int64_t na, const double* b, int64_t nb, const double* cc,
          int64_t nc, double* out) {
  double* part = tmp_buf<double>(c, slot, 3 * kSumBlocks);
  hipLaunchKernelGGL(k_sum3_part, dim3(kSumBlocks), dim3(256), 0, c->stream, a, na, b, nb, cc, nc, part);
  hipLaunchKernelGGL(k_sum3_fin, dim3(1), dim3(256), 0, c->stream, part, out);
  NGPB_HIP(hipGetLastError());
}

void upload_tri_table(hipStream_t s) {
  static signed char tab[256][16];
  static bool ready = false;
  if (!ready) {
    for (int c = 0; c < 256; ++c) {
      const char* t = kMcTriangles[c];
      const int len = (int)std::strlen(t);
      for (int i = 0; i < 16; ++i) tab[c][i] = -1;
      for (int i = 0; i < len && i < 15; ++i) tab[c][i] = (signed char)(t[i] <= '9' ? t[i] - '0' : t[i] - 'a' + 10);
    }
    ready = true;
  }
  NGPB_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(kTri), tab, sizeof(tab), 0, hipMemcpyHostToDevice, s));
}

}  // namespace

namespace {

// Flux and triangle-vertex sources of the border leaves in the reference's sweep
// order, compacted: fd[4 i] = x y z flux, vd[8 i] = x y z nx ny nz phi_sup area.
// The arrays live in the context's energy scratch slots.
struct Sources {
  double* fd = nullptr;
  double* vd = nullptr;
  double* leaf_charge = nullptr;
  int64_t* counts = nullptr;  // [0] flux edges, [1] triangles
  int64_t nf = 0, nv = 0, nb = 0;
};

void build_sources(ngpb_ctx* c, bool ionic, Sources& S) {
  const int64_t nb = c->n_border;
  S.nb = nb;
  S.counts = tmp_buf<int64_t>(c, kTmpEnergy + 13, 2);
  NGPB_HIP(hipMemsetAsync(S.counts, 0, 2 * sizeof(int64_t), c->stream));
  S.leaf_charge = tmp_buf<double>(c, kTmpEnergy + 12, (size_t)std::max<int64_t>(nb, 1));
  if (nb == 0) return;
  upload_tri_table(c->stream);
  const int64_t nf = kFluxSlots * nb, nv = kTriSlots * nb;
  struct D {
    double* p;
  };
  struct U {
    uint8_t* p;
  };
  D fsrc{tmp_buf<double>(c, kTmpEnergy + 0, (size_t)kFS * nf)}, vsrc{tmp_buf<double>(c, kTmpEnergy + 1, (size_t)kVS * nv)},
      leaf_last{tmp_buf<double>(c, kTmpEnergy + 2, (size_t)3 * nb)};
  U fvalid{tmp_buf<uint8_t>(c, kTmpEnergy + 4, (size_t)nf)}, vvalid{tmp_buf<uint8_t>(c, kTmpEnergy + 5, (size_t)nv)},
      vneed{tmp_buf<uint8_t>(c, kTmpEnergy + 6, (size_t)nv)}, leaf_has{tmp_buf<uint8_t>(c, kTmpEnergy + 7, (size_t)nb)};
  NGPB_HIP(hipMemsetAsync(vneed.p, 0, nv, c->stream));
  EnergyArgs ea;
  ea.m = mesh_view_of(c);
  ea.nb = nb;
  ea.border = c->border.p;
  ea.eps = c->eps.p;
  ea.phi = c->x.p;
  ea.frac = c->frac.p;
  ea.nrm = c->frac_normal.p;
  ea.found = c->frac_found.p;
  ea.eps_out = c->pc.eps_out;
  ea.ionic = ionic ? 1 : 0;
  hipLaunchKernelGGL(k_sources, dim3(blocks_for(nb, 64)), dim3(64), 0, c->stream, ea, fsrc.p, fvalid.p, vsrc.p, vvalid.p,
                     vneed.p, S.leaf_charge, leaf_last.p, leaf_has.p, S.counts);
  NGPB_HIP(hipGetLastError());
  if (ionic) {
    hipLaunchKernelGGL(k_apply_carry, dim3(blocks_for(nv, 256)), dim3(256), 0, c->stream, nv, vvalid.p, vneed.p,
                       leaf_has.p, leaf_last.p, vsrc.p);
    NGPB_HIP(hipGetLastError());
  }
  // compact valid sources (slot order is the reference's summation order)
  int32_t* fidx = tmp_buf<int32_t>(c, kTmpEnergy + 8, (size_t)nf);
  int32_t* vidx = tmp_buf<int32_t>(c, kTmpEnergy + 9, (size_t)nv);
  S.nf = compact_flags(c, fvalid.p, nf, fidx);
  S.nv = ionic ? compact_flags(c, vvalid.p, nv, vidx) : 0;
  S.fd = tmp_buf<double>(c, kTmpEnergy + 10, (size_t)kFS * std::max<int64_t>(S.nf, 1));
  S.vd = tmp_buf<double>(c, kTmpEnergy + 11, (size_t)kVS * std::max<int64_t>(S.nv, 1));
  if (S.nf)
    hipLaunchKernelGGL(k_gather_src<kFS>, dim3(blocks_for(S.nf, 256)), dim3(256), 0, c->stream, S.nf, fidx, fsrc.p,
                       S.fd);
  if (S.nv)
    hipLaunchKernelGGL(k_gather_src<kVS>, dim3(blocks_for(S.nv, 256)), dim3(256), 0, c->stream, S.nv, vidx, vsrc.p,
                       S.vd);
  NGPB_HIP(hipGetLastError());
}

// ---- potentials and fields at the atoms (pot_field / pot_field_fast) ----------------
// Each block takes kNB atoms (one per thread) and one chunk of sources (blockIdx.y),
// staged through LDS in source order; per-atom partial sums of the chunk go to
// part[(chunk * A + atom) * 4 + k] and are added over chunks in chunk order, so the
// result does not depend on scheduling.

// flux sources: S = sum flux / r, G = sum (a - V) flux / r^3
template <bool FIELD>
__global__ void __launch_bounds__(kNB) k_pf_flux(int64_t ns, const double* __restrict__ src, int64_t chunk, int64_t A,
                                                 const double* __restrict__ axyz, double* __restrict__ part) {
  __shared__ double4 s_src[kNB];
  const int64_t a = (int64_t)blockIdx.x * kNB + threadIdx.x;
  const bool act = a < A;
  const double ax = act ? axyz[3 * a] : 0.0, ay = act ? axyz[3 * a + 1] : 0.0, az = act ? axyz[3 * a + 2] : 0.0;
  const int64_t s0 = (int64_t)blockIdx.y * chunk, s1 = min(ns, s0 + chunk);
  double S = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  for (int64_t t0 = s0; t0 < s1; t0 += kNB) {
    __syncthreads();
    if (t0 + threadIdx.x < s1) {
      const double* v = src + kFS * (t0 + threadIdx.x);
      s_src[threadIdx.x] = make_double4(v[0], v[1], v[2], v[3]);
    }
    __syncthreads();
    const int cnt = (int)min((int64_t)kNB, s1 - t0);
    for (int j = 0; j < cnt; ++j) {
      const double4 v = s_src[j];
      const double dx = ax - v.x, dy = ay - v.y, dz = az - v.z;
      const double ir = inv_sqrt(dx * dx + dy * dy + dz * dz);
      S += v.w * ir;
      if (FIELD) {
        const double w = v.w * (ir * ir * ir);
        gx += dx * w;
        gy += dy * w;
        gz += dz * w;
      }
    }
  }
  if (act) {
    double* o = part + ((int64_t)blockIdx.y * A + a) * 4;
    o[0] = S;
    o[1] = gx;
    o[2] = gy;
    o[3] = gz;
  }
}

// triangle-vertex sources (weight w = phi_sup area / 3 / 4 pi, d = V - a):
// T = sum w (d . n) / r^3, H = sum w (n / r^3 - 3 d (d . n) / r^5)
template <bool FIELD>
__global__ void __launch_bounds__(kNB) k_pf_vert(int64_t ns, const double* __restrict__ src, int64_t chunk, int64_t A,
                                                 const double* __restrict__ axyz, double* __restrict__ part) {
  __shared__ double4 s_p[kNB];
  __shared__ double4 s_n[kNB];
  const double inv_4pi = 1.0 / (4.0 * kPi);
  const int64_t a = (int64_t)blockIdx.x * kNB + threadIdx.x;
  const bool act = a < A;
  const double ax = act ? axyz[3 * a] : 0.0, ay = act ? axyz[3 * a + 1] : 0.0, az = act ? axyz[3 * a + 2] : 0.0;
  const int64_t s0 = (int64_t)blockIdx.y * chunk, s1 = min(ns, s0 + chunk);
  double T = 0.0, hx = 0.0, hy = 0.0, hz = 0.0;
  for (int64_t t0 = s0; t0 < s1; t0 += kNB) {
    __syncthreads();
    if (t0 + threadIdx.x < s1) {
      const double* v = src + kVS * (t0 + threadIdx.x);
      s_p[threadIdx.x] = make_double4(v[0], v[1], v[2], v[6] * inv_4pi * v[7] / 3.0);
      s_n[threadIdx.x] = make_double4(v[3], v[4], v[5], 0.0);
    }
    __syncthreads();
    const int cnt = (int)min((int64_t)kNB, s1 - t0);
    for (int j = 0; j < cnt; ++j) {
      const double4 v = s_p[j];
      const double4 nv = s_n[j];
      const double dx = v.x - ax, dy = v.y - ay, dz = v.z - az;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double ir = inv_sqrt(r2);
      const double ir3 = ir * ir * ir;
      const double dot = dx * nv.x + dy * nv.y + dz * nv.z;
      T += v.w * dot * ir3;
      if (FIELD) {
        const double t5 = 3.0 * dot * ir3 * ir * ir;
        hx += v.w * (nv.x * ir3 - dx * t5);
        hy += v.w * (nv.y * ir3 - dy * t5);
        hz += v.w * (nv.z * ir3 - dz * t5);
      }
    }
  }
  if (act) {
    double* o = part + ((int64_t)blockIdx.y * A + a) * 4;
    o[0] = T;
    o[1] = hx;
    o[2] = hy;
    o[3] = hz;
  }
}

// Coulomb terms of atom i over every other atom j in index order (src/pb_class.cpp:
// 3776-3813 accumulates phi_c[i] over j < i first, then j > i, which is the same
// order), with the reference's sqrt and divisions; ce[i] = its share of the Coulomb
// energy (pairs j > i).
__global__ void __launch_bounds__(kNB) k_pf_coulomb(int64_t A, const double* __restrict__ axyz,
                                                    const double* __restrict__ q, double den_in,
                                                    double* __restrict__ table, double* __restrict__ ce) {
  __shared__ double4 s_at[kNB];
  const int64_t i = (int64_t)blockIdx.x * kNB + threadIdx.x;
  const bool act = i < A;
  const double xi = act ? axyz[3 * i] : 0.0, yi = act ? axyz[3 * i + 1] : 0.0, zi = act ? axyz[3 * i + 2] : 0.0;
  const double qi = act ? q[i] : 0.0;
  double ph = 0.0, fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  for (int64_t t0 = 0; t0 < A; t0 += kNB) {
    __syncthreads();
    if (t0 + threadIdx.x < A) {
      const int64_t j = t0 + threadIdx.x;
      s_at[threadIdx.x] = make_double4(axyz[3 * j], axyz[3 * j + 1], axyz[3 * j + 2], q[j]);
    }
    __syncthreads();
    const int cnt = (int)min((int64_t)kNB, A - t0);
    for (int jj = 0; jj < cnt; ++jj) {
      const int64_t j = t0 + jj;
      if (j == i) continue;
      const double4 at = s_at[jj];
      const double dx = xi - at.x, dy = yi - at.y, dz = zi - at.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double r = sqrt(r2);
      const double inv_r3 = 1.0 / (r2 * r);
      ph += at.w / r * den_in;
      fx += dx * inv_r3 * at.w * den_in;
      fy += dy * inv_r3 * at.w * den_in;
      fz += dz * inv_r3 * at.w * den_in;
      if (j > i) e += (qi * at.w) / r * den_in;
    }
  }
  if (act) {
    double* row = table + 12 * i;
    row[0] = ph;
    row[3] = fx;
    row[4] = fy;
    row[5] = fz;
    ce[i] = e;
  }
}

// per atom: chunk sums, constants, and the write-out corrections of the ionic columns
__global__ void k_pf_finish(int64_t A, int64_t nfc, const double* __restrict__ fpart, int64_t nvc,
                            const double* __restrict__ vpart, int fieldp, int ionic, int fieldi, double cp, double cr,
                            const double* __restrict__ q, double* __restrict__ table, double* __restrict__ e1,
                            double* __restrict__ e2) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= A) return;
  double S = 0.0, G[3] = {0.0, 0.0, 0.0}, T = 0.0, H[3] = {0.0, 0.0, 0.0};
  for (int64_t k = 0; k < nfc; ++k) {
    const double* o = fpart + (k * A + a) * 4;
    S += o[0];
    for (int d = 0; d < 3; ++d) G[d] += o[1 + d];
  }
  for (int64_t k = 0; k < nvc; ++k) {
    const double* o = vpart + (k * A + a) * 4;
    T += o[0];
    for (int d = 0; d < 3; ++d) H[d] += o[1 + d];
  }
  double* row = table + 12 * a;
  const double phi_p = S * cp;
  row[1] = phi_p;
  row[2] = ionic ? T - phi_p / cp * cr : 0.0;
  for (int d = 0; d < 3; ++d) {
    const double fp = G[d] * cp;
    row[6 + d] = fieldp ? fp : 0.0;
    row[9 + d] = fieldi ? H[d] - fp / cp * cr : 0.0;
  }
  e1[a] = q[a] * S;
  e2[a] = q[a] * T;
}

// source chunks so that atom tiles x chunks is about 2048 blocks
int64_t pf_chunks(int64_t A, int64_t ns) {
  if (ns <= 0) return 0;
  const int64_t tiles = (A + kNB - 1) / kNB;
  const int64_t want = std::max<int64_t>(1, 2048 / std::max<int64_t>(tiles, 1));
  return std::min<int64_t>(want, (ns + kNB - 1) / kNB);
}

}  // namespace

int pot_field_branch(int P, int F, double k) {
  if ((F == 1 || (F == 2 && k < 1.e-5)) && P < 2) return 1;
  if ((P == 1 || (P == 2 && k < 1.e-5)) && F == 0) return 2;
  if (F == 2 && k > 1.e-5) return 3;
  if (P == 2 && k > 1.e-5 && F == 0) return 4;
  if (P == 2 && k > 1.e-5 && F == 1) return 5;
  return 0;
}

void run_pot_field(ngpb_ctx* c, int P, int F, double* table_host, double out4[4], int* branch_out) {
  const int64_t A = c->A;
  const int branch = pot_field_branch(P, F, c->pc.k);
  *branch_out = branch;
  const bool fieldp = branch == 1 || branch == 3 || branch == 5;
  const bool ionic = branch >= 3;
  const bool fieldi = branch == 3;
  out4[0] = out4[1] = out4[2] = out4[3] = 0.0;
  if (A == 0) return;
  struct D {
    double* p;
  };
  D table{tmp_buf<double>(c, kTmpPotField + 0, (size_t)12 * A)}, ce{tmp_buf<double>(c, kTmpPotField + 1, (size_t)A)},
      e1{tmp_buf<double>(c, kTmpPotField + 2, (size_t)A)}, e2{tmp_buf<double>(c, kTmpPotField + 3, (size_t)A)},
      fin{tmp_buf<double>(c, kTmpPotField + 4, 8)};
  NGPB_HIP(hipMemsetAsync(table.p, 0, sizeof(double) * 12 * A, c->stream));
  NGPB_HIP(hipMemsetAsync(e1.p, 0, sizeof(double) * A, c->stream));
  NGPB_HIP(hipMemsetAsync(e2.p, 0, sizeof(double) * A, c->stream));
  const unsigned tiles = (unsigned)((A + kNB - 1) / kNB);
  hipLaunchKernelGGL(k_pf_coulomb, dim3(tiles), dim3(kNB), 0, c->stream, A, c->atom_xyz.p, c->atom_q.p,
                     1.0 / c->pc.eps_in, table.p, ce.p);
  NGPB_HIP(hipGetLastError());
  Sources S;
  if (branch != 0) {
    build_sources(c, ionic, S);
    const int64_t nfc = pf_chunks(A, S.nf), nvc = pf_chunks(A, S.nv);
    D fpart{tmp_buf<double>(c, kTmpPotField + 5, (size_t)4 * A * std::max<int64_t>(nfc, 1))},
        vpart{tmp_buf<double>(c, kTmpPotField + 6, (size_t)4 * A * std::max<int64_t>(nvc, 1))};
    if (nfc) {
      const int64_t chunk = (S.nf + nfc - 1) / nfc;
      if (fieldp)
        hipLaunchKernelGGL(k_pf_flux<true>, dim3(tiles, (unsigned)nfc), dim3(kNB), 0, c->stream, S.nf, S.fd, chunk, A,
                           c->atom_xyz.p, fpart.p);
      else
        hipLaunchKernelGGL(k_pf_flux<false>, dim3(tiles, (unsigned)nfc), dim3(kNB), 0, c->stream, S.nf, S.fd, chunk, A,
                           c->atom_xyz.p, fpart.p);
    }
    if (nvc) {
      const int64_t chunk = (S.nv + nvc - 1) / nvc;
      if (fieldi)
        hipLaunchKernelGGL(k_pf_vert<true>, dim3(tiles, (unsigned)nvc), dim3(kNB), 0, c->stream, S.nv, S.vd, chunk, A,
                           c->atom_xyz.p, vpart.p);
      else
        hipLaunchKernelGGL(k_pf_vert<false>, dim3(tiles, (unsigned)nvc), dim3(kNB), 0, c->stream, S.nv, S.vd, chunk, A,
                           c->atom_xyz.p, vpart.p);
    }
    NGPB_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pf_finish, dim3(blocks_for(A, 256)), dim3(256), 0, c->stream, A, nfc, fpart.p, nvc, vpart.p,
                       fieldp ? 1 : 0, ionic ? 1 : 0, fieldi ? 1 : 0, c->pc.constant_pol, c->pc.constant_react,
                       c->atom_q.p, table.p, e1.p, e2.p);
    NGPB_HIP(hipGetLastError());
    sum3(c, kTmpEnergy + 3, e1.p, A, e2.p, A, S.leaf_charge, S.nb, fin.p);
  }
  sum3(c, kTmpEnergy + 3, ce.p, A, ce.p, 0, ce.p, 0, fin.p + 4);
  NGPB_HIP(hipGetLastError());
  double h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (branch != 0) NGPB_HIP(hipMemcpyAsync(h, fin.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipMemcpyAsync(h + 4, fin.p + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (table_host)
    NGPB_HIP(hipMemcpyAsync(table_host, table.p, sizeof(double) * 12 * A, hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  const double first_int = h[0], second_int = h[1];
  if (branch != 0) out4[0] = 0.5 * c->pc.constant_pol * first_int;
  if (ionic) out4[1] = 0.5 * (second_int - first_int * c->pc.constant_react);
  out4[2] = h[4];
  out4[3] = h[2];
}

void run_energy(ngpb_ctx* c, ngpb_energy_out* out) {
  const int64_t nb = c->n_border;
  const bool ionic = c->physics.calc_energy == 2 && c->pc.k > 1.e-5;
  const bool polar_only = c->physics.calc_energy == 1 || (c->physics.calc_energy == 2 && c->pc.k < 1.e-5);
  std::memset(out, 0, sizeof(*out));
  out->n_border = nb;
  const bool coulomb = c->physics.calc_energy > 0 && c->physics.calc_coulombic == 1;
  if (!coulomb && (nb == 0 || (!ionic && !polar_only))) return;
  // charged atoms, in order
  double4* atoms = tmp_buf<double4>(c, kTmpEnergy + 16, (size_t)std::max<int64_t>(c->A, 1));
  int64_t nq = 0;
  if (c->A > 0) {
    uint8_t* aflag = tmp_buf<uint8_t>(c, kTmpEnergy + 14, (size_t)c->A);
    int32_t* aidx = tmp_buf<int32_t>(c, kTmpEnergy + 15, (size_t)c->A);
    hipLaunchKernelGGL(k_charged_flags, dim3(blocks_for(c->A, 256)), dim3(256), 0, c->stream, c->A, c->atom_q.p, aflag);
    nq = compact_flags(c, aflag, c->A, aidx);
    if (nq)
      hipLaunchKernelGGL(k_gather_atoms, dim3(blocks_for(nq, 256)), dim3(256), 0, c->stream, nq, aidx, c->atom_xyz.p,
                         c->atom_q.p, atoms);
  }
  if (coulomb && nq > 1) {
    double* rows = tmp_buf<double>(c, kTmpEnergy + 18, (size_t)nq);
    hipLaunchKernelGGL(k_coul_rows, dim3(blocks_for(nq, kNB)), dim3(kNB), 0, c->stream, nq, atoms, rows);
    c->en_out.resize(4);
    sum3(c, kTmpEnergy + 3, rows, nq, rows, 0, rows, 0, c->en_out.p);
    NGPB_HIP(hipGetLastError());
    double h = 0.0;
    NGPB_HIP(hipMemcpyAsync(&h, c->en_out.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    NGPB_HIP(hipStreamSynchronize(c->stream));
    out->coul_energy = h * (1.0 / c->pc.eps_in);
  }
  if (nb == 0 || (!ionic && !polar_only)) return;
  Sources S;
  build_sources(c, ionic, S);
  const int64_t nfv = S.nf, nvv = S.nv;
  double* facc = tmp_buf<double>(c, kTmpEnergy + 17, (size_t)std::max<int64_t>(nfv, 1));
  double* vacc = tmp_buf<double>(c, kTmpEnergy + 18, (size_t)std::max<int64_t>(nvv, 1));
  if (nfv) {
    if (ionic)
      hipLaunchKernelGGL(k_nbody<1>, dim3(blocks_for(nfv, kNB * kNS)), dim3(kNB), 0, c->stream, nfv, S.fd, nq, atoms,
                         facc);
    else
      hipLaunchKernelGGL(k_nbody<0>, dim3(blocks_for(nfv, kNB * kNS)), dim3(kNB), 0, c->stream, nfv, S.fd, nq, atoms,
                         facc);
  }
  if (nvv)
    hipLaunchKernelGGL(k_nbody<2>, dim3(blocks_for(nvv, kNB * kNS)), dim3(kNB), 0, c->stream, nvv, S.vd, nq, atoms,
                       vacc);
  NGPB_HIP(hipGetLastError());
  c->en_out.resize(4);
  sum3(c, kTmpEnergy + 3, facc, nfv, vacc, nvv, S.leaf_charge, nb, c->en_out.p);
  NGPB_HIP(hipGetLastError());
  double h[3];
  int64_t hc[2];
  NGPB_HIP(hipMemcpyAsync(h, c->en_out.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipMemcpyAsync(hc, S.counts, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  const double first_int = h[0], second_int = h[1];
  out->charge_pol = h[2];
  out->n_flux_edges = hc[0];
  out->n_triangles = hc[1];
  out->energy_pol = 0.5 * c->pc.constant_pol * first_int;
  out->energy_react = ionic ? 0.5 * (second_int - first_int * c->pc.constant_react) : 0.0;
}

}  // namespace ngpb
