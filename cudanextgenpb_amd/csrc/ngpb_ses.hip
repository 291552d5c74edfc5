// Axis rays against the solvent-excluded surface (SES) of a union of balls, on the
// device. The reference's default surface is the SES with probe radius 1.4 A
// (surface_type 0, src/pb_class.cpp:956-963; data/options.prm:136-140), cast by
// NanoShaper (src/raytracer.cpp:337-376), which is not vendored. This follows the
// textbook definition (Richards 1977; Connolly 1983): with SAS balls B(a_i, R_i),
// R_i = r_i + rp, and probe centres C = { q : |q - a_i| >= R_i for all i } (interior
// cavities included), p is inside the SES iff dist(p, C) > rp. Inside an atom's own
// ball p is inside; outside every SAS ball p is outside; otherwise the nearest point
// of C is an exposed SAS-surface candidate: the radial projection of p on an SAS
// sphere, the nearest point of an SAS circle (pair), or an SAS vertex (triple).
//
// Rays are the distinct vertex lines of the mesh (as for the sphere union). Along a
// line the SES boundary lies on an atom sphere, a torus (tube rp around an SAS
// circle) or a probe sphere (radius rp around an exposed SAS vertex); every such
// intersection is a candidate, the candidates are sorted by (t, kind, id), and a
// crossing is kept where the classification of the midpoints between consecutive
// distinct candidates changes. Torus roots: the quartic in tau = t - c_axis is
// sampled at 65 points on [-T, T] and each sign change bisected 64 times; candidates
// are normalised by + 0.0 (no negative zero, so equal values have equal keys). Normals
// point out of the molecule: (p - a)/r on atoms, (q - p)/rp towards the probe centre q
// on tori and probe spheres. Only IEEE double + - * / sqrt are used, so the results
// equal the CPU restatement's bit for bit (the circles' exposure aside, see below).
//
// Work split: the SAS pairs (circles) and exposed SAS vertices are built on the host
// from a cell grid of the atoms (ordered by atom indices; buried circles dropped, the
// exposure of each circle tested by one wave on the device, whose arc angles come from
// the device's acos and atan2);
// per axis, every line gathers the atoms, pairs and vertices within reach into
// compact lists sorted by axial coordinate, one thread per list item counts and
// writes its candidates, the candidates are sorted per line (segmented radix sorts:
// by (kind, id), then stably by t), one thread per candidate classifies the interval
// after its group of equal t (within an axial window of the line's lists), and the
// crossings are the group starts whose state differs from the previous group's.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "ngpb_internal.h"

namespace ngpb {

namespace {

struct SesPairD {
  double c[3], n[3], rho;
  int32_t i, j;
};

struct SesGrid2 {            // 2D bins of items by their transverse coordinates
  const int32_t* ptr;        // gn0 * gn1 + 1
  const int32_t* item;
  double lo0, lo1, cell;
  int gn0, gn1;
};

struct SesView {
  const double* xyz;
  const double* r;
  int64_t na;
  double rp;
  const SesPairD* pairs;
  const double* verts;       // 3 per exposed vertex
  SesGrid2 ga, gp, gv;       // atoms, pairs (by circle centre), vertices
  double wa, wp, wv;         // axial half-widths of the classification windows
};

__device__ __forceinline__ uint64_t ses_bits(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__device__ __forceinline__ double ses_unbits(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffULL) : ~k;
  return __longlong_as_double((long long)b);
}

__device__ __forceinline__ double sq3(const double* a, const double* b) {
  const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
  return x * x + y * y + z * z;
}

template <class F>
__device__ __forceinline__ void ses_visit(const SesGrid2& G, double u, double v, F&& f) {
  const int b0 = (int)floor((u - G.lo0) / G.cell), b1 = (int)floor((v - G.lo1) / G.cell);
  for (int y = b1 - 1; y <= b1 + 1; ++y)
    for (int x = b0 - 1; x <= b0 + 1; ++x) {
      if (x < 0 || y < 0 || x >= G.gn0 || y >= G.gn1) continue;
      const int64_t bin = (int64_t)y * G.gn0 + x;
      for (int32_t k = G.ptr[bin]; k < G.ptr[bin + 1]; ++k) f(G.item[k]);
    }
}

// reach tests of the line lists (a small margin keeps every item whose predicates
// can hold somewhere on the line)
__device__ __forceinline__ bool atom_in_reach(const SesView& S, int a, double u, double v, int d1, int d2) {
  const double du = u - S.xyz[3 * a + d1], dv = v - S.xyz[3 * a + d2];
  const double reach = S.r[a] + 2.0 * S.rp + 1e-6;
  return du * du + dv * dv < reach * reach;
}
__device__ __forceinline__ bool pair_in_reach(const SesView& S, int k, double u, double v, int d1, int d2) {
  const SesPairD& P = S.pairs[k];
  const double du = u - P.c[d1], dv = v - P.c[d2];
  const double reach = P.rho + S.rp + 1e-6;
  return du * du + dv * dv < reach * reach;
}
__device__ __forceinline__ bool vert_in_reach(const SesView& S, int k, double u, double v, int d1, int d2) {
  const double du = u - S.verts[3 * k + d1], dv = v - S.verts[3 * k + d2];
  const double reach = S.rp + 1e-6;
  return du * du + dv * dv < reach * reach;
}

__global__ void k_ses_list_count(int64_t nl, const uint64_t* __restrict__ lu, const uint64_t* __restrict__ lv,
                                 SesView S, int d1, int d2, int64_t* __restrict__ ca, int64_t* __restrict__ cp,
                                 int64_t* __restrict__ cv) {
  const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nl) return;
  const double u = ses_unbits(lu[li]), v = ses_unbits(lv[li]);
  int64_t na = 0, np = 0, nv = 0;
  ses_visit(S.ga, u, v, [&](int a) { na += atom_in_reach(S, a, u, v, d1, d2); });
  ses_visit(S.gp, u, v, [&](int k) { np += pair_in_reach(S, k, u, v, d1, d2); });
  ses_visit(S.gv, u, v, [&](int k) { nv += vert_in_reach(S, k, u, v, d1, d2); });
  ca[li] = na;
  cp[li] = np;
  cv[li] = nv;
}

// fills the lists with the items' axial coordinates as keys (sorted per line next)
__global__ void k_ses_list_fill(int64_t nl, const uint64_t* __restrict__ lu, const uint64_t* __restrict__ lv,
                                SesView S, int axis, int d1, int d2, const int64_t* __restrict__ oa,
                                const int64_t* __restrict__ op, const int64_t* __restrict__ ov,
                                int32_t* __restrict__ la, int32_t* __restrict__ lp, int32_t* __restrict__ lvv,
                                uint64_t* __restrict__ ka, uint64_t* __restrict__ kp, uint64_t* __restrict__ kv) {
  const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nl) return;
  const double u = ses_unbits(lu[li]), v = ses_unbits(lv[li]);
  int64_t ia = oa[li], ip = op[li], iv = ov[li];
  ses_visit(S.ga, u, v, [&](int a) {
    if (!atom_in_reach(S, a, u, v, d1, d2)) return;
    ka[ia] = ses_bits(S.xyz[3 * a + axis]);
    la[ia++] = a;
  });
  ses_visit(S.gp, u, v, [&](int k) {
    if (!pair_in_reach(S, k, u, v, d1, d2)) return;
    kp[ip] = ses_bits(S.pairs[k].c[axis]);
    lp[ip++] = k;
  });
  ses_visit(S.gv, u, v, [&](int k) {
    if (!vert_in_reach(S, k, u, v, d1, d2)) return;
    kv[iv] = ses_bits(S.verts[3 * k + axis]);
    lvv[iv++] = k;
  });
}

struct LineLists {
  const int32_t* la;
  const int32_t* lp;
  const int32_t* lv;
  const int64_t* oa;  // nl + 1 each
  const int64_t* op;
  const int64_t* ov;
  const uint64_t* ka;  // axial coordinate keys, ascending within each line
  const uint64_t* kp;
  const uint64_t* kv;
};

// positions [lo, hi) of the keys within [b, e) that lie in [t - w, t + w]
__device__ __forceinline__ void key_window(const uint64_t* k, int64_t b, int64_t e, double t, double w, int64_t* lo,
                                           int64_t* hi) {
  const uint64_t x0 = ses_bits(t - w), x1 = ses_bits(t + w);
  int64_t l = b, h = e;
  while (l < h) {
    const int64_t m = (l + h) >> 1;
    if (k[m] < x0) l = m + 1;
    else h = m;
  }
  *lo = l;
  h = e;
  while (l < h) {
    const int64_t m = (l + h) >> 1;
    if (k[m] <= x1) l = m + 1;
    else h = m;
  }
  *hi = l;
}

// the torus of pair P against the line (u, v): calls f(t) for every root, in increasing t
template <class F>
__device__ __forceinline__ void torus_roots(const SesPairD& P, double rp, int axis, int d1, int d2, double u, double v,
                                            F&& f) {
  const double wu = u - P.c[d1], wv = v - P.c[d2];
  const double D2 = wu * wu + wv * wv;
  const double T2 = (P.rho + rp) * (P.rho + rp) - D2;
  if (!(T2 > 0.0)) return;
  const double T = sqrt(T2);
  const double s0 = P.n[d1] * wu + P.n[d2] * wv;
  const double na = P.n[axis];
  const double K = D2 + P.rho * P.rho - rp * rp;
  const double fr = 4.0 * P.rho * P.rho;
  auto g = [&](double tau) {
    const double a = tau * tau + K;
    const double b = na * tau + s0;
    return a * a - fr * (tau * tau + D2 - b * b);
  };
  double t0 = -T, g0 = g(t0);
  for (int m = 1; m <= 64; ++m) {
    const double t1 = -T + (2.0 * T) * ((double)m / 64.0);
    const double g1 = g(t1);
    if (g0 == 0.0) {
      f(P.c[axis] + t0 + 0.0);
    } else if ((g0 < 0.0) != (g1 < 0.0) && g1 != 0.0) {
      double lo = t0, hi = t1, glo = g0;
      for (int it = 0; it < 64; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double gm = g(mid);
        if ((gm < 0.0) == (glo < 0.0)) {
          lo = mid;
          glo = gm;
        } else {
          hi = mid;
        }
      }
      f(P.c[axis] + 0.5 * (lo + hi) + 0.0);
    }
    t0 = t1;
    g0 = g1;
  }
  if (g0 == 0.0) f(P.c[axis] + t0 + 0.0);
}

// nearest point of the SAS circle of P to p
__device__ __forceinline__ void circle_point(const SesPairD& P, const double* p, double* q) {
  double w[3] = {p[0] - P.c[0], p[1] - P.c[1], p[2] - P.c[2]};
  const double s = w[0] * P.n[0] + w[1] * P.n[1] + w[2] * P.n[2];
  double u[3] = {w[0] - s * P.n[0], w[1] - s * P.n[1], w[2] - s * P.n[2]};
  double L = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  if (!(L > 0.0)) {
    const int a = fabs(P.n[0]) < 0.9 ? 0 : 1;
    double e[3] = {0, 0, 0};
    e[a] = 1.0;
    const double ds = e[0] * P.n[0] + e[1] * P.n[1] + e[2] * P.n[2];
    for (int k = 0; k < 3; ++k) u[k] = e[k] - ds * P.n[k];
    L = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  }
  for (int k = 0; k < 3; ++k) q[k] = P.c[k] + P.rho * (u[k] / L);
}

__device__ __forceinline__ bool exposed_local(const SesView& S, const LineLists& L, int64_t a0, int64_t a1,
                                              const double* q, int e0, int e1) {
  for (int64_t k = a0; k < a1; ++k) {
    const int m = L.la[k];
    if (m == e0 || m == e1) continue;
    const double Rm = S.r[m] + S.rp;
    if (sq3(q, S.xyz + 3 * m) < Rm * Rm) return false;
  }
  return true;
}

// 1 inside the SES, 0 outside, for the point p at t on line li. Only the items whose
// axial coordinate lies within a window around t can satisfy any test here (an atom
// containing p, a covered candidate within rp of p, a circle or vertex within rp).
__device__ int ses_classify(const SesView& S, const LineLists& L, int64_t li, const double* p, double t) {
  int64_t a0, a1, p0, p1, v0, v1;
  key_window(L.ka, L.oa[li], L.oa[li + 1], t, S.wa, &a0, &a1);
  bool in_sas = false;
  for (int64_t k = a0; k < a1; ++k) {
    const int i = L.la[k];
    const double dd = sq3(p, S.xyz + 3 * i);
    if (dd < S.r[i] * S.r[i]) return 1;
    const double Ri = S.r[i] + S.rp;
    if (dd < Ri * Ri) in_sas = true;
  }
  if (!in_sas) return 0;
  const double rp2 = S.rp * S.rp;
  for (int64_t k = a0; k < a1; ++k) {
    const int i = L.la[k];
    const double* a = S.xyz + 3 * i;
    const double dd = sqrt(sq3(p, a));
    const double Ri = S.r[i] + S.rp;
    const double gap = Ri - dd;
    if (!(dd > 0.0) || fabs(gap) > S.rp) continue;
    double q[3];
    for (int c = 0; c < 3; ++c) q[c] = a[c] + Ri * ((p[c] - a[c]) / dd);
    if (exposed_local(S, L, a0, a1, q, i, -1)) return 0;
  }
  key_window(L.kp, L.op[li], L.op[li + 1], t, S.wp, &p0, &p1);
  for (int64_t k = p0; k < p1; ++k) {
    const SesPairD& P = S.pairs[L.lp[k]];
    double q[3];
    circle_point(P, p, q);
    if (sq3(p, q) <= rp2 && exposed_local(S, L, a0, a1, q, P.i, P.j)) return 0;
  }
  key_window(L.kv, L.ov[li], L.ov[li + 1], t, S.wv, &v0, &v1);
  for (int64_t k = v0; k < v1; ++k)
    if (sq3(p, S.verts + 3 * L.lv[k]) <= rp2) return 0;
  return 1;
}

// ---- item-parallel candidates and classification (one thread per list item or per
// candidate, which balances lines of very different lengths)

// line owning position k of a list whose lines start at off[0 .. nl] (off[nl] = total)
__device__ __forceinline__ int64_t line_of(const int64_t* off, int64_t nl, int64_t k) {
  int64_t l = 0, h = nl;  // largest l with off[l] <= k
  while (h - l > 1) {
    const int64_t m = (l + h) >> 1;
    if (off[m] <= k) l = m;
    else h = m;
  }
  return l;
}

// kind 0 atoms, 1 pairs (tori), 2 vertices: candidates of list item k (count, or
// written at out[..] when out is not null)
template <int KIND>
__global__ void k_ses_item_cands(int64_t n_items, int64_t nl, const uint64_t* __restrict__ lu,
                                 const uint64_t* __restrict__ lv, SesView S, LineLists L, int axis, int d1, int d2,
                                 int64_t* __restrict__ cnt, const int64_t* __restrict__ pos,
                                 uint64_t* __restrict__ tkey, uint32_t* __restrict__ kind_id) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_items) return;
  const int64_t* offs = KIND == 0 ? L.oa : KIND == 1 ? L.op : L.ov;
  const int64_t li = line_of(offs, nl, k);
  const double u = ses_unbits(lu[li]), v = ses_unbits(lv[li]);
  int64_t o = pos ? pos[k] : 0, c = 0;
  auto emit = [&](double t, int id) {
    if (pos) {
      tkey[o] = ses_bits(t);
      kind_id[o] = ((uint32_t)KIND << 28) | (uint32_t)id;
      ++o;
    }
    ++c;
  };
  if (KIND == 0) {
    const int a = L.la[k];
    const double du = u - S.xyz[3 * a + d1], dv = v - S.xyz[3 * a + d2];
    const double q = S.r[a] * S.r[a] - (du * du + dv * dv);
    if (q > 0.0) {
      const double sq = sqrt(q);
      emit(S.xyz[3 * a + axis] - sq + 0.0, a);
      emit(S.xyz[3 * a + axis] + sq + 0.0, a);
    }
  } else if (KIND == 2) {
    const int id = L.lv[k];
    const double* cc = S.verts + 3 * id;
    const double du = u - cc[d1], dv = v - cc[d2];
    const double q = S.rp * S.rp - (du * du + dv * dv);
    if (q > 0.0) {
      const double sq = sqrt(q);
      emit(cc[axis] - sq + 0.0, id);
      emit(cc[axis] + sq + 0.0, id);
    }
  } else {
    const int id = L.lp[k];
    if (pos) torus_roots(S.pairs[id], S.rp, axis, d1, d2, u, v, [&](double t) { emit(t, id); });
    else torus_roots(S.pairs[id], S.rp, axis, d1, d2, u, v, [&](double) { ++c; });
  }
  if (!pos) cnt[k] = c;
}

// per line: its candidate total from the items' scanned counts (atoms | vertices | tori)
__global__ void k_ses_line_total(int64_t nl, LineLists L, const int64_t* __restrict__ sa,
                                 const int64_t* __restrict__ sp, const int64_t* __restrict__ sv,
                                 int64_t* __restrict__ tot) {
  const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nl) return;
  tot[li] = (sa[L.oa[li + 1]] - sa[L.oa[li]]) + (sv[L.ov[li + 1]] - sv[L.ov[li]]) + (sp[L.op[li + 1]] - sp[L.op[li]]);
}

// output position of every item's first candidate: the line's block holds the atoms'
// candidates, then the vertices', then the tori's
template <int KIND>
__global__ void k_ses_item_pos(int64_t n_items, int64_t nl, LineLists L, const int64_t* __restrict__ coff,
                               const int64_t* __restrict__ sa, const int64_t* __restrict__ sp,
                               const int64_t* __restrict__ sv, int64_t* __restrict__ pos) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_items) return;
  const int64_t* offs = KIND == 0 ? L.oa : KIND == 1 ? L.op : L.ov;
  const int64_t li = line_of(offs, nl, k);
  const int64_t A = sa[L.oa[li + 1]] - sa[L.oa[li]];
  const int64_t V = sv[L.ov[li + 1]] - sv[L.ov[li]];
  int64_t p = coff[li];
  if (KIND == 0) p += sa[k] - sa[L.oa[li]];
  else if (KIND == 2) p += A + (sv[k] - sv[L.ov[li]]);
  else p += A + V + (sp[k] - sp[L.op[li]]);
  pos[k] = p;
}

// state of the open interval after each group of equal t (2 marks positions that do
// not start a group): 1 inside the SES, 0 outside
__global__ void k_ses_state(int64_t nc, int64_t nl, const uint64_t* __restrict__ lu, const uint64_t* __restrict__ lv,
                            SesView S, LineLists L, int axis, int d1, int d2, const int64_t* __restrict__ coff,
                            const uint64_t* __restrict__ tkey, uint8_t* __restrict__ state) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nc) return;
  const int64_t li = line_of(coff, nl, k);
  const int64_t b = coff[li], e = coff[li + 1];
  if (k > b && tkey[k] == tkey[k - 1]) {
    state[k] = 2;
    return;
  }
  int64_t h = k + 1;
  while (h < e && tkey[h] == tkey[k]) ++h;
  if (h >= e) {
    state[k] = 0;
    return;
  }
  double p[3];
  p[axis] = 0.5 * (ses_unbits(tkey[k]) + ses_unbits(tkey[h]));
  p[d1] = ses_unbits(lu[li]);
  p[d2] = ses_unbits(lv[li]);
  state[k] = (uint8_t)ses_classify(S, L, li, p, p[axis]);
}

// a crossing at every group start whose state differs from the previous group's
__global__ void k_ses_flag(int64_t nc, int64_t nl, const int64_t* __restrict__ coff, const uint8_t* __restrict__ state,
                           uint8_t* __restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nc) return;
  if (state[k] == 2) {
    flag[k] = 0;
    return;
  }
  const int64_t li = line_of(coff, nl, k);
  const int64_t b = coff[li];
  int before = 0;
  if (k > b) {
    int64_t j = k - 1;
    while (state[j] == 2) --j;  // the previous group's start (b starts a group)
    before = state[j];
  }
  flag[k] = state[k] != before ? 1 : 0;
}

__global__ void k_ses_line_cross(int64_t nl, const int64_t* __restrict__ coff, const uint8_t* __restrict__ flag,
                                 int64_t* __restrict__ cnt) {
  const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nl) return;
  int64_t c = 0;
  for (int64_t k = coff[li]; k < coff[li + 1]; ++k) c += flag[k];
  cnt[li] = c;
}

__global__ void k_ses_nonzero(int64_t n, const int64_t* __restrict__ c, uint8_t* __restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = c[i] > 0 ? 1 : 0;
}

__global__ void k_ses_ray_keys(int64_t nr, const int32_t* __restrict__ kept, const uint64_t* __restrict__ lu,
                               const uint64_t* __restrict__ lv, const int64_t* __restrict__ cnt,
                               double* __restrict__ key, int64_t* __restrict__ ptr) {
  const int64_t rr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (rr >= nr) return;
  const int64_t li = kept[rr];
  key[2 * rr] = ses_unbits(lu[li]);
  key[2 * rr + 1] = ses_unbits(lv[li]);
  ptr[rr] = cnt[li];
}

__global__ void k_ses_emit(int64_t nr, const int32_t* __restrict__ kept, const int64_t* __restrict__ ptr,
                           const int64_t* __restrict__ off, const uint64_t* __restrict__ tkey,
                           const uint32_t* __restrict__ kind_id, const uint8_t* __restrict__ flag,
                           const double* __restrict__ key, SesView S, int axis, int d1, int d2,
                           double* __restrict__ inters, double* __restrict__ normals) {
  const int64_t rr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (rr >= nr) return;
  const int64_t li = kept[rr];
  int64_t o = ptr[rr];
  for (int64_t k = off[li]; k < off[li + 1]; ++k) {
    if (!flag[k]) continue;
    const double t = ses_unbits(tkey[k]);
    const int kind = (int)(kind_id[k] >> 28);
    const int id = (int)(kind_id[k] & 0x0fffffffu);
    double p[3], n[3];
    p[axis] = t;
    p[d1] = key[2 * rr];
    p[d2] = key[2 * rr + 1];
    if (kind == 0) {
      const double* a = S.xyz + 3 * id;
      for (int c = 0; c < 3; ++c) n[c] = (p[c] - a[c]) / S.r[id];
    } else if (kind == 2) {
      const double* a = S.verts + 3 * id;
      for (int c = 0; c < 3; ++c) n[c] = (a[c] - p[c]) / S.rp;
    } else {
      double q[3];
      circle_point(S.pairs[id], p, q);
      for (int c = 0; c < 3; ++c) n[c] = (q[c] - p[c]) / S.rp;
    }
    inters[o] = t;
    for (int c = 0; c < 3; ++c) normals[3 * o + c] = n[c];
    ++o;
  }
}

// ------------------------------------------------- SAS circle exposure on the device
// One wave per candidate circle (pair i < j, in the host's order) runs the test of
// circle_exposed below on the balls of atom i, deepest first, 64 at a time: a ball
// covering the whole circle (S >= A) buries it; every other ball that reaches the
// circle adds the arc it covers, as the same [start, end] angle intervals, to the
// wave's part of LDS; after each group of balls the arcs so far are tested for a gap,
// and a circle they cover is buried (more arcs only cover more). Every other step is
// the host's arithmetic (IEEE + - * / sqrt in the same order); the arc angles use the
// device's acos and atan2, so an arc end can differ from the host's in the last bit,
// which changes a decision only where two arc ends meet to within rounding.
// out: 0 buried, 1 exposed, 2 more arcs than the wave holds (decided on the host).
constexpr int kCxWaves = 2;
constexpr int kCxCap = 1024;  // arc intervals per wave (16 KB of LDS)

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whether the n intervals [s_k, e_k] leave part of [0, 2 pi) open. The sorted sweep of
// circle_exposed finds a gap exactly when some v in {0} and the ends below 2 pi lies in
// no interval as s <= v < e, so each lane tests its share of those v against all.
__device__ bool arcs_leave_gap(const double* s, const double* e, int n, double two_pi) {
  bool gap = false;
  for (int q = (int)(threadIdx.x & 63); q <= n; q += 64) {
    const double v = q == n ? 0.0 : e[q];
    if (!(v < two_pi)) continue;
    bool held = false;
    for (int k = 0; k < n && !held; ++k) held = s[k] <= v && v < e[k];
    gap = gap || !held;
  }
  return __any(gap) != 0;
}

__global__ void __launch_bounds__(64 * kCxWaves) k_circle_exposed(int64_t ncand, const int32_t* __restrict__ ci,
                                                                  const int32_t* __restrict__ cj,
                                                                  const double* __restrict__ xyz,
                                                                  const double* __restrict__ R,
                                                                  const int64_t* __restrict__ dptr,
                                                                  const int32_t* __restrict__ didx,
                                                                  uint8_t* __restrict__ out) {
  __shared__ double s_lo[kCxWaves][kCxCap], s_hi[kCxWaves][kCxCap];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t pid = (int64_t)blockIdx.x * kCxWaves + wv;
  if (pid >= ncand) return;  // the whole wave; no block-wide barrier follows
  const int i = ci[pid], j = cj[pid];
  const double* ai = xyz + 3 * i;
  const double* aj = xyz + 3 * j;
  // the circle of the pair, as the host computes it
  const double dx = ai[0] - aj[0], dy = ai[1] - aj[1], dz = ai[2] - aj[2];
  const double dd = sqrt(dx * dx + dy * dy + dz * dz);
  const double Ri = R[i], Rj = R[j];
  const double t = (dd * dd + Ri * Ri - Rj * Rj) / (2.0 * dd);
  const double rho = sqrt(Ri * Ri - t * t);
  double n[3], c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    n[k] = (aj[k] - ai[k]) / dd;
    c[k] = ai[k] + t * n[k];
  }
  // the circle's (u, w) frame, as circle_exposed builds it
  const int ax = fabs(n[0]) < 0.9 ? 0 : 1;
  double u[3] = {0.0, 0.0, 0.0};
  u[ax] = 1.0;
  const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) u[k] = u[k] - un * n[k];
  const double ul = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) u[k] = u[k] / ul;
  const double w[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
  const double two_pi = 2.0 * 3.14159265358979323846;
  double* lo = s_lo[wv];
  double* hi = s_hi[wv];
  int nv = 0;  // intervals so far (the same in every lane)
  const int64_t b1 = dptr[i + 1];
  for (int64_t b = dptr[i]; b < b1; b += 64) {
    const int64_t k = b + lane;
    const int m = k < b1 ? didx[k] : j;
    bool cov = false, arc = false, two = false;
    double s0 = 0.0, e0 = 0.0;
    if (m != j) {
      const double* am = xyz + 3 * m;
      const double d[3] = {c[0] - am[0], c[1] - am[1], c[2] - am[2]};
      const double far = R[m] * (1.0 + 1e-9) + rho;
      if (!(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > far * far)) {
        const double P = 2.0 * rho * (d[0] * u[0] + d[1] * u[1] + d[2] * u[2]);
        const double Q = 2.0 * rho * (d[0] * w[0] + d[1] * w[1] + d[2] * w[2]);
        const double S = R[m] * R[m] - (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - rho * rho;
        const double A = sqrt(P * P + Q * Q);
        cov = S >= A;
        if (!cov && !(S <= -A)) {
          arc = true;
          const double beta = acos(S / A);
          s0 = atan2(Q, P) + beta;
          if (s0 < 0.0) s0 += two_pi;
          if (s0 >= two_pi) s0 -= two_pi;
          e0 = s0 + (two_pi - 2.0 * beta);
          two = e0 > two_pi;
        }
      }
    }
    if (__any(cov)) {
      if (lane == 0) out[pid] = 0;
      return;
    }
    const int cnt = arc ? (two ? 2 : 1) : 0;
    int pre = cnt;  // inclusive prefix over the lanes
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(pre, off, 64);
      if (lane >= off) pre += v;
    }
    const int total = __shfl(pre, 63, 64);
    if (nv + total > kCxCap) {
      if (lane == 0) out[pid] = 2;
      return;
    }
    if (arc) {
      const int q = nv + pre - cnt;
      if (two) {
        lo[q] = s0;
        hi[q] = two_pi;
        lo[q + 1] = 0.0;
        hi[q + 1] = e0 - two_pi;
      } else {
        lo[q] = s0;
        hi[q] = e0;
      }
    }
    nv += total;
    if (total > 0) {
      wave_lds_sync();
      const bool gap = arcs_leave_gap(lo, hi, nv, two_pi);
      wave_lds_sync();
      if (!gap) {
        if (lane == 0) out[pid] = 0;
        return;
      }
    }
  }
  // every ball taken: a gap remains, or no ball reaches the circle
  if (lane == 0) out[pid] = 1;
}

// ---------------------------------------------------------------- host side
struct SesHost {
  std::vector<SesPairD> pairs;
  std::vector<double> verts;
};

double sqd(const double* a, const double* b) {
  const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
  return x * x + y * y + z * z;
}

// Whether some arc of the SAS circle (c, n, rho) of pair (i, j) lies outside the other
// SAS balls overlapping ball i: each covers the arc where P cos th + Q sin th < S in the
// circle's (u, w) frame; the covered open arcs are swept in angle for a gap. Buried
// circles carry no SES patch and are dropped (an exposed SAS vertex lies on three
// exposed circles). Host arithmetic, with the C library's atan2 and acos.
//
// A circle is buried iff one ball covers all of it (S >= A) or the union of the arcs
// does, and that union only grows with more balls. So the balls are taken in the order
// given (deepest over ball i's centre first), and the arcs gathered so far are swept
// after 16, 64, 256 ... balls: in dense packings most circles are covered early, and
// only circles that stay uncovered are swept with every arc. The decision is the one a
// single sweep over every ball makes.
bool circle_exposed(const double* c, const double* n, double rho, int j, const std::vector<int>& balls,
                    const double* xyz, const std::vector<double>& R) {
  const int ax = std::fabs(n[0]) < 0.9 ? 0 : 1;
  double u[3] = {0, 0, 0};
  u[ax] = 1.0;
  const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
  for (int k = 0; k < 3; ++k) u[k] = u[k] - un * n[k];
  const double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int k = 0; k < 3; ++k) u[k] = u[k] / ul;
  const double w[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
  const double two_pi = 2.0 * 3.14159265358979323846;
  thread_local std::vector<std::pair<double, double>> iv;
  iv.clear();
  // whether the arcs in iv leave no gap
  auto covered = [&]() {
    if (iv.empty()) return false;
    std::sort(iv.begin(), iv.end());
    double cur = 0.0;
    for (const auto& a : iv) {
      if (a.first > cur) return false;
      cur = std::max(cur, a.second);
    }
    return !(cur < two_pi);
  };
  size_t seen = 0, next = 16;
  for (int m : balls) {  // every ball overlapping ball i
    if (m == j) continue;
    const double* am = xyz + 3 * m;
    const double d[3] = {c[0] - am[0], c[1] - am[1], c[2] - am[2]};
    // a ball farther than R_m (1 + 1e-9) from every point of the circle covers no arc:
    // S + A (= R_m^2 - its squared distance to the circle) is then below -2e-9 R_m^2,
    // far outside rounding, so the sweep would skip it too
    const double far = R[(size_t)m] * (1.0 + 1e-9) + rho;
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > far * far) continue;
    const double P = 2.0 * rho * (d[0] * u[0] + d[1] * u[1] + d[2] * u[2]);
    const double Q = 2.0 * rho * (d[0] * w[0] + d[1] * w[1] + d[2] * w[2]);
    const double S = R[(size_t)m] * R[(size_t)m] - (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - rho * rho;
    const double A = std::sqrt(P * P + Q * Q);
    if (S >= A) return false;
    if (++seen == next) {
      if (covered()) return false;
      next *= 4;
    }
    if (S <= -A) continue;
    const double beta = std::acos(S / A);
    double s0 = std::atan2(Q, P) + beta;
    if (s0 < 0.0) s0 += two_pi;
    if (s0 >= two_pi) s0 -= two_pi;
    const double e0 = s0 + (two_pi - 2.0 * beta);
    if (e0 > two_pi) {
      iv.push_back({s0, two_pi});
      iv.push_back({0.0, e0 - two_pi});
    } else {
      iv.push_back({s0, e0});
    }
  }
  return !covered();
}

// SAS circles of overlapping pairs (i < j, in index order) that are not buried, and
// exposed SAS vertices (per kept pair in that order, third atom k > j ascending, +z
// before -z)
// f(k) for k in [0, n) on the host's hardware threads (each k handled once)
template <class F>
void host_parallel(int64_t n, F&& f) {
  // the hardware threads, at most OMP_NUM_THREADS when that is set (the CPU share a
  // process is given) and at most 64
  int64_t T = std::min<int64_t>((int64_t)std::thread::hardware_concurrency(), 64);
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const long v = std::strtol(e, nullptr, 10);
    if (v > 0) T = std::min<int64_t>(T, v);
  }
  T = std::max<int64_t>(T, 1);
  if (n < 256 || T == 1) {
    for (int64_t k = 0; k < n; ++k) f(k);
    return;
  }
  std::vector<std::thread> th;
  for (int64_t t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      for (int64_t k = t; k < n; k += T) f(k);
    });
  for (auto& x : th) x.join();
}

SesHost ses_build(ngpb_ctx* c, int64_t na, const double* xyz, const double* r, double rp) {
  SesHost H;
  std::vector<double> R((size_t)na);
  double Rmax = 0.0;
  for (int64_t i = 0; i < na; ++i) {
    R[(size_t)i] = r[i] + rp;
    Rmax = std::max(Rmax, R[(size_t)i]);
  }
  if (na == 0) return H;
  // cell grid of edge 2 Rmax: SAS balls that overlap lie in adjacent cells
  const double cell = std::max(2.0 * Rmax, 1e-3);
  double lo[3] = {1e300, 1e300, 1e300};
  for (int64_t i = 0; i < na; ++i)
    for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], xyz[3 * i + d]);
  std::vector<std::array<int, 3>> ci((size_t)na);
  int gn[3] = {1, 1, 1};
  for (int64_t i = 0; i < na; ++i)
    for (int d = 0; d < 3; ++d) {
      ci[(size_t)i][d] = (int)std::floor((xyz[3 * i + d] - lo[d]) / cell);
      gn[d] = std::max(gn[d], ci[(size_t)i][d] + 1);
    }
  std::vector<std::vector<int>> grid((size_t)gn[0] * gn[1] * gn[2]);
  for (int64_t i = 0; i < na; ++i)
    grid[((size_t)ci[(size_t)i][2] * gn[1] + ci[(size_t)i][1]) * gn[0] + ci[(size_t)i][0]].push_back((int)i);
  std::vector<std::vector<int>> over((size_t)na), nb((size_t)na);  // overlapping SAS balls; circle pairs
  host_parallel(na, [&](int64_t i) {
    const auto& c = ci[(size_t)i];
    for (int z = c[2] - 1; z <= c[2] + 1; ++z)
      for (int y = c[1] - 1; y <= c[1] + 1; ++y)
        for (int x = c[0] - 1; x <= c[0] + 1; ++x) {
          if (x < 0 || y < 0 || z < 0 || x >= gn[0] || y >= gn[1] || z >= gn[2]) continue;
          for (int j : grid[((size_t)z * gn[1] + y) * gn[0] + x]) {
            if (j == i) continue;
            const double dd = std::sqrt(sqd(xyz + 3 * i, xyz + 3 * j));
            if (dd < R[(size_t)i] + R[(size_t)j]) over[(size_t)i].push_back(j);
          }
        }
    std::sort(over[(size_t)i].begin(), over[(size_t)i].end());
  });
  // the same balls per atom, deepest over its centre first (|a_m - a_i| - R_m
  // ascending): the order in which circle_exposed tries them
  std::vector<std::vector<int>> deep((size_t)na);
  host_parallel(na, [&](int64_t i) {
    std::vector<std::pair<double, int>> k;
    k.reserve(over[(size_t)i].size());
    for (int m : over[(size_t)i]) k.push_back({std::sqrt(sqd(xyz + 3 * i, xyz + 3 * m)) - R[(size_t)m], m});
    std::sort(k.begin(), k.end());
    auto& v = deep[(size_t)i];
    v.reserve(k.size());
    for (const auto& e : k) v.push_back(e.second);
  });
  // the circle of a pair (i < j) of overlapping balls, if it has one
  auto circle_of = [&](int64_t i, int j, SesPairD& P) {
    const double* ai = xyz + 3 * i;
    const double* aj = xyz + 3 * j;
    const double dd = std::sqrt(sqd(ai, aj));
    if (!(dd < R[(size_t)i] + R[(size_t)j]) || !(dd > std::fabs(R[(size_t)i] - R[(size_t)j]))) return false;
    const double t = (dd * dd + R[(size_t)i] * R[(size_t)i] - R[(size_t)j] * R[(size_t)j]) / (2.0 * dd);
    const double rho2 = R[(size_t)i] * R[(size_t)i] - t * t;
    if (!(rho2 > 0.0)) return false;
    P.i = (int32_t)i;
    P.j = (int32_t)j;
    for (int k = 0; k < 3; ++k) {
      P.n[k] = (aj[k] - ai[k]) / dd;
      P.c[k] = ai[k] + t * P.n[k];
    }
    P.rho = std::sqrt(rho2);
    return true;
  };
  // candidate circles in (i, j) order; their exposure is tested on the device
  std::vector<std::vector<int32_t>> cand_j((size_t)na);
  host_parallel(na, [&](int64_t i) {
    SesPairD P;
    for (int j : over[(size_t)i])
      if (j > i && circle_of(i, j, P)) cand_j[(size_t)i].push_back(j);
  });
  std::vector<int32_t> pair_i, pair_j, didx;
  std::vector<int64_t> dptr((size_t)na + 1, 0);
  for (int64_t i = 0; i < na; ++i) {
    pair_i.insert(pair_i.end(), cand_j[(size_t)i].size(), (int32_t)i);
    pair_j.insert(pair_j.end(), cand_j[(size_t)i].begin(), cand_j[(size_t)i].end());
    dptr[(size_t)i + 1] = dptr[(size_t)i] + (int64_t)deep[(size_t)i].size();
  }
  didx.reserve((size_t)dptr[(size_t)na]);
  for (const auto& v : deep) didx.insert(didx.end(), v.begin(), v.end());
  const int64_t ncand = (int64_t)pair_i.size();
  std::vector<uint8_t> flag((size_t)ncand, 0);
  if (ncand > 0) {
    hipStream_t st = c->stream;
    DBuf<double> d_xyz(st), d_R(st);
    DBuf<int64_t> d_ptr(st);
    DBuf<int32_t> d_idx(st), d_ci(st), d_cj(st);
    DBuf<uint8_t> d_out(st);
    d_xyz.upload(xyz, (size_t)(3 * na), st);
    d_R.upload(R.data(), R.size(), st);
    d_ptr.upload(dptr.data(), dptr.size(), st);
    d_idx.upload(didx.empty() ? nullptr : didx.data(), std::max<size_t>(didx.size(), 1), st);
    d_ci.upload(pair_i.data(), pair_i.size(), st);
    d_cj.upload(pair_j.data(), pair_j.size(), st);
    d_out.resize((size_t)ncand);
    hipLaunchKernelGGL(k_circle_exposed, dim3((unsigned)((ncand + kCxWaves - 1) / kCxWaves)), dim3(64 * kCxWaves), 0,
                       st, ncand, d_ci.p, d_cj.p, d_xyz.p, d_R.p, d_ptr.p, d_idx.p, d_out.p);
    NGPB_HIP(hipGetLastError());
    NGPB_HIP(hipMemcpyAsync(flag.data(), d_out.p, (size_t)ncand, hipMemcpyDeviceToHost, st));
    NGPB_HIP(hipStreamSynchronize(st));
  }
  std::vector<std::vector<SesPairD>> pairs_of((size_t)na);
  std::vector<int64_t> first((size_t)na + 1, 0);
  for (int64_t i = 0; i < na; ++i) first[(size_t)i + 1] = first[(size_t)i] + (int64_t)cand_j[(size_t)i].size();
  host_parallel(na, [&](int64_t i) {
    for (int64_t q = first[(size_t)i]; q < first[(size_t)i + 1]; ++q) {
      if (flag[(size_t)q] == 0) continue;
      SesPairD P;
      circle_of(i, pair_j[(size_t)q], P);
      // more arcs than a wave holds: the host sweep decides
      if (flag[(size_t)q] == 2 && !circle_exposed(P.c, P.n, P.rho, P.j, deep[(size_t)i], xyz, R)) continue;
      pairs_of[(size_t)i].push_back(P);
    }
  });
  for (int64_t i = 0; i < na; ++i)
    for (const SesPairD& P : pairs_of[(size_t)i]) {
      H.pairs.push_back(P);
      nb[(size_t)P.i].push_back(P.j);
      nb[(size_t)P.j].push_back(P.i);
    }
  for (auto& v : nb) std::sort(v.begin(), v.end());
  auto exposed = [&](const double* q, int e0, int e1, int e2) {
    for (int m : over[(size_t)e0]) {  // any ball covering q overlaps ball e0
      if (m == e1 || m == e2) continue;
      if (sqd(q, xyz + 3 * m) < R[(size_t)m] * R[(size_t)m]) return false;
    }
    return true;
  };
  std::vector<std::vector<double>> verts_of(H.pairs.size());
  host_parallel((int64_t)H.pairs.size(), [&](int64_t pk) {
    const SesPairD& P = H.pairs[(size_t)pk];
    for (int k : nb[(size_t)P.j]) {
      if (k <= P.j || !std::binary_search(nb[(size_t)P.i].begin(), nb[(size_t)P.i].end(), k)) continue;
      const double* ai = xyz + 3 * P.i;
      const double* ak = xyz + 3 * k;
      double ex[3] = {P.n[0], P.n[1], P.n[2]};
      double w[3] = {ak[0] - ai[0], ak[1] - ai[1], ak[2] - ai[2]};
      const double ii = ex[0] * w[0] + ex[1] * w[1] + ex[2] * w[2];
      double ey[3] = {w[0] - ii * ex[0], w[1] - ii * ex[1], w[2] - ii * ex[2]};
      const double ey_n = std::sqrt(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
      if (!(ey_n > 0.0)) continue;
      for (int q = 0; q < 3; ++q) ey[q] = ey[q] / ey_n;
      const double ez[3] = {ex[1] * ey[2] - ex[2] * ey[1], ex[2] * ey[0] - ex[0] * ey[2], ex[0] * ey[1] - ex[1] * ey[0]};
      const double jj = ey[0] * w[0] + ey[1] * w[1] + ey[2] * w[2];
      const double dd = std::sqrt(sqd(ai, xyz + 3 * P.j));
      const double Ri = R[(size_t)P.i], Rj = R[(size_t)P.j], Rk = R[(size_t)k];
      const double x = (Ri * Ri - Rj * Rj + dd * dd) / (2.0 * dd);
      const double y = (Ri * Ri - Rk * Rk + ii * ii + jj * jj) / (2.0 * jj) - (ii / jj) * x;
      const double z2 = Ri * Ri - x * x - y * y;
      if (!(z2 > 0.0)) continue;
      const double z = std::sqrt(z2);
      for (int sgn = 0; sgn < 2; ++sgn) {
        const double zz = sgn == 0 ? z : -z;
        double v[3];
        for (int q = 0; q < 3; ++q) v[q] = ai[q] + x * ex[q] + y * ey[q] + zz * ez[q];
        if (exposed(v, P.i, P.j, k)) verts_of[(size_t)pk].insert(verts_of[(size_t)pk].end(), v, v + 3);
      }
    }
  });
  for (const auto& v : verts_of) H.verts.insert(H.verts.end(), v.begin(), v.end());
  return H;
}

// 2D bins of n items with transverse coordinates (c1[k], c2[k]), items ascending
void bin2(int64_t n, const std::vector<double>& c1, const std::vector<double>& c2, double cell,
          std::vector<int32_t>& ptr, std::vector<int32_t>& item, double* lo, int* gn) {
  lo[0] = lo[1] = 0.0;
  double hi[2] = {0.0, 0.0};
  if (n > 0) {
    lo[0] = hi[0] = c1[0];
    lo[1] = hi[1] = c2[0];
  }
  for (int64_t k = 0; k < n; ++k) {
    lo[0] = std::min(lo[0], c1[(size_t)k]);
    hi[0] = std::max(hi[0], c1[(size_t)k]);
    lo[1] = std::min(lo[1], c2[(size_t)k]);
    hi[1] = std::max(hi[1], c2[(size_t)k]);
  }
  gn[0] = (int)std::floor((hi[0] - lo[0]) / cell) + 1;
  gn[1] = (int)std::floor((hi[1] - lo[1]) / cell) + 1;
  ptr.assign((size_t)gn[0] * gn[1] + 1, 0);
  std::vector<int32_t> bin((size_t)std::max<int64_t>(n, 1));
  for (int64_t k = 0; k < n; ++k) {
    const int b0 = (int)std::floor((c1[(size_t)k] - lo[0]) / cell);
    const int b1 = (int)std::floor((c2[(size_t)k] - lo[1]) / cell);
    bin[(size_t)k] = b1 * gn[0] + b0;
    ++ptr[(size_t)bin[(size_t)k] + 1];
  }
  for (size_t i = 1; i < ptr.size(); ++i) ptr[i] += ptr[i - 1];
  item.assign((size_t)std::max<int64_t>(n, 1), 0);
  std::vector<int32_t> cur(ptr.begin(), ptr.end() - 1);
  for (int64_t k = 0; k < n; ++k) item[(size_t)cur[(size_t)bin[(size_t)k]]++] = (int32_t)k;
}

}  // namespace

// Device SES data of one molecule (built once per ngpb_cast_rays call).
struct SesDevice {
  std::vector<double> rmax_of;  // atom radii (host)
  DBuf<double> xyz, r, verts;
  DBuf<SesPairD> pairs;
  int64_t na = 0, npairs = 0, nverts = 0;
  double rp = 0.0, reach_cell = 1.0;
  explicit SesDevice(hipStream_t s) : xyz(s), r(s), verts(s), pairs(s) {}
};

SesDevice* ses_prepare(ngpb_ctx* c, int64_t na, const double* xyz, const double* r, double rp) {
  SesHost H = ses_build(c, na, xyz, r, rp);
  auto* D = new SesDevice(c->stream);
  D->rmax_of.assign(r, r + na);
  D->na = na;
  D->rp = rp;
  D->npairs = (int64_t)H.pairs.size();
  D->nverts = (int64_t)H.verts.size() / 3;
  double rmax = 0.0;
  for (int64_t a = 0; a < na; ++a) rmax = std::max(rmax, r[a]);
  D->reach_cell = std::max(rmax + 2.0 * rp + 1e-3, 1.0);  // every reach is below it
  D->xyz.upload(na ? xyz : nullptr, (size_t)(3 * na), c->stream);
  D->r.upload(na ? r : nullptr, (size_t)na, c->stream);
  if (!na) {
    D->xyz.resize(3);
    D->r.resize(1);
  }
  if (D->npairs) D->pairs.upload(H.pairs.data(), H.pairs.size(), c->stream);
  else D->pairs.resize(1);
  if (D->nverts) D->verts.upload(H.verts.data(), H.verts.size(), c->stream);
  else D->verts.resize(3);
  NGPB_HIP(hipStreamSynchronize(c->stream));
  return D;
}

void ses_release(SesDevice* D) { delete D; }

void ses_counts(const SesDevice* D, int64_t* npairs, int64_t* nverts) {
  *npairs = D->npairs;
  *nverts = D->nverts;
}

// SES rays of the nl distinct lines (lu, lv) of one axis into c->rays[axis]
void ses_cast_lines(ngpb_ctx* c, const SesDevice* D, int axis, int64_t nl, const uint64_t* lu, const uint64_t* lv,
                    const double* h_xyz) {
  const int d1 = axis == 0 ? 1 : 0;
  const int d2 = axis == 2 ? 1 : 2;
  const double cell = D->reach_cell;
  // bins of atoms, pairs and vertices in the (d1, d2) plane
  std::vector<double> a1((size_t)D->na), a2((size_t)D->na);
  for (int64_t a = 0; a < D->na; ++a) {
    a1[(size_t)a] = h_xyz[3 * a + d1];
    a2[(size_t)a] = h_xyz[3 * a + d2];
  }
  std::vector<SesPairD> hp((size_t)D->npairs);
  std::vector<double> hv((size_t)(3 * D->nverts));
  if (D->npairs) D->pairs.download(hp.data(), hp.size(), c->stream);
  if (D->nverts) D->verts.download(hv.data(), hv.size(), c->stream);
  NGPB_HIP(hipStreamSynchronize(c->stream));
  std::vector<double> p1((size_t)D->npairs), p2((size_t)D->npairs), v1((size_t)D->nverts), v2((size_t)D->nverts);
  for (int64_t k = 0; k < D->npairs; ++k) {
    p1[(size_t)k] = hp[(size_t)k].c[d1];
    p2[(size_t)k] = hp[(size_t)k].c[d2];
  }
  for (int64_t k = 0; k < D->nverts; ++k) {
    v1[(size_t)k] = hv[(size_t)(3 * k + d1)];
    v2[(size_t)k] = hv[(size_t)(3 * k + d2)];
  }
  std::vector<int32_t> bpa, bia, bpp, bip, bpv, biv;
  double loa[2], lop[2], lov[2];
  int gna[2], gnp[2], gnv[2];
  bin2(D->na, a1, a2, cell, bpa, bia, loa, gna);
  bin2(D->npairs, p1, p2, cell, bpp, bip, lop, gnp);
  bin2(D->nverts, v1, v2, cell, bpv, biv, lov, gnv);
  DBuf<int32_t> dbpa(c->stream), dbia(c->stream), dbpp(c->stream), dbip(c->stream), dbpv(c->stream), dbiv(c->stream);
  dbpa.upload(bpa.data(), bpa.size(), c->stream);
  dbia.upload(bia.data(), bia.size(), c->stream);
  dbpp.upload(bpp.data(), bpp.size(), c->stream);
  dbip.upload(bip.data(), bip.size(), c->stream);
  dbpv.upload(bpv.data(), bpv.size(), c->stream);
  dbiv.upload(biv.data(), biv.size(), c->stream);
  SesView S;
  double rmax = 0.0, rhomax = 0.0;
  for (int64_t a = 0; a < D->na; ++a) rmax = std::max(rmax, D->rmax_of[(size_t)a]);
  for (int64_t k = 0; k < D->npairs; ++k) rhomax = std::max(rhomax, hp[(size_t)k].rho);
  S.wa = rmax + 2.0 * D->rp + 1e-6;  // R + rp bounds every atom test
  S.wp = rhomax + D->rp + 1e-6;
  S.wv = D->rp + 1e-6;
  S.xyz = D->xyz.p;
  S.r = D->r.p;
  S.na = D->na;
  S.rp = D->rp;
  S.pairs = D->pairs.p;
  S.verts = D->verts.p;
  S.ga = SesGrid2{dbpa.p, dbia.p, loa[0], loa[1], cell, gna[0], gna[1]};
  S.gp = SesGrid2{dbpp.p, dbip.p, lop[0], lop[1], cell, gnp[0], gnp[1]};
  S.gv = SesGrid2{dbpv.p, dbiv.p, lov[0], lov[1], cell, gnv[0], gnv[1]};
  const dim3 B(128), G((unsigned)blocks_for(nl, 128));
  // per-line lists of the atoms, pairs and vertices within reach
  DBuf<int64_t> oa(c->stream), op(c->stream), ov(c->stream);
  oa.resize((size_t)nl + 1);
  op.resize((size_t)nl + 1);
  ov.resize((size_t)nl + 1);
  hipLaunchKernelGGL(k_ses_list_count, G, B, 0, c->stream, nl, lu, lv, S, d1, d2, oa.p, op.p, ov.p);
  NGPB_HIP(hipGetLastError());
  const int64_t ta = exclusive_scan(c, oa.p, oa.p, nl);
  const int64_t tp = exclusive_scan(c, op.p, op.p, nl);
  const int64_t tv = exclusive_scan(c, ov.p, ov.p, nl);
  NGPB_HIP(hipMemcpyAsync(oa.p + nl, &ta, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  NGPB_HIP(hipMemcpyAsync(op.p + nl, &tp, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  NGPB_HIP(hipMemcpyAsync(ov.p + nl, &tv, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  DBuf<int32_t> la(c->stream), lpp(c->stream), lvv(c->stream), la2(c->stream), lpp2(c->stream), lvv2(c->stream);
  DBuf<uint64_t> ka(c->stream), kp(c->stream), kv(c->stream), ka2(c->stream), kp2(c->stream), kv2(c->stream);
  for (auto* b : {&la, &la2}) b->resize((size_t)std::max<int64_t>(ta, 1));
  for (auto* b : {&lpp, &lpp2}) b->resize((size_t)std::max<int64_t>(tp, 1));
  for (auto* b : {&lvv, &lvv2}) b->resize((size_t)std::max<int64_t>(tv, 1));
  for (auto* b : {&ka, &ka2}) b->resize((size_t)std::max<int64_t>(ta, 1));
  for (auto* b : {&kp, &kp2}) b->resize((size_t)std::max<int64_t>(tp, 1));
  for (auto* b : {&kv, &kv2}) b->resize((size_t)std::max<int64_t>(tv, 1));
  hipLaunchKernelGGL(k_ses_list_fill, G, B, 0, c->stream, nl, lu, lv, S, axis, d1, d2, oa.p, op.p, ov.p, la.p, lpp.p,
                     lvv.p, ka.p, kp.p, kv.p);
  NGPB_HIP(hipGetLastError());
  // each line's lists by axial coordinate
  auto seg_sort = [&](int64_t n, DBuf<uint64_t>& k_in, DBuf<uint64_t>& k_out, DBuf<int32_t>& v_in,
                      DBuf<int32_t>& v_out, const int64_t* offs) {
    if (n <= 0) return;
    size_t tb = 0;
    NGPB_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tb, k_in.p, k_out.p, v_in.p, v_out.p, (int)n,
                                                         (int)nl, offs, offs + 1, 0, 64, c->stream));
    DBuf<uint8_t> tmp(c->stream);
    tmp.resize(std::max<size_t>(tb, 1));
    NGPB_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp.p, tb, k_in.p, k_out.p, v_in.p, v_out.p, (int)n, (int)nl,
                                                         offs, offs + 1, 0, 64, c->stream));
  };
  NGPB_REQUIRE(ta < INT32_MAX && tp < INT32_MAX && tv < INT32_MAX, NGPB_E_ARG, "SES line lists too long");
  seg_sort(ta, ka, ka2, la, la2, oa.p);
  seg_sort(tp, kp, kp2, lpp, lpp2, op.p);
  seg_sort(tv, kv, kv2, lvv, lvv2, ov.p);
  const LineLists L{ta > 0 ? la2.p : la.p, tp > 0 ? lpp2.p : lpp.p, tv > 0 ? lvv2.p : lvv.p, oa.p, op.p, ov.p,
                    ta > 0 ? ka2.p : ka.p, tp > 0 ? kp2.p : kp.p, tv > 0 ? kv2.p : kv.p};
  // candidates: counted and written per list item, then sorted per line by (t, kind, id)
  auto items = [&](int64_t n) { return dim3((unsigned)blocks_for(std::max<int64_t>(n, 1), 128)); };
  DBuf<int64_t> sa(c->stream), sp(c->stream), sv(c->stream);
  sa.resize((size_t)ta + 1);
  sp.resize((size_t)tp + 1);
  sv.resize((size_t)tv + 1);
  if (ta) hipLaunchKernelGGL(k_ses_item_cands<0>, items(ta), B, 0, c->stream, ta, nl, lu, lv, S, L, axis, d1, d2, sa.p,
                             nullptr, nullptr, nullptr);
  if (tp) hipLaunchKernelGGL(k_ses_item_cands<1>, items(tp), B, 0, c->stream, tp, nl, lu, lv, S, L, axis, d1, d2, sp.p,
                             nullptr, nullptr, nullptr);
  if (tv) hipLaunchKernelGGL(k_ses_item_cands<2>, items(tv), B, 0, c->stream, tv, nl, lu, lv, S, L, axis, d1, d2, sv.p,
                             nullptr, nullptr, nullptr);
  NGPB_HIP(hipGetLastError());
  const int64_t na_c = ta ? exclusive_scan(c, sa.p, sa.p, ta) : 0;
  const int64_t np_c = tp ? exclusive_scan(c, sp.p, sp.p, tp) : 0;
  const int64_t nv_c = tv ? exclusive_scan(c, sv.p, sv.p, tv) : 0;
  NGPB_HIP(hipMemcpyAsync(sa.p + ta, &na_c, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  NGPB_HIP(hipMemcpyAsync(sp.p + tp, &np_c, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  NGPB_HIP(hipMemcpyAsync(sv.p + tv, &nv_c, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  DBuf<int64_t> coff(c->stream);
  coff.resize((size_t)nl + 1);
  hipLaunchKernelGGL(k_ses_line_total, G, B, 0, c->stream, nl, L, sa.p, sp.p, sv.p, coff.p);
  NGPB_HIP(hipGetLastError());
  const int64_t nc = exclusive_scan(c, coff.p, coff.p, nl);
  NGPB_HIP(hipMemcpyAsync(coff.p + nl, &nc, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  NGPB_REQUIRE(nc < INT32_MAX, NGPB_E_ARG, "too many SES candidates for one axis");
  const size_t ncs = (size_t)std::max<int64_t>(nc, 1);
  DBuf<uint64_t> tk(c->stream), tk2(c->stream);
  DBuf<uint32_t> kid(c->stream), kid2(c->stream);
  tk.resize(ncs);
  tk2.resize(ncs);
  kid.resize(ncs);
  kid2.resize(ncs);
  {
    DBuf<int64_t> pos(c->stream);
    pos.resize((size_t)std::max<int64_t>(std::max(ta, std::max(tp, tv)), 1));
    if (ta) {
      hipLaunchKernelGGL(k_ses_item_pos<0>, items(ta), B, 0, c->stream, ta, nl, L, coff.p, sa.p, sp.p, sv.p, pos.p);
      hipLaunchKernelGGL(k_ses_item_cands<0>, items(ta), B, 0, c->stream, ta, nl, lu, lv, S, L, axis, d1, d2, nullptr,
                         pos.p, tk.p, kid.p);
    }
    if (tv) {
      hipLaunchKernelGGL(k_ses_item_pos<2>, items(tv), B, 0, c->stream, tv, nl, L, coff.p, sa.p, sp.p, sv.p, pos.p);
      hipLaunchKernelGGL(k_ses_item_cands<2>, items(tv), B, 0, c->stream, tv, nl, lu, lv, S, L, axis, d1, d2, nullptr,
                         pos.p, tk.p, kid.p);
    }
    if (tp) {
      hipLaunchKernelGGL(k_ses_item_pos<1>, items(tp), B, 0, c->stream, tp, nl, L, coff.p, sa.p, sp.p, sv.p, pos.p);
      hipLaunchKernelGGL(k_ses_item_cands<1>, items(tp), B, 0, c->stream, tp, nl, lu, lv, S, L, axis, d1, d2, nullptr,
                         pos.p, tk.p, kid.p);
    }
    NGPB_HIP(hipGetLastError());
  }
  if (nc > 0) {
    size_t tb1 = 0, tb2 = 0;
    NGPB_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tb1, kid.p, kid2.p, tk.p, tk2.p, (int)nc, (int)nl,
                                                         coff.p, coff.p + 1, 0, 32, c->stream));
    NGPB_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tb2, tk2.p, tk.p, kid2.p, kid.p, (int)nc, (int)nl,
                                                         coff.p, coff.p + 1, 0, 64, c->stream));
    DBuf<uint8_t> tmp(c->stream);
    tmp.resize(std::max(tb1, tb2));
    // by (kind, id), then stably by t
    NGPB_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp.p, tb1, kid.p, kid2.p, tk.p, tk2.p, (int)nc, (int)nl,
                                                         coff.p, coff.p + 1, 0, 32, c->stream));
    NGPB_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp.p, tb2, tk2.p, tk.p, kid2.p, kid.p, (int)nc, (int)nl,
                                                         coff.p, coff.p + 1, 0, 64, c->stream));
  }
  DBuf<uint8_t> flag(c->stream);
  flag.resize(ncs);
  DBuf<int64_t> xcnt(c->stream);
  xcnt.resize((size_t)nl);
  {
    DBuf<uint8_t> state(c->stream);
    state.resize(ncs);
    if (nc > 0) {
      hipLaunchKernelGGL(k_ses_state, items(nc), B, 0, c->stream, nc, nl, lu, lv, S, L, axis, d1, d2, coff.p, tk.p,
                         state.p);
      hipLaunchKernelGGL(k_ses_flag, items(nc), B, 0, c->stream, nc, nl, coff.p, state.p, flag.p);
    }
    hipLaunchKernelGGL(k_ses_line_cross, G, B, 0, c->stream, nl, coff.p, flag.p, xcnt.p);
    NGPB_HIP(hipGetLastError());
  }
  // rays = lines with crossings
  DBuf<uint8_t> keep(c->stream);
  keep.resize((size_t)nl);
  hipLaunchKernelGGL(k_ses_nonzero, G, B, 0, c->stream, nl, xcnt.p, keep.p);
  DBuf<int32_t> kept(c->stream);
  kept.resize((size_t)nl);
  const int64_t nr = compact_flags(c, keep.p, nl, kept.p);
  auto& R = c->rays[axis];
  R.n = nr;
  R.key.resize((size_t)2 * std::max<int64_t>(nr, 1));
  R.ptr.resize((size_t)nr + 1);
  if (nr > 0) {
    hipLaunchKernelGGL(k_ses_ray_keys, dim3((unsigned)blocks_for(nr, 256)), dim3(256), 0, c->stream, nr, kept.p, lu,
                       lv, xcnt.p, R.key.p, R.ptr.p);
    R.k = exclusive_scan(c, R.ptr.p, R.ptr.p, nr);
  } else {
    R.k = 0;
  }
  NGPB_HIP(hipMemcpyAsync(R.ptr.p + nr, &R.k, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  R.inters.resize((size_t)std::max<int64_t>(R.k, 1));
  R.normals.resize((size_t)3 * std::max<int64_t>(R.k, 1));
  if (nr > 0)
    hipLaunchKernelGGL(k_ses_emit, dim3((unsigned)blocks_for(nr, 128)), dim3(128), 0, c->stream, nr, kept.p, R.ptr.p,
                       coff.p, tk.p, kid.p, flag.p, R.key.p, S, axis, d1, d2, R.inters.p, R.normals.p);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace ngpb
