// Synthetic workload generator for the linearized Poisson-Boltzmann path.
//
// Produces, on the host, exactly the inputs that the reference's driver hands to
// its discretization stage:
//   * atoms (positions, radii, charges) drawn from seeded distributions;
//   * a 2:1-balanced linear octree over a cubic/box domain (leaves in Morton order,
//     corner->node map with p4est corner numbering bit0=x, bit1=y, bit2=z, hanging
//     corners with their 2 or 4 parents), i.e. what tmesh_3d exposes through
//     p(d,i), gt(i), is_hanging(i), num_parents(i), gparent(j,i)
//     (reference: src/pb_class.cpp:1838-1852, 2024-2036);
//   * sorted surface crossings along axis-aligned grid lines in the form of the
//     reference's ray cache (include/raytracer.h:40-54, src/raytracer.cpp:22-63):
//     for axis d a ray is keyed by the two other coordinates in increasing axis
//     order, and holds its crossings in increasing order with unit normals;
//   * the owning leaf of every atom with the half-open rule of
//     src/pb_class.cpp:3633-3676 (l - tol < x <= r - tol on every axis).
//
// The molecular surface is the union of atomic spheres (van der Waals surface);
// crossings are analytic. This file is workload generation, not the hot path.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------- random numbers
// splitmix64 stream; uniform doubles on [0,1) with 53 random bits; normals by the
// Box-Muller transform (both outputs used in turn).
struct Rng {
  uint64_t s;
  bool have_spare = false;
  double spare = 0.0;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double normal() {
    if (have_spare) { have_spare = false; return spare; }
    double u1, u2;
    do { u1 = uniform(); } while (u1 <= 0.0);
    u2 = uniform();
    const double rad = std::sqrt(-2.0 * std::log(u1));
    const double th = 2.0 * M_PI * u2;
    spare = rad * std::sin(th);
    have_spare = true;
    return rad * std::cos(th);
  }
};

// ---------------------------------------------------------------- atom hash grid
struct AtomGrid {
  double lo[3], cell;
  int n[3];
  std::vector<int> head, next;
  void build(int na, const double* xyz, double cell_size) {
    cell = cell_size;
    double hi[3];
    for (int d = 0; d < 3; ++d) { lo[d] = 1e300; hi[d] = -1e300; }
    for (int a = 0; a < na; ++a)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], xyz[3 * a + d]);
        hi[d] = std::max(hi[d], xyz[3 * a + d]);
      }
    if (na == 0) for (int d = 0; d < 3; ++d) { lo[d] = 0; hi[d] = 0; }
    for (int d = 0; d < 3; ++d) n[d] = std::max(1, (int)std::floor((hi[d] - lo[d]) / cell) + 1);
    head.assign((size_t)n[0] * n[1] * n[2], -1);
    next.assign(na, -1);
    for (int a = na - 1; a >= 0; --a) {
      size_t c = index_of(xyz + 3 * a);
      next[a] = head[c];
      head[c] = a;
    }
  }
  int clampi(int v, int d) const { return v < 0 ? 0 : (v >= n[d] ? n[d] - 1 : v); }
  size_t index_of(const double* p) const {
    int c[3];
    for (int d = 0; d < 3; ++d) c[d] = clampi((int)std::floor((p[d] - lo[d]) / cell), d);
    return ((size_t)c[2] * n[1] + c[1]) * n[0] + c[0];
  }
};

}  // namespace

extern "C" {

const char* ngpbw_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Atoms by rejection sampling inside an axis-aligned ellipsoid centred at c with
// semi-axes ax, ay, az and a minimum centre-to-centre separation min_sep.
// Draw order: candidate positions (3 uniforms each) until n are accepted, then n
// radii uniform in [rmin, rmax], then n charges N(0, qsigma) shifted by a common
// constant so that the net charge is qnet (no shift when qnet is NaN).
// Returns 0 on success, 1 if max_tries candidates did not yield n atoms.
int ngpbw_atoms_ellipsoid(uint64_t seed, int n, const double* c, double ax, double ay,
                          double az, double min_sep, double rmin, double rmax, double qsigma,
                          double qnet, long long max_tries, double* xyz, double* r,
                          double* q) {
  Rng rng(seed);
  // any cell >= min_sep finds every conflicting neighbour in the 3x3x3 block; the
  // floor keeps the grid small when min_sep is tiny or zero
  const double cell = std::max(min_sep, 1.0);
  const int gn[3] = {std::max(1, (int)std::ceil(2 * ax / cell)) + 1,
                     std::max(1, (int)std::ceil(2 * ay / cell)) + 1,
                     std::max(1, (int)std::ceil(2 * az / cell)) + 1};
  std::vector<int> grid((size_t)gn[0] * gn[1] * gn[2], -1);
  std::vector<int> nxt;
  nxt.reserve(n);
  const double lo[3] = {c[0] - ax, c[1] - ay, c[2] - az};
  const double ext[3] = {ax, ay, az};
  const double sep2 = min_sep * min_sep;
  int accepted = 0;
  long long tries = 0;
  while (accepted < n) {
    if (tries++ >= max_tries) {
      g_err = "atom rejection sampling exhausted max_tries (" + std::to_string(accepted) +
              " of " + std::to_string(n) + " placed)";
      return 1;
    }
    double p[3];
    for (int d = 0; d < 3; ++d) p[d] = lo[d] + 2.0 * ext[d] * rng.uniform();
    double e = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double t = (p[d] - c[d]) / ext[d];
      e += t * t;
    }
    if (e > 1.0) continue;
    int gc[3];
    for (int d = 0; d < 3; ++d)
      gc[d] = std::min(gn[d] - 1, std::max(0, (int)std::floor((p[d] - lo[d]) / cell)));
    bool ok = true;
    for (int dz = -1; dz <= 1 && ok; ++dz)
      for (int dy = -1; dy <= 1 && ok; ++dy)
        for (int dx = -1; dx <= 1 && ok; ++dx) {
          const int x = gc[0] + dx, y = gc[1] + dy, z = gc[2] + dz;
          if (x < 0 || y < 0 || z < 0 || x >= gn[0] || y >= gn[1] || z >= gn[2]) continue;
          for (int a = grid[((size_t)z * gn[1] + y) * gn[0] + x]; a >= 0; a = nxt[a]) {
            const double ddx = xyz[3 * a] - p[0], ddy = xyz[3 * a + 1] - p[1],
                         ddz = xyz[3 * a + 2] - p[2];
            if (ddx * ddx + ddy * ddy + ddz * ddz < sep2) { ok = false; break; }
          }
        }
    if (!ok) continue;
    for (int d = 0; d < 3; ++d) xyz[3 * accepted + d] = p[d];
    const size_t gi = ((size_t)gc[2] * gn[1] + gc[1]) * gn[0] + gc[0];
    nxt.push_back(grid[gi]);
    grid[gi] = accepted;
    ++accepted;
  }
  for (int a = 0; a < n; ++a) r[a] = rmin + (rmax - rmin) * rng.uniform();
  double sum = 0.0;
  for (int a = 0; a < n; ++a) {
    q[a] = qsigma * rng.normal();
    sum += q[a];
  }
  if (n > 0 && !std::isnan(qnet)) {
    const double shift = (qnet - sum) / n;
    for (int a = 0; a < n; ++a) q[a] += shift;
  }
  return 0;
}

}  // extern "C"

// ============================================================================
// Octree mesh
// ============================================================================
namespace {

struct Leaf {
  int32_t x, y, z;  // anchor in finest-grid units
  int32_t s;        // log2 of the edge length in finest-grid units
};

static inline uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}
static inline uint64_t morton(const Leaf& l) {
  return spread3((uint64_t)l.x) | (spread3((uint64_t)l.y) << 1) | (spread3((uint64_t)l.z) << 2);
}

}  // namespace

struct ngpbw_mesh {
  double origin[3];
  double hf;
  int64_t n[3];  // finest cells per axis
  int smax;
  std::vector<Leaf> leaves;
  std::vector<int32_t> leaf_corner;  // 8 per leaf: >=0 node id, <0 -> hanging id -1-k
  std::vector<double> node_xyz;      // 3 per node
  std::vector<int64_t> node_ijk;     // 3 per node (finest-grid vertex coordinates)
  std::vector<double> hang_xyz;      // 3 per hanging node
  std::vector<int64_t> hang_ijk;
  std::vector<int32_t> hang_np;      // 2 or 4
  std::vector<int32_t> hang_par;     // 4 per hanging node, -1 padded
  std::vector<int32_t> bnd_nodes;    // node ids on the domain boundary
  std::vector<int32_t> cell_leaf;    // finest cell -> leaf id
  double coord(int d, int64_t i) const { return origin[d] + (double)i * hf; }
  size_t cell_index(int64_t x, int64_t y, int64_t z) const {
    return ((size_t)z * n[1] + y) * n[0] + x;
  }
};

namespace {

// Signed distance lower/upper helper: f(x) = min_a (|x - c_a| - r_a) is the distance
// from x to the union of atomic balls for points outside it, negative inside.
struct SurfaceDistance {
  int na;
  const double* xyz;
  const double* r;
  AtomGrid grid;
  double rmax = 0.0;
  double bc[3] = {0, 0, 0};
  double brad = 0.0;  // bounding sphere radius of atom surfaces
  void init(int na_, const double* xyz_, const double* r_) {
    na = na_; xyz = xyz_; r = r_;
    for (int a = 0; a < na; ++a) rmax = std::max(rmax, r[a]);
    grid.build(na, xyz, 4.0);
    if (na) {
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (int a = 0; a < na; ++a)
        for (int d = 0; d < 3; ++d) {
          lo[d] = std::min(lo[d], xyz[3 * a + d]);
          hi[d] = std::max(hi[d], xyz[3 * a + d]);
        }
      for (int d = 0; d < 3; ++d) bc[d] = 0.5 * (lo[d] + hi[d]);
      for (int a = 0; a < na; ++a) {
        double dd = 0;
        for (int d = 0; d < 3; ++d) dd += (xyz[3 * a + d] - bc[d]) * (xyz[3 * a + d] - bc[d]);
        brad = std::max(brad, std::sqrt(dd) + r[a]);
      }
    }
  }
  // true iff f(p) < thr
  bool below(const double* p, double thr) const {
    if (na == 0) return false;
    double dd = 0;
    for (int d = 0; d < 3; ++d) dd += (p[d] - bc[d]) * (p[d] - bc[d]);
    if (std::sqrt(dd) - brad >= thr) return false;
    const double rr = std::max(thr + rmax, 0.0);
    int c0[3], c1[3];
    for (int d = 0; d < 3; ++d) {
      c0[d] = grid.clampi((int)std::floor((p[d] - rr - grid.lo[d]) / grid.cell), d);
      c1[d] = grid.clampi((int)std::floor((p[d] + rr - grid.lo[d]) / grid.cell), d);
    }
    for (int z = c0[2]; z <= c1[2]; ++z)
      for (int y = c0[1]; y <= c1[1]; ++y)
        for (int x = c0[0]; x <= c1[0]; ++x)
          for (int a = grid.head[((size_t)z * grid.n[1] + y) * grid.n[0] + x]; a >= 0;
               a = grid.next[a]) {
            const double dx = p[0] - xyz[3 * a], dy = p[1] - xyz[3 * a + 1],
                         dz = p[2] - xyz[3 * a + 2];
            if (std::sqrt(dx * dx + dy * dy + dz * dz) - r[a] < thr) return true;
          }
    return false;
  }
};

}  // namespace

extern "C" {

// Builds a 2:1 balanced (face, edge and corner) octree over the box
// [origin, origin + n*hf] with leaves of edge 2^s * hf, s in [0, smax].
// Root cells have s = smax and tile the box (n[d] must be multiples of 2^smax).
// A cell of level s >= 1 is split when the lower bound f(centre) - half_diagonal of
// the surface distance over the cell is below band[s-1]; band must be
// non-decreasing. The result is then 2:1 balanced by splitting offending leaves.
// Nodes are numbered lexicographically (x fastest, then y, then z) by their
// finest-grid coordinates; hanging nodes likewise in their own table.
ngpbw_mesh* ngpbw_mesh_octree(const double* origin, double hf, const int64_t* n, int smax,
                              const double* band, int na, const double* xyz, const double* r) {
  try {
    if (smax < 0 || smax > 10 || hf <= 0) { g_err = "bad smax/hf"; return nullptr; }
    const int64_t R = (int64_t)1 << smax;
    for (int d = 0; d < 3; ++d)
      if (n[d] <= 0 || n[d] % R != 0) { g_err = "n must be positive multiples of 2^smax"; return nullptr; }
    for (int s = 1; s < smax; ++s)
      if (band[s] < band[s - 1]) { g_err = "band must be non-decreasing"; return nullptr; }
    auto m = std::make_unique<ngpbw_mesh>();
    for (int d = 0; d < 3; ++d) { m->origin[d] = origin[d]; m->n[d] = n[d]; }
    m->hf = hf;
    m->smax = smax;
    SurfaceDistance sd;
    sd.init(na, xyz, r);

    // ---- top-down refinement
    std::vector<Leaf> work, leaves;
    for (int64_t z = 0; z < n[2]; z += R)
      for (int64_t y = 0; y < n[1]; y += R)
        for (int64_t x = 0; x < n[0]; x += R) work.push_back({(int32_t)x, (int32_t)y, (int32_t)z, smax});
    while (!work.empty()) {
      std::vector<char> split(work.size(), 0);
#pragma omp parallel for schedule(dynamic, 256)
      for (long long i = 0; i < (long long)work.size(); ++i) {
        const Leaf& l = work[i];
        if (l.s == 0) continue;
        const double sz = (double)((int64_t)1 << l.s) * hf;
        const double c[3] = {m->coord(0, l.x) + 0.5 * sz, m->coord(1, l.y) + 0.5 * sz,
                             m->coord(2, l.z) + 0.5 * sz};
        const double half_diag = 0.5 * std::sqrt(3.0) * sz;
        split[i] = sd.below(c, band[l.s - 1] + half_diag) ? 1 : 0;
      }
      std::vector<Leaf> nw;
      for (size_t i = 0; i < work.size(); ++i) {
        const Leaf& l = work[i];
        if (!split[i]) { leaves.push_back(l); continue; }
        const int32_t h = (int32_t)1 << (l.s - 1);
        for (int k = 0; k < 8; ++k)
          nw.push_back({l.x + ((k & 1) ? h : 0), l.y + ((k & 2) ? h : 0), l.z + ((k & 4) ? h : 0), l.s - 1});
      }
      work.swap(nw);
    }

    // ---- 2:1 balance on the finest-cell level map
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    std::vector<int8_t> lev(ncell);
    auto paint = [&](const Leaf& l) {
      const int32_t S = (int32_t)1 << l.s;
      for (int32_t z = l.z; z < l.z + S; ++z)
        for (int32_t y = l.y; y < l.y + S; ++y)
          for (int32_t x = l.x; x < l.x + S; ++x) lev[m->cell_index(x, y, z)] = (int8_t)l.s;
    };
    for (const Leaf& l : leaves) paint(l);
    for (int iter = 0; iter < 64; ++iter) {
      std::vector<char> split(leaves.size(), 0);
      long long nsplit = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : nsplit)
      for (long long i = 0; i < (long long)leaves.size(); ++i) {
        const Leaf& l = leaves[i];
        if (l.s < 2) continue;
        const int32_t S = (int32_t)1 << l.s;
        int mn = 127;
        for (int32_t z = l.z - 1; z <= l.z + S && mn >= l.s - 1; ++z) {
          if (z < 0 || z >= n[2]) continue;
          const bool zin = (z >= l.z && z < l.z + S);
          for (int32_t y = l.y - 1; y <= l.y + S; ++y) {
            if (y < 0 || y >= n[1]) continue;
            const bool yin = (y >= l.y && y < l.y + S);
            if (zin && yin) {
              if (l.x - 1 >= 0) mn = std::min(mn, (int)lev[m->cell_index(l.x - 1, y, z)]);
              if (l.x + S < n[0]) mn = std::min(mn, (int)lev[m->cell_index(l.x + S, y, z)]);
            } else {
              for (int32_t x = std::max(0, l.x - 1); x <= std::min<int64_t>(n[0] - 1, l.x + S); ++x)
                mn = std::min(mn, (int)lev[m->cell_index(x, y, z)]);
            }
          }
        }
        if (mn < l.s - 1) { split[i] = 1; ++nsplit; }
      }
      if (nsplit == 0) break;
      std::vector<Leaf> nl;
      nl.reserve(leaves.size() + 7 * nsplit);
      for (size_t i = 0; i < leaves.size(); ++i) {
        const Leaf& l = leaves[i];
        if (!split[i]) { nl.push_back(l); continue; }
        const int32_t h = (int32_t)1 << (l.s - 1);
        for (int k = 0; k < 8; ++k) {
          Leaf c{l.x + ((k & 1) ? h : 0), l.y + ((k & 2) ? h : 0), l.z + ((k & 4) ? h : 0), l.s - 1};
          nl.push_back(c);
          paint(c);
        }
      }
      leaves.swap(nl);
    }

    // ---- Morton order
    {
      std::vector<std::pair<uint64_t, Leaf>> keyed(leaves.size());
      for (size_t i = 0; i < leaves.size(); ++i) keyed[i] = {morton(leaves[i]), leaves[i]};
      std::sort(keyed.begin(), keyed.end(),
                [](const std::pair<uint64_t, Leaf>& a, const std::pair<uint64_t, Leaf>& b) { return a.first < b.first; });
      for (size_t i = 0; i < leaves.size(); ++i) leaves[i] = keyed[i].second;
    }
    const int64_t L = (int64_t)leaves.size();
    m->cell_leaf.assign(ncell, -1);
    for (int64_t i = 0; i < L; ++i) {
      const Leaf& l = leaves[i];
      const int32_t S = (int32_t)1 << l.s;
      for (int32_t z = l.z; z < l.z + S; ++z)
        for (int32_t y = l.y; y < l.y + S; ++y)
          for (int32_t x = l.x; x < l.x + S; ++x) m->cell_leaf[m->cell_index(x, y, z)] = (int32_t)i;
    }
    std::vector<int8_t>().swap(lev);

    // ---- vertices: 0 unused, 1 regular, 2 hanging
    const int64_t vn[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
    auto vidx = [&](int64_t x, int64_t y, int64_t z) { return ((size_t)z * vn[1] + y) * vn[0] + x; };
    std::vector<int32_t> vmap((size_t)vn[0] * vn[1] * vn[2], 0);
    for (const Leaf& l : leaves) {
      const int32_t S = (int32_t)1 << l.s;
      for (int k = 0; k < 8; ++k)
        vmap[vidx(l.x + ((k & 1) ? S : 0), l.y + ((k & 2) ? S : 0), l.z + ((k & 4) ? S : 0))] = 1;
    }
    // hanging: edge midpoints and face centres of leaves with s >= 1 that are
    // corners of some (smaller) leaf.
    std::unordered_map<size_t, std::array<size_t, 5>> hpar;  // vertex -> {np, parents...}
    static const int edge_nodes[12][2] = {{0, 1}, {1, 3}, {2, 3}, {0, 2}, {4, 5}, {5, 7},
                                          {6, 7}, {4, 6}, {0, 4}, {1, 5}, {3, 7}, {2, 6}};
    static const int face_nodes[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                         {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
    for (const Leaf& l : leaves) {
      if (l.s == 0) continue;
      const int64_t S = (int64_t)1 << l.s;
      int64_t cx[8], cy[8], cz[8];
      for (int k = 0; k < 8; ++k) {
        cx[k] = l.x + ((k & 1) ? S : 0);
        cy[k] = l.y + ((k & 2) ? S : 0);
        cz[k] = l.z + ((k & 4) ? S : 0);
      }
      for (int e = 0; e < 12; ++e) {
        const int a = edge_nodes[e][0], b = edge_nodes[e][1];
        const size_t v = vidx((cx[a] + cx[b]) / 2, (cy[a] + cy[b]) / 2, (cz[a] + cz[b]) / 2);
        if (vmap[v] != 0) {
          vmap[v] = 2;
          if (!hpar.count(v)) hpar[v] = {2, vidx(cx[a], cy[a], cz[a]), vidx(cx[b], cy[b], cz[b]), 0, 0};
        }
      }
      for (int f = 0; f < 6; ++f) {
        const int* fn = face_nodes[f];
        const size_t v = vidx((cx[fn[0]] + cx[fn[3]]) / 2, (cy[fn[0]] + cy[fn[3]]) / 2, (cz[fn[0]] + cz[fn[3]]) / 2);
        if (vmap[v] != 0) {
          vmap[v] = 2;
          if (!hpar.count(v))
            hpar[v] = {4, vidx(cx[fn[0]], cy[fn[0]], cz[fn[0]]), vidx(cx[fn[1]], cy[fn[1]], cz[fn[1]]),
                       vidx(cx[fn[2]], cy[fn[2]], cz[fn[2]]), vidx(cx[fn[3]], cy[fn[3]], cz[fn[3]])};
        }
      }
    }
    // numbering
    int64_t nreg = 0, nhang = 0;
    for (int64_t z = 0; z < vn[2]; ++z)
      for (int64_t y = 0; y < vn[1]; ++y)
        for (int64_t x = 0; x < vn[0]; ++x) {
          int32_t& v = vmap[vidx(x, y, z)];
          if (v == 1) {
            m->node_ijk.insert(m->node_ijk.end(), {x, y, z});
            v = (int32_t)(nreg++);
          } else if (v == 2) {
            m->hang_ijk.insert(m->hang_ijk.end(), {x, y, z});
            v = (int32_t)(-1 - nhang++);
          } else {
            v = INT32_MIN;
          }
        }
    if (nreg >= INT32_MAX / 2) { g_err = "too many nodes for int32 ids"; return nullptr; }
    m->node_xyz.resize(3 * nreg);
    for (int64_t i = 0; i < nreg; ++i)
      for (int d = 0; d < 3; ++d) m->node_xyz[3 * i + d] = m->coord(d, m->node_ijk[3 * i + d]);
    m->hang_xyz.resize(3 * nhang);
    m->hang_np.resize(nhang);
    m->hang_par.assign(4 * nhang, -1);
    for (int64_t k = 0; k < nhang; ++k) {
      for (int d = 0; d < 3; ++d) m->hang_xyz[3 * k + d] = m->coord(d, m->hang_ijk[3 * k + d]);
      const size_t v = vidx(m->hang_ijk[3 * k], m->hang_ijk[3 * k + 1], m->hang_ijk[3 * k + 2]);
      const auto& hp = hpar.at(v);
      m->hang_np[k] = (int32_t)hp[0];
      for (int j = 0; j < (int)hp[0]; ++j) {
        const int32_t pid = vmap[hp[1 + j]];
        if (pid < 0) { g_err = "hanging node with a hanging parent (mesh not 2:1 balanced)"; return nullptr; }
        m->hang_par[4 * k + j] = pid;
      }
    }
    m->leaf_corner.resize(8 * L);
    for (int64_t i = 0; i < L; ++i) {
      const Leaf& l = leaves[i];
      const int64_t S = (int64_t)1 << l.s;
      for (int k = 0; k < 8; ++k)
        m->leaf_corner[8 * i + k] = vmap[vidx(l.x + ((k & 1) ? S : 0), l.y + ((k & 2) ? S : 0), l.z + ((k & 4) ? S : 0))];
    }
    for (int64_t i = 0; i < nreg; ++i) {
      bool b = false;
      for (int d = 0; d < 3; ++d) b = b || m->node_ijk[3 * i + d] == 0 || m->node_ijk[3 * i + d] == n[d];
      if (b) m->bnd_nodes.push_back((int32_t)i);
    }
    m->leaves.swap(leaves);
    return m.release();
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory in mesh generation";
    return nullptr;
  }
}

void ngpbw_mesh_free(ngpbw_mesh* m) { delete m; }

// sizes[0..4] = nodes, hanging nodes, leaves, boundary nodes, leaves per level... (5 values)
void ngpbw_mesh_sizes(const ngpbw_mesh* m, int64_t* sizes) {
  sizes[0] = (int64_t)m->node_xyz.size() / 3;
  sizes[1] = (int64_t)m->hang_np.size();
  sizes[2] = (int64_t)m->leaves.size();
  sizes[3] = (int64_t)m->bnd_nodes.size();
}

void ngpbw_mesh_copy(const ngpbw_mesh* m, double* node_xyz, double* hang_xyz, int32_t* hang_np,
                     int32_t* hang_par, int32_t* leaf_corner, int32_t* bnd_nodes, int32_t* leaf_level) {
  if (node_xyz) std::memcpy(node_xyz, m->node_xyz.data(), m->node_xyz.size() * sizeof(double));
  if (hang_xyz) std::memcpy(hang_xyz, m->hang_xyz.data(), m->hang_xyz.size() * sizeof(double));
  if (hang_np) std::memcpy(hang_np, m->hang_np.data(), m->hang_np.size() * sizeof(int32_t));
  if (hang_par) std::memcpy(hang_par, m->hang_par.data(), m->hang_par.size() * sizeof(int32_t));
  if (leaf_corner) std::memcpy(leaf_corner, m->leaf_corner.data(), m->leaf_corner.size() * sizeof(int32_t));
  if (bnd_nodes) std::memcpy(bnd_nodes, m->bnd_nodes.data(), m->bnd_nodes.size() * sizeof(int32_t));
  if (leaf_level)
    for (size_t i = 0; i < m->leaves.size(); ++i) leaf_level[i] = m->leaves[i].s;
}

// Owner leaf of every atom: the leaf whose corner-0/corner-7 box satisfies
// l - tol < x <= r - tol on all three axes (src/pb_class.cpp:3668-3673), -1 if none.
int ngpbw_locate_atoms(const ngpbw_mesh* m, int na, const double* xyz, double tol, int32_t* owner) {
  const int64_t L = (int64_t)m->leaves.size();
  for (int a = 0; a < na; ++a) {
    owner[a] = -1;
    int64_t k[3];
    bool inside = true;
    for (int d = 0; d < 3; ++d) {
      const double t = std::ceil((xyz[3 * a + d] - m->origin[d] + tol) / m->hf) - 1.0;
      k[d] = (int64_t)t;
      if (t < 0 || t >= (double)m->n[d]) { inside = false; }
    }
    // candidates: the estimated finest cell and its neighbours (rounding guard)
    for (int dz = -1; dz <= 1 && owner[a] < 0; ++dz)
      for (int dy = -1; dy <= 1 && owner[a] < 0; ++dy)
        for (int dx = -1; dx <= 1 && owner[a] < 0; ++dx) {
          const int64_t c[3] = {k[0] + dx, k[1] + dy, k[2] + dz};
          bool ok = true;
          for (int d = 0; d < 3; ++d) ok = ok && c[d] >= 0 && c[d] < m->n[d];
          if (!ok) continue;
          const int32_t li = m->cell_leaf[m->cell_index(c[0], c[1], c[2])];
          const int32_t* lc = &m->leaf_corner[8 * (size_t)li];
          auto cxyz = [&](int32_t id, int d) {
            return id >= 0 ? m->node_xyz[3 * (size_t)id + d] : m->hang_xyz[3 * (size_t)(-1 - id) + d];
          };
          bool in = true;
          for (int d = 0; d < 3; ++d) {
            const double lo = cxyz(lc[0], d), hi = cxyz(lc[7], d), x = xyz[3 * a + d];
            in = in && (x > lo - tol) && (x <= hi - tol);
          }
          if (in) owner[a] = li;
        }
    (void)inside;
    (void)L;
  }
  return 0;
}

}  // extern "C"

// ============================================================================
// Axis rays against a union of spheres
// ============================================================================
struct ngpbw_rays {
  std::vector<double> key;      // 2 per ray
  std::vector<int64_t> ptr;     // R+1
  std::vector<double> inters;   // K
  std::vector<double> normals;  // 3K
};

extern "C" {

// Casts every grid line of axis `axis` that passes through a mesh vertex (regular
// or hanging) against the union of balls (xyz, r) and keeps those with at least
// one crossing. Crossings are the end points of the merged covering intervals
// along the line, in increasing order; the normal at a crossing is the outward unit
// normal (p - c_a)/r_a of the sphere that owns that end point. A line that only
// touches a sphere (distance == radius) does not cross it.
ngpbw_rays* ngpbw_cast_rays(const ngpbw_mesh* m, int axis, int na, const double* xyz,
                            const double* r) {
  try {
    const int d1 = axis == 0 ? 1 : 0;
    const int d2 = axis == 2 ? 1 : 2;
    // distinct lines through vertices, sorted by (i1, i2)
    std::vector<uint64_t> lines;
    lines.reserve(m->node_ijk.size() / 3 + m->hang_ijk.size() / 3);
    auto add = [&](const std::vector<int64_t>& ijk) {
      for (size_t i = 0; i < ijk.size() / 3; ++i)
        lines.push_back(((uint64_t)ijk[3 * i + d1] << 32) | (uint64_t)ijk[3 * i + d2]);
    };
    add(m->node_ijk);
    add(m->hang_ijk);
    std::sort(lines.begin(), lines.end());
    lines.erase(std::unique(lines.begin(), lines.end()), lines.end());
    double rmax = 0;
    for (int a = 0; a < na; ++a) rmax = std::max(rmax, r[a]);
    // 2D bins of atoms in the (d1, d2) plane
    const double cell = std::max(2.0 * rmax, 1.0);
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (int a = 0; a < na; ++a) {
      lo[0] = std::min(lo[0], xyz[3 * a + d1]); hi[0] = std::max(hi[0], xyz[3 * a + d1]);
      lo[1] = std::min(lo[1], xyz[3 * a + d2]); hi[1] = std::max(hi[1], xyz[3 * a + d2]);
    }
    if (na == 0) { lo[0] = lo[1] = hi[0] = hi[1] = 0; }
    const int gn0 = (int)std::floor((hi[0] - lo[0]) / cell) + 1, gn1 = (int)std::floor((hi[1] - lo[1]) / cell) + 1;
    std::vector<std::vector<int>> bins((size_t)gn0 * gn1);
    for (int a = 0; a < na; ++a) {
      const int b0 = (int)std::floor((xyz[3 * a + d1] - lo[0]) / cell);
      const int b1 = (int)std::floor((xyz[3 * a + d2] - lo[1]) / cell);
      bins[(size_t)b1 * gn0 + b0].push_back(a);
    }
    const int64_t nl = (int64_t)lines.size();
    std::vector<std::vector<double>> lx(nl), ln(nl);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t li = 0; li < nl; ++li) {
      const int64_t i1 = (int64_t)(lines[li] >> 32), i2 = (int64_t)(lines[li] & 0xffffffffULL);
      const double u = m->coord(d1, i1), v = m->coord(d2, i2);
      const int b0 = (int)std::floor((u - lo[0]) / cell), b1 = (int)std::floor((v - lo[1]) / cell);
      struct Iv { double a, b; int ia, ib; };
      std::vector<Iv> iv;
      for (int y = b1 - 1; y <= b1 + 1; ++y)
        for (int x = b0 - 1; x <= b0 + 1; ++x) {
          if (x < 0 || y < 0 || x >= gn0 || y >= gn1) continue;
          for (int a : bins[(size_t)y * gn0 + x]) {
            const double du = u - xyz[3 * a + d1], dv = v - xyz[3 * a + d2];
            const double q = r[a] * r[a] - (du * du + dv * dv);
            if (q <= 0) continue;
            const double s = std::sqrt(q);
            const double c = xyz[3 * a + axis];
            iv.push_back({c - s, c + s, a, a});
          }
        }
      if (iv.empty()) continue;
      std::sort(iv.begin(), iv.end(), [](const Iv& p, const Iv& q) {
        return p.a < q.a || (p.a == q.a && (p.b < q.b || (p.b == q.b && p.ia < q.ia)));
      });
      std::vector<Iv> mg;
      for (const Iv& t : iv) {
        if (!mg.empty() && t.a <= mg.back().b) {
          if (t.b > mg.back().b) { mg.back().b = t.b; mg.back().ib = t.ib; }
        } else {
          mg.push_back(t);
        }
      }
      for (const Iv& t : mg) {
        for (int e = 0; e < 2; ++e) {
          const double tc = e == 0 ? t.a : t.b;
          const int a = e == 0 ? t.ia : t.ib;
          double p[3];
          p[axis] = tc; p[d1] = u; p[d2] = v;
          lx[li].push_back(tc);
          for (int d = 0; d < 3; ++d) ln[li].push_back((p[d] - xyz[3 * a + d]) / r[a]);
        }
      }
    }
    auto rays = std::make_unique<ngpbw_rays>();
    rays->ptr.push_back(0);
    for (int64_t li = 0; li < nl; ++li) {
      if (lx[li].empty()) continue;
      const int64_t i1 = (int64_t)(lines[li] >> 32), i2 = (int64_t)(lines[li] & 0xffffffffULL);
      rays->key.push_back(m->coord(d1, i1));
      rays->key.push_back(m->coord(d2, i2));
      rays->inters.insert(rays->inters.end(), lx[li].begin(), lx[li].end());
      rays->normals.insert(rays->normals.end(), ln[li].begin(), ln[li].end());
      rays->ptr.push_back((int64_t)rays->inters.size());
    }
    return rays.release();
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory in ray casting";
    return nullptr;
  }
}

void ngpbw_rays_sizes(const ngpbw_rays* r, int64_t* sizes) {
  sizes[0] = (int64_t)r->key.size() / 2;
  sizes[1] = (int64_t)r->inters.size();
}
void ngpbw_rays_copy(const ngpbw_rays* r, double* key, int64_t* ptr, double* inters, double* normals) {
  std::memcpy(key, r->key.data(), r->key.size() * sizeof(double));
  std::memcpy(ptr, r->ptr.data(), r->ptr.size() * sizeof(int64_t));
  std::memcpy(inters, r->inters.data(), r->inters.size() * sizeof(double));
  std::memcpy(normals, r->normals.data(), r->normals.size() * sizeof(double));
}
void ngpbw_rays_free(ngpbw_rays* r) { delete r; }

}  // extern "C"
