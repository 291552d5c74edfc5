// ORACLE — test infrastructure only.
//
// CPU restatement of NextGenPB's linearized Poisson-Boltzmann hot path, used as the
// checker for the HIP implementation. Only the test suite, the smoke check and the
// CPU-baseline leg of the benchmark load this library; the product path never does.
//
// Every function follows the reference's sequential semantics (one MPI rank):
//   orc_constants      include/pb_class.h:47-54, src/pb_class.cpp:1800-1806, 2803-2814
//   orc_markers        src/pb_class.cpp:768-805 (is_in_ns_surf), 1792-1941 (create_markers)
//   orc_fractions      src/pb_class.cpp:2422-2532 (cube_fraction_intersection)
//   orc_locate         src/pb_class.cpp:3633-3729 (controlla_coordinate, cerca_atomo, search_points)
//   orc_rhs            src/pb_class.cpp:1943-2050 (create_density_map) + bim3a_rhs (bimpp)
//   orc_assemble       src/pb_class.cpp:2052-2159 + bimpp bim3a_laplacian_frac /
//                      bim3a_reaction_frac / bim3a_dirichlet_bc (conventions below)
//   orc_pcg            src/pb_class.cpp:2243-2329 + LIS 2.1.6 CG with Jacobi (-i cg -p jacobi)
//   orc_krylov         the other option strings the reference passes to LIS (default
//                      -i cgs -p ssor, src/pb_class.cpp:970): CG / CGS / BiCGSTAB with
//                      no, Jacobi or SSOR preconditioning (published algorithms, below)
//   orc_energy         src/pb_class.cpp:2534-2788 (helpers), 2790-3063 (energy)
//   orc_phi_on_atoms   src/pb_class.cpp:2333-2368 (write_potential_on_atoms_fast)
//   orc_pot_field      src/pb_class.cpp:3731-5060 (pot_field_fast / pot_field)
//   orc_refine_surface src/pb_class.cpp:1289-1296 (init_tmesh), 1467-1745 (refine_surface)
//
// Pinning (how each is checked without the reference's libraries, which are absent):
//   constants  closed forms; markers: brute-force distance test on sphere unions +
//   exhaustive rules; fractions: closed form t = c +- sqrt(r^2-d^2); locate: brute
//   force; rhs: weights sum to one, sum rhs = 4 pi sum q; assembly: uniform 7-point
//   closed form, symmetry, zero interior row sums at kappa=0, exact reproduction of
//   linear fields; pcg and krylov: dense direct solves of small systems, SSOR sweeps
//   against M z = r with M a product of sparse factors; energy: Born-sphere
//   analytic energy and the Gauss flux-charge check; phi on atoms: exact
//   reproduction of linear fields (trilinear interpolation); pot_field: every field
//   column is minus the gradient of its potential column with respect to the atom
//   position (central differences), half the charge-weighted sum of phi_p + phi_i
//   is the energy above, and the Coulomb columns match direct pair sums.
// The bimpp conventions (hanging-node algebra, reaction lumping, Dirichlet
// elimination) cannot be pinned against bimpp itself, which is not available:
// "parity unpinned" against bimpp; pinned by the invariants above.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../cudanextgenpb_amd/csrc/mc_table.inc"

namespace {

constexpr double kE0 = 8.85418781762e-12;  // F/m
constexpr double kKb = 1.380649e-23;       // J/K
constexpr double kE = 1.602176634e-19;     // C
constexpr double kNav = 6.022e23;          // 1/mol
constexpr double kAngs = 1e-10;            // m
constexpr double kPi = 3.14159265358979323846;

// corner pairs of the 12 cube edges (include/pb_class.h:193-207) and their axes (:190-191)
constexpr int kEdgeNodes[24] = {0, 1, 1, 3, 2, 3, 0, 2, 4, 5, 5, 7, 6, 7, 4, 6, 0, 4, 1, 5, 3, 7, 2, 6};
constexpr int kEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};

}  // namespace

extern "C" {

struct orc_mesh {
  int64_t n_nodes, n_hang, n_leaves;
  const double* node_xyz;
  const double* hang_xyz;
  const int32_t* hang_np;
  const int32_t* hang_par;
  const int32_t* leaf_corner;
};

struct orc_rays {
  int64_t n_rays;
  const double* key;
  const int64_t* ptr;
  const double* inters;
  const double* normals;
};

// out[0..5] = eps_in_hat, eps_out_hat, k2, k, constant_pol, constant_react
void orc_constants(double e_in, double e_out, double T, double I, double* out) {
  const double eps_in = 4.0 * kPi * kE0 * e_in * kKb * T * kAngs / (kE * kE);
  const double eps_out = 4.0 * kPi * kE0 * e_out * kKb * T * kAngs / (kE * kE);
  const double C0 = 1.0e3 * kNav * I;
  const double k2 = 2.0 * C0 * kAngs * kAngs * kE * kE / (kE0 * e_out * kKb * T);
  const double inv_4pi = 1.0 / (4.0 * kPi);
  out[0] = eps_in;
  out[1] = eps_out;
  out[2] = k2;
  out[3] = std::sqrt(k2);
  out[4] = (1.0 / eps_out - 1.0 / eps_in) * inv_4pi;
  out[5] = (1.0 / eps_out) * inv_4pi;
}

}  // extern "C"

namespace {

// ray cache restated as one ordered map per axis keyed by the two transverse
// coordinates (src/raytracer.cpp:22-63; exact key comparison)
struct Ray {
  const double* inters;
  const double* normals;
  int64_t n;
};
using RayMap = std::map<std::array<double, 2>, Ray>;

std::array<RayMap, 3> build_maps(const orc_rays* rays) {
  std::array<RayMap, 3> m;
  for (int d = 0; d < 3; ++d) {
    const orc_rays& r = rays[d];
    for (int64_t i = 0; i < r.n_rays; ++i) {
      const int64_t b = r.ptr[i], e = r.ptr[i + 1];
      m[d][{r.key[2 * i], r.key[2 * i + 1]}] = Ray{r.inters + b, r.normals + 3 * b, e - b};
    }
  }
  return m;
}

struct MeshView {
  const orc_mesh* m;
  bool hanging(int64_t l, int i) const { return m->leaf_corner[8 * l + i] < 0; }
  int32_t gt(int64_t l, int i) const { return m->leaf_corner[8 * l + i]; }
  int32_t hid(int64_t l, int i) const { return -1 - m->leaf_corner[8 * l + i]; }
  double p(int64_t l, int d, int i) const {
    const int32_t c = m->leaf_corner[8 * l + i];
    return c >= 0 ? m->node_xyz[3 * (int64_t)c + d] : m->hang_xyz[3 * (int64_t)(-1 - c) + d];
  }
  int num_parents(int64_t l, int i) const { return m->hang_np[hid(l, i)]; }
  int32_t gparent(int64_t l, int j, int i) const { return m->hang_par[4 * (int64_t)hid(l, i) + j]; }
};

// is_in_ns_surf (src/pb_class.cpp:768-805): parity of #{crossings <= x3} along the
// axis-`dir` ray through the point; no ray or no crossing -> outside.
int is_in_surf(const std::array<RayMap, 3>& maps, double x, double y, double z, int dir) {
  double x1 = x, x2 = y, x3 = z;
  if (dir == 0) { x1 = y; x2 = z; x3 = x; }
  else if (dir == 1) { x1 = x; x2 = z; x3 = y; }
  auto it = maps[dir].find({x1, x2});
  if (it == maps[dir].end()) return 0;
  const Ray& ct = it->second;
  int64_t i = 0;
  if (ct.n == 0 || x3 < ct.inters[0]) return 0;
  while (i < ct.n && x3 >= ct.inters[i]) ++i;
  return (int)(i % 2);
}

// normal_intersection (src/pb_class.cpp:2534-2568): last crossing t with x1 <= t <= x2
// on the axis ray through corner i1 of edge `edge`; frac defaults to 0.5 and the
// normal is only written when a crossing is found. Returns whether one was found.
bool normal_intersection(const MeshView& mv, const std::array<RayMap, 3>& maps, int64_t l,
                         int edge, double* norm, double* frac) {
  const int dir = kEdgeAxis[edge];
  const int i1 = kEdgeNodes[2 * edge], i2 = kEdgeNodes[2 * edge + 1];
  const double x1 = mv.p(l, dir, i1), x2 = mv.p(l, dir, i2);
  std::array<double, 2> ray;
  int k = 0;
  for (int d = 0; d < 3; ++d)
    if (d != dir) ray[k++] = mv.p(l, d, i1);
  *frac = 0.5;
  auto it = maps[dir].find(ray);
  if (it == maps[dir].end()) return false;
  bool found = false;
  for (int64_t ii = 0; ii < it->second.n; ++ii) {
    const double t = it->second.inters[ii];
    if (t >= x1 && t <= x2) {
      norm[0] = it->second.normals[3 * ii + 0];
      norm[1] = it->second.normals[3 * ii + 1];
      norm[2] = it->second.normals[3 * ii + 2];
      *frac = (t - x1) / (x2 - x1);
      found = true;
    }
  }
  return found;
}

double wha(double eps1, double eps2, double frac) { return 1.0 / (frac / eps1 + (1 - frac) / eps2); }

double phi0(double eps1, double eps2, double phi1, double phi2, double frac) {
  return phi1 + frac * eps2 * (phi2 - phi1) / (eps2 * frac + eps1 * (1 - frac));
}

double area_triangle(const double t[3][3]) {
  double ab[3], ac[3];
  for (int i = 0; i < 3; ++i) {
    ab[i] = t[0][i] - t[1][i];
    ac[i] = t[0][i] - t[2][i];
  }
  return 0.5 * std::hypot(ab[1] * ac[2] - ab[2] * ac[1], ab[0] * ac[2] - ab[2] * ac[0],
                          ab[1] * ac[0] - ab[0] * ac[1]);
}

int mc_digit(char c) { return c <= '9' ? c - '0' : c - 'a' + 10; }

}  // namespace

extern "C" {

// create_markers (src/pb_class.cpp:1792-1941), one rank, one cycle, no Stern layer.
// leaf_marker: 0 = inside (0.0), 1 = border (0.5), 2 = outside (1.0).
int orc_markers(const orc_mesh* m, const orc_rays* rays, double eps_in, double eps_out,
                double k2, double* eps_node, double* reac_node, uint8_t* leaf_marker,
                int32_t* border, int64_t* n_border) {
  const auto maps = build_maps(rays);
  MeshView mv{m};
  for (int64_t i = 0; i < m->n_nodes; ++i) {
    eps_node[i] = eps_out;
    reac_node[i] = eps_out * k2;
  }
  int64_t nb = 0;
  for (int64_t l = 0; l < m->n_leaves; ++l) {
    int num_int = 0, num_hang = 0;
    for (int ii = 0; ii < 8; ++ii) {
      const double x = mv.p(l, 0, ii), y = mv.p(l, 1, ii), z = mv.p(l, 2, ii);
      if (!mv.hanging(l, ii)) {
        if (is_in_surf(maps, x, y, z, 2) > 0) {
          ++num_int;
          eps_node[mv.gt(l, ii)] = eps_in;
          reac_node[mv.gt(l, ii)] = 0.0;
        }
      } else {
        ++num_hang;
      }
    }
    uint8_t mk = 0;
    if (num_int == 0) mk = 2;
    else if (num_int < 8 - num_hang) { mk = 1; border[nb++] = (int32_t)l; }
    leaf_marker[l] = mk;
  }
  *n_border = nb;
  return 0;
}

// cube_fraction_intersection (src/pb_class.cpp:2422-2532) for every leaf: 0.5 unless
// the leaf is a border leaf and a crossing lies on the edge. found[] reports whether
// a crossing was found (the reference does not keep this; it drives the normal).
int orc_fractions(const orc_mesh* m, const orc_rays* rays, const uint8_t* leaf_marker,
                  double* frac, double* normal, uint8_t* found) {
  const auto maps = build_maps(rays);
  MeshView mv{m};
  for (int64_t l = 0; l < m->n_leaves; ++l) {
    for (int e = 0; e < 12; ++e) {
      frac[12 * l + e] = 0.5;
      if (found) found[12 * l + e] = 0;
      if (normal) for (int d = 0; d < 3; ++d) normal[36 * l + 3 * e + d] = 0.0;
      if (leaf_marker[l] != 1) continue;
      double nrm[3] = {0, 0, 0}, f = 0.5;
      const bool fd = normal_intersection(mv, maps, l, e, nrm, &f);
      frac[12 * l + e] = f;
      if (found) found[12 * l + e] = fd ? 1 : 0;
      if (normal && fd) for (int d = 0; d < 3; ++d) normal[36 * l + 3 * e + d] = nrm[d];
    }
  }
  return 0;
}

// search_points / cerca_atomo / controlla_coordinate (src/pb_class.cpp:3633-3729):
// the first leaf (in leaf order) with l - tol < x <= r - tol on every axis, where l
// and r are the coordinates of corners 0 and 7; -1 when no leaf qualifies.
int orc_locate(const orc_mesh* m, int64_t na, const double* xyz, double tol, int32_t* owner) {
  MeshView mv{m};
  for (int64_t a = 0; a < na; ++a) {
    owner[a] = -1;
    for (int64_t l = 0; l < m->n_leaves; ++l) {
      bool in = true;
      for (int d = 0; d < 3 && in; ++d) {
        const double lo = mv.p(l, d, 0), hi = mv.p(l, d, 7), x = xyz[3 * a + d];
        in = (x > lo - tol) && (x <= hi - tol);
      }
      if (in) { owner[a] = (int32_t)l; break; }
    }
  }
  return 0;
}

// create_density_map (src/pb_class.cpp:1943-2050) followed by bim3a_rhs(const_ones,
// rho) (src/pb_class.cpp:2085). bim3a_rhs assembles sum over leaves of
// f_leaf * g(node) * V/8 per corner, a hanging corner contributing V/(8 np) to each
// parent with the parent's value; vol_patch is the same sum with g = 1.
// write_potential_on_atoms_fast (src/pb_class.cpp:2333-2368): trilinear interpolation
// of phi inside each atom's owner leaf with the charge-projection weights; a hanging
// corner takes the mean of its parents (sum of phi/np). Atoms outside the domain
// (owner -1) get NaN.
int orc_phi_on_atoms(const orc_mesh* m, int64_t na, const double* xyz, const int32_t* owner,
                     const double* phi, double* out) {
  MeshView mv{m};
  for (int64_t a = 0; a < na; ++a) {
    const int64_t l = owner[a];
    if (l < 0) {
      out[a] = std::nan("");
      continue;
    }
    const double volume = (mv.p(l, 0, 7) - mv.p(l, 0, 0)) * (mv.p(l, 1, 7) - mv.p(l, 1, 0)) *
                          (mv.p(l, 2, 7) - mv.p(l, 2, 0));
    double phi_on_atom = 0.0;
    for (int ii = 0; ii < 8; ++ii) {
      const double weight = std::abs((xyz[3 * a + 0] - mv.p(l, 0, 7 - ii)) *
                                     (xyz[3 * a + 1] - mv.p(l, 1, 7 - ii)) *
                                     (xyz[3 * a + 2] - mv.p(l, 2, 7 - ii))) / volume;
      if (!mv.hanging(l, ii)) {
        phi_on_atom += phi[mv.gt(l, ii)] * weight;
      } else {
        double ph = 0.0;
        for (int jj = 0; jj < mv.num_parents(l, ii); ++jj) ph += phi[mv.gparent(l, jj, ii)] / mv.num_parents(l, ii);
        phi_on_atom += ph * weight;
      }
    }
    out[a] = phi_on_atom;
  }
  return 0;
}

int orc_rhs(const orc_mesh* m, int64_t na, const double* xyz, const double* q,
            const int32_t* owner, double* rhs) {
  MeshView mv{m};
  const int64_t N = m->n_nodes;
  std::vector<double> vol_patch(N, 0.0), rho(N, 0.0);
  auto vol8 = [&](int64_t l) {
    return (mv.p(l, 0, 7) - mv.p(l, 0, 0)) * (mv.p(l, 1, 7) - mv.p(l, 1, 0)) *
           (mv.p(l, 2, 7) - mv.p(l, 2, 0)) / 8.0;
  };
  for (int64_t l = 0; l < m->n_leaves; ++l) {
    const double v8 = vol8(l);
    for (int ii = 0; ii < 8; ++ii) {
      if (!mv.hanging(l, ii)) vol_patch[mv.gt(l, ii)] += v8;
      else
        for (int jj = 0; jj < mv.num_parents(l, ii); ++jj)
          vol_patch[mv.gparent(l, jj, ii)] += v8 / mv.num_parents(l, ii);
    }
  }
  // lookup_table is a std::map<int, quadrant>: atoms in increasing index order
  for (int64_t a = 0; a < na; ++a) {
    const int64_t l = owner[a];
    if (l < 0) continue;
    const double volume = (mv.p(l, 0, 7) - mv.p(l, 0, 0)) * (mv.p(l, 1, 7) - mv.p(l, 1, 0)) *
                          (mv.p(l, 2, 7) - mv.p(l, 2, 0));
    for (int ii = 0; ii < 8; ++ii) {
      const double weight = std::abs((xyz[3 * a + 0] - mv.p(l, 0, 7 - ii)) *
                                     (xyz[3 * a + 1] - mv.p(l, 1, 7 - ii)) *
                                     (xyz[3 * a + 2] - mv.p(l, 2, 7 - ii))) / volume;
      if (!mv.hanging(l, ii)) {
        const int32_t g = mv.gt(l, ii);
        rho[g] += q[a] * 4.0 * kPi * weight / vol_patch[g];
      } else {
        for (int jj = 0; jj < mv.num_parents(l, ii); ++jj) {
          const int32_t g = mv.gparent(l, jj, ii);
          const double denom = mv.num_parents(l, ii) * vol_patch[g];
          rho[g] += q[a] * 4.0 * kPi * weight / denom;
        }
      }
    }
  }
  for (int64_t i = 0; i < N; ++i) rhs[i] = 0.0;
  for (int64_t l = 0; l < m->n_leaves; ++l) {
    const double v8 = vol8(l);
    for (int ii = 0; ii < 8; ++ii) {
      if (!mv.hanging(l, ii)) { const int32_t g = mv.gt(l, ii); rhs[g] += v8 * rho[g]; }
      else
        for (int jj = 0; jj < mv.num_parents(l, ii); ++jj) {
          const int32_t g = mv.gparent(l, jj, ii);
          rhs[g] += v8 / mv.num_parents(l, ii) * rho[g];
        }
    }
  }
  return 0;
}

// ---------------------------------------------------------------- assembly
struct orc_sys {
  int64_t n = 0;
  std::vector<int32_t> rowptr, col;
  std::vector<double> val, rhs;
};

// assemple_system_matrix (src/pb_class.cpp:2052-2159) with map-based rows as in
// bimpp's sparse matrix. Conventions (bimpp's source is absent; these are this
// project's statement of them):
//  * laplacian_frac: per leaf and per edge (i1, i2) along axis a, conductance
//    c = wha(eps(i1), eps(i2), f) * h_b * h_c / (4 h_a), with f the edge fraction
//    measured from i1 and eps at a hanging corner the mean of its parents
//    (sum of eps[parent]/np in parent order, as src/pb_class.cpp:2635-2647). The
//    leaf matrix c*[[1,-1],[-1,1]] on (i1, i2) is scattered as P^T A P, a hanging
//    corner's row/column spread to its parents with weight 1/np. Within an edge
//    the entries are visited (i1,i1), (i1,i2), (i2,i1), (i2,i2), each over parent
//    pairs in parent order.
//  * reaction_frac: lumped mass reac(node) * V/8 on the diagonal per corner; a
//    hanging corner adds reac(parent) * V/(8 np) to each parent's diagonal.
//  * dirichlet: symmetric elimination. A Dirichlet row keeps only its assembled
//    diagonal d with rhs = d * g; every other row moves a_ij * g_j to its rhs and
//    drops column j. phi is the same as with any other elimination.
orc_sys* orc_assemble(const orc_mesh* m, const double* eps_node, const double* reac_node,
                      const double* frac, int64_t nb, const int32_t* bnd_nodes,
                      const double* bnd_vals, const double* rhs_charge) {
  MeshView mv{m};
  const int64_t N = m->n_nodes;
  std::vector<std::map<int32_t, double>> A(N);
  for (int64_t l = 0; l < m->n_leaves; ++l) {
    double h[3];
    for (int d = 0; d < 3; ++d) h[d] = mv.p(l, d, 7) - mv.p(l, d, 0);
    double eps[8];
    int np[8];
    int32_t par[8][4];
    for (int ii = 0; ii < 8; ++ii) {
      if (!mv.hanging(l, ii)) {
        eps[ii] = eps_node[mv.gt(l, ii)];
        np[ii] = 1;
        par[ii][0] = mv.gt(l, ii);
      } else {
        np[ii] = mv.num_parents(l, ii);
        eps[ii] = 0.0;
        for (int jj = 0; jj < np[ii]; ++jj) {
          par[ii][jj] = mv.gparent(l, jj, ii);
          eps[ii] += eps_node[par[ii][jj]] / np[ii];
        }
      }
    }
    for (int e = 0; e < 12; ++e) {
      const int a = kEdgeAxis[e];
      const int b = (a + 1) % 3, c = (a + 2) % 3;
      const int i1 = kEdgeNodes[2 * e], i2 = kEdgeNodes[2 * e + 1];
      const double cond = wha(eps[i1], eps[i2], frac[12 * l + e]) * (h[b] * h[c] / (4.0 * h[a]));
      const int loc[2] = {i1, i2};
      for (int r = 0; r < 2; ++r)
        for (int s = 0; s < 2; ++s) {
          const double sgn = (r == s) ? 1.0 : -1.0;
          const int ir = loc[r], is = loc[s];
          for (int pr = 0; pr < np[ir]; ++pr)
            for (int ps = 0; ps < np[is]; ++ps) {
              const double w = sgn / (double)(np[ir] * np[is]);
              A[par[ir][pr]][par[is][ps]] += cond * w;
            }
        }
    }
  }
  for (int64_t l = 0; l < m->n_leaves; ++l) {
    const double v8 = (mv.p(l, 0, 7) - mv.p(l, 0, 0)) * (mv.p(l, 1, 7) - mv.p(l, 1, 0)) *
                      (mv.p(l, 2, 7) - mv.p(l, 2, 0)) * 0.125;
    for (int ii = 0; ii < 8; ++ii) {
      if (!mv.hanging(l, ii)) {
        const int32_t g = mv.gt(l, ii);
        A[g][g] += reac_node[g] * v8;
      } else {
        const int npi = mv.num_parents(l, ii);
        for (int jj = 0; jj < npi; ++jj) {
          const int32_t g = mv.gparent(l, jj, ii);
          A[g][g] += reac_node[g] * v8 * (1.0 / npi);
        }
      }
    }
  }
  std::vector<char> isb(N, 0);
  std::vector<double> g(N, 0.0);
  for (int64_t k = 0; k < nb; ++k) { isb[bnd_nodes[k]] = 1; g[bnd_nodes[k]] = bnd_vals[k]; }
  auto S = std::make_unique<orc_sys>();
  S->n = N;
  S->rowptr.assign(N + 1, 0);
  S->rhs.assign(rhs_charge, rhs_charge + N);
  for (int64_t i = 0; i < N; ++i) {
    if (isb[i]) {
      const double d = A[i][(int32_t)i];
      S->col.push_back((int32_t)i);
      S->val.push_back(d);
      S->rhs[i] = d * g[i];
    } else {
      for (const auto& kv : A[i]) {
        if (isb[kv.first]) { S->rhs[i] -= kv.second * g[kv.first]; continue; }
        S->col.push_back(kv.first);
        S->val.push_back(kv.second);
      }
    }
    S->rowptr[i + 1] = (int32_t)S->col.size();
  }
  return S.release();
}

void orc_sys_sizes(const orc_sys* s, int64_t* n, int64_t* nnz) {
  *n = s->n;
  *nnz = (int64_t)s->col.size();
}
void orc_sys_copy(const orc_sys* s, int32_t* rowptr, int32_t* col, double* val, double* rhs) {
  std::memcpy(rowptr, s->rowptr.data(), s->rowptr.size() * sizeof(int32_t));
  std::memcpy(col, s->col.data(), s->col.size() * sizeof(int32_t));
  std::memcpy(val, s->val.data(), s->val.size() * sizeof(double));
  std::memcpy(rhs, s->rhs.data(), s->rhs.size() * sizeof(double));
}
void orc_sys_free(orc_sys* s) { delete s; }

// ---------------------------------------------------------------- Krylov
// y = A x with each row summed left to right as a fused multiply-add chain.
void orc_spmv(int64_t n, const int32_t* rowptr, const int32_t* col, const double* val,
              const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s = std::fma(val[k], x[col[k]], s);
    y[i] = s;
  }
}

namespace {
double dot(int64_t n, const double* a, const double* b, int nth) {
  double s = 0.0;
  if (nth <= 1) {
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  }
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(nth)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
}  // namespace

// LIS 2.1.6 CG with the Jacobi preconditioner as called from
// lis_compute_electric_potential (src/pb_class.cpp:2243-2329) with x0 = 0:
//   r = b, z = D^-1 r, p = z, rho = (r, z)
//   repeat: q = A p; alpha = rho / (p, q); x += alpha p; r -= alpha q;
//           stop when ||r||_2 <= tol ||b||_2;  z = D^-1 r; rho' = (r, z);
//           p = z + (rho'/rho) p
// iters = number of matrix-vector products. nthreads > 1 runs the vector work
// with OpenMP (the CPU baseline); nthreads <= 1 is sequential.
// Returns 0 on convergence, 4 if maxiter was reached, 5 on a non-positive diagonal.
int orc_pcg(int64_t n, const int32_t* rowptr, const int32_t* col, const double* val,
            const double* b, double tol, int maxiter, int nthreads, double* x, int* iters,
            double* relres) {
  const int nth = nthreads < 1 ? 1 : nthreads;
#ifdef _OPENMP
  const int saved = omp_get_max_threads();
  omp_set_num_threads(nth);
#endif
  std::vector<double> dinv(n), r(n), z(n), p(n), q(n);
  int status = 0;
  for (int64_t i = 0; i < n && status == 0; ++i) {
    double d = 0.0;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (col[k] == i) d = val[k];
    if (!(d > 0.0)) status = 5;
    dinv[i] = 1.0 / d;
  }
  *iters = 0;
  *relres = 0.0;
  if (status) goto done;
  {
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int64_t i = 0; i < n; ++i) {
      x[i] = 0.0;
      r[i] = b[i];
      z[i] = dinv[i] * r[i];
      p[i] = z[i];
    }
    const double bnrm = std::sqrt(dot(n, b, b, nth));
    if (bnrm == 0.0) goto done;
    double rho = dot(n, r.data(), z.data(), nth);
    status = 4;
    for (int it = 1; it <= maxiter; ++it) {
      orc_spmv(n, rowptr, col, val, p.data(), q.data());
      const double pq = dot(n, p.data(), q.data(), nth);
      const double alpha = rho / pq;
#pragma omp parallel for schedule(static) num_threads(nth)
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
      }
      const double rn = std::sqrt(dot(n, r.data(), r.data(), nth));
      *iters = it;
      *relres = rn / bnrm;
      if (rn <= tol * bnrm) { status = 0; break; }
#pragma omp parallel for schedule(static) num_threads(nth)
      for (int64_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
      const double rho_new = dot(n, r.data(), z.data(), nth);
      const double beta = rho_new / rho;
      rho = rho_new;
#pragma omp parallel for schedule(static) num_threads(nth)
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
  }
done:
#ifdef _OPENMP
  omp_set_num_threads(saved);
#endif
  return status;
}

// ---------------------------------------------------------------- other LIS solvers
// The reference's default solver options are "-p ssor -ssor_omega 0.51 -i cgs
// -tol 1.e-6" (src/pb_class.cpp:970, data/options.prm:174), handed to LIS 2.1.6
// through lis_solver_set_option (src/pb_class.cpp:2294-2315). LIS is not vendored in
// the reference; what follows restates the published algorithms it names:
//   SSOR preconditioner (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.,
//   sec. 10.2): with A = L + D + U (strict lower, diagonal, strict upper),
//   M = (D + w L) D^-1 (D + w U). Applying M^-1 is a forward sweep
//   y_i = (r_i - w sum_{j<i} a_ij y_j) / d_i followed by a backward sweep
//   z_i = y_i - w (sum_{j>i} a_ij z_j) / d_i. The variant (D/w + L)(D/w)^-1(D/w + U)
//   differs from it by the scalar 1/w, which leaves every Krylov iterate unchanged.
//   CGS (Sonneveld 1989) and BiCGSTAB (van der Vorst 1992) in the preconditioned forms
//   of Barrett et al., Templates for the Solution of Linear Systems (1994), sec. 2.3.7
//   and 2.3.8, with shadow residual r~ = r0 = b (x0 = 0) and the stop test
//   ||r||_2 <= tol ||b||_2 after each iteration (BiCGSTAB also on s at mid-step).
// Arithmetic order (the device follows it exactly in the sweeps): within a row the
// entries are visited in ascending column order, t = t - (w * a_ij) * y_j in the
// forward sweep, t = t + a_ij * z_j in the backward sweep, d_i^-1 = 1.0 / d_i.
// Parity against LIS itself is unpinned (LIS is absent); the sweeps are pinned by
// M z = r with M formed as a product of sparse factors, the solvers by direct solves.
namespace {
struct SsorSplit {
  std::vector<int32_t> diag;  // position of the diagonal entry of each row
  std::vector<double> dinv;
};

int ssor_split(int64_t n, const int32_t* rowptr, const int32_t* col, const double* val, SsorSplit& S) {
  S.diag.assign(n, -1);
  S.dinv.assign(n, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (col[k] == i) S.diag[i] = k;
    if (S.diag[i] < 0 || !(val[S.diag[i]] > 0.0)) return 5;
    S.dinv[i] = 1.0 / val[S.diag[i]];
  }
  return 0;
}

void ssor_apply(int64_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                const SsorSplit& S, double w, const double* r, double* z) {
  for (int64_t i = 0; i < n; ++i) {
    double t = r[i];
    for (int32_t k = rowptr[i]; k < S.diag[i]; ++k) t = t - (w * val[k]) * z[col[k]];
    z[i] = t * S.dinv[i];
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    double t = 0.0;
    for (int32_t k = S.diag[i] + 1; k < rowptr[i + 1]; ++k) t = t + val[k] * z[col[k]];
    z[i] = z[i] - (w * t) * S.dinv[i];
  }
}

double dot1(int64_t n, const double* a, const double* b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
}  // namespace

// z = M^-1 r for the SSOR preconditioner with relaxation w. Returns 5 if a diagonal
// entry is missing or not positive.
int orc_ssor_apply(int64_t n, const int32_t* rowptr, const int32_t* col, const double* val, double w,
                   const double* r, double* z) {
  SsorSplit S;
  if (ssor_split(n, rowptr, col, val, S)) return 5;
  ssor_apply(n, rowptr, col, val, S, w, r, z);
  return 0;
}

// solver: 0 CG, 1 CGS, 2 BiCGSTAB; precon: 0 none, 1 Jacobi, 2 SSOR (relaxation w).
// x0 = 0; iters = iterations done; relres = ||r||_2 / ||b||_2 of the last residual
// tested. Returns 0 on convergence, 4 at maxiter or on a breakdown (a zero rho or
// denominator), 5 on a missing or non-positive diagonal. Sequential.
int orc_krylov(int64_t n, const int32_t* rowptr, const int32_t* col, const double* val,
               const double* b, double tol, int maxiter, int solver, int precon, double w,
               double* x, int* iters, double* relres) {
  SsorSplit S;
  *iters = 0;
  *relres = 0.0;
  if (ssor_split(n, rowptr, col, val, S)) return 5;
  auto psolve = [&](const double* in, double* out) {
    if (precon == 2) ssor_apply(n, rowptr, col, val, S, w, in, out);
    else if (precon == 1) for (int64_t i = 0; i < n; ++i) out[i] = S.dinv[i] * in[i];
    else for (int64_t i = 0; i < n; ++i) out[i] = in[i];
  };
  std::vector<double> r(b, b + n), rt(b, b + n), p(n, 0.0), q(n, 0.0), u(n, 0.0), v(n, 0.0),
      ph(n, 0.0), uh(n, 0.0), t(n, 0.0), z(n, 0.0);
  for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  const double bnrm = std::sqrt(dot1(n, b, b));
  if (bnrm == 0.0) return 0;
  const double stop = tol * bnrm;
  if (solver == 0) {  // preconditioned CG
    psolve(r.data(), z.data());
    p = z;
    double rho = dot1(n, r.data(), z.data());
    for (int it = 1; it <= maxiter; ++it) {
      orc_spmv(n, rowptr, col, val, p.data(), q.data());
      const double pq = dot1(n, p.data(), q.data());
      if (pq == 0.0) return 4;
      const double alpha = rho / pq;
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
      }
      const double rn = std::sqrt(dot1(n, r.data(), r.data()));
      *iters = it;
      *relres = rn / bnrm;
      if (rn <= stop) return 0;
      psolve(r.data(), z.data());
      const double rho1 = dot1(n, r.data(), z.data());
      const double beta = rho1 / rho;
      rho = rho1;
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    return 4;
  }
  if (solver == 1) {  // CGS
    double rho_old = 1.0;
    for (int it = 1; it <= maxiter; ++it) {
      const double rho = dot1(n, rt.data(), r.data());
      if (rho == 0.0) return 4;
      const double beta = it == 1 ? 0.0 : rho / rho_old;
      for (int64_t i = 0; i < n; ++i) {
        u[i] = r[i] + beta * q[i];
        p[i] = u[i] + beta * (q[i] + beta * p[i]);
      }
      psolve(p.data(), ph.data());
      orc_spmv(n, rowptr, col, val, ph.data(), v.data());
      const double rv = dot1(n, rt.data(), v.data());
      if (rv == 0.0) return 4;
      const double alpha = rho / rv;
      for (int64_t i = 0; i < n; ++i) {
        q[i] = u[i] - alpha * v[i];
        t[i] = u[i] + q[i];
      }
      psolve(t.data(), uh.data());
      for (int64_t i = 0; i < n; ++i) x[i] += alpha * uh[i];
      orc_spmv(n, rowptr, col, val, uh.data(), t.data());
      for (int64_t i = 0; i < n; ++i) r[i] -= alpha * t[i];
      const double rn = std::sqrt(dot1(n, r.data(), r.data()));
      *iters = it;
      *relres = rn / bnrm;
      if (rn <= stop) return 0;
      rho_old = rho;
    }
    return 4;
  }
  // BiCGSTAB
  double rho_old = 1.0, alpha = 1.0, omega = 1.0;
  std::vector<double>& s = u;
  std::vector<double>& sh = uh;
  for (int it = 1; it <= maxiter; ++it) {
    const double rho = dot1(n, rt.data(), r.data());
    if (rho == 0.0) return 4;
    const double beta = it == 1 ? 0.0 : (rho / rho_old) * (alpha / omega);
    for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    psolve(p.data(), ph.data());
    orc_spmv(n, rowptr, col, val, ph.data(), v.data());
    const double rv = dot1(n, rt.data(), v.data());
    if (rv == 0.0) return 4;
    alpha = rho / rv;
    for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    const double sn = std::sqrt(dot1(n, s.data(), s.data()));
    *iters = it;
    if (sn <= stop) {
      for (int64_t i = 0; i < n; ++i) x[i] += alpha * ph[i];
      *relres = sn / bnrm;
      return 0;
    }
    psolve(s.data(), sh.data());
    orc_spmv(n, rowptr, col, val, sh.data(), t.data());
    const double ts = dot1(n, t.data(), s.data()), tt = dot1(n, t.data(), t.data());
    if (tt == 0.0 || ts == 0.0) return 4;
    omega = ts / tt;
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = s[i] - omega * t[i];
    }
    const double rn = std::sqrt(dot1(n, r.data(), r.data()));
    *relres = rn / bnrm;
    if (rn <= stop) return 0;
    rho_old = rho;
  }
  return 4;
}

// Timing helper for the CPU baseline: runs `iters` PCG iterations (no stopping
// test) on n unknowns with nthreads OpenMP threads and returns seconds elapsed.
double orc_pcg_time_iters(int64_t n, const int32_t* rowptr, const int32_t* col,
                          const double* val, const double* b, int iters, int nthreads) {
  std::vector<double> x(n);
  int it = 0;
  double rr = 0;
#ifdef _OPENMP
  const double t0 = omp_get_wtime();
#endif
  orc_pcg(n, rowptr, col, val, b, 0.0, iters, nthreads, x.data(), &it, &rr);
#ifdef _OPENMP
  return omp_get_wtime() - t0;
#else
  return -1.0;
#endif
}

// ---------------------------------------------------------------- Coulomb energy
// calc_coulombic == 1 in energy / energy_fast (src/pb_class.cpp:2842-2857 and
// 3116-3131): over the charged atoms (|q| > 1e-5, :2827-2832) in input order, one
// running sum of (q_i q_j) / r_ij over pairs i < j, then times 1 / eps_in_hat.
double orc_coul_energy(int64_t na, const double* xyz, const double* q, double eps_in) {
  std::vector<int64_t> idx;
  for (int64_t a = 0; a < na; ++a)
    if (std::fabs(q[a]) > 1.e-5) idx.push_back(a);
  double e = 0.0;
  for (size_t i = 0; i < idx.size(); ++i) {
    const double* ri = xyz + 3 * idx[i];
    const double qi = q[idx[i]];
    for (size_t j = i + 1; j < idx.size(); ++j) {
      const double* rj = xyz + 3 * idx[j];
      const double dx = ri[0] - rj[0], dy = ri[1] - rj[1], dz = ri[2] - rj[2];
      const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
      e += qi * q[idx[j]] / r;
    }
  }
  return e * (1.0 / eps_in);
}

// ---------------------------------------------------------------- energy
// energy (src/pb_class.cpp:2790-3063) for one rank, with classifyCube
// (:2570-2598), classifyCube_flux (:2627-2661), getTriangles (:2771-2788), phi0
// (:2700-2704), areaTriangle (:2706-2722). The normal used at a triangle vertex is
// the one most recently written by normal_intersection; it starts at zero.
// out[0] = energy_pol, out[1] = energy_react, out[2] = charge_pol (flux charge * 4 pi)
int orc_energy(const orc_mesh* m, const orc_rays* rays, const double* eps_node,
               const double* phi, int64_t n_border, const int32_t* border, int64_t na,
               const double* xyz, const double* q, double e_in, double e_out, double T,
               double I, int calc_energy, double* out) {
  const auto maps = build_maps(rays);
  MeshView mv{m};
  double cst[6];
  orc_constants(e_in, e_out, T, I, cst);
  const double eps_out = cst[1];
  const double k = cst[3];
  const double inv_4pi = 1.0 / (4.0 * kPi);
  const double constant_pol = cst[4], constant_react = cst[5];
  std::vector<double> qa;
  std::vector<std::array<double, 3>> pa;
  for (int64_t a = 0; a < na; ++a)
    if (std::fabs(q[a]) > 1.e-5) {
      qa.push_back(q[a]);
      pa.push_back({xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]});
    }
  const size_t nq = qa.size();
  double first_int = 0.0, second_int = 0.0, charge_pol = 0.0;
  double N[3] = {0.0, 0.0, 0.0};
  double energy_pol = 0.0, energy_react = 0.0;
  const bool ionic = (calc_energy == 2 && k > 1.e-5);
  const bool polar_only = (calc_energy == 1 || (calc_energy == 2 && k < 1.e-5));
  for (int64_t bi = 0; bi < n_border && (ionic || polar_only); ++bi) {
    const int64_t l = border[bi];
    double h[3], area_h[3];
    for (int d = 0; d < 3; ++d) h[d] = mv.p(l, d, 7) - mv.p(l, d, 0);
    area_h[0] = h[1] * h[2] / h[0] * 0.25;
    area_h[1] = h[0] * h[2] / h[1] * 0.25;
    area_h[2] = h[0] * h[1] / h[2] * 0.25;
    // classifyCube_flux
    double te[8], tp[8];
    for (int ii = 0; ii < 8; ++ii) {
      if (!mv.hanging(l, ii)) {
        te[ii] = eps_node[mv.gt(l, ii)];
        tp[ii] = phi[mv.gt(l, ii)];
      } else {
        te[ii] = 0.0;
        tp[ii] = 0.0;
        const int npi = mv.num_parents(l, ii);
        for (int jj = 0; jj < npi; ++jj) {
          te[ii] += eps_node[mv.gparent(l, jj, ii)] / npi;
          tp[ii] += phi[mv.gparent(l, jj, ii)] / npi;
        }
      }
    }
    int cubeindex = -1;
    if (ionic) {
      // classifyCube (hanging corners: mean of parents)
      constexpr double EPS = 1e-10;
      static const int order[8] = {0, 1, 3, 2, 4, 5, 7, 6};
      int ci = 0, bit = 1;
      for (int o = 0; o < 8; ++o) {
        const int ii = order[o];
        double v;
        if (!mv.hanging(l, ii)) v = eps_node[mv.gt(l, ii)];
        else {
          v = 0.0;
          const int npi = mv.num_parents(l, ii);
          for (int jj = 0; jj < npi; ++jj) v += eps_node[mv.gparent(l, jj, ii)] / npi;
        }
        if (v < eps_out - EPS) ci |= bit;
        bit *= 2;
      }
      cubeindex = (ci == 0 || ci == 255) ? -1 : ci;
    }
    for (int e = 0; e < 12; ++e) {
      const int i1 = kEdgeNodes[2 * e], i2 = kEdgeNodes[2 * e + 1];
      double fl;
      if (te[i1] < te[i2]) fl = 1;
      else if (te[i1] > te[i2]) fl = -1;
      else continue;
      const int axis = kEdgeAxis[e];
      double fract = 0.0;
      normal_intersection(mv, maps, l, e, N, &fract);
      double V[3] = {mv.p(l, 0, i1), mv.p(l, 1, i1), mv.p(l, 2, i1)};
      V[axis] += fract * h[axis];
      const double tmp_flux = -(tp[i2] - tp[i1]) * wha(te[i1], te[i2], fract) * fl * area_h[axis];
      charge_pol += tmp_flux;
      for (size_t ia = 0; ia < nq; ++ia) {
        const double dx = pa[ia][0] - V[0], dy = pa[ia][1] - V[1], dz = pa[ia][2] - V[2];
        const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (polar_only) {
          const double qflux = tmp_flux / r;
          first_int += qa[ia] * qflux;
        } else {
          first_int += qa[ia] * tmp_flux / r;
        }
      }
    }
    if (ionic && cubeindex >= 0) {
      const char* tri = kMcTriangles[cubeindex];
      const int ntri = (int)std::strlen(tri) / 3;
      for (int t = 0; t < ntri; ++t) {
        double vt[3][3], nv[3][3], ps[3];
        for (int jj = 0; jj < 3; ++jj) {
          const int edge = mc_digit(tri[3 * t + jj]);
          const int axis = kEdgeAxis[edge];
          const int i1 = kEdgeNodes[2 * edge], i2 = kEdgeNodes[2 * edge + 1];
          double fract = 0.0;
          normal_intersection(mv, maps, l, edge, N, &fract);
          double V[3] = {mv.p(l, 0, i1), mv.p(l, 1, i1), mv.p(l, 2, i1)};
          V[axis] += fract * h[axis];
          for (int d = 0; d < 3; ++d) { vt[jj][d] = V[d]; nv[jj][d] = N[d]; }
          ps[jj] = phi0(te[i1], te[i2], tp[i1], tp[i2], fract);
        }
        const double area = area_triangle(vt);
        for (size_t ia = 0; ia < nq; ++ia) {
          const double qi = qa[ia];
          for (int kk = 0; kk < 3; ++kk) {
            const double dv[3] = {vt[kk][0] - pa[ia][0], vt[kk][1] - pa[ia][1], vt[kk][2] - pa[ia][2]};
            const double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            const double r = std::sqrt(r2);
            const double inv_r3 = 1.0 / (r2 * r);
            const double dt = dv[0] * nv[kk][0] + dv[1] * nv[kk][1] + dv[2] * nv[kk][2];
            second_int += qi * ps[kk] * dt * inv_r3 * inv_4pi * area / 3.0;
          }
        }
      }
    }
  }
  if (polar_only) energy_pol = 0.5 * constant_pol * first_int;
  if (ionic) {
    energy_pol = 0.5 * constant_pol * first_int;
    energy_react = 0.5 * (second_int - first_int * constant_react);
  }
  out[0] = energy_pol;
  out[1] = energy_react;
  out[2] = charge_pol;
  return 0;
}

// ---------------------------------------------------------------- potentials and fields at atoms
// pot_field (src/pb_class.cpp:4389-5060; pot_field_fast :3731-4387 is the same on
// meshes without hanging nodes, where every leaf has the first border leaf's h).
// Branch, from calc_potential_term P and calc_field_term F (:3857-4156):
//   1  (F == 1 || (F == 2 && k < 1e-5)) && P < 2   phi_p, field_p
//   2  (P == 1 || (P == 2 && k < 1e-5)) && F == 0   phi_p
//   3  F == 2 && k > 1e-5                            phi_p, phi_i, field_p, field_i
//   4  P == 2 && k > 1e-5 && F == 0                  phi_p, phi_i
//   5  P == 2 && k > 1e-5 && F == 1                  phi_p, phi_i, field_p
//   0  none of these                                 Coulomb terms only
// table: n_atoms rows of 12 columns phi_c phi_p phi_i Ex_c Ey_c Ez_c Ex_p Ey_p Ez_p
// Ex_i Ey_i Ez_i, with the write-out corrections applied (:4286-4290: phi_i and
// field_i minus phi_p and field_p times constant_react / constant_pol); columns a
// branch does not compute are 0. out[0..3] = energy_pol, energy_react, coul_energy,
// charge_pol.
int orc_pot_field_branch(int P, int F, double k) {
  if ((F == 1 || (F == 2 && k < 1.e-5)) && P < 2) return 1;
  if ((P == 1 || (P == 2 && k < 1.e-5)) && F == 0) return 2;
  if (F == 2 && k > 1.e-5) return 3;
  if (P == 2 && k > 1.e-5 && F == 0) return 4;
  if (P == 2 && k > 1.e-5 && F == 1) return 5;
  return 0;
}

int orc_pot_field(const orc_mesh* m, const orc_rays* rays, const double* eps_node, const double* phi,
                  int64_t n_border, const int32_t* border, int64_t na, const double* xyz, const double* q,
                  double e_in, double e_out, double T, double I, int P, int F, double* table, double* out) {
  const auto maps = build_maps(rays);
  MeshView mv{m};
  double cst[6];
  orc_constants(e_in, e_out, T, I, cst);
  const double eps_in = cst[0], eps_out = cst[1], k = cst[3];
  const double inv_4pi = 1.0 / (4.0 * kPi);
  const double den_in = 1.0 / eps_in;
  const double constant_pol = cst[4], constant_react = cst[5];
  const int branch = orc_pot_field_branch(P, F, k);
  std::vector<double> phi_c(na, 0.0), fcx(na, 0.0), fcy(na, 0.0), fcz(na, 0.0);
  std::vector<double> phi_p(na, 0.0), fpx(na, 0.0), fpy(na, 0.0), fpz(na, 0.0);
  std::vector<double> phi_i(na, 0.0), fix(na, 0.0), fiy(na, 0.0), fiz(na, 0.0);
  auto pos = [&](int64_t a, int d) { return xyz[3 * a + d]; };
  double coul_energy = 0.0;
  // Coulomb pairs (:3776-3813)
  for (int64_t i = 0; i < na; ++i) {
    const double qi = q[i];
    for (int64_t j = i + 1; j < na; ++j) {
      const double qj = q[j];
      const double dx = pos(i, 0) - pos(j, 0), dy = pos(i, 1) - pos(j, 1), dz = pos(i, 2) - pos(j, 2);
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double r = std::sqrt(r2);
      const double inv_r3 = 1.0 / (r2 * r);
      coul_energy += (qi * qj) / r * den_in;
      phi_c[i] += qj / r * den_in;
      phi_c[j] += qi / r * den_in;
      fcx[i] += dx * inv_r3 * qj * den_in;
      fcy[i] += dy * inv_r3 * qj * den_in;
      fcz[i] += dz * inv_r3 * qj * den_in;
      fcx[j] += -dx * inv_r3 * qi * den_in;
      fcy[j] += -dy * inv_r3 * qi * den_in;
      fcz[j] += -dz * inv_r3 * qi * den_in;
    }
  }
  const bool fieldp = branch == 1 || branch == 3 || branch == 5;
  const bool ionic = branch >= 3;
  const bool fieldi = branch == 3;
  double first_int = 0.0, second_int = 0.0, charge_pol = 0.0;
  double N[3] = {0.0, 0.0, 0.0};
  for (int64_t bi = 0; bi < n_border && branch != 0; ++bi) {
    const int64_t l = border[bi];
    double h[3], area_h[3];
    for (int d = 0; d < 3; ++d) h[d] = mv.p(l, d, 7) - mv.p(l, d, 0);
    area_h[0] = h[1] * h[2] / h[0] * 0.25;
    area_h[1] = h[0] * h[2] / h[1] * 0.25;
    area_h[2] = h[0] * h[1] / h[2] * 0.25;
    double te[8], tp[8];
    for (int ii = 0; ii < 8; ++ii) {
      if (!mv.hanging(l, ii)) {
        te[ii] = eps_node[mv.gt(l, ii)];
        tp[ii] = phi[mv.gt(l, ii)];
      } else {
        te[ii] = 0.0;
        tp[ii] = 0.0;
        const int npi = mv.num_parents(l, ii);
        for (int jj = 0; jj < npi; ++jj) {
          te[ii] += eps_node[mv.gparent(l, jj, ii)] / npi;
          tp[ii] += phi[mv.gparent(l, jj, ii)] / npi;
        }
      }
    }
    int cubeindex = -1;
    if (ionic) {
      static const int order[8] = {0, 1, 3, 2, 4, 5, 7, 6};
      int ci = 0, bit = 1;
      for (int o = 0; o < 8; ++o) {
        if (te[order[o]] < eps_out - 1e-10) ci |= bit;
        bit *= 2;
      }
      cubeindex = (ci == 0 || ci == 255) ? -1 : ci;
    }
    for (int e = 0; e < 12; ++e) {
      const int i1 = kEdgeNodes[2 * e], i2 = kEdgeNodes[2 * e + 1];
      double fl;
      if (te[i1] < te[i2]) fl = 1;
      else if (te[i1] > te[i2]) fl = -1;
      else continue;
      const int axis = kEdgeAxis[e];
      double fract = 0.0;
      normal_intersection(mv, maps, l, e, N, &fract);
      double V[3] = {mv.p(l, 0, i1), mv.p(l, 1, i1), mv.p(l, 2, i1)};
      V[axis] += fract * h[axis];
      const double tmp_flux = -(tp[i2] - tp[i1]) * wha(te[i1], te[i2], fract) * fl * area_h[axis];
      charge_pol += tmp_flux;
      for (int64_t ia = 0; ia < na; ++ia) {
        const double dx = pos(ia, 0) - V[0], dy = pos(ia, 1) - V[1], dz = pos(ia, 2) - V[2];
        const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
        const double inv_r3 = 1.0 / (r * r * r);
        if (branch <= 2) {
          const double qflux = tmp_flux / r;
          first_int += q[ia] * qflux;
          phi_p[ia] += qflux * constant_pol;
        } else {
          first_int += q[ia] * tmp_flux / r;
          phi_p[ia] += tmp_flux / r * constant_pol;
        }
        if (fieldp) {
          fpx[ia] += dx * inv_r3 * tmp_flux * constant_pol;
          fpy[ia] += dy * inv_r3 * tmp_flux * constant_pol;
          fpz[ia] += dz * inv_r3 * tmp_flux * constant_pol;
        }
      }
    }
    if (ionic && cubeindex >= 0) {
      const char* tri = kMcTriangles[cubeindex];
      const int ntri = (int)std::strlen(tri) / 3;
      for (int t = 0; t < ntri; ++t) {
        double vt[3][3], nv[3][3], ps[3];
        for (int jj = 0; jj < 3; ++jj) {
          const int edge = mc_digit(tri[3 * t + jj]);
          const int axis = kEdgeAxis[edge];
          const int i1 = kEdgeNodes[2 * edge], i2 = kEdgeNodes[2 * edge + 1];
          double fract = 0.0;
          normal_intersection(mv, maps, l, edge, N, &fract);
          double V[3] = {mv.p(l, 0, i1), mv.p(l, 1, i1), mv.p(l, 2, i1)};
          V[axis] += fract * h[axis];
          for (int d = 0; d < 3; ++d) { vt[jj][d] = V[d]; nv[jj][d] = N[d]; }
          ps[jj] = phi0(te[i1], te[i2], tp[i1], tp[i2], fract);
        }
        const double area = area_triangle(vt);
        for (int64_t ia = 0; ia < na; ++ia) {
          const double qi = q[ia];
          for (int kk = 0; kk < 3; ++kk) {
            const double dv[3] = {vt[kk][0] - pos(ia, 0), vt[kk][1] - pos(ia, 1), vt[kk][2] - pos(ia, 2)};
            const double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            const double r = std::sqrt(r2);
            const double inv_r3 = 1.0 / (r2 * r);
            const double inv_r5 = inv_r3 / r2;
            const double dt = dv[0] * nv[kk][0] + dv[1] * nv[kk][1] + dv[2] * nv[kk][2];
            const double factor = ps[kk] * inv_4pi * area / 3.0;
            second_int += qi * ps[kk] * dt * inv_r3 * inv_4pi * area / 3.0;
            phi_i[ia] += ps[kk] * dt * inv_r3 * inv_4pi * area / 3.0;
            if (fieldi) {
              fix[ia] += factor * (-3 * dv[0] * inv_r5 * dt + nv[kk][0] * inv_r3);
              fiy[ia] += factor * (-3 * dv[1] * inv_r5 * dt + nv[kk][1] * inv_r3);
              fiz[ia] += factor * (-3 * dv[2] * inv_r5 * dt + nv[kk][2] * inv_r3);
            }
          }
        }
      }
    }
  }
  double energy_pol = 0.0, energy_react = 0.0;
  if (branch != 0) energy_pol = 0.5 * constant_pol * first_int;
  if (ionic) energy_react = 0.5 * (second_int - first_int * constant_react);
  for (int64_t i = 0; i < na; ++i) {
    if (ionic) phi_i[i] -= phi_p[i] / constant_pol * constant_react;
    if (fieldi) {
      fix[i] -= fpx[i] / constant_pol * constant_react;
      fiy[i] -= fpy[i] / constant_pol * constant_react;
      fiz[i] -= fpz[i] / constant_pol * constant_react;
    }
    double* row = table + 12 * i;
    const double vals[12] = {phi_c[i], phi_p[i], phi_i[i], fcx[i], fcy[i], fcz[i],
                             fpx[i],   fpy[i],   fpz[i],   fix[i], fiy[i], fiz[i]};
    for (int c = 0; c < 12; ++c) row[c] = vals[c];
  }
  out[0] = energy_pol;
  out[1] = energy_react;
  out[2] = coul_energy;
  out[3] = charge_pol;
  return branch;
}

// Born-sphere potential used for bc_type 3 (src/pb_class.cpp:3549-3571): q = +1 and
// rs = 2 are fixed there.
double orc_analytic_solution(double x, double y, double z, double e_in, double e_out,
                             double T, double I) {
  double c[6];
  orc_constants(e_in, e_out, T, I, c);
  const double eps_in = c[0], eps_out = c[1], k = c[3];
  const double rs = 2.0;
  const double pippo_out = eps_out * (1.0 + k * rs);
  const double pippo_in = eps_in * rs;
  const double dist = std::hypot(x, y, z);
  if (dist <= rs) return 1.0 / (pippo_out * rs) + (rs - dist) / (pippo_in * dist);
  return std::exp(k * (rs - dist)) / (dist * pippo_out);
}

}  // extern "C"

// ---------------------------------------------------------------- SES rays
// The reference's default surface is the solvent-excluded surface with probe radius
// 1.4 A (surface_type 0, src/pb_class.cpp:956-963, data/options.prm:136-140), whose
// axis-ray crossings NanoShaper supplies (src/raytracer.cpp:337-376). NanoShaper is not
// vendored; this is the textbook definition restated (Richards 1977; Connolly 1983):
// with SAS balls B_i = B(a_i, R_i), R_i = r_i + rp, the probe-accessible centres are
// C = { q : |q - a_i| >= R_i for all i } (interior cavities included), and a point p
// lies inside the SES iff no probe ball B(q, rp), q in C, reaches it:
// dist(p, C) > rp. For p inside some B_i the nearest point of C lies on the SAS
// surface, at one of three kinds of candidates, each counted only where it is
// exposed (inside no other SAS ball):
//   the radial projection of p onto SAS sphere i          (convex atom patches),
//   the point of the SAS circle of a pair (i, j) nearest p  (re-entrant tori),
//   an SAS vertex of a triple (i, j, k)                     (re-entrant probe spheres).
// Circles buried in the union of the other SAS balls are dropped beforehand
// (circle_exposed), and SAS vertices are taken from triples of kept circles.
// A point inside an atom's own ball (|p - a_i| < r_i) is inside; a point outside
// every SAS ball is outside. Along an axis ray the SES boundary can only lie on an
// atom sphere (radius r_i), a torus (tube rp around an SAS circle) or a probe sphere
// (radius rp around an exposed SAS vertex); the ray's crossings are those candidate
// parameters across which the classification of the midpoints changes.
// Torus roots: the quartic g(tau) = (tau^2 + K)^2 - 4 rho^2 (tau^2 + D^2 - (n_a tau +
// s)^2) on [-T, T] (T^2 = (rho + rp)^2 - D^2, tau = t - c_axis) is sampled at 65
// points; every sign change is bisected 64 times (two roots closer than 2T/64 are not
// resolved); every candidate is normalised by + 0.0 (no negative zero). Normals point
// out of the molecule: (p - a_i)/r_i on atoms, (q - p)/rp
// towards the probe centre q on tori and probe spheres. All arithmetic is plain IEEE
// double (+ - * / sqrt), so the device caster reproduces it bit for bit.
namespace {
struct SesPair {
  int i, j;
  double c[3], n[3], rho;
};
struct SesModel {
  int64_t na;
  const double* xyz;
  const double* r;
  double rp;
  std::vector<double> R;
  std::vector<SesPair> pairs;
  std::vector<std::array<double, 3>> verts;
};

double d2(const double* a, const double* b) {
  const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
  return x * x + y * y + z * z;
}

bool ses_exposed(const SesModel& M, const double* q, int e0, int e1, int e2) {
  for (int64_t m = 0; m < M.na; ++m) {
    if (m == e0 || m == e1 || m == e2) continue;
    if (d2(q, M.xyz + 3 * m) < M.R[m] * M.R[m]) return false;
  }
  return true;
}

// Whether some arc of the SAS circle (c, n, rho) of pair (i, j) lies outside every
// other SAS ball m that overlaps ball i (only those can cover its points): ball m covers
// the circle points q(th) = c + rho (cos th u + sin th w) with P cos th + Q sin th < S,
// P = 2 rho (c - a_m).u, Q = 2 rho (c - a_m).w, S = R_m^2 - |c - a_m|^2 - rho^2, an arc
// of half-width pi - acos(S / A) (A = |(P, Q)|) about the direction opposite (P, Q).
// The union of the covered open arcs is swept in angle; a circle whose arcs leave no
// gap is buried and contributes no SES patch (an exposed SAS vertex lies on three
// exposed circles). Host-only arithmetic (atan2, acos of the C library).
bool circle_exposed(const double* c, const double* n, double rho, int i, int j, int64_t na, const double* xyz,
                    const std::vector<double>& R) {
  const int ax = std::fabs(n[0]) < 0.9 ? 0 : 1;
  double u[3] = {0, 0, 0};
  u[ax] = 1.0;
  const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
  for (int k = 0; k < 3; ++k) u[k] = u[k] - un * n[k];
  const double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int k = 0; k < 3; ++k) u[k] = u[k] / ul;
  const double w[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
  const double two_pi = 2.0 * kPi;
  std::vector<std::pair<double, double>> iv;
  for (int64_t m = 0; m < na; ++m) {
    if (m == i || m == j) continue;
    const double* am = xyz + 3 * m;
    if (!(std::sqrt(d2(xyz + 3 * i, am)) < R[i] + R[m])) continue;
    const double d[3] = {c[0] - am[0], c[1] - am[1], c[2] - am[2]};
    const double P = 2.0 * rho * (d[0] * u[0] + d[1] * u[1] + d[2] * u[2]);
    const double Q = 2.0 * rho * (d[0] * w[0] + d[1] * w[1] + d[2] * w[2]);
    const double S = R[m] * R[m] - (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - rho * rho;
    const double A = std::sqrt(P * P + Q * Q);
    if (S >= A) return false;  // the whole circle
    if (S <= -A) continue;     // none of it
    const double beta = std::acos(S / A);
    double s0 = std::atan2(Q, P) + beta;  // covered: angles in (s0, s0 + 2 pi - 2 beta)
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
  if (iv.empty()) return true;
  std::sort(iv.begin(), iv.end());
  double cur = 0.0;
  for (const auto& a : iv) {
    if (a.first > cur) return true;  // an uncovered gap
    cur = std::max(cur, a.second);
  }
  return cur < two_pi;
}

void ses_build(SesModel& M) {
  M.R.resize(M.na);
  for (int64_t i = 0; i < M.na; ++i) M.R[i] = M.r[i] + M.rp;
  for (int i = 0; i < (int)M.na; ++i)
    for (int j = i + 1; j < (int)M.na; ++j) {
      const double* ai = M.xyz + 3 * i;
      const double* aj = M.xyz + 3 * j;
      const double dd = std::sqrt(d2(ai, aj));
      if (!(dd < M.R[i] + M.R[j]) || !(dd > std::fabs(M.R[i] - M.R[j]))) continue;
      const double t = (dd * dd + M.R[i] * M.R[i] - M.R[j] * M.R[j]) / (2.0 * dd);
      const double rho2 = M.R[i] * M.R[i] - t * t;
      if (!(rho2 > 0.0)) continue;
      SesPair P;
      P.i = i;
      P.j = j;
      for (int k = 0; k < 3; ++k) {
        P.n[k] = (aj[k] - ai[k]) / dd;
        P.c[k] = ai[k] + t * P.n[k];
      }
      P.rho = std::sqrt(rho2);
      if (circle_exposed(P.c, P.n, P.rho, i, j, M.na, M.xyz, M.R)) M.pairs.push_back(P);
    }
  // SAS vertices: the two points at distance R_i, R_j, R_k of a triple whose SAS
  // spheres meet pairwise, kept where no other SAS ball covers them
  const size_t np = M.pairs.size();
  std::vector<std::vector<int>> nb(M.na);
  for (const SesPair& P : M.pairs) {
    nb[P.i].push_back(P.j);
    nb[P.j].push_back(P.i);
  }
  for (auto& v : nb) std::sort(v.begin(), v.end());
  (void)np;
  for (const SesPair& P : M.pairs) {
    for (int k : nb[P.j]) {
      if (k <= P.j || !std::binary_search(nb[P.i].begin(), nb[P.i].end(), k)) continue;
      const double* ai = M.xyz + 3 * P.i;
      const double* ak = M.xyz + 3 * k;
      // trilateration in the frame (ex = n, ey, ez)
      double ex[3] = {P.n[0], P.n[1], P.n[2]};
      double w[3] = {ak[0] - ai[0], ak[1] - ai[1], ak[2] - ai[2]};
      const double ii = ex[0] * w[0] + ex[1] * w[1] + ex[2] * w[2];
      double ey[3] = {w[0] - ii * ex[0], w[1] - ii * ex[1], w[2] - ii * ex[2]};
      const double ey_n = std::sqrt(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
      if (!(ey_n > 0.0)) continue;
      for (int q = 0; q < 3; ++q) ey[q] = ey[q] / ey_n;
      const double ez[3] = {ex[1] * ey[2] - ex[2] * ey[1], ex[2] * ey[0] - ex[0] * ey[2], ex[0] * ey[1] - ex[1] * ey[0]};
      const double jj = ey[0] * w[0] + ey[1] * w[1] + ey[2] * w[2];
      const double dd = std::sqrt(d2(ai, M.xyz + 3 * P.j));
      const double x = (M.R[P.i] * M.R[P.i] - M.R[P.j] * M.R[P.j] + dd * dd) / (2.0 * dd);
      const double y = (M.R[P.i] * M.R[P.i] - M.R[k] * M.R[k] + ii * ii + jj * jj) / (2.0 * jj) - (ii / jj) * x;
      const double z2 = M.R[P.i] * M.R[P.i] - x * x - y * y;
      if (!(z2 > 0.0)) continue;
      const double z = std::sqrt(z2);
      for (int sgn = 0; sgn < 2; ++sgn) {
        const double zz = sgn == 0 ? z : -z;
        std::array<double, 3> v;
        for (int q = 0; q < 3; ++q) v[q] = ai[q] + x * ex[q] + y * ey[q] + zz * ez[q];
        if (ses_exposed(M, v.data(), P.i, P.j, k)) M.verts.push_back(v);
      }
    }
  }
}

// nearest point of the SAS circle of P to p (any circle point when p is on the axis)
void ses_circle_point(const SesPair& P, const double* p, double* q) {
  double w[3] = {p[0] - P.c[0], p[1] - P.c[1], p[2] - P.c[2]};
  const double s = w[0] * P.n[0] + w[1] * P.n[1] + w[2] * P.n[2];
  double u[3] = {w[0] - s * P.n[0], w[1] - s * P.n[1], w[2] - s * P.n[2]};
  double L = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  if (!(L > 0.0)) {  // a fixed direction perpendicular to n
    const int a = std::fabs(P.n[0]) < 0.9 ? 0 : 1;
    double e[3] = {0, 0, 0};
    e[a] = 1.0;
    const double ds = e[0] * P.n[0] + e[1] * P.n[1] + e[2] * P.n[2];
    for (int k = 0; k < 3; ++k) u[k] = e[k] - ds * P.n[k];
    L = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  }
  for (int k = 0; k < 3; ++k) q[k] = P.c[k] + P.rho * (u[k] / L);
}

// 1 inside the SES, 0 outside (on the surface counts as outside)
int ses_classify(const SesModel& M, const double* p) {
  bool in_sas = false;
  for (int64_t i = 0; i < M.na; ++i) {
    const double dd = d2(p, M.xyz + 3 * i);
    if (dd < M.r[i] * M.r[i]) return 1;
    if (dd < M.R[i] * M.R[i]) in_sas = true;
  }
  if (!in_sas) return 0;
  const double rp2 = M.rp * M.rp;
  for (int64_t i = 0; i < M.na; ++i) {
    const double* a = M.xyz + 3 * i;
    const double dd = std::sqrt(d2(p, a));
    const double gap = M.R[i] - dd;
    if (!(dd > 0.0) || std::fabs(gap) > M.rp) continue;
    double q[3];
    for (int k = 0; k < 3; ++k) q[k] = a[k] + M.R[i] * ((p[k] - a[k]) / dd);
    if (ses_exposed(M, q, (int)i, -1, -1)) return 0;
  }
  for (const SesPair& P : M.pairs) {
    double q[3];
    ses_circle_point(P, p, q);
    if (d2(p, q) <= rp2 && ses_exposed(M, q, P.i, P.j, -1)) return 0;
  }
  for (const auto& v : M.verts)
    if (d2(p, v.data()) <= rp2) return 0;
  return 1;
}

struct SesCand {
  double t;
  int type;  // 0 atom, 1 torus, 2 probe sphere
  int id;
};

void ses_ray(const SesModel& M, int axis, double u, double v, std::vector<double>& xs, std::vector<double>& ns) {
  const int d1 = axis == 0 ? 1 : 0;
  const int d2_ = axis == 2 ? 1 : 2;
  std::vector<SesCand> cand;
  auto at = [&](double t, double* p) {
    p[axis] = t;
    p[d1] = u;
    p[d2_] = v;
  };
  for (int64_t i = 0; i < M.na; ++i) {
    const double* a = M.xyz + 3 * i;
    const double du = u - a[d1], dv = v - a[d2_];
    const double q = M.r[i] * M.r[i] - (du * du + dv * dv);
    if (!(q > 0.0)) continue;
    const double s = std::sqrt(q);
    cand.push_back({a[axis] - s + 0.0, 0, (int)i});
    cand.push_back({a[axis] + s + 0.0, 0, (int)i});
  }
  for (size_t k = 0; k < M.verts.size(); ++k) {
    const double* c = M.verts[k].data();
    const double du = u - c[d1], dv = v - c[d2_];
    const double q = M.rp * M.rp - (du * du + dv * dv);
    if (!(q > 0.0)) continue;
    const double s = std::sqrt(q);
    cand.push_back({c[axis] - s + 0.0, 2, (int)k});
    cand.push_back({c[axis] + s + 0.0, 2, (int)k});
  }
  for (size_t k = 0; k < M.pairs.size(); ++k) {
    const SesPair& P = M.pairs[k];
    const double wu = u - P.c[d1], wv = v - P.c[d2_];
    const double D2 = wu * wu + wv * wv;
    const double T2 = (P.rho + M.rp) * (P.rho + M.rp) - D2;
    if (!(T2 > 0.0)) continue;
    const double T = std::sqrt(T2);
    const double s0 = P.n[d1] * wu + P.n[d2_] * wv;
    const double na_ = P.n[axis];
    const double K = D2 + P.rho * P.rho - M.rp * M.rp;
    const double f = 4.0 * P.rho * P.rho;
    auto g = [&](double tau) {
      const double a = tau * tau + K;
      const double b = na_ * tau + s0;
      return a * a - f * (tau * tau + D2 - b * b);
    };
    double t0 = -T, g0 = g(t0);
    for (int m = 1; m <= 64; ++m) {
      const double t1 = -T + (2.0 * T) * ((double)m / 64.0);
      const double g1 = g(t1);
      if (g0 == 0.0) {
        cand.push_back({P.c[axis] + t0 + 0.0, 1, (int)k});
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
        cand.push_back({P.c[axis] + 0.5 * (lo + hi) + 0.0, 1, (int)k});
      }
      t0 = t1;
      g0 = g1;
    }
    if (g0 == 0.0) cand.push_back({P.c[axis] + t0 + 0.0, 1, (int)k});
  }
  std::sort(cand.begin(), cand.end(), [](const SesCand& a, const SesCand& b) {
    return a.t < b.t || (a.t == b.t && (a.type < b.type || (a.type == b.type && a.id < b.id)));
  });
  // groups of equal t; classification of the open interval after each group
  std::vector<size_t> gs;
  for (size_t k = 0; k < cand.size(); ++k)
    if (k == 0 || cand[k].t != cand[k - 1].t) gs.push_back(k);
  int before = 0;
  for (size_t g = 0; g < gs.size(); ++g) {
    int after = 0;
    if (g + 1 < gs.size()) {
      double p[3];
      at(0.5 * (cand[gs[g]].t + cand[gs[g + 1]].t), p);
      after = ses_classify(M, p);
    }
    if (after != before) {
      const SesCand& c = cand[gs[g]];
      double p[3], nrm[3];
      at(c.t, p);
      if (c.type == 0) {
        const double* a = M.xyz + 3 * c.id;
        for (int k = 0; k < 3; ++k) nrm[k] = (p[k] - a[k]) / M.r[c.id];
      } else if (c.type == 2) {
        const double* a = M.verts[c.id].data();
        for (int k = 0; k < 3; ++k) nrm[k] = (a[k] - p[k]) / M.rp;
      } else {
        double q[3];
        ses_circle_point(M.pairs[c.id], p, q);
        for (int k = 0; k < 3; ++k) nrm[k] = (q[k] - p[k]) / M.rp;
      }
      xs.push_back(c.t);
      for (int k = 0; k < 3; ++k) ns.push_back(nrm[k]);
    }
    before = after;
  }
}
}  // namespace

extern "C" {

// SES inside test of n points (1 inside, 0 outside) for balls (xyz, r) and probe rp.
int orc_ses_classify(int64_t na, const double* xyz, const double* r, double rp, int64_t npts, const double* pts,
                     int32_t* inside) {
  SesModel M{na, xyz, r, rp, {}, {}, {}};
  ses_build(M);
  for (int64_t k = 0; k < npts; ++k) inside[k] = ses_classify(M, pts + 3 * k);
  return 0;
}

// SES crossings of axis lines (uv: 2 transverse coordinates per line, ordered as the
// ray keys are). count[l] = crossings of line l; at most `cap` are written per line
// into inters[cap l ...] and normals[3 cap l ...]. Returns the number of SAS pairs
// and stores the number of exposed SAS vertices in *nvert.
int64_t orc_ses_rays(int64_t na, const double* xyz, const double* r, double rp, int axis, int64_t nlines,
                     const double* uv, int cap, int32_t* count, double* inters, double* normals, int64_t* nvert) {
  SesModel M{na, xyz, r, rp, {}, {}, {}};
  ses_build(M);
  if (nvert) *nvert = (int64_t)M.verts.size();
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t l = 0; l < nlines; ++l) {
    std::vector<double> xs, ns;
    ses_ray(M, axis, uv[2 * l], uv[2 * l + 1], xs, ns);
    count[l] = (int32_t)xs.size();
    for (size_t k = 0; k < xs.size() && (int)k < cap; ++k) {
      inters[(int64_t)cap * l + (int64_t)k] = xs[k];
      for (int q = 0; q < 3; ++q) normals[3 * ((int64_t)cap * l + (int64_t)k) + q] = ns[3 * k + q];
    }
  }
  return (int64_t)M.pairs.size();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// refine_surface (src/pb_class.cpp:1467-1745) after init_tmesh (:1289-1296), on a
// dense map of the finest cells holding the level of their leaf (a leaf of level s is
// an aligned cube of 2^s cells; its anchor is the cell whose coordinates are multiples
// of 2^s). Flags are is_in_ns_surf along z (:769-806): the number of crossings at or
// below the vertex on its z-line, modulo 2. Steps of a cycle, each decided on the
// state before it:
//   refine   leaves of level s >= 1 with regular corners on both sides split once;
//   balance  leaves of level s >= 2 with a cell of level < s - 1 in the layer around
//            them split, repeated until none does;
//   coarsen  complete families of 8 sibling leaves of level s < smax whose members all
//            have their regular corners on one side merge;
//   balance.
// A corner is regular when it is a corner of every leaf around it. The all-members
// rule for families is this restatement's reading of the tmesh_3d coarsening callback
// (bimpp, absent from the reference): parity unpinned against bimpp; pinned by the
// closed forms and the independent Python restatement in the tests.
namespace {

struct SurfGrid {
  int64_t n[3];
  double origin[3], hf;
  std::vector<int8_t> lev;
  std::map<std::pair<int64_t, int64_t>, std::vector<double>> zlines;  // (i, j) -> crossings

  int8_t& at(int64_t x, int64_t y, int64_t z) { return lev[(size_t)((z * n[1] + y) * n[0] + x)]; }

  int flag(int64_t x, int64_t y, int64_t z) const {
    const auto it = zlines.find({x, y});
    if (it == zlines.end()) return 0;
    const double zc = origin[2] + (double)z * hf;
    int cnt = 0;
    for (double t : it->second)
      if (zc >= t) ++cnt;
    return cnt % 2;
  }

  bool regular(int64_t px, int64_t py, int64_t pz) {
    for (int64_t z = pz - 1; z <= pz; ++z)
      for (int64_t y = py - 1; y <= py; ++y)
        for (int64_t x = px - 1; x <= px; ++x) {
          if (x < 0 || y < 0 || z < 0 || x >= n[0] || y >= n[1] || z >= n[2]) continue;
          const int64_t S = (int64_t)1 << at(x, y, z);
          if (px % S != 0 || py % S != 0 || pz % S != 0) return false;
        }
    return true;
  }

  // true when the regular corners of cube (x, y, z, edge S) are all on one side
  bool uniform(int64_t x, int64_t y, int64_t z, int64_t S) {
    int seen0 = 0, seen1 = 0;
    for (int k = 0; k < 8; ++k) {
      const int64_t px = x + ((k & 1) ? S : 0), py = y + ((k & 2) ? S : 0), pz = z + ((k & 4) ? S : 0);
      if (!regular(px, py, pz)) continue;
      if (flag(px, py, pz)) seen1 = 1;
      else seen0 = 1;
    }
    return !(seen0 && seen1);
  }

  void set_cube(int64_t x, int64_t y, int64_t z, int64_t S, int8_t v) {
    for (int64_t c = z; c < z + S; ++c)
      for (int64_t b = y; b < y + S; ++b)
        for (int64_t a = x; a < x + S; ++a) at(a, b, c) = v;
  }

  template <class F>
  void each_leaf(F&& f) {  // anchor and level of every leaf
    for (int64_t z = 0; z < n[2]; ++z)
      for (int64_t y = 0; y < n[1]; ++y)
        for (int64_t x = 0; x < n[0]; ++x) {
          const int s = at(x, y, z);
          const int64_t S = (int64_t)1 << s;
          if (x % S == 0 && y % S == 0 && z % S == 0) f(x, y, z, s);
        }
  }

  void balance() {
    for (;;) {
      std::vector<std::array<int64_t, 4>> split;
      each_leaf([&](int64_t x, int64_t y, int64_t z, int s) {
        if (s < 2) return;
        const int64_t S = (int64_t)1 << s;
        bool bad = false;
        for (int64_t c = z - 1; c <= z + S && !bad; ++c)
          for (int64_t b = y - 1; b <= y + S && !bad; ++b)
            for (int64_t a = x - 1; a <= x + S && !bad; ++a) {
              if (a < 0 || b < 0 || c < 0 || a >= n[0] || b >= n[1] || c >= n[2]) continue;
              if (a >= x && a < x + S && b >= y && b < y + S && c >= z && c < z + S) continue;
              if (at(a, b, c) < s - 1) bad = true;
            }
        if (bad) split.push_back({x, y, z, s});
      });
      if (split.empty()) return;
      for (const auto& l : split) set_cube(l[0], l[1], l[2], (int64_t)1 << l[3], (int8_t)(l[3] - 1));
    }
  }
};

}  // namespace

extern "C" {

// Leaves of the refine_surface mesh of the box origin + [0, n] hf: z-line rays (nr
// lines, key[2r] = x, key[2r+1] = y of line r, crossings inters[ptr[r] .. ptr[r+1]));
// out[4 l ..] = anchor x, y, z (finest cells) and level of leaf l, in anchor order
// (x fastest); returns the leaf count (out may be NULL to count), -1 for bad input.
int64_t orc_refine_surface(const int64_t* n, const double* origin, double hf, int smax, int s_uni, int cycles,
                           int64_t nr, const double* key, const int64_t* ptr, const double* inters, int32_t* out) {
  if (smax < 0 || s_uni < 0 || s_uni > smax || hf <= 0) return -1;
  SurfGrid G;
  for (int d = 0; d < 3; ++d) {
    G.n[d] = n[d];
    G.origin[d] = origin[d];
    if (n[d] <= 0 || n[d] % ((int64_t)1 << smax) != 0) return -1;
  }
  G.hf = hf;
  for (int64_t r = 0; r < nr; ++r) {
    const int64_t i = std::llround((key[2 * r] - origin[0]) / hf), j = std::llround((key[2 * r + 1] - origin[1]) / hf);
    G.zlines[{i, j}] = std::vector<double>(inters + ptr[r], inters + ptr[r + 1]);
  }
  G.lev.assign((size_t)(n[0] * n[1] * n[2]), (int8_t)s_uni);
  for (int cyc = 0; cyc < cycles; ++cyc) {
    std::vector<std::array<int64_t, 4>> split;
    G.each_leaf([&](int64_t x, int64_t y, int64_t z, int s) {
      if (s > 0 && !G.uniform(x, y, z, (int64_t)1 << s)) split.push_back({x, y, z, s});
    });
    for (const auto& l : split) G.set_cube(l[0], l[1], l[2], (int64_t)1 << l[3], (int8_t)(l[3] - 1));
    G.balance();
    std::vector<std::array<int64_t, 4>> merge;
    G.each_leaf([&](int64_t x, int64_t y, int64_t z, int s) {
      if (s >= smax) return;
      const int64_t S = (int64_t)1 << s, P = 2 * S;
      if (x % P || y % P || z % P) return;
      for (int k = 0; k < 8; ++k) {
        const int64_t a = x + ((k & 1) ? S : 0), b = y + ((k & 2) ? S : 0), c = z + ((k & 4) ? S : 0);
        if (G.at(a, b, c) != s || !G.uniform(a, b, c, S)) return;
      }
      merge.push_back({x, y, z, s});
    });
    for (const auto& l : merge) G.set_cube(l[0], l[1], l[2], (int64_t)2 << l[3], (int8_t)(l[3] + 1));
    G.balance();
  }
  int64_t L = 0;
  G.each_leaf([&](int64_t x, int64_t y, int64_t z, int s) {
    if (out) {
      out[4 * L] = (int32_t)x;
      out[4 * L + 1] = (int32_t)y;
      out[4 * L + 2] = (int32_t)z;
      out[4 * L + 3] = s;
    }
    ++L;
  });
  return L;
}

}  // extern "C"
