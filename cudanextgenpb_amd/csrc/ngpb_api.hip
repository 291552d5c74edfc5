// C ABI of the linearized-PB library (declarations and reference mapping in
// include/ngpb.h). Host-side argument checks, state ordering and error reporting;
// all numerical work runs in the kernels of the other translation units.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <sstream>
#include <string>
#include <vector>

#include "ngpb_internal.h"

namespace ngpb {

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
  // clear the thread's last-error slot, so that a failed allocation (which leaves the
  // context usable) is not reported again by the next launch check
  (void)hipGetLastError();
  std::ostringstream os;
  os << "HIP error " << hipGetErrorName(e) << " (" << hipGetErrorString(e) << ") in " << what << " at " << file << ":"
     << line;
  throw Error(e == hipErrorOutOfMemory ? NGPB_E_NOMEM : NGPB_E_HIP, os.str());
}

PhysConst physical_constants(const ngpb_physics& p) {
  // include/pb_class.h:47-54; src/pb_class.cpp:1800-1806 and 2803-2814
  constexpr double e0 = 8.85418781762e-12, kb = 1.380649e-23, e = 1.602176634e-19, Nav = 6.022e23, Angs = 1e-10,
                   pi = 3.14159265358979323846;
  PhysConst c;
  c.eps_in = 4.0 * pi * e0 * p.e_in * kb * p.T * Angs / (e * e);
  c.eps_out = 4.0 * pi * e0 * p.e_out * kb * p.T * Angs / (e * e);
  const double C0 = 1.0e3 * Nav * p.ionic_strength;
  c.k2 = 2.0 * C0 * Angs * Angs * e * e / (e0 * p.e_out * kb * p.T);
  c.k = std::sqrt(c.k2);
  const double inv_4pi = 1.0 / (4.0 * pi);
  c.constant_pol = (1.0 / c.eps_out - 1.0 / c.eps_in) * inv_4pi;
  c.constant_react = (1.0 / c.eps_out) * inv_4pi;
  return c;
}

void run_markers(ngpb_ctx* c);

// the mesh arrays (node_xyz, hang_*, leaf_corner) and the Dirichlet node list
// (bnd_node, is_bnd) were written on the device; records the sizes as ngpb_set_mesh
// and ngpb_set_dirichlet would, with the boundary values still to come
void install_device_mesh(ngpb_ctx* c, int64_t N, int64_t H, int64_t L, int64_t NB) {
  c->N = N;
  c->H = H;
  c->L = L;
  c->NB = NB;
  if (!NB) c->bnd_node.resize(1);
  c->gval.resize((size_t)N);
  NGPB_HIP(hipMemsetAsync(c->gval.p, 0, (size_t)N * sizeof(double), c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  c->have_mesh = true;
  c->have_dirichlet = true;
  c->have_bnd_values = NB == 0;
  c->bnd_bc_type = 0;
  c->have_charges = false;
  c->have_markers = c->have_frac = c->have_symbolic = c->have_system = c->have_compact = c->have_phi = false;
}

}  // namespace ngpb

int64_t ngpb_ctx::device_bytes() const {
  int64_t b = 0;
  b += node_xyz.bytes() + hang_xyz.bytes() + hang_np.bytes() + hang_par.bytes() + leaf_corner.bytes();
  for (const auto& r : rays) b += r.key.bytes() + r.inters.bytes() + r.normals.bytes() + r.ptr.bytes();
  b += eps.bytes() + reac.bytes() + leaf_marker.bytes() + border.bytes() + border_of.bytes();
  b += frac.bytes() + frac_normal.bytes() + frac_found.bytes();
  b += atom_xyz.bytes() + atom_q.bytes() + atom_owner.bytes();
  b += bnd_node.bytes() + bnd_val.bytes() + is_bnd.bytes() + gval.bytes();
  b += em_id.bytes() + em_code.bytes() + slot_rel.bytes() + blk_em.bytes() + diag_slot.bytes() + bc_rowptr.bytes() +
       bc_col.bytes() + bc_ptr.bytes() + bc_id.bytes() + bc_code.bytes();
  b += rowptr.bytes() + col.bytes() + val.bytes() + rhs.bytes() + rhs_charge.bytes() + c_edge.bytes() + v8.bytes() +
       vol_patch.bytes() + rho.bytes();
  b += slice_width.bytes() + slice_base_ptr.bytes() + slice_wide_ptr.bytes() + slice_ov_ptr.bytes() +
       sell_val.bytes() + sell_off.bytes() + sell_base.bytes() + sell_col32.bytes() + sell_ov_cnt.bytes() +
       sell_ov_val.bytes() + sell_ov_col.bytes() + dscale.bytes() + sell_pk_val.bytes() + sell_pk_idx.bytes() +
       sell_pk_lp.bytes() + sell_pk_dict.bytes() + sell_pk_hdr.bytes() + sell_cval.bytes() + sell_list_delta.bytes() +
       sell_list_packed.bytes() + sell_list_offset.bytes() + sell_dpair.bytes() + sell_pk_hdr_list.bytes();
  b += x.bytes() + r.bytes() + z.bytes() + p.bytes() + q.bytes() + dinv.bytes();
  b += lis_rt.bytes() + lis_u.bytes() + lis_v.bytes() + lis_ph.bytes() + lis_uh.bytes() + lis_t.bytes() +
       diag_pos.bytes() + lev_rows_f.bytes() + lev_rows_b.bytes();
  b += xs.bytes() + rs.bytes() + ps.bytes() + qs.bytes() + ks.bytes();
  b += red.partials.bytes() + scan_tmp.bytes();
  for (const auto& t : tmp) b += t.bytes();
  return b;
}

namespace {

using ngpb::Error;

// at the end of an API call the context's stream keeps a bounded set of cached blocks
struct CallEnd {
  hipStream_t s;
  ~CallEnd() {
    if (s) ngpb::stream_cache_end_call(s);
  }
};

template <class F>
int guarded(ngpb_ctx* c, F&& f) {
  if (!c) return NGPB_E_ARG;
  c->prof_kind = -1;  // no SpMV timing left open by a call that failed inside one
  CallEnd end{c->stream};
  try {
    NGPB_HIP(hipSetDevice(c->device));
    f();
    c->err.clear();
    return NGPB_OK;
  } catch (const Error& e) {
    c->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    c->err = "host allocation failed";
    return NGPB_E_NOMEM;
  } catch (const std::exception& e) {
    c->err = e.what();
    return NGPB_E_HIP;
  }
}

void invalidate_after_mesh(ngpb_ctx* c) {
  c->have_markers = c->have_frac = c->have_symbolic = c->have_system = c->have_compact = c->have_phi = false;
}

void need(bool cond, const char* what) {
  if (!cond) throw Error(NGPB_E_STATE, what);
}

// a context in a batch lends its operator and vectors to the batch's kernels
void not_in_batch(const ngpb_ctx* c) {
  if (c->in_batch) throw Error(NGPB_E_STATE, "context is in a batch (ngpb_batch_remove it first)");
}

}  // namespace

namespace ngpb {

// ---- per-stream cache of released stream-ordered blocks
// Bounds: a stream caches at most kCacheMax bytes during a call; at the end of an API
// call it keeps at most kCacheKeep bytes, and none once the streams of all contexts
// together keep more than kGlobalKeep (512 contexts of a batch each keeping blocks
// between calls would hold the device's memory).
namespace {
constexpr size_t kCacheMax = size_t(1) << 30;
constexpr size_t kCacheKeep = size_t(256) << 20;
constexpr size_t kGlobalKeep = size_t(1) << 30;
struct StreamBlocks {
  std::multimap<size_t, void*> free;  // by size
  size_t bytes = 0;
};
std::mutex g_cache_mu;
std::unordered_map<hipStream_t, StreamBlocks> g_cache;
size_t g_cached = 0;  // bytes cached over all streams
}  // namespace

void* stream_cache_take(hipStream_t s, size_t bytes, size_t* cap) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  auto it = g_cache.find(s);
  if (it == g_cache.end()) return nullptr;
  StreamBlocks& b = it->second;
  auto f = b.free.lower_bound(bytes);
  // a block at most twice (plus 1 MiB) the request: no large block for a small buffer
  if (f == b.free.end() || f->first > 2 * bytes + (size_t(1) << 20)) return nullptr;
  void* p = f->second;
  *cap = f->first;
  b.bytes -= f->first;
  g_cached -= f->first;
  b.free.erase(f);
  return p;
}

bool stream_cache_put(hipStream_t s, void* p, size_t cap) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  StreamBlocks& b = g_cache[s];
  if (b.bytes + cap > kCacheMax) return false;
  b.free.emplace(cap, p);
  b.bytes += cap;
  g_cached += cap;
  return true;
}

void stream_cache_trim(hipStream_t s, size_t keep) {
  std::vector<void*> out;
  {
    std::lock_guard<std::mutex> g(g_cache_mu);
    auto it = g_cache.find(s);
    if (it == g_cache.end()) return;
    StreamBlocks& b = it->second;
    if (keep > 0 && g_cached > kGlobalKeep) keep = 0;
    while (b.bytes > keep && !b.free.empty()) {  // largest first
      auto f = std::prev(b.free.end());
      out.push_back(f->second);
      b.bytes -= f->first;
      g_cached -= f->first;
      b.free.erase(f);
    }
  }
  for (void* p : out) (void)hipFreeAsync(p, s);
}

void stream_cache_end_call(hipStream_t s) { stream_cache_trim(s, kCacheKeep); }

void stream_cache_drop(hipStream_t s) {
  stream_cache_trim(s, 0);
  std::lock_guard<std::mutex> g(g_cache_mu);
  g_cache.erase(s);
}

}  // namespace ngpb

extern "C" {

const char* ngpb_version(void) { return "ngpb-gfx950 0.1"; }

int ngpb_device_count(int* count) {
  if (!count) return NGPB_E_ARG;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    return NGPB_E_NODEV;
  }
  *count = n;
  return NGPB_OK;
}

int ngpb_create(ngpb_ctx** out, int device) {
  if (!out) return NGPB_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return NGPB_E_NODEV;
  }
  if (device < 0 || device >= n) return NGPB_E_ARG;
  ngpb_ctx* c = new (std::nothrow) ngpb_ctx();
  if (!c) return NGPB_E_NOMEM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return NGPB_E_HIP;
  }
  ngpb_physics ph{2.0, 80.0, 298.15, 0.145, 2};
  c->physics = ph;
  c->pc = ngpb::physical_constants(ph);
  *out = c;
  return NGPB_OK;
}

int ngpb_destroy(ngpb_ctx* c) {
  if (!c) return NGPB_OK;
  if (c->in_batch) {
    c->err = "context is in a batch (ngpb_batch_remove it first)";
    return NGPB_E_STATE;
  }
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->ks_host) (void)hipHostFree(c->ks_host);
  for (auto& p : c->kprof)
    for (hipEvent_t e : p.ev) (void)hipEventDestroy(e);
  c->ks_host = nullptr;
  if (c->stream) {
    ngpb::stream_cache_drop(c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  delete c;  // device buffers free themselves
  return NGPB_OK;
}

const char* ngpb_last_error(const ngpb_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ngpb_constants(const ngpb_physics* p, double* out6) {
  if (!p || !out6) return NGPB_E_ARG;
  const ngpb::PhysConst c = ngpb::physical_constants(*p);
  out6[0] = c.eps_in;
  out6[1] = c.eps_out;
  out6[2] = c.k2;
  out6[3] = c.k;
  out6[4] = c.constant_pol;
  out6[5] = c.constant_react;
  return NGPB_OK;
}

int ngpb_device_bytes(const ngpb_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return NGPB_E_ARG;
  *bytes = c->device_bytes();
  return NGPB_OK;
}

int ngpb_set_physics(ngpb_ctx* c, const ngpb_physics* p) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(p, NGPB_E_ARG, "physics is NULL");
    NGPB_REQUIRE(p->e_in > 0 && p->e_out > 0 && p->T > 0 && p->ionic_strength >= 0, NGPB_E_ARG,
                 "permittivities and temperature must be positive, ionic strength non-negative");
    NGPB_REQUIRE(p->calc_energy >= 0 && p->calc_energy <= 2, NGPB_E_ARG, "calc_energy must be 0, 1 or 2");
    NGPB_REQUIRE(p->calc_coulombic == 0 || p->calc_coulombic == 1, NGPB_E_ARG, "calc_coulombic must be 0 or 1");
    c->physics = *p;
    c->pc = ngpb::physical_constants(*p);
    c->have_physics = true;
    c->have_markers = c->have_system = c->have_compact = c->have_phi = false;
    if (c->bnd_bc_type) c->have_bnd_values = false;  // recomputed at assembly
  });
}

int ngpb_set_mesh(ngpb_ctx* c, int64_t n_nodes, const double* node_xyz, int64_t n_hang, const double* hang_xyz,
                  const int32_t* hang_np, const int32_t* hang_par, int64_t n_leaves, const int32_t* leaf_corner) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(n_nodes > 0 && n_leaves > 0 && n_hang >= 0, NGPB_E_ARG, "mesh sizes must be positive");
    NGPB_REQUIRE(n_nodes < INT32_MAX && n_hang < INT32_MAX && n_leaves < INT32_MAX, NGPB_E_ARG, "mesh too large");
    NGPB_REQUIRE(node_xyz && leaf_corner, NGPB_E_ARG, "NULL mesh array");
    NGPB_REQUIRE(n_hang == 0 || (hang_xyz && hang_np && hang_par), NGPB_E_ARG, "NULL hanging-node array");
    for (int64_t k = 0; k < n_hang; ++k) {
      NGPB_REQUIRE(hang_np[k] == 2 || hang_np[k] == 4, NGPB_E_ARG, "hanging nodes need 2 or 4 parents");
      for (int j = 0; j < hang_np[k]; ++j)
        NGPB_REQUIRE(hang_par[4 * k + j] >= 0 && hang_par[4 * k + j] < n_nodes, NGPB_E_ARG, "hanging parent out of range");
    }
    std::vector<uint8_t> seen((size_t)n_nodes, 0);
    for (int64_t i = 0; i < 8 * n_leaves; ++i) {
      const int32_t v = leaf_corner[i];
      NGPB_REQUIRE(v < n_nodes && v >= -n_hang, NGPB_E_ARG, "leaf corner index out of range");
      if (v >= 0) seen[v] = 1;
    }
    for (int64_t i = 0; i < n_nodes; ++i)
      NGPB_REQUIRE(seen[i], NGPB_E_ARG, "every node must be a (non-hanging) corner of some leaf");
    for (int64_t i = 0; i < 3 * n_nodes; ++i) NGPB_REQUIRE(std::isfinite(node_xyz[i]), NGPB_E_ARG, "non-finite node coordinate");
    c->N = n_nodes;
    c->H = n_hang;
    c->L = n_leaves;
    c->node_xyz.upload(node_xyz, (size_t)(3 * n_nodes), c->stream);
    if (n_hang) {
      c->hang_xyz.upload(hang_xyz, (size_t)(3 * n_hang), c->stream);
      c->hang_np.upload(hang_np, (size_t)n_hang, c->stream);
      c->hang_par.upload(hang_par, (size_t)(4 * n_hang), c->stream);
    } else {
      c->hang_xyz.resize(3);
      c->hang_np.resize(1);
      c->hang_par.resize(4);
    }
    c->leaf_corner.upload(leaf_corner, (size_t)(8 * n_leaves), c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
    c->have_mesh = true;
    c->have_dirichlet = c->have_bnd_values = c->have_charges = false;
    c->bnd_bc_type = 0;
    invalidate_after_mesh(c);
  });
}

int ngpb_build_octree(ngpb_ctx* c, const double* origin, double hf, const int64_t* n, int smax, const double* band,
                      int64_t n_balls, const double* xyz, const double* r, int64_t* sizes) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(origin && n, NGPB_E_ARG, "NULL origin or n");
    NGPB_REQUIRE(smax >= 0 && smax <= 10, NGPB_E_ARG, "smax must be in 0..10");
    NGPB_REQUIRE(std::isfinite(hf) && hf > 0, NGPB_E_ARG, "hf must be finite and > 0");
    NGPB_REQUIRE(smax == 0 || band, NGPB_E_ARG, "NULL band");
    for (int d = 0; d < 3; ++d) {
      NGPB_REQUIRE(std::isfinite(origin[d]), NGPB_E_ARG, "non-finite origin");
      NGPB_REQUIRE(n[d] > 0 && n[d] <= ((int64_t)1 << 21) && n[d] % ((int64_t)1 << smax) == 0, NGPB_E_ARG,
                   "n must be positive multiples of 2^smax, at most 2^21");
    }
    NGPB_REQUIRE((double)(n[0] + 1) * (double)(n[1] + 1) * (double)(n[2] + 1) < 4294967296.0, NGPB_E_ARG,
                 "at most 2^32 - 1 grid vertices");
    for (int s = 0; s < smax; ++s) NGPB_REQUIRE(std::isfinite(band[s]), NGPB_E_ARG, "non-finite band");
    for (int s = 1; s < smax; ++s) NGPB_REQUIRE(band[s] >= band[s - 1], NGPB_E_ARG, "band must be non-decreasing");
    NGPB_REQUIRE(n_balls >= 0 && n_balls < INT32_MAX, NGPB_E_ARG, "bad ball count");
    NGPB_REQUIRE(n_balls == 0 || (xyz && r), NGPB_E_ARG, "NULL ball array");
    for (int64_t a = 0; a < n_balls; ++a)
      NGPB_REQUIRE(std::isfinite(r[a]) && r[a] >= 0 && std::isfinite(xyz[3 * a]) && std::isfinite(xyz[3 * a + 1]) &&
                       std::isfinite(xyz[3 * a + 2]),
                   NGPB_E_ARG, "ball centres and radii must be finite, radii >= 0");
    // a failed build leaves no mesh behind
    c->have_mesh = c->have_dirichlet = c->have_bnd_values = c->have_charges = false;
    invalidate_after_mesh(c);
    const ngpb::OctreeSizes z = ngpb::run_build_octree(c, origin, hf, n, smax, band, n_balls, xyz, r);
    if (sizes) {
      sizes[0] = z.N;
      sizes[1] = z.H;
      sizes[2] = z.L;
      sizes[3] = z.NB;
    }
  });
}

int ngpb_build_octree_surface(ngpb_ctx* c, const double* origin, double hf, const int64_t* n, int smax, int s_uni,
                              int cycles, int64_t n_balls, const double* xyz, const double* r, double probe,
                              int64_t* sizes) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(origin && n, NGPB_E_ARG, "NULL origin or n");
    NGPB_REQUIRE(smax >= 0 && smax <= 10, NGPB_E_ARG, "smax must be in 0..10");
    NGPB_REQUIRE(s_uni >= 0 && s_uni <= smax, NGPB_E_ARG, "s_uni must be in 0..smax");
    NGPB_REQUIRE(cycles <= 64, NGPB_E_ARG, "at most 64 cycles");
    NGPB_REQUIRE(std::isfinite(hf) && hf > 0, NGPB_E_ARG, "hf must be finite and > 0");
    NGPB_REQUIRE(std::isfinite(probe) && probe >= 0, NGPB_E_ARG, "probe must be finite and >= 0");
    for (int d = 0; d < 3; ++d) {
      NGPB_REQUIRE(std::isfinite(origin[d]), NGPB_E_ARG, "non-finite origin");
      NGPB_REQUIRE(n[d] > 0 && n[d] <= ((int64_t)1 << 21) && n[d] % ((int64_t)1 << smax) == 0, NGPB_E_ARG,
                   "n must be positive multiples of 2^smax, at most 2^21");
    }
    NGPB_REQUIRE((double)(n[0] + 1) * (double)(n[1] + 1) * (double)(n[2] + 1) < 4294967296.0, NGPB_E_ARG,
                 "at most 2^32 - 1 grid vertices");
    NGPB_REQUIRE((double)(n[0] + 1) * (double)(n[1] + 1) < 2147483647.0, NGPB_E_ARG, "too many grid z-lines");
    NGPB_REQUIRE(n_balls >= 0 && n_balls < INT32_MAX, NGPB_E_ARG, "bad ball count");
    NGPB_REQUIRE(n_balls == 0 || (xyz && r), NGPB_E_ARG, "NULL ball array");
    for (int64_t a = 0; a < n_balls; ++a)
      NGPB_REQUIRE(std::isfinite(r[a]) && r[a] >= 0 && std::isfinite(xyz[3 * a]) && std::isfinite(xyz[3 * a + 1]) &&
                       std::isfinite(xyz[3 * a + 2]),
                   NGPB_E_ARG, "ball centres and radii must be finite, radii >= 0");
    c->have_mesh = c->have_dirichlet = c->have_bnd_values = c->have_charges = false;
    invalidate_after_mesh(c);
    const int cyc = cycles < 0 ? std::max(s_uni, smax - s_uni) : cycles;
    const ngpb::OctreeSizes z =
        ngpb::run_build_octree_surface(c, origin, hf, n, smax, s_uni, cyc, n_balls, xyz, r, probe);
    if (sizes) {
      sizes[0] = z.N;
      sizes[1] = z.H;
      sizes[2] = z.L;
      sizes[3] = z.NB;
    }
  });
}

int ngpb_get_mesh(ngpb_ctx* c, int64_t* sizes, double* node_xyz, double* hang_xyz, int32_t* hang_np,
                  int32_t* hang_par, int32_t* leaf_corner, int32_t* bnd_nodes) {
  return guarded(c, [&] {
    need(c->have_mesh, "no mesh");
    NGPB_REQUIRE(sizes, NGPB_E_ARG, "NULL sizes");
    const int64_t nb = c->have_dirichlet ? c->NB : 0;
    sizes[0] = c->N;
    sizes[1] = c->H;
    sizes[2] = c->L;
    sizes[3] = nb;
    if (node_xyz) c->node_xyz.download(node_xyz, (size_t)(3 * c->N), c->stream);
    if (c->H) {
      if (hang_xyz) c->hang_xyz.download(hang_xyz, (size_t)(3 * c->H), c->stream);
      if (hang_np) c->hang_np.download(hang_np, (size_t)c->H, c->stream);
      if (hang_par) c->hang_par.download(hang_par, (size_t)(4 * c->H), c->stream);
    }
    if (leaf_corner) c->leaf_corner.download(leaf_corner, (size_t)(8 * c->L), c->stream);
    if (bnd_nodes && nb) c->bnd_node.download(bnd_nodes, (size_t)nb, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
  });
}

int ngpb_set_rays(ngpb_ctx* c, int axis, int64_t n_rays, const double* key, const int64_t* ptr, const double* inters,
                  const double* normals) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(axis >= 0 && axis < 3, NGPB_E_ARG, "axis must be 0, 1 or 2");
    NGPB_REQUIRE(n_rays >= 0, NGPB_E_ARG, "negative ray count");
    auto& r = c->rays[axis];
    if (n_rays == 0) {
      r.n = r.k = 0;
      r.key.resize(2);
      r.ptr.resize(1);
      r.inters.resize(1);
      r.normals.resize(3);
      const int64_t zero = 0;
      r.ptr.upload(&zero, 1, c->stream);
    } else {
      NGPB_REQUIRE(key && ptr, NGPB_E_ARG, "NULL ray array");
      NGPB_REQUIRE(ptr[0] == 0, NGPB_E_ARG, "ray pointers must start at 0");
      for (int64_t i = 0; i < n_rays; ++i)
        NGPB_REQUIRE(ptr[i + 1] >= ptr[i], NGPB_E_ARG, "ray pointers must be non-decreasing");
      const int64_t K = ptr[n_rays];
      NGPB_REQUIRE(K == 0 || (inters && normals), NGPB_E_ARG, "NULL crossing array");
      for (int64_t i = 0; i < n_rays; ++i) {
        if (i > 0)
          NGPB_REQUIRE(key[2 * i - 2] < key[2 * i] || (key[2 * i - 2] == key[2 * i] && key[2 * i - 1] < key[2 * i + 1]),
                       NGPB_E_ARG, "ray keys must be strictly increasing");
        for (int64_t k = ptr[i] + 1; k < ptr[i + 1]; ++k)
          NGPB_REQUIRE(inters[k - 1] <= inters[k], NGPB_E_ARG, "crossings must be sorted within a ray");
      }
      r.n = n_rays;
      r.k = K;
      r.key.upload(key, (size_t)(2 * n_rays), c->stream);
      r.ptr.upload(ptr, (size_t)(n_rays + 1), c->stream);
      if (K) {
        r.inters.upload(inters, (size_t)K, c->stream);
        r.normals.upload(normals, (size_t)(3 * K), c->stream);
      } else {
        r.inters.resize(1);
        r.normals.resize(3);
      }
    }
    NGPB_HIP(hipStreamSynchronize(c->stream));
    c->have_markers = c->have_frac = c->have_system = c->have_compact = c->have_phi = false;
  });
}

int ngpb_cast_rays(ngpb_ctx* c, int64_t n_balls, const double* xyz, const double* r) {
  return guarded(c, [&] {
    not_in_batch(c);
    need(c->have_mesh, "ngpb_set_mesh must come before ngpb_cast_rays");
    NGPB_REQUIRE(n_balls >= 0 && n_balls < INT32_MAX, NGPB_E_ARG, "bad ball count");
    NGPB_REQUIRE(n_balls == 0 || (xyz && r), NGPB_E_ARG, "NULL ball array");
    for (int64_t a = 0; a < n_balls; ++a)
      NGPB_REQUIRE(std::isfinite(r[a]) && r[a] >= 0 && std::isfinite(xyz[3 * a]) && std::isfinite(xyz[3 * a + 1]) &&
                       std::isfinite(xyz[3 * a + 2]),
                   NGPB_E_ARG, "ball centres and radii must be finite, radii >= 0");
    ngpb::run_cast_rays(c, n_balls, xyz, r, 0.0);
    c->have_markers = c->have_frac = c->have_system = c->have_compact = c->have_phi = false;
  });
}

int ngpb_cast_rays_ses(ngpb_ctx* c, int64_t n_balls, const double* xyz, const double* r, double probe_radius) {
  return guarded(c, [&] {
    not_in_batch(c);
    need(c->have_mesh, "ngpb_set_mesh must come before ngpb_cast_rays_ses");
    NGPB_REQUIRE(n_balls >= 0 && n_balls < (int64_t)1 << 28, NGPB_E_ARG, "bad ball count");
    NGPB_REQUIRE(std::isfinite(probe_radius) && probe_radius > 0.0, NGPB_E_ARG, "probe radius must be finite and > 0");
    NGPB_REQUIRE(n_balls == 0 || (xyz && r), NGPB_E_ARG, "NULL ball array");
    for (int64_t a = 0; a < n_balls; ++a)
      NGPB_REQUIRE(std::isfinite(r[a]) && r[a] >= 0 && std::isfinite(xyz[3 * a]) && std::isfinite(xyz[3 * a + 1]) &&
                       std::isfinite(xyz[3 * a + 2]),
                   NGPB_E_ARG, "ball centres and radii must be finite, radii >= 0");
    ngpb::run_cast_rays(c, n_balls, xyz, r, probe_radius);
    c->have_markers = c->have_frac = c->have_system = c->have_compact = c->have_phi = false;
  });
}

int ngpb_locate_atoms(ngpb_ctx* c, int64_t n_atoms, const double* xyz, double tol, int32_t* owner_leaf) {
  return guarded(c, [&] {
    need(c->have_mesh, "ngpb_set_mesh must come before ngpb_locate_atoms");
    NGPB_REQUIRE(n_atoms >= 0 && (n_atoms == 0 || (xyz && owner_leaf)), NGPB_E_ARG, "bad atom arrays");
    NGPB_REQUIRE(std::isfinite(tol) && tol >= 0, NGPB_E_ARG, "tolerance must be finite and >= 0");
    ngpb::run_locate_atoms(c, n_atoms, xyz, tol, owner_leaf);
  });
}

int ngpb_get_rays(ngpb_ctx* c, int axis, int64_t* sizes, double* key, int64_t* ptr, double* inters, double* normals) {
  return guarded(c, [&] {
    NGPB_REQUIRE(axis >= 0 && axis < 3 && sizes, NGPB_E_ARG, "bad arguments");
    const auto& R = c->rays[axis];
    sizes[0] = R.n;
    sizes[1] = R.k;
    if (key && R.n) R.key.download(key, (size_t)(2 * R.n), c->stream);
    if (ptr) R.ptr.download(ptr, (size_t)(R.n + 1), c->stream);
    if (inters && R.k) R.inters.download(inters, (size_t)R.k, c->stream);
    if (normals && R.k) R.normals.download(normals, (size_t)(3 * R.k), c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
  });
}

int ngpb_create_markers(ngpb_ctx* c) {
  return guarded(c, [&] {
    not_in_batch(c);
    need(c->have_mesh, "ngpb_set_mesh must come before ngpb_create_markers");
    for (int d = 0; d < 3; ++d)
      if (!c->rays[d].ptr.p) {
        const int64_t zero = 0;
        c->rays[d].ptr.upload(&zero, 1, c->stream);
        c->rays[d].key.resize(2);
        c->rays[d].inters.resize(1);
        c->rays[d].normals.resize(3);
      }
    ngpb::run_markers(c);
    c->have_markers = true;
    c->have_frac = c->have_system = c->have_compact = c->have_phi = false;
  });
}

int ngpb_set_markers(ngpb_ctx* c, const double* eps_node, const double* reac_node, const uint8_t* leaf_marker) {
  return guarded(c, [&] {
    not_in_batch(c);
    need(c->have_mesh, "ngpb_set_mesh must come before ngpb_set_markers");
    NGPB_REQUIRE(eps_node && reac_node && leaf_marker, NGPB_E_ARG, "NULL marker array");
    for (int64_t i = 0; i < c->N; ++i)
      NGPB_REQUIRE(eps_node[i] > 0 && reac_node[i] >= 0 && std::isfinite(eps_node[i]) && std::isfinite(reac_node[i]),
                   NGPB_E_ARG, "eps must be positive and reaction non-negative");
    for (int64_t l = 0; l < c->L; ++l) NGPB_REQUIRE(leaf_marker[l] <= 2, NGPB_E_ARG, "leaf marker must be 0, 1 or 2");
    c->eps.upload(eps_node, (size_t)c->N, c->stream);
    c->reac.upload(reac_node, (size_t)c->N, c->stream);
    c->leaf_marker.upload(leaf_marker, (size_t)c->L, c->stream);
    for (int d = 0; d < 3; ++d)
      if (!c->rays[d].ptr.p) {
        const int64_t zero = 0;
        c->rays[d].ptr.upload(&zero, 1, c->stream);
        c->rays[d].key.resize(2);
        c->rays[d].inters.resize(1);
        c->rays[d].normals.resize(3);
      }
    ngpb::run_border_from_markers(c);
    c->have_markers = true;
    c->have_frac = c->have_system = c->have_compact = c->have_phi = false;
  });
}

int ngpb_get_markers(ngpb_ctx* c, double* eps_node, double* reac_node, uint8_t* leaf_marker, int64_t* n_border,
                     int32_t* border) {
  return guarded(c, [&] {
    need(c->have_markers, "markers are not available");
    if (eps_node) c->eps.download(eps_node, (size_t)c->N, c->stream);
    if (reac_node) c->reac.download(reac_node, (size_t)c->N, c->stream);
    if (leaf_marker) c->leaf_marker.download(leaf_marker, (size_t)c->L, c->stream);
    if (n_border) *n_border = c->n_border;
    if (border && c->n_border) c->border.download(border, (size_t)c->n_border, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
  });
}

int ngpb_get_fractions(ngpb_ctx* c, double* frac) {
  return guarded(c, [&] {
    NGPB_REQUIRE(frac, NGPB_E_ARG, "NULL output");
    need(c->have_markers, "markers are not available");
    if (!c->have_frac) ngpb::run_fractions(c);
    std::vector<double> fb((size_t)12 * c->n_border);
    std::vector<int32_t> bl((size_t)c->n_border);
    if (c->n_border) {
      c->frac.download(fb.data(), fb.size(), c->stream);
      c->border.download(bl.data(), bl.size(), c->stream);
    }
    NGPB_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < 12 * c->L; ++i) frac[i] = 0.5;
    for (int64_t b = 0; b < c->n_border; ++b)
      for (int e = 0; e < 12; ++e) frac[12 * (int64_t)bl[b] + e] = fb[12 * b + e];
  });
}

int ngpb_set_charges(ngpb_ctx* c, int64_t n_atoms, const double* xyz, const double* q, const int32_t* owner) {
  return guarded(c, [&] {
    not_in_batch(c);
    need(c->have_mesh, "ngpb_set_mesh must come before ngpb_set_charges");
    NGPB_REQUIRE(n_atoms >= 0, NGPB_E_ARG, "negative atom count");
    NGPB_REQUIRE(n_atoms == 0 || (xyz && q && owner), NGPB_E_ARG, "NULL atom array");
    for (int64_t a = 0; a < n_atoms; ++a) {
      NGPB_REQUIRE(owner[a] >= -1 && owner[a] < c->L, NGPB_E_ARG, "owner leaf out of range");
      NGPB_REQUIRE(std::isfinite(q[a]) && std::isfinite(xyz[3 * a]) && std::isfinite(xyz[3 * a + 1]) &&
                       std::isfinite(xyz[3 * a + 2]),
                   NGPB_E_ARG, "non-finite atom data");
    }
    c->A = n_atoms;
    if (n_atoms) {
      c->atom_xyz.upload(xyz, (size_t)(3 * n_atoms), c->stream);
      c->atom_q.upload(q, (size_t)n_atoms, c->stream);
      c->atom_owner.upload(owner, (size_t)n_atoms, c->stream);
    } else {
      c->atom_xyz.resize(3);
      c->atom_q.resize(1);
      c->atom_owner.resize(1);
    }
    NGPB_HIP(hipStreamSynchronize(c->stream));
    c->have_charges = true;
    c->have_system = c->have_compact = c->have_phi = false;
    if (c->bnd_bc_type) c->have_bnd_values = false;  // recomputed at assembly
  });
}

int ngpb_set_dirichlet(ngpb_ctx* c, int64_t n_bnd, const int32_t* nodes, const double* values) {
  return guarded(c, [&] {
    not_in_batch(c);
    need(c->have_mesh, "ngpb_set_mesh must come before ngpb_set_dirichlet");
    NGPB_REQUIRE(n_bnd >= 0 && (n_bnd == 0 || nodes), NGPB_E_ARG, "bad Dirichlet list");
    std::vector<uint8_t> isb((size_t)c->N, 0);
    std::vector<double> g((size_t)c->N, 0.0);
    for (int64_t k = 0; k < n_bnd; ++k) {
      NGPB_REQUIRE(nodes[k] >= 0 && nodes[k] < c->N, NGPB_E_ARG, "Dirichlet node out of range");
      NGPB_REQUIRE(!isb[nodes[k]], NGPB_E_ARG, "duplicate Dirichlet node");
      isb[nodes[k]] = 1;
      if (values) {
        NGPB_REQUIRE(std::isfinite(values[k]), NGPB_E_ARG, "non-finite Dirichlet value");
        g[nodes[k]] = values[k];
      }
    }
    c->NB = n_bnd;
    c->bnd_node.upload(n_bnd ? nodes : nullptr, (size_t)n_bnd, c->stream);
    if (!n_bnd) c->bnd_node.resize(1);
    c->is_bnd.upload(isb.data(), isb.size(), c->stream);
    c->gval.upload(g.data(), g.size(), c->stream);
    if (values && n_bnd) c->bnd_val.upload(values, (size_t)n_bnd, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
    c->have_dirichlet = true;
    c->have_bnd_values = values != nullptr || n_bnd == 0;
    c->bnd_bc_type = 0;
    c->have_symbolic = c->have_system = c->have_compact = c->have_phi = false;
  });
}

int ngpb_dirichlet_values(ngpb_ctx* c, int bc_type) {
  return guarded(c, [&] {
    not_in_batch(c);
    need(c->have_dirichlet, "ngpb_set_dirichlet must come before ngpb_dirichlet_values");
    ngpb::run_dirichlet_values(c, bc_type);
    c->bnd_bc_type = bc_type;
    c->have_system = c->have_compact = c->have_phi = false;
  });
}

int ngpb_assemble(ngpb_ctx* c) {
  return guarded(c, [&] {
    not_in_batch(c);
    need(c->have_mesh, "no mesh");
    need(c->have_markers, "markers must be created or set before assembly");
    if (!c->have_dirichlet) {
      c->NB = 0;
      std::vector<uint8_t> isb((size_t)c->N, 0);
      std::vector<double> g((size_t)c->N, 0.0);
      c->is_bnd.upload(isb.data(), isb.size(), c->stream);
      c->gval.upload(g.data(), g.size(), c->stream);
      c->bnd_node.resize(1);
      c->have_dirichlet = c->have_bnd_values = true;
      c->bnd_bc_type = 0;
    }
    // computed boundary values follow the current charges and physics
    if (!c->have_bnd_values && c->bnd_bc_type) ngpb::run_dirichlet_values(c, c->bnd_bc_type);
    need(c->have_bnd_values, "Dirichlet values missing (ngpb_dirichlet_values)");
    if (!c->have_frac) ngpb::run_fractions(c);
    // after ngpb_set_csr the CSR arrays no longer hold the mesh's pattern
    if (!c->have_symbolic || c->external_csr) {
      ngpb::run_symbolic(c);
      c->have_levels = false;
    }
    ngpb::run_numeric(c);
    c->external_csr = false;
    c->have_phi = false;
  });
}

int ngpb_assemble_many(ngpb_ctx* const* cs, int n) {
  if (n < 0 || (n > 0 && !cs)) return NGPB_E_ARG;
  if (n == 0) return NGPB_OK;
  std::vector<ngpb_ctx*> v(cs, cs + n);
  for (ngpb_ctx* c : v)
    if (!c) return NGPB_E_ARG;
  {
    std::vector<ngpb_ctx*> u(v);
    std::sort(u.begin(), u.end());
    if (std::adjacent_find(u.begin(), u.end()) != u.end()) {
      for (ngpb_ctx* c : v) c->err = "a context appears twice in ngpb_assemble_many";
      return NGPB_E_ARG;
    }
  }
  // each context's own preparation, as in ngpb_assemble
  std::vector<char> sym((size_t)n);
  for (int i = 0; i < n; ++i) {
    ngpb_ctx* c = v[i];
    const int rc = guarded(c, [&] {
      NGPB_REQUIRE(c->device == v[0]->device, NGPB_E_ARG, "ngpb_assemble_many needs contexts on one device");
      not_in_batch(c);
      need(c->have_mesh, "no mesh");
      need(c->have_markers, "markers must be created or set before assembly");
      if (!c->have_dirichlet) {
        c->NB = 0;
        std::vector<uint8_t> isb((size_t)c->N, 0);
        std::vector<double> g((size_t)c->N, 0.0);
        c->is_bnd.upload(isb.data(), isb.size(), c->stream);
        c->gval.upload(g.data(), g.size(), c->stream);
        c->bnd_node.resize(1);
        c->have_dirichlet = c->have_bnd_values = true;
        c->bnd_bc_type = 0;
      }
      if (!c->have_bnd_values && c->bnd_bc_type) ngpb::run_dirichlet_values(c, c->bnd_bc_type);
      need(c->have_bnd_values, "Dirichlet values missing (ngpb_dirichlet_values)");
      if (!c->have_frac) ngpb::run_fractions(c);
      sym[i] = !c->have_symbolic || c->external_csr;
      // a system that fails below is left without one
      c->have_system = c->have_compact = c->have_phi = false;
      if (sym[i]) c->have_symbolic = false;
    });
    if (rc != NGPB_OK) return rc;
  }
  auto fail_all = [&](int code, const std::string& msg) {
    for (ngpb_ctx* c : v) c->err = msg;
    return code;
  };
  try {
    NGPB_HIP(hipSetDevice(v[0]->device));
    ngpb::run_assemble_many(v);
  } catch (const ngpb::Error& e) {
    return fail_all(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail_all(NGPB_E_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail_all(NGPB_E_HIP, e.what());
  }
  for (int i = 0; i < n; ++i) {
    ngpb_ctx* c = v[i];
    if (sym[i]) c->have_levels = false;
    c->external_csr = false;
    c->have_phi = false;
    c->err.clear();
  }
  return NGPB_OK;
}

int ngpb_system_size(ngpb_ctx* c, int64_t* n, int64_t* nnz) {
  return guarded(c, [&] {
    NGPB_REQUIRE(n && nnz, NGPB_E_ARG, "NULL output");
    need(c->have_system, "no assembled system");
    *n = c->n;
    *nnz = c->nnz;
  });
}

int ngpb_export_csr(ngpb_ctx* c, int32_t* rowptr, int32_t* col, double* val, double* rhs) {
  return guarded(c, [&] {
    need(c->have_system, "no assembled system");
    if (rowptr) c->rowptr.download(rowptr, (size_t)c->n + 1, c->stream);
    if (col) c->col.download(col, (size_t)c->nnz, c->stream);
    if (val) c->val.download(val, (size_t)c->nnz, c->stream);
    if (rhs) c->rhs.download(rhs, (size_t)c->n, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
  });
}

int ngpb_set_csr(ngpb_ctx* c, int64_t n, int64_t nnz, const int32_t* rowptr, const int32_t* col, const double* val,
                 const double* rhs) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(n > 0 && nnz > 0 && n < INT32_MAX && nnz < INT32_MAX, NGPB_E_ARG, "bad CSR sizes");
    NGPB_REQUIRE(rowptr && col && val && rhs, NGPB_E_ARG, "NULL CSR array");
    NGPB_REQUIRE(rowptr[0] == 0 && rowptr[n] == nnz, NGPB_E_ARG, "row pointers must run from 0 to nnz");
    for (int64_t i = 0; i < n; ++i) {
      NGPB_REQUIRE(rowptr[i + 1] >= rowptr[i], NGPB_E_ARG, "row pointers must be non-decreasing");
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        NGPB_REQUIRE(col[k] >= 0 && col[k] < n, NGPB_E_ARG, "column index out of range");
        if (k > rowptr[i]) NGPB_REQUIRE(col[k] > col[k - 1], NGPB_E_ARG, "columns must be strictly ascending");
        NGPB_REQUIRE(std::isfinite(val[k]), NGPB_E_NUMERIC, "matrix holds NaN or Inf entries");
      }
      NGPB_REQUIRE(std::isfinite(rhs[i]), NGPB_E_NUMERIC, "right-hand side holds NaN or Inf entries");
    }
    c->n = n;
    c->nnz = nnz;
    c->rowptr.upload(rowptr, (size_t)n + 1, c->stream);
    c->col.upload(col, (size_t)nnz, c->stream);
    c->val.upload(val, (size_t)nnz, c->stream);
    c->rhs.upload(rhs, (size_t)n, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
    c->have_system = true;
    // the symbolic tables describe the mesh's pattern, not this one: the next
    // ngpb_assemble rebuilds them
    c->have_symbolic = false;
    c->external_csr = true;
    c->have_levels = false;
    c->have_compact = c->have_phi = false;
  });
}

int ngpb_parse_lis_options(const char* options, ngpb_solve_opts* o) {
  if (!options || !o) return NGPB_E_ARG;
  o->precision = NGPB_PREC_FP64;
  o->tol = 1e-12;  // LIS default
  o->maxiter = 1000;  // LIS default
  o->drop = 1e-3;
  o->check_every = 32;
  o->solver = NGPB_SOLVER_CG;     // LIS default -i cg
  o->precon = NGPB_PRECON_NONE;   // LIS default -p none
  o->omega = 1.0;                 // LIS default -ssor_omega
  std::istringstream is(options);
  std::vector<std::string> tok;
  for (std::string t; is >> t;) tok.push_back(t);
  for (size_t i = 0; i < tok.size(); ++i) {
    const std::string& k = tok[i];
    auto arg = [&](std::string* v) {
      if (i + 1 >= tok.size()) return false;
      *v = tok[++i];
      return true;
    };
    std::string v;
    if (k == "-i") {
      if (!arg(&v)) return NGPB_E_ARG;
      if (v == "cg" || v == "1") o->solver = NGPB_SOLVER_CG;
      else if (v == "cgs" || v == "3") o->solver = NGPB_SOLVER_CGS;
      else if (v == "bicgstab" || v == "4") o->solver = NGPB_SOLVER_BICGSTAB;
      else return NGPB_E_ARG;
    } else if (k == "-p") {
      if (!arg(&v)) return NGPB_E_ARG;
      if (v == "none" || v == "0") o->precon = NGPB_PRECON_NONE;
      else if (v == "jacobi" || v == "1") o->precon = NGPB_PRECON_JACOBI;
      else if (v == "ssor" || v == "3") o->precon = NGPB_PRECON_SSOR;
      else return NGPB_E_ARG;
    } else if (k == "-ssor_omega") {
      if (!arg(&v)) return NGPB_E_ARG;
      o->omega = std::atof(v.c_str());
      if (!(o->omega > 0.0 && o->omega < 2.0)) return NGPB_E_ARG;
    } else if (k == "-tol") {
      if (!arg(&v)) return NGPB_E_ARG;
      o->tol = std::atof(v.c_str());
    } else if (k == "-maxiter") {
      if (!arg(&v)) return NGPB_E_ARG;
      o->maxiter = std::atoi(v.c_str());
    } else if (k == "-precision" || k == "-f") {
      if (!arg(&v)) return NGPB_E_ARG;
      if (v == "mixed" || v == "fp32") o->precision = NGPB_PREC_MIXED;
      else if (v == "double" || v == "fp64") o->precision = NGPB_PREC_FP64;
      else return NGPB_E_ARG;
    } else if (k == "-print" || k == "-conv_cond" || k == "-tol_w" || k == "-initx_zeros" || k == "-scale") {
      if (!arg(&v)) return NGPB_E_ARG;  // accepted, no effect (x0 = 0, ||r|| <= tol ||b||)
    } else {
      return NGPB_E_ARG;
    }
  }
  return (o->tol > 0 && o->maxiter > 0) ? NGPB_OK : NGPB_E_ARG;
}

int ngpb_solve(ngpb_ctx* c, const ngpb_solve_opts* o, double* phi_out, ngpb_solve_stats* st) {
  ngpb_solve_stats local{};
  ngpb_solve_stats& s = st ? *st : local;
  std::memset(&s, 0, sizeof(s));
  const int rc = guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(o, NGPB_E_ARG, "NULL options");
    NGPB_REQUIRE(o->tol > 0 && o->maxiter > 0, NGPB_E_ARG, "tol and maxiter must be positive");
    need(c->have_system, "no assembled system");
    ngpb::check_lis_options(*o);
    const bool cg_jacobi = o->solver == NGPB_SOLVER_CG && o->precon == NGPB_PRECON_JACOBI;
    if (o->precision == NGPB_PREC_FP64) {
      if (cg_jacobi) ngpb::solve_fp64(c, *o, s);
      else ngpb::solve_lis(c, *o, s);
    } else if (o->precision == NGPB_PREC_MIXED) {
      NGPB_REQUIRE(cg_jacobi, NGPB_E_ARG, "mixed precision runs CG with Jacobi scaling only");
      ngpb::solve_mixed(c, *o, s);
    } else {
      throw Error(NGPB_E_ARG, "unknown precision");
    }
    if (phi_out) {
      c->x.download(phi_out, (size_t)c->n, c->stream);
      NGPB_HIP(hipStreamSynchronize(c->stream));
    }
  });
  if (rc != NGPB_OK) return rc;
  if (s.status != NGPB_OK) {
    c->err = "Krylov solver stopped at maxiter (or stagnated) before reaching the tolerance";
    return s.status;
  }
  return NGPB_OK;
}

int ngpb_precond_apply(ngpb_ctx* c, int precon, double omega, const double* r, double* z) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(r && z, NGPB_E_ARG, "NULL vector");
    NGPB_REQUIRE(precon == NGPB_PRECON_JACOBI || precon == NGPB_PRECON_SSOR, NGPB_E_ARG,
                 "preconditioner must be Jacobi or SSOR");
    NGPB_REQUIRE(precon != NGPB_PRECON_SSOR || (omega > 0.0 && omega < 2.0), NGPB_E_ARG,
                 "SSOR relaxation must lie in (0, 2)");
    need(c->have_system, "no assembled system");
    ngpb::DBuf<double> dr, dz;
    dr.upload(r, (size_t)c->n, c->stream);
    dz.resize((size_t)c->n);
    ngpb::precond_apply(c, precon, omega, dr.p, dz.p);
    dz.download(z, (size_t)c->n, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
  });
}

int ngpb_solve_csr_noalloc(int device, void* stream, int64_t n, int64_t nnz, const int32_t* rowptr,
                           const int32_t* col, const double* val, const double* rhs, double* x,
                           const ngpb_solve_opts* opts, ngpb_solve_stats* stats) {
  ngpb_solve_stats local{};
  ngpb_solve_stats& st = stats ? *stats : local;
  std::memset(&st, 0, sizeof(st));
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return NGPB_E_NODEV;
  }
  if (device < 0 || device >= count || device >= 64) return NGPB_E_ARG;
  if (!rowptr || !col || !val || !rhs || !x || !opts || n <= 0 || nnz <= 0 || nnz >= INT32_MAX) return NGPB_E_ARG;
  if (n > ngpb::noalloc_max_rows()) return NGPB_E_ARG;
  if (!(opts->tol > 0) || opts->maxiter <= 0 || opts->precision != NGPB_PREC_FP64 ||
      opts->solver != NGPB_SOLVER_CG || opts->precon != NGPB_PRECON_JACOBI)
    return NGPB_E_ARG;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) {
    (void)hipGetLastError();
    prev = -1;
  }
  int rc = NGPB_OK;
  try {
    NGPB_HIP(hipSetDevice(device));
    ngpb::solve_csr_noalloc(device, static_cast<hipStream_t>(stream), n, nnz, rowptr, col, val, rhs, x, opts->tol,
                            opts->maxiter, opts->check_every, st);
    rc = st.status;
  } catch (const Error& e) {
    rc = e.code;
  } catch (const std::exception&) {
    rc = NGPB_E_HIP;
  }
  if (prev >= 0) (void)hipSetDevice(prev);
  return rc;
}

int ngpb_get_phi(ngpb_ctx* c, double* phi) {
  return guarded(c, [&] {
    NGPB_REQUIRE(phi, NGPB_E_ARG, "NULL output");
    need(c->have_phi, "no solution");
    c->x.download(phi, (size_t)c->n, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
  });
}

int ngpb_set_phi(ngpb_ctx* c, const double* phi) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(phi, NGPB_E_ARG, "NULL input");
    need(c->have_mesh, "no mesh");
    c->x.upload(phi, (size_t)c->N, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
    c->have_phi = true;
    c->phi_len = c->N;
  });
}

int ngpb_spmv(ngpb_ctx* c, int precision, const double* x, double* y) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(x && y, NGPB_E_ARG, "NULL vector");
    need(c->have_system, "no assembled system");
    ngpb::DBuf<double> dx, dy;
    dx.upload(x, (size_t)c->n, c->stream);
    dy.resize((size_t)c->n);
    if (precision == NGPB_PREC_FP64) ngpb::spmv_fp64(c, dx.p, dy.p);
    else if (precision == NGPB_PREC_MIXED) ngpb::spmv_compact(c, dx.p, dy.p);
    else throw Error(NGPB_E_ARG, "unknown precision");
    dy.download(y, (size_t)c->n, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
  });
}

int ngpb_phi_on_atoms(ngpb_ctx* c, double* phi_atoms) {
  return guarded(c, [&] {
    NGPB_REQUIRE(phi_atoms, NGPB_E_ARG, "NULL output");
    need(c->have_phi, "no potential (solve or ngpb_set_phi first)");
    need(c->have_mesh && c->phi_len == c->N, "the potential does not hold one value per mesh node");
    need(c->have_charges, "no atoms (ngpb_set_charges)");
    if (c->A == 0) return;
    ngpb::DBuf<double> d;
    d.resize((size_t)c->A);
    ngpb::run_phi_on_atoms(c, d.p);
    d.download(phi_atoms, (size_t)c->A, c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
  });
}

int ngpb_energy(ngpb_ctx* c, ngpb_energy_out* out) {
  return guarded(c, [&] {
    NGPB_REQUIRE(out, NGPB_E_ARG, "NULL output");
    need(c->have_markers, "markers are not available");
    need(c->have_phi, "no potential (solve or ngpb_set_phi first)");
    need(c->phi_len == c->N, "the potential does not hold one value per mesh node");
    if (!c->have_frac) ngpb::run_fractions(c);
    ngpb::run_energy(c, out);
  });
}

int ngpb_pot_field(ngpb_ctx* c, int calc_potential_term, int calc_field_term, double* table,
                   ngpb_pot_field_out* out) {
  return guarded(c, [&] {
    NGPB_REQUIRE(out, NGPB_E_ARG, "NULL output");
    NGPB_REQUIRE(calc_potential_term >= 0 && calc_potential_term <= 2 && calc_field_term >= 0 &&
                     calc_field_term <= 2,
                 NGPB_E_ARG, "calc_potential_term and calc_field_term must be 0, 1 or 2");
    need(c->have_charges, "no atoms (ngpb_set_charges)");
    std::memset(out, 0, sizeof(*out));
    out->branch = ngpb::pot_field_branch(calc_potential_term, calc_field_term, c->pc.k);
    if (out->branch != 0) {
      need(c->have_markers, "markers are not available");
      need(c->have_phi, "no potential (solve or ngpb_set_phi first)");
      need(c->phi_len == c->N, "the potential does not hold one value per mesh node");
      if (!c->have_frac) ngpb::run_fractions(c);
    }
    double e[4];
    int br = 0;
    ngpb::run_pot_field(c, calc_potential_term, calc_field_term, table, e, &br);
    out->energy_pol = e[0];
    out->energy_react = e[1];
    out->coul_energy = e[2];
    out->charge_pol = e[3];
  });
}

int ngpb_time_spmv(ngpb_ctx* c, int precision, int reps, double* ms, double* bytes) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(ms && bytes && reps > 0, NGPB_E_ARG, "bad arguments");
    need(c->have_system, "no assembled system");
    if (precision == NGPB_PREC_FP64) {
      *ms = ngpb::time_spmv_fp64(c, reps);
      *bytes = ngpb::spmv_fp64_bytes(c);
    } else if (precision == NGPB_PREC_MIXED) {
      *ms = ngpb::time_spmv_compact(c, reps);
      *bytes = ngpb::spmv_compact_bytes(c);
    } else {
      throw Error(NGPB_E_ARG, "unknown precision");
    }
  });
}

int ngpb_set_profiling(ngpb_ctx* c, int enable) {
  return guarded(c, [&] {
    c->profile = enable != 0;
    for (auto& p : c->kprof) {
      p.used = 0;
      p.launches = 0;
      p.ms = 0.0;
    }
  });
}

int ngpb_kernel_stats(ngpb_ctx* c, int precision, int64_t* launches, double* total_ms, double* bytes_per_launch) {
  return guarded(c, [&] {
    NGPB_REQUIRE(launches && total_ms && bytes_per_launch, NGPB_E_ARG, "NULL output");
    NGPB_REQUIRE(precision == NGPB_PREC_FP64 || precision == NGPB_PREC_MIXED, NGPB_E_ARG, "unknown precision");
    const auto& p = c->kprof[precision];
    *launches = p.launches;
    *total_ms = p.ms;
    *bytes_per_launch = !c->have_system ? 0.0
                        : precision == NGPB_PREC_FP64 ? ngpb::spmv_fp64_bytes(c)
                        : (c->have_compact ? ngpb::spmv_compact_bytes(c) : 0.0);
  });
}

int ngpb_operator_info(ngpb_ctx* c, int64_t* info) {
  return guarded(c, [&] {
    if (c->in_batch) need(c->have_compact, "context is in a batch");  // read-only then
    NGPB_REQUIRE(info, NGPB_E_ARG, "NULL output");
    need(c->have_system, "no assembled system");
    if (!c->have_compact) ngpb::build_compact(c);
    std::vector<int32_t> w((size_t)c->n_slices);
    c->slice_width.download(w.data(), w.size(), c->stream);
    NGPB_HIP(hipStreamSynchronize(c->stream));
    int64_t nd = 0;
    for (int32_t v : w) nd += (v < 0 && v > -(64 + 1)) ? 1 : 0;  // packed slices have width <= -65
    info[0] = c->n_slices;
    info[1] = nd;
    info[2] = c->sell_delta_entries;
    info[3] = c->sell_entries - c->sell_delta_entries - c->sell_const_slots * 64;
    info[4] = c->sell_wide_entries;
    info[5] = c->sell_ov_entries;
    info[6] = (int64_t)ngpb::spmv_compact_bytes(c);
    info[7] = (int64_t)ngpb::spmv_fp64_bytes(c);
    info[8] = c->sell_pk_slices;
    info[9] = c->sell_pk_entries;
    info[10] = c->sell_const_slices;
    info[11] = c->sell_const_slots;
  });
}

int ngpb_set_compact_options(ngpb_ctx* c, int constant_slices) {
  return guarded(c, [&] {
    not_in_batch(c);
    const bool v = constant_slices != 0;
    if (v != c->sell_allow_const) {
      c->sell_allow_const = v;
      c->have_compact = false;  // rebuilt with the new forms on next use
    }
  });
}

int ngpb_set_row_order(ngpb_ctx* c, int band) {
  return guarded(c, [&] {
    not_in_batch(c);
    if (band < 0 || band > 4096) throw Error(NGPB_E_ARG, "band must be 0 (node order) or 1 .. 4096");
    if (band != c->row_band) {
      c->row_band = band;
      c->have_compact = false;  // rebuilt in the new order on next use
    }
  });
}

int ngpb_row_order(ngpb_ctx* c, int* band) {
  return guarded(c, [&] {
    NGPB_REQUIRE(band, NGPB_E_ARG, "NULL output");
    if (c->in_batch) need(c->have_compact, "context is in a batch");
    need(c->have_system, "no assembled system");
    if (!c->have_compact) ngpb::build_compact(c);
    *band = c->sell_banded ? c->sell_band : 0;
  });
}

int ngpb_set_spmv_schedule(ngpb_ctx* c, int schedule) {
  return guarded(c, [&] {
    not_in_batch(c);
    if (schedule < 0 || schedule > 6) throw Error(NGPB_E_ARG, "schedule must be 0 .. 6");
    if (schedule != c->spmv_sched) c->sell_resident = 0;  // grid follows the schedule's kernel
    c->spmv_sched = schedule;
  });
}

int ngpb_time_copy(ngpb_ctx* c, int64_t bytes, int reps, double* ms_per_launch) {
  return guarded(c, [&] {
    not_in_batch(c);
    NGPB_REQUIRE(ms_per_launch && reps > 0 && bytes >= 16 && bytes % 16 == 0, NGPB_E_ARG,
                 "bytes must be a positive multiple of 16");
    *ms_per_launch = ngpb::time_copy(c, bytes, reps);
  });
}

}  // extern "C"
