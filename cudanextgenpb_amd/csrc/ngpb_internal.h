// Internal declarations of the gfx950 linearized-PB library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ngpb.h"

namespace ngpb {

// --------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define NGPB_HIP(call)                                               \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) ::ngpb::throw_hip(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define NGPB_REQUIRE(cond, code, msg)                 \
  do {                                                \
    if (!(cond)) throw ::ngpb::Error((code), (msg));  \
  } while (0)

// --------------------------------------------------------------- device buffers
// Blocks of stream-ordered buffers released on a stream are kept for later buffers on
// the same stream (stream order makes the reuse safe: every use of the old buffer was
// enqueued before every use of the new one), so a call's repeated temporaries (the
// per-axis arrays of a ray cast, the per-level arrays of an octree build) are not
// handed back to the driver and fetched again: a hipFreeAsync of a 64 MB block has
// held the host for ~5 ms. No events are involved, so no thread ever waits on work
// of another thread's stream. stream_cache_take returns a block of at least `bytes`
// (and not much more) with its size in *cap, or nullptr; stream_cache_put returns
// false when the stream's cache is full (the caller frees the block);
// stream_cache_trim frees blocks until at most `keep` bytes stay cached;
// stream_cache_end_call trims to the bound kept between API calls; and
// stream_cache_drop frees them all (before the stream is destroyed).
void* stream_cache_take(hipStream_t s, size_t bytes, size_t* cap);
bool stream_cache_put(hipStream_t s, void* p, size_t cap);
void stream_cache_trim(hipStream_t s, size_t keep);
void stream_cache_end_call(hipStream_t s);
void stream_cache_drop(hipStream_t s);

// A buffer constructed with a stream (or given one by order_on) is allocated and
// freed in that stream's order (hipMallocAsync / hipFreeAsync, through the stream's
// block cache): freeing it waits neither for the device nor for other streams, which
// matters for the temporaries of one call while other contexts' work runs on the
// same GPU. The others use hipMalloc / hipFree (the resident arrays of a context).
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  size_t cap = 0;  // bytes of the block (stream-ordered buffers)
  hipStream_t os = nullptr;
  bool ordered = false;
  DBuf() = default;
  explicit DBuf(hipStream_t s) : os(s), ordered(true) {}
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void order_on(hipStream_t s) {
    release();
    os = s;
    ordered = true;
  }
  void release() {
    if (p) {
      if (!ordered) (void)hipFree(p);
      else if (!stream_cache_put(os, p, cap)) (void)hipFreeAsync(p, os);
    }
    p = nullptr;
    n = 0;
    cap = 0;
  }
  // (re)allocates when the size differs; contents are undefined afterwards
  void resize(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    const size_t bytes = count * sizeof(T);
    if (ordered && (p = static_cast<T*>(stream_cache_take(os, bytes, &cap)))) {
      n = count;
      return;
    }
    cap = bytes;
    hipError_t e = ordered ? hipMallocAsync((void**)&p, bytes, os) : hipMalloc((void**)&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw Error(NGPB_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    n = count;
  }
  // exchanges the blocks of two buffers of the same allocation kind and stream (the
  // block sizes go with them: a stream-ordered block returns to the cache with its own)
  void swap(DBuf& o) {
    if (ordered != o.ordered || os != o.os) throw Error(NGPB_E_STATE, "buffer swap across allocation kinds");
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(cap, o.cap);
  }
  // grows only
  void reserve(size_t count) {
    if (count > n) resize(count);
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    resize(count);
    if (count) NGPB_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count) NGPB_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  size_t bytes() const { return n * sizeof(T); }
};

// --------------------------------------------------------------- physics
struct PhysConst {
  double eps_in, eps_out, k2, k, constant_pol, constant_react;
};
PhysConst physical_constants(const ngpb_physics& p);

// --------------------------------------------------------------- device views
struct RayView {
  int64_t n = 0;
  const double* key = nullptr;     // 2 per ray
  const int64_t* ptr = nullptr;    // n+1
  const double* inters = nullptr;  // K
  const double* normals = nullptr; // 3K
};

struct MeshView {
  int64_t N = 0, H = 0, L = 0;
  const double* node_xyz = nullptr;
  const double* hang_xyz = nullptr;
  const int32_t* hang_np = nullptr;
  const int32_t* hang_par = nullptr;
  const int32_t* leaf_corner = nullptr;
};

// per-block partial sums of the Krylov reductions (up to 4 per block)
struct GridReduce {
  DBuf<double> partials;
};

// Scalar state of a Krylov solve on the device. Inside an iteration the scalars that
// one kernel writes while other blocks of the same kernel still read the previous
// value are double-buffered by the parity of the host's launch index (slot [par] is
// read, slot [par ^ 1] written).
struct KrylovScalars {
  double rho;       // (r, z) of the current iterate (outside the iteration loop)
  double pq;        // (p, A p)
  double alpha, beta;
  double rr;        // (r, r) of the current iterate
  double stop_rr;   // stop when rr <= stop_rr
  double bb;        // (b, b)
  double aux[4];
  double rho2[2];   // (r, z) entering the iteration of parity par
  double alpha2[2]; // alpha of the iteration of parity par
  int iter2[2];     // iterations completed before the iteration of parity par
  int stop_par;     // parity of the iteration that raised the stop flag
  int iter;         // iterations done in this solve
  int done;         // 1 = stop flag raised
  int maxiter;      // iteration limit for the current cycle
  int status;       // 0 converged, 4 maxiter
};

// Grow-only device scratch slots held by the context and reused across calls, so that
// the per-step path (markers, compact-operator build, energy) allocates and frees
// nothing once its sizes have been reached (hipFree also waits for the device). Each
// slot has one user; no slot is held across a call that uses it itself.
enum TmpSlot : int {
  kTmpFlagCount = 0,              // compact_flags: per-element counts
  kTmpInside,                     // markers: node inside flags
  kTmpBorderFlag,                 // markers: border leaf flags
  kTmpCompact,                    // build_compact: 14 per-slice arrays
  kTmpForms = kTmpCompact + 14,   // build_compact: 3 form flag arrays
  kTmpEnergy = kTmpForms + 3,     // energy sources and sums: 19 arrays
  kTmpPotField = kTmpEnergy + 19, // pot_field: 7 arrays
  kTmpCount = kTmpPotField + 7
};

// ----------------------------------------------------------------- context
}  // namespace ngpb

struct ngpb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  bool have_physics = false;
  ngpb_physics physics{};
  ngpb::PhysConst pc{};

  // mesh
  int64_t N = 0, H = 0, L = 0;
  ngpb::DBuf<double> node_xyz, hang_xyz;
  ngpb::DBuf<int32_t> hang_np, hang_par, leaf_corner;
  bool have_mesh = false;

  // rays
  struct Rays {
    int64_t n = 0, k = 0;
    ngpb::DBuf<double> key, inters, normals;
    ngpb::DBuf<int64_t> ptr;
  } rays[3];

  // markers
  bool have_markers = false;
  ngpb::DBuf<double> eps, reac;
  ngpb::DBuf<uint8_t> leaf_marker;
  ngpb::DBuf<int32_t> border;      // border leaf ids
  ngpb::DBuf<int32_t> border_of;   // leaf -> border index or -1
  int64_t n_border = 0;

  // fractions of border leaves
  bool have_frac = false;
  ngpb::DBuf<double> frac, frac_normal;  // 12 / 36 per border leaf
  ngpb::DBuf<uint8_t> frac_found;        // 12 per border leaf

  // charges
  int64_t A = 0;
  ngpb::DBuf<double> atom_xyz, atom_q;
  ngpb::DBuf<int32_t> atom_owner;
  bool have_charges = false;

  // Dirichlet data
  int64_t NB = 0;
  ngpb::DBuf<int32_t> bnd_node;
  ngpb::DBuf<double> bnd_val;
  ngpb::DBuf<uint8_t> is_bnd;   // N
  ngpb::DBuf<double> gval;      // N (value on boundary nodes, 0 elsewhere)
  bool have_dirichlet = false, have_bnd_values = false;
  // boundary values computed by ngpb_dirichlet_values (not supplied): they depend on
  // the charges and the physics, so changing either recomputes them at assembly
  int bnd_bc_type = 0;  // 0: supplied by the caller, else the bc_type they were computed for

  // symbolic assembly (depends on mesh + Dirichlet set)
  bool have_symbolic = false;
  // The sorted emission lists are cut into slots, then repacked in slot order with
  // only what the numeric phase reads (source id and weight code, 5 bytes); the sort
  // keys are released. Slot s of a row in 8-row block b starts at
  // blk_em[b] + slot_rel[s] and ends where slot s + 1 starts (blk_em[b + 1] for the
  // block's last slot).
  ngpb::DBuf<uint32_t> em_id;     // slot emissions: source id (edge 12 l + e, corner 12 L + 8 l + i)
  ngpb::DBuf<uint8_t> em_code;    //                 weight code
  ngpb::DBuf<uint32_t> slot_rel;  // nnz: first emission of the slot, relative to its block
  ngpb::DBuf<int64_t> blk_em;     // ceil(N / 8) + 1: first emission of each 8-row block
  ngpb::DBuf<int32_t> diag_slot;  // N
  ngpb::DBuf<int32_t> bc_rowptr;  // N+1
  ngpb::DBuf<int32_t> bc_col;
  ngpb::DBuf<int64_t> bc_ptr;     // couplings + 1: emission range of each Dirichlet coupling
  ngpb::DBuf<uint32_t> bc_id;     // coupling emissions: source id
  ngpb::DBuf<uint8_t> bc_code;    //                     weight code

  // assembled system
  bool have_system = false;
  bool external_csr = false;  // the system came from ngpb_set_csr, not from the mesh
  int64_t n = 0, nnz = 0;
  ngpb::DBuf<int32_t> rowptr, col;
  ngpb::DBuf<double> val, rhs, rhs_charge;
  ngpb::DBuf<double> c_edge, v8;  // per leaf-edge conductance, per leaf V/8
  ngpb::DBuf<double> vol_patch, rho;

  // compact mixed-precision operator
  bool have_compact = false;
  int64_t n_slices = 0, sell_entries = 0, sell_delta_entries = 0, sell_wide_entries = 0, sell_bases = 0,
          sell_ov_entries = 0;
  ngpb::DBuf<int32_t> slice_width;    // per slice: main slots
  ngpb::DBuf<int64_t> slice_base_ptr; // per slice: first slot (entries start at 64x this)
  ngpb::DBuf<int64_t> slice_wide_ptr; // per slice: first 32-bit column (or -1)
  ngpb::DBuf<int64_t> slice_ov_ptr;   // per slice: first overflow entry (or -1)
  ngpb::DBuf<float> sell_val;
  ngpb::DBuf<uint16_t> sell_off;
  ngpb::DBuf<int32_t> sell_base, sell_col32;
  ngpb::DBuf<uint8_t> sell_ov_cnt;    // per row: overflow entries
  ngpb::DBuf<float> sell_ov_val;
  ngpb::DBuf<int32_t> sell_ov_col;
  int64_t sell_pk_slices = 0, sell_pk_entries = 0, sell_pk_dict_entries = 0, sell_pk_words = 0;
  ngpb::DBuf<float> sell_pk_val;      // packed slices: values
  ngpb::DBuf<uint32_t> sell_pk_idx;   //                dictionary indices, 4 slots per word
  ngpb::DBuf<uint16_t> sell_pk_lp;    //                per sorted position: lane | count << 8
  ngpb::DBuf<int32_t> sell_pk_dict;   //                column deltas
  ngpb::DBuf<int64_t> sell_pk_hdr;    //                32-byte header per slice (4 x int64)
  ngpb::DBuf<float> sell_cval;        // constant delta slices: one value per slot
  int64_t sell_const_slices = 0, sell_const_slots = 0;
  bool sell_allow_const = true;       // build constant delta slices where they apply
  // slices of each form, for the form-specialised SpMV kernels
  int64_t n_list_delta = 0, n_list_packed = 0, n_list_offset = 0;
  ngpb::DBuf<int32_t> sell_list_delta, sell_list_packed, sell_list_offset;
  // pipelined SpMV: per pair of delta-list entries its slice ids, slot pointers, widths
  // and pairing (8 x int32), and the packed slices' headers in list order
  ngpb::DBuf<int32_t> sell_dpair;
  // row order of the compact operator: row_band 0 node order, > 0 bands of that many
  // finest-grid spacings in y; sell_banded: the built operator's rows are nodes
  // row_perm[0 .. n-1]
  int row_band = 0;
  bool sell_banded = false;
  int sell_band = 0;  // band of the built operator when sell_banded
  ngpb::DBuf<int32_t> row_perm;
  ngpb::DBuf<int64_t> sell_pk_hdr_list;
  ngpb::DBuf<double> dscale;         // 1/sqrt(diag)
  int spmv_sched = 6;                // SpMV work schedule (0 .. 6)
  int sell_resident = 0;             // resident blocks of the compact SpMV (0: not yet queried)
  int csr_resident = 0;              // resident blocks of the FP64 CSR SpMV (0: not yet queried)

  // solver state
  ngpb::DBuf<double> x, r, z, p, q, dinv;
  // other LIS solvers (ngpb_lis.hip): work vectors and the SSOR level sets, built
  // once per sparsity pattern
  ngpb::DBuf<double> lis_rt, lis_u, lis_v, lis_ph, lis_uh, lis_t;
  bool have_levels = false;
  ngpb::DBuf<int32_t> diag_pos;              // position of each row's diagonal entry
  ngpb::DBuf<int32_t> lev_rows_f, lev_rows_b;  // rows by forward / backward level
  std::vector<int64_t> lev_ptr_f, lev_ptr_b;   // level offsets (host)
  ngpb::DBuf<float> xs, rs, ps, qs;
  ngpb::DBuf<ngpb::KrylovScalars> ks;
  ngpb::KrylovScalars* ks_host = nullptr;  // pinned mirror
  ngpb::GridReduce red;
  bool have_phi = false;
  int64_t phi_len = 0;    // entries of x that hold the potential (N after an assembled solve)
  bool in_batch = false;  // solver vectors are referenced by a batch: no reallocation

  // energy scratch
  ngpb::DBuf<double> en_src, en_acc;
  ngpb::DBuf<double> en_out;
  ngpb::DBuf<int64_t> en_count;

  // scan scratch
  ngpb::DBuf<int64_t> scan_tmp;
  ngpb::DBuf<uint8_t> tmp[ngpb::kTmpCount];  // grow-only scratch slots (TmpSlot)

  // optional HIP-event timing of SpMV launches (kind 0: FP64 CSR, 1: compact FP32)
  bool profile = false;
  struct KProf {
    std::vector<hipEvent_t> ev;  // start/stop pairs, one per timed kernel
    std::vector<int64_t> tag;    // launch index of the SpMV each pair belongs to
    std::vector<char> opens;     // 1 for the first kernel of that SpMV
    size_t used = 0;             // events attached since the last collection
    int64_t launches = 0;
    double ms = 0.0;
  } kprof[2];
  int prof_kind = -1;   // SpMV being timed: its kind, or -1
  int64_t prof_tag = 0;
  bool prof_first = false;

  int64_t device_bytes() const;
};

namespace ngpb {

// ----------------------------------------------------------------- kernels (host wrappers)
// scan
int64_t exclusive_scan(ngpb_ctx* c, const int64_t* d_in, int64_t* d_out, int64_t n);
// Grouped launches run one kernel body over many independent ranges (systems,
// segments): range i owns the blocks [blk0[i], blk0[i + 1]) of the launch and sees
// them as its blocks 0, 1, ... . group_of returns the range of block b (empty ranges
// are skipped: the last i with blk0[i] <= b).
__device__ __forceinline__ int group_of(const int64_t* __restrict__ blk0, int n, int64_t b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
struct ScanSeg {
  const int64_t* in;
  int64_t* out;  // may alias in
  int64_t n;
};
// exclusive scans of many segments in three launches; totals (host, one per segment)
// receive the segment sums. Synchronizes the stream.
void exclusive_scan_many(hipStream_t st, const std::vector<ScanSeg>& segs, int64_t* totals);
// compaction: indices i in [0, n) with flag[i] != 0, in order; returns count
int64_t compact_flags(ngpb_ctx* c, const uint8_t* d_flag, int64_t n, int32_t* d_out);

// markers
void run_markers(ngpb_ctx* c);
void run_border_from_markers(ngpb_ctx* c);
void run_fractions(ngpb_ctx* c);

// assembly
void run_symbolic(ngpb_ctx* c);
void run_numeric(ngpb_ctx* c);
// run_symbolic (where a context needs it) and run_numeric of many contexts on one
// device, each stage in one launch over all of them, on the first context's stream;
// every context's results equal its own run_symbolic / run_numeric bit for bit
void run_assemble_many(const std::vector<ngpb_ctx*>& cs);
void run_dirichlet_values(ngpb_ctx* c, int bc_type);
void run_phi_on_atoms(ngpb_ctx* c, double* d_out);
void run_cast_rays(ngpb_ctx* c, int64_t na, const double* xyz, const double* r, double probe);
// rays of every grid line along z through (origin0 + i hf, origin1 + j hf), i <= n0,
// j <= n1, into c->rays[2] (sphere union, or SES for probe > 0)
void cast_grid_zlines(ngpb_ctx* c, const double* origin, double hf, int64_t n0, int64_t n1, int64_t na,
                      const double* xyz, const double* r, double probe);
void run_locate_atoms(ngpb_ctx* c, int64_t A, const double* h_xyz, double tol, int32_t* h_owner);
struct OctreeSizes {
  int64_t N, H, L, NB;
};
OctreeSizes run_build_octree(ngpb_ctx* c, const double* origin, double hf, const int64_t* n, int smax,
                             const double* band, int64_t na, const double* h_xyz, const double* h_r);
OctreeSizes run_build_octree_surface(ngpb_ctx* c, const double* origin, double hf, const int64_t* n, int smax,
                                     int s_uni, int cycles, int64_t na, const double* h_xyz, const double* h_r,
                                     double probe);

// Krylov
unsigned vec_grid(int64_t n);
void check_operator(ngpb_ctx* c);
void ensure_reduce(ngpb_ctx* c, int64_t blocks);
// FP64 CSR SpMV y = A x that also writes per-block (x, y) partials and returns at once
// once the stop flag of c->ks is raised; grid size of its launches
void spmv_fp64_dot(ngpb_ctx* c, const double* x, double* y, double* partials);
unsigned spmv_fp64_grid(ngpb_ctx* c);
void check_operator(ngpb_ctx* c);
// other LIS solvers (ngpb_lis.hip)
void check_lis_options(const ngpb_solve_opts& o);
void solve_lis(ngpb_ctx* c, const ngpb_solve_opts& o, ngpb_solve_stats& st);
void precond_apply(ngpb_ctx* c, int precon, double omega, const double* r, double* z);
// CG-Jacobi on caller device arrays with the library's static scratch (no allocation)
int64_t noalloc_max_rows();
void solve_csr_noalloc(int device, hipStream_t s, int64_t n, int64_t nnz, const int32_t* rowptr, const int32_t* col,
                       const double* val, const double* b, double* x, double tol, int maxiter, int check_every,
                       ngpb_solve_stats& st);
void spmv_fp64(ngpb_ctx* c, const double* x, double* y);
void solve_fp64(ngpb_ctx* c, const ngpb_solve_opts& o, ngpb_solve_stats& st);
double spmv_fp64_bytes(const ngpb_ctx* c);
double time_spmv_fp64(ngpb_ctx* c, int reps);

// mixed
void build_compact(ngpb_ctx* c, bool allow_band = true);  // allow_band: banded row order may apply
float* spmv_input(ngpb_ctx* c);  // FP32 SpMV input vector (zero guard entries either side)
void spmv_compact(ngpb_ctx* c, const double* x, double* y);
void solve_mixed(ngpb_ctx* c, const ngpb_solve_opts& o, ngpb_solve_stats& st);
double spmv_compact_bytes(const ngpb_ctx* c);
double time_spmv_compact(ngpb_ctx* c, int reps);

// energy
void run_energy(ngpb_ctx* c, ngpb_energy_out* out);
int pot_field_branch(int P, int F, double k);
void run_pot_field(ngpb_ctx* c, int P, int F, double* table_host, double out4[4], int* branch);

// SpMV launch timing: every kProfEvery-th SpMV of a solve is timed. Between prof_begin
// and prof_end, the kernels launched through spmv_launch carry a start/stop event pair
// of their own (hipExtLaunchKernelGGL), which times the kernel itself: the gap between
// kernels and the end-of-kernel cache release are not counted.
constexpr int64_t kProfEvery = 8;
void prof_begin(ngpb_ctx* c, int kind, int64_t tag);
void prof_end(ngpb_ctx* c, int kind);
void prof_collect(ngpb_ctx* c, int kind, int64_t limit);
hipEvent_t* prof_pair(ngpb_ctx* c);  // nullptr unless an SpMV is being timed

template <typename... K, typename... A>
void spmv_launch(ngpb_ctx* c, void (*k)(K...), dim3 grid, dim3 block, A... args) {
  if (hipEvent_t* e = prof_pair(c))
    hipExtLaunchKernelGGL(k, grid, block, 0, c->stream, e[0], e[1], 0, args...);
  else
    hipLaunchKernelGGL(k, grid, block, 0, c->stream, args...);
}
double time_copy(ngpb_ctx* c, int64_t bytes, int reps);

// launch helper
inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// typed view of scratch slot `slot` with room for `count` elements (grows only)
template <class T>
T* tmp_buf(ngpb_ctx* c, int slot, size_t count) {
  DBuf<uint8_t>& b = c->tmp[slot];
  b.reserve(std::max<size_t>(count, 1) * sizeof(T));
  return reinterpret_cast<T*>(b.p);
}

}  // namespace ngpb

// ------------------------------------------------------------------- device helpers
namespace ngpb {
namespace dev {

__device__ __forceinline__ double corner_coord(const MeshView& m, int32_t c, int d) {
  return c >= 0 ? m.node_xyz[3 * (int64_t)c + d] : m.hang_xyz[3 * (int64_t)(-1 - c) + d];
}

// index of the ray with key (a, b) or -1 (exact comparison, lexicographic order)
__device__ __forceinline__ int64_t ray_find(const RayView& r, double a, double b) {
  int64_t lo = 0, hi = r.n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const double ka = r.key[2 * mid], kb = r.key[2 * mid + 1];
    if (ka < a || (ka == a && kb < b)) lo = mid + 1;
    else hi = mid;
  }
  if (lo < r.n && r.key[2 * lo] == a && r.key[2 * lo + 1] == b) return lo;
  return -1;
}

// 32-bit value of lane (lane ^ j), j a power of two below 64 (a constant once the
// caller's loop is unrolled): DPP lane moves inside 16-lane rows for j <= 8 (quad
// permutations for 1 and 2, a half-row mirror then a quad reversal for 4, a rotation
// by 8 for 8), the LDS permute across rows for 16 and 32.
__device__ __forceinline__ int lane_xor_i32(int v, int j) {
  switch (j) {
    case 1: return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true);
    case 2: return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true);
    case 4: return __builtin_amdgcn_update_dpp(0, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true), 0x1B, 0xF,
                                               0xF, true);
    case 8: return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, true);
    default: return __shfl_xor(v, j, 64);
  }
}

__device__ __forceinline__ double lane_xor_f64(double v, int j) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_xor_i32((int)(unsigned)(b & 0xffffffffLL), j);
  const int hi = lane_xor_i32((int)(b >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// wave64 sum (fixed butterfly order: partners at distance 32, 16, ..., 1); every lane
// gets the sum. The lane moves do not change the arithmetic, only how values travel.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += lane_xor_f64(v, off);
  return v;
}

// block sum for blockDim.x == BS (multiple of 64), result valid in all threads.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* lds /* >= BS/64 */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < BS / 64; ++w) t += lds[w];
  __syncthreads();
  return t;
}

// Fixed-order sum of `count` partials (stride `stride`) by one block of BS threads.
// The partials were written by an earlier kernel on the same stream.
template <int BS>
__device__ __forceinline__ double sum_partials(const double* part, int count, int stride, double* lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += BS) s += part[(int64_t)i * stride];
  return block_sum<BS>(s, lds);
}

}  // namespace dev
}  // namespace ngpb
