// Finite-volume assembly on the device (src/pb_class.cpp:1943-2159 and the bimpp
// operators it calls: bim3a_rhs, bim3a_laplacian_frac, bim3a_reaction_frac,
// bim3a_dirichlet_bc).
//
// Every matrix entry is a sum of "emissions": one per (leaf edge, parent pair) of the
// Laplacian and one per (leaf corner, parent) of the kappa^2 mass. The symbolic phase
// (once per mesh and Dirichlet set) writes the emissions of each row, sorts them by
// (column, emission id, sub-index) and cuts them into CSR slots; the numeric phase
// computes per-leaf conductances and sums each slot's emissions in that order. The
// order equals the sequential leaf sweep of the reference, so the sums are
// reproducible and equal to a sequential CPU assembly, and no floating-point atomics
// are used anywhere.
//
// Emission ids: edge e of leaf l -> 12 l + e; corner i of leaf l -> 12 L + 8 l + i.
// Weight code: bit 7 = negative, bits 0..6 = e with |w| = 2^-e (products of 1/np).
#include <algorithm>
#include <cstdlib>
#include <list>

#include "ngpb_internal.h"

namespace ngpb {

MeshView mesh_view_of(const ngpb_ctx* c);

namespace {

__constant__ int kEN[24] = {0, 1, 1, 3, 2, 3, 0, 2, 4, 5, 5, 7, 6, 7, 4, 6, 0, 4, 1, 5, 3, 7, 2, 6};
__constant__ int kEA[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};

constexpr double kPi = 3.14159265358979323846;
constexpr int kSvRows = 8;  // rows per block of the repacked slot emissions

__device__ __forceinline__ int ilog2_small(int v) { return v == 1 ? 0 : v == 2 ? 1 : v == 4 ? 2 : v == 8 ? 3 : 4; }

__device__ __forceinline__ void corner_parents(const MeshView& m, int64_t l, int i, int& np, int32_t par[4]) {
  const int32_t c = m.leaf_corner[8 * l + i];
  if (c >= 0) {
    np = 1;
    par[0] = c;
  } else {
    const int64_t h = -1 - (int64_t)c;
    np = m.hang_np[h];
    for (int j = 0; j < 4; ++j) par[j] = m.hang_par[4 * h + j];
  }
}

// Emissions (row, column, id, k, code), the contributions the numeric phase sums per
// matrix slot in (column, id, k) order:
//   edge terms: for leaf l, edge e = (loc[0], loc[1]) (kEN order) and each pair
//     (r, s) of its end corners, each parent pr of corner loc[r] (the row) and parent
//     ps of corner loc[s] (the column) give id = 12 l + e, k = the emission's place
//     in the edge's order (r, s) = (0,0), (0,1), (1,0), (1,1), then pr, then ps, and
//     code = (r != s ? 0x80 : 0) | log2(np[loc[r]] np[loc[s]]);
//   mass terms: corner i of leaf l gives, for each parent j, row = column = par_j,
//     id = 12 L + 8 l + i, k = j, code = log2 np[i].
// (id, k) is unique, so sorting a row by (column, id, k) fixes its order whatever the
// order of writing. One thread per (leaf, corner i) writes, for each parent p of
// corner i, every emission whose row is p -- the (i, i) and (i, j) terms of the 3
// edges through i and the mass term of i -- into one chunk reserved with a single
// atomic per (leaf, corner, parent); pass 0 only counts.
template <int PASS>
__device__ __forceinline__ void k_emit_rows_body(int64_t bid, MeshView m, unsigned* __restrict__ cnt, const int64_t* __restrict__ em_rowptr,
                            uint64_t* __restrict__ em_key, uint16_t* __restrict__ em_kw) {
  const int64_t t = bid * blockDim.x + threadIdx.x;
  if (t >= 8 * m.L) return;
  const int64_t l = t >> 3;
  const int i = (int)(t & 7);
  int npi;
  int32_t pi[4];
  corner_parents(m, l, i, npi, pi);
  // the 3 edges through corner i: edge id, position r of i in it, partner parents
  int eid[3], rpos[3], npj[3];
  int32_t pj[3][4];
  int ne = 0;
  for (int e = 0; e < 12; ++e) {
    const int a = kEN[2 * e], b = kEN[2 * e + 1];
    if (a != i && b != i) continue;
    eid[ne] = e;
    rpos[ne] = a == i ? 0 : 1;
    corner_parents(m, l, a == i ? b : a, npj[ne], pj[ne]);
    ++ne;
  }
  unsigned per_row = 1;
  for (int q = 0; q < 3; ++q) per_row += (unsigned)(npi + npj[q]);
  for (int pr = 0; pr < npi; ++pr) {
    const int32_t row = pi[pr];
    if (PASS == 0) {
      atomicAdd(&cnt[row], per_row);
      continue;
    }
    int64_t pos = em_rowptr[row] + atomicAdd(&cnt[row], per_row);
    for (int q = 0; q < 3; ++q) {
      const int r = rpos[q];
      const int npa = r == 0 ? npi : npj[q], npb = r == 0 ? npj[q] : npi;  // corners loc[0], loc[1]
      const uint32_t id = (uint32_t)(12 * l + eid[q]);
      for (int sgn = 0; sgn < 2; ++sgn) {
        // pairs before (r, sgn) in the order (0,0), (0,1), (1,0), (1,1)
        const int ns = sgn == 0 ? npa : npb;
        int off = 0;
        if (r == 1) off += npa * npa + npa * npb;
        if (sgn == 1) off += (r == 0 ? npa : npb) * npa;
        const bool self = sgn == r;  // column corner is i itself
        const int code = (self ? 0 : 0x80) | ilog2_small(npi * ns);
        for (int ps = 0; ps < ns; ++ps) {
          const int32_t colv = self ? pi[ps] : pj[q][ps];
          em_key[pos] = ((uint64_t)(uint32_t)colv << 32) | (uint64_t)id;
          em_kw[pos] = (uint16_t)(((off + pr * ns + ps) << 8) | code);
          ++pos;
        }
      }
    }
    em_key[pos] = ((uint64_t)(uint32_t)row << 32) | (uint64_t)(uint32_t)(12 * m.L + 8 * l + i);
    em_kw[pos] = (uint16_t)((pr << 8) | ilog2_small(npi));
  }
}

template <int PASS>
__global__ void k_emit_rows(MeshView m, unsigned* __restrict__ cnt, const int64_t* __restrict__ em_rowptr,
                            uint64_t* __restrict__ em_key, uint16_t* __restrict__ em_kw) {
  k_emit_rows_body<PASS>(blockIdx.x, m, cnt, em_rowptr, em_key, em_kw);
}

__device__ __forceinline__ void k_u32_to_i64_body(int64_t bid, int64_t n, const unsigned* __restrict__ a, int64_t* __restrict__ b) {
  const int64_t i = bid * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_u32_to_i64(int64_t n, const unsigned* __restrict__ a, int64_t* __restrict__ b) {
  k_u32_to_i64_body(blockIdx.x, n, a, b);
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ bool em_less(uint64_t ka, uint16_t wa, uint64_t kb, uint16_t wb) {
  return ka < kb || (ka == kb && (wa >> 8) < (wb >> 8));
}

// Each row's emissions are sorted by (column, id, k) -- unique within a row -- and its
// distinct columns counted as matrix slots or Dirichlet couplings. Rows of at most 64
// emissions: one wave per row (k_row_sort_wave). Longer rows (parents of hanging
// nodes) are listed and sorted by one block per row with up to kSortBigCap entries
// in LDS; beyond that, an insertion sort by one lane.
constexpr int kSortBS = 256;
constexpr int kSortBigCap = 2048;  // 20 KB of LDS per block

// class of a row's emission with column col: 0 a matrix slot, 1 a coupling to a
// Dirichlet column (moved to the right-hand side), 2 dropped (off-diagonal of a
// Dirichlet row)
__device__ __forceinline__ int em_class(int64_t row, bool rb, int64_t col, const uint8_t* __restrict__ is_bnd) {
  if (rb) return col == row ? 0 : 2;
  return is_bnd[col] ? 1 : 0;
}

// per-row counts of a sorted run: slots and couplings (distinct columns) and the
// emissions that belong to each
struct RowCounts {
  int64_t slots = 0, couplings = 0, slot_em = 0, coupling_em = 0;
  __device__ void add(int cls, bool new_col) {
    if (cls == 0) {
      ++slot_em;
      slots += new_col ? 1 : 0;
    } else if (cls == 1) {
      ++coupling_em;
      couplings += new_col ? 1 : 0;
    }
  }
};

__device__ __forceinline__ bool em_before(uint64_t ka, int wa, uint64_t kb, int wb) {
  return ka < kb || (ka == kb && wa < wb);
}

// v of lane (lane ^ j) for j in {1, 2, 4, 8, 16, 32}: DPP lane moves inside each row of
// 16 lanes (quad permutations for 1 and 2, a half-row mirror then a quad reversal for
// 4, a rotation by 8 for 8), the LDS permute across rows (16, 32). With the sort loops
// unrolled, j is a constant and the switch folds away.
__device__ __forceinline__ unsigned lane_xor(unsigned v, int j) {
  switch (j) {
    case 1: return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);   // quad_perm [1,0,3,2]
    case 2: return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);   // quad_perm [2,3,0,1]
    case 4: {
      const int m = __builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, true);         // row_half_mirror
      return (unsigned)__builtin_amdgcn_update_dpp(0, m, 0x1B, 0xF, 0xF, true);            // quad_perm [3,2,1,0]
    }
    case 8: return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, true);  // row_ror:8
    default: return __shfl_xor(v, j, 64);
  }
}

// v of lane i - 1 (DPP wave shift); lane 0 keeps its own
__device__ __forceinline__ uint64_t lane_prev_u64(uint64_t v) {
  const unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  const unsigned plo = (unsigned)__builtin_amdgcn_update_dpp((int)lo, (int)lo, 0x138, 0xF, 0xF, false);
  const unsigned phi = (unsigned)__builtin_amdgcn_update_dpp((int)hi, (int)hi, 0x138, 0xF, 0xF, false);
  return ((uint64_t)phi << 32) | plo;
}

// v of lane 63
__device__ __forceinline__ uint64_t lane_last_u64(uint64_t v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t lane_xor_u64(uint64_t v, int j) {
  return ((uint64_t)lane_xor((unsigned)(v >> 32), j) << 32) | lane_xor((unsigned)v, j);
}

// Rows of at most 64 emissions: one wave per row, one emission per lane, sorted by a
// bitonic network over the lanes (unused lanes hold keys that sort last). Longer rows
// are listed for k_row_sort_block.
__device__ __forceinline__ void k_row_sort_wave_body(int64_t bid, int64_t N, const int64_t* __restrict__ em_rowptr,
                                                           uint64_t* __restrict__ key, uint16_t* __restrict__ kw,
                                                           const uint8_t* __restrict__ is_bnd, int64_t* __restrict__ cnt4,
                                                           int32_t* __restrict__ big_rows, int* __restrict__ n_big) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row = bid * (kSortBS / 64) + wv;
  if (row >= N) return;  // the whole wave
  const int64_t b = em_rowptr[row], e = em_rowptr[row + 1];
  const int E = (int)(e - b);
  if (E > 64) {
    if (lane == 0) big_rows[atomicAdd(n_big, 1)] = (int32_t)row;
    return;
  }
  uint64_t kk = lane < E ? key[b + lane] : ~0ULL;
  unsigned ww = lane < E ? (unsigned)kw[b + lane] : 0xffffu;
#pragma unroll
  for (int w = 2; w <= 64; w <<= 1)
#pragma unroll
    for (int j = w >> 1; j > 0; j >>= 1) {
      const uint64_t ok = lane_xor_u64(kk, j);
      const unsigned ow = lane_xor(ww, j);
      const bool keep_min = ((lane & j) == 0) == ((lane & w) == 0);
      const bool take = keep_min ? em_before(ok, (int)(ow >> 8), kk, (int)(ww >> 8))
                                 : em_before(kk, (int)(ww >> 8), ok, (int)(ow >> 8));
      if (take) {
        kk = ok;
        ww = ow;
      }
    }
  const uint64_t prev = lane_prev_u64(kk);
  RowCounts rc;
  if (lane < E) {
    key[b + lane] = kk;
    kw[b + lane] = (uint16_t)ww;
    const int64_t col = (int64_t)(kk >> 32);
    rc.add(em_class(row, is_bnd[row] != 0, col, is_bnd), lane == 0 || (int64_t)(prev >> 32) != col);
  }
  // the four counts of a lane are 0 or 1 and a row has at most 64 emissions, so they
  // are summed over the wave as four 16-bit fields of one word
  uint64_t packed = (uint64_t)rc.slots | ((uint64_t)rc.couplings << 16) | ((uint64_t)rc.slot_em << 32) |
                    ((uint64_t)rc.coupling_em << 48);
#pragma unroll
  for (int j = 32; j > 0; j >>= 1) packed += lane_xor_u64(packed, j);
  if (lane < 4) cnt4[(int64_t)lane * N + row] = (int64_t)((packed >> (16 * lane)) & 0xffffu);
}

__global__ void __launch_bounds__(kSortBS) k_row_sort_wave(int64_t N, const int64_t* __restrict__ em_rowptr,
                                                           uint64_t* __restrict__ key, uint16_t* __restrict__ kw,
                                                           const uint8_t* __restrict__ is_bnd, int64_t* __restrict__ cnt4,
                                                           int32_t* __restrict__ big_rows, int* __restrict__ n_big) {
  k_row_sort_wave_body(blockIdx.x, N, em_rowptr, key, kw, is_bnd, cnt4, big_rows, n_big);
}

// one block per listed long row
__device__ __forceinline__ void k_row_sort_block_body(int64_t bid, const int32_t* __restrict__ rows,
                                                            const int64_t* __restrict__ em_rowptr,
                                                            uint64_t* __restrict__ key, uint16_t* __restrict__ kw,
                                                            const uint8_t* __restrict__ is_bnd, int64_t N,
                                                            int64_t* __restrict__ cnt4) {
  extern __shared__ uint64_t s_dyn[];  // key[cap], then kw[cap]
  __shared__ int64_t s_red[4][kSortBS / 64];
  const int64_t row = rows[bid];
  const int64_t b = em_rowptr[row], e = em_rowptr[row + 1];
  const int E = (int)(e - b);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const bool rb = is_bnd[row] != 0;
  RowCounts rc;
  if (E <= kSortBigCap) {
    // bitonic network over the next power of two, padded with keys that sort last
    int P = 64;
    while (P < E) P <<= 1;
    uint64_t* sk = s_dyn;
    uint16_t* sw = reinterpret_cast<uint16_t*>(s_dyn + kSortBigCap);
    for (int i = tid; i < P; i += kSortBS) {
      sk[i] = i < E ? key[b + i] : ~0ULL;
      sw[i] = i < E ? kw[b + i] : (uint16_t)0xffff;
    }
    __syncthreads();
    for (int w = 2; w <= P; w <<= 1)
      for (int j = w >> 1; j > 0; j >>= 1) {
        for (int i = tid; i < P; i += kSortBS) {
          const int x = i ^ j;
          if (x <= i) continue;
          const uint64_t ka = sk[i], kb = sk[x];
          const uint16_t wa = sw[i], wb = sw[x];
          if (em_before(kb, wb >> 8, ka, wa >> 8) == ((i & w) == 0)) {
            sk[i] = kb;
            sk[x] = ka;
            sw[i] = wb;
            sw[x] = wa;
          }
        }
        __syncthreads();
      }
    for (int i = tid; i < E; i += kSortBS) {
      const uint64_t kk = sk[i];
      key[b + i] = kk;
      kw[b + i] = sw[i];
      const int64_t col = (int64_t)(kk >> 32);
      rc.add(em_class(row, rb, col, is_bnd), i == 0 || (int64_t)(sk[i - 1] >> 32) != col);
    }
  } else if (tid == 0) {
    for (int64_t i = b + 1; i < e; ++i) {
      const uint64_t kk = key[i];
      const uint16_t ww = kw[i];
      int64_t j = i - 1;
      while (j >= b && em_less(kk, ww, key[j], kw[j])) {
        key[j + 1] = key[j];
        kw[j + 1] = kw[j];
        --j;
      }
      key[j + 1] = kk;
      kw[j + 1] = ww;
    }
    int64_t prev = -1;
    for (int64_t i = b; i < e; ++i) {
      const int64_t col = (int64_t)(key[i] >> 32);
      rc.add(em_class(row, rb, col, is_bnd), col != prev);
      prev = col;
    }
  }
  const int64_t v[4] = {wave_sum_i64(rc.slots), wave_sum_i64(rc.couplings), wave_sum_i64(rc.slot_em),
                        wave_sum_i64(rc.coupling_em)};
  if (lane == 0)
    for (int q = 0; q < 4; ++q) s_red[q][wv] = v[q];
  __syncthreads();
  if (tid < 4) {
    int64_t a = 0;
    for (int w = 0; w < kSortBS / 64; ++w) a += s_red[tid][w];
    cnt4[(int64_t)tid * N + row] = a;
  }
}

__global__ void __launch_bounds__(kSortBS) k_row_sort_block(const int32_t* __restrict__ rows,
                                                            const int64_t* __restrict__ em_rowptr,
                                                            uint64_t* __restrict__ key, uint16_t* __restrict__ kw,
                                                            const uint8_t* __restrict__ is_bnd, int64_t N,
                                                            int64_t* __restrict__ cnt4) {
  k_row_sort_block_body(blockIdx.x, rows, em_rowptr, key, kw, is_bnd, N, cnt4);
}

// One wave per row over its sorted emissions, 64 at a time: the matrix slots (col,
// diag_slot, sptr = first emission of the slot in the repacked list) and the couplings
// to Dirichlet columns (bc_col, bc_ptr), and the emissions of each, repacked to 5 bytes
// (source id, weight code) in slot order: each class compacted by ballot ranks from
// the row's offsets (prefix sums of the per-row counts).
constexpr int kSlotBS = 256;

__device__ __forceinline__ void k_row_slots_body(int64_t bid, 
    int64_t N, const int64_t* __restrict__ em_rowptr, const uint64_t* __restrict__ key,
    const uint16_t* __restrict__ kw, const uint8_t* __restrict__ is_bnd, const int64_t* __restrict__ off4,
    int32_t* __restrict__ rowptr, int32_t* __restrict__ col, int32_t* __restrict__ diag_slot,
    int32_t* __restrict__ bc_rowptr, int32_t* __restrict__ bc_col, int64_t* __restrict__ sptr,
    uint32_t* __restrict__ em_id, uint8_t* __restrict__ em_code, int64_t* __restrict__ bc_ptr,
    uint32_t* __restrict__ bc_id, uint8_t* __restrict__ bc_code, int64_t n_slots, int64_t n_coup, int64_t n_slot_em,
    int64_t n_coup_em) {
  const int lane = threadIdx.x & 63;
  const int64_t row = bid * (kSlotBS / 64) + (threadIdx.x >> 6);
  if (row >= N) return;  // the whole wave
  const int64_t b = em_rowptr[row], e = em_rowptr[row + 1];
  const bool rb = is_bnd[row] != 0;
  int64_t fs = off4[row], bs = off4[N + row], fe = off4[2 * N + row], be = off4[3 * N + row];
  if (lane == 0) {
    rowptr[row] = (int32_t)fs;
    bc_rowptr[row] = (int32_t)bs;
    if (row == N - 1) {
      rowptr[N] = (int32_t)n_slots;
      bc_rowptr[N] = (int32_t)n_coup;
      sptr[n_slots] = n_slot_em;
      bc_ptr[n_coup] = n_coup_em;
    }
  }
  const uint64_t below = (1ULL << lane) - 1;
  int64_t lastcol = -1;
  for (int64_t c0 = b; c0 < e; c0 += 64) {
    const int64_t i = c0 + lane;
    const bool valid = i < e;
    const uint64_t kk = valid ? key[i] : 0;
    const int64_t cl = valid ? (int64_t)(kk >> 32) : -1;
    const int64_t up = (int64_t)lane_prev_u64((uint64_t)cl);
    const int64_t prev = lane == 0 ? lastcol : up;
    const bool start = valid && cl != prev;
    const int cls = valid ? em_class(row, rb, cl, is_bnd) : 3;
    const uint64_t mf = __ballot(cls == 0), mb = __ballot(cls == 1);
    const uint64_t sf = __ballot(start && cls == 0), sb = __ballot(start && cls == 1);
    if (cls == 0) {
      const int64_t o = fe + __popcll(mf & below);
      em_id[o] = (uint32_t)(kk & 0xffffffffULL);
      em_code[o] = (uint8_t)(kw[i] & 0xff);
      if (start) {
        const int64_t sl = fs + __popcll(sf & below);
        col[sl] = (int32_t)cl;
        sptr[sl] = o;
        if (cl == row) diag_slot[row] = (int32_t)sl;
      }
    } else if (cls == 1) {
      const int64_t o = be + __popcll(mb & below);
      bc_id[o] = (uint32_t)(kk & 0xffffffffULL);
      bc_code[o] = (uint8_t)(kw[i] & 0xff);
      if (start) {
        const int64_t sl = bs + __popcll(sb & below);
        bc_col[sl] = (int32_t)cl;
        bc_ptr[sl] = o;
      }
    }
    fe += __popcll(mf);
    be += __popcll(mb);
    fs += __popcll(sf);
    bs += __popcll(sb);
    lastcol = (int64_t)lane_last_u64((uint64_t)cl);
  }
}

__global__ void __launch_bounds__(kSlotBS) k_row_slots(
    int64_t N, const int64_t* __restrict__ em_rowptr, const uint64_t* __restrict__ key,
    const uint16_t* __restrict__ kw, const uint8_t* __restrict__ is_bnd, const int64_t* __restrict__ off4,
    int32_t* __restrict__ rowptr, int32_t* __restrict__ col, int32_t* __restrict__ diag_slot,
    int32_t* __restrict__ bc_rowptr, int32_t* __restrict__ bc_col, int64_t* __restrict__ sptr,
    uint32_t* __restrict__ em_id, uint8_t* __restrict__ em_code, int64_t* __restrict__ bc_ptr,
    uint32_t* __restrict__ bc_id, uint8_t* __restrict__ bc_code, int64_t n_slots, int64_t n_coup, int64_t n_slot_em,
    int64_t n_coup_em) {
  k_row_slots_body(blockIdx.x, N, em_rowptr, key, kw, is_bnd, off4, rowptr, col, diag_slot, bc_rowptr, bc_col, sptr, em_id, em_code, bc_ptr, bc_id, bc_code, n_slots, n_coup, n_slot_em, n_coup_em);
}

// blk_em[b] = first emission of the first slot of rows kSvRows b ..; blk_em[nblk] = total
__device__ __forceinline__ void k_blk_em_body(int64_t bid, int64_t nblk, int64_t N, const int32_t* __restrict__ rowptr, const int64_t* __restrict__ sptr,
                         int64_t* __restrict__ blk_em) {
  const int64_t b = bid * blockDim.x + threadIdx.x;
  if (b <= nblk) blk_em[b] = sptr[rowptr[min(N, b * kSvRows)]];
}

__global__ void k_blk_em(int64_t nblk, int64_t N, const int32_t* __restrict__ rowptr, const int64_t* __restrict__ sptr,
                         int64_t* __restrict__ blk_em) {
  k_blk_em_body(blockIdx.x, nblk, N, rowptr, sptr, blk_em);
}

__device__ __forceinline__ void k_slot_rel_body(int64_t bid, int64_t N, const int32_t* __restrict__ rowptr, const int64_t* __restrict__ sptr,
                           const int64_t* __restrict__ blk_em, uint32_t* __restrict__ rel) {
  const int64_t row = bid * blockDim.x + threadIdx.x;
  if (row >= N) return;
  const int64_t base = blk_em[row / kSvRows];
  for (int64_t sl = rowptr[row]; sl < rowptr[row + 1]; ++sl) rel[sl] = (uint32_t)(sptr[sl] - base);
}

__global__ void k_slot_rel(int64_t N, const int32_t* __restrict__ rowptr, const int64_t* __restrict__ sptr,
                           const int64_t* __restrict__ blk_em, uint32_t* __restrict__ rel) {
  k_slot_rel_body(blockIdx.x, N, rowptr, sptr, blk_em, rel);
}

// ------------------------------------------------------------------------ numeric
__device__ __forceinline__ double wha(double eps1, double eps2, double frac) {
  return 1.0 / (frac / eps1 + (1 - frac) / eps2);
}

__device__ __forceinline__ void k_leaf_coeffs_body(int64_t bid, MeshView m, const double* __restrict__ eps, const int32_t* __restrict__ border_of,
                              const double* __restrict__ frac, double* __restrict__ c_edge,
                              double* __restrict__ v8) {
  const int64_t l = bid * blockDim.x + threadIdx.x;
  if (l >= m.L) return;
  const int32_t c0 = m.leaf_corner[8 * l], c7 = m.leaf_corner[8 * l + 7];
  double h[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) h[d] = dev::corner_coord(m, c7, d) - dev::corner_coord(m, c0, d);
  double ec[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int32_t c = m.leaf_corner[8 * l + i];
    if (c >= 0) {
      ec[i] = eps[c];
    } else {
      const int64_t hh = -1 - (int64_t)c;
      const int np = m.hang_np[hh];
      double s = 0.0;
      for (int j = 0; j < np; ++j) s += eps[m.hang_par[4 * hh + j]] / np;
      ec[i] = s;
    }
  }
  const int32_t bi = border_of[l];
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const int a = kEA[e], b = (a + 1) % 3, cc = (a + 2) % 3;
    const double f = bi >= 0 ? frac[12 * (int64_t)bi + e] : 0.5;
    c_edge[12 * l + e] = wha(ec[kEN[2 * e]], ec[kEN[2 * e + 1]], f) * (h[b] * h[cc] / (4.0 * h[a]));
  }
  v8[l] = h[0] * h[1] * h[2] * 0.125;
}

__global__ void k_leaf_coeffs(MeshView m, const double* __restrict__ eps, const int32_t* __restrict__ border_of,
                              const double* __restrict__ frac, double* __restrict__ c_edge,
                              double* __restrict__ v8) {
  k_leaf_coeffs_body(blockIdx.x, m, eps, border_of, frac, c_edge, v8);
}

__device__ __forceinline__ double decode_code(uint8_t code) {
  const double mag = ldexp(1.0, -(code & 0x7f));
  return (code & 0x80) ? -mag : mag;
}

__device__ __forceinline__ double emission_value(uint32_t id, uint8_t code, int64_t row, int64_t L,
                                                 const double* c_edge, const double* reac, const double* v8) {
  const double w = decode_code(code);
  if ((int64_t)id < 12 * L) return c_edge[id] * w;
  const int64_t leaf = ((int64_t)id - 12 * L) >> 3;
  return reac[row] * v8[leaf] * w;
}

// Matrix values. A block (one wave) owns kSvRows consecutive rows (the blocks of the
// repacked slot emissions): their emissions (contiguous, in slot order) are loaded
// coalesced and evaluated in parallel into LDS, then one thread per CSR slot adds its
// run of values in order, so every sum equals the sequential sum of the reference's
// leaf sweep. Blocks whose rows hold more than kSvCap emissions evaluate straight from
// global memory. Small blocks with 4 KB of LDS keep many waves resident; on the 8M
// globular mesh (8 rows, 512 entries, 64 threads) the numeric assembly takes 4.6 ms
// against 6.4 ms with 32-row blocks of 256 threads and 32 KB of LDS, with identical
// results.
constexpr int kSvCap = 512;
constexpr int kSvBS = 64;

// last r in [0, nr) with a[r] <= v (a[0] <= v)
__device__ __forceinline__ int last_le(const uint32_t* a, int nr, uint32_t v) {
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void k_slot_values_body(int64_t bid, int64_t N, int64_t L, const int32_t* __restrict__ rowptr,
                                                       const int64_t* __restrict__ blk_em,
                                                       const uint32_t* __restrict__ slot_rel,
                                                       const uint32_t* __restrict__ em_id,
                                                       const uint8_t* __restrict__ em_code,
                                                       const double* __restrict__ c_edge,
                                                       const double* __restrict__ reac,
                                                       const double* __restrict__ v8, double* __restrict__ val) {
  __shared__ double s_v[kSvCap];
  __shared__ uint32_t s_re[kSvRows + 1];  // per row: first emission (block-relative)
  __shared__ uint32_t s_rs[kSvRows + 1];  // per row: first slot (block-relative)
  const int64_t blk = bid;
  const int64_t r0 = blk * kSvRows;
  const int64_t r1 = min(r0 + kSvRows, N);
  const int nr = (int)(r1 - r0);
  const int64_t base = blk_em[blk];
  const uint32_t nem = (uint32_t)(blk_em[blk + 1] - base);
  const int64_t sl0 = rowptr[r0], sl1 = rowptr[r1];
  for (int i = threadIdx.x; i < nr; i += kSvBS) {
    const int64_t sl = rowptr[r0 + i];
    s_re[i] = sl < sl1 ? slot_rel[sl] : nem;
    s_rs[i] = (uint32_t)(sl - sl0);
  }
  if (threadIdx.x == 0) {
    s_re[nr] = nem;
    s_rs[nr] = (uint32_t)(sl1 - sl0);
  }
  __syncthreads();
  const bool staged = nem <= (uint32_t)kSvCap;
  if (staged) {
    for (uint32_t i = threadIdx.x; i < nem; i += kSvBS) {
      const uint32_t id = em_id[base + i];
      const double w = decode_code(em_code[base + i]);
      double v;
      if ((int64_t)id < 12 * L) {
        v = c_edge[id] * w;
      } else {
        const int r = last_le(s_re, nr, i);  // lumped mass: the row's own kappa^2
        v = reac[r0 + r] * v8[((int64_t)id - 12 * L) >> 3] * w;
      }
      s_v[i] = v;
    }
  }
  __syncthreads();
  for (int64_t sl = sl0 + threadIdx.x; sl < sl1; sl += kSvBS) {
    const uint32_t b = slot_rel[sl], e = sl + 1 < sl1 ? slot_rel[sl + 1] : nem;
    double v = 0.0;
    if (staged) {
      for (uint32_t i = b; i < e; ++i) v += s_v[i];
    } else {
      const int64_t row = r0 + last_le(s_rs, nr, (uint32_t)(sl - sl0));
      for (uint32_t i = b; i < e; ++i) v += emission_value(em_id[base + i], em_code[base + i], row, L, c_edge, reac, v8);
    }
    val[sl] = v;
  }
}

__global__ void __launch_bounds__(kSvBS) k_slot_values(int64_t N, int64_t L, const int32_t* __restrict__ rowptr,
                                                       const int64_t* __restrict__ blk_em,
                                                       const uint32_t* __restrict__ slot_rel,
                                                       const uint32_t* __restrict__ em_id,
                                                       const uint8_t* __restrict__ em_code,
                                                       const double* __restrict__ c_edge,
                                                       const double* __restrict__ reac,
                                                       const double* __restrict__ v8, double* __restrict__ val) {
  k_slot_values_body(blockIdx.x, N, L, rowptr, blk_em, slot_rel, em_id, em_code, c_edge, reac, v8, val);
}

// Charge projection (src/pb_class.cpp:2002-2039). Contribution t = 32 a + 4 ii + jj
// (atom a, leaf corner ii, parent jj); each valid one is pushed onto the list of its
// node.
__device__ __forceinline__ void k_rho_push_body(int64_t bid, MeshView m, int64_t A, const int32_t* __restrict__ owner, int32_t* __restrict__ cnode,
                           int32_t* __restrict__ cnext, int32_t* __restrict__ head) {
  const int64_t t = bid * blockDim.x + threadIdx.x;
  if (t >= 32 * A) return;
  const int64_t a = t >> 5;
  const int ii = (int)((t >> 2) & 7), jj = (int)(t & 3);
  cnode[t] = -1;
  const int64_t l = owner[a];
  if (l < 0) return;
  const int32_t c = m.leaf_corner[8 * l + ii];
  int32_t node;
  if (c >= 0) {
    if (jj != 0) return;
    node = c;
  } else {
    const int64_t h = -1 - (int64_t)c;
    if (jj >= m.hang_np[h]) return;
    node = m.hang_par[4 * h + jj];
  }
  cnode[t] = node;
  cnext[t] = atomicExch(&head[node], (int32_t)t);
}

__global__ void k_rho_push(MeshView m, int64_t A, const int32_t* __restrict__ owner, int32_t* __restrict__ cnode,
                           int32_t* __restrict__ cnext, int32_t* __restrict__ head) {
  k_rho_push_body(blockIdx.x, m, A, owner, cnode, cnext, head);
}

// One thread per touched node (the head of its list): the patch volume (the node's
// lumped-mass emissions with unit coefficient, in order), the density
// rho = sum over its contributions in increasing t of q 4 pi w / vol_patch
// (/ (np vol_patch) through a hanging corner), and the right-hand side
// sum over the same emissions of (V/8 w) rho. Untouched nodes keep rhs 0.
__device__ __forceinline__ void k_rho_node_body(int64_t bid, MeshView m, int64_t C, int64_t L, const double* __restrict__ axyz,
                           const double* __restrict__ aq, const int32_t* __restrict__ owner,
                           const int32_t* __restrict__ cnode, const int32_t* __restrict__ cnext,
                           const int32_t* __restrict__ head, const int32_t* __restrict__ diag_slot,
                           int64_t N, const int32_t* __restrict__ rowptr, const int64_t* __restrict__ blk_em,
                           const uint32_t* __restrict__ slot_rel, const uint32_t* __restrict__ em_id,
                           const uint8_t* __restrict__ em_code, const double* __restrict__ v8,
                           double* __restrict__ rho, double* __restrict__ rhs_charge) {
  const int64_t t = bid * blockDim.x + threadIdx.x;
  if (t >= C) return;
  const int32_t node = cnode[t];
  if (node < 0 || head[node] != (int32_t)t) return;
  const int64_t ds = diag_slot[node];
  const int64_t blk = node / kSvRows;
  const int64_t last_slot = rowptr[min(N, (blk + 1) * kSvRows)];
  const int64_t b = blk_em[blk] + slot_rel[ds];
  const int64_t e = ds + 1 < last_slot ? blk_em[blk] + slot_rel[ds + 1] : blk_em[blk + 1];
  double vp = 0.0;
  for (int64_t i = b; i < e; ++i) {
    const uint32_t id = em_id[i];
    if ((int64_t)id < 12 * L) continue;
    vp += v8[((int64_t)id - 12 * L) >> 3] * decode_code(em_code[i]);
  }
  double s = 0.0;
  int64_t last = -1;
  while (true) {
    int64_t best = INT64_MAX;
    for (int32_t u = head[node]; u >= 0; u = cnext[u])
      if ((int64_t)u > last && (int64_t)u < best) best = u;
    if (best == INT64_MAX) break;
    last = best;
    const int64_t a = best >> 5;
    const int ii = (int)((best >> 2) & 7);
    const int64_t l = owner[a];
    const int32_t c = m.leaf_corner[8 * l + ii];
    const int32_t c0 = m.leaf_corner[8 * l], c7 = m.leaf_corner[8 * l + 7];
    const double volume = (dev::corner_coord(m, c7, 0) - dev::corner_coord(m, c0, 0)) *
                          (dev::corner_coord(m, c7, 1) - dev::corner_coord(m, c0, 1)) *
                          (dev::corner_coord(m, c7, 2) - dev::corner_coord(m, c0, 2));
    const int32_t co = m.leaf_corner[8 * l + 7 - ii];
    const double weight = fabs((axyz[3 * a] - dev::corner_coord(m, co, 0)) *
                               (axyz[3 * a + 1] - dev::corner_coord(m, co, 1)) *
                               (axyz[3 * a + 2] - dev::corner_coord(m, co, 2))) / volume;
    if (c >= 0) {
      s += aq[a] * 4.0 * kPi * weight / vp;
    } else {
      const int npi = m.hang_np[-1 - (int64_t)c];
      const double denom = npi * vp;
      s += aq[a] * 4.0 * kPi * weight / denom;
    }
  }
  rho[node] = s;
  double r = 0.0;
  for (int64_t i = b; i < e; ++i) {
    const uint32_t id = em_id[i];
    if ((int64_t)id < 12 * L) continue;
    r += v8[((int64_t)id - 12 * L) >> 3] * decode_code(em_code[i]) * s;
  }
  rhs_charge[node] = r;
}

__global__ void k_rho_node(MeshView m, int64_t C, int64_t L, const double* __restrict__ axyz,
                           const double* __restrict__ aq, const int32_t* __restrict__ owner,
                           const int32_t* __restrict__ cnode, const int32_t* __restrict__ cnext,
                           const int32_t* __restrict__ head, const int32_t* __restrict__ diag_slot,
                           int64_t N, const int32_t* __restrict__ rowptr, const int64_t* __restrict__ blk_em,
                           const uint32_t* __restrict__ slot_rel, const uint32_t* __restrict__ em_id,
                           const uint8_t* __restrict__ em_code, const double* __restrict__ v8,
                           double* __restrict__ rho, double* __restrict__ rhs_charge) {
  k_rho_node_body(blockIdx.x, m, C, L, axyz, aq, owner, cnode, cnext, head, diag_slot, N, rowptr, blk_em, slot_rel, em_id, em_code, v8, rho, rhs_charge);
}

__device__ __forceinline__ void k_dirichlet_rhs_body(int64_t bid, int64_t N, int64_t L, const uint8_t* __restrict__ is_bnd,
                                const double* __restrict__ gval, const int32_t* __restrict__ diag_slot,
                                const double* __restrict__ val, const int32_t* __restrict__ bc_rowptr,
                                const int32_t* __restrict__ bc_col, const int64_t* __restrict__ bc_ptr,
                                const uint32_t* __restrict__ bc_id, const uint8_t* __restrict__ bc_code,
                                const double* __restrict__ c_edge, const double* __restrict__ reac,
                                const double* __restrict__ v8, const double* __restrict__ rhs_charge,
                                double* __restrict__ rhs) {
  const int64_t row = bid * blockDim.x + threadIdx.x;
  if (row >= N) return;
  if (is_bnd[row]) {
    rhs[row] = val[diag_slot[row]] * gval[row];
    return;
  }
  double r = rhs_charge[row];
  for (int64_t s = bc_rowptr[row]; s < bc_rowptr[row + 1]; ++s) {
    double a = 0.0;
    for (int64_t i = bc_ptr[s]; i < bc_ptr[s + 1]; ++i) a += emission_value(bc_id[i], bc_code[i], row, L, c_edge, reac, v8);
    r -= a * gval[bc_col[s]];
  }
  rhs[row] = r;
}

__global__ void k_dirichlet_rhs(int64_t N, int64_t L, const uint8_t* __restrict__ is_bnd,
                                const double* __restrict__ gval, const int32_t* __restrict__ diag_slot,
                                const double* __restrict__ val, const int32_t* __restrict__ bc_rowptr,
                                const int32_t* __restrict__ bc_col, const int64_t* __restrict__ bc_ptr,
                                const uint32_t* __restrict__ bc_id, const uint8_t* __restrict__ bc_code,
                                const double* __restrict__ c_edge, const double* __restrict__ reac,
                                const double* __restrict__ v8, const double* __restrict__ rhs_charge,
                                double* __restrict__ rhs) {
  k_dirichlet_rhs_body(blockIdx.x, N, L, is_bnd, gval, diag_slot, val, bc_rowptr, bc_col, bc_ptr, bc_id, bc_code, c_edge, reac, v8, rhs_charge, rhs);
}

// ------------------------------------------------------------------ Dirichlet data
__global__ void k_bc_values(int64_t nb, const int32_t* __restrict__ nodes, const double* __restrict__ xyz, int bc_type,
                            double eps_in, double eps_out, double k, int64_t A, const double* __restrict__ axyz,
                            const double* __restrict__ aq, double* __restrict__ bval, double* __restrict__ gval) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb) return;
  const int64_t nd = nodes[t];
  const double x = xyz[3 * nd], y = xyz[3 * nd + 1], z = xyz[3 * nd + 2];
  double v = 0.0;
  if (bc_type == 2) {
    // coulomb_boundary_conditions (src/pb_class.cpp:3530-3547) over the charges
    for (int64_t a = 0; a < A; ++a) {
      const double dx = axyz[3 * a] - x, dy = axyz[3 * a + 1] - y, dz = axyz[3 * a + 2] - z;
      const double dist = sqrt(dx * dx + dy * dy + dz * dz);
      v += aq[a] * exp(-k * dist) / (dist * eps_out);
    }
  } else if (bc_type == 3) {
    // analytic_solution (src/pb_class.cpp:3549-3571): q = +1, rs = 2
    const double rs = 2.0;
    const double pout = eps_out * (1.0 + k * rs), pin = eps_in * rs;
    const double dist = sqrt(x * x + y * y + z * z);
    v = dist <= rs ? 1.0 / (pout * rs) + (rs - dist) / (pin * dist) : exp(k * (rs - dist)) / (dist * pout);
  }
  bval[t] = v;
  gval[nd] = v;
}

// ------------------------------------------------------------- system tables
// One launch per assembly stage over many systems: system s owns the blocks
// [blk0[s], blk0[s + 1]) and runs the single-system kernel's body on them as its
// blocks 0, 1, ..., so its results are those of its own launches, bit for bit.
struct AsmSys {
  MeshView m;
  int64_t N = 0, L = 0, A = 0;
  // symbolic temporaries and counts
  unsigned* cnt = nullptr;
  int64_t* em_rowptr = nullptr;  // N + 1
  uint64_t* em_key = nullptr;
  uint16_t* em_kw = nullptr;
  int64_t* cnt4 = nullptr;  // 4 N
  int32_t* big = nullptr;   // N
  int* nbig = nullptr;
  int64_t* sptr = nullptr;  // nnz + 1
  int64_t n_em = 0, n_slots = 0, n_coup = 0, n_slot_em = 0, n_coup_em = 0;
  // the system's own arrays
  const uint8_t* is_bnd = nullptr;
  int32_t *rowptr = nullptr, *col = nullptr, *diag_slot = nullptr, *bc_rowptr = nullptr, *bc_col = nullptr;
  int64_t *bc_ptr = nullptr, *blk_em = nullptr;
  uint32_t *em_id = nullptr, *bc_id = nullptr, *slot_rel = nullptr;
  uint8_t *em_code = nullptr, *bc_code = nullptr;
  // numeric
  const double *eps = nullptr, *frac = nullptr, *reac = nullptr, *gval = nullptr, *axyz = nullptr, *aq = nullptr;
  const int32_t *border_of = nullptr, *owner = nullptr;
  double *c_edge = nullptr, *v8 = nullptr, *val = nullptr, *rho = nullptr, *rhs_charge = nullptr, *rhs = nullptr;
  int32_t *cnode = nullptr, *cnext = nullptr, *head = nullptr;
};

#define NGPB_SYS                                      \
  const int s_ = group_of(blk0, ns, blockIdx.x);     \
  const AsmSys& S = T[s_];                           \
  const int64_t bid = (int64_t)blockIdx.x - blk0[s_]

template <int PASS>
__global__ void kg_emit_rows(const AsmSys* __restrict__ T, const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_emit_rows_body<PASS>(bid, S.m, S.cnt, S.em_rowptr, S.em_key, S.em_kw);
}

__global__ void kg_u32_to_i64(const AsmSys* __restrict__ T, const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_u32_to_i64_body(bid, S.N, S.cnt, S.em_rowptr);
}

// one thread per system: the end of its emission row pointers
__global__ void kg_em_end(const AsmSys* __restrict__ T, int ns) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < ns) T[s].em_rowptr[T[s].N] = T[s].n_em;
}

__global__ void __launch_bounds__(kSortBS) kg_row_sort_wave(const AsmSys* __restrict__ T,
                                                            const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_row_sort_wave_body(bid, S.N, S.em_rowptr, S.em_key, S.em_kw, S.is_bnd, S.cnt4, S.big, S.nbig);
}

__global__ void __launch_bounds__(kSortBS) kg_row_sort_block(const AsmSys* __restrict__ T,
                                                             const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_row_sort_block_body(bid, S.big, S.em_rowptr, S.em_key, S.em_kw, S.is_bnd, S.N, S.cnt4);
}

__global__ void __launch_bounds__(kSlotBS) kg_row_slots(const AsmSys* __restrict__ T,
                                                        const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  const int64_t N = S.N;
  k_row_slots_body(bid, N, S.em_rowptr, S.em_key, S.em_kw, S.is_bnd, S.cnt4, S.rowptr, S.col, S.diag_slot,
                   S.bc_rowptr, S.bc_col, S.sptr, S.em_id, S.em_code, S.bc_ptr, S.bc_id, S.bc_code, S.n_slots,
                   S.n_coup, S.n_slot_em, S.n_coup_em);
}

__global__ void kg_blk_em(const AsmSys* __restrict__ T, const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_blk_em_body(bid, (S.N + kSvRows - 1) / kSvRows, S.N, S.rowptr, S.sptr, S.blk_em);
}

__global__ void kg_slot_rel(const AsmSys* __restrict__ T, const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_slot_rel_body(bid, S.N, S.rowptr, S.sptr, S.blk_em, S.slot_rel);
}

__global__ void kg_leaf_coeffs(const AsmSys* __restrict__ T, const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_leaf_coeffs_body(bid, S.m, S.eps, S.border_of, S.frac, S.c_edge, S.v8);
}

__global__ void __launch_bounds__(kSvBS) kg_slot_values(const AsmSys* __restrict__ T,
                                                        const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_slot_values_body(bid, S.N, S.L, S.rowptr, S.blk_em, S.slot_rel, S.em_id, S.em_code, S.c_edge, S.reac, S.v8,
                     S.val);
}

// rhs_charge = 0 (the nodes no atom touches keep it)
__global__ void kg_zero_charge(const AsmSys* __restrict__ T, const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  const int64_t i = bid * blockDim.x + threadIdx.x;
  if (i < S.N) S.rhs_charge[i] = 0.0;
}

__global__ void kg_rho_push(const AsmSys* __restrict__ T, const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_rho_push_body(bid, S.m, S.A, S.owner, S.cnode, S.cnext, S.head);
}

__global__ void kg_rho_node(const AsmSys* __restrict__ T, const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_rho_node_body(bid, S.m, 32 * S.A, S.L, S.axyz, S.aq, S.owner, S.cnode, S.cnext, S.head, S.diag_slot, S.N,
                  S.rowptr, S.blk_em, S.slot_rel, S.em_id, S.em_code, S.v8, S.rho, S.rhs_charge);
}

__global__ void kg_dirichlet_rhs(const AsmSys* __restrict__ T, const int64_t* __restrict__ blk0, int ns) {
  NGPB_SYS;
  k_dirichlet_rhs_body(bid, S.N, S.L, S.is_bnd, S.gval, S.diag_slot, S.val, S.bc_rowptr, S.bc_col, S.bc_ptr,
                       S.bc_id, S.bc_code, S.c_edge, S.reac, S.v8, S.rhs_charge, S.rhs);
}

#undef NGPB_SYS

}  // namespace

// Potential at the atoms (write_potential_on_atoms_fast, src/pb_class.cpp:2333-2368):
// trilinear interpolation inside the owner leaf with the charge-projection weights,
// hanging corners taking the mean of their parents; NaN for atoms outside the domain.
__global__ void k_phi_atoms(MeshView m, int64_t A, const double* __restrict__ axyz, const int32_t* __restrict__ owner,
                            const double* __restrict__ phi, double* __restrict__ out) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= A) return;
  const int64_t l = owner[a];
  if (l < 0) {
    out[a] = __builtin_nan("");
    return;
  }
  const int32_t c0 = m.leaf_corner[8 * l], c7 = m.leaf_corner[8 * l + 7];
  const double volume = (dev::corner_coord(m, c7, 0) - dev::corner_coord(m, c0, 0)) *
                        (dev::corner_coord(m, c7, 1) - dev::corner_coord(m, c0, 1)) *
                        (dev::corner_coord(m, c7, 2) - dev::corner_coord(m, c0, 2));
  double acc = 0.0;
  for (int ii = 0; ii < 8; ++ii) {
    const int32_t co = m.leaf_corner[8 * l + 7 - ii];
    const double weight = fabs((axyz[3 * a + 0] - dev::corner_coord(m, co, 0)) *
                               (axyz[3 * a + 1] - dev::corner_coord(m, co, 1)) *
                               (axyz[3 * a + 2] - dev::corner_coord(m, co, 2))) / volume;
    const int32_t c = m.leaf_corner[8 * l + ii];
    if (c >= 0) {
      acc += phi[c] * weight;
    } else {
      const int64_t h = -1 - (int64_t)c;
      const int np = m.hang_np[h];
      double ph = 0.0;
      for (int j = 0; j < np; ++j) ph += phi[m.hang_par[4 * h + j]] / np;
      acc += ph * weight;
    }
  }
  out[a] = acc;
}

void run_symbolic(ngpb_ctx* c) {
  const MeshView m = mesh_view_of(c);
  NGPB_REQUIRE(20 * m.L < (int64_t)UINT32_MAX, NGPB_E_ARG, "too many leaves for 32-bit emission ids");
  const int64_t N = c->N;
  DBuf<unsigned> cnt(c->stream);
  cnt.resize((size_t)N);
  NGPB_HIP(hipMemsetAsync(cnt.p, 0, N * sizeof(unsigned), c->stream));
  hipLaunchKernelGGL(k_emit_rows<0>, dim3(blocks_for(8 * m.L, 128)), dim3(128), 0, c->stream, m, cnt.p, nullptr,
                     nullptr, nullptr);
  // per-row emission lists, sorted keys: needed only until the slots are cut
  DBuf<int64_t> em_rowptr(c->stream);
  DBuf<uint64_t> em_key(c->stream);  // (col << 32) | id
  DBuf<uint16_t> em_kw(c->stream);  // (k << 8) | weight code
  em_rowptr.resize((size_t)N + 1);
  hipLaunchKernelGGL(k_u32_to_i64, dim3(blocks_for(N, 256)), dim3(256), 0, c->stream, N, cnt.p, em_rowptr.p);
  NGPB_HIP(hipGetLastError());
  const int64_t total = exclusive_scan(c, em_rowptr.p, em_rowptr.p, N);
  NGPB_HIP(hipMemcpyAsync(em_rowptr.p + N, &total, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  em_key.resize((size_t)total);
  em_kw.resize((size_t)total);
  NGPB_HIP(hipMemsetAsync(cnt.p, 0, N * sizeof(unsigned), c->stream));
  hipLaunchKernelGGL(k_emit_rows<1>, dim3(blocks_for(8 * m.L, 128)), dim3(128), 0, c->stream, m, cnt.p, em_rowptr.p,
                     em_key.p, em_kw.p);
  NGPB_HIP(hipGetLastError());
  // per row: slots, couplings, slot emissions, coupling emissions (4 planes of N)
  DBuf<int64_t> cnt4(c->stream);
  cnt4.resize((size_t)4 * N);
  {
    DBuf<int32_t> big(c->stream);
    DBuf<int> nbig(c->stream);
    big.resize((size_t)N);
    nbig.resize(1);
    NGPB_HIP(hipMemsetAsync(nbig.p, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_row_sort_wave, dim3(blocks_for(N, kSortBS / 64)), dim3(kSortBS), 0, c->stream, N,
                       em_rowptr.p, em_key.p, em_kw.p, c->is_bnd.p, cnt4.p, big.p, nbig.p);
    NGPB_HIP(hipGetLastError());
    int hbig = 0;
    NGPB_HIP(hipMemcpyAsync(&hbig, nbig.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    NGPB_HIP(hipStreamSynchronize(c->stream));
    if (hbig > 0) {
      const size_t lds = (size_t)kSortBigCap * (8 + 2);
      hipLaunchKernelGGL(k_row_sort_block, dim3((unsigned)hbig), dim3(kSortBS), lds, c->stream, big.p, em_rowptr.p,
                         em_key.p, em_kw.p, c->is_bnd.p, N, cnt4.p);
      NGPB_HIP(hipGetLastError());
    }
  }
  int64_t tot[4];
  for (int q = 0; q < 4; ++q) tot[q] = exclusive_scan(c, cnt4.p + (int64_t)q * N, cnt4.p + (int64_t)q * N, N);
  const int64_t nnz = tot[0], nbc = tot[1];
  NGPB_REQUIRE(nnz < (int64_t)INT32_MAX && nbc < (int64_t)INT32_MAX, NGPB_E_ARG, "nnz exceeds 32-bit CSR");
  c->n = N;
  c->nnz = nnz;
  c->rowptr.resize((size_t)N + 1);
  c->col.resize((size_t)std::max<int64_t>(nnz, 1));
  c->diag_slot.resize((size_t)N);
  c->bc_rowptr.resize((size_t)N + 1);
  c->bc_col.resize((size_t)std::max<int64_t>(nbc, 1));
  c->bc_ptr.resize((size_t)nbc + 1);
  c->em_id.resize((size_t)std::max<int64_t>(tot[2], 1));
  c->em_code.resize((size_t)std::max<int64_t>(tot[2], 1));
  c->bc_id.resize((size_t)std::max<int64_t>(tot[3], 1));
  c->bc_code.resize((size_t)std::max<int64_t>(tot[3], 1));
  // slots, couplings and the repacked emissions (5 bytes each: source id, weight
  // code) in slot order, then block-relative slot offsets
  DBuf<int64_t> sptr(c->stream);
  sptr.resize((size_t)nnz + 1);
  hipLaunchKernelGGL(k_row_slots, dim3(blocks_for(N, kSlotBS / 64)), dim3(kSlotBS), 0, c->stream, N, em_rowptr.p,
                     em_key.p, em_kw.p, c->is_bnd.p, cnt4.p, c->rowptr.p, c->col.p, c->diag_slot.p, c->bc_rowptr.p,
                     c->bc_col.p, sptr.p, c->em_id.p, c->em_code.p, c->bc_ptr.p, c->bc_id.p, c->bc_code.p, nnz, nbc,
                     tot[2], tot[3]);
  NGPB_HIP(hipGetLastError());
  const int64_t nblk = (N + kSvRows - 1) / kSvRows;
  c->blk_em.resize((size_t)nblk + 1);
  hipLaunchKernelGGL(k_blk_em, dim3(blocks_for(nblk + 1, 256)), dim3(256), 0, c->stream, nblk, N, c->rowptr.p, sptr.p,
                     c->blk_em.p);
  c->slot_rel.resize((size_t)std::max<int64_t>(nnz, 1));
  hipLaunchKernelGGL(k_slot_rel, dim3(blocks_for(N, 256)), dim3(256), 0, c->stream, N, c->rowptr.p, sptr.p,
                     c->blk_em.p, c->slot_rel.p);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipStreamSynchronize(c->stream));
  c->have_symbolic = true;
}

void run_numeric(ngpb_ctx* c) {
  const MeshView m = mesh_view_of(c);
  const int64_t N = c->N, L = c->L;
  c->c_edge.resize((size_t)12 * L);
  c->v8.resize((size_t)L);
  c->val.resize((size_t)std::max<int64_t>(c->nnz, 1));
  c->rho.resize((size_t)N);
  c->rhs_charge.resize((size_t)N);
  c->rhs.resize((size_t)N);
  hipLaunchKernelGGL(k_leaf_coeffs, dim3(blocks_for(L, 128)), dim3(128), 0, c->stream, m, c->eps.p, c->border_of.p,
                     c->frac.p, c->c_edge.p, c->v8.p);
  hipLaunchKernelGGL(k_slot_values, dim3(blocks_for(N, kSvRows)), dim3(kSvBS), 0, c->stream, N, L, c->rowptr.p,
                     c->blk_em.p, c->slot_rel.p, c->em_id.p, c->em_code.p, c->c_edge.p, c->reac.p, c->v8.p, c->val.p);
  NGPB_HIP(hipGetLastError());
  // charge density and its right-hand side, on the nodes the atoms touch
  NGPB_HIP(hipMemsetAsync(c->rhs_charge.p, 0, N * sizeof(double), c->stream));
  if (c->have_charges && c->A > 0) {
    const int64_t C = 32 * c->A;
    DBuf<int32_t> cnode(c->stream), cnext(c->stream), head(c->stream);
    cnode.resize((size_t)C);
    cnext.resize((size_t)C);
    head.resize((size_t)N);
    NGPB_HIP(hipMemsetAsync(head.p, 0xff, N * sizeof(int32_t), c->stream));
    hipLaunchKernelGGL(k_rho_push, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, m, c->A, c->atom_owner.p, cnode.p,
                       cnext.p, head.p);
    hipLaunchKernelGGL(k_rho_node, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, m, C, L, c->atom_xyz.p,
                       c->atom_q.p, c->atom_owner.p, cnode.p, cnext.p, head.p, c->diag_slot.p, N, c->rowptr.p,
                       c->blk_em.p, c->slot_rel.p, c->em_id.p, c->em_code.p, c->v8.p, c->rho.p, c->rhs_charge.p);
    NGPB_HIP(hipGetLastError());
    NGPB_HIP(hipStreamSynchronize(c->stream));
  }
  hipLaunchKernelGGL(k_dirichlet_rhs, dim3(blocks_for(N, 128)), dim3(128), 0, c->stream, N, L, c->is_bnd.p,
                     c->gval.p, c->diag_slot.p, c->val.p, c->bc_rowptr.p, c->bc_col.p, c->bc_ptr.p, c->bc_id.p,
                     c->bc_code.p, c->c_edge.p, c->reac.p, c->v8.p, c->rhs_charge.p, c->rhs.p);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipStreamSynchronize(c->stream));
  c->have_system = true;
  c->have_compact = false;
}

void run_phi_on_atoms(ngpb_ctx* c, double* d_out) {
  const MeshView m = mesh_view_of(c);
  if (c->A > 0)
    hipLaunchKernelGGL(k_phi_atoms, dim3(blocks_for(c->A, 256)), dim3(256), 0, c->stream, m, c->A, c->atom_xyz.p,
                       c->atom_owner.p, c->x.p, d_out);
  NGPB_HIP(hipGetLastError());
}

void run_dirichlet_values(ngpb_ctx* c, int bc_type) {
  NGPB_REQUIRE(bc_type >= 1 && bc_type <= 3, NGPB_E_ARG, "bc_type must be 1, 2 or 3");
  NGPB_REQUIRE(bc_type != 2 || c->have_charges, NGPB_E_STATE, "Coulomb boundary data need the charges");
  c->bnd_val.resize((size_t)std::max<int64_t>(c->NB, 1));
  if (c->NB > 0)
    hipLaunchKernelGGL(k_bc_values, dim3(blocks_for(c->NB, 128)), dim3(128), 0, c->stream, c->NB, c->bnd_node.p,
                       c->node_xyz.p, bc_type, c->pc.eps_in, c->pc.eps_out, c->pc.k, bc_type == 2 ? c->A : 0,
                       c->atom_xyz.p, c->atom_q.p, c->bnd_val.p, c->gval.p);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipStreamSynchronize(c->stream));
  c->have_bnd_values = true;
}

// ------------------------------------------------------- assembly of many systems
namespace {

// The launches of one grouped stage: block ranges and system tables uploaded on the
// stream stay allocated, with their host copies, until the run ends.
struct GroupRun {
  hipStream_t st;
  std::list<std::vector<int64_t>> host_ranges;
  std::list<DBuf<int64_t>> ranges;
  std::list<std::vector<AsmSys>> host_tables;
  std::list<DBuf<AsmSys>> tables;
  explicit GroupRun(hipStream_t s) : st(s) {}

  const AsmSys* table(const std::vector<AsmSys>& t) {
    host_tables.push_back(t);
    tables.emplace_back(st);
    tables.back().upload(host_tables.back().data(), t.size(), st);
    return tables.back().p;
  }

  // one launch of kernel over every system; nb[s] = blocks of system s
  template <typename K>
  void launch(K kernel, const std::vector<int64_t>& nb, int bs, size_t lds, const AsmSys* T) {
    host_ranges.emplace_back(nb.size() + 1, 0);
    std::vector<int64_t>& h = host_ranges.back();
    for (size_t i = 0; i < nb.size(); ++i) h[i + 1] = h[i] + nb[i];
    const int64_t total = h.back();
    if (total == 0) return;
    NGPB_REQUIRE(total * bs < ((int64_t)1 << 32), NGPB_E_ARG, "grouped launch exceeds 2^32 work-items");
    ranges.emplace_back(st);
    ranges.back().upload(h.data(), h.size(), st);
    hipLaunchKernelGGL(kernel, dim3((unsigned)total), dim3(bs), lds, st, T, ranges.back().p, (int)nb.size());
    NGPB_HIP(hipGetLastError());
  }
};

template <typename F>
std::vector<int64_t> per_system(const std::vector<AsmSys>& T, F f) {
  std::vector<int64_t> v(T.size());
  for (size_t i = 0; i < T.size(); ++i) v[i] = f(T[i]);
  return v;
}

// run_symbolic of every context, stage by stage in one launch each
void symbolic_many(hipStream_t st, const std::vector<ngpb_ctx*>& cs) {
  GroupRun g(st);
  const int ns = (int)cs.size();
  std::vector<AsmSys> T((size_t)ns);
  int64_t sumN = 0;
  for (int i = 0; i < ns; ++i) {
    const ngpb_ctx* c = cs[i];
    T[i].m = mesh_view_of(c);
    NGPB_REQUIRE(20 * T[i].m.L < (int64_t)UINT32_MAX, NGPB_E_ARG, "too many leaves for 32-bit emission ids");
    T[i].N = c->N;
    T[i].L = c->L;
    T[i].is_bnd = c->is_bnd.p;
    sumN += c->N;
  }
  DBuf<unsigned> cnt(st);
  DBuf<int64_t> emrp(st), cnt4(st);
  DBuf<int32_t> big(st);
  DBuf<int> nbig(st);
  cnt.resize((size_t)sumN);
  emrp.resize((size_t)(sumN + ns));
  cnt4.resize((size_t)(4 * sumN));
  big.resize((size_t)sumN);
  nbig.resize((size_t)ns);
  for (int64_t i = 0, o = 0; i < ns; o += T[i].N, ++i) {
    T[i].cnt = cnt.p + o;
    T[i].em_rowptr = emrp.p + o + i;
    T[i].cnt4 = cnt4.p + 4 * o;
    T[i].big = big.p + o;
    T[i].nbig = nbig.p + i;
  }
  NGPB_HIP(hipMemsetAsync(cnt.p, 0, (size_t)sumN * sizeof(unsigned), st));
  NGPB_HIP(hipMemsetAsync(nbig.p, 0, (size_t)ns * sizeof(int), st));
  const AsmSys* d = g.table(T);
  const auto nb_emit = per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for(8 * a.L, 128); });
  g.launch(kg_emit_rows<0>, nb_emit, 128, 0, d);
  g.launch(kg_u32_to_i64, per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for(a.N, 256); }), 256, 0, d);
  std::vector<ScanSeg> segs((size_t)ns);
  for (int i = 0; i < ns; ++i) segs[i] = {T[i].em_rowptr, T[i].em_rowptr, T[i].N};
  std::vector<int64_t> n_em((size_t)ns);
  exclusive_scan_many(st, segs, n_em.data());
  int64_t sumE = 0;
  for (int i = 0; i < ns; ++i) sumE += n_em[i];
  DBuf<uint64_t> key(st);
  DBuf<uint16_t> kw(st);
  key.resize((size_t)std::max<int64_t>(sumE, 1));
  kw.resize((size_t)std::max<int64_t>(sumE, 1));
  for (int64_t i = 0, o = 0; i < ns; o += n_em[i], ++i) {
    T[i].em_key = key.p + o;
    T[i].em_kw = kw.p + o;
    T[i].n_em = n_em[i];
  }
  d = g.table(T);
  hipLaunchKernelGGL(kg_em_end, dim3(blocks_for(ns, 64)), dim3(64), 0, st, d, ns);
  NGPB_HIP(hipMemsetAsync(cnt.p, 0, (size_t)sumN * sizeof(unsigned), st));
  g.launch(kg_emit_rows<1>, nb_emit, 128, 0, d);
  g.launch(kg_row_sort_wave, per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for(a.N, kSortBS / 64); }),
           kSortBS, 0, d);
  std::vector<int> hbig((size_t)ns);
  NGPB_HIP(hipMemcpyAsync(hbig.data(), nbig.p, (size_t)ns * sizeof(int), hipMemcpyDeviceToHost, st));
  NGPB_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> nb_big((size_t)ns);
  for (int i = 0; i < ns; ++i) nb_big[i] = hbig[i];
  g.launch(kg_row_sort_block, nb_big, kSortBS, (size_t)kSortBigCap * (8 + 2), d);
  std::vector<ScanSeg> segs4((size_t)4 * ns);
  for (int i = 0; i < ns; ++i)
    for (int q = 0; q < 4; ++q) segs4[4 * i + q] = {T[i].cnt4 + q * T[i].N, T[i].cnt4 + q * T[i].N, T[i].N};
  std::vector<int64_t> tot((size_t)4 * ns);
  exclusive_scan_many(st, segs4, tot.data());
  int64_t sumS = 0;
  for (int i = 0; i < ns; ++i) {
    ngpb_ctx* c = cs[i];
    const int64_t N = c->N, nnz = tot[4 * i], nbc = tot[4 * i + 1], nse = tot[4 * i + 2], nce = tot[4 * i + 3];
    NGPB_REQUIRE(nnz < (int64_t)INT32_MAX && nbc < (int64_t)INT32_MAX, NGPB_E_ARG, "nnz exceeds 32-bit CSR");
    c->n = N;
    c->nnz = nnz;
    c->rowptr.resize((size_t)N + 1);
    c->col.resize((size_t)std::max<int64_t>(nnz, 1));
    c->diag_slot.resize((size_t)N);
    c->bc_rowptr.resize((size_t)N + 1);
    c->bc_col.resize((size_t)std::max<int64_t>(nbc, 1));
    c->bc_ptr.resize((size_t)nbc + 1);
    c->em_id.resize((size_t)std::max<int64_t>(nse, 1));
    c->em_code.resize((size_t)std::max<int64_t>(nse, 1));
    c->bc_id.resize((size_t)std::max<int64_t>(nce, 1));
    c->bc_code.resize((size_t)std::max<int64_t>(nce, 1));
    c->blk_em.resize((size_t)(N + kSvRows - 1) / kSvRows + 1);
    c->slot_rel.resize((size_t)std::max<int64_t>(nnz, 1));
    AsmSys& a = T[i];
    a.n_slots = nnz;
    a.n_coup = nbc;
    a.n_slot_em = nse;
    a.n_coup_em = nce;
    a.rowptr = c->rowptr.p;
    a.col = c->col.p;
    a.diag_slot = c->diag_slot.p;
    a.bc_rowptr = c->bc_rowptr.p;
    a.bc_col = c->bc_col.p;
    a.bc_ptr = c->bc_ptr.p;
    a.em_id = c->em_id.p;
    a.em_code = c->em_code.p;
    a.bc_id = c->bc_id.p;
    a.bc_code = c->bc_code.p;
    a.blk_em = c->blk_em.p;
    a.slot_rel = c->slot_rel.p;
    sumS += nnz + 1;
  }
  DBuf<int64_t> sptr(st);
  sptr.resize((size_t)sumS);
  for (int64_t i = 0, o = 0; i < ns; o += T[i].n_slots + 1, ++i) T[i].sptr = sptr.p + o;
  d = g.table(T);
  g.launch(kg_row_slots, per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for(a.N, kSlotBS / 64); }),
           kSlotBS, 0, d);
  g.launch(kg_blk_em,
           per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for((a.N + kSvRows - 1) / kSvRows + 1, 256); }),
           256, 0, d);
  g.launch(kg_slot_rel, per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for(a.N, 256); }), 256, 0, d);
  NGPB_HIP(hipStreamSynchronize(st));
  for (ngpb_ctx* c : cs) c->have_symbolic = true;
}

// run_numeric of every context, stage by stage in one launch each
void numeric_many(hipStream_t st, const std::vector<ngpb_ctx*>& cs) {
  GroupRun g(st);
  const int ns = (int)cs.size();
  std::vector<AsmSys> T((size_t)ns);
  int64_t sumC = 0, sumH = 0;
  for (int i = 0; i < ns; ++i) {
    ngpb_ctx* c = cs[i];
    const int64_t N = c->N, L = c->L;
    c->c_edge.resize((size_t)12 * L);
    c->v8.resize((size_t)L);
    c->val.resize((size_t)std::max<int64_t>(c->nnz, 1));
    c->rho.resize((size_t)N);
    c->rhs_charge.resize((size_t)N);
    c->rhs.resize((size_t)N);
    AsmSys& a = T[i];
    a.m = mesh_view_of(c);
    a.N = N;
    a.L = L;
    a.A = c->have_charges ? c->A : 0;
    a.is_bnd = c->is_bnd.p;
    a.rowptr = c->rowptr.p;
    a.diag_slot = c->diag_slot.p;
    a.bc_rowptr = c->bc_rowptr.p;
    a.bc_col = c->bc_col.p;
    a.bc_ptr = c->bc_ptr.p;
    a.blk_em = c->blk_em.p;
    a.slot_rel = c->slot_rel.p;
    a.em_id = c->em_id.p;
    a.em_code = c->em_code.p;
    a.bc_id = c->bc_id.p;
    a.bc_code = c->bc_code.p;
    a.eps = c->eps.p;
    a.border_of = c->border_of.p;
    a.frac = c->frac.p;
    a.reac = c->reac.p;
    a.gval = c->gval.p;
    a.axyz = c->atom_xyz.p;
    a.aq = c->atom_q.p;
    a.owner = c->atom_owner.p;
    a.c_edge = c->c_edge.p;
    a.v8 = c->v8.p;
    a.val = c->val.p;
    a.rho = c->rho.p;
    a.rhs_charge = c->rhs_charge.p;
    a.rhs = c->rhs.p;
    if (a.A > 0) {
      sumC += 32 * a.A;
      sumH += N;
    }
  }
  // charge lists of the systems with atoms: contributions 32 A, list heads N
  DBuf<int32_t> cnode(st), cnext(st), head(st);
  if (sumC > 0) {
    cnode.resize((size_t)sumC);
    cnext.resize((size_t)sumC);
    head.resize((size_t)sumH);
    NGPB_HIP(hipMemsetAsync(head.p, 0xff, (size_t)sumH * sizeof(int32_t), st));
    for (int64_t i = 0, oc = 0, oh = 0; i < ns; ++i) {
      if (T[i].A == 0) continue;
      T[i].cnode = cnode.p + oc;
      T[i].cnext = cnext.p + oc;
      T[i].head = head.p + oh;
      oc += 32 * T[i].A;
      oh += T[i].N;
    }
  }
  const AsmSys* d = g.table(T);
  g.launch(kg_leaf_coeffs, per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for(a.L, 128); }), 128, 0, d);
  g.launch(kg_slot_values, per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for(a.N, kSvRows); }), kSvBS,
           0, d);
  g.launch(kg_zero_charge, per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for(a.N, 256); }), 256, 0, d);
  const auto nb_charge = per_system(T, [](const AsmSys& a) { return (int64_t)(a.A > 0 ? blocks_for(32 * a.A, 256) : 0); });
  g.launch(kg_rho_push, nb_charge, 256, 0, d);
  g.launch(kg_rho_node, nb_charge, 256, 0, d);
  g.launch(kg_dirichlet_rhs, per_system(T, [](const AsmSys& a) { return (int64_t)blocks_for(a.N, 128); }), 128, 0,
           d);
  NGPB_HIP(hipStreamSynchronize(st));
  for (ngpb_ctx* c : cs) {
    c->have_system = true;
    c->have_compact = false;
  }
}

// systems per grouped launch: at most this many leaves (the symbolic temporaries take
// about 0.6 KB per leaf) so that the temporaries and the grids stay bounded; the
// environment variable NGPB_GROUP_LEAVES sets another bound
constexpr int64_t kGroupLeaves = (int64_t)16 << 20;

int64_t group_leaves() {
  if (const char* e = std::getenv("NGPB_GROUP_LEAVES")) {
    const long long v = std::strtoll(e, nullptr, 10);
    if (v > 0) return (int64_t)v;
  }
  return kGroupLeaves;
}

template <typename F>
void in_chunks(const std::vector<ngpb_ctx*>& cs, F f) {
  const int64_t cap = group_leaves();
  size_t i = 0;
  while (i < cs.size()) {
    size_t j = i;
    int64_t L = 0;
    while (j < cs.size() && (j == i || L + cs[j]->L <= cap)) L += cs[j++]->L;
    f(std::vector<ngpb_ctx*>(cs.begin() + i, cs.begin() + j));
    i = j;
  }
}

}  // namespace

void run_assemble_many(const std::vector<ngpb_ctx*>& cs) {
  if (cs.empty()) return;
  hipStream_t st = cs[0]->stream;
  for (ngpb_ctx* c : cs) NGPB_HIP(hipStreamSynchronize(c->stream));
  std::vector<ngpb_ctx*> sym;
  for (ngpb_ctx* c : cs)
    if (!c->have_symbolic || c->external_csr) sym.push_back(c);
  in_chunks(sym, [&](const std::vector<ngpb_ctx*>& part) { symbolic_many(st, part); });
  in_chunks(cs, [&](const std::vector<ngpb_ctx*>& part) { numeric_many(st, part); });
}

}  // namespace ngpb
