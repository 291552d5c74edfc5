// Mixed-precision solve: FP32 compact operator and FP32 inner CG, FP64 residual
// replacement, FP64 reductions.
//
// Operator. With D = diag(A) and S = D^-1/2, CG on Ahat = S A S (unit diagonal) is
// Jacobi-preconditioned CG on A. Ahat's off-diagonal part is stored in slices of 64
// consecutive rows (one wavefront, one row per lane), each slice in one of two forms:
//   * delta slice: the slice's off-diagonals take at most 16 distinct values of
//     (column - row) -- the stencil of a uniform region in lexicographic node order;
//     slot k holds, per row, the FP32 value at column row + delta_k (0 if absent) and
//     only the values and the slot's delta are stored;
//   * offset slice: w "main" slots per row, w chosen to minimise bytes (padding of
//     short rows against overflow of long rows); slot k holds the k-th off-diagonal
//     (ascending column) as an FP32 value and a 16-bit offset from the slot's 32-bit
//     base column (32-bit columns when a slot spans 2^16 or more); off-diagonals
//     beyond w go to the slice's overflow list (FP32 value + 32-bit column, rows in
//     lane order, an 8-bit count per row).
// The form with fewer bytes is chosen per slice. The diagonal is not stored. Row
// sums accumulate in FP64 and are rounded once. Each wavefront walks many slices.
//
// Solve (reliable updates in the sense of Sleijpen & van der Vorst, 1996):
//   x, b, the true residual r = b - A x and all dot products are FP64; the inner
//   iterate xhat, residual rhat, direction p and q = Ahat p are FP32. An inner cycle
//   runs CG until (rhat, rhat) falls below drop^2 times its value at the last
//   replacement (or below the convergence target); then x += S xhat, the FP64
//   residual is recomputed with the FP64 CSR operator, rhat = S r, and CG continues
//   with the same direction p. The solve stops when ||b - A x||_2 <= tol ||b||_2 in
//   FP64. An iteration is three kernels (SpMV, update, direction); each combines
//   the previous kernel's per-block partial sums itself, in the same fixed order in
//   every block.
#include <algorithm>
#include <cmath>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "ngpb_internal.h"

namespace ngpb {

void spmv_fp64(ngpb_ctx* c, const double* x, double* y);

namespace {

constexpr int kBS = 256;
constexpr int kW = 64;
constexpr int kFin = 1024;

#include "ngpb_sell.inc"


__global__ void k_scale(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                        const double* __restrict__ val, double* __restrict__ dscale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 1.0;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    if (col[k] == i) d = val[k];
  dscale[i] = 1.0 / sqrt(d);
}

// position of the diagonal in row i (or -1) and the number of off-diagonals
__device__ __forceinline__ void row_info(const int32_t* rowptr, const int32_t* col, int64_t i, int64_t* dpos, int* cnt) {
  const int64_t b = rowptr[i], e = rowptr[i + 1];
  int64_t d = -1;
  for (int64_t t = b; t < e; ++t)
    if (col[t] == i) { d = t; break; }
  *dpos = d;
  *cnt = (int)(e - b) - (d >= 0 ? 1 : 0);
}

// CSR index of the k-th off-diagonal of a row
__device__ __forceinline__ int64_t offdiag_at(int64_t b, int64_t dpos, int k) {
  const int64_t idx = b + k;
  return (dpos >= 0 && dpos <= idx) ? idx + 1 : idx;
}

constexpr int kDMax = 16;  // most distinct column deltas of a delta slice

// Walks the distinct values of (column - row) over the off-diagonals of the slice's
// rows in increasing order (each lane's off-diagonals are sorted by column). For the
// k-th distinct delta `m`, visit(k, m, e, pos) is called with e = this lane's CSR
// index of the entry with that delta (or -1) and pos = that entry's rank among the
// lane's off-diagonals. Stops after LIMIT + 1 deltas; returns the count.
template <int LIMIT, class F>
__device__ __forceinline__ int for_each_delta(const int32_t* col, int64_t row, bool valid, int64_t b, int64_t dpos,
                                              int cnt, F&& visit) {
  int pos = 0, nd = 0;
  while (true) {
    const int64_t e = (valid && pos < cnt) ? offdiag_at(b, dpos, pos) : -1;
    const int d = e >= 0 ? (int)((int64_t)col[e] - row) : INT32_MAX;
    const int m = wave_min_i(d);
    if (m == INT32_MAX || nd > LIMIT) break;
    visit(nd, m, d == m ? e : (int64_t)-1, pos);
    ++nd;
    if (d == m) ++pos;
  }
  return nd;
}

// One wave per slice: chooses the slice form.
//   delta   at most kDMax distinct deltas: 4 bytes per slot entry, gathers contiguous
//   packed  at most 256 distinct deltas and at most 64 off-diagonals per row: 5 bytes
//           per entry (no padding) + 2 bytes per row + 4 per dictionary entry
//   offset  otherwise: main width minimising padding + overflow bytes
// The delta form is taken when it moves no more than 1.25x the packed form's bytes
// (its loads are fewer and wider); the offset form only when neither fits.
__global__ void __launch_bounds__(kBS) k_sell_shape(int64_t n, int64_t nslices, const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ col, const double* __restrict__ val,
                                                    const double* __restrict__ dscale, int allow_const,
                                                    int64_t* __restrict__ cslots, int32_t* __restrict__ width,
                                                    int64_t* __restrict__ wcount, int64_t* __restrict__ wide_count,
                                                    int64_t* __restrict__ ov_count, int64_t* __restrict__ dslots,
                                                    int64_t* __restrict__ pk_count, int64_t* __restrict__ pk_dcount,
                                                    int64_t* __restrict__ pk_gcount, int* __restrict__ bad) {
  const int64_t s = (int64_t)blockIdx.x * (kBS / kW) + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const int64_t row = s * kW + lane_id();
  int cnt = 0;
  int64_t dpos = -1;
  if (row < n) row_info(rowptr, col, row, &dpos, &cnt);
  const int64_t b = row < n ? rowptr[row] : 0;
  const int mx = wave_max_i(cnt);
  const int tot = wave_sum_i(cnt);
  const int nd = for_each_delta<kPackedMaxDeltas>(col, row, row < n, b, dpos, cnt, [](int, int, int64_t, int) {});
  const bool packable = nd <= kPackedMaxDeltas && mx <= kW;
  const long long pk_cost = (long long)tot * 5 + 2 * kW + (long long)nd * 4;
  const bool use_delta = nd <= kDMax && (!packable || (long long)nd * kW * 4 * 4 <= pk_cost * 5);
  const bool use_packed = !use_delta && packable;
  // constant delta slice: every row of the slice has every delta, with the same FP32
  // value (bit for bit) in each slot (lane 0's row always exists)
  bool cst = allow_const && use_delta && nd > 0;
  if (cst) {
    const double di = row < n ? dscale[row] : 0.0;
    for_each_delta<kDMax>(col, row, row < n, b, dpos, cnt, [&](int, int, int64_t e, int) {
      const uint32_t bits = e >= 0 ? __float_as_uint((float)(val[e] * di * dscale[col[e]])) : 0u;
      const uint32_t b0 = (uint32_t)__shfl((int)bits, 0, 64);
      if (__ballot(row < n && (e < 0 || bits != b0))) cst = false;
    });
  }
  // index words of a packed slice: one per row taking part in each group of 4 slots
  int words = 0;
  if (use_packed)
    for (int k = 0; k < mx; k += 4) words += __popcll(__ballot(cnt > k));
  // offset form: main slots (6 bytes per entry) against overflow entries, which cost
  // 8 bytes but are walked serially per row (two dependent loads each), so they are
  // charged 32 bytes apiece
  int best_w = mx;
  bool wide = false;
  if (!use_delta && !use_packed) {
    long long best = (long long)mx * kW * 6;
    for (int w = mx - 1; w >= 0; --w) {
      const int extra = wave_sum_i(max(0, cnt - w));
      const long long cost = (long long)w * kW * 6 + (long long)extra * 32 + kW;
      if (cost < best) {
        best = cost;
        best_w = w;
      }
    }
    for (int k = 0; k < best_w; ++k) {
      const int cc = (row < n && k < cnt) ? col[offdiag_at(b, dpos, k)] : INT32_MAX;
      const int cmin = wave_min_i(cc);
      const int cmax = wave_max_i(cc == INT32_MAX ? INT32_MIN : cc);
      if (cmin != INT32_MAX && (int64_t)cmax - (int64_t)cmin > 65535) wide = true;
    }
    if (cnt - best_w > 255) atomicOr(bad, 1);
  }
  const bool offset = !use_delta && !use_packed;
  const int extra = offset ? wave_sum_i(max(0, cnt - best_w)) : 0;
  if (lane_id() == 0) {
    // a slice without off-diagonals (rows of Dirichlet nodes) is a delta slice of
    // zero slots, marked -kPackedBase since -0 would read as an offset slice
    width[s] = use_delta ? (nd == 0 ? -kPackedBase : (cst ? -(kConstBase + nd) : -nd))
                         : (use_packed ? -(kPackedBase + nd) : best_w);
    wcount[s] = use_delta ? nd : (use_packed ? 0 : best_w);
    wide_count[s] = (offset && wide) ? (int64_t)best_w * kW : 0;
    ov_count[s] = extra;
    dslots[s] = (use_delta && !cst) ? nd : 0;
    cslots[s] = cst ? nd : 0;
    pk_count[s] = use_packed ? tot : 0;
    pk_dcount[s] = use_packed ? nd : 0;
    pk_gcount[s] = use_packed ? words : 0;
  }
}

__global__ void __launch_bounds__(kBS) k_sell_fill(int64_t n, int64_t nslices, const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, const double* __restrict__ val,
                                                   const double* __restrict__ dscale, const int32_t* __restrict__ width,
                                                   const int64_t* __restrict__ base_ptr,
                                                   const int64_t* __restrict__ wide_off,
                                                   const int64_t* __restrict__ wide_cnt,
                                                   const int64_t* __restrict__ ov_off, const int64_t* __restrict__ ov_cnt_s,
                                                   float* __restrict__ sval, uint16_t* __restrict__ soff,
                                                   int32_t* __restrict__ sbase, int32_t* __restrict__ scol32,
                                                   int64_t* __restrict__ wide_ptr, int64_t* __restrict__ ov_ptr,
                                                   uint8_t* __restrict__ ov_cnt, float* __restrict__ ov_val,
                                                   int32_t* __restrict__ ov_col, const int64_t* __restrict__ pk_off,
                                                   const int64_t* __restrict__ pk_doff,
                                                   const int64_t* __restrict__ pk_goff, float* __restrict__ pk_val,
                                                   uint8_t* __restrict__ pk_idx_bytes, uint16_t* __restrict__ pk_lp,
                                                   int32_t* __restrict__ pk_dict, PkHdr* __restrict__ pk_hdr,
                                                   float* __restrict__ cval) {
  const int64_t s = (int64_t)blockIdx.x * (kBS / kW) + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const int lane = lane_id();
  const int64_t row = s * kW + lane;
  const int w = width[s];
  const int64_t bp = base_ptr[s];
  const int64_t sp = bp * kW;
  const bool packed = w <= -(kPackedBase + 1);
  const bool wide = wide_cnt[s] > 0;
  const int64_t wp = wide ? wide_off[s] : -1;
  const bool has_ov = ov_cnt_s[s] > 0;
  if (lane == 0) {
    wide_ptr[s] = packed ? pk_off[s] : wp;
    ov_ptr[s] = packed ? pk_doff[s] : (has_ov ? ov_off[s] : -1);
  }
  int cnt = 0;
  int64_t dpos = -1, b = 0;
  if (row < n) {
    row_info(rowptr, col, row, &dpos, &cnt);
    b = rowptr[row];
  }
  const double di = row < n ? dscale[row] : 0.0;
  if (packed) {
    // sorted position of this row: off-diagonal count descending, ties by lane
    int rank = 0;
    for (int j = 0; j < kW; ++j) {
      const int cj = __shfl(cnt, j, 64);
      rank += (cj > cnt || (cj == cnt && j < lane)) ? 1 : 0;
    }
    pk_lp[s * kW + rank] = (uint16_t)(lane | (cnt << 8));
    // lane k holds the start of slot k: m_0 + ... + m_{k-1}, m_j = rows with > j entries
    const int mx = wave_max_i(cnt);
    int my_m = 0;
    for (int k = 0; k < mx; ++k) {
      const int mk = __popcll(__ballot(cnt > k));
      if (lane == k) my_m = mk;
    }
    const int slot_start = wave_excl_scan_i(my_m);
    // lane g < 16 holds the first index word of slot group g. The permute runs in
    // every lane: a lane whose source lane is inactive would read 0.
    const int m_of_group = __shfl(my_m, (4 * lane) & (kW - 1), 64);
    const int my_mg = lane < 16 ? m_of_group : 0;
    const int group_start = wave_excl_scan_i(my_mg);
    const int64_t E = pk_off[s], D = pk_doff[s], G = pk_goff[s];
    if (lane < 16) pk_hdr[s].m[lane] = (uint8_t)my_m;
    if (lane == 0) {
      pk_hdr[s].E = (int32_t)E;
      pk_hdr[s].G = (int32_t)G;
      pk_hdr[s].D = (int32_t)D;
      pk_hdr[s].nd = (uint16_t)(-w - kPackedBase);
      pk_hdr[s].maxlen = (uint8_t)mx;
      pk_hdr[s].pad = 0;
    }
    for_each_delta<kPackedMaxDeltas>(col, row, row < n, b, dpos, cnt, [&](int k, int m, int64_t e, int pos) {
      const int pk = pos < kW ? pos : 0;
      const int st = __shfl(slot_start, pk, 64);
      const int gst = __shfl(group_start, pk >> 2, 64);
      if (lane == 0) pk_dict[D + k] = m;
      if (e >= 0) {
        pk_val[E + st + rank] = (float)(val[e] * di * dscale[col[e]]);
        pk_idx_bytes[(G + gst + rank) * 4 + (pos & 3)] = (uint8_t)k;
      }
    });
    if (row < n) ov_cnt[row] = 0;
    return;
  }
  if (w <= -(kConstBase + 1) && w >= -(kConstBase + kDMax)) {
    // constant delta slice: the slot value once (every row holds the same)
    for_each_delta<kDMax>(col, row, row < n, b, dpos, cnt, [&](int k, int m, int64_t e, int) {
      if (lane == 0) {
        sbase[bp + k] = m;
        cval[bp + k] = (float)(val[e] * di * dscale[col[e]]);
      }
    });
    if (row < n) ov_cnt[row] = 0;
    return;
  }
  if (w < 0) {
    // delta slice: slot k holds the entry of column row + delta_k (0 if absent)
    for_each_delta<kDMax>(col, row, row < n, b, dpos, cnt, [&](int k, int m, int64_t e, int) {
      const int64_t idx = sp + (int64_t)k * kW + lane;
      sval[idx] = e >= 0 ? (float)(val[e] * di * dscale[col[e]]) : 0.f;
      soff[idx] = 0;
      if (lane == 0) sbase[bp + k] = m;
    });
    if (row < n) ov_cnt[row] = 0;
    return;
  }
  for (int k = 0; k < w; ++k) {
    const int64_t e = (row < n && k < cnt) ? offdiag_at(b, dpos, k) : -1;
    const int cc = e >= 0 ? col[e] : INT32_MAX;
    const int cmin = wave_min_i(cc);
    const int base = (cmin == INT32_MAX) ? 0 : cmin;
    const int64_t idx = sp + (int64_t)k * kW + lane;
    float v = 0.f;
    if (e >= 0) v = (float)(val[e] * di * dscale[cc]);
    sval[idx] = v;
    soff[idx] = (uint16_t)(e >= 0 && !wide ? (cc - base) : 0);
    if (lane == 0) sbase[bp + k] = base;
    if (wide) scol32[wp + (int64_t)k * kW + lane] = e >= 0 ? cc : base;
  }
  const int extra = max(0, cnt - w);
  if (row < n) ov_cnt[row] = (uint8_t)extra;
  if (has_ov) {
    const int64_t start = ov_off[s] + wave_excl_scan_i(extra);
    for (int k = 0; k < extra; ++k) {
      const int64_t e = offdiag_at(b, dpos, w + k);
      const int cc = col[e];
      ov_val[start + k] = (float)(val[e] * di * dscale[cc]);
      ov_col[start + k] = cc;
    }
  }
}

// q = Ahat p with per-block (p, q) when DOT: one kernel for every slice form. Wave w
// takes the pairs of adjacent slices w, w + W, w + 2W, ... (W waves in the grid).
template <bool DOT>
__global__ void __launch_bounds__(kBS) k_sell_spmv(Sell A, const float* __restrict__ x, float* __restrict__ y,
                                                   const KrylovScalars* ks, double* __restrict__ partials) {
  __shared__ double s_red[kBS / 64];
  if (DOT && ks->done) return;
  const int64_t npairs = (A.nslices + 1) / 2;
  const int64_t step = (int64_t)gridDim.x * (kBS / kW);
  int64_t t = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBS / kW) + (threadIdx.x >> 6)));
  double part = 0.0;
  // two delta slices with the same slot count are processed together so that twice
  // as many loads are in flight; the next pair's widths and slot pointers are
  // requested before the current pair is processed
  int w0n = 0, w1n = 0;
  int64_t b0n = 0, b1n = 0;
  if (t < npairs) {
    w0n = A.width[2 * t];
    b0n = A.base_ptr[2 * t];
    if (2 * t + 1 < A.nslices) {
      w1n = A.width[2 * t + 1];
      b1n = A.base_ptr[2 * t + 1];
    }
  }
  while (t < npairs) {
    const int w0 = w0n, w1 = w1n;
    const int64_t b0 = b0n, b1 = b1n;
    const int64_t s0 = 2 * t;
    const bool has1 = s0 + 1 < A.nslices;
    const int64_t t2 = t + step;
    if (t2 < npairs) {
      w0n = A.width[2 * t2];
      b0n = A.base_ptr[2 * t2];
      if (2 * t2 + 1 < A.nslices) {
        w1n = A.width[2 * t2 + 1];
        b1n = A.base_ptr[2 * t2 + 1];
      }
    }
    bool done = false;
    if (has1 && w0 == w1) {
      switch (w0) {
        case -4: delta_pair<4>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -5: delta_pair<5>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -6: delta_pair<6>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -7: delta_pair<7>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -8: delta_pair<8>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -36: delta_pair<4, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -37: delta_pair<5, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -38: delta_pair<6, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -39: delta_pair<7, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -40: delta_pair<8, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        default: break;
      }
    }
    if (!done) {
      one_slice(A, s0, w0, b0, x, y, &part);
      if (has1) one_slice(A, s0 + 1, w1, b1, x, y, &part);
    }
    t = t2;
  }
  if (DOT) {
    part = dev::block_sum<kBS>(part, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = part;
  }
}

// Form-specialised SpMV (the default): the delta slices, the packed slices and the
// offset slices each have a kernel of their own, walking a list of their slices, so
// each kernel holds only the registers its form needs. Each writes its per-block
// (p, q) partials into its own range of the partial array.

// delta slices: list entries taken two at a time; two adjacent slices with the same
// slot count go through the paired path; the next pair's slice ids, widths and slot
// pointers are requested before the current pair is processed
// delta work of wave `wave` among `nwaves`
__device__ __forceinline__ void delta_work(const Sell& A, const int32_t* __restrict__ list, int64_t cnt, int64_t wave,
                                           int64_t nwaves, const float* __restrict__ x, float* __restrict__ y,
                                           double* part_out) {
  const int64_t npairs = (cnt + 1) / 2;
  const int64_t step = nwaves;
  int64_t t = wave;
  double part = 0.0;
  int64_t s0n = 0, s1n = -1, b0n = 0, b1n = 0;
  int w0n = 0, w1n = 0;
  if (t < npairs) {
    s0n = list[2 * t];
    s1n = 2 * t + 1 < cnt ? list[2 * t + 1] : -1;
    w0n = A.width[s0n];
    b0n = A.base_ptr[s0n];
    if (s1n >= 0) {
      w1n = A.width[s1n];
      b1n = A.base_ptr[s1n];
    }
  }
  while (t < npairs) {
    const int64_t s0 = s0n, s1 = s1n, b0 = b0n, b1 = b1n;
    const int w0 = w0n, w1 = w1n;
    const int64_t t2 = t + step;
    if (t2 < npairs) {
      s0n = list[2 * t2];
      s1n = 2 * t2 + 1 < cnt ? list[2 * t2 + 1] : -1;
      w0n = A.width[s0n];
      b0n = A.base_ptr[s0n];
      if (s1n >= 0) {
        w1n = A.width[s1n];
        b1n = A.base_ptr[s1n];
      }
    }
    bool done = false;
    if (s1 == s0 + 1 && w0 == w1) {
      switch (w0) {
        case -4: delta_pair<4>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -5: delta_pair<5>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -6: delta_pair<6>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -7: delta_pair<7>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -8: delta_pair<8>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -36: delta_pair<4, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -37: delta_pair<5, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -38: delta_pair<6, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -39: delta_pair<7, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        case -40: delta_pair<8, true>(A, s0, b0, b1, x, y, &part); done = true; break;
        default: break;
      }
    }
    if (!done) {
      delta_slice(A, s0, w0, b0, x, y, &part);
      if (s1 >= 0) delta_slice(A, s1, w1, b1, x, y, &part);
    }
    t = t2;
  }
  *part_out += part;
}

template <bool DOT>
__global__ void __launch_bounds__(kBS) k_sell_delta(Sell A, const int32_t* __restrict__ list, int64_t cnt,
                                                    const float* __restrict__ x, float* __restrict__ y,
                                                    const KrylovScalars* ks, double* __restrict__ partials) {
  __shared__ double s_red[kBS / 64];
  if (DOT && ks->done) return;
  double part = 0.0;
  delta_work(A, list, cnt, __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBS / kW) + (threadIdx.x >> 6))),
             (int64_t)gridDim.x * (kBS / kW), x, y, &part);
  if (DOT) {
    part = dev::block_sum<kBS>(part, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = part;
  }
}

// packed work of wave `wave` among `nwaves`: one slice per visit; the next slice's
// 32-byte header is requested before the current slice is processed
__device__ __forceinline__ void packed_work(const Sell& A, const int32_t* __restrict__ list, int64_t cnt,
                                            int64_t wave, int64_t nwaves, const float* __restrict__ x,
                                            float* __restrict__ y, double* part_out) {
  int64_t i = wave;
  double part = 0.0;
  int64_t sn = i < cnt ? list[i] : 0;
  PkHdr hn;
  if (i < cnt) hn = A.pk_hdr[sn];
  while (i < cnt) {
    const int64_t s = sn;
    const PkHdr h = hn;
    const int64_t i2 = i + nwaves;
    if (i2 < cnt) {
      sn = list[i2];
      hn = A.pk_hdr[sn];
    }
    if (h.nd <= kW) packed_slice_h<1>(A, s, h, x, y, &part);
    else if (h.nd <= 2 * kW) packed_slice_h<2>(A, s, h, x, y, &part);
    else packed_slice_h<4>(A, s, h, x, y, &part);
    i = i2;
  }
  *part_out += part;
}

template <bool DOT>
__global__ void __launch_bounds__(kBS) k_sell_packed(Sell A, const int32_t* __restrict__ list, int64_t cnt,
                                                     const float* __restrict__ x, float* __restrict__ y,
                                                     const KrylovScalars* ks, double* __restrict__ partials) {
  __shared__ double s_red[kBS / 64];
  if (DOT && ks->done) return;
  double part = 0.0;
  packed_work(A, list, cnt, __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBS / kW) + (threadIdx.x >> 6))),
              (int64_t)gridDim.x * (kBS / kW), x, y, &part);
  if (DOT) {
    part = dev::block_sum<kBS>(part, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = part;
  }
}

// Delta and packed slices in one launch, interleaved within every wave: each wave
// owns an equal share of both lists (items w, w + W, ...) and alternates one delta
// pair with one packed slice, so that at any time the round-trip-bound packed work
// and the streaming delta work of the CU's waves overlap, and no wave idles while
// the other list still has work.
//
// XCD = true: the hardware deals blocks to the 8 XCDs round-robin (block b on XCD
// b % 8), so the blocks of one XCD take one contiguous eighth of each list and
// interleave within it; every XCD's L2 then sees the input vector entries of its own
// row range instead of entries from all over the vector. Needs a grid that is a
// multiple of 8.
template <bool DOT, bool XCD, bool NT = false>
__global__ void __launch_bounds__(kBS) k_sell_both(Sell A, const int32_t* __restrict__ dlist, int64_t dcnt,
                                                   const int32_t* __restrict__ plist, int64_t pcnt,
                                                   const float* __restrict__ x, float* __restrict__ y,
                                                   const KrylovScalars* ks, double* __restrict__ partials) {
  __shared__ double s_red[kBS / 64];
  if (DOT && ks->done) return;
  const int64_t npairs = (dcnt + 1) / 2;
  int64_t W, wave, dlo = 0, dhi = npairs, plo = 0, phi = pcnt;
  if (XCD) {
    const int64_t xcd = blockIdx.x % 8u, local = blockIdx.x / 8u;
    W = (int64_t)(gridDim.x / 8u) * (kBS / kW);
    wave = __builtin_amdgcn_readfirstlane((int)(local * (kBS / kW) + (threadIdx.x >> 6)));
    dlo = xcd * npairs / 8;
    dhi = (xcd + 1) * npairs / 8;
    plo = xcd * pcnt / 8;
    phi = (xcd + 1) * pcnt / 8;
  } else {
    W = (int64_t)gridDim.x * (kBS / kW);
    wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBS / kW) + (threadIdx.x >> 6)));
  }
  double part = 0.0;
  int64_t td = dlo + wave, tp = plo + wave;
  while (td < dhi || tp < phi) {
    if (td < dhi) {
      const int64_t s0 = dlist[2 * td];
      const int64_t s1 = 2 * td + 1 < dcnt ? dlist[2 * td + 1] : -1;
      const int w0 = A.width[s0];
      const int64_t b0 = A.base_ptr[s0];
      const int w1 = s1 >= 0 ? A.width[s1] : 0;
      const int64_t b1 = s1 >= 0 ? A.base_ptr[s1] : 0;
      bool done = false;
      if (s1 == s0 + 1 && w0 == w1) {
        switch (w0) {
          case -4: delta_pair<4, false, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          case -5: delta_pair<5, false, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          case -6: delta_pair<6, false, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          case -7: delta_pair<7, false, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          case -8: delta_pair<8, false, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          case -36: delta_pair<4, true, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          case -37: delta_pair<5, true, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          case -38: delta_pair<6, true, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          case -39: delta_pair<7, true, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          case -40: delta_pair<8, true, NT>(A, s0, b0, b1, x, y, &part); done = true; break;
          default: break;
        }
      }
      if (!done) {
        delta_slice<NT>(A, s0, w0, b0, x, y, &part);
        if (s1 >= 0) delta_slice<NT>(A, s1, w1, b1, x, y, &part);
      }
      td += W;
    }
    if (tp < phi) {
      const int64_t s = plist[tp];
      const PkHdr h = A.pk_hdr[s];
      if (h.nd <= kW) packed_slice_h<1, NT>(A, s, h, x, y, &part);
      else if (h.nd <= 2 * kW) packed_slice_h<2, NT>(A, s, h, x, y, &part);
      else packed_slice_h<4, NT>(A, s, h, x, y, &part);
      tp += W;
    }
  }
  if (DOT) {
    part = dev::block_sum<kBS>(part, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = part;
  }
}

// ---- pipelined SpMV (schedule 6)
// The work of k_sell_both<DOT, true> (same lists, same XCD segmentation, same items per
// wave in the same order, same per-row arithmetic, so the same results bit for bit),
// with the metadata of a wave's items requested ahead so that an item's operator values
// and input gathers go out together in one memory round trip:
//   item i + 2  the delta pair's record (slice ids, slot pointers, widths, pairing) and
//               the packed slice's id and 32-byte header (scalar loads);
//   item i + 1  the slot deltas (and constant-slice values) of both delta slices, one
//               per lane (lanes 0-15 the first slice, 16-31 the second), and the packed
//               slice's dictionary, row map and dictionary-index words (vector loads);
//   item i      values and gathers; deltas come from the prefetched lanes by readlane,
//               packed columns from the prefetched index words and dictionary.
struct DPair {
  int32_t s0, s1;  // slice ids (s1 = -1: none)
  int32_t b0, b1;  // slot pointers (base_ptr)
  int32_t w0, w1;  // widths
  int32_t paired;  // 1: s1 = s0 + 1 with the same width of 4 .. 8 slots (plain or constant)
  int32_t pad;
};
static_assert(sizeof(DPair) == 32, "DPair is 32 bytes");

__device__ __forceinline__ int delta_slots(int w) { return w >= -16 ? -w : (w >= -48 ? -w - kConstBase : 0); }
__device__ __forceinline__ bool delta_const(int w) { return w <= -(kConstBase + 1) && w >= -(kConstBase + 16); }

__global__ void k_dpair_fill(int64_t npairs, const int32_t* __restrict__ list, int64_t cnt,
                             const int32_t* __restrict__ width, const int64_t* __restrict__ base_ptr,
                             DPair* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  DPair m;
  m.s0 = list[2 * t];
  m.s1 = 2 * t + 1 < cnt ? list[2 * t + 1] : -1;
  m.w0 = width[m.s0];
  m.b0 = (int32_t)base_ptr[m.s0];
  m.w1 = m.s1 >= 0 ? width[m.s1] : 0;
  m.b1 = m.s1 >= 0 ? (int32_t)base_ptr[m.s1] : 0;
  const int sw = m.w0 <= -(kConstBase + 1) ? -m.w0 - kConstBase : -m.w0;
  m.paired = (m.s1 == m.s0 + 1 && m.w0 == m.w1 && sw >= 4 && sw <= 8) ? 1 : 0;
  m.pad = 0;
  out[t] = m;
}

__global__ void k_pk_hdr_list(int64_t cnt, const int32_t* __restrict__ list, const PkHdr* __restrict__ hdr,
                              PkHdr* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) out[i] = hdr[list[i]];
}

// ---- lean SpMV (schedule 6)
// The work of k_sell_both<DOT, true> (same lists and XCD segmentation, the same items
// per wave in the same order, the same per-row arithmetic) with fewer vector
// instructions per entry:
//   * the input vector has kGuard zero entries on either side (spmv_input), so gathers
//     need no clamping: every gathered index is a column of some row of the same
//     slice plus a row offset below 64 (absent slots and rows past n included);
//   * delta slots address x through a scalar base (x + slice row 0 + delta) and the
//     lane index, values through a scalar base and an immediate slot offset;
//   * packed entries of rows that do not take part in a slot read the operator's one
//     +0 entry (the value the select of k_sell_both gives them), so no select follows;
//   * a pair's slice ids, slot pointers and widths come from one 32-byte record, a
//     packed slice's id and header from two independent loads.
// Results equal those of k_sell_both except possibly for the sign of an exact zero.
constexpr int kGuard = 64;

__device__ __forceinline__ void lean_dpair(const Sell& A, const DPair& m, const float* __restrict__ x,
                                           float* __restrict__ y, double* part) {
  if (m.paired) {
    switch (m.w0) {
      case -4: lean_pair<4, false>(A, m.s0, m.b0, m.b1, x, y, part); return;
      case -5: lean_pair<5, false>(A, m.s0, m.b0, m.b1, x, y, part); return;
      case -6: lean_pair<6, false>(A, m.s0, m.b0, m.b1, x, y, part); return;
      case -7: lean_pair<7, false>(A, m.s0, m.b0, m.b1, x, y, part); return;
      case -8: lean_pair<8, false>(A, m.s0, m.b0, m.b1, x, y, part); return;
      case -36: lean_pair<4, true>(A, m.s0, m.b0, m.b1, x, y, part); return;
      case -37: lean_pair<5, true>(A, m.s0, m.b0, m.b1, x, y, part); return;
      case -38: lean_pair<6, true>(A, m.s0, m.b0, m.b1, x, y, part); return;
      case -39: lean_pair<7, true>(A, m.s0, m.b0, m.b1, x, y, part); return;
      case -40: lean_pair<8, true>(A, m.s0, m.b0, m.b1, x, y, part); return;
      default: break;
    }
  }
  lean_delta(A, m.s0, m.w0, m.b0, x, y, part);
  if (m.s1 >= 0) lean_delta(A, m.s1, m.w1, m.b1, x, y, part);
}

// 72 VGPRs without spills (7 waves per SIMD against 6 at the compiler's 79): in-solve
// SpMV of the 8.04M-unknown globular system 54.6 -> 53.2 us; 63.3M rows unchanged.
template <bool DOT>
__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(7))) k_sell_lean(Sell A, const DPair* __restrict__ dp, int64_t npairs,
                                                   const int32_t* __restrict__ plist, const PkHdr* __restrict__ hl,
                                                   int64_t pcnt, const float* __restrict__ x, float* __restrict__ y,
                                                   const KrylovScalars* ks, double* __restrict__ partials) {
  __shared__ double s_red[kBS / 64];
  if (DOT && ks->done) return;
  const int64_t xcd = blockIdx.x % 8u, local = blockIdx.x / 8u;
  const int64_t W = (int64_t)(gridDim.x / 8u) * (kBS / kW);
  const int64_t wave = __builtin_amdgcn_readfirstlane((int)(local * (kBS / kW) + (threadIdx.x >> 6)));
  const int64_t dhi = (xcd + 1) * npairs / 8, phi = (xcd + 1) * pcnt / 8;
  int64_t td = xcd * npairs / 8 + wave, tp = xcd * pcnt / 8 + wave;
  double part = 0.0;
  while (td < dhi || tp < phi) {
    if (td < dhi) {
      lean_dpair(A, dp[td], x, y, &part);
      td += W;
    }
    if (tp < phi) {
      const int64_t s = plist[tp];
      const PkHdr h = hl[tp];
      if (h.nd <= kW) lean_packed<1>(A, s, h, x, y, &part);
      else if (h.nd <= 2 * kW) lean_packed<2>(A, s, h, x, y, &part);
      else lean_packed<4>(A, s, h, x, y, &part);
      tp += W;
    }
  }
  if (DOT) {
    part = dev::block_sum<kBS>(part, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = part;
  }
}

// zero the guard entries of the SpMV input vector (p points kGuard entries before x_0)
__global__ void k_zero_guard(float* __restrict__ p, int64_t n) {
  const int t = threadIdx.x;
  if (t < kGuard) p[t] = 0.f;
  else if (t < 2 * kGuard) p[kGuard + n + (t - kGuard)] = 0.f;
}

// offset slices (rows too irregular for the other forms; rare)
template <bool DOT>
__global__ void __launch_bounds__(kBS) k_sell_offset(Sell A, const int32_t* __restrict__ list, int64_t cnt,
                                                     const float* __restrict__ x, float* __restrict__ y,
                                                     const KrylovScalars* ks, double* __restrict__ partials) {
  __shared__ double s_red[kBS / 64];
  if (DOT && ks->done) return;
  const int64_t step = (int64_t)gridDim.x * (kBS / kW);
  double part = 0.0;
  for (int64_t i = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBS / kW) + (threadIdx.x >> 6))); i < cnt;
       i += step) {
    const int64_t s = list[i];
    offset_slice(A, s, A.width[s], A.base_ptr[s], x, y, &part);
  }
  if (DOT) {
    part = dev::block_sum<kBS>(part, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = part;
  }
}

// slice form flags from the widths (0 delta, 1 packed, 2 offset)
__global__ void k_sell_forms(int64_t ns, const int32_t* __restrict__ width, uint8_t* __restrict__ fd,
                             uint8_t* __restrict__ fp, uint8_t* __restrict__ fo) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int w = width[s];
  fp[s] = w <= -(kPackedBase + 1) ? 1 : 0;
  fd[s] = (w < 0 && w > -(kPackedBase + 1)) ? 1 : 0;
  fo[s] = w >= 0 ? 1 : 0;
}

// Iteration of parity par: every block forms (p, q) from the SpMV partials in the
// same order, alpha = rho / (p, q); rhat -= alpha q; per block (rhat, rhat).
// (xhat += alpha p is done by the direction kernel of the same iteration, or by the
// accumulation after the iteration that raised the stop flag.)
__global__ void __launch_bounds__(kBS) k_mx_update(int64_t n, float* __restrict__ rs, const float* __restrict__ qs,
                                                   KrylovScalars* ks, const double* __restrict__ pq_part, int npq,
                                                   double* __restrict__ partials, int par) {
  __shared__ double s_red[kBS / 64];
  if (ks->done) return;
  const double pq = dev::sum_partials<kBS>(pq_part, npq, 1, s_red);
  const double alpha_d = ks->rho2[par] / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ks->pq = pq;
    ks->alpha2[par] = alpha_d;
  }
  const float alpha = (float)alpha_d;
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    // q is dead after this read: a non-temporal load leaves the last-level cache to the
    // operator (with the non-temporal rhat read of k_mx_dir_x: in-solve SpMV of the
    // 8M-unknown globular system 62.7 -> 57.0 us, mixed solve 60.9 -> 58.1 ms)
    const float ri = fmaf(-alpha, __builtin_nontemporal_load(&qs[i]), rs[i]);
    rs[i] = ri;
    rr += (double)ri * (double)ri;
  }
  rr = dev::block_sum<kBS>(rr, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x] = rr;
}

// p = rhat + beta p (first direction of a cycle)
__global__ void __launch_bounds__(kBS) k_mx_dir(int64_t n, const float* __restrict__ rs, float* __restrict__ ps,
                                                const KrylovScalars* ks) {
  if (ks->done) return;
  const float beta = (float)ks->beta;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS)
    ps[i] = fmaf(beta, ps[i], rs[i]);
}

// Stop test on (rhat, rhat) (every block decides alike); otherwise xhat += alpha p and
// p = rhat + beta p. Block 0 records the scalars for the next iteration and the host.
__global__ void __launch_bounds__(kBS) k_mx_dir_x(int64_t n, const float* __restrict__ rs, float* __restrict__ ps,
                                                  float* __restrict__ xs, KrylovScalars* ks,
                                                  const double* __restrict__ rr_part, int npr, int par) {
  __shared__ double s_red[kBS / 64];
  if (ks->done) return;
  const double RR = dev::sum_partials<kBS>(rr_part, npr, 1, s_red);
  const int it = ks->iter2[par] + 1;
  const bool stop = RR <= ks->stop_rr || it >= ks->maxiter;
  const double rho = ks->rho2[par];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ks->rr = RR;
    ks->iter = it;
    ks->iter2[par ^ 1] = it;
    if (stop) {
      ks->done = 1;
      ks->stop_par = par;
    } else {
      ks->beta = RR / rho;
      ks->rho = RR;
      ks->rho2[par ^ 1] = RR;
    }
  }
  if (stop) return;
  const float alpha = (float)ks->alpha2[par];
  const float beta = (float)(RR / rho);
  // xhat is next touched a whole iteration later, after every other stream of the
  // iteration has passed through the caches, so it is loaded and stored non-temporally
  // and leaves the last-level cache to the operator and the other vectors (measured:
  // mixed solve of the 8M-unknown globular system 78.2 -> 74.7 ms)
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const float pi = ps[i];
    __builtin_nontemporal_store(fmaf(alpha, pi, __builtin_nontemporal_load(&xs[i])), &xs[i]);
    ps[i] = fmaf(beta, pi, __builtin_nontemporal_load(&rs[i]));  // rhat: next read after the SpMV
  }
}

// x += S (xhat + alpha p) for the iteration that raised the stop flag; xhat = 0
// perm (NULL: identity): FP32 row i is FP64 row perm[i] (banded row order)
__global__ void __launch_bounds__(kBS) k_mx_accum(int64_t n, const double* __restrict__ dscale, double* __restrict__ x,
                                                  float* __restrict__ xs, const float* __restrict__ ps,
                                                  const KrylovScalars* ks, const int32_t* __restrict__ perm) {
  const float alpha = (float)ks->alpha2[ks->stop_par];
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const int64_t o = perm ? (int64_t)perm[i] : i;
    x[o] += dscale[i] * (double)fmaf(alpha, ps[i], xs[i]);
    xs[i] = 0.f;
  }
}

// r = b - t (t = A x in FP64); rhat = S r; per block (r, r), (rhat, rhat).
// first: r = b and x = 0, xhat = 0, p = 0.
// (with perm: FP32 row i takes FP64 row perm[i]; the sums run in FP32 row order)
__global__ void __launch_bounds__(kBS) k_mx_resid(int64_t n, const double* __restrict__ b, const double* __restrict__ t,
                                                  const double* __restrict__ dscale, float* __restrict__ rs,
                                                  double* __restrict__ x, float* __restrict__ xs, float* __restrict__ ps,
                                                  int first, double* __restrict__ partials,
                                                  const int32_t* __restrict__ perm) {
  __shared__ double s_red[kBS / 64];
  double rr = 0.0, ss = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const int64_t o = perm ? (int64_t)perm[i] : i;
    double ri;
    if (first) {
      ri = b[o];
      x[o] = 0.0;
      xs[i] = 0.f;
      ps[i] = 0.f;
    } else {
      ri = b[o] - t[o];
    }
    const float si = (float)(dscale[i] * ri);
    rs[i] = si;
    rr += ri * ri;
    ss += (double)si * (double)si;
  }
  rr = dev::block_sum<kBS>(rr, s_red);
  ss = dev::block_sum<kBS>(ss, s_red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = rr;
    partials[2 * blockIdx.x + 1] = ss;
  }
}

__global__ void __launch_bounds__(kFin) k_mx_fin_resid(KrylovScalars* ks, const double* __restrict__ partials, int np,
                                                       int first) {
  __shared__ double s_red[kFin / 64];
  const double RR = dev::sum_partials<kFin>(partials, np, 2, s_red);
  const double SS = dev::sum_partials<kFin>(partials + 1, np, 2, s_red);
  if (threadIdx.x == 0) {
    if (first) {
      ks->bb = RR;
      ks->beta = 0.0;
      ks->iter = 0;
    } else {
      ks->beta = ks->rho > 0.0 ? SS / ks->rho : 0.0;
    }
    ks->aux[0] = RR;  // true FP64 residual norm^2
    ks->aux[1] = SS;  // scaled residual norm^2 after replacement
    ks->rho = SS;
    ks->rr = SS;
  }
}

__global__ void k_to_scaled(int64_t n, const double* __restrict__ x, const double* __restrict__ dscale,
                            float* __restrict__ xs, const int32_t* __restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xs[i] = (float)(x[perm ? (int64_t)perm[i] : i] / dscale[i]);
}
__global__ void k_from_scaled(int64_t n, const float* __restrict__ ys, const double* __restrict__ dscale,
                              double* __restrict__ y, const int32_t* __restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[perm ? (int64_t)perm[i] : i] = (double)ys[i] / dscale[i];
}

// ---- banded row order of the compact operator
// Mesh nodes come in lexicographic order (x fastest, then y, then z), so a row's z
// neighbours are a whole z-plane of rows away; where planes hold hundreds of
// thousands of nodes, the input entries a plane's rows gather are evicted from L2
// before the next plane but one gathers them again. The banded order cuts y into bands
// of `band` finest-grid spacings and takes the bands one after another, each in the
// node order, so that a row and its z neighbour are one band-plane apart. It is a
// stable partition of the node order by band; the FP32 solve runs in this order and
// maps rows to nodes where it meets FP64 vectors (residual, accumulation).
__global__ void k_leaf_edge(int64_t L, const int32_t* __restrict__ corner, const double* __restrict__ node_xyz,
                            const double* __restrict__ hang_xyz, double* __restrict__ out) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= L) return;
  const int32_t a = corner[8 * l], b = corner[8 * l + 7];  // opposite corners (z-order 0 and 7)
  const double xa = a >= 0 ? node_xyz[3 * (int64_t)a] : hang_xyz[3 * (int64_t)(-1 - a)];
  const double xb = b >= 0 ? node_xyz[3 * (int64_t)b] : hang_xyz[3 * (int64_t)(-1 - b)];
  out[l] = fabs(xb - xa);
}

// per-block minimum and maximum of a[off + stride * i], i < n
__global__ void __launch_bounds__(256) k_minmax(int64_t n, const double* __restrict__ a, int64_t stride, int64_t off,
                                                double* __restrict__ part) {
  __shared__ double smin[256], smax[256];
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double v = a[off + stride * i];
    mn = fmin(mn, v);
    mx = fmax(mx, v);
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

__global__ void k_band_key(int64_t n, const double* __restrict__ node_xyz, double y0, double inv_w,
                           uint32_t* __restrict__ key, int32_t* __restrict__ id) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = (uint32_t)floor((node_xyz[3 * i + 1] - y0) * inv_w);
  id[i] = (int32_t)i;
}

__global__ void k_inv_perm(int64_t n, const int32_t* __restrict__ perm, int32_t* __restrict__ inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[perm[i]] = (int32_t)i;
}

__global__ void k_perm_len(int64_t n, const int32_t* __restrict__ perm, const int32_t* __restrict__ rowptr,
                           int64_t* __restrict__ len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) len[i] = rowptr[perm[i] + 1] - rowptr[perm[i]];
}

// row i of P A P^T: row perm[i] of A with columns renumbered by inv, sorted ascending
__global__ void k_perm_fill(int64_t n, const int32_t* __restrict__ perm, const int32_t* __restrict__ inv,
                            const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                            const double* __restrict__ val, const int64_t* __restrict__ off,
                            int32_t* __restrict__ rp2, int32_t* __restrict__ col2, double* __restrict__ val2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t o = perm[i];
  const int64_t b = off[i];
  const int32_t k0 = rowptr[o], k1 = rowptr[o + 1];
  for (int32_t k = k0; k < k1; ++k) {
    const int32_t cn = inv[col[k]];
    const double v = val[k];
    int64_t j = b + (k - k0);
    while (j > b && col2[j - 1] > cn) {
      col2[j] = col2[j - 1];
      val2[j] = val2[j - 1];
      --j;
    }
    col2[j] = cn;
    val2[j] = v;
  }
  rp2[i] = (int32_t)b;
  if (i == n - 1) rp2[n] = (int32_t)(b + (k1 - k0));
}

Sell sell_view(const ngpb_ctx* c) {
  Sell A;
  A.n = c->n;
  A.nslices = c->n_slices;
  A.width = c->slice_width.p;
  A.base_ptr = c->slice_base_ptr.p;
  A.wide_ptr = c->slice_wide_ptr.p;
  A.ov_ptr = c->slice_ov_ptr.p;
  A.val = c->sell_val.p;
  A.off = c->sell_off.p;
  A.base = c->sell_base.p;
  A.col32 = c->sell_col32.p;
  A.ov_cnt = c->sell_ov_cnt.p;
  A.ov_val = c->sell_ov_val.p;
  A.ov_col = c->sell_ov_col.p;
  A.pk_val = c->sell_pk_val.p;
  A.pk_zero = (uint32_t)c->sell_pk_entries;
  A.pk_idx = c->sell_pk_idx.p;
  A.pk_lp = c->sell_pk_lp.p;
  A.pk_hdr = reinterpret_cast<const PkHdr*>(c->sell_pk_hdr.p);
  A.pk_dict = c->sell_pk_dict.p;
  A.cval = c->sell_cval.p;
  return A;
}

// blocks of the form-specialised SpMV kernels resident on the device at once
unsigned sell_resident(ngpb_ctx* c) {
  if (c->sell_resident > 0) return (unsigned)c->sell_resident;
  int cus = 0, per_cu = 0;
  NGPB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  if (c->spmv_sched == 6)
    NGPB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sell_lean<true>, kBS, 0));
  else
    NGPB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sell_both<true, false>, kBS, 0));
  const int r = std::max(8, (std::max(cus, 1) * std::max(per_cu, 1)) / 8 * 8);
  c->sell_resident = r;
  return (unsigned)r;
}

// blocks for `units` wave work units (4 waves per block): at most 2048 (schedule 0)
// or the resident block count (schedule 2); 0 when there is no work
unsigned form_grid(ngpb_ctx* c, int64_t units) {
  if (units <= 0) return 0;
  const int64_t need = (units + (kBS / kW) - 1) / (kBS / kW);
  const int64_t cap = c->spmv_sched == 2 ? (int64_t)sell_resident(c) : 2048;
  return (unsigned)std::min<int64_t>(need, cap);
}

// number of (p, q) partials the SpMV of the current schedule writes
int sell_partials(ngpb_ctx* c) {
  if (c->spmv_sched >= 3)  // combined delta/packed grid, then the offset kernel's blocks
    return (int)(std::max<int64_t>(3, sell_resident(c)) + form_grid(c, c->n_list_offset));
  if (c->spmv_sched == 1)
    return (int)std::max<int64_t>(1, std::min<int64_t>(((c->n_slices + 1) / 2 + (kBS / kW) - 1) / (kBS / kW), 2048));
  return (int)(form_grid(c, (c->n_list_delta + 1) / 2) + form_grid(c, c->n_list_packed) +
               form_grid(c, c->n_list_offset));
}

template <bool DOT>
void launch_sell_t(ngpb_ctx* c, const float* x, float* y, double* part) {
  const KrylovScalars* ks = DOT ? c->ks.p : nullptr;
  const Sell A = sell_view(c);
  if (c->spmv_sched == 1) {
    const unsigned g = (unsigned)sell_partials(c);
    spmv_launch(c, k_sell_spmv<DOT>, dim3(g), dim3(kBS), A, x, y, ks, part);
    return;
  }
  if (c->spmv_sched >= 3) {
    const unsigned g = (unsigned)std::max<int64_t>(3, sell_resident(c));
    if (c->spmv_sched == 4) {
      NGPB_REQUIRE(g % 8 == 0, NGPB_E_STATE, "XCD-segmented SpMV needs a grid that is a multiple of 8");
      spmv_launch(c, k_sell_both<DOT, true>, dim3(g), dim3(kBS), A, c->sell_list_delta.p,
                  c->n_list_delta, c->sell_list_packed.p, c->n_list_packed, x, y, ks, part);
    } else if (c->spmv_sched == 6) {
      NGPB_REQUIRE(g % 8 == 0, NGPB_E_STATE, "XCD-segmented SpMV needs a grid that is a multiple of 8");
      spmv_launch(c, k_sell_lean<DOT>, dim3(g), dim3(kBS), A,
                  reinterpret_cast<const DPair*>(c->sell_dpair.p), (c->n_list_delta + 1) / 2,
                  c->sell_list_packed.p, reinterpret_cast<const PkHdr*>(c->sell_pk_hdr_list.p),
                  c->n_list_packed, x, y, ks, part);
    } else if (c->spmv_sched == 5) {
      NGPB_REQUIRE(g % 8 == 0, NGPB_E_STATE, "XCD-segmented SpMV needs a grid that is a multiple of 8");
      spmv_launch(c, k_sell_both<DOT, true, true>, dim3(g), dim3(kBS), A, c->sell_list_delta.p,
                  c->n_list_delta, c->sell_list_packed.p, c->n_list_packed, x, y, ks, part);
    } else {
      spmv_launch(c, k_sell_both<DOT, false>, dim3(g), dim3(kBS), A, c->sell_list_delta.p,
                  c->n_list_delta, c->sell_list_packed.p, c->n_list_packed, x, y, ks, part);
    }
    const unsigned go = form_grid(c, c->n_list_offset);
    if (go)
      spmv_launch(c, k_sell_offset<DOT>, dim3(go), dim3(kBS), A, c->sell_list_offset.p,
                  c->n_list_offset, x, y, ks, DOT ? part + g : nullptr);
    return;
  }
  const unsigned gd = form_grid(c, (c->n_list_delta + 1) / 2);
  const unsigned gp = form_grid(c, c->n_list_packed);
  const unsigned go = form_grid(c, c->n_list_offset);
  if (gd)
    spmv_launch(c, k_sell_delta<DOT>, dim3(gd), dim3(kBS), A, c->sell_list_delta.p, c->n_list_delta,
                x, y, ks, part);
  if (gp)
    spmv_launch(c, k_sell_packed<DOT>, dim3(gp), dim3(kBS), A, c->sell_list_packed.p,
                c->n_list_packed, x, y, ks, DOT ? part + gd : nullptr);
  if (go)
    spmv_launch(c, k_sell_offset<DOT>, dim3(go), dim3(kBS), A, c->sell_list_offset.p,
                c->n_list_offset, x, y, ks, DOT ? part + gd + gp : nullptr);
}

void launch_sell(ngpb_ctx* c, bool dot, const float* x, float* y, double* part) {
  if (dot) launch_sell_t<true>(c, x, y, part);
  else launch_sell_t<false>(c, x, y, nullptr);
}

}  // namespace

namespace {
// Decides the row order of the compact operator and, for the banded order, computes
// c->row_perm (FP32 row -> node). Banded needs the context's own mesh and assembly.
bool band_rows(ngpb_ctx* c, bool allow) {
  c->sell_banded = false;
  if (!allow || !c->have_mesh || c->external_csr || c->n != c->N || c->L <= 0 || c->n >= INT32_MAX) return false;
  const int band = c->row_band;
  if (band <= 0) return false;
  const int64_t n = c->n, L = c->L;
  DBuf<double> edge;
  edge.resize((size_t)L);
  hipLaunchKernelGGL(k_leaf_edge, dim3(blocks_for(L, 256)), dim3(256), 0, c->stream, L, c->leaf_corner.p,
                     c->node_xyz.p, c->hang_xyz.p, edge.p);
  const unsigned gb = (unsigned)std::min<int64_t>(1024, blocks_for(std::max(L, n), 256));
  DBuf<double> part;
  part.resize((size_t)4 * gb);
  hipLaunchKernelGGL(k_minmax, dim3(gb), dim3(256), 0, c->stream, L, edge.p, (int64_t)1, (int64_t)0, part.p);
  hipLaunchKernelGGL(k_minmax, dim3(gb), dim3(256), 0, c->stream, n, c->node_xyz.p, (int64_t)3, (int64_t)1,
                     part.p + 2 * gb);
  NGPB_HIP(hipGetLastError());
  std::vector<double> hp((size_t)4 * gb);
  NGPB_HIP(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  double hmin = INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (unsigned k = 0; k < gb; ++k) {
    hmin = std::min(hmin, hp[2 * k]);
    y0 = std::min(y0, hp[2 * gb + 2 * k]);
    y1 = std::max(y1, hp[2 * gb + 2 * k + 1]);
  }
  if (!(hmin > 0.0) || !std::isfinite(y0) || !std::isfinite(y1)) return false;
  const double w = band * hmin;
  const double nb = std::floor((y1 - y0) / w) + 1.0;
  if (nb < 2.0 || nb > 1e6) return false;
  int bits = 1;
  while ((double)(1u << bits) < nb + 1.0) ++bits;
  DBuf<uint32_t> key, key2;
  DBuf<int32_t> id;
  key.resize((size_t)n);
  key2.resize((size_t)n);
  id.resize((size_t)n);
  c->row_perm.resize((size_t)n);
  hipLaunchKernelGGL(k_band_key, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, c->node_xyz.p, y0, 1.0 / w,
                     key.p, id.p);
  NGPB_HIP(hipGetLastError());
  size_t tb = 0;
  NGPB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.p, key2.p, id.p, c->row_perm.p, (int)n, 0, bits,
                                              c->stream));
  DBuf<uint8_t> tmp;
  tmp.resize(std::max<size_t>(tb, 1));
  NGPB_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.p, key2.p, id.p, c->row_perm.p, (int)n, 0, bits,
                                              c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  c->sell_banded = true;
  c->sell_band = band;
  return true;
}
}  // namespace

void build_compact_from(ngpb_ctx* c, const int32_t* rowptr, const int32_t* colidx, const double* vals);

void build_compact(ngpb_ctx* c, bool allow_band) {
  if (!band_rows(c, allow_band)) {
    build_compact_from(c, c->rowptr.p, c->col.p, c->val.p);
    return;
  }
  const int64_t n = c->n, nnz = c->nnz;
  DBuf<int32_t> inv, rp2, col2;
  DBuf<int64_t> len, off;
  DBuf<double> val2;
  inv.resize((size_t)n);
  len.resize((size_t)n);
  off.resize((size_t)n);
  rp2.resize((size_t)n + 1);
  col2.resize((size_t)std::max<int64_t>(nnz, 1));
  val2.resize((size_t)std::max<int64_t>(nnz, 1));
  const unsigned g = (unsigned)blocks_for(n, 256);
  hipLaunchKernelGGL(k_inv_perm, dim3(g), dim3(256), 0, c->stream, n, c->row_perm.p, inv.p);
  hipLaunchKernelGGL(k_perm_len, dim3(g), dim3(256), 0, c->stream, n, c->row_perm.p, c->rowptr.p, len.p);
  NGPB_HIP(hipGetLastError());
  const int64_t tot = exclusive_scan(c, len.p, off.p, n);
  NGPB_REQUIRE(tot == nnz, NGPB_E_STATE, "banded row order: entry count mismatch");
  hipLaunchKernelGGL(k_perm_fill, dim3(g), dim3(256), 0, c->stream, n, c->row_perm.p, inv.p, c->rowptr.p, c->col.p,
                     c->val.p, off.p, rp2.p, col2.p, val2.p);
  NGPB_HIP(hipGetLastError());
  build_compact_from(c, rp2.p, col2.p, val2.p);
}

void build_compact_from(ngpb_ctx* c, const int32_t* rowptr, const int32_t* colidx, const double* vals) {
  const int64_t n = c->n;
  const int64_t ns = (n + kW - 1) / kW;
  c->n_slices = ns;
  c->dscale.resize((size_t)n);
  hipLaunchKernelGGL(k_scale, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, rowptr, colidx, vals,
                     c->dscale.p);
  c->slice_width.resize((size_t)ns);
  // per-slice counts and offsets in the context's scratch slots
  struct Arr {
    int64_t* p;
  };
  int slot = kTmpCompact;
  auto arr = [&]() { return Arr{tmp_buf<int64_t>(c, slot++, (size_t)ns)}; };
  Arr wcount = arr(), widec = arr(), wide_off = arr(), ovc = arr(), ov_off = arr(), dslots = arr(), cslots = arr(),
      pkc = arr(), pk_off = arr(), pkd = arr(), pk_doff = arr(), pkg = arr(), pk_goff = arr();
  struct IArr {
    int* p;
  };
  IArr bad{tmp_buf<int>(c, slot++, 1)};
  NGPB_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), c->stream));
  const unsigned g = (unsigned)((ns + (kBS / kW) - 1) / (kBS / kW));
  hipLaunchKernelGGL(k_sell_shape, dim3(g), dim3(kBS), 0, c->stream, n, ns, rowptr, colidx, vals,
                     c->dscale.p, c->sell_allow_const ? 1 : 0, cslots.p, c->slice_width.p, wcount.p, widec.p, ovc.p,
                     dslots.p, pkc.p, pkd.p, pkg.p, bad.p);
  NGPB_HIP(hipGetLastError());
  int hb = 0;
  NGPB_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  NGPB_REQUIRE(hb == 0, NGPB_E_ARG, "a row has more than 255 off-diagonal overflow entries");
  const int64_t nbases = exclusive_scan(c, wcount.p, wcount.p, ns);  // wcount -> base_ptr
  const int64_t nwide = exclusive_scan(c, widec.p, wide_off.p, ns);
  const int64_t nov = exclusive_scan(c, ovc.p, ov_off.p, ns);
  const int64_t ndslots = exclusive_scan(c, dslots.p, dslots.p, ns);
  const int64_t ncslots = exclusive_scan(c, cslots.p, cslots.p, ns);
  const int64_t npk = exclusive_scan(c, pkc.p, pk_off.p, ns);
  const int64_t npkd = exclusive_scan(c, pkd.p, pk_doff.p, ns);
  const int64_t npkg = exclusive_scan(c, pkg.p, pk_goff.p, ns);
  NGPB_REQUIRE(npk < INT32_MAX && npkd < INT32_MAX && npkg < INT32_MAX, NGPB_E_ARG,
               "operator too large for 32-bit packed-slice offsets");
  {
    // packed slices = slices with a dictionary (every packed slice has >= 1 delta or
    // rows without off-diagonals only, which are counted through their width below)
    std::vector<int32_t> wh((size_t)ns);
    NGPB_HIP(hipMemcpyAsync(wh.data(), c->slice_width.p, ns * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    NGPB_HIP(hipStreamSynchronize(c->stream));
    int64_t k = 0;
    int64_t kc = 0;
    for (int32_t w : wh) {
      k += w <= -(kPackedBase + 1) ? 1 : 0;
      kc += (w <= -(kConstBase + 1) && w >= -(kConstBase + kDMax)) ? 1 : 0;
    }
    c->sell_pk_slices = k;
    c->sell_const_slices = kc;
  }
  c->sell_pk_entries = npk;
  c->sell_pk_dict_entries = npkd;
  // one more value, +0, read by the lean SpMV for rows that do not take part in a slot
  c->sell_pk_val.resize((size_t)npk + 1);
  NGPB_HIP(hipMemsetAsync(c->sell_pk_val.p + npk, 0, sizeof(float), c->stream));
  c->sell_pk_words = npkg;
  c->sell_pk_idx.resize((size_t)std::max<int64_t>(npkg, 1));
  // bytes of rows that end inside a group stay 0, a valid dictionary index
  NGPB_HIP(hipMemsetAsync(c->sell_pk_idx.p, 0, c->sell_pk_idx.bytes(), c->stream));
  c->sell_pk_lp.resize((size_t)ns * kW);
  c->sell_pk_hdr.resize((size_t)ns * (sizeof(PkHdr) / sizeof(int64_t)));
  c->sell_pk_dict.resize((size_t)std::max<int64_t>(npkd, 1));
  c->sell_bases = nbases;
  c->sell_entries = nbases * kW;
  c->sell_delta_entries = ndslots * kW;
  c->sell_const_slots = ncslots;
  c->sell_cval.resize((size_t)std::max<int64_t>(nbases, 1));
  c->sell_wide_entries = nwide;
  c->sell_ov_entries = nov;
  c->slice_base_ptr.resize((size_t)ns);
  NGPB_HIP(hipMemcpyAsync(c->slice_base_ptr.p, wcount.p, ns * sizeof(int64_t), hipMemcpyDeviceToDevice, c->stream));
  c->slice_wide_ptr.resize((size_t)ns);
  c->slice_ov_ptr.resize((size_t)ns);
  c->sell_val.resize((size_t)std::max<int64_t>(c->sell_entries, 1));
  c->sell_off.resize((size_t)std::max<int64_t>(c->sell_entries, 1));
  c->sell_base.resize((size_t)std::max<int64_t>(nbases, 1));
  c->sell_col32.resize((size_t)std::max<int64_t>(nwide, 1));
  c->sell_ov_cnt.resize((size_t)n);
  c->sell_ov_val.resize((size_t)std::max<int64_t>(nov, 1));
  c->sell_ov_col.resize((size_t)std::max<int64_t>(nov, 1));
  hipLaunchKernelGGL(k_sell_fill, dim3(g), dim3(kBS), 0, c->stream, n, ns, rowptr, colidx, vals, c->dscale.p,
                     c->slice_width.p, c->slice_base_ptr.p, wide_off.p, widec.p, ov_off.p, ovc.p, c->sell_val.p,
                     c->sell_off.p, c->sell_base.p, c->sell_col32.p, c->slice_wide_ptr.p, c->slice_ov_ptr.p,
                     c->sell_ov_cnt.p, c->sell_ov_val.p, c->sell_ov_col.p, pk_off.p, pk_doff.p, pk_goff.p,
                     c->sell_pk_val.p, reinterpret_cast<uint8_t*>(c->sell_pk_idx.p), c->sell_pk_lp.p, c->sell_pk_dict.p,
                     reinterpret_cast<PkHdr*>(c->sell_pk_hdr.p), c->sell_cval.p);
  NGPB_HIP(hipGetLastError());
  {
    struct BArr {
      uint8_t* p;
    };
    BArr fd{tmp_buf<uint8_t>(c, kTmpForms, (size_t)ns)}, fp{tmp_buf<uint8_t>(c, kTmpForms + 1, (size_t)ns)},
        fo{tmp_buf<uint8_t>(c, kTmpForms + 2, (size_t)ns)};
    hipLaunchKernelGGL(k_sell_forms, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, ns, c->slice_width.p, fd.p,
                       fp.p, fo.p);
    c->sell_list_delta.resize((size_t)ns);
    c->sell_list_packed.resize((size_t)ns);
    c->sell_list_offset.resize((size_t)ns);
    c->n_list_delta = compact_flags(c, fd.p, ns, c->sell_list_delta.p);
    c->n_list_packed = compact_flags(c, fp.p, ns, c->sell_list_packed.p);
    c->n_list_offset = compact_flags(c, fo.p, ns, c->sell_list_offset.p);
    const int64_t npairs = (c->n_list_delta + 1) / 2;
    NGPB_REQUIRE(nbases < INT32_MAX, NGPB_E_ARG, "operator too large for 32-bit slot pointers");
    c->sell_dpair.resize((size_t)std::max<int64_t>(npairs, 1) * (sizeof(DPair) / sizeof(int32_t)));
    if (npairs > 0)
      hipLaunchKernelGGL(k_dpair_fill, dim3(blocks_for(npairs, 256)), dim3(256), 0, c->stream, npairs,
                         c->sell_list_delta.p, c->n_list_delta, c->slice_width.p, c->slice_base_ptr.p,
                         reinterpret_cast<DPair*>(c->sell_dpair.p));
    c->sell_pk_hdr_list.resize((size_t)std::max<int64_t>(c->n_list_packed, 1) * (sizeof(PkHdr) / sizeof(int64_t)));
    if (c->n_list_packed > 0)
      hipLaunchKernelGGL(k_pk_hdr_list, dim3(blocks_for(c->n_list_packed, 256)), dim3(256), 0, c->stream,
                         c->n_list_packed, c->sell_list_packed.p, reinterpret_cast<const PkHdr*>(c->sell_pk_hdr.p),
                         reinterpret_cast<PkHdr*>(c->sell_pk_hdr_list.p));
    NGPB_HIP(hipGetLastError());
  }
  NGPB_HIP(hipStreamSynchronize(c->stream));
  c->have_compact = true;
}

// The FP32 SpMV input vector of the context (the scaled direction p of the mixed
// solve): x_0 .. x_n-1 with kGuard zero entries on either side, which the lean SpMV
// gathers from in place of clamped indices. Every kernel writes only x_0 .. x_n-1.
float* spmv_input(ngpb_ctx* c) {
  c->ps.reserve((size_t)(c->n + 2 * kGuard));
  hipLaunchKernelGGL(k_zero_guard, dim3(1), dim3(2 * kGuard), 0, c->stream, c->ps.p, c->n);
  NGPB_HIP(hipGetLastError());
  return c->ps.p + kGuard;
}

double spmv_compact_bytes(const ngpb_ctx* c) {
  // delta-slice entries (4: value), offset-slice entries (4 value + 2 offset), slot
  // bases/deltas (4), constant-slice slot values (4), 32-bit columns of wide slices, overflow entries (4 + 4),
  // packed values (4), packed index words (4, four slots of a row each), packed rows
  // (permutation + count, 1 + 1), dictionary entries (4), packed-slice headers (32),
  // slice metadata (width 4 + three 8-byte offsets), x read once and y written
  const double off_entries = (double)(c->sell_entries - c->sell_delta_entries - c->sell_const_slots * kW);
  return (double)c->sell_delta_entries * 4.0 + off_entries * 6.0 + (double)c->sell_bases * 4.0 +
         (double)c->sell_const_slots * 4.0 +
         (double)c->sell_wide_entries * 4.0 + (double)c->sell_ov_entries * 8.0 + (double)c->sell_pk_entries * 4.0 +
         (double)c->sell_pk_words * 4.0 + (double)c->sell_pk_slices * (kW * 2.0 + 32.0) +
         (double)c->sell_pk_dict_entries * 4.0 +
         (double)c->n_slices * 28.0 + (double)c->n * (4.0 + 4.0);
}

void spmv_compact(ngpb_ctx* c, const double* x, double* y) {
  if (!c->have_compact) build_compact(c);
  const int64_t n = c->n;
  float* pin = spmv_input(c);
  c->qs.reserve((size_t)n);
  const int32_t* perm = c->sell_banded ? c->row_perm.p : nullptr;
  hipLaunchKernelGGL(k_to_scaled, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, x, c->dscale.p, pin, perm);
  launch_sell(c, false, pin, c->qs.p, nullptr);
  hipLaunchKernelGGL(k_from_scaled, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, c->qs.p, c->dscale.p, y,
                     perm);
  NGPB_HIP(hipGetLastError());
}

double time_spmv_compact(ngpb_ctx* c, int reps) {
  if (!c->have_compact) build_compact(c);
  float* pin = spmv_input(c);
  c->qs.reserve((size_t)c->n);
  NGPB_HIP(hipMemsetAsync(pin, 0, c->n * sizeof(float), c->stream));
  hipEvent_t e0, e1;
  NGPB_HIP(hipEventCreate(&e0));
  NGPB_HIP(hipEventCreate(&e1));
  launch_sell(c, false, pin, c->qs.p, nullptr);
  NGPB_HIP(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps; ++i) launch_sell(c, false, pin, c->qs.p, nullptr);
  NGPB_HIP(hipEventRecord(e1, c->stream));
  NGPB_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  NGPB_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  NGPB_HIP(hipGetLastError());
  return (double)ms / std::max(reps, 1);
}

void solve_mixed(ngpb_ctx* c, const ngpb_solve_opts& o, ngpb_solve_stats& st) {
  const int64_t n = c->n;
  check_operator(c);
  if (!c->have_compact) build_compact(c);
  c->x.resize((size_t)n);
  c->q.reserve((size_t)n);
  c->xs.reserve((size_t)n);
  c->rs.reserve((size_t)n);
  float* const pv = spmv_input(c);
  const int32_t* const perm = c->sell_banded ? c->row_perm.p : nullptr;  // FP32 row -> node
  c->qs.reserve((size_t)n);
  const unsigned gv = vec_grid(n);
  const unsigned gs = (unsigned)sell_partials(c);
  ensure_reduce(c, std::max<int64_t>(gs, 2048) + 2 * (int64_t)gv);
  double* part = c->red.partials.p;
  double* ppq = part;                                  // SpMV (p, q) partials
  double* pup = part + std::max<int64_t>(gs, 2048);    // update (rhat, rhat) partials
  int64_t launch = 0;                                  // iteration launches (parity)
  KrylovScalars* ks = c->ks.p;
  KrylovScalars* h = c->ks_host;
  const double drop = (o.drop > 0.0 && o.drop < 1.0) ? o.drop : 1e-3;
  const int chunk = std::max(1, o.check_every);

  hipEvent_t e0, e1;
  NGPB_HIP(hipEventCreate(&e0));
  NGPB_HIP(hipEventCreate(&e1));
  NGPB_HIP(hipEventRecord(e0, c->stream));
  hipLaunchKernelGGL(k_mx_resid, dim3(gv), dim3(kBS), 0, c->stream, n, c->rhs.p, (const double*)nullptr, c->dscale.p,
                     c->rs.p, c->x.p, c->xs.p, pv, 1, part, perm);
  hipLaunchKernelGGL(k_mx_fin_resid, dim3(1), dim3(kFin), 0, c->stream, ks, part, (int)gv, 1);
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipMemcpyAsync(h, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c->stream));
  NGPB_HIP(hipStreamSynchronize(c->stream));
  const double bb = h->bb;
  const double target = o.tol * o.tol * bb;
  int status = 0, iters = 0, repl = 0;
  double true_rr = bb;
  if (bb > 0.0) {
    status = NGPB_E_NOCONV;
    double prev_true = bb;
    int stagnant = 0;
    while (iters < o.maxiter) {
      // cycle: stop at drop^2 of the current scaled residual or at the scaled image
      // of the FP64 target (ratio of the two norms taken at the last replacement)
      const double ss = h->aux[1], tt = h->aux[0];
      const double ratio = tt > 0.0 ? ss / tt : 1.0;
      h->stop_rr = std::max(drop * drop * ss, 0.25 * target * ratio);
      h->maxiter = o.maxiter;
      h->iter = iters;
      h->done = 0;
      h->rho2[launch & 1] = h->rho;
      h->iter2[launch & 1] = iters;
      NGPB_HIP(hipMemcpyAsync(ks, h, sizeof(KrylovScalars), hipMemcpyHostToDevice, c->stream));
      const int64_t cycle_start = launch;
      hipLaunchKernelGGL(k_mx_dir, dim3(gv), dim3(kBS), 0, c->stream, n, c->rs.p, pv, ks);
      while (true) {
        for (int k = 0; k < chunk; ++k, ++launch) {
          const int par = (int)(launch & 1);
          prof_begin(c, 1, launch);
          launch_sell(c, true, pv, c->qs.p, ppq);
          prof_end(c, 1);
          hipLaunchKernelGGL(k_mx_update, dim3(gv), dim3(kBS), 0, c->stream, n, c->rs.p, c->qs.p, ks, ppq, (int)gs, pup,
                             par);
          hipLaunchKernelGGL(k_mx_dir_x, dim3(gv), dim3(kBS), 0, c->stream, n, c->rs.p, pv, c->xs.p, ks, pup,
                             (int)gv, par);
        }
        NGPB_HIP(hipGetLastError());
        NGPB_HIP(hipMemcpyAsync(h, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c->stream));
        NGPB_HIP(hipStreamSynchronize(c->stream));
        prof_collect(c, 1, cycle_start + (h->iter - iters));
        if (h->done) break;
      }
      iters = h->iter;
      // FP64 residual replacement
      hipLaunchKernelGGL(k_mx_accum, dim3(gv), dim3(kBS), 0, c->stream, n, c->dscale.p, c->x.p, c->xs.p, pv, ks, perm);
      prof_begin(c, 0, -1);
      spmv_fp64(c, c->x.p, c->q.p);
      prof_end(c, 0);
      hipLaunchKernelGGL(k_mx_resid, dim3(gv), dim3(kBS), 0, c->stream, n, c->rhs.p, c->q.p, c->dscale.p, c->rs.p, c->x.p,
                         c->xs.p, pv, 0, part, perm);
      hipLaunchKernelGGL(k_mx_fin_resid, dim3(1), dim3(kFin), 0, c->stream, ks, part, (int)gv, 0);
      NGPB_HIP(hipGetLastError());
      NGPB_HIP(hipMemcpyAsync(h, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c->stream));
      NGPB_HIP(hipStreamSynchronize(c->stream));
      prof_collect(c, 0, -1);
      ++repl;
      true_rr = h->aux[0];
      if (true_rr <= target) {
        status = 0;
        break;
      }
      // the FP32 inner solve no longer reduces the FP64 residual
      if (true_rr > 0.81 * prev_true) {
        if (++stagnant >= 3) break;
      } else {
        stagnant = 0;
      }
      prev_true = std::min(prev_true, true_rr);
    }
  }
  NGPB_HIP(hipEventRecord(e1, c->stream));
  NGPB_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  NGPB_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  st.status = status;
  st.iterations = iters;
  st.replacements = repl;
  st.relres = bb > 0 ? std::sqrt(true_rr / bb) : 0.0;
  st.ms_solve = ms;
  // SpMV + update (rhat, q read; rhat written) + direction (rhat, p, xhat read; p, xhat written)
  st.bytes_per_iter = spmv_compact_bytes(c) + (double)n * 4.0 * (3.0 + 5.0);
  c->have_phi = true;
  c->phi_len = c->n;
}

}  // namespace ngpb
