// The LIS option strings beyond "-i cg -p jacobi": CG, CGS and BiCGSTAB with no,
// Jacobi or SSOR preconditioning, in FP64 on the device. The reference's default is
// "-p ssor -ssor_omega 0.51 -i cgs -tol 1.e-6" (src/pb_class.cpp:970,
// data/options.prm:174), handed to LIS through lis_solver_set_option
// (src/pb_class.cpp:2294-2315).
//
// Algorithms (LIS is not vendored with the reference; these are the published forms):
//   SSOR (Saad, Iterative Methods for Sparse Linear Systems, sec. 10.2): with
//   A = L + D + U, M = (D + w L) D^-1 (D + w U). M^-1 r is a forward sweep
//   y_i = (r_i - w sum_{j<i} a_ij y_j) / d_i, then a backward sweep
//   z_i = y_i - w (sum_{j>i} a_ij z_j) / d_i, each row's entries in ascending column
//   order with the arithmetic t = t - (w a_ij) y_j and t = t + a_ij z_j, so the result
//   is the oracle's bit for bit.
//   CGS and BiCGSTAB as in Barrett et al., Templates (1994), sec. 2.3.7-2.3.8, x0 = 0,
//   shadow residual r~ = b, stop when ||r||_2 <= tol ||b||_2 (BiCGSTAB also tests s).
//
// The sweeps are level-scheduled: a row's forward level is one more than the largest
// level among its lower neighbours (backward: upper neighbours), so the rows of one
// level are independent and one launch per level processes them, one thread per row.
// The levels depend only on the sparsity pattern and are built once per pattern on
// the host from the CSR (rows sorted by level, ascending row within a level).
//
// Reductions follow the FP64 CG: per-block partials, summed in a fixed order by a
// one-block finalize kernel that also updates the scalars and the stop flag; every
// kernel launched after the stop flag returns at once, and the host polls the flag
// every check_every iterations.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ngpb_internal.h"

namespace ngpb {
namespace {

constexpr int kBS = 256;
constexpr int kFin = 1024;

// KrylovScalars fields used here: rho, pq (denominator), alpha, beta, rr, stop_rr, bb,
// aux[0] omega, aux[1] rho of the previous iteration, aux[2] (s, s), aux[3] 1 once
// BiCGSTAB's s met the tolerance; iter, maxiter, done, status.
enum Fin : int { kFinRho = 0, kFinAlpha = 1, kFinResid = 2, kFinBicgS = 3, kFinOmega = 4, kFinBicgResid = 5,
                 kFinCgBeta = 6, kFinCgRho0 = 7 };

__global__ void __launch_bounds__(kBS) k_lis_init(int64_t n, const double* __restrict__ b, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ rt,
                                                  double* __restrict__ p, double* __restrict__ q,
                                                  double* __restrict__ v, double* __restrict__ part) {
  __shared__ double s_red[kBS / 64];
  double bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    rt[i] = bi;
    p[i] = 0.0;
    q[i] = 0.0;
    v[i] = 0.0;
    bb += bi * bi;
  }
  bb = dev::block_sum<kBS>(bb, s_red);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = bb;
}

__global__ void __launch_bounds__(kFin) k_lis_fin_init(KrylovScalars* ks, const double* __restrict__ part, int np,
                                                       double tol2, int maxiter) {
  __shared__ double s_red[kFin / 64];
  const double BB = dev::sum_partials<kFin>(part, np, 2, s_red);
  if (threadIdx.x == 0) {
    ks->bb = BB;
    ks->rr = BB;
    ks->stop_rr = tol2 * BB;
    ks->rho = 0.0;
    ks->pq = 0.0;
    ks->alpha = 1.0;
    ks->beta = 0.0;
    ks->aux[0] = 1.0;
    ks->aux[1] = 1.0;
    ks->aux[2] = 0.0;
    ks->aux[3] = 0.0;
    ks->iter = 0;
    ks->maxiter = maxiter;
    ks->status = 0;
    ks->done = (BB == 0.0) ? 1 : 0;
  }
}

// per-block partials of (a, b) and, with TWO, (c, d); stride 2
template <bool TWO>
__global__ void __launch_bounds__(kBS) k_dots(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                              const double* __restrict__ c, const double* __restrict__ d,
                                              const KrylovScalars* ks, double* __restrict__ part) {
  __shared__ double s_red[kBS / 64];
  if (ks->done) return;
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    s1 += a[i] * b[i];
    if (TWO) s2 += c[i] * d[i];
  }
  s1 = dev::block_sum<kBS>(s1, s_red);
  if (TWO) s2 = dev::block_sum<kBS>(s2, s_red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    if (TWO) part[2 * blockIdx.x + 1] = s2;
  }
}

// One block: sums the partials in a fixed order and advances the scalar state.
__global__ void __launch_bounds__(kFin) k_lis_fin(KrylovScalars* ks, const double* __restrict__ part, int np,
                                                  int stride, int what, int solver) {
  __shared__ double s_red[kFin / 64];
  if (ks->done) return;
  const double S0 = dev::sum_partials<kFin>(part, np, stride, s_red);
  const double S1 = (what == kFinOmega) ? dev::sum_partials<kFin>(part + 1, np, stride, s_red) : 0.0;
  if (threadIdx.x != 0) return;
  const bool sconv = ks->aux[3] != 0.0;
  switch (what) {
    case kFinRho:  // rho = (r~, r) at the start of a CGS / BiCGSTAB iteration
      if (S0 == 0.0) {
        ks->done = 1;
        ks->status = NGPB_E_NOCONV;
        break;
      }
      ks->rho = S0;
      if (ks->iter == 0) ks->beta = 0.0;
      else if (solver == NGPB_SOLVER_CGS) ks->beta = S0 / ks->aux[1];
      else ks->beta = (S0 / ks->aux[1]) * (ks->alpha / ks->aux[0]);
      break;
    case kFinAlpha:  // alpha = rho / denominator
      ks->pq = S0;
      if (S0 == 0.0) {
        ks->done = 1;
        ks->status = NGPB_E_NOCONV;
        break;
      }
      ks->alpha = ks->rho / S0;
      break;
    case kFinResid:  // CG / CGS: stop test on the updated residual
      ks->rr = S0;
      ks->iter += 1;
      if (S0 <= ks->stop_rr) {
        ks->done = 1;
        ks->status = 0;
      } else if (ks->iter >= ks->maxiter) {
        ks->done = 1;
        ks->status = NGPB_E_NOCONV;
      } else {
        ks->aux[1] = ks->rho;
      }
      break;
    case kFinBicgS:  // BiCGSTAB: s = r - alpha v met the tolerance?
      ks->aux[2] = S0;
      ks->iter += 1;
      if (S0 <= ks->stop_rr) ks->aux[3] = 1.0;
      break;
    case kFinOmega:  // omega = (t, s) / (t, t)
      if (sconv) break;
      if (S0 == 0.0 || S1 == 0.0) {
        ks->done = 1;
        ks->status = NGPB_E_NOCONV;
        break;
      }
      ks->aux[0] = S0 / S1;
      break;
    case kFinBicgResid:
      if (sconv) {
        ks->rr = ks->aux[2];
        ks->done = 1;
        ks->status = 0;
        break;
      }
      ks->rr = S0;
      if (S0 <= ks->stop_rr) {
        ks->done = 1;
        ks->status = 0;
      } else if (ks->iter >= ks->maxiter) {
        ks->done = 1;
        ks->status = NGPB_E_NOCONV;
      } else {
        ks->aux[1] = ks->rho;
      }
      break;
    case kFinCgBeta: {  // CG: rho' = (r, z); beta = rho' / rho
      ks->beta = S0 / ks->rho;
      ks->rho = S0;
      break;
    }
    case kFinCgRho0:
      ks->rho = S0;
      break;
  }
}

// CGS: u = r + beta q; p = u + beta (q + beta p)
__global__ void __launch_bounds__(kBS) k_cgs_up(int64_t n, const double* __restrict__ r, const double* __restrict__ q,
                                                double* __restrict__ u, double* __restrict__ p, const KrylovScalars* ks) {
  if (ks->done) return;
  const double beta = ks->beta;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double ui = r[i] + beta * q[i];
    u[i] = ui;
    p[i] = ui + beta * (q[i] + beta * p[i]);
  }
}

// CGS: q = u - alpha v; t = u + q
__global__ void __launch_bounds__(kBS) k_cgs_q(int64_t n, const double* __restrict__ u, const double* __restrict__ v,
                                               double* __restrict__ q, double* __restrict__ t, const KrylovScalars* ks) {
  if (ks->done) return;
  const double alpha = ks->alpha;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double qi = u[i] - alpha * v[i];
    q[i] = qi;
    t[i] = u[i] + qi;
  }
}

// CG and CGS: x += alpha d; r -= alpha w; per block (r, r)
__global__ void __launch_bounds__(kBS) k_xr_update(int64_t n, double* __restrict__ x, const double* __restrict__ d,
                                                   double* __restrict__ r, const double* __restrict__ w,
                                                   const KrylovScalars* ks, double* __restrict__ part) {
  __shared__ double s_red[kBS / 64];
  if (ks->done) return;
  const double alpha = ks->alpha;
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    x[i] += alpha * d[i];
    const double ri = r[i] - alpha * w[i];
    r[i] = ri;
    rr += ri * ri;
  }
  rr = dev::block_sum<kBS>(rr, s_red);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = rr;
}

// BiCGSTAB: p = r + beta (p - omega v)
__global__ void __launch_bounds__(kBS) k_bicg_p(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                                                double* __restrict__ p, const KrylovScalars* ks) {
  if (ks->done) return;
  const double beta = ks->beta, omega = ks->aux[0];
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS)
    p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

// BiCGSTAB: s = r - alpha v; per block (s, s)
__global__ void __launch_bounds__(kBS) k_bicg_s(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                                                double* __restrict__ s, const KrylovScalars* ks,
                                                double* __restrict__ part) {
  __shared__ double s_red[kBS / 64];
  if (ks->done) return;
  const double alpha = ks->alpha;
  double ss = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double si = r[i] - alpha * v[i];
    s[i] = si;
    ss += si * si;
  }
  ss = dev::block_sum<kBS>(ss, s_red);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = ss;
}

// BiCGSTAB: x += alpha ph + omega sh; r = s - omega t; per block (r, r). Once s met
// the tolerance only x += alpha ph is applied.
__global__ void __launch_bounds__(kBS) k_bicg_x(int64_t n, double* __restrict__ x, const double* __restrict__ ph,
                                                const double* __restrict__ sh, double* __restrict__ r,
                                                const double* __restrict__ s, const double* __restrict__ t,
                                                const KrylovScalars* ks, double* __restrict__ part) {
  __shared__ double s_red[kBS / 64];
  if (ks->done) return;
  const double alpha = ks->alpha, omega = ks->aux[0];
  const bool sconv = ks->aux[3] != 0.0;
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    if (sconv) {
      x[i] += alpha * ph[i];
    } else {
      x[i] += alpha * ph[i] + omega * sh[i];
      const double ri = s[i] - omega * t[i];
      r[i] = ri;
      rr += ri * ri;
    }
  }
  rr = dev::block_sum<kBS>(rr, s_red);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = rr;
}

// CG: p = z + beta p (beta = 0 copies z)
__global__ void __launch_bounds__(kBS) k_cg_p(int64_t n, const double* __restrict__ z, double* __restrict__ p,
                                              const KrylovScalars* ks, int first) {
  if (ks->done) return;
  const double beta = first ? 0.0 : ks->beta;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS)
    p[i] = first ? z[i] : z[i] + beta * p[i];
}

__global__ void __launch_bounds__(kBS) k_jacobi_apply(int64_t n, const double* __restrict__ dinv,
                                                      const double* __restrict__ in, double* __restrict__ out,
                                                      const KrylovScalars* ks) {
  if (ks->done) return;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS)
    out[i] = dinv[i] * in[i];
}

// One level of the SSOR forward sweep: z_i = (r_i - w sum_{j<i} a_ij z_j) d_i^-1.
__global__ void __launch_bounds__(kBS) k_ssor_fwd(const int32_t* __restrict__ rows, int cnt,
                                                  const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                  const double* __restrict__ val, const int32_t* __restrict__ diag,
                                                  const double* __restrict__ dinv, double w,
                                                  const double* __restrict__ r, double* z, const KrylovScalars* ks) {
  if (ks->done) return;
  const int k = blockIdx.x * kBS + threadIdx.x;
  if (k >= cnt) return;
  const int32_t i = rows[k];
  double t = r[i];
  for (int32_t e = rowptr[i], de = diag[i]; e < de; ++e) t = t - (w * val[e]) * z[col[e]];
  z[i] = t * dinv[i];
}

// One level of the backward sweep: z_i = z_i - (w sum_{j>i} a_ij z_j) d_i^-1.
__global__ void __launch_bounds__(kBS) k_ssor_bwd(const int32_t* __restrict__ rows, int cnt,
                                                  const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                  const double* __restrict__ val, const int32_t* __restrict__ diag,
                                                  const double* __restrict__ dinv, double w, double* z,
                                                  const KrylovScalars* ks) {
  if (ks->done) return;
  const int k = blockIdx.x * kBS + threadIdx.x;
  if (k >= cnt) return;
  const int32_t i = rows[k];
  double t = 0.0;
  for (int32_t e = diag[i] + 1, ee = rowptr[i + 1]; e < ee; ++e) t = t + val[e] * z[col[e]];
  z[i] = z[i] - (w * t) * dinv[i];
}

// rows sorted by level (counting sort, ascending row within a level)
void sort_by_level(const std::vector<int32_t>& lev, int32_t nlev, std::vector<int32_t>& rows,
                   std::vector<int64_t>& ptr) {
  ptr.assign((size_t)nlev + 1, 0);
  for (int32_t l : lev) ++ptr[(size_t)l + 1];
  for (int32_t l = 0; l < nlev; ++l) ptr[(size_t)l + 1] += ptr[(size_t)l];
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  rows.resize(lev.size());
  for (size_t i = 0; i < lev.size(); ++i) rows[(size_t)fill[(size_t)lev[i]]++] = (int32_t)i;
}

}  // namespace

// Level sets of the SSOR sweeps and the diagonal positions, from the CSR pattern.
void build_ssor_levels(ngpb_ctx* c) {
  const int64_t n = c->n, nnz = c->nnz;
  std::vector<int32_t> rp((size_t)n + 1), col((size_t)nnz);
  c->rowptr.download(rp.data(), rp.size(), c->stream);
  c->col.download(col.data(), col.size(), c->stream);
  NGPB_HIP(hipStreamSynchronize(c->stream));
  std::vector<int32_t> diag((size_t)n), lf((size_t)n, 0), lb((size_t)n, 0);
  int32_t nf = 0, nb = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* b = col.data() + rp[(size_t)i];
    const int32_t* e = col.data() + rp[(size_t)i + 1];
    const int32_t* d = std::lower_bound(b, e, (int32_t)i);
    NGPB_REQUIRE(d != e && *d == i, NGPB_E_NUMERIC, "operator has a row without a diagonal entry");
    diag[(size_t)i] = (int32_t)(d - col.data());
    int32_t l = 0;
    for (const int32_t* k = b; k < d; ++k) l = std::max(l, lf[(size_t)*k] + 1);
    lf[(size_t)i] = l;
    nf = std::max(nf, l + 1);
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    int32_t l = 0;
    for (int32_t k = diag[(size_t)i] + 1; k < rp[(size_t)i + 1]; ++k) l = std::max(l, lb[(size_t)col[(size_t)k]] + 1);
    lb[(size_t)i] = l;
    nb = std::max(nb, l + 1);
  }
  std::vector<int32_t> rows;
  sort_by_level(lf, nf, rows, c->lev_ptr_f);
  c->lev_rows_f.upload(rows.data(), rows.size(), c->stream);
  sort_by_level(lb, nb, rows, c->lev_ptr_b);
  c->lev_rows_b.upload(rows.data(), rows.size(), c->stream);
  c->diag_pos.upload(diag.data(), diag.size(), c->stream);
  NGPB_HIP(hipStreamSynchronize(c->stream));
  c->have_levels = true;
}

namespace {

// z = M^-1 in; returns the vector that holds the result (`in` itself without a
// preconditioner). c->dinv must hold the inverse diagonal.
const double* psolve(ngpb_ctx* c, int precon, double w, const double* in, double* out) {
  const int64_t n = c->n;
  const KrylovScalars* ks = c->ks.p;
  if (precon == NGPB_PRECON_NONE) return in;
  if (precon == NGPB_PRECON_JACOBI) {
    hipLaunchKernelGGL(k_jacobi_apply, dim3(vec_grid(n)), dim3(kBS), 0, c->stream, n, c->dinv.p, in, out, ks);
    return out;
  }
  const size_t nf = c->lev_ptr_f.size() - 1, nb = c->lev_ptr_b.size() - 1;
  for (size_t l = 0; l < nf; ++l) {
    const int64_t o = c->lev_ptr_f[l], cnt = c->lev_ptr_f[l + 1] - o;
    hipLaunchKernelGGL(k_ssor_fwd, dim3((unsigned)((cnt + kBS - 1) / kBS)), dim3(kBS), 0, c->stream,
                       c->lev_rows_f.p + o, (int)cnt, c->rowptr.p, c->col.p, c->val.p, c->diag_pos.p, c->dinv.p, w,
                       in, out, ks);
  }
  for (size_t l = 0; l < nb; ++l) {
    const int64_t o = c->lev_ptr_b[l], cnt = c->lev_ptr_b[l + 1] - o;
    hipLaunchKernelGGL(k_ssor_bwd, dim3((unsigned)((cnt + kBS - 1) / kBS)), dim3(kBS), 0, c->stream,
                       c->lev_rows_b.p + o, (int)cnt, c->rowptr.p, c->col.p, c->val.p, c->diag_pos.p, c->dinv.p, w,
                       out, ks);
  }
  return out;
}

void prepare_precon(ngpb_ctx* c, int precon) {
  check_operator(c);  // inverse diagonal, NaN/Inf and diagonal checks
  if (precon == NGPB_PRECON_SSOR && !c->have_levels) build_ssor_levels(c);
}

}  // namespace

double ssor_omega(const ngpb_solve_opts& o) { return o.omega > 0.0 ? o.omega : 1.0; }

void check_lis_options(const ngpb_solve_opts& o) {
  NGPB_REQUIRE(o.solver == NGPB_SOLVER_CG || o.solver == NGPB_SOLVER_CGS || o.solver == NGPB_SOLVER_BICGSTAB,
               NGPB_E_ARG, "solver must be CG, CGS or BiCGSTAB");
  NGPB_REQUIRE(o.precon == NGPB_PRECON_JACOBI || o.precon == NGPB_PRECON_NONE || o.precon == NGPB_PRECON_SSOR,
               NGPB_E_ARG, "preconditioner must be Jacobi, none or SSOR");
  NGPB_REQUIRE(o.precon != NGPB_PRECON_SSOR || (ssor_omega(o) > 0.0 && ssor_omega(o) < 2.0), NGPB_E_ARG,
               "SSOR relaxation must lie in (0, 2)");
}

void precond_apply(ngpb_ctx* c, int precon, double omega, const double* r, double* z) {
  prepare_precon(c, precon);
  ensure_reduce(c, 2048);
  KrylovScalars zero;
  std::memset(&zero, 0, sizeof(zero));
  NGPB_HIP(hipMemcpyAsync(c->ks.p, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
  const double* res = psolve(c, precon, omega, r, z);
  if (res != z) NGPB_HIP(hipMemcpyAsync(z, res, (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  NGPB_HIP(hipGetLastError());
  NGPB_HIP(hipStreamSynchronize(c->stream));
}

void solve_lis(ngpb_ctx* c, const ngpb_solve_opts& o, ngpb_solve_stats& st) {
  check_lis_options(o);
  const int64_t n = c->n;
  const double w = ssor_omega(o);
  prepare_precon(c, o.precon);
  c->x.resize((size_t)n);
  c->r.resize((size_t)n);
  c->z.reserve((size_t)n);
  c->p.reserve((size_t)n);
  c->q.reserve((size_t)n);
  for (auto* b : {&c->lis_rt, &c->lis_u, &c->lis_v, &c->lis_ph, &c->lis_uh, &c->lis_t}) b->reserve((size_t)n);
  const unsigned gv = vec_grid(n);
  const int64_t gs_region = std::max<int64_t>(spmv_fp64_grid(c), 2048);
  ensure_reduce(c, gs_region + 2 * (int64_t)gv);
  double* spart = c->red.partials.p;             // SpMV (x, y) partials, stride 1
  double* dpart = c->red.partials.p + gs_region;  // vector partials, stride 2
  KrylovScalars* ks = c->ks.p;
  double *x = c->x.p, *r = c->r.p, *p = c->p.p, *q = c->q.p, *z = c->z.p;
  double *rt = c->lis_rt.p, *u = c->lis_u.p, *v = c->lis_v.p, *ph = c->lis_ph.p, *uh = c->lis_uh.p, *t = c->lis_t.p;
  const dim3 G(gv), B(kBS), F(1), BF(kFin);
  const int np = (int)gv, nsp = (int)spmv_fp64_grid(c);
  auto fin = [&](const double* part, int cnt, int stride, int what) {
    hipLaunchKernelGGL(k_lis_fin, F, BF, 0, c->stream, ks, part, cnt, stride, what, o.solver);
  };
  hipEvent_t e0, e1;
  NGPB_HIP(hipEventCreate(&e0));
  NGPB_HIP(hipEventCreate(&e1));
  NGPB_HIP(hipEventRecord(e0, c->stream));
  hipLaunchKernelGGL(k_lis_init, G, B, 0, c->stream, n, c->rhs.p, x, r, rt, p, q, v, dpart);
  hipLaunchKernelGGL(k_lis_fin_init, F, BF, 0, c->stream, ks, dpart, np, o.tol * o.tol, o.maxiter);
  if (o.solver == NGPB_SOLVER_CG) {
    const double* zz = psolve(c, o.precon, w, r, z);
    hipLaunchKernelGGL(k_dots<false>, G, B, 0, c->stream, n, r, zz, nullptr, nullptr, ks, dpart);
    fin(dpart, np, 2, kFinCgRho0);
    hipLaunchKernelGGL(k_cg_p, G, B, 0, c->stream, n, zz, p, ks, 1);
  }
  NGPB_HIP(hipGetLastError());
  const int chunk = std::max(1, o.check_every);
  while (true) {
    for (int k = 0; k < chunk; ++k) {
      if (o.solver == NGPB_SOLVER_CG) {
        spmv_fp64_dot(c, p, q, spart);  // q = A p with (p, q) partials
        fin(spart, nsp, 1, kFinAlpha);
        hipLaunchKernelGGL(k_xr_update, G, B, 0, c->stream, n, x, p, r, q, ks, dpart);
        fin(dpart, np, 2, kFinResid);
        const double* zz = psolve(c, o.precon, w, r, z);
        hipLaunchKernelGGL(k_dots<false>, G, B, 0, c->stream, n, r, zz, nullptr, nullptr, ks, dpart);
        fin(dpart, np, 2, kFinCgBeta);
        hipLaunchKernelGGL(k_cg_p, G, B, 0, c->stream, n, zz, p, ks, 0);
      } else if (o.solver == NGPB_SOLVER_CGS) {
        hipLaunchKernelGGL(k_dots<false>, G, B, 0, c->stream, n, rt, r, nullptr, nullptr, ks, dpart);
        fin(dpart, np, 2, kFinRho);
        hipLaunchKernelGGL(k_cgs_up, G, B, 0, c->stream, n, r, q, u, p, ks);
        const double* pp = psolve(c, o.precon, w, p, ph);
        spmv_fp64_dot(c, pp, v, spart);
        hipLaunchKernelGGL(k_dots<false>, G, B, 0, c->stream, n, rt, v, nullptr, nullptr, ks, dpart);
        fin(dpart, np, 2, kFinAlpha);
        hipLaunchKernelGGL(k_cgs_q, G, B, 0, c->stream, n, u, v, q, t, ks);
        const double* uu = psolve(c, o.precon, w, t, uh);
        spmv_fp64_dot(c, uu, ph, spart);  // A uh into ph, free once v = A ph was formed
        hipLaunchKernelGGL(k_xr_update, G, B, 0, c->stream, n, x, uu, r, ph, ks, dpart);
        fin(dpart, np, 2, kFinResid);
      } else {  // BiCGSTAB; s lives in u, its preconditioned copy in uh
        hipLaunchKernelGGL(k_dots<false>, G, B, 0, c->stream, n, rt, r, nullptr, nullptr, ks, dpart);
        fin(dpart, np, 2, kFinRho);
        hipLaunchKernelGGL(k_bicg_p, G, B, 0, c->stream, n, r, v, p, ks);
        const double* pp = psolve(c, o.precon, w, p, ph);
        spmv_fp64_dot(c, pp, v, spart);
        hipLaunchKernelGGL(k_dots<false>, G, B, 0, c->stream, n, rt, v, nullptr, nullptr, ks, dpart);
        fin(dpart, np, 2, kFinAlpha);
        hipLaunchKernelGGL(k_bicg_s, G, B, 0, c->stream, n, r, v, u, ks, dpart);
        fin(dpart, np, 2, kFinBicgS);
        const double* ss = psolve(c, o.precon, w, u, uh);
        spmv_fp64_dot(c, ss, t, spart);
        hipLaunchKernelGGL(k_dots<true>, G, B, 0, c->stream, n, t, u, t, t, ks, dpart);
        fin(dpart, np, 2, kFinOmega);
        hipLaunchKernelGGL(k_bicg_x, G, B, 0, c->stream, n, x, pp, ss, r, u, t, ks, dpart);
        fin(dpart, np, 2, kFinBicgResid);
      }
    }
    NGPB_HIP(hipGetLastError());
    NGPB_HIP(hipMemcpyAsync(c->ks_host, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c->stream));
    NGPB_HIP(hipStreamSynchronize(c->stream));
    if (c->ks_host->done) break;
  }
  NGPB_HIP(hipEventRecord(e1, c->stream));
  NGPB_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  NGPB_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const KrylovScalars& h = *c->ks_host;
  st.status = h.status;
  st.iterations = h.iter;
  st.replacements = 0;
  st.relres = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
  st.ms_solve = ms;
  // matrix-vector products and preconditioner applications of one iteration, plus
  // the vector streams (8 bytes each)
  const double spmv = spmv_fp64_bytes(c);
  const double pre = o.precon == NGPB_PRECON_SSOR ? spmv + (double)n * 8.0 * 3.0
                     : o.precon == NGPB_PRECON_JACOBI ? (double)n * 24.0 : 0.0;
  const double vec = o.solver == NGPB_SOLVER_CG ? 14.0 : o.solver == NGPB_SOLVER_CGS ? 22.0 : 24.0;
  const double nmv = o.solver == NGPB_SOLVER_CG ? 1.0 : 2.0;
  st.bytes_per_iter = nmv * (spmv + pre) + (double)n * 8.0 * vec;
  c->have_phi = true;
  c->phi_len = c->n;
}

}  // namespace ngpb
