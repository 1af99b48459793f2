// CCSD ground-state multipliers: the right-hand side eta (CC_ETA,
// cc/cc_fop.F:2955), the left equations lambda A + eta = 0 that CC_FOP
// solves for the list 'L0' (cc/cc_fop.F:274) with CCEQ_SOL
// (cc/cceq_solv.F:20) from the left, and the unrelaxed one-electron density
// of CC_D1AO (cc/cc_d1ao.F:20), here in the MO basis.
//
// Conventions are those of the left transformation (lhtr.hip): singles
// [i][a], doubles packed over ai <= bj, and the Lagrangian
//   L(t, lambda; h) = E(t; h) + sum_p lambda_p Omega_p(t; h)
// with packed dot products, so eta_p = dE/dt_p.  For canonical Hartree-Fock
// orbitals (F_ia = 0)
//   eta1[i][a] = 2 sum_jb L[(i,a),(j,b)] t1[j][b],
//   eta2[ai<=bj] = 2 L[(i,a),(j,b)] (ai < bj),  L[(i,a),(i,a)] (ai = bj),
// with L_iajb = 2 (ia|jb) - (ib|ja).
//
// Solver: reduced space, as CCEQ_SOL for ISIDE = -1.  Orthonormal trial
// vectors b_k and their transforms sigma_k = b_k A stay resident in HBM
// (two blocks of maxred packed vectors); the reduced matrix
// G_kl = b_k . sigma_l and right-hand side -b_k . eta grow by one row and
// column per iteration (two multi-dot launches), the K x K system is solved
// on the host, and lambda = sum c_k b_k, r = sum c_k sigma_k + eta are
// formed on the device.  The next trial vector is -r / Delta (Delta the
// orbital-energy differences of each packed element), orthogonalised twice
// against the space.  A full space collapses onto lambda (its transform is
// r - eta, so the collapse costs no transformation).
//
// With the BF16x3 bulk arithmetic the solve is an iterative refinement: the
// forward intermediates are formed once in FP64; each step takes the FP64
// residual r64 = lambda A + eta (one FP64 transformation) and solves the
// correction delta A + r64 = 0 to 1e-4 |r64| in a space of BF16x3
// transformations, which is kept from step to step (its matrix does not
// change, only the right-hand side), then lambda += delta.  The converged
// lambda satisfies the FP64 equations; the BF16x3 rounding (about 1e-5
// relative) only slows the contraction of each step.
//
// Density: D_pq = dL/dh_pq.  One left transformation of lambda leaves in Fb
// the cotangent of h~ after the first half of the T1 dressing
// (h~[o+a][q] = h[o+a][q] - sum_l t(l,a) h[l][q]); its adjoint gives
//   D[l][q] = Fb[l][q] - sum_a t(l,a) Fb[o+a][q]  (l < o),  Fb[p][q] (p >= o),
// and the energy adds 2 delta_ij (reference) and 2 t(i,a) at D[i][o+a]
// (the 2 F_ia t_ai term).
#include <algorithm>
#include <cmath>
#include <vector>

#include "ccsd.h"
#include "kernels.h"
#include "ops.h"

namespace ccgpu {

namespace {

// x[e] = -r[e] / Delta_e over the packed vector [singles; doubles] (x may be r).
__global__ void k_precond_packed(double* x, const double* r, const double* __restrict__ eps, int o, int v, long nov,
                                 long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double d;
    if (e < nov) {
      const long i = e / v, a = e - i * v;
      d = eps[o + a] - eps[i];
    } else {
      const long p = e - nov;
      long bj = (long)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
      while (bj * (bj + 1) / 2 > p) --bj;
      while ((bj + 1) * (bj + 2) / 2 <= p) ++bj;
      const long ai = p - bj * (bj + 1) / 2;
      const long i = ai / v, a = ai - i * v, j = bj / v, b = bj - j * v;
      d = eps[o + a] - eps[i] + eps[o + b] - eps[j];
    }
    x[e] = -r[e] / d;
  }
}

// out = a x + b y (y may be null; out may alias x or y)
__global__ void k_axpby(double* out, double a, const double* x, double b, const double* y, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    out[e] = a * x[e] + (y ? b * y[e] : 0.0);
}

// D[i][i] += f for i < o
__global__ void k_add_occ_diag2(double* D, long N, int o, double f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < o) D[(long)i * N + i] += f;
}

unsigned grid_of(long n) { return (unsigned)std::max<long>(1, std::min<long>(65536, (n + 255) / 256)); }

void axpby(double* out, double a, const double* x, double b, const double* y, long n, hipStream_t st) {
  check_span(out, 0, n - 1, sizeof(double), "axpby out");
  check_span(x, 0, n - 1, sizeof(double), "axpby x");
  if (y) check_span(y, 0, n - 1, sizeof(double), "axpby y");
  CCGPU_LAUNCH(k_axpby, dim3(grid_of(n)), dim3(256), 0, st, out, a, x, b, y, n);
}

// Solves G c = g (K x K, row-major) by Gaussian elimination with partial
// pivoting; false if a pivot vanishes.
bool small_solve(std::vector<double> G, std::vector<double> g, int K, std::vector<double>& c) {
  for (int col = 0; col < K; ++col) {
    int piv = col;
    for (int r = col + 1; r < K; ++r)
      if (std::fabs(G[r * K + col]) > std::fabs(G[piv * K + col])) piv = r;
    if (G[piv * K + col] == 0.0) return false;
    if (piv != col) {
      for (int k = 0; k < K; ++k) std::swap(G[col * K + k], G[piv * K + k]);
      std::swap(g[col], g[piv]);
    }
    for (int r = col + 1; r < K; ++r) {
      const double f = G[r * K + col] / G[col * K + col];
      for (int k = col; k < K; ++k) G[r * K + k] -= f * G[col * K + k];
      g[r] -= f * g[col];
    }
  }
  c.assign(K, 0.0);
  for (int r = K - 1; r >= 0; --r) {
    double s = g[r];
    for (int k = r + 1; k < K; ++k) s -= G[r * K + k] * c[k];
    c[r] = s / G[r * K + r];
  }
  return true;
}

}  // namespace

// y = x A for packed device vectors [x1; x2] (Om_ and bufX2_ as work space).
void CCSD::lhtr_packed(const double* x, double* y) {
  const long nov = dm_.nov;
  unpack_sym(Om_.p, x + nov, nov, st_);
  scale_diag(Om_.p, nov, nov, 2.0, st_);  // Xb = Lf + diag(Lf)
  lhtr_apply(x, Om_.p, y, bufX2_.p);
  pack_sym(y + nov, bufX2_.p, nov, st_);
}

void CCSD::eta_device(double* e) {
  const long nov = dm_.nov, nov2 = dm_.nov2;
  check_span(e, 0, dm_.ntamp - 1, sizeof(double), "eta");
  gemv_mc(nov, nov, 2.0, L_.p, nov, t1_.p, 0.0, e, st_);
  perm(bufX2_.p, L_.p, 1, 1, 1, nov2, 0, 0, 0, 1, 2.0, 0.0, st_);
  scale_diag(bufX2_.p, nov, nov, 0.5, st_);
  pack_sym(e + nov, bufX2_.p, nov, st_);
}

void CCSD::eta_host(double* e1, double* e2p, bool dalton) {
  CCGPU_REQUIRE(iter_ >= 1, "ccgpu_eta: no amplitudes (run the CCSD iterations or set amplitudes first)");
  ensure_work();
  DBuf<double> e((size_t)dm_.ntamp);
  eta_device(e.p);
  if (dalton) scale_pdiag(e.p + dm_.nov, dm_.nov, 2.0, st_);
  CCGPU_HIPX(hipMemcpyAsync(e1, e.p, dm_.nov * sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipMemcpyAsync(e2p, e.p + dm_.nov, (dm_.ntamp - dm_.nov) * sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
}

int CCSD::lambda_solve(double thr, int maxit, int maxred, double* l1, double* l2p, double* resnorm, int* niter,
                       bool dalton) {
  CCGPU_REQUIRE(iter_ >= 1, "ccgpu_lambda_solve: no amplitudes (run the CCSD iterations or set amplitudes first)");
  CCGPU_REQUIRE(thr > 0.0 && maxit > 0 && maxred >= 2, "ccgpu_lambda_solve: bad threshold, iteration or space limit");
  const long n = dm_.ntamp, nov = dm_.nov;
  const GemmPrec user = bulk_;
  const bool refine = user != GemmPrec::F64;
  maxred = std::min(maxred, 64);  // dots and coefficients staged in 128-entry device arrays
  // forward intermediates in FP64 for both modes: the BF16x3 transforms of
  // the refinement use them with BF16x3 bulk products
  set_precision(GemmPrec::F64);
  lhtr_prepare();
  ensure_lhtr();
  // space for eta, r, w, lambda, the FP64 residual, the correction and 2 x
  // maxred vectors, within the free memory
  constexpr long kFixed = 6;
  if (!dryrun()) {
    size_t fr = 0, tot = 0;
    CCGPU_HIP(hipMemGetInfo(&fr, &tot));
    const long fit = (long)((0.9 * (double)fr / (8.0 * (double)n) - kFixed) / 2.0);
    maxred = (int)std::min<long>(maxred, fit);
    CCGPU_REQUIRE(maxred >= 2, "ccgpu_lambda_solve: not enough device memory for the reduced space");
  }
  DBuf<double> work((size_t)(kFixed + 2 * (long)maxred) * n), sm(256);
  double* eta = work.p;
  double* r = work.p + n;       // residual of the reduced-space solution
  double* w = work.p + 2 * n;   // Gram-Schmidt work
  double* lam = work.p + 3 * n;
  double* r64 = work.p + 4 * n;  // FP64 residual lambda A + eta (refinement)
  double* dl = work.p + 5 * n;   // correction of one refinement step
  double* Bs = work.p + kFixed * n;
  double* Ss = Bs + (long)maxred * n;
  double* dd = sm.p;        // dots
  double* cd = sm.p + 128;  // coefficients
  eta_device(eta);

  auto dots = [&](const double* e, const double* E, int nvec) {
    std::vector<double> h(nvec, 0.0);
    if (nvec == 0) return h;
    multidot(dd, e, E, n, nvec, n, st_);
    CCGPU_HIPX(hipMemcpyAsync(h.data(), dd, nvec * sizeof(double), hipMemcpyDeviceToHost, st_));
    CCGPU_HIPX(hipStreamSynchronize(st_));
    return h;
  };
  auto combine = [&](double* out, const double* S, const std::vector<double>& c) {
    CCGPU_HIPX(hipMemcpyAsync(cd, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice, st_));
    lincomb(out, S, n, cd, (int)c.size(), n, st_);
  };
  // (a dry run reads no values: unit norms keep its control flow going)
  auto norm = [&](const double* x) { return dryrun() ? 1.0 : std::sqrt(std::max(0.0, dots(x, x, 1)[0])); };
  int it = 0;
  // y = x A in the given arithmetic (the forward intermediates stay FP64)
  lam_counts_[0] = lam_counts_[1] = 0;
  auto transform = [&](const double* x, double* y, GemmPrec p) {
    set_precision(p);
    lhtr_packed(x, y);
    set_precision(GemmPrec::F64);
    ++it;
    ++lam_counts_[p == GemmPrec::F64 ? 0 : 1];
  };

  // Reduced space: G_kl = b_k . sigma_l; the right-hand side is rebuilt from
  // the vector f of the equations x A + f = 0 (f = eta, or the FP64 residual).
  std::vector<double> G((size_t)maxred * maxred);
  int K = 0;
  const GemmPrec sp = refine ? user : GemmPrec::F64;  // arithmetic of the space's transforms
  auto add = [&](double* x) {  // appends x (overwritten): two Gram-Schmidt passes, normalise, transform
    for (int pass = 0; pass < 2 && K > 0; ++pass) {
      std::vector<double> d = dots(x, Bs, K);
      for (double& z : d) z = -z;
      combine(w, Bs, d);
      axpy(x, 1.0, w, n, st_);
    }
    const double nx = norm(x);
    if (!(nx > 1e-14)) return false;
    axpby(Bs + (long)K * n, 1.0 / nx, x, 0.0, nullptr, n, st_);
    transform(Bs + (long)K * n, Ss + (long)K * n, sp);
    const std::vector<double> row = dots(Bs + (long)K * n, Ss, K + 1);  // b_K . sigma_l
    const std::vector<double> col = dots(Ss + (long)K * n, Bs, K);      // b_l . sigma_K
    for (int l = 0; l <= K; ++l) G[(size_t)K * maxred + l] = row[l];
    for (int l = 0; l < K; ++l) G[(size_t)l * maxred + K] = col[l];
    ++K;
    return true;
  };
  // Solves x A + f = 0 in the space (x into out, r its residual) until
  // |r| < tol, the space gives no new direction or the transformation limit
  // is reached; returns |r|.  stall: also leave once |r| stops halving every
  // two steps.  A full space collapses onto x (FP64 solve) or ends the call
  // (refinement step; the caller restarts the space).
  auto solve_space = [&](const double* f, double tol, bool stall, double* out) {
    std::vector<double> hist;
    if (K == 0) {
      CCGPU_LAUNCH(k_precond_packed, dim3(grid_of(n)), dim3(256), 0, st_, r, f, eps_.p, dm_.o, dm_.v, nov, n);
      if (!add(r)) {
        fill(out, 0.0, n, st_);
        return 0.0;
      }
    }
    while (true) {
      std::vector<double> Gk((size_t)K * K), gk = dots(f, Bs, K), c;
      for (double& z : gk) z = -z;
      for (int a = 0; a < K; ++a)
        for (int b = 0; b < K; ++b) Gk[(size_t)a * K + b] = G[(size_t)a * maxred + b];
      CCGPU_REQUIRE(small_solve(Gk, gk, K, c) || dryrun(), "ccgpu_lambda_solve: singular reduced matrix");
      if (dryrun()) c.assign(K, 0.0);
      combine(out, Bs, c);
      combine(r, Ss, c);
      axpy(r, 1.0, f, n, st_);
      const double rn = norm(r);
      hist.push_back(rn);
      const bool stalled = stall && hist.size() >= 3 && rn > 0.5 * hist[hist.size() - 3];
      if (rn < tol || stalled || it >= maxit || dryrun()) return rn;
      if (K == maxred) {
        if (refine) return rn;  // the outer step takes this correction; the space restarts
        {  // collapse onto x; its transform is r - f
          const double nl = norm(out);
          axpby(Bs, 1.0 / nl, out, 0.0, nullptr, n, st_);
          axpby(Ss, 1.0 / nl, r, -1.0 / nl, f, n, st_);
          G[0] = dots(Bs, Ss, 1)[0];
          K = 1;
        }
      }
      CCGPU_LAUNCH(k_precond_packed, dim3(grid_of(n)), dim3(256), 0, st_, r, r, eps_.p, dm_.o, dm_.v, nov, n);
      if (!add(r)) return rn;
    }
  };

  double rn = 0.0;
  if (!refine) {
    rn = solve_space(eta, thr, false, lam);
  } else {
    // iterative refinement: r64 = lambda A + eta in FP64, the correction
    // delta A + r64 = 0 solved to 1e-4 |r64| in the BF16x3 space (kept across
    // steps: its matrix does not change), lambda += delta
    fill(lam, 0.0, n, st_);
    axpby(r64, 1.0, eta, 0.0, nullptr, n, st_);
    rn = norm(r64);
    while (rn >= thr && it < maxit && !dryrun()) {
      solve_space(r64, 1e-4 * rn, true, dl);
      axpy(lam, 1.0, dl, n, st_);
      if (K == maxred) K = 0;
      transform(lam, r64, GemmPrec::F64);
      axpy(r64, 1.0, eta, n, st_);
      rn = norm(r64);
    }
    if (dryrun()) {  // one pass of each step for the span checks
      solve_space(r64, 1.0, true, dl);
      axpy(lam, 1.0, dl, n, st_);
      transform(lam, r64, GemmPrec::F64);
    }
  }
  const bool conv = rn < thr || dryrun();
  set_precision(user);
  if (dalton) scale_pdiag(lam + nov, nov, 2.0, st_);
  if (l1) CCGPU_HIPX(hipMemcpyAsync(l1, lam, nov * sizeof(double), hipMemcpyDeviceToHost, st_));
  if (l2p) CCGPU_HIPX(hipMemcpyAsync(l2p, lam + nov, (n - nov) * sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
  CCGPU_HIPX(hipGetLastError());
  if (resnorm) *resnorm = rn;
  if (niter) *niter = it;
  if (!conv) throw Error(4, "ccgpu_lambda_solve: residual norm above the threshold after maxit transformations");
  return it;
}

void CCSD::onedens_host(const double* l1, const double* l2p, double* D, bool dalton) {
  CCGPU_REQUIRE(iter_ >= 1, "ccgpu_onedens: no amplitudes (run the CCSD iterations or set amplitudes first)");
  const long n = dm_.ntamp, nov = dm_.nov, N = dm_.N, NN = N * N;
  const int o = dm_.o, v = dm_.v;
  lhtr_prepare();
  ensure_lhtr();
  DBuf<double> x((size_t)2 * n), Dd((size_t)NN);
  CCGPU_HIPX(hipMemcpyAsync(x.p, l1, nov * sizeof(double), hipMemcpyHostToDevice, st_));
  CCGPU_HIPX(hipMemcpyAsync(x.p + nov, l2p, (n - nov) * sizeof(double), hipMemcpyHostToDevice, st_));
  if (dalton) scale_pdiag(x.p + nov, nov, 0.5, st_);
  lhtr_packed(x.p, x.p + n);  // leaves the cotangent of the half-dressed h~ in Fb
  CCGPU_HIPX(hipMemcpyAsync(Dd.p, Fb_.p, NN * sizeof(double), hipMemcpyDeviceToDevice, st_));
  g(o, N, v, -1.0, t1_.p, v, 1, Fb_.p + o * N, N, 1, 1.0, Dd.p, N, 1, st_);
  add_strided3(Dd.p + o, N, 1, 0, t1_.p, o, v, 1, 2.0, st_);
  CCGPU_LAUNCH(k_add_occ_diag2, dim3(cdiv(o, 256)), dim3(256), 0, st_, Dd.p, N, o, 2.0);
  CCGPU_HIPX(hipMemcpyAsync(D, Dd.p, NN * sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
  CCGPU_HIPX(hipGetLastError());
}

}  // namespace ccgpu
