// Closed-shell DF-CCSD residual and iteration on one GPU.
//
// Layouts (row-major): ov pair index (i,a) -> i*v + a (Dalton's NAI order).
//   T, U, Tx, Om, X2, C, D : [(i,a)][(j,b)]  (o*v x o*v, symmetric matrices)
//   B, Bt : [P][p][q] (p,q over N = o + v, occupied first)
//   Bov [P][(k,c)], Bvv [P][a][c], Boo [P][k][i], BvoT [P][(i,a)] = Bt[P][a][i]
// Every large contraction is an MFMA GEMM (gemm.hip); permutations and
// updates are the bandwidth kernels of kernels.hip.
#include "ccsd.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "kernels.h"
#include "ops.h"

namespace ccgpu {

namespace {

constexpr int kMaxVec = 80;  // Dalton MAXVEC, cc/ccsd_diis.F

// Gamma [(k,l),(i,j)] (the residual's hole-ladder matrix) into Dalton's
// packed NGAMMA order: pair ki = k + o i, element (ki, lj) with ki <= lj at
// lj (lj + 1) / 2 + ki (cc/ccsd_rhs.F:5158-5167).  Thread x = ki, y = lj.
__global__ void k_pack_gamma(double* __restrict__ out, const double* __restrict__ Gam, int o) {
  const long n = (long)o * o;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
  if (p > q || q >= n) return;
  const long k = p % o, i = p / o, l = q % o, j = q / o;
  out[q * (q + 1) / 2 + p] = Gam[((k * o + l) * o + i) * o + j];
}

}  // namespace

CCSD::CCSD(int o, int v, int naux, const double* eps_host, const double* B_host, int mxdiis)
    : mxdiis_(mxdiis) {
  CCGPU_REQUIRE(o > 0 && v > 0 && naux > 0, "ccgpu: bad dimensions");
  CCGPU_REQUIRE(mxdiis >= 1 && mxdiis <= kMaxVec, "ccgpu: MXDIIS out of range (1..80)");
  dm_.o = o; dm_.v = v; dm_.naux = naux;
  dm_.N = o + v;
  dm_.nov = (long)o * v;
  dm_.nov2 = dm_.nov * dm_.nov;
  dm_.npo = (long)o * (o + 1) / 2;
  dm_.npv = (long)v * (v + 1) / 2;
  dm_.ntamp = dm_.nov + dm_.nov * (dm_.nov + 1) / 2;
  CCGPU_REQUIRE(dm_.nov <= 65535 * 8, "ccgpu: o*v too large");
  CCGPU_HIPX(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  const long N = dm_.N, nov = dm_.nov, nov2 = dm_.nov2;

  eps_.alloc(N);
  B_.alloc((size_t)naux * N * N);
  CCGPU_HIPX(hipMemcpyAsync(eps_.p, eps_host, N * sizeof(double), hipMemcpyHostToDevice, st_));
  CCGPU_HIPX(hipMemcpyAsync(B_.p, B_host, B_.bytes(), hipMemcpyHostToDevice, st_));

  // implicit one-electron operator: h = diag(eps) - sum_k [2(pq|kk) - (pk|kq)]
  h_.alloc(N * N);
  Jd_.alloc(naux);
  coulomb_diag(Jd_.p, B_.p, naux, (int)N, o, st_);
  gemv_mc(N * N, naux, -1.0, B_.p, N * N, Jd_.p, 0.0, h_.p, st_);
  g(N, N, (long)naux * o, 1.0, B_.p, N, 1, B_.p, N, 1, 1.0, h_.p, N, 1, st_, 1, 0, 0, 0, o, N * N, N * N);
  if (!dryrun()) {
    std::vector<double> hh(N * N);
    CCGPU_HIPX(hipMemcpyAsync(hh.data(), h_.p, N * N * sizeof(double), hipMemcpyDeviceToHost, st_));
    CCGPU_HIPX(hipStreamSynchronize(st_));
    for (long p = 0; p < N; ++p) hh[p * N + p] += eps_host[p];
    CCGPU_HIP(hipMemcpy(h_.p, hh.data(), N * N * sizeof(double), hipMemcpyHostToDevice));
  }

  // ovov constants: J = (kc|ld), K = X(J) = (kd|lc), L = 2J - K, Joovv[(k,l),(c,d)]
  Bov_.alloc((size_t)naux * nov);
  perm(Bov_.p, B_.p + o, 1, naux, o, v, 0, N * N, N, 1, 1.0, 0.0, st_);
  J_.alloc(nov2);
  K_.alloc(nov2);
  L_.alloc(nov2);
  Joovv_.alloc(nov2);
  g(nov, nov, naux, 1.0, Bov_.p, 1, nov, Bov_.p, nov, 1, 0.0, J_.p, nov, 1, st_);
  perm(K_.p, J_.p, o, v, o, v, v, nov, (long)v * nov, 1, 1.0, 0.0, st_);
  perm(L_.p, J_.p, 1, 1, 1, nov2, 0, 0, 0, 1, 2.0, 0.0, st_);
  perm(L_.p, K_.p, 1, 1, 1, nov2, 0, 0, 0, 1, -1.0, 1.0, st_);
  perm(Joovv_.p, J_.p, o, o, v, v, (long)v * nov, v, nov, 1, 1.0, 0.0, st_);

  t1_.alloc(nov);
  T_.alloc(nov2);
  om1_.alloc(nov);
  Om_.alloc(nov2);
  dots_.alloc(kMaxVec + 1);
  coef_.alloc(kMaxVec + 1);
  Amat_.alloc((kMaxVec + 1) * (kMaxVec + 1));
  diiswork_.alloc((kMaxVec + 2) * (kMaxVec + 2));
  ctr_.alloc(4);
  scal_.alloc(4);
  if (const char* e = std::getenv("CCGPU_LADDER_BUDGET_MB")) {
    const double mb = std::atof(e);
    if (mb > 0.0) ladder_budget_ = mb * 1.0e6;
  }
  CCGPU_HIPX(hipStreamSynchronize(st_));
}

void CCSD::ensure_work() {
  if (work_) return;
  pt_release();
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long N = dm_.N, nov = dm_.nov, nov2 = dm_.nov2;
  (void)N;
  // per-iteration buffers
  Bt_.alloc((size_t)naux * N * N);
  Bvv_.alloc((size_t)naux * v * v);
  Boo_.alloc((size_t)naux * o * o);
  BvoT_.alloc((size_t)naux * nov);
  F_.alloc(N * N);
  ht_.alloc(N * N);
  hx_.alloc(nov);
  G1_.alloc((size_t)o * o * o * o);
  Gam_.alloc((size_t)o * o * o * o);
  Z_.alloc((size_t)naux * nov);
  W_.alloc((size_t)o * o * nov);
  E1_.alloc((size_t)v * v);
  E2_.alloc((size_t)o * o);
  bufR_.alloc(std::max<long>(nov2, (long)v * v * o * o));
  bufX2_.alloc(std::max<long>(nov2, dm_.ntamp));
  bufTx_.alloc(nov2);
  bufU_.alloc(nov2);
  bufC_.alloc(nov2);
  bufD_.alloc(nov2);
  taup_.alloc(dm_.npo * dm_.npv);
  taum_.alloc(dm_.npo * dm_.npv);

  // particle-ladder batches over a: V+/- rows (a,b>=a) x npv and the
  // Y[(a,c),(b,d)] block must fit a fixed budget; rows < 65536 (grid.y).
  // bytes per V+/- / Y buffer (10 GB unless set; CCGPU_LADDER_BUDGET_MB overrides)
  const double budget = ladder_budget_;
  long maxrows = std::max<long>(1, std::min<long>(65535, (long)(budget / (8.0 * dm_.npv))));
  // Rows per batch in whole waves of the S product's 256 x 256 tiles over
  // the CUs (V+ and V- each ceil(npo/256) tiles wide), so that no batch ends
  // in a nearly empty wave; the grouping of rows changes no result bit.
  {
    int dev = 0, ncu = 0;
    if (!dryrun() && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0) {
      const long wide = 2 * ((dm_.npo + 255) / 256);
      const long blocks_per_wave = ncu / wide;
      const long step = blocks_per_wave * 256;
      if (blocks_per_wave > 0 && maxrows >= step) maxrows = maxrows / step * step;
    }
  }
  lad_batches_.clear();
  lad_batches_.push_back(0);
  long rows = 0, maxr = 0, maxy = 0;
  int a0 = 0;
  for (int a = 0; a < v; ++a) {
    const long add = v - a;
    const long ysz = (long)(a + 1 - a0) * v * (long)(v - a0) * v;
    if (rows > 0 && (rows + add > maxrows || 8.0 * ysz > budget)) {
      lad_batches_.push_back(a);
      maxr = std::max(maxr, rows);
      maxy = std::max(maxy, (long)(a - a0) * v * (long)(v - a0) * v);
      a0 = a;
      rows = 0;
    }
    rows += add;
  }
  lad_batches_.push_back(v);
  maxr = std::max(maxr, rows);
  maxy = std::max(maxy, (long)(v - a0) * v * (long)(v - a0) * v);
  Y_.alloc(maxy);
  Vp_.alloc(maxr * dm_.npv);
  Vm_.alloc(maxr * dm_.npv);
  Sp_.alloc(maxr * dm_.npo);
  Sm_.alloc(maxr * dm_.npo);

  work_ = true;
}

// DIIS history (Dalton's DIIST/DIISS files).  Released by the left
// transformation; reallocating it restarts the DIIS cycle.
void CCSD::ensure_diis() {
  if (diisS_.p) return;
  diisS_.alloc((size_t)mxdiis_ * dm_.ntamp);
  diisE_.alloc((size_t)mxdiis_ * dm_.ntamp);
  diis_off_ = iter_ > 0 ? iter_ - 1 : 0;
  reset_cycle();
}

// The next iteration starts a new DIIS cycle (device counter = 1).
void CCSD::reset_cycle() {
  const int c[4] = {1, 0, 0, 0};
  CCGPU_HIPX(hipMemcpyAsync(ctr_.p, c, sizeof c, hipMemcpyHostToDevice, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
}

void CCSD::release_work() {
  if (!work_) return;
  for (DBuf<double>* b : {&Bt_, &Bvv_, &Boo_, &BvoT_, &F_, &ht_, &hx_, &G1_, &Gam_, &Z_, &W_, &E1_, &E2_,
                          &bufR_, &bufX2_, &bufTx_, &bufU_, &bufC_, &bufD_, &taup_, &taum_, &Y_, &Vp_,
                          &Vm_, &Sp_, &Sm_, &S2_, &diisS_, &diisE_, &Tb_, &Ub_, &Txb_, &Db_, &Btb_, &Bvvb_, &Boob_,
                          &BvoTb_, &Zb_, &Wb_, &Fb_, &E1b_, &E2b_, &Jdb_, &Gamb_, &G1b_, &taupb_, &taumb_, &vecb_,
                          &lst_, &rt1_, &rT_})
    b->release();
  for (DBuf<unsigned short>* b : {&lpBh_, &lpBl_, &lpTh_, &lpTl_, &lpVh_, &lpVl_}) b->release();
  work_ = false;
  lhtr_ready_ = false;
  diis_off_ = iter_ > 0 ? iter_ - 1 : 0;  // the DIIS history is gone: restart the cycle
}

// Second stream (default priority) with two pairs of events, made once:
// staging copies of the left transformation, (T) energies beside the W passes.
void CCSD::side_stream() {
  if (cst_ || dryrun()) return;
  CCGPU_HIP(hipStreamCreateWithFlags(&cst_, hipStreamNonBlocking));
  for (int q = 0; q < 2; ++q) {
    CCGPU_HIP(hipEventCreateWithFlags(&ev_in_[q], hipEventDisableTiming));
    CCGPU_HIP(hipEventCreateWithFlags(&ev_done_[q], hipEventDisableTiming));
  }
}

CCSD::~CCSD() {
  if (dryrun()) return;
  if (cst_) {
    (void)hipStreamSynchronize(cst_);
    for (int q = 0; q < 2; ++q) {
      if (ev_in_[q]) (void)hipEventDestroy(ev_in_[q]);
      if (ev_done_[q]) (void)hipEventDestroy(ev_done_[q]);
    }
    (void)hipStreamDestroy(cst_);
  }
  if (st_) {
    (void)hipStreamSynchronize(st_);
    release_scratch(st_);
    (void)hipStreamDestroy(st_);
  }
}

void CCSD::mark(const char* name) { prof_phase(name, st_); }

// T1 transformation of B and h, then the dressed Fock matrix
// (LAMMAT/LAMDA1 cc/ccsd_rhs.F:5887,6030; CC_AOFOCK/CC_FCKMO :3388,2724).
void CCSD::dress(const double* t1) {
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long N = dm_.N, NN = N * N, nov = dm_.nov;
  CCGPU_HIPX(hipMemcpyAsync(Bt_.p, B_.p, B_.bytes(), hipMemcpyDeviceToDevice, st_));
  // Bt[P][o+a][q] -= sum_l t(l,a) B[P][l][q]
  g(v, N, o, -1.0, t1, 1, v, B_.p, N, 1, 1.0, Bt_.p + o * N, N, 1, st_, naux, 0, NN, NN);
  // Bt[P][p][i] += sum_a Bt[P][p][o+a] t(i,a)
  g(N, o, v, 1.0, Bt_.p + o, N, 1, t1, 1, v, 1.0, Bt_.p, N, 1, st_, naux, NN, 0, NN);
  CCGPU_HIPX(hipMemcpyAsync(ht_.p, h_.p, NN * sizeof(double), hipMemcpyDeviceToDevice, st_));
  g(v, N, o, -1.0, t1, 1, v, h_.p, N, 1, 1.0, ht_.p + o * N, N, 1, st_);
  g(N, o, v, 1.0, ht_.p + o, N, 1, t1, 1, v, 1.0, ht_.p, N, 1, st_);
  // F = ht + sum_P Bt[P] * 2 sum_k Bt[P,k,k] - sum_{P,k} Bt[P,p,k] Bt[P,k,q]
  coulomb_diag(Jd_.p, Bt_.p, naux, (int)N, o, st_);
  CCGPU_HIPX(hipMemcpyAsync(F_.p, ht_.p, NN * sizeof(double), hipMemcpyDeviceToDevice, st_));
  gemv_mc(NN, naux, 1.0, Bt_.p, NN, Jd_.p, 1.0, F_.p, st_);
  g(N, N, (long)naux * o, -1.0, Bt_.p, N, 1, Bt_.p, N, 1, 1.0, F_.p, N, 1, st_, 1, 0, 0, 0, o, NN, NN);
  // contiguous blocks
  perm(Bvv_.p, Bt_.p + o * N + o, 1, naux, v, v, 0, NN, N, 1, 1.0, 0.0, st_);
  perm(Boo_.p, Bt_.p, 1, naux, o, o, 0, NN, N, 1, 1.0, 0.0, st_);
  perm(BvoT_.p, Bt_.p + o * N, 1, naux, o, v, 0, NN, 1, N, 1.0, 0.0, st_);
  (void)nov;
}

// Particle-particle ladder sum_cd t_ij^cd (ac~|bd~) in the symmetric /
// antisymmetric scheme (CC_BF/CC_BF1/CCRHS_IPM, cc/ccsd_rhs.F:9129,9319,7571):
// V+/-_{ab,cd} = (ac|bd) +/- (ad|bc) for a<=b, c<=d, formed per a-batch from
// the dressed DF factor (v^4 never stored), contracted with tau+/- on MFMA.
void CCSD::ladder(const double* T, double* Om) {
  (void)T;
  if (bulk_ == GemmPrec::BF16X3) {
    ladder_mixed(Om);
    return;
  }
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long vv = (long)v * v;
  const Gemm bulk{bulk_, st_};
  for (size_t bi = 0; bi + 1 < lad_batches_.size(); ++bi) {
    const int a0 = lad_batches_[bi], a1 = lad_batches_[bi + 1];
    long rows = 0;
    for (int a = a0; a < a1; ++a) rows += v - a;
    const long M = (long)(a1 - a0) * v, Nn = (long)(v - a0) * v;
    bulk(M, Nn, naux, 1.0, Bvv_.p + (long)a0 * v, 1, vv, Bvv_.p + (long)a0 * v, vv, 1, 0.0, Y_.p, Nn, 1);
    ladder_vpm(Vp_.p, Vm_.p, Y_.p, Nn, a0, a1, v, st_);
    bulk(rows, dm_.npo, dm_.npv, 1.0, Vp_.p, dm_.npv, 1, taup_.p, 1, dm_.npv, 0.0, Sp_.p, dm_.npo, 1);
    bulk(rows, dm_.npo, dm_.npv, 1.0, Vm_.p, dm_.npv, 1, taum_.p, 1, dm_.npv, 0.0, Sm_.p, dm_.npo, 1);
    ladder_scatter(Om, Sp_.p, Sm_.p, a0, a1, o, v, st_);
  }
}

// The same ladder with BF16x3 products.  The dressed factor Bvv and tau+/-
// are split into bf16 hi/lo once per call (not per batch); each batch forms
// Y in FP32 straight from the packed factor, and V+/- goes from Y directly
// into the packed operand of one batched (+ and -) S GEMM.
void CCSD::ladder_mixed(double* Om) {
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long vv = (long)v * v, npo = dm_.npo, npv = dm_.npv;
  long maxr = 0;
  for (size_t bi = 0; bi + 1 < lad_batches_.size(); ++bi) {
    long rows = 0;
    for (int a = lad_batches_[bi]; a < lad_batches_[bi + 1]; ++a) rows += v - a;
    maxr = std::max(maxr, rows);
  }
  // rows of the packed factor: every batch reads a0*v + rpad(...) < vv + (row padding) rows
  const long RpB = bf16x3_rpad(vv + bf16x3_rpad(1) - 1), KpB = bf16x3_kpad(naux);
  const long RpT = bf16x3_rpad(npo), KpT = bf16x3_kpad(npv);
  const long RpV = bf16x3_rpad(maxr);
  lpBh_.ensure((size_t)RpB * KpB);
  lpBl_.ensure((size_t)RpB * KpB);
  lpTh_.ensure((size_t)2 * RpT * KpT);
  lpTl_.ensure((size_t)2 * RpT * KpT);
  lpVh_.ensure((size_t)2 * RpV * KpT);
  lpVl_.ensure((size_t)2 * RpV * KpT);
  S2_.ensure((size_t)2 * maxr * npo);
  bf16x3_pack(Bvv_.p, vv, naux, 1, vv, 0, 1, RpB, KpB, 1.0f, lpBh_.p, lpBl_.p, st_);
  bf16x3_pack(taup_.p, npo, npv, npv, 1, 0, 1, RpT, KpT, 1.0f, lpTh_.p, lpTl_.p, st_);
  bf16x3_pack(taum_.p, npo, npv, npv, 1, 0, 1, RpT, KpT, 1.0f, lpTh_.p + RpT * KpT, lpTl_.p + RpT * KpT, st_);
  float* Y32 = reinterpret_cast<float*>(Y_.p);
  double* Sp = S2_.p;
  double* Sm = S2_.p + maxr * npo;
  for (size_t bi = 0; bi + 1 < lad_batches_.size(); ++bi) {
    const int a0 = lad_batches_[bi], a1 = lad_batches_[bi + 1];
    long rows = 0;
    for (int a = a0; a < a1; ++a) rows += v - a;
    const long M = (long)(a1 - a0) * v, Nn = (long)(v - a0) * v;
    CCGPU_REQUIRE((size_t)(M * Nn) <= 2 * Y_.n, "ladder_mixed: Y block exceeds its buffer");
    const BTerm ty{lpBh_.p + a0 * v * KpB, lpBl_.p + a0 * v * KpB, lpBh_.p + a0 * v * KpB,
                   lpBl_.p + a0 * v * KpB, KpB, 0, 0};
    bf16x3_multi(&ty, 1, M, Nn, 1, Y32, true, Nn, 1, 0, 1.0, 0.0, 2.0 * M * Nn * naux, st_);
    const long Rp = bf16x3_rpad(rows);
    ladder_vpm_pack(lpVh_.p, lpVl_.p, Y32, Nn, a0, a1, v, Rp, KpT, st_);
    const BTerm ts{lpVh_.p, lpVl_.p, lpTh_.p, lpTl_.p, KpT, Rp * KpT, RpT * KpT};
    bf16x3_multi(&ts, 1, rows, npo, 2, Sp, false, npo, 1, maxr * npo, 1.0, 0.0, 4.0 * rows * npo * npv, st_);
    ladder_scatter(Om, Sp, Sm, a0, a1, o, v, st_);
  }
}

// CCSD vector function (CCRHSN, cc/ccsd_rhs.F:20-2278), T1-dressed form.
void CCSD::residual(const double* t1, const double* T, double* om1, double* Om) {
  ensure_work();
  lhtr_ready_ = false;
  const Gemm bulk{bulk_, st_};
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long N = dm_.N, nov = dm_.nov, nov2 = dm_.nov2, oo = (long)o * o, vv = (long)v * v;
  mark("start");
  dress(t1);
  mark("dress");

  // A2 integral part: (ai~|bj~)
  bulk(nov, nov, naux, 1.0, BvoT_.p, 1, nov, BvoT_.p, nov, 1, 0.0, Om, nov, 1);
  mark("A2_int");

  // Too[(i,j),(a,b)] (into bufR), tau+/-, hole-hole ladder (CCRHS_A :2505)
  double* Too = bufR_.p;
  perm(Too, T, o, o, v, v, (long)v * nov, v, nov, 1, 1.0, 0.0, st_);
  ladder_tau(taup_.p, taum_.p, Too, o, v, st_);
  g(oo, oo, naux, 1.0, Boo_.p, 1, oo, Boo_.p, oo, 1, 0.0, G1_.p, oo, 1, st_);  // [(k,i),(l,j)]
  perm(Gam_.p, G1_.p, o, o, o, o, oo * o, o, oo, 1, 1.0, 0.0, st_);               // [(k,l),(i,j)]
  bulk(oo, oo, vv, 1.0, Joovv_.p, vv, 1, Too, 1, vv, 1.0, Gam_.p, oo, 1);
  double* LADh = bufX2_.p;
  bulk(oo, vv, oo, 1.0, Gam_.p, 1, oo, Too, vv, 1, 0.0, LADh, vv, 1);           // [(i,j),(a,b)]
  perm(Om, LADh, o, v, o, v, (long)o * vv, v, vv, 1, 1.0, 1.0, st_);
  mark("B2_hole_ladder");

  ladder(T, Om);
  mark("A2_ladder");

  // ring terms (CCRHS_C/CIO :4272,7656; CCRHS_D/DIO :3525,8023)
  double* Tx = bufTx_.p;
  double* U = bufU_.p;
  perm(Tx, T, o, v, o, v, v, nov, (long)v * nov, 1, 1.0, 0.0, st_);
  perm(U, T, 1, 1, 1, nov2, 0, 0, 0, 1, 2.0, 0.0, st_);
  perm(U, Tx, 1, 1, 1, nov2, 0, 0, 0, 1, -1.0, 1.0, st_);
  double* R = bufR_.p;
  bulk(vv, oo, naux, 1.0, Bvv_.p, 1, vv, Boo_.p, oo, 1, 0.0, R, oo, 1);  // [(a,c),(k,i)] = (ki~|ac~)
  double* C = bufC_.p;
  double* D = bufD_.p;
  perm(C, R, o, v, o, v, 1, (long)v * oo, o, oo, 1.0, 0.0, st_);   // C[(i,a),(k,c)]
  perm(D, R, o, v, o, v, 1, (long)v * oo, o, oo, -1.0, 0.0, st_);  // D = -(ac~|ki~)
  mark("CD_int");
  bulk(nov, nov, nov, -0.5, Tx, nov, 1, K_.p, nov, 1, 1.0, C, nov, 1);
  bulk(nov, nov, naux, 2.0, BvoT_.p, 1, nov, Bov_.p, nov, 1, 1.0, D, nov, 1);
  bulk(nov, nov, nov, 0.5, U, nov, 1, L_.p, nov, 1, 1.0, D, nov, 1);
  mark("CD_intermediates");
  double* P = bufR_.p;
  double* X2 = bufX2_.p;
  bulk(nov, nov, nov, 1.0, C, nov, 1, Tx, nov, 1, 0.0, P, nov, 1);
  perm(X2, P, 1, 1, 1, nov2, 0, 0, 0, 1, -0.5, 0.0, st_);
  perm(X2, P, o, v, o, v, v, nov, (long)v * nov, 1, -1.0, 1.0, st_);
  bulk(nov, nov, nov, 0.5, D, nov, 1, U, nov, 1, 1.0, X2, nov, 1);
  mark("CD_contract");

  // E terms (CCRHS_EI/EFCK/E :3158,8604,2280)
  perm(E1_.p, F_.p + o * N + o, 1, 1, v, v, 0, 0, N, 1, 1.0, 0.0, st_);
  g(v, v, (long)o * nov, -1.0, U, nov, 1, J_.p, 1, nov, 1.0, E1_.p, v, 1, st_, 1, 0, 0, 0, nov,
    (long)v * nov, (long)v * nov);
  perm(E2_.p, F_.p, 1, 1, o, o, 0, 0, N, 1, 1.0, 0.0, st_);
  g(o, o, (long)v * nov, 1.0, J_.p, (long)v * nov, 1, U, 1, (long)v * nov, 1.0, E2_.p, o, 1, st_, 1, 0,
    0, 0, nov, nov, nov);
  g(oo * v, v, v, 1.0, T, v, 1, E1_.p, 1, v, 1.0, X2, v, 1, st_);
  g(o, v, o, -1.0, E2_.p, 1, o, T, v, 1, 1.0, X2, v, 1, st_, (int)nov, 0, nov, nov);
  mark("E");

  // Om += X2 + X2^T  (P_ij^ab)
  perm(Om, X2, 1, 1, 1, nov2, 0, 0, 0, 1, 1.0, 1.0, st_);
  perm(Om, X2, 1, 1, nov, nov, 0, 0, 1, nov, 1.0, 1.0, st_);
  mark("sym");

  // singles (CCRHS_G/H/I/J :2986,2858,2450,2686)
  bulk(naux, nov, nov, 1.0, Bov_.p, nov, 1, U, nov, 1, 0.0, Z_.p, nov, 1);
  g(o, v, (long)naux * v, 1.0, Z_.p, v, 1, Bvv_.p, 1, v, 0.0, om1, v, 1, st_, 1, 0, 0, 0, v, nov, vv);
  g(oo, nov, naux, 1.0, Boo_.p, 1, oo, Bov_.p, nov, 1, 0.0, W_.p, nov, 1, st_);
  g(o, v, (long)o * nov, -1.0, W_.p, nov, 1, U, 1, nov, 1.0, om1, v, 1, st_, 1, 0, 0, 0, nov,
    (long)o * nov, (long)v * nov);
  perm(hx_.p, F_.p + o, 1, 1, o, v, 0, 0, N, 1, 1.0, 0.0, st_);  // F_kc
  gemv_mc(nov, nov, 1.0, U, nov, hx_.p, 1.0, om1, st_);
  perm(om1, F_.p + o * N, 1, 1, o, v, 0, 0, 1, N, 1.0, 1.0, st_);  // F~_ai
  mark("singles");
}

double CCSD::mp2_start() {
  const long nov = dm_.nov;
  fill(t1_.p, 0.0, nov, st_);
  mp2_guess(T_.p, J_.p, eps_.p, dm_.o, dm_.v, st_);
  ccsd_energy(scal_.p + 1, L_.p, T_.p, t1_.p, nov, st_);
  double e = 0.0;
  CCGPU_HIPX(hipMemcpyAsync(&e, scal_.p + 1, sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
  iter_ = 1;
  diis_off_ = 0;
  lhtr_ready_ = false;
  reset_cycle();
  return e;
}

double CCSD::energy_current() {
  CCGPU_REQUIRE(iter_ >= 1, "ccgpu: no amplitudes");
  ccsd_energy(scal_.p + 1, L_.p, T_.p, t1_.p, dm_.nov, st_);
  double e = 0.0;
  CCGPU_HIPX(hipMemcpyAsync(&e, scal_.p + 1, sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
  return e;
}

// CCSD_ECCSD (cc/ccsd_energy.F:1964) on host amplitudes.
double CCSD::energy_host(const double* t1, const double* t2p) {
  const long nov = dm_.nov, np2 = dm_.ntamp - nov;
  DBuf<double> d((size_t)dm_.ntamp);
  CCGPU_HIPX(hipMemcpyAsync(d.p, t1, nov * sizeof(double), hipMemcpyHostToDevice, st_));
  CCGPU_HIPX(hipMemcpyAsync(d.p + nov, t2p, np2 * sizeof(double), hipMemcpyHostToDevice, st_));
  ccsd_energy_packed(scal_.p + 1, L_.p, d.p + nov, d.p, nov, st_);
  double e = 0.0;
  CCGPU_HIPX(hipMemcpyAsync(&e, scal_.p + 1, sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
  return e;
}

// CCSD_DIIS (cc/ccsd_diis.F:251-365) on host vectors: slot, B-matrix row and
// bordered solve as in iterate(), with the cycle position from niter.
bool CCSD::diis_host(const double* s, double* t, int niter) {
  CCGPU_REQUIRE(niter >= 1, "ccgpu_diis: niter >= 1");
  const long nt = dm_.ntamp;
  ensure_diis();
  const int c[2] = {niter, 0};
  CCGPU_HIPX(hipMemcpyAsync(ctr_.p, c, sizeof c, hipMemcpyHostToDevice, st_));
  DBuf<double> d((size_t)(2 * nt));
  CCGPU_HIPX(hipMemcpyAsync(d.p, s, nt * sizeof(double), hipMemcpyHostToDevice, st_));
  CCGPU_HIPX(hipMemcpyAsync(d.p + nt, t, nt * sizeof(double), hipMemcpyHostToDevice, st_));
  const int mx = mxdiis_;
  pack_diis_slot_flat(diisS_.p, diisE_.p, ctr_.p, mx, d.p, d.p + nt, nt, st_);
  multidot_slot(dots_.p, diisE_.p, nt, ctr_.p, mx, nt, st_);
  diis_solve(Amat_.p, kMaxVec + 1, dots_.p, coef_.p, ctr_.p, mx, diiswork_.p, st_);
  lincomb_slot(d.p + nt, diisS_.p, nt, coef_.p, ctr_.p, mx, nt, st_);
  int flag[2] = {0, 0};
  CCGPU_HIPX(hipMemcpyAsync(t, d.p + nt, nt * sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipMemcpyAsync(flag, ctr_.p, sizeof flag, hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
  lhtr_ready_ = false;
  return flag[1] == 0;
}

// One pass of the CCSD_ENERGY loop (cc/ccsd_energy.F:742-982): residual,
// CCSD_NXTAM, CCSD_DIIS (cc/ccsd_diis.F:251-365), CCSD_ECCSD, all on the
// device: the DIIS slot, the B-matrix extension and the bordered solve take
// the cycle count from the device counter.
void CCSD::prepare_iteration() {
  CCGPU_REQUIRE(iter_ >= 1, "ccgpu: call mp2_start (or set amplitudes) before iterating");
  ensure_work();
  ensure_diis();
}

void CCSD::enqueue_iteration(double* report) {
  const int o = dm_.o, v = dm_.v;
  const long nov = dm_.nov, nt = dm_.ntamp;
  const int mx = mxdiis_;
  residual(t1_.p, T_.p, om1_.p, Om_.p);
  nxtam(om1_.p, Om_.p, t1_.p, T_.p, eps_.p, o, v, scal_.p, st_);
  pack_diis_slot(diisS_.p, diisE_.p, ctr_.p, mx, om1_.p, Om_.p, t1_.p, T_.p, nov, st_);
  multidot_slot(dots_.p, diisE_.p, nt, ctr_.p, mx, nt, st_);
  diis_solve(Amat_.p, kMaxVec + 1, dots_.p, coef_.p, ctr_.p, mx, diiswork_.p, st_);
  double* newv = bufX2_.p;
  lincomb_slot(newv, diisS_.p, nt, coef_.p, ctr_.p, mx, nt, st_);
  CCGPU_HIPX(hipMemcpyAsync(t1_.p, newv, nov * sizeof(double), hipMemcpyDeviceToDevice, st_));
  unpack_sym(T_.p, newv + nov, nov, st_);
  ccsd_energy(scal_.p + 1, L_.p, T_.p, t1_.p, nov, st_);
  ctr_advance(ctr_.p, st_);
  if (report) report4(report, scal_.p, ctr_.p, st_);
}

double CCSD::finish_iteration(double* omnm) {
  double sc[2] = {0.0, 0.0};
  int flag[2] = {0, 0};
  CCGPU_HIPX(hipMemcpyAsync(sc, scal_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipMemcpyAsync(flag, ctr_.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
  note_iteration();
  if (flag[1]) throw Error(3, "DIIS: ZERO PIVOT - MATRIX SINGULAR");
  if (omnm) *omnm = std::sqrt(sc[0]);
  return sc[1];
}

double CCSD::iterate(double* omnm) {
  prepare_iteration();
  enqueue_iteration(nullptr);
  return finish_iteration(omnm);
}

void CCSD::set_amplitudes(const double* t1_host, const double* t2p_host) {
  const long nov = dm_.nov;
  CCGPU_HIPX(hipMemcpyAsync(t1_.p, t1_host, nov * sizeof(double), hipMemcpyDefault, st_));
  CCGPU_HIPX(hipMemcpyAsync(Om_.p, t2p_host, (dm_.ntamp - nov) * sizeof(double),
                           hipMemcpyDefault, st_));
  unpack_sym(T_.p, Om_.p, nov, st_);
  CCGPU_HIPX(hipStreamSynchronize(st_));
  iter_ = 1;
  diis_off_ = 0;
  lhtr_ready_ = false;
  reset_cycle();
}

void CCSD::get_amplitudes(double* t1_host, double* t2p_host) {
  const long nov = dm_.nov;
  if (t1_host)
    CCGPU_HIPX(hipMemcpyAsync(t1_host, t1_.p, nov * sizeof(double), hipMemcpyDefault, st_));
  if (t2p_host) {
    pack_sym(Om_.p, T_.p, nov, st_);
    CCGPU_HIPX(hipMemcpyAsync(t2p_host, Om_.p, (dm_.ntamp - nov) * sizeof(double),
                             hipMemcpyDefault, st_));
  }
  CCGPU_HIPX(hipStreamSynchronize(st_));
}

void CCSD::e_intermediates_host(double* e1, double* e2) {
  CCGPU_REQUIRE(iter_ >= 1, "ccgpu: no amplitudes (run the CCSD iterations or set amplitudes first)");
  const int o = dm_.o, v = dm_.v;
  residual(t1_.p, T_.p, om1_.p, Om_.p);  // E1_ [b][c], E2_ [k][j] at the current amplitudes
  // transposed into Dalton's column-major order (the leading index fastest)
  perm(bufR_.p, E1_.p, 1, 1, v, v, 0, 0, 1, v, 1.0, 0.0, st_);
  perm(bufR_.p + (long)v * v, E2_.p, 1, 1, o, o, 0, 0, 1, o, 1.0, 0.0, st_);
  if (e1) CCGPU_HIPX(hipMemcpyAsync(e1, bufR_.p, (size_t)v * v * sizeof(double), hipMemcpyDeviceToHost, st_));
  if (e2)
    CCGPU_HIPX(hipMemcpyAsync(e2, bufR_.p + (long)v * v, (size_t)o * o * sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
}

void CCSD::gamma_host(double* gamma) {
  CCGPU_REQUIRE(iter_ >= 1, "ccgpu: no amplitudes (run the CCSD iterations or set amplitudes first)");
  const int o = dm_.o;
  const long n = (long)o * o, ng = n * (n + 1) / 2;
  residual(t1_.p, T_.p, om1_.p, Om_.p);  // Gam_ at the current amplitudes
  DBuf<double> packed((size_t)ng);
  check_span(Gam_.p, 0, n * n - 1, sizeof(double), "gamma source");
  CCGPU_REQUIRE(n <= 65535, "gamma: occupied space too large");
  CCGPU_LAUNCH(k_pack_gamma, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, st_, packed.p, Gam_.p, o);
  CCGPU_HIPX(hipGetLastError());
  CCGPU_HIPX(hipMemcpyAsync(gamma, packed.p, (size_t)ng * sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
}

void CCSD::residual_host(const double* t1, const double* t2p, double* om1, double* om2p) {
  // The vector function of the given amplitudes only: they are staged in
  // buffers of their own, so the context's amplitudes, iteration count and
  // DIIS history and cycle are left as they were (CCRHSN has no side effects
  // on the caller's iteration state either).
  const long nov = dm_.nov, np2 = dm_.ntamp - nov;
  ensure_work();
  rt1_.ensure((size_t)nov);
  rT_.ensure((size_t)dm_.nov2);
  CCGPU_HIPX(hipMemcpyAsync(rt1_.p, t1, nov * sizeof(double), hipMemcpyDefault, st_));
  CCGPU_HIPX(hipMemcpyAsync(bufX2_.p, t2p, np2 * sizeof(double), hipMemcpyDefault, st_));
  unpack_sym(rT_.p, bufX2_.p, nov, st_);
  residual(rt1_.p, rT_.p, om1_.p, Om_.p);
  CCGPU_HIPX(hipMemcpyAsync(om1, om1_.p, nov * sizeof(double), hipMemcpyDefault, st_));
  pack_sym(bufX2_.p, Om_.p, nov, st_);
  CCGPU_HIPX(hipMemcpyAsync(om2p, bufX2_.p, np2 * sizeof(double), hipMemcpyDefault, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
}

}  // namespace ccgpu
