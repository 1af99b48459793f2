// Left transformation rho = lambda A of the closed-shell CCSD Jacobian
// (CC_LHTR, cc/cc_lhtr.F:20), written as the hand-derived adjoint (reverse
// mode) of the T1-dressed residual in ccsd.hip: every forward contraction
// C = A B contributes the two products A^T C_bar and C_bar B^T, in reverse
// order, on the same MFMA GEMM kernels and layouts.  Amplitudes and trial
// vectors use Dalton's packed storage, so
//   rho_mu = sum_nu lambda_nu d Omega_nu / d t_mu
// with packed dot products on both sides: the packed doubles cotangent
// becomes the full matrix  Lt = (Lf + diag(Lf)) / 2  (Lf the symmetric
// unpacking), and the full-matrix amplitude gradient Tb folds back as
// rho2[ai<=bj] = Tb[ai][bj] + Tb[bj][ai] (ai < bj), Tb[ai][ai] (diagonal).
//
// Cotangent arrays carry the suffix b and the forward layouts:
//   Xb = 2 Lt = Lt + Lt^T (the cotangent of X2, since Om = ... + X2 + X2^T)
//   Tb, Ub, Txb [(i,a),(j,b)]; Bvvb [P][(a,c)]; Boob [P][(k,i)];
//   BvoTb [P][(i,a)]; Btb [P][p][q]; Fb [p][q] (also the cotangent of h~).
#include <algorithm>

#include "ccsd.h"
#include "kernels.h"
#include "ops.h"

namespace ccgpu {

void CCSD::ensure_lhtr() {
  if (Tb_.p) return;
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long N = dm_.N, nov = dm_.nov, nov2 = dm_.nov2;
  Tb_.alloc(nov2);
  Ub_.alloc(nov2);
  Txb_.alloc(nov2);
  Db_.alloc(nov2);
  Btb_.alloc((size_t)naux * N * N);
  Bvvb_.alloc((size_t)naux * v * v);
  Boob_.alloc((size_t)naux * o * o);
  BvoTb_.alloc((size_t)naux * nov);
  Zb_.alloc((size_t)naux * nov);
  Wb_.alloc((size_t)o * o * nov);
  Fb_.alloc(N * N);
  E1b_.alloc((size_t)v * v);
  E2b_.alloc((size_t)o * o);
  Jdb_.alloc(naux);
  Gamb_.alloc((size_t)o * o * o * o);
  G1b_.alloc((size_t)o * o * o * o);
  taupb_.alloc(dm_.npo * dm_.npv);
  taumb_.alloc(dm_.npo * dm_.npv);
  vecb_.alloc(2 * nov);
}

// Forward pass at the current amplitudes: leaves Bt, Bvv, Boo, BvoT, F, h~,
// Jd, Tx, U, C, D, Gam, E1, E2, Z, W, F_kc and tau+/- on the device.
void CCSD::lhtr_prepare() {
  if (lhtr_ready_ && lhtr_prec_ == bulk_) return;
  if (diisS_.p && !Tb_.p) {
    // The DIIS history stays when the cotangent buffers and the trial-vector
    // staging fit beside it (with 4 GB to spare); otherwise it is released,
    // as the converged ground state needs none (a later iterate restarts
    // the DIIS cycle).
    const int o = dm_.o, v = dm_.v, naux = dm_.naux;
    const double N = (double)dm_.N, nov = (double)dm_.nov, oo = (double)o * o;
    const double need = 8.0 * (4.0 * dm_.nov2 + naux * (N * N + (double)v * v + oo + 2.0 * nov + 1.0) +
                               oo * nov + N * N + (double)v * v + oo + 2.0 * oo * oo +
                               2.0 * dm_.npo * dm_.npv + 2.0 * nov + 4.0 * dm_.ntamp) + 4.0e9;
    size_t fr = 0, tot = 0;
    if (!dryrun()) CCGPU_HIP(hipMemGetInfo(&fr, &tot));
    if (!dryrun() && (double)fr < need) {
      diisS_.release();
      diisE_.release();
    }
  }
  residual(t1_.p, T_.p, om1_.p, Om_.p);
  ensure_lhtr();
  lhtr_ready_ = true;
  lhtr_prec_ = bulk_;
}

// Adjoint of the particle ladder (CCSD::ladder): per a-batch the V+/- block
// is rebuilt, S+/-_bar gathered from the doubles cotangent, then
//   tau+/-_bar += S_bar^T V,  V_bar = S_bar tau,  Y_bar = adjoint(V+/-),
//   Bvv_bar += Bvv Y_bar^T + Bvv Y_bar  (both factors of Y = Bvv^T Bvv).
// Xb holds 2 Lt, hence the factors 1/2.
void CCSD::lhtr_ladder(const double* Xb) {
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long vv = (long)v * v, npo = dm_.npo, npv = dm_.npv;
  const Gemm bulk{bulk_, st_};
  fill(taupb_.p, 0.0, npo * npv, st_);
  fill(taumb_.p, 0.0, npo * npv, st_);
  for (size_t bi = 0; bi + 1 < lad_batches_.size(); ++bi) {
    const int a0 = lad_batches_[bi], a1 = lad_batches_[bi + 1];
    long rows = 0;
    for (int a = a0; a < a1; ++a) rows += v - a;
    const long M = (long)(a1 - a0) * v, Nn = (long)(v - a0) * v;
    const double* Bb = Bvv_.p + (long)a0 * v;
    bulk(M, Nn, naux, 1.0, Bb, 1, vv, Bb, vv, 1, 0.0, Y_.p, Nn, 1);
    ladder_vpm(Vp_.p, Vm_.p, Y_.p, Nn, a0, a1, v, st_);
    ladder_gather(Sp_.p, Sm_.p, Xb, a0, a1, o, v, st_);
    bulk(npo, npv, rows, 0.5, Sp_.p, 1, npo, Vp_.p, npv, 1, 1.0, taupb_.p, npv, 1);
    bulk(npo, npv, rows, 0.5, Sm_.p, 1, npo, Vm_.p, npv, 1, 1.0, taumb_.p, npv, 1);
    bulk(rows, npv, npo, 0.5, Sp_.p, npo, 1, taup_.p, npv, 1, 0.0, Vp_.p, npv, 1);
    bulk(rows, npv, npo, 0.5, Sm_.p, npo, 1, taum_.p, npv, 1, 0.0, Vm_.p, npv, 1);
    fill(Y_.p, 0.0, M * Nn, st_);  // entries with b < a stay zero
    ladder_vpm_adj(Y_.p, Nn, Vp_.p, Vm_.p, a0, a1, v, st_);
    double* Bbb = Bvvb_.p + (long)a0 * v;
    bulk(naux, M, Nn, 1.0, Bb, vv, 1, Y_.p, 1, Nn, 1.0, Bbb, vv, 1);
    bulk(naux, Nn, M, 1.0, Bb, vv, 1, Y_.p, Nn, 1, 1.0, Bbb, vv, 1);
  }
}

// One trial vector: l1 [i][a] and Xb = 2 Lt on the device; rho1 [i][a] and
// the symmetric full-matrix doubles result rho2s (Tb + Tb^T with the
// diagonal taken once) are written.
void CCSD::lhtr_apply(const double* l1, double* Xb, double* rho1, double* rho2s) {
  const Gemm bulk{bulk_, st_};
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long N = dm_.N, NN = N * N, nov = dm_.nov, nov2 = dm_.nov2, oo = (long)o * o, vv = (long)v * v;
  const double* T = T_.p;
  const double* t1 = t1_.p;
  const double* Tx = bufTx_.p;
  const double* U = bufU_.p;
  const double* C = bufC_.p;
  const double* D = bufD_.p;
  double *Tb = Tb_.p, *Ub = Ub_.p, *Txb = Txb_.p, *Fb = Fb_.p;
  double* fovb = vecb_.p + nov;
  mark("lhtr_start");

  // singles: om1 = F~_ai + sum U F_kc + Bvv . Z - W . U  (CCRHS_G/H/I/J)
  fill(Fb, 0.0, NN, st_);
  add_strided3(Fb + o * N, 1, N, 0, l1, o, v, 1, 1.0, st_);                    // F~_ai
  g(nov, nov, 1, 1.0, l1, 1, 0, hx_.p, 0, 1, 0.0, Ub, nov, 1, st_);            // Ub = l1 (x) F_kc
  gemv_mc(nov, nov, 1.0, U, nov, l1, 0.0, fovb, st_);                          // F_kc bar = U^T l1
  add_strided3(Fb + o, N, 1, 0, fovb, o, v, 1, 1.0, st_);
  g(v, nov, o, -1.0, l1, 1, v, W_.p, nov, 1, 1.0, Ub, nov, 1, st_, o, 0, (long)o * nov, (long)v * nov);
  g(o, nov, v, -1.0, l1, v, 1, U, nov, 1, 0.0, Wb_.p, nov, 1, st_, o, 0, (long)v * nov, (long)o * nov);
  g(naux, oo, nov, 1.0, Bov_.p, nov, 1, Wb_.p, 1, nov, 0.0, Boob_.p, oo, 1, st_);  // W = Boo^T Bov
  g(o, v, v, 1.0, l1, v, 1, Bvv_.p, v, 1, 0.0, Zb_.p, v, 1, st_, naux, 0, vv, nov);
  g(v, v, o, 1.0, l1, 1, v, Z_.p, v, 1, 0.0, Bvvb_.p, v, 1, st_, naux, 0, nov, vv);
  bulk(nov, nov, naux, 1.0, Bov_.p, 1, nov, Zb_.p, nov, 1, 1.0, Ub, nov, 1);    // Z = Bov U
  mark("lhtr_singles");

  // E terms: X2 += T E1^T - E2^T T  (CCRHS_E)
  g(nov * o, v, v, 1.0, Xb, v, 1, E1_.p, v, 1, 0.0, Tb, v, 1, st_);
  g(v, v, nov * o, 1.0, Xb, 1, v, T, v, 1, 0.0, E1b_.p, v, 1, st_);
  g(o, v, o, -1.0, E2_.p, o, 1, Xb, v, 1, 1.0, Tb, v, 1, st_, (int)nov, 0, nov, nov);
  g(o, o, nov * v, -1.0, T, v, 1, Xb, 1, v, 0.0, E2b_.p, o, 1, st_, 1, 0, 0, 0, v, nov, nov);
  add_strided3(Fb + o * N + o, N, 1, 0, E1b_.p, v, v, 1, 1.0, st_);
  g(v, nov, v, -1.0, E1b_.p, v, 1, J_.p, nov, 1, 1.0, Ub, nov, 1, st_, o, 0, (long)v * nov, (long)v * nov);
  add_strided3(Fb, N, 1, 0, E2b_.p, o, o, 1, 1.0, st_);
  g(o, (long)v * nov, o, 1.0, E2b_.p, 1, o, J_.p, (long)v * nov, 1, 1.0, Ub, (long)v * nov, 1, st_);
  mark("lhtr_E");

  // rings: P = C Tx, X2 = -P/2 - X(P) + D U / 2, C = Cint - Tx K / 2,
  // D = -Cint + 2 BvoT^T Bov + U L / 2  (CCRHS_C/D)
  double* Pb = bufR_.p;
  double* Cb = bufX2_.p;
  double* Db = Db_.p;
  perm(Pb, Xb, 1, 1, 1, nov2, 0, 0, 0, 1, -0.5, 0.0, st_);
  perm(Pb, Xb, o, v, o, v, v, nov, (long)v * nov, 1, -1.0, 1.0, st_);
  bulk(nov, nov, nov, 1.0, Pb, nov, 1, Tx, nov, 1, 0.0, Cb, nov, 1);
  bulk(nov, nov, nov, 1.0, C, 1, nov, Pb, nov, 1, 0.0, Txb, nov, 1);
  bulk(nov, nov, nov, -0.5, Cb, nov, 1, K_.p, nov, 1, 1.0, Txb, nov, 1);
  bulk(nov, nov, nov, 0.5, Xb, nov, 1, U, nov, 1, 0.0, Db, nov, 1);
  bulk(nov, nov, nov, 0.5, D, 1, nov, Xb, nov, 1, 1.0, Ub, nov, 1);
  bulk(naux, nov, nov, 2.0, Bov_.p, nov, 1, Db, 1, nov, 0.0, BvoTb_.p, nov, 1);
  bulk(nov, nov, nov, 0.5, Db, nov, 1, L_.p, nov, 1, 1.0, Ub, nov, 1);
  // Cint[(i,a),(k,c)] = R[(a,c),(k,i)], R = Bvv^T Boo
  double* Rb = bufR_.p;
  perm(Rb, Cb, v, v, o, o, nov, 1, v, (long)v * nov, 1.0, 0.0, st_);
  perm(Rb, Db, v, v, o, o, nov, 1, v, (long)v * nov, -1.0, 1.0, st_);
  bulk(naux, vv, oo, 1.0, Boo_.p, oo, 1, Rb, 1, oo, 1.0, Bvvb_.p, vv, 1);
  bulk(naux, oo, vv, 1.0, Bvv_.p, vv, 1, Rb, oo, 1, 1.0, Boob_.p, oo, 1);
  // U = 2 T - Tx, Tx = X(T)
  axpy(Tb, 2.0, Ub, nov2, st_);
  axpy(Txb, -1.0, Ub, nov2, st_);
  perm(Tb, Txb, o, v, o, v, v, nov, (long)v * nov, 1, 1.0, 1.0, st_);
  mark("lhtr_rings");

  // A2 integral term Om = BvoT^T BvoT
  bulk(naux, nov, nov, 1.0, BvoT_.p, nov, 1, Xb, nov, 1, 1.0, BvoTb_.p, nov, 1);
  mark("lhtr_A2_int");

  lhtr_ladder(Xb);
  double* Toob = Db_.p;
  fill(Toob, 0.0, nov2, st_);
  ladder_tau_adj(Toob, taupb_.p, taumb_.p, o, v, st_);
  mark("lhtr_ladder");

  // hole ladder: Gam = perm(Boo^T Boo) + Joovv Too^T, LADh = Gam^T Too
  double* Too = bufR_.p;
  double* LADb = bufX2_.p;
  perm(Too, T, o, o, v, v, (long)v * nov, v, nov, 1, 1.0, 0.0, st_);
  perm(LADb, Xb, o, o, v, v, (long)v * nov, v, nov, 1, 0.5, 0.0, st_);
  bulk(oo, oo, vv, 1.0, Too, vv, 1, LADb, 1, vv, 0.0, Gamb_.p, oo, 1);
  bulk(oo, vv, oo, 1.0, Gam_.p, oo, 1, LADb, vv, 1, 1.0, Toob, vv, 1);
  bulk(oo, vv, oo, 1.0, Gamb_.p, 1, oo, Joovv_.p, vv, 1, 1.0, Toob, vv, 1);
  perm(G1b_.p, Gamb_.p, o, o, o, o, (long)o * oo, o, oo, 1, 1.0, 0.0, st_);  // [(k,i),(l,j)]
  g(naux, oo, oo, 1.0, Boo_.p, oo, 1, G1b_.p, 1, oo, 1.0, Boob_.p, oo, 1, st_);
  g(naux, oo, oo, 1.0, Boo_.p, oo, 1, G1b_.p, oo, 1, 1.0, Boob_.p, oo, 1, st_);
  perm(Tb, Toob, o, v, o, v, (long)o * vv, v, vv, 1, 1.0, 1.0, st_);
  mark("lhtr_hole_ladder");

  // dressed Fock matrix F = h~ + sum_P Jd[P] Bt[P] - sum_{P,k} Bt[P][p][k] Bt[P][k][q]
  double* Btb = Btb_.p;
  const double* Bt = Bt_.p;
  g(naux, NN, 1, 1.0, Jd_.p, 1, 0, Fb, 0, 1, 0.0, Btb, NN, 1, st_);
  g(naux, 1, NN, 1.0, Bt, NN, 1, Fb, 1, 0, 0.0, Jdb_.p, 1, 0, st_);
  add_occ_diag(Btb, Jdb_.p, naux, (int)N, o, 2.0, st_);
  g(N, o, N, -1.0, Fb, N, 1, Bt, 1, N, 1.0, Btb, N, 1, st_, naux, 0, NN, NN);
  g(o, N, N, -1.0, Bt, 1, N, Fb, N, 1, 1.0, Btb, N, 1, st_, naux, NN, 0, NN);
  add_strided3(Btb + o * N + o, NN, N, 1, Bvvb_.p, naux, v, v, 1.0, st_);
  add_strided3(Btb, NN, N, 1, Boob_.p, naux, o, o, 1.0, st_);
  add_strided3(Btb + o * N, NN, 1, N, BvoTb_.p, naux, o, v, 1.0, st_);

  // T1 dressing (LAMMAT): Bt[P][o+a][q] -= sum_l t(l,a) B[P][l][q], then
  // Bt[P][p][i] += sum_a Bt[P][p][o+a] t(i,a); the same for h.
  g(o, v, (long)naux * N, 1.0, Btb, 1, N, Bt + o, N, 1, 0.0, rho1, v, 1, st_);
  g((long)naux * N, v, o, 1.0, Btb, N, 1, t1, v, 1, 1.0, Btb + o, N, 1, st_);
  g(o, v, (long)naux * N, -1.0, B_.p, N, 1, Btb + o * N, 1, N, 1.0, rho1, v, 1, st_, 1, 0, 0, 0, N, NN, NN);
  g(o, v, N, 1.0, Fb, 1, N, ht_.p + o, N, 1, 1.0, rho1, v, 1, st_);
  g(N, v, o, 1.0, Fb, N, 1, t1, v, 1, 1.0, Fb + o, N, 1, st_);
  g(o, v, N, -1.0, h_.p, N, 1, Fb + o * N, 1, N, 1.0, rho1, v, 1, st_);
  mark("lhtr_fock_dressing");

  // packed-storage fold of the doubles gradient
  perm(rho2s, Tb, 1, 1, 1, nov2, 0, 0, 0, 1, 1.0, 0.0, st_);
  perm(rho2s, Tb, 1, 1, nov, nov, 0, 0, 1, nov, 1.0, 1.0, st_);
  scale_diag(rho2s, nov, nov, 0.5, st_);
  mark("lhtr_fold");
}

// Vectors are pipelined: while vector k is transformed on the compute
// stream, vector k+1 is copied in and the result of vector k-1 copied out
// on a copy stream, through two sets of device staging buffers.
void CCSD::lhtr_host(int nvec, const double* l1, const double* l2p, double* rho1, double* rho2p, bool dalton) {
  CCGPU_REQUIRE(nvec >= 0, "ccgpu_lhtr: nvec < 0");
  CCGPU_REQUIRE(iter_ >= 1, "ccgpu_lhtr: no amplitudes (run the CCSD iterations or set amplitudes first)");
  if (nvec == 0) return;
  const long nov = dm_.nov, np2 = dm_.ntamp - nov;
  lhtr_prepare();
  if (!lst_.p) lst_.alloc(4 * (size_t)(np2 + nov));  // [in 0, in 1, out 0, out 1] x (l1/rho1, l2/rho2)
  side_stream();
  const long slot = np2 + nov;
  auto in1 = [&](int q) { return lst_.p + q * slot; };
  auto in2 = [&](int q) { return lst_.p + q * slot + nov; };
  auto out1 = [&](int q) { return lst_.p + (2 + q) * slot; };
  auto out2 = [&](int q) { return lst_.p + (2 + q) * slot + nov; };
  auto copy_in = [&](int iv) {
    const int q = iv & 1;
    CCGPU_HIPX(hipMemcpyAsync(in1(q), l1 + iv * nov, nov * sizeof(double), hipMemcpyDefault, cst_));
    CCGPU_HIPX(hipMemcpyAsync(in2(q), l2p + iv * np2, np2 * sizeof(double), hipMemcpyDefault, cst_));
    CCGPU_HIPX(hipEventRecord(ev_in_[q], cst_));
  };
  auto compute = [&](int iv) {
    const int q = iv & 1;
    CCGPU_HIPX(hipStreamWaitEvent(st_, ev_in_[q], 0));
    unpack_sym(Om_.p, in2(q), nov, st_);
    // Xb = 2 Lt: Lf + diag(Lf) for packed dots; the squared vector itself in
    // Dalton's convention (S halves the diagonal, then the same 2 Lt)
    if (!dalton) scale_diag(Om_.p, nov, nov, 2.0, st_);
    lhtr_apply(in1(q), Om_.p, out1(q), bufX2_.p);
    if (dalton) scale_diag(bufX2_.p, nov, nov, 2.0, st_);  // D on the result
    pack_sym(out2(q), bufX2_.p, nov, st_);
    CCGPU_HIPX(hipEventRecord(ev_done_[q], st_));
  };
  auto copy_out = [&](int iv) {
    const int q = iv & 1;
    CCGPU_HIPX(hipStreamWaitEvent(cst_, ev_done_[q], 0));
    CCGPU_HIPX(hipMemcpyAsync(rho1 + iv * nov, out1(q), nov * sizeof(double), hipMemcpyDefault, cst_));
    CCGPU_HIPX(hipMemcpyAsync(rho2p + iv * np2, out2(q), np2 * sizeof(double), hipMemcpyDefault, cst_));
    CCGPU_HIPX(hipStreamSynchronize(cst_));  // slot q (in and out) is free again
  };
  copy_in(0);
  compute(0);
  for (int iv = 0; iv < nvec; ++iv) {
    if (iv + 1 < nvec) {
      copy_in(iv + 1);  // slot of iv+1 was released by copy_out(iv-1)
      compute(iv + 1);
    }
    copy_out(iv);
  }
  CCGPU_HIPX(hipStreamSynchronize(st_));
}

}  // namespace ccgpu
