// Closed-shell DF-CCSD on one GPU: device-resident amplitudes, T1-dressed
// residual as MFMA GEMMs + layout kernels, Jacobi step, DIIS with Dalton's
// restart rule, energy.
#pragma once
#include <vector>

#include "common.h"

namespace ccgpu {

// When a CCSD loop in reduced precision (BF16x3 or FP32 bulk contractions)
// moves to FP64: once |dE| < kSwitch and the residual norm < sqrt(kSwitch).
// Convergence is accepted only after at least two FP64 iterations, which
// take the amplitudes to the FP64 fixed point; an FP64 loop accepts it at
// once.  (The reduced-precision iteration converges to the fixed point of
// a Hamiltonian perturbed by the rounding of its operands, about 3e-7 Eh
// from the FP64 one at o=64 v=512: iterating it further, until its own test
// is met, only costs time, and the FP64 iterations needed afterwards do not
// get fewer.)
struct PrecisionSwitch {
  static constexpr double kSwitch = 1e-6;
  bool fp64;
  bool reduced;
  int nfp64 = 0;
  explicit PrecisionSwitch(bool start_fp64) : fp64(start_fp64), reduced(!start_fp64) {}
  // After an iteration (run in FP64 iff fp64 was set) with energy change dE,
  // residual norm nrm and convergence test conv: true when the loop is
  // converged.  May set fp64 for the next iteration.
  bool after(bool conv, double dE, double nrm) {
    if (fp64) ++nfp64;
    if (conv && (!reduced || nfp64 >= 2)) return true;
    if (!fp64 && dE < kSwitch && nrm * nrm < kSwitch) fp64 = true;
    return false;
  }
};

struct Dims {
  int o = 0, v = 0, naux = 0;
  long N = 0, nov = 0, nov2 = 0;
  long npo = 0, npv = 0;  // o(o+1)/2, v(v+1)/2
  long ntamp = 0;         // nov + nov(nov+1)/2 (Dalton NT1AMX + NT2AMX)
};

class CCSD {
 public:
  CCSD(int o, int v, int naux, const double* eps_host, const double* B_host, int mxdiis);
  ~CCSD();

  const Dims& dims() const { return dm_; }
  hipStream_t stream() const { return st_; }

  // Arithmetic of the bulk contractions (ladder, rings, hole ladder, DF
  // integrals, (T) W): FP64, or BF16x3 split products for the mixed path.
  void set_precision(GemmPrec p) { bulk_ = p; }
  GemmPrec precision() const { return bulk_; }

  // CCRHSN mirror on device arrays: t1 [i][a], T [(i,a),(j,b)] -> om1, Om.
  void residual(const double* t1, const double* T, double* om1, double* Om);
  // residual of the current amplitudes into the internal Omega buffers
  void residual_current() { residual(t1_.p, T_.p, om1_.p, Om_.p); }

  // Device-resident iteration (CCSD_ENERGY loop body).
  double mp2_start();  // t1 = 0, t2 = MP2; returns E_MP2 (correlation)
  // One iteration: residual, Jacobi step, DIIS, energy.  Returns the new
  // correlation energy; *omnm = 2-norm of the incoming residual.
  double iterate(double* omnm);
  // The same iteration split for graph capture: enqueue issues every launch
  // on stream() with no host synchronisation or host-dependent argument
  // (report, if given, receives [norm^2, energy, singular flag, cycle count]
  // on the device); finish_iteration reads the result back afterwards.
  // prepare_iteration sizes every buffer an iteration needs (call eagerly).
  void prepare_iteration();
  void enqueue_iteration(double* report);
  double finish_iteration(double* omnm);
  // Host bookkeeping after an iteration enqueued elsewhere (graph replay).
  void note_iteration() { ++iter_; lhtr_ready_ = false; }
  // Bytes per particle-ladder batch buffer (V+/-, Y); applies from the next
  // workspace allocation.
  void set_ladder_budget(double bytes) { ladder_budget_ = bytes; }
  int iteration() const { return iter_; }

  // Host <-> device amplitude exchange in Dalton's layouts.
  void set_amplitudes(const double* t1_host, const double* t2_packed_host);
  void get_amplitudes(double* t1_host, double* t2_packed_host);

  // Correlation energy of the current amplitudes (CCSD_ECCSD).
  double energy_current();
  // CCSD_ECCSD and CCSD_DIIS on host arrays in Dalton's layouts, for a
  // caller that keeps its own iteration loop: the energy of (t1, t2p); one
  // DIIS step from the new estimate s to the extrapolated vector, written
  // over t (NTAMP values each), with the history held on the device (shared
  // with iterate()).  diis_host returns false if the DIIS matrix was
  // singular (t then holds the newest estimate).
  double energy_host(const double* t1, const double* t2p);
  bool diis_host(const double* s, double* t, int niter);

  // residual with host arrays (Dalton layouts), used by ccgpu_rhs; leaves the
  // context's amplitudes, iteration count and DIIS state untouched
  void residual_host(const double* t1, const double* t2p, double* om1, double* om2p);

  // The E intermediates of the residual at the current amplitudes
  // (CCRHS_EFCK/CCRHS_E, cc/ccsd_rhs.F:8604,2280) in Dalton's column-major
  // order: e1[b + v*c] = E1(b,c) (NMATAB values), e2[k + o*j] = E2(k,j)
  // (NMATIJ values).  Runs one residual; the amplitudes are not changed.
  void e_intermediates_host(double* e1, double* e2);
  // The hole-ladder matrix Gamma of CCRHS_A at the current amplitudes, in
  // Dalton's packed NGAMMA order (CC_GAMIM); runs one residual.
  void gamma_host(double* gamma);

  // (T) energy over the restricted triple list i >= j >= k (i = j = k
  // excluded), enumerated i-major; triples first, first+stride, ... (count
  // of them).  Adds E4/E5 of those triples; *flops gets their algorithmic
  // flop count.  Releases the iteration workspace on first use.
  long pt_ntriples() const;
  void pt_energy(long first, long count, long stride, double* e4, double* e5, double* flops);
  void pt_release();
  void release_work();

  // Left transformation rho = lambda A at the current amplitudes (CC_LHTR):
  // nvec trial vectors in Dalton's layouts (host arrays, vector-major).  The
  // forward intermediates are formed once per amplitude set and precision
  // and reused for every vector; the DIIS history is released first unless
  // the transformation's buffers fit beside it.
  // dalton: Dalton's packed conventions for left vectors (CC_TRDRV,
  // cc/cc_trdrv.F:484-486, and the left product of CCLR_DUMTRR on the
  // finite-difference Jacobian of CCLR_FDJAC, cc/cclr_util.F:1228,1343):
  // the packed doubles diagonal of the trial vector enters the squared
  // vector unscaled and the result's diagonal is twice the packed-gradient
  // one, rho_D(L) = D rho(S L) with S (D) halving (doubling) the doubles
  // diagonal; otherwise rho_mu = sum_nu l_nu dOmega_nu/dt_mu with packed
  // dot products on both sides.
  void lhtr_host(int nvec, const double* l1, const double* l2p, double* rho1, double* rho2p, bool dalton = false);

  // Ground-state multipliers (lambda.hip): eta = dE/dt (CC_ETA), the
  // reduced-space solution of lambda A + eta = 0 (CCEQ_SOL from the left;
  // returns the iteration count, also in *niter, and throws code 4 if not
  // converged, lambda and *niter still set) and the
  // unrelaxed MO one-electron density D[p][q] = dL/dh[p][q] (CC_D1AO).
  // Host arrays in the layouts of lhtr_host.
  // dalton: eta_D = D eta (CC_ETA, cc/cc_fop.F:2955), lambda_D = D lambda
  // (the solution of rho_D(lambda_D) + eta_D = 0), onedens reads lambda_D.
  void eta_host(double* e1, double* e2p, bool dalton = false);
  int lambda_solve(double thr, int maxit, int maxred, double* l1, double* l2p, double* resnorm, int* niter,
                   bool dalton = false);
  void onedens_host(const double* l1, const double* l2p, double* D, bool dalton = false);
  // Transformations of the last lambda_solve: [0] FP64, [1] BF16x3.
  const int* lambda_counts() const { return lam_counts_; }

  const double* t1() const { return t1_.p; }
  const double* T() const { return T_.p; }
  const double* eps() const { return eps_.p; }
  const double* B() const { return B_.p; }
  const double* J() const { return J_.p; }

 private:
  void dress(const double* t1);
  void ladder(const double* T, double* Om);
  void ladder_mixed(double* Om);
  void mark(const char* name);
  void ensure_work();
  // (T) drivers: per-triple launches of the pre-packed three-pass W build and
  // blocked energy (default); the same as one persistent kernel; the first
  // per-triple driver (twelve GEMMs, per-triple packing, transposed W copies)
  enum class PtPath { Launch = 0, Persistent = 1 };
  void pt_prepare(GemmPrec mode, PtPath path);
  void pt_energy_persistent(long first, long n, long stride, double* e4, double* e5);
  void pt_energy_launches(long first, long n, long stride, double* e4, double* e5);
  void pt_descs(int i, int j, int k, int slot, unsigned char* out);
  size_t pt_desc_bytes() const;
  void side_stream();
  void ensure_diis();
  void reset_cycle();
  void ensure_lhtr();
  void lhtr_prepare();
  void lhtr_apply(const double* l1, double* Xb, double* rho1, double* rho2s);
  void lhtr_ladder(const double* Xb);
  void lhtr_packed(const double* x, double* y);
  void eta_device(double* e);

  Dims dm_;
  int mxdiis_;
  int iter_ = 0;
  int diis_off_ = 0;  // DIIS cycle restarts after the history was released
  bool work_ = false;
  GemmPrec bulk_ = GemmPrec::F64;
  bool pt_ready_ = false;
  std::vector<int> tri_i_, tri_j_, tri_k_;
  hipStream_t st_ = nullptr;

  // inputs / constants
  DBuf<double> eps_, B_, h_, Bov_, J_, K_, L_, Joovv_;
  // dressing
  DBuf<double> Bt_, Bvv_, Boo_, BvoT_, F_, ht_, Jd_, G1_, Gam_, Z_, W_, E1_, E2_, hx_;
  // amplitudes / residual
  DBuf<double> t1_, T_, om1_, Om_;
  DBuf<double> rt1_, rT_;  // amplitudes handed to residual_host (ccgpu_rhs)
  DBuf<double> bufR_, bufX2_, bufTx_, bufU_, bufC_, bufD_;
  // ladder
  DBuf<double> taup_, taum_, Y_, Vp_, Vm_, Sp_, Sm_;
  std::vector<int> lad_batches_;
  // mixed-precision ladder: bf16 hi/lo of Bvv [(a,c)][P], tau+/- [s][ij][cd],
  // V+/- [s][r][cd]; S+/- [s][r][ij]
  DBuf<unsigned short> lpBh_, lpBl_, lpTh_, lpTl_, lpVh_, lpVl_;
  DBuf<double> S2_;
  // DIIS
  DBuf<double> diisS_, diisE_, dvec_, dots_, coef_, Amat_, diiswork_;
  DBuf<int> ctr_;  // [0] iteration of the DIIS cycle, [1] singular-matrix flag
  DBuf<double> scal_;  // [0] squared residual norm, [1] energy
  double ladder_budget_ = 10.0e9;
  // (T)
  GemmPrec pt_mode_ = GemmPrec::F64;
  DBuf<double> G_, Gp_, Tvvo_, Q_, Wt_, ptpart_, etrip_;
  // (T), mixed path: (yd|zk) packed as bf16 hi/lo [k][(y,z)][d] in both (y,z)
  // orders, per-triple packed slices, FP32 W
  DBuf<unsigned short> GTh_, GTl_, GpTh_, GpTl_;
  DBuf<float> W32_;  // four W slots (FP32)
  // (T), persistent kernel: bf16 hi/lo of every T(p,q) [x][d] and Q_rq [z][l]
  // slice and of -Tvvo_p in both (x,y) orders (mixed), Tvvo with (x,y)
  // swapped (FP64); four W slots (in W32_ / Wt_); launch tables
  PtPath pt_path_ = PtPath::Launch;
  DBuf<unsigned short> ppH_[4], ppL_[4];
  DBuf<double> TvvoS_;
  DBuf<int> pttri_, ptctab_;
  DBuf<unsigned long long> ptclaim_;
  DBuf<unsigned char> ptdesc_;
  // left transformation: cotangents (suffix b) of the residual's intermediates
  int lam_counts_[2] = {0, 0};
  bool lhtr_ready_ = false;
  GemmPrec lhtr_prec_ = GemmPrec::F64;
  DBuf<double> Tb_, Ub_, Txb_, Db_, Btb_, Bvvb_, Boob_, BvoTb_, Zb_, Wb_, Fb_, E1b_, E2b_, Jdb_, Gamb_, G1b_,
      taupb_, taumb_, vecb_;
  // trial-vector staging (two in, two out slots) and the copy stream
  DBuf<double> lst_;
  hipStream_t cst_ = nullptr;
  hipEvent_t ev_in_[2] = {nullptr, nullptr}, ev_done_[2] = {nullptr, nullptr};
};

}  // namespace ccgpu
