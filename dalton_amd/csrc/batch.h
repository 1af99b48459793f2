// Independent closed-shell CCSD systems converged together, one iteration of
// all active systems per HIP graph replay.
#pragma once
#include <memory>
#include <vector>

#include "ccsd.h"

namespace ccgpu {

class Batch {
 public:
  // ladder_budget: bytes per particle-ladder batch buffer of each system
  // (<= 0: 128 v^3 bytes, room for 16 a-rows, clamped to 256 MB .. 10 GB).
  Batch(int n, const int* o, const int* v, const int* naux, const double* const* eps, const double* const* B,
        int mxdiis, double ladder_budget);
  ~Batch();

  // Converge every system from its MP2 start with Dalton's test
  // (|dE| < threnr and norm < thrvec), as ccgpu_ccsd_energy does for one.
  // use_graph: replay a captured graph per iteration (else eager launches on
  // the same per-system streams).  tune: autotune the FP64 GEMM shapes first.
  void solve(int maxit, double threnr, double thrvec, GemmPrec prec, bool use_graph, bool tune, double* e_mp2,
             double* e_ccsd, int* niter);

  // Autotuning pass alone (the one solve() runs first when tune is set).
  void tune();

  int size() const { return (int)sys_.size(); }
  CCSD& system(int k) { return *sys_[k]; }
  int captures() const { return captures_; }
  int replays() const { return replays_; }

 private:
  void enqueue_active(const std::vector<int>& active);
  void capture(const std::vector<int>& active);
  void drop_graph();

  std::vector<std::unique_ptr<CCSD>> sys_;
  hipStream_t master_ = nullptr;
  std::vector<hipEvent_t> events_;  // [0] fork, [k+1] join of system k
  DBuf<double> report_;             // [k][norm^2, energy, singular flag, cycle]
  std::vector<double> host_report_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  int captures_ = 0, replays_ = 0;
};

}  // namespace ccgpu
