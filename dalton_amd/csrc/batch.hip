// Many independent closed-shell CCSD systems of uneven size converged
// together.  Each system keeps its own CCSD state and stream.  One iteration
// of every still-active system is captured once into a HIP graph: each system
// is a branch forked from and joined back to the master stream, so the GPU
// overlaps the small systems' kernels.  The graph is replayed once per
// iteration; its only host traffic is one copy of the per-system
// [norm^2, energy, singular flag, cycle] reports.  When systems converge (or
// switch precision) the graph is captured again for the remaining set.
//
// Before the first capture an autotuning pass times the FP64 GEMM launch
// configurations of every residual shape (throwaway results, then a fresh MP2
// start); the captured launches use the fastest configuration per shape.
#include "batch.h"

#include <algorithm>
#include <cmath>

#include "kernels.h"

namespace ccgpu {

Batch::Batch(int n, const int* o, const int* v, const int* naux, const double* const* eps,
             const double* const* B, int mxdiis, double ladder_budget) {
  CCGPU_REQUIRE(n > 0, "ccgpu_batch: no systems");
  sys_.reserve(n);
  for (int k = 0; k < n; ++k) {
    sys_.emplace_back(new CCSD(o[k], v[k], naux[k], eps[k], B[k], mxdiis));
    // default: room for 16 a-rows of the ladder tile block, 256 MB .. 10 GB
    const double vk = v[k];
    const double auto_budget = std::min(10.0e9, std::max(2.56e8, 128.0 * vk * vk * vk));
    sys_.back()->set_ladder_budget(ladder_budget > 0.0 ? ladder_budget : auto_budget);
  }
  CCGPU_HIPX(hipStreamCreateWithFlags(&master_, hipStreamNonBlocking));
  report_.alloc(4 * (size_t)n);
  host_report_.assign(4 * (size_t)n, 0.0);
  events_.resize(n + 1, nullptr);
  if (!dryrun())
    for (auto& e : events_) CCGPU_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

Batch::~Batch() {
  drop_graph();
  if (dryrun()) return;
  (void)hipStreamSynchronize(master_);
  for (auto& e : events_)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(master_);
}

void Batch::drop_graph() {
  if (dryrun()) return;
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

// Fork every active system from the master stream, enqueue one iteration on
// its own stream, join back.
void Batch::enqueue_active(const std::vector<int>& active) {
  CCGPU_HIPX(hipEventRecord(events_[0], master_));
  for (int k : active) {
    CCSD& s = *sys_[k];
    CCGPU_HIPX(hipStreamWaitEvent(s.stream(), events_[0], 0));
    s.enqueue_iteration(report_.p + 4 * (long)k);
    CCGPU_HIPX(hipEventRecord(events_[k + 1], s.stream()));
    CCGPU_HIPX(hipStreamWaitEvent(master_, events_[k + 1], 0));
  }
}

void Batch::capture(const std::vector<int>& active) {
  drop_graph();
  if (dryrun()) return;
  CCGPU_HIP(hipStreamBeginCapture(master_, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_active(active);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(master_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  CCGPU_HIP(hipStreamEndCapture(master_, &graph_));
  CCGPU_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  ++captures_;
}

// Autotuning pass: one throwaway FP64 residual per system at its MP2
// amplitudes, every new GEMM shape timed in each launch configuration.
void Batch::tune() {
  autotune_enable(true);
  try {
    for (auto& s : sys_) {
      s->set_precision(GemmPrec::F64);
      s->mp2_start();
      s->residual_current();
    }
  } catch (...) {
    autotune_enable(false);
    throw;
  }
  autotune_enable(false);
  CCGPU_HIPX(hipDeviceSynchronize());
}

void Batch::solve(int maxit, double threnr, double thrvec, GemmPrec prec, bool use_graph, bool tune,
                  double* e_mp2, double* e_ccsd, int* niter) {
  const int n = (int)sys_.size();
  CCGPU_REQUIRE(maxit > 0, "ccgpu_batch_energy: maxit must be positive");
  captures_ = 0;
  replays_ = 0;
  if (tune) this->tune();
  std::vector<double> en(n), emp2(n);
  std::vector<int> it(n, 0);
  std::vector<PrecisionSwitch> ps(n, PrecisionSwitch(prec == GemmPrec::F64));
  std::vector<char> done(n, 0);
  for (int k = 0; k < n; ++k) {
    emp2[k] = sys_[k]->mp2_start();
    en[k] = emp2[k];
    sys_[k]->prepare_iteration();
    if (e_mp2) e_mp2[k] = emp2[k];
  }
  CCGPU_HIPX(hipDeviceSynchronize());
  // Reduced precision moves to FP64 per system as in ccgpu_ccsd_energy.
  std::vector<int> active;
  for (int k = 0; k < n; ++k) active.push_back(k);
  bool dirty = true;
  // A round that launches shapes or precisions not seen before runs eagerly:
  // it sizes the per-stream scratch, which must not grow inside a capture.
  bool fresh = true;
  for (int round = 1; !active.empty(); ++round) {
    for (int k : active) sys_[k]->set_precision(ps[k].fp64 ? GemmPrec::F64 : prec);
    if (use_graph && !dryrun() && !fresh) {
      if (dirty || !exec_) capture(active);
      dirty = false;
      CCGPU_HIP(hipGraphLaunch(exec_, master_));
      ++replays_;
    } else {
      enqueue_active(active);
      dirty = true;
    }
    fresh = false;
    CCGPU_HIPX(hipMemcpyAsync(host_report_.data(), report_.p, report_.bytes(), hipMemcpyDeviceToHost, master_));
    CCGPU_HIPX(hipStreamSynchronize(master_));
    std::vector<int> next;
    for (int k : active) {
      sys_[k]->note_iteration();
      const double* r = host_report_.data() + 4 * (long)k;
      if (!dryrun() && r[2] != 0.0)
        throw Error(3, "DIIS: ZERO PIVOT - MATRIX SINGULAR (system " + std::to_string(k) + ")");
      const double e_old = en[k];
      en[k] = dryrun() ? 0.0 : r[1];
      const double nrm = dryrun() ? 0.0 : std::sqrt(r[0]);
      it[k] = round;
      const bool was_fp64 = ps[k].fp64;
      const double de = std::fabs(en[k] - e_old);
      if (ps[k].after(de < threnr && nrm < thrvec, de, nrm)) {
        done[k] = 1;
        dirty = true;
        continue;
      }
      if (ps[k].fp64 != was_fp64) {  // new precision: re-capture after an eager round
        dirty = true;
        fresh = true;
      }
      if (round + 1 > maxit) throw Error(4, "CC equations not converged (system " + std::to_string(k) + ")");
      next.push_back(k);
    }
    active.swap(next);
  }
  drop_graph();
  for (int k = 0; k < n; ++k) {
    if (e_ccsd) e_ccsd[k] = en[k];
    if (niter) niter[k] = it[k];
  }
}

}  // namespace ccgpu
