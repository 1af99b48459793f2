// extern "C" entry points of libccgpu (declared in include/ccgpu.h).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ccgpu.h"
#include "batch.h"
#include "ccsd.h"

struct ccgpu_ctx {
  std::unique_ptr<ccgpu::CCSD> cc;
};

struct ccgpu_batch {
  std::unique_ptr<ccgpu::Batch> b;
};

namespace {
thread_local std::string g_err;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const ccgpu::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
}

ccgpu::GemmPrec prec_of(int precision) {
  CCGPU_REQUIRE(precision == CCGPU_FP64 || precision == CCGPU_MIXED || precision == CCGPU_FP32,
                "unknown precision selector");
  return precision == CCGPU_MIXED  ? ccgpu::GemmPrec::BF16X3
         : precision == CCGPU_FP32 ? ccgpu::GemmPrec::F32
                                   : ccgpu::GemmPrec::F64;
}
// (T) runs its W passes in FP64 or as BF16x3 products only.
ccgpu::GemmPrec pt_prec_of(int precision) {
  CCGPU_REQUIRE(precision == CCGPU_FP64 || precision == CCGPU_MIXED,
                "(T): precision must be CCGPU_FP64 or CCGPU_MIXED");
  return prec_of(precision);
}
// A non-blocking stream for work outside a context, with its scratch released at the end.
struct SideStream {
  hipStream_t s = nullptr;
  SideStream() {
    if (!ccgpu::dryrun()) CCGPU_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  ~SideStream() {
    if (ccgpu::dryrun() || !s) return;
    (void)hipStreamSynchronize(s);
    ccgpu::release_scratch(s);
    (void)hipStreamDestroy(s);
  }
};

// CCGPU_DRYRUN=1 (read once, at the first context or batch creation)
bool dryrun_env() {
  static const bool dry = [] {
    const char* e = std::getenv("CCGPU_DRYRUN");
    return e != nullptr && std::atoi(e) != 0;
  }();
  ccgpu::dryrun() = dry;
  return dry;
}
}  // namespace

extern "C" {

int ccgpu_version(void) { return 1; }

const char* ccgpu_last_error(void) { return g_err.c_str(); }

int ccgpu_init(int nocc, int nvir, int naux, const double* eps, const double* B, int mxdiis,
               ccgpu_ctx** ctx) {
  return guarded([&] {
    const bool dry = dryrun_env();
    CCGPU_REQUIRE(ctx != nullptr && (dry || (eps != nullptr && B != nullptr)), "ccgpu_init: null argument");
    *ctx = nullptr;
    auto c = std::make_unique<ccgpu_ctx>();
    c->cc = std::make_unique<ccgpu::CCSD>(nocc, nvir, naux, eps, B, mxdiis);
    *ctx = c.release();
  });
}

int ccgpu_df_factor_ao(int nbas, int nmo, const double* cmo, const double* eri_ao, double thr, long maxvec,
                       double* B, long* naux) {
  return guarded([&] {
    const bool dry = dryrun_env();
    CCGPU_REQUIRE(naux != nullptr && maxvec > 0 && (dry || (cmo != nullptr && eri_ao != nullptr)),
                  "ccgpu_df_factor_ao: null argument or no vector room");
    SideStream ss;
    bool conv = false;
    const std::vector<double> f = ccgpu::df_factor_ao(nbas, nmo, cmo, eri_ao, thr, maxvec, naux, &conv, ss.s);
    if (B != nullptr && !dry) std::memcpy(B, f.data(), f.size() * sizeof(double));
    if (!conv) throw ccgpu::Error(4, "ccgpu_df_factor_ao: threshold not reached within maxvec vectors");
  });
}

int ccgpu_df_factor_ao_cols(int nbas, int nmo, const double* cmo, const double* diag, ccgpu_ao_columns_fn columns,
                            void* user, double thr, long maxvec, double* B, long* naux) {
  return guarded([&] {
    const bool dry = dryrun_env();
    CCGPU_REQUIRE(naux != nullptr && maxvec > 0 && diag != nullptr && columns != nullptr && (dry || cmo != nullptr),
                  "ccgpu_df_factor_ao_cols: null argument or no vector room");
    SideStream ss;
    bool conv = false;
    const std::vector<double> f =
        ccgpu::df_factor_ao_cols(nbas, nmo, cmo, diag, columns, user, thr, maxvec, naux, &conv, ss.s);
    if (B != nullptr && !dry) std::memcpy(B, f.data(), f.size() * sizeof(double));
    if (!conv && !dry) throw ccgpu::Error(4, "ccgpu_df_factor_ao_cols: threshold not reached within maxvec vectors");
  });
}

int ccgpu_init_ao(int nocc, int nvir, int nbas, const double* eps, const double* cmo, const double* eri_ao,
                  double thr, int mxdiis, ccgpu_ctx** ctx, long* naux) {
  return guarded([&] {
    const bool dry = dryrun_env();
    CCGPU_REQUIRE(ctx != nullptr && naux != nullptr &&
                      (dry || (eps != nullptr && cmo != nullptr && eri_ao != nullptr)),
                  "ccgpu_init_ao: null argument");
    *ctx = nullptr;
    std::vector<double> f;
    {
      SideStream ss;
      bool conv = false;
      f = ccgpu::df_factor_ao(nbas, nocc + nvir, cmo, eri_ao, thr, 0, naux, &conv, ss.s);
    }
    CCGPU_REQUIRE(*naux > 0 && *naux < (1L << 31), "ccgpu_init_ao: no Cholesky vectors above the threshold");
    auto c = std::make_unique<ccgpu_ctx>();
    c->cc = std::make_unique<ccgpu::CCSD>(nocc, nvir, (int)*naux, eps, f.data(), mxdiis);
    *ctx = c.release();
  });
}

int ccgpu_finalize(ccgpu_ctx* ctx) {
  return guarded([&] { delete ctx; });
}

int ccgpu_rhs(ccgpu_ctx* ctx, const double* t1am, const double* t2am, double* omega1, double* omega2,
              int precision) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && t1am && t2am && omega1 && omega2, "ccgpu_rhs: null argument");
    ctx->cc->set_precision(prec_of(precision));
    ctx->cc->residual_host(t1am, t2am, omega1, omega2);
  });
}

int ccgpu_lhtr(ccgpu_ctx* ctx, int precision, int nsimtr, const double* ctr1, const double* ctr2, double* rho1,
               double* rho2) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx, "ccgpu_lhtr: null context");
    CCGPU_REQUIRE(nsimtr == 0 || (ctr1 && ctr2 && rho1 && rho2), "ccgpu_lhtr: null argument");
    ctx->cc->set_precision(prec_of(precision));
    ctx->cc->lhtr_host(nsimtr, ctr1, ctr2, rho1, rho2, true);
  });
}

int ccgpu_eta(ccgpu_ctx* ctx, double* eta1, double* eta2) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && eta1 && eta2, "ccgpu_eta: null argument");
    ctx->cc->eta_host(eta1, eta2, true);
  });
}

int ccgpu_lambda_solve(ccgpu_ctx* ctx, int precision, double thr, int maxit, int maxred, double* l1, double* l2,
                       int* niter, double* resnorm) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && l1 && l2, "ccgpu_lambda_solve: null argument");
    ctx->cc->set_precision(prec_of(precision));
    ctx->cc->lambda_solve(thr, maxit, maxred, l1, l2, resnorm, niter, true);
  });
}

int ccgpu_lambda_stats(ccgpu_ctx* ctx, int* n_fp64, int* n_mixed) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && n_fp64 && n_mixed, "ccgpu_lambda_stats: null argument");
    *n_fp64 = ctx->cc->lambda_counts()[0];
    *n_mixed = ctx->cc->lambda_counts()[1];
  });
}

int ccgpu_onedens(ccgpu_ctx* ctx, int precision, const double* l1, const double* l2, double* dens) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && l1 && l2 && dens, "ccgpu_onedens: null argument");
    ctx->cc->set_precision(prec_of(precision));
    ctx->cc->onedens_host(l1, l2, dens, true);
  });
}

int ccgpu_ccsd_guess(ccgpu_ctx* ctx, double* e_mp2) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx, "ccgpu_ccsd_guess: null context");
    const double e = ctx->cc->mp2_start();
    if (e_mp2) *e_mp2 = e;
  });
}

int ccgpu_ccsd_iterate(ccgpu_ctx* ctx, int precision, double* e_corr, double* omega_norm) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx, "ccgpu_ccsd_iterate: null context");
    ctx->cc->set_precision(prec_of(precision));
    double nrm = 0.0;
    const double e = ctx->cc->iterate(&nrm);
    if (e_corr) *e_corr = e;
    if (omega_norm) *omega_norm = nrm;
  });
}

// The CCSD_ENERGY loop from the context's current amplitudes, whose energy
// is en2 (MP2 start or restart).
static void ccsd_loop(ccgpu_ctx* ctx, int maxit, double threnr, double thrvec, int precision, double en2, double* e_ccsd,
                      int* niter, double* history) {
  const ccgpu::GemmPrec bulk = prec_of(precision);
  ccgpu::PrecisionSwitch ps(bulk == ccgpu::GemmPrec::F64);
  for (int it = 1;; ++it) {
    const double en1 = en2;
    double nrm = 0.0;
    ctx->cc->set_precision(ps.fp64 ? ccgpu::GemmPrec::F64 : bulk);
    en2 = ctx->cc->iterate(&nrm);
    if (history) history[it - 1] = en2;
    if (niter) *niter = it;
    if (e_ccsd) *e_ccsd = en2;
    const double de = std::fabs(en2 - en1);
    if (ps.after(de < threnr && nrm < thrvec, de, nrm)) break;
    if (it + 1 > maxit) throw ccgpu::Error(4, "CC equations not converged.");
  }
}

int ccgpu_ccsd_energy(ccgpu_ctx* ctx, int maxit, double threnr, double thrvec, int precision,
                      double* e_mp2, double* e_ccsd, int* niter, double* history) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && maxit > 0, "ccgpu_ccsd_energy: bad argument");
    const double en2 = ctx->cc->mp2_start();
    if (e_mp2) *e_mp2 = en2;
    ccsd_loop(ctx, maxit, threnr, thrvec, precision, en2, e_ccsd, niter, history);
  });
}

int ccgpu_ccsd_resume(ccgpu_ctx* ctx, int maxit, double threnr, double thrvec, int precision, double* e_start,
                      double* e_ccsd, int* niter, double* history) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && maxit > 0, "ccgpu_ccsd_resume: bad argument");
    const double en2 = ctx->cc->energy_current();
    if (e_start) *e_start = en2;
    ccsd_loop(ctx, maxit, threnr, thrvec, precision, en2, e_ccsd, niter, history);
  });
}

int ccgpu_energy(ccgpu_ctx* ctx, const double* t1, const double* t2, double* e_corr) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && t1 && t2 && e_corr, "ccgpu_energy: null argument");
    *e_corr = ctx->cc->energy_host(t1, t2);
  });
}

int ccgpu_diis(ccgpu_ctx* ctx, const double* s, double* t, int niter) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && s && t, "ccgpu_diis: null argument");
    if (!ctx->cc->diis_host(s, t, niter)) throw ccgpu::Error(3, "DIIS: ZERO PIVOT - MATRIX SINGULAR");
  });
}

int ccgpu_rsp_filename(const char* list, int isym, int idx, char* name) {
  return guarded([&] {
    CCGPU_REQUIRE(name, "ccgpu_rsp_filename: null argument");
    const std::string n = ccgpu::rsp_filename(list, isym, idx);
    std::memcpy(name, n.c_str(), n.size() + 1);
  });
}

int ccgpu_rsp_write(const char* path, const char* model, int intsize, long nallai, long nt1, long nt2,
                    const double* vec0, const double* vec1, const double* vec2) {
  return guarded([&] {
    CCGPU_REQUIRE(path, "ccgpu_rsp_write: null path");
    ccgpu::rsp_write(path, model, intsize, nallai, nt1, nt2, vec0, vec1, vec2);
  });
}

int ccgpu_rsp_read(const char* path, long nallai, long nt1, long nt2, double* vec0, double* vec1, double* vec2,
                   char* model, int* nwrite) {
  return guarded([&] {
    CCGPU_REQUIRE(path, "ccgpu_rsp_read: null path");
    const int nw = ccgpu::rsp_read(path, nallai, nt1, nt2, vec0, vec1, vec2, model);
    if (nwrite) *nwrite = nw;
    if ((vec1 && nw < 2) || (vec2 && nw < 3))
      throw ccgpu::Error(4, "ccgpu_rsp_read: the file holds fewer parts than requested");
  });
}

int ccgpu_write_e_intermediates(ccgpu_ctx* ctx, int precision, const char* e1_path, const char* e2_path,
                                double* e1, double* e2) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx, "ccgpu_write_e_intermediates: null context");
    const auto& d = ctx->cc->dims();
    std::vector<double> b1((size_t)d.v * d.v), b2((size_t)d.o * d.o);
    ctx->cc->set_precision(prec_of(precision));
    ctx->cc->e_intermediates_host(b1.data(), b2.data());
    if (e1_path) ccgpu::write_one_record(e1_path, b1.data(), (long)b1.size());
    if (e2_path) ccgpu::write_one_record(e2_path, b2.data(), (long)b2.size());
    if (e1) std::memcpy(e1, b1.data(), b1.size() * sizeof(double));
    if (e2) std::memcpy(e2, b2.data(), b2.size() * sizeof(double));
  });
}

int ccgpu_write_gamma_intermediate(ccgpu_ctx* ctx, int precision, const char* path, double* gamma) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx, "ccgpu_write_gamma_intermediate: null context");
    const long n = (long)ctx->cc->dims().o * ctx->cc->dims().o;
    std::vector<double> g((size_t)(n * (n + 1) / 2));
    ctx->cc->set_precision(prec_of(precision));
    ctx->cc->gamma_host(g.data());
    if (path) ccgpu::write_one_record(path, g.data(), (long)g.size());
    if (gamma) std::memcpy(gamma, g.data(), g.size() * sizeof(double));
  });
}

int ccgpu_save_amplitudes(ccgpu_ctx* ctx, const char* path, int intsize) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && path, "ccgpu_save_amplitudes: null argument");
    const auto& d = ctx->cc->dims();
    std::vector<double> t1((size_t)d.nov), t2((size_t)(d.ntamp - d.nov));
    ctx->cc->get_amplitudes(t1.data(), t2.data());
    ccgpu::rsp_write(path, "CCSD", intsize, d.nov, d.nov, d.ntamp - d.nov, nullptr, t1.data(), t2.data());
  });
}

int ccgpu_restart_amplitudes(ccgpu_ctx* ctx, const char* path, double* e_corr) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && path, "ccgpu_restart_amplitudes: null argument");
    const auto& d = ctx->cc->dims();
    std::vector<double> t1((size_t)d.nov), t2((size_t)(d.ntamp - d.nov));
    char model[11];
    const int nw = ccgpu::rsp_read(path, d.nov, d.nov, d.ntamp - d.nov, nullptr, t1.data(), t2.data(), model);
    if (nw < 3) throw ccgpu::Error(4, "ccgpu_restart_amplitudes: the file holds no doubles");
    ctx->cc->set_amplitudes(t1.data(), t2.data());
    const double e = ctx->cc->energy_current();
    if (e_corr) *e_corr = e;
  });
}

int ccgpu_batch_init(int nsys, const int* nocc, const int* nvir, const int* naux, const double* const* eps,
                     const double* const* B, int mxdiis, double ladder_budget_bytes, ccgpu_batch** batch) {
  return guarded([&] {
    CCGPU_REQUIRE(batch && nsys > 0 && nocc && nvir && naux && eps && B, "ccgpu_batch_init: bad argument");
    for (int k = 0; k < nsys; ++k)
      CCGPU_REQUIRE(dryrun_env() || (eps[k] && B[k]), "ccgpu_batch_init: null system arrays");
    auto* h = new ccgpu_batch;
    try {
      h->b.reset(new ccgpu::Batch(nsys, nocc, nvir, naux, eps, B, mxdiis, ladder_budget_bytes));
    } catch (...) {
      delete h;
      throw;
    }
    *batch = h;
  });
}

int ccgpu_batch_energy(ccgpu_batch* batch, int maxit, double threnr, double thrvec, int precision, int flags,
                       double* e_mp2, double* e_ccsd, int* niter) {
  return guarded([&] {
    CCGPU_REQUIRE(batch, "ccgpu_batch_energy: null batch");
    if (flags & CCGPU_BATCH_TUNE_ONLY) {
      batch->b->tune();
      return;
    }
    const bool graph = !(flags & CCGPU_BATCH_EAGER), tune = !(flags & CCGPU_BATCH_NO_TUNE);
    batch->b->solve(maxit, threnr, thrvec, prec_of(precision), graph, tune, e_mp2, e_ccsd, niter);
  });
}

int ccgpu_batch_get_amplitudes(ccgpu_batch* batch, int k, double* t1am, double* t2am) {
  return guarded([&] {
    CCGPU_REQUIRE(batch && k >= 0 && k < batch->b->size(), "ccgpu_batch_get_amplitudes: bad argument");
    batch->b->system(k).get_amplitudes(t1am, t2am);
  });
}

int ccgpu_batch_stats(ccgpu_batch* batch, int* captures, int* replays) {
  return guarded([&] {
    CCGPU_REQUIRE(batch, "ccgpu_batch_stats: null batch");
    if (captures) *captures = batch->b->captures();
    if (replays) *replays = batch->b->replays();
  });
}

int ccgpu_batch_finalize(ccgpu_batch* batch) {
  return guarded([&] { delete batch; });
}

int ccgpu_get_amplitudes(ccgpu_ctx* ctx, double* t1am, double* t2am) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx, "ccgpu_get_amplitudes: null context");
    ctx->cc->get_amplitudes(t1am, t2am);
  });
}

int ccgpu_set_amplitudes(ccgpu_ctx* ctx, const double* t1am, const double* t2am) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx && t1am && t2am, "ccgpu_set_amplitudes: null argument");
    ctx->cc->set_amplitudes(t1am, t2am);
  });
}

int ccgpu_pt(ccgpu_ctx* ctx, int precision, double* e4, double* e5) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx, "ccgpu_pt: null context");
    ctx->cc->set_precision(pt_prec_of(precision));
    ctx->cc->pt_energy(0, ctx->cc->pt_ntriples(), 1, e4, e5, nullptr);
  });
}

int ccgpu_pt_partial(ccgpu_ctx* ctx, int precision, long long first, long long count, long long stride,
                     double* e4, double* e5, double* flops, long long* ntriples) {
  return guarded([&] {
    CCGPU_REQUIRE(ctx, "ccgpu_pt_partial: null context");
    ctx->cc->set_precision(pt_prec_of(precision));
    if (ntriples) *ntriples = ctx->cc->pt_ntriples();
    ctx->cc->pt_energy(first, count, stride, e4, e5, flops);
  });
}

int ccgpu_profile_start(void) {
  return guarded([&] { ccgpu::prof_start(); });
}

int ccgpu_profile_stop(char* buf, int len) {
  std::string js;
  const int rc = guarded([&] { js = ccgpu::prof_stop_json(); });
  if (rc != 0) return -rc;
  if (buf && len > 0) {
    const size_t n = std::min<size_t>(js.size(), (size_t)len - 1);
    std::memcpy(buf, js.data(), n);
    buf[n] = '\0';
  }
  return static_cast<int>(js.size()) + 1;
}

namespace {
// Column-major BLAS operands (DGEMM's argument list) as a GemmDesc; the
// element counts of the three arrays are returned in na, nb, nc.
ccgpu::GemmDesc blas_desc(char transa, char transb, int m, int n, int k, double alpha, const double* a, int lda,
                          const double* b, int ldb, double beta, double* c, int ldc, size_t& na, size_t& nb,
                          size_t& nc) {
  const bool ta = transa == 'T' || transa == 't' || transa == 'C' || transa == 'c';
  const bool tb = transb == 'T' || transb == 't' || transb == 'C' || transb == 'c';
  CCGPU_REQUIRE(m >= 0 && n >= 0 && k >= 0, "ccgpu_dgemm: negative dimension");
  CCGPU_REQUIRE(lda >= std::max(1, ta ? k : m) && ldb >= std::max(1, tb ? n : k) && ldc >= std::max(1, m),
                "ccgpu_dgemm: leading dimension too small");
  na = (size_t)lda * (ta ? m : k);
  nb = (size_t)ldb * (tb ? k : n);
  nc = (size_t)ldc * n;
  ccgpu::GemmDesc d;
  d.M = m; d.N = n; d.K = k; d.alpha = alpha; d.beta = beta;
  d.A = a; d.B = b; d.C = c;
  if (!ta) { d.a_m = 1; d.a_k = lda; } else { d.a_m = lda; d.a_k = 1; }
  if (!tb) { d.b_k = 1; d.b_n = ldb; } else { d.b_k = ldb; d.b_n = 1; }
  d.c_m = 1; d.c_n = ldc;
  return d;
}

int dgemm_impl(ccgpu::GemmPrec prec, char transa, char transb, int m, int n, int k, double alpha,
               const double* a, int lda, const double* b, int ldb, double beta, double* c, int ldc) {
  return guarded([&] {
    size_t na = 0, nb = 0, nc = 0;
    ccgpu::GemmDesc d = blas_desc(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, na, nb, nc);
    if (m == 0 || n == 0) return;
    ccgpu::DBuf<double> A(na), B(nb), C(nc);
    CCGPU_HIPX(hipMemcpy(A.p, a, na * sizeof(double), hipMemcpyHostToDevice));
    CCGPU_HIPX(hipMemcpy(B.p, b, nb * sizeof(double), hipMemcpyHostToDevice));
    CCGPU_HIPX(hipMemcpy(C.p, c, nc * sizeof(double), hipMemcpyHostToDevice));
    d.A = A.p; d.B = B.p; d.C = C.p;
    ccgpu::gemm(d, prec, nullptr);
    CCGPU_HIPX(hipDeviceSynchronize());
    CCGPU_HIPX(hipMemcpy(c, C.p, nc * sizeof(double), hipMemcpyDeviceToHost));
  });
}

}  // namespace

int ccgpu_mfma_ceiling(int precision, double* tflops) {
  return guarded([&] {
    CCGPU_REQUIRE(tflops, "ccgpu_mfma_ceiling: null argument");
    CCGPU_REQUIRE(precision == CCGPU_FP64 || precision == CCGPU_MIXED,
                  "ccgpu_mfma_ceiling: precision must be CCGPU_FP64 or CCGPU_MIXED");
    CCGPU_REQUIRE(!dryrun_env(), "ccgpu_mfma_ceiling: nothing to measure in a dry run");
    *tflops = ccgpu::mfma_only_tflops(precision == CCGPU_FP64);
  });
}

int ccgpu_dgemm(char transa, char transb, int m, int n, int k, double alpha, const double* a, int lda,
                const double* b, int ldb, double beta, double* c, int ldc) {
  return dgemm_impl(ccgpu::GemmPrec::F64, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

int ccgpu_dgemm_mixed(char transa, char transb, int m, int n, int k, double alpha, const double* a, int lda,
                      const double* b, int ldb, double beta, double* c, int ldc) {
  return dgemm_impl(ccgpu::GemmPrec::BF16X3, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

int ccgpu_dgemm_mixed_noalloc(char transa, char transb, int m, int n, int k, double alpha, const double* a,
                              int lda, const double* b, int ldb, double beta, double* c, int ldc, void* stream) {
  return guarded([&] {
    size_t na = 0, nb = 0, nc = 0;
    const ccgpu::GemmDesc d = blas_desc(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, na, nb, nc);
    if (m == 0 || n == 0) return;
    CCGPU_REQUIRE(c != nullptr && (k == 0 || (a != nullptr && b != nullptr)),
                  "ccgpu_dgemm_mixed_noalloc: null device array");
    const void* spans[3] = {a, b, c};
    const size_t bytes[3] = {na * sizeof(double), nb * sizeof(double), nc * sizeof(double)};
    ccgpu::ExternalSpans ext(spans, bytes, 3);
    ccgpu::gemm_bf16x3_noalloc(d, static_cast<hipStream_t>(stream));
  });
}

}  // extern "C"
