// (T) energy on one GPU, occupied-driven over i >= j >= k
// (mlcc/mlccsdpt_e_calc.F90:71, loop :326-411): the integral and amplitude
// layouts shared by both drivers, and the driver choice.
//
// Per triple, W_ijk^abc (v^3) gets twelve terms: six "virtual" terms
// sum_d t_pq^{xd} (yd|zr) (K = v) and six "occupied" terms
// -sum_l t_pl^{xy} (zr|lq) (K = o), one per simultaneous permutation of
// (ai),(bj),(ck) (w_calc :966).  They run as three four-term passes grouped
// by output orientation (pt_persistent.hip), reading the integrals (yd|zr)
// in two layouts, G[r][d][y][z] and Gp[r][d][z][y] (packed bf16 hi/lo
// transposes of both in the mixed path), t2 re-laid as Tvvo[p][x][y][l] =
// t_pl^{xy} and (zr|lq) as Q[r][q][l][z], all built once from the DF factor
// and the amplitudes.  The energy evaluates the permutation-symmetrised
// functional of w_prod (:1547-1640) per orbit {a,b,c}: E4 takes V = W, E5
// the t1 part of V (v_calc :1345, :1483-1527).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

#include "ccsd.h"
#include "kernels.h"
#include "ops.h"

namespace ccgpu {

namespace {

// p[0..n) = 0 (grid-stride; sizes past 2^31 bytes)
__global__ void k_zero_u16(unsigned short* p, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = 0;
}

}  // namespace

long CCSD::pt_ntriples() const {
  const long o = dm_.o;
  return o * (o + 1) * (o + 2) / 6 - o;  // i >= j >= k without i = j = k
}

void CCSD::pt_release() {
  for (DBuf<double>* b : {&G_, &Gp_, &Tvvo_, &Q_, &Wt_, &ptpart_, &etrip_}) b->release();
  for (DBuf<unsigned short>* b : {&GTh_, &GTl_, &GpTh_, &GpTl_}) b->release();
  for (int q = 0; q < 4; ++q) {
    ppH_[q].release();
    ppL_[q].release();
  }
  for (DBuf<int>* b : {&pttri_, &ptctab_}) b->release();
  ptclaim_.release();
  TvvoS_.release();
  ptdesc_.release();
  W32_.release();
  pt_ready_ = false;
}

// Integral and amplitude layouts for the triple loop (built once per call
// sequence; the iteration workspace is released first to make room): every
// per-triple operand pre-packed, and four W slots (the persistent kernel's
// pipeline; the per-triple launches use the first).
void CCSD::pt_prepare(GemmPrec mode, PtPath path) {
  if (pt_ready_ && pt_mode_ == mode && pt_path_ == path) return;
  pt_release();
  release_work();
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long N = dm_.N, NN = N * N, nov = dm_.nov, vv = (long)v * v, vvv = vv * v;
  const bool mixed = mode == GemmPrec::BF16X3;
  Tvvo_.alloc((size_t)o * vv * o);
  Q_.alloc((size_t)o * o * o * v);
  ptpart_.alloc(std::max<size_t>((size_t)(vv + v) + 2, 8 * 2048));
  {
    // undressed virtual-virtual and occupied-occupied blocks of B
    DBuf<double> Bvv0((size_t)naux * vv), Boo0((size_t)naux * o * o), R((size_t)nov * o * o);
    perm(Bvv0.p, B_.p + o * N + o, 1, naux, v, v, 0, NN, N, 1, 1.0, 0.0, st_);
    perm(Boo0.p, B_.p, 1, naux, o, o, 0, NN, N, 1, 1.0, 0.0, st_);
    if (!mixed) {
      G_.alloc((size_t)o * vvv);
      Gp_.alloc((size_t)o * vvv);
      Wt_.alloc(4 * (size_t)vvv);
      TvvoS_.alloc((size_t)o * vv * o);
      // G[k][d][y][z] = (yd|zk) = sum_P B[P][d][y] B[P][k][z], batched over k
      g(vv, v, naux, 1.0, Bvv0.p, 1, vv, Bov_.p, nov, 1, 0.0, G_.p, v, 1, st_, o, 0, v, vvv);
      // Gp[k][d][y][z] = G[k][d][z][y]
      perm(Gp_.p, G_.p, 1, (long)o * v, v, v, 0, vv, 1, v, 1.0, 0.0, st_);
    } else {
      // packed transposes GT[k][(y,z)][d] = G[k][d][y][z] and GpT = the same of Gp,
      // built one k at a time; 256 zero rows at the end absorb tile over-reads
      const long Kpv = bf16x3_kpad(v);
      const size_t ng = ((size_t)o * vv + 256) * Kpv;
      GTh_.alloc(ng);
      GTl_.alloc(ng);
      GpTh_.alloc(ng);
      GpTl_.alloc(ng);
      W32_.alloc(4 * (size_t)vvv);
      DBuf<double> Gk(vvv), Gpk(vvv);
      for (int k = 0; k < o; ++k) {
        g(vv, v, naux, 1.0, Bvv0.p, 1, vv, Bov_.p + (long)k * v, nov, 1, 0.0, Gk.p, v, 1, st_);
        perm(Gpk.p, Gk.p, 1, v, v, v, 0, vv, 1, v, 1.0, 0.0, st_);
        const long rp = (k == o - 1) ? vv + 256 : vv;
        const size_t off = (size_t)k * vv * Kpv;
        // element (r=(y,z), d) of the operand at Gk[d*vv + r]
        bf16x3_pack(Gk.p, vv, v, 1, vv, 0, 1, rp, Kpv, 1.0f, GTh_.p + off, GTl_.p + off, st_);
        bf16x3_pack(Gpk.p, vv, v, 1, vv, 0, 1, rp, Kpv, 1.0f, GpTh_.p + off, GpTl_.p + off, st_);
      }
      CCGPU_HIPX(hipStreamSynchronize(st_));
    }
    // R[(k,c)][(l,j)] = (ck|lj);  Q[k][j][l][c] = R
    g(nov, (long)o * o, naux, 1.0, Bov_.p, 1, nov, Boo0.p, (long)o * o, 1, 0.0, R.p, (long)o * o, 1, st_);
    perm(Q_.p, R.p, o, o, o, v, (long)v * o * o, 1, o, (long)o * o, 1.0, 0.0, st_);
    CCGPU_HIPX(hipStreamSynchronize(st_));
  }
  // Tvvo[i][a][b][l] = t_il^ab = T[(i,a),(l,b)]
  perm(Tvvo_.p, T_.p, o, v, v, o, (long)v * nov, nov, 1, v, 1.0, 0.0, st_);
  {
    if (!mixed) {
      // TvvoS[p][x][y][l] = Tvvo[p][y][x][l]
      perm(TvvoS_.p, Tvvo_.p, o, v, v, o, vv * o, o, (long)v * o, 1, 1.0, 0.0, st_);
    } else {
      const long Kpv = bf16x3_kpad(v), Kpo = bf16x3_kpad(o), Rv = bf16x3_rpad(v), Rvv = bf16x3_rpad(vv) + 256;
      CCGPU_REQUIRE(o <= 255 && v <= 65535, "pt: occupied or virtual space too large for the packed slices");
      const size_t n0 = (size_t)o * o * Rv * Kpv, n1 = (size_t)o * o * Rv * Kpo, n2 = (size_t)o * Rvv * Kpo;
      const size_t nn[4] = {n0, n1, n2, n2};
      for (int q = 0; q < 4; ++q) {
        ppH_[q].alloc(nn[q]);
        ppL_[q].alloc(nn[q]);
      }
      // slot tails past v*v rows of the swapped -Tvvo are read by edge tiles only: keep them finite
      for (DBuf<unsigned short>* b : {&ppH_[3], &ppL_[3]})
        CCGPU_LAUNCH(k_zero_u16, dim3(4096), dim3(256), 0, st_, b->p, (long)b->n);
      for (int p = 0; p < o; ++p) {
        // T(p,q)[x][d] = t_pq^{xd} for every q: [(p,q)][x][d]
        const size_t ot = (size_t)p * o * Rv * Kpv;
        bf16x3_pack(T_.p + (long)p * v * nov, v, v, nov, 1, v, o, Rv, Kpv, 1.0f, ppH_[0].p + ot, ppL_[0].p + ot,
                    st_);
        // [x][c][l] = -Tvvo_p[c][x][l]
        const size_t os = (size_t)p * Rvv * Kpo;
        bf16x3_pack(Tvvo_.p + (long)p * vv * o, v, o, (long)v * o, 1, o, v, v, Kpo, -1.0f, ppH_[3].p + os,
                    ppL_[3].p + os, st_);
      }
      // Q_rq[z][l] = (zr|lq) for every (r,q)
      bf16x3_pack(Q_.p, v, o, 1, v, (long)o * v, o * o, Rv, Kpo, 1.0f, ppH_[1].p, ppL_[1].p, st_);
      // [(x,y)][l] = -Tvvo_p[x][y][l]
      bf16x3_pack(Tvvo_.p, vv, o, o, 1, vv * o, o, Rvv, Kpo, -1.0f, ppH_[2].p, ppL_[2].p, st_);
    }
    // canonical 8x8x8 blocks A <= B <= C of the virtual cube, A | B << 10 | C << 20
    const long nb = cdiv(v, 8);
    CCGPU_REQUIRE(nb <= 1023, "pt: virtual space too large for the block table");
    std::vector<int> ct;
    for (long a = 0; a < nb; ++a)
      for (long b = a; b < nb; ++b)
        for (long c = b; c < nb; ++c) ct.push_back((int)(a | (b << 10) | (c << 20)));
    ptctab_.alloc(ct.size());
    CCGPU_HIPX(hipMemcpyAsync(ptctab_.p, ct.data(), ct.size() * sizeof(int), hipMemcpyHostToDevice, st_));
    CCGPU_HIPX(hipStreamSynchronize(st_));
  }
  tri_i_.clear();
  tri_j_.clear();
  tri_k_.clear();
  for (int i = 0; i < o; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= j; ++k) {
        if (i == j && j == k) continue;
        tri_i_.push_back(i);
        tri_j_.push_back(j);
        tri_k_.push_back(k);
      }
  pt_mode_ = mode;
  pt_path_ = path;
  pt_ready_ = true;
}

// (T) driver from CCGPU_PT_PATH: "launch" (default: per-triple launches of
// the pass kernels) or "persistent" (one cooperative kernel per chunk of
// triples).
static int pt_path_from_env() {
  const char* e = std::getenv("CCGPU_PT_PATH");
  if (e && std::string(e) == "persistent") return 1;
  return 0;
}

void CCSD::pt_energy(long first, long count, long stride, double* e4, double* e5, double* flops) {
  const bool mixed = bulk_ == GemmPrec::BF16X3;
  const PtPath path = static_cast<PtPath>(pt_path_from_env());
  pt_prepare(mixed ? GemmPrec::BF16X3 : GemmPrec::F64, path);
  const int o = dm_.o, v = dm_.v;
  const long vv = (long)v * v, vvv = vv * v;
  const long ntr = pt_ntriples();
  CCGPU_REQUIRE(first >= 0 && stride >= 1 && count >= 0, "pt: bad triple range");
  long n = 0;
  for (long t = first; t < ntr && n < count; t += stride) ++n;
  if (flops) *flops = (double)n * (12.0 * (double)vv * vv + 12.0 * (double)o * vvv);
  if (path == PtPath::Persistent)
    pt_energy_persistent(first, n, stride, e4, e5);
  else
    pt_energy_launches(first, n, stride, e4, e5);
}

}  // namespace ccgpu
