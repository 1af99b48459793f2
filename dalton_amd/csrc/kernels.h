// Bandwidth-bound kernels of the CC path: layout permutations, packing in
// Dalton's ai<=bj order, the amplitude update, DIIS vector algebra, energy
// and norm reductions, T1 dressing helpers and the ladder +/- helpers.
#pragma once
#include "common.h"

namespace ccgpu {

// out[j0][j1][j2][j3] = alpha * in[j0*s0 + j1*s1 + j2*s2 + j3*s3] + beta * out
// (row-major out of dims n0..n3; beta == 0 never reads out).
void permute4(double* out, const double* in, const long n[4], const long s[4], double alpha,
              double beta, hipStream_t st);

// y[m] = alpha * sum_k A[k*lda + m] x[k] + beta * y[m]   (A column-contiguous in m)
void gemv_mc(long M, long K, double alpha, const double* A, long lda, const double* x, double beta,
             double* y, hipStream_t st);

// Dalton packed symmetric order: packed[bj*(bj+1)/2 + ai] = F[bj][ai], ai <= bj.
void pack_sym(double* packed, const double* F, long n, hipStream_t st);
// F = full symmetric matrix from packed.
void unpack_sym(double* F, const double* packed, long n, hipStream_t st);

// Jacobi step (CCSD_NXTAM): S = T + Om / D, written over Om; eps = (o occ, v vir).
// Singles layout [i][a]; doubles [(i,a),(j,b)].  Also returns (device) the
// packed squared norm of the incoming Om1 and Om2 into norm2[0].
void nxtam(double* om1, double* om2, const double* t1, const double* t2, const double* eps, int o,
           int v, double* norm2, hipStream_t st);

// svec = [S1; pack(S2)], evec = svec - [t1; pack(t2)].
void pack_diis(double* svec, double* evec, const double* s1, const double* s2, const double* t1,
               const double* t2, long nov, hipStream_t st);

// dots[i] = <e, E_i> for i < nvec, E_i = E + i*ld   (deterministic two-pass).
void multidot(double* dots, const double* e, const double* E, long ld, int nvec, long n,
              hipStream_t st);
// out = sum_i c[i] * S_i  (c on device)
void lincomb(double* out, const double* S, long ld, const double* c, int nvec, long n, hipStream_t st);

// The same energy from packed doubles t2p[bj(bj+1)/2 + ai] (ai <= bj).
void ccsd_energy_packed(double* out, const double* L, const double* t2p, const double* t1, long nov,
                        hipStream_t st);
// E = sum L[(i,a),(j,b)] * (T + t1 (x) t1)   -> out[0]
void ccsd_energy(double* out, const double* L, const double* t2, const double* t1, long nov,
                 hipStream_t st);

// tau+/- for the symmetric/antisymmetric particle ladder (rows ij with i<=j,
// columns cd with c<=d, p = d(d+1)/2 + c) from Too[(i,j)][(c,d)].
void ladder_tau(double* taup, double* taum, const double* Too, int o, int v, hipStream_t st);
// V+/-[r][p] from Y[(a,c),(b,d)] for a in [a0,a1), b in [a,v).
void ladder_vpm(double* vp, double* vm, const double* Y, long ldY, int a0, int a1, int v,
                hipStream_t st);
// Om2[(i,a),(j,b)] += S+ +/- S- (4-way scatter of the batch rows)
void ladder_scatter(double* om2, const double* sp, const double* sm, int a0, int a1, int o, int v,
                    hipStream_t st);

// Jd[P] = 2 * sum_k Bt[P][k][k]
void coulomb_diag(double* Jd, const double* Bt, long naux, int N, int o, hipStream_t st);

// small helpers
void fill(double* p, double val, long n, hipStream_t st);
void axpy(double* y, double a, const double* x, long n, hipStream_t st);

}  // namespace ccgpu

namespace ccgpu {
// MP2 guess (CCSD_GUESS): T = J / (e_i + e_j - e_a - e_b)
void mp2_guess(double* T, const double* J, const double* eps, int o, int v, hipStream_t st);
}  // namespace ccgpu

namespace ccgpu {
// ---- left transformation (adjoint) helpers
// M[i][i] *= f for i < n (row stride ld)
void scale_diag(double* M, long n, long ld, double f, hipStream_t st);
// t2p[ai(ai+3)/2] *= f for ai < nov: the packed doubles diagonal.
void scale_pdiag(double* t2p, long nov, double f, hipStream_t st);
// Bt[P][k][k] += f * x[P] for k < o
void add_occ_diag(double* Bt, const double* x, long naux, int N, int o, double f, hipStream_t st);
// S+/-bar[r][q] gathered from the full cotangent L (adjoint of ladder_scatter)
void ladder_gather(double* spb, double* smb, const double* L, int a0, int a1, int o, int v, hipStream_t st);
// Ybar[(a,c),(b,d)] from V+/-bar (adjoint of ladder_vpm); entries with b < a untouched
void ladder_vpm_adj(double* Yb, long ldY, const double* vpb, const double* vmb, int a0, int a1, int v,
                    hipStream_t st);
// Too_bar[(i,j)][(c,d)] += adjoint of ladder_tau for i <= j
void ladder_tau_adj(double* Toob, const double* taupb, const double* taumb, int o, int v, hipStream_t st);
}  // namespace ccgpu

namespace ccgpu {
// out[j0*os0 + j1*os1 + j2*os2] += alpha * in[(j0*n1 + j1)*n2 + j2]
void add_strided3(double* out, long os0, long os1, long os2, const double* in, long n0, long n1, long n2,
                  double alpha, hipStream_t st);
}  // namespace ccgpu

namespace ccgpu {
// ---- device-resident DIIS (CCSD_DIIS, cc/ccsd_diis.F:251-365).  ctr[0] is the
// iteration count of the current DIIS cycle (1-based), so every launch is
// independent of host state and an iteration can be captured in a graph.
// Slot ivec = it for it <= mx, else it - ((it-1)/mx)*mx; nloop = ivec.
// svec/evec of slot ivec from S = [s1; pack(s2)] and T = [t1; pack(t2)]
void pack_diis_slot(double* diisS, double* diisE, const int* ctr, int mx, const double* s1, const double* s2,
                    const double* t1, const double* t2, long nov, hipStream_t st);
// slot ivec of flat vectors: S_ivec = s, E_ivec = s - t (n values each)
void pack_diis_slot_flat(double* diisS, double* diisE, const int* ctr, int mx, const double* s, const double* t,
                         long n, hipStream_t st);
// dots[i] = <E_ivec, E_i> for i < nloop (deterministic two-pass)
void multidot_slot(double* dots, const double* diisE, long ld, const int* ctr, int mx, long n, hipStream_t st);
// Extends the B matrix (Amat, leading dimension ld) by the new dots and
// solves Dalton's bordered system with partial pivoting; coef[0..nloop).
// A zero pivot sets ctr[1] = 1 and takes the newest vector alone.
void diis_solve(double* Amat, int ld, const double* dots, double* coef, int* ctr, int mx, double* work,
                hipStream_t st);
// out = sum_{i < nloop} coef[i] S_i
void lincomb_slot(double* out, const double* S, long ld, const double* coef, const int* ctr, int mx, long n,
                  hipStream_t st);
// ctr[0] += 1
void ctr_advance(int* ctr, hipStream_t st);
// out[0..3] = scal[0], scal[1], ctr[1], ctr[0]
void report4(double* out, const double* scal, const int* ctr, hipStream_t st);
}  // namespace ccgpu
