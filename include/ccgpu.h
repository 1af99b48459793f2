/*
 * ccgpu: MI355X (gfx950) closed-shell coupled-cluster hot path, C ABI.
 *
 * Drop-in boundary for Dalton's closed-shell CCSD / (T) path.  All arrays
 * are caller-owned host arrays in Dalton's own layouts:
 *   - singles T1AM/OMEGA1: (a,i) with a fastest, NAI = a + NVIR*(i-1)
 *     (cc/ccsd_energy.F, e.g. :1685-1690);
 *   - doubles T2AM/OMEGA2: packed over ai <= bj, INDEX(ai,bj) =
 *     bj*(bj-1)/2 + ai (1-based; statement function at cc/ccsd_energy.F:1666),
 *     NT2AMX = NT1AMX*(NT1AMX+1)/2 elements (C1 symmetry);
 *   - orbital energies: occupied first, then virtual (FOCKD in Dalton);
 *   - two-electron integrals: a density-fitting / Cholesky factor B of shape
 *     [naux][N][N] (row-major, N = nocc + nvir, symmetric in the last two
 *     indices) with (pq|rs) = sum_P B[P][p][q] B[P][r][s].  Dalton's CC code
 *     reads AO integral distributions instead (CCRDAO, cc/ccsd_rhs.F:5652);
 *     the factor replaces that source.
 * Every function returns 0 on success and a nonzero code otherwise
 * (1 = bad argument, 2 = HIP runtime error, 3 = numerical failure such as a
 * singular DIIS matrix, 4 = not converged); ccgpu_last_error() gives the
 * message.  A Fortran caller maps nonzero codes to QUIT (gp/gptrygve.F:256),
 * which is Dalton's only error convention.
 * Calls on one context are synchronous and must not overlap.
 * ccgpu_rhs, ccgpu_lhtr, ccgpu_set_amplitudes and ccgpu_get_amplitudes also
 * accept device arrays (allocations on the context's GPU, e.g. torch
 * tensors): the copy direction is taken from the addresses, and the call
 * returns once the results are in place.  Work the caller queued on its own
 * streams that writes their inputs must be complete before the call.
 */
#ifndef CCGPU_H
#define CCGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ccgpu_ctx ccgpu_ctx;

/* precision selectors */
#define CCGPU_FP64 0  /* all contractions in FP64 on MFMA */
#define CCGPU_MIXED 1 /* bulk contractions BF16x3 on MFMA (FP32 accumulation), FP64 update/DIIS/energy */
#define CCGPU_FP32 2  /* bulk contractions FP32 on MFMA (v_mfma_f32_16x16x4f32), FP64 update/DIIS/energy;
                         the CCSD iterations, the left transformation and the multipliers, not (T) */

int ccgpu_version(void);
const char* ccgpu_last_error(void);

/* Registers the Hamiltonian (f = diag(eps), integrals from B) and allocates
 * device memory for one closed-shell system.  Replaces the setup done in
 * CCSD_ENERGY before its loop (orbital energies read from SIRIFC,
 * cc/ccsd_energy.F:504-518; (ia|jb) built by CCSD_IAJB, cc/ccsd_energy.F:2328).
 * mxdiis: DIIS subspace size (Dalton .MXDIIS, default 8, cc/ccsd_init0.F:179-186). */
int ccgpu_init(int nocc, int nvir, int naux, const double* eps, const double* B, int mxdiis,
               ccgpu_ctx** ctx);
int ccgpu_finalize(ccgpu_ctx* ctx);

/* CCSD vector function: replaces
 *   SUBROUTINE CCRHSN(OMEGA1,OMEGA2,T1AM,T2AM,WORK,LWORK,APROXR12)
 * cc/ccsd_rhs.F:20 (plain closed-shell CCSD, APROXR12 blank).  T1AM/T2AM
 * are read only (Dalton squares T2AM in place and restores it; no caller
 * workspace is needed here).  Like CCRHSN it has no effect on the iteration
 * state: the context's amplitudes, iteration count and DIIS history and
 * cycle are unchanged, so it can be mixed with ccgpu_ccsd_iterate or called
 * for finite-difference Jacobians (CCLR_FDJAC, cc/cclr_util.F:946). */
int ccgpu_rhs(ccgpu_ctx* ctx, const double* t1am, const double* t2am, double* omega1,
              double* omega2, int precision);

/* Left-hand transformation of nsimtr trial vectors at the context's current
 * amplitudes: replaces CC_LHTR (cc/cc_lhtr.F:20) for MODEL = CCSD
 * (closed shell, C1 symmetry, no R12/CC3/PE terms), whose CTR1/CTR2 trial
 * vectors and RHO1/RHO2 results are held NSIMTR at a time.  Dalton's packed
 * convention for left vectors: the doubles part of a trial vector is squared
 * without scaling its diagonal (CC_TRDRV scales only right vectors,
 * cc/cc_trdrv.F:484-486), and the result is the left product with Dalton's
 * Jacobian as CCLR_FDJAC defines it (cc/cclr_util.F:1228,1343: diagonal
 * doubles rows halved, diagonal doubles columns doubled; applied from the
 * left by CCLR_DUMTRR, :1624).  With J_mu,nu = dOmega_mu/dt_nu on packed
 * arrays (Omega the vector function of ccgpu_rhs) and S, D the diagonal
 * matrices that halve, respectively double, the packed doubles diagonal,
 *   rho = D J^T S ctr.
 * ctr1/rho1 are nsimtr x NT1AMX, ctr2/rho2 nsimtr x NT2AMX, vector-major.  The forward
 * intermediates at the current amplitudes are built on the first call and
 * reused until the amplitudes or the precision change.  The DIIS history is
 * kept when the transformation's buffers fit beside it in device memory,
 * and released otherwise (a later iterate then restarts the DIIS cycle). */
int ccgpu_lhtr(ccgpu_ctx* ctx, int precision, int nsimtr, const double* ctr1, const double* ctr2, double* rho1,
               double* rho2);

/* Correlation energy of host amplitudes t1 (NT1AMX) and packed t2 (NT2AMX):
 * CCSD_ECCSD (cc/ccsd_energy.F:1964) for canonical Hartree-Fock orbitals,
 * E = sum_{ai,bj} L_iajb (t_aibj + t_ai t_bj), L_iajb = 2(ia|jb) - (ib|ja).
 * The context's own amplitudes are not touched. */
int ccgpu_energy(ccgpu_ctx* ctx, const double* t1, const double* t2, double* e_corr);

/* One DIIS step for a caller that keeps Dalton's iteration loop: CCSD_DIIS
 * (cc/ccsd_diis.F:251), with s the new estimate and t the previous
 * amplitudes, both [t1; packed t2] (NTAMP = NT1AMX + NT2AMX values); t
 * receives the extrapolated vector.  niter is the loop's ITER: the cycle
 * position niter - ((niter - 1)/MXDIIS)*MXDIIS picks the slot, and the
 * subspace restarts every MXDIIS calls as in CCSD_DIIS.  The history lives
 * in device memory of the context, shared with ccgpu_ccsd_iterate (use one
 * or the other within a cycle).  Returns 3 on a singular DIIS matrix (t then
 * holds s). */
int ccgpu_diis(ccgpu_ctx* ctx, const double* s, double* t, int niter);

/* CCSD iterations from the context's current amplitudes (a restart, as
 * CCSD_ENERGY with .RESTAR, cc/ccsd_energy.F:558, after
 * ccgpu_restart_amplitudes or ccgpu_set_amplitudes): the loop and arguments
 * of ccgpu_ccsd_energy, with e_start the energy of the starting amplitudes in
 * place of the MP2 energy.  The DIIS history starts empty. */
int ccgpu_ccsd_resume(ccgpu_ctx* ctx, int maxit, double threnr, double thrvec, int precision, double* e_start,
                      double* e_ccsd, int* niter, double* history);

/* Dalton's CC response-vector files.  ccgpu_rsp_filename writes (at least 11
 * bytes) the name CC_RWPRE (cc/cc_rdrsp.F:354) gives list `list` ("R0",
 * "L0", "R1", ...), symmetry isym and file index idx: "CCR0_1___0" for the
 * ground-state amplitudes, "CCL0_1___0" for the multipliers. */
int ccgpu_rsp_filename(const char* list, int isym, int idx, char* name);

/* Writes a closed-shell singlet response vector in the records of CC_WRRSP
 * (cc/cc_wrrsp.F:20; CC_WVRSP :811) for a model without R12 or CC3 parts:
 * header (NWRITE, MODFIL), orbital part (2*nallai values of vec0, zeros when
 * vec0 is NULL), singles (nt1), doubles (nt2, Dalton packed; NWRITE = 2 and
 * no doubles when nt2 = 0), then six empty records.  Fortran unformatted
 * sequential framing with 4-byte length markers as gfortran writes it
 * (records over 2147483639 bytes split into subrecords).  intsize is the byte
 * size of the Dalton build's default INTEGER (4, or 8 for 64-bit integers);
 * model is the CC model label ("CCSD"), at most 10 characters. */
int ccgpu_rsp_write(const char* path, const char* model, int intsize, long nallai, long nt1, long nt2,
                    const double* vec0, const double* vec1, const double* vec2);

/* Reads such a file (CC_RDRSP, cc/cc_rdrsp.F:20; CC_RVRSP :1463).  Either
 * integer size is recognised from the header.  vec0/vec1/vec2 may be NULL to
 * skip a part; a part that is read must have exactly 2*nallai, nt1 or nt2
 * values.  model (11 bytes, may be NULL) receives the label, nwrite (may be
 * NULL) the number of leading parts the file holds.  Returns 4 if vec1 or
 * vec2 is requested and the file does not hold that part. */
int ccgpu_rsp_read(const char* path, long nallai, long nt1, long nt2, double* vec0, double* vec1, double* vec2,
                   char* model, int* nwrite);

/* The E intermediates that CCRHSN writes for the response code when RSPIM is
 * set (cc/ccsd_rhs.F:1898-1915), at the context's current amplitudes (one
 * vector-function evaluation in the given precision; the amplitudes and the
 * iteration state are not changed): E1(b,c) (CCRHS_EFCK, Fock included,
 * NMATAB = v*v values, b fastest) and E2(k,j) (NMATIJ = o*o values, k
 * fastest), as CCRHS_E contracts them (cc/ccsd_rhs.F:2280).  Each non-NULL
 * path receives a Fortran unformatted sequential file of one record, the
 * CC_E1IM / CC_E2IM files that CC_LHTR, CC_RHTR and CC_BMAT read
 * (cc/cc_lhtr.F:1678, cc/cc_rhtr.F:1553, cc/cc_bmat.F:4803); e1 and e2 (may
 * be NULL) receive the same values. */
int ccgpu_write_e_intermediates(ccgpu_ctx* ctx, int precision, const char* e1_path, const char* e2_path,
                                double* e1, double* e2);

/* CCRHSN's RSPIM Gamma intermediate (CC_GAMIM, cc/ccsd_rhs.F:1603-1612),
 * the hole-ladder matrix CCRHS_A contracts and CC_LHTR reads back for its
 * LT22A term (cc/cc_lhtr.F:1732-1747): Gamma(ki,lj) = (ki~|lj~) +
 * sum_cd t_ij^cd (kc|ld) at the context's current amplitudes (one residual
 * evaluation; the amplitudes are not changed).  Packed as Dalton's NGAMMA =
 * NMATIJ (NMATIJ + 1) / 2 vector: pair ki = k + o i (NMATIJ order), element
 * (ki, lj) with ki <= lj at lj (lj + 1) / 2 + ki, 0-based
 * (cc/ccsd_rhs.F:5158-5167).  Written as one Fortran record to path (if not
 * NULL) and copied to gamma (if not NULL). */
int ccgpu_write_gamma_intermediate(ccgpu_ctx* ctx, int precision, const char* path, double* gamma);

/* Writes the context's amplitudes as the 'R0' file of a CCSD run
 * (CC_WRRSP('R0', ...) at the end of CCSD_ENERGY; orbital part zero). */
int ccgpu_save_amplitudes(ccgpu_ctx* ctx, const char* path, int intsize);

/* Reads an 'R0' file (any default-integer size) into the context's
 * amplitudes, as CCSD_ENERGY does for .RESTAR (cc/ccsd_energy.F:558);
 * e_corr (may be NULL) receives their correlation energy.  Continue with
 * ccgpu_ccsd_resume.  Returns 4 if the file holds no doubles. */
int ccgpu_restart_amplitudes(ccgpu_ctx* ctx, const char* path, double* e_corr);

/* Right-hand side of the ground-state multiplier equations at the context's
 * current amplitudes: replaces CC_ETA (cc/cc_fop.F:2955) for MODEL = CCSD
 * with canonical Hartree-Fock orbitals.  eta = dE/dt in the packed storage
 * of ccgpu_lhtr (eta1 NT1AMX, eta2 NT2AMX values), as CC_ETA forms it
 * (2 F^_ia and 2 TCME of (ia|jb) over every packed element, :3019-3069):
 *   eta1[(a,i)] = 2 sum_bj L_iajb t_bj,
 *   eta2[ai<=bj] = 2 L_iajb,  L_iajb = 2(ia|jb) - (ib|ja),
 * which is D dE/dt in the notation of ccgpu_lhtr. */
int ccgpu_eta(ccgpu_ctx* ctx, double* eta1, double* eta2);

/* Solves the ground-state multiplier (L0) equations lambda A + eta = 0 at
 * the context's current amplitudes: replaces the CCEQ_SOL call for the list
 * 'L0' in CC_FOP (cc/cc_fop.F:274; cc/cceq_solv.F:20, ISIDE = -1) by a
 * reduced-space solver whose trial vectors stay in device memory.  lambda A
 * is ccgpu_lhtr and eta ccgpu_eta, both in Dalton's left-vector convention,
 * so lambda = D lambda_p with lambda_p the multiplier of
 * L = E + sum_p lambda_p,p Omega_p in plain packed dot products.  Converged
 * when the 2-norm of J^T lambda_p + dE/dt (the packed-gradient form of the
 * equations, at most 2x smaller than that of lambda A + eta) is below thr
 * (Dalton's THRLEQ); at most maxit
 * transformations, maxred trial vectors (capped at 64 and by free device
 * memory; a full space collapses onto the current solution).  precision 1
 * (BF16x3 bulk) runs an iterative refinement: the FP64 residual of lambda is
 * formed once per step and the correction solved with BF16x3 transformations,
 * so the result meets thr in FP64 arithmetic; niter counts both kinds.
 * l1 (NT1AMX) and l2 (NT2AMX) receive lambda; niter and resnorm (may be
 * NULL) the transformation count and final residual norm.  Returns 4 if thr
 * is not reached within maxit (the last lambda is still returned). */
int ccgpu_lambda_solve(ccgpu_ctx* ctx, int precision, double thr, int maxit, int maxred, double* l1, double* l2,
                       int* niter, double* resnorm);

/* Transformations made by the context's last ccgpu_lambda_solve, by
 * arithmetic: n_fp64 FP64, n_mixed BF16x3 (the sum is its niter).  No
 * counterpart in the reference (CCEQ_SOL prints its iteration count). */
int ccgpu_lambda_stats(ccgpu_ctx* ctx, int* n_fp64, int* n_mixed);

/* Unrelaxed one-electron density in the MO basis from the amplitudes and
 * multipliers lambda (l1, l2 as returned by ccgpu_lambda_solve, Dalton's
 * left-vector convention): CC_D1AO
 * (cc/cc_d1ao.F:20) without its back-transformation to the AO basis.
 * dens[p*N + q] = dL/dh_pq, N = nocc + nvir, reference part 2 delta_ij
 * included, so that an unrelaxed first-order property (CC_FOP) of a
 * one-electron operator X is sum_pq dens[p*N + q] X_pq (times the particle
 * charge, -1 for electrons).  The density is not symmetrised. */
int ccgpu_onedens(ccgpu_ctx* ctx, int precision, const double* l1, const double* l2, double* dens);

/* The density-fitting factor from atomic-orbital integrals, the first step of
 * the in-Dalton drop-in: Dalton's CC code reads AO integral distributions
 * (CCRDAO, cc/ccsd_rhs.F:5652) and MO coefficients (SIRIFC, CMO), and its
 * Cholesky module decomposes the AO integrals (CC_CHOLESKY,
 * choles/cc_cholesky.F:2).  Here: pivoted Cholesky decomposition on the GPU
 * of the AO pair supermatrix V[(ab)][(cd)] = (ab|cd) until the largest
 * remaining diagonal is below thr, then B[P][p][q] = sum C[a][p] L_P[ab] C[b][q].
 * eri_ao: the packed lower triangle of V over pairs (ab) = a(a+1)/2 + b,
 * a >= b, element (P, Q), P >= Q, at P(P+1)/2 + Q (nbas(nbas+1)/2 pairs);
 * cmo: the MO coefficients column by column, C[a][p] at cmo[p*nbas + a]
 * (Dalton's CMO array for one symmetry block), nmo <= nbas columns.
 * B (may be NULL) receives naux x nmo x nmo values; at most maxvec vectors
 * are made (returns 4 if the threshold is not reached within them). */
int ccgpu_df_factor_ao(int nbas, int nmo, const double* cmo, const double* eri_ao, double thr, long maxvec,
                       double* B, long* naux);

/* The factor of ccgpu_df_factor_ao with the AO integrals supplied column by
 * column instead of as the whole supermatrix, as the reference's
 * CC_CHOLESKY computes the integral columns of a set of qualified diagonals
 * at a time (choles/cc_cholesky.F:2) from its AO integral routines.  diag
 * holds the nbas(nbas+1)/2 pair diagonals (ab|ab) (pairs ordered as for
 * ccgpu_df_factor_ao).  Per panel the library picks the (at most 64) largest
 * remaining diagonals that are at least max(thr, 0.01 x the largest) and
 * calls columns(user, ncol, pairs, cols), which writes the integral columns
 * cols[j*M + (cd)] = (cd|ab) for every pair (cd) < M = nbas(nbas+1)/2 and
 * (ab) = pairs[j], and returns 0 (nonzero aborts the call with code 1).  The
 * panel's pivots are chosen among those columns; the decomposition stops
 * when every remaining diagonal is below thr (returns 4 if maxvec vectors do
 * not reach it).  Device memory is O(maxvec x M): no M x M array is formed.
 * B and naux as for ccgpu_df_factor_ao. */
typedef int (*ccgpu_ao_columns_fn)(void* user, int ncol, const long* pairs, double* cols);
int ccgpu_df_factor_ao_cols(int nbas, int nmo, const double* cmo, const double* diag, ccgpu_ao_columns_fn columns,
                            void* user, double thr, long maxvec, double* B, long* naux);

/* ccgpu_init from AO integrals: the factor of ccgpu_df_factor_ao (no vector
 * limit) with nmo = nocc + nvir, then the context as ccgpu_init; *naux
 * receives the number of Cholesky vectors. */
int ccgpu_init_ao(int nocc, int nvir, int nbas, const double* eps, const double* cmo, const double* eri_ao,
                  double thr, int mxdiis, ccgpu_ctx** ctx, long* naux);

/* MP2 starting guess and its energy: CCSD_GUESS (cc/ccsd_energy.F:1654) followed
 * by CCSD_ECCSD (cc/ccsd_energy.F:1964).  e_mp2 is the correlation energy. */
int ccgpu_ccsd_guess(ccgpu_ctx* ctx, double* e_mp2);

/* One pass of the CCSD_ENERGY iteration (cc/ccsd_energy.F:742-982): CCRHSN,
 * the 2-norm of the vector function, CCSD_NXTAM (:1814), CCSD_DIIS
 * (cc/ccsd_diis.F:251, restart every MXDIIS vectors) and CCSD_ECCSD.
 * e_corr: new correlation energy; omega_norm: norm of the residual that was
 * used for this step (Dalton's OMNM). */
int ccgpu_ccsd_iterate(ccgpu_ctx* ctx, int precision, double* e_corr, double* omega_norm);

/* The whole loop: guess, then iterate until |dE| < threnr and norm < thrvec
 * (cc/ccsd_energy.F:946-949) or maxit (returns 4, "CC equations not
 * converged", :952-956).  history (may be NULL, else room for maxit values)
 * receives the energy of each iteration made, at most maxit. */
int ccgpu_ccsd_energy(ccgpu_ctx* ctx, int maxit, double threnr, double thrvec, int precision,
                      double* e_mp2, double* e_ccsd, int* niter, double* history);

/* Many independent closed-shell systems converged together (no reference
 * counterpart: it runs the CCSD_ENERGY loop of cc/ccsd_energy.F:742-982 for
 * nsys systems at once).  System k has nocc[k], nvir[k], naux[k] and host
 * arrays eps[k], B[k] as for ccgpu_init.  ladder_budget_bytes bounds each
 * system's particle-ladder batch buffers (<= 0: 128 v^3 bytes per system,
 * clamped to 256 MB .. 10 GB).  ccgpu_batch_energy converges every system from its MP2 start with
 * Dalton's test (|dE| < threnr and norm < thrvec; mixed precision switches to
 * FP64 as in ccgpu_ccsd_energy): one iteration of all still-active systems is
 * one HIP graph replay, systems running as concurrent graph branches, after
 * an autotuning pass over the FP64 GEMM shapes.  e_mp2/e_ccsd/niter receive
 * nsys values.  flags: CCGPU_BATCH_EAGER issues the same work without a
 * graph, CCGPU_BATCH_NO_TUNE skips the autotuning pass, CCGPU_BATCH_TUNE_ONLY
 * runs only the autotuning pass (outputs untouched; a later call with
 * CCGPU_BATCH_NO_TUNE then uses its choices). */
typedef struct ccgpu_batch ccgpu_batch;
#define CCGPU_BATCH_EAGER 1
#define CCGPU_BATCH_NO_TUNE 2
#define CCGPU_BATCH_TUNE_ONLY 4
int ccgpu_batch_init(int nsys, const int* nocc, const int* nvir, const int* naux, const double* const* eps,
                     const double* const* B, int mxdiis, double ladder_budget_bytes, ccgpu_batch** batch);
int ccgpu_batch_energy(ccgpu_batch* batch, int maxit, double threnr, double thrvec, int precision, int flags,
                       double* e_mp2, double* e_ccsd, int* niter);
/* Converged amplitudes of system k (layouts of ccgpu_get_amplitudes). */
int ccgpu_batch_get_amplitudes(ccgpu_batch* batch, int k, double* t1am, double* t2am);
/* Number of graph captures and replays of the last ccgpu_batch_energy. */
int ccgpu_batch_stats(ccgpu_batch* batch, int* captures, int* replays);
int ccgpu_batch_finalize(ccgpu_batch* batch);

/* Amplitudes in Dalton's layouts (CC_WRRSP/CC_RDRSP 'R0' content, cc/cc_wrrsp.F:20). */
int ccgpu_get_amplitudes(ccgpu_ctx* ctx, double* t1am, double* t2am);
int ccgpu_set_amplitudes(ccgpu_ctx* ctx, const double* t1am, const double* t2am);

/* (T) energy correction from the context's current amplitudes, returned as
 * Dalton's two parts: E4 ("doubles and triples") and E5 ("singles and
 * triples"), E_(T) = E4 + E5 (cc/ccsd_energy.F:1239-1243, printed :1468-1499).
 * The triples are evaluated occupied-driven over i >= j >= k as in
 * mlccsdpt_e_calc (mlcc/mlccsdpt_e_calc.F90:71, loop :326-411).  Releases
 * the iteration workspace (a later iterate restarts the DIIS cycle). */
int ccgpu_pt(ccgpu_ctx* ctx, int precision, double* e4, double* e5);

/* Part of the (T) sum: the restricted triples (i >= j >= k, i = j = k
 * excluded, enumerated i-major then j then k) with list positions first,
 * first+stride, ... (count of them).  flops receives their algorithmic flop
 * count (12 v^4 + 12 o v^3 per triple), ntriples the length of the full list. */
int ccgpu_pt_partial(ccgpu_ctx* ctx, int precision, long long first, long long count, long long stride,
                     double* e4, double* e5, double* flops, long long* ntriples);

/* Per-kernel and per-term timing, the counterpart of CCRHSN's per-term
 * timers (cc/ccsd_rhs.F:2161-2232, printed at IPRINT > 9).  Between start and
 * stop every MFMA GEMM launch and every residual phase is bracketed by HIP
 * events on the stream it runs on.  stop synchronises and writes a JSON
 * summary {"kernels": {name: {count, ms, flops}}, "phases": {name: ms}} into
 * buf (NUL-terminated, truncated to len); it returns the full length needed,
 * or a negative error code. */
int ccgpu_profile_start(void);
int ccgpu_profile_stop(char* buf, int len);

/* Sustained matrix-core rate of this GPU, the ceiling the MFMA kernels are
 * read against (no counterpart in the reference: a measurement helper).
 * Runs an MFMA-only loop (no memory traffic) for about 20 ms, four waves per
 * SIMD, on operands with random bits, and returns its rate in TFLOP/s:
 * CCGPU_MIXED the bf16 16x16x32 instruction (each BF16x3 product needs three),
 * CCGPU_FP64 the FP64 16x16x4 one.  Not available in a dry run. */
int ccgpu_mfma_ceiling(int precision, double* tflops);

/* DGEMM with BLAS semantics (column-major, host arrays), the operation
 * Dalton's hot path spends its time in (pdpack/gp_blas.F DGEMM, e.g. the
 * ladder DGEMMs at cc/ccsd_rhs.F:9589-9597), run on the FP64 MFMA kernel. */
int ccgpu_dgemm(char transa, char transb, int m, int n, int k, double alpha, const double* a,
                int lda, const double* b, int ldb, double beta, double* c, int ldc);

/* The same operation with the mixed-precision arithmetic of CCGPU_MIXED:
 * operands rounded to FP32 and split into bf16 hi + lo, products
 * hi*hi + hi*lo + lo*hi on the bf16 matrix cores, FP32 accumulation,
 * alpha/beta applied in FP64 (no reference counterpart: the precision
 * variant of the DGEMM drop-in above). */
int ccgpu_dgemm_mixed(char transa, char transb, int m, int n, int k, double alpha, const double* a,
                      int lda, const double* b, int ldb, double beta, double* c, int ldc);

/* ccgpu_dgemm_mixed for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  a, b and c are DEVICE arrays (same BLAS
 * column-major semantics); the call is stream-ordered on `stream` (a
 * hipStream_t; NULL = the null stream) and returns once the work is
 * enqueued, without synchronising.  It calls no hipMalloc/hipFree: the
 * packed bf16 hi/lo operands live in a fixed per-GPU scratch array of
 * 64 MiB, so 4 * (Mp + Np) * Kp bytes must fit, with Mp, Np = m, n rounded
 * up to 256 (128 under CCGPU_BF16X3_VARIANT=0) and Kp = k rounded up to 64
 * (e.g. m = n = k = 2048 fits); larger products return 1 with a message.
 * Every call on one GPU uses that GPU's array: calls that overlap in time on
 * one GPU (different streams, or threads) race, calls on one stream do not.
 * Results are bit-identical to ccgpu_dgemm_mixed. */
int ccgpu_dgemm_mixed_noalloc(char transa, char transb, int m, int n, int k, double alpha, const double* a,
                              int lda, const double* b, int ldb, double beta, double* c, int ldc, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CCGPU_H */
