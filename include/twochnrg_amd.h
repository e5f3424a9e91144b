/*
 * twochnrg_amd.h -- C-ABI of the MI355X (gfx950) NRG iterative-diagonalisation
 * hot path.  Plain pointers and sizes only.  Every entry point returns 0 on
 * success and a negative code on error; nrg_last_error() gives the message
 * (thread-local).  No exception crosses this boundary.  Calls are synchronous
 * at the boundary: host buffers passed in are read/written before return.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference tree lgds/TwoChNRG).
 */
#ifndef TWOCHNRG_AMD_H
#define TWOCHNRG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRG_OK 0
#define NRG_ERR_ARG (-1)
#define NRG_ERR_HIP (-2)
#define NRG_ERR_NUMERIC (-3)
#define NRG_ERR_STATE (-4)

typedef struct nrg_ctx nrg_ctx;

/* One context per host thread: owns the HIP stream, device workspaces and
 * per-step scratch.  device = HIP ordinal. */
int nrg_ctx_create(int device, nrg_ctx **out);
void nrg_ctx_destroy(nrg_ctx *ctx);
const char *nrg_last_error(void);
/* Number of HIP devices visible (0 on a machine without a GPU). */
int nrg_device_count(void);

/*
 * Batched symmetric eigensolve of nb uneven blocks, FP64, all vectors.
 * Replaces CNRGmatrix::DiagBlock (src/NRGmatrixClass.cpp:566-696,
 * include/NRGclasses.hpp:598-600) -> dspev_/dsyev_.
 *   n[b]      block dimension
 *   h_off[b]  offset of block b in H: packed != 0 -> n(n+1)/2 doubles, row-wise
 *             upper (ij2rUpTr, src/ij2r.cpp:63-68, what DiagHN builds with
 *             UpperTriangular=true); packed == 0 -> n*n row-major full.
 *   W         sum n doubles, ascending per block, blocks in input order
 *   Z         sum n*n doubles, block b at the prefix sum of n^2, row i =
 *             eigenvector i (the dEigVec layout, src/NRGmatrixClass.cpp:647-654)
 *   info[b]   0 ok, 1 if the QL iteration limit was hit (may be NULL)
 * Host pointers.
 */
int nrg_diag_blocks(nrg_ctx *ctx, int nb, const int *n, const int64_t *h_off, const double *H, int packed,
                    double *W, double *Z, int *info);

/* Eigensolver used by every entry point of this context:
 *   2 (default) Householder tridiagonalisation (one workgroup per block up to
 *     n=256, 256-column slabs above) + divide and conquer + compact-WY
 *     back-transformation, every stage batched over all blocks;
 *   1 one workgroup per block: Householder + implicit QL + back-transformation. */
int nrg_set_eig_impl(nrg_ctx *ctx, int impl);

/* on != 0: "plain" launch mode -- the same kernels, but one symmetry block per
 * eigensolve launch sequence and one block pair per rotation launch (no
 * batching across blocks); the baseline the batched mode is measured against. */
int nrg_set_plain(nrg_ctx *ctx, int on);

/* Streams of the blocked Householder reduction (impl 2, blocks above n=256):
 * the blocks are split into n groups of about equal sum n^2, each reduced by
 * its own chain of column launches on its own stream, so that one group's
 * latency-bound launches overlap another's bandwidth-bound ones (results do
 * not depend on n).  1..4, default 3. */
int nrg_set_red_streams(nrg_ctx *ctx, int n);

/* Two-stage reduction (impl 2) of blocks with n >= min_n (0: none, the
 * default): dense -> band of width 16 (TSQR panels and two GEMMs per panel)
 * -> tridiagonal (pipelined bulge chase in LDS), and a grouped compact-WY
 * back-transformation of the chase before that of the panels.  Replaces the
 * one-column-at-a-time reduction of those blocks inside DiagBlock's
 * eigensolve (reference src/NRGmatrixClass.cpp:566-696, dsyev_/dspev_);
 * results agree with the one-stage path to rounding.  Blocks above
 * 16384 + 16 rows keep the one-stage reduction. */
int nrg_set_two_stage(nrg_ctx *ctx, int min_n);

/* hipGraph mode (impl 2): each eigensolve's launch sequences (the reduction,
 * divide and conquer and back-transformation, with their auxiliary streams)
 * are captured into a graph after their descriptor uploads, instantiated and
 * launched once -- the per-step capture of the NRG iteration's dominant launch
 * sequence.  Two-stage blocks run uncaptured (cooperative launch).  Results
 * are identical to the eager launches.  No reference counterpart (the
 * reference runs DiagBlock on the host). */
int nrg_set_graph(nrg_ctx *ctx, int on);
/* host milliseconds spent capturing, instantiating and launching graphs, the
 * graph nodes and graphs since the last call (then reset) */
int nrg_graph_stats(nrg_ctx *ctx, double *host_ms, long *nodes, long *graphs);

/* Autotuner of the eigensolver's launch configuration for one step shape
 * (impl 2).  Random symmetric blocks of sizes n[0..nb-1] are generated on the
 * device and one eigensolve (every eigenvalue, krow_frac of each block's
 * eigenvector rows: the kept rows of a truncated step) is timed per candidate,
 * best of reps after a warm-up run: the reduction's stream count 1..3, then
 * the two-stage thresholds 2048 and 1024 (where blocks that large exist) with
 * the best stream count, then graph launches with the best of those; a
 * candidate replaces the default (3 streams, one-stage, no graph) only when
 * it is more than 2 % faster.  The choice is applied to ctx (as
 * nrg_set_red_streams / nrg_set_two_stage / nrg_set_graph would) and every
 * candidate's time is written to table
 * (*n_entries of them, at most max_entries stored), the choice to *chosen.  No reference counterpart
 * (LAPACK's dspev_/dsyev_ have no launch configuration). */
typedef struct {
  int red_streams, ts_min, graph, pad;
  double ms;
} nrg_tune_entry;
int nrg_autotune_eig(nrg_ctx *ctx, int nb, const int *n, double krow_frac, int reps, nrg_tune_entry *table,
                     int max_entries, int *n_entries, nrg_tune_entry *chosen);

/* Diagnostics: run the batched solver (impl 2) on one n x n matrix and return
 * the stage outputs -- the tear-modified tridiagonal diagonal d (after the
 * divide-and-conquer set-up), the untorn off-diagonal e, the sorted
 * eigenvalues dd and vector rows qv of the tridiagonal problem, and the final
 * eigenvalues w and eigenvector rows z.  Any output pointer may be NULL. */
int nrg_eig2_stages(nrg_ctx *ctx, int n, const double *A, double *d_out, double *e_out, double *dd_out,
                    double *qv_out, double *w_out, double *z_out);

/* Device-resident variant on the context's stream: dA holds full n*n blocks at
 * a_off (overwritten by eigenvector rows), eigenvalues to dW at w_off. */
int nrg_diag_blocks_dev(nrg_ctx *ctx, int nb, const int *n, const int64_t *a_off, double *dA,
                        const int64_t *w_off, double *dW, int *info);

/*
 * Truncation decision with the reference's exact semantics:
 * CNRGarray::SetDegen (src/NRGarrayClass.cpp:244-268) on en in place, then
 * CNRGarray::Ecut (:661-691) and CNRGarray::SetKept (:461-491).
 *   kept[i] = en[i] <= Ecut; *ecut, *nkept returned.  Host pointers.
 */
int nrg_truncate(int nstates, double *en, int ncut, double myeps, uint8_t *kept, double *ecut, int *nkept);

/*
 * CutStates' eigenvector compaction (src/CutStates.cpp:130-215: RemoveBlock
 * and CNRGarray::RemoveStatesFromBlock, src/NRGarrayClass.cpp:1467-1548).
 * Block b of dZ holds n[b] eigenvector rows of length n[b] at z_off[b] (the
 * dEigVec layout, row i = eigenvector i, energies ascending within the block);
 * its first keep[b] rows (the kept states of an energy-ordered block are its
 * lowest) are copied, in block order, to dZcut; blocks with keep[b] = 0 are
 * dropped.  zcut_off[b] (may be NULL) receives the block's offset in dZcut, or
 * -1 when dropped, *zcut_len (may be NULL) the total length.  dZ and dZcut are
 * device arrays that must not overlap; dZcut = NULL only reports the offsets.
 * Synchronous on the context's stream.
 */
int nrg_compact(nrg_ctx *ctx, int nb, const int *n, const int64_t *z_off, const int *keep, const double *dZ,
                double *dZcut, int64_t *zcut_off, int64_t *zcut_len);

/* CNRGarray::Ecut alone (src/NRGarrayClass.cpp:661-691). */
int nrg_ecut(int nstates, const double *en, int ncut, double *ecut);

/* One rotation problem of UpdateMatrices (src/UpdateMatrices.cpp:481-494):
 * out(ki x kj) = Zi(ki x ni) * F(ni x nj) * Zj(kj x nj)^T, offsets in doubles. */
typedef struct {
  int ni, nj, ki, kj;
  int64_t zi_off, zj_off, f_off, out_off;
} nrg_pair_desc;

/* Batched operator rotation (host buffers).  precision 0 = FP64.
 * Replaces the cblas_dgemm pair of UpdateMatrices and the uBLAS
 * prod(trans(Z),prod(M,Z)) of RotateBlock2BLAS (src/NRGmatrixClass.cpp:1471-1512). */
int nrg_rotate_ops(nrg_ctx *ctx, int npairs, const nrg_pair_desc *pairs, const double *Z, int64_t zlen,
                   const double *F, int64_t flen, double *out, int64_t outlen, int precision);

/* No counterpart in the reference.  nrg_rotate_ops on device buffers that
 * allocates nothing, for callers that cannot call hipMalloc (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * dZ, dF, dOut are device pointers on the current HIP device (zlen, flen,
 * outlen doubles, pair offsets as in nrg_rotate_ops); stream is a hipStream_t
 * (NULL: the null stream); the call returns after the stream has finished it.
 * dOut is zero wherever no pair writes, as nrg_rotate_ops leaves it.  The
 * intermediate products (sum over pairs of kj*ni doubles) and the GEMM
 * descriptors live in a fixed scratch area of NRG_NOALLOC_SCRATCH_BYTES per
 * device, which the library holds until the process ends; a call that needs
 * more fails with NRG_ERR_ARG before any work.  Every call on a device shares
 * that device's scratch area: calls that overlap on one device race. */
#define NRG_NOALLOC_SCRATCH_BYTES (64ll << 20)
int nrg_rotate_ops_noalloc(void *stream, int npairs, const nrg_pair_desc *pairs, const double *dZ, int64_t zlen,
                           const double *dF, int64_t flen, double *dOut, int64_t outlen, int precision);

/* ---------------- per-kernel timing ---------------- */

/* Kernel classes timed with HIP events on the context stream. */
#define NRG_K_EIG 0      /* batched eigensolve (eig_block_kernel) */
#define NRG_K_HN 1       /* H_N assembly */
#define NRG_K_FBASIS 2   /* operator matrix in the product basis */
#define NRG_K_GEMM 3     /* rotation GEMMs */
#define NRG_K_SYMV 4     /* y = A v of the blocked reduction of blocks with n > 256 (one launch per column) */
#define NRG_NKSTAT 5

typedef struct {
  double ms[NRG_NKSTAT];        /* summed kernel time */
  long launches[NRG_NKSTAT];
  double flops[NRG_NKSTAT];     /* algorithmic flops of those launches */
  double bytes[NRG_NKSTAT];     /* algorithmic bytes of those launches */
} nrg_stats;

/* Turn event timing on/off (off by default) and read/reset the counters. */
int nrg_set_timing(nrg_ctx *ctx, int on);
int nrg_get_stats(nrg_ctx *ctx, nrg_stats *out);
int nrg_reset_stats(nrg_ctx *ctx);

/* ---------------- whole-chain driver (NRG_main loop) ---------------- */

/* Per-step record (same layout as the CPU checker's record). */
typedef struct {
  int nshell, nq, nblocks;
  const double *qn;
  const int *bbe;
  int nstates;
  const double *en;
  const unsigned char *kept;
  double emin, ecut;
  int nkept;
  const double *hpacked;
  const double *evec;
  double thermo_T[4], suscep[4], entropy[4];
  double T_avg;
  int nops_avg;
  double opavg[8];
  int cut_nblocks;
  const double *cut_qn;
  const int *cut_bbe;
  const int *cut_bbe_bc;
} nrg_step;

typedef void (*nrg_step_cb)(void *user, const nrg_step *st);

typedef struct {
  double U, Gamma, ed, Lambda;
  int nsitesmax, ncut, calcdens, disc_scheme;
  double myeps, betabar;
  int record_blocks;
} nrg_onechqs_params;

/*
 * OneChQS Anderson model through the NRG_main loop
 * (src/Main/NRG_main.cpp:282-427; ModelSwitch src/CodeHandler_ModelSwitch.cpp:121-293):
 * BuildBasis (host) -> H_N assembly + batched eigensolve (device) ->
 * SetKept/CalcThermo/CutStates (host decisions on device results) ->
 * UpdateMatrices rotations (device) -> CalcStuff.  cb may be NULL.
 * timing_ms (may be NULL) receives the wall time of the step loop.
 */
int nrg_run_onechqs(nrg_ctx *ctx, const nrg_onechqs_params *p, nrg_step_cb cb, void *user, double *timing_ms);

/*
 * Impurity thermodynamics of NRG_main's CalcThermo (src/NRGcodehandlerClass.cpp:627-652)
 * with its chain subtraction (CNRGthermo::AddValue / SaveNValue,
 * src/NRGthermoClass.cpp:10-151): two device chains with the parameters of p,
 * calcdens 2 -- the chain run with the impurity decoupled (U = ed = Gamma = 0),
 * whose values minus the free-orbital 1/8 and 2 ln 2 (dImpValue,
 * src/CodeHandler_ModelSwitch.cpp:217-220) are what the chain files
 * SuscepChain1Ch_QS*.dat / EntropyChain1Ch_QS*.dat hold -- and the impurity run.
 * Writes 4 rows per step N = 0 .. nsitesmax-1 (betabar_nc, nc = 1..4; row
 * 4N+nc-1 = the reference's Nvalue) of 7 columns:
 *   T, Tchi_tot, Tchi_chain, Tchi_imp, S_tot, S_chain, S_imp
 * i.e. the four columns of SuscepImp1ChQS_Anderson*.dat and of
 * EntropyImp1ChQS_Anderson*.dat.  out may be NULL to query *nrows.
 */
int nrg_onechqs_thermo(nrg_ctx *ctx, const nrg_onechqs_params *p, double *out, int cap_rows, int *nrows);

/*
 * Many independent OneChQS instances in lockstep (configuration of 64 seeded
 * Anderson impurities): per step all instances' H_N blocks go into one batched
 * eigensolve.  p is an array of ninst parameter sets (same nsitesmax); cb
 * receives the instance index.  Each instance's records equal those of
 * nrg_run_onechqs with the same parameters.
 */
typedef void (*nrg_multi_cb)(void *user, int instance, const nrg_step *st);
int nrg_run_onechqs_multi(nrg_ctx *ctx, int ninst, const nrg_onechqs_params *p, nrg_multi_cb cb, void *user,
                          double *timing_ms);

typedef struct {
  double U, Gamma1, Gamma2, ed, Lambda;
  int nsitesmax, ncut;
  double betabar;
  int record_blocks;
} nrg_twochqs_params;

/*
 * TwoChQS (two-channel, SU(2)xU(1)) Anderson model through the legacy
 * standalone loop (src/TwoChQS/TwoChQS.cpp:204-495; H_0 from
 * TwoChQS_SetH0Anderson src/TwoChQS/TwoChQS_SetH0Anderson.cpp:15-345, H_N
 * rules of TwoChQS_DiagHN src/TwoChQS/TwoChQS_DiagHN.cpp:15-289, operator
 * update TwoChQS_UpdateMatrixAfterCutting src/TwoChQS/TwoChQS_UpdateMatrices.cpp:277-627,
 * by-value CutStates src/CutStates.cpp:103-160).  Steps N = 0 .. nsitesmax;
 * one betabar per N for the thermodynamics.  Returns NRG_ERR_ARG when ncut is
 * below the N=0 basis size.
 */
int nrg_run_twochqs(nrg_ctx *ctx, const nrg_twochqs_params *p, nrg_step_cb cb, void *user, double *timing_ms);

/*
 * TwoChQSz Anderson through its standalone driver (src/TwoChQSz/TwoChQSz.cpp:103-425,
 * the code after the self-test exit at :100; NRG_main's ModelSwitch has no case
 * for it): (Q, Sz) blocks with no SU(2) (QSz_BuildBasis src/QSz_BuildBasis.cpp:9-117,
 * TwoChQSz_DiagHN src/TwoChQSz/TwoChQSz_DiagHN.cpp:14-133, TwoChQSz_SetH0Anderson
 * src/TwoChQSz/TwoChQSz_SetH0Anderson.cpp:15-251, TwoChQSz_UpdateQm1fQ /
 * _UpdateMatrixAfterCutting src/TwoChQSz/TwoChQSz_UpdateMatrices.cpp:16-341).
 * Parameters and step loop as nrg_run_twochqs; Ncut counts states; T*chi is
 * <Sz^2> (CalcSuscep with totalS false, src/CalcThermo.cpp:16-78).
 */
int nrg_run_twochqsz(nrg_ctx *ctx, const nrg_twochqs_params *p, nrg_step_cb cb, void *user, double *timing_ms);

/* ---------------- derivative of the fixed-truncation iteration ---------------- */

/*
 * No counterpart in the reference (which has no derivative; its "gradient" is
 * a finite difference of whole NRG_main runs).  Forward-mode derivative of the
 * OneChQS NRG_main loop of nrg_run_onechqs with respect to the Anderson
 * parameters, every step's Kept set frozen:
 *   kept + kept_off[N] (N = 0 .. nsitesmax-1; kept_off has nsitesmax+1
 *   entries) are step N's Kept flags in Aeig state order (the `kept` of the
 *   nrg_step records of a plain run), or kept = NULL for the usual SetKept;
 *   tangent t is the direction dpar[3t .. 3t+2] = (dU, dGamma, ded).
 * Per step and tangent: d(rescaled E) of Aeig (SetE0zero's shift included),
 * d(T chi) and dS at the 4 CalcThermo betabar values, d<O> of the CalcOpAvg
 * operators (n_d, n_d^2, S_d^2 with calcdens 0/2).  With ntan = 0 it is
 * nrg_run_onechqs with the given Kept sets (finite differences).  cb receives
 * the usual record and the tangent record.  Needs Gamma > 0 when ntan > 0.
 */
#define NRG_MAX_TAN 4
typedef struct {
  int nshell, ntan, nstates;
  const double *den;                         /* ntan x nstates */
  double dsuscep[NRG_MAX_TAN][4], dentropy[NRG_MAX_TAN][4];
  int nops_avg;
  double dopavg[NRG_MAX_TAN][8];
} nrg_tangent_step;
typedef void (*nrg_tangent_cb)(void *user, const nrg_step *st, const nrg_tangent_step *dt);
int nrg_onechqs_jvp(nrg_ctx *ctx, const nrg_onechqs_params *p, const unsigned char *kept, const int64_t *kept_off,
                    int ntan, const double *dpar, nrg_tangent_cb cb, void *user);

/* ---------------- DM-NRG (density-matrix NRG spectral function) ---------------- */

typedef struct {
  double U, Gamma, ed, Lambda;     /* OneChQS Anderson chain, run with calcdens 1 (f_N, c_d) */
  int nsitesmax, ncut, disc_scheme;
  double myeps, betabar_nrg;       /* NRG run: MyEPS, betabar (0 -> 1e-18, 0.727) */
  int Mtemp;                       /* DM_NRG -M: T = D_Mtemp / betabar */
  double betabar;                  /* DM_NRG -b */
  double twindow, broadtemp, bbroad;   /* -w -B -e (0 -> 2, 0.8, ln(Lambda)/2) */
  int Nw;                          /* -n: 0 -> fixed omegas (even/odd shells), >0 -> Nw omegas per shell */
  int method;                      /* 0: DM-NRG (UseCFS 0); 1: CFS (-C); 2: FDM (-F) */
  int update_bef_cut;              /* NRG run UpdateBefCut (src/Main/NRG_main.cpp:334-343,393-396);
                                      CFS and FDM need 1: the discarded states and their operator
                                      elements are kept per shell */
  int model;                       /* 0: OneChQS Anderson (U, Gamma, ed); 1: TwoChQSz two-channel
                                      Anderson (U, Gamma, Gamma2, ed; the chain of nrg_run_twochqsz),
                                      run with d_up kept per shell and the sweep
                                      over (Q, Sz) blocks.  Model 1 needs method 1 or 2 and no
                                      derivatives. */
  double Gamma2;                   /* model 1: hybridisation of the second channel */
  double fdm_site_dim;             /* FDM states per later site in w_N = Z_N^disc d^{Nmax-N}/Z_tot and
                                      rho_N's Z_tot/d^{Nmax-N}: 0 -> 4 for model 0 (the reference's
                                      value, src/DM_NRG/DM_NRG.cpp:244-247,365-366) and 16 for model 1;
                                      any other positive value is used as given */
  int fdm_fix;                     /* FDM only: NRG_FDM_FIX_* bits switch the reference's FDM quirks off
                                      one by one; 0 = the reference's FDM_SpecDens_M_STL path,
                                      NRG_FDM_EXACT = the full density matrix (Weichselbaum and von Delft),
                                      whose CalcNorm is the sum rule <{B, B^dag}> = 1 */
} nrg_dmnrg_params;

/* Quirks of the reference's FDM path, each switched off by one bit of fdm_fix:
 *   RHOMAX   rho_Nmax's exponent at the scale of Z_tot and w_Nmax (reference: Temp/D_Mtemp,
 *            src/DM_NRG/DM_NRG.cpp:102-118,273-297)
 *   DN       shell N's Boltzmann factors at D_N/Temp (reference: D_{N-1}, src/DM_NRG/DM_NRG.cpp:203-214,240,349)
 *   LAST     the last shell as an FDM shell with rho_Nmax = e^{-bE}/Z_tot (reference: the Costi form
 *            with its own 1/Z, src/SpecFunc_allSTL.cpp:1270-1277)
 *   PAIRS    block pairs with no kept state keep their discarded-discarded terms (reference skips
 *            them, src/SpecFunc_allSTL.cpp:1375-1381)
 *   ONESIDED kept-discarded terms with rho on the discarded side also when the other block has no
 *            discarded state (reference: only inside its both-blocks-discarded branch)
 *   NORM     every pole carries its whole weight at both signs of omega and CalcNorm counts it once
 *            (reference: the discarded-discarded terms sit in Positive and Negative, so CalcNorm counts
 *            them twice, src/SpecFunc_allSTL.cpp:2326-2348) */
#define NRG_FDM_FIX_RHOMAX 1
#define NRG_FDM_FIX_DN 2
#define NRG_FDM_FIX_LAST 4
#define NRG_FDM_FIX_PAIRS 8
#define NRG_FDM_FIX_ONESIDED 16
#define NRG_FDM_FIX_NORM 32
#define NRG_FDM_EXACT 63

typedef struct nrg_dm nrg_dm;

/*
 * DM_NRG main (src/DM_NRG/DM_NRG.cpp:28-531; QS symmetry; DM-NRG, CFS or FDM)
 * on the device: the OneChQS chain (as nrg_run_onechqs with calcdens 1) keeping
 * what SaveArrays writes per shell (src/NRGcodehandlerClass.cpp:887-933), then
 * rho_Nmax (DM_NRG_SetRhoNmax src/DM_NRG/DM_NRG_SetRhoNmax.cpp:9-126; FDM
 * weights :203-297) and the backward sweep (RotateMatrix src/RotateMatrix.cpp:17-99
 * + DM_NRG_CalcRhoN_withSU2 src/DM_NRG/DM_NRG_CalcRhoN.cpp:252-449).  The handle
 * keeps device memory until nrg_dm_destroy; ctx must outlive it.
 */
int nrg_dm_create(nrg_ctx *ctx, const nrg_dmnrg_params *p, nrg_dm **out);
void nrg_dm_destroy(nrg_dm *dm);
/*
 * No counterpart in the reference: nrg_dm_create with every NRG step's Kept set
 * frozen (kept/kept_off as in nrg_onechqs_jvp, from a calcdens-1 run with the
 * same parameters; NULL: SetKept) and ntan tangents along dpar = (dU, dGamma,
 * ded) carried through the chain, rho_Nmax and the backward sweep.  Tangents
 * need method 0 (DM-NRG), update_bef_cut 0 and Gamma > 0; ntan = 0 with kept
 * is the frozen-Kept run of finite differences.  The handle works with every
 * nrg_dm_* call.
 */
int nrg_dm_create_jvp(nrg_ctx *ctx, const nrg_dmnrg_params *p, const unsigned char *kept, const int64_t *kept_off,
                      int ntan, const double *dpar, nrg_dm **out);
/* d/dtheta_t of nrg_dm_spec (the shell sums, Costi form at the last shell):
 * out[i * ntan + t] for the n (shell, omegabar) pairs. */
int nrg_dm_spec_jvp(nrg_dm *dm, int n, const int *shells, const double *omegabar, double *out);
/* NshellMax of the sweep, CSpecFunction::NshellMax and the temperature Temp. */
int nrg_dm_info(const nrg_dm *dm, int *nshell_max, int *spec_nshell_max, double *temp);
/* Wall times (ms): forward chain, backward sweep, pole weights, broadening. */
int nrg_dm_timing(const nrg_dm *dm, double *ms4);
/* No counterpart in the reference: device bytes held by the context's buffers
 * when the sweep starts (every shell's eigenvectors and operators: the most
 * nrg_dm_create holds, besides one shell's temporaries) and now. */
int nrg_dm_memory(const nrg_dm *dm, uint64_t *before_sweep, uint64_t *now);
/* rho_N blocks (k_b x k_b row-major, Acut_N block order) and its trace
 * sum (2S+1) rho_ii (the value CalcRhoN prints); out may be NULL. */
int nrg_dm_rho(nrg_dm *dm, int N, double *out, int64_t cap, int *nblocks, double *trace);
/* The shell sum at omegabar for n (shell, omegabar) pairs: DMNRG_SpecDens_M_STL
 * (src/SpecFunc_allSTL.cpp:223-692), CFS_SpecDens_M_STL (:702-1240) or
 * FDM_SpecDens_M_STL (:1253-2354) by method; the last shell is
 * CalcSpecCosti_T_N (src/SpecFuncClass.cpp:417-662). */
int nrg_dm_spec(nrg_dm *dm, int n, const int *shells, const double *omegabar, double *out);
/* CSpecFunction::CalcSpecDM_NRG(omega) (src/SpecFuncClass.cpp:1827-1885) for n
 * omegas: the sum over the shell window [N_w-8, N_w+14) of A_N(omega/D_N)/D_N.
 * CFS and FDM only (NRG_ERR_ARG for DM-NRG, whose reference path for it is the
 * unused non-STL DMNRG_SpecDens_M). */
int nrg_dm_spec_omega(nrg_dm *dm, int n, const double *omega, double *out);
/* CSpecFunction::CalcNorm (src/SpecFuncClass.cpp:1887-1930), the "NormCFS" that
 * DM_NRG_CalcSpecFunc_ij prints (src/DM_NRG/DM_NRG_CalcSpecFuncs.cpp:77-79); CFS/FDM. */
int nrg_dm_norm(nrg_dm *dm, double *norm);
/* FDM weights w_N (N = 0 .. NshellMax) and Z_tot (src/DM_NRG/DM_NRG.cpp:203-271);
 * zeros unless FDM.  wn may be NULL. */
int nrg_dm_weights(const nrg_dm *dm, double *wn, int cap, double *ztot);
/* rho_0_0 on the grid of DM_NRG_CalcSpecFunc_ij (src/DM_NRG/DM_NRG_CalcSpecFuncs.cpp:73-84):
 * CalcSpecDM_NRG_FixedOmegas (src/SpecFuncClass.cpp:1932-2233) or CalcSpec_ManyOmegas
 * (:2239-2341) -- the two columns SaveOmegaRhow writes.  *n receives the count;
 * omega/rho may be NULL to query it. */
int nrg_dm_spectrum(nrg_dm *dm, double *omega, double *rho, int cap, int *n);

/* ---------------- DM-NRG over the caller's shells ---------------- */

/*
 * One NRG shell N as SaveArrays writes it and DM_NRG reads it back
 * (src/NRGcodehandlerClass.cpp:887-933, src/DM_NRG/DM_NRG.cpp:177-298), as
 * plain arrays in the layout of the reference's objects (host pointers):
 *   Acut_N (CNRGbasisarray): nblocks blocks of nq = 2 quantum numbers ((Q, S),
 *     or (Q, Sz) for model 1) in qn; BlockBegEnd in bbe and BlockBegEndBC in
 *     bbe_bc (inclusive pairs, 2*nblocks); dEn in en (nstates); Kept in kept
 *     (nstates bytes, NULL: every state kept); dEigVec in evec with
 *     BlockBegEndEigVec in bbe_ev: block b's rows (bsize x bsize_bc, row i =
 *     eigenvector i over the block's pre-cut basis) start at bbe_ev[2b];
 *     evec_len doubles in all;
 *   Abasis_N (CNRGbasisarray): basis_nblocks blocks, basis_qn, basis_bbe, and
 *     per basis state basis_itype (site state) and basis_stcf (StCameFrom: the
 *     Acut_{N-1} state it came from; unused at N = 0);
 *   the spectral operator at N (the Mat0 file, CNRGmatrix): op_nblocks block
 *     pairs MatBlockMap in op_map, MatBlockBegEnd in op_mbe (inclusive), row-major
 *     k1 x k2 blocks over Acut_N in op (MatEl, op_len doubles).
 */
typedef struct {
  int nblocks, nq, nstates;
  const double *qn;
  const int *bbe, *bbe_bc, *bbe_ev;
  const double *en;
  const unsigned char *kept;
  const double *evec;
  int64_t evec_len;
  int basis_nblocks, basis_nstates;
  const double *basis_qn;
  const int *basis_bbe, *basis_itype, *basis_stcf;
  int op_nblocks;
  const int *op_map, *op_mbe;
  const double *op;
  int64_t op_len;
} nrg_dm_shell;

/*
 * DM_NRG on shells the caller owns: DM_NRG_SetRhoNmax, DM_NRG_CalcRhoN_withSU2 /
 * DM_NRG_CalcRhoN (include/DM_NRG.hpp:40-55) and the CSpecFunction shell sums
 * (include/SpecFuncClass.hpp:107-149) over shells[0 .. nshells-1] = N = 0 ..
 * nshells-1 (the reference's Nsites0 = 0; nshells = its Nsitesmax).  p gives
 * the DM_NRG switches (Lambda, Mtemp, betabar, twindow, broadtemp, bbroad, Nw,
 * method, fdm_site_dim, fdm_fix) and the site through model (0: the OneChQS
 * site with SU(2) Clebsch-Gordan weights; 1: the two-channel (Q, Sz) site,
 * plain partial traces); the chain parameters are not used.  The arrays are
 * read during the call only.  The handle works with every nrg_dm_* call.
 */
int nrg_dm_create_from(nrg_ctx *ctx, const nrg_dmnrg_params *p, int nshells, const nrg_dm_shell *shells,
                       nrg_dm **out);

/*
 * The reference's .bin hand-off files, byte for byte (int32, 1-byte bool,
 * double, little-endian; CNRGarray::SaveBin/ReadBin src/NRGarrayClass.cpp:778-945,
 * CNRGbasisarray::SaveBin/ReadBin :1947-2075, CNRGmatrix::SaveBin/ReadBin
 * src/NRGmatrixClass.cpp:957-1077):
 * nrg_save_arrays runs the chain of p as nrg_dm_create does and writes, per shell
 * N, what CNRGCodeHandler::SaveArrays (src/NRGcodehandlerClass.cpp:887-933) writes:
 * <dir>/Abasis_<id>_N<N>.bin, <dir>/Acut_<id>_N<N>.bin and <dir>/Mat0_<id>_N<N>.bin
 * (the spectral operator: c_d, or d_up for model 1); *nshells receives the count.
 * nrg_dm_create_from_files reads those files for N = 0 .. nshells-1 (as
 * src/DM_NRG/DM_NRG.cpp:177-298 and :312-434 do) and calls nrg_dm_create_from.
 * nrg_dm_save_rho writes the handle's rho_N as <dir>/rhoDM_<id>_N<N>.bin (the
 * files DM_NRG saves for rho_Nmax and after each CalcRhoN, src/DM_NRG/DM_NRG.cpp:286-297,372-394: packed upper blocks).
 */
int nrg_save_arrays(nrg_ctx *ctx, const nrg_dmnrg_params *p, const char *dir, const char *id, int *nshells);
int nrg_dm_create_from_files(nrg_ctx *ctx, const nrg_dmnrg_params *p, const char *dir, const char *id, int nshells,
                             nrg_dm **out);
int nrg_dm_save_rho(nrg_dm *dm, const char *dir, const char *id);

/*
 * The omega-grid drivers on their own, for shell sums computed elsewhere:
 * CalcSpecDM_NRG_FixedOmegas (Nw = 0, one omega per shell as DM_NRG_CalcSpecFunc_ij
 * calls it; src/SpecFuncClass.cpp:1932-2233) and CalcSpec_ManyOmegas (Nw > 0;
 * :2239-2341) over nshell spectral shells (CSpecFunction::NshellMax); windowed != 0
 * for CFS/FDM, which start two shells later.
 * nrg_spectrum_plan: the n evaluation points in the reference's recording order:
 *   shell[i], omegabar[i] (DM-NRG: the shell sum is evaluated there, then / D_shell),
 *   omega[i] (the recorded omega; CFS/FDM: CalcSpecDM_NRG(omega) is evaluated there),
 *   parity[i] (FixedOmegas: 0 even / 1 odd shell set; ManyOmegas: -1).
 *   All outputs NULL queries *n.
 * nrg_spectrum_combine: the two SaveOmegaRhow columns from value[i] at those points
 *   (FixedOmegas: even and odd sets interpolated linearly on the sorted union and
 *   averaged, gsl_interp_linear as :2112-2222; ManyOmegas: the values as recorded).
 *   omega_out and rho_out hold n doubles.
 */
int nrg_spectrum_plan(double Lambda, int nshell, int windowed, int Nw, int cap, int *n, int *shell, double *omegabar,
                      double *omega, int *parity);
int nrg_spectrum_combine(int n, const double *omega, const int *parity, const double *value, int Nw,
                         double *omega_out, double *rho_out);

#ifdef __cplusplus
}
#endif
#endif
