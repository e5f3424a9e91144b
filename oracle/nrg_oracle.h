/*
 * nrg_oracle.h -- CPU restatement (FP64, single thread) of the TwoChNRG
 * iterative-diagonalisation path.  TEST INFRASTRUCTURE ONLY: this code is the
 * checker that the HIP path is compared against, and the timed "port" CPU
 * baseline.  Nothing in the product library links or calls it.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef NRG_ORACLE_H
#define NRG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One NRG step as seen after DiagHN + SetKept (Aeig) and CutStates (AeigCut). */
typedef struct {
  int nshell;          /* N */
  int nq;              /* quantum numbers per block */
  int nblocks;         /* Aeig blocks */
  const double *qn;    /* nblocks*nq */
  const int *bbe;      /* BlockBegEnd, 2*nblocks (inclusive) */
  int nstates;
  const double *en;    /* rescaled energies (after SetE0zero + SetDegen) */
  const unsigned char *kept; /* Kept flags (nstates) */
  double emin;         /* Emin subtracted by SetE0zero */
  double ecut;         /* Ecut used by SetKept */
  int nkept;           /* kept states of Aeig */
  /* optional per-block data: packed upper H (row-wise) and eigvec rows */
  const double *hpacked;   /* sum n(n+1)/2, NULL if not recorded */
  const double *evec;      /* sum n*n, NULL if not recorded */
  /* thermodynamics on Aeig at the 4 betabar values of CalcThermo */
  double thermo_T[4], suscep[4], entropy[4];
  /* CalcStuff: op averages on AeigCut at betabar (T = DN/betabar) */
  double T_avg;
  int nops_avg;
  double opavg[8];
  /* AeigCut structure */
  int cut_nblocks;
  const double *cut_qn;
  const int *cut_bbe;
  const int *cut_bbe_bc;
} orc_step;

typedef void (*orc_step_cb)(void *user, const orc_step *st);

/* OneChQS Anderson through the NRG_main driver
 * (src/Main/NRG_main.cpp:45-427, ModelSwitch src/CodeHandler_ModelSwitch.cpp:121-293). */
typedef struct {
  double U, Gamma, ed, Lambda;
  int nsitesmax;       /* loop runs Nsites = 0 .. nsitesmax-1 */
  int ncut;
  int calcdens;        /* 0: n_d,n_d^2,S_d^2 op averages; 2: thermo + ops; 1: f_N + c_d */
  int disc_scheme;     /* 1: no sqrt(A_Lambda) factor on chi_{-1} (NRG_main -D 1) */
  double myeps;        /* CNRGarray::MyEPS (default 1e-18) */
  double betabar;      /* default 0.727 */
  int record_blocks;   /* pass H and eigvec blocks to the callback */
} orc_onechqs_params;

int orc_run_onechqs(const orc_onechqs_params *p, orc_step_cb cb, void *user);

/* Derivatives of the same loop with every step's Kept set frozen (forward
 * mode; no counterpart in the reference).  kept + kept_off[N] (N = 0 ..
 * nsitesmax-1) gives Aeig's Kept flags of step N, or kept = NULL for the
 * usual SetKept.  Tangent t is the direction dpar[3t..3t+2] = (dU, dGamma,
 * ded).  The callback receives the usual record and the tangent record:
 * d(rescaled E) of Aeig, d(T chi) and dS at the 4 CalcThermo betabar, and
 * d<O> of the CalcOpAvg operators.  Needs Gamma > 0 when ntan > 0. */
#define ORC_MAX_TAN 4
typedef struct {
  int nshell, ntan, nstates;
  const double *den;                  /* ntan x nstates */
  double dsuscep[ORC_MAX_TAN][4], dentropy[ORC_MAX_TAN][4];
  int nops_avg;
  double dopavg[ORC_MAX_TAN][8];
} orc_tan_step;
typedef void (*orc_tan_cb)(void *user, const orc_step *st, const orc_tan_step *dt);
int orc_run_onechqs_jvp(const orc_onechqs_params *p, const unsigned char *kept, const long long *kept_off, int ntan,
                        const double *dpar, orc_tan_cb cb, void *user);

/* TwoChQS Anderson through the legacy standalone driver
 * (src/TwoChQS/TwoChQS.cpp:23-504, ModelNo 0): Gamma~ = Gamma*(2/pi)/HLF^2,
 * chi_{-1} = sqrt(Gamma~/Lambda) per channel (no A_Lambda), one betabar per N.
 * Loop Nsites = 1 .. nsitesmax after the N=0 set-up (impurity + first site). */
typedef struct {
  double U, Gamma1, Gamma2, ed, Lambda;
  int nsitesmax;
  int ncut;
  double betabar;
  int record_blocks;
} orc_twochqs_params;

int orc_run_twochqs(const orc_twochqs_params *p, orc_step_cb cb, void *user);

/* TwoChQSz Anderson through its standalone driver (src/TwoChQSz/TwoChQSz.cpp:103-425,
 * the part after the self-test exit at :100): (Q, Sz) blocks, same parameters
 * and conventions as orc_run_twochqs; T*chi = <Sz^2>, entropy with degeneracy 1. */
int orc_run_twochqsz(const orc_twochqs_params *p, orc_step_cb cb, void *user);

/* DM-NRG on the OneChQS chain (NRG_main with calcdens 1, UpdateBefCut 0 or 1)
 * followed by the DM_NRG post-processing (src/DM_NRG/DM_NRG.cpp:28-531):
 * plain DM-NRG (UseCFS 0), CFS (-C) or FDM (-F); spectral function of c_d. */
typedef struct {
  double U, Gamma, ed, Lambda;
  int nsitesmax, ncut, disc_scheme;
  double myeps, betabar_nrg;  /* NRG run: MyEPS, betabar (0 -> defaults) */
  int Mtemp;                  /* DM_NRG -M */
  double betabar;             /* DM_NRG -b */
  double twindow, broadtemp, bbroad;   /* -w, -B, -e (0 -> 2, 0.8, ln(Lambda)/2) */
  int Nw;                     /* -n: 0 -> CalcSpecDM_NRG_FixedOmegas, >0 -> CalcSpec_ManyOmegas */
  int method;                 /* 0: DM-NRG (UseCFS 0); 1: CFS (-C); 2: FDM (-F) */
  int update_bef_cut;         /* NRG run UpdateBefCut (src/Main/NRG_main.cpp:334-343,393-396) */
  int model;                  /* 0: OneChQS (NRG_main); 1: TwoChQSz (standalone loop, Gamma1 = Gamma), which
                                 the reference never hands to DM_NRG: every shell keeps all its states and
                                 d_up on them, Sz blocks (no CG factors) */
  double Gamma2;              /* TwoChQSz second channel */
  double fdm_site_dim;        /* FDM environment count per later site: 0 -> 4 for model 0 (the reference's
                                 value, src/DM_NRG/DM_NRG.cpp:244,365) and 16 for model 1 (two channels) */
  int fdm_fix;                /* FDM only: ORC_FDM_FIX_* bits switch the reference's FDM quirks off one by
                                 one; 0 = the reference, ORC_FDM_EXACT = the full density matrix of
                                 Weichselbaum and von Delft (sum rule exact) */
} orc_dmnrg_params;

/* Quirks of the reference's FDM path, each switched off by one bit of fdm_fix. */
#define ORC_FDM_FIX_RHOMAX 1    /* rho_Nmax's exponent at the same shell scale as Z_tot and w_Nmax
                                   (reference: Temp/D_Mtemp, src/DM_NRG/DM_NRG.cpp:102-118,273-297) */
#define ORC_FDM_FIX_DN 2        /* shell N's Boltzmann factors at D_N/Temp (reference: D_{N-1},
                                   src/DM_NRG/DM_NRG.cpp:203-214,340) */
#define ORC_FDM_FIX_LAST 4      /* the last shell as an FDM shell with rho_Nmax = e^{-bE}/Z_tot (reference:
                                   the Costi form with its own 1/Z, src/SpecFunc_allSTL.cpp:1270-1277) */
#define ORC_FDM_FIX_PAIRS 8     /* block pairs without a kept state keep their discarded-discarded terms
                                   (reference skips them, src/SpecFunc_allSTL.cpp:1375-1381) */
#define ORC_FDM_FIX_ONESIDED 16 /* kept-discarded terms with rho on the discarded side also when the other
                                   block has no discarded state (reference: only inside the
                                   both-blocks-discarded branch) */
#define ORC_FDM_FIX_NORM 32     /* every pole carries its whole weight at both signs of omega and CalcNorm
                                   counts it once (reference: DD terms in Positive and Negative, so CalcNorm
                                   counts them twice, src/SpecFunc_allSTL.cpp:2326-2348) */
#define ORC_FDM_EXACT 63

typedef struct orc_dm orc_dm;

typedef struct {
  int nblocks, nq, nstates;
  const double *qn; const int *bbe; const double *en;    /* Acut_N */
  const int *bbe_bc; const double *ev; const int *bbe_ev;
  const double *rho; const int *rho_mbe;                 /* rho_N packed upper, block diagonal */
  double trace;                                          /* sum (2S+1) rho_ii */
  int op_nblocks; const int *op_map; const int *op_mbe; const double *op;   /* c_d blocks */
  int basis_nstates, basis_nblocks; const int *basis_stcf, *basis_itype;     /* Abasis_N */
  const double *basis_qn; const int *basis_bbe;
  const int *kept;                                       /* Acut_N Kept flags (nstates) */
  double wn;                                             /* FDM weight w_N (0 unless FDM) */
} orc_dm_shell_view;

orc_dm *orc_dm_run(const orc_dmnrg_params *p);
/* The same run with every NRG step's Kept set given (as orc_run_onechqs_jvp;
 * kept = NULL: SetKept) and ntan tangents along dpar (dU, dGamma, ded) carried
 * through the chain, rho_Nmax and the backward sweep (DM-NRG, method 0,
 * UpdateBefCut 0 only; NULL otherwise).  No counterpart in the reference. */
orc_dm *orc_dm_run_jvp(const orc_dmnrg_params *p, const unsigned char *kept, const long long *kept_off, int ntan,
                       const double *dpar);
/* d/dtheta of orc_dm_spec_shell(D, N, omegabar) along each tangent (out[ntan]) */
int orc_dm_spec_shell_jvp(const orc_dm *D, int N, double omegabar, double *out);
void orc_dm_free(orc_dm *D);
int orc_dm_info(const orc_dm *D, int *nshell_max, int *spec_nshell_max, double *temp);
int orc_dm_shell(const orc_dm *D, int N, orc_dm_shell_view *v);
/* Shell sum at omegabar: DMNRG_SpecDens_M_STL (src/SpecFunc_allSTL.cpp:223-692),
 * CFS_SpecDens_M_STL (:702-1240) or FDM_SpecDens_M_STL (:1253-2354) by method;
 * the last shell is CalcSpecCosti_T_N (src/SpecFuncClass.cpp:417-662). */
double orc_dm_spec_shell(const orc_dm *D, int N, double omegabar);
/* CSpecFunction::CalcSpecDM_NRG(omega) src/SpecFuncClass.cpp:1827-1885 (CFS/FDM):
 * sum over the shell window [N_w-8, N_w+14) of A_N(omega/D_N)/D_N */
double orc_dm_spec_omega(const orc_dm *D, double omega);
/* CSpecFunction::CalcNorm src/SpecFuncClass.cpp:1887-1930 (CFS/FDM; NaN for DM-NRG) */
double orc_dm_norm(const orc_dm *D);
/* Z_tot of the FDM weights (src/DM_NRG/DM_NRG.cpp:203-268; 0 unless FDM) */
double orc_dm_ztot(const orc_dm *D);
/* rho_0_0 on the DM_NRG_CalcSpecFunc_ij grid (evaluated on the first call) */
int orc_dm_spectrum(orc_dm *D, const double **omega, const double **rho);
/* The omega-grid drivers split from the shell sums: the evaluation points of
 * CalcSpecDM_NRG_FixedOmegas (Nw = 0; src/SpecFuncClass.cpp:1932-2233) or
 * CalcSpec_ManyOmegas (Nw > 0; :2239-2341) for nshell spectral shells
 * (windowed = CFS/FDM), and the spectrum from the values at those points
 * (DM-NRG: shell sum at (shell, wbar) / D_shell; windowed: CalcSpecDM_NRG(omega)). */
int orc_grid_plan(double L, int nshell, int windowed, int Nw, int *shell, double *wbar, double *omega, int *parity);
int orc_grid_combine(int n, const double *omega, const int *parity, const double *val, int Nw, double *w_out,
                     double *rho_out);

/* ---- kernel-level restatements ---- */

/* CGordan  src/CGordan.cpp:34-120 */
double orc_cgordan(double J1, double M1, double J2, double M2, double J, double M);

/* DiagBlock -> dspev_('V','L') on row-wise packed upper storage
 * (src/NRGmatrixClass.cpp:566-654).  W ascending, Z rows = eigenvectors. */
void orc_eigh_packed(int n, const double *ap, double *w, double *z);
/* same, full row-major symmetric input (dsyev_ path, :655-680) */
void orc_eigh_full(int n, const double *a, double *w, double *z);

/* CNRGarray::SetDegen src/NRGarrayClass.cpp:244-268 (in place) */
void orc_setdegen(int n, double *en, double myeps);
/* CNRGarray::Ecut src/NRGarrayClass.cpp:661-691 */
double orc_ecut(int n, const double *en, int ncut);

/* The two cblas_dgemm of UpdateMatrices (src/UpdateMatrices.cpp:481-494):
 * out(ki x kj) = Zi(ki x ni) * F(ni x nj) * Zj(kj x nj)^T */
void orc_rotate(int ni, int nj, int ki, int kj, const double *zi,
                const double *f, const double *zj, double *out);

/* chi_N of SetChainWilson src/NRGchainClass.cpp:132-155 */
double orc_chin(int n, double Lambda);

#ifdef __cplusplus
}
#endif
#endif
