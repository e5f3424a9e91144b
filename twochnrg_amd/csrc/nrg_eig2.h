// nrg_eig2.h -- descriptors of the batched eigensolver (Householder
// tridiagonalisation, divide and conquer, compact-WY back-transformation).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// one symmetric block; all offsets in doubles (ints for i_off)
struct E2Block {
  int n;
  int krow;        // rows of Z the back-transformation forms and the output writes (a prefix)
  int64_t a_off;   // caller's matrix (in: full symmetric, out: eigenvector rows)
  int64_t w_off;   // caller's eigenvalue slot
  int64_t q_off;   // n*n  eigenvectors of T, vector-major
  int64_t k_off;   // n*n  gathered rows for the merge GEMM
  int64_t u_off;   // n*n  rank-one eigenvectors
  int64_t dl_off;  // n*n  d_l - lambda_j
  int64_t t_off;   // n*n  merge output / reorder scratch
  int64_t p_off;   // panel scratch of the reduction: 3 x E2_PANEL rows of n
  int64_t v_off;   // 20*n doubles of vectors
  int64_t i_off;   // 8*n ints
  int ts;          // 1: two-stage reduction (dense -> band of width E2_TSB -> tridiagonal)
  int ncref;       // two-stage: columns 0..ncref-1 carry a stage-1 reflector (stored in their rows)
  int64_t s_off;   // two-stage: stage-1 workspace (doubles) in the stage buffer
  int64_t v2_off;  // two-stage: band-reduction reflectors in the reflector buffer
  int64_t f_off;   // two-stage: progress flags (ints) of the bulge chase
};

// ---------------------------------------------------------------------------
// two-stage reduction of large blocks (nrg_eig2s.hip)
// ---------------------------------------------------------------------------
#define E2_TSB 16          // band width (panel width of stage 1)
#define E2_TS_LEAF 256     // target rows of a TSQR leaf
#define E2_TS_LR 512       // most rows of a leaf or of the stacked R factors
#define E2_TS_QMAX 32      // most leaves per panel
#define E2_TS_NK 512       // K chunk of the split-K product X = A22 (Y T)
#define E2_SB_P 8          // sweeps per bulge-chase batch (one workgroup, P x 16 threads)
#define E2_SB_WIN 256      // band columns held by a batch (LDS ring, a power of 2); >= 2 P TSB - P + 1
#define E2_SB_LDB 32       // stored diagonals per band column (band + bulge)
#define E2_SB_MAXB 32      // two-stage blocks per chase launch

// stage-1 workspace of one block (offsets in doubles); rows of YW/WY/U/X/XP
// are relative to the first row below the band of the current panel
struct E2TsOff {
  int64_t yw, wy, u, x, xp, mp, lt, rs, cb, sm, tot;
};
__host__ __device__ inline E2TsOff e2_ts_off(int64_t s_off, int n) {
  const int64_t nn = n, nk = (nn + E2_TS_NK - 1) / E2_TS_NK, nt = (nn + 63) / 64;
  E2TsOff o;
  o.yw = s_off;
  o.wy = o.yw + 32 * nn;
  o.u = o.wy + 32 * nn;
  o.x = o.u + 16 * nn;
  o.xp = o.x + 16 * nn;
  o.mp = o.xp + nk * 16 * nn;
  o.lt = o.mp + nt * 256;
  o.rs = o.lt + E2_TS_QMAX * 16;
  o.cb = o.rs + E2_TS_QMAX * 256;
  o.sm = o.cb + E2_TS_QMAX * 256;   // T (256), U (256), L (256), signs S (16)
  o.tot = o.sm + 800 - s_off;
  return o;
}
// bulge-chase reflectors of block with n rows: sweep s, position j at
// ((j * n + s) * E2_TSB), tau in slot 0, v[1..] after it (v[0] = 1)
__host__ __device__ inline int64_t e2_v2_len(int n) { return ((int64_t)(n + E2_TSB - 2) / E2_TSB + 1) * n * E2_TSB; }
__host__ __device__ inline int e2_sb_nbatch(int n) { return n > 2 ? (n - 2 + E2_SB_P - 1) / E2_SB_P : 0; }

// one panel of one block in a stage-1 panel step
struct E2TsStep {
  int blk, k, m, q;   // panel columns k..k+E2_TSB-1, m = n - k - E2_TSB rows below the band, q leaves
};
// leaf i (or 64-row tile i) of step `step`
struct E2TsItem {
  int step, i;
};
// the two-stage blocks of one bulge-chase launch: workgroup w works on
// block blk[w / G], batches (w % G), (w % G) + G, ...
struct E2SbGrid {
  int nb, G;
  int blk[E2_SB_MAXB];
};
// one tick of the grouped back-transformation of the bulge-chase reflectors:
// workgroup x of block i (cum[i] <= x < cum[i+1]) applies group g = glo[i] +
// (x - cum[i]) / nch[i] to row chunk (x - cum[i]) % nch[i]
struct E2Q2Grid {
  int nb, tick;
  int cum[E2_SB_MAXB + 1];
  int blk[E2_SB_MAXB];
  int glo[E2_SB_MAXB];
  int nch[E2_SB_MAXB];
};

struct E2Leaf {
  int blk, lo, hi, pad;
};

struct E2Merge {
  int blk, lo, mid, hi;
};

// per-merge results of the deflation pass
struct E2MState {
  int k, nd, ngiv, pad;
  double rho, tol;
};

struct E2Gemm {
  int m, n, k;
  int dyn;           // >= 0: m = k = state[dyn].k
  int64_t a_off, b_off, c_off;
  int64_t as0, as1, bs0, bs1, cs0;
  double alpha;
  int accumulate;
  int pad;
};

struct E2Tile {
  int prob, tm, tn, pad;
};

// one compact-WY panel of the back-transformation: reflectors k0..k0+nb-1 of
// block blk (rows of the caller's matrix), the ns partial Gram matrices and
// T^T at g_off / t_off of the auxiliary buffer (64 x 64 each)
struct E2BtPanel {
  int blk, k0, nb, ns;
  int64_t g_off, t_off;
  int voff;   // the reflectors' stored part starts at column k0 + voff (1: one-stage, E2_TSB: two-stage)
  int pad;
};

// one 64-row tile of W2 = (sum of the ns partial products W_s) T^T in a
// back-transformation round: rows row0.. of the krow rows, partials at wp_off
// (krow x 64 each), T^T at t_off, W2 at w2_off (krow x 64)
struct E2BtRows {
  int row0, krow, nb, ns;
  int64_t wp_off, t_off, w2_off;
};

// The live blocks of one reduction launch, passed by value: workgroup x works
// on block blk[i] for cum[i] <= x < cum[i+1], so no workgroup is launched for
// a block that is smaller than the largest or already reduced.
#ifndef E2_TDB_MAXB
#define E2_TDB_MAXB 128
#endif
#define E2_RS_MAXN 208     // largest matrix the register-resident reduction holds
#ifndef E2_PANEL
#define E2_PANEL 16        // columns per panel of the large blocks' reduction
#endif
// a large block's column chain stops at panel boundary k0 (n - k0 <= E2_RS_MAXN)
// and the resident reduction finishes its trailing matrix
struct E2Hand {
  int blk;
  int k0;
};
struct E2TdbGrid {
  int nb;
  int cum[E2_TDB_MAXB + 1];
  int blk[E2_TDB_MAXB];
};
// the panel updates of one launch: problem p owns workgroups cum[p] .. cum[p+1]-1
#define E2_PANEL_MAXP 32
struct E2PanelGrid {
  int np;
  int cum[E2_PANEL_MAXP + 1];
};

extern "C" {
hipError_t e2_launch_tridiag_small(const E2Block* b, const int* list, int nlist, double* A, double* WS, hipStream_t s);
hipError_t e2_launch_tdb_norm(const E2Block* b, const int* list, int nlist, int nmax, int k, double* A, double* WS,
                              hipStream_t s);
// symv(k) and fin(k) over the large blocks: blk_idx / blk_n (host arrays) are
// each block's index in b and its order
hipError_t e2_launch_tdb_symv(const E2Block* b, const int* blk_idx, const int* blk_n, int nlist, int k, double* A,
                              double* WS, hipStream_t s);
// rows s0..s0+E2_PANEL-1 of every block still being reduced rebuilt from their columns
hipError_t e2_launch_tdb_refresh(const E2Block* b, const int* blk_idx, const int* blk_n, int nlist, int s0, double* A,
                                 hipStream_t s);
hipError_t e2_launch_tdb_fin(const E2Block* b, const int* blk_idx, const int* blk_n, int nlist, int k, double* A,
                             double* WS, hipStream_t s);
hipError_t e2_launch_gemm3(const E2Gemm* probs, const E2Tile* tiles, int ntiles, const double* A, const double* B,
                           double* C, hipStream_t s);
// the products of the reduction's panel updates: each problem's m x m result
// over its lower triangle and a band of two 64 x 64 tiles above it; dprobs:
// the np problems on the device, hprobs their host copy (for the launch sizes)
hipError_t e2_launch_panel_gemm(const E2Gemm* dprobs, const E2Gemm* hprobs, int np, const double* A, const double* B,
                                double* C, hipStream_t s);
hipError_t e2_launch_tridiag_hand(const E2Block* b, const E2Hand* hands, int nh, double* A, double* WS, hipStream_t s);
hipError_t e2_launch_dc_init(const E2Block* b, int nb, int nmax, const E2Merge* merges, int nmerge, const int* mbeg,
                             double* WS, hipStream_t s);
hipError_t e2_launch_dc_leaf(const E2Block* b, const E2Leaf* leaves, int nleaf, double* WS, hipStream_t s);
hipError_t e2_launch_dc_level(const E2Block* b, const E2Merge* merges, int nm, int mmax, E2MState* ms, double* WS,
                              int* IW, const E2Gemm* gprobs, const E2Tile* gtiles, int ntiles, hipStream_t s);
// back-transformation: clear the part of each panel's reflector rows that lies
// on or left of the diagonal (so the rows are the reflectors exactly), T^T of
// every panel from its partial Gram matrices, and W2 of one round's row tiles
hipError_t e2_launch_bt_zero(const E2Block* b, const E2BtPanel* pans, int npan, double* A, hipStream_t s);
hipError_t e2_launch_bt_tmat(const E2Block* b, const E2BtPanel* pans, int npan, const double* WS, double* aux,
                             hipStream_t s);
hipError_t e2_launch_bt_w2(const E2BtRows* rows, int nrows, double* aux, hipStream_t s);
// what: 1 eigenvalues to W, 2 the first krow eigenvector rows to A, 3 both
hipError_t e2_launch_output(const E2Block* b, int nb, int nmax, const double* WS, double* A, double* W, int what,
                            hipStream_t s);
// two-stage reduction (nrg_eig2s.hip)
hipError_t e2s_launch_init(const E2Block* b, const int* list, int nlist, int nmax, double* WS, int* FL, hipStream_t s);
hipError_t e2s_launch_leaf(const E2Block* b, const E2TsStep* st, const E2TsItem* leaves, int nleaf, double* A,
                           double* S, hipStream_t s);
hipError_t e2s_launch_root(const E2Block* b, const E2TsStep* st, int nstep, double* A, double* S, double* WS,
                           hipStream_t s);
hipError_t e2s_launch_form(const E2Block* b, const E2TsStep* st, const E2TsItem* leaves, int nleaf, double* A,
                           double* S, hipStream_t s);
hipError_t e2s_launch_fin1(const E2Block* b, const E2TsStep* st, const E2TsItem* tiles, int ntile, double* S,
                           hipStream_t s);
hipError_t e2s_launch_fin2(const E2Block* b, const E2TsStep* st, const E2TsItem* tiles, int ntile, double* S,
                           hipStream_t s);
hipError_t e2s_launch_band(const E2Block* b, const int* list, int nlist, int nmax, const double* A, double* WS,
                           hipStream_t s);
// the bulge chase of the blocks in G (cooperative launch), then d and e to the
// blocks' vector workspace
hipError_t e2s_launch_chase(const E2Block* b, const E2SbGrid& G, double* WS, double* V2, int* FL, hipStream_t s);
int e2s_chase_max_groups();
hipError_t e2s_launch_tri(const E2Block* b, const int* list, int nlist, int nmax, double* WS, hipStream_t s);
hipError_t e2s_launch_q2(const E2Block* b, const E2Q2Grid& G, double* WS, const double* V2, hipStream_t s);
hipError_t e2_launch_gemm(const E2Gemm* probs, const E2Tile* tiles, int ntiles, const double* A, const double* B,
                          double* C, hipStream_t s);
}
