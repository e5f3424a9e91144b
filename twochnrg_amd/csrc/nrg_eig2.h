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
  int64_t p_off;   // panel scratch of the reduction: 3 x 32 rows of n
  int64_t v_off;   // 20*n doubles of vectors
  int64_t i_off;   // 8*n ints
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
#define E2_TDB_MAXB 32
struct E2TdbGrid {
  int nb;
  int cum[E2_TDB_MAXB + 1];
  int blk[E2_TDB_MAXB];
};

extern "C" {
hipError_t e2_launch_tridiag_small(const E2Block* b, const int* list, int nlist, double* A, double* WS, hipStream_t s);
hipError_t e2_launch_tdb_norm(const E2Block* b, const int* list, int nlist, int nmax, int k, double* A, double* WS,
                              hipStream_t s);
// symv(k) and fin(k) over the large blocks: blk_idx / blk_n (host arrays) are
// each block's index in b and its order
hipError_t e2_launch_tdb_symv(const E2Block* b, const int* blk_idx, const int* blk_n, int nlist, int k, double* A,
                              double* WS, hipStream_t s);
// rows s0..s0+31 of every block still being reduced rebuilt from their columns
hipError_t e2_launch_tdb_refresh(const E2Block* b, const int* blk_idx, const int* blk_n, int nlist, int s0, double* A,
                                 hipStream_t s);
hipError_t e2_launch_tdb_fin(const E2Block* b, const int* blk_idx, const int* blk_n, int nlist, int k, double* A,
                             double* WS, hipStream_t s);
hipError_t e2_launch_gemm3(const E2Gemm* probs, const E2Tile* tiles, int ntiles, const double* A, const double* B,
                           double* C, hipStream_t s);
hipError_t e2_launch_tridiag_tail(const E2Block* b, const int* list, int nlist, const double* A, double* WS,
                                  hipStream_t s);
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
hipError_t e2_launch_gemm(const E2Gemm* probs, const E2Tile* tiles, int ntiles, const double* A, const double* B,
                          double* C, hipStream_t s);
}
