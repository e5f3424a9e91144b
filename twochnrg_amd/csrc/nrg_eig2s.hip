// nrg_eig2s.hip -- two-stage reduction of large symmetric blocks to
// tridiagonal form, and the back-transformation of its second stage.
//
// Stage 1, dense -> band of width E2_TSB, in panels of E2_TSB columns.  For
// the panel P = A[k+b:, k:k+b] (b = E2_TSB, m rows):
//   * TSQR: the rows are split into q <= 32 leaves of <= 512 rows, each leaf
//     is QR-factored in LDS by one workgroup (Householder), and the stacked
//     b x b R factors once more by one workgroup per panel.
//   * Householder vectors rebuilt from the explicit Q1 = Q [I; 0] (m x b):
//     Q1 - diag(S) = Y U by an LU without pivoting whose signs S are chosen
//     so that |U_jj| >= 1; then P = (I - Y T Y^T) [S R; 0] with
//     T = -U S Y1^{-T} (Y1 the unit lower top of Y).  One panel then costs a
//     handful of launches instead of two per column.
//   * Trailing update A22 -= Y W^T + W Y^T with X = A22 (Y T),
//     W = X - Y (T^T (Y^T X)) / 2: X by a split-K FP64 MFMA GEMM, the update
//     as one rank-2b MFMA GEMM (the grouped GEMM of nrg_eig2.hip).
// Stage 2, band -> tridiagonal, by bulge chasing: sweep s eliminates column
// s with a reflector on rows s+1..s+b and chases the bulge down in steps of
// b rows (task j of sweep s works on rows s+1+jb..s+(j+1)b and the block
// left of them).  Task (s, j) may run once (s-1, j+1) is done, so tasks with
// equal 2s+j are independent.  A workgroup takes a batch of E2_SB_P
// consecutive sweeps, one 16-lane group per sweep, and runs them in that
// tick order on a ring of E2_SB_WIN band columns in LDS: columns are loaded
// from the band in HBM ahead of the batch's first sweep and written back
// behind its last.  Batch q loads a column only after batch q-1 has written
// it back (a progress flag per batch, release/acquire at agent scope), so
// consecutive batches run as a pipeline on different workgroups.  Every
// wait is bounded: a wait that exceeds its time limit sets the block's abort
// flag, every workgroup of that block leaves, and the host reports an error.
// Back-transformation of stage 2: the reflectors of 16 consecutive sweeps at
// one chase position form a compact-WY block (31 x 16 staircase); blocks are
// applied to the kept eigenvector rows along the wavefront of their
// dependencies, one launch per wavefront step.  (Stage 1's reflectors are
// applied by the compact-WY back-transformation of nrg_eig2.hip.)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrg_eig2.h"
#include "nrg_wave.h"

#define TB E2_TSB
#define WAVE 64
#define LRW (E2_TS_LR)

namespace e2s {

using nrgw::dpp_d;

// sum over the 16 lanes of a DPP row; every lane of the row gets the same value
__device__ __forceinline__ double sum16(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  return v;
}
// lane t of this thread's 16-lane row (DPP row_newbcast: a VALU move, no LDS round trip)
template <int T>
__device__ __forceinline__ double bc16(double v) {
  return dpp_d<0x150 + T>(v);
}

// ---------------------------------------------------------------------------
// panel QR in LDS.  The panel is held column-major, P[c][r] (row pitch
// LRW + 1), so that the 64 lanes of a wavefront read consecutive rows of one
// column.  1024 threads: wavefront c owns column c (lanes r = lane (mod 64)),
// so a column's sums are wavefront reductions.
// qr_cm: P (rows x TB) <- [R; V] in LAPACK form (beta on the diagonal, v
// below it, v_j = 1 implicit), tau[j] (0 where no reflector).
// ---------------------------------------------------------------------------
#define QT 1024
#define LDP (LRW + 1)
__device__ __forceinline__ double wsum(double v) { return nrgw::wave_sum(v); }

__device__ void qr_cm(double (*P)[LDP], int rows, double* tau_s, double* scal_s, double* beta_s) {
  const int tid = threadIdx.x, c = tid >> 6, ln = tid & 63;
  const int nref = rows < TB ? rows : TB;
  for (int j = 0; j < nref; j++) {
    if (c == j) {
      double sg = 0.0;
      for (int r = j + 1 + ln; r < rows; r += 64) sg += P[j][r] * P[j][r];
      sg = wsum(sg);
      if (ln == 0) {
        const double alpha = P[j][j];
        double tau = 0.0, scal = 0.0, beta = alpha;
        if (sg > 0.0) {
          beta = -copysign(sqrt(alpha * alpha + sg), alpha);
          tau = (beta - alpha) / beta;
          scal = 1.0 / (alpha - beta);
        }
        tau_s[j] = tau;
        scal_s[j] = scal;
        beta_s[j] = beta;
      }
    }
    __syncthreads();
    const double tau = tau_s[j], scal = scal_s[j];
    if (c > j && tau != 0.0) {
      double dot = 0.0;
      for (int r = j + 1 + ln; r < rows; r += 64) dot += P[j][r] * P[c][r];
      dot = wsum(dot) * scal + P[c][j];
      const double f = tau * dot, fs = f * scal;
      for (int r = j + 1 + ln; r < rows; r += 64) P[c][r] -= fs * P[j][r];
      if (ln == 0) P[c][j] -= f;
    }
    __syncthreads();
  }
  // column j below the diagonal still holds x: scale it to v; beta on the diagonal
  if (c < nref) {
    const double scal = scal_s[c];
    for (int r = c + 1 + ln; r < rows; r += 64) P[c][r] *= scal;
    if (ln == 0) P[c][c] = beta_s[c];
  }
  if (tid >= nref && tid < TB) tau_s[tid] = 0.0;
  __syncthreads();
}

// X (rows x TB, column-major in LDS) <- H_0 H_1 ... H_{nref-1} X with the
// reflectors of V (LAPACK form: the diagonal read as 1) and tau
__device__ void apply_q_cm(const double (*V)[LDP], const double* tau_s, double (*X)[LDP], int rows) {
  const int tid = threadIdx.x, c = tid >> 6, ln = tid & 63;
  const int nref = rows < TB ? rows : TB;
  for (int j = nref - 1; j >= 0; j--) {
    const double tau = tau_s[j];
    if (tau != 0.0) {
      double dot = 0.0;
      for (int r = j + 1 + ln; r < rows; r += 64) dot += V[j][r] * X[c][r];
      dot = wsum(dot) + X[c][j];
      const double f = tau * dot;
      for (int r = j + 1 + ln; r < rows; r += 64) X[c][r] -= f * V[j][r];
      if (ln == 0) X[c][j] -= f;
    }
    __syncthreads();
  }
}

// row r (r < 16) of T for 16 reflectors with Gram matrix G = V^T V and tau,
// in registers: T[r][i] = -tau_i sum_{c=r}^{i-1} T[r][c] G[c][i]
__device__ __forceinline__ void tmat_row(const double (*G)[TB + 1], const double* tau_s, int r, double* trow) {
#pragma unroll
  for (int i = 0; i < TB; i++) {
    double w = 0.0;
#pragma unroll
    for (int c = 0; c < i; c++)
      if (c >= r) w += trow[c] * G[c][i];
    trow[i] = (i < r) ? 0.0 : (i == r ? tau_s[i] : -tau_s[i] * w);
  }
}

__device__ __forceinline__ void leaf_rows(const E2TsStep& st, int i, int& lo, int& hi) {
  lo = (int)(((int64_t)i * st.m) / st.q);
  hi = (int)(((int64_t)(i + 1) * st.m) / st.q);
}

// ---------------------------------------------------------------------------
// stage 1 kernels
// ---------------------------------------------------------------------------
// zero the stage-1 tau slots (read by the back-transformation for every
// column) and the chase flags of every two-stage block
__global__ void init_kernel(const E2Block* __restrict__ blks, const int* __restrict__ list, double* __restrict__ WS,
                            int* __restrict__ FL) {
  const E2Block B = blks[list[blockIdx.y]];
  const int n = B.n;
  double* tau = WS + B.v_off + 2 * (size_t)n;
  int* fl = FL + B.f_off;
  const int nf = 1 + e2_sb_nbatch(n);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) tau[i] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x) fl[i] = 0;
}

// one leaf of one panel: Householder QR in LDS; V and R back in place (the
// panel rows of A), tau and R to the stage workspace
__global__ __launch_bounds__(QT) void leaf_kernel(const E2Block* __restrict__ blks, const E2TsStep* __restrict__ steps,
                                                  const E2TsItem* __restrict__ leaves, double* __restrict__ A0,
                                                  double* __restrict__ S) {
  __shared__ double P[TB][LDP];
  __shared__ double tau_s[TB], scal_s[TB], beta_s[TB];
  const E2TsItem it = leaves[blockIdx.x];
  const E2TsStep st = steps[it.step];
  const E2Block B = blks[st.blk];
  const int n = B.n, r0 = st.k + TB, tid = threadIdx.x;
  int lo, hi;
  leaf_rows(st, it.i, lo, hi);
  const int rows = hi - lo;
  double* A = A0 + B.a_off;
  const E2TsOff o = e2_ts_off(B.s_off, n);
  for (int idx = tid; idx < rows * TB; idx += QT) {
    const int r = idx / TB, c = idx % TB;
    P[c][r] = A[(size_t)(r0 + lo + r) * n + st.k + c];
  }
  __syncthreads();
  qr_cm(P, rows, tau_s, scal_s, beta_s);
  for (int idx = tid; idx < rows * TB; idx += QT) {
    const int r = idx / TB, c = idx % TB;
    A[(size_t)(r0 + lo + r) * n + st.k + c] = P[c][r];
  }
  if (tid < TB) S[o.lt + it.i * TB + tid] = tau_s[tid];
  if (tid < 256) {
    const int r = tid / TB, c = tid % TB;   // one 16 x 16 R
    S[o.rs + it.i * 256 + tid] = (r <= c && r < rows) ? P[c][r] : 0.0;
  }
}

// one panel: T of its compact-WY form.  One leaf: Y = the leaf's reflectors,
// T by the forward recurrence from Y^T Y.  Several leaves: QR of the stacked
// R factors, C = Q_root [I; 0], the top of Q1 = Q_leaf0 [C_0; 0], the signed
// LU of that top (S, U, L), T = -U S L^{-T}, and S R written over the band
// part of the panel.  tau_j = T_jj goes to the block's tau slots.
__global__ __launch_bounds__(QT) void root_kernel(const E2Block* __restrict__ blks, const E2TsStep* __restrict__ steps,
                                                  double* __restrict__ A0, double* __restrict__ S,
                                                  double* __restrict__ WS) {
  __shared__ double P[TB][LDP];
  __shared__ double X[TB][LDP];
  __shared__ double tau_s[TB], scal_s[TB], beta_s[TB], sg[TB];
  __shared__ double M[TB][TB + 1], Tm[TB][TB + 1], Li[TB][TB + 1], Rr[TB][TB + 1];
  const E2TsStep st = steps[blockIdx.x];
  const E2Block B = blks[st.blk];
  const int n = B.n, r0 = st.k + TB, tid = threadIdx.x, q = st.q, m = st.m;
  const bool sq = tid < 256;   // the threads of the 16 x 16 work
  const int ti = (tid / TB) & (TB - 1), tj = tid % TB;
  double* A = A0 + B.a_off;
  const E2TsOff o = e2_ts_off(B.s_off, n);
  double* taucol = WS + B.v_off + 2 * (size_t)n;
  if (q == 1) {
    for (int idx = tid; idx < m * TB; idx += QT) {
      const int r = idx / TB, c = idx % TB;
      P[c][r] = r > c ? A[(size_t)(r0 + r) * n + st.k + c] : (r == c ? 1.0 : 0.0);
    }
    if (tid < TB) { tau_s[tid] = S[o.lt + tid]; sg[tid] = 1.0; }
    __syncthreads();
    if (sq) {
      double gsum = 0.0;
      for (int r = 0; r < m; r++) gsum += P[ti][r] * P[tj][r];
      M[ti][tj] = gsum;   // Gram matrix Y^T Y
    }
    __syncthreads();
    if (tid < TB) {
      double trow[TB];
      tmat_row(M, tau_s, tid, trow);
#pragma unroll
      for (int i = 0; i < TB; i++) Tm[tid][i] = trow[i];
    }
    __syncthreads();
  } else {
    for (int idx = tid; idx < q * 256; idx += QT) {
      const int lf = idx / 256, r = (idx / TB) % TB, c = idx % TB;
      P[c][lf * TB + r] = S[o.rs + idx];
    }
    __syncthreads();
    qr_cm(P, q * TB, tau_s, scal_s, beta_s);
    if (sq) Rr[ti][tj] = ti <= tj ? P[tj][ti] : 0.0;
    for (int idx = tid; idx < q * TB * TB; idx += QT) {
      const int r = idx / TB, c = idx % TB;
      X[c][r] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    apply_q_cm(P, tau_s, X, q * TB);   // X = C = Q_root [I; 0]
    for (int idx = tid; idx < q * TB * TB; idx += QT) {
      const int r = idx / TB, c = idx % TB;
      S[o.cb + idx] = X[c][r];
    }
    __syncthreads();
    // the top of Q1: leaf 0's reflectors applied to [C_0; 0]
    int lo0, hi0;
    leaf_rows(st, 0, lo0, hi0);
    const int rows0 = hi0 - lo0;
    for (int idx = tid; idx < rows0 * TB; idx += QT) {
      const int r = idx / TB, c = idx % TB;
      P[c][r] = A[(size_t)(r0 + r) * n + st.k + c];
      if (r >= TB) X[c][r] = 0.0;
    }
    if (tid < TB) tau_s[tid] = S[o.lt + tid];
    __syncthreads();
    apply_q_cm(P, tau_s, X, rows0);
    // signed LU of the top TB x TB: Q1_top - diag(S) = L U
    if (sq) M[ti][tj] = X[tj][ti];
    __syncthreads();
    for (int j = 0; j < TB; j++) {
      if (tid == 0) {
        const double sj = M[j][j] >= 0.0 ? -1.0 : 1.0;
        sg[j] = sj;
        M[j][j] -= sj;
      }
      __syncthreads();
      if (sq && ti > j && tj == j) M[ti][j] /= M[j][j];
      __syncthreads();
      if (sq && ti > j && tj > j) M[ti][tj] -= M[ti][j] * M[j][tj];
      __syncthreads();
    }
    // Li = L^{-1} (unit lower), column tid
    if (tid < TB) {
      const int c = tid;
      for (int r = 0; r < TB; r++) {
        double v = (r == c) ? 1.0 : 0.0;
        if (r > c) {
          v = 0.0;
          for (int t = c; t < r; t++) v -= M[r][t] * Li[t][c];
        }
        Li[r][c] = v;
      }
    }
    __syncthreads();
    // T[i][j] = -sum_{k=i}^{j} U[i][k] S_k Li[j][k]
    if (sq) {
      double t = 0.0;
      if (tj >= ti)
        for (int k2 = ti; k2 <= tj; k2++) t -= M[ti][k2] * sg[k2] * Li[tj][k2];
      Tm[ti][tj] = tj >= ti ? t : 0.0;
    }
    __syncthreads();
    // S R over the band part of the panel (upper triangle of its top TB rows)
    if (sq) {
      if (ti <= tj) A[(size_t)(r0 + ti) * n + st.k + tj] = sg[ti] * Rr[ti][tj];
      S[o.sm + 256 + tid] = ti <= tj ? M[ti][tj] : 0.0;                         // U
      S[o.sm + 512 + tid] = ti > tj ? M[ti][tj] : (ti == tj ? 1.0 : 0.0);       // L
    }
  }
  if (sq) S[o.sm + tid] = Tm[ti][tj];
  if (tid < TB) {
    S[o.sm + 768 + tid] = sg[tid];
    taucol[st.k + tid] = Tm[tid][tid];
  }
}

// one leaf: its rows of Y (the leaf's reflectors applied to [C_i; 0], then
// (Q1 - S) U^{-1}; one leaf: the reflectors themselves), U = Y T, Y into the
// stage workspace and, transposed, into the rows k..k+TB-1 of A right of the
// band (where the back-transformation reads the reflectors)
__global__ __launch_bounds__(QT) void form_kernel(const E2Block* __restrict__ blks, const E2TsStep* __restrict__ steps,
                                                  const E2TsItem* __restrict__ leaves, double* __restrict__ A0,
                                                  double* __restrict__ S) {
  __shared__ double P[TB][LDP];
  __shared__ double X[TB][LDP];
  __shared__ double tau_s[TB];
  __shared__ double Ts[TB][TB + 1], Us[TB][TB + 1], Ls[TB][TB + 1];
  const E2TsItem it = leaves[blockIdx.x];
  const E2TsStep st = steps[it.step];
  const E2Block B = blks[st.blk];
  const int n = B.n, r0 = st.k + TB, tid = threadIdx.x, q = st.q;
  int lo, hi;
  leaf_rows(st, it.i, lo, hi);
  const int rows = hi - lo;
  double* A = A0 + B.a_off;
  const E2TsOff o = e2_ts_off(B.s_off, n);
  if (tid < 256) {
    const int ti = tid / TB, tj = tid % TB;
    Ts[ti][tj] = S[o.sm + tid];
    Us[ti][tj] = S[o.sm + 256 + tid];
    Ls[ti][tj] = S[o.sm + 512 + tid];
  }
  if (q == 1) {
    for (int idx = tid; idx < rows * TB; idx += QT) {
      const int r = idx / TB, c = idx % TB;
      X[c][r] = r > c ? A[(size_t)(r0 + lo + r) * n + st.k + c] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
  } else {
    for (int idx = tid; idx < rows * TB; idx += QT) {
      const int r = idx / TB, c = idx % TB;
      P[c][r] = A[(size_t)(r0 + lo + r) * n + st.k + c];
      X[c][r] = r < TB ? S[o.cb + (size_t)(it.i * TB + r) * TB + c] : 0.0;
    }
    if (tid < TB) tau_s[tid] = S[o.lt + it.i * TB + tid];
    __syncthreads();
    apply_q_cm(P, tau_s, X, rows);   // this leaf's rows of Q1
    // rows of Y: (Q1 - S) U^{-1}; leaf 0's first TB rows are L
    for (int r = tid; r < rows; r += QT) {
      double y[TB];
      if (it.i == 0 && r < TB) {
#pragma unroll
        for (int c = 0; c < TB; c++) y[c] = Ls[r][c];
      } else {
#pragma unroll
        for (int c = 0; c < TB; c++) {
          double v = X[c][r];
#pragma unroll
          for (int t = 0; t < c; t++) v -= y[t] * Us[t][c];
          y[c] = v / Us[c][c];
        }
      }
#pragma unroll
      for (int c = 0; c < TB; c++) X[c][r] = y[c];
    }
    __syncthreads();
  }
  // U = Y T; Y and U out
  for (int r = tid; r < rows; r += QT) {
    const size_t row = (size_t)(lo + r);
    double y[TB];
#pragma unroll
    for (int c = 0; c < TB; c++) y[c] = X[c][r];
#pragma unroll
    for (int c = 0; c < TB; c++) {
      double u = 0.0;
#pragma unroll
      for (int t = 0; t <= c; t++) u += y[t] * Ts[t][c];
      S[o.u + row * TB + c] = u;
      S[o.yw + row * 32 + c] = y[c];
    }
  }
  for (int idx = tid; idx < rows * TB; idx += QT) {
    const int c = idx / rows, r = idx % rows;
    A[(size_t)(st.k + c) * n + r0 + lo + r] = X[c][r];
  }
}

// 64-row tile of X = sum of the split-K partial products; partial Y^T X
__global__ __launch_bounds__(256) void fin1_kernel(const E2Block* __restrict__ blks, const E2TsStep* __restrict__ steps,
                                                   const E2TsItem* __restrict__ tiles, double* __restrict__ S) {
  __shared__ double xs[64][TB + 1], ys[64][TB + 1];
  const E2TsItem it = tiles[blockIdx.x];
  const E2TsStep st = steps[it.step];
  const E2Block B = blks[st.blk];
  const int n = B.n, m = st.m, tid = threadIdx.x;
  const E2TsOff o = e2_ts_off(B.s_off, n);
  const int nk = (m + E2_TS_NK - 1) / E2_TS_NK;
  const int r = tid >> 2, c0 = (tid & 3) * 4, row = it.i * 64 + r;
  double x[4] = {0.0, 0.0, 0.0, 0.0};
  if (row < m)
    for (int s = 0; s < nk; s++)
#pragma unroll
      for (int u = 0; u < 4; u++) x[u] += S[o.xp + ((size_t)s * n + row) * TB + c0 + u];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    xs[r][c0 + u] = row < m ? x[u] : 0.0;
    ys[r][c0 + u] = row < m ? S[o.yw + (size_t)row * 32 + c0 + u] : 0.0;
    if (row < m) S[o.x + (size_t)row * TB + c0 + u] = x[u];
  }
  __syncthreads();
  const int i = tid / TB, j = tid % TB;
  double acc = 0.0;
  for (int rr = 0; rr < 64; rr++) acc += ys[rr][i] * xs[rr][j];
  S[o.mp + (size_t)it.i * 256 + tid] = acc;
}

// 64-row tile of W = X - Y (T^T (Y^T X)) / 2, then [Y W] and [W Y] for the update
__global__ __launch_bounds__(256) void fin2_kernel(const E2Block* __restrict__ blks, const E2TsStep* __restrict__ steps,
                                                   const E2TsItem* __restrict__ tiles, double* __restrict__ S) {
  __shared__ double Ms[TB][TB + 1], Ts[TB][TB + 1], Zs[TB][TB + 1];
  const E2TsItem it = tiles[blockIdx.x];
  const E2TsStep st = steps[it.step];
  const E2Block B = blks[st.blk];
  const int n = B.n, m = st.m, tid = threadIdx.x;
  const E2TsOff o = e2_ts_off(B.s_off, n);
  const int nt = (m + 63) / 64;
  const int i = tid / TB, j = tid % TB;
  double acc = 0.0;
  for (int t = 0; t < nt; t++) acc += S[o.mp + (size_t)t * 256 + tid];
  Ms[i][j] = acc;
  Ts[i][j] = S[o.sm + tid];
  __syncthreads();
  double z = 0.0;
  for (int l = 0; l < TB; l++) z += Ts[l][i] * Ms[l][j];
  Zs[i][j] = z;
  __syncthreads();
  const int r = tid >> 2, c0 = (tid & 3) * 4, row = it.i * 64 + r;
  if (row < m) {
    double y[TB];
#pragma unroll
    for (int c = 0; c < TB; c++) y[c] = S[o.yw + (size_t)row * 32 + c];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int c = c0 + u;
      double w = 0.0;
#pragma unroll
      for (int t = 0; t < TB; t++) w += y[t] * Zs[t][c];
      w = S[o.x + (size_t)row * TB + c] - 0.5 * w;
      S[o.yw + (size_t)row * 32 + TB + c] = w;
      S[o.wy + (size_t)row * 32 + c] = w;
      S[o.wy + (size_t)row * 32 + TB + c] = y[c];
    }
  }
}

// the band after stage 1, column-major with E2_SB_LDB diagonals per column
// (AB[c][d] = A(c+d, c), zero beyond the band: room for the bulge), in the
// block's vector-matrix workspace
__global__ void band_kernel(const E2Block* __restrict__ blks, const int* __restrict__ list,
                            const double* __restrict__ A0, double* __restrict__ WS) {
  const E2Block B = blks[list[blockIdx.y]];
  const int n = B.n;
  const double* A = A0 + B.a_off;
  double* AB = WS + B.q_off;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)n * E2_SB_LDB;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx / E2_SB_LDB), d = (int)(idx % E2_SB_LDB);
    AB[idx] = (d <= TB && c + d < n) ? A[(size_t)(c + d) * n + c] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// stage 2: bulge chase
// ---------------------------------------------------------------------------
#define SBP E2_SB_P
#define SBW E2_SB_WIN
#define LDB E2_SB_LDB
#define SB_T (SBP * 16)
#ifndef E2_SB_PUB
#define E2_SB_PUB 4   // progress published every E2_SB_PUB ticks (a power of 2)
#endif

__device__ __forceinline__ int ld_acq(int* p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st_rel(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_rlx(int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st_rlx(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// last chase position of sweep s (tasks j = 0..J(s))
__device__ __forceinline__ int jlast(int s, int n) { return (n - 1 - s + TB - 1) / TB - 1; }

#define UNROLL16(M) M(0) M(1) M(2) M(3) M(4) M(5) M(6) M(7) M(8) M(9) M(10) M(11) M(12) M(13) M(14) M(15)

// LDS writes of this wavefront's lanes visible to its other lanes (LDS
// operations of a wavefront complete in order; the fence keeps the compiler
// from moving accesses across it)
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// band columns that tick tk of a batch starting at sweep s0 brings in:
// [s0, s0+17) at tick 0, [s0+16tk+1, s0+16tk+17) after (clipped at n)
__device__ __forceinline__ void tick_cols(int s0, int n, int tk, int& lo, int& hi) {
  lo = tk == 0 ? s0 : s0 + TB * tk + 1;
  hi = s0 + TB * (tk + 1) + 1;
  lo = min(lo, n);
  hi = min(hi, n);
}

// One workgroup = one batch of SBP sweeps at a time; group g (16 lanes) runs
// sweep s0+g, lane t owning column t of the task's blocks.  A tick: the
// columns prefetched two ticks earlier go into the ring, every sweep does its
// task, the columns no later task of the batch touches go back to the band,
// and every fourth tick the progress is published (each wavefront fences its
// write-backs first).  Loads of tick tk+2's columns are issued at tick tk,
// after the predecessor batch has published them.
__global__ __launch_bounds__(SB_T) void chase_kernel(const E2Block* __restrict__ blks, const E2SbGrid G,
                                                     double* __restrict__ WS, double* __restrict__ V2,
                                                     int* __restrict__ FL) {
  // the ring: band column c at slot c mod SBW, RS = LDB + 2 doubles per slot
  // (a task's lanes walk a column, stride 1, or a diagonal (c + t, d - t),
  // stride RS - 1: both odd, so 16 lanes hit 16 different banks); after it
  // one dummy double per lane, the target of masked stores (so that the task
  // code has no per-element branches)
  constexpr int RS = LDB + 2;
  __shared__ double Wf[SBW * RS + 64];
  __shared__ int sh_abort;
  const int bi = blockIdx.x / G.G, gq = blockIdx.x % G.G;
  if (bi >= G.nb) return;
  const E2Block B = blks[G.blk[bi]];
  const int n = B.n, tid = threadIdx.x, grp = tid >> 4, t = tid & 15;
  const int DUM = SBW * RS + (tid & 63);
  double* AB = WS + B.q_off;
  double* VB = V2 + B.v2_off;
  int* abortf = FL + B.f_off;
  int* ret = abortf + 1;
  const int nbatch = e2_sb_nbatch(n);
  constexpr int PF = (TB * LDB) / SB_T;   // prefetched doubles per thread per tick
#define WA(c, d) ((((c) & (SBW - 1)) * RS) + (d))
  for (int q = gq; q < nbatch; q += G.G) {
    const int s0 = q * SBP, s1 = min(s0 + SBP, n - 2);
    const int s = s0 + grp;   // this group's sweep
    int retired_lo = s0, published = s0;
    int known = q > 0 ? 0 : n;   // thread 0: predecessor's progress last read
    // thread 0 waits until the predecessor has written back every column below hi
    auto wait_pred = [&](int hi) -> int {
      if (q == 0 || known >= hi) return 0;
      known = ld_acq(&ret[q - 1]);
      const uint64_t t0 = wall_clock64();
      while (known < hi) {
        if (ld_rlx(abortf)) return 1;
        if (wall_clock64() - t0 > 200000000ull) { st_rlx(abortf, 1); return 1; }   // ~2 s at 100 MHz
        __builtin_amdgcn_s_sleep(1);
        known = ld_acq(&ret[q - 1]);
      }
      return 0;
    };
    double pf[2][PF];
    // prologue: tick 0's columns straight into the ring, ticks 1 and 2 into the prefetch buffers
    {
      int lo, hi;
      tick_cols(s0, n, 2, lo, hi);
      if (tid == 0) sh_abort = wait_pred(hi);
      __syncthreads();
      if (sh_abort) return;
      tick_cols(s0, n, 0, lo, hi);
      for (int idx = tid; idx < (hi - lo) * LDB; idx += SB_T) {
        const int c = lo + idx / LDB, d = idx % LDB;
        Wf[WA(c, d)] = AB[(size_t)c * LDB + d];
      }
#pragma unroll
      for (int k = 1; k <= 2; k++) {
        tick_cols(s0, n, k, lo, hi);
#pragma unroll
        for (int u = 0; u < PF; u++) {
          const int idx = tid + u * SB_T, c = lo + idx / LDB;
          pf[k & 1][u] = c < hi ? AB[(size_t)c * LDB + idx % LDB] : 0.0;
        }
      }
    }
    double vfull[TB];   // the sweep's last reflector, every lane holds all of it
#pragma unroll
    for (int l = 0; l < TB; l++) vfull[l] = 0.0;
    double tprev = 0.0;
    for (int tk = 0;; tk++) {
      if (tk >= 1) {
        int lo, hi;
        tick_cols(s0, n, tk, lo, hi);
        if (hi - retired_lo > SBW) {   // the ring cannot hold the span (not reached: span <= 2 P TSB - P + 1)
          if (tid == 0) st_rlx(abortf, 1);
          return;
        }
#pragma unroll
        for (int u = 0; u < PF; u++) {
          const int idx = tid + u * SB_T, c = lo + idx / LDB;
          if (c < hi) Wf[WA(c, idx % LDB)] = pf[tk & 1][u];
        }
        tick_cols(s0, n, tk + 2, lo, hi);   // its wait was done at the end of the previous tick
#pragma unroll
        for (int u = 0; u < PF; u++) {
          const int idx = tid + u * SB_T, c = lo + idx / LDB;
          pf[tk & 1][u] = c < hi ? AB[(size_t)c * LDB + idx % LDB] : 0.0;
        }
      }
      __syncthreads();
      // ---- this tick's task of each sweep: (s, j) with 2 (s - s0) + j = tk
      const int j = tk - 2 * grp;
      if (s < s1 && j >= 0 && j <= jlast(s, n)) {
        const int lo = s + 1 + j * TB, L = min(TB, n - lo);
        const bool tin = t < L;
        double x[TB];      // lane 0: the vector to eliminate
        double bcol[TB];   // lane t: column t of B (j >= 1)
        // column t of B = A[lo.., lo-TB+t]: contiguous from bb (d = TB - t + l <= 31)
        const int bb = WA(lo - TB + t, TB - t);
        if (j == 0) {
          const int xb = WA(s, 1);
#pragma unroll
          for (int l = 0; l < TB; l++) {
            const double v = Wf[xb + l];
            x[l] = (l < L) ? v : 0.0;
          }
        } else {
          // row t of B times the previous reflector, back to the ring
          double b[TB], d4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int u = 0; u < TB; u++) {
            b[u] = Wf[WA(lo - TB + u, TB + t - u)];
            d4[u & 3] += b[u] * vfull[u];
          }
          const double f = tprev * ((d4[0] + d4[1]) + (d4[2] + d4[3]));
#pragma unroll
          for (int u = 0; u < TB; u++) Wf[tin ? WA(lo - TB + u, TB + t - u) : DUM] = b[u] - f * vfull[u];
          wave_lds_fence();
#pragma unroll
          for (int l = 0; l < TB; l++) {
            const double v = Wf[bb + l];
            bcol[l] = (l < L) ? v : 0.0;
            x[l] = bcol[l];
          }
        }
        // reflector of lane 0's x, broadcast to the group
        double tau = 0.0, beta = x[0], vv[TB];
        {
          double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int l = 1; l < TB; l++) s4[l & 3] += x[l] * x[l];
          const double sig = (s4[0] + s4[1]) + (s4[2] + s4[3]);
          const double alpha = x[0];
          const double bnew = -copysign(sqrt(alpha * alpha + sig), alpha);
          const bool nz = sig > 0.0;
          beta = nz ? bnew : alpha;
          tau = nz ? (bnew - alpha) / bnew : 0.0;
          const double scal = nz ? 1.0 / (alpha - bnew) : 0.0;
          vv[0] = 1.0;
#pragma unroll
          for (int l = 1; l < TB; l++) vv[l] = x[l] * scal;
        }
        tau = bc16<0>(tau);
        beta = bc16<0>(beta);
#define VB16(l) vfull[l] = bc16<0>(vv[l]);
        UNROLL16(VB16)
#undef VB16
        double vown = 0.0;
#pragma unroll
        for (int l = 0; l < TB; l++) vown = (l == t) ? vfull[l] : vown;
        if (j == 0) {
          Wf[tin ? WA(s, 1 + t) : DUM] = (t == 0) ? beta : 0.0;
        } else if (L >= 2) {
          // column t of (I - tau v v^T) B; column 0 becomes (beta, 0, ...)
          double d4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int l = 0; l < TB; l++) d4[l & 3] += vfull[l] * bcol[l];
          const double f = tau * ((d4[0] + d4[1]) + (d4[2] + d4[3]));
#pragma unroll
          for (int l = 0; l < TB; l++) {
            const double nv = (t == 0) ? (l == 0 ? beta : 0.0) : bcol[l] - f * vfull[l];
            Wf[(l < L) ? bb + l : DUM] = nv;
          }
        }
        // D = A[lo..lo+L-1, same] <- H D H; lane t holds column t:
        // (l, t) at column lo+t, d = l-t for l >= t, at column lo+l, d = t-l above
        const int bt0 = WA(lo + t, 0) - t;
        double dr[TB];
#pragma unroll
        for (int l = 0; l < TB; l++) {
          const int ad = (l >= t) ? bt0 + l : WA(lo + l, t - l);
          const double v = Wf[ad];
          dr[l] = (l < L && tin) ? v : 0.0;
        }
        double p4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int l = 0; l < TB; l++) p4[l & 3] += dr[l] * vfull[l];
        const double p = tau * ((p4[0] + p4[1]) + (p4[2] + p4[3]));
        const double vp = sum16(vown * p);
        const double w = p - 0.5 * tau * vp * vown;
        double wfull[TB];
#define WB16(l) wfull[l] = bc16<l>(w);
        UNROLL16(WB16)
#undef WB16
#pragma unroll
        for (int l = 0; l < TB; l++)
          Wf[(l >= t && l < L && tin) ? bt0 + l : DUM] = dr[l] - (vfull[l] * w + wfull[l] * vown);
        VB[((size_t)j * n + s) * TB + t] = (t == 0) ? tau : vown;
        tprev = tau;
      }
      __syncthreads();
      // ---- write back the columns no later task of this batch touches
      int low = n, live = 0;
      for (int gg = 0; gg < SBP; gg++) {
        const int ss = s0 + gg;
        if (ss >= s1) break;
        const int jj = tk + 1 - 2 * gg;
        if (jj > jlast(ss, n)) continue;
        live = 1;
        const int mc = (jj <= 0) ? ss : ss + 1 + (jj - 1) * TB;
        low = min(low, mc);
      }
      int lo_, hi_;
      tick_cols(s0, n, tk, lo_, hi_);   // loaded so far: columns below hi_
      const int upto = live ? min(low, hi_) : hi_;   // finished: every loaded column (hi_ = n by then)
      for (int idx = tid; idx < (upto - retired_lo) * LDB; idx += SB_T) {
        const int c = retired_lo + idx / LDB, d = idx % LDB;
        AB[(size_t)c * LDB + d] = Wf[WA(c, d)];
      }
      retired_lo = max(retired_lo, upto);
      const bool pub = !live || ((tk & (E2_SB_PUB - 1)) == E2_SB_PUB - 1 && retired_lo > published);
      if (pub) __threadfence();   // each wavefront's write-backs complete and visible at agent scope
      if (tid == 0) {
        int lo3, hi3;
        tick_cols(s0, n, tk + 3, lo3, hi3);
        sh_abort = live ? wait_pred(hi3) : 0;
      }
      __syncthreads();
      if (pub && tid == 0) st_rel(&ret[q], live ? retired_lo : n);
      published = pub ? retired_lo : published;
      if (sh_abort) return;
      if (!live) break;
    }
  }
#undef WA
}

// d and e of the tridiagonal matrix from the chased band
__global__ void tri_kernel(const E2Block* __restrict__ blks, const int* __restrict__ list, double* __restrict__ WS) {
  const E2Block B = blks[list[blockIdx.y]];
  const int n = B.n;
  const double* AB = WS + B.q_off;
  double* d = WS + B.v_off;
  double* e = d + n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    d[i] = AB[(size_t)i * LDB];
    e[i] = (i + 1 < n) ? AB[(size_t)i * LDB + 1] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// back-transformation of the chase: group (g, j) = the reflectors of sweeps
// 16g..16g+15 at position j, rows r0 = 16g+1+16j .. r0+30 (a staircase V of
// 31 x 16, reflector c in rows c..c+15), composite H_{16g} ... H_{16g+15} =
// I - V T V^T; applied to 64 eigenvector rows Z (row form: Z <- Z - Z V T^T V^T)
// ---------------------------------------------------------------------------
#define Q2H (2 * TB - 1)
__global__ __launch_bounds__(256) void q2_kernel(const E2Block* __restrict__ blks, const E2Q2Grid G,
                                                 double* __restrict__ WS, const double* __restrict__ V2) {
  __shared__ double V[Q2H][TB + 1];
  __shared__ double Gm[TB][TB + 1], Tm[TB][TB + 1];
  __shared__ double tau_s[TB];
  __shared__ double Zs[64][Q2H + 2];
  __shared__ double P1[64][TB + 1];
  int bi = 0;
  while (bi + 1 < G.nb && (int)blockIdx.x >= G.cum[bi + 1]) bi++;
  const E2Block B = blks[G.blk[bi]];
  const int n = B.n, krow = B.krow, tid = threadIdx.x;
  const int g = G.glo[bi] + (blockIdx.x - G.cum[bi]);
  const int ng = (n - 2 + TB - 1) / TB;
  const int j = G.tick - (ng - 1 - g);
  const int s0 = g * TB, r0 = s0 + 1 + j * TB;
  if (j < 0 || r0 >= n) return;
  const int h = min(Q2H, n - r0);
  const double* VB = V2 + B.v2_off;
  for (int idx = tid; idx < Q2H * TB; idx += blockDim.x) V[idx / TB][idx % TB] = 0.0;
  __syncthreads();
  if (tid < TB) {
    const int c = tid, s = s0 + c;
    double tau = 0.0;
    if (s <= n - 3 && s + 1 + j * TB < n) {
      const int len = min(TB, n - (s + 1 + j * TB));
      const double* v = VB + ((size_t)j * n + s) * TB;
      tau = v[0];
      V[c][c] = 1.0;
      for (int e = 1; e < len; e++) V[c + e][c] = v[e];
    }
    tau_s[c] = tau;
  }
  __syncthreads();
  const int ti = tid / TB, tj = tid % TB;
  {
    double a = 0.0;
    for (int r = 0; r < h; r++) a += V[r][ti] * V[r][tj];
    Gm[ti][tj] = a;
  }
  __syncthreads();
  if (tid < TB) {
    double trow[TB];
    tmat_row(Gm, tau_s, tid, trow);
#pragma unroll
    for (int i = 0; i < TB; i++) Tm[tid][i] = trow[i];
  }
  __syncthreads();
  // every 64-row chunk of the kept rows: Z <- Z - (Z V) T^T V^T on columns r0 .. r0+h-1
  double* Z = WS + B.q_off;
  for (int e0 = 0; e0 < krow; e0 += 64) {
    const int ne = min(64, krow - e0);
    for (int idx = tid; idx < 64 * Q2H; idx += blockDim.x) {
      const int e = idx / Q2H, r = idx % Q2H;
      Zs[e][r] = (e < ne && r < h) ? Z[(size_t)(e0 + e) * n + r0 + r] : 0.0;
    }
    __syncthreads();
    const int e = tid >> 2, c0 = (tid & 3) * 4;
    {
      double a[4] = {0.0, 0.0, 0.0, 0.0};
      for (int r = 0; r < h; r++) {
        const double z = Zs[e][r];
#pragma unroll
        for (int u = 0; u < 4; u++) a[u] += z * V[r][c0 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) P1[e][c0 + u] = a[u];
    }
    __syncthreads();
    double p2[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      double a = 0.0;
      for (int t = 0; t < TB; t++) a += P1[e][t] * Tm[c0 + u][t];
      p2[u] = a;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; u++) P1[e][c0 + u] = p2[u];
    __syncthreads();
    for (int idx = tid; idx < 64 * Q2H; idx += blockDim.x) {
      const int ee = idx / Q2H, r = idx % Q2H;
      if (ee < ne && r < h) {
        double a = Zs[ee][r];
        for (int c = 0; c < TB; c++) a -= P1[ee][c] * V[r][c];
        Z[(size_t)(e0 + ee) * n + r0 + r] = a;
      }
    }
    __syncthreads();
  }
}

}  // namespace e2s

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
extern "C" {

hipError_t e2s_launch_init(const E2Block* b, const int* list, int nlist, int nmax, double* WS, int* FL,
                           hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  int gx = (nmax + 255) / 256;
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(e2s::init_kernel, dim3(gx, nlist), dim3(256), 0, s, b, list, WS, FL);
  return hipGetLastError();
}

hipError_t e2s_launch_leaf(const E2Block* b, const E2TsStep* st, const E2TsItem* leaves, int nleaf, double* A,
                           double* S, hipStream_t s) {
  if (nleaf <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2s::leaf_kernel, dim3(nleaf), dim3(QT), 0, s, b, st, leaves, A, S);
  return hipGetLastError();
}

hipError_t e2s_launch_root(const E2Block* b, const E2TsStep* st, int nstep, double* A, double* S, double* WS,
                           hipStream_t s) {
  if (nstep <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2s::root_kernel, dim3(nstep), dim3(QT), 0, s, b, st, A, S, WS);
  return hipGetLastError();
}

hipError_t e2s_launch_form(const E2Block* b, const E2TsStep* st, const E2TsItem* leaves, int nleaf, double* A,
                           double* S, hipStream_t s) {
  if (nleaf <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2s::form_kernel, dim3(nleaf), dim3(QT), 0, s, b, st, leaves, A, S);
  return hipGetLastError();
}

hipError_t e2s_launch_fin1(const E2Block* b, const E2TsStep* st, const E2TsItem* tiles, int ntile, double* S,
                           hipStream_t s) {
  if (ntile <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2s::fin1_kernel, dim3(ntile), dim3(256), 0, s, b, st, tiles, S);
  return hipGetLastError();
}

hipError_t e2s_launch_fin2(const E2Block* b, const E2TsStep* st, const E2TsItem* tiles, int ntile, double* S,
                           hipStream_t s) {
  if (ntile <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2s::fin2_kernel, dim3(ntile), dim3(256), 0, s, b, st, tiles, S);
  return hipGetLastError();
}

hipError_t e2s_launch_band(const E2Block* b, const int* list, int nlist, int nmax, const double* A, double* WS,
                           hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  int gx = (int)(((size_t)nmax * E2_SB_LDB + 255) / 256);
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(e2s::band_kernel, dim3(gx, nlist), dim3(256), 0, s, b, list, A, WS);
  return hipGetLastError();
}

// workgroups of the chase kernel that can be resident at once on this device
int e2s_chase_max_groups() {
  static int cached = -1;
  if (cached >= 0) return cached;
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)e2s::chase_kernel, SB_T, 0) != hipSuccess)
    return 0;
  cached = per * ncu;
  return cached;
}

hipError_t e2s_launch_chase(const E2Block* b, const E2SbGrid& G, double* WS, double* V2, int* FL, hipStream_t s) {
  if (G.nb <= 0) return hipSuccess;
  E2SbGrid g = G;
  const E2Block* bb = b;
  void* args[] = {(void*)&bb, (void*)&g, (void*)&WS, (void*)&V2, (void*)&FL};
  return hipLaunchCooperativeKernel((const void*)e2s::chase_kernel, dim3(G.nb * G.G), dim3(SB_T), args, 0, s);
}

hipError_t e2s_launch_tri(const E2Block* b, const int* list, int nlist, int nmax, double* WS, hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  int gx = (nmax + 255) / 256;
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(e2s::tri_kernel, dim3(gx, nlist), dim3(256), 0, s, b, list, WS);
  return hipGetLastError();
}

hipError_t e2s_launch_q2(const E2Block* b, const E2Q2Grid& G, double* WS, const double* V2, hipStream_t s) {
  if (G.nb <= 0 || G.cum[G.nb] <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2s::q2_kernel, dim3(G.cum[G.nb]), dim3(256), 0, s, b, G, WS, V2);
  return hipGetLastError();
}

}  // extern "C"
