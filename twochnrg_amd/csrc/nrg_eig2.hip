// nrg_eig2.hip -- batched FP64 symmetric eigensolver for uneven symmetry
// blocks, built from three stages that each run over all blocks at once:
//
//   1. Householder reduction to tridiagonal form.  Blocks with n <= 256: one
//      1024-thread workgroup per block, the matrix held in registers.  Larger
//      blocks: blocked reduction in panels of E2_PANEL columns -- per column
//      y = A v over the lower-triangle tiles of the symmetric trailing matrix
//      with the panel corrections, then a finish launch (w, the next column
//      and its partial norms); per panel one rank-2 E2_PANEL GEMM update of
//      the lower triangle; the last <= E2_RS_MAXN rows of each block go to the
//      register-resident kernel.
//   2. Divide and conquer on the tridiagonal matrix (Cuppen's tearing): leaves
//      of <= 16 rows by implicit QL; merges by z-vector deflation (small
//      components, and Givens rotation of nearly equal poles), secular-equation
//      roots by safeguarded Newton in the distance to the nearer pole,
//      Gu-Eisenstat recomputation of z for orthogonal vectors, and a GEMM for
//      the new vectors.
//   3. Back-transformation with blocked (compact WY) Householder panels:
//      three GEMMs per 64-reflector panel, restricted to the columns the
//      panel's reflectors touch.
//
// Storage: eigenvectors are kept vector-major (row i = eigenvector i), the
// layout of dEigVec in the reference (src/NRGmatrixClass.cpp:647-654), so the
// last stage writes the output rows directly.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrg_eig2.h"
#include "nrg_wave.h"

#define E2_T 256
#define WAVE 64
#define LEAF 16
#define PANEL E2_PANEL   // reduction panel
#define BTP 64     // back-transformation panel

namespace e2 {

__device__ __forceinline__ double wave_sum(double v) { return nrgw::wave_sum(v); }
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x / WAVE); i++) s += red[i];
  return s;
}

// ---------------------------------------------------------------------------
// stage 1a: whole reduction inside one workgroup (n <= 256)
// ---------------------------------------------------------------------------
#define SM_G 4              // row groups of the non-resident small-block reduction
#define SM_T (256 * SM_G)    // its workgroup: column c = tid % 256, row group g = tid / 256
// resident path: row r = w + R2_W * i belongs to wave w, column c = lane + 64 * j to the lane
#define R2_W (SM_T / WAVE)   // 16 waves = row groups
#define R2_NI 13             // row slots per thread (largest variant)
#define R2_NP 16             // row-slot stride of the by-row vectors (even: slot pairs are 16-byte aligned;
                             // slots NI .. 15 stay zero)
#define RS_MAXN (R2_W * R2_NI)   // largest block the resident path holds (208)
static_assert(RS_MAXN == E2_RS_MAXN, "the host's hand-off size");

// LDS of the small-block kernel: the resident path's arrays (the small, often
// read ones first, inside the 64 KiB reach of a DS instruction's offset) and
// the non-resident path's, which are never live together
struct alignas(16) R2Lds {
  double vt[R2_W][R2_NP];               // v by row: vt[w][i] = v[w + 16 i]
  double vwt[R2_W][R2_NP][2];           // (v, w) by row
  double vn[256];                       // v by column
  double vw[256][2];                    // (v, w) by column
  double pw[R2_W];                      // per-wave partials of v . (A v)
  double ycol[R2_W][4];                 // (A v) at the corner columns: four 16-lane partials per corner row
  double fold[R2_W][WAVE];              // corner rows' elements at the next reflector's column (its lane)
  double ts[2];                         // t of the current reflector
  double qp[R2_W / 2][256][2];          // per-wave partial sums of (A v)_c, waves 2h and 2h+1 side by side
};
struct NrLds {
  double red[SM_T / WAVE];
  double vsh[256], wsh[256], vnew[256], ash[256];
  double qpart[SM_G][256];
};
union alignas(16) SmallLds {
  R2Lds r2;
  NrLds nr;
};

__device__ __forceinline__ double2 ld2(const double* p) { return *reinterpret_cast<const double2*>(p); }
__device__ __forceinline__ void st2(double* p, double x, double y) {
  *reinterpret_cast<double2*>(p) = make_double2(x, y);
}
// the thread index as a value the compiler cannot carry across this point:
// per-lane LDS addresses are recomputed per phase (two integer operations)
// instead of being kept in registers over the column loop
__device__ __forceinline__ int r2_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}
__device__ __forceinline__ double r2_sum16(double v) {   // 16-lane row sum, the same bits in each lane
  v += nrgw::dpp_d<0xB1>(v);
  v += nrgw::dpp_d<0x4E>(v);
  v += nrgw::dpp_d<0x141>(v);
  v += nrgw::dpp_d<0x140>(v);
  return v;
}

// One variant of the resident reduction: NI row slots (n <= 16 NI) and NJ
// column slots of 64 columns, held in registers.  With CORNER the block also
// has the 16 columns 64 NJ .. 64 NJ + 15 (NI = 4 NJ + 1: rows 64 NJ .. are the
// last row slot): rows below 64 NJ keep no copy of them -- by symmetry they are
// the last row slot's rows, element (r, c) = (c, r) -- and the last slot's rows
// keep their 16 x 16 corner in one register (lane m: column 64 NJ + m).  The
// vectors v and w see these columns as column slot NJ, lanes 16 .. 63 past n;
// (A v)_c there is the row c of the last slot times v, one wave sum by the wave
// holding it, and the wave forming reflector kk reads row kk's entries there
// from the holders' elements (c, kk), published before the barrier.
template <int NI, int NJ, bool CORNER>
struct R2 {
  static constexpr int NJL = NJ + (CORNER ? 1 : 0);   // column slots of the vectors (column lane + 64 j)
  static constexpr int CF = WAVE * NJ;                 // first corner column
  static constexpr int IC = NI - 1;                    // the corner rows' slot
  static_assert(!CORNER || R2_W * IC == CF, "the corner rows are the columns past the column slots");
  typedef double Tile[NI][NJ];

  // q_j += a[i][j] v_(w + 16 i) over slots I .. NI-1 (I even: v of two slots per 16-byte read)
  template <int I, int I0, int J0>
  static __device__ __forceinline__ void symv(const Tile& a, const R2Lds& S, int w, double (&q)[NJL]) {
    if constexpr (I < NI) {
      const double2 vv = ld2(&S.vt[w][I]);
      if constexpr (I >= I0) {
#pragma unroll
        for (int j = J0; j < NJ; j++) q[j] = __fma_rn(a[I][j], vv.x, q[j]);
      }
      if constexpr (I + 1 < NI) {
#pragma unroll
        for (int j = J0; j < NJ; j++) q[j] = __fma_rn(a[I + 1][j], vv.y, q[j]);
      }
      symv<I + 2, I0, J0>(a, S, w, q);
    }
  }
  // a[i][j] -= v_r w_c + w_r v_c over slots I .. NI-1
  template <int I, int J0>
  static __device__ __forceinline__ void update(Tile& a, const R2Lds& S, int w, const double2 (&cw)[NJ]) {
    if constexpr (I < NI) {
      const double2 rw = ld2(S.vwt[w][I]);   // (v_r, w_r)
#pragma unroll
      for (int j = J0; j < NJ; j++) {
        a[I][j] = __fma_rn(-rw.x, cw[j].y, a[I][j]);
        a[I][j] = __fma_rn(-rw.y, cw[j].x, a[I][j]);
      }
      if constexpr ((I & 1) == 1) __builtin_amdgcn_sched_barrier(0);   // (v_r, w_r) of two rows in flight at most
      update<I + 1, J0>(a, S, w, cw);
    }
  }

  // reflector kk from row kk of the matrix (x: column slots of this wave's row,
  // all lanes), published by column (vn), by row (vt), in row kk of A, and as
  // d, e, tau, t
  static __device__ __forceinline__ void reflector(const int n, const int ld, const int kk, const double (&x)[NJL],
                                                   R2Lds& S, double* __restrict__ A, double* __restrict__ d,
                                                   double* __restrict__ e, double* __restrict__ tau, int l) {
    double ss = 0.0, xk1 = 0.0, xkk = 0.0;
#pragma unroll
    for (int j = 0; j < NJL; j++) {
      const int c = l + WAVE * j;
      if (c > kk) ss = __fma_rn(x[j], x[j], ss);
      if (j == (kk + 1) / WAVE) xk1 = x[j];
      if (j == kk / WAVE) xkk = x[j];
    }
    const double nrm2 = nrgw::wave_sum_bc(ss);
    const double x0 = nrgw::readlane_d(xk1, (kk + 1) & (WAVE - 1));   // element (kk, kk+1)
    const double akk = nrgw::readlane_d(xkk, kk & (WAVE - 1));        // element (kk, kk)
    double t = 0.0, v0 = x0, alpha = 0.0;
    if (nrm2 != 0.0) {
      const double nrm = sqrt(nrm2);
      alpha = (x0 > 0.0) ? -nrm : nrm;
      v0 = x0 - alpha;
      // t = 2 / |v|^2 = 1 / (nrm (nrm + |x0|)): hardware reciprocal and one
      // Newton step (within an ulp or so, the rounding of any other t)
      const double den = nrm * (nrm + fabs(x0));
      const double r0 = __builtin_amdgcn_rcp(den);
      t = __fma_rn(r0, __fma_rn(-den, r0, 1.0), r0);
    }
#pragma unroll
    for (int j = 0; j < NJL; j++) {
      const int c = l + WAVE * j;
      const double vc = (c > kk + 1) ? x[j] : (c == kk + 1 ? v0 : 0.0);
      S.vn[c] = vc;
      if (c < R2_W * NI) S.vt[c & (R2_W - 1)][c / R2_W] = vc;
      if (c > kk && c < n) A[(size_t)kk * ld + c] = vc;   // reflector kk in row kk
    }
    if (l == 0) {
      d[kk] = akk;
      e[kk] = alpha;
      tau[kk] = t;
      S.ts[0] = t;
    }
  }

  // row kk = k+1 (slot I of this wave) after the update of reflector k; then
  // reflector kk from it, or, after the last column, d and e of the trailing
  // 2 x 2 block
  template <int I, int J0>
  static __device__ __forceinline__ void next_row(const int n, const int ld, const int k, Tile& a, double& cr, R2Lds& S,
                                                  double* __restrict__ A, double* __restrict__ d,
                                                  double* __restrict__ e, double* __restrict__ tau, int l, double vr,
                                                  double wr, const double (&vc)[NJL], const double (&wc)[NJL]) {
    const int kk = k + 1;
    double x[NJL];
#pragma unroll
    for (int j = J0; j < NJ; j++) {
      a[I][j] = __fma_rn(-vr, wc[j], a[I][j]);
      a[I][j] = __fma_rn(-wr, vc[j], a[I][j]);
    }
#pragma unroll
    for (int j = 0; j < NJ; j++) x[j] = a[I][j];
    if constexpr (CORNER) {
      if constexpr (I == IC) {   // the row's own corner
        cr = __fma_rn(-vr, wc[NJ], cr);
        cr = __fma_rn(-wr, vc[NJ], cr);
        x[NJ] = cr;
      } else {   // elements (CF + m, kk) of the corner rows, updated as their holders will
        double y = l < 16 ? S.fold[l][kk & (WAVE - 1)] : 0.0;
        y = __fma_rn(-vc[NJ], wr, y);
        y = __fma_rn(-wc[NJ], vr, y);
        x[NJ] = y;
      }
    }
    if (kk + 2 < n) {
      reflector(n, ld, kk, x, S, A, d, e, tau, l);
    } else {
#pragma unroll
      for (int j = 0; j < NJL; j++) {
        const int c = l + WAVE * j;
        if (c == n - 2) d[n - 2] = x[j];
        if (c == n - 1) e[n - 2] = x[j];
      }
    }
  }

  // columns k of stage I0 (k in [16 I0, 16 I0 + 16), k + 2 < n): rows k and
  // k + 1 are in slots I0 and I0 or I0 + 1, every row of slots below I0 is
  // finished (v = w = 0 there, so slot I0's finished rows and any column <= k
  // are left unchanged, exactly), and column slots below J0 hold only finished
  // columns.  Reflector k is published when the stage's column k begins.
  template <int I0>
  static __device__ __forceinline__ void stage(const int n, const int ld, Tile& a, double& cr, R2Lds& S,
                                               double* __restrict__ A,
                                               double* __restrict__ d, double* __restrict__ e,
                                               double* __restrict__ tau, int w) {
    constexpr int J0 = ((R2_W * I0 + 1) / WAVE < NJL) ? (R2_W * I0 + 1) / WAVE : NJL - 1;
    const int k1 = min(R2_W * (I0 + 1), n - 2);
#pragma unroll 1
    for (int k = R2_W * I0; k < k1; k++) {
      // 1. this wave's partial (A v)_c over its rows, and its partial of v . (A v)
      {
        const int l = r2_tid() & (WAVE - 1);
        double q[NJL];
#pragma unroll
        for (int j = 0; j < NJL; j++) q[j] = 0.0;
        symv<I0 & ~1, I0, J0>(a, S, w, q);
        double ss = 0.0;
#pragma unroll
        for (int j = J0; j < NJ; j++) {
          const int c = l + WAVE * j;
          S.qp[w >> 1][c][w & 1] = q[j];
          ss = __fma_rn(S.vn[c], q[j], ss);   // v_c = 0 for c <= k and c >= n
        }
        if constexpr (CORNER) {
          // (A v) at column CF + w is row CF + w (this wave's corner row) times
          // v: four 16-lane partials for the wave forming w, and its term
          // v_(CF+w) (A v)_(CF+w) of v.(A v) lane by lane
          double yr = 0.0;
#pragma unroll
          for (int j = J0; j < NJ; j++) yr = __fma_rn(a[IC][j], S.vn[l + WAVE * j], yr);
          yr = __fma_rn(cr, S.vn[CF + l], yr);   // lanes 16 .. 63: cr = 0
          ss = __fma_rn(S.vn[CF + w], yr, ss);
          yr = r2_sum16(yr);
          if ((l & 15) == 0) S.ycol[w][l >> 4] = yr;
          // element (CF + w, k + 1) for the wave forming reflector k + 1
          const int kk = k + 1;
          if (kk < CF && l == (kk & (WAVE - 1))) {
            double x = 0.0;
#pragma unroll
            for (int j = 0; j < NJ; j++)
              if (j == kk / WAVE) x = a[IC][j];
            S.fold[w][l] = x;
          }
        }
        const double pv = nrgw::wave_sum_bc(ss);
        if (l == 0) S.pw[w] = pv;
      }
      __syncthreads();
      // 2. the wave holding row k+1: (A v)_c from the 16 partials, w = t A v -
      // (t^2/2)(v.A v) v published by column and by row (zero for row k+1,
      // which it updates itself), then row k+1 and reflector k+1
      if (w == ((k + 1) & (R2_W - 1))) {
        const int l = r2_tid() & (WAVE - 1);
        const double vav = r2_sum16(S.pw[l & (R2_W - 1)]);
        const double t = S.ts[0];
        const double K = 0.5 * t * (t * vav);
        double vc[NJL], wc[NJL];
#pragma unroll
        for (int j = 0; j < NJL; j++) {
          vc[j] = 0.0;
          wc[j] = 0.0;
        }
#pragma unroll
        for (int j = J0; j < NJL; j++) {
          const int c = l + WAVE * j;
          double y;
          if (j < NJ) {
            double p[R2_W / 2];
#pragma unroll
            for (int h = 0; h < R2_W / 2; h++) {   // waves 2h + 2h+1, then pairs of those, ...
              const double2 pr = ld2(S.qp[h][c]);
              p[h] = pr.x + pr.y;
            }
#pragma unroll
            for (int m = 1; m < R2_W / 2; m *= 2)
#pragma unroll
              for (int h = 0; h < R2_W / 2; h += 2 * m) p[h] += p[h + m];
            y = p[0];
          } else {
            const double* yc = S.ycol[l & (R2_W - 1)];   // lanes 16 .. 63 are past n
            y = (yc[0] + yc[1]) + (yc[2] + yc[3]);
          }
          vc[j] = S.vn[c];
          wc[j] = (c > k && c < n) ? t * y - K * vc[j] : 0.0;
          st2(S.vw[c], vc[j], wc[j]);
          if (c < R2_W * NI) {
            const bool mine = c == k + 1;
            st2(S.vwt[c & (R2_W - 1)][c / R2_W], mine ? 0.0 : vc[j], mine ? 0.0 : wc[j]);
          }
          __builtin_amdgcn_sched_barrier(0);   // one column slot's partials in flight at a time
        }
        double vr = 0.0, wr = 0.0;   // v and w at k+1
#pragma unroll
        for (int j = J0; j < NJL; j++)
          if (j == (k + 1) / WAVE) {
            vr = vc[j];
            wr = wc[j];
          }
        vr = nrgw::readlane_d(vr, (k + 1) & (WAVE - 1));
        wr = nrgw::readlane_d(wr, (k + 1) & (WAVE - 1));
        if ((k + 1) / R2_W == I0) {
          next_row<I0, J0>(n, ld, k, a, cr, S, A, d, e, tau, l, vr, wr, vc, wc);
        } else {
          if constexpr (I0 + 1 < NI) next_row<I0 + 1, J0>(n, ld, k, a, cr, S, A, d, e, tau, l, vr, wr, vc, wc);
        }
      }
      __syncthreads();
      // 3. A -= v w^T + w v^T on this thread's slots I0 .. NI-1
      {
        const int l = r2_tid() & (WAVE - 1);
        double2 cw[NJ];
#pragma unroll
        for (int j = J0; j < NJ; j++) cw[j] = ld2(S.vw[l + WAVE * j]);
        update<I0, J0>(a, S, w, cw);
        if constexpr (CORNER) {   // lanes 16 .. 63 read (v, w) = 0 and keep cr = 0
          const double2 rw = ld2(S.vwt[w][IC]);
          const double2 c2 = ld2(S.vw[CF + l]);
          cr = __fma_rn(-rw.x, c2.y, cr);
          cr = __fma_rn(-rw.y, c2.x, cr);
        }
      }
    }
    if (k1 == n - 2) {   // the last stage: row n-1 (slot I0 or I0 + 1) is final
      const int l = r2_tid() & (WAVE - 1);
      trailing<I0>(n, a, cr, d, w, l);
      if constexpr (I0 + 1 < NI) trailing<I0 + 1>(n, a, cr, d, w, l);
    }
  }
  // d[n-1] from row n-1's owner, if slot I holds it
  template <int I>
  static __device__ __forceinline__ void trailing(const int n, const Tile& a, double cr, double* __restrict__ d, int w,
                                                  int l) {
    if (w == ((n - 1) & (R2_W - 1)) && (n - 1) / R2_W == I) {
#pragma unroll
      for (int j = 0; j < NJ; j++)
        if (l + WAVE * j == n - 1) d[n - 1] = a[I][j];
      if constexpr (CORNER && I == IC)
        if (CF + l == n - 1) d[n - 1] = cr;
    }
  }
  template <int I0>
  static __device__ __forceinline__ void stages(const int n, const int ld, Tile& a, double& cr, R2Lds& S,
                                                double* __restrict__ A,
                                                double* __restrict__ d, double* __restrict__ e,
                                                double* __restrict__ tau, int w) {
    if constexpr (I0 < NI) {
      if (R2_W * I0 + 2 < n) {
        stage<I0>(n, ld, a, cr, S, A, d, e, tau, w);
        stages<I0 + 1>(n, ld, a, cr, S, A, d, e, tau, w);
      }
    }
  }

  // the n x n matrix at A with row stride ld: in full (lower = false) or its
  // lower triangle only, the upper triangle read from it by symmetry
  static __device__ __forceinline__ double load(const double* __restrict__ A, int ld, bool lower, int r, int c) {
    return (lower && r < c) ? A[(size_t)c * ld + r] : A[(size_t)r * ld + c];
  }
  static __device__ __forceinline__ void run(const int n, double* __restrict__ A, const int ld, const bool lower,
                                             double* __restrict__ d, double* __restrict__ e, double* __restrict__ tau,
                                             R2Lds& S) {
    const int tid = threadIdx.x, l = tid & (WAVE - 1);
    const int w = __builtin_amdgcn_readfirstlane(tid / WAVE);
    if (n <= 2) {   // already tridiagonal
      if (tid == 0) {
        d[0] = A[0];
        e[n - 1] = 0.0;
        if (n == 2) {
          e[0] = load(A, ld, lower, 0, 1);
          d[1] = A[ld + 1];
        }
        for (int k = 0; k < n; k++) tau[k] = 0.0;
      }
      return;
    }
    Tile a;
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int r = w + R2_W * i;
#pragma unroll
      for (int j = 0; j < NJ; j++) {
        const int c = l + WAVE * j;
        a[i][j] = (r < n && c < n) ? load(A, ld, lower, r, c) : 0.0;
      }
      __builtin_amdgcn_sched_barrier(0);   // one row of loads in flight at a time: the registers hold the matrix
    }
    double cr = 0.0;   // corner: element (CF + w, CF + l), lanes 0 .. 15
    if (CORNER && CF + w < n && l < 16 && CF + l < n) cr = load(A, ld, lower, CF + w, CF + l);
    // v and w by row: the padding slots and the rows >= 16 NI are read (as zeros) but never written
    for (int q = tid; q < R2_W * R2_NP; q += SM_T) {
      (&S.vt[0][0])[q] = 0.0;
      st2(&S.vwt[0][0][0] + 2 * q, 0.0, 0.0);
    }
    if (CORNER && l == 0) S.fold[w][0] = a[IC][0];   // element (CF + w, 0) for reflector 0
    __syncthreads();
    if (w == 0) {   // reflector 0 from row 0 (slot 0 of wave 0)
      double x[NJL];
#pragma unroll
      for (int j = 0; j < NJ; j++) x[j] = a[0][j];
      if constexpr (CORNER) x[NJ] = l < 16 ? S.fold[l][0] : 0.0;
      reflector(n, ld, 0, x, S, A, d, e, tau, l);
    }
    __syncthreads();
    stages<0>(n, ld, a, cr, S, A, d, e, tau, w);   // the last one also writes the trailing 2 x 2 block
    if (tid == 0) {
      e[n - 1] = 0.0;
      tau[n - 2] = 0.0;
      tau[n - 1] = 0.0;
    }
  }
};

// Resident Householder reduction of one block with n <= RS_MAXN (one 1024-thread
// workgroup).  The whole matrix stays in registers: wave w holds rows
// w + 16 i, each lane columns lane + 64 j (past 192 rows, the last 16 columns
// by symmetry, with a 16 x 16 corner), so that every value of v or w read from
// LDS serves all of a thread's columns.  Reflector k is
// published before column k begins; per column k, two barriers:
//   1. every wave forms its rows' partial (A v)_c and its partial of v.(A v);
//   2. the wave holding row k+1 adds the 16 partials of each column, forms
//      w = t A v - (t^2/2)(v.A v) v and publishes (v, w) by column and by row,
//      updates row k+1 and forms and publishes reflector k+1 (norm by a wave
//      reduction);
// then each thread applies A -= v w^T + w v^T to its other elements (two
// fused multiply-adds each).  The column loop runs in stages of 16 columns,
// each compiled for the row and column slots that can still change, so no
// per-slot condition is evaluated at run time; six variants sized to n (up
// to 64, 96, 128, 160, 192, 208 rows) keep the slots past n out of the work.
// The reflector is written to row k of A (the back-transformation reads it
// there), with d, e and tau.
__device__ __forceinline__ void tridiag_resident(const int n, double* __restrict__ A, const int ld, const bool lower,
                                                 double* __restrict__ d, double* __restrict__ e,
                                                 double* __restrict__ tau, R2Lds& S) {
  if (n <= 64) R2<4, 1, false>::run(n, A, ld, lower, d, e, tau, S);
  else if (n <= 96) R2<6, 2, false>::run(n, A, ld, lower, d, e, tau, S);
  else if (n <= 128) R2<8, 2, false>::run(n, A, ld, lower, d, e, tau, S);
  else if (n <= 160) R2<10, 3, false>::run(n, A, ld, lower, d, e, tau, S);
  else if (n <= 192) R2<12, 3, false>::run(n, A, ld, lower, d, e, tau, S);
  else R2<R2_NI, 3, true>::run(n, A, ld, lower, d, e, tau, S);
}

__global__ __launch_bounds__(SM_T) void tridiag_small_kernel(const E2Block* __restrict__ blks,
                                                              const int* __restrict__ list, double* __restrict__ A0,
                                                              double* __restrict__ WS) {
  // Householder reduction, one workgroup per block.  Column c is owned by the
  // SM_G threads (c, g); for each column k the rows below k are split into SM_G
  // contiguous ranges, one per row group, so that SM_G times more loads are in
  // flight per block, and the row groups' partial sums of (A v)_c are added in
  // group order.  The rank-2 update of column k-1 (A -= v w^T + w v^T) is
  // applied inside the pass of column k that also forms A v: every trailing
  // element is read once, updated, written and used in the product.  The
  // update of element (j, i) uses the same operations as that of (i, j).
  __shared__ SmallLds sm;
  double* red = sm.nr.red;
  double* vsh = sm.nr.vsh;
  double* wsh = sm.nr.wsh;
  double* vnew = sm.nr.vnew;
  double* ash = sm.nr.ash;
  double(*qpart)[256] = sm.nr.qpart;
  const E2Block B = blks[list[blockIdx.x]];
  const int n = B.n, tid = threadIdx.x;
  const int c = tid & 255, g = tid >> 8;
  double* A = A0 + B.a_off;
  double* d = WS + B.v_off;
  double* e = d + n;
  double* tau = e + n;
  if (n <= RS_MAXN) {
    tridiag_resident(n, A, n, false, d, e, tau, sm.r2);
    return;
  }
  bool pending = false;   // vsh/wsh hold the update of the previous column
  double vi_old = 0.0, wi_old = 0.0;
  if (g == 0) { vsh[c] = 0.0; wsh[c] = 0.0; }
  __syncthreads();
  for (int k = 0; k + 2 < n; k++) {
    // row k of the current matrix (row group 0)
    double ak = 0.0;
    if (g == 0 && c >= k && c < n) {
      ak = A[(size_t)k * n + c];
      if (pending) {
        const double vj = vsh[k], qj = wsh[k];
        ak -= (k <= c) ? __fma_rn(vj, wi_old, vi_old * qj) : __fma_rn(vi_old, qj, vj * wi_old);
      }
    }
    if (g == 0) ash[c] = ak;
    const double x = (g == 0 && c > k && c < n) ? ak : 0.0;
    const double nrm2 = block_sum(x * x, red);   // its barriers order the ash stores
    const double x0 = ash[k + 1];
    double t = 0.0, v0 = x0, alpha = 0.0;
    if (nrm2 != 0.0) {
      const double nrm = sqrt(nrm2);
      alpha = (x0 > 0.0) ? -nrm : nrm;
      v0 = x0 - alpha;
      t = 2.0 / (nrm2 - x0 * x0 + v0 * v0);
    }
    const double ac = ash[c];
    const double vi = (c > k + 1 && c < n) ? ac : (c == k + 1 ? v0 : 0.0);
    if (g == 0) vnew[c] = vi;
    __syncthreads();
    // this group's rows > k: apply the pending update, write back, accumulate A v
    double qi = 0.0;
    if (c > k && c < n) {
      const int L = (n - k - 1 + SM_G - 1) / SM_G;
      int j = k + 1 + g * L;
      const int j1 = min(n, j + L);
      if (pending) {
        for (; j + 8 <= j1; j += 8) {
          double a[8];
#pragma unroll
          for (int u = 0; u < 8; u++) a[u] = A[(size_t)(j + u) * n + c];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int jj = j + u;
            const double vj = vsh[jj], qj = wsh[jj];
            a[u] -= (jj <= c) ? __fma_rn(vj, wi_old, vi_old * qj) : __fma_rn(vi_old, qj, vj * wi_old);
            qi += a[u] * vnew[jj];
          }
#pragma unroll
          for (int u = 0; u < 8; u++) A[(size_t)(j + u) * n + c] = a[u];
        }
        for (; j < j1; j++) {
          const double vj = vsh[j], qj = wsh[j];
          double a = A[(size_t)j * n + c];
          a -= (j <= c) ? __fma_rn(vj, wi_old, vi_old * qj) : __fma_rn(vi_old, qj, vj * wi_old);
          A[(size_t)j * n + c] = a;
          qi += a * vnew[j];
        }
      } else {
        for (; j + 8 <= j1; j += 8) {
          double a[8];
#pragma unroll
          for (int u = 0; u < 8; u++) a[u] = A[(size_t)(j + u) * n + c];
#pragma unroll
          for (int u = 0; u < 8; u++) qi += a[u] * vnew[j + u];
        }
        for (; j < j1; j++) qi += A[(size_t)j * n + c] * vnew[j];
      }
    }
    qpart[g][c] = qi;
    __syncthreads();
    qi = 0.0;
#pragma unroll
    for (int h = 0; h < SM_G; h++) qi += qpart[h][c];
    qi *= t;
    const double vp = block_sum((g == 0 && c > k && c < n) ? qi * vi : 0.0, red);   // barriers: qpart read
    const double K = 0.5 * t * vp;
    const double wi = (c > k && c < n) ? qi - K * vi : 0.0;
    if (g == 0) { vsh[c] = vi; wsh[c] = wi; }
    vi_old = vi;
    wi_old = wi;
    pending = true;
    if (g == 0 && c == k) { d[k] = ak; e[k] = alpha; tau[k] = t; }
    if (g == 0 && c > k && c < n) A[(size_t)k * n + c] = vi;   // reflector k in row k
    __syncthreads();
  }
  // the last update reaches the trailing 2 x 2 block
  if (g == 0 && n >= 2 && pending && c >= n - 2 && c < n) {
    for (int j = n - 2; j < n; j++) {
      const double vj = vsh[j], qj = wsh[j];
      A[(size_t)j * n + c] -= (j <= c) ? __fma_rn(vj, wi_old, vi_old * qj) : __fma_rn(vi_old, qj, vj * wi_old);
    }
  }
  __syncthreads();
  if (tid == 0) {
    if (n >= 2) {
      d[n - 2] = A[(size_t)(n - 2) * n + (n - 2)];
      e[n - 2] = A[(size_t)(n - 2) * n + (n - 1)];
    }
    d[n - 1] = A[(size_t)(n - 1) * n + (n - 1)];
    e[n - 1] = 0.0;
    for (int k = (n >= 2 ? n - 2 : 0); k < n; k++) tau[k] = 0.0;
  }
}

// The trailing matrix of a large block whose column chain stopped at panel
// boundary k0 (n - k0 <= RS_MAXN): its lower triangle is current there (the
// panel products keep it so) and the resident reduction continues from it,
// the upper triangle read by symmetry, writing reflectors k0 .. into their rows
// and d, e, tau from k0 on.
__global__ __launch_bounds__(SM_T) void tridiag_hand_kernel(const E2Block* __restrict__ blks,
                                                             const E2Hand* __restrict__ hands, double* __restrict__ A0,
                                                             double* __restrict__ WS) {
  __shared__ R2Lds S;
  const E2Hand H = hands[blockIdx.x];
  const E2Block B = blks[H.blk];
  const int n = B.n, k0 = H.k0;
  double* A = A0 + B.a_off + (size_t)k0 * n + k0;
  double* d = WS + B.v_off;
  tridiag_resident(n - k0, A, n, true, d + k0, d + n + k0, d + 2 * (size_t)n + k0, S);
}

// ---------------------------------------------------------------------------
// stage 1b: blocked reduction of the large blocks (n > 256), panels of PANEL
// columns.  Inside a panel the trailing matrix is left as it was at the panel
// start; column k is brought up to date from the panel's reflectors V and the
// vectors W (A_k -= V W_k^T + W V_k^T), and the product y = A v is corrected
// the same way (y -= V (W^T v) + W (V^T v)).  At the panel end one GEMM applies
// A -= [V W] [W V]^T to the trailing matrix.  Two launches per column:
//   symv(k): the reflector of column k (from partial norms), y = A v over the
//            lower-triangle tiles of the (symmetric) trailing matrix, each
//            off-diagonal tile read once for its column and its row products,
//            v^T A v per tile, and the panel dot products V^T v, W^T v over
//            column chunks;
//   fin(k):  w = tau y + alpha v per column chunk, the reflector stored in
//            row k, and (inside a panel) column k+1 brought up to date with
//            its partial norms.
//   norm(k): partial norms of column k at a panel start (after the GEMM).
// Panel buffer X (3 PANEL rows of n): rows [0,P) V^T, [P,2P) W^T, [2P,3P) V^T,
// so that rows [0,2P) and [P,3P) are the two operands of the panel GEMM.
// ---------------------------------------------------------------------------
#define SYMV_R 128   // rows and columns of a symv tile
#define CHUNK 256    // threads of a symv tile / elements per chunk
#ifndef E2_FIN_E
#define E2_FIN_E 128
#endif
#define FIN_E E2_FIN_E   // elements per finish workgroup, and per partial norm of a column
static_assert(CHUNK % FIN_E == 0 && FIN_E % WAVE == 0, "partial-norm chunks");

struct TdbPtr {
  double *A, *X, *d, *e, *tau, *refl, *npart, *part, *tpart, *cpart;
};
__device__ __forceinline__ TdbPtr tdb_ptrs(const E2Block& B, double* A0, double* WS) {
  const size_t n = B.n;
  TdbPtr p;
  p.A = A0 + B.a_off;
  p.X = WS + B.p_off;
  p.d = WS + B.v_off;
  p.e = p.d + n;
  p.tau = p.e + n;
  p.refl = WS + B.v_off + 8 * n;    // t, v0, alpha
  p.npart = WS + B.v_off + 9 * n;   // partial norms, one per chunk
  p.part = WS + B.k_off;            // row tiles x n partial y
  p.tpart = WS + B.u_off;           // one v^T A v partial per tile
  p.cpart = WS + B.dl_off;          // column chunk x 2P panel dot products
  return p;
}

// the reflector of column k from its partial norms; called by whole waves
// (every lane gets the result): lane l adds partials l, l + 64, ... and the
// wave sums the lanes, so the loads are issued together
__device__ __forceinline__ void tdb_reflector(const TdbPtr& P, int k, int n, double& t, double& v0, double& alpha) {
  const int nch = (n - k + FIN_E - 1) / FIN_E;
  double part = 0.0;
  for (int c = (int)(threadIdx.x & (WAVE - 1)); c < nch; c += WAVE) part += P.npart[c];
  const double nrm2 = wave_sum(part);
  const double x0 = P.A[(size_t)k * n + k + 1];
  if (nrm2 == 0.0) { t = 0.0; v0 = 0.0; alpha = 0.0; return; }
  const double nrm = sqrt(nrm2);
  alpha = (x0 > 0.0) ? -nrm : nrm;
  v0 = x0 - alpha;
  t = 2.0 / (nrm2 - x0 * x0 + v0 * v0);
}

// partial norms of column k (entries i > k) at a panel start, one per FIN_E
// entries (the finish kernel's granularity); for the last, partial panel
// also clear the unused panel rows in the two columns the final GEMM reads
__global__ __launch_bounds__(CHUNK) void tdb_norm_kernel(const E2Block* __restrict__ blks,
                                                         const int* __restrict__ list, int k, double* __restrict__ A0,
                                                         double* __restrict__ WS) {
  __shared__ double red[CHUNK / WAVE];
  const E2Block B = blks[list[blockIdx.y]];
  const int n = B.n;
  if (k + 2 >= n || (k % PANEL) != 0) return;
  const int c = blockIdx.x;
  if (k + c * CHUNK >= n) return;
  const TdbPtr P = tdb_ptrs(B, A0, WS);
  const int i = k + c * CHUNK + threadIdx.x;
  const double a = (i > k && i < n) ? P.A[(size_t)k * n + i] : 0.0;
  const double ws = wave_sum(a * a);
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = ws;
  __syncthreads();
  constexpr int SUB = CHUNK / FIN_E, WPS = FIN_E / WAVE;
  if (threadIdx.x < SUB && k + c * CHUNK + (int)threadIdx.x * FIN_E < n) {
    double s = 0.0;
    for (int w = 0; w < WPS; w++) s += red[threadIdx.x * WPS + w];
    P.npart[c * SUB + threadIdx.x] = s;
  }
  const int nbp = min(PANEL, n - 2 - k);
  if (c == 0 && nbp < PANEL && threadIdx.x < 3 * PANEL) {
    const int p = threadIdx.x % PANEL, sect = threadIdx.x / PANEL;
    if (p >= nbp) {
      P.X[(size_t)(sect * PANEL + p) * n + n - 2] = 0.0;
      P.X[(size_t)(sect * PANEL + p) * n + n - 1] = 0.0;
    }
  }
}

// the entry of G that workgroup x belongs to: the last i with cum[i] <= x
__device__ __forceinline__ int tdb_find(const E2TdbGrid& G, int x) {
  int lo = 0, hi = G.nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (G.cum[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// lower-triangle tile x of MT(MT+1)/2 in row order: row tile I >= column tile J
__device__ __forceinline__ void tri_tile(int x, int& I, int& J) {
  int i = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= x) i++;
  while (i * (i + 1) / 2 > x) i--;
  I = i;
  J = x - i * (i + 1) / 2;
}

// y = A v over the trailing matrix A[k+1:, k+1:], which is symmetric and still
// as it was at the panel start.  Tiles of SYMV_R x SYMV_R with row tile
// I >= column tile J; each is read once.  A diagonal tile gives the column
// products sum_r A[r][c] v_r of its elements (slot I of their partials); an
// off-diagonal tile also stands for its mirror tile (J, I) and gives the row
// products sum_c A[r][c] v_c of its row elements (slot J).  Element i of row
// tile K so receives exactly one partial per slot 0..MT-1: K from the diagonal
// tile, I > K from tiles below it, J < K from tiles left of it.
// Threads: wave w, lane (cl = lane % 16, rl = lane / 16) reads rows
// w*32 + rl + 4s (s < 8) at columns cl + 16g (g < 8): every load instruction
// covers four rows x 128 contiguous bytes.
__global__ __launch_bounds__(CHUNK) void tdb_symv_kernel(const E2Block* __restrict__ blks, const E2TdbGrid G, int k,
                                                         double* __restrict__ A0, double* __restrict__ WS) {
  constexpr int T = SYMV_R, LP = T + 4;   // LP: padded LDS row of the partials
  __shared__ double red[CHUNK / WAVE];
  __shared__ double vs[CHUNK];
  __shared__ double colp[16][T], rowp[16][LP];
  const int gi = tdb_find(G, blockIdx.x);
  const E2Block B = blks[G.blk[gi]];
  const int n = B.n;
  if (k + 2 >= n) return;
  const int m = n - k - 1;
  const int CT = (m + CHUNK - 1) / CHUNK, MT = (m + T - 1) / T, NT = MT * (MT + 1) / 2;
  const int x = blockIdx.x - G.cum[gi];
  if (x >= NT + CT) return;
  const TdbPtr P = tdb_ptrs(B, A0, WS);
  const double* rowk = P.A + (size_t)k * n;
  const int tid = threadIdx.x;
  // workgroups [0, CT): the panel dot products, one column chunk each (they
  // are dispatched first); [CT, CT + NT): the lower-triangle tiles
  if (x >= CT) {
    const int xt = x - CT;
    int I, J;
    tri_tile(xt, I, J);
    const int r0 = k + 1 + I * T, c0 = k + 1 + J * T;
    const int lane = tid & (WAVE - 1), w = tid / WAVE;
    const int cl = lane & 15, rl = lane >> 4;
    const int rb = w * 32 + rl;
    // the tile's loads first: they depend on nothing computed in this kernel
    double a[8][8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
      const int r = r0 + rb + 4 * s;
      const double* row = P.A + (size_t)r * n + c0 + cl;
#pragma unroll
      for (int g = 0; g < 8; g++) a[s][g] = (r < n && c0 + cl + 16 * g < n) ? row[16 * g] : 0.0;
    }
    // vs[0,T): v over the tile's rows, vs[T,2T): v over its columns; element
    // k+1 (in column tile 0 only) is v0, which needs the reflector
    const int q = (tid < T) ? r0 + tid : c0 + tid - T;
    double vq = (q < n) ? rowk[q] : 0.0;
    if (J == 0) {
      double t, v0, alpha;
      tdb_reflector(P, k, n, t, v0, alpha);
      if (q == k + 1) vq = v0;
      if (xt == 0 && tid == 0) {
        P.d[k] = rowk[k];
        P.e[k] = alpha;
        P.tau[k] = t;
        P.refl[0] = t; P.refl[1] = v0; P.refl[2] = alpha;
      }
    }
    vs[tid] = vq;
    __syncthreads();
    double cs[8], rs[8];
#pragma unroll
    for (int g = 0; g < 8; g++) cs[g] = 0.0;
#pragma unroll
    for (int s = 0; s < 8; s++) {
      const double vr = vs[rb + 4 * s];
      double acc = 0.0;
#pragma unroll
      for (int g = 0; g < 8; g++) {
        cs[g] += a[s][g] * vr;
        acc += a[s][g] * vs[T + cl + 16 * g];
      }
      rs[s] = acc;
    }
#pragma unroll
    for (int g = 0; g < 8; g++) colp[w * 4 + rl][cl + 16 * g] = cs[g];
#pragma unroll
    for (int s = 0; s < 8; s++) rowp[cl][rb + 4 * s] = rs[s];
    __syncthreads();
    double tv = 0.0;
    if (tid < T) {
      double y = 0.0;
#pragma unroll
      for (int q = 0; q < 16; q++) y += colp[q][tid];
      if (c0 + tid < n) P.part[(size_t)I * n + c0 + tid] = y;
      tv = y * vs[T + tid];
    } else if (I != J) {
      const int e = tid - T;
      double y = 0.0;
#pragma unroll
      for (int q = 0; q < 16; q++) y += rowp[q][e];
      if (r0 + e < n) P.part[(size_t)J * n + r0 + e] = y;
      tv = y * vs[e];
    }
    tv = block_sum(tv, red);
    if (tid == 0) P.tpart[xt] = tv;
  } else {
    // panel dot products over one column chunk: cpart[c][p] = W_p . v,
    // cpart[c][P+p] = V_p . v; wave w takes rows p = w, w + 4, ... and lane l
    // elements l, l + 64, ... of the chunk.  The panel rows are loaded before
    // the reflector and v are known, all in one round trip.
    constexpr int NW = CHUNK / WAVE, PW = (PANEL + NW - 1) / NW, EW = CHUNK / WAVE;
    const int c = x;
    const int j = k % PANEL;
    const int c0 = k + 1 + c * CHUNK;
    const int lane = tid & (WAVE - 1), wv = tid / WAVE;
    double xw[PW][EW], xv[PW][EW];
#pragma unroll
    for (int r = 0; r < PW; r++) {
      const int p = wv + r * NW;
#pragma unroll
      for (int u = 0; u < EW; u++) {
        const int ii = c0 + lane + u * WAVE;
        const bool ok = p < j && ii < n;
        xw[r][u] = ok ? P.X[(size_t)(PANEL + p) * n + ii] : 0.0;
        xv[r][u] = ok ? P.X[(size_t)p * n + ii] : 0.0;
      }
    }
    double t, v0, alpha;
    tdb_reflector(P, k, n, t, v0, alpha);
    const int i = c0 + tid;
    vs[tid] = (i < n) ? ((i == k + 1) ? v0 : rowk[i]) : 0.0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PW; r++) {
      const int p = wv + r * NW;
      if (p >= j) break;
      double a1 = 0.0, a2 = 0.0;
#pragma unroll
      for (int u = 0; u < EW; u++) {
        a1 += xw[r][u] * vs[lane + u * WAVE];
        a2 += xv[r][u] * vs[lane + u * WAVE];
      }
      a1 = wave_sum(a1);
      a2 = wave_sum(a2);
      if (lane == 0) { P.cpart[(size_t)c * 2 * PANEL + p] = a1; P.cpart[(size_t)c * 2 * PANEL + PANEL + p] = a2; }
    }
  }
}

// At a panel start s0 (after the panel GEMM, which brings only the lower
// triangle and a band above it up to date), rows s0..s0+PANEL-1 are rebuilt
// from their columns: A[r][j] = A[j][r] for j > r.  Those rows are the
// reflector sources of the panel and must be current over their whole length.
// One workgroup per 64 columns j of a block: the 64 x PANEL strip A[j][s0..]
// staged through LDS and stored as PANEL rows of 64 contiguous entries.
__global__ __launch_bounds__(256) void tdb_refresh_kernel(const E2Block* __restrict__ blks, const E2TdbGrid G,
                                                          int s0, double* __restrict__ A0) {
  __shared__ double t[64][PANEL + 1];
  const int gi = tdb_find(G, blockIdx.x);
  const E2Block B = blks[G.blk[gi]];
  const int n = B.n;
  double* A = A0 + B.a_off;
  const int j0 = s0 + 64 * (blockIdx.x - G.cum[gi]);
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 64 * PANEL; idx += 256) {
    const int jj = idx / PANEL, r = idx % PANEL;
    const int j = j0 + jj;
    t[jj][r] = (j < n && s0 + r < n) ? A[(size_t)j * n + s0 + r] : 0.0;
  }
  __syncthreads();
  for (int idx = tid; idx < 64 * PANEL; idx += 256) {
    const int r = idx / 64, jj = idx % 64;
    const int j = j0 + jj, row = s0 + r;
    if (row < n && j < n && j > row) A[(size_t)row * n + j] = t[jj][r];
  }
}

// One workgroup finishes FIN_E consecutive elements (128 by default): the
// kernel reads about RT + 2j doubles per element, and at large m a
// 256-element chunk is bound by the bandwidth of the one CU that runs it, so
// short chunks spread a column's reads over more CUs.
// FG thread groups share a chunk's elements: group g sums the row-tile
// partials rt = g (mod FG) and the panel terms p = g (mod FG) of each element;
// group 0 adds the four sums in a fixed order and finishes the element.  Every
// global load that depends on nothing computed here (the partials, the panel
// rows, the reflector entries, row k+1, the per-tile and per-chunk partial dot
// products) is issued before the first barrier, so the kernel waits on memory
// about once.
#define FG 4
#define FIN_PR 12   // row-tile partials per thread loaded at once (all of them while RT <= FG * FIN_PR)
__global__ __launch_bounds__(FIN_E * FG) void tdb_fin_kernel(const E2Block* __restrict__ blks, const E2TdbGrid G,
                                                             int k, double* __restrict__ A0, double* __restrict__ WS) {
  constexpr int NQ = PANEL / FG;   // panel terms per thread
  constexpr int NTH = FIN_E * FG;
  static_assert(NTH >= CHUNK, "the panel dot products are summed by CHUNK threads");
  __shared__ double red[NTH / WAVE];
  __shared__ double c1s[PANEL], c2s[PANEL], vk1[PANEL], wk1[PANEL];
  __shared__ double ys[FG][FIN_E], as[FG][FIN_E];
  __shared__ double cred[2][CHUNK / PANEL][PANEL];
  const int gi = tdb_find(G, blockIdx.x);
  const E2Block B = blks[G.blk[gi]];
  const int n = B.n;
  if (k + 2 >= n) return;
  const int m = n - k - 1;
  // CT: the symv's column chunks (of its panel dot products); FT: this kernel's chunks
  const int CT = (m + CHUNK - 1) / CHUNK, RT = (m + SYMV_R - 1) / SYMV_R, NT = RT * (RT + 1) / 2;
  const int FT = (m + FIN_E - 1) / FIN_E;
  const int c = blockIdx.x - G.cum[gi];
  if (c >= FT) return;
  const TdbPtr P = tdb_ptrs(B, A0, WS);
  const int tid = threadIdx.x, j = k % PANEL;
  const int e = tid % FIN_E, g = tid / FIN_E;
  const double* rowk = P.A + (size_t)k * n;
  const int i = k + 1 + c * FIN_E + e;
  const bool panel_end = (j == PANEL - 1) || (k + 3 >= n);
  const bool in = i < n;
  // ---- loads
  double pr[FIN_PR];
#pragma unroll
  for (int u = 0; u < FIN_PR; u++) {
    const int rt = g + FG * u;
    pr[u] = (in && rt < RT) ? P.part[(size_t)rt * n + i] : 0.0;
  }
  double x1r[NQ], x2r[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int p = g + FG * q;
    x1r[q] = (in && p < j) ? P.X[(size_t)p * n + i] : 0.0;
    x2r[q] = (in && p < j) ? P.X[(size_t)(PANEL + p) * n + i] : 0.0;
  }
  const double vraw = (in && g == 0) ? rowk[i] : 0.0;
  const double arow = (in && g == 0 && !panel_end) ? P.A[(size_t)(k + 1) * n + i] : 0.0;
  // wave 1 also gathers the partials of element k+1 for W(k+1, j), which every chunk needs
  const bool wk = !panel_end && tid >= WAVE && tid < 2 * WAVE;
  double yk1 = 0.0;
  if (wk)
    for (int rt = tid - WAVE; rt < RT; rt += WAVE) yk1 += P.part[(size_t)rt * n + k + 1];
  if (!panel_end && tid < j) {
    vk1[tid] = P.X[(size_t)tid * n + k + 1];
    wk1[tid] = P.X[(size_t)(PANEL + tid) * n + k + 1];
  }
  // panel dot products summed over the column chunks, split 8 ways over the first CHUNK threads
  if (tid < CHUNK) {
    const int pp = tid % PANEL, gg = tid / PANEL;
    double a1 = 0.0, a2 = 0.0;
    if (pp < j)
      for (int cc = gg; cc < CT; cc += CHUNK / PANEL) {
        a1 += P.cpart[(size_t)cc * 2 * PANEL + pp];
        a2 += P.cpart[(size_t)cc * 2 * PANEL + PANEL + pp];
      }
    cred[0][gg][pp] = a1;
    cred[1][gg][pp] = a2;
  }
  double tv = 0.0;
  for (int q = tid; q < NT; q += NTH) tv += P.tpart[q];   // one per lower-triangle tile
  const double t = P.refl[0], v0 = P.refl[1];
  // this group's share of y_i = (A v)_i
  double yp = 0.0;
#pragma unroll
  for (int u = 0; u < FIN_PR; u++) yp += pr[u];
  for (int rt = g + FG * FIN_PR; in && rt < RT; rt += FG) yp += P.part[(size_t)rt * n + i];
  // ---- reductions
  const double vav = block_sum(tv, red);   // also orders the cred, vk1, wk1 stores
  if (tid < j) {
    double a1 = 0.0, a2 = 0.0;
    for (int gg = 0; gg < CHUNK / PANEL; gg++) { a1 += cred[0][gg][tid]; a2 += cred[1][gg][tid]; }
    c1s[tid] = a1;   // W_p . v
    c2s[tid] = a2;   // V_p . v
  }
  __syncthreads();
  double cc12 = 0.0;
  for (int p = 0; p < j; p++) cc12 += c1s[p] * c2s[p];
  const double alpha2 = -0.5 * t * t * (vav - 2.0 * cc12);
  if (wk) {
    // W(k+1, j) = t (y_{k+1} - sum_p (V_p c1_p + W_p c2_p)_{k+1}) + alpha2 v0
    const int l = tid - WAVE;
    double y = yk1;
    if (l < j) y -= vk1[l] * c1s[l] + wk1[l] * c2s[l];
    y = wave_sum(y);
    if (l == 0) {
      wk1[j] = t * y + alpha2 * v0;
      vk1[j] = v0;
    }
  }
  // y_i -= sum_p (X_p c1_p + X_{P+p} c2_p) and, inside a panel, the correction
  // -sum_p (X_p wk1_p + X_{P+p} vk1_p) of column k+1 from the same X values
  double ap = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int p = g + FG * q;
    if (p < j) {
      yp -= x1r[q] * c1s[p] + x2r[q] * c2s[p];
      if (!panel_end) ap -= x1r[q] * wk1[p] + x2r[q] * vk1[p];
    }
  }
  ys[g][e] = yp;
  as[g][e] = ap;
  __syncthreads();   // the group sums and wk1[j] precede the stores below; every read of the panel rows precedes them
  double a = 0.0;
  if (g == 0 && in) {
    const double vi = (i == k + 1) ? v0 : vraw;
    const double y = (ys[0][e] + ys[1][e]) + (ys[2][e] + ys[3][e]);
    const double wi = t * y + alpha2 * vi;
    P.X[(size_t)j * n + i] = vi;
    P.X[(size_t)(2 * PANEL + j) * n + i] = vi;
    P.X[(size_t)(PANEL + j) * n + i] = wi;
    if (i == k + 1) P.A[(size_t)k * n + k + 1] = v0;
    if (!panel_end) {
      // column k+1 brought up to date with the panel's columns 0..j
      a = arow + ((as[0][e] + as[1][e]) + (as[2][e] + as[3][e]));
      a -= vi * wk1[j] + wi * vk1[j];
      P.A[(size_t)(k + 1) * n + i] = a;
    }
  }
  if (panel_end) return;
  const double sq = block_sum((g == 0 && i > k + 1 && in) ? a * a : 0.0, red);
  if (tid == 0) P.npart[c] = sq;
}

// ---------------------------------------------------------------------------
// stage 2: divide and conquer
// ---------------------------------------------------------------------------
// per-block vector workspace layout (doubles, offsets in units of n):
//   0 d | 1 e | 2 tau | 3 q | 4 part+refl | 5 dd (current eigenvalues) | 6 z
//   7 ds | 8 zs | 9 dk | 10 zk | 11 lam | 12 zhat | 13 dval | 14 gc | 15 gs | 16 erow (raw e)
// per-block int workspace (ints, offsets in units of n):
//   0 perm | 1 krow | 2 drow | 3 giv_a | 4 giv_b | 5 orig | 6 src
#define VD(k) (WS + B.v_off + (size_t)(k) * n)
#define VI(k) (IW + B.i_off + (size_t)(k) * n)

// zero the vector matrix and tear at every split point (sequential per block)
__global__ void dc_init_kernel(const E2Block* __restrict__ blks, const E2Merge* __restrict__ merges, int nmerge_total,
                               const int* __restrict__ mbeg, double* __restrict__ WS) {
  const int bi = blockIdx.y;
  const E2Block B = blks[bi];
  const int n = B.n;
  double* Qv = WS + B.q_off;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)n * n;
       idx += (size_t)gridDim.x * blockDim.x)
    Qv[idx] = 0.0;
  if (blockIdx.x == 0) {
    // the raw e saved, then the tears: split points are at least LEAF/2 apart,
    // so no two merges touch the same d or e entry and they run in parallel
    double* d = VD(0);
    double* e = VD(1);
    double* erow = VD(16);
    for (int i = threadIdx.x; i < n; i += blockDim.x) erow[i] = e[i];
    __syncthreads();
    for (int m = mbeg[bi] + threadIdx.x; m < mbeg[bi + 1]; m += blockDim.x) {
      const int mid = merges[m].mid;
      const double r = fabs(erow[mid - 1]);
      d[mid - 1] -= r;
      d[mid] -= r;
      e[mid - 1] = 0.0;
    }
  }
  (void)nmerge_total;
}

// implicit QL on one leaf [lo,hi): thread 0 chases, the workgroup applies the
// rotations to the leaf's vector rows; result sorted ascending
__global__ __launch_bounds__(64) void dc_leaf_kernel(const E2Block* __restrict__ blks, const E2Leaf* __restrict__ leaves,
                                                     double* __restrict__ WS) {
  __shared__ double d[LEAF], e[LEAF], cs[LEAF], sn[LEAF], z[LEAF][LEAF + 1];
  __shared__ int sh_lo, sh_hi, sh_done;
  __shared__ int rank[LEAF];
  const E2Leaf L = leaves[blockIdx.x];
  const E2Block B = blks[L.blk];
  const int n = B.n, lo = L.lo, m = L.hi - L.lo, tid = threadIdx.x;
  const double* dg = VD(0);
  const double* eg = VD(1);
  for (int i = tid; i < m; i += blockDim.x) { d[i] = dg[lo + i]; e[i] = (i + 1 < m) ? eg[lo + i] : 0.0; }
  for (int i = tid; i < m * m; i += blockDim.x) z[i / m][i % m] = (i / m == i % m) ? 1.0 : 0.0;
  __syncthreads();
  int l = 0, iter = 0;
  for (;;) {
    if (tid == 0) {
      int done = 0, lo_r = 0, hi_r = -1;
      for (;;) {
        if (l >= m) { done = 1; break; }
        int mm;
        for (mm = l; mm + 1 < m; mm++) {
          const double dd = fabs(d[mm]) + fabs(d[mm + 1]);
          if (fabs(e[mm]) <= 1e-300 || fabs(e[mm]) + dd == dd) break;
        }
        if (mm == l || iter >= 60) { l++; iter = 0; continue; }
        iter++;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = sqrt(g * g + 1.0);
        g = d[mm] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
        double s = 1.0, c = 1.0, p = 0.0;
        int i, under = 0;
        for (i = mm - 1; i >= l; i--) {
          const double f = s * e[i], b = c * e[i];
          r = hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) { d[i + 1] -= p; e[mm] = 0.0; under = 1; break; }
          s = f / r; c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          cs[i] = c; sn[i] = s;
        }
        if (!under) { d[l] -= p; e[l] = g; e[mm] = 0.0; lo_r = l; }
        else lo_r = i + 1;
        hi_r = mm;
        break;
      }
      sh_done = done; sh_lo = lo_r; sh_hi = hi_r;
    }
    __syncthreads();
    if (sh_done) break;
    const int a = sh_lo, b = sh_hi;
    if (b > a && tid < m) {
      double carry = z[b][tid];
      for (int i = b - 1; i >= a; i--) {
        const double zi = z[i][tid], c = cs[i], s = sn[i];
        z[i + 1][tid] = s * zi + c * carry;
        carry = c * zi - s * carry;
      }
      z[a][tid] = carry;
    }
    __syncthreads();
  }
  if (tid < m) {
    int r = 0;
    for (int i = 0; i < m; i++) r += (d[i] < d[tid]) || (d[i] == d[tid] && i < tid);
    rank[tid] = r;
  }
  __syncthreads();
  double* dd = VD(5);
  double* Qv = WS + B.q_off;
  for (int i = 0; i < m; i++) {
    const int r = rank[i];
    if (tid == 0) dd[lo + r] = d[i];
    if (tid < m) Qv[(size_t)(lo + r) * n + lo + tid] = z[i][tid];
  }
}

// merge preparation: z vector, merge of the two sorted halves (parallel, by
// binary search), deflation scan (sequential, staged through LDS), Givens
// rotations of the deflated pairs
__device__ __forceinline__ int count_less(const double* a, int n, double x) {   // #{a_i < x}, a ascending
  int lo = 0, hi = n;
  while (lo < hi) { const int mid = (lo + hi) >> 1; if (a[mid] < x) lo = mid + 1; else hi = mid; }
  return lo;
}
__device__ __forceinline__ int count_leq(const double* a, int n, double x) {    // #{a_i <= x}, a ascending
  int lo = 0, hi = n;
  while (lo < hi) { const int mid = (lo + hi) >> 1; if (a[mid] <= x) lo = mid + 1; else hi = mid; }
  return lo;
}

__global__ __launch_bounds__(E2_T) void dc_prep_kernel(const E2Block* __restrict__ blks,
                                                       const E2Merge* __restrict__ merges, E2MState* __restrict__ ms,
                                                       double* __restrict__ WS, int* __restrict__ IW) {
  __shared__ double red[E2_T / WAVE];
  __shared__ double cds[E2_T], czs[E2_T];
  __shared__ int cpm[E2_T];
  const E2Merge M = merges[blockIdx.x];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, mid = M.mid, hi = M.hi, m = hi - lo, n1 = mid - lo, tid = threadIdx.x;
  double* Qv = WS + B.q_off;
  const double* dd = VD(5);
  double* z = VD(6) + lo;
  double* ds = VD(7) + lo;
  double* zs = VD(8) + lo;
  double* dk = VD(9) + lo;
  double* zk = VD(10) + lo;
  double* dval = VD(13) + lo;
  double* gc = VD(14) + lo;
  double* gs = VD(15) + lo;
  const double er = VD(16)[mid - 1];
  int* perm = VI(0) + lo;
  int* krow = VI(1) + lo;
  int* drow = VI(2) + lo;
  int* ga = VI(3) + lo;
  int* gb = VI(4) + lo;
  const double theta = (er >= 0.0) ? 1.0 : -1.0, beta = fabs(er);
  double part = 0.0;
  for (int i = tid; i < m; i += blockDim.x) {
    const double zi = (i < n1) ? Qv[(size_t)(lo + i) * n + (mid - 1)] : theta * Qv[(size_t)(lo + i) * n + mid];
    z[i] = zi;
    part += zi * zi;
  }
  const double zn2 = block_sum(part, red);
  const double zn = sqrt(zn2);
  const double rho = beta * zn2;
  // stable merge of the ascending runs [0,n1) and [n1,m): ties keep the left run first
  double amax = 0.0;
  for (int i = tid; i < m; i += blockDim.x) {
    const double x = dd[lo + i];
    const int pos = (i < n1) ? i + count_less(dd + lo + n1, m - n1, x) : (i - n1) + count_leq(dd + lo, n1, x);
    perm[pos] = i;
    ds[pos] = x;
    zs[pos] = (zn > 0.0) ? z[i] / zn : 0.0;
    amax = fmax(amax, fabs(x));
  }
  {
    double v = amax;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
    const int lane = tid & (WAVE - 1), wid = tid / WAVE;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    amax = 0.0;
    for (int w = 0; w < (int)(blockDim.x / WAVE); w++) amax = fmax(amax, red[w]);
  }
  const double tol = 8.0 * 2.220446049250313e-16 * fmax(amax, rho);
  // deflation scan in sorted order: thread 0 walks each LDS-staged chunk
  int k = 0, nd = 0, ng = 0, pend = -1, pperm = 0;
  double pds = 0.0, pzs = 0.0;
  for (int c0 = 0; c0 < m; c0 += E2_T) {
    const int cnt = min(E2_T, m - c0);
    __syncthreads();
    if (tid < cnt) { cds[tid] = ds[c0 + tid]; czs[tid] = zs[c0 + tid]; cpm[tid] = perm[c0 + tid]; }
    __syncthreads();
    if (tid == 0) {
      for (int q = 0; q < cnt; q++) {
        double di = cds[q], zi = czs[q];
        const int pi = cpm[q];
        if (rho * fabs(zi) <= tol) { dval[nd] = di; drow[nd] = pi; nd++; continue; }
        if (pend < 0) { pend = c0 + q; pds = di; pzs = zi; pperm = pi; continue; }
        // the test |gap c s| <= tol with c = zi/t, s = -pzs/t, t^2 = zi^2 + pzs^2,
        // multiplied through by t^2 (z is normalised, so t^2 <= 1 does not overflow)
        const double t2 = zi * zi + pzs * pzs;
        const double gap = di - pds;
        if (fabs(gap * zi * pzs) <= tol * t2) {
          const double tv = sqrt(t2);
          const double cc = zi / tv, sv = -pzs / tv;
          ga[ng] = pperm; gb[ng] = pi; gc[ng] = cc; gs[ng] = sv; ng++;
          const double dp = pds * cc * cc + di * sv * sv;
          di = pds * sv * sv + di * cc * cc;
          zi = tv;
          dval[nd] = dp; drow[nd] = pperm; nd++;
        } else {
          dk[k] = pds; zk[k] = pzs; krow[k] = pperm; k++;
        }
        pend = c0 + q; pds = di; pzs = zi; pperm = pi;
      }
    }
  }
  if (tid == 0) {
    if (pend >= 0) { dk[k] = pds; zk[k] = pzs; krow[k] = pperm; k++; }
    E2MState st;
    st.k = k; st.nd = nd; st.ngiv = ng; st.pad = 0; st.rho = rho; st.tol = tol;
    ms[blockIdx.x] = st;
  }
}

// Givens rotations of the deflated pairs of each merge, in the order the
// deflation scan found them: q_a' = c q_a + s q_b, q_b' = -s q_a + c q_b on
// rows a, b of the eigenvector matrix over the merge range.  Every thread owns
// one column and applies the whole sequence to it, GV rotations at a time with
// their loads issued together.  The scan makes row b of a rotation one not
// rotated before, and row a either such a row or the b row of the rotation
// just before (a chain, carried in a register); an a row is never rotated
// again.  So the values loaded ahead are current.
#define GV 16
__global__ __launch_bounds__(E2_T) void dc_givens_kernel(const E2Block* __restrict__ blks,
                                                         const E2Merge* __restrict__ merges,
                                                         const E2MState* __restrict__ ms, double* __restrict__ WS,
                                                         int* __restrict__ IW) {
  const E2Merge M = merges[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, m = M.hi - M.lo;
  const int ng = ms[blockIdx.y].ngiv;
  if (ng == 0 || (int)blockIdx.x * E2_T >= m) return;
  const int i = blockIdx.x * E2_T + threadIdx.x;
  const bool act = i < m;
  double* col = WS + B.q_off + lo + (act ? i : 0);   // element (row r, column lo+i) at col[(lo+r)*n]
  const int* ga = VI(3) + lo;
  const int* gb = VI(4) + lo;
  const double* gc = VD(14) + lo;
  const double* gs = VD(15) + lo;
  for (int g0 = 0; g0 < ng; g0 += GV) {
    const int cnt = min(GV, ng - g0);
    int ra[GV], rb[GV];
    double xa[GV], xb[GV];
#pragma unroll
    for (int u = 0; u < GV; u++) {
      ra[u] = (u < cnt) ? ga[g0 + u] : 0;
      rb[u] = (u < cnt) ? gb[g0 + u] : 0;
      xa[u] = (act && u < cnt) ? col[(size_t)(lo + ra[u]) * n] : 0.0;
      xb[u] = (act && u < cnt) ? col[(size_t)(lo + rb[u]) * n] : 0.0;
    }
    int carry = -1;
    double xc = 0.0;
#pragma unroll
    for (int u = 0; u < GV; u++) {
      if (u < cnt) {
        const double c = gc[g0 + u], sn = gs[g0 + u];
        double a = xa[u];
        if (ra[u] == carry) a = xc;
        else if (carry >= 0 && act) col[(size_t)(lo + carry) * n] = xc;
        if (act) col[(size_t)(lo + ra[u]) * n] = c * a + sn * xb[u];
        xc = -sn * a + c * xb[u];
        carry = rb[u];
      }
    }
    if (carry >= 0 && act) col[(size_t)(lo + carry) * n] = xc;
  }
}

#define SEC_LDS 2048   // merges up to this many poles keep them in LDS for the root iterations
// secular equation root j: f(lambda) = 1 + rho sum_l zk_l^2/(dk_l - lambda), one
// wavefront per root.  lambda = dk[org] + tau with the nearer pole as origin;
// iteration: the two-pole rational model of the left and right sums (matching
// value and slope at the iterate), safeguarded by bisection of the bracket;
// delta(j,l) = dk_l - lambda_j
__global__ __launch_bounds__(E2_T) void dc_secular_kernel(const E2Block* __restrict__ blks,
                                                          const E2Merge* __restrict__ merges,
                                                          const E2MState* __restrict__ ms, double* __restrict__ WS,
                                                          int* __restrict__ IW) {
  // the poles and weights of the merge, staged once per workgroup (its
  // wavefronts solve roots of the same merge); larger merges read them from
  // global memory
  __shared__ double sdk[SEC_LDS], szk[SEC_LDS];
  const E2Merge M = merges[blockIdx.y];
  const E2MState S = ms[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, k = S.k;
  const int j0 = blockIdx.x * (E2_T / WAVE);
  if (j0 >= k) return;
  const int j = j0 + threadIdx.x / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  const double* gdk = VD(9) + lo;
  const double* gzk = VD(10) + lo;
  const bool staged = k <= SEC_LDS;
  if (staged) {
    for (int l = threadIdx.x; l < k; l += E2_T) { sdk[l] = gdk[l]; szk[l] = gzk[l]; }
    __syncthreads();
  }
  if (j >= k) return;
  const double* dk = staged ? sdk : gdk;
  const double* zk = staged ? szk : gzk;
  double* lam = VD(11) + lo;
  int* org = VI(5) + lo;
  const int m = M.hi - M.lo;
  double* delta = WS + B.dl_off + (size_t)lo * n;   // k rows of stride m, row j = delta(j, .)
  const double rho = S.rho;
  const double eps = 2.220446049250313e-16;
  int o;
  double tlo, thi;
  const bool last = (j == k - 1);
  if (!last) {
    const double gap = dk[j + 1] - dk[j], half = 0.5 * gap;
    double f = 0.0;
    for (int l = lane; l < k; l += WAVE) f += zk[l] * zk[l] / ((dk[l] - dk[j]) - half);
    f = 1.0 + rho * wave_sum(f);
    if (f >= 0.0) { o = j; tlo = 0.0; thi = half; }
    else { o = j + 1; tlo = -(gap - half); thi = 0.0; }
  } else {
    double zz = 0.0;
    for (int l = lane; l < k; l += WAVE) zz += zk[l] * zk[l];
    o = j; tlo = 0.0; thi = rho * wave_sum(zz);
  }
  const double dorg = dk[o];
  const double DL = dk[j] - dorg;                  // left pole, relative to the origin
  const double DR = last ? 0.0 : dk[j + 1] - dorg; // right pole
  double tau = 0.5 * (tlo + thi);
  for (int it = 0; it < 64; it++) {
    double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
    for (int l = lane; l < k; l += WAVE) {
      const double del = (dk[l] - dorg) - tau;
      const double w = zk[l] / del;
      if (l <= j) { psi += zk[l] * w; dpsi += w * w; }
      else { phi += zk[l] * w; dphi += w * w; }
    }
    psi = rho * wave_sum(psi); dpsi = rho * wave_sum(dpsi);
    phi = rho * wave_sum(phi); dphi = rho * wave_sum(dphi);
    const double f = 1.0 + psi + phi;
    if (f == 0.0) break;
    if (f < 0.0) tlo = tau; else thi = tau;
    // rational model in the step s: g(s) = C + a/(dl - s) + b/(dr - s)
    const double dl = DL - tau;
    const double a = dpsi * dl * dl, A = psi - dpsi * dl;
    double s;
    bool ok = true;
    if (last) {
      const double C = 1.0 + A;
      if (C > 0.0) s = dl + a / C; else ok = false;
    } else {
      const double dr = DR - tau;
      const double b = dphi * dr * dr, Bc = phi - dphi * dr;
      const double C = 1.0 + A + Bc;
      const double qa = C, qb = -(C * (dl + dr) + a + b), qc = f * dl * dr;
      const double disc = qb * qb - 4.0 * qa * qc;
      if (disc < 0.0 || qa == 0.0) ok = false;
      else {
        const double q = -0.5 * (qb + copysign(sqrt(disc), qb));
        const double s1 = q / qa, s2 = (q != 0.0) ? qc / q : 0.0;
        const double blo = tlo - tau, bhi = thi - tau;
        const bool in1 = s1 > blo && s1 < bhi, in2 = s2 > blo && s2 < bhi;
        if (in2 && (!in1 || fabs(s2) < fabs(s1))) s = s2;
        else if (in1) s = s1;
        else ok = false;
      }
    }
    double tn = ok ? tau + s : 0.5 * (tlo + thi);
    // a model step below the resolution of tau is convergence: tested before
    // the bracket check, which rejects tn == tau once the bracket has closed on
    // tau and would send the iterate back to the bracket's far half
    if (ok && fabs(tn - tau) <= 2.0 * eps * fabs(tau)) { tau = tn; break; }
    if (!(tn > tlo && tn < thi)) tn = 0.5 * (tlo + thi);
    const bool done = fabs(tn - tau) <= 2.0 * eps * fabs(tn) || (thi - tlo) <= 4.0 * eps * fmax(fabs(tlo), fabs(thi));
    tau = tn;
    if (done) break;
  }
  if (lane == 0) { lam[j] = dorg + tau; org[j] = o; }
  double* dj = delta + (size_t)j * m;
  for (int l = lane; l < k; l += WAVE) dj[l] = (dk[l] - dorg) - tau;
}

// Gu-Eisenstat: zhat_l^2 = prod_j (lambda_j - d_l) / (rho prod_{j!=l} (d_j - d_l)).
// ZG threads per element l, each multiplying the factors of a contiguous range
// of j in order; the ZG partial products are then multiplied in range order
#define ZG 8
__global__ __launch_bounds__(E2_T) void dc_zhat_kernel(const E2Block* __restrict__ blks,
                                                       const E2Merge* __restrict__ merges,
                                                       const E2MState* __restrict__ ms, double* __restrict__ WS) {
  __shared__ double part[E2_T];
  const E2Merge M = merges[blockIdx.y];
  const E2MState S = ms[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, k = S.k;
  if ((int)blockIdx.x * (E2_T / ZG) >= k) return;
  const int tid = threadIdx.x, g = tid % ZG;
  const int l = blockIdx.x * (E2_T / ZG) + tid / ZG;
  const double* dk = VD(9) + lo;
  const double* zk = VD(10) + lo;
  double* zh = VD(12) + lo;
  const int m = M.hi - M.lo;
  const double* delta = WS + B.dl_off + (size_t)lo * n;
  double p = 1.0;
  if (l < k) {
    const int c = (k + ZG - 1) / ZG, j0 = g * c, j1 = min(k, j0 + c);
    const double dl = dk[l];
    for (int j = j0; j < j1; j++) {
      if (j == l) continue;
      p *= -delta[(size_t)j * m + l] / (dk[j] - dl);
    }
  }
  part[tid] = p;
  __syncthreads();
  if (g == 0 && l < k) {
    double q = -delta[(size_t)l * m + l] / S.rho;
    for (int h = 0; h < ZG; h++) q *= part[tid + h];
    const double a = sqrt(fabs(q));
    zh[l] = (zk[l] >= 0.0) ? a : -a;
  }
}

// eigenvector j of the rank-one problem: u_j(l) = zhat_l / (d_l - lambda_j),
// normalised; one wavefront per vector
__global__ __launch_bounds__(E2_T) void dc_vec_kernel(const E2Block* __restrict__ blks,
                                                      const E2Merge* __restrict__ merges,
                                                      const E2MState* __restrict__ ms, double* __restrict__ WS) {
  const E2Merge M = merges[blockIdx.y];
  const E2MState S = ms[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, k = S.k;
  const int j = blockIdx.x * (E2_T / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  if (j >= k) return;
  const int m = M.hi - M.lo;
  const double* zh = VD(12) + lo;
  const double* dj = WS + B.dl_off + (size_t)lo * n + (size_t)j * m;
  double* uj = WS + B.u_off + (size_t)lo * n + (size_t)j * m;
  double s = 0.0;
  for (int l = lane; l < k; l += WAVE) {
    const double u = zh[l] / dj[l];
    uj[l] = u;
    s += u * u;
  }
  const double inv = 1.0 / sqrt(wave_sum(s));
  for (int l = lane; l < k; l += WAVE) uj[l] *= inv;
}

// gather the non-deflated vector rows (components [lo,hi)) into Qk (k x m) and
// the deflated rows into rows k..m-1 of Qt (the GEMM writes rows 0..k-1)
__global__ void dc_gather_kernel(const E2Block* __restrict__ blks, const E2Merge* __restrict__ merges,
                                 const E2MState* __restrict__ ms, double* __restrict__ WS, const int* __restrict__ IW) {
  const E2Merge M = merges[blockIdx.y];
  const E2MState S = ms[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, m = M.hi - M.lo, k = S.k;
  const int* krow = VI(1) + lo;
  const int* drow = VI(2) + lo;
  const double* Qv = WS + B.q_off;
  double* Qk = WS + B.k_off + (size_t)lo * n;
  double* Qt = WS + B.t_off + (size_t)lo * n;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)m * m;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / m), c = (int)(idx % m);
    if (r < k) Qk[idx] = Qv[(size_t)(lo + krow[r]) * n + lo + c];
    else Qt[idx] = Qv[(size_t)(lo + drow[r - k]) * n + lo + c];
  }
}

// final order of one merge: the deflated values are sorted by rank (ties by
// index), then merged with the ascending roots by binary search (roots first
// on ties); writes the eigenvalues and the source row of every output row
// rank of each deflated value (ties by index): element i goes to position
// #{y < x_i} + #{y == x_i, index < i} of the sorted deflated list; one thread
// per element, E2_T elements per workgroup, every value staged through LDS
__global__ __launch_bounds__(E2_T) void dc_rank_kernel(const E2Block* __restrict__ blks,
                                                       const E2Merge* __restrict__ merges,
                                                       const E2MState* __restrict__ ms, double* __restrict__ WS,
                                                       int* __restrict__ IW) {
  __shared__ double cv[E2_T];
  const E2Merge M = merges[blockIdx.y];
  const E2MState S = ms[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, nd = S.nd, tid = threadIdx.x;
  if ((int)blockIdx.x * E2_T >= nd) return;
  const double* dval = VD(13) + lo;
  double* dvs = VD(7) + lo;   // sorted deflated values (the ds scratch of the merge)
  int* dix = VI(0) + lo;      // their indices (the perm scratch)
  const int i = blockIdx.x * E2_T + tid;
  const double xv = (i < nd) ? dval[i] : 0.0;
  int rk = 0;
  for (int c0 = 0; c0 < nd; c0 += E2_T) {
    const int cnt = min(E2_T, nd - c0);
    __syncthreads();
    if (tid < cnt) cv[tid] = dval[c0 + tid];
    __syncthreads();
    if (i < nd)
      for (int t = 0; t < cnt; t++) {
        const double y = cv[t];
        rk += (y < xv) || (y == xv && c0 + t < i);
      }
  }
  if (i < nd) { dvs[rk] = xv; dix[rk] = i; }
}

// final order of a merge: the roots and the sorted deflated values merged by
// binary search (ties: roots first)
__global__ __launch_bounds__(E2_T) void dc_order_kernel(const E2Block* __restrict__ blks,
                                                        const E2Merge* __restrict__ merges,
                                                        const E2MState* __restrict__ ms, double* __restrict__ WS,
                                                        int* __restrict__ IW) {
  const E2Merge M = merges[blockIdx.x];
  const E2MState S = ms[blockIdx.x];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, k = S.k, nd = S.nd, tid = threadIdx.x;
  const double* lam = VD(11) + lo;
  double* dd = VD(5);
  const double* dvs = VD(7) + lo;
  const int* dix = VI(0) + lo;
  int* src = VI(6) + lo;
  for (int a = tid; a < k; a += E2_T) {
    const int pos = a + count_less(dvs, nd, lam[a]);
    dd[lo + pos] = lam[a];
    src[pos] = a;
  }
  for (int b = tid; b < nd; b += E2_T) {
    const int pos = b + count_leq(lam, k, dvs[b]);
    dd[lo + pos] = dvs[b];
    src[pos] = k + dix[b];
  }
}

// output rows of one merge: Qv row lo+o <- Qt row src[o] (Qt rows 0..k-1 are
// the GEMM output, k..m-1 the deflated rows)
__global__ void dc_copy_kernel(const E2Block* __restrict__ blks, const E2Merge* __restrict__ merges,
                               double* __restrict__ WS, const int* __restrict__ IW) {
  const E2Merge M = merges[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, m = M.hi - M.lo;
  const int* src = VI(6) + lo;
  double* Qv = WS + B.q_off;
  const double* Qt = WS + B.t_off + (size_t)lo * n;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)m * m;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int o = (int)(idx / m), c = (int)(idx % m);
    Qv[(size_t)(lo + o) * n + lo + c] = Qt[(size_t)src[o] * m + c];
  }
}

// ---------------------------------------------------------------------------
// stage 3: compact-WY panels.  Reflector i of a panel is row k0+i of the
// caller's matrix from column k0+1 on (zero up to and including the diagonal),
// T from the Gram matrix G = V V^T (T_ii = tau_i, T[0:i,i] = -tau_i T[0:i,0:i]
// G[0:i,i]), and each round applies Z[:, k0+1:] -= (Z V^T) T^T V.
// ---------------------------------------------------------------------------
// the entries of reflector row k0+a left of where its reflector starts: zero
// in the reflector; the reduction left other values there
__global__ void bt_zero_kernel(const E2Block* __restrict__ blks, const E2BtPanel* __restrict__ pans,
                               double* __restrict__ A0) {
  const E2BtPanel Pn = pans[blockIdx.x];
  const E2Block B = blks[Pn.blk];
  const int n = B.n;
  double* A = A0 + B.a_off;
  // one-stage: reflector k0+a starts at column k0+a+1; two-stage (voff = E2_TSB):
  // reflector k0+a starts at column k0+E2_TSB*(a/E2_TSB+1) (its stage-1 panel's band end)
  for (int idx = threadIdx.x; idx < BTP * BTP; idx += blockDim.x) {
    const int a = idx / BTP, jj = idx % BTP;
    const int nz = Pn.voff == 1 ? a : E2_TSB * (a / E2_TSB);
    if (a < Pn.nb && jj < nz) A[(size_t)(Pn.k0 + a) * n + Pn.k0 + Pn.voff + jj] = 0.0;
  }
}

// T of the compact-WY form, one workgroup per panel: G summed from its ns
// partial products, T built in LDS (T_ii = tau_i, T[0:i,i] = -tau_i
// T[0:i,0:i] G[0:i,i]) and stored transposed
__global__ __launch_bounds__(4 * BTP) void bt_tmat_kernel(const E2Block* __restrict__ blks,
                                                          const E2BtPanel* __restrict__ pans,
                                                          const double* __restrict__ WS, double* __restrict__ aux) {
  __shared__ double T[BTP][BTP + 1];
  __shared__ double Gs[BTP][BTP + 1];   // the Gram matrix, staged once: the recurrence below reads it nb times
  __shared__ double ts[BTP];
  const E2BtPanel Pn = pans[blockIdx.x];
  const E2Block B = blks[Pn.blk];
  const int n = B.n, k0 = Pn.k0, nb = Pn.nb;
  const double* tau = WS + B.v_off + 2 * (size_t)n;
  const double* G = aux + Pn.g_off;
  double* Tg = aux + Pn.t_off;   // BTP x BTP, T^T row-major
  const int tid = threadIdx.x;
  for (int i = tid; i < BTP * BTP; i += blockDim.x) {
    const int r = i / BTP, c = i % BTP;
    double g = 0.0;
    if (r < nb && c < nb)
      for (int q = 0; q < Pn.ns; q++) g += G[(size_t)q * BTP * BTP + i];
    T[r][c] = 0.0;
    Gs[r][c] = g;
  }
  if (tid < nb) ts[tid] = tau[k0 + tid];
  __syncthreads();
  // column i of T: T(0:i-1, i) = -tau_i T(0:i-1, 0:i-1) G(0:i-1, i); four
  // neighbouring lanes share row r (every fourth term each), summed by shuffles
  const int r = tid >> 2, q = tid & 3;
  for (int i = 0; i < nb; i++) {
    const double ti = ts[i];
    double w = 0.0;
    if (r < i)
      for (int c = r + q; c < i; c += 4) w += T[r][c] * Gs[c][i];
    w += __shfl_xor(w, 1, WAVE);
    w += __shfl_xor(w, 2, WAVE);
    __syncthreads();
    if (q == 0 && r < i) T[r][i] = -ti * w;
    if (tid == 0) T[i][i] = ti;
    __syncthreads();
  }
  for (int i = tid; i < BTP * BTP; i += blockDim.x) {
    const int r = i / BTP, c = i % BTP;
    Tg[i] = T[c][r];   // T^T
  }
}

// W2 = (sum_s W_s) T^T for one 64-row tile: the partial products of the split
// K range summed in a fixed order, columns >= nb (no reflector) taken as zero;
// each thread forms a 4 x 4 block of the tile
__global__ __launch_bounds__(256) void bt_w2_kernel(const E2BtRows* __restrict__ rows, double* __restrict__ aux) {
  __shared__ double Ws[BTP][BTP + 1];
  __shared__ double Ts[BTP][BTP + 1];
  constexpr int EPT = BTP * BTP / 256;   // elements of the tile per thread
  const E2BtRows R = rows[blockIdx.x];
  const int tid = threadIdx.x;
  const double* Tg = aux + R.t_off;
  // the partials summed in order q = 0, 1, ...; the loads of two partials of
  // all this thread's elements are issued together
  double w[EPT];
  bool ok[EPT];
  size_t off[EPT];
#pragma unroll
  for (int u = 0; u < EPT; u++) {
    const int i = tid + 256 * u, r = i / BTP, c = i % BTP, row = R.row0 + r;
    w[u] = 0.0;
    ok[u] = row < R.krow && c < R.nb;
    off[u] = R.wp_off + (size_t)row * BTP + c;
  }
  const size_t qstride = (size_t)R.krow * BTP;
  int q = 0;
  for (; q + 2 <= R.ns; q += 2) {
    double x0[EPT], x1[EPT];
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      x0[u] = ok[u] ? aux[off[u] + q * qstride] : 0.0;
      x1[u] = ok[u] ? aux[off[u] + (q + 1) * qstride] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < EPT; u++) w[u] = (w[u] + x0[u]) + x1[u];
  }
  if (q < R.ns) {
#pragma unroll
    for (int u = 0; u < EPT; u++) w[u] += ok[u] ? aux[off[u] + q * qstride] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < EPT; u++) {
    const int i = tid + 256 * u;
    Ws[i / BTP][i % BTP] = w[u];
    Ts[i / BTP][i % BTP] = Tg[i];
  }
  __syncthreads();
  const int r0 = (tid / 16) * 4, c0 = (tid % 16) * 4;
  double acc[4][4] = {};
  for (int l = 0; l < BTP; l++) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; u++) { a[u] = Ws[r0 + u][l]; b[u] = Ts[l][c0 + u]; }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) acc[u][v] += a[u] * b[v];
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int row = R.row0 + r0 + u;
    if (row < R.krow)
#pragma unroll
      for (int v = 0; v < 4; v++) aux[R.w2_off + (size_t)row * BTP + c0 + v] = acc[u][v];
  }
}

// ---------------------------------------------------------------------------
// batched grouped GEMM, FP64 on the matrix cores, strided operands:
//   C(m x n) = alpha * sum_k A(i,k) B(k,j) (+ C if accumulate)
//   A(i,k) = A[a_off + i*as0 + k*as1],  B(k,j) = B[b_off + k*bs0 + j*bs1]
// 64x64 tile per 256-thread workgroup; wavefront w owns the 32x32 quadrant
// (w/2, w%2) as 2x2 blocks of v_mfma_f64_16x16x4.  Operand lanes: a = A(l%16,
// l/16), b = B(l/16, l%16); accumulator r of lane l is C(l/16 + 4r, l%16).
// K advances in steps of 16 through LDS (k-major), the next step's global
// loads issued before the current step's MFMAs.  Tile loads pick the thread
// mapping that keeps the contiguous dimension along the wavefront.
// dyn >= 0: m and k are read from dyns[dyn] (sizes only known on the device).
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
#define GT 64
#define GK 16
#define GP 4   // LDS row padding (doubles)
__device__ __forceinline__ void gemm_tile(const E2Gemm& P, int M, int K, int m0, int n0,
                                          const double* __restrict__ Abase, const double* __restrict__ Bbase,
                                          double* __restrict__ Cbase, double (&As)[GK][GT + GP],
                                          double (&Bs)[GK][GT + GP]) {
  if (m0 >= M || n0 >= P.n) return;
  const double* A = Abase + P.a_off;
  const double* Bm = Bbase + P.b_off;
  double* C = Cbase + P.c_off;
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  const bool a_kcontig = (P.as1 == 1);
  const bool b_ncontig = (P.bs1 == 1);
  // this thread's four tile-load slots
  int arow[4], akk[4], bcol[4], bkk[4];
#pragma unroll
  for (int rr = 0; rr < 4; rr++) {
    if (a_kcontig) { arow[rr] = (tid >> 4) + 16 * rr; akk[rr] = tid & 15; }
    else { arow[rr] = tid & 63; akk[rr] = (tid >> 6) + 4 * rr; }
    if (b_ncontig) { bcol[rr] = tid & 63; bkk[rr] = (tid >> 6) + 4 * rr; }
    else { bcol[rr] = (tid >> 4) + 16 * rr; bkk[rr] = tid & 15; }
  }
  double ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const int gi = m0 + arow[rr], gk = k0 + akk[rr];
      ra[rr] = (gi < M && gk < K) ? A[(int64_t)gi * P.as0 + (int64_t)gk * P.as1] : 0.0;
      const int gj = n0 + bcol[rr], gk2 = k0 + bkk[rr];
      rb[rr] = (gj < P.n && gk2 < K) ? Bm[(int64_t)gk2 * P.bs0 + (int64_t)gj * P.bs1] : 0.0;
    }
  };
  v4d acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
  load(0);
  for (int k0 = 0; k0 < K; k0 += GK) {
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      As[akk[rr]][arow[rr]] = ra[rr];
      Bs[bkk[rr]][bcol[rr]] = rb[rr];
    }
    __syncthreads();
    if (k0 + GK < K) load(k0 + GK);
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const int k = kk + (lane >> 4), c16 = lane & 15;
      const double a0 = As[k][wr + c16], a1 = As[k][wr + 16 + c16];
      const double b0 = Bs[k][wc + c16], b1 = Bs[k][wc + 16 + c16];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi = m0 + wr + 16 * x + (lane >> 4) + 4 * r;
        const int gj = n0 + wc + 16 * y + (lane & 15);
        if (gi < M && gj < P.n) {
          double v = P.alpha * acc[x][y][r];
          double* cp = C + (int64_t)gi * P.cs0 + gj;
          if (P.accumulate) v += *cp;
          *cp = v;
        }
      }
}

__global__ __launch_bounds__(256) void gemm_kernel(const E2Gemm* __restrict__ probs, const E2Tile* __restrict__ tiles,
                                                   const double* __restrict__ Abase, const double* __restrict__ Bbase,
                                                   double* __restrict__ Cbase, const E2MState* __restrict__ dyns) {
  __shared__ double As[GK][GT + GP];
  __shared__ double Bs[GK][GT + GP];
  const E2Tile Tl = tiles[blockIdx.x];
  const E2Gemm P = probs[Tl.prob];
  int M = P.m, K = P.k;
  if (P.dyn >= 0) { M = dyns[P.dyn].k; K = M; }
  gemm_tile(P, M, K, Tl.tm * GT, Tl.tn * GT, Abase, Bbase, Cbase, As, Bs);
}

// The panel updates of the large blocks' reduction: problem p of the launch
// (an m x m result, m = probs[p].m) covers tile rows a < tt = ceil(m / 64) and,
// in row a, tile columns 0 .. min(a + 2, tt - 1): the lower triangle and the
// band of two tiles above it.  Row a then has min(a + 3, tt) tiles, rows
// a < tt - 2 have a + 3 and start at tile a (a + 5) / 2, the last two rows
// (or all rows when tt <= 2) have tt.  Workgroup x of problem p (cum[p] <= x <
// cum[p + 1]) takes tile x - cum[p] in row order, so no tile list is built.
__device__ __forceinline__ void band_tile(int tt, int t, int& a, int& b) {
  const int a0 = tt > 2 ? tt - 2 : 0;      // first full row
  const int s0 = a0 * (a0 + 5) / 2;        // tiles before it
  if (t >= s0) {
    a = a0 + (t - s0) / tt;
    b = (t - s0) % tt;
    return;
  }
  int r = (int)((sqrt(25.0 + 8.0 * t) - 5.0) * 0.5);   // largest r with r (r + 5) / 2 <= t
  while ((r + 1) * (r + 6) / 2 <= t) r++;
  while (r * (r + 5) / 2 > t) r--;
  a = r;
  b = t - r * (r + 5) / 2;
}
__global__ __launch_bounds__(256) void panel_gemm_kernel(const E2Gemm* __restrict__ probs, const E2PanelGrid G,
                                                         const double* __restrict__ Abase,
                                                         const double* __restrict__ Bbase, double* __restrict__ Cbase) {
  __shared__ double As[GK][GT + GP];
  __shared__ double Bs[GK][GT + GP];
  const int x = blockIdx.x;
  int p = 0;
#pragma unroll
  for (int i = 1; i < E2_PANEL_MAXP; i++) p += (i < G.np && G.cum[i] <= x) ? 1 : 0;
  const E2Gemm P = probs[p];
  int a, b;
  band_tile((P.m + GT - 1) / GT, x - G.cum[p], a, b);
  gemm_tile(P, P.m, P.k, a * GT, b * GT, Abase, Bbase, Cbase, As, Bs);
}

// copy eigenvalues (what & 1) and the first krow eigenvector rows (what & 2)
// to the caller's buffers
__global__ void e2_output_kernel(const E2Block* __restrict__ blks, const double* __restrict__ WS,
                                 double* __restrict__ A0, double* __restrict__ W0, int what) {
  const E2Block B = blks[blockIdx.y];
  const int n = B.n;
  const double* Qv = WS + B.q_off;
  double* A = A0 + B.a_off;
  if (what & 2)
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)B.krow * n;
         idx += (size_t)gridDim.x * blockDim.x)
      A[idx] = Qv[idx];
  if ((what & 1) && blockIdx.x == 0)
    for (int i = threadIdx.x; i < n; i += blockDim.x) W0[B.w_off + i] = WS[B.v_off + 5 * (size_t)n + i];
}

}  // namespace e2

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
// the launches of one reduction kernel over the blocks still being reduced at
// column k, E2_TDB_MAXB blocks per launch; per(m) = workgroups of a block with
// an m x m trailing matrix
template <class Per, class Launch>
static hipError_t tdb_launch_all(const int* blk_idx, const int* blk_n, int nlist, int k, Per per, Launch launch) {
  E2TdbGrid G;
  G.nb = 0;
  G.cum[0] = 0;
  for (int i = 0; i <= nlist; i++) {
    if (G.nb == E2_TDB_MAXB || (i == nlist && G.nb > 0)) {
      launch(G);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      G.nb = 0;
    }
    if (i == nlist) break;
    if (k + 2 >= blk_n[i]) continue;
    G.blk[G.nb] = blk_idx[i];
    G.cum[G.nb + 1] = G.cum[G.nb] + per(blk_n[i] - k - 1);
    G.nb++;
  }
  return hipSuccess;
}

extern "C" {

hipError_t e2_launch_tridiag_small(const E2Block* b, const int* list, int nlist, double* A, double* WS, hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::tridiag_small_kernel, dim3(nlist), dim3(SM_T), 0, s, b, list, A, WS);
  return hipGetLastError();
}

hipError_t e2_launch_tridiag_hand(const E2Block* b, const E2Hand* hands, int nh, double* A, double* WS, hipStream_t s) {
  if (nh <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::tridiag_hand_kernel, dim3(nh), dim3(SM_T), 0, s, b, hands, A, WS);
  return hipGetLastError();
}

hipError_t e2_launch_tdb_norm(const E2Block* b, const int* list, int nlist, int nmax, int k, double* A, double* WS,
                              hipStream_t s) {
  if (nlist <= 0 || k + 2 >= nmax) return hipSuccess;
  const int gx = (nmax - k + CHUNK - 1) / CHUNK;
  hipLaunchKernelGGL(e2::tdb_norm_kernel, dim3(gx, nlist), dim3(CHUNK), 0, s, b, list, k, A, WS);
  return hipGetLastError();
}

hipError_t e2_launch_tdb_symv(const E2Block* b, const int* blk_idx, const int* blk_n, int nlist, int k, double* A,
                              double* WS, hipStream_t s) {
  auto per = [](int m) {
    const int MT = (m + SYMV_R - 1) / SYMV_R;
    return MT * (MT + 1) / 2 + (m + CHUNK - 1) / CHUNK;
  };
  return tdb_launch_all(blk_idx, blk_n, nlist, k, per, [&](const E2TdbGrid& G) {
    hipLaunchKernelGGL(e2::tdb_symv_kernel, dim3(G.cum[G.nb]), dim3(CHUNK), 0, s, b, G, k, A, WS);
  });
}

hipError_t e2_launch_tdb_refresh(const E2Block* b, const int* blk_idx, const int* blk_n, int nlist, int s0, double* A,
                                 hipStream_t s) {
  // the blocks that reduce column s0 (n > s0 + 2), ceil((n - s0) / 64) workgroups each (m = n - s0 - 1)
  auto per = [](int m) { return (m + 64) / 64; };
  return tdb_launch_all(blk_idx, blk_n, nlist, s0, per, [&](const E2TdbGrid& G) {
    hipLaunchKernelGGL(e2::tdb_refresh_kernel, dim3(G.cum[G.nb]), dim3(256), 0, s, b, G, s0, A);
  });
}

hipError_t e2_launch_tdb_fin(const E2Block* b, const int* blk_idx, const int* blk_n, int nlist, int k, double* A,
                             double* WS, hipStream_t s) {
  auto per = [](int m) { return (m + FIN_E - 1) / FIN_E; };
  return tdb_launch_all(blk_idx, blk_n, nlist, k, per, [&](const E2TdbGrid& G) {
    hipLaunchKernelGGL(e2::tdb_fin_kernel, dim3(G.cum[G.nb]), dim3(FIN_E * FG), 0, s, b, G, k, A, WS);
  });
}

// tiles of the band product of an m x m result (see band_tile)
static int band_tiles(int m) {
  const int tt = (m + GT - 1) / GT, a0 = tt > 2 ? tt - 2 : 0;
  return a0 * (a0 + 5) / 2 + (tt - a0) * tt;
}

hipError_t e2_launch_panel_gemm(const E2Gemm* dprobs, const E2Gemm* hprobs, int np, const double* A, const double* B,
                                double* C, hipStream_t s) {
  for (int p0 = 0; p0 < np; p0 += E2_PANEL_MAXP) {
    E2PanelGrid G = {};
    G.np = np - p0 < E2_PANEL_MAXP ? np - p0 : E2_PANEL_MAXP;
    G.cum[0] = 0;
    for (int i = 0; i < G.np; i++) G.cum[i + 1] = G.cum[i] + band_tiles(hprobs[p0 + i].m);
    if (G.cum[G.np] <= 0) continue;
    hipLaunchKernelGGL(e2::panel_gemm_kernel, dim3(G.cum[G.np]), dim3(256), 0, s, dprobs + p0, G, A, B, C);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t e2_launch_gemm3(const E2Gemm* probs, const E2Tile* tiles, int ntiles, const double* A, const double* B,
                           double* C, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::gemm_kernel, dim3(ntiles), dim3(256), 0, s, probs, tiles, A, B, C,
                     (const E2MState*)nullptr);
  return hipGetLastError();
}


hipError_t e2_launch_dc_init(const E2Block* b, int nb, int nmax, const E2Merge* merges, int nmerge, const int* mbeg,
                             double* WS, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  size_t tot = (size_t)nmax * nmax;
  int gx = (int)((tot + 255) / 256); if (gx > 1024) gx = 1024; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(e2::dc_init_kernel, dim3(gx, nb), dim3(256), 0, s, b, merges, nmerge, mbeg, WS);
  return hipGetLastError();
}

hipError_t e2_launch_dc_leaf(const E2Block* b, const E2Leaf* leaves, int nleaf, double* WS, hipStream_t s) {
  if (nleaf <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::dc_leaf_kernel, dim3(nleaf), dim3(64), 0, s, b, leaves, WS);
  return hipGetLastError();
}

hipError_t e2_launch_dc_level(const E2Block* b, const E2Merge* merges, int nm, int mmax, E2MState* ms, double* WS,
                              int* IW, const E2Gemm* gprobs, const E2Tile* gtiles, int ntiles, hipStream_t s) {
  if (nm <= 0) return hipSuccess;
  hipError_t e;
  hipLaunchKernelGGL(e2::dc_prep_kernel, dim3(nm), dim3(E2_T), 0, s, b, merges, ms, WS, IW);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int gw = (mmax + E2_T / WAVE - 1) / (E2_T / WAVE);   // one wavefront per root / vector
  const int gy = (mmax + E2_T - 1) / E2_T;                   // one thread per component
  hipLaunchKernelGGL(e2::dc_givens_kernel, dim3(gy, nm), dim3(E2_T), 0, s, b, merges, ms, WS, IW);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(e2::dc_secular_kernel, dim3(gw, nm), dim3(E2_T), 0, s, b, merges, ms, WS, IW);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(e2::dc_zhat_kernel, dim3((mmax + E2_T / ZG - 1) / (E2_T / ZG), nm), dim3(E2_T), 0, s, b, merges,
                     ms, WS);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(e2::dc_vec_kernel, dim3(gw, nm), dim3(E2_T), 0, s, b, merges, ms, WS);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  int gx = (int)(((size_t)mmax * mmax + 255) / 256); if (gx > 1024) gx = 1024; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(e2::dc_gather_kernel, dim3(gx, nm), dim3(256), 0, s, b, merges, ms, WS, IW);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (ntiles > 0) {
    hipLaunchKernelGGL(e2::gemm_kernel, dim3(ntiles), dim3(256), 0, s, gprobs, gtiles, WS, WS, WS, ms);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(e2::dc_rank_kernel, dim3(gy, nm), dim3(E2_T), 0, s, b, merges, ms, WS, IW);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(e2::dc_order_kernel, dim3(nm), dim3(E2_T), 0, s, b, merges, ms, WS, IW);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(e2::dc_copy_kernel, dim3(gx, nm), dim3(256), 0, s, b, merges, WS, IW);
  return hipGetLastError();
}

hipError_t e2_launch_bt_zero(const E2Block* b, const E2BtPanel* pans, int npan, double* A, hipStream_t s) {
  if (npan <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::bt_zero_kernel, dim3(npan), dim3(256), 0, s, b, pans, A);
  return hipGetLastError();
}

hipError_t e2_launch_bt_tmat(const E2Block* b, const E2BtPanel* pans, int npan, const double* WS, double* aux,
                             hipStream_t s) {
  if (npan <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::bt_tmat_kernel, dim3(npan), dim3(4 * BTP), 0, s, b, pans, WS, aux);
  return hipGetLastError();
}

hipError_t e2_launch_bt_w2(const E2BtRows* rows, int nrows, double* aux, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::bt_w2_kernel, dim3(nrows), dim3(256), 0, s, rows, aux);
  return hipGetLastError();
}

hipError_t e2_launch_output(const E2Block* b, int nb, int nmax, const double* WS, double* A, double* W, int what,
                            hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  int gx = (what & 2) ? (int)(((size_t)nmax * nmax + 255) / 256) : 1; if (gx > 1024) gx = 1024; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(e2::e2_output_kernel, dim3(gx, nb), dim3(256), 0, s, b, WS, A, W, what);
  return hipGetLastError();
}

hipError_t e2_launch_gemm(const E2Gemm* probs, const E2Tile* tiles, int ntiles, const double* A, const double* B,
                          double* C, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::gemm_kernel, dim3(ntiles), dim3(256), 0, s, probs, tiles, A, B, C,
                     (const E2MState*)nullptr);
  return hipGetLastError();
}

}  // extern "C"
