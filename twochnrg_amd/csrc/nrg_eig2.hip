// nrg_eig2.hip -- batched FP64 symmetric eigensolver for uneven symmetry
// blocks, built from three stages that each run over all blocks at once:
//
//   1. Householder reduction to tridiagonal form.  Blocks with n <= 256: one
//      256-thread workgroup per block.  Larger blocks: two launches per column
//      step, each workgroup owning a 256-column slab (symv, then rank-2 update);
//      a thread owns one column, so every row access of the wavefront is one
//      contiguous 512-byte segment.
//   2. Divide and conquer on the tridiagonal matrix (Cuppen's tearing): leaves
//      of <= 32 rows by implicit QL; merges by z-vector deflation (small
//      components, and Givens rotation of nearly equal poles), secular-equation
//      roots by safeguarded Newton in the distance to the nearer pole,
//      Gu-Eisenstat recomputation of z for orthogonal vectors, and a GEMM for
//      the new vectors.
//   3. Back-transformation with blocked (compact WY) Householder panels:
//      three GEMMs per 32-reflector panel.
//
// Storage: eigenvectors are kept vector-major (row i = eigenvector i), the
// layout of dEigVec in the reference (src/NRGmatrixClass.cpp:647-654), so the
// last stage writes the output rows directly.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrg_eig2.h"

#define E2_T 256
#define WAVE 64
#define LEAF 32
#define PANEL 32

namespace e2 {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
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
__global__ __launch_bounds__(E2_T) void tridiag_small_kernel(const E2Block* __restrict__ blks,
                                                              const int* __restrict__ list, double* __restrict__ A0,
                                                              double* __restrict__ WS) {
  __shared__ double red[E2_T / WAVE];
  __shared__ double vsh[256];
  __shared__ double qsh[256];
  const E2Block B = blks[list[blockIdx.x]];
  const int n = B.n, tid = threadIdx.x;
  double* A = A0 + B.a_off;
  double* d = WS + B.v_off;
  double* e = d + n;
  double* tau = e + n;
  for (int k = 0; k + 2 < n; k++) {
    const double* rowk = A + (size_t)k * n;
    double x = (tid > k && tid < n) ? rowk[tid] : 0.0;
    const double nrm2 = block_sum(x * x, red);
    const double x0 = rowk[k + 1];
    if (nrm2 == 0.0) {
      if (tid == 0) { d[k] = rowk[k]; e[k] = 0.0; tau[k] = 0.0; }
      __syncthreads();
      continue;
    }
    const double nrm = sqrt(nrm2), alpha = (x0 > 0.0) ? -nrm : nrm, v0 = x0 - alpha;
    const double t = 2.0 / (nrm2 - x0 * x0 + v0 * v0);
    if (tid < n) vsh[tid] = (tid > k + 1) ? rowk[tid] : (tid == k + 1 ? v0 : 0.0);
    __syncthreads();
    // q_i = t * sum_j A[j][i] v_j  (thread i owns column i)
    double qi = 0.0;
    if (tid > k && tid < n) {
      for (int j = k + 1; j < n; j++) qi += A[(size_t)j * n + tid] * vsh[j];
      qi *= t;
    }
    const double vp = block_sum((tid > k && tid < n) ? qi * vsh[tid] : 0.0, red);
    const double K = 0.5 * t * vp;
    if (tid < n) qsh[tid] = (tid > k) ? qi - K * vsh[tid] : 0.0;
    __syncthreads();
    if (tid > k && tid < n) {
      const double vi = vsh[tid], q2 = qsh[tid];
      for (int j = k + 1; j < n; j++) {
        const double vj = vsh[j], qj = qsh[j];
        const double upd = (j <= tid) ? __fma_rn(vj, q2, vi * qj) : __fma_rn(vi, qj, vj * q2);
        A[(size_t)j * n + tid] -= upd;
      }
    }
    __syncthreads();
    if (tid == 0) { A[(size_t)k * n + k + 1] = v0; d[k] = A[(size_t)k * n + k]; e[k] = alpha; tau[k] = t; }
    __syncthreads();
  }
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

// ---------------------------------------------------------------------------
// stage 1b: one column step k for every large block.  Launch 1 (symv):
// every workgroup recomputes the reflector of row k (O(n)), then q over its
// column slab and a partial v.q; launch 2 (update): rank-2 update of its slab.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void reflector(const double* rowk, int k, int n, double* red, double& nrm2, double& x0) {
  double part = 0.0;
  for (int j = k + 1 + threadIdx.x; j < n; j += blockDim.x) { double x = rowk[j]; part += x * x; }
  nrm2 = block_sum(part, red);
  x0 = rowk[k + 1];
}

__global__ __launch_bounds__(E2_T) void tridiag_symv_kernel(const E2Block* __restrict__ blks,
                                                             const int* __restrict__ list, int k,
                                                             double* __restrict__ A0, double* __restrict__ WS) {
  __shared__ double red[E2_T / WAVE];
  const E2Block B = blks[list[blockIdx.y]];
  const int n = B.n;
  if (k + 2 >= n) return;
  const int c0 = blockIdx.x * E2_T;
  if (c0 >= n) return;
  double* A = A0 + B.a_off;
  double* d = WS + B.v_off;
  double* q = d + 3 * (size_t)n;        // q[n]
  double* part = q + n;                 // partial sums, one per slab
  double* refl = part + (n + E2_T - 1) / E2_T;  // (t, v0, alpha)
  const double* rowk = A + (size_t)k * n;
  double nrm2, x0;
  reflector(rowk, k, n, red, nrm2, x0);
  if (nrm2 == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) { refl[0] = 0.0; refl[1] = 0.0; refl[2] = 0.0; }
    return;
  }
  const double nrm = sqrt(nrm2), alpha = (x0 > 0.0) ? -nrm : nrm, v0 = x0 - alpha;
  const double t = 2.0 / (nrm2 - x0 * x0 + v0 * v0);
  const int i = c0 + threadIdx.x;
  double qi = 0.0, vi = 0.0;
  if (i > k && i < n) {
    // v_j = A[k][j] for j > k+1, v_{k+1} = v0 (row k is read by every lane: broadcast)
    qi = A[(size_t)(k + 1) * n + i] * v0;
    for (int j = k + 2; j < n; j++) qi += A[(size_t)j * n + i] * rowk[j];
    qi *= t;
    q[i] = qi;
    vi = (i == k + 1) ? v0 : rowk[i];
  }
  const double vp = block_sum(qi * vi, red);
  if (threadIdx.x == 0) part[blockIdx.x] = vp;
  if (blockIdx.x == 0 && threadIdx.x == 0) { refl[0] = t; refl[1] = v0; refl[2] = alpha; }
}

__global__ __launch_bounds__(E2_T) void tridiag_update_kernel(const E2Block* __restrict__ blks,
                                                               const int* __restrict__ list, int k,
                                                               double* __restrict__ A0, double* __restrict__ WS) {
  const E2Block B = blks[list[blockIdx.y]];
  const int n = B.n;
  if (k + 2 >= n) return;
  const int c0 = blockIdx.x * E2_T;
  if (c0 >= n) return;
  double* A = A0 + B.a_off;
  double* d = WS + B.v_off;
  double* e = d + n;
  double* tau = e + n;
  double* q = tau + n;
  const int nslab = (n + E2_T - 1) / E2_T;
  double* part = q + n;
  double* refl = part + nslab;
  const double t = refl[0], v0 = refl[1], alpha = refl[2];
  double* rowk = A + (size_t)k * n;
  if (t == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) { d[k] = rowk[k]; e[k] = rowk[k + 1]; tau[k] = 0.0; }
    return;
  }
  double vp = 0.0;
  for (int s = 0; s < nslab; s++) vp += part[s];
  const double K = 0.5 * t * vp;
  const int i = c0 + threadIdx.x;
  if (i > k && i < n) {
    const double vi = (i == k + 1) ? v0 : rowk[i];
    const double q2 = q[i] - K * vi;
    for (int j = k + 1; j < n; j++) {
      const double vj = (j == k + 1) ? v0 : rowk[j];
      const double qj = q[j] - K * vj;
      const double upd = (j <= i) ? __fma_rn(vj, q2, vi * qj) : __fma_rn(vi, qj, vj * q2);
      A[(size_t)j * n + i] -= upd;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    d[k] = rowk[k]; e[k] = alpha; tau[k] = t;
  }
}

// the reflector head v0 goes into row k only after every slab of step k has
// read the old A[k][k+1]
__global__ void tridiag_store_kernel(const E2Block* __restrict__ blks, const int* __restrict__ list, int k,
                                     double* __restrict__ A0, const double* __restrict__ WS) {
  const E2Block B = blks[list[blockIdx.x]];
  const int n = B.n;
  if (k + 2 >= n || threadIdx.x != 0) return;
  const double* refl = WS + B.v_off + 4 * (size_t)n + (n + E2_T - 1) / E2_T;
  if (refl[0] != 0.0) A0[B.a_off + (size_t)k * n + k + 1] = refl[1];
}

__global__ void tridiag_tail_kernel(const E2Block* __restrict__ blks, const int* __restrict__ list,
                                    const double* __restrict__ A0, double* __restrict__ WS) {
  const E2Block B = blks[list[blockIdx.x]];
  if (threadIdx.x != 0) return;
  const int n = B.n;
  const double* A = A0 + B.a_off;
  double* d = WS + B.v_off;
  double* e = d + n;
  double* tau = e + n;
  if (n >= 2) {
    d[n - 2] = A[(size_t)(n - 2) * n + (n - 2)];
    e[n - 2] = A[(size_t)(n - 2) * n + (n - 1)];
    tau[n - 2] = 0.0;
  }
  d[n - 1] = A[(size_t)(n - 1) * n + (n - 1)];
  e[n - 1] = 0.0;
  tau[n - 1] = 0.0;
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
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double* d = VD(0);
    double* e = VD(1);
    double* erow = VD(16);
    for (int i = 0; i < n; i++) erow[i] = e[i];
    for (int m = mbeg[bi]; m < mbeg[bi + 1]; m++) {
      const int mid = merges[m].mid;
      const double r = e[mid - 1];
      d[mid - 1] -= fabs(r);
      d[mid] -= fabs(r);
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

// merge preparation: z vector, sort of the two halves, deflation, Givens
__global__ __launch_bounds__(E2_T) void dc_prep_kernel(const E2Block* __restrict__ blks,
                                                       const E2Merge* __restrict__ merges, E2MState* __restrict__ ms,
                                                       double* __restrict__ WS, int* __restrict__ IW) {
  __shared__ double red[E2_T / WAVE];
  __shared__ int sh_ngiv;
  const E2Merge M = merges[blockIdx.x];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, mid = M.mid, hi = M.hi, m = hi - lo, n1 = mid - lo, tid = threadIdx.x;
  double* Qv = WS + B.q_off;
  double* dd = VD(5);
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
  if (tid == 0) {
    // merge the two ascending runs of eigenvalues
    int a = 0, b = n1, s = 0;
    while (a < n1 || b < m) {
      int take;
      if (a >= n1) take = b++;
      else if (b >= m) take = a++;
      else take = (dd[lo + b] < dd[lo + a]) ? b++ : a++;
      perm[s] = take;
      ds[s] = dd[lo + take];
      zs[s] = (zn > 0.0) ? z[take] / zn : 0.0;
      s++;
    }
    double dmax = 0.0;
    for (int i = 0; i < m; i++) dmax = fmax(dmax, fabs(ds[i]));
    const double tol = 8.0 * 2.220446049250313e-16 * fmax(dmax, rho);
    int k = 0, nd = 0, ng = 0, pend = -1;
    for (int i = 0; i < m; i++) {
      if (rho * fabs(zs[i]) <= tol) { dval[nd] = ds[i]; drow[nd] = perm[i]; nd++; continue; }
      if (pend < 0) { pend = i; continue; }
      const double zp = zs[pend], zc = zs[i];
      const double tv = hypot(zc, zp);
      const double c = zc / tv, sv = -zp / tv;
      const double gap = ds[i] - ds[pend];
      if (fabs(gap * c * sv) <= tol) {
        ga[ng] = perm[pend]; gb[ng] = perm[i]; gc[ng] = c; gs[ng] = sv; ng++;
        zs[i] = tv; zs[pend] = 0.0;
        const double dp = ds[pend] * c * c + ds[i] * sv * sv;
        ds[i] = ds[pend] * sv * sv + ds[i] * c * c;
        ds[pend] = dp;
        dval[nd] = dp; drow[nd] = perm[pend]; nd++;
        pend = i;
      } else {
        dk[k] = ds[pend]; zk[k] = zs[pend]; krow[k] = perm[pend]; k++;
        pend = i;
      }
    }
    if (pend >= 0) { dk[k] = ds[pend]; zk[k] = zs[pend]; krow[k] = perm[pend]; k++; }
    E2MState st;
    st.k = k; st.nd = nd; st.ngiv = ng; st.pad = 0; st.rho = rho; st.tol = tol;
    ms[blockIdx.x] = st;
    sh_ngiv = ng;
  }
  __syncthreads();
  // Givens rotations of the deflated pairs, in order: q_a' = c q_a + s q_b, q_b' = -s q_a + c q_b
  const int ng = sh_ngiv;
  for (int g = 0; g < ng; g++) {
    double* ra = Qv + (size_t)(lo + ga[g]) * n + lo;
    double* rb = Qv + (size_t)(lo + gb[g]) * n + lo;
    const double c = gc[g], s = gs[g];
    for (int i = tid; i < m; i += blockDim.x) {
      const double xa = ra[i], xb = rb[i];
      ra[i] = c * xa + s * xb;
      rb[i] = -s * xa + c * xb;
    }
    __syncthreads();
  }
}

// secular equation root j: f(lambda) = 1 + rho sum_l zk_l^2/(dk_l - lambda);
// lambda = dk[org] + tau with the nearer pole as origin; delta(j,l) = dk_l - lambda_j
__global__ __launch_bounds__(E2_T) void dc_secular_kernel(const E2Block* __restrict__ blks,
                                                          const E2Merge* __restrict__ merges,
                                                          const E2MState* __restrict__ ms, double* __restrict__ WS,
                                                          int* __restrict__ IW) {
  const E2Merge M = merges[blockIdx.y];
  const E2MState S = ms[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, k = S.k;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const double* dk = VD(9) + lo;
  const double* zk = VD(10) + lo;
  double* lam = VD(11) + lo;
  int* org = VI(5) + lo;
  const int m = M.hi - M.lo;
  double* delta = WS + B.dl_off + (size_t)lo * n;   // k rows of stride m, row j = delta(j, .)
  const double rho = S.rho;
  int o;
  double tlo, thi;
  if (j < k - 1) {
    const double gap = dk[j + 1] - dk[j], half = 0.5 * gap;
    double f = 1.0;
    for (int l = 0; l < k; l++) f += rho * zk[l] * zk[l] / ((dk[l] - dk[j]) - half);
    if (f >= 0.0) { o = j; tlo = 0.0; thi = half; }
    else { o = j + 1; tlo = -(gap - half); thi = 0.0; }
  } else {
    double zz = 0.0;
    for (int l = 0; l < k; l++) zz += zk[l] * zk[l];
    o = j; tlo = 0.0; thi = rho * zz;
  }
  const double dorg = dk[o];
  double tau = 0.5 * (tlo + thi);
  for (int it = 0; it < 120; it++) {
    double f = 1.0, fp = 0.0;
    for (int l = 0; l < k; l++) {
      const double del = (dk[l] - dorg) - tau;
      const double w = zk[l] / del;
      f += rho * zk[l] * w;
      fp += rho * w * w;
    }
    if (f == 0.0) break;
    if (f < 0.0) tlo = tau; else thi = tau;
    double tn = tau - f / fp;
    if (!(tn > tlo && tn < thi)) tn = 0.5 * (tlo + thi);
    if (tn == tau || fabs(thi - tlo) <= 4.0 * 2.220446049250313e-16 * fabs(tau)) { tau = tn; break; }
    tau = tn;
  }
  lam[j] = dorg + tau;
  org[j] = o;
  double* dj = delta + (size_t)j * m;
  for (int l = 0; l < k; l++) dj[l] = (dk[l] - dorg) - tau;
}

// Gu-Eisenstat: zhat_l^2 = prod_j (lambda_j - d_l) / (rho prod_{j!=l} (d_j - d_l))
__global__ __launch_bounds__(E2_T) void dc_zhat_kernel(const E2Block* __restrict__ blks,
                                                       const E2Merge* __restrict__ merges,
                                                       const E2MState* __restrict__ ms, double* __restrict__ WS) {
  const E2Merge M = merges[blockIdx.y];
  const E2MState S = ms[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, k = S.k;
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= k) return;
  const double* dk = VD(9) + lo;
  const double* zk = VD(10) + lo;
  double* zh = VD(12) + lo;
  const int m = M.hi - M.lo;
  const double* delta = WS + B.dl_off + (size_t)lo * n;
  double p = -delta[(size_t)l * m + l] / S.rho;
  for (int j = 0; j < k; j++) {
    if (j == l) continue;
    p *= -delta[(size_t)j * m + l] / (dk[j] - dk[l]);
  }
  const double a = sqrt(fabs(p));
  zh[l] = (zk[l] >= 0.0) ? a : -a;
}

// eigenvector j of the rank-one problem: u_j(l) = zhat_l / (d_l - lambda_j), normalised
__global__ __launch_bounds__(E2_T) void dc_vec_kernel(const E2Block* __restrict__ blks,
                                                      const E2Merge* __restrict__ merges,
                                                      const E2MState* __restrict__ ms, double* __restrict__ WS) {
  const E2Merge M = merges[blockIdx.y];
  const E2MState S = ms[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, k = S.k;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const int m = M.hi - M.lo;
  const double* zh = VD(12) + lo;
  const double* dj = WS + B.dl_off + (size_t)lo * n + (size_t)j * m;
  double* uj = WS + B.u_off + (size_t)lo * n + (size_t)j * m;
  double s = 0.0;
  for (int l = 0; l < k; l++) {
    const double u = zh[l] / dj[l];
    uj[l] = u;
    s += u * u;
  }
  const double inv = 1.0 / sqrt(s);
  for (int l = 0; l < k; l++) uj[l] *= inv;
}

// gather the non-deflated vector rows (components [lo,hi)) into Qk (k x m)
__global__ void dc_gather_kernel(const E2Block* __restrict__ blks, const E2Merge* __restrict__ merges,
                                 const E2MState* __restrict__ ms, double* __restrict__ WS, const int* __restrict__ IW) {
  const E2Merge M = merges[blockIdx.y];
  const E2MState S = ms[blockIdx.y];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, m = M.hi - M.lo, k = S.k;
  const int* krow = VI(1) + lo;
  const double* Qv = WS + B.q_off;
  double* Qk = WS + B.k_off + (size_t)lo * n;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)k * m;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / m), c = (int)(idx % m);
    Qk[idx] = Qv[(size_t)(lo + krow[r]) * n + lo + c];
  }
}

// final order: merge the ascending non-deflated roots (rows of Qnew) with the
// deflated eigenvalues (rows of Qv), write eigenvalues and rows back
__global__ __launch_bounds__(E2_T) void dc_assemble_kernel(const E2Block* __restrict__ blks,
                                                           const E2Merge* __restrict__ merges,
                                                           const E2MState* __restrict__ ms, double* __restrict__ WS,
                                                           int* __restrict__ IW) {
  const E2Merge M = merges[blockIdx.x];
  const E2MState S = ms[blockIdx.x];
  const E2Block B = blks[M.blk];
  const int n = B.n, lo = M.lo, m = M.hi - M.lo, k = S.k, nd = S.nd, tid = threadIdx.x;
  double* Qv = WS + B.q_off;
  double* Qt = WS + B.t_off + (size_t)lo * n;     // rows 0..k-1 hold the new vectors (GEMM output)
  const double* lam = VD(11) + lo;
  const double* dval = VD(13) + lo;
  double* dd = VD(5);
  const int* drow = VI(2) + lo;
  int* src = VI(6) + lo;
  // deflated rows -> Qt rows k..k+nd-1
  for (size_t idx = tid; idx < (size_t)nd * m; idx += blockDim.x) {
    const int r = (int)(idx / m), c = (int)(idx % m);
    Qt[(size_t)(k + r) * m + c] = Qv[(size_t)(lo + drow[r]) * n + lo + c];
  }
  if (tid == 0) {
    // deflated values are ascending up to rotation-level perturbations: insertion sort keeps it exact
    int a = 0, b = 0, o = 0;
    double* dvs = VD(7) + lo;   // reuse ds as scratch for the sorted deflated values
    int* dix = VI(0) + lo;      // reuse perm
    for (int i = 0; i < nd; i++) { dvs[i] = dval[i]; dix[i] = i; }
    for (int i = 1; i < nd; i++) {
      double x = dvs[i]; int xi = dix[i]; int j = i - 1;
      while (j >= 0 && dvs[j] > x) { dvs[j + 1] = dvs[j]; dix[j + 1] = dix[j]; j--; }
      dvs[j + 1] = x; dix[j + 1] = xi;
    }
    while (a < k || b < nd) {
      if (b >= nd || (a < k && lam[a] <= dvs[b])) { dd[lo + o] = lam[a]; src[o] = a; a++; }
      else { dd[lo + o] = dvs[b]; src[o] = k + dix[b]; b++; }
      o++;
    }
  }
  __syncthreads();
  for (size_t idx = tid; idx < (size_t)m * m; idx += blockDim.x) {
    const int o = (int)(idx / m), c = (int)(idx % m);
    Qv[(size_t)(lo + o) * n + lo + c] = Qt[(size_t)src[o] * m + c];
  }
}

// ---------------------------------------------------------------------------
// stage 3: compact-WY panels.  Vt = masked reflector rows, T from the Gram
// matrix G = Vt Vt^T (T_ii = tau_i, T[0:i,i] = -tau_i T[0:i,0:i] G[0:i,i]).
// ---------------------------------------------------------------------------
__global__ void bt_vt_kernel(const E2Block* __restrict__ blks, const int* __restrict__ list, const int* __restrict__ p0,
                             const double* __restrict__ A0, double* __restrict__ WS) {
  const E2Block B = blks[list[blockIdx.y]];
  const int n = B.n, k0 = p0[blockIdx.y];
  if (k0 < 0) return;
  const int nb = min(PANEL, n - 2 - k0);
  const double* A = A0 + B.a_off;
  double* Vt = WS + B.p_off;  // PANEL x n
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)PANEL * n;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int a = (int)(idx / n), j = (int)(idx % n);
    Vt[idx] = (a < nb && j > k0 + a) ? A[(size_t)(k0 + a) * n + j] : 0.0;
  }
}

__global__ __launch_bounds__(64) void bt_tmat_kernel(const E2Block* __restrict__ blks, const int* __restrict__ list,
                                                     const int* __restrict__ p0, double* __restrict__ WS) {
  __shared__ double T[PANEL][PANEL + 1];
  __shared__ double G[PANEL][PANEL + 1];
  const E2Block B = blks[list[blockIdx.x]];
  const int n = B.n, k0 = p0[blockIdx.x];
  if (k0 < 0) return;
  const int nb = min(PANEL, n - 2 - k0);
  const double* tau = WS + B.v_off + 2 * (size_t)n;
  const double* Gg = WS + B.p_off + (size_t)PANEL * n;       // PANEL x PANEL Gram matrix (GEMM output)
  double* Tg = WS + B.p_off + (size_t)PANEL * n + PANEL * PANEL;  // PANEL x PANEL, T^T stored row-major
  const int tid = threadIdx.x;
  for (int i = tid; i < PANEL * PANEL; i += blockDim.x) {
    G[i / PANEL][i % PANEL] = Gg[i];
    T[i / PANEL][i % PANEL] = 0.0;
  }
  __syncthreads();
  for (int i = 0; i < nb; i++) {
    const double ti = tau[k0 + i];
    // w_r = sum_{c<i} T[r][c] G[c][i], r < i
    double w = 0.0;
    if (tid < i)
      for (int c = tid; c < i; c++) w += T[tid][c] * G[c][i];
    __syncthreads();
    if (tid < i) T[tid][i] = -ti * w;
    if (tid == 0) T[i][i] = ti;
    __syncthreads();
  }
  for (int i = tid; i < PANEL * PANEL; i += blockDim.x) {
    const int r = i / PANEL, c = i % PANEL;
    Tg[i] = T[c][r];   // T^T
  }
}

// ---------------------------------------------------------------------------
// batched grouped GEMM, FP64, strided operands:
//   C(m x n) = alpha * sum_k A(i,k) B(k,j) (+ C if accumulate)
//   A(i,k) = A[a_off + i*as0 + k*as1],  B(k,j) = B[b_off + k*bs0 + j*bs1]
// 64x64 tile per 256-thread workgroup, 4x4 per thread; tile loads choose the
// thread mapping that keeps the contiguous dimension along the wavefront.
// dyn >= 0: m and k are read from dynk[dyn] (sizes only known on the device).
// ---------------------------------------------------------------------------
#define GT 64
#define GK 16
__global__ __launch_bounds__(256) void gemm_kernel(const E2Gemm* __restrict__ probs, const E2Tile* __restrict__ tiles,
                                                   const double* __restrict__ Abase, const double* __restrict__ Bbase,
                                                   double* __restrict__ Cbase, const E2MState* __restrict__ dyns) {
  __shared__ double As[GK][GT + 1];
  __shared__ double Bs[GK][GT + 1];
  const E2Tile Tl = tiles[blockIdx.x];
  const E2Gemm P = probs[Tl.prob];
  int M = P.m, K = P.k;
  if (P.dyn >= 0) { M = dyns[P.dyn].k; K = M; }
  const int m0 = Tl.tm * GT, n0 = Tl.tn * GT;
  if (m0 >= M || n0 >= P.n) return;
  const double* A = Abase + P.a_off;
  const double* Bm = Bbase + P.b_off;
  double* C = Cbase + P.c_off;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[r][c] = 0.0;
  const bool a_kcontig = (P.as1 == 1);
  const bool b_ncontig = (P.bs1 == 1);
  for (int k0 = 0; k0 < K; k0 += GK) {
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      int row, kk;
      if (a_kcontig) { row = ty + 16 * rr; kk = tx; }
      else { row = tid & 63; kk = (tid >> 6) + 4 * rr; }
      const int gi = m0 + row, gk = k0 + kk;
      As[kk][row] = (gi < M && gk < K) ? A[(int64_t)gi * P.as0 + (int64_t)gk * P.as1] : 0.0;
      int col;
      if (b_ncontig) { col = tid & 63; kk = (tid >> 6) + 4 * rr; }
      else { col = ty + 16 * rr; kk = tx; }
      const int gj = n0 + col, gk2 = k0 + kk;
      Bs[kk][col] = (gj < P.n && gk2 < K) ? Bm[(int64_t)gk2 * P.bs0 + (int64_t)gj * P.bs1] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk++) {
      double av[4], bv[4];
#pragma unroll
      for (int r = 0; r < 4; r++) av[r] = As[kk][ty * 4 + r];
#pragma unroll
      for (int c = 0; c < 4; c++) bv[c] = Bs[kk][tx * 4 + c];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[r][c] = __fma_rn(av[r], bv[c], acc[r][c]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int gi = m0 + ty * 4 + r;
    if (gi >= M) continue;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int gj = n0 + tx * 4 + c;
      if (gj >= P.n) continue;
      double v = P.alpha * acc[r][c];
      double* cp = C + (int64_t)gi * P.cs0 + gj;
      if (P.accumulate) v += *cp;
      *cp = v;
    }
  }
}

// copy eigenvalues and eigenvector rows to the caller's buffers
__global__ void e2_output_kernel(const E2Block* __restrict__ blks, const double* __restrict__ WS,
                                 double* __restrict__ A0, double* __restrict__ W0) {
  const E2Block B = blks[blockIdx.y];
  const int n = B.n;
  const double* Qv = WS + B.q_off;
  double* A = A0 + B.a_off;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)n * n;
       idx += (size_t)gridDim.x * blockDim.x)
    A[idx] = Qv[idx];
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < n; i += blockDim.x) W0[B.w_off + i] = WS[B.v_off + 5 * (size_t)n + i];
}

}  // namespace e2

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
extern "C" {

hipError_t e2_launch_tridiag_small(const E2Block* b, const int* list, int nlist, double* A, double* WS, hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::tridiag_small_kernel, dim3(nlist), dim3(E2_T), 0, s, b, list, A, WS);
  return hipGetLastError();
}

hipError_t e2_launch_tridiag_step(const E2Block* b, const int* list, int nlist, int nslab, int k, double* A,
                                  double* WS, hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::tridiag_symv_kernel, dim3(nslab, nlist), dim3(E2_T), 0, s, b, list, k, A, WS);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(e2::tridiag_update_kernel, dim3(nslab, nlist), dim3(E2_T), 0, s, b, list, k, A, WS);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(e2::tridiag_store_kernel, dim3(nlist), dim3(64), 0, s, b, list, k, A, WS);
  return hipGetLastError();
}

hipError_t e2_launch_tridiag_tail(const E2Block* b, const int* list, int nlist, const double* A, double* WS,
                                  hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  hipLaunchKernelGGL(e2::tridiag_tail_kernel, dim3(nlist), dim3(64), 0, s, b, list, A, WS);
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
  const int gy = (mmax + E2_T - 1) / E2_T;
  hipLaunchKernelGGL(e2::dc_secular_kernel, dim3(gy, nm), dim3(E2_T), 0, s, b, merges, ms, WS, IW);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(e2::dc_zhat_kernel, dim3(gy, nm), dim3(E2_T), 0, s, b, merges, ms, WS);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(e2::dc_vec_kernel, dim3(gy, nm), dim3(E2_T), 0, s, b, merges, ms, WS);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  int gx = (int)(((size_t)mmax * mmax + 255) / 256); if (gx > 512) gx = 512; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(e2::dc_gather_kernel, dim3(gx, nm), dim3(256), 0, s, b, merges, ms, WS, IW);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (ntiles > 0) {
    hipLaunchKernelGGL(e2::gemm_kernel, dim3(ntiles), dim3(256), 0, s, gprobs, gtiles, WS, WS, WS, ms);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(e2::dc_assemble_kernel, dim3(nm), dim3(E2_T), 0, s, b, merges, ms, WS, IW);
  return hipGetLastError();
}

hipError_t e2_launch_bt_panel(const E2Block* b, const int* list, const int* p0, int nlist, int nmax, const double* A,
                              double* WS, const E2Gemm* gprobs, const E2Tile* tiles_g, int nt_g,
                              const E2Tile* tiles_rest, int nt_w, int nt_w2, int nt_z, hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  hipError_t e;
  int gx = (int)(((size_t)PANEL * nmax + 255) / 256); if (gx > 256) gx = 256; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(e2::bt_vt_kernel, dim3(gx, nlist), dim3(256), 0, s, b, list, p0, A, WS);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (nt_g > 0) hipLaunchKernelGGL(e2::gemm_kernel, dim3(nt_g), dim3(256), 0, s, gprobs, tiles_g, WS, WS, WS,
                                   (const E2MState*)nullptr);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(e2::bt_tmat_kernel, dim3(nlist), dim3(64), 0, s, b, list, p0, WS);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const E2Tile* t = tiles_rest;
  if (nt_w > 0) hipLaunchKernelGGL(e2::gemm_kernel, dim3(nt_w), dim3(256), 0, s, gprobs, t, WS, WS, WS,
                                   (const E2MState*)nullptr);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  t += nt_w;
  if (nt_w2 > 0) hipLaunchKernelGGL(e2::gemm_kernel, dim3(nt_w2), dim3(256), 0, s, gprobs, t, WS, WS, WS,
                                    (const E2MState*)nullptr);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  t += nt_w2;
  if (nt_z > 0) hipLaunchKernelGGL(e2::gemm_kernel, dim3(nt_z), dim3(256), 0, s, gprobs, t, WS, WS, WS,
                                   (const E2MState*)nullptr);
  return hipGetLastError();
}

hipError_t e2_launch_output(const E2Block* b, int nb, int nmax, const double* WS, double* A, double* W,
                            hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  int gx = (int)(((size_t)nmax * nmax + 255) / 256); if (gx > 1024) gx = 1024; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(e2::e2_output_kernel, dim3(gx, nb), dim3(256), 0, s, b, WS, A, W);
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
