// nrg_kernels.hip -- HIP kernels for gfx950 (MI355X) of the NRG
// iterative-diagonalisation step:
//   * eig_block_kernel   : batched FP64 symmetric eigensolver over uneven
//                          symmetry blocks (one workgroup per block, LPT order)
//   * hn_assemble_kernel : H_N block assembly from the previous step's reduced
//                          f-operator and per-(S, site type) coupling tables
//   * fbasis_kernel      : operator matrix in the product basis (F of
//                          UpdateMatrices) for the rotation GEMMs
//   * unpack/compact helpers
//
// Layouts follow the reference's storage contract (include/NRGclasses.hpp):
// blocks row-major, eigenvector i = row i (src/NRGmatrixClass.cpp:647-654),
// packed upper row-wise = ij2rUpTr (src/ij2r.cpp:63-68).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrg_kernels.h"
#include "nrg_wave.h"

#define EIG_THREADS 256
#define WAVE 64

namespace nrgk {

__device__ __forceinline__ double wave_sum(double v) { return nrgw::wave_sum(v); }

// block-wide sum over EIG_THREADS threads; result broadcast to all threads
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < EIG_THREADS / WAVE; i++) s += red[i];
  return s;
}

// ---------------------------------------------------------------------------
// Batched symmetric eigensolver.
//   A (n x n, row-major, symmetric) is reduced in place to tridiagonal form by
//   Householder reflectors (reflector k stored in row k, columns k+1..n-1),
//   the tridiagonal eigenproblem is solved by implicit QL with Wilkinson
//   shifts -- thread 0 runs the scalar bulge chase, the whole workgroup applies
//   each sweep's Givens rotations to Zt (column k owned by one thread, so
//   accesses are coalesced) -- then the reflectors are applied back to the
//   eigenvectors and the pairs are written in ascending order, eigenvector i
//   as row i of A.  Replaces dspev_/dsyev_ of DiagBlock
//   (src/NRGmatrixClass.cpp:566-696).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(EIG_THREADS) void eig_block_kernel(const EigDesc* __restrict__ desc,
                                                                 const int* __restrict__ order,
                                                                 double* __restrict__ Abuf,
                                                                 double* __restrict__ Wbuf,
                                                                 double* __restrict__ WS,
                                                                 int* __restrict__ info, int lds_nmax) {
  extern __shared__ double lds[];
  __shared__ double red[EIG_THREADS / WAVE];
  __shared__ int sh_lo, sh_hi, sh_done;
  const int bi = order[blockIdx.x];
  const EigDesc D = desc[bi];
  const int n = D.n;
  const int tid = threadIdx.x;
  if (n <= 0) return;
  double* A = Abuf + D.a_off;
  double* W = Wbuf + D.w_off;
  double* Zt = WS + D.ws_off;           // n*n
  double* tau = WS + D.vec_off;         // n
  double* q = tau + n;                  // n
  double* d;
  double* e;
  double* cs;
  double* sn;
  if (n <= lds_nmax) {
    d = lds; e = d + n; cs = e + n; sn = cs + n;
  } else {
    d = q + n; e = d + n; cs = e + n; sn = cs + n;   // global fallback (vec area = 7n)
  }
  if (n == 1) {
    if (tid == 0) { W[0] = A[0]; A[0] = 1.0; if (info) info[bi] = 0; }
    return;
  }
  // ---------------- Householder tridiagonalisation ----------------
  for (int k = 0; k + 2 < n; k++) {
    double* rowk = A + (size_t)k * n;
    double part = 0.0;
    for (int j = k + 1 + tid; j < n; j += EIG_THREADS) { double x = rowk[j]; part += x * x; }
    const double nrm2 = block_sum(part, red);
    const double x0 = rowk[k + 1];
    if (nrm2 == 0.0) {
      if (tid == 0) { d[k] = rowk[k]; e[k] = 0.0; tau[k] = 0.0; }
      __syncthreads();
      continue;
    }
    const double nrm = sqrt(nrm2);
    const double alpha = (x0 > 0.0) ? -nrm : nrm;
    const double v0 = x0 - alpha;
    const double vtv = nrm2 - x0 * x0 + v0 * v0;
    const double t = 2.0 / vtv;
    __syncthreads();
    if (tid == 0) { rowk[k + 1] = v0; d[k] = rowk[k]; e[k] = alpha; tau[k] = t; }
    __syncthreads();
    // q_i = t * sum_j A[i][j] v_j  (one wave per row, lanes along the row)
    const int lane = tid & (WAVE - 1), wid = tid / WAVE;
    double vp_part = 0.0;
    for (int i = k + 1 + wid; i < n; i += EIG_THREADS / WAVE) {
      const double* ri = A + (size_t)i * n;
      double s = 0.0;
      for (int j = k + 1 + lane; j < n; j += WAVE) s += ri[j] * rowk[j];
      s = wave_sum(s) * t;
      if (lane == 0) q[i] = s;
      vp_part += (lane == 0) ? s * rowk[i] : 0.0;
    }
    const double vp = block_sum(vp_part, red);
    const double K = 0.5 * t * vp;
    for (int i = k + 1 + tid; i < n; i += EIG_THREADS) q[i] -= K * rowk[i];
    __syncthreads();
    // A[i][j] -= v_i q_j + v_j q_i on the trailing block; the expression is
    // evaluated with (min,max) ordering so both triangles stay bit-identical
    const int m = n - k - 1;
    for (int i = k + 1 + wid; i < n; i += EIG_THREADS / WAVE) {
      double* ri = A + (size_t)i * n;
      const double vi = rowk[i], qi = q[i];
      for (int j = k + 1 + lane; j < n; j += WAVE) {
        const double vj = rowk[j], qj = q[j];
        double upd = (i <= j) ? __fma_rn(vi, qj, vj * qi) : __fma_rn(vj, qi, vi * qj);
        ri[j] -= upd;
      }
    }
    (void)m;
    __syncthreads();
  }
  if (tid == 0) {
    d[n - 2] = A[(size_t)(n - 2) * n + (n - 2)];
    e[n - 2] = A[(size_t)(n - 2) * n + (n - 1)];
    d[n - 1] = A[(size_t)(n - 1) * n + (n - 1)];
    e[n - 1] = 0.0;
  }
  // Zt = I
  for (size_t idx = tid; idx < (size_t)n * n; idx += EIG_THREADS) {
    const int r = (int)(idx / n), c = (int)(idx % n);
    Zt[idx] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  // ---------------- implicit QL ----------------
  int l = 0, iter = 0, bad = 0;
  for (;;) {
    if (tid == 0) {
      int done = 0, lo = 0, hi = -1;
      for (;;) {
        if (l >= n) { done = 1; break; }
        int m;
        for (m = l; m + 1 < n; m++) {
          const double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= 1e-300 || fabs(e[m]) + dd == dd) break;
        }
        if (m == l || iter >= 60) { if (m != l) bad = 1; l++; iter = 0; continue; }
        iter++;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        int under = 0;
        for (i = m - 1; i >= l; i--) {
          const double f = s * e[i], b = c * e[i];
          r = hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; under = 1; break; }
          s = f / r; c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          cs[i] = c; sn[i] = s;
        }
        if (!under) { d[l] -= p; e[l] = g; e[m] = 0.0; lo = l; }
        else lo = i + 1;
        hi = m;
        break;
      }
      sh_done = done; sh_lo = lo; sh_hi = hi;
    }
    __syncthreads();
    if (sh_done) break;
    const int lo = sh_lo, hi = sh_hi;
    if (hi > lo) {
      for (int kcol = tid; kcol < n; kcol += EIG_THREADS) {
        double carry = Zt[(size_t)hi * n + kcol];
        for (int i = hi - 1; i >= lo; i--) {
          const double zi = Zt[(size_t)i * n + kcol];
          const double c = cs[i], s = sn[i];
          Zt[(size_t)(i + 1) * n + kcol] = s * zi + c * carry;
          carry = c * zi - s * carry;
        }
        Zt[(size_t)lo * n + kcol] = carry;
      }
    }
    __syncthreads();
  }
  // ---------------- back-transform: x_j <- H_0 ... H_{n-3} x_j ----------------
  {
    const int lane = tid & (WAVE - 1), wid = tid / WAVE;
    for (int k = n - 3; k >= 0; k--) {
      const double tk = tau[k];
      if (tk == 0.0) continue;
      const double* vk = A + (size_t)k * n;
      for (int j = wid; j < n; j += EIG_THREADS / WAVE) {
        double* xj = Zt + (size_t)j * n;
        double s = 0.0;
        for (int i = k + 1 + lane; i < n; i += WAVE) s += vk[i] * xj[i];
        s = wave_sum(s) * tk;
        for (int i = k + 1 + lane; i < n; i += WAVE) xj[i] -= s * vk[i];
      }
      __syncthreads();
    }
  }
  // ---------------- ascending order, row i = eigenvector i ----------------
  for (int j = tid; j < n; j += EIG_THREADS) {
    const double dj = d[j];
    int rank = 0;
    for (int i = 0; i < n; i++) {
      const double di = d[i];
      rank += (di < dj) || (di == dj && i < j);
    }
    q[j] = (double)rank;
    W[rank] = dj;
  }
  __syncthreads();
  for (int j = 0; j < n; j++) {
    const int rank = (int)q[j];
    const double* src = Zt + (size_t)j * n;
    double* dst = A + (size_t)rank * n;
    for (int i = tid; i < n; i += EIG_THREADS) dst[i] = src[i];
  }
  if (tid == 0 && info) {
    // bad is only tracked by thread 0
    info[bi] = bad;
  }
}

// packed row-wise upper (ij2rUpTr) -> full symmetric row-major
__global__ void unpack_upper_kernel(const EigDesc* __restrict__ desc, const int64_t* __restrict__ p_off,
                                    const double* __restrict__ P, double* __restrict__ A) {
  const EigDesc D = desc[blockIdx.y];
  const int n = D.n;
  const double* src = P + p_off[blockIdx.y];
  double* dst = A + D.a_off;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)n * n;
       idx += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(idx / n), j = (int)(idx % n);
    if (i > j) { int t = i; i = j; j = t; }
    dst[idx] = src[(size_t)i * n - (size_t)i * (i - 1) / 2 + (j - i)];
  }
}

// ---------------------------------------------------------------------------
// Old-operator element lookup: state r of the previous kept basis lives in
// block ob[r] at position op[r]; mtab[ob_i*nbo+ob_j] is the offset of that
// operator block (row-major, bsz[ob_j] columns) or -1.  GetMatEl of
// src/NRGmatrixClass.cpp:160-170 with O(1) lookups instead of linear search.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double old_el(const OldOp& O, int r, int rp) {
  const int bi = O.ob[r], bj = O.ob[rp];
  const int64_t off = O.mtab[(int64_t)bi * O.nbo + bj];
  if (off < 0) return 0.0;
  return O.mat[off + (int64_t)O.op[r] * O.bsz[bj] + O.op[rp]];
}

// ---------------------------------------------------------------------------
// H_N assembly (DiagHN src/DiagHN.cpp:6-119 with the OneChQS/TwoChQS
// element rules).  Element (i,j), i<j, of block b:
//   sum_ch chi_ch * <stcf_i|f_ch|stcf_j> * C_ch[S_b][t_i][t_j]
// where the element is taken from the (stcf_j, stcf_i) entry with swapped
// site types when the (i,j) one is zero (src/OpMatRules.cpp:1477-1488); the
// coupling table C already carries the fermion sign.  Diagonal:
//   sqrt(Lambda) * E_parent + eN * (Q_site + 1).
// Both triangles are written (full storage for the eigensolver).
// ---------------------------------------------------------------------------
__global__ void hn_assemble_kernel(const HnArgs a) {
  const int b = blockIdx.y;
  const int n = a.bsize[b];
  const int b0 = a.bbeg[b];
  double* H = a.H + a.h_off[b];
  const int sidx = a.blk_sidx[b];
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)n * n;
       idx += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(idx / n), j = (int)(idx % n);
    if (i > j) { int t = i; i = j; j = t; }
    const int ist = b0 + i, jst = b0 + j;
    double val;
    if (i == j) {
      val = a.sqrtL * a.en_par[ist] + a.eN * (a.site_q[a.itype[ist]] + 1.0);
    } else {
      val = 0.0;
      const int ri = a.stcf[ist], rj = a.stcf[jst];
      const int ti = a.itype[ist], tj = a.itype[jst];
      for (int ch = 0; ch < a.nch; ch++) {
        double old = old_el(a.f[ch], ri, rj);
        int tp = ti, tt = tj;
        if (fabs(old) < 1e-20) { old = old_el(a.f[ch], rj, ri); tp = tj; tt = ti; }
        const double c = a.coef[ch][((int64_t)sidx * a.ntypes + tp) * a.ntypes + tt];
        val += a.chi[ch] * old * c;
      }
    }
    H[idx] = val;
  }
}

// ---------------------------------------------------------------------------
// F of UpdateMatrices (src/UpdateMatrices.cpp:390-430) for one (operator,
// block pair): F[ist][jst] = G[t_i][t_j] * (need_old ? O_old(stcf_i,stcf_j)
//                                                   : delta(stcf_i,stcf_j)).
// G is the pair's site-type table (rule value for the pair's (S_i,S_j)).
// ---------------------------------------------------------------------------
__global__ void fbasis_kernel(const FArgs a) {
  const FPair P = a.pairs[blockIdx.y];
  const int ni = P.ni_bc, nj = P.nj_bc;
  double* F = a.F + P.f_off;
  const double* G = a.G + P.g_off;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)ni * nj;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / nj), j = (int)(idx % nj);
    const int ist = P.ibeg_bc + i, jst = P.jbeg_bc + j;
    const double g = G[a.itype[ist] * a.ntypes + a.itype[jst]];
    double v = 0.0;
    if (g != 0.0) {
      if (P.need_old) v = g * old_el(a.old[P.op], a.stcf[ist], a.stcf[jst]);
      else v = (a.stcf[ist] == a.stcf[jst]) ? g : 0.0;
    }
    F[idx] = v;
  }
}

// copy the first k rows (length n) of each block -- CutStates' eigenvector
// compaction (src/NRGarrayClass.cpp:1467-1548): kept states are the lowest
// rows of each energy-ordered block.
__global__ void compact_rows_kernel(const CompactDesc* __restrict__ cd, const double* __restrict__ src,
                                    double* __restrict__ dst) {
  const CompactDesc C = cd[blockIdx.y];
  const size_t tot = (size_t)C.k * C.n;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < tot;
       idx += (size_t)gridDim.x * blockDim.x)
    dst[C.dst_off + idx] = src[C.src_off + idx];
}

}  // namespace nrgk

// ---------------------------------------------------------------------------
// launch wrappers (host)
// ---------------------------------------------------------------------------
// diagonals of operator blocks (for averages over the kept states)
__global__ void diag_gather_kernel(const DiagJob* __restrict__ jobs, double* __restrict__ out) {
  const DiagJob J = jobs[blockIdx.x];
  for (int i = threadIdx.x; i < J.n; i += blockDim.x) out[J.dst + i] = J.src[(size_t)i * J.n + i];
}

extern "C" {

hipError_t nrgk_launch_diag_gather(const DiagJob* d_jobs, int njobs, double* out, hipStream_t s) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(diag_gather_kernel, dim3(njobs), dim3(256), 0, s, d_jobs, out);
  return hipGetLastError();
}

hipError_t nrgk_launch_eig(const EigDesc* d_desc, const int* d_order, int nb, int nmax, double* A,
                           double* W, double* WS, int* d_info, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  int lds_nmax = 1536;
  int nl = nmax < lds_nmax ? nmax : lds_nmax;
  size_t lds = (size_t)4 * nl * sizeof(double);
  hipLaunchKernelGGL(nrgk::eig_block_kernel, dim3(nb), dim3(EIG_THREADS), lds, s, d_desc, d_order, A, W, WS,
                     d_info, nl);
  return hipGetLastError();
}

hipError_t nrgk_launch_unpack(const EigDesc* d_desc, const int64_t* d_poff, const double* P, double* A, int nb,
                              int nmax, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  size_t tot = (size_t)nmax * nmax;
  int gx = (int)((tot + 255) / 256); if (gx > 1024) gx = 1024; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(nrgk::unpack_upper_kernel, dim3(gx, nb), dim3(256), 0, s, d_desc, d_poff, P, A);
  return hipGetLastError();
}

hipError_t nrgk_launch_hn(const HnArgs* a, int nb, int nmax, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  size_t tot = (size_t)nmax * nmax;
  int gx = (int)((tot + 255) / 256); if (gx > 2048) gx = 2048; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(nrgk::hn_assemble_kernel, dim3(gx, nb), dim3(256), 0, s, *a);
  return hipGetLastError();
}

hipError_t nrgk_launch_fbasis(const FArgs* a, int npairs, int64_t maxel, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  int gx = (int)((maxel + 255) / 256); if (gx > 2048) gx = 2048; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(nrgk::fbasis_kernel, dim3(gx, npairs), dim3(256), 0, s, *a);
  return hipGetLastError();
}


hipError_t nrgk_launch_compact(const CompactDesc* d_cd, int nb, int64_t maxel, const double* src, double* dst,
                               hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  int gx = (int)((maxel + 255) / 256); if (gx > 1024) gx = 1024; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(nrgk::compact_rows_kernel, dim3(gx, nb), dim3(256), 0, s, d_cd, src, dst);
  return hipGetLastError();
}

}  // extern "C"
