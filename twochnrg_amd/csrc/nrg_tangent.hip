// nrg_tangent.hip -- kernels of the derivative (tangent) pass of the
// fixed-truncation NRG iteration (no counterpart in the reference).
//
// For one symmetric block with eigenvector rows Z and ascending E, a
// perturbation dH of H_N moves the eigenpairs to first order as
//   G = Z dH Z^T,  dE_i = G_ii,  dz_i = sum_{j != i} G_ij / (E_i - E_j) z_j.
// The kept rows are only moved towards the discarded ones (rotations inside
// the kept space change nothing downstream), so the next step's kept-space
// Hamiltonian tangent is G on the kept rows and columns and the kept rows
// move by dZ_k = C Z_d with C = G_kd / (E_k - E_d).  The products Z dH,
// (Z dH) Z^T and C Z_d run as grouped FP64 MFMA GEMMs; these kernels build
// dH, take dE and split G.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrg_kernels.h"
#include "nrg_wave.h"

#define WAVE 64

namespace nrgt {

// dH element (i, j) of block b along tangent t:
//   sqrt(Lambda) th(parent_i, parent_j)        equal site types, same old block
// + C[S][t_p][t] (chi df + dchi f)(a, b)       the coupling rule of H_N, in the
//                                              orientation f stores the block pair
__global__ void tan_hn_kernel(const TanHnArgs a) {
  const int b = blockIdx.y, t = blockIdx.z;
  const int n = a.bsize[b];
  const int b0 = a.bbeg[b];
  double* H = a.dH + t * a.dh_stride + a.h_off[b];
  const double* th = a.th + t * a.th_stride;
  const double* df = a.df + t * a.df_stride;
  const double dchi = a.dchi[t];
  const int sidx = a.blk_sidx[b];
  const OldOp& F = a.f;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)n * n;
       idx += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(idx / n), j = (int)(idx % n);
    if (i > j) { int x = i; i = j; j = x; }
    const int ri = a.stcf[b0 + i], rj = a.stcf[b0 + j];
    const int ti = a.itype[b0 + i], tj = a.itype[b0 + j];
    const int bi = F.ob[ri], bj = F.ob[rj];
    double v = 0.0;
    if (ti == tj && bi == bj) v = a.sqrtL * th[a.th_off[bi] + (int64_t)F.op[ri] * F.bsz[bi] + F.op[rj]];
    if (i != j) {
      int64_t off = F.mtab[(int64_t)bi * F.nbo + bj];
      int pi = F.op[ri], pj = F.op[rj], bc = bj, tp = ti, tt = tj;
      if (off < 0) {
        off = F.mtab[(int64_t)bj * F.nbo + bi];
        pi = F.op[rj]; pj = F.op[ri]; bc = bi; tp = tj; tt = ti;
      }
      if (off >= 0) {
        const double c = a.coef[((int64_t)sidx * a.ntypes + tp) * a.ntypes + tt];
        const int64_t e = off + (int64_t)pi * F.bsz[bc] + pj;
        v += c * (a.chi * df[e] + dchi * F.mat[e]);
      }
    }
    H[idx] = v;
  }
}

// dE (before the SetE0zero shift) of every state: kept rows read the diagonal
// of G, the other rows the dot product of their rows of Y = Z dH and Z.
// One wavefront per row.
__global__ __launch_bounds__(256) void tan_diag_kernel(const TanBlk* __restrict__ blk, const double* __restrict__ Z,
                                                       const double* __restrict__ Y, int64_t y_stride,
                                                       const double* __restrict__ G, int64_t g_stride,
                                                       double* __restrict__ dE, int64_t e_stride) {
  const TanBlk B = blk[blockIdx.y];
  const int t = blockIdx.z;
  const int lane = threadIdx.x & (WAVE - 1);
  const int row = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (row >= B.n) return;
  double v;
  if (row < B.k) {
    v = G[t * g_stride + B.goff + (int64_t)row * B.n + row];
  } else {
    const double* y = Y + t * y_stride + B.zoff + (int64_t)row * B.n;
    const double* z = Z + B.zoff + (int64_t)row * B.n;
    v = 0.0;
    for (int j = lane; j < B.n; j += WAVE) v += y[j] * z[j];
    v = nrgw::wave_sum(v);
  }
  if (lane == 0) dE[t * e_stride + B.e0 + row] = v;
}

// th' = G_kk - dEmin I and C = G_kd / (E_k - E_d) of every kept block;
// dEmin = dE of the ground state gs
__global__ void tan_cut_kernel(const TanCut* __restrict__ cut, const double* __restrict__ G, int64_t g_stride,
                               const double* __restrict__ E, const double* __restrict__ dE, int64_t e_stride,
                               int64_t gs, double* __restrict__ th, int64_t th_stride, double* __restrict__ C,
                               int64_t c_stride) {
  const TanCut B = cut[blockIdx.y];
  const int t = blockIdx.z;
  const int n = B.n, k = B.k;
  const double s = dE[t * e_stride + gs];
  const double* g = G + t * g_stride + B.goff;
  const double* e = E + B.e0;
  double* h = th + t * th_stride + B.thoff;
  double* c = C + t * c_stride + B.coff;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)k * n;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / n), j = (int)(idx % n);
    const double v = g[idx];
    if (j < k) {
      h[(int64_t)i * k + j] = v - (i == j ? s : 0.0);
    } else {
      const double d = e[i] - e[j];
      c[(int64_t)i * (n - k) + (j - k)] = d != 0.0 ? v / d : 0.0;
    }
  }
}

}  // namespace nrgt

extern "C" {

hipError_t nrgk_launch_tan_hn(const TanHnArgs* a, int nb, int nmax, int ntan, hipStream_t s) {
  if (nb <= 0 || ntan <= 0) return hipSuccess;
  size_t tot = (size_t)nmax * nmax;
  int gx = (int)((tot + 255) / 256); if (gx > 2048) gx = 2048; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(nrgt::tan_hn_kernel, dim3(gx, nb, ntan), dim3(256), 0, s, *a);
  return hipGetLastError();
}

hipError_t nrgk_launch_tan_diag(const TanBlk* d_blk, int nb, int nmax, int ntan, const double* Z, const double* Y,
                                int64_t y_stride, const double* G, int64_t g_stride, double* dE, int64_t e_stride,
                                hipStream_t s) {
  if (nb <= 0 || ntan <= 0) return hipSuccess;
  int gx = (nmax + 3) / 4; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(nrgt::tan_diag_kernel, dim3(gx, nb, ntan), dim3(256), 0, s, d_blk, Z, Y, y_stride, G, g_stride,
                     dE, e_stride);
  return hipGetLastError();
}

hipError_t nrgk_launch_tan_cut(const TanCut* d_cut, int nbc, int64_t maxel, int ntan, const double* G,
                               int64_t g_stride, const double* E, const double* dE, int64_t e_stride, int64_t gs,
                               double* th, int64_t th_stride, double* C, int64_t c_stride, hipStream_t s) {
  if (nbc <= 0 || ntan <= 0) return hipSuccess;
  int gx = (int)((maxel + 255) / 256); if (gx > 2048) gx = 2048; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(nrgt::tan_cut_kernel, dim3(gx, nbc, ntan), dim3(256), 0, s, d_cut, G, g_stride, E, dE,
                     e_stride, gs, th, th_stride, C, c_stride);
  return hipGetLastError();
}

}  // extern "C"
