// nrg_dm.hip -- DM-NRG kernels: the child trace-out of the rotated density
// matrix (src/DM_NRG/DM_NRG_CalcRhoN.cpp:252-449, FDM diagonal :393-398), the
// pole weights of the shell spectral sums (DM-NRG src/SpecFunc_allSTL.cpp:223-692,
// CFS :702-1240, FDM :1253-2354, Costi src/SpecFuncClass.cpp:417-662) and their
// broadening (src/BroadDelta.cpp:12-132, MijxBDeltaEij src/SpecFunc_allSTL.cpp:36-219).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "nrg_dm.h"
#include "nrg_wave.h"
#include "../../include/twochnrg_amd.h"

namespace dmk {

// rho_N block = sum over its child segments of cg * (segment of the rotated
// rho_{N+1}) on the kept prefix (the states with children), zero elsewhere; FDM:
// e^{-bb E_p} / zdisc on the diagonal of the discarded states p >= kd
__global__ void gather_kernel(const DmGBlk* __restrict__ blks, const DmSeg* __restrict__ segs,
                              const double* __restrict__ rot, const double* __restrict__ en, double bb, double zdisc,
                              double* __restrict__ out) {
  const DmGBlk B = blks[blockIdx.y];
  if (B.compact) {
    const int64_t kk = (int64_t)B.kc * B.kc, tot = kk + (B.k - B.kc);
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
      if (idx < kk) {
        const int p = (int)(idx / B.kc), q = (int)(idx % B.kc);
        double s = 0.0;
        for (int g = B.seg_beg; g < B.seg_end; g++) {
          const DmSeg S = segs[g];
          s += S.cg * rot[S.rot_off + (int64_t)p * S.ld + q];
        }
        out[B.out_off + idx] = s;
      } else {
        const int p = B.kc + (int)(idx - kk);
        out[B.d_off + (p - B.kc)] = (p >= B.kd) ? exp(-bb * en[B.e_off + p]) / zdisc : 0.0;
      }
    }
    return;
  }
  const int64_t kk = (int64_t)B.k * B.k;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < kk; idx += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(idx / B.k), q = (int)(idx % B.k);
    double s = 0.0;
    if (p < B.kc && q < B.kc)
      for (int g = B.seg_beg; g < B.seg_end; g++) {
        const DmSeg S = segs[g];
        s += S.cg * rot[S.rot_off + (int64_t)p * S.ld + q];
      }
    if (p == q && p >= B.kd) s = exp(-bb * en[B.e_off + p]) / zdisc;
    out[B.out_off + idx] = s;
  }
}

// a matrix times the discarded diagonal of a compact rho, from the left (rows)
// or from the right (columns)
__global__ void scale_kernel(const DmScale* __restrict__ blks, const double* __restrict__ d,
                             const double* __restrict__ a, double* __restrict__ out) {
  const DmScale B = blks[blockIdx.y];
  const int w = B.n1 - B.n0;
  const int64_t tot = (int64_t)(B.m1 - B.m0) * w;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = B.m0 + (int)(idx / w), j = B.n0 + (int)(idx % w);
    const int64_t e = (int64_t)i * B.ld + j;
    out[B.o_off + e] = a[B.a_off + e] * d[B.d_off + (B.mode ? j - B.n0 : i - B.m0)];
  }
}

// pole weights (omega >= 0: Wp, omega < 0: Wn) and energies of one operator
// block pair, with w1 = (rho1 A)_ij A_ij (Termo1) and w2 = A_ij (A rho2)_ij (Termo2):
//   DM-NRG   Wp = Wn = cg (w1 + w2)
//   Costi    Wp = Wn = cg A_ij^2 (e^{-bb E1_i} + e^{-bb E2_j}) / Z
//   CFS      (i kept, j disc): Wp = cg w1;  (i disc, j kept): Wn = cg w2
//   FDM      (i kept, j disc): Wp = cg (w1 + [block 1 has discarded] w2)
//            (i disc, j kept): Wn = cg (w2 + [block 2 has discarded] w1)
//            (i disc, j disc): Wp = Wn = cg (w1 + w2)
//            the ONESIDED bit drops the [block has discarded] conditions; the
//            NORM bit gives every pole its weight at both signs (Wp = Wn)
//   dE = E2_j - E1_i
__global__ void weights_kernel(const DmPair* __restrict__ pairs, const double* __restrict__ ops,
                               const double* __restrict__ prod, const double* __restrict__ en,
                               double* __restrict__ Wp, double* __restrict__ Wn, double* __restrict__ dE) {
  const DmPair P = pairs[blockIdx.y];
  const int64_t n = (int64_t)P.k1 * P.k2;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / P.k2), j = (int)(idx % P.k2);
    const double a = ops[P.a_off + idx];
    const double e1 = en[P.e1_off + i], e2 = en[P.e2_off + j];
    double wp = 0.0, wn = 0.0;
    if (P.form == DM_FORM_COSTI) {
      wp = wn = P.cg * a * a * (exp(-P.bb * e1) + exp(-P.bb * e2)) * P.invZ;
    } else {
      const double w1 = prod[P.p1_off + idx] * a, w2 = a * prod[P.p2_off + idx];
      if (P.form == DM_FORM_DMNRG) {
        wp = wn = P.cg * (w1 + w2);
      } else {
        const bool ik = i < P.kc1, jk = j < P.kc2;
        if (P.form == DM_FORM_CFS) {
          if (ik && !jk) wp = P.cg * w1;
          else if (!ik && jk) wn = P.cg * w2;
        } else {
          const bool ons = (P.fix & NRG_FDM_FIX_ONESIDED) != 0;
          const bool d1 = ons || P.kc1 < P.k1, d2 = ons || P.kc2 < P.k2;
          if (ik && !jk) wp = P.cg * (w1 + (d1 ? w2 : 0.0));
          else if (!ik && jk) wn = P.cg * (w2 + (d2 ? w1 : 0.0));
          else if (!ik && !jk) wp = wn = P.cg * (w1 + w2);
          if (P.fix & NRG_FDM_FIX_NORM) wp = wn = (ik && !jk) ? wp : wn;
        }
      }
    }
    Wp[P.pole_off + idx] = wp;
    Wn[P.pole_off + idx] = wn;
    dE[P.pole_off + idx] = e2 - e1;
  }
}

// BDelta(w, x, dBroad), or BDeltaTemp(w, x, dBroadTemp) inside the window:
//   DM_BD_BROAD_LORENTZ   BroadDelta / LorentzDelta        (DM-NRG)
//   DM_BD_BROAD_ANDERS    BroadDelta / LorentzDeltaAnders  (CFS)
//   DM_BD_LOGGAUSS_GAUSS  LogGaussDelta / GaussDelta       (FDM)
__device__ __forceinline__ double bdelta(int bd, double w, double x, double twin, double bbroad, double btemp) {
  if (fabs(x) < twin) {
    const double d = w - x;
    if (bd == DM_BD_LOGGAUSS_GAUSS) {
      const double t = d / btemp;
      return (1.0 / (btemp * sqrt(M_PI))) * exp(-t * t);
    }
    const double l = btemp / (d * d + btemp * btemp);
    return (bd == DM_BD_BROAD_ANDERS) ? (1.0 / M_PI) * l : (1.0 / (2.0 * M_PI)) * l;
  }
  if (x == 0.0) x = 1.0e-100;
  if (w * x < 0.0) return 0.0;   // log-Gaussians: same-sign poles only
  const double a1 = log(fabs(w / x)) / bbroad;
  if (bd == DM_BD_LOGGAUSS_GAUSS) {
    const double a0 = 1.0 / (bbroad * fabs(x) * sqrt(M_PI));
    const double t = a1 - bbroad / 4.0;
    return a0 * exp(-t * t);
  }
  const double a0 = exp(-bbroad * bbroad / 4.0) / (bbroad * fabs(x) * sqrt(M_PI));
  return a0 * exp(-a1 * a1);
}

#define BT 256
#define BCH (BT * 8)   // poles per workgroup chunk

// the shell sums of a tile of omegas over one slab of the shell's poles: each
// pole is read once and broadened at every omega of the tile (the omegas of
// one shell share its poles, so the pole traffic falls by the tile width).
// A pole on the other side of zero than every omega of the tile, outside the
// BDeltaTemp window, contributes the exact 0 of bdelta's sign test and is
// skipped.
__global__ __launch_bounds__(BT) void broaden_tiles_kernel(const DmTile* __restrict__ tiles,
                                                           const double* __restrict__ wlist,
                                                           const double* __restrict__ Wp,
                                                           const double* __restrict__ Wn,
                                                           const double* __restrict__ dE, int bd, double bbroad,
                                                           double btemp, double* __restrict__ partial) {
  __shared__ double red[BT / 64][DM_TILE_W];
  const DmTile T = tiles[blockIdx.y];
  const int slab = blockIdx.x;
  if (slab >= T.nslab) return;
  const int64_t b0 = T.pole_beg + (int64_t)slab * T.slab_len;
  const int64_t b1 = min(b0 + T.slab_len, T.pole_end);
  const double* __restrict__ W = T.wsel ? Wn : Wp;
  double w[DM_TILE_W], acc[DM_TILE_W];
#pragma unroll
  for (int j = 0; j < DM_TILE_W; j++) {
    w[j] = (j < T.ntask) ? wlist[T.task_beg + j] : 0.0;
    acc[j] = 0.0;
  }
  if (T.norm) {
    for (int64_t p = b0 + threadIdx.x; p < b1; p += BT) acc[0] += W[p];
  } else {
    // outside the window both broadenings are pref/(b|x| sqrt(pi)) exp(-t^2) with
    // t = ln|w/x|/b - shift = u_w - c_x: ln|w|/b once per omega, ln|x|/b once per
    // pole, so each (pole, omega) costs one exp (the log is shared by the tile)
    const bool lg = (bd == DM_BD_LOGGAUSS_GAUSS);
    const double shift = lg ? 0.25 * bbroad : 0.0;
    const double pref = (lg ? 1.0 : exp(-0.25 * bbroad * bbroad)) / (bbroad * sqrt(M_PI));
    double u[DM_TILE_W];
#pragma unroll
    for (int j = 0; j < DM_TILE_W; j++) u[j] = log(fabs(w[j])) / bbroad;
    for (int64_t p = b0 + threadIdx.x; p < b1; p += BT) {
      const double wt = W[p];
      if (wt == 0.0) continue;
      const double x = dE[p];
      if (fabs(x) < T.twin) {
#pragma unroll
        for (int j = 0; j < DM_TILE_W; j++)
          if (j < T.ntask) acc[j] += wt * bdelta(bd, w[j], x, T.twin, bbroad, btemp);
        continue;
      }
      if ((T.sign > 0 && x < 0.0) || (T.sign < 0 && x >= 0.0)) continue;
      const double xx = (x == 0.0) ? 1.0e-100 : x;
      const double c = log(fabs(xx)) / bbroad + shift;
      const double f = wt * pref / fabs(xx);
#pragma unroll
      for (int j = 0; j < DM_TILE_W; j++)
        if (j < T.ntask && !(w[j] * xx < 0.0)) {
          const double t = u[j] - c;
          acc[j] += f * exp(-t * t);
        }
    }
  }
#pragma unroll
  for (int j = 0; j < DM_TILE_W; j++) {
    const double v = nrgw::wave_sum(acc[j]);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][j] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < T.ntask) {
    const int j = threadIdx.x;
    partial[T.part_beg + (int64_t)j * T.nslab + slab] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
  }
}

// per tile omega: sum of its slab partials in slab order
__global__ void reduce_tiles_kernel(const DmTile* __restrict__ tiles, int ntile, const double* __restrict__ partial,
                                    double* __restrict__ out) {
  const int t = blockIdx.x;
  if (t >= ntile) return;
  const DmTile T = tiles[t];
  for (int j = threadIdx.x; j < T.ntask; j += blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < T.nslab; c++) s += partial[T.part_beg + (int64_t)j * T.nslab + c];
    out[T.task_beg + j] = s;
  }
}

// per task: sum of its chunk partials in chunk order
__global__ void reduce_kernel(const DmTask* __restrict__ tasks, int ntask, const double* __restrict__ partial,
                              double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntask) return;
  const DmTask T = tasks[t];
  double s = 0.0;
  for (int c = 0; c < T.nchunk; c++) s += partial[T.chunk_beg + c];
  out[t] = s;
}

// ---------------------------------------------------------------------------
// derivative pass of the DM-NRG shell sums (no counterpart in the reference)
// ---------------------------------------------------------------------------
// change of frame of one block: M = th / (E_i - E_j) for non-degenerate pairs
// (the rotation to the perturbed eigenbasis), Tc = th inside degenerate
// clusters (|E_i - E_j| <= tol, the diagonal included)
__global__ void frame_kernel(const DmFrame* __restrict__ blks, const double* __restrict__ th, int64_t th_stride,
                             const double* __restrict__ en, double tol, double* __restrict__ M,
                             double* __restrict__ Tc, int64_t stride) {
  const DmFrame B = blks[blockIdx.y];
  const int t = blockIdx.z;
  const int64_t n = (int64_t)B.k * B.k;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / B.k), j = (int)(idx % B.k);
    const double d = en[B.e_off + i] - en[B.e_off + j];
    const double h = th[t * th_stride + B.off + idx];
    const bool deg = fabs(d) <= tol;
    M[t * stride + B.off + idx] = deg ? 0.0 : h / d;
    Tc[t * stride + B.off + idx] = deg ? h : 0.0;
  }
}

// W1 = cg (dA o Q + A o dQ), W2 = cg Q o (A Tc2 - Tc1 A): the weights of
// delta(w, E_j - E_i) and of its derivative in the pole energy
__global__ void tan_weights_kernel(const DmTanPair* __restrict__ pairs, const double* __restrict__ ops,
                                   const double* __restrict__ dAe, int64_t a_stride, const double* __restrict__ Q,
                                   const double* __restrict__ dQ, const double* __restrict__ cl, int64_t q_stride,
                                   double* __restrict__ W1, double* __restrict__ W2, int64_t w_stride) {
  const DmTanPair P = pairs[blockIdx.y];
  const int t = blockIdx.z;
  const int64_t n = (int64_t)P.k1 * P.k2;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    const double a = ops[P.a_off + idx], q = Q[P.q_off + idx];
    const double da = dAe[t * a_stride + P.a_off + idx];
    const double dq = dQ[t * q_stride + P.q_off + idx], c = cl[t * q_stride + P.q_off + idx];
    W1[t * w_stride + P.pole_off + idx] = P.cg * (da * q + a * dq);
    W2[t * w_stride + P.pole_off + idx] = P.cg * q * c;
  }
}

// BroadDelta / LorentzDelta of the DM-NRG method and their derivatives in the
// pole energy x:  d/dx BroadDelta = BroadDelta (2 ln|w/x| / b^2 - 1) / x,
// d/dx LorentzDelta = b (w - x) / (pi ((w - x)^2 + b^2)^2)
__device__ __forceinline__ double bdelta_dx(double w, double x, double twin, double bbroad, double btemp,
                                            double* gp) {
  if (fabs(x) < twin) {
    const double d = w - x, q = d * d + btemp * btemp;
    *gp = (1.0 / (2.0 * M_PI)) * btemp * 2.0 * d / (q * q);
    return (1.0 / (2.0 * M_PI)) * (btemp / q);
  }
  if (x == 0.0) x = 1.0e-100;
  if (w * x < 0.0) { *gp = 0.0; return 0.0; }
  const double a1 = log(fabs(w / x)) / bbroad;
  const double g = exp(-bbroad * bbroad / 4.0) / (bbroad * fabs(x) * sqrt(M_PI)) * exp(-a1 * a1);
  *gp = g * (2.0 * a1 / bbroad - 1.0) / x;
  return g;
}

__global__ __launch_bounds__(BT) void broaden_jvp_kernel(const DmTask* __restrict__ tasks,
                                                         const double* __restrict__ W1,
                                                         const double* __restrict__ W2, int64_t w_stride,
                                                         const double* __restrict__ dE, double bbroad, double btemp,
                                                         double* __restrict__ partial, int64_t p_stride) {
  __shared__ double red[BT / 64];
  const DmTask T = tasks[blockIdx.y];
  const int c = blockIdx.x, t = blockIdx.z;
  if (c >= T.nchunk) return;
  const int64_t b0 = T.pole_beg + (int64_t)c * BCH;
  const int64_t b1 = min(b0 + (int64_t)BCH, T.pole_end);
  const double* __restrict__ w1 = W1 + t * w_stride;
  const double* __restrict__ w2 = W2 + t * w_stride;
  double s = 0.0;
  for (int64_t p = b0 + threadIdx.x; p < b1; p += BT) {
    double gp;
    const double g = bdelta_dx(T.w, dE[p], T.twin, bbroad, btemp, &gp);
    s += w1[p] * g + w2[p] * gp;
  }
  s = nrgw::wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[t * p_stride + T.chunk_beg + c] = (red[0] + red[1]) + (red[2] + red[3]);
}

}  // namespace dmk

extern "C" {

hipError_t dm_launch_frame(const DmFrame* blks, int nblk, int kkmax, int ntan, const double* th, int64_t th_stride,
                           const double* en, double tol, double* M, double* Tc, int64_t stride, hipStream_t s) {
  if (nblk <= 0 || ntan <= 0) return hipSuccess;
  int gx = (kkmax + 255) / 256; if (gx > 1024) gx = 1024; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(dmk::frame_kernel, dim3(gx, nblk, ntan), dim3(256), 0, s, blks, th, th_stride, en, tol, M, Tc,
                     stride);
  return hipGetLastError();
}

hipError_t dm_launch_tan_weights(const DmTanPair* pairs, int npair, int kkmax, int ntan, const double* ops,
                                 const double* dAe, int64_t a_stride, const double* Q, const double* dQ,
                                 const double* cl, int64_t q_stride, double* W1, double* W2, int64_t w_stride,
                                 hipStream_t s) {
  if (npair <= 0 || ntan <= 0) return hipSuccess;
  int gx = (kkmax + 255) / 256; if (gx > 1024) gx = 1024; if (gx < 1) gx = 1;
  hipLaunchKernelGGL(dmk::tan_weights_kernel, dim3(gx, npair, ntan), dim3(256), 0, s, pairs, ops, dAe, a_stride, Q,
                     dQ, cl, q_stride, W1, W2, w_stride);
  return hipGetLastError();
}

hipError_t dm_launch_broaden_jvp(const DmTask* tasks, int ntask, int maxchunk, int ntan, const double* W1,
                                 const double* W2, int64_t w_stride, const double* dE, double bbroad, double btemp,
                                 double* partial, int64_t p_stride, double* out, hipStream_t s) {
  if (ntask <= 0 || ntan <= 0) return hipSuccess;
  if (maxchunk > 0) {
    hipLaunchKernelGGL(dmk::broaden_jvp_kernel, dim3(maxchunk, ntask, ntan), dim3(BT), 0, s, tasks, W1, W2, w_stride,
                       dE, bbroad, btemp, partial, p_stride);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  for (int t = 0; t < ntan; t++) {
    hipLaunchKernelGGL(dmk::reduce_kernel, dim3((ntask + 255) / 256), dim3(256), 0, s, tasks, ntask,
                       partial + t * p_stride, out + (int64_t)t * ntask);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}


hipError_t dm_launch_gather(const DmGBlk* blks, int nblk, int kmax, const DmSeg* segs, const double* rot,
                            const double* en, double bb, double zdisc, double* out, hipStream_t s) {
  if (nblk <= 0) return hipSuccess;
  int gx = (int)(((int64_t)kmax * kmax + 255) / 256);
  if (gx > 256) gx = 256;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(dmk::gather_kernel, dim3(gx, nblk), dim3(256), 0, s, blks, segs, rot, en, bb, zdisc, out);
  return hipGetLastError();
}

hipError_t dm_launch_weights(const DmPair* pairs, int npair, int kkmax, const double* ops, const double* prod,
                             const double* en, double* Wp, double* Wn, double* dE, hipStream_t s) {
  if (npair <= 0) return hipSuccess;
  int gx = (kkmax + 255) / 256;
  if (gx > 256) gx = 256;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(dmk::weights_kernel, dim3(gx, npair), dim3(256), 0, s, pairs, ops, prod, en, Wp, Wn, dE);
  return hipGetLastError();
}

hipError_t dm_launch_scale(const DmScale* blks, int nblk, int64_t maxel, const double* d, const double* a, double* out,
                           hipStream_t s) {
  if (nblk <= 0) return hipSuccess;
  int64_t gx = (maxel + 255) / 256;
  if (gx > 1024) gx = 1024;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(dmk::scale_kernel, dim3((int)gx, nblk), dim3(256), 0, s, blks, d, a, out);
  return hipGetLastError();
}

hipError_t dm_launch_broaden_tiles(const DmTile* tiles, int ntile, int maxslab, const double* w, const double* Wp,
                                   const double* Wn, const double* dE, int bd, double bbroad, double btemp,
                                   double* partial, double* out, hipStream_t s) {
  if (ntile <= 0) return hipSuccess;
  if (maxslab > 0) {
    hipLaunchKernelGGL(dmk::broaden_tiles_kernel, dim3(maxslab, ntile), dim3(BT), 0, s, tiles, w, Wp, Wn, dE, bd,
                       bbroad, btemp, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(dmk::reduce_tiles_kernel, dim3(ntile), dim3(64), 0, s, tiles, ntile, partial, out);
  return hipGetLastError();
}

}  // extern "C"
