// nrg_dm.hip -- DM-NRG kernels: the child trace-out of the rotated density
// matrix (src/DM_NRG/DM_NRG_CalcRhoN.cpp:252-449), the pole weights of the
// shell spectral sums (src/SpecFunc_allSTL.cpp:223-692, src/SpecFuncClass.cpp:
// 417-662) and their broadening (src/BroadDelta.cpp:12-73, MijxBDeltaEij
// src/SpecFunc_allSTL.cpp:36-114).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "nrg_dm.h"

namespace dmk {

// rho_N block = sum over its child segments of cg * (segment of the rotated rho_{N+1})
__global__ void gather_kernel(const DmGBlk* __restrict__ blks, const DmSeg* __restrict__ segs,
                              const double* __restrict__ rot, double* __restrict__ out) {
  const DmGBlk B = blks[blockIdx.y];
  const int64_t kk = (int64_t)B.k * B.k;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < kk; idx += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(idx / B.k), q = (int)(idx % B.k);
    double s = 0.0;
    for (int g = B.seg_beg; g < B.seg_end; g++) {
      const DmSeg S = segs[g];
      s += S.cg * rot[S.rot_off + (int64_t)p * S.ld + q];
    }
    out[B.out_off + idx] = s;
  }
}

// pole weights and energies of one operator block pair:
//   DM-NRG form  W = cg (rho1 A o A + A o A rho2)   (Termo1 + Termo2)
//   Costi form   W = cg A o A (e^{-bb E1_i} + e^{-bb E2_j}) / Z
//   dE = E2_j - E1_i
__global__ void weights_kernel(const DmPair* __restrict__ pairs, const double* __restrict__ ops,
                               const double* __restrict__ prod, const double* __restrict__ en,
                               double* __restrict__ W, double* __restrict__ dE) {
  const DmPair P = pairs[blockIdx.y];
  const int64_t n = (int64_t)P.k1 * P.k2;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / P.k2), j = (int)(idx % P.k2);
    const double a = ops[P.a_off + idx];
    const double e1 = en[P.e1_off + i], e2 = en[P.e2_off + j];
    double w;
    if (P.costi) w = P.cg * a * a * (exp(-P.bb * e1) + exp(-P.bb * e2)) * P.invZ;
    else w = P.cg * (prod[P.p1_off + idx] * a + a * prod[P.p2_off + idx]);
    W[P.pole_off + idx] = w;
    dE[P.pole_off + idx] = e2 - e1;
  }
}

__device__ __forceinline__ double bdelta(double w, double x, double twin, double bbroad, double btemp) {
  if (fabs(x) < twin) {   // LorentzDelta with the broadening dBroadTemp
    const double d = w - x;
    return (1.0 / (2.0 * M_PI)) * (btemp / (d * d + btemp * btemp));
  }
  if (x == 0.0) x = 1.0e-100;
  if (w * x < 0.0) return 0.0;   // BroadDelta: log-Gaussian, same-sign poles only
  const double a0 = exp(-bbroad * bbroad / 4.0) / (bbroad * fabs(x) * sqrt(M_PI));
  const double a1 = log(fabs(w / x)) / bbroad;
  return a0 * exp(-a1 * a1);
}

#define BT 256
#define BCH (BT * 8)   // poles per workgroup chunk
__global__ __launch_bounds__(BT) void broaden_kernel(const DmTask* __restrict__ tasks, const double* __restrict__ W,
                                                     const double* __restrict__ dE, double bbroad, double btemp,
                                                     double* __restrict__ partial) {
  __shared__ double red[BT / 64];
  const DmTask T = tasks[blockIdx.y];
  const int c = blockIdx.x;
  if (c >= T.nchunk) return;
  const int64_t b0 = T.pole_beg + (int64_t)c * BCH;
  const int64_t b1 = min(b0 + (int64_t)BCH, T.pole_end);
  double s = 0.0;
  for (int64_t p = b0 + threadIdx.x; p < b1; p += BT) {
    const double w = W[p];
    if (w != 0.0) s += w * bdelta(T.w, dE[p], T.twin, bbroad, btemp);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[T.chunk_beg + c] = (red[0] + red[1]) + (red[2] + red[3]);
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

}  // namespace dmk

extern "C" {

hipError_t dm_launch_gather(const DmGBlk* blks, int nblk, int kmax, const DmSeg* segs, const double* rot,
                            double* out, hipStream_t s) {
  if (nblk <= 0) return hipSuccess;
  int gx = (int)(((int64_t)kmax * kmax + 255) / 256);
  if (gx > 256) gx = 256;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(dmk::gather_kernel, dim3(gx, nblk), dim3(256), 0, s, blks, segs, rot, out);
  return hipGetLastError();
}

hipError_t dm_launch_weights(const DmPair* pairs, int npair, int kkmax, const double* ops, const double* prod,
                             const double* en, double* W, double* dE, hipStream_t s) {
  if (npair <= 0) return hipSuccess;
  int gx = (kkmax + 255) / 256;
  if (gx > 256) gx = 256;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(dmk::weights_kernel, dim3(gx, npair), dim3(256), 0, s, pairs, ops, prod, en, W, dE);
  return hipGetLastError();
}

hipError_t dm_launch_broaden(const DmTask* tasks, int ntask, int maxchunk, const double* W, const double* dE,
                             double bbroad, double btemp, double* partial, double* out, hipStream_t s) {
  if (ntask <= 0) return hipSuccess;
  if (maxchunk > 0) {
    hipLaunchKernelGGL(dmk::broaden_kernel, dim3(maxchunk, ntask), dim3(BT), 0, s, tasks, W, dE, bbroad, btemp,
                       partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(dmk::reduce_kernel, dim3((ntask + 255) / 256), dim3(256), 0, s, tasks, ntask, partial, out);
  return hipGetLastError();
}

}  // extern "C"
