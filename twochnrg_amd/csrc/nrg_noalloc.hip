// nrg_noalloc.hip -- the operator rotations of nrg_rotate_ops on caller-owned
// device buffers, with no device allocation: the intermediate products and the
// GEMM descriptors live in one fixed scratch array per device, set up once.
// For callers that cannot allocate (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/twochnrg_amd.h"
#include "nrg_eig2.h"
#include "nrg_internal.h"

// The scratch area: every device the library runs on holds one copy until the
// process ends.  Layout of a call: GEMM problems, tiles (each section padded
// to 256 bytes), then the intermediate products auxT = Zj F^T.  Every byte a
// call reads from it is written by that call first (descriptors by the copy,
// auxT by the first GEMM pass, which overwrites), so nothing is cleared.
// Calls that overlap on one device race on it.
static __device__ __align__(256) char scratch[NRG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_scratch[kMaxDevices] = {};   // device address of scratch, per device

// the device address of this device's scratch array (looked up once per device,
// with that device current)
int scratch_ptr(int dev, char** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return nrg::fail(NRG_ERR_HIP, "nrg_rotate_ops_noalloc: no device address for the scratch area");
    g_scratch[dev] = (char*)p;
  }
  *out = g_scratch[dev];
  return NRG_OK;
}

__global__ void zero_fill_kernel(double* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

int64_t pad256(int64_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" int nrg_rotate_ops_noalloc(void* stream, int npairs, const nrg_pair_desc* pairs, const double* dZ,
                                      int64_t zlen, const double* dF, int64_t flen, double* dOut, int64_t outlen,
                                      int precision) {
  using namespace nrg;
  if (npairs < 0 || outlen < 0 || (npairs > 0 && (!pairs || !dZ || !dF || !dOut)) || (outlen > 0 && !dOut))
    return fail(NRG_ERR_ARG, "nrg_rotate_ops_noalloc: bad arguments");
  if (precision != 0) return fail(NRG_ERR_ARG, "nrg_rotate_ops_noalloc: only precision 0 (FP64) is available");
  std::vector<RotJob> jobs;
  int64_t aux = 0;
  int rc = rot_jobs(npairs, pairs, zlen, flen, outlen, jobs, aux, "nrg_rotate_ops_noalloc");
  if (rc) return rc;
  std::vector<E2Gemm> probs;
  std::vector<E2Tile> tiles;
  size_t n1tiles = 0;
  const int n1 = npairs > 0 ? rot_plan(jobs, probs, tiles, n1tiles) : 0;
  const int64_t b_probs = pad256((int64_t)(sizeof(E2Gemm) * probs.size()));
  const int64_t b_tiles = pad256((int64_t)(sizeof(E2Tile) * tiles.size()));
  const int64_t need = b_probs + b_tiles + (int64_t)sizeof(double) * aux;
  if (need > (int64_t)NRG_NOALLOC_SCRATCH_BYTES)
    return fail(NRG_ERR_ARG, "nrg_rotate_ops_noalloc: the call needs " + std::to_string(need) +
                                 " bytes of scratch, more than the " + std::to_string((long long)NRG_NOALLOC_SCRATCH_BYTES) +
                                 " available");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(NRG_ERR_ARG, "nrg_rotate_ops_noalloc: device ordinal out of range");
  char* base = nullptr;
  if ((rc = scratch_ptr(dev, &base))) return rc;
  hipStream_t s = (hipStream_t)stream;
  // the output as nrg_rotate_ops leaves it: zero wherever no pair writes
  if (outlen > 0) {
    const int64_t gx = std::min<int64_t>((outlen + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)gx), dim3(256), 0, s, dOut, outlen);
    HIPCHK(hipGetLastError());
  }
  if (n1 > 0) {
    E2Gemm* dp = (E2Gemm*)base;
    E2Tile* dt = (E2Tile*)(base + b_probs);
    double* daux = (double*)(base + b_probs + b_tiles);
    HIPCHK(hipMemcpyAsync(dp, probs.data(), sizeof(E2Gemm) * probs.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dt, tiles.data(), sizeof(E2Tile) * tiles.size(), hipMemcpyHostToDevice, s));
    HIPCHK(e2_launch_gemm3(dp, dt, (int)n1tiles, dZ, dF, daux, s));
    HIPCHK(e2_launch_gemm3(dp, dt + n1tiles, (int)(tiles.size() - n1tiles), dZ, daux, dOut, s));
  }
  // synchronous at the boundary, as every entry point (the host descriptor
  // vectors are the copies' sources)
  HIPCHK(hipStreamSynchronize(s));
  return NRG_OK;
}
