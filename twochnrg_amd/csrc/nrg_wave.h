// nrg_wave.h -- wavefront (64-lane) reductions of FP64 values on gfx950.
//
// wave_sum: a butterfly inside each row of 16 lanes through DPP moves (xor 1,
// xor 2 by quad permutation, then the half-row and row mirrors), which leaves
// every lane of a row with the row's sum, then the four row sums read from
// lanes 0, 16, 32, 48 and added as (r0 + r1) + (r2 + r3).  Every lane returns
// the same value, in a fixed order of additions.  A DPP move costs a few
// cycles where a cross-lane shuffle (ds_bpermute) is an LDS round trip; the
// reductions sit on the critical path of the per-column eigensolver kernels.
// Called with every lane of the wavefront active.
#pragma once
#include <hip/hip_runtime.h>

namespace nrgw {

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_d<0xB1>(v);    // quad_perm [1,0,3,2]: lane ^ 1
  v += dpp_d<0x4E>(v);    // quad_perm [2,3,0,1]: lane ^ 2
  v += dpp_d<0x141>(v);   // row_half_mirror: lane i <-> 7 - i within 8
  v += dpp_d<0x140>(v);   // row_mirror: lane i <-> 15 - i within 16
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// The same sum with the cross-row steps done by DPP row broadcasts (row_bcast:15
// into rows 1 and 3, then row_bcast:31 into rows 2 and 3), so lane 63 holds
// (r2 + r3) + (r0 + r1) -- bitwise the value wave_sum returns, addition being
// commutative -- and two lane reads instead of eight hand it to every lane.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_rows_d(double v) {   // rows outside ROWS read 0
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_bc(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);                 // every lane: its row's sum
  v += dpp_rows_d<0x142, 0xA>(v);       // rows 1, 3: + lane 15 of rows 0, 2
  v += dpp_rows_d<0x143, 0xC>(v);       // rows 2, 3: + lane 31 (rows 0 + 1)
  return readlane_d(v, 63);
}

}  // namespace nrgw
