// nrg_dm.h -- DM-NRG on the device: what the forward chain keeps per shell
// (the data SaveArrays writes, src/NRGcodehandlerClass.cpp:887-933) and the
// kernels of the backward density-matrix sweep and the spectral sums.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

// rho_N(p,q) += cg * rot[rot_off + p*ld + q] for one (site type, N+1 block) child segment
struct DmSeg {
  int64_t rot_off;
  int ld;
  int pad;
  double cg;
};
// one Acut_N block of the gather: output k x k at out_off, segments [seg_beg, seg_end)
struct DmGBlk {
  int64_t out_off;
  int k, seg_beg, seg_end, pad;
};
// one operator block pair (b1 -> b2) at a shell: k1 x k2 poles
struct DmPair {
  int64_t a_off;    // operator block (k1 x k2) in the shell's operator buffer
  int64_t p1_off;   // rho1 A (k1 x k2) in the product buffer (DM-NRG form)
  int64_t p2_off;   // A rho2 (k1 x k2)
  int64_t pole_off; // first pole of the pair
  int64_t e1_off, e2_off;   // energies of the two blocks in the shell's energy array
  int k1, k2;
  double cg;
  int costi;        // 1: last shell, Costi weights
  int pad;
  double bb, invZ;  // Costi betabar and 1/Z
};
// one (shell, omega) task of the broadening sum
struct DmTask {
  int64_t pole_beg, pole_end;
  double w;         // omegabar
  double twin;      // Lorentzian window |dE| < twindow*T/D_N
  int chunk_beg;    // first partial-sum slot
  int nchunk;
};

struct DmShell {
  int N = 0;
  std::vector<double> qn, en;          // Acut_N: (Q,S) per block, kept energies
  std::vector<int> bbe, bbe_bc;        // block limits, pre-cut limits
  std::vector<int64_t> zoff;           // kept rows (k_b x n_b) of block b in zbuf
  std::string zbuf;
  std::vector<int> stcf, itype, bas_bbe;   // Abasis_N
  std::vector<double> bas_qn;
  std::vector<int> op_map, op_mbe;     // c_d blocks over Acut_N
  std::string opbuf;
  int nblocks() const { return (int)qn.size() / 2; }
  int bsize(int b) const { return bbe[2 * b + 1] - bbe[2 * b] + 1; }
  int bsize_bc(int b) const { return bbe_bc[2 * b + 1] - bbe_bc[2 * b] + 1; }
};

extern "C" {
hipError_t dm_launch_gather(const DmGBlk* blks, int nblk, int kmax, const DmSeg* segs, const double* rot,
                            double* out, hipStream_t s);
hipError_t dm_launch_weights(const DmPair* pairs, int npair, int kkmax, const double* ops, const double* prod,
                             const double* en, double* W, double* dE, hipStream_t s);
hipError_t dm_launch_broaden(const DmTask* tasks, int ntask, int maxchunk, const double* W, const double* dE,
                             double bbroad, double btemp, double* partial, double* out, hipStream_t s);
}
