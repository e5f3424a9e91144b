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
// over the leading kc x kc (the states with children: the kept prefix); FDM puts
// e^{-bb E}/Z_disc on the diagonal of the discarded states p >= kd
// compact (no tangents): only the kc x kc kept block at out_off and the k - kc
// discarded diagonal at d_off are stored (the rest of rho is zero)
struct DmGBlk {
  int64_t out_off;
  int64_t e_off;    // energies of the block in the shell's energy array
  int64_t d_off;
  int k, kc, kd, seg_beg, seg_end, compact;
};
// out[i][j] = a[i][j] * d[i - m0] (rows, mode 0) or d[j - n0] (columns, mode 1)
// for i in [m0, m1), j in [n0, n1); out and a share the leading dimension ld:
// products of a matrix with the discarded diagonal of a compact rho
struct DmScale {
  int64_t o_off, a_off, d_off;
  int m0, m1, n0, n1, ld, mode;
};
// one operator block pair (b1 -> b2) at a shell: k1 x k2 poles, each with a weight
// for omega >= 0 (Wp) and one for omega < 0 (Wn)
enum { DM_FORM_DMNRG = 0, DM_FORM_CFS = 1, DM_FORM_FDM = 2, DM_FORM_COSTI = 3 };
struct DmPair {
  int64_t a_off;    // operator block (k1 x k2) in the shell's operator buffer
  int64_t p1_off;   // rho1 A (k1 x k2) in the product buffer
  int64_t p2_off;   // A rho2 (k1 x k2)
  int64_t pole_off; // first pole of the pair
  int64_t e1_off, e2_off;   // energies of the two blocks in the shell's energy array
  int k1, k2;
  int kc1, kc2;     // kept prefixes of the two blocks (CFS/FDM classes)
  double cg;
  int form;         // DM_FORM_*
  int fix;          // FDM: NRG_FDM_FIX_* bits in effect (ONESIDED and NORM act on the weights)
  double bb, invZ;  // Costi betabar and 1/Z
};
// the broadening of a method (BDelta / BDeltaTemp, src/DM_NRG/DM_NRG.cpp:474-482)
enum { DM_BD_BROAD_LORENTZ = 0, DM_BD_BROAD_ANDERS = 1, DM_BD_LOGGAUSS_GAUSS = 2 };
// one (shell, omega) task of the broadening sum
struct DmTask {
  int64_t pole_beg, pole_end;
  double w;         // omegabar (its sign selects Wp / Wn)
  double twin;      // BDeltaTemp window |dE| < twindow*T/D_N
  int chunk_beg;    // first partial-sum slot
  int nchunk;
  int norm;         // 1: CalcNorm, the broadening is replaced by 1
  int pad;
};

// up to DM_TILE_W (shell, omega) sums that read the same poles: one shell, one
// weight array (Wp for omega >= 0, Wn below); the shell's poles are cut into
// nslab slabs of slab_len, one workgroup per (slab, tile)
#define DM_TILE_W 16
struct DmTile {
  int64_t pole_beg, pole_end;   // the shell's poles in the arena
  int64_t slab_len;
  int task_beg, ntask;          // omegas task_beg .. task_beg + ntask - 1
  int nslab, part_beg;          // partial sums at part_beg + j * nslab + slab
  int wsel;                     // 0: Wp, 1: Wn
  int norm;                     // CalcNorm: the broadening is replaced by 1 (ntask 1)
  int sign;                     // +1 / -1: every omega of the tile > 0 / < 0; 0 otherwise
  int pad;
  double twin;                  // BDeltaTemp window |dE| < twindow*T/D_N
};

struct DmShell {
  int N = 0;
  std::vector<double> qn, en;          // Acut_N: (Q,S) per block, energies of its states
  std::vector<unsigned char> kept;     // Kept flags (all 1 unless UpdateBefCut 1)
  std::vector<int> bbe, bbe_bc;        // block limits, pre-cut limits
  std::vector<int64_t> zoff;           // kept rows (k_b x n_b) of block b in zbuf
  std::string zbuf;
  std::vector<int> stcf, itype, bas_bbe;   // Abasis_N
  std::vector<double> bas_qn;
  std::vector<int> op_map, op_mbe;     // c_d blocks over Acut_N
  std::string opbuf;
  // derivative pass: per tangent, the kept-space Hamiltonian tangent (k x k per
  // block at th_off), the kept rows' tangent (k x n at dz_off) and the c_d
  // tangent (layout of opbuf); tangent t at t * (th_tot | dz_tot | op_cnt)
  int ntan = 0;
  std::vector<int64_t> th_off, dz_off;
  int64_t th_tot = 0, dz_tot = 0, op_cnt = 0;
  std::string thbuf, dzbuf, dopbuf;
  bool sz = false;                     // (Q, Sz) blocks: degeneracy 1, no Clebsch-Gordan factors
  int nblocks() const { return (int)qn.size() / 2; }
  double deg(int b) const { return sz ? 1.0 : 2.0 * qn[2 * b + 1] + 1.0; }
  int bsize(int b) const { return bbe[2 * b + 1] - bbe[2 * b] + 1; }
  int bsize_bc(int b) const { return bbe_bc[2 * b + 1] - bbe_bc[2 * b] + 1; }
};

// derivative pass: one Acut_N block of the change to the perturbed eigenbasis
struct DmFrame {
  int64_t off;      // k x k block (th, M, Tc)
  int64_t e_off;
  int k, pad;
};
// one operator block pair of the derivative pole weights
struct DmTanPair {
  int64_t a_off;    // operator block (k1 x k2)
  int64_t q_off;    // Q = rho1 A + A rho2 (k1 x k2)
  int64_t pole_off;
  int k1, k2;
  double cg;
};

extern "C" {
hipError_t dm_launch_frame(const DmFrame* blks, int nblk, int kkmax, int ntan, const double* th, int64_t th_stride,
                           const double* en, double tol, double* M, double* Tc, int64_t stride, hipStream_t s);
hipError_t dm_launch_tan_weights(const DmTanPair* pairs, int npair, int kkmax, int ntan, const double* ops,
                                 const double* dAe, int64_t a_stride, const double* Q, const double* dQ,
                                 const double* cl, int64_t q_stride, double* W1, double* W2, int64_t w_stride,
                                 hipStream_t s);
hipError_t dm_launch_broaden_jvp(const DmTask* tasks, int ntask, int maxchunk, int ntan, const double* W1,
                                 const double* W2, int64_t w_stride, const double* dE, double bbroad, double btemp,
                                 double* partial, int64_t p_stride, double* out, hipStream_t s);
hipError_t dm_launch_scale(const DmScale* blks, int nblk, int64_t maxel, const double* d, const double* a, double* out,
                           hipStream_t s);
hipError_t dm_launch_gather(const DmGBlk* blks, int nblk, int kmax, const DmSeg* segs, const double* rot,
                            const double* en, double bb, double zdisc, double* out, hipStream_t s);
hipError_t dm_launch_weights(const DmPair* pairs, int npair, int kkmax, const double* ops, const double* prod,
                             const double* en, double* Wp, double* Wn, double* dE, hipStream_t s);
// the shell sums of ntile tiles: out[task_beg + j] for the omegas w[task_beg + j]
hipError_t dm_launch_broaden_tiles(const DmTile* tiles, int ntile, int maxslab, const double* w, const double* Wp,
                                   const double* Wn, const double* dE, int bd, double bbroad, double btemp,
                                   double* partial, double* out, hipStream_t s);
}
