// nrg_kernels.h -- device-side descriptors shared by the HIP kernels and the
// host code that launches them.  Plain structs of offsets and device
// pointers; no torch types.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NRG_MAX_CH 4

// one symmetric block for the eigensolver
struct EigDesc {
  int n;            // block dimension
  int pad;
  int64_t a_off;    // offset of the n*n matrix (in: H, out: eigenvector rows)
  int64_t w_off;    // offset of the n eigenvalues
  int64_t ws_off;   // workspace offset, n*n doubles
  int64_t vec_off;  // workspace offset, 7*n doubles
};

// a block-sparse operator over the previous kept basis (reduced elements)
struct OldOp {
  const double* mat;     // MatEl
  const int64_t* mtab;   // nbo*nbo: offset of block (bi,bj) or -1
  const int* ob;         // block of each old state
  const int* op;         // position of each old state in its block
  const int* bsz;        // size of each old block
  int nbo;               // number of old blocks
  int pad;
};

struct HnArgs {
  double* H;                 // output blocks (full n*n, row-major)
  const int64_t* h_off;      // per block
  const int* bbeg;           // first state of each block
  const int* bsize;          // block sizes
  const int* blk_sidx;       // coupling-table row (S index) of each block
  const int* stcf;           // StCameFrom of each basis state
  const int* itype;          // iType of each basis state
  const double* en_par;      // energy of the parent (kept) state
  const double* site_q;      // Q of each site type (eN term)
  const double* coef[NRG_MAX_CH];   // [nS][ntypes][ntypes] coupling tables
  OldOp f[NRG_MAX_CH];
  double chi[NRG_MAX_CH];
  double sqrtL, eN;
  int nch, ntypes;
};

struct FPair {
  int ni_bc, nj_bc;     // pre-cut block sizes (rows, cols of F)
  int ibeg_bc, jbeg_bc; // first basis state of the two blocks
  int64_t f_off;        // F output offset
  int64_t g_off;        // site-type table offset (ntypes*ntypes)
  int need_old;
  int op;               // index into FArgs::old
};

struct FArgs {
  double* F;
  const FPair* pairs;
  const double* G;
  const int* stcf;
  const int* itype;
  OldOp old[8];
  int ntypes;
  int pad;
};


struct CompactDesc {
  int k, n;
  int64_t src_off, dst_off;
};

// ---- derivative (tangent) pass of the fixed-truncation iteration ----
#define NRGK_MAX_TAN 4

// dH of every Abasis block per tangent t (full n x n at t*dh_stride + h_off[b])
struct TanHnArgs {
  double* dH;
  int64_t dh_stride;
  const int64_t* h_off;
  const int* bbeg;
  const int* bsize;
  const int* blk_sidx;
  const int* stcf;
  const int* itype;
  const double* th;          // kept-space Hamiltonian tangent of the previous step
  const int64_t* th_off;     // per previous kept block
  int64_t th_stride;
  const double* coef;        // coupling table [S index][typep][type]
  OldOp f;                   // f of the previous step (its block map and kept-basis view)
  const double* df;          // tangent of f, layout of f.mat
  int64_t df_stride;
  double chi, sqrtL;
  double dchi[NRGK_MAX_TAN];
  int ntypes, pad;
};

// one Aeig block for the per-state dE
struct TanBlk {
  int n, k;                  // block size, kept rows
  int64_t zoff;              // Z (and dH, Y) offset
  int64_t goff;              // G kept rows (k x n) offset
  int64_t e0;                // first state
};

// one AeigCut block: new kept-space tangent and kept-discarded mixing C
struct TanCut {
  int n, k;
  int64_t goff, e0;
  int64_t thoff;             // k x k
  int64_t coff;              // k x (n-k)
};

// the diagonal of one n x n operator block, to out[dst .. dst+n)
struct DiagJob {
  const double* src;
  int64_t dst;
  int n, pad;
};

extern "C" {
hipError_t nrgk_launch_diag_gather(const DiagJob* d_jobs, int njobs, double* out, hipStream_t s);
hipError_t nrgk_launch_tan_hn(const TanHnArgs* a, int nb, int nmax, int ntan, hipStream_t s);
hipError_t nrgk_launch_tan_diag(const TanBlk* d_blk, int nb, int nmax, int ntan, const double* Z, const double* Y,
                                int64_t y_stride, const double* G, int64_t g_stride, double* dE, int64_t e_stride,
                                hipStream_t s);
hipError_t nrgk_launch_tan_cut(const TanCut* d_cut, int nbc, int64_t maxel, int ntan, const double* G,
                               int64_t g_stride, const double* E, const double* dE, int64_t e_stride, int64_t gs,
                               double* th, int64_t th_stride, double* C, int64_t c_stride, hipStream_t s);
hipError_t nrgk_launch_eig(const EigDesc* d_desc, const int* d_order, int nb, int nmax, double* A, double* W,
                           double* WS, int* d_info, hipStream_t s);
hipError_t nrgk_launch_unpack(const EigDesc* d_desc, const int64_t* d_poff, const double* P, double* A, int nb,
                              int nmax, hipStream_t s);
hipError_t nrgk_launch_hn(const HnArgs* a, int nb, int nmax, hipStream_t s);
hipError_t nrgk_launch_fbasis(const FArgs* a, int npairs, int64_t maxel, hipStream_t s);
hipError_t nrgk_launch_compact(const CompactDesc* d_cd, int nb, int64_t maxel, const double* src, double* dst,
                               hipStream_t s);
}
