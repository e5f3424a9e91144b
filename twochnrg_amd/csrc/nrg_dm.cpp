// nrg_dm.cpp -- DM-NRG on the device (DM_NRG main, src/DM_NRG/DM_NRG.cpp:28-531,
// QS symmetry; DM-NRG, CFS or FDM): the forward OneChQS chain keeps each shell's
// eigenvector rows (kept, or every state with UpdateBefCut 1), c_d blocks and
// child map; rho_Nmax = e^{-betabar E}/Z (DM_NRG_SetRhoNmax
// src/DM_NRG/DM_NRG_SetRhoNmax.cpp:9-126; FDM: Z_tot of the shell weights,
// src/DM_NRG/DM_NRG.cpp:203-297); the backward sweep rotates rho_{N+1} into the
// pre-cut basis with two batched GEMMs per shell (RotateMatrix
// src/RotateMatrix.cpp:17-99) and traces out the site as CG-weighted sums of
// child segments (DM_NRG_CalcRhoN_withSU2 src/DM_NRG/DM_NRG_CalcRhoN.cpp:252-449,
// FDM discarded diagonal :393-398); the shell spectral sums (DMNRG_SpecDens_M_STL
// src/SpecFunc_allSTL.cpp:223-692, CFS_SpecDens_M_STL :702-1240,
// FDM_SpecDens_M_STL :1253-2354, last shell CalcSpecCosti_T_N
// src/SpecFuncClass.cpp:417-662) become pole weights from two GEMMs per operator
// block pair and one broadening reduction over every (shell, omega); CFS/FDM
// sum a window of shells per omega (CalcSpecDM_NRG src/SpecFuncClass.cpp:1827-1885);
// the omega grids follow CalcSpecDM_NRG_FixedOmegas (src/SpecFuncClass.cpp:1932-2233)
// and CalcSpec_ManyOmegas (:2239-2341).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "nrg_bin.h"
#include "nrg_dm.h"
#include "nrg_eig2.h"
#include "nrg_internal.h"

namespace nrg {
int run_onechqs_dm(nrg_ctx* c, const nrg_onechqs_params& p, std::vector<DmShell>& dm, int befcut);
int run_twochqsz_dm(nrg_ctx* c, const nrg_twochqs_params& p, std::vector<DmShell>& dm);
int run_onechqs_dm_jvp(nrg_ctx* c, const nrg_onechqs_params& p, std::vector<DmShell>& dm, const unsigned char* kept,
                       const int64_t* kept_off, int ntan, const double* dpar);
double cgordan(double J1, double M1, double J2, double M2, double J, double M);  // nrg_driver.cpp
}

struct nrg_dm {
  nrg_ctx* c = nullptr;
  std::string pfx;                             // names of the handle's own buffers in the context
  nrg_dmnrg_params p{};
  std::vector<DmShell> sh;
  int NM = 0, specNM = 0;
  int method = 0;                              // spec1.UseCFS: 0 DM-NRG, 1 CFS, 2 FDM
  double temp = 0.0, betabar_spec = 0.0;
  double ztot = 0.0;                           // FDM Z_tot
  std::vector<double> wn;                      // FDM w_N
  int fix = 0;                                 // FDM: NRG_FDM_FIX_* bits in effect (0 unless FDM)
  double bb_last = 0.0;                        // FDM with NRG_FDM_FIX_LAST: 1/TempBar of the last shell
  // rho_N per shell and block in <pfx>R<N>: the k x k block at rho_off, or when
  // compact (no tangents) the kc x kc kept block at rho_off and the k - kc
  // discarded diagonal at rho_doff (rho has nothing else)
  bool compact = false;
  std::vector<std::vector<int64_t>> rho_off, rho_doff;
  std::vector<int64_t> rho_len;
  // poles (W, dE) of shells [ar_lo, ar_hi) in the arena, at most arena_max at
  // once; pole_beg/pole_end: every shell's range over all shells (the arena
  // holds them from ar_base on)
  bool layout_ready = false;
  std::vector<std::vector<DmPair>> lay_pairs;
  std::vector<std::vector<std::pair<int, int>>> lay_blks;
  std::vector<int64_t> pole_beg, pole_end;
  int ar_lo = -1, ar_hi = -1;
  int64_t ar_base = 0, arena_max = 0;
  // derivative pass (DM-NRG method): tangents carried through the chain and the
  // sweep (dmdR<N>), pole weights of delta and of its derivative (dm_W1, dm_W2)
  int ntan = 0;
  bool tpoles_ready = false;
  int64_t tpole_tot = 0;
  double ms_forward = 0, ms_sweep = 0, ms_poles = 0, ms_broaden = 0;
  uint64_t mem_peak = 0;
  ~nrg_dm() {   // the handle's buffers go back to the device with it
    if (!c) return;   // detached: the context and its buffers are gone already
    c->dms.erase(std::remove(c->dms.begin(), c->dms.end(), this), c->dms.end());
    if (pfx.empty()) return;
    for (auto it = c->bufs.begin(); it != c->bufs.end();)
      if (it->first.compare(0, pfx.size(), pfx) == 0) it = c->bufs.erase(it);
      else ++it;
  }
};

void nrg::dm_detach_all(nrg_ctx* c) {
  for (nrg_dm* d : c->dms) d->c = nullptr;
  c->dms.clear();
}

namespace {

using namespace nrg;

double calc_dn(double L, int N) {   // CalcDN src/CalcDN.cpp:4-15, z = 1
  const double HLF = (L > 0.0) ? 0.5 * (1.0 + 1.0 / L) : 1.0;
  return HLF * std::pow(L, -(N - 1) / 2.0);
}
double partition_func(const DmShell& s, double bb) {   // src/NRGarrayClass.cpp:997-1030
  double z = 0.0;
  for (int b = 0; b < s.nblocks(); b++)
    for (int i = s.bbe[2 * b]; i <= s.bbe[2 * b + 1]; i++) z += s.deg(b) * std::exp(-bb * s.en[i]);
  return z;
}
double partition_func_disc(const DmShell& s, double bb) {   // src/NRGarrayClass.cpp:1032-1058
  if (s.kept.empty()) return -1.0;
  double z = 0.0;
  for (int b = 0; b < s.nblocks(); b++)
    for (int i = s.bbe[2 * b]; i <= s.bbe[2 * b + 1]; i++)
      if (!s.kept[i]) z += s.deg(b) * std::exp(-bb * s.en[i]);
  return z;
}
int kept_prefix(const DmShell& s, int b, int& kc) {   // Kept states lead each energy-ordered block
  const int i0 = s.bbe[2 * b], k = s.bsize(b);
  kc = 0;
  while (kc < k && (s.kept.empty() || s.kept[i0 + kc])) kc++;
  for (int p = kc; p < k; p++)
    if (s.kept[i0 + p]) return fail(NRG_ERR_STATE, "DM-NRG: kept states are not a prefix of their block");
  return NRG_OK;
}
double site_S(int t) { return (t == 1 || t == 2) ? 0.5 : 0.0; }   // OneChQS site: types 0..3 = 0, up, down, updown

int gemm_groups(nrg_ctx* c, const char* tag, const std::vector<E2Gemm>& probs, const double* A, const double* B,
                double* C) {
  if (probs.empty()) return NRG_OK;
  std::vector<E2Tile> tiles;
  for (size_t q = 0; q < probs.size(); q++) {
    const int tm = (probs[q].m + 63) / 64, tn = (probs[q].n + 63) / 64;
    for (int a = 0; a < tm; a++)
      for (int b = 0; b < tn; b++) tiles.push_back(E2Tile{(int)q, a, b, 0});
  }
  E2Gemm* dp;
  E2Tile* dt;
  int rc;
  if ((rc = upload(c, (std::string(tag) + "_p").c_str(), probs, &dp))) return rc;
  if ((rc = upload(c, (std::string(tag) + "_t").c_str(), tiles, &dt))) return rc;
  HIPCHK(e2_launch_gemm(dp, dt, (int)tiles.size(), A, B, C, c->stream));
  return NRG_OK;
}

E2Gemm gemm(int m, int n, int k, int64_t a_off, int64_t as0, int64_t as1, int64_t b_off, int64_t bs0, int64_t bs1,
            int64_t c_off, int64_t cs0) {
  E2Gemm g{};
  g.m = m; g.n = n; g.k = k; g.dyn = -1;
  g.a_off = a_off; g.as0 = as0; g.as1 = as1;
  g.b_off = b_off; g.bs0 = bs0; g.bs1 = bs1;
  g.c_off = c_off; g.cs0 = cs0;
  g.alpha = 1.0; g.accumulate = 0;
  return g;
}
E2Gemm acc(E2Gemm g, double alpha = 1.0) {   // C += alpha A B
  g.accumulate = 1;
  g.alpha = alpha;
  return g;
}
double* dbuf(nrg_ctx* c, const std::string& n) { return (double*)c->buf(n).p; }
// divided difference of exp(-b E) between Ei and Ej
double dd_exp(double b, double Ei, double Ej) {
  const double d = Ei - Ej, xj = std::exp(-b * Ej);
  return d != 0.0 ? xj * std::expm1(-b * d) / d : -b * xj;
}
// pairs of the derivative of the shell sums closer than this are one degenerate cluster
const double kDegTol = 1e-9;

// rho_{N} from rho_{N+1}: rotation into the pre-cut basis, then the child trace-out;
// fdm: e^{-E/TempBar}/zdisc on the diagonal of the discarded states (zdisc != 0)
int sweep_shell(nrg_dm* D, int N, bool fdm, double TempBar, double zdisc) {
  nrg_ctx* c = D->c;
  const DmShell& a = D->sh[N];       // Acut_N
  const DmShell& b = D->sh[N + 1];   // Acut_{N+1} with Abasis_{N+1}
  int rc;
  const int nbb = b.nblocks();
  const bool full = D->ntan > 0;   // the derivative pass gathers from the whole rotated blocks
  // Abasis_{N+1} block -> surviving Acut_{N+1} block (matched by its first pre-cut state)
  const int nbas = (int)b.bas_qn.size() / 2;
  std::vector<int> bas2cut(nbas, -1);
  for (int q = 0; q < nbb; q++) {
    const int first = b.bbe_bc[2 * q];
    for (int ab = 0; ab < nbas; ab++)
      if (b.bas_bbe[2 * ab] == first) { bas2cut[ab] = q; break; }
  }
  std::vector<int> st_blk((size_t)b.stcf.size());
  for (int ab = 0; ab < nbas; ab++)
    for (int i = b.bas_bbe[2 * ab]; i <= b.bas_bbe[2 * ab + 1]; i++) st_blk[i] = ab;
  std::vector<std::vector<int>> children(a.en.size());
  for (size_t cidx = 0; cidx < b.stcf.size(); cidx++) {
    if (b.stcf[cidx] < 0 || b.stcf[cidx] >= (int)a.en.size())
      return fail(NRG_ERR_STATE, "DM-NRG: StCameFrom outside the parent shell");
    children[b.stcf[cidx]].push_back((int)cidx);
  }
  // T_b = rho_b Z_b (k x n) per Acut_{N+1} block; the whole rotation Z_b^T T_b
  // (n x n) only for the derivative pass
  std::vector<int64_t> toff(nbb), roff(nbb);
  int64_t ttot = 0, rtot = 0;
  for (int q = 0; q < nbb; q++) {
    const int k = b.bsize(q), n = b.bsize_bc(q);
    toff[q] = ttot; ttot += (int64_t)k * n;
    roff[q] = rtot; rtot += full ? (int64_t)n * n : 0;
  }
  // segments: children of a block's first state; the child of state p of the
  // block with the same type is the segment start + p.  Without tangents each
  // segment's kc x kc diagonal block of Z^T T is its own small product (only
  // those blocks survive the trace over the site), stored compactly.
  std::vector<DmGBlk> gb;
  std::vector<DmSeg> segs;
  std::vector<E2Gemm> gs;
  int kmax = 1;
  const int nba = a.nblocks();
  D->rho_off[N].assign(nba, 0);
  D->rho_doff[N].assign(nba, 0);
  int64_t otot = 0, stot = 0;
  const bool fdm_diag = fdm && std::fabs(zdisc - 0.0) > 1e-20;
  for (int B = 0; B < nba; B++) {
    const int i0 = a.bbe[2 * B], k = a.bsize(B);
    const double Sold = a.qn[2 * B + 1];
    int kc;
    if ((rc = kept_prefix(a, B, kc))) return rc;
    for (int p = 0; p < k; p++)   // the kept states are the parents of the N+1 basis
      if (children[i0 + p].empty() != (p >= kc))
        return fail(NRG_ERR_STATE, "DM-NRG: child states do not match the kept states");
    DmGBlk G{};
    G.out_off = otot; G.e_off = i0; G.k = k; G.kc = kc; G.kd = fdm_diag ? kc : k; G.seg_beg = (int)segs.size();
    G.compact = D->compact ? 1 : 0;
    D->rho_off[N][B] = otot;
    if (D->compact) {
      otot += (int64_t)kc * kc;
      G.d_off = otot;
      D->rho_doff[N][B] = otot;
      otot += k - kc;
    } else {
      otot += (int64_t)k * k;
    }
    kmax = std::max(kmax, k);
    if (kc > 0) for (int c0 : children[i0]) {
      const int t = b.itype[c0], ab = st_blk[c0];
      for (int p = 1; p < kc; p++) {
        const int cp = c0 + p;
        if (cp >= (int)b.stcf.size() || b.stcf[cp] != i0 + p || b.itype[cp] != t || st_blk[cp] != ab)
          return fail(NRG_ERR_STATE, "DM-NRG: child states of a block are not one contiguous segment per type");
      }
      const int q = bas2cut[ab];
      if (q < 0) continue;   // block cut at N+1: its rotated rho is zero
      const double Si = b.bas_qn[2 * ab + 1], St = site_S(t);
      double cg = b.sz ? 1.0 : 0.0;   // (Q, Sz) blocks: the plain partial trace
      if (!b.sz)
        for (double Sz = -Si; Sz <= Si; Sz += 1.0) {
          const double x = cgordan(Sold, Sold, St, Sz - Sold, Si, Sz);
          cg += x * x;
        }
      const int n = b.bsize_bc(q), kq = b.bsize(q);
      const int pc = c0 - b.bas_bbe[2 * ab];
      if (pc < 0 || pc + kc > n) return fail(NRG_ERR_STATE, "DM-NRG: a child segment leaves its block");
      DmSeg S{};
      if (full) {
        S.rot_off = roff[q] + (int64_t)pc * n + pc;
        S.ld = n;
      } else {
        // (Z^T T)[pc+p][pc+p'] = sum_i Z[i][pc+p] T[i][pc+p']
        S.rot_off = stot;
        S.ld = kc;
        gs.push_back(gemm(kc, kc, kq, b.zoff[q] + pc, 1, n, toff[q] + pc, n, 1, stot, kc));
        stot += (int64_t)kc * kc;
      }
      S.cg = cg;
      segs.push_back(S);
    }
    G.seg_end = (int)segs.size();
    gb.push_back(G);
  }
  DBuf& tb = c->buf("dm_T");
  DBuf& rb = c->buf("dm_rot");
  if ((rc = tb.reserve(sizeof(double) * (size_t)(ttot + 1)))) return rc;
  if ((rc = rb.reserve(sizeof(double) * (size_t)((full ? rtot : stot) + 1)))) return rc;
  double* dR1 = (double*)c->buf(D->pfx + "R" + std::to_string(N + 1)).p;
  double* dZ1 = (double*)c->buf(b.zbuf).p;
  if (full) {
    std::vector<E2Gemm> g1, g2;
    for (int q = 0; q < nbb; q++) {
      const int k = b.bsize(q), n = b.bsize_bc(q);
      g1.push_back(gemm(k, n, k, D->rho_off[N + 1][q], k, 1, b.zoff[q], n, 1, toff[q], n));
      g2.push_back(gemm(n, n, k, b.zoff[q], 1, n, toff[q], n, 1, roff[q], n));
    }
    if ((rc = gemm_groups(c, "dm_g1", g1, dR1, dZ1, (double*)tb.p))) return rc;
    if ((rc = gemm_groups(c, "dm_g2", g2, dZ1, (double*)tb.p, (double*)rb.p))) return rc;
  } else {
    // rho_{N+1} is its kept block plus a diagonal on the discarded states (the
    // gather writes nothing else): T = [rho_KK Z_K ; diag(rho_DD) Z_D]
    std::vector<E2Gemm> g1;
    std::vector<DmScale> sc;
    int64_t smax = 1;
    for (int q = 0; q < nbb; q++) {
      const int k = b.bsize(q), n = b.bsize_bc(q);
      int kcq;
      if ((rc = kept_prefix(b, q, kcq))) return rc;
      if (kcq > 0) g1.push_back(gemm(kcq, n, kcq, D->rho_off[N + 1][q], kcq, 1, b.zoff[q], n, 1, toff[q], n));
      if (kcq < k) {
        sc.push_back(DmScale{toff[q], b.zoff[q], D->rho_doff[N + 1][q], kcq, k, 0, n, n, 0});
        smax = std::max<int64_t>(smax, (int64_t)(k - kcq) * n);
      }
    }
    if ((rc = gemm_groups(c, "dm_g1", g1, dR1, dZ1, (double*)tb.p))) return rc;
    if (!sc.empty()) {
      DmScale* dsc;
      if ((rc = upload(c, "dm_sc", sc, &dsc))) return rc;
      HIPCHK(dm_launch_scale(dsc, (int)sc.size(), smax, dR1, dZ1, (double*)tb.p, c->stream));
    }
    if ((rc = gemm_groups(c, "dm_g2", gs, dZ1, (double*)tb.p, (double*)rb.p))) return rc;
  }
  D->rho_len[N] = otot;
  DBuf& ob = c->buf(D->pfx + "R" + std::to_string(N));
  if ((rc = ob.reserve(sizeof(double) * (size_t)(otot + 1), true))) return rc;
  DmGBlk* dgb;
  DmSeg* dsg;
  double* den;
  if ((rc = upload(c, "dm_gb", gb, &dgb))) return rc;
  if ((rc = upload(c, "dm_sg", segs, &dsg))) return rc;
  if ((rc = upload(c, "dm_gen", a.en, &den))) return rc;
  HIPCHK(dm_launch_gather(dgb, (int)gb.size(), kmax, dsg, (double*)rb.p, den, fdm_diag ? 1.0 / TempBar : 0.0,
                          fdm_diag ? zdisc : 1.0, (double*)ob.p, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (D->ntan > 0) {
    // derivative, in the chain's kept-space frame: the rotation of rho_{N+1}
    // moves by Z^T (drho Z + rho dZ) + dZ^T (rho Z); the trace-out is linear
    const int T = D->ntan;
    std::vector<E2Gemm> h1, h2, h3, h4;
    for (int q = 0; q < nbb; q++) {
      const int k = b.bsize(q), n = b.bsize_bc(q);
      for (int t = 0; t < T; t++) {
        const int64_t tt = (int64_t)t * ttot + toff[q], tr = (int64_t)t * rtot + roff[q];
        const int64_t dz = (int64_t)t * b.dz_tot + b.dz_off[q];
        h1.push_back(gemm(k, n, k, (int64_t)t * D->rho_len[N + 1] + D->rho_off[N + 1][q], k, 1, b.zoff[q], n, 1, tt,
                          n));                                                                  // drho Z
        h2.push_back(acc(gemm(k, n, k, D->rho_off[N + 1][q], k, 1, dz, n, 1, tt, n)));          // + rho dZ
        h3.push_back(gemm(n, n, k, b.zoff[q], 1, n, tt, n, 1, tr, n));                          // Z^T (.)
        h4.push_back(acc(gemm(n, n, k, dz, 1, n, toff[q], n, 1, tr, n)));                       // + dZ^T (rho Z)
      }
    }
    DBuf& ttb = c->buf("dm_tT");
    DBuf& trb = c->buf("dm_trot");
    DBuf& dob = c->buf(D->pfx + "dR" + std::to_string(N));
    if ((rc = ttb.reserve(sizeof(double) * (size_t)(T * ttot + 1)))) return rc;
    if ((rc = trb.reserve(sizeof(double) * (size_t)(T * rtot + 1)))) return rc;
    if ((rc = dob.reserve(sizeof(double) * (size_t)(T * otot + 1), true))) return rc;
    double* dR1t = dbuf(c, D->pfx + "dR" + std::to_string(N + 1));
    double* dZ1t = dbuf(c, b.dzbuf);
    if ((rc = gemm_groups(c, "dm_h1", h1, dR1t, dZ1, (double*)ttb.p))) return rc;
    if ((rc = gemm_groups(c, "dm_h2", h2, dR1, dZ1t, (double*)ttb.p))) return rc;
    if ((rc = gemm_groups(c, "dm_h3", h3, dZ1, (double*)ttb.p, (double*)trb.p))) return rc;
    if ((rc = gemm_groups(c, "dm_h4", h4, dZ1t, (double*)tb.p, (double*)trb.p))) return rc;
    for (int t = 0; t < T; t++)
      HIPCHK(dm_launch_gather(dgb, (int)gb.size(), kmax, dsg, (double*)trb.p + (int64_t)t * rtot, den, 0.0, 1.0,
                              (double*)dob.p + (int64_t)t * otot, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return NRG_OK;
}

double op_cgfactor(double S1, double S2) {
  double cg = 0.0;
  for (double Sz = -S1; Sz <= S1; Sz += 1.0) {
    const double x = cgordan(S1, Sz, 0.5, 0.5, S2, Sz + 0.5);
    cg += x * x;
  }
  return cg;
}

// pole weights (Wp for omega >= 0, Wn for omega < 0) and energies of every shell
// 0 .. specNM-1 (omega independent).  CFS/FDM: a block pair with no kept state
// on either side, or no discarded state on either side, has no poles
// (src/SpecFunc_allSTL.cpp:1375-1381, :824-830).
int pole_layout(nrg_dm* D) {
  if (D->layout_ready) return NRG_OK;
  int rc;
  const int ns = D->specNM;   // shells with spectral weight: 0 .. specNM-1
  const double bb = (std::fabs(D->temp) < 1e-20) ? 1.0e10 : D->betabar_spec;
  D->lay_pairs.assign(ns, std::vector<DmPair>());
  D->lay_blks.assign(ns, std::vector<std::pair<int, int>>());
  D->pole_beg.assign(ns, 0);
  D->pole_end.assign(ns, 0);
  int64_t ptot = 0;
  for (int N = 0; N < ns; N++) {
    const DmShell& s = D->sh[N];
    const bool costi = (N == D->specNM - 1);
    const int form = costi ? DM_FORM_COSTI : D->method;
    // the last shell: the Costi form with its own 1/Z, or (NRG_FDM_FIX_LAST) the
    // FDM shell of rho_Nmax = e^{-b_last E}/Z_tot, the same sum with other scalars
    const bool fdm_last = costi && (D->fix & NRG_FDM_FIX_LAST);
    const double bbN = fdm_last ? D->bb_last : bb;
    const double invZ = costi ? (fdm_last ? 1.0 / D->ztot : 1.0 / partition_func(s, bb)) : 0.0;
    D->pole_beg[N] = ptot;
    for (size_t q = 0; q + 1 < s.op_map.size(); q += 2) {
      const int b1 = s.op_map[q], b2 = s.op_map[q + 1];
      const int k1 = s.bsize(b1), k2 = s.bsize(b2);
      DmPair P{};
      P.a_off = s.op_mbe[q];
      P.k1 = k1; P.k2 = k2;
      P.e1_off = s.bbe[2 * b1]; P.e2_off = s.bbe[2 * b2];
      P.cg = s.sz ? 1.0 : op_cgfactor(s.qn[2 * b1 + 1], s.qn[2 * b2 + 1]);
      P.form = form;
      P.fix = D->fix;
      P.bb = bbN; P.invZ = invZ;
      if ((rc = kept_prefix(s, b1, P.kc1))) return rc;
      if ((rc = kept_prefix(s, b2, P.kc2))) return rc;
      if (form == DM_FORM_CFS || form == DM_FORM_FDM) {
        const bool nokept = P.kc1 == 0 && P.kc2 == 0, nodisc = P.kc1 == k1 && P.kc2 == k2;
        if (nodisc || (nokept && !(D->fix & NRG_FDM_FIX_PAIRS))) continue;
      }
      P.pole_off = ptot;
      ptot += (int64_t)k1 * k2;
      D->lay_pairs[N].push_back(P);
      D->lay_blks[N].emplace_back(b1, b2);
    }
    D->pole_end[N] = ptot;
  }
  // the arena holds at most this many poles (three doubles each) at once: 40%
  // of the device memory free now, which leaves the per-shell products room
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  D->arena_max = std::max<int64_t>((int64_t)(0.4 * (double)fr / 24.0), 1 << 20);
  D->layout_ready = true;
  return NRG_OK;
}

// pole weights and energies of shells N0 .. N1-1 into the arena (offsets from
// the first pole of N0).  Costi at the last shell; CFS/FDM: a block pair with
// no kept state on either side, or no discarded state on either side, has no
// poles (src/SpecFunc_allSTL.cpp:1375-1381, :824-830).
int build_pole_range(nrg_dm* D, int N0, int N1) {
  nrg_ctx* c = D->c;
  int rc;
  if ((rc = pole_layout(D))) return rc;
  if (D->ar_lo == N0 && D->ar_hi == N1) return NRG_OK;
  auto t0 = std::chrono::steady_clock::now();
  const int64_t base = D->pole_beg[N0], np = D->pole_end[N1 - 1] - base;
  DBuf& wpb = c->buf(D->pfx + "Wp");
  DBuf& wnb = c->buf(D->pfx + "Wn");
  DBuf& eb = c->buf(D->pfx + "dE");
  D->ar_lo = D->ar_hi = -1;
  if ((rc = wpb.reserve(sizeof(double) * (size_t)(np + 1), true))) return rc;
  if ((rc = wnb.reserve(sizeof(double) * (size_t)(np + 1), true))) return rc;
  if ((rc = eb.reserve(sizeof(double) * (size_t)(np + 1), true))) return rc;
  for (int N = N0; N < N1; N++) {
    const DmShell& s = D->sh[N];
    std::vector<DmPair> pairs = D->lay_pairs[N];
    if (pairs.empty()) continue;
    const bool costi = (N == D->specNM - 1);
    std::vector<E2Gemm> g1, g2;
    int64_t prod = 0;
    int kkmax = 1;
    // rho1 A (k1 x k2) and A rho2 (k1 x k2) over the whole blocks: rho has no
    // kept-discarded elements, so the products of the reference's kept/discarded
    // sub-blocks are sub-blocks of these
    std::vector<DmScale> sc;
    int64_t smax = 1;
    for (size_t q = 0; q < pairs.size(); q++) {
      DmPair& P = pairs[q];
      P.pole_off -= base;
      if (!costi) {
        const int b1 = D->lay_blks[N][q].first, b2 = D->lay_blks[N][q].second;
        P.p1_off = prod;
        P.p2_off = prod + (int64_t)P.k1 * P.k2;
        prod += 2 * (int64_t)P.k1 * P.k2;
        if (!D->compact) {
          g1.push_back(gemm(P.k1, P.k2, P.k1, D->rho_off[N][b1], P.k1, 1, P.a_off, P.k2, 1, P.p1_off, P.k2));
          g2.push_back(gemm(P.k1, P.k2, P.k2, P.a_off, P.k2, 1, D->rho_off[N][b2], P.k2, 1, P.p2_off, P.k2));
        } else {
          // rho1 A: kept rows through rho1's kept block, discarded rows scaled by its diagonal;
          // A rho2: the same by columns
          const int c1 = P.kc1, c2 = P.kc2;
          if (c1 > 0) g1.push_back(gemm(c1, P.k2, c1, D->rho_off[N][b1], c1, 1, P.a_off, P.k2, 1, P.p1_off, P.k2));
          if (c1 < P.k1)
            sc.push_back(DmScale{P.p1_off, P.a_off, D->rho_doff[N][b1], c1, P.k1, 0, P.k2, P.k2, 0});
          if (c2 > 0) g2.push_back(gemm(P.k1, c2, c2, P.a_off, P.k2, 1, D->rho_off[N][b2], c2, 1, P.p2_off, P.k2));
          if (c2 < P.k2)
            sc.push_back(DmScale{P.p2_off, P.a_off, D->rho_doff[N][b2], 0, P.k1, c2, P.k2, P.k2, 1});
          smax = std::max<int64_t>(smax, (int64_t)P.k1 * P.k2);
        }
      }
      kkmax = std::max(kkmax, P.k1 * P.k2);
    }
    DBuf& pb = c->buf("dm_prod");
    if ((rc = pb.reserve(sizeof(double) * (size_t)(prod + 1)))) return rc;
    double* dO = (double*)c->buf(s.opbuf).p;
    if (!costi) {
      double* dR = (double*)c->buf(D->pfx + "R" + std::to_string(N)).p;
      if ((rc = gemm_groups(c, "dm_p1", g1, dR, dO, (double*)pb.p))) return rc;
      if ((rc = gemm_groups(c, "dm_p2", g2, dO, dR, (double*)pb.p))) return rc;
      if (!sc.empty()) {
        DmScale* dsc;
        if ((rc = upload(c, "dm_psc", sc, &dsc))) return rc;
        HIPCHK(dm_launch_scale(dsc, (int)sc.size(), smax, dR, dO, (double*)pb.p, c->stream));
      }
    }
    double* dEn;
    DmPair* dpairs;
    if ((rc = upload(c, "dm_en", s.en, &dEn))) return rc;
    if ((rc = upload(c, "dm_pairs", pairs, &dpairs))) return rc;
    HIPCHK(dm_launch_weights(dpairs, (int)pairs.size(), kkmax, dO, (double*)pb.p, dEn, (double*)wpb.p,
                             (double*)wnb.p, (double*)eb.p, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  D->ar_lo = N0;
  D->ar_hi = N1;
  D->ar_base = base;
  D->ms_poles += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return NRG_OK;
}

// every shell's poles at once (the derivative pass reads them all)
int build_poles(nrg_dm* D) {
  int rc;
  if ((rc = pole_layout(D))) return rc;
  const int ns = D->specNM;
  if (D->pole_end[ns - 1] > D->arena_max)
    return fail(NRG_ERR_HIP, "DM-NRG: the poles of all shells do not fit in device memory");
  return build_pole_range(D, 0, ns);
}

// Derivative pole weights of every shell (DM-NRG method; the last shell in the
// Costi form).  The shell's tangents live in the chain's kept-space frame; the
// perturbed eigenbasis is reached with M = th / (E_i - E_j) on non-degenerate
// pairs (dX -> dX + M1 X - X M2), while inside degenerate clusters th (Tc)
// moves the poles:  W1 = cg (dA o Q + A o dQ), W2 = cg Q o (A Tc2 - Tc1 A),
// Q = rho1 A + A rho2, and the shell sum's derivative is
// sum_p W1_p delta(w, dE_p) + W2_p d delta / d dE_p.
int build_poles_jvp(nrg_dm* D) {
  nrg_ctx* c = D->c;
  int rc;
  const int T = D->ntan, ns = D->specNM;
  const double bb = (std::fabs(D->temp) < 1e-20) ? 1.0e10 : D->betabar_spec;
  int64_t ptot = 0;
  for (int N = 0; N < ns; N++)
    for (size_t q = 0; q + 1 < D->sh[N].op_map.size(); q += 2)
      ptot += (int64_t)D->sh[N].bsize(D->sh[N].op_map[q]) * D->sh[N].bsize(D->sh[N].op_map[q + 1]);
  if (ptot != D->pole_end[ns - 1]) return fail(NRG_ERR_STATE, "DM-NRG derivative: pole layout differs");
  DBuf& w1b = c->buf(D->pfx + "W1");
  DBuf& w2b = c->buf(D->pfx + "W2");
  if ((rc = w1b.reserve(sizeof(double) * (size_t)(T * ptot + 1)))) return rc;
  if ((rc = w2b.reserve(sizeof(double) * (size_t)(T * ptot + 1)))) return rc;
  int64_t pole = 0;
  for (int N = 0; N < ns; N++) {
    const DmShell& s = D->sh[N];
    const bool costi = (N == ns - 1);
    const int nb = s.nblocks();
    const int64_t btot = s.th_tot, opc = s.op_cnt;
    std::vector<int64_t> boff(s.th_off);
    if ((int)boff.size() != nb) return fail(NRG_ERR_STATE, "DM-NRG derivative: block layout");
    std::vector<DmFrame> fr(nb);
    int kkmax = 1;
    for (int b = 0; b < nb; b++) {
      const int k = s.bsize(b);
      if (!costi && D->rho_off[N][b] != boff[b]) return fail(NRG_ERR_STATE, "DM-NRG derivative: rho layout");
      fr[b] = DmFrame{boff[b], (int64_t)s.bbe[2 * b], k, 0};
      kkmax = std::max(kkmax, k * k);
    }
    DmFrame* dfr;
    double* den;
    if ((rc = upload(c, "dmt_fr", fr, &dfr))) return rc;
    if ((rc = upload(c, "dmt_en", s.en, &den))) return rc;
    DBuf& mb = c->buf("dmt_M");
    DBuf& tcb = c->buf("dmt_Tc");
    DBuf& drb = c->buf("dmt_dRe");
    if ((rc = mb.reserve(sizeof(double) * (size_t)(T * btot + 1)))) return rc;
    if ((rc = tcb.reserve(sizeof(double) * (size_t)(T * btot + 1)))) return rc;
    if ((rc = drb.reserve(sizeof(double) * (size_t)(T * btot + 1)))) return rc;
    HIPCHK(dm_launch_frame(dfr, nb, kkmax, T, dbuf(c, s.thbuf), s.th_tot, den, kDegTol, (double*)mb.p,
                           (double*)tcb.p, btot, c->stream));
    double* R;
    if (costi) {
      // rho = exp(-bb H)/Z on the kept space; in the eigenbasis its tangent lives
      // in the degenerate clusters: dd_exp(bb) th / Z, minus rho dZ / Z
      std::vector<double> th((size_t)T * btot + 1);
      HIPCHK(hipMemcpyAsync(th.data(), dbuf(c, s.thbuf), sizeof(double) * (size_t)(T * btot), hipMemcpyDeviceToHost,
                            c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      const double Z = partition_func(s, bb);
      std::vector<double> r((size_t)btot + 1, 0.0), dr((size_t)T * btot + 1, 0.0);
      for (int b = 0; b < nb; b++) {
        const int k = s.bsize(b), i0 = s.bbe[2 * b];
        for (int i = 0; i < k; i++) r[boff[b] + (int64_t)i * k + i] = std::exp(-bb * s.en[i0 + i]) / Z;
      }
      for (int t = 0; t < T; t++) {
        double dZ = 0.0;
        for (int b = 0; b < nb; b++) {
          const int k = s.bsize(b), i0 = s.bbe[2 * b];
          const double dg = s.deg(b);
          for (int i = 0; i < k; i++)
            dZ += dg * th[(size_t)t * btot + boff[b] + (int64_t)i * k + i] * (-bb * std::exp(-bb * s.en[i0 + i]));
        }
        for (int b = 0; b < nb; b++) {
          const int k = s.bsize(b), i0 = s.bbe[2 * b];
          for (int i = 0; i < k; i++)
            for (int j = 0; j < k; j++) {
              const double Ei = s.en[i0 + i], Ej = s.en[i0 + j];
              const int64_t e = boff[b] + (int64_t)i * k + j;
              double v = std::fabs(Ei - Ej) <= kDegTol ? dd_exp(bb, Ei, Ej) * th[(size_t)t * btot + e] / Z : 0.0;
              if (i == j) v -= r[e] * dZ / Z;
              dr[(size_t)t * btot + e] = v;
            }
        }
      }
      double* dRc;
      if ((rc = upload(c, "dmt_Rc", r, &dRc))) return rc;
      R = dRc;
      HIPCHK(hipMemcpyAsync(drb.p, dr.data(), sizeof(double) * (size_t)(T * btot), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    } else {
      R = dbuf(c, D->pfx + "R" + std::to_string(N));
      HIPCHK(hipMemcpyAsync(drb.p, dbuf(c, D->pfx + "dR" + std::to_string(N)), sizeof(double) * (size_t)(T * btot),
                            hipMemcpyDeviceToDevice, c->stream));
      std::vector<E2Gemm> g1, g2;   // drho += M rho - rho M
      for (int b = 0; b < nb; b++) {
        const int k = s.bsize(b);
        for (int t = 0; t < T; t++) {
          const int64_t o = (int64_t)t * btot + boff[b];
          g1.push_back(acc(gemm(k, k, k, o, k, 1, boff[b], k, 1, o, k)));
          g2.push_back(acc(gemm(k, k, k, boff[b], k, 1, o, k, 1, o, k), -1.0));
        }
      }
      if ((rc = gemm_groups(c, "dmt_r1", g1, (double*)mb.p, R, (double*)drb.p))) return rc;
      if ((rc = gemm_groups(c, "dmt_r2", g2, R, (double*)mb.p, (double*)drb.p))) return rc;
    }
    // operator block pairs
    std::vector<DmTanPair> tp;
    int64_t qtot = 0;
    int kk2 = 1;
    for (size_t q = 0; q + 1 < s.op_map.size(); q += 2) {
      const int b1 = s.op_map[q], b2 = s.op_map[q + 1];
      DmTanPair P{};
      P.a_off = s.op_mbe[q]; P.k1 = s.bsize(b1); P.k2 = s.bsize(b2);
      P.q_off = qtot; P.pole_off = pole;
      P.cg = s.sz ? 1.0 : op_cgfactor(s.qn[2 * b1 + 1], s.qn[2 * b2 + 1]);
      qtot += (int64_t)P.k1 * P.k2;
      pole += (int64_t)P.k1 * P.k2;
      kk2 = std::max(kk2, P.k1 * P.k2);
      tp.push_back(P);
    }
    if (tp.empty()) continue;
    DBuf& dab = c->buf("dmt_dAe");
    DBuf& qb = c->buf("dmt_Q");
    DBuf& dqb = c->buf("dmt_dQ");
    DBuf& clb = c->buf("dmt_cl");
    if ((rc = dab.reserve(sizeof(double) * (size_t)(T * opc + 1)))) return rc;
    if ((rc = qb.reserve(sizeof(double) * (size_t)(qtot + 1)))) return rc;
    if ((rc = dqb.reserve(sizeof(double) * (size_t)(T * qtot + 1)))) return rc;
    if ((rc = clb.reserve(sizeof(double) * (size_t)(T * qtot + 1)))) return rc;
    HIPCHK(hipMemcpyAsync(dab.p, dbuf(c, s.dopbuf), sizeof(double) * (size_t)(T * opc), hipMemcpyDeviceToDevice,
                          c->stream));
    double* A = dbuf(c, s.opbuf);
    std::vector<E2Gemm> ga1, ga2, gq1, gq2, gd1, gd2, gd3, gd4, gc1, gc2;
    for (size_t u = 0; u < tp.size(); u++) {
      const DmTanPair& P = tp[u];
      const int b1 = s.op_map[2 * u], b2 = s.op_map[2 * u + 1];
      const int k1 = P.k1, k2 = P.k2;
      gq1.push_back(gemm(k1, k2, k1, boff[b1], k1, 1, P.a_off, k2, 1, P.q_off, k2));             // rho1 A
      gq2.push_back(acc(gemm(k1, k2, k2, P.a_off, k2, 1, boff[b2], k2, 1, P.q_off, k2)));        // + A rho2
      for (int t = 0; t < T; t++) {
        const int64_t m1 = (int64_t)t * btot + boff[b1], m2 = (int64_t)t * btot + boff[b2];
        const int64_t da = (int64_t)t * opc + P.a_off, dq = (int64_t)t * qtot + P.q_off;
        ga1.push_back(acc(gemm(k1, k2, k1, m1, k1, 1, P.a_off, k2, 1, da, k2)));              // dA += M1 A
        ga2.push_back(acc(gemm(k1, k2, k2, P.a_off, k2, 1, m2, k2, 1, da, k2), -1.0));        //    -= A M2
        gd1.push_back(gemm(k1, k2, k1, m1, k1, 1, P.a_off, k2, 1, dq, k2));                   // dQ = drho1 A
        gd2.push_back(acc(gemm(k1, k2, k1, boff[b1], k1, 1, da, k2, 1, dq, k2)));             //  + rho1 dA
        gd3.push_back(acc(gemm(k1, k2, k2, da, k2, 1, boff[b2], k2, 1, dq, k2)));             //  + dA rho2
        gd4.push_back(acc(gemm(k1, k2, k2, P.a_off, k2, 1, m2, k2, 1, dq, k2)));              //  + A drho2
        gc1.push_back(gemm(k1, k2, k2, P.a_off, k2, 1, m2, k2, 1, dq, k2));                   // cl = A Tc2
        gc2.push_back(acc(gemm(k1, k2, k1, m1, k1, 1, P.a_off, k2, 1, dq, k2), -1.0));        //    - Tc1 A
      }
    }
    double *M = (double*)mb.p, *Tc = (double*)tcb.p, *dRe = (double*)drb.p, *dAe = (double*)dab.p;
    if ((rc = gemm_groups(c, "dmt_a1", ga1, M, A, dAe))) return rc;
    if ((rc = gemm_groups(c, "dmt_a2", ga2, A, M, dAe))) return rc;
    if ((rc = gemm_groups(c, "dmt_q1", gq1, R, A, (double*)qb.p))) return rc;
    if ((rc = gemm_groups(c, "dmt_q2", gq2, A, R, (double*)qb.p))) return rc;
    if ((rc = gemm_groups(c, "dmt_d1", gd1, dRe, A, (double*)dqb.p))) return rc;
    if ((rc = gemm_groups(c, "dmt_d2", gd2, R, dAe, (double*)dqb.p))) return rc;
    if ((rc = gemm_groups(c, "dmt_d3", gd3, dAe, R, (double*)dqb.p))) return rc;
    if ((rc = gemm_groups(c, "dmt_d4", gd4, A, dRe, (double*)dqb.p))) return rc;
    if ((rc = gemm_groups(c, "dmt_c1", gc1, A, Tc, (double*)clb.p))) return rc;
    if ((rc = gemm_groups(c, "dmt_c2", gc2, Tc, A, (double*)clb.p))) return rc;
    DmTanPair* dtp;
    if ((rc = upload(c, "dmt_pairs", tp, &dtp))) return rc;
    HIPCHK(dm_launch_tan_weights(dtp, (int)tp.size(), kk2, T, A, dAe, opc, (double*)qb.p, (double*)dqb.p,
                                 (double*)clb.p, qtot, (double*)w1b.p, (double*)w2b.p, ptot, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  D->tpole_tot = ptot;
  D->tpoles_ready = true;
  return NRG_OK;
}

int bd_kind(const nrg_dm* D) {   // BDelta / BDeltaTemp by method (src/DM_NRG/DM_NRG.cpp:474-482)
  return D->method == 2 ? DM_BD_LOGGAUSS_GAUSS : (D->method == 1 ? DM_BD_BROAD_ANDERS : DM_BD_BROAD_LORENTZ);
}

// shell sums at (shell, omegabar); norm: CalcNorm (broadening 1, Wp + Wn below
// the last shell, the Costi sum once at the last)
int spec_tasks(nrg_dm* D, int n, const int* shells, const double* w, double* out, bool norm = false) {
  nrg_ctx* c = D->c;
  int rc;
  if ((rc = pole_layout(D))) return rc;
  const int ns = D->specNM;
  // the (shell, omegabar) sums by shell and weight array: tasks[N][sel] = slots
  struct Task { double w; int slot; };
  std::vector<std::vector<Task>> tk[2];
  tk[0].assign(ns, std::vector<Task>());
  tk[1].assign(ns, std::vector<Task>());
  for (int i = 0; i < n; i++) {
    out[i] = 0.0;
    const int N = shells[i];
    if (N < 0 || N > ns - 1) continue;
    if (norm) {   // CalcNorm: Wp + Wn below the last shell, the Costi sum once at the last
      tk[0][N].push_back(Task{1.0, i});   // (NRG_FDM_FIX_NORM: Wp = Wn, every pole once)
      if (N < ns - 1 && !(D->fix & NRG_FDM_FIX_NORM)) tk[1][N].push_back(Task{-1.0, i});
    } else {
      tk[w[i] >= 0.0 ? 0 : 1][N].push_back(Task{w[i], i});
    }
  }
  // shell groups whose poles fit the arena together, in shell order; the
  // arena's current range first when it covers every shell asked for
  std::vector<std::pair<int, int>> groups;
  {
    int lo = ns, hi = -1;
    for (int N = 0; N < ns; N++)
      if (!tk[0][N].empty() || !tk[1][N].empty()) { lo = std::min(lo, N); hi = std::max(hi, N); }
    if (hi < 0) return NRG_OK;
    if (D->ar_lo >= 0 && D->ar_lo <= lo && hi < D->ar_hi) {
      groups.emplace_back(D->ar_lo, D->ar_hi);
    } else if (D->pole_end[ns - 1] <= D->arena_max) {
      groups.emplace_back(0, ns);
    } else {
      int g0 = lo;
      while (g0 <= hi) {
        int g1 = g0 + 1;
        if (D->pole_end[g0] - D->pole_beg[g0] > D->arena_max)
          return fail(NRG_ERR_HIP, "DM-NRG: the poles of shell " + std::to_string(g0) + " do not fit in device memory");
        while (g1 <= hi && D->pole_end[g1] - D->pole_beg[g0] <= D->arena_max) g1++;
        groups.emplace_back(g0, g1);
        g0 = g1;
      }
    }
  }
  for (const auto& g : groups) {
    std::vector<DmTile> tiles;
    std::vector<double> wl;
    std::vector<int> slot;
    int64_t part = 0;
    int maxslab = 0;
    for (int N = g.first; N < g.second; N++)
      for (int sel = 0; sel < 2; sel++) {
        const std::vector<Task>& L = tk[sel][N];
        if (L.empty()) continue;
        const int64_t np = D->pole_end[N] - D->pole_beg[N];
        // slabs of at least 64 poles per lane and at most 1024 per shell
        int64_t nslab = std::max<int64_t>(1, std::min<int64_t>((np + 16383) / 16384, 1024));
        int64_t len = (np + nslab - 1) / nslab;
        len = std::max<int64_t>(256, (len + 255) / 256 * 256);
        nslab = std::max<int64_t>(1, (np + len - 1) / len);
        const int step = norm ? 1 : DM_TILE_W;
        for (size_t q0 = 0; q0 < L.size(); q0 += step) {
          DmTile T{};
          T.pole_beg = D->pole_beg[N] - D->pole_beg[g.first];
          T.pole_end = D->pole_end[N] - D->pole_beg[g.first];
          T.slab_len = len;
          T.nslab = (int)nslab;
          T.task_beg = (int)wl.size();
          T.ntask = (int)std::min<size_t>(step, L.size() - q0);
          T.part_beg = (int)part;
          T.wsel = sel;
          T.norm = norm ? 1 : 0;
          T.twin = D->p.twindow * D->temp / calc_dn(D->p.Lambda, N);
          bool pos = true, neg = true;
          for (int j = 0; j < T.ntask; j++) {
            const double x = L[q0 + j].w;
            wl.push_back(x);
            slot.push_back(L[q0 + j].slot);
            pos = pos && x > 0.0;
            neg = neg && x < 0.0;
          }
          T.sign = norm ? 0 : (pos ? 1 : (neg ? -1 : 0));
          part += (int64_t)T.ntask * nslab;
          if (part > (int64_t)1 << 30) return fail(NRG_ERR_ARG, "DM-NRG: too many omegas in one call");
          maxslab = std::max(maxslab, T.nslab);
          tiles.push_back(T);
        }
      }
    if (tiles.empty()) continue;
    if ((rc = build_pole_range(D, g.first, g.second))) return rc;
    auto t0 = std::chrono::steady_clock::now();
    DBuf& pb = c->buf("dm_partial");
    DBuf& ob = c->buf("dm_out");
    if ((rc = pb.reserve(sizeof(double) * (size_t)(part + 1)))) return rc;
    if ((rc = ob.reserve(sizeof(double) * (wl.size() + 1)))) return rc;
    DmTile* dt;
    double* dw;
    if ((rc = upload(c, "dm_tiles", tiles, &dt))) return rc;
    if ((rc = upload(c, "dm_wl", wl, &dw))) return rc;
    HIPCHK(dm_launch_broaden_tiles(dt, (int)tiles.size(), maxslab, dw, dbuf(c, D->pfx + "Wp"),
                                   dbuf(c, D->pfx + "Wn"), dbuf(c, D->pfx + "dE"), bd_kind(D), D->p.bbroad,
                                   D->p.broadtemp, (double*)pb.p, (double*)ob.p, c->stream));
    std::vector<double> res(wl.size());
    HIPCHK(hipMemcpyAsync(res.data(), ob.p, sizeof(double) * res.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < res.size(); q++) out[slot[q]] += res[q];
    D->ms_broaden += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return NRG_OK;
}

// derivative of the shell sums at (shell, omegabar): out[i * ntan + t]
int spec_tasks_jvp(nrg_dm* D, int n, const int* shells, const double* w, double* out) {
  nrg_ctx* c = D->c;
  int rc;
  const int T = D->ntan;
  if ((rc = build_poles(D))) return rc;
  if (!D->tpoles_ready) {
    auto t0 = std::chrono::steady_clock::now();
    if ((rc = build_poles_jvp(D))) return rc;
    D->ms_poles += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  auto t0 = std::chrono::steady_clock::now();
  const int BCH = 256 * 8;
  std::vector<DmTask> tasks;
  std::vector<int> slot;
  int cb = 0, maxchunk = 0;
  for (int i = 0; i < n; i++) {
    for (int t = 0; t < T; t++) out[(size_t)i * T + t] = 0.0;
    const int N = shells[i];
    if (N < 0 || N > D->specNM - 1) continue;
    DmTask Tk{};
    Tk.pole_beg = D->pole_beg[N];
    Tk.pole_end = D->pole_end[N];
    Tk.w = w[i];
    Tk.twin = D->p.twindow * D->temp / calc_dn(D->p.Lambda, N);
    Tk.nchunk = (int)((Tk.pole_end - Tk.pole_beg + BCH - 1) / BCH);
    Tk.chunk_beg = cb;
    cb += Tk.nchunk;
    maxchunk = std::max(maxchunk, Tk.nchunk);
    tasks.push_back(Tk);
    slot.push_back(i);
  }
  if (tasks.empty()) return NRG_OK;
  const int nt = (int)tasks.size();
  DBuf& pb = c->buf("dmt_partial");
  DBuf& ob = c->buf("dmt_out");
  if ((rc = pb.reserve(sizeof(double) * (size_t)(T * cb + 1)))) return rc;
  if ((rc = ob.reserve(sizeof(double) * (size_t)(T * nt + 1)))) return rc;
  DmTask* dt;
  if ((rc = upload(c, "dmt_tasks", tasks, &dt))) return rc;
  HIPCHK(dm_launch_broaden_jvp(dt, nt, maxchunk, T, dbuf(c, D->pfx + "W1"), dbuf(c, D->pfx + "W2"), D->tpole_tot,
                               dbuf(c, D->pfx + "dE"), D->p.bbroad, D->p.broadtemp, (double*)pb.p, cb, (double*)ob.p,
                               c->stream));
  std::vector<double> res((size_t)T * nt);
  HIPCHK(hipMemcpyAsync(res.data(), ob.p, sizeof(double) * res.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int q = 0; q < nt; q++)
    for (int t = 0; t < T; t++) out[(size_t)slot[q] * T + t] += res[(size_t)t * nt + q];
  D->ms_broaden += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return NRG_OK;
}

// CalcNfromOmega  src/SpecFuncClass.cpp:41-49 (z = 1)
int n_from_omega(double L, double w) {
  const double HLF = 0.5 * (1.0 + 1.0 / L);
  int i = (int)(2.0 * std::log(2.0 * std::fabs(w) / HLF) / std::log(L));
  i = 1 - i;
  return i < 0 ? 0 : i;
}

// CalcSpecDM_NRG(omega) for n omegas: every (omega, shell in its window) is one
// task of a single broadening launch; the window sums run on the host in the
// reference's shell order
int spec_omegas(nrg_dm* D, int n, const double* omega, double* out) {
  std::vector<int> sh, owner;
  std::vector<double> wb;
  for (int i = 0; i < n; i++) {
    const int No = n_from_omega(D->p.Lambda, omega[i]);
    const int N0 = (No - 8 > 0) ? No - 8 : 0;
    const int N1 = (No + 14 < D->specNM) ? No + 14 : D->specNM;
    for (int N = N0; N < N1; N++) {
      sh.push_back(N);
      wb.push_back(omega[i] / calc_dn(D->p.Lambda, N));
      owner.push_back(i);
    }
  }
  std::vector<double> val(sh.size());
  int rc;
  if ((rc = spec_tasks(D, (int)sh.size(), sh.data(), wb.data(), val.data()))) return rc;
  for (int i = 0; i < n; i++) out[i] = 0.0;
  for (size_t q = 0; q < sh.size(); q++) out[owner[q]] += val[q] / calc_dn(D->p.Lambda, sh[q]);
  return NRG_OK;
}

// linear interpolation on ascending knots (the gsl_interp_linear the reference calls)
double lin_interp(const std::vector<double>& x, const std::vector<double>& y, double t) {
  int lo = 0, hi = (int)x.size() - 1;
  while (hi > lo + 1) {
    const int m = (lo + hi) / 2;
    if (x[m] > t) hi = m; else lo = m;
  }
  const double dx = x[hi] - x[lo];
  return dx > 0.0 ? y[lo] + (t - x[lo]) * ((y[hi] - y[lo]) / dx) : y[lo];
}

}  // namespace

static int dm_fill(nrg_ctx* c, const nrg_dmnrg_params* p, const unsigned char* kept, const int64_t* kept_off,
                   int ntan, const double* dpar, nrg_dm* D);
static int dm_post(nrg_ctx* c, const nrg_dmnrg_params* p, int Nsitesmax, nrg_dm* D);

static int dm_create(nrg_ctx* c, const nrg_dmnrg_params* p, const unsigned char* kept, const int64_t* kept_off,
                     int ntan, const double* dpar, nrg_dm** out) {
  if (!c || !p || !out) return fail(NRG_ERR_ARG, "nrg_dm_create: NULL argument");
  if (ntan < 0 || ntan > NRG_MAX_TAN || (ntan > 0 && !dpar) || (kept && !kept_off))
    return fail(NRG_ERR_ARG, "nrg_dm_create_jvp: bad tangent or kept arguments");
  if (ntan > 0 && (p->method != 0 || p->update_bef_cut || !(p->Gamma > 0.0)))
    return fail(NRG_ERR_ARG, "nrg_dm_create_jvp: derivatives need the DM-NRG method, UpdateBefCut 0 and Gamma > 0");
  if (p->nsitesmax < 2 || p->ncut < 1 || p->Lambda <= 1.0 || p->betabar <= 0.0 || p->Mtemp < 1 ||
      p->method < 0 || p->method > 2)
    return fail(NRG_ERR_ARG, "nrg_dm_create: bad parameters");
  if (!(p->fdm_site_dim >= 0.0)) return fail(NRG_ERR_ARG, "nrg_dm_create: fdm_site_dim must be 0 or positive");
  if (p->fdm_fix < 0 || p->fdm_fix > NRG_FDM_EXACT || (p->fdm_fix != 0 && p->method != 2))
    return fail(NRG_ERR_ARG, "nrg_dm_create: fdm_fix takes NRG_FDM_FIX_* bits and needs FDM (method 2)");
  if (p->model < 0 || p->model > 1) return fail(NRG_ERR_ARG, "nrg_dm_create: model must be 0 (OneChQS) or 1 (TwoChQSz)");
  if (p->model == 1 && (ntan > 0 || kept))
    return fail(NRG_ERR_ARG, "nrg_dm_create: the TwoChQSz chain has no derivative pass");
  HIPCHK(hipSetDevice(c->device));
  // the handle owns its device buffers: on any error they go back with it
  std::unique_ptr<nrg_dm> D(new nrg_dm());
  const int rc = dm_fill(c, p, kept, kept_off, ntan, dpar, D.get());
  if (rc) return rc;
  *out = D.release();
  return NRG_OK;
}

static int dm_fill(nrg_ctx* c, const nrg_dmnrg_params* p, const unsigned char* kept, const int64_t* kept_off,
                   int ntan, const double* dpar, nrg_dm* D) {
  D->c = c;
  c->dms.push_back(D);
  D->pfx = "dm#" + std::to_string(c->dm_seq++) + ":";
  D->p = *p;
  D->method = p->method;
  const bool fdm = (p->method == 2);
  if (D->p.twindow <= 0) D->p.twindow = 2.0;
  if (D->p.broadtemp <= 0) D->p.broadtemp = 0.8;
  if (D->p.bbroad <= 0) D->p.bbroad = 0.5 * std::log(p->Lambda);
  nrg_onechqs_params q{};
  q.U = p->U; q.Gamma = p->Gamma; q.ed = p->ed; q.Lambda = p->Lambda; q.nsitesmax = p->nsitesmax;
  q.ncut = p->ncut; q.calcdens = 1; q.disc_scheme = p->disc_scheme; q.myeps = p->myeps; q.betabar = p->betabar_nrg;
  int rc;
  auto t0 = std::chrono::steady_clock::now();
  D->ntan = ntan;
  if (p->model == 1) {
    // every state of each shell is kept for the sweep (UpdateBefCut 1 is implied)
    nrg_twochqs_params q2{};
    q2.U = p->U; q2.Gamma1 = p->Gamma; q2.Gamma2 = p->Gamma2; q2.ed = p->ed; q2.Lambda = p->Lambda;
    q2.nsitesmax = p->nsitesmax; q2.ncut = p->ncut; q2.betabar = p->betabar_nrg;
    if ((rc = run_twochqsz_dm(c, q2, D->sh))) return rc;
    for (DmShell& s : D->sh) s.sz = true;
  } else if (kept || ntan > 0) {
    if (p->update_bef_cut) return fail(NRG_ERR_ARG, "nrg_dm_create_jvp: frozen Kept sets need UpdateBefCut 0");
    if ((rc = run_onechqs_dm_jvp(c, q, D->sh, kept, kept_off, ntan, dpar))) return rc;
  } else if ((rc = run_onechqs_dm(c, q, D->sh, p->update_bef_cut ? 1 : 0))) {
    return rc;
  }
  auto t1 = std::chrono::steady_clock::now();
  // the shells' buffers become the handle's own
  for (DmShell& s : D->sh)
    for (std::string* nm : {&s.zbuf, &s.opbuf, &s.thbuf, &s.dzbuf, &s.dopbuf})
      if (!nm->empty()) {
        c->rename(*nm, D->pfx + *nm);
        *nm = D->pfx + *nm;
      }
  // the chain's work buffers are not needed any more: what the context holds
  // now (every shell's eigenvectors and operators) is the most it will hold
  c->trim("dm#");
  D->mem_peak = c->held();
  // the TwoChQSz loop diagonalises N = 0 .. nsitesmax: one shell more than NRG_main's sites
  rc = dm_post(c, p, (p->model == 1) ? (int)D->sh.size() : p->nsitesmax, D);
  auto t2 = std::chrono::steady_clock::now();
  D->ms_forward = std::chrono::duration<double, std::milli>(t1 - t0).count();
  D->ms_sweep = std::chrono::duration<double, std::milli>(t2 - t1).count();
  return rc;
}

// DM_NRG main after its shells are in place (src/DM_NRG/DM_NRG.cpp:102-434):
// temperature, FDM weights, rho_Nmax and the backward sweep over D->sh, whose
// eigenvector rows are released shell by shell
static int dm_post(nrg_ctx* c, const nrg_dmnrg_params* p, int Nsitesmax, nrg_dm* D) {
  const bool fdm = (p->method == 2);
  const int ntan = D->ntan;
  int rc;
  // temperature from Mtemp (src/DM_NRG/DM_NRG.cpp:102-118); FDM keeps every shell (:105)
  double DM;
  if (p->Mtemp < Nsitesmax - 1) {
    if (!fdm) Nsitesmax = p->Mtemp + 1;
    DM = calc_dn(p->Lambda, p->Mtemp);
    D->temp = DM / p->betabar;
  } else {
    D->temp = calc_dn(p->Lambda, Nsitesmax + 199) / p->betabar;
    DM = calc_dn(p->Lambda, Nsitesmax - 1);
  }
  D->betabar_spec = p->betabar;
  if ((int)D->sh.size() < Nsitesmax) return fail(NRG_ERR_STATE, "nrg_dm_create: fewer shells than NRG sites");
  // FDM's count of the discarded states' environment: the site dimension per
  // later site (4 for one channel, as src/DM_NRG/DM_NRG.cpp:244; 16 for two)
  const double site_dim = p->fdm_site_dim > 0.0 ? p->fdm_site_dim : ((p->model == 1) ? 16.0 : 4.0);
  D->NM = Nsitesmax - 1;
  D->specNM = Nsitesmax;
  const int NM = D->NM;
  D->rho_off.assign(NM + 1, std::vector<int64_t>());
  D->rho_doff.assign(NM + 1, std::vector<int64_t>());
  D->rho_len.assign(NM + 1, 0);
  D->compact = (ntan == 0);
  D->fix = fdm ? p->fdm_fix : 0;
  // shell N's Boltzmann factors at Temp/D_{N-1} (src/DM_NRG/DM_NRG.cpp:203-214,240,349),
  // or at Temp/D_N with NRG_FDM_FIX_DN
  const int dsh = (D->fix & NRG_FDM_FIX_DN) ? 0 : 1;
  // Z_N at Temp/D_{NM-1} (src/DM_NRG/DM_NRG.cpp:203-214); FDM: Z_tot = Z_N +
  // sum_N Z_N^disc 4^{NM-N} and w_N (:221-271)
  const double ZN = partition_func(D->sh[NM], 1.0 / (D->temp / calc_dn(p->Lambda, NM - dsh)));
  D->bb_last = 1.0 / (D->temp / calc_dn(p->Lambda, NM - dsh));
  D->wn.assign(NM + 1, 0.0);
  double ZTot = ZN;
  std::vector<double> wraw(NM + 1, 0.0);
  wraw[NM] = ZN;
  for (int N = NM - 1; N >= 0; N--) {
    const double zd = partition_func_disc(D->sh[N], 1.0 / (D->temp / calc_dn(p->Lambda, N - dsh))) * std::pow(site_dim, NM - N);
    wraw[N] = zd;
    ZTot += zd;
  }
  if (fdm) {
    for (int N = 0; N <= NM; N++) D->wn[N] = wraw[N] / ZTot;
    D->ztot = ZTot;
  }
  // rho_Nmax: diag e^{-E/TempBar}/Z with TempBar = Temp/DM of :102-118 and
  // Z = Z_N (DM-NRG, CFS) or Z_tot (FDM) (src/DM_NRG/DM_NRG.cpp:273-297);
  // NRG_FDM_FIX_RHOMAX: TempBar at the scale of Z_tot and w_Nmax
  {
    const DmShell& s = D->sh[NM];
    const double TempBar = (D->fix & NRG_FDM_FIX_RHOMAX) ? 1.0 / D->bb_last : D->temp / DM;
    const double bb = (std::fabs(TempBar) < 1e-20) ? 1.0e10 : 1.0 / TempBar;
    const double Z = fdm ? ZTot : ZN;
    std::vector<double> r;
    D->rho_off[NM].assign(s.nblocks(), 0);
    D->rho_doff[NM].assign(s.nblocks(), 0);
    for (int b = 0; b < s.nblocks(); b++) {
      const int k = s.bsize(b), i0 = s.bbe[2 * b];
      int kc = k;
      if (D->compact && (rc = kept_prefix(s, b, kc))) return rc;
      D->rho_off[NM][b] = (int64_t)r.size();
      for (int i = 0; i < kc; i++)
        for (int j = 0; j < kc; j++) r.push_back(i == j ? std::exp(-bb * s.en[i0 + i]) / Z : 0.0);
      D->rho_doff[NM][b] = (int64_t)r.size();
      for (int i = kc; i < k; i++) r.push_back(std::exp(-bb * s.en[i0 + i]) / Z);
    }
    D->rho_len[NM] = (int64_t)r.size();
    DBuf& rb = c->buf(D->pfx + "R" + std::to_string(NM));
    if ((rc = rb.reserve(sizeof(double) * (r.size() + 1), true))) return rc;
    HIPCHK(hipMemcpyAsync(rb.p, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ntan > 0) {
      // its derivative in the kept-space frame of the chain: exp(-bb H)/Z_N with
      // H = diag(E) + th and Z_N = sum (2S+1) exp(-bz E) at the reference's other shell
      const double bz = 1.0 / (D->temp / calc_dn(p->Lambda, NM - 1));
      const int64_t tht = s.th_tot, rl = D->rho_len[NM];
      std::vector<double> th((size_t)ntan * tht + 1), dr((size_t)ntan * rl + 1, 0.0);
      HIPCHK(hipMemcpyAsync(th.data(), c->buf(s.thbuf).p, sizeof(double) * (size_t)(ntan * tht),
                            hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      for (int t = 0; t < ntan; t++) {
        double dZ = 0.0;
        for (int b = 0; b < s.nblocks(); b++) {
          const int k = s.bsize(b), i0 = s.bbe[2 * b];
          const double dg = s.deg(b);
          for (int i = 0; i < k; i++)
            dZ += dg * th[(size_t)t * tht + s.th_off[b] + (int64_t)i * k + i] * (-bz * std::exp(-bz * s.en[i0 + i]));
        }
        for (int b = 0; b < s.nblocks(); b++) {
          const int k = s.bsize(b), i0 = s.bbe[2 * b];
          for (int i = 0; i < k; i++)
            for (int j = 0; j < k; j++) {
              const double Ei = s.en[i0 + i], Ej = s.en[i0 + j];
              double v = dd_exp(bb, Ei, Ej) * th[(size_t)t * tht + s.th_off[b] + (int64_t)i * k + j] / Z;
              if (i == j) v -= std::exp(-bb * Ei) / Z * dZ / Z;
              dr[(size_t)t * rl + D->rho_off[NM][b] + (int64_t)i * k + j] = v;
            }
        }
      }
      DBuf& drb = c->buf(D->pfx + "dR" + std::to_string(NM));
      if ((rc = drb.reserve(sizeof(double) * dr.size(), true))) return rc;
      HIPCHK(hipMemcpyAsync(drb.p, dr.data(), sizeof(double) * dr.size(), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    }
  }
  // the sweep: TempBar = Temp/D_{N-1}, ZNdisc = Z_tot/4^{NM-N} (:361-366)
  // the eigenvectors of shell N+1 (and their tangents) serve the step to N
  // only, and go back to the device after it
  auto drop_z = [&](DmShell& s) {
    if (!s.zbuf.empty()) c->drop(s.zbuf);
    if (!s.dzbuf.empty()) c->drop(s.dzbuf);
    s.zbuf.clear();
    s.dzbuf.clear();
  };
  for (int N = NM - 1; N >= 0; N--) {
    if ((rc = sweep_shell(D, N, fdm, D->temp / calc_dn(p->Lambda, N - dsh), ZTot / std::pow(site_dim, NM - N)))) {
      return rc;
    }
    drop_z(D->sh[N + 1]);
  }
  // what is left is the handle's own: rho_N, the operators, the energies' tangents
  for (DmShell& s : D->sh) drop_z(s);
  c->trim("dm#");
  return NRG_OK;
}

extern "C" int nrg_dm_create(nrg_ctx* c, const nrg_dmnrg_params* p, nrg_dm** out) {
  return dm_create(c, p, nullptr, nullptr, 0, nullptr, out);
}

extern "C" int nrg_dm_create_jvp(nrg_ctx* c, const nrg_dmnrg_params* p, const unsigned char* kept,
                                 const int64_t* kept_off, int ntan, const double* dpar, nrg_dm** out) {
  return dm_create(c, p, kept, kept_off, ntan, dpar, out);
}

extern "C" int nrg_dm_spec_jvp(nrg_dm* D, int n, const int* shells, const double* omegabar, double* out) {
  if (!D || n < 0 || (n > 0 && (!shells || !omegabar || !out))) return fail(NRG_ERR_ARG, "nrg_dm_spec_jvp: bad arguments");
  if (!D->c) return fail(NRG_ERR_STATE, "nrg_dm_spec_jvp: the handle's context was destroyed");
  if (D->ntan <= 0) return fail(NRG_ERR_STATE, "nrg_dm_spec_jvp: the handle carries no tangents");
  HIPCHK(hipSetDevice(D->c->device));
  return spec_tasks_jvp(D, n, shells, omegabar, out);
}

extern "C" void nrg_dm_destroy(nrg_dm* D) { delete D; }

extern "C" int nrg_dm_info(const nrg_dm* D, int* nshell_max, int* spec_nshell_max, double* temp) {
  if (!D) return fail(NRG_ERR_ARG, "nrg_dm_info: NULL handle");
  if (nshell_max) *nshell_max = D->NM;
  if (spec_nshell_max) *spec_nshell_max = D->specNM;
  if (temp) *temp = D->temp;
  return NRG_OK;
}

extern "C" int nrg_dm_memory(const nrg_dm* D, uint64_t* before_sweep, uint64_t* now) {
  if (!D) return fail(NRG_ERR_ARG, "nrg_dm_memory: NULL handle");
  if (!D->c) return fail(NRG_ERR_STATE, "nrg_dm_memory: the handle's context was destroyed");
  if (before_sweep) *before_sweep = D->mem_peak;
  if (now) *now = D->c->held();
  return NRG_OK;
}

extern "C" int nrg_dm_timing(const nrg_dm* D, double* ms) {
  if (!D || !ms) return fail(NRG_ERR_ARG, "nrg_dm_timing: NULL argument");
  ms[0] = D->ms_forward; ms[1] = D->ms_sweep; ms[2] = D->ms_poles; ms[3] = D->ms_broaden;
  return NRG_OK;
}

extern "C" int nrg_dm_rho(nrg_dm* D, int N, double* out, int64_t cap, int* nblocks, double* trace) {
  if (!D || N < 0 || N > D->NM) return fail(NRG_ERR_ARG, "nrg_dm_rho: bad shell");
  if (!D->c) return fail(NRG_ERR_STATE, "nrg_dm_rho: the handle's context was destroyed");
  const DmShell& s = D->sh[N];
  const int nb = s.nblocks();
  if (nblocks) *nblocks = nb;
  const int64_t len = D->rho_len[N];
  std::vector<double> r(len);
  HIPCHK(hipMemcpyAsync(r.data(), D->c->buf(D->pfx + "R" + std::to_string(N)).p, sizeof(double) * len,
                        hipMemcpyDeviceToHost, D->c->stream));
  HIPCHK(hipStreamSynchronize(D->c->stream));
  // the k x k blocks (a compact rho is its kept block plus the discarded diagonal)
  int64_t full = 0;
  for (int b = 0; b < nb; b++) full += (int64_t)s.bsize(b) * s.bsize(b);
  std::vector<double> f;
  if (D->compact) {
    f.assign(full, 0.0);
    int64_t o = 0;
    for (int b = 0; b < nb; b++) {
      const int k = s.bsize(b);
      int kc = k;
      int rc = kept_prefix(s, b, kc);
      if (rc) return rc;
      for (int i = 0; i < kc; i++)
        for (int j = 0; j < kc; j++) f[o + (int64_t)i * k + j] = r[D->rho_off[N][b] + (int64_t)i * kc + j];
      for (int i = kc; i < k; i++) f[o + (int64_t)i * k + i] = r[D->rho_doff[N][b] + (i - kc)];
      o += (int64_t)k * k;
    }
  } else {
    f.swap(r);
  }
  if (trace) {
    double t = 0.0;
    int64_t o = 0;
    for (int b = 0; b < nb; b++) {
      const int k = s.bsize(b);
      for (int i = 0; i < k; i++) t += s.deg(b) * f[o + (int64_t)i * k + i];
      o += (int64_t)k * k;
    }
    *trace = t;
  }
  if (out) {
    if (cap < full) return fail(NRG_ERR_ARG, "nrg_dm_rho: output too small");
    std::copy(f.begin(), f.end(), out);
  }
  return NRG_OK;
}

extern "C" int nrg_dm_spec(nrg_dm* D, int n, const int* shells, const double* omegabar, double* out) {
  if (!D || n < 0 || (n > 0 && (!shells || !omegabar || !out))) return fail(NRG_ERR_ARG, "nrg_dm_spec: bad arguments");
  if (!D->c) return fail(NRG_ERR_STATE, "nrg_dm_spec: the handle's context was destroyed");
  HIPCHK(hipSetDevice(D->c->device));
  return spec_tasks(D, n, shells, omegabar, out);
}

extern "C" int nrg_dm_spec_omega(nrg_dm* D, int n, const double* omega, double* out) {
  if (!D || n < 0 || (n > 0 && (!omega || !out))) return fail(NRG_ERR_ARG, "nrg_dm_spec_omega: bad arguments");
  if (!D->c) return fail(NRG_ERR_STATE, "nrg_dm_spec_omega: the handle's context was destroyed");
  if (D->method == 0) return fail(NRG_ERR_ARG, "nrg_dm_spec_omega: CFS or FDM only");
  HIPCHK(hipSetDevice(D->c->device));
  return spec_omegas(D, n, omega, out);
}

extern "C" int nrg_dm_norm(nrg_dm* D, double* norm) {
  if (!D || !norm) return fail(NRG_ERR_ARG, "nrg_dm_norm: NULL argument");
  if (!D->c) return fail(NRG_ERR_STATE, "nrg_dm_norm: the handle's context was destroyed");
  if (D->method == 0) return fail(NRG_ERR_ARG, "nrg_dm_norm: CFS or FDM only");
  HIPCHK(hipSetDevice(D->c->device));
  std::vector<int> sh(D->specNM);
  std::vector<double> w(D->specNM, 0.0), val(D->specNM);
  for (int N = 0; N < D->specNM; N++) sh[N] = N;
  int rc;
  if ((rc = spec_tasks(D, D->specNM, sh.data(), w.data(), val.data(), true))) return rc;
  double s = 0.0;
  for (int N = 0; N < D->specNM; N++) s += val[N];
  *norm = s;
  return NRG_OK;
}

extern "C" int nrg_dm_weights(const nrg_dm* D, double* wn, int cap, double* ztot) {
  if (!D) return fail(NRG_ERR_ARG, "nrg_dm_weights: NULL handle");
  if (wn) {
    if (cap < D->NM + 1) return fail(NRG_ERR_ARG, "nrg_dm_weights: output too small");
    std::copy(D->wn.begin(), D->wn.end(), wn);
  }
  if (ztot) *ztot = D->ztot;
  return NRG_OK;
}

// The omega-grid drivers apart from the shell sums (CalcSpecDM_NRG_FixedOmegas
// src/SpecFuncClass.cpp:1932-2233 with NwEachN = 1, CalcSpec_ManyOmegas :2239-2341):
// the evaluation points in the reference's recording order, and the spectrum
// from the values at them.
namespace {
struct GridPoint {
  int shell;
  double wbar;   // DM-NRG: the shell sum is evaluated at this omegabar
  double w;      // recorded omega (CFS/FDM: the window sum is evaluated here)
  int parity;    // FixedOmegas: 0 even, 1 odd shells; ManyOmegas: -1
};
std::vector<GridPoint> grid_plan(double L, int nshell, bool windowed, int Nw) {
  const double fac = std::pow(L, 1.25);
  std::vector<GridPoint> g;
  if (Nw > 0) {
    // recorded w = +-Lambda^(-i/(2Nw)) Lambda^1.25 D_N; the DM-NRG shell sum at
    // +-Lambda^1.25 Lambda^(+i/(2Nw)), as the reference evaluates it
    const int N0 = windowed ? 2 : 0, Nmax = nshell - 1;
    for (int N = N0; N <= Nmax; N++)
      for (int i = 0; i < Nw; i++) {
        const double rf = (double)i / (2.0 * Nw);
        g.push_back({N, -fac * std::pow(L, rf), -std::pow(L, -rf) * fac * calc_dn(L, N), -1});
      }
    for (int N = Nmax; N >= N0; N--)
      for (int i = Nw - 1; i >= 0; i--) {
        const double rf = (double)i / (2.0 * Nw);
        g.push_back({N, fac * std::pow(L, rf), std::pow(L, -rf) * fac * calc_dn(L, N), -1});
      }
  } else {
    // even shells then odd ones: -w with N going up, +w coming back down;
    // CFS/FDM two shells later (:1954)
    for (int par = 0; par < 2; par++) {
      const int n0 = par + (windowed ? 2 : 0);
      int last = 0;
      for (int N = n0; N < nshell; N += 2) {
        g.push_back({N, -fac, -fac * calc_dn(L, N), par});
        last = N;
      }
      for (int N = last; N >= n0; N -= 2) g.push_back({N, fac, fac * calc_dn(L, N), par});
    }
  }
  return g;
}
// ManyOmegas records the values as they are; FixedOmegas interpolates the even
// and the odd set linearly on the sorted union of their omegas and averages, a
// point outside one set's range taking the other's (outside both: -100)
// (src/SpecFuncClass.cpp:2112-2222)
void grid_combine(const std::vector<GridPoint>& g, const std::vector<double>& val, int Nw, std::vector<double>& W,
                  std::vector<double>& R) {
  W.clear();
  R.clear();
  if (Nw > 0) {
    for (size_t i = 0; i < g.size(); i++) { W.push_back(g[i].w); R.push_back(val[i]); }
    return;
  }
  std::vector<double> ex, ey, ox, oy;
  for (size_t i = 0; i < g.size(); i++) {
    if (g[i].parity == 0) { ex.push_back(g[i].w); ey.push_back(val[i]); }
    else { ox.push_back(g[i].w); oy.push_back(val[i]); }
    W.push_back(g[i].w);
  }
  std::sort(W.begin(), W.end());
  for (double w : W) {
    const bool inE = !(ex.empty() || w < ex.front() || w > ex.back());
    const bool inO = !(ox.empty() || w < ox.front() || w > ox.back());
    double rE = inE ? lin_interp(ex, ey, w) : 0.0, rO = inO ? lin_interp(ox, oy, w) : 0.0;
    if (!inE) { if (!inO) { rE = -100.0; rO = -100.0; } else rE = rO; }
    else if (!inO) rO = rE;
    R.push_back(0.5 * (rE + rO));
  }
}
}  // namespace

extern "C" int nrg_spectrum_plan(double Lambda, int nshell, int windowed, int Nw, int cap, int* n, int* shell,
                                 double* omegabar, double* omega, int* parity) {
  if (!n || !(Lambda > 1.0) || nshell < 1 || Nw < 0) return fail(NRG_ERR_ARG, "nrg_spectrum_plan: bad arguments");
  const std::vector<GridPoint> g = grid_plan(Lambda, nshell, windowed != 0, Nw);
  *n = (int)g.size();
  if (!shell && !omegabar && !omega && !parity) return NRG_OK;
  if (cap < (int)g.size()) return fail(NRG_ERR_ARG, "nrg_spectrum_plan: output too small");
  for (size_t i = 0; i < g.size(); i++) {
    if (shell) shell[i] = g[i].shell;
    if (omegabar) omegabar[i] = g[i].wbar;
    if (omega) omega[i] = g[i].w;
    if (parity) parity[i] = g[i].parity;
  }
  return NRG_OK;
}

extern "C" int nrg_spectrum_combine(int n, const double* omega, const int* parity, const double* value, int Nw,
                                    double* omega_out, double* rho_out) {
  if (n < 0 || (n > 0 && (!omega || !parity || !value || !omega_out || !rho_out)) || Nw < 0)
    return fail(NRG_ERR_ARG, "nrg_spectrum_combine: bad arguments");
  std::vector<GridPoint> g(n);
  for (int i = 0; i < n; i++) g[i] = GridPoint{0, 0.0, omega[i], parity[i]};
  std::vector<double> W, R;
  grid_combine(g, std::vector<double>(value, value + n), Nw, W, R);
  std::copy(W.begin(), W.end(), omega_out);
  std::copy(R.begin(), R.end(), rho_out);
  return NRG_OK;
}

extern "C" int nrg_dm_spectrum(nrg_dm* D, double* omega, double* rho, int cap, int* nout) {
  if (!D || !nout) return fail(NRG_ERR_ARG, "nrg_dm_spectrum: NULL argument");
  if (!D->c) return fail(NRG_ERR_STATE, "nrg_dm_spectrum: the handle's context was destroyed");
  HIPCHK(hipSetDevice(D->c->device));
  const double L = D->p.Lambda;
  const bool win = D->method != 0;   // CFS/FDM: CalcSpecDM_NRG(w) at the recorded omegas
  const std::vector<GridPoint> g = grid_plan(L, D->specNM, win, D->p.Nw);
  std::vector<int> sh;
  std::vector<double> wbar, wout;
  for (const GridPoint& q : g) { sh.push_back(q.shell); wbar.push_back(q.wbar); wout.push_back(q.w); }
  std::vector<double> val(g.size());
  int rc;
  if (win) {
    if ((rc = spec_omegas(D, (int)wout.size(), wout.data(), val.data()))) return rc;
  } else {
    if ((rc = spec_tasks(D, (int)sh.size(), sh.data(), wbar.data(), val.data()))) return rc;
    for (size_t i = 0; i < sh.size(); i++) val[i] /= calc_dn(L, sh[i]);
  }
  std::vector<double> W, R;
  grid_combine(g, val, D->p.Nw, W, R);
  *nout = (int)W.size();
  if (omega || rho) {
    if (cap < (int)W.size()) return fail(NRG_ERR_ARG, "nrg_dm_spectrum: output too small");
    if (omega) std::copy(W.begin(), W.end(), omega);
    if (rho) std::copy(R.begin(), R.end(), rho);
  }
  return NRG_OK;
}

// ---------------------------------------------------------------------------
// DM_NRG over the caller's shells and the reference's .bin hand-off files
// ---------------------------------------------------------------------------

// the DM_NRG switches of p (the chain parameters are not used here)
static int dm_check_switches(const nrg_dmnrg_params* p, const char* who) {
  if (!(p->Lambda > 1.0) || !(p->betabar > 0.0) || p->Mtemp < 1 || p->method < 0 || p->method > 2 || p->Nw < 0)
    return fail(NRG_ERR_ARG, std::string(who) + ": bad parameters");
  if (!(p->fdm_site_dim >= 0.0)) return fail(NRG_ERR_ARG, std::string(who) + ": fdm_site_dim must be 0 or positive");
  if (p->fdm_fix < 0 || p->fdm_fix > NRG_FDM_EXACT || (p->fdm_fix != 0 && p->method != 2))
    return fail(NRG_ERR_ARG, std::string(who) + ": fdm_fix takes NRG_FDM_FIX_* bits and needs FDM (method 2)");
  if (p->model < 0 || p->model > 1) return fail(NRG_ERR_ARG, std::string(who) + ": model must be 0 or 1");
  return NRG_OK;
}

// inclusive [beg, end] pairs that tile [0, n) in order (each block non-empty)
static bool tiles(const int* bbe, int nb, int64_t n) {
  int64_t next = 0;
  for (int b = 0; b < nb; b++) {
    if (bbe[2 * b] != next || bbe[2 * b + 1] < bbe[2 * b]) return false;
    next = (int64_t)bbe[2 * b + 1] + 1;
  }
  return next == n;
}

// one caller shell into a DmShell; its eigenvector rows and operator blocks go
// to the device (buffers named for the handle).  Every index the device reads
// is checked here.
static int shell_from(nrg_dm* D, int N, const nrg_dm_shell& v, int prev_nstates, DmShell& s) {
  nrg_ctx* c = D->c;
  const std::string who = "nrg_dm_create_from: shell " + std::to_string(N);
  if (v.nq != 2) return fail(NRG_ERR_ARG, who + ": two quantum numbers per block are expected");
  if (v.nblocks < 1 || v.nstates < 1 || !v.qn || !v.bbe || !v.bbe_bc || !v.bbe_ev || !v.en || !v.evec)
    return fail(NRG_ERR_ARG, who + ": missing Acut arrays");
  if (!tiles(v.bbe, v.nblocks, v.nstates)) return fail(NRG_ERR_ARG, who + ": BlockBegEnd does not tile the states");
  if (v.basis_nblocks < 1 || v.basis_nstates < 1 || !v.basis_qn || !v.basis_bbe || !v.basis_itype || !v.basis_stcf)
    return fail(NRG_ERR_ARG, who + ": missing Abasis arrays");
  if (!tiles(v.basis_bbe, v.basis_nblocks, v.basis_nstates))
    return fail(NRG_ERR_ARG, who + ": Abasis BlockBegEnd does not tile its states");
  s.N = N;
  s.sz = (D->p.model == 1);
  s.qn.assign(v.qn, v.qn + 2 * v.nblocks);
  s.bbe.assign(v.bbe, v.bbe + 2 * v.nblocks);
  s.bbe_bc.assign(v.bbe_bc, v.bbe_bc + 2 * v.nblocks);
  s.en.assign(v.en, v.en + v.nstates);
  s.kept.assign(v.nstates, 1);
  if (v.kept) s.kept.assign(v.kept, v.kept + v.nstates);
  s.zoff.assign(v.nblocks, 0);
  for (int b = 0; b < v.nblocks; b++) {
    const int lo = v.bbe_bc[2 * b], hi = v.bbe_bc[2 * b + 1];
    if (lo < 0 || hi < lo || hi >= v.basis_nstates) return fail(NRG_ERR_ARG, who + ": BlockBegEndBC outside Abasis");
    const int64_t k = s.bsize(b), n = s.bsize_bc(b);
    if (k > n) return fail(NRG_ERR_ARG, who + ": a block has more states than its pre-cut basis");
    s.zoff[b] = v.bbe_ev[2 * b];
    if (s.zoff[b] < 0 || s.zoff[b] + k * n > v.evec_len)
      return fail(NRG_ERR_ARG, who + ": eigenvector rows outside evec");
  }
  s.stcf.assign(v.basis_stcf, v.basis_stcf + v.basis_nstates);
  s.itype.assign(v.basis_itype, v.basis_itype + v.basis_nstates);
  const int ntypes = D->p.model == 1 ? 16 : 4;
  for (int i = 0; i < v.basis_nstates; i++) {
    if (N > 0 && (s.stcf[i] < 0 || s.stcf[i] >= prev_nstates))
      return fail(NRG_ERR_ARG, who + ": StCameFrom outside the previous shell");
    if (s.itype[i] < 0 || s.itype[i] >= ntypes) return fail(NRG_ERR_ARG, who + ": iType outside the site");
  }
  s.bas_bbe.assign(v.basis_bbe, v.basis_bbe + 2 * v.basis_nblocks);
  s.bas_qn.assign(v.basis_qn, v.basis_qn + 2 * v.basis_nblocks);
  if (v.op_nblocks < 0 || (v.op_nblocks > 0 && (!v.op_map || !v.op_mbe || !v.op)))
    return fail(NRG_ERR_ARG, who + ": missing operator arrays");
  s.op_map.assign(v.op_map, v.op_map + 2 * v.op_nblocks);
  s.op_mbe.assign(v.op_mbe, v.op_mbe + 2 * v.op_nblocks);
  for (int q = 0; q < v.op_nblocks; q++) {
    const int b1 = s.op_map[2 * q], b2 = s.op_map[2 * q + 1];
    if (b1 < 0 || b1 >= v.nblocks || b2 < 0 || b2 >= v.nblocks) return fail(NRG_ERR_ARG, who + ": MatBlockMap outside Acut");
    const int64_t lo = s.op_mbe[2 * q], hi = s.op_mbe[2 * q + 1];
    if (lo < 0 || hi - lo + 1 != (int64_t)s.bsize(b1) * s.bsize(b2) || hi >= v.op_len)
      return fail(NRG_ERR_ARG, who + ": MatBlockBegEnd does not match the block sizes");
  }
  int rc;
  s.zbuf = D->pfx + "dmZ" + std::to_string(N);
  s.opbuf = D->pfx + "dmO" + std::to_string(N);
  DBuf& zb = c->buf(s.zbuf);
  DBuf& ob = c->buf(s.opbuf);
  if ((rc = zb.reserve(sizeof(double) * (size_t)(v.evec_len + 1), true))) return rc;
  if ((rc = ob.reserve(sizeof(double) * (size_t)(v.op_len + 1), true))) return rc;
  HIPCHK(hipMemcpy(zb.p, v.evec, sizeof(double) * (size_t)v.evec_len, hipMemcpyHostToDevice));
  if (v.op_len > 0) HIPCHK(hipMemcpy(ob.p, v.op, sizeof(double) * (size_t)v.op_len, hipMemcpyHostToDevice));
  return NRG_OK;
}

extern "C" int nrg_dm_create_from(nrg_ctx* c, const nrg_dmnrg_params* p, int nshells, const nrg_dm_shell* shells,
                                  nrg_dm** out) {
  if (!c || !p || !out || nshells < 2 || !shells) return fail(NRG_ERR_ARG, "nrg_dm_create_from: bad arguments");
  int rc;
  if ((rc = dm_check_switches(p, "nrg_dm_create_from"))) return rc;
  HIPCHK(hipSetDevice(c->device));
  std::unique_ptr<nrg_dm> D(new nrg_dm());
  D->c = c;
  c->dms.push_back(D.get());
  D->pfx = "dm#" + std::to_string(c->dm_seq++) + ":";
  D->p = *p;
  D->method = p->method;
  if (D->p.twindow <= 0) D->p.twindow = 2.0;
  if (D->p.broadtemp <= 0) D->p.broadtemp = 0.8;
  if (D->p.bbroad <= 0) D->p.bbroad = 0.5 * std::log(p->Lambda);
  auto t0 = std::chrono::steady_clock::now();
  D->sh.resize(nshells);
  for (int N = 0; N < nshells; N++)
    if ((rc = shell_from(D.get(), N, shells[N], N > 0 ? shells[N - 1].nstates : 0, D->sh[N]))) return rc;
  D->mem_peak = c->held();
  auto t1 = std::chrono::steady_clock::now();
  if ((rc = dm_post(c, p, nshells, D.get()))) return rc;
  auto t2 = std::chrono::steady_clock::now();
  D->ms_forward = std::chrono::duration<double, std::milli>(t1 - t0).count();
  D->ms_sweep = std::chrono::duration<double, std::milli>(t2 - t1).count();
  *out = D.release();
  return NRG_OK;
}

// CNRGarray's quantum-number bookkeeping for a shell's arrays
static void bin_qn_part(BinArray& a, int N, const std::vector<double>& qn, const std::vector<int>& bbe, bool sz) {
  a.nshell = N;
  a.nq = 2;
  a.qn = qn;
  a.bbe = bbe;
  a.totalS = sz ? 0 : 1;
  a.sqn.clear();
  if (!sz) a.sqn.push_back(1);   // total S is quantum number 1
}

extern "C" int nrg_save_arrays(nrg_ctx* c, const nrg_dmnrg_params* p, const char* dir, const char* id,
                               int* nshells) {
  if (!c || !p || !id || !nshells) return fail(NRG_ERR_ARG, "nrg_save_arrays: NULL argument");
  if (p->nsitesmax < 2 || p->ncut < 1 || !(p->Lambda > 1.0) || p->model < 0 || p->model > 1)
    return fail(NRG_ERR_ARG, "nrg_save_arrays: bad parameters");
  HIPCHK(hipSetDevice(c->device));
  std::vector<DmShell> sh;
  int rc;
  if (p->model == 1) {
    nrg_twochqs_params q2{};
    q2.U = p->U; q2.Gamma1 = p->Gamma; q2.Gamma2 = p->Gamma2; q2.ed = p->ed; q2.Lambda = p->Lambda;
    q2.nsitesmax = p->nsitesmax; q2.ncut = p->ncut; q2.betabar = p->betabar_nrg;
    rc = run_twochqsz_dm(c, q2, sh);
  } else {
    nrg_onechqs_params q{};
    q.U = p->U; q.Gamma = p->Gamma; q.ed = p->ed; q.Lambda = p->Lambda; q.nsitesmax = p->nsitesmax;
    q.ncut = p->ncut; q.calcdens = 1; q.disc_scheme = p->disc_scheme; q.myeps = p->myeps;
    q.betabar = p->betabar_nrg;
    rc = run_onechqs_dm(c, q, sh, p->update_bef_cut ? 1 : 0);
  }
  // the captured shells' device buffers go back whatever happens below
  struct Drop {
    nrg_ctx* c;
    std::vector<DmShell>& sh;
    ~Drop() {
      for (DmShell& s : sh)
        for (const std::string& nm : {s.zbuf, s.opbuf})
          if (!nm.empty()) c->drop(nm);
      c->trim("dm#");
    }
  } drop{c, sh};
  if (rc) return rc;
  const bool sz = (p->model == 1);
  std::string err;
  for (size_t N = 0; N < sh.size(); N++) {
    const DmShell& s = sh[N];
    const int nb = s.nblocks();
    // Acut_N: kept (or, with UpdateBefCut 1 and model 1, every) state, eigenvector rows
    BinBasis a;
    bin_qn_part(a, (int)N, s.qn, s.bbe, sz);
    a.en = s.en;
    a.kept = s.kept;
    int64_t zlen = 0;
    std::vector<int> bbe_ev(2 * nb);
    for (int b = 0; b < nb; b++) {
      const int64_t len = (int64_t)s.bsize(b) * s.bsize_bc(b);
      if (s.zoff[b] != zlen) return fail(NRG_ERR_STATE, "nrg_save_arrays: eigenvector rows are not contiguous");
      bbe_ev[2 * b] = (int)zlen;
      bbe_ev[2 * b + 1] = (int)(zlen + len - 1);
      zlen += len;
    }
    if (zlen > INT32_MAX) return fail(NRG_ERR_ARG, "nrg_save_arrays: a shell's eigenvectors exceed the int32 counts of the format");
    a.evec.resize(zlen);
    HIPCHK(hipMemcpy(a.evec.data(), c->buf(s.zbuf).p, sizeof(double) * (size_t)zlen, hipMemcpyDeviceToHost));
    a.bbe_bc = s.bbe_bc;
    a.bbe_ev = bbe_ev;
    a.stcf.resize(s.en.size());   // SetSCFfromdEn: each state's index before the cut
    for (int b = 0; b < nb; b++)
      for (int i = 0; i < s.bsize(b); i++) a.stcf[s.bbe[2 * b] + i] = s.bbe_bc[2 * b] + i;
    if (nrg::bin_write(nrg::bin_name(dir, "Acut", id, (int)N), a, err)) return fail(NRG_ERR_ARG, "nrg_save_arrays: " + err);
    // Abasis_N
    BinBasis ab;
    bin_qn_part(ab, (int)N, s.bas_qn, s.bas_bbe, sz);
    ab.itype = s.itype;
    ab.stcf = s.stcf;
    if (nrg::bin_write(nrg::bin_name(dir, "Abasis", id, (int)N), ab, err)) return fail(NRG_ERR_ARG, "nrg_save_arrays: " + err);
    // Mat0: the spectral operator over Acut_N
    BinMatrix m;
    bin_qn_part(m, (int)N, s.qn, s.bbe, sz);
    m.kept = s.kept;
    m.map = s.op_map;
    m.mbe = s.op_mbe;
    const int64_t nel = s.op_mbe.empty() ? 0 : (int64_t)s.op_mbe.back() + 1;
    m.matel.resize(nel);
    if (nel) HIPCHK(hipMemcpy(m.matel.data(), c->buf(s.opbuf).p, sizeof(double) * (size_t)nel, hipMemcpyDeviceToHost));
    m.need_old = 1;
    m.wigner_eckart_l = 0.5;
    if (nrg::bin_write(nrg::bin_name(dir, "Mat0", id, (int)N), m, err)) return fail(NRG_ERR_ARG, "nrg_save_arrays: " + err);
  }
  *nshells = (int)sh.size();
  return NRG_OK;
}

extern "C" int nrg_dm_create_from_files(nrg_ctx* c, const nrg_dmnrg_params* p, const char* dir, const char* id,
                                        int nshells, nrg_dm** out) {
  if (!c || !p || !id || !out || nshells < 2) return fail(NRG_ERR_ARG, "nrg_dm_create_from_files: bad arguments");
  std::vector<BinBasis> acut(nshells), abas(nshells);
  std::vector<BinMatrix> mat(nshells);
  std::vector<nrg_dm_shell> v(nshells);
  std::string err;
  for (int N = 0; N < nshells; N++) {
    if (nrg::bin_read(nrg::bin_name(dir, "Acut", id, N), acut[N], err) ||
        nrg::bin_read(nrg::bin_name(dir, "Abasis", id, N), abas[N], err) ||
        nrg::bin_read(nrg::bin_name(dir, "Mat0", id, N), mat[N], err))
      return fail(NRG_ERR_ARG, "nrg_dm_create_from_files: " + err);
    const BinBasis& a = acut[N];
    const BinBasis& b = abas[N];
    const BinMatrix& m = mat[N];
    if (a.nq != 2 || b.nq != 2 || a.qn.size() % 2 || b.qn.size() % 2 || a.bbe.size() != a.qn.size() ||
        b.bbe.size() != b.qn.size() || a.bbe_bc.size() != a.bbe.size() || a.bbe_ev.size() != a.bbe.size() ||
        (!a.kept.empty() && a.kept.size() != a.en.size()) || b.itype.size() != b.stcf.size() ||
        m.map.size() != m.mbe.size() || m.is_complex)
      return fail(NRG_ERR_ARG, "nrg_dm_create_from_files: shell " + std::to_string(N) +
                                   ": arrays of inconsistent sizes (or complex operators)");
    nrg_dm_shell& s = v[N];
    s.nblocks = (int)a.qn.size() / 2;
    s.nq = 2;
    s.nstates = (int)a.en.size();
    s.qn = a.qn.data(); s.bbe = a.bbe.data(); s.bbe_bc = a.bbe_bc.data(); s.bbe_ev = a.bbe_ev.data();
    s.en = a.en.data();
    s.kept = a.kept.empty() ? nullptr : a.kept.data();
    s.evec = a.evec.data(); s.evec_len = (int64_t)a.evec.size();
    s.basis_nblocks = (int)b.qn.size() / 2;
    s.basis_nstates = (int)b.stcf.size();
    s.basis_qn = b.qn.data(); s.basis_bbe = b.bbe.data(); s.basis_itype = b.itype.data();
    s.basis_stcf = b.stcf.data();
    s.op_nblocks = (int)m.map.size() / 2;
    s.op_map = m.map.data(); s.op_mbe = m.mbe.data(); s.op = m.matel.data(); s.op_len = (int64_t)m.matel.size();
  }
  return nrg_dm_create_from(c, p, nshells, v.data(), out);
}

extern "C" int nrg_dm_save_rho(nrg_dm* D, const char* dir, const char* id) {
  if (!D || !id) return fail(NRG_ERR_ARG, "nrg_dm_save_rho: NULL argument");
  if (!D->c) return fail(NRG_ERR_STATE, "nrg_dm_save_rho: the handle's context was destroyed");
  HIPCHK(hipSetDevice(D->c->device));
  std::string err;
  for (int N = 0; N <= D->NM; N++) {
    const DmShell& s = D->sh[N];
    const int nb = s.nblocks();
    int64_t full = 0;
    for (int b = 0; b < nb; b++) full += (int64_t)s.bsize(b) * s.bsize(b);
    std::vector<double> f(full + 1);
    int nbo = 0;
    int rc = nrg_dm_rho(D, N, f.data(), full + 1, &nbo, nullptr);
    if (rc) return rc;
    // CNRGmatrix with UpperTriangular blocks (the layout DM_NRG_SetRhoNmax and
    // CalcRhoN build): each diagonal block's upper triangle row by row
    BinMatrix m;
    bin_qn_part(m, N, s.qn, s.bbe, s.sz);
    m.kept = s.kept;
    m.upper = 1;
    int64_t o = 0;
    for (int b = 0; b < nb; b++) {
      const int k = s.bsize(b);
      m.map.push_back(b); m.map.push_back(b);
      m.mbe.push_back((int)m.matel.size());
      for (int i = 0; i < k; i++)
        for (int j = i; j < k; j++) m.matel.push_back(f[o + (int64_t)i * k + j]);
      m.mbe.push_back((int)m.matel.size() - 1);
      o += (int64_t)k * k;
    }
    if (nrg::bin_write(nrg::bin_name(dir, "rhoDM", id, N), m, err)) return fail(NRG_ERR_ARG, "nrg_dm_save_rho: " + err);
  }
  return NRG_OK;
}
