// nrg_eig2_host.cpp -- host orchestration of the batched eigensolver: builds
// the per-block workspace layout, the divide-and-conquer trees, the grouped
// GEMM descriptors of every merge level and back-transformation panel, uploads
// them once, and enqueues every launch on the context stream (no host
// synchronisation inside).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <vector>

#include "nrg_eig2.h"
#include "nrg_internal.h"

namespace nrg {

static const int kLeaf = 16;   // = LEAF in nrg_eig2.hip
static const int kPanel = 32;  // = PANEL in nrg_eig2.hip
static const int kBtp = 64;    // = BTP in nrg_eig2.hip
static const int kSmall = 256; // single-workgroup reduction up to this size
static const int kBtKs = 256;  // K chunk of the back-transformation products Z V^T
static const int kBtKsG = 512; // K chunk of the Gram products V V^T

static int64_t e2_panel_len(int64_t nn) { return 3 * (int64_t)kPanel * nn; }

// The workspace layout of every block of one eigensolve (offsets into e2_ws / e2_iw)
static void e2_layout(int nb, const int* n, const int64_t* a_off, const int64_t* w_off, std::vector<E2Block>& blk,
                      int64_t& ws, int64_t& iw) {
  blk.assign(nb, E2Block{});
  ws = 0;
  iw = 0;
  for (int b = 0; b < nb; b++) {
    const int64_t nn = n[b], n2 = nn * nn;
    E2Block& B = blk[b];
    B.n = n[b];
    B.krow = n[b];
    B.a_off = a_off[b];
    B.w_off = w_off[b];
    B.q_off = ws; ws += n2;
    B.k_off = ws; ws += n2;
    B.u_off = ws; ws += n2;
    B.dl_off = ws; ws += n2;
    B.t_off = ws; ws += n2;
    // panel scratch: the reduction's 3 x 32 panel rows
    B.p_off = ws; ws += e2_panel_len(nn);
    B.v_off = ws; ws += 20 * nn + 64;
    B.i_off = iw; iw += 8 * nn + 8;
  }
}

// Phase A on blocks blk (offsets into the whole workspace): Householder
// reduction to tridiagonal form, divide and conquer, eigenvalues to dW.  The
// eigenvector matrix of T stays in the workspace (q_off) for phase B.
static int e2_phase_a(nrg_ctx* c, const std::vector<E2Block>& blk, double* dA, double* dW, double* WS, int* IW) {
  const int nb = (int)blk.size();
  if (nb <= 0) return NRG_OK;
  int rc;
  int nmax = 0;
  for (const E2Block& B : blk) nmax = std::max(nmax, B.n);
  // ---- stage 1 lists
  std::vector<int> small, large;
  for (int b = 0; b < nb; b++) {
    if (blk[b].n <= 0) continue;
    if (blk[b].n <= kSmall) small.push_back(b);
    else large.push_back(b);
  }
  // ---- stage 1b groups: the large blocks split into groups of about equal
  // sum n^2 (largest first, each to the lightest group); each group's chain of
  // column launches runs on a stream of its own, so that one group's
  // latency-bound launches overlap another's bandwidth-bound products
  std::sort(large.begin(), large.end(), [&](int x, int y) { return blk[x].n > blk[y].n; });
  const int ngrp = std::min<int>(c->plain ? 1 : c->red_streams, (int)large.size());
  std::vector<std::vector<int>> grp(ngrp);
  {
    std::vector<double> load(ngrp, 0.0);
    for (int b : large) {
      const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      grp[g].push_back(b);
      load[g] += (double)blk[b].n * blk[b].n;
    }
  }
  // ---- stage 1b panel GEMMs: A[s:,s:] -= [V W][W V]^T at each panel end of each large block,
  // on the 64 x 64 tiles that hold the lower triangle or lie within 127 of the
  // diagonal: the product y = A v reads nothing else, and the rows of the next
  // panel are rebuilt from their columns before it starts.  Per group: the
  // launch at column k covers the group's blocks whose panel ends there.
  std::vector<E2Gemm> tprobs;
  std::vector<E2Tile> ttiles;
  std::vector<std::vector<int>> tev_k(ngrp), tev_pbeg(ngrp), tev_tbeg(ngrp), tev_tcnt(ngrp);
  for (int g = 0; g < ngrp; g++) {
    const int nmax_g = grp[g].empty() ? 0 : blk[grp[g][0]].n;
    for (int k = 0; k + 2 < nmax_g; k++) {
      const int j = k % kPanel, k0 = k - j;
      const int pbase = (int)tprobs.size(), tbase = (int)ttiles.size();
      for (int b : grp[g]) {
        const int nb_ = blk[b].n;
        if (k + 2 >= nb_ || !(j == kPanel - 1 || k + 3 >= nb_)) continue;
        const int s0 = k0 + std::min(kPanel, nb_ - 2 - k0);
        const int64_t nn = nb_;
        E2Gemm gm{};
        gm.m = nb_ - s0; gm.n = nb_ - s0; gm.k = 2 * kPanel; gm.dyn = -1;
        gm.a_off = blk[b].p_off + s0; gm.as0 = 1; gm.as1 = nn;
        gm.b_off = blk[b].p_off + (int64_t)kPanel * nn + s0; gm.bs0 = nn; gm.bs1 = 1;
        gm.c_off = blk[b].a_off + (int64_t)s0 * nn + s0; gm.cs0 = nn;
        gm.alpha = -1.0; gm.accumulate = 1;
        const int pi = (int)tprobs.size() - pbase;
        tprobs.push_back(gm);
        const int tt = (gm.m + 63) / 64;
        for (int a = 0; a < tt; a++)
          for (int bb = 0; bb < tt && bb <= a + 2; bb++) ttiles.push_back(E2Tile{pi, a, bb, 0});
      }
      if ((int)tprobs.size() > pbase) {
        tev_k[g].push_back(k); tev_pbeg[g].push_back(pbase); tev_tbeg[g].push_back(tbase);
        tev_tcnt[g].push_back((int)ttiles.size() - tbase);
      }
    }
  }
  // ---- stage 2 trees
  std::vector<E2Leaf> leaves;
  std::vector<E2Merge> all_merges;
  std::vector<int> mbeg(nb + 1, 0);
  std::vector<std::vector<E2Merge>> by_h;
  for (int b = 0; b < nb; b++) {
    mbeg[b] = (int)all_merges.size();
    if (blk[b].n <= 0) continue;
    std::function<int(int, int)> build = [&](int lo, int hi) -> int {
      if (hi - lo <= kLeaf) { leaves.push_back(E2Leaf{b, lo, hi, 0}); return 0; }
      const int mid = lo + (hi - lo) / 2;
      const int h = std::max(build(lo, mid), build(mid, hi)) + 1;
      E2Merge M{b, lo, mid, hi};
      all_merges.push_back(M);
      if ((int)by_h.size() < h) by_h.resize(h);
      by_h[h - 1].push_back(M);
      return h;
    };
    build(0, blk[b].n);
  }
  mbeg[nb] = (int)all_merges.size();
  std::vector<E2Merge> lvl_merges;
  std::vector<int> lvl_beg, lvl_mmax;
  std::vector<E2Gemm> gprobs;
  std::vector<E2Tile> gtiles;
  std::vector<int> lvl_pbeg, lvl_tbeg, lvl_tcnt;
  for (auto& L : by_h) {
    lvl_beg.push_back((int)lvl_merges.size());
    lvl_pbeg.push_back((int)gprobs.size());
    lvl_tbeg.push_back((int)gtiles.size());
    int mm = 0, pbase = (int)gprobs.size();
    for (size_t q = 0; q < L.size(); q++) {
      const E2Merge& M = L[q];
      const E2Block& B = blk[M.blk];
      const int msz = M.hi - M.lo;
      mm = std::max(mm, msz);
      E2Gemm g{};
      g.m = msz; g.n = msz; g.k = msz; g.dyn = (int)q;
      g.a_off = B.u_off + (int64_t)M.lo * B.n; g.as0 = msz; g.as1 = 1;
      g.b_off = B.k_off + (int64_t)M.lo * B.n; g.bs0 = msz; g.bs1 = 1;
      g.c_off = B.t_off + (int64_t)M.lo * B.n; g.cs0 = msz;
      g.alpha = 1.0;
      const int pi = (int)gprobs.size() - pbase;
      gprobs.push_back(g);
      const int tt = (msz + 63) / 64;
      for (int a = 0; a < tt; a++)
        for (int bb = 0; bb < tt; bb++) gtiles.push_back(E2Tile{pi, a, bb, 0});
      lvl_merges.push_back(M);
    }
    lvl_mmax.push_back(mm);
    lvl_tcnt.push_back((int)gtiles.size() - lvl_tbeg.back());
  }
  const int nlev = (int)by_h.size();
  // ---- uploads
  E2Block* d_blk;
  int *d_small, *d_large, *d_mbeg;
  E2Leaf* d_leaf;
  E2Merge *d_all, *d_lvl;
  E2Gemm* d_gp;
  E2Tile* d_gt;
  std::vector<E2MState> ms0(lvl_merges.size() + 1);
  E2MState* d_ms;
  if ((rc = upload(c, "e2_blk", blk, &d_blk))) return rc;
  if ((rc = upload(c, "e2_small", small, &d_small))) return rc;
  std::vector<int> glist, goff(ngrp + 1, 0);
  for (int g = 0; g < ngrp; g++) {
    glist.insert(glist.end(), grp[g].begin(), grp[g].end());
    goff[g + 1] = (int)glist.size();
  }
  if ((rc = upload(c, "e2_large", glist, &d_large))) return rc;
  if ((rc = upload(c, "e2_leaf", leaves, &d_leaf))) return rc;
  if ((rc = upload(c, "e2_all", all_merges, &d_all))) return rc;
  if ((rc = upload(c, "e2_mbeg", mbeg, &d_mbeg))) return rc;
  if ((rc = upload(c, "e2_lvl", lvl_merges, &d_lvl))) return rc;
  if ((rc = upload(c, "e2_ms", ms0, &d_ms))) return rc;
  if ((rc = upload(c, "e2_gp", gprobs, &d_gp))) return rc;
  if ((rc = upload(c, "e2_gt", gtiles, &d_gt))) return rc;
  E2Gemm* d_tp;
  E2Tile* d_tt;
  if ((rc = upload(c, "e2_tp", tprobs, &d_tp))) return rc;
  if ((rc = upload(c, "e2_tt", ttiles, &d_tt))) return rc;
  // ---- launches
  hipStream_t s = c->stream;
  // the small blocks' single-workgroup reductions run beside the large groups'
  // chains, on a stream of their own when there are large blocks
  const int nfork = ngrp > 0 ? (ngrp - 1) + (small.empty() ? 0 : 1) : 0;   // aux streams used below
  if (nfork > 0 && (rc = c->fork(nfork))) return rc;   // after everything enqueued so far on s
  const bool small_aside = ngrp > 0 && !small.empty();   // then aux[ngrp - 1] is the small blocks' stream
  HIPCHK(e2_launch_tridiag_small(d_blk, d_small, (int)small.size(), dA, WS, small_aside ? c->aux[ngrp - 1] : s));
  if (ngrp > 0) {
    std::vector<hipStream_t> gs(ngrp, s);
    for (int g = 1; g < ngrp; g++) gs[g] = c->aux[g - 1];
    std::vector<std::vector<int>> gn(ngrp);
    std::vector<int> nmax_g(ngrp);
    std::vector<size_t> ev(ngrp, 0);
    int kmax = 0;
    for (int g = 0; g < ngrp; g++) {
      for (int b : grp[g]) gn[g].push_back(blk[b].n);
      nmax_g[g] = gn[g][0];
      kmax = std::max(kmax, nmax_g[g]);
      HIPCHK(e2_launch_tdb_norm(d_blk, d_large + goff[g], (int)grp[g].size(), nmax_g[g], 0, dA, WS, gs[g]));
    }
    for (int k = 0; k + 2 < kmax; k++) {
      for (int g = 0; g < ngrp; g++) {
        if (k + 2 >= nmax_g[g]) continue;
        const hipStream_t sg = gs[g];
        const int nl = (int)grp[g].size();
        const int* bl = grp[g].data();
        const int* bn = gn[g].data();
        if (k % 8 == 0) {
          // every 8th column is timed (the events cost more than a short launch):
          // algorithmic bytes = the lower triangle (diagonal included) of the
          // symmetric trailing matrix of every active large block of the group
          double by = 0.0;
          for (int b : grp[g])
            if (k + 2 < blk[b].n) {
              const double mm = (double)(blk[b].n - k - 1);
              by += 8.0 * mm * (mm + 1.0) * 0.5;
            }
          PendingTimer tm;
          if ((rc = c->timer_begin(tm, NRG_K_SYMV, 0.0, by, sg))) return rc;
          HIPCHK(e2_launch_tdb_symv(d_blk, bl, bn, nl, k, dA, WS, sg));
          if ((rc = c->timer_end(tm, sg))) return rc;
        } else {
          HIPCHK(e2_launch_tdb_symv(d_blk, bl, bn, nl, k, dA, WS, sg));
        }
        HIPCHK(e2_launch_tdb_fin(d_blk, bl, bn, nl, k, dA, WS, sg));
        size_t& e = ev[g];
        if (e < tev_k[g].size() && tev_k[g][e] == k) {
          HIPCHK(e2_launch_gemm3(d_tp + tev_pbeg[g][e], d_tt + tev_tbeg[g][e], tev_tcnt[g][e], WS, WS, dA, sg));
          e++;
        }
        if ((k + 1) % kPanel == 0) {
          HIPCHK(e2_launch_tdb_refresh(d_blk, bl, bn, nl, k + 1, dA, sg));
          HIPCHK(e2_launch_tdb_norm(d_blk, d_large + goff[g], nl, nmax_g[g], k + 1, dA, WS, sg));
        }
      }
    }
    for (int g = 0; g < ngrp; g++)
      HIPCHK(e2_launch_tridiag_tail(d_blk, d_large + goff[g], (int)grp[g].size(), dA, WS, gs[g]));
  }
  // s continues after every group's chain and the small blocks
  if (nfork > 0 && (rc = c->join(nfork))) return rc;
  HIPCHK(e2_launch_dc_init(d_blk, nb, nmax, d_all, (int)all_merges.size(), d_mbeg, WS, s));
  HIPCHK(e2_launch_dc_leaf(d_blk, d_leaf, (int)leaves.size(), WS, s));
  for (int h = 0; h < nlev; h++)
    HIPCHK(e2_launch_dc_level(d_blk, d_lvl + lvl_beg[h], (int)by_h[h].size(), lvl_mmax[h], d_ms + lvl_beg[h], WS, IW,
                              d_gp + lvl_pbeg[h], d_gt + lvl_tbeg[h], lvl_tcnt[h], s));
  HIPCHK(e2_launch_output(d_blk, nb, nmax, WS, dA, dW, 1, s));
  return NRG_OK;
}

// Phase B on blocks blk: the compact-WY back-transformation of the first
// blk[b].krow eigenvector rows of T (the panel reflectors stored by phase A in
// the block's lower triangle), then those rows to dA.
static int e2_phase_b(nrg_ctx* c, const std::vector<E2Block>& blk, double* dA, double* WS) {
  int rc;
  std::vector<int> act;
  int nmax = 0;
  for (int b = 0; b < (int)blk.size(); b++)
    if (blk[b].n > 0 && blk[b].krow > 0) { act.push_back(b); nmax = std::max(nmax, blk[b].n); }
  if (act.empty()) return NRG_OK;
  // round r applies panel P_b-1-r of every block that has it; the reflectors of
  // a panel starting at k0 vanish in columns <= k0, so every product runs over
  // the L = n-k0-1 columns after it; only the krow rows are formed.  The K
  // range L of the products Z V^T and V V^T is split into chunks (partial
  // products summed afterwards in a fixed order), so that a round is not
  // carried by the few workgroups of its 64-column output.
  const int nb = (int)blk.size();
  std::vector<int> npan(nb, 0), pan_beg(nb, 0);
  int maxP = 0;
  for (int b : act) {
    npan[b] = blk[b].n > 2 ? (blk[b].n - 2 + kBtp - 1) / kBtp : 0;
    maxP = std::max(maxP, npan[b]);
  }
  // the auxiliary buffer: partial Gram matrices and T^T per panel, then the
  // partial products W_s and W2 per block
  std::vector<E2BtPanel> pans;
  int64_t aux = 0;
  for (int b : act) {
    pan_beg[b] = (int)pans.size();
    for (int p = 0; p < npan[b]; p++) {
      E2BtPanel P{};
      P.blk = b;
      P.k0 = p * kBtp;
      P.nb = std::min(kBtp, blk[b].n - 2 - P.k0);
      P.ns = (blk[b].n - P.k0 - 1 + kBtKsG - 1) / kBtKsG;
      P.g_off = aux;
      aux += (int64_t)P.ns * kBtp * kBtp;
      pans.push_back(P);
    }
  }
  for (E2BtPanel& P : pans) { P.t_off = aux; aux += kBtp * kBtp; }
  std::vector<int64_t> wp_off(nb, 0), w2_off(nb, 0);
  for (int b : act) {
    wp_off[b] = aux;
    aux += (int64_t)blk[b].krow * kBtp * ((blk[b].n - 1 + kBtKs - 1) / kBtKs);
    w2_off[b] = aux;
    aux += (int64_t)blk[b].krow * kBtp;
  }
  // Gram products of every panel: G_s = V[:, chunk s] V[:, chunk s]^T
  std::vector<E2Gemm> gprobs;
  std::vector<E2Tile> gtiles;
  for (const E2BtPanel& P : pans) {
    const E2Block& B = blk[P.blk];
    const int64_t nn = B.n;
    const int L = B.n - P.k0 - 1;
    for (int q = 0; q < P.ns; q++) {
      E2Gemm g{};
      g.dyn = -1; g.alpha = 1.0;
      g.m = P.nb; g.n = P.nb; g.k = std::min(kBtKsG, L - q * kBtKsG);
      g.a_off = B.a_off + (int64_t)P.k0 * nn + P.k0 + 1 + (int64_t)q * kBtKsG; g.as0 = nn; g.as1 = 1;
      g.b_off = g.a_off; g.bs0 = 1; g.bs1 = nn;
      g.c_off = P.g_off + (int64_t)q * kBtp * kBtp; g.cs0 = kBtp;
      gtiles.push_back(E2Tile{(int)gprobs.size(), 0, 0, 0});
      gprobs.push_back(g);
    }
  }
  // the rounds: W_s = Z[:, chunk s] V[:, chunk s]^T, W2 = (sum W_s) T^T,
  // Z[:, k0+1:] -= W2 V
  std::vector<E2Gemm> rprobs;
  std::vector<E2Tile> rtiles;
  std::vector<E2BtRows> rrows;
  std::vector<int> r_pbeg, r_wt, r_zt, r_tbeg, r_rbeg, r_nr;
  for (int r = 0; r < maxP; r++) {
    r_pbeg.push_back((int)rprobs.size());
    r_tbeg.push_back((int)rtiles.size());
    r_rbeg.push_back((int)rrows.size());
    const int pbase = (int)rprobs.size();
    std::vector<E2Tile> tw, tz;
    for (int b : act) {
      if (npan[b] <= r) continue;
      const E2Block& B = blk[b];
      const E2BtPanel& P = pans[pan_beg[b] + npan[b] - 1 - r];
      const int64_t nn = B.n;
      const int m = B.krow, L = B.n - P.k0 - 1, ns = (L + kBtKs - 1) / kBtKs;
      const int64_t vrow = B.a_off + (int64_t)P.k0 * nn + P.k0 + 1;   // V(i, j) = A[vrow + i*n + j]
      for (int q = 0; q < ns; q++) {
        E2Gemm g{};
        g.dyn = -1; g.alpha = 1.0;
        g.m = m; g.n = P.nb; g.k = std::min(kBtKs, L - q * kBtKs);
        g.a_off = B.q_off + P.k0 + 1 + (int64_t)q * kBtKs; g.as0 = nn; g.as1 = 1;
        g.b_off = vrow + (int64_t)q * kBtKs; g.bs0 = 1; g.bs1 = nn;
        g.c_off = wp_off[b] + (int64_t)q * m * kBtp; g.cs0 = kBtp;
        const int pi = (int)rprobs.size() - pbase;
        rprobs.push_back(g);
        for (int a = 0; a < (m + 63) / 64; a++) tw.push_back(E2Tile{pi, a, 0, 0});
      }
      for (int a = 0; a < (m + 63) / 64; a++) {
        E2BtRows R{};
        R.row0 = a * 64; R.krow = m; R.nb = P.nb; R.ns = ns;
        R.wp_off = wp_off[b]; R.t_off = P.t_off; R.w2_off = w2_off[b];
        rrows.push_back(R);
      }
      E2Gemm gz{};
      gz.dyn = -1; gz.alpha = -1.0; gz.accumulate = 1;
      gz.m = m; gz.n = L; gz.k = P.nb;
      gz.a_off = w2_off[b]; gz.as0 = kBtp; gz.as1 = 1;
      gz.b_off = vrow; gz.bs0 = nn; gz.bs1 = 1;
      gz.c_off = B.q_off + P.k0 + 1; gz.cs0 = nn;
      const int pz = (int)rprobs.size() - pbase;
      rprobs.push_back(gz);
      for (int a = 0; a < (m + 63) / 64; a++)
        for (int bb = 0; bb < (L + 63) / 64; bb++) tz.push_back(E2Tile{pz, a, bb, 0});
    }
    r_wt.push_back((int)tw.size());
    r_zt.push_back((int)tz.size());
    r_nr.push_back((int)rrows.size() - r_rbeg.back());
    rtiles.insert(rtiles.end(), tw.begin(), tw.end());
    rtiles.insert(rtiles.end(), tz.begin(), tz.end());
  }
  E2Block* d_blk;
  E2BtPanel* d_pan;
  E2Gemm *d_gp, *d_rp;
  E2Tile *d_gt, *d_rt;
  E2BtRows* d_rr;
  if ((rc = upload(c, "e2b_blk", blk, &d_blk))) return rc;
  if ((rc = upload(c, "e2b_pan", pans, &d_pan))) return rc;
  if ((rc = upload(c, "e2b_gp", gprobs, &d_gp))) return rc;
  if ((rc = upload(c, "e2b_gt", gtiles, &d_gt))) return rc;
  if ((rc = upload(c, "e2b_rp", rprobs, &d_rp))) return rc;
  if ((rc = upload(c, "e2b_rt", rtiles, &d_rt))) return rc;
  if ((rc = upload(c, "e2b_rr", rrows, &d_rr))) return rc;
  DBuf& daux = c->buf("e2b_aux");
  if ((rc = daux.reserve(sizeof(double) * (size_t)(aux + 1)))) return rc;
  double* AX = (double*)daux.p;
  hipStream_t s = c->stream;
  const int np = (int)pans.size();
  HIPCHK(e2_launch_bt_zero(d_blk, d_pan, np, dA, s));
  HIPCHK(e2_launch_gemm3(d_gp, d_gt, (int)gtiles.size(), dA, dA, AX, s));
  HIPCHK(e2_launch_bt_tmat(d_blk, d_pan, np, WS, AX, s));
  for (int r = 0; r < maxP; r++) {
    const E2Gemm* pr = d_rp + r_pbeg[r];
    const E2Tile* t0 = d_rt + r_tbeg[r];
    HIPCHK(e2_launch_gemm3(pr, t0, r_wt[r], WS, dA, AX, s));
    HIPCHK(e2_launch_bt_w2(d_rr + r_rbeg[r], r_nr[r], AX, s));
    HIPCHK(e2_launch_gemm3(pr, t0 + r_wt[r], r_zt[r], AX, dA, WS, s));
  }
  HIPCHK(e2_launch_output(d_blk, nb, nmax, WS, dA, nullptr, 2, s));
  return NRG_OK;
}

// Phase A of every block (batched, or one block per launch sequence in plain
// mode); the blocks wait in the context for eig2_finish.
int eig2_begin(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA, const int64_t* w_off, double* dW) {
  int rc;
  c->e2p.pending = false;
  if (nb <= 0) return NRG_OK;
  int64_t ws = 0, iw = 0;
  e2_layout(nb, n, a_off, w_off, c->e2p.blk, ws, iw);
  DBuf& dws = c->buf("e2_ws");
  DBuf& diw = c->buf("e2_iw");
  if ((rc = dws.reserve(sizeof(double) * (size_t)(ws + 1)))) return rc;
  if ((rc = diw.reserve(sizeof(int) * (size_t)(iw + 1)))) return rc;
  double* WS = (double*)dws.p;
  int* IW = (int*)diw.p;
  if (c->plain) {
    for (int b = 0; b < nb; b++)
      if ((rc = e2_phase_a(c, std::vector<E2Block>(1, c->e2p.blk[b]), dA, dW, WS, IW))) return rc;
  } else if ((rc = e2_phase_a(c, c->e2p.blk, dA, dW, WS, IW))) {
    return rc;
  }
  c->e2p.dA = dA;
  c->e2p.pending = true;
  return NRG_OK;
}

// Phase B of the pending blocks: krow[b] eigenvector rows of block b (a prefix
// in ascending eigenvalue order; NULL: every row).  Rows past krow[b] of the
// block's slice in dA are left undefined.
int eig2_finish(nrg_ctx* c, const int* krow) {
  if (!c->e2p.pending) return NRG_OK;
  c->e2p.pending = false;
  std::vector<E2Block>& blk = c->e2p.blk;
  for (size_t b = 0; b < blk.size(); b++)
    blk[b].krow = krow ? std::max(0, std::min(krow[b], blk[b].n)) : blk[b].n;
  double* WS = (double*)c->buf("e2_ws").p;
  int rc;
  if (c->plain) {
    for (size_t b = 0; b < blk.size(); b++)
      if ((rc = e2_phase_b(c, std::vector<E2Block>(1, blk[b]), c->e2p.dA, WS))) return rc;
    return NRG_OK;
  }
  return e2_phase_b(c, blk, c->e2p.dA, WS);
}

int eig2_dev(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA, const int64_t* w_off, double* dW) {
  int rc;
  if ((rc = eig2_begin(c, nb, n, a_off, dA, w_off, dW))) return rc;
  return eig2_finish(c, nullptr);
}

}  // namespace nrg

// ---------------------------------------------------------------------------
// stage outputs of the batched solver for one matrix (diagnostics)
// ---------------------------------------------------------------------------
extern "C" int nrg_eig2_stages(nrg_ctx* c, int n, const double* A, double* d_out, double* e_out, double* dd_out,
                               double* qv_out, double* w_out, double* z_out) {
  using namespace nrg;
  if (!c || n <= 0 || !A) return fail(NRG_ERR_ARG, "nrg_eig2_stages: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  DBuf& dA = c->buf("dbg_A");
  DBuf& dW = c->buf("dbg_W");
  int rc;
  if ((rc = dA.reserve(sizeof(double) * (size_t)n * n))) return rc;
  if ((rc = dW.reserve(sizeof(double) * (size_t)n))) return rc;
  HIPCHK(hipMemcpy(dA.p, A, sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice));
  int64_t a0 = 0, w0 = 0;
  if ((rc = eig2_dev(c, 1, &n, &a0, (double*)dA.p, &w0, (double*)dW.p))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  // the single block's workspace starts at offset 0: q,k,u,dl,t (5 n^2), panel, vectors
  const int64_t nn = n, q_off = 0;
  const int64_t p_off = 5 * nn * nn;
  const int64_t v_off = p_off + e2_panel_len(nn);
  double* WS = (double*)c->buf("e2_ws").p;
  if (d_out) HIPCHK(hipMemcpy(d_out, WS + v_off, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (e_out) HIPCHK(hipMemcpy(e_out, WS + v_off + 16 * nn, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (dd_out) HIPCHK(hipMemcpy(dd_out, WS + v_off + 5 * nn, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (qv_out) HIPCHK(hipMemcpy(qv_out, WS + q_off, sizeof(double) * nn * nn, hipMemcpyDeviceToHost));
  if (w_out) HIPCHK(hipMemcpy(w_out, dW.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (z_out) HIPCHK(hipMemcpy(z_out, dA.p, sizeof(double) * nn * nn, hipMemcpyDeviceToHost));
  return NRG_OK;
}
