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
    // panel scratch: back-transformation Vt, G, T^T, W, W2 (the reduction's
    // 3 x 32 panel rows fit inside)
    B.p_off = ws; ws += kBtp * nn + 2 * kBtp * kBtp + 2 * nn * kBtp;
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
  int nmax_large = 0;
  for (int b = 0; b < nb; b++) {
    if (blk[b].n <= 0) continue;
    if (blk[b].n <= kSmall) small.push_back(b);
    else { large.push_back(b); nmax_large = std::max(nmax_large, blk[b].n); }
  }
  // ---- stage 1b panel GEMMs: A[s:,s:] -= [V W][W V]^T at each panel end of each large block
  std::vector<E2Gemm> tprobs;
  std::vector<E2Tile> ttiles;
  std::vector<int> tev_k, tev_pbeg, tev_tbeg, tev_tcnt;
  for (int k = 0; k + 2 < nmax_large; k++) {
    const int j = k % kPanel, k0 = k - j;
    const int pbase = (int)tprobs.size(), tbase = (int)ttiles.size();
    for (int b : large) {
      const int nb_ = blk[b].n;
      if (k + 2 >= nb_ || !(j == kPanel - 1 || k + 3 >= nb_)) continue;
      const int s0 = k0 + std::min(kPanel, nb_ - 2 - k0);
      const int64_t nn = nb_;
      E2Gemm g{};
      g.m = nb_ - s0; g.n = nb_ - s0; g.k = 2 * kPanel; g.dyn = -1;
      g.a_off = blk[b].p_off + s0; g.as0 = 1; g.as1 = nn;
      g.b_off = blk[b].p_off + (int64_t)kPanel * nn + s0; g.bs0 = nn; g.bs1 = 1;
      g.c_off = blk[b].a_off + (int64_t)s0 * nn + s0; g.cs0 = nn;
      g.alpha = -1.0; g.accumulate = 1;
      const int pi = (int)tprobs.size() - pbase;
      tprobs.push_back(g);
      const int tt = (g.m + 63) / 64;
      for (int a = 0; a < tt; a++)
        for (int bb = 0; bb < tt; bb++) ttiles.push_back(E2Tile{pi, a, bb, 0});
    }
    if ((int)tprobs.size() > pbase) {
      tev_k.push_back(k); tev_pbeg.push_back(pbase); tev_tbeg.push_back(tbase);
      tev_tcnt.push_back((int)ttiles.size() - tbase);
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
  if ((rc = upload(c, "e2_large", large, &d_large))) return rc;
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
  HIPCHK(e2_launch_tridiag_small(d_blk, d_small, (int)small.size(), dA, WS, s));
  if (!large.empty()) {
    const int nl = (int)large.size();
    HIPCHK(e2_launch_tdb_norm(d_blk, d_large, nl, nmax_large, 0, dA, WS, s));
    size_t ev = 0;
    for (int k = 0; k + 2 < nmax_large; k++) {
      if (k % 8 == 0) {
        // every 8th column is timed (the events cost more than a short launch):
        // algorithmic bytes = the lower triangle (diagonal included) of the
        // symmetric trailing matrix of every active large block
        double by = 0.0;
        for (int b : large)
          if (k + 2 < blk[b].n) {
            const double mm = (double)(blk[b].n - k - 1);
            by += 8.0 * mm * (mm + 1.0) * 0.5;
          }
        PendingTimer tm;
        if ((rc = c->timer_begin(tm, NRG_K_SYMV, 0.0, by))) return rc;
        HIPCHK(e2_launch_tdb_symv(d_blk, d_large, nl, nmax_large, k, dA, WS, s));
        if ((rc = c->timer_end(tm))) return rc;
      } else {
        HIPCHK(e2_launch_tdb_symv(d_blk, d_large, nl, nmax_large, k, dA, WS, s));
      }
      HIPCHK(e2_launch_tdb_fin(d_blk, d_large, nl, nmax_large, k, dA, WS, s));
      if (ev < tev_k.size() && tev_k[ev] == k) {
        HIPCHK(e2_launch_gemm3(d_tp + tev_pbeg[ev], d_tt + tev_tbeg[ev], tev_tcnt[ev], WS, WS, dA, s));
        ev++;
      }
      if ((k + 1) % kPanel == 0) HIPCHK(e2_launch_tdb_norm(d_blk, d_large, nl, nmax_large, k + 1, dA, WS, s));
    }
    HIPCHK(e2_launch_tridiag_tail(d_blk, d_large, (int)large.size(), dA, WS, s));
  }
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
  // the L = n-k0-1 columns after it; only the krow rows are formed
  const int nb = (int)blk.size();
  std::vector<int> npan(nb, 0);
  int maxP = 0;
  for (int b : act) {
    npan[b] = blk[b].n > 2 ? (blk[b].n - 2 + kBtp - 1) / kBtp : 0;
    maxP = std::max(maxP, npan[b]);
  }
  std::vector<int> bt_list, bt_p0, r_lbeg, r_lcnt, r_pbeg, r_tbeg, r_ng, r_nw, r_nw2, r_nz;
  std::vector<E2Gemm> bprobs;
  std::vector<E2Tile> btiles;
  int nmax_bt = 0;
  for (int r = 0; r < maxP; r++) {
    r_lbeg.push_back((int)bt_list.size());
    r_pbeg.push_back((int)bprobs.size());
    r_tbeg.push_back((int)btiles.size());
    std::vector<int> bl;
    for (int b : act)
      if (npan[b] > r) bl.push_back(b);
    const int pbase = (int)bprobs.size();
    std::vector<E2Tile> tg, tw, tw2, tz;
    for (int b : bl) {
      const E2Block& B = blk[b];
      const int64_t nn = B.n;
      const int m = B.krow;
      const int k0 = (npan[b] - 1 - r) * kBtp;
      const int L = B.n - k0 - 1;
      bt_list.push_back(b);
      bt_p0.push_back(k0);
      nmax_bt = std::max(nmax_bt, B.n);
      const int64_t vt = B.p_off, G = vt + kBtp * nn, Tg = G + kBtp * kBtp, W = Tg + kBtp * kBtp,
                    W2 = W + nn * kBtp;
      E2Gemm g{};
      g.dyn = -1; g.alpha = 1.0;
      // Gram G = Vt Vt^T over columns k0+1..n-1
      E2Gemm gg = g;
      gg.m = kBtp; gg.n = kBtp; gg.k = L;
      gg.a_off = vt + k0 + 1; gg.as0 = nn; gg.as1 = 1; gg.b_off = vt + k0 + 1; gg.bs0 = 1; gg.bs1 = nn;
      gg.c_off = G; gg.cs0 = kBtp;
      int pg = (int)bprobs.size() - pbase;
      bprobs.push_back(gg);
      tg.push_back(E2Tile{pg, 0, 0, 0});
      // W = Qv[:m, k0+1:] Vt[:, k0+1:]^T
      E2Gemm gw = g;
      gw.m = m; gw.n = kBtp; gw.k = L;
      gw.a_off = B.q_off + k0 + 1; gw.as0 = nn; gw.as1 = 1; gw.b_off = vt + k0 + 1; gw.bs0 = 1; gw.bs1 = nn;
      gw.c_off = W; gw.cs0 = kBtp;
      int pw = (int)bprobs.size() - pbase;
      bprobs.push_back(gw);
      for (int a = 0; a < (m + 63) / 64; a++) tw.push_back(E2Tile{pw, a, 0, 0});
      // W2 = W T^T
      E2Gemm g2 = g;
      g2.m = m; g2.n = kBtp; g2.k = kBtp;
      g2.a_off = W; g2.as0 = kBtp; g2.as1 = 1; g2.b_off = Tg; g2.bs0 = kBtp; g2.bs1 = 1; g2.c_off = W2;
      g2.cs0 = kBtp;
      int p2 = (int)bprobs.size() - pbase;
      bprobs.push_back(g2);
      for (int a = 0; a < (m + 63) / 64; a++) tw2.push_back(E2Tile{p2, a, 0, 0});
      // Qv[:m, k0+1:] -= W2 Vt[:, k0+1:]
      E2Gemm gz = g;
      gz.m = m; gz.n = L; gz.k = kBtp;
      gz.a_off = W2; gz.as0 = kBtp; gz.as1 = 1; gz.b_off = vt + k0 + 1; gz.bs0 = nn; gz.bs1 = 1;
      gz.c_off = B.q_off + k0 + 1; gz.cs0 = nn; gz.alpha = -1.0; gz.accumulate = 1;
      int pz = (int)bprobs.size() - pbase;
      bprobs.push_back(gz);
      for (int a = 0; a < (m + 63) / 64; a++)
        for (int bb = 0; bb < (L + 63) / 64; bb++) tz.push_back(E2Tile{pz, a, bb, 0});
    }
    r_lcnt.push_back((int)bl.size());
    r_ng.push_back((int)tg.size()); r_nw.push_back((int)tw.size());
    r_nw2.push_back((int)tw2.size()); r_nz.push_back((int)tz.size());
    btiles.insert(btiles.end(), tg.begin(), tg.end());
    btiles.insert(btiles.end(), tw.begin(), tw.end());
    btiles.insert(btiles.end(), tw2.begin(), tw2.end());
    btiles.insert(btiles.end(), tz.begin(), tz.end());
  }
  E2Block* d_blk;
  int *d_btl, *d_btp;
  E2Gemm* d_bp;
  E2Tile* d_bt;
  if ((rc = upload(c, "e2b_blk", blk, &d_blk))) return rc;
  if ((rc = upload(c, "e2_btl", bt_list, &d_btl))) return rc;
  if ((rc = upload(c, "e2_btp", bt_p0, &d_btp))) return rc;
  if ((rc = upload(c, "e2_bp", bprobs, &d_bp))) return rc;
  if ((rc = upload(c, "e2_bt", btiles, &d_bt))) return rc;
  hipStream_t s = c->stream;
  for (int r = 0; r < maxP; r++) {
    const E2Tile* t0 = d_bt + r_tbeg[r];
    HIPCHK(e2_launch_bt_panel(d_blk, d_btl + r_lbeg[r], d_btp + r_lbeg[r], r_lcnt[r], nmax_bt, dA, WS,
                              d_bp + r_pbeg[r], t0, r_ng[r], t0 + r_ng[r], r_nw[r], r_nw2[r], r_nz[r], s));
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
  const int64_t v_off = p_off + kBtp * nn + 2 * kBtp * kBtp + 2 * nn * kBtp;
  double* WS = (double*)c->buf("e2_ws").p;
  if (d_out) HIPCHK(hipMemcpy(d_out, WS + v_off, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (e_out) HIPCHK(hipMemcpy(e_out, WS + v_off + 16 * nn, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (dd_out) HIPCHK(hipMemcpy(dd_out, WS + v_off + 5 * nn, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (qv_out) HIPCHK(hipMemcpy(qv_out, WS + q_off, sizeof(double) * nn * nn, hipMemcpyDeviceToHost));
  if (w_out) HIPCHK(hipMemcpy(w_out, dW.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (z_out) HIPCHK(hipMemcpy(z_out, dA.p, sizeof(double) * nn * nn, hipMemcpyDeviceToHost));
  return NRG_OK;
}
