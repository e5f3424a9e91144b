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

#include <chrono>

namespace nrg {

// Graph mode: a launch sequence enqueued on c->stream (and on auxiliary
// streams forked from it and joined back) is captured, instantiated and
// launched once.  Uploads come before begin(), so the graph holds kernels and
// the fork/join dependencies only.  An error path ends the capture and drops
// the graph.
struct GraphScope {
  nrg_ctx* c;
  bool on = false;
  std::chrono::steady_clock::time_point t0;
  GraphScope(nrg_ctx* c_, bool want) : c(c_) {
    if (!want) return;
    t0 = std::chrono::steady_clock::now();
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess) return;
    on = true;
    c->capturing = true;
  }
  int finish() {
    if (!on) return NRG_OK;
    on = false;
    c->capturing = false;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t ge = nullptr;
    size_t nn = 0;
    HIPCHK(hipGraphGetNodes(g, nullptr, &nn));
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipError_t e2 = hipGraphDestroy(g);
    if (e != hipSuccess) return fail(NRG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    HIPCHK(e2);
    HIPCHK(hipGraphLaunch(ge, c->stream));
    c->gexec.push_back(ge);
    // graphs of earlier eigensolves have finished (each step reads its results back)
    while (c->gexec.size() > 4) {
      HIPCHK(hipGraphExecDestroy(c->gexec.front()));
      c->gexec.erase(c->gexec.begin());
    }
    c->graph_nodes += (long)nn;
    c->graph_count++;
    c->graph_host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return NRG_OK;
  }
  ~GraphScope() {
    if (!on) return;
    c->capturing = false;
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(c->stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
  }
};

static const int kLeaf = 16;   // = LEAF in nrg_eig2.hip
static const int kPanel = E2_PANEL;
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

// Stage 1 and stage 2 of the two-stage blocks tsl (enqueued on stream st):
// per panel step (column k = 16 p of every block still being reduced) the
// leaf QRs, the root (T, S R), the Y rows and U = Y T, X = A22 U (split-K
// GEMM), W, and A22 -= Y W^T + W Y^T; then the band, the bulge chase and d, e.
static int e2s_phase_a(nrg_ctx* c, const std::vector<E2Block>& blk, const std::vector<int>& tsl, const E2Block* d_blk,
                       double* dA, double* WS, hipStream_t st) {
  int rc;
  double* S = (double*)c->buf("e2s_ws").p;
  double* V2 = (double*)c->buf("e2s_v2").p;
  int* FL = (int*)c->buf("e2s_fl").p;
  int nmax = 0;
  for (int b : tsl) nmax = std::max(nmax, blk[b].n);
  std::vector<E2TsStep> steps;
  std::vector<E2TsItem> leaves, tiles;
  std::vector<E2Gemm> gp;
  std::vector<E2Tile> gt;
  struct PStep { int st0, nst, lf0, nlf, tl0, ntl, sp0, st_tiles0, nst_tiles, rp0, rt0, nrt; };
  std::vector<PStep> ps;
  for (int k = 0; k + E2_TSB + 1 < nmax; k += E2_TSB) {
    PStep P{};
    P.st0 = (int)steps.size(); P.lf0 = (int)leaves.size(); P.tl0 = (int)tiles.size();
    for (int b : tsl) {
      const int n = blk[b].n;
      if (k + E2_TSB + 1 >= n) continue;
      const int m = n - k - E2_TSB;
      const int q = std::min(E2_TS_QMAX, (m + E2_TS_LEAF - 1) / E2_TS_LEAF);
      const int si = (int)steps.size() - P.st0;
      steps.push_back(E2TsStep{b, k, m, q});
      for (int i = 0; i < q; i++) leaves.push_back(E2TsItem{si, i});
      for (int t = 0; t < (m + 63) / 64; t++) tiles.push_back(E2TsItem{si, t});
    }
    P.nst = (int)steps.size() - P.st0; P.nlf = (int)leaves.size() - P.lf0; P.ntl = (int)tiles.size() - P.tl0;
    // X_s = A22[:, chunk s] U[chunk s]
    P.sp0 = (int)gp.size(); P.st_tiles0 = (int)gt.size();
    for (int i = 0; i < P.nst; i++) {
      const E2TsStep& T = steps[P.st0 + i];
      const E2Block& B = blk[T.blk];
      const int64_t nn = B.n, r0 = T.k + E2_TSB;
      const E2TsOff o = e2_ts_off(B.s_off, B.n);
      const int nk = (T.m + E2_TS_NK - 1) / E2_TS_NK;
      for (int q = 0; q < nk; q++) {
        E2Gemm g{};
        g.dyn = -1; g.alpha = 1.0;
        g.m = T.m; g.n = E2_TSB; g.k = std::min(E2_TS_NK, T.m - q * E2_TS_NK);
        g.a_off = B.a_off + r0 * nn + r0 + (int64_t)q * E2_TS_NK; g.as0 = nn; g.as1 = 1;
        g.b_off = o.u + (int64_t)q * E2_TS_NK * E2_TSB; g.bs0 = E2_TSB; g.bs1 = 1;
        g.c_off = o.xp + (int64_t)q * nn * E2_TSB; g.cs0 = E2_TSB;
        const int pi = (int)gp.size() - P.sp0;
        gp.push_back(g);
        for (int a = 0; a < (T.m + 63) / 64; a++) gt.push_back(E2Tile{pi, a, 0, 0});
      }
    }
    P.nst_tiles = (int)gt.size() - P.st_tiles0;
    // A22 -= [Y W] [W Y]^T
    P.rp0 = (int)gp.size(); P.rt0 = (int)gt.size();
    for (int i = 0; i < P.nst; i++) {
      const E2TsStep& T = steps[P.st0 + i];
      const E2Block& B = blk[T.blk];
      const int64_t nn = B.n, r0 = T.k + E2_TSB;
      const E2TsOff o = e2_ts_off(B.s_off, B.n);
      E2Gemm g{};
      g.dyn = -1; g.alpha = -1.0; g.accumulate = 1;
      g.m = T.m; g.n = T.m; g.k = 2 * E2_TSB;
      g.a_off = o.yw; g.as0 = 2 * E2_TSB; g.as1 = 1;
      g.b_off = o.wy; g.bs0 = 1; g.bs1 = 2 * E2_TSB;
      g.c_off = B.a_off + r0 * nn + r0; g.cs0 = nn;
      const int pi = (int)gp.size() - P.rp0;
      gp.push_back(g);
      const int tt = (T.m + 63) / 64;
      for (int a = 0; a < tt; a++)
        for (int bb = 0; bb < tt; bb++) gt.push_back(E2Tile{pi, a, bb, 0});
    }
    P.nrt = (int)gt.size() - P.rt0;
    ps.push_back(P);
  }
  E2TsStep* d_st;
  E2TsItem *d_lf, *d_tl;
  E2Gemm* d_gp;
  E2Tile* d_gt;
  int* d_tsl;
  if ((rc = upload(c, "e2s_st", steps, &d_st))) return rc;
  if ((rc = upload(c, "e2s_lf", leaves, &d_lf))) return rc;
  if ((rc = upload(c, "e2s_tl", tiles, &d_tl))) return rc;
  if ((rc = upload(c, "e2s_gp", gp, &d_gp))) return rc;
  if ((rc = upload(c, "e2s_gt", gt, &d_gt))) return rc;
  if ((rc = upload(c, "e2s_tsl", tsl, &d_tsl))) return rc;
  const int nts = (int)tsl.size();
  HIPCHK(e2s_launch_init(d_blk, d_tsl, nts, nmax, WS, FL, st));
  for (const PStep& P : ps) {
    const E2TsStep* sp = d_st + P.st0;
    HIPCHK(e2s_launch_leaf(d_blk, sp, d_lf + P.lf0, P.nlf, dA, S, st));
    HIPCHK(e2s_launch_root(d_blk, sp, P.nst, dA, S, WS, st));
    HIPCHK(e2s_launch_form(d_blk, sp, d_lf + P.lf0, P.nlf, dA, S, st));
    HIPCHK(e2_launch_gemm3(d_gp + P.sp0, d_gt + P.st_tiles0, P.nst_tiles, dA, S, S, st));
    HIPCHK(e2s_launch_fin1(d_blk, sp, d_tl + P.tl0, P.ntl, S, st));
    HIPCHK(e2s_launch_fin2(d_blk, sp, d_tl + P.tl0, P.ntl, S, st));
    HIPCHK(e2_launch_gemm3(d_gp + P.rp0, d_gt + P.rt0, P.nrt, S, S, dA, st));
  }
  HIPCHK(e2s_launch_band(d_blk, d_tsl, nts, nmax, dA, WS, st));
  // the chase: G workgroups per block, every workgroup resident (cooperative launch)
  const int cap = e2s_chase_max_groups();
  for (int t0 = 0; t0 < nts; t0 += E2_SB_MAXB) {
    E2SbGrid G{};
    G.nb = std::min(E2_SB_MAXB, nts - t0);
    int nbm = 0;
    for (int i = 0; i < G.nb; i++) {
      G.blk[i] = tsl[t0 + i];
      nbm = std::max(nbm, e2_sb_nbatch(blk[tsl[t0 + i]].n));
    }
    // batches in flight: a batch trails its predecessor by about 2 P ticks and lasts n/16 + 2 P ticks
    int want = std::min(nbm, nmax / (2 * E2_SB_P * E2_TSB) + 4);
    want = std::min(want, 32);
    if (cap > 0) want = std::min(want, cap / G.nb);
    G.G = std::max(1, want);
    if (cap > 0 && G.G * G.nb > cap) return fail(NRG_ERR_HIP, "two-stage chase: too many blocks for one launch");
    HIPCHK(e2s_launch_chase(d_blk, G, WS, V2, FL, st));
  }
  HIPCHK(e2s_launch_tri(d_blk, d_tsl, nts, nmax, WS, st));
  // the blocks' abort flags, read back in eig2_finish
  HIPCHK(hipMemcpyAsync(c->e2p.abort_host, FL, sizeof(int) * (size_t)c->e2p.nflag, hipMemcpyDeviceToHost, st));
  return NRG_OK;
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
  std::vector<int> small, large, tsl;
  for (int b = 0; b < nb; b++) {
    if (blk[b].n <= 0) continue;
    if (blk[b].ts) tsl.push_back(b);
    else if (blk[b].n <= kSmall) small.push_back(b);
    else large.push_back(b);
  }
  // ---- stage 1b groups: the large blocks split into groups of about equal
  // sum n^2 (largest first, each to the lightest group); each group's chain of
  // column launches runs on a stream of its own, so that one group's
  // latency-bound launches overlap another's bandwidth-bound products
  std::sort(large.begin(), large.end(), [&](int x, int y) { return blk[x].n > blk[y].n; });
  // each large block's column chain stops at the panel boundary khand, from
  // which its trailing matrix (at most E2_RS_MAXN rows) is reduced by the
  // register-resident kernel: its last columns would otherwise be two
  // latency-bound launches each.  khand grows with n, so in a group sorted by
  // n the blocks still on the chain at column k are a prefix of the group.
  std::vector<int> khand(nb, 0);
  for (int b : large) khand[b] = kPanel * ((blk[b].n - E2_RS_MAXN + kPanel - 1) / kPanel);
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
  std::vector<std::vector<int>> tev_k(ngrp), tev_pbeg(ngrp), tev_np(ngrp);
  for (int g = 0; g < ngrp; g++) {
    const int kend_g = grp[g].empty() ? 0 : khand[grp[g][0]];
    for (int k = 0; k < kend_g; k++) {
      const int j = k % kPanel, k0 = k - j;
      const int pbase = (int)tprobs.size();
      for (int b : grp[g]) {
        const int nb_ = blk[b].n;
        if (k >= khand[b] || j != kPanel - 1) continue;   // the panel ending at khand - 1 is the last
        const int s0 = k0 + std::min(kPanel, nb_ - 2 - k0);
        const int64_t nn = nb_;
        E2Gemm gm{};
        gm.m = nb_ - s0; gm.n = nb_ - s0; gm.k = 2 * kPanel; gm.dyn = -1;
        gm.a_off = blk[b].p_off + s0; gm.as0 = 1; gm.as1 = nn;
        gm.b_off = blk[b].p_off + (int64_t)kPanel * nn + s0; gm.bs0 = nn; gm.bs1 = 1;
        gm.c_off = blk[b].a_off + (int64_t)s0 * nn + s0; gm.cs0 = nn;
        gm.alpha = -1.0; gm.accumulate = 1;
        tprobs.push_back(gm);
      }
      if ((int)tprobs.size() > pbase) {
        tev_k[g].push_back(k); tev_pbeg[g].push_back(pbase); tev_np[g].push_back((int)tprobs.size() - pbase);
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
  std::vector<E2Hand> hands;
  for (int b : glist) hands.push_back(E2Hand{b, khand[b]});
  E2Hand* d_hand;
  if ((rc = upload(c, "e2_hand", hands, &d_hand))) return rc;
  if ((rc = upload(c, "e2_leaf", leaves, &d_leaf))) return rc;
  if ((rc = upload(c, "e2_all", all_merges, &d_all))) return rc;
  if ((rc = upload(c, "e2_mbeg", mbeg, &d_mbeg))) return rc;
  if ((rc = upload(c, "e2_lvl", lvl_merges, &d_lvl))) return rc;
  if ((rc = upload(c, "e2_ms", ms0, &d_ms))) return rc;
  if ((rc = upload(c, "e2_gp", gprobs, &d_gp))) return rc;
  if ((rc = upload(c, "e2_gt", gtiles, &d_gt))) return rc;
  E2Gemm* d_tp;
  if ((rc = upload(c, "e2_tp", tprobs, &d_tp))) return rc;
  // ---- launches
  hipStream_t s = c->stream;
  GraphScope gs_(c, c->graph && tsl.empty());
  // lanes of concurrent work: the two-stage blocks' chain, the large groups'
  // column chains, the small blocks' single-workgroup reductions; lane 0 is
  // the context stream, lane i the auxiliary stream i-1
  const int lane_ts = tsl.empty() ? -1 : 0;
  const int lane_g0 = tsl.empty() ? 0 : 1;
  const int lane_small = small.empty() ? -1 : lane_g0 + ngrp;
  const int nlanes = lane_g0 + ngrp + (small.empty() ? 0 : 1);
  const int nfork = std::max(0, nlanes - 1);
  if (nfork > 0 && (rc = c->fork(nfork))) return rc;   // after everything enqueued so far on s
  auto lane_stream = [&](int l) { return l <= 0 ? s : c->aux[l - 1]; };
  HIPCHK(e2_launch_tridiag_small(d_blk, d_small, (int)small.size(), dA, WS, lane_stream(lane_small)));
  if (!tsl.empty() && (rc = e2s_phase_a(c, blk, tsl, d_blk, dA, WS, lane_stream(lane_ts)))) return rc;
  if (ngrp > 0) {
    std::vector<hipStream_t> gs(ngrp, s);
    for (int g = 0; g < ngrp; g++) gs[g] = lane_stream(lane_g0 + g);
    std::vector<std::vector<int>> gn(ngrp);
    std::vector<int> nmax_g(ngrp), kend_g(ngrp);
    std::vector<size_t> ev(ngrp, 0);
    int kmax = 0;
    for (int g = 0; g < ngrp; g++) {
      for (int b : grp[g]) gn[g].push_back(blk[b].n);
      nmax_g[g] = gn[g][0];
      kend_g[g] = khand[grp[g][0]];
      kmax = std::max(kmax, kend_g[g]);
      HIPCHK(e2_launch_tdb_norm(d_blk, d_large + goff[g], (int)grp[g].size(), nmax_g[g], 0, dA, WS, gs[g]));
    }
    // blocks of group g still on the chain at column k: a prefix of the group
    auto live = [&](int g, int k) {
      int nl = 0;
      while (nl < (int)grp[g].size() && khand[grp[g][nl]] > k) nl++;
      return nl;
    };
    for (int k = 0; k < kmax; k++) {
      for (int g = 0; g < ngrp; g++) {
        if (k >= kend_g[g]) continue;
        const hipStream_t sg = gs[g];
        const int nl = live(g, k);
        const int* bl = grp[g].data();
        const int* bn = gn[g].data();
        if (k % 8 == 0) {
          // every 8th column is timed (the events cost more than a short launch):
          // algorithmic bytes = the lower triangle (diagonal included) of the
          // symmetric trailing matrix of every block of the group on the chain
          double by = 0.0;
          for (int q = 0; q < nl; q++) {
            const double mm = (double)(blk[bl[q]].n - k - 1);
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
          HIPCHK(e2_launch_panel_gemm(d_tp + tev_pbeg[g][e], tprobs.data() + tev_pbeg[g][e], tev_np[g][e], WS, WS, dA,
                                      sg));
          e++;
        }
        if ((k + 1) % kPanel == 0) {
          const int nl2 = live(g, k + 1);   // blocks that go on with the next panel
          if (nl2 > 0) {
            HIPCHK(e2_launch_tdb_refresh(d_blk, bl, bn, nl2, k + 1, dA, sg));
            HIPCHK(e2_launch_tdb_norm(d_blk, d_large + goff[g], nl2, nmax_g[g], k + 1, dA, WS, sg));
          }
        }
      }
    }
    for (int g = 0; g < ngrp; g++)
      HIPCHK(e2_launch_tridiag_hand(d_blk, d_hand + goff[g], (int)grp[g].size(), dA, WS, gs[g]));
  }
  // s continues after every group's chain and the small blocks
  if (nfork > 0 && (rc = c->join(nfork))) return rc;
  HIPCHK(e2_launch_dc_init(d_blk, nb, nmax, d_all, (int)all_merges.size(), d_mbeg, WS, s));
  HIPCHK(e2_launch_dc_leaf(d_blk, d_leaf, (int)leaves.size(), WS, s));
  for (int h = 0; h < nlev; h++)
    HIPCHK(e2_launch_dc_level(d_blk, d_lvl + lvl_beg[h], (int)by_h[h].size(), lvl_mmax[h], d_ms + lvl_beg[h], WS, IW,
                              d_gp + lvl_pbeg[h], d_gt + lvl_tbeg[h], lvl_tcnt[h], s));
  HIPCHK(e2_launch_output(d_blk, nb, nmax, WS, dA, dW, 1, s));
  return gs_.finish();
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
  // reflector columns: one-stage 0..n-3 (stored from column k+1 of row k);
  // two-stage 0..ncref-1 (stage-1 panels, stored from the end of their band)
  auto nref_of = [&](int b) { return blk[b].ts ? blk[b].ncref : std::max(0, blk[b].n - 2); };
  auto voff_of = [&](int b) { return blk[b].ts ? E2_TSB : 1; };
  for (int b : act) {
    npan[b] = (nref_of(b) + kBtp - 1) / kBtp;
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
      P.nb = std::min(kBtp, nref_of(b) - P.k0);
      P.voff = voff_of(b);
      P.ns = (blk[b].n - P.k0 - P.voff + kBtKsG - 1) / kBtKsG;
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
    const int L = B.n - P.k0 - P.voff;
    for (int q = 0; q < P.ns; q++) {
      E2Gemm g{};
      g.dyn = -1; g.alpha = 1.0;
      g.m = P.nb; g.n = P.nb; g.k = std::min(kBtKsG, L - q * kBtKsG);
      g.a_off = B.a_off + (int64_t)P.k0 * nn + P.k0 + P.voff + (int64_t)q * kBtKsG; g.as0 = nn; g.as1 = 1;
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
      const int m = B.krow, L = B.n - P.k0 - P.voff, ns = (L + kBtKs - 1) / kBtKs;
      const int64_t vrow = B.a_off + (int64_t)P.k0 * nn + P.k0 + P.voff;   // V(i, j) = A[vrow + i*n + j]
      for (int q = 0; q < ns; q++) {
        E2Gemm g{};
        g.dyn = -1; g.alpha = 1.0;
        g.m = m; g.n = P.nb; g.k = std::min(kBtKs, L - q * kBtKs);
        g.a_off = B.q_off + P.k0 + P.voff + (int64_t)q * kBtKs; g.as0 = nn; g.as1 = 1;
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
      gz.c_off = B.q_off + P.k0 + P.voff; gz.cs0 = nn;
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
  bool any_ts = false;
  for (int b : act) any_ts = any_ts || blk[b].ts;
  GraphScope gs_(c, c->graph && !any_ts);
  // the bulge-chase reflectors of the two-stage blocks first (x = Q1 Q2 z)
  {
    std::vector<int> tsa;
    for (int b : act)
      if (blk[b].ts) tsa.push_back(b);
    const double* V2 = tsa.empty() ? nullptr : (const double*)c->buf("e2s_v2").p;
    for (size_t t0 = 0; t0 < tsa.size(); t0 += E2_SB_MAXB) {
      const int nbl = (int)std::min<size_t>(E2_SB_MAXB, tsa.size() - t0);
      // group g (sweeps 16g..16g+15) at position j runs at tick (ng-1-g) + j
      int nt = 0;
      for (int i = 0; i < nbl; i++) {
        const int n = blk[tsa[t0 + i]].n, ng = (n - 2 + E2_TSB - 1) / E2_TSB;
        for (int g = 0; g < ng; g++) nt = std::max(nt, (ng - 1 - g) + (n - 1 - E2_TSB * g + E2_TSB - 1) / E2_TSB);
      }
      for (int t = 0; t < nt; t++) {
        E2Q2Grid G{};
        G.nb = nbl;
        G.tick = t;
        G.cum[0] = 0;
        for (int i = 0; i < nbl; i++) {
          const E2Block& B = blk[tsa[t0 + i]];
          const int n = B.n, ng = (n - 2 + E2_TSB - 1) / E2_TSB;
          int glo = -1, ghi = -2;
          for (int g = std::max(0, ng - 1 - t); g < ng; g++) {
            const int j = t - (ng - 1 - g);
            if (E2_TSB * g + 1 + j * E2_TSB >= n) break;   // r0 past the end: larger g only larger r0
            if (glo < 0) glo = g;
            ghi = g;
          }
          G.blk[i] = tsa[t0 + i];
          G.glo[i] = glo < 0 ? 0 : glo;
          G.nch[i] = 1;   // one workgroup per group covers every kept row
          G.cum[i + 1] = G.cum[i] + (glo < 0 ? 0 : (ghi - glo + 1));
          (void)B;
        }
        HIPCHK(e2s_launch_q2(d_blk, G, WS, V2, s));
      }
    }
  }
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
  return gs_.finish();
}

// Phase A of every block (batched, or one block per launch sequence in plain
// mode); the blocks wait in the context for eig2_finish.
int eig2_begin(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA, const int64_t* w_off, double* dW) {
  int rc;
  c->e2p.pending = false;
  if (nb <= 0) return NRG_OK;
  int64_t ws = 0, iw = 0;
  e2_layout(nb, n, a_off, w_off, c->e2p.blk, ws, iw);
  // two-stage blocks: n >= ts_min (0: none), within the TSQR limits
  int64_t sws = 0, sv2 = 0, sfl = 0;
  c->e2p.ts_foff.clear();
  for (E2Block& B : c->e2p.blk) {
    B.ts = 0;
    if (c->ts_min <= 0 || B.n < c->ts_min || B.n < 2 * E2_TSB + 2 ||
        B.n - E2_TSB > E2_TS_QMAX * E2_TS_LR)
      continue;
    B.ts = 1;
    B.ncref = E2_TSB * ((B.n - E2_TSB - 2) / E2_TSB + 1);   // panels k = 0, 16, .. with k + 17 < n
    B.s_off = sws; sws += e2_ts_off(0, B.n).tot;
    B.v2_off = sv2; sv2 += e2_v2_len(B.n);
    B.f_off = sfl; sfl += 1 + e2_sb_nbatch(B.n);
    c->e2p.ts_foff.push_back((int)B.f_off);
  }
  DBuf& dws = c->buf("e2_ws");
  DBuf& diw = c->buf("e2_iw");
  if ((rc = dws.reserve(sizeof(double) * (size_t)(ws + 1)))) return rc;
  if ((rc = diw.reserve(sizeof(int) * (size_t)(iw + 1)))) return rc;
  c->e2p.nflag = (int)sfl;
  if (sfl > 0) {
    if ((rc = c->buf("e2s_ws").reserve(sizeof(double) * (size_t)(sws + 1)))) return rc;
    if ((rc = c->buf("e2s_v2").reserve(sizeof(double) * (size_t)(sv2 + 1)))) return rc;
    if ((rc = c->buf("e2s_fl").reserve(sizeof(int) * (size_t)(sfl + 1)))) return rc;
    if ((size_t)sfl > c->e2p.abort_cap) {
      if (c->e2p.abort_host) HIPCHK(hipHostFree(c->e2p.abort_host));
      c->e2p.abort_host = nullptr;
      HIPCHK(hipHostMalloc((void**)&c->e2p.abort_host, sizeof(int) * (size_t)sfl));
      c->e2p.abort_cap = (size_t)sfl;
    }
  }
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
  if (!c->e2p.ts_foff.empty()) {
    // a bulge chase that gave up waiting (its bounded wait ran out) left garbage
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int f : c->e2p.ts_foff)
      if (c->e2p.abort_host[f]) return fail(NRG_ERR_NUMERIC, "two-stage reduction: bulge chase aborted (wait limit)");
  }
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
