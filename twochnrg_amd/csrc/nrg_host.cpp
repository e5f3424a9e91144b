// nrg_host.cpp -- host side of the MI355X NRG hot path: the C-ABI, the
// context (stream + device workspaces), exact truncation bookkeeping, and the
// kernel-level entry points.  The whole-chain driver lives in nrg_driver.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/twochnrg_amd.h"
#include "nrg_eig2.h"
#include "nrg_internal.h"

using namespace nrg;

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int nrg_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int nrg_ctx_create(int device, nrg_ctx** out) {
  if (!out) return fail(NRG_ERR_ARG, "nrg_ctx_create: out is NULL");
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0)
    return fail(NRG_ERR_HIP, "nrg_ctx_create: no HIP device available");
  if (device < 0 || device >= nd) return fail(NRG_ERR_ARG, "nrg_ctx_create: bad device ordinal");
  nrg_ctx* c = new nrg_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return fail(NRG_ERR_HIP, "nrg_ctx_create: hipSetDevice/hipStreamCreate failed");
  }
  *out = c;
  return NRG_OK;
}

extern "C" void nrg_ctx_destroy(nrg_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  nrg::dm_detach_all(c);
  c->resolve();
  for (auto e : c->event_pool) hipEventDestroy(e);
  for (auto a : c->aux) {
    hipStreamSynchronize(a);
    hipStreamDestroy(a);
  }
  for (auto e : c->sync_ev) hipEventDestroy(e);
  c->bufs.clear();
  if (c->ring) hipHostFree(c->ring);
  if (c->e2p.abort_host) hipHostFree(c->e2p.abort_host);
  for (auto g : c->gexec) hipGraphExecDestroy(g);
  hipStreamDestroy(c->stream);
  delete c;
}

DBuf::~DBuf() {
  if (p) hipFree(p);
}
int DBuf::reserve(size_t bytes, bool exact) {
  if (bytes <= cap) return NRG_OK;
  if (p) hipFree(p);
  p = nullptr;
  cap = 0;
  size_t b = exact ? bytes : bytes + bytes / 4 + 256;
  if (hipMalloc(&p, b) != hipSuccess) { p = nullptr; return fail(NRG_ERR_HIP, "hipMalloc failed (" + std::to_string(b) + " bytes)"); }
  cap = b;
  return NRG_OK;
}

hipEvent_t nrg_ctx::get_event() {
  if (!event_pool.empty()) { hipEvent_t e = event_pool.back(); event_pool.pop_back(); return e; }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int nrg_ctx::timer_begin(PendingTimer& t, int kind, double flops, double bytes, hipStream_t s) {
  t.kind = kind; t.flops = flops; t.bytes = bytes;
  t.a = t.b = nullptr;
  if (!timing || capturing) return 0;
  t.a = get_event();
  t.b = get_event();
  if (!t.a || !t.b) return fail(NRG_ERR_HIP, "hipEventCreate failed");
  HIPCHK(hipEventRecord(t.a, s ? s : stream));
  return 0;
}

int nrg_ctx::timer_end(PendingTimer& t, hipStream_t s) {
  if (!timing || !t.a) return 0;
  HIPCHK(hipEventRecord(t.b, s ? s : stream));
  pending.push_back(t);
  return 0;
}

int nrg_ctx::fork(int n) {
  while ((int)aux.size() < n) {
    hipStream_t a = nullptr;
    HIPCHK(hipStreamCreate(&a));
    aux.push_back(a);
  }
  while ((int)sync_ev.size() < n + 1) {
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    sync_ev.push_back(e);
  }
  HIPCHK(hipEventRecord(sync_ev[0], stream));
  for (int i = 0; i < n; i++) HIPCHK(hipStreamWaitEvent(aux[i], sync_ev[0], 0));
  return 0;
}

int nrg_ctx::join(int n) {
  for (int i = 0; i < n; i++) {
    HIPCHK(hipEventRecord(sync_ev[i + 1], aux[i]));
    HIPCHK(hipStreamWaitEvent(stream, sync_ev[i + 1], 0));
  }
  return 0;
}

int nrg_ctx::resolve() {
  for (auto& t : pending) {
    HIPCHK(hipEventSynchronize(t.b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
    KStat& k = stat[t.kind];
    k.ms += ms; k.launches++; k.flops += t.flops; k.bytes += t.bytes;
    event_pool.push_back(t.a);
    event_pool.push_back(t.b);
  }
  pending.clear();
  return 0;
}

extern "C" int nrg_set_eig_impl(nrg_ctx* c, int impl) {
  if (!c || (impl != 1 && impl != 2)) return fail(NRG_ERR_ARG, "nrg_set_eig_impl: impl must be 1 or 2");
  c->eig_impl = impl;
  return NRG_OK;
}

extern "C" int nrg_set_plain(nrg_ctx* c, int on) {
  if (!c) return fail(NRG_ERR_ARG, "nrg_set_plain: NULL ctx");
  c->plain = on != 0;
  return NRG_OK;
}

extern "C" int nrg_set_red_streams(nrg_ctx* c, int n) {
  if (!c) return fail(NRG_ERR_ARG, "nrg_set_red_streams: NULL ctx");
  if (n < 1 || n > 4) return fail(NRG_ERR_ARG, "nrg_set_red_streams: n must be 1..4");
  c->red_streams = n;
  return NRG_OK;
}

extern "C" int nrg_set_graph(nrg_ctx* c, int on) {
  if (!c) return fail(NRG_ERR_ARG, "nrg_set_graph: NULL ctx");
  c->graph = on != 0;
  return NRG_OK;
}

extern "C" int nrg_graph_stats(nrg_ctx* c, double* host_ms, long* nodes, long* graphs) {
  if (!c) return fail(NRG_ERR_ARG, "nrg_graph_stats: NULL ctx");
  if (host_ms) *host_ms = c->graph_host_ms;
  if (nodes) *nodes = c->graph_nodes;
  if (graphs) *graphs = c->graph_count;
  c->graph_host_ms = 0.0;
  c->graph_nodes = c->graph_count = 0;
  return NRG_OK;
}

extern "C" int nrg_set_two_stage(nrg_ctx* c, int min_n) {
  if (!c) return fail(NRG_ERR_ARG, "nrg_set_two_stage: NULL ctx");
  if (min_n < 0) return fail(NRG_ERR_ARG, "nrg_set_two_stage: min_n must be >= 0");
  c->ts_min = min_n;
  return NRG_OK;
}

extern "C" int nrg_set_timing(nrg_ctx* c, int on) {
  if (!c) return fail(NRG_ERR_ARG, "nrg_set_timing: NULL ctx");
  c->timing = on != 0;
  return NRG_OK;
}

extern "C" int nrg_get_stats(nrg_ctx* c, nrg_stats* out) {
  if (!c || !out) return fail(NRG_ERR_ARG, "nrg_get_stats: NULL argument");
  int rc = c->resolve();
  if (rc) return rc;
  for (int k = 0; k < NRG_NKSTAT; k++) {
    out->ms[k] = c->stat[k].ms; out->launches[k] = c->stat[k].launches;
    out->flops[k] = c->stat[k].flops; out->bytes[k] = c->stat[k].bytes;
  }
  return NRG_OK;
}

extern "C" int nrg_reset_stats(nrg_ctx* c) {
  if (!c) return fail(NRG_ERR_ARG, "nrg_reset_stats: NULL ctx");
  int rc = c->resolve();
  if (rc) return rc;
  for (int k = 0; k < NRG_NKSTAT; k++) c->stat[k] = KStat();
  return NRG_OK;
}

DBuf& nrg_ctx::buf(const std::string& name) {
  auto it = bufs.find(name);
  if (it == bufs.end()) it = bufs.emplace(name, std::unique_ptr<DBuf>(new DBuf())).first;
  return *it->second;
}
void nrg_ctx::drop(const std::string& name) { bufs.erase(name); }
int nrg_ctx::stage_h2d(void* dst, const void* src, size_t bytes) {
  const size_t kRing = (size_t)64 << 20;
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (need > kRing / 4) {
    // large copies: from the caller's memory, finished before it may be released
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return NRG_OK;
  }
  if (!ring) {
    void* p = nullptr;
    HIPCHK(hipHostMalloc(&p, kRing, hipHostMallocDefault));
    ring = (char*)p;
    ring_cap = kRing;
    ring_head = 0;
  }
  if (ring_head + need > ring_cap) {   // every copy out of the ring has completed after this
    HIPCHK(hipStreamSynchronize(stream));
    ring_head = 0;
  }
  std::memcpy(ring + ring_head, src, bytes);
  HIPCHK(hipMemcpyAsync(dst, ring + ring_head, bytes, hipMemcpyHostToDevice, stream));
  ring_head += need;
  return NRG_OK;
}
uint64_t nrg_ctx::held() const {
  uint64_t b = 0;
  for (const auto& x : bufs) b += x.second->cap;
  return b;
}
void nrg_ctx::trim(const std::string& keep) {
  for (auto it = bufs.begin(); it != bufs.end();)
    if (it->first.compare(0, keep.size(), keep) != 0) it = bufs.erase(it);
    else ++it;
}
void nrg_ctx::rename(const std::string& from, const std::string& to) {
  if (from == to) return;
  auto it = bufs.find(from);
  if (it == bufs.end()) return;
  std::unique_ptr<DBuf> b = std::move(it->second);
  bufs.erase(it);
  bufs[to] = std::move(b);
}

// ---------------------------------------------------------------------------
// eigensolve over blocks already resident in dA
// ---------------------------------------------------------------------------
// algorithmic work of an eigensolve: (10/3) n^3 flops; 8*(n(n+1)/2 + n^2 + n) bytes per block
static void eig_work(int nb, const int* n, double& fl, double& by) {
  fl = by = 0.0;
  for (int b = 0; b < nb; b++) {
    double x = n[b];
    fl += 10.0 / 3.0 * x * x * x;
    by += 8.0 * (x * (x + 1) / 2 + x * x + x);
  }
}

int nrg::eig_begin(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA, const int64_t* w_off,
                   double* dW) {
  if (nb <= 0) return NRG_OK;
  if (c->eig_impl != 2) return eig_dev(c, nb, n, a_off, dA, w_off, dW, nullptr);
  for (int b = 0; b < nb; b++)
    if (n[b] < 0) return fail(NRG_ERR_ARG, "negative block size");
  double fl, by;
  eig_work(nb, n, fl, by);
  PendingTimer tm;
  int rc;
  if ((rc = c->timer_begin(tm, NRG_K_EIG, fl, by))) return rc;
  if ((rc = eig2_begin(c, nb, n, a_off, dA, w_off, dW))) return rc;
  return c->timer_end(tm);
}

int nrg::eig_finish(nrg_ctx* c, const int* krow) {
  if (!c->e2p.pending) return NRG_OK;
  PendingTimer tm;
  int rc;
  if ((rc = c->timer_begin(tm, NRG_K_EIG, 0.0, 0.0))) return rc;
  if ((rc = eig2_finish(c, krow))) return rc;
  return c->timer_end(tm);
}

int nrg::eig_dev(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA, const int64_t* w_off,
                 double* dW, int* info_host) {
  if (nb <= 0) return NRG_OK;
  if (c->eig_impl == 2) {
    int rc;
    if ((rc = eig_begin(c, nb, n, a_off, dA, w_off, dW))) return rc;
    if ((rc = eig_finish(c, nullptr))) return rc;
    if (info_host) for (int b = 0; b < nb; b++) info_host[b] = 0;
    return NRG_OK;
  }
  std::vector<EigDesc> desc(nb);
  int64_t ws = 0, vs = 0;
  int nmax = 0;
  for (int b = 0; b < nb; b++) {
    if (n[b] < 0) return fail(NRG_ERR_ARG, "negative block size");
    desc[b].n = n[b];
    desc[b].a_off = a_off[b];
    desc[b].w_off = w_off[b];
    desc[b].ws_off = ws;
    ws += (int64_t)n[b] * n[b];
    nmax = std::max(nmax, n[b]);
  }
  for (int b = 0; b < nb; b++) { desc[b].vec_off = ws + vs; vs += 7 * (int64_t)n[b]; }
  std::vector<int> order(nb);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
  DBuf& dws = c->buf("eig_ws");
  DBuf& ddesc = c->buf("eig_desc");
  DBuf& dord = c->buf("eig_order");
  DBuf& dinfo = c->buf("eig_info");
  int rc;
  if ((rc = dws.reserve(sizeof(double) * (size_t)(ws + vs + 1)))) return rc;
  if ((rc = ddesc.reserve(sizeof(EigDesc) * nb))) return rc;
  if ((rc = dord.reserve(sizeof(int) * nb))) return rc;
  if ((rc = dinfo.reserve(sizeof(int) * nb))) return rc;
  HIPCHK(hipMemcpyAsync(ddesc.p, desc.data(), sizeof(EigDesc) * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dord.p, order.data(), sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double fl, by;
  eig_work(nb, n, fl, by);
  PendingTimer tm;
  if ((rc = c->timer_begin(tm, NRG_K_EIG, fl, by))) return rc;
  HIPCHK(nrgk_launch_eig((EigDesc*)ddesc.p, (int*)dord.p, nb, nmax, dA, dW, (double*)dws.p, (int*)dinfo.p,
                         c->stream));
  if ((rc = c->timer_end(tm))) return rc;
  if (info_host) {
    HIPCHK(hipMemcpyAsync(info_host, dinfo.p, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return NRG_OK;
}

extern "C" int nrg_diag_blocks_dev(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA,
                                   const int64_t* w_off, double* dW, int* info) {
  if (!c || nb < 0 || (nb > 0 && (!n || !a_off || !dA || !w_off || !dW)))
    return fail(NRG_ERR_ARG, "nrg_diag_blocks_dev: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  return eig_dev(c, nb, n, a_off, dA, w_off, dW, info);
}

extern "C" int nrg_diag_blocks(nrg_ctx* c, int nb, const int* n, const int64_t* h_off, const double* H,
                               int packed, double* W, double* Z, int* info) {
  if (!c || nb < 0 || (nb > 0 && (!n || !h_off || !H || !W || !Z)))
    return fail(NRG_ERR_ARG, "nrg_diag_blocks: bad arguments");
  if (nb == 0) return NRG_OK;
  HIPCHK(hipSetDevice(c->device));
  std::vector<int64_t> a_off(nb), w_off(nb);
  int64_t za = 0, wa = 0, hlen = 0;
  int nmax = 0;
  for (int b = 0; b < nb; b++) {
    if (n[b] < 0) return fail(NRG_ERR_ARG, "nrg_diag_blocks: negative block size");
    a_off[b] = za; w_off[b] = wa;
    za += (int64_t)n[b] * n[b]; wa += n[b];
    int64_t len = packed ? (int64_t)n[b] * (n[b] + 1) / 2 : (int64_t)n[b] * n[b];
    hlen = std::max(hlen, h_off[b] + len);
    nmax = std::max(nmax, n[b]);
  }
  DBuf& dA = c->buf("diag_A");
  DBuf& dW = c->buf("diag_W");
  int rc;
  if ((rc = dA.reserve(sizeof(double) * (size_t)(za + 1)))) return rc;
  if ((rc = dW.reserve(sizeof(double) * (size_t)(wa + 1)))) return rc;
  if (packed) {
    DBuf& dP = c->buf("diag_P");
    DBuf& dpo = c->buf("diag_poff");
    DBuf& dd = c->buf("diag_desc");
    if ((rc = dP.reserve(sizeof(double) * (size_t)(hlen + 1)))) return rc;
    if ((rc = dpo.reserve(sizeof(int64_t) * nb))) return rc;
    if ((rc = dd.reserve(sizeof(EigDesc) * nb))) return rc;
    std::vector<EigDesc> desc(nb);
    for (int b = 0; b < nb; b++) { desc[b] = EigDesc{}; desc[b].n = n[b]; desc[b].a_off = a_off[b]; }
    HIPCHK(hipMemcpyAsync(dP.p, H, sizeof(double) * hlen, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dpo.p, h_off, sizeof(int64_t) * nb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dd.p, desc.data(), sizeof(EigDesc) * nb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(nrgk_launch_unpack((EigDesc*)dd.p, (int64_t*)dpo.p, (double*)dP.p, (double*)dA.p, nb, nmax, c->stream));
  } else {
    for (int b = 0; b < nb; b++)
      HIPCHK(hipMemcpyAsync((double*)dA.p + a_off[b], H + h_off[b], sizeof(double) * (size_t)n[b] * n[b],
                            hipMemcpyHostToDevice, c->stream));
  }
  std::vector<int> inf(nb, 0);
  if ((rc = eig_dev(c, nb, n, a_off.data(), (double*)dA.p, w_off.data(), (double*)dW.p, inf.data()))) return rc;
  HIPCHK(hipMemcpyAsync(W, dW.p, sizeof(double) * wa, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(Z, dA.p, sizeof(double) * za, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (info) std::memcpy(info, inf.data(), sizeof(int) * nb);
  return NRG_OK;
}

extern "C" int nrg_compact(nrg_ctx* c, int nb, const int* n, const int64_t* z_off, const int* keep, const double* dZ,
                           double* dZcut, int64_t* zcut_off, int64_t* zcut_len) {
  if (!c || nb < 0 || (nb > 0 && (!n || !z_off || !keep))) return fail(NRG_ERR_ARG, "nrg_compact: bad arguments");
  std::vector<CompactDesc> cd;
  int64_t off = 0, maxel = 0;
  for (int b = 0; b < nb; b++) {
    if (n[b] < 0 || keep[b] < 0 || keep[b] > n[b] || z_off[b] < 0)
      return fail(NRG_ERR_ARG, "nrg_compact: block " + std::to_string(b) + " has bad sizes");
    if (keep[b] == 0) {
      if (zcut_off) zcut_off[b] = -1;
      continue;
    }
    if (zcut_off) zcut_off[b] = off;
    cd.push_back(CompactDesc{keep[b], n[b], z_off[b], off});
    const int64_t el = (int64_t)keep[b] * n[b];
    off += el;
    maxel = std::max(maxel, el);
  }
  if (zcut_len) *zcut_len = off;
  if (!dZcut || cd.empty()) return NRG_OK;
  if (!dZ) return fail(NRG_ERR_ARG, "nrg_compact: dZ is NULL");
  HIPCHK(hipSetDevice(c->device));
  CompactDesc* d_cd;
  int rc = upload(c, "compact_desc", cd, &d_cd);
  if (rc) return rc;
  HIPCHK(nrgk_launch_compact(d_cd, (int)cd.size(), maxel, dZ, dZcut, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return NRG_OK;
}

// ---------------------------------------------------------------------------
// rotations: out = Zi * F * Zj^T as two NT GEMMs through auxT = Zj * F^T
// ---------------------------------------------------------------------------
// The two grouped GEMM passes of a batch of rotations, on the FP64 MFMA tile
// kernel of the eigensolver (B^T read through its strides):
//   pass 1: auxT(kj x ni) = Zj(kj x nj) * F(ni x nj)^T
//   pass 2: out(ki x kj)  = Zi(ki x ni) * auxT(kj x ni)^T
// probs holds pass 1's problems then pass 2's; tiles likewise, the first n1tiles
// of them pass 1's.  Returns the number of pass-1 problems (0: nothing to do).
int nrg::rot_plan(const std::vector<RotJob>& jobs, std::vector<E2Gemm>& pall, std::vector<E2Tile>& tall,
                  size_t& n1tiles) {
  pall.clear();
  tall.clear();
  for (size_t q = 0; q < jobs.size(); q++) {
    const RotJob& J = jobs[q];
    if (J.ki == 0 || J.kj == 0) continue;
    E2Gemm g{};
    g.dyn = -1; g.alpha = 1.0;
    g.m = J.kj; g.n = J.ni; g.k = J.nj;
    g.a_off = J.zj_off; g.as0 = J.ldzj; g.as1 = 1;
    g.b_off = J.f_off; g.bs0 = 1; g.bs1 = J.nj;
    g.c_off = J.aux_off; g.cs0 = J.ni;
    pall.push_back(g);
  }
  const int n1 = (int)pall.size();
  for (size_t q = 0; q < jobs.size(); q++) {
    const RotJob& J = jobs[q];
    if (J.ki == 0 || J.kj == 0) continue;
    E2Gemm h{};
    h.dyn = -1; h.alpha = 1.0;
    h.m = J.ki; h.n = J.kj; h.k = J.ni;
    h.a_off = J.zi_off; h.as0 = J.ldzi; h.as1 = 1;
    h.b_off = J.aux_off; h.bs0 = 1; h.bs1 = J.ni;
    h.c_off = J.out_off; h.cs0 = J.kj;
    pall.push_back(h);
  }
  n1tiles = 0;
  for (int q = 0; q < (int)pall.size(); q++) {
    const int tm = (pall[q].m + 63) / 64, tn = (pall[q].n + 63) / 64;
    for (int a = 0; a < tm; a++)
      for (int b = 0; b < tn; b++) tall.push_back(E2Tile{q, a, b, 0});
    if (q == n1 - 1) n1tiles = tall.size();
  }
  return n1;
}

int nrg::rotate_dev(nrg_ctx* c, const std::vector<RotJob>& jobs, const double* dZ, const double* dF, double* dAux,
                    double* dOut) {
  if (jobs.empty()) return NRG_OK;
  if (c->plain && jobs.size() > 1) {
    for (size_t q = 0; q < jobs.size(); q++) {
      std::vector<RotJob> one(1, jobs[q]);
      int rc = rotate_dev(c, one, dZ, dF, dAux, dOut);
      if (rc) return rc;
    }
    return NRG_OK;
  }
  std::vector<E2Gemm> pall;
  std::vector<E2Tile> tall;
  size_t n1t = 0;
  if (rot_plan(jobs, pall, tall, n1t) == 0) return NRG_OK;
  const size_t n1tiles = n1t;
  const size_t nt = tall.size();
  int rc;
  // through the pinned ring, with no wait: a later call's copies into the same
  // buffers run after these launches on the stream
  E2Gemm* d_p;
  E2Tile* d_t;
  if ((rc = upload(c, "rot_probs", pall, &d_p))) return rc;
  if ((rc = upload(c, "rot_tiles", tall, &d_t))) return rc;
  double fl = 0.0, by = 0.0;
  for (auto& g : pall) {
    fl += 2.0 * g.m * (double)g.n * g.k;
    by += 8.0 * ((double)g.m * g.k + (double)g.n * g.k + (double)g.m * g.n);
  }
  PendingTimer tm;
  if ((rc = c->timer_begin(tm, NRG_K_GEMM, fl, by))) return rc;
  HIPCHK(e2_launch_gemm3(d_p, d_t, (int)n1tiles, dZ, dF, dAux, c->stream));
  HIPCHK(e2_launch_gemm3(d_p, d_t + n1tiles, (int)(nt - n1tiles), dZ, dAux, dOut, c->stream));
  if ((rc = c->timer_end(tm))) return rc;
  return NRG_OK;
}

int nrg::rot_jobs(int npairs, const nrg_pair_desc* pairs, int64_t zlen, int64_t flen, int64_t outlen,
                  std::vector<RotJob>& jobs, int64_t& aux, const char* who) {
  jobs.assign(npairs, RotJob{});
  aux = 0;
  for (int q = 0; q < npairs; q++) {
    const nrg_pair_desc& P = pairs[q];
    if (P.ni < 0 || P.nj < 0 || P.ki < 0 || P.kj < 0 || P.ki > P.ni || P.kj > P.nj)
      return fail(NRG_ERR_ARG, std::string(who) + ": bad pair shape");
    if (P.zi_off < 0 || P.zj_off < 0 || P.f_off < 0 || P.out_off < 0 ||
        P.zi_off + (int64_t)P.ki * P.ni > zlen || P.zj_off + (int64_t)P.kj * P.nj > zlen ||
        P.f_off + (int64_t)P.ni * P.nj > flen || P.out_off + (int64_t)P.ki * P.kj > outlen)
      return fail(NRG_ERR_ARG, std::string(who) + ": pair exceeds buffer");
    RotJob& J = jobs[q];
    J.ni = P.ni; J.nj = P.nj; J.ki = P.ki; J.kj = P.kj;
    J.zi_off = P.zi_off; J.zj_off = P.zj_off; J.ldzi = P.ni; J.ldzj = P.nj;
    J.f_off = P.f_off; J.out_off = P.out_off; J.aux_off = aux;
    aux += (int64_t)P.kj * P.ni;
  }
  return NRG_OK;
}

extern "C" int nrg_rotate_ops(nrg_ctx* c, int npairs, const nrg_pair_desc* pairs, const double* Z, int64_t zlen,
                              const double* F, int64_t flen, double* out, int64_t outlen, int precision) {
  if (!c || npairs < 0 || (npairs > 0 && (!pairs || !Z || !F || !out)))
    return fail(NRG_ERR_ARG, "nrg_rotate_ops: bad arguments");
  if (precision != 0) return fail(NRG_ERR_ARG, "nrg_rotate_ops: only precision 0 (FP64) is available");
  HIPCHK(hipSetDevice(c->device));
  std::vector<RotJob> jobs;
  int64_t aux = 0;
  int rc0 = rot_jobs(npairs, pairs, zlen, flen, outlen, jobs, aux, "nrg_rotate_ops");
  if (rc0) return rc0;
  DBuf& dZ = c->buf("rot_Z");
  DBuf& dF = c->buf("rot_F");
  DBuf& dA = c->buf("rot_aux");
  DBuf& dO = c->buf("rot_out");
  int rc;
  if ((rc = dZ.reserve(sizeof(double) * (size_t)(zlen + 1)))) return rc;
  if ((rc = dF.reserve(sizeof(double) * (size_t)(flen + 1)))) return rc;
  if ((rc = dA.reserve(sizeof(double) * (size_t)(aux + 1)))) return rc;
  if ((rc = dO.reserve(sizeof(double) * (size_t)(outlen + 1)))) return rc;
  HIPCHK(hipMemcpyAsync(dZ.p, Z, sizeof(double) * zlen, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dF.p, F, sizeof(double) * flen, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(dO.p, 0, sizeof(double) * (size_t)(outlen + 1), c->stream));
  if ((rc = rotate_dev(c, jobs, (double*)dZ.p, (double*)dF.p, (double*)dA.p, (double*)dO.p))) return rc;
  HIPCHK(hipMemcpyAsync(out, dO.p, sizeof(double) * outlen, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return NRG_OK;
}
