// nrg_internal.h -- shared host-side internals of the HIP library.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/twochnrg_amd.h"
#include "nrg_kernels.h"
#include "nrg_eig2.h"
#include "nrg_trunc.h"

namespace nrg {
int fail(int code, const std::string& msg);
// every live DM-NRG handle of c forgets its context (nrg_ctx_destroy before nrg_dm_destroy)
void dm_detach_all(nrg_ctx* c);
}

#define HIPCHK(x)                                                                               \
  do {                                                                                          \
    hipError_t _e = (x);                                                                        \
    if (_e != hipSuccess)                                                                       \
      return nrg::fail(NRG_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(_e));            \
  } while (0)

// grow-only device allocation (exact: no growth slack, for buffers written once)
struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf();
  int reserve(size_t bytes, bool exact = false);
  template <class T>
  T* as() { return (T*)p; }
};

// per-kernel timing with HIP events on the context stream
struct KStat {
  double ms = 0.0;       // summed kernel time
  long launches = 0;
  double flops = 0.0;    // algorithmic flops of the timed launches
  double bytes = 0.0;    // algorithmic bytes of the timed launches
};

struct PendingTimer {
  hipEvent_t a, b;
  int kind;
  double flops, bytes;
};

struct nrg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<std::string, std::unique_ptr<DBuf>> bufs;
  DBuf& buf(const std::string& name);
  // release one named buffer / every buffer whose name does not start with keep
  void drop(const std::string& name);
  void trim(const std::string& keep);
  // move a buffer to another name (replacing what that name held)
  void rename(const std::string& from, const std::string& to);
  // an eigensolve whose eigenvalues are out and whose vectors wait for the
  // truncation (eig_begin / eig_finish)
  struct E2Pending {
    bool pending = false;
    std::vector<E2Block> blk;
    double* dA = nullptr;
    std::vector<int> ts_foff;    // abort-flag slots of the two-stage blocks
    int nflag = 0;
    int* abort_host = nullptr;   // pinned copy of the chase flags
    size_t abort_cap = 0;
  } e2p;
  int ts_min = 0;      // blocks with n >= ts_min use the two-stage reduction (0: none)
  // hipGraph mode: the eigensolver's launch sequences are captured into a graph
  // (after their descriptor uploads), instantiated and launched once
  bool graph = false;
  bool capturing = false;
  std::vector<hipGraphExec_t> gexec;   // launched graphs, released two eigensolves later
  double graph_host_ms = 0.0;          // host time spent capturing + instantiating + launching
  long graph_nodes = 0, graph_count = 0;
  int dm_seq = 0;      // DM-NRG handles created on this context (buffer name prefixes)
  std::vector<nrg_dm*> dms;   // live DM-NRG handles: detached when the context goes first
  // small host-to-device copies go through a pinned ring on the stream, with no
  // wait: the ring is reused only after a stream synchronisation
  char* ring = nullptr;
  size_t ring_cap = 0, ring_head = 0;
  int stage_h2d(void* dst, const void* src, size_t bytes);
  uint64_t held() const;   // bytes of every named buffer
  bool timing = false;
  int eig_impl = 2;   // 1: one workgroup per block (Householder + QL); 2: tridiag + D&C + WY
  bool plain = false; // one block (eigensolve) / one block pair (rotation) per launch sequence
  int red_streams = 3; // streams (block groups) of the blocked reduction
  KStat stat[NRG_NKSTAT];
  std::vector<PendingTimer> pending;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t get_event();
  // record an event pair around a launch (on stream s, default the context
  // stream); resolve() after a stream sync
  int timer_begin(PendingTimer& t, int kind, double flops, double bytes, hipStream_t s = nullptr);
  int timer_end(PendingTimer& t, hipStream_t s = nullptr);
  // auxiliary streams (default priority, created on first use): fork(n) makes
  // aux[0..n) wait for the work enqueued on stream so far, join(n) makes
  // stream wait for theirs
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> sync_ev;
  int fork(int n);
  int join(int n);
  int resolve();
};

namespace nrg {

// one rotation problem out(ki x kj) = Zi(ki x ni) F(ni x nj) Zj(kj x nj)^T
struct RotJob {
  int ni, nj, ki, kj;
  int ldzi, ldzj;
  int64_t zi_off, zj_off, f_off, aux_off, out_off;
};

int eig_dev(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA, const int64_t* w_off, double* dW,
            int* info_host);
int eig2_dev(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA, const int64_t* w_off, double* dW);
int eig2_begin(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA, const int64_t* w_off, double* dW);
int eig2_finish(nrg_ctx* c, const int* krow);
// eig_dev in two halves: eig_begin writes the eigenvalues (ascending per block)
// to dW; eig_finish writes the first krow[b] eigenvector rows of block b to dA
// (krow NULL: every row).  Between the two nothing may run another eigensolve
// on the context.  With the one-workgroup solver (impl 1) eig_begin does it all.
int eig_begin(nrg_ctx* c, int nb, const int* n, const int64_t* a_off, double* dA, const int64_t* w_off, double* dW);
int eig_finish(nrg_ctx* c, const int* krow);
int rotate_dev(nrg_ctx* c, const std::vector<RotJob>& jobs, const double* dZ, const double* dF, double* dAux,
               double* dOut);
// the descriptors of rotate_dev's two GEMM passes (see nrg_host.cpp)
int rot_plan(const std::vector<RotJob>& jobs, std::vector<E2Gemm>& pall, std::vector<E2Tile>& tall,
             size_t& n1tiles);
// RotJobs of nrg_rotate_ops' pairs after its bounds checks; aux_len receives the
// doubles of the intermediate products
int rot_jobs(int npairs, const nrg_pair_desc* pairs, int64_t zlen, int64_t flen, int64_t outlen,
             std::vector<RotJob>& jobs, int64_t& aux_len, const char* who);
void set_degen(std::vector<double>& en, double eps);
double ecut_of(const std::vector<double>& en, int ncut);

// upload helper
template <class T>
int upload(nrg_ctx* c, const char* name, const std::vector<T>& v, T** dptr) {
  DBuf& b = c->buf(name);
  int rc = b.reserve(sizeof(T) * (v.size() + 1));
  if (rc) return rc;
  if (!v.empty() && (rc = c->stage_h2d(b.p, v.data(), sizeof(T) * v.size()))) return rc;
  *dptr = (T*)b.p;
  return 0;
}

}  // namespace nrg
