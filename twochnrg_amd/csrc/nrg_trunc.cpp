// nrg_trunc.cpp -- the truncation of an NRG step on the host (SetDegen, Ecut,
// SetKept) and the library's error string.  Plain C++ with no HIP: it is also
// built on its own under AddressSanitizer/UBSan by the tests.
#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/twochnrg_amd.h"
#include "nrg_trunc.h"

namespace nrg {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// v sorted by comp, on up to 8 host threads when it is long (an NRG step of
// tens of thousands of states): each thread sorts a contiguous chunk, then
// neighbouring chunks are merged.  std::merge takes from the left range on
// ties, so with stable chunk sorts the result is std::stable_sort's; for a
// plain sort of values the sorted sequence is the same whatever the method.
template <class T, class C>
static void par_sort(std::vector<T>& v, C comp, bool stable) {
  const size_t n = v.size();
  const int hw = (int)std::thread::hardware_concurrency();
  const int nt = std::min(8, hw > 0 ? hw : 1);
  auto serial = [&](size_t a, size_t b) {
    if (stable) std::stable_sort(v.begin() + a, v.begin() + b, comp);
    else std::sort(v.begin() + a, v.begin() + b, comp);
  };
  if (n < 16384 || nt < 2) {
    serial(0, n);
    return;
  }
  std::vector<size_t> bd(nt + 1);
  for (int t = 0; t <= nt; t++) bd[t] = n * t / nt;
  std::vector<std::thread> th;
  int spawned = 0;
  try {
    for (int t = 1; t < nt; t++, spawned++) th.emplace_back(serial, bd[t], bd[t + 1]);
  } catch (...) {
  }
  serial(bd[0], bd[1]);
  for (int t = 1 + spawned; t < nt; t++) serial(bd[t], bd[t + 1]);   // chunks of threads that could not start
  for (auto& x : th) x.join();
  std::vector<T> tmp(n);
  for (int w = 1; w < nt; w *= 2) {
    for (int i = 0; i < nt; i += 2 * w) {
      const size_t a = bd[i], m = bd[std::min(i + w, nt)], b = bd[std::min(i + 2 * w, nt)];
      std::merge(v.begin() + a, v.begin() + m, v.begin() + m, v.begin() + b, tmp.begin() + a, comp);
    }
    v.swap(tmp);
  }
}

// ---------------------------------------------------------------------------
// SetDegen (src/NRGarrayClass.cpp:244-268): pairs (ii<jj) in index order with
// |E_jj - E_ii| < eps are snapped to the smaller value, the comparisons using
// values already snapped.  Values only ever move to another value of the same
// cluster (maximal run of sorted values with gaps < eps), and values of
// different clusters stay >= eps apart, so the quadratic loop restricted to
// each cluster (members in index order) gives the identical result.
// ---------------------------------------------------------------------------
void set_degen(std::vector<double>& en, double eps) {
  const int n = (int)en.size();
  if (n < 2) return;
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  par_sort(idx, [&](int a, int b) { return en[a] < en[b]; }, true);
  int s = 0;
  while (s < n) {
    int e = s + 1;
    while (e < n && std::fabs(en[idx[e]] - en[idx[e - 1]]) < eps) e++;
    if (e - s > 1) {
      std::vector<int> mem(idx.begin() + s, idx.begin() + e);
      std::sort(mem.begin(), mem.end());
      for (size_t a = 0; a < mem.size(); a++)
        for (size_t b = a + 1; b < mem.size(); b++) {
          double& x = en[mem[a]];
          double& y = en[mem[b]];
          if (std::fabs(y - x) < eps) {
            if (y > x) y = x; else x = y;
          }
        }
    }
    s = e;
  }
}

// CNRGarray::Ecut (src/NRGarrayClass.cpp:661-691).  Its 1% window reads
// Eaux[Ncut+ii] bounded only by ii<size; an index past the end is taken here
// as a level far above the cut, which keeps every remaining state.
double ecut_of(const std::vector<double>& en, int ncut) {
  std::vector<double> e(en);
  par_sort(e, std::less<double>(), false);
  const int n = (int)e.size();
  if (n == 0) return 0.0;
  if (ncut >= n - 1) return e[n - 1];
  double aux = e[ncut - 1], aux2 = e[ncut];
  int ii = 1;
  while (aux2 > 0.0 && std::fabs((aux2 - aux) / aux) < 0.01 && ii < n) {
    if (ncut + ii >= n) { aux2 = HUGE_VAL; ii++; break; }
    aux2 = e[ncut + ii];
    ii++;
  }
  if (ii > 1) aux = (ncut + ii - 2 < n) ? e[ncut + ii - 2] : e[n - 1];
  return aux;
}

}  // namespace nrg

using namespace nrg;

extern "C" const char* nrg_last_error(void) { return g_err.c_str(); }

extern "C" int nrg_ecut(int nstates, const double* en, int ncut, double* ecut) {
  if (nstates <= 0 || !en || !ecut || ncut < 1) return fail(NRG_ERR_ARG, "nrg_ecut: bad arguments");
  std::vector<double> e(en, en + nstates);
  *ecut = ecut_of(e, ncut);
  return NRG_OK;
}

extern "C" int nrg_truncate(int nstates, double* en, int ncut, double myeps, uint8_t* kept, double* ecut,
                            int* nkept) {
  if (nstates <= 0 || !en || !kept || ncut < 1) return fail(NRG_ERR_ARG, "nrg_truncate: bad arguments");
  std::vector<double> e(en, en + nstates);
  set_degen(e, myeps);
  double ec = ecut_of(e, ncut);
  int nk = 0;
  for (int i = 0; i < nstates; i++) { kept[i] = e[i] <= ec; nk += kept[i]; en[i] = e[i]; }
  if (ecut) *ecut = ec;
  if (nkept) *nkept = nk;
  return NRG_OK;
}
