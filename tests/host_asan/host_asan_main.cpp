// Host logic under AddressSanitizer and UBSan (built by the Makefile beside
// this file with g++, no HIP): the truncation (SetDegen / Ecut / SetKept) and
// the reference's .bin hand-off reader and writer.
//  * nrg_truncate against a literal restatement of the reference's pairwise
//    SetDegen loop (src/NRGarrayClass.cpp:244-268) and the Ecut rule, on random
//    spectra with exact and near degeneracies, sizes 1 .. 3000, Ncut from 1 past
//    the end;
//  * .bin files: random CNRGbasisarray / CNRGmatrix contents written and read
//    back equal; every truncation of a written file, and files whose counts are
//    negative or larger than the file, are refused without reading out of
//    bounds.
// Exit status 0 when every check passes; a sanitizer report aborts the run.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/twochnrg_amd.h"
#include "../../twochnrg_amd/csrc/nrg_bin.h"

static int failures = 0;
#define CHECK(cond, ...)                                \
  do {                                                  \
    if (!(cond)) {                                      \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);               \
      std::fprintf(stderr, "\n");                       \
      failures++;                                       \
    }                                                   \
  } while (0)

// the reference's SetDegen loop, literally: every pair in index order
static void set_degen_pairwise(std::vector<double>& e, double eps) {
  for (size_t i = 0; i < e.size(); i++)
    for (size_t j = i + 1; j < e.size(); j++)
      if (std::fabs(e[j] - e[i]) < eps) {
        if (e[j] > e[i]) e[j] = e[i]; else e[i] = e[j];
      }
}

static void check_truncate(std::mt19937_64& rng) {
  std::uniform_real_distribution<double> u(0.0, 10.0);
  const int sizes[] = {1, 2, 3, 17, 100, 401, 1000, 3000};
  for (int n : sizes)
    for (int kind = 0; kind < 4; kind++)
      for (int ncut : {1, n / 2 + 1, n - 1, n, n + 5}) {
        if (ncut < 1) continue;
        std::vector<double> e(n);
        for (int i = 0; i < n; i++) e[i] = u(rng);
        if (kind == 1)   // exact multiplets
          for (int i = 1; i < n; i += 3) e[i] = e[i - 1];
        if (kind == 2)   // near degeneracies below and above eps
          for (int i = 1; i < n; i += 2) e[i] = e[i - 1] + ((i & 2) ? 0.5e-10 : 2e-10);
        if (kind == 3)   // everything equal
          for (int i = 0; i < n; i++) e[i] = 1.5;
        const double eps = 1e-10;
        std::vector<double> lib(e), ref(e);
        std::vector<uint8_t> kept(n);
        double ecut = 0.0;
        int nk = 0;
        CHECK(nrg_truncate(n, lib.data(), ncut, eps, kept.data(), &ecut, &nk) == NRG_OK, "nrg_truncate rc");
        if (n <= 1000) {   // the quadratic restatement
          set_degen_pairwise(ref, eps);
          for (int i = 0; i < n; i++) CHECK(lib[i] == ref[i], "SetDegen n=%d kind=%d i=%d", n, kind, i);
        }
        int cnt = 0;
        for (int i = 0; i < n; i++) {
          CHECK(kept[i] == (lib[i] <= ecut), "Kept flag n=%d i=%d", n, i);
          cnt += kept[i];
        }
        CHECK(cnt == nk, "kept count");
        CHECK(nk >= std::min(ncut, n), "at least Ncut kept (n=%d ncut=%d nk=%d)", n, ncut, nk);
        double e2 = 0.0;
        CHECK(nrg_ecut(n, lib.data(), ncut, &e2) == NRG_OK && e2 == ecut, "nrg_ecut agrees");
      }
  // argument checks
  double x = 1.0, ec;
  uint8_t k;
  CHECK(nrg_truncate(0, &x, 1, 1e-10, &k, &ec, nullptr) != NRG_OK, "n = 0 refused");
  CHECK(nrg_truncate(1, &x, 0, 1e-10, &k, &ec, nullptr) != NRG_OK, "Ncut = 0 refused");
  CHECK(std::strlen(nrg_last_error()) > 0, "error string set");
}

template <class T>
static void fill(std::vector<T>& v, std::mt19937_64& rng, int n) {
  v.resize(n);
  for (auto& x : v) x = (T)(rng() % 1000) - (T)300;
}

static void random_array(nrg::BinArray& a, std::mt19937_64& rng) {
  const int nb = 1 + (int)(rng() % 7);
  a.nshell = (int)(rng() % 50);
  a.nq = 2;
  fill(a.qn, rng, 2 * nb);
  fill(a.en, rng, 3 * nb);
  fill(a.evec, rng, (int)(rng() % 40));
  fill(a.cevec, rng, 2 * (int)(rng() % 3));
  fill(a.bbe, rng, 2 * nb);
  fill(a.idegen, rng, (int)(rng() % 5));
  a.kept.resize(a.en.size());
  for (auto& b : a.kept) b = rng() & 1;
  a.totalS = 1;
  a.sqn = {1};
}

static std::vector<unsigned char> slurp(const std::string& p) {
  std::vector<unsigned char> b;
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) return b;
  int c;
  while ((c = std::fgetc(f)) != EOF) b.push_back((unsigned char)c);
  std::fclose(f);
  return b;
}
static void spit(const std::string& p, const std::vector<unsigned char>& b, size_t n) {
  FILE* f = std::fopen(p.c_str(), "wb");
  if (n) std::fwrite(b.data(), 1, n, f);
  std::fclose(f);
}

static void check_bin(const std::string& dir, std::mt19937_64& rng) {
  std::string err;
  for (int rep = 0; rep < 20; rep++) {
    nrg::BinBasis a;
    random_array(a, rng);
    fill(a.itype, rng, (int)a.en.size());
    fill(a.stcf, rng, (int)a.en.size());
    fill(a.bbe_bc, rng, (int)a.bbe.size());
    fill(a.bbe_ev, rng, (int)a.bbe.size());
    a.nq_stcf = 2;
    fill(a.stcf_qn, rng, 4);
    nrg::BinMatrix m;
    random_array(m, rng);
    fill(m.matel, rng, (int)(rng() % 60));
    fill(m.map, rng, 4);
    fill(m.mbe, rng, 4);
    m.need_old = 1;
    m.wigner_eckart_l = 0.5;
    const std::string pa = dir + "/asan_basis.bin", pm = dir + "/asan_mat.bin", pt = dir + "/asan_cut.bin";
    CHECK(nrg::bin_write(pa, a, err) == 0, "write basis: %s", err.c_str());
    CHECK(nrg::bin_write(pm, m, err) == 0, "write matrix: %s", err.c_str());
    nrg::BinBasis a2;
    nrg::BinMatrix m2;
    CHECK(nrg::bin_read(pa, a2, err) == 0, "read basis: %s", err.c_str());
    CHECK(nrg::bin_read(pm, m2, err) == 0, "read matrix: %s", err.c_str());
    CHECK(a2.qn == a.qn && a2.en == a.en && a2.evec == a.evec && a2.bbe == a.bbe && a2.kept == a.kept &&
              a2.itype == a.itype && a2.stcf == a.stcf && a2.bbe_bc == a.bbe_bc && a2.bbe_ev == a.bbe_ev &&
              a2.stcf_qn == a.stcf_qn && a2.cevec == a.cevec && a2.sqn == a.sqn,
          "basis round trip");
    CHECK(m2.matel == m.matel && m2.map == m.map && m2.mbe == m.mbe && m2.need_old == m.need_old &&
              m2.wigner_eckart_l == m.wigner_eckart_l && m2.en == m.en,
          "matrix round trip");
    // every proper prefix of the files is refused
    for (const std::string& p : {pa, pm}) {
      const std::vector<unsigned char> bytes = slurp(p);
      for (size_t len = 0; len < bytes.size(); len++) {
        spit(pt, bytes, len);
        nrg::BinBasis xb;
        nrg::BinMatrix xm;
        const int rc = (p == pa) ? nrg::bin_read(pt, xb, err) : nrg::bin_read(pt, xm, err);
        CHECK(rc != 0, "truncated file of %zu/%zu bytes accepted", len, bytes.size());
      }
      // a corrupt count: negative, then larger than the file
      std::vector<unsigned char> bad(bytes);
      const int neg = -5, huge = 1 << 30;
      std::memcpy(bad.data() + 8, &neg, 4);   // the count of qn
      spit(pt, bad, bad.size());
      nrg::BinMatrix xm;
      CHECK(nrg::bin_read(pt, xm, err) != 0, "negative count accepted");
      std::memcpy(bad.data() + 8, &huge, 4);
      spit(pt, bad, bad.size());
      CHECK(nrg::bin_read(pt, xm, err) != 0, "oversized count accepted");
    }
  }
  nrg::BinBasis none;
  CHECK(nrg::bin_read(dir + "/does_not_exist.bin", none, err) != 0, "missing file refused");
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  std::mt19937_64 rng(20261204);
  check_truncate(rng);
  check_bin(dir, rng);
  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("host checks passed\n");
  return 0;
}
