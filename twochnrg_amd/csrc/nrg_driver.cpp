// nrg_driver.cpp -- chain drivers of the MI355X NRG path:
//   * OneChQS Anderson through the NRG_main loop (src/Main/NRG_main.cpp:282-427)
//   * TwoChQS Anderson through the legacy standalone loop
//     (src/TwoChQS/TwoChQS.cpp:386-495)
//
// Host objects keep the reference's data-model semantics
// (include/NRGclasses.hpp:47-759): QNumbers per block, BlockBegEnd inclusive
// pairs, dEn per state, Kept flags, iType/StCameFrom/BlockBegEndBC for basis
// arrays, MatBlockMap/MatBlockBegEnd for operators.  Eigenvectors and
// operator elements stay resident in HBM; the host holds the integer
// bookkeeping (basis build, block maps, truncation decisions), the per-model
// coupling tables (Clebsch-Gordan sums, computed once per spin value) and the
// O(n) FP64 Boltzmann sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

#include "nrg_dm.h"
#include "nrg_eig2.h"
#include "nrg_internal.h"

namespace nrg {

// ---------------------------------------------------------------------------
// data model
// ---------------------------------------------------------------------------
struct CNRGarray {
  int Nshell = 0, NQNumbers = 0;
  std::vector<double> QNumbers, dEn;
  std::vector<int> BlockBegEnd;
  std::vector<unsigned char> Kept;
  int NumBlocks() const { return NQNumbers ? (int)QNumbers.size() / NQNumbers : 0; }
  int GetBlockLimit(int b, int w) const { return BlockBegEnd[2 * b + w]; }
  int GetBlockSize(int b) const { return BlockBegEnd[2 * b + 1] - BlockBegEnd[2 * b] + 1; }
  int Nstates() const { return BlockBegEnd.empty() ? 0 : BlockBegEnd.back() - BlockBegEnd.front() + 1; }
  double GetQNumber(int b, int q) const { return QNumbers[(size_t)b * NQNumbers + q]; }
};

struct CNRGbasisarray : CNRGarray {
  std::vector<int> iType, StCameFrom, BlockBegEndBC;
  // device handle of the kept eigenvector rows of each block: offset of row 0
  // in the step's eigenvector buffer and the row length (pre-cut block size)
  std::vector<int64_t> zoff;
  std::vector<int> zld;
};

// an operator over a kept basis; elements on the device, block map on the host
struct CNRGmatrix {
  bool NeedOld = false, CalcAvg = false;
  std::vector<int> MatBlockMap, MatBlockBegEnd;
  std::string buf[2];  // device buffer names (double-buffered across steps)
  int cur = 0;
};

static long key2(double x) { return std::lround(2.0 * x); }
static bool deq(double a, double b) { return std::fabs(a - b) < 1e-20; }  // dEqual src/dEqual.cpp:6-10

// ---------------------------------------------------------------------------
// Clebsch-Gordan (Racah formula) -- src/CGordan.cpp:34-120
// ---------------------------------------------------------------------------
static double factd(int n) {
  static std::vector<double> t;
  if (t.empty()) { t.resize(171); t[0] = 1.0; for (int i = 1; i < 171; i++) t[i] = t[i - 1] * (double)i; }
  return (n < 0 || n > 170) ? 0.0 : t[n];
}
static bool triangle(double a, double b, double c) {  // TriangIneq src/CGordan.cpp:13-17
  return !((c - (a + b) > 1.0e-10) || (c - std::fabs(a - b) < -1.0e-10));
}
double cgordan(double J1, double M1, double J2, double M2, double J, double M) {
  if (J1 < 0 || J2 < 0 || J < 0) return 0.0;
  if (std::fabs(M1) > J1 || std::fabs(M2) > J2 || std::fabs(M) > J) return 0.0;
  if (std::fabs(M - (M1 + M2)) > 1.0e-10) return 0.0;
  if (J > J1 + J2 || J < std::fabs(J1 - J2)) return 0.0;
  double rho2 = factd((int)(J1 + J2 - J)) * factd((int)(J1 - J2 + J)) * factd((int)(-J1 + J2 + J)) * (2 * J + 1) /
                factd((int)(J1 + J2 + J + 1));
  double rho = rho2 > 0 ? std::sqrt(rho2) : 0.0;
  double sg2 = factd((int)(J + M)) * factd((int)(J - M)) * factd((int)(J1 + M1)) * factd((int)(J1 - M1)) *
               factd((int)(J2 + M2)) * factd((int)(J2 - M2));
  double sg = sg2 > 0 ? std::sqrt(sg2) : 0.0;
  int rmax = std::min({(int)(J1 - M1), (int)(J2 + M2), (int)(J1 + J2 - J)});
  int rmin = std::max({-(int)(J - J2 + M1), -(int)(J - J1 - M2), 0});
  double tau = 0.0;
  for (int r = rmin; r <= rmax; r++) {
    double den = factd((int)(J1 - M1) - r) * factd((int)(J2 + M2) - r) * factd((int)(J - J2 + M1) + r) *
                 factd((int)(J - J1 - M2) + r) * factd((int)(J1 + J2 - J) - r) * factd(r);
    tau += ((r & 1) ? -1.0 : 1.0) / den;
  }
  return rho * sg * tau;
}

// ---------------------------------------------------------------------------
// single-site description: site blocks (Q,S,Sz) with their state ranges;
// a state's index is its iType
// ---------------------------------------------------------------------------
struct SiteDef {
  std::vector<double> bQ, bS, bSz;
  std::vector<int> bbeg, bend, blk;
  int ntypes() const { return (int)blk.size(); }
  int nblocks() const { return (int)bQ.size(); }
  double Q(int t) const { return bQ[blk[t]]; }
  double S(int t) const { return bS[blk[t]]; }
  double Sz(int t) const { return bSz[blk[t]]; }
  // GetBlockFromQNumbers: first block with all three QNs dEqual, else -1
  int find(double q, double s, double sz) const {
    for (int b = 0; b < nblocks(); b++)
      if (deq(bQ[b], q) && deq(bS[b], s) && deq(bSz[b], sz)) return b;
    return -1;
  }
  // state range of a site block; GetBlockLimit(-1,.) returns 0
  void range(int b, int& lo, int& hi) const {
    if (b < 0) { lo = hi = 0; return; }
    lo = bbeg[b]; hi = bend[b];
  }
  void add_block(double q, double s, double sz, int nstates) {
    int lo = ntypes();
    bQ.push_back(q); bS.push_back(s); bSz.push_back(sz);
    bbeg.push_back(lo); bend.push_back(lo + nstates - 1);
    for (int i = 0; i < nstates; i++) blk.push_back(nblocks() - 1);
  }
};

// ---------------------------------------------------------------------------
// OneChQS plugin: OneChQS_SetSingleSite, OneCh_fd_table (src/OpTables.cpp:8-33),
// OneChQS_HN_MatEl / OneChQS_fN_MatEl / OneChQS_cd_MatEl
// (src/OpMatRules.cpp:1167-1535)
// ---------------------------------------------------------------------------
static SiteDef onechqs_site() {
  SiteDef s;
  s.add_block(-1, 0, 0, 1);
  s.add_block(0, 0.5, 0.5, 1);
  s.add_block(0, 0.5, -0.5, 1);
  s.add_block(1, 0, 0, 1);
  return s;
}
static double onech_fd(int sigma, int ti, int tj) {
  if (sigma == 1) return (ti == 1 && tj == 0) ? 1.0 : ((ti == 3 && tj == 2) ? 1.0 : 0.0);
  if (sigma == -1) return (ti == 2 && tj == 0) ? 1.0 : ((ti == 3 && tj == 1) ? -1.0 : 0.0);
  return 0.0;
}
// CoefMatEl * FermiSign of OneChQS_HN_MatEl (block spin Si, site types typep, type)
static double onech_hn_coef(const SiteDef& st, double Si, int tp, int tt) {
  double Stp = st.S(tp), Sztp = st.Sz(tp), Qt = st.Q(tt), St = st.S(tt), Szt = st.Sz(tt);
  double Scfp = Si - Sztp, Scf = Si - Szt;
  double fsign = deq(Qt, 0.0) ? -1.0 : 1.0;
  double coef = 0.0;
  for (double sz = -St; sz <= St; sz += 1.0) {
    int ss = st.find(Qt, St, sz);
    for (int sg = -1; sg <= 1; sg += 2) {
      double ds = 0.5 * sg, szcf = Si - sz, szcfp = szcf - ds, sztp = sz + ds;
      int ssp = st.find(st.Q(tp), Stp, sz + ds);
      coef += cgordan(Scf, szcf, St, sz, Si, Si) * cgordan(Scfp, szcfp, Stp, sztp, Si, Si) *
              cgordan(Scfp, szcfp, 0.5, ds, Scf, szcf) * onech_fd(sg, ssp, ss);
    }
  }
  return coef * fsign;
}
static double onech_fn_el(const SiteDef& st, double Si, double Sj, int ti, int tj) {
  double Sti = st.S(ti), Szti = st.Sz(ti), Stj = st.S(tj), Sztj = st.Sz(tj);
  double Scfi = Si - Szti, Scfj = Sj - Sztj, sigma = Sj - Si, acc = 0.0;
  for (double sz = -Stj; sz <= Stj; sz += 1.0) {
    int ssj = st.find(st.Q(tj), Stj, sz), ssi = st.find(st.Q(ti), Sti, sz - sigma);
    acc += onech_fd((int)(sigma / 0.5), ssj, ssi) * cgordan(Scfi, Si - sz + sigma, Sti, sz - sigma, Si, Si) *
           cgordan(Scfj, Sj - sz, Stj, sz, Sj, Sj);
  }
  double cgr = cgordan(Si, Si, 0.5, sigma, Sj, Sj);
  return !deq(cgr, 0.0) ? acc / cgr : 0.0;
}
static double onech_cd_el(const SiteDef& st, double Si, double Sj, int ti, int tj) {
  double Sti = st.S(ti), Szti = st.Sz(ti), Qtj = st.Q(tj), Stj = st.S(tj), Sztj = st.Sz(tj);
  double Scfi = Si - Szti, Scfj = Sj - Sztj, sigma = Sj - Si;
  double cg3 = cgordan(Si, Si, 0.5, sigma, Sj, Sj);
  if (deq(cg3, 0.0)) {
    sigma = Si - Sj;
    cg3 = cgordan(Si, Si, 0.5, sigma, Sj, Sj);
    if (deq(cg3, 0.0)) return 0.0;
  }
  double fb = 0.0;
  for (double szcfj = -Scfj; szcfj <= Scfj; szcfj += 1.0) {
    double szcfi = szcfj - sigma, sztj = Sj - szcfj, szti = Si - szcfi;
    int ssj = st.find(Qtj, Stj, sztj), ssi = st.find(st.Q(ti), Sti, szti);
    double fs = deq(Qtj, 0.0) ? -1.0 : 1.0;
    if (ssi == ssj)
      fb += cgordan(Scfi, szcfi, Sti, szti, Si, Si) * cgordan(Scfj, szcfj, Stj, sztj, Sj, Sj) *
            cgordan(Scfi, szcfi, 0.5, sigma, Scfj, szcfj) * fs;
  }
  return std::fabs(cg3) > 1e-20 ? fb / cg3 : 0.0;
}

// ---------------------------------------------------------------------------
// TwoChQS plugin: TwoChQS_SetSingleSite (src/TwoChQS/TwoChQS_SetSingleSite.cpp),
// TwoChQS_fd_table / TwoChQS_SpSm_table (src/TwoChQS/TwoChQS_OpTables.cpp:19-183),
// H_N rules of TwoChQS_DiagHN (src/TwoChQS/TwoChQS_DiagHN.cpp:62-250) and
// TwoChQS_SetH0Anderson (src/TwoChQS/TwoChQS_SetH0Anderson.cpp:95-190),
// f_ch rule of TwoChQS_UpdateMatrixAfterCutting (src/TwoChQS/TwoChQS_UpdateMatrices.cpp:380-470)
// ---------------------------------------------------------------------------
static SiteDef twochqs_site() {
  SiteDef s;
  s.add_block(-2, 0, 0, 1);
  s.add_block(-1, 0.5, 0.5, 2);
  s.add_block(-1, 0.5, -0.5, 2);
  s.add_block(0, 0, 0, 3);
  s.add_block(0, 1, -1, 1);
  s.add_block(0, 1, 0, 1);
  s.add_block(0, 1, 1, 1);
  s.add_block(1, 0.5, 0.5, 2);
  s.add_block(1, 0.5, -0.5, 2);
  s.add_block(2, 0, 0, 1);
  return s;
}
struct TwoChTables {
  double fd[2][2][16][16];  // [channel][sigma>0][i][j] of f^dagger
  TwoChTables() {
    std::memset(fd, 0, sizeof(fd));
    const double r = 1.0 / std::sqrt(2.0);
    auto set = [&](int ch, int up, std::initializer_list<std::tuple<int, int, double>> L) {
      for (auto& x : L) fd[ch][up][std::get<0>(x)][std::get<1>(x)] = std::get<2>(x);
    };
    set(0, 1, {{2, 0, 1}, {10, 1, 1}, {9, 3, r}, {11, 6, 1}, {5, 3, r}, {7, 4, 1}, {12, 9, r}, {12, 5, -r},
               {14, 8, 1}, {15, 13, 1}});
    set(0, 0, {{4, 0, 1}, {9, 1, r}, {5, 1, -r}, {8, 3, 1}, {13, 6, 1}, {7, 2, -1}, {12, 10, -1}, {14, 9, -r},
               {14, 5, -r}, {15, 11, -1}});
    set(1, 1, {{1, 0, 1}, {6, 3, 1}, {10, 2, -1}, {11, 9, -r}, {11, 5, -r}, {9, 4, -r}, {5, 4, r}, {13, 8, -1},
               {12, 7, 1}, {15, 14, 1}});
    set(1, 0, {{3, 0, 1}, {6, 1, -1}, {9, 2, -r}, {11, 10, 1}, {5, 2, -r}, {8, 4, -1}, {13, 9, r}, {13, 5, -r},
               {14, 7, 1}, {15, 12, -1}});
  }
  double f(int ch, int sigma, int ti, int tj) const {
    if (ch < 1 || ch > 2 || (sigma != 1 && sigma != -1) || ti < 0 || ti > 15 || tj < 0 || tj > 15) return 0.0;
    return fd[ch - 1][sigma == 1][ti][tj];
  }
  static double spsm(int ti, int tj) {
    if (ti == tj) return 1.0;
    int a = std::min(ti, tj), b = std::max(ti, tj);
    return ((a == 1 && b == 3) || (a == 2 && b == 4) || (a == 11 && b == 13) || (a == 12 && b == 14) ||
            (a == 8 && b == 9) || (a == 9 && b == 10) || (a == 8 && b == 10))
               ? 1.0
               : 0.0;
  }
};
static bool dle(double a, double b) { return (a < b) || deq(a, b); }  // dLEqual
// h0 = 1: TwoChQS_SetH0Anderson fermion sign (Q = +-1 and Sz = +-1/2)
static double twoch_hn_coef(const SiteDef& st, const TwoChTables& T, int ch, double Si, int typep, int type, int h0) {
  double Qtp = st.Q(typep), Stp = st.S(typep), Sztp = st.Sz(typep);
  double Qt = st.Q(type), St = st.S(type), Szt = st.Sz(type);
  double Soldp = Si - Sztp, Sold = Si - Szt;
  double fs = 1.0;
  if (h0) { if ((deq(Qt, 1.0) || deq(Qt, -1.0)) && (deq(Szt, 0.5) || deq(Szt, -0.5))) fs = -1.0; }
  else if (deq(Szt, 0.5) || deq(Szt, -0.5)) fs = -1.0;
  double tot = 0.0;
  for (int sg = -1; sg <= 1; sg += 2) {
    double ds = 0.5 * sg;
    for (double szo = Sold; szo >= -Sold; szo -= 1.0) {
      double szop = szo - ds, szt = Si - szo, sztp = szt + ds;
      int sb = st.find(Qt, St, szt), sbp = st.find(Qtp, Stp, sztp);
      if (dle(std::fabs(sztp), Stp) && dle(std::fabs(szo), Sold) && dle(std::fabs(szop), Soldp)) {
        double c = cgordan(Sold, szo, St, szt, Si, Si) * cgordan(Soldp, szop, Stp, sztp, Si, Si) *
                   cgordan(Soldp, szop, 0.5, ds, Sold, szo);
        int p0, p1, q0, q1;
        st.range(sbp, p0, p1);
        st.range(sb, q0, q1);
        for (int ssp = p0; ssp <= p1; ssp++)
          for (int ss = q0; ss <= q1; ss++)
            tot += c * T.f(ch, sg, ssp, ss) * fs * TwoChTables::spsm(ss, type) * TwoChTables::spsm(ssp, typep);
      }
    }
  }
  return tot;
}
static double twoch_fn_el(const SiteDef& st, const TwoChTables& T, int ch, double Si, double Sj, int ti, int tj) {
  double Qti = st.Q(ti), Sti = st.S(ti), Szti = st.Sz(ti), Qtj = st.Q(tj), Stj = st.S(tj), Sztj = st.Sz(tj);
  double Scfi = Si - Szti, Scfj = Sj - Sztj;
  double cgm = cgordan(Si, Si, 0.5, -0.5, Sj, Sj), cgp = cgordan(Si, Si, 0.5, 0.5, Sj, Sj);
  int sigma = 0;
  double norm = 1.0;
  if (std::fabs(cgm) > 1e-20) { sigma = -1; norm = cgm; }
  else if (std::fabs(cgp) > 1e-20) { sigma = 1; norm = cgp; }
  if (sigma == 0) return 0.0;
  double acc = 0.0;
  for (double szc = Scfj; szc >= -Scfj; szc -= 1.0) {
    double szi = Si - szc, szj = Sj - szc;
    int bsi = st.find(Qti, Sti, szi), bsj = st.find(Qtj, Stj, szj);
    double c = cgordan(Scfi, szc, Sti, szi, Si, Si) * cgordan(Scfj, szc, Stj, szj, Sj, Sj);
    int a0, a1, b0, b1;
    st.range(bsi, a0, a1);
    st.range(bsj, b0, b1);
    for (int ssi = a0; ssi <= a1; ssi++)
      for (int ssj = b0; ssj <= b1; ssj++)
        acc += T.f(ch, sigma, ssj, ssi) * c * TwoChTables::spsm(ssi, ti) * TwoChTables::spsm(ssj, tj);
  }
  return acc / norm;
}

// ---------------------------------------------------------------------------
// TwoChQSz plugin: TwoChQSz_SetSingleSite (src/TwoChQSz/TwoChQSz_SetSingleSite.cpp:5-85,
// 16 one-state (Q, Sz) blocks), fd_table (src/TwoChQSz/TwoChQSz_OpTables.cpp:3-78),
// H_N rules of TwoChQSz_DiagHN (src/TwoChQSz/TwoChQSz_DiagHN.cpp:74-97) and
// TwoChQSz_SetH0Anderson (src/TwoChQSz/TwoChQSz_SetH0Anderson.cpp:180-215), f_ch
// rule of TwoChQSz_UpdateMatrixAfterCutting (src/TwoChQSz/TwoChQSz_UpdateMatrices.cpp:288-289)
// ---------------------------------------------------------------------------
static SiteDef twochqsz_site() {
  static const double q[16][2] = {{-2, 0}, {-1, 0.5}, {-1, -0.5}, {0, 0}, {-1, 0.5}, {0, 1}, {0, 0}, {1, 0.5},
                                  {-1, -0.5}, {0, 0}, {0, -1}, {1, -0.5}, {0, 0}, {1, 0.5}, {1, -0.5}, {2, 0}};
  SiteDef s;
  for (int t = 0; t < 16; t++) s.add_block(q[t][0], std::fabs(q[t][1]), q[t][1], 1);
  return s;
}
// <t_i| f^dagger_{ch,sigma} |t_j>; state bits 1 = ch2 up, 2 = ch2 down, 4 = ch1 up,
// 8 = ch1 down; the table's signs: ch1 down -1 past ch1 up, ch2 up -1 past an odd
// number of ch1 modes, ch2 down that sign and -1 past ch2 up
static double qsz_fd(int ch, int sigma, int ti, int tj) {
  if (ti < 0 || ti > 15 || tj < 0 || tj > 15) return 0.0;
  const int bit = ch == 1 ? (sigma == 1 ? 4 : (sigma == -1 ? 8 : 0)) : (ch == 2 ? (sigma == 1 ? 1 : (sigma == -1 ? 2 : 0)) : 0);
  if (!bit || (tj & bit) || ti != (tj | bit)) return 0.0;
  if (bit == 4) return 1.0;
  if (bit == 8) return (tj & 4) ? -1.0 : 1.0;
  const double sg = (((tj >> 2) & 1) ^ ((tj >> 3) & 1)) ? -1.0 : 1.0;
  if (bit == 1) return sg;
  return (tj & 1) ? -sg : sg;
}
// H_N coupling: sum over sigma of fd(ch, sigma, typep, type) times the fermion sign
// of the site state `type` (-1 for Sz = +-1/2; at N = 0 also Q = +-1)
static double qsz_hn_coef(const SiteDef& st, int ch, int typep, int type, int h0) {
  const double Qt = st.Q(type), Szt = st.Sz(type);
  double fs = 1.0;
  if (h0) { if ((deq(Qt, 1.0) || deq(Qt, -1.0)) && (deq(Szt, 0.5) || deq(Szt, -0.5))) fs = -1.0; }
  else if (deq(Szt, 0.5) || deq(Szt, -0.5)) fs = -1.0;
  double v = 0.0;
  for (int sg = -1; sg <= 1; sg += 2) v += qsz_fd(ch, sg, typep, type) * fs;
  return v;
}
static double qsz_fn_el(int ch, int ti, int tj) { return qsz_fd(ch, -1, tj, ti) + qsz_fd(ch, 1, tj, ti); }

// QSz_BuildBasis (src/QSz_BuildBasis.cpp:9-117): blocks (Q_old+Q_site, Sz_old+Sz_site)
// in first-appearance order; per new block: old block, site block, old state
static void build_basis_qsz(const CNRGbasisarray& Ac, CNRGbasisarray& Ab, const SiteDef& site, int befcut) {
  Ab = CNRGbasisarray();
  Ab.Nshell = Ac.Nshell + 1;
  Ab.NQNumbers = 2;
  std::map<std::pair<long, long>, int> seen;
  std::vector<std::pair<double, double>> qns;
  for (int b = 0; b < Ac.NumBlocks(); b++)
    for (int s = 0; s < site.nblocks(); s++) {
      const double q = Ac.GetQNumber(b, 0) + site.bQ[s], sz = Ac.GetQNumber(b, 1) + site.bSz[s];
      if (seen.emplace(std::make_pair(key2(q), key2(sz)), (int)qns.size()).second) qns.push_back({q, sz});
    }
  int ist = 0;
  for (auto& qs : qns) {
    Ab.QNumbers.push_back(qs.first);
    Ab.QNumbers.push_back(qs.second);
    Ab.BlockBegEnd.push_back(ist);
    for (int bo = 0; bo < Ac.NumBlocks(); bo++) {
      const double Qo = Ac.GetQNumber(bo, 0), So = Ac.GetQNumber(bo, 1);
      for (int s = 0; s < site.nblocks(); s++) {
        if (key2(qs.first) != key2(Qo + site.bQ[s]) || key2(qs.second) != key2(So + site.bSz[s])) continue;
        for (int io = Ac.GetBlockLimit(bo, 0); io <= Ac.GetBlockLimit(bo, 1); io++)
          for (int is = site.bbeg[s]; is <= site.bend[s]; is++) {
            Ab.iType.push_back(is);
            Ab.dEn.push_back(Ac.dEn[io]);
            Ab.StCameFrom.push_back(befcut == 1 ? Ac.StCameFrom[io] : io);
            ist++;
          }
      }
    }
    Ab.BlockBegEnd.push_back(ist - 1);
  }
}

// ---------------------------------------------------------------------------
// basis construction.  New blocks in first-appearance order of (Q_old+Q_site,
// S from S_old+S_site down to |S_old-S_site|) (FilterQNumbers keeps first
// occurrences); a state (old state o, site state s) enters block (Q,S) when
// Q = Q_o + Q_s, S = S_o + Sz_s and (S_o, S_s, S) obey the triangle rule.
//   order 0: BuildBasis      (src/BuildBasis.cpp:8-432): old block, site block, old state, site state
//   order 1: QS_BuildBasis   (src/QS_BuildBasis.cpp:9-169): old block, site block, site state, old state
// ---------------------------------------------------------------------------
static void build_basis(const CNRGbasisarray& Ac, CNRGbasisarray& Ab, const SiteDef& site, int order, int befcut) {
  Ab = CNRGbasisarray();
  Ab.Nshell = Ac.Nshell + 1;
  Ab.NQNumbers = 2;
  std::map<std::pair<long, long>, int> seen;
  std::vector<std::pair<double, double>> qns;
  for (int b = 0; b < Ac.NumBlocks(); b++) {
    double Qo = Ac.GetQNumber(b, 0), So = Ac.GetQNumber(b, 1);
    for (int s = 0; s < site.nblocks(); s++)
      for (double sn = So + site.bS[s]; sn >= std::fabs(So - site.bS[s]); sn -= 1.0) {
        auto k = std::make_pair(key2(Qo + site.bQ[s]), key2(sn));
        if (seen.emplace(k, (int)qns.size()).second) qns.push_back({Qo + site.bQ[s], sn});
      }
  }
  int ist = 0;
  for (auto& qs : qns) {
    Ab.QNumbers.push_back(qs.first);
    Ab.QNumbers.push_back(qs.second);
    Ab.BlockBegEnd.push_back(ist);
    for (int bo = 0; bo < Ac.NumBlocks(); bo++) {
      double Qo = Ac.GetQNumber(bo, 0), So = Ac.GetQNumber(bo, 1);
      for (int s = 0; s < site.nblocks(); s++) {
        if (key2(qs.first) != key2(site.bQ[s] + Qo)) continue;
        if (key2(qs.second) != key2(site.bSz[s] + So) || !triangle(So, site.bS[s], qs.second)) continue;
        const int o0 = Ac.GetBlockLimit(bo, 0), o1 = Ac.GetBlockLimit(bo, 1);
        auto push = [&](int io, int is) {
          Ab.iType.push_back(is);
          Ab.dEn.push_back(Ac.dEn[io]);
          Ab.StCameFrom.push_back(befcut == 1 ? Ac.StCameFrom[io] : io);
          ist++;
        };
        if (order == 0) {
          for (int io = o0; io <= o1; io++)
            for (int is = site.bbeg[s]; is <= site.bend[s]; is++) push(io, is);
        } else {
          for (int is = site.bbeg[s]; is <= site.bend[s]; is++)
            for (int io = o0; io <= o1; io++) push(io, is);
        }
      }
    }
    Ab.BlockBegEnd.push_back(ist - 1);
  }
}

// ---------------------------------------------------------------------------
// device views
// ---------------------------------------------------------------------------
struct BasisDev {  // per-state block/position and per-block size of a kept basis
  int* ob = nullptr;
  int* op = nullptr;
  int* bsz = nullptr;
  int nb = 0;
};

static int make_basis_dev(nrg_ctx* c, const CNRGbasisarray& A, const char* tag, BasisDev& bd) {
  std::vector<int> ob, op, bsz;
  for (int b = 0; b < A.NumBlocks(); b++) {
    int n = A.GetBlockSize(b);
    bsz.push_back(n);
    for (int i = 0; i < n; i++) { ob.push_back(b); op.push_back(i); }
  }
  std::string t(tag);
  int rc;
  if ((rc = upload(c, (t + "_ob").c_str(), ob, &bd.ob))) return rc;
  if ((rc = upload(c, (t + "_op").c_str(), op, &bd.op))) return rc;
  if ((rc = upload(c, (t + "_bsz").c_str(), bsz, &bd.bsz))) return rc;
  bd.nb = A.NumBlocks();
  return 0;
}

static int make_oldop(nrg_ctx* c, const CNRGmatrix& M, const BasisDev& bd, const std::string& tag, OldOp& o) {
  std::vector<int64_t> mtab((size_t)bd.nb * bd.nb, -1);
  for (size_t q = 0; q + 1 < M.MatBlockMap.size(); q += 2) {
    int64_t& slot = mtab[(size_t)M.MatBlockMap[q] * bd.nb + M.MatBlockMap[q + 1]];
    if (slot < 0) slot = M.MatBlockBegEnd[q];  // FindMatBlock returns the first match
  }
  int64_t* dm = nullptr;
  int rc = upload(c, (tag + "_mtab").c_str(), mtab, &dm);
  if (rc) return rc;
  o.mat = (const double*)c->buf(M.buf[M.cur]).p;
  o.mtab = dm;
  o.ob = bd.ob;
  o.op = bd.op;
  o.bsz = bd.bsz;
  o.nbo = bd.nb;
  return 0;
}

// ---------------------------------------------------------------------------
// thermodynamics (src/CalcThermo.cpp:16-121)
// ---------------------------------------------------------------------------
static double calc_suscep(double bb, const CNRGarray& A) {
  double s0 = 0, s1 = 0;
  for (int b = 0; b < A.NumBlocks(); b++) {
    double aS = 2.0 * A.GetQNumber(b, 1) + 1.0, szsq = (1.0 / 12.0) * aS * (aS * aS - 1.0);
    for (int i = A.GetBlockLimit(b, 0); i <= A.GetBlockLimit(b, 1); i++) {
      double x = std::exp(-bb * A.dEn[i]);
      s0 += aS * x; s1 += szsq * x;
    }
  }
  return s1 / s0;
}
// CalcSuscep / CalcEntropy with totalS false (src/CalcThermo.cpp:16-121): <Sz^2>
// with Sz the block's second quantum number, degeneracy 1 per state
static double calc_suscep_sz(double bb, const CNRGarray& A) {
  double s0 = 0, s1 = 0;
  for (int b = 0; b < A.NumBlocks(); b++) {
    const double Sz = A.GetQNumber(b, 1);
    for (int i = A.GetBlockLimit(b, 0); i <= A.GetBlockLimit(b, 1); i++) {
      const double x = std::exp(-bb * A.dEn[i]);
      s0 += x; s1 += Sz * Sz * x;
    }
  }
  return s1 / s0;
}
static double calc_entropy_sz(double bb, const CNRGarray& A) {
  double Z = 0, s = 0;
  for (size_t i = 0; i < A.dEn.size(); i++) {
    const double E = A.dEn[i], x = std::exp(-bb * E);
    s += E * x; Z += x;
  }
  return bb * (s / Z) + std::log(Z);
}
static double calc_entropy(double bb, const CNRGarray& A) {
  double Z = 0, s = 0;
  for (int b = 0; b < A.NumBlocks(); b++) {
    double dg = 2.0 * A.GetQNumber(b, 1) + 1.0;
    for (int i = A.GetBlockLimit(b, 0); i <= A.GetBlockLimit(b, 1); i++) {
      double E = A.dEn[i], x = std::exp(-bb * E);
      s += dg * E * x; Z += dg * x;
    }
  }
  return bb * (s / Z) + std::log(Z);
}
// calc_suscep and calc_entropy in one pass: each Boltzmann factor is evaluated
// once and feeds both sets of sums, which are accumulated in the same order as
// in the two functions (so the results are theirs, bit for bit)
static void calc_suscep_entropy(double bb, const CNRGarray& A, double& sus, double& ent) {
  double s0 = 0, s1 = 0, Z = 0, s = 0;
  for (int b = 0; b < A.NumBlocks(); b++) {
    const double aS = 2.0 * A.GetQNumber(b, 1) + 1.0, szsq = (1.0 / 12.0) * aS * (aS * aS - 1.0);
    for (int i = A.GetBlockLimit(b, 0); i <= A.GetBlockLimit(b, 1); i++) {
      const double E = A.dEn[i], x = std::exp(-bb * E);
      s0 += aS * x; s1 += szsq * x;
      s += aS * E * x; Z += aS * x;
    }
  }
  sus = s1 / s0;
  ent = bb * (s / Z) + std::log(Z);
}

// operator update specification
enum PairRule { PAIR_DIAG = 0, PAIR_QS_UP = 1, PAIR_QS_UP_UPPER = 2, PAIR_QSZ_UP = 3 };
struct OpSpec {
  bool need_old = false;
  bool calc_avg = false;
  PairRule rule = PAIR_DIAG;
  std::function<double(double, double, int, int)> g;  // site-type value for (S_i, S_j, t_i, t_j)
  // site-type tables already evaluated, keyed by (2S_i, 2S_j): the rules depend
  // on the block spins only, so each table is computed once per run
  std::map<std::pair<long, long>, std::vector<double>> cache;
  const std::vector<double>& table(double Si, double Sj, int ntypes) {
    auto key = std::make_pair(key2(Si), key2(Sj));
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<double> t((size_t)ntypes * ntypes);
    for (int a = 0; a < ntypes; a++)
      for (int b = 0; b < ntypes; b++) t[(size_t)a * ntypes + b] = g(Si, Sj, a, b);
    return cache.emplace(key, std::move(t)).first->second;
  }
};

// CheckForMatEl rules: Diag_check (src/OpMatRules.cpp:12-24), OneChQS_cd_check
// (:106-125); the two-channel update only visits jbl > ibl
// (src/TwoChQS/TwoChQS_UpdateMatrices.cpp:333)
static bool pair_ok(PairRule r, const CNRGarray& A, int i, int j) {
  if (r == PAIR_DIAG) return i == j;
  if (r == PAIR_QSZ_UP)   // d_up on (Q, Sz) blocks: (Q, Sz) -> (Q+1, Sz+1/2)
    return deq(A.GetQNumber(j, 0), A.GetQNumber(i, 0) + 1.0) && deq(A.GetQNumber(j, 1), A.GetQNumber(i, 1) + 0.5);
  if (r == PAIR_QS_UP_UPPER && j <= i) return false;
  double Qi = A.GetQNumber(i, 0), Si = A.GetQNumber(i, 1), Qj = A.GetQNumber(j, 0), Sj = A.GetQNumber(j, 1);
  return deq(Qj, Qi + 1.0) && (deq(Sj, Si + 0.5) || deq(Sj, Si - 0.5));
}

// ---------------------------------------------------------------------------
// the step machinery shared by the chain drivers
// ---------------------------------------------------------------------------
struct Chain {
  nrg_ctx* c;
  std::string tag;   // prefix of this chain's persistent device buffers
  SiteDef site;
  int ntypes = 4;
  CNRGarray Aeig;
  CNRGbasisarray AeigCut, Abasis;
  std::vector<CNRGmatrix> mats;
  std::vector<OpSpec> specs;
  std::vector<int64_t> Aeig_zoff;  // offset of each Aeig block in the Z buffer
  double* dZ = nullptr;
  double* dW = nullptr;
  BasisDev oldb;                   // device view of the basis the current ops live on
  int zpar = 0;
  bool record = false;
  bool sz_only = false;            // (Q, Sz) blocks: one coupling table (no spin index)
  // the eigenvectors are formed after the truncation, kept rows only (nothing
  // downstream reads a discarded state's eigenvector); finish_vectors forms them
  bool defer = false;
  bool vec_pending = false;
  std::vector<double> hrec, zrec;

  // rows of each Aeig block that AeigCut keeps (a prefix of the block), at krow[base + b]
  void kept_rows(std::vector<int>& krow, int base) const {
    const int nb = Aeig.NumBlocks();
    for (int b = 0; b < nb; b++) krow[base + b] = 0;
    for (int q = 0; q < AeigCut.NumBlocks(); q++) {
      const int first = AeigCut.BlockBegEndBC[2 * q];
      for (int b = 0; b < nb; b++)
        if (Aeig.GetBlockLimit(b, 0) == first) { krow[base + b] = AeigCut.GetBlockSize(q); break; }
    }
  }
  // the deferred eigenvectors of this chain's last eigensolve: the kept rows, or every row
  int finish_vectors(bool all) {
    if (!vec_pending) return 0;
    vec_pending = false;
    if (all) return eig_finish(c, nullptr);
    std::vector<int> krow(Aeig.NumBlocks());
    kept_rows(krow, 0);
    return eig_finish(c, krow.data());
  }

  void init_ops(int nm) {
    mats.assign(nm, CNRGmatrix());
    for (int m = 0; m < nm; m++) {
      mats[m].buf[0] = tag + "op" + std::to_string(m) + "_a";
      mats[m].buf[1] = tag + "op" + std::to_string(m) + "_b";
      mats[m].NeedOld = specs[m].need_old;
      mats[m].CalcAvg = specs[m].calc_avg;
    }
  }

  int set_op(int m, const std::vector<int>& map, const std::vector<int>& be, const std::vector<double>& el) {
    CNRGmatrix& M = mats[m];
    M.MatBlockMap = map;
    M.MatBlockBegEnd = be;
    M.cur = 0;
    DBuf& b = c->buf(M.buf[0]);
    int rc = b.reserve(sizeof(double) * (el.size() + 1));
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(b.p, el.data(), sizeof(double) * el.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  }

  // H_N assembly on the device for the blocks of Abasis, then the batched
  // eigensolve; Aeig receives the raw (unshifted) eigenvalues.  Three phases so
  // that several chains can share one eigensolve: layout (block sizes and
  // offsets in a shared Z / W buffer), assemble (H into Z), collect (eigen-
  // values back, Z rows stay on the device).
  //   fops[ch]  operator used as f_ch;  coef[ch]  [S index][typep][type] table
  std::vector<int> L_bsize;
  std::vector<int64_t> L_hoff, L_woff;
  int64_t L_ztot = 0, L_wtot = 0;
  int L_nmax = 0;

  void layout(int64_t zbase, int64_t wbase) {
    const int nb = Abasis.NumBlocks();
    L_bsize.assign(nb, 0); L_hoff.assign(nb, 0); L_woff.assign(nb, 0);
    int64_t z = 0;
    L_nmax = 0;
    for (int b = 0; b < nb; b++) {
      L_bsize[b] = Abasis.GetBlockSize(b);
      L_hoff[b] = zbase + z;
      L_woff[b] = wbase + Abasis.GetBlockLimit(b, 0);
      z += (int64_t)L_bsize[b] * L_bsize[b];
      L_nmax = std::max(L_nmax, L_bsize[b]);
    }
    L_ztot = z;
    L_wtot = Abasis.Nstates();
  }

  int assemble(int nch, const int* fops, const double* chi, const std::vector<double>* coef, double sqrtL, double eN,
               double* dZn) {
    int rc;
    const int nb = Abasis.NumBlocks();
    std::vector<int> bbeg(nb), sidx(nb);
    for (int b = 0; b < nb; b++) {
      bbeg[b] = Abasis.GetBlockLimit(b, 0);
      sidx[b] = sz_only ? 0 : (int)key2(Abasis.GetQNumber(b, 1));
    }
    int *d_bbeg, *d_bsize, *d_sidx, *d_stcf, *d_itype;
    int64_t* d_hoff;
    double *d_en, *d_siteq;
    if ((rc = upload(c, "hb_bbeg", bbeg, &d_bbeg))) return rc;
    if ((rc = upload(c, "hb_bsize", L_bsize, &d_bsize))) return rc;
    if ((rc = upload(c, "hb_sidx", sidx, &d_sidx))) return rc;
    if ((rc = upload(c, "hb_hoff", L_hoff, &d_hoff))) return rc;
    if ((rc = upload(c, "hb_stcf", Abasis.StCameFrom, &d_stcf))) return rc;
    if ((rc = upload(c, "hb_itype", Abasis.iType, &d_itype))) return rc;
    if ((rc = upload(c, "hb_en", Abasis.dEn, &d_en))) return rc;
    std::vector<double> siteq(ntypes);
    for (int t = 0; t < ntypes; t++) siteq[t] = site.Q(t);
    if ((rc = upload(c, "hb_siteq", siteq, &d_siteq))) return rc;
    HnArgs ha{};
    ha.H = dZn; ha.h_off = d_hoff; ha.bbeg = d_bbeg; ha.bsize = d_bsize; ha.blk_sidx = d_sidx;
    ha.stcf = d_stcf; ha.itype = d_itype; ha.en_par = d_en; ha.site_q = d_siteq;
    for (int ch = 0; ch < nch; ch++) {
      double* dcoef;
      std::string nm = "hb_coef" + std::to_string(ch);
      if ((rc = upload(c, nm.c_str(), coef[ch], &dcoef))) return rc;
      ha.coef[ch] = dcoef;
      if ((rc = make_oldop(c, mats[fops[ch]], oldb, "hf" + std::to_string(ch), ha.f[ch]))) return rc;
      ha.chi[ch] = chi[ch];
    }
    ha.sqrtL = sqrtL;
    ha.eN = eN;
    ha.nch = nch;
    ha.ntypes = ntypes;
    {
      double by = 0.0;
      for (int b = 0; b < nb; b++) by += 8.0 * (double)L_bsize[b] * L_bsize[b];
      PendingTimer tm;
      if ((rc = c->timer_begin(tm, NRG_K_HN, 0.0, by))) return rc;
      HIPCHK(nrgk_launch_hn(&ha, nb, L_nmax, c->stream));
      if ((rc = c->timer_end(tm))) return rc;
    }
    hrec.clear();
    if (record) {
      std::vector<double> hfull(L_ztot);
      HIPCHK(hipMemcpyAsync(hfull.data(), dZn + L_hoff[0], sizeof(double) * L_ztot, hipMemcpyDeviceToHost,
                            c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      for (int b = 0; b < nb; b++) {
        int n = L_bsize[b];
        const int64_t o = L_hoff[b] - L_hoff[0];
        for (int i = 0; i < n; i++)
          for (int j = i; j < n; j++) hrec.push_back(hfull[o + (int64_t)i * n + j]);
      }
    }
    return 0;
  }

  // the eigenvalues from a host copy of the whole eigenvalue buffer (wall[i] =
  // dWn[i]); no device call, so instances can do this on host threads
  void collect_host(const double* wall, double* dZn, double* dWn) {
    const int ns = Abasis.Nstates();
    Aeig = CNRGarray();
    Aeig.Nshell = Abasis.Nshell; Aeig.NQNumbers = 2;
    Aeig.QNumbers = Abasis.QNumbers;
    Aeig.BlockBegEnd = Abasis.BlockBegEnd;
    const int64_t w0 = L_woff.empty() ? 0 : L_woff[0] - Abasis.GetBlockLimit(0, 0);
    Aeig.dEn.assign(wall + w0, wall + w0 + ns);
    zrec.clear();
    Aeig_zoff = L_hoff;
    dZ = dZn;
    dW = dWn;
  }
  int collect(double* dZn, double* dWn) {
    const int ns = Abasis.Nstates();
    Aeig = CNRGarray();
    Aeig.Nshell = Abasis.Nshell; Aeig.NQNumbers = 2;
    Aeig.QNumbers = Abasis.QNumbers;
    Aeig.BlockBegEnd = Abasis.BlockBegEnd;
    Aeig.dEn.resize(ns);
    const int64_t w0 = L_woff.empty() ? 0 : L_woff[0] - Abasis.GetBlockLimit(0, 0);
    HIPCHK(hipMemcpyAsync(Aeig.dEn.data(), dWn + w0, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
    zrec.clear();
    if (record) {
      zrec.resize(L_ztot);
      HIPCHK(hipMemcpyAsync(zrec.data(), dZn + L_hoff[0], sizeof(double) * L_ztot, hipMemcpyDeviceToHost,
                            c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    Aeig_zoff = L_hoff;
    dZ = dZn;
    dW = dWn;
    return 0;
  }

  int diag_step(int nch, const int* fops, const double* chi, const std::vector<double>* coef, double sqrtL,
                double eN) {
    int rc;
    layout(0, 0);
    zpar ^= 1;
    DBuf& zb = c->buf(tag + (zpar ? "Z1" : "Z2"));
    DBuf& wb = c->buf(tag + "W");
    if ((rc = zb.reserve(sizeof(double) * (size_t)(L_ztot + 1)))) return rc;
    if ((rc = wb.reserve(sizeof(double) * (size_t)(L_wtot + 1)))) return rc;
    double* dZn = (double*)zb.p;
    double* dWn = (double*)wb.p;
    if ((rc = assemble(nch, fops, chi, coef, sqrtL, eN, dZn))) return rc;
    const int nb = Abasis.NumBlocks();
    if (defer && !record) {
      if ((rc = eig_begin(c, nb, L_bsize.data(), L_hoff.data(), dZn, L_woff.data(), dWn))) return rc;
      vec_pending = true;
    } else if ((rc = eig_dev(c, nb, L_bsize.data(), L_hoff.data(), dZn, L_woff.data(), dWn, nullptr))) {
      return rc;
    }
    return collect(dZn, dWn);
  }

  // CutStates (src/CutStates.cpp:5-96; by-value form :103-160): AeigCut keeps,
  // per surviving block, the energy-ordered prefix up to the first level above
  // Ecut; blocks whose lowest level is above Ecut are dropped.  Eigenvector
  // rows are not moved: AeigCut records where each block's kept rows start in
  // the Z buffer.  legacy=false repeats SetKept(Ncut) (SetDegen + Ecut) on the
  // copy as CutStates does; the by-value form of the two-channel driver does not.
  //   fixed != nullptr: the given Kept set (Aeig state order; each block keeps as
  //   many of its lowest states as it has kept flags) instead of SetKept.
  void cut_states(int ncut, bool legacy, double myeps, const unsigned char* fixed = nullptr) {
    CNRGbasisarray A;
    A.Nshell = Aeig.Nshell; A.NQNumbers = Aeig.NQNumbers;
    int nb = Aeig.NumBlocks();
    const bool notrunc = !fixed && ncut >= Aeig.Nstates();
    std::vector<double> en = Aeig.dEn;
    std::vector<unsigned char> kept = Aeig.Kept;
    double ec = HUGE_VAL;
    if (fixed) {
      kept.assign(fixed, fixed + en.size());
    } else if (!notrunc) {
      ec = ecut_of(Aeig.dEn, ncut);
      if (!legacy) {
        set_degen(en, myeps);
        double ec2 = ecut_of(en, ncut);
        kept.assign(en.size(), 0);
        for (size_t i = 0; i < en.size(); i++) kept[i] = en[i] <= ec2;
      }
    }
    int ist = 0;
    for (int b = 0; b < nb; b++) {
      int b0 = Aeig.GetBlockLimit(b, 0), b1 = Aeig.GetBlockLimit(b, 1);
      int nbc = b1 - b0 + 1;
      int k = nbc;
      if (fixed) {
        k = 0;
        for (int i = b0; i <= b1; i++) k += kept[i] != 0;
        if (k == 0) continue;
      } else if (!notrunc) {
        double emin = *std::min_element(en.begin() + b0, en.begin() + b1 + 1);
        if (emin > ec) continue;
        double emax = *std::max_element(en.begin() + b0, en.begin() + b1 + 1);
        if (emax > ec) {
          int ip = b0;
          while (ip <= b1 && en[ip] <= ec) ip++;
          k = ip - b0;
          if (k == 0) continue;  // RemoveStatesFromBlock of a one-state block removes it
        }
      }
      A.QNumbers.push_back(Aeig.GetQNumber(b, 0));
      A.QNumbers.push_back(Aeig.GetQNumber(b, 1));
      A.BlockBegEnd.push_back(ist);
      A.BlockBegEnd.push_back(ist + k - 1);
      A.BlockBegEndBC.push_back(b0);
      A.BlockBegEndBC.push_back(b1);
      for (int i = 0; i < k; i++) {
        A.dEn.push_back(en[b0 + i]);
        A.Kept.push_back(b0 + i < (int)kept.size() ? kept[b0 + i] : 1);
        A.StCameFrom.push_back(b0 + i);
      }
      A.zoff.push_back(Aeig_zoff[b]);
      A.zld.push_back(nbc);
      ist += k;
    }
    AeigCut = std::move(A);
  }

  // UpdateMatrices (src/UpdateMatrices.cpp:272-568): for each (ib, jb) of
  // AeigCut in loop order and each operator whose pair rule holds,
  // F = G[t_i][t_j] * (NeedOld ? O_old(stcf_i,stcf_j) : delta(stcf_i,stcf_j))
  // over the pre-cut blocks, then O_new = Z_i F Z_j^T on the device.
  //   plan_update: pair list, F on the device; finish_update: the rotations and
  //   the switch of every operator to AeigCut.  Between the two, F (upd_F), the
  //   jobs and the previous operators with their block maps are all valid.
  std::vector<FPair> up_pairs;
  std::vector<RotJob> up_jobs;
  std::vector<int> up_job_op;
  std::vector<std::vector<int>> up_nmap, up_nbe;
  std::vector<int64_t> up_icount;
  int64_t up_ftot = 0, up_auxtot = 0;
  FArgs up_fa{};
  int update_ops() {
    int rc;
    if ((rc = plan_update())) return rc;
    return finish_update();
  }
  int plan_update() {
    int rc;
    const int nm = (int)mats.size();
    std::vector<FPair>& pairs = up_pairs;
    std::vector<double> G;
    std::vector<RotJob>& jobs = up_jobs;
    std::vector<int>& job_op = up_job_op;
    pairs.clear(); jobs.clear(); job_op.clear();
    up_nmap.assign(nm, {}); up_nbe.assign(nm, {});
    up_icount.assign(nm, 0);
    std::vector<std::vector<int>>& nmap = up_nmap;
    std::vector<std::vector<int>>& nbe = up_nbe;
    std::vector<int64_t>& icount = up_icount;
    int64_t ftot = 0, auxtot = 0, maxel = 1;
    const int nbc = AeigCut.NumBlocks();
    for (int ib = 0; ib < nbc; ib++) {
      int ki = AeigCut.GetBlockSize(ib);
      int ibc_beg = AeigCut.BlockBegEndBC[2 * ib];
      int ni = AeigCut.BlockBegEndBC[2 * ib + 1] - ibc_beg + 1;
      double Si = AeigCut.GetQNumber(ib, 1);
      for (int jb = 0; jb < nbc; jb++) {
        int kj = AeigCut.GetBlockSize(jb);
        int jbc_beg = AeigCut.BlockBegEndBC[2 * jb];
        int nj = AeigCut.BlockBegEndBC[2 * jb + 1] - jbc_beg + 1;
        double Sj = AeigCut.GetQNumber(jb, 1);
        for (int m = 0; m < nm; m++) {
          if (!pair_ok(specs[m].rule, AeigCut, ib, jb)) continue;
          nmap[m].push_back(ib); nmap[m].push_back(jb);
          nbe[m].push_back((int)icount[m]);
          int64_t out_off = icount[m];
          icount[m] += (int64_t)ki * kj;
          nbe[m].push_back((int)icount[m] - 1);
          FPair P{};
          P.ni_bc = ni; P.nj_bc = nj; P.ibeg_bc = ibc_beg; P.jbeg_bc = jbc_beg;
          P.f_off = ftot; P.g_off = (int64_t)G.size(); P.need_old = specs[m].need_old ? 1 : 0; P.op = m;
          const std::vector<double>& gt = specs[m].table(Si, Sj, ntypes);
          G.insert(G.end(), gt.begin(), gt.end());
          pairs.push_back(P);
          RotJob J{};
          J.ni = ni; J.nj = nj; J.ki = ki; J.kj = kj;
          J.ldzi = AeigCut.zld[ib]; J.ldzj = AeigCut.zld[jb];
          J.zi_off = AeigCut.zoff[ib]; J.zj_off = AeigCut.zoff[jb];
          J.f_off = ftot; J.aux_off = auxtot; J.out_off = out_off;
          jobs.push_back(J);
          job_op.push_back(m);
          ftot += (int64_t)ni * nj;
          auxtot += (int64_t)kj * ni;
          maxel = std::max(maxel, (int64_t)ni * nj);
        }
      }
    }
    DBuf& fb = c->buf("upd_F");
    DBuf& ab = c->buf("upd_aux");
    if ((rc = fb.reserve(sizeof(double) * (size_t)(ftot + 1)))) return rc;
    if ((rc = ab.reserve(sizeof(double) * (size_t)(auxtot + 1)))) return rc;
    if (!pairs.empty()) {
      FPair* d_pairs;
      double* d_G;
      int *d_stcf, *d_itype;
      if ((rc = upload(c, "upd_pairs", pairs, &d_pairs))) return rc;
      if ((rc = upload(c, "upd_G", G, &d_G))) return rc;
      if ((rc = upload(c, "upd_stcf", Abasis.StCameFrom, &d_stcf))) return rc;
      if ((rc = upload(c, "upd_itype", Abasis.iType, &d_itype))) return rc;
      FArgs fa{};
      fa.F = (double*)fb.p; fa.pairs = d_pairs; fa.G = d_G; fa.stcf = d_stcf; fa.itype = d_itype;
      fa.ntypes = ntypes;
      for (int m = 0; m < nm && m < 8; m++)
        if (specs[m].need_old)
          if ((rc = make_oldop(c, mats[m], oldb, "uo" + std::to_string(m), fa.old[m]))) return rc;
      PendingTimer tm;
      if ((rc = c->timer_begin(tm, NRG_K_FBASIS, 0.0, 8.0 * (double)ftot))) return rc;
      HIPCHK(nrgk_launch_fbasis(&fa, (int)pairs.size(), maxel, c->stream));
      if ((rc = c->timer_end(tm))) return rc;
      up_fa = fa;
    }
    up_ftot = ftot;
    up_auxtot = auxtot;
    return 0;
  }
  int finish_update() {
    int rc;
    const int nm = (int)mats.size();
    DBuf& fb = c->buf("upd_F");
    DBuf& ab = c->buf("upd_aux");
    for (int m = 0; m < nm; m++) {
      std::vector<RotJob> jm;
      for (size_t q = 0; q < up_jobs.size(); q++)
        if (up_job_op[q] == m) jm.push_back(up_jobs[q]);
      CNRGmatrix& M = mats[m];
      int nxt = M.cur ^ 1;
      DBuf& ob = c->buf(M.buf[nxt]);
      if ((rc = ob.reserve(sizeof(double) * (size_t)(up_icount[m] + 1)))) return rc;
      if ((rc = rotate_dev(c, jm, dZ, (double*)fb.p, (double*)ab.p, (double*)ob.p))) return rc;
      M.cur = nxt;
      M.MatBlockMap = up_nmap[m];
      M.MatBlockBegEnd = up_nbe[m];
    }
    // no wait here: what reads the new operators (the next step's f-basis, the
    // averages' gather) is ordered after the rotations on the stream
    if ((rc = c->resolve())) return rc;
    // the new operators live on AeigCut
    return make_basis_dev(c, AeigCut, (tag + "oldb").c_str(), oldb);
  }

  // CalcOpAvg (src/CalcOpAvg.cpp:13-84) on AeigCut for the operators with CalcAvg.
  // Only the diagonals of the diagonal blocks enter; they are gathered on the
  // device for every such operator and come back in one transfer.  In two
  // halves for many chains at once: op_avgs_issue gathers them into ddst
  // (op_avgs_len() doubles, stream-ordered), op_avgs_sum adds them up from a
  // host copy.
  std::vector<int> oa_avg;
  std::vector<unsigned char> oa_has;
  int64_t op_avgs_len() const {
    int na = 0;
    for (size_t m = 0; m < mats.size(); m++) na += mats[m].CalcAvg ? 1 : 0;
    return AeigCut.dEn.empty() ? 0 : (int64_t)na * (int64_t)AeigCut.dEn.size();
  }
  int op_avgs_issue(double* ddst) {
    const int nb = AeigCut.NumBlocks();
    const int64_t ns = AeigCut.dEn.empty() ? 0 : (int64_t)AeigCut.dEn.size();
    std::vector<int>& avg = oa_avg;
    avg.clear();
    for (size_t m = 0; m < mats.size(); m++)
      if (mats[m].CalcAvg) avg.push_back((int)m);
    oa_has.assign((size_t)avg.size() * nb, 0);
    if (avg.empty()) return 0;
    std::vector<DiagJob> jobs;
    for (size_t a = 0; a < avg.size(); a++) {
      const CNRGmatrix& M = mats[avg[a]];
      const double* base = (const double*)c->buf(M.buf[M.cur]).p;
      for (size_t q = 0; q + 1 < M.MatBlockMap.size(); q += 2) {
        const int b = M.MatBlockMap[q];
        if (b != M.MatBlockMap[q + 1] || b < 0 || b >= nb || oa_has[a * nb + b]) continue;
        oa_has[a * nb + b] = 1;   // the first (b, b) entry of the map, as a linear search finds it
        jobs.push_back(DiagJob{base + M.MatBlockBegEnd[q], (int64_t)a * ns + AeigCut.GetBlockLimit(b, 0),
                               AeigCut.GetBlockSize(b), 0});
      }
    }
    if (!jobs.empty() && ns > 0) {
      DiagJob* dj;
      int rc = upload(c, (tag + "opavg_jobs").c_str(), jobs, &dj);
      if (rc) return rc;
      HIPCHK(nrgk_launch_diag_gather(dj, (int)jobs.size(), ddst, c->stream));
    }
    return 0;
  }
  // h: the op_avgs_len() doubles op_avgs_issue gathered (entries no job wrote are not read)
  void op_avgs_sum(double bbar, const double* h, std::vector<double>& out) const {
    out.clear();
    const int nb = AeigCut.NumBlocks();
    const int64_t ns = AeigCut.dEn.empty() ? 0 : (int64_t)AeigCut.dEn.size();
    for (size_t a = 0; a < oa_avg.size(); a++) {
      double s = 0, Z = 0;
      for (int b = 0; b < nb; b++) {
        const double dg = 2.0 * AeigCut.GetQNumber(b, 1) + 1.0;
        const int n = AeigCut.GetBlockSize(b), i0 = AeigCut.GetBlockLimit(b, 0);
        for (int i = 0; i < n; i++) {
          const double x = std::exp(-bbar * AeigCut.dEn[i0 + i]);
          const double el = oa_has[a * nb + b] ? h[(size_t)a * ns + i0 + i] : 0.0;
          s += dg * el * x; Z += dg * x;
        }
      }
      out.push_back(std::fabs(Z) > 1e-20 ? s / Z : s);
    }
  }
  int op_avgs(double bbar, std::vector<double>& out) {
    const int64_t len = op_avgs_len();
    DBuf& db = c->buf("opavg_diag");
    int rc;
    if ((rc = db.reserve(sizeof(double) * (size_t)(len + 1)))) return rc;
    if ((rc = op_avgs_issue((double*)db.p))) return rc;
    std::vector<double> h((size_t)len + 1, 0.0);
    if (len > 0 && !oa_avg.empty()) {
      HIPCHK(hipMemcpyAsync(h.data(), db.p, sizeof(double) * (size_t)len, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    }
    op_avgs_sum(bbar, h.data(), out);
    return 0;
  }

  void emit(nrg_step_cb cb, void* user, double emin, double ecut, int nkept, const double* th_T, const double* sus,
            const double* ent, double Tavg, const std::vector<double>& ops) {
    if (!cb) return;
    nrg_step st;
    std::memset(&st, 0, sizeof(st));
    st.nshell = Aeig.Nshell; st.nq = Aeig.NQNumbers; st.nblocks = Aeig.NumBlocks();
    st.qn = Aeig.QNumbers.data(); st.bbe = Aeig.BlockBegEnd.data();
    st.nstates = (int)Aeig.dEn.size(); st.en = Aeig.dEn.data();
    std::vector<unsigned char> kp(Aeig.dEn.size(), 1);
    for (size_t i = 0; i < kp.size() && i < Aeig.Kept.size(); i++) kp[i] = Aeig.Kept[i];
    st.kept = kp.data();
    st.emin = emin; st.ecut = ecut; st.nkept = nkept;
    st.hpacked = (record && !hrec.empty()) ? hrec.data() : nullptr;
    st.evec = (record && !zrec.empty()) ? zrec.data() : nullptr;
    for (int i = 0; i < 4; i++) {
      st.thermo_T[i] = th_T ? th_T[i] : 0; st.suscep[i] = sus ? sus[i] : 0; st.entropy[i] = ent ? ent[i] : 0;
    }
    st.T_avg = Tavg;
    st.nops_avg = (int)ops.size();
    for (size_t i = 0; i < ops.size() && i < 8; i++) st.opavg[i] = ops[i];
    st.cut_nblocks = AeigCut.NumBlocks();
    st.cut_qn = AeigCut.QNumbers.data();
    st.cut_bbe = AeigCut.BlockBegEnd.data();
    st.cut_bbe_bc = AeigCut.BlockBegEndBC.data();
    cb(user, &st);
  }
};

// per-S coupling table of a model rule, [S index = 2S][typep][type]; rows
// for spins already evaluated are kept and only new spin values are computed
struct CoefTable {
  std::function<double(double, int, int)> f;
  int ntypes = 0;
  int done = -1;
  std::vector<double> t;
  const std::vector<double>& upto(int smax) {
    if (smax > done) {
      t.resize((size_t)(smax + 1) * ntypes * ntypes, 0.0);
      for (int si = done + 1; si <= smax; si++)
        for (int a = 0; a < ntypes; a++)
          for (int b = 0; b < ntypes; b++) t[((size_t)si * ntypes + a) * ntypes + b] = f(0.5 * si, a, b);
      done = smax;
    }
    return t;
  }
};
static int max_sidx(const CNRGarray& A) {
  int m = 0;
  for (int b = 0; b < A.NumBlocks(); b++) m = std::max(m, (int)key2(A.GetQNumber(b, 1)));
  return m;
}

// What SaveArrays (src/NRGcodehandlerClass.cpp:887-933) writes after step N:
// Abasis_N, Acut_N (here with its kept eigenvector rows copied out of the
// step's double-buffered Z into a buffer of their own) and the operator opidx.
static int dm_capture(Chain& ch, int N, int opidx, DmShell& s) {
  nrg_ctx* c = ch.c;
  const CNRGbasisarray& A = ch.AeigCut;
  int rc;
  s.N = N;
  s.qn = A.QNumbers; s.en = A.dEn; s.bbe = A.BlockBegEnd; s.bbe_bc = A.BlockBegEndBC;
  s.kept = A.Kept;
  s.stcf = ch.Abasis.StCameFrom; s.itype = ch.Abasis.iType;
  s.bas_bbe = ch.Abasis.BlockBegEnd; s.bas_qn = ch.Abasis.QNumbers;
  const int nb = A.NumBlocks();
  s.zoff.assign(nb, 0);
  int64_t tot = 0;
  for (int b = 0; b < nb; b++) { s.zoff[b] = tot; tot += (int64_t)A.GetBlockSize(b) * A.zld[b]; }
  s.zbuf = "dmZ" + std::to_string(N);
  DBuf& zb = c->buf(s.zbuf);
  if ((rc = zb.reserve(sizeof(double) * (size_t)(tot + 1), true))) return rc;
  for (int b = 0; b < nb; b++)   // kept rows = the first k rows of the block's n x n slice
    HIPCHK(hipMemcpyAsync((double*)zb.p + s.zoff[b], ch.dZ + A.zoff[b],
                          sizeof(double) * (size_t)A.GetBlockSize(b) * A.zld[b], hipMemcpyDeviceToDevice, c->stream));
  const CNRGmatrix& M = ch.mats[opidx];
  s.op_map = M.MatBlockMap; s.op_mbe = M.MatBlockBegEnd;
  const int64_t nel = M.MatBlockBegEnd.empty() ? 0 : (int64_t)M.MatBlockBegEnd.back() + 1;
  s.opbuf = "dmO" + std::to_string(N);
  DBuf& ob = c->buf(s.opbuf);
  if ((rc = ob.reserve(sizeof(double) * (size_t)(nel + 1), true))) return rc;
  if (nel)
    HIPCHK(hipMemcpyAsync(ob.p, c->buf(M.buf[M.cur]).p, sizeof(double) * (size_t)nel, hipMemcpyDeviceToDevice,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// What DM_NRG would read of a TwoChQSz shell with UpdateBefCut 1 (the reference
// has no such path): every state of Aeig with its Kept flags, all
// eigenvector rows, Abasis with StCameFrom as indices of the previous shell's
// states before its cut, and d_up on every pair of Aeig blocks that its rule
// connects, Z_i F Z_j^T over all rows with F from the previous step's d_up
// (which lives on the previous AeigCut: call before Chain::update_ops).
static int dm_capture_all(Chain& ch, int N, int opidx, DmShell& s) {
  nrg_ctx* c = ch.c;
  const CNRGarray& A = ch.Aeig;
  const CNRGbasisarray& Ab = ch.Abasis;
  int rc;
  s.N = N;
  s.qn = A.QNumbers; s.en = A.dEn; s.bbe = A.BlockBegEnd; s.bbe_bc = A.BlockBegEnd;
  s.kept = A.Kept;
  s.itype = Ab.iType;
  s.stcf.resize(Ab.StCameFrom.size());
  for (size_t i = 0; i < Ab.StCameFrom.size(); i++) s.stcf[i] = ch.AeigCut.StCameFrom[Ab.StCameFrom[i]];
  s.bas_bbe = Ab.BlockBegEnd; s.bas_qn = Ab.QNumbers;
  const int nb = A.NumBlocks();
  s.zoff.assign(nb, 0);
  int64_t ztot = 0;
  for (int b = 0; b < nb; b++) { s.zoff[b] = ztot; ztot += (int64_t)A.GetBlockSize(b) * A.GetBlockSize(b); }
  s.zbuf = "dmZ" + std::to_string(N);
  DBuf& zb = c->buf(s.zbuf);
  if ((rc = zb.reserve(sizeof(double) * (size_t)(ztot + 1), true))) return rc;
  for (int b = 0; b < nb; b++)
    HIPCHK(hipMemcpyAsync((double*)zb.p + s.zoff[b], ch.dZ + ch.Aeig_zoff[b],
                          sizeof(double) * (size_t)A.GetBlockSize(b) * A.GetBlockSize(b), hipMemcpyDeviceToDevice,
                          c->stream));
  // pairs and product-basis F of the operator over every Aeig block pair
  OpSpec& sp = ch.specs[opidx];
  std::vector<FPair> pairs;
  std::vector<double> G;
  std::vector<RotJob> jobs;
  s.op_map.clear(); s.op_mbe.clear();
  int64_t ftot = 0, aux = 0, cnt = 0, maxel = 1;
  for (int ib = 0; ib < nb; ib++)
    for (int jb = 0; jb < nb; jb++) {
      if (!pair_ok(sp.rule, A, ib, jb)) continue;
      const int ni = A.GetBlockSize(ib), nj = A.GetBlockSize(jb);
      s.op_map.push_back(ib); s.op_map.push_back(jb);
      s.op_mbe.push_back((int)cnt);
      FPair P{};
      P.ni_bc = ni; P.nj_bc = nj; P.ibeg_bc = A.GetBlockLimit(ib, 0); P.jbeg_bc = A.GetBlockLimit(jb, 0);
      P.f_off = ftot; P.g_off = (int64_t)G.size(); P.need_old = 1; P.op = opidx;
      const std::vector<double>& gt = sp.table(A.GetQNumber(ib, 1), A.GetQNumber(jb, 1), ch.ntypes);
      G.insert(G.end(), gt.begin(), gt.end());
      pairs.push_back(P);
      RotJob J{};
      J.ni = ni; J.nj = nj; J.ki = ni; J.kj = nj; J.ldzi = ni; J.ldzj = nj;
      J.zi_off = ch.Aeig_zoff[ib]; J.zj_off = ch.Aeig_zoff[jb];
      J.f_off = ftot; J.aux_off = aux; J.out_off = cnt;
      jobs.push_back(J);
      ftot += (int64_t)ni * nj;
      aux += (int64_t)nj * ni;
      cnt += (int64_t)ni * nj;
      if (cnt > INT32_MAX) return fail(NRG_ERR_ARG, "DM-NRG: d_up of one shell exceeds 2^31 elements");
      s.op_mbe.push_back((int)cnt - 1);
      maxel = std::max(maxel, (int64_t)ni * nj);
    }
  s.opbuf = "dmO" + std::to_string(N);
  DBuf& ob = c->buf(s.opbuf);
  if ((rc = ob.reserve(sizeof(double) * (size_t)(cnt + 1), true))) return rc;
  if (!pairs.empty()) {
    DBuf& fb = c->buf("dmc_F");
    DBuf& ab = c->buf("dmc_aux");
    if ((rc = fb.reserve(sizeof(double) * (size_t)(ftot + 1)))) return rc;
    if ((rc = ab.reserve(sizeof(double) * (size_t)(aux + 1)))) return rc;
    FPair* d_pairs;
    double* d_G;
    int *d_stcf, *d_itype;
    if ((rc = upload(c, "dmc_pairs", pairs, &d_pairs))) return rc;
    if ((rc = upload(c, "dmc_G", G, &d_G))) return rc;
    if ((rc = upload(c, "dmc_stcf", Ab.StCameFrom, &d_stcf))) return rc;
    if ((rc = upload(c, "dmc_itype", Ab.iType, &d_itype))) return rc;
    FArgs fa{};
    fa.F = (double*)fb.p; fa.pairs = d_pairs; fa.G = d_G; fa.stcf = d_stcf; fa.itype = d_itype; fa.ntypes = ch.ntypes;
    if ((rc = make_oldop(c, ch.mats[opidx], ch.oldb, "dmc_old", fa.old[opidx]))) return rc;
    HIPCHK(nrgk_launch_fbasis(&fa, (int)pairs.size(), maxel, c->stream));
    if ((rc = rotate_dev(c, jobs, ch.dZ, (double*)fb.p, (double*)ab.p, (double*)ob.p))) return rc;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return c->resolve();
}

// ---------------------------------------------------------------------------
// Derivative of the fixed-truncation OneChQS loop with respect to (U, Gamma,
// ed), forward mode (no counterpart in the reference).  Each step, for every
// tangent t:
//   dH  = sqrt(Lambda) th_{N-1} (x) 1 + chi df_{N-1} + dchi f_{N-1}   (tan_hn_kernel)
//   G   = Z dH Z^T on the kept rows (two MFMA GEMMs), dE = diag G      (tan_diag_kernel)
//   th_N = G_kk - dEmin I,  C = G_kd / (E_k - E_d)                    (tan_cut_kernel)
//   dZ_k = C Z_d                                                       (MFMA GEMM)
//   dO  = dZ_i F Z_j^T + Z_i F dZ_j^T + Z_i dF Z_j^T per operator block pair
// with the kept-space Hamiltonian tangent th kept as a full k x k block per
// kept block (rotations inside the kept space are not undone).  T chi, S and
// <O> derivatives are host sums over dE and th (Daleckii-Krein divided
// differences of exp(-beta E) for the off-diagonal part).
// ---------------------------------------------------------------------------
static E2Gemm gemm_desc(int m, int n, int k, int64_t a_off, int64_t as0, int64_t as1, int64_t b_off, int64_t bs0,
                        int64_t bs1, int64_t c_off, int64_t cs0, int acc) {
  E2Gemm g{};
  g.m = m; g.n = n; g.k = k; g.dyn = -1;
  g.a_off = a_off; g.as0 = as0; g.as1 = as1;
  g.b_off = b_off; g.bs0 = bs0; g.bs1 = bs1;
  g.c_off = c_off; g.cs0 = cs0;
  g.alpha = 1.0; g.accumulate = acc;
  return g;
}
// one launch of the grouped FP64 MFMA GEMM over problems sharing A, B, C bases
static int gemm_batch(nrg_ctx* c, const std::vector<E2Gemm>& P, const double* A, const double* B, double* C,
                      const std::string& tag) {
  std::vector<E2Tile> T;
  double fl = 0.0, by = 0.0;
  for (size_t q = 0; q < P.size(); q++) {
    const int tm = (P[q].m + 63) / 64, tn = (P[q].n + 63) / 64;
    for (int a = 0; a < tm; a++)
      for (int b = 0; b < tn; b++) T.push_back(E2Tile{(int)q, a, b, 0});
    fl += 2.0 * P[q].m * (double)P[q].n * P[q].k;
    by += 8.0 * ((double)P[q].m * P[q].k + (double)P[q].k * P[q].n + (double)P[q].m * P[q].n);
  }
  if (T.empty()) return 0;
  E2Gemm* dp;
  E2Tile* dt;
  int rc;
  if ((rc = upload(c, (tag + "_p").c_str(), P, &dp))) return rc;
  if ((rc = upload(c, (tag + "_t").c_str(), T, &dt))) return rc;
  PendingTimer tm;
  if ((rc = c->timer_begin(tm, NRG_K_GEMM, fl, by))) return rc;
  HIPCHK(e2_launch_gemm(dp, dt, (int)T.size(), A, B, C, c->stream));
  return c->timer_end(tm);
}

struct Tangent {
  nrg_ctx* c = nullptr;
  int T = 0;
  double dpar[NRGK_MAX_TAN][3] = {};
  double dchi[NRGK_MAX_TAN] = {};
  int cur = 0;                          // th / operator tangents: buffer "tg_th<cur>", "tg_op<cur>"
  std::vector<int64_t> th_off;          // per current kept block
  int64_t th_tot = 0;
  std::vector<int64_t> op_cnt, op_base; // per operator: elements, start (T copies follow)
  // per step
  std::vector<int> kb;                  // kept rows of each Aeig block
  std::vector<int64_t> goff;
  int64_t gtot = 0;
  int64_t gs = 0;                       // SetE0zero's ground state
  std::vector<double> den;              // T x nstates, after the shift
  std::vector<int64_t> dzoff;           // per kept block: its k x n rows of dZ
  int64_t dztot = 0, ctot = 0;
  std::vector<double> dsus, dent, dops;

  double* buf(const std::string& n) { return (double*)c->buf(n).p; }
  std::string thb(int w) const { return w ? "tg_th1" : "tg_th0"; }
  std::string opb(int w) const { return w ? "tg_op1" : "tg_op0"; }

  // H_{-1}: one-state blocks (src/Main/Main_SetH0.cpp:1847-1849); f and the
  // impurity operators are constants
  int init(nrg_ctx* cc, int ntan, const double* dp, const Chain& ch, double Lambda, double HLF, double chi_m1,
           double Gamma) {
    c = cc; T = ntan;
    for (int t = 0; t < T; t++)
      for (int q = 0; q < 3; q++) dpar[t][q] = dp[3 * t + q];
    const double sc = 1.0 / (Lambda * HLF);
    th_off = {0, 1, 2};
    th_tot = 3;
    std::vector<double> th((size_t)T * 3 + 1);
    den.assign((size_t)T * 3, 0.0);
    for (int t = 0; t < T; t++) {
      const double dU = dpar[t][0], ded = dpar[t][2];
      const double e3[3] = {0.5 * dU * sc, (ded + 0.5 * dU) * sc, (2.0 * ded + 1.5 * dU) * sc};
      for (int b = 0; b < 3; b++) th[(size_t)t * 3 + b] = den[(size_t)t * 3 + b] = e3[b];
      dchi[t] = chi_m1 * dpar[t][1] / (2.0 * Gamma);
    }
    int rc;
    DBuf& tb = c->buf(thb(cur));
    if ((rc = tb.reserve(sizeof(double) * th.size()))) return rc;
    HIPCHK(hipMemcpyAsync(tb.p, th.data(), sizeof(double) * th.size(), hipMemcpyHostToDevice, c->stream));
    const int nm = (int)ch.mats.size();
    op_cnt.assign(nm, 0); op_base.assign(nm, 0);
    int64_t tot = 0;
    for (int m = 0; m < nm; m++) {
      op_cnt[m] = ch.mats[m].MatBlockBegEnd.empty() ? 0 : ch.mats[m].MatBlockBegEnd.back() + 1;
      op_base[m] = tot;
      tot += (int64_t)T * op_cnt[m];
    }
    DBuf& ob = c->buf(opb(cur));
    if ((rc = ob.reserve(sizeof(double) * (size_t)(tot + 1)))) return rc;
    HIPCHK(hipMemsetAsync(ob.p, 0, sizeof(double) * (size_t)(tot + 1), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  }

  // dH of every Abasis block, right after the base H_N assembly
  int assemble(Chain& ch, const std::vector<double>& coef, double chi, double sqrtL) {
    int rc;
    const CNRGbasisarray& Ab = ch.Abasis;
    const int nb = Ab.NumBlocks();
    std::vector<int> bbeg(nb), sidx(nb);
    for (int b = 0; b < nb; b++) {
      bbeg[b] = Ab.GetBlockLimit(b, 0);
      sidx[b] = (int)key2(Ab.GetQNumber(b, 1));
    }
    TanHnArgs a{};
    int *d_bbeg, *d_bsize, *d_sidx, *d_stcf, *d_itype;
    int64_t *d_hoff, *d_thoff;
    double* d_coef;
    if ((rc = upload(c, "tg_bbeg", bbeg, &d_bbeg))) return rc;
    if ((rc = upload(c, "tg_bsize", ch.L_bsize, &d_bsize))) return rc;
    if ((rc = upload(c, "tg_sidx", sidx, &d_sidx))) return rc;
    if ((rc = upload(c, "tg_hoff", ch.L_hoff, &d_hoff))) return rc;
    if ((rc = upload(c, "tg_stcf", Ab.StCameFrom, &d_stcf))) return rc;
    if ((rc = upload(c, "tg_itype", Ab.iType, &d_itype))) return rc;
    if ((rc = upload(c, "tg_thoff", th_off, &d_thoff))) return rc;
    if ((rc = upload(c, "tg_coef", coef, &d_coef))) return rc;
    if ((rc = make_oldop(c, ch.mats[0], ch.oldb, "tg_f", a.f))) return rc;
    DBuf& hb = c->buf("tg_dH");
    if ((rc = hb.reserve(sizeof(double) * (size_t)(T * ch.L_ztot + 1)))) return rc;
    a.dH = (double*)hb.p; a.dh_stride = ch.L_ztot;
    a.h_off = d_hoff; a.bbeg = d_bbeg; a.bsize = d_bsize; a.blk_sidx = d_sidx; a.stcf = d_stcf; a.itype = d_itype;
    a.th = buf(thb(cur)); a.th_off = d_thoff; a.th_stride = th_tot;
    a.coef = d_coef;
    a.df = buf(opb(cur)) + op_base[0]; a.df_stride = op_cnt[0];
    a.chi = chi; a.sqrtL = sqrtL;
    for (int t = 0; t < T; t++) a.dchi[t] = dchi[t];
    a.ntypes = ch.ntypes;
    HIPCHK(nrgk_launch_tan_hn(&a, nb, ch.L_nmax, T, c->stream));
    return 0;
  }

  // after the Kept decision: G on the kept rows, dE of every state (host copy, shifted)
  int diag(Chain& ch) {
    int rc;
    const CNRGarray& A = ch.Aeig;
    const int nb = A.NumBlocks(), ns = A.Nstates();
    const int64_t zt = ch.L_ztot;
    kb.assign(nb, 0); goff.assign(nb, 0);
    gtot = 0;
    gs = std::min_element(A.dEn.begin(), A.dEn.end()) - A.dEn.begin();
    std::vector<TanBlk> tb(nb);
    std::vector<E2Gemm> py, pg;
    for (int b = 0; b < nb; b++) {
      const int n = ch.L_bsize[b], b0 = A.GetBlockLimit(b, 0);
      int k = 0;
      for (int i = 0; i < n; i++) k += A.Kept[b0 + i] != 0;
      kb[b] = k; goff[b] = gtot;
      tb[b] = TanBlk{n, k, ch.L_hoff[b], gtot, b0};
      for (int t = 0; t < T; t++) {
        const int64_t h = (int64_t)t * zt + ch.L_hoff[b];
        py.push_back(gemm_desc(n, n, n, ch.L_hoff[b], n, 1, h, n, 1, h, n, 0));      // Y = Z dH
      }
      gtot += (int64_t)k * n;
    }
    for (int b = 0; b < nb; b++) {
      const int n = ch.L_bsize[b], k = kb[b];
      for (int t = 0; t < T; t++)   // G_k = Y_k Z^T
        pg.push_back(gemm_desc(k, n, n, (int64_t)t * zt + ch.L_hoff[b], n, 1, ch.L_hoff[b], 1, n,
                               (int64_t)t * gtot + goff[b], n, 0));
    }
    DBuf& yb = c->buf("tg_Y");
    DBuf& gb = c->buf("tg_G");
    DBuf& eb = c->buf("tg_dE");
    if ((rc = yb.reserve(sizeof(double) * (size_t)(T * zt + 1)))) return rc;
    if ((rc = gb.reserve(sizeof(double) * (size_t)(T * gtot + 1)))) return rc;
    if ((rc = eb.reserve(sizeof(double) * (size_t)(T * (int64_t)ns + 1)))) return rc;
    if ((rc = gemm_batch(c, py, ch.dZ, buf("tg_dH"), (double*)yb.p, "tg_gy"))) return rc;
    if ((rc = gemm_batch(c, pg, (double*)yb.p, ch.dZ, (double*)gb.p, "tg_gg"))) return rc;
    TanBlk* d_tb;
    if ((rc = upload(c, "tg_blk", tb, &d_tb))) return rc;
    HIPCHK(nrgk_launch_tan_diag(d_tb, nb, ch.L_nmax, T, ch.dZ, (double*)yb.p, zt, (double*)gb.p, gtot,
                                (double*)eb.p, ns, c->stream));
    den.resize((size_t)T * ns);
    HIPCHK(hipMemcpyAsync(den.data(), eb.p, sizeof(double) * den.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int t = 0; t < T; t++) {
      const double s = den[(size_t)t * ns + gs];
      for (int i = 0; i < ns; i++) den[(size_t)t * ns + i] -= s;
    }
    return c->resolve();
  }

  // CalcSuscep / CalcEntropy (src/CalcThermo.cpp:16-121) differentiated along dE
  void thermo(const CNRGarray& A, const double* bN4) {
    const int ns = A.Nstates();
    dsus.assign((size_t)T * 4, 0.0); dent.assign((size_t)T * 4, 0.0);
    for (int t = 0; t < T; t++)
      for (int nc = 0; nc < 4; nc++) {
        const double bb = bN4[nc];
        double s0 = 0, s1 = 0, d0 = 0, d1 = 0, se = 0, dse = 0;
        for (int b = 0; b < A.NumBlocks(); b++) {
          const double S = A.GetQNumber(b, 1), aS = 2.0 * S + 1.0, szsq = (1.0 / 12.0) * aS * (aS * aS - 1.0);
          for (int i = A.GetBlockLimit(b, 0); i <= A.GetBlockLimit(b, 1); i++) {
            const double E = A.dEn[i], dE = den[(size_t)t * ns + i], x = std::exp(-bb * E), dx = -bb * dE * x;
            s0 += aS * x; s1 += szsq * x; d0 += aS * dx; d1 += szsq * dx;
            se += aS * E * x; dse += aS * (dE * x + E * dx);
          }
        }
        dsus[(size_t)t * 4 + nc] = (d1 * s0 - s1 * d0) / (s0 * s0);
        dent[(size_t)t * 4 + nc] = bb * (dse * s0 - se * d0) / (s0 * s0) + d0 / s0;
      }
  }

  // after CutStates: th of the new kept blocks, C, dZ_k = C Z_d
  int cut(Chain& ch) {
    int rc;
    const CNRGarray& A = ch.Aeig;
    const CNRGbasisarray& Ac = ch.AeigCut;
    const int nbc = Ac.NumBlocks(), ns = A.Nstates();
    std::map<int, int> from_beg;
    for (int b = 0; b < A.NumBlocks(); b++) from_beg[A.GetBlockLimit(b, 0)] = b;
    std::vector<TanCut> tc(nbc);
    std::vector<int64_t> nth(nbc), coff(nbc);
    dzoff.assign(nbc, 0);
    int64_t tht = 0, maxel = 1;
    ctot = 0; dztot = 0;
    for (int cb = 0; cb < nbc; cb++) {
      const int be = from_beg.at(Ac.BlockBegEndBC[2 * cb]);
      const int n = ch.L_bsize[be], k = Ac.GetBlockSize(cb);
      if (k != kb[be]) return fail(NRG_ERR_STATE, "tangent: kept rows differ from the cut");
      nth[cb] = tht; coff[cb] = ctot; dzoff[cb] = dztot;
      tc[cb] = TanCut{n, k, goff[be], A.GetBlockLimit(be, 0), tht, ctot};
      tht += (int64_t)k * k; ctot += (int64_t)k * (n - k); dztot += (int64_t)k * n;
      maxel = std::max(maxel, (int64_t)k * n);
    }
    double* d_E;
    TanCut* d_tc;
    if ((rc = upload(c, "tg_E", A.dEn, &d_E))) return rc;
    if ((rc = upload(c, "tg_cut", tc, &d_tc))) return rc;
    DBuf& thn = c->buf(thb(cur ^ 1));
    DBuf& cbf = c->buf("tg_C");
    DBuf& dzb = c->buf("tg_dZ");
    if ((rc = thn.reserve(sizeof(double) * (size_t)(T * tht + 1)))) return rc;
    if ((rc = cbf.reserve(sizeof(double) * (size_t)(T * ctot + 1)))) return rc;
    if ((rc = dzb.reserve(sizeof(double) * (size_t)(T * dztot + 1)))) return rc;
    HIPCHK(nrgk_launch_tan_cut(d_tc, nbc, maxel, T, buf("tg_G"), gtot, d_E, buf("tg_dE"), ns, gs, (double*)thn.p, tht,
                               (double*)cbf.p, ctot, c->stream));
    std::vector<E2Gemm> pz;
    for (int cb = 0; cb < nbc; cb++) {
      const int n = tc[cb].n, k = tc[cb].k;
      const int64_t z = Ac.zoff[cb];
      for (int t = 0; t < T; t++)   // dZ_k = C Z[k:n, :]  (K = 0 writes the zeros of a fully kept block)
        pz.push_back(gemm_desc(k, n, n - k, (int64_t)t * ctot + coff[cb], n - k, 1, z + (int64_t)k * n, n, 1,
                               (int64_t)t * dztot + dzoff[cb], n, 0));
    }
    if ((rc = gemm_batch(c, pz, (double*)cbf.p, ch.dZ, (double*)dzb.p, "tg_gz"))) return rc;
    th_off = nth;   // th of the new kept basis is in buffer cur^1 from here on
    th_tot = tht;
    return 0;
  }

  // operator tangents, between Chain::plan_update and Chain::finish_update
  int ops(Chain& ch) {
    int rc;
    const int nm = (int)ch.mats.size();
    const CNRGbasisarray& Ac = ch.AeigCut;
    std::map<int64_t, int> cut_of_z;
    for (int cb = 0; cb < Ac.NumBlocks(); cb++) cut_of_z[Ac.zoff[cb]] = cb;
    std::vector<int64_t> ncnt(nm), nbase(nm);
    int64_t tot = 0;
    for (int m = 0; m < nm; m++) { ncnt[m] = ch.up_icount[m]; nbase[m] = tot; tot += (int64_t)T * ncnt[m]; }
    DBuf& onew = c->buf(opb(cur ^ 1));
    if ((rc = onew.reserve(sizeof(double) * (size_t)(tot + 1)))) return rc;
    const std::vector<RotJob>& J = ch.up_jobs;
    std::vector<int64_t> poff(J.size()), qoff(J.size());
    int64_t ptot = 0, qtot = 0;
    std::vector<E2Gemm> pp, pq, po1, po2;
    for (size_t q = 0; q < J.size(); q++) {
      const RotJob& j = J[q];
      poff[q] = ptot; qoff[q] = qtot;
      ptot += (int64_t)j.ni * j.kj; qtot += (int64_t)j.ki * j.nj;
      pp.push_back(gemm_desc(j.ni, j.kj, j.nj, j.f_off, j.nj, 1, j.zj_off, 1, j.ldzj, poff[q], j.kj, 0));   // F Z_j^T
      pq.push_back(gemm_desc(j.ki, j.nj, j.ni, j.zi_off, j.ldzi, 1, j.f_off, j.nj, 1, qoff[q], j.nj, 0));   // Z_i F
      const int m = ch.up_job_op[q];
      const int ci = cut_of_z.at(j.zi_off), cj = cut_of_z.at(j.zj_off);
      for (int t = 0; t < T; t++) {
        const int64_t o = nbase[m] + (int64_t)t * ncnt[m] + j.out_off;
        po1.push_back(gemm_desc(j.ki, j.kj, j.ni, (int64_t)t * dztot + dzoff[ci], j.ni, 1, poff[q], j.kj, 1, o, j.kj,
                                0));   // dZ_i (F Z_j^T)
        po2.push_back(gemm_desc(j.ki, j.kj, j.nj, qoff[q], j.nj, 1, (int64_t)t * dztot + dzoff[cj], 1, j.nj, o, j.kj,
                                1));   // + (Z_i F) dZ_j^T
      }
    }
    DBuf& pb = c->buf("tg_P");
    DBuf& qb = c->buf("tg_Q");
    if ((rc = pb.reserve(sizeof(double) * (size_t)(ptot + 1)))) return rc;
    if ((rc = qb.reserve(sizeof(double) * (size_t)(qtot + 1)))) return rc;
    double* F = buf("upd_F");
    if ((rc = gemm_batch(c, pp, F, ch.dZ, (double*)pb.p, "tg_gp"))) return rc;
    if ((rc = gemm_batch(c, pq, ch.dZ, F, (double*)qb.p, "tg_gq"))) return rc;
    if ((rc = gemm_batch(c, po1, buf("tg_dZ"), (double*)pb.p, (double*)onew.p, "tg_go1"))) return rc;
    if ((rc = gemm_batch(c, po2, (double*)qb.p, buf("tg_dZ"), (double*)onew.p, "tg_go2"))) return rc;
    // + Z_i dF Z_j^T for the operators built on their previous values
    std::vector<FPair> sub;
    std::vector<size_t> subq;
    int64_t maxel = 1;
    for (size_t q = 0; q < J.size(); q++)
      if (ch.specs[ch.up_job_op[q]].need_old) {
        sub.push_back(ch.up_pairs[q]);
        subq.push_back(q);
        maxel = std::max(maxel, (int64_t)J[q].ni * J[q].nj);
      }
    if (!sub.empty()) {
      FPair* d_sub;
      if ((rc = upload(c, "tg_pairs", sub, &d_sub))) return rc;
      const int64_t ftot = ch.up_ftot;
      DBuf& dfb = c->buf("tg_dF");
      if ((rc = dfb.reserve(sizeof(double) * (size_t)(T * ftot + 1)))) return rc;
      int64_t rtot = 0;
      std::vector<int64_t> roff(subq.size());
      for (size_t u = 0; u < subq.size(); u++) { roff[u] = rtot; rtot += (int64_t)J[subq[u]].ni * J[subq[u]].kj; }
      DBuf& rb = c->buf("tg_R");
      if ((rc = rb.reserve(sizeof(double) * (size_t)(T * rtot + 1)))) return rc;
      std::vector<E2Gemm> pr, po3;
      for (int t = 0; t < T; t++) {
        FArgs fa = ch.up_fa;
        fa.F = (double*)dfb.p + (int64_t)t * ftot;
        fa.pairs = d_sub;
        for (int m = 0; m < nm && m < 8; m++)
          if (ch.specs[m].need_old) fa.old[m].mat = buf(opb(cur)) + op_base[m] + (int64_t)t * op_cnt[m];
        HIPCHK(nrgk_launch_fbasis(&fa, (int)sub.size(), maxel, c->stream));
        for (size_t u = 0; u < subq.size(); u++) {
          const RotJob& j = J[subq[u]];
          const int m = ch.up_job_op[subq[u]];
          const int64_t r = (int64_t)t * rtot + roff[u];
          pr.push_back(gemm_desc(j.ni, j.kj, j.nj, (int64_t)t * ftot + j.f_off, j.nj, 1, j.zj_off, 1, j.ldzj, r,
                                 j.kj, 0));   // dF Z_j^T
          po3.push_back(gemm_desc(j.ki, j.kj, j.ni, j.zi_off, j.ldzi, 1, r, j.kj, 1,
                                  nbase[m] + (int64_t)t * ncnt[m] + j.out_off, j.kj, 1));   // + Z_i (dF Z_j^T)
        }
      }
      if ((rc = gemm_batch(c, pr, (double*)dfb.p, ch.dZ, (double*)rb.p, "tg_gr"))) return rc;
      if ((rc = gemm_batch(c, po3, ch.dZ, (double*)rb.p, (double*)onew.p, "tg_go3"))) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    op_cnt = ncnt;
    op_base = nbase;
    cur ^= 1;   // th (from cut) and the operator tangents now live in buffer cur
    return c->resolve();
  }

  // CalcOpAvg (src/CalcOpAvg.cpp:13-84) differentiated: with the kept-space
  // Hamiltonian diag(E) + th, sum_i O_ii x_i becomes Tr[O exp(-beta H)]
  int op_avgs(Chain& ch, double bb) {
    const CNRGbasisarray& A = ch.AeigCut;
    const int nm = (int)ch.mats.size();
    dops.assign((size_t)T * 8, 0.0);
    std::vector<double> th((size_t)T * th_tot + 1);
    HIPCHK(hipMemcpyAsync(th.data(), buf(thb(cur)), sizeof(double) * (size_t)(T * th_tot), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int q = 0;
    for (int m = 0; m < nm; m++) {
      if (!ch.mats[m].CalcAvg) continue;
      const CNRGmatrix& M = ch.mats[m];
      const int64_t cnt = op_cnt[m];
      std::vector<double> h(cnt + 1), dh((size_t)T * cnt + 1);
      if (cnt > 0) {
        HIPCHK(hipMemcpyAsync(h.data(), c->buf(M.buf[M.cur]).p, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipMemcpyAsync(dh.data(), buf(opb(cur)) + op_base[m], sizeof(double) * (size_t)(T * cnt),
                              hipMemcpyDeviceToHost, c->stream));
      }
      HIPCHK(hipStreamSynchronize(c->stream));
      for (int t = 0; t < T; t++) {
        double s = 0, Z = 0, ds = 0, dZ = 0;
        for (int b = 0; b < A.NumBlocks(); b++) {
          const double dg = 2.0 * A.GetQNumber(b, 1) + 1.0;
          const int k = A.GetBlockSize(b), b0 = A.GetBlockLimit(b, 0);
          int64_t off = -1;
          for (size_t r = 0; r + 1 < M.MatBlockMap.size(); r += 2)
            if (M.MatBlockMap[r] == b && M.MatBlockMap[r + 1] == b) { off = M.MatBlockBegEnd[r]; break; }
          const double* tb = th.data() + (size_t)t * th_tot + th_off[b];
          for (int i = 0; i < k; i++) {
            const double x = std::exp(-bb * A.dEn[b0 + i]);
            Z += dg * x;
            dZ += dg * tb[(int64_t)i * k + i] * (-bb * x);
            if (off < 0) continue;
            const double* O = h.data() + off;
            s += dg * O[(int64_t)i * k + i] * x;
            ds += dg * dh[(size_t)t * cnt + off + (int64_t)i * k + i] * x;
            for (int j = 0; j < k; j++) {
              const double d = A.dEn[b0 + i] - A.dEn[b0 + j], xj = std::exp(-bb * A.dEn[b0 + j]);
              const double phi = d != 0.0 ? xj * std::expm1(-bb * d) / d : -bb * xj;
              ds += dg * O[(int64_t)j * k + i] * tb[(int64_t)i * k + j] * phi;
            }
          }
        }
        dops[(size_t)t * 8 + q] = (ds * Z - s * dZ) / (Z * Z);
      }
      q++;
    }
    return 0;
  }

  // What the DM-NRG derivative needs of shell N, next to dm_capture's data:
  // th (k x k per kept block), dZ_k (k x n per kept block) and the c_d tangent
  int capture(DmShell& s, int opidx) {
    s.ntan = T;
    s.th_off = th_off; s.th_tot = th_tot;
    s.dz_off = dzoff; s.dz_tot = dztot;
    s.op_cnt = op_cnt[opidx];
    const std::string n = std::to_string(s.N);
    s.thbuf = "dmTH" + n; s.dzbuf = "dmdZ" + n; s.dopbuf = "dmdO" + n;
    const std::pair<std::string, std::pair<const double*, int64_t>> cp[3] = {
        {s.thbuf, {buf(thb(cur)), T * th_tot}},
        {s.dzbuf, {buf("tg_dZ"), T * dztot}},
        {s.dopbuf, {buf(opb(cur)) + op_base[opidx], T * op_cnt[opidx]}}};
    for (const auto& x : cp) {
      DBuf& b = c->buf(x.first);
      int rc = b.reserve(sizeof(double) * (size_t)(x.second.second + 1), true);
      if (rc) return rc;
      if (x.second.second > 0)
        HIPCHK(hipMemcpyAsync(b.p, x.second.first, sizeof(double) * (size_t)x.second.second, hipMemcpyDeviceToDevice,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  }

  void fill(nrg_tangent_step& dt, int N, int ns, int nops) const {
    std::memset(&dt, 0, sizeof(dt));
    dt.nshell = N; dt.ntan = T; dt.nstates = ns; dt.nops_avg = nops;
    dt.den = den.data();
    for (int t = 0; t < T; t++) {
      for (int i = 0; i < 4; i++) {
        dt.dsuscep[t][i] = dsus.empty() ? 0.0 : dsus[(size_t)t * 4 + i];
        dt.dentropy[t][i] = dent.empty() ? 0.0 : dent[(size_t)t * 4 + i];
      }
      for (int i = 0; i < nops && i < 8; i++) dt.dopavg[t][i] = dops.empty() ? 0.0 : dops[(size_t)t * 8 + i];
    }
  }
};

// ---------------------------------------------------------------------------
// OneChQS Anderson: NRG_main loop
// ---------------------------------------------------------------------------
// One OneChQS instance of the NRG_main loop, in phases so that several
// instances can step in lockstep around one shared eigensolve.
struct OneChRun {
  Chain ch;
  nrg_onechqs_params p{};
  double bbar = 0.727, HLF = 1.0, Lambda = 2.0, par_chi = 0.0;
  int nm = 2;
  int befcut = 0;   // UpdateBefCut (src/Main/NRG_main.cpp:334-343,393-396)
  CoefTable hn_tab;
  std::vector<double> ops;

  int init(nrg_ctx* c, const nrg_onechqs_params& pp, const std::string& tag, nrg_step_cb cb, void* user) {
    p = pp;
    Lambda = p.Lambda;
    bbar = p.betabar > 0 ? p.betabar : 0.727;
    if (p.myeps <= 0) p.myeps = 1e-18;
    HLF = 0.5 * (1.0 + 1.0 / Lambda);
    int rc;
    ch.c = c;
    ch.tag = tag;
    ch.site = onechqs_site();
    ch.ntypes = 4;
    ch.record = p.record_blocks != 0;
    const SiteDef* st = &ch.site;
    // ModelSwitch (src/CodeHandler_ModelSwitch.cpp:121-293): [0] f_N, then
    // n_d, n_d^2, S_d^2 (calcdens 0/2, ImpOnly_MatEl x old) or c_d (calcdens 1)
    nm = (p.calcdens == 1) ? 2 : 4;
    ch.specs.resize(nm);
    ch.specs[0].rule = PAIR_QS_UP;
    ch.specs[0].g = [st](double Si, double Sj, int ti, int tj) { return onech_fn_el(*st, Si, Sj, ti, tj); };
    for (int m = 1; m < nm; m++) {
      ch.specs[m].need_old = true;
      if (p.calcdens == 1) {
        ch.specs[m].rule = PAIR_QS_UP;
        ch.specs[m].g = [st](double Si, double Sj, int ti, int tj) { return onech_cd_el(*st, Si, Sj, ti, tj); };
      } else {
        ch.specs[m].rule = PAIR_DIAG;
        ch.specs[m].calc_avg = true;
        ch.specs[m].g = [](double, double, int ti, int tj) { return ti == tj ? 1.0 : 0.0; };
      }
    }
    ch.init_ops(nm);
    // H_{-1}: OneChQS_SetAnderson_Hm1 (src/Main/Main_SetH0.cpp:1794-1969)
    double chi_m1 = std::sqrt(2.0 * p.Gamma / M_PI) / (std::sqrt(Lambda) * HLF);
    if (p.disc_scheme != 1) chi_m1 *= std::sqrt(0.5 * std::log(Lambda) * (Lambda + 1) / (Lambda - 1));
    ch.Aeig.Nshell = -1; ch.Aeig.NQNumbers = 2;
    ch.Aeig.QNumbers = {-1, 0, 0, 0.5, 1, 0};
    ch.Aeig.BlockBegEnd = {0, 0, 1, 1, 2, 2};
    ch.Aeig.dEn = {0.5 * p.U / (Lambda * HLF), (p.ed + 0.5 * p.U) / (Lambda * HLF),
                   (2.0 * p.ed + 1.5 * p.U) / (Lambda * HLF)};
    ch.Aeig_zoff = {0, 1, 2};
    {
      DBuf& zb = c->buf(ch.tag + "Z0");
      if ((rc = zb.reserve(sizeof(double) * 8))) return rc;
      double ones[3] = {1, 1, 1};
      HIPCHK(hipMemcpyAsync(zb.p, ones, sizeof(ones), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      ch.dZ = (double*)zb.p;
    }
    std::vector<double> f = {1.0, -std::sqrt(2.0)};
    if ((rc = ch.set_op(0, {0, 1, 1, 2}, {0, 0, 1, 1}, f))) return rc;
    for (int m = 1; m < nm; m++) {
      if (p.calcdens == 1) {
        if ((rc = ch.set_op(m, {0, 1, 1, 2}, {0, 0, 1, 1}, f))) return rc;
      } else {
        std::vector<double> v(3);
        for (int b = 0; b < 3; b++) {
          double nd = ch.Aeig.GetQNumber(b, 0) + 1.0, Sd = ch.Aeig.GetQNumber(b, 1);
          v[b] = (m == 1) ? nd : (m == 2 ? nd * nd : Sd * (Sd + 1.0));
        }
        if ((rc = ch.set_op(m, {0, 0, 1, 1, 2, 2}, {0, 0, 1, 1, 2, 2}, v))) return rc;
      }
    }
    ch.cut_states(p.ncut, false, p.myeps);
    if ((rc = make_basis_dev(c, ch.AeigCut, (ch.tag + "oldb").c_str(), ch.oldb))) return rc;
    if ((rc = ch.op_avgs(bbar, ops))) return rc;
    double DN = HLF * std::pow(Lambda, -(0 - 2) / 2.0);
    ch.emit(cb, user, 0.0, HUGE_VAL, ch.Aeig.Nstates(), nullptr, nullptr, nullptr, DN / bbar, ops);
    par_chi = chi_m1;
    hn_tab.ntypes = 4;
    hn_tab.f = [st](double S, int a, int b) { return onech_hn_coef(*st, S, a, b); };
    return 0;
  }

  void basis() { build_basis(ch.AeigCut, ch.Abasis, ch.site, 0, befcut); }

  int assemble(double* dZ) {
    const std::vector<double>& coef = hn_tab.upto(max_sidx(ch.Abasis));
    int fop = 0;
    return ch.assemble(1, &fop, &par_chi, &coef, std::sqrt(Lambda), 0.0, dZ);
  }

  // after the eigensolve: SetE0zero, SetKept (SetDegen + Ecut), CalcThermo at
  // 4 betabar values, CutStates, UpdateMatrices, CalcStuff
  //   tg: carry the derivative pass along; fixed: this step's Kept set (Aeig order)
  int post(int N, nrg_step_cb cb, void* user, std::vector<DmShell>* dm, Tangent* tg = nullptr,
           const unsigned char* fixed = nullptr) {
    int rc;
    if ((rc = post_cut(N, tg, fixed))) return rc;
    // the deferred eigenvectors: every row with UpdateBefCut 1, else the kept
    // ones; the host work that does not decide them runs while they are formed
    if ((rc = ch.finish_vectors(befcut != 0))) return rc;
    if ((rc = post_thermo(N, tg))) return rc;
    return post_rest(N, cb, user, dm, tg);
  }

  // the step's results between post_cut and post_rest
  double s_DN = 0, s_emin = 0, s_ecut = 0, s_thT[4] = {}, s_sus[4] = {}, s_ent[4] = {};
  int s_nkept = 0;

  // SetE0zero, SetKept, CalcThermo, CutStates: everything the truncation decides
  int post_cut(int N, Tangent* tg, const unsigned char* fixed) {
    int rc;
    const double DN = HLF * std::pow(Lambda, -(N - 1) / 2.0);
    double emin = *std::min_element(ch.Aeig.dEn.begin(), ch.Aeig.dEn.end());
    for (auto& x : ch.Aeig.dEn) x -= emin;
    set_degen(ch.Aeig.dEn, p.myeps);
    const int ns = (int)ch.Aeig.dEn.size();
    double ecut;
    int nkept = 0;
    if (fixed) {   // the given Kept set; Ecut reported as its highest level
      ch.Aeig.Kept.assign(fixed, fixed + ns);
      ecut = 0.0;
      for (int i = 0; i < ns; i++)
        if (ch.Aeig.Kept[i]) { nkept++; ecut = std::max(ecut, ch.Aeig.dEn[i]); }
    } else {
      ecut = ecut_of(ch.Aeig.dEn, p.ncut);
      ch.Aeig.Kept.assign(ns, 0);
      for (int i = 0; i < ns; i++) { ch.Aeig.Kept[i] = ch.Aeig.dEn[i] <= ecut; nkept += ch.Aeig.Kept[i]; }
    }
    if (tg && (rc = tg->diag(ch))) return rc;
    // UpdateBefCut 1: the operators are rotated over every eigenstate of the copy
    // AeigCut = Aeig with its Kept flags (CutStates with 2 Nstates), SaveArrays
    // writes that copy, and the real cut follows CalcStuff; the operators stay on
    // the pre-cut blocks, which the next basis indexes through StCameFrom
    ch.cut_states(befcut ? 2 * ns : p.ncut, false, p.myeps, fixed);
    s_DN = DN; s_emin = emin; s_ecut = ecut; s_nkept = nkept;
    return 0;
  }

  // CalcThermo at the 4 betabar of the step: host work that the eigenvector
  // formation does not wait for
  int post_thermo(int N, Tangent* tg) {
    const double DN = s_DN;
    double thT[4], sus[4], ent[4], bN4[4];
    for (int nc = 1; nc <= 4; nc++) {
      double bN = bbar * std::pow(Lambda, (nc - 4) / 8.0);
      bN4[nc - 1] = bN;
      thT[nc - 1] = DN / bN;
      calc_suscep_entropy(bN, ch.Aeig, sus[nc - 1], ent[nc - 1]);
    }
    if (tg) tg->thermo(ch.Aeig, bN4);
    for (int q = 0; q < 4; q++) { s_thT[q] = thT[q]; s_sus[q] = sus[q]; s_ent[q] = ent[q]; }
    (void)N;
    return 0;
  }

  // UpdateMatrices, the DM-NRG capture, CalcStuff, the record, chi for the next site
  int post_rest(int N, nrg_step_cb cb, void* user, std::vector<DmShell>* dm, Tangent* tg) {
    int rc;
    if ((rc = rest_update(N, dm, tg))) return rc;
    if ((rc = ch.op_avgs(bbar, ops))) return rc;
    return rest_emit(N, cb, user, tg);
  }
  // post_rest before the operator averages: UpdateMatrices and the DM-NRG capture
  int rest_update(int N, std::vector<DmShell>* dm, Tangent* tg) {
    int rc;
    if (tg) {
      if ((rc = tg->cut(ch))) return rc;
      if ((rc = ch.plan_update())) return rc;
      if ((rc = tg->ops(ch))) return rc;
      if ((rc = ch.finish_update())) return rc;
    } else if ((rc = ch.update_ops())) {
      return rc;
    }
    if (dm && p.calcdens == 1) {
      dm->emplace_back();
      if ((rc = dm_capture(ch, N, 1, dm->back()))) return rc;
      if (tg && (rc = tg->capture(dm->back(), 1))) return rc;
    }
    return 0;
  }
  // post_rest after the operator averages (in ops): the record, chi for the next site
  int rest_emit(int N, nrg_step_cb cb, void* user, Tangent* tg) {
    int rc;
    const double DN = s_DN, emin = s_emin, ecut = s_ecut;
    const int nkept = s_nkept;
    double* thT = s_thT;
    double* sus = s_sus;
    double* ent = s_ent;
    if (tg && (rc = tg->op_avgs(ch, bbar))) return rc;
    if (befcut) ch.cut_states(p.ncut, false, p.myeps);
    ch.emit(cb, user, emin, ecut, nkept, thT, sus, ent, DN / bbar, ops);
    if (tg)
      for (int t = 0; t < NRGK_MAX_TAN; t++) tg->dchi[t] = 0.0;   // Gamma enters chi_{-1} only
    // chi for the next site: SetChainWilson (src/NRGchainClass.cpp:132-155)
    double d0 = 1.0 - std::pow(Lambda, -((double)N + 1.0));
    double d1 = std::sqrt(1.0 - std::pow(Lambda, -(2.0 * (double)N + 1.0)));
    double d2 = std::sqrt(1.0 - std::pow(Lambda, -(2.0 * (double)N + 3.0)));
    par_chi = d0 / (d1 * d2);
    return 0;
  }
};

// ---------------------------------------------------------------------------
// OneChQS Anderson: NRG_main loop (src/Main/NRG_main.cpp:282-427)
// ---------------------------------------------------------------------------
static int run_onechqs(nrg_ctx* c, nrg_onechqs_params p, nrg_step_cb cb, void* user, double* timing_ms,
                       std::vector<DmShell>* dm = nullptr, int befcut = 0) {
  int rc;
  OneChRun R;
  R.befcut = befcut;
  if ((rc = R.init(c, p, "", cb, user))) return rc;
  // UpdateBefCut 1 rotates the operators over every eigenstate; otherwise only
  // the kept rows are formed, after the truncation
  R.ch.defer = (befcut == 0) && !R.ch.record;
  auto t0 = std::chrono::steady_clock::now();
  for (int N = 0; N < p.nsitesmax; N++) {
    R.basis();
    R.ch.layout(0, 0);
    R.ch.zpar ^= 1;
    DBuf& zb = c->buf(R.ch.zpar ? "Z1" : "Z2");
    DBuf& wb = c->buf("W");
    if ((rc = zb.reserve(sizeof(double) * (size_t)(R.ch.L_ztot + 1)))) return rc;
    if ((rc = wb.reserve(sizeof(double) * (size_t)(R.ch.L_wtot + 1)))) return rc;
    if ((rc = R.assemble((double*)zb.p))) return rc;
    if (R.ch.defer) {
      if ((rc = eig_begin(c, R.ch.Abasis.NumBlocks(), R.ch.L_bsize.data(), R.ch.L_hoff.data(), (double*)zb.p,
                          R.ch.L_woff.data(), (double*)wb.p)))
        return rc;
      R.ch.vec_pending = true;
    } else if ((rc = eig_dev(c, R.ch.Abasis.NumBlocks(), R.ch.L_bsize.data(), R.ch.L_hoff.data(), (double*)zb.p,
                             R.ch.L_woff.data(), (double*)wb.p, nullptr))) {
      return rc;
    }
    if ((rc = R.ch.collect((double*)zb.p, (double*)wb.p))) return rc;
    if ((rc = R.post(N, cb, user, dm))) return rc;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  auto t1 = std::chrono::steady_clock::now();
  if (timing_ms) *timing_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return 0;
}

// Several independent OneChQS instances in lockstep: per step every instance
// builds its basis and assembles H_N into one shared buffer, one eigensolve
// covers the blocks of all instances, then each instance truncates, rotates
// its operators and reports.  All instances run the same number of sites.
// f(0), ..., f(n-1) on up to 16 host threads (instance i on thread i mod T);
// the first nonzero result, or 0.  For host work that touches only the
// instance's own arrays.
template <class F>
static int host_par_for(int n, F f) {
  const int hw = (int)std::thread::hardware_concurrency();
  const int nt = std::max(1, std::min(n, std::min(hw > 0 ? hw : 1, 16)));
  std::vector<int> res(nt, 0);
  auto part = [&](int t) {
    for (int i = t; i < n && !res[t]; i += nt) res[t] = f(i);
  };
  std::vector<std::thread> th;
  int spawned = 0;
  try {
    for (int t = 1; t < nt; t++, spawned++) th.emplace_back(part, t);
  } catch (...) {
  }
  part(0);
  for (int t = 1 + spawned; t < nt; t++) part(t);   // shares of threads that could not start
  for (auto& x : th) x.join();
  for (int r : res)
    if (r) return r;
  return 0;
}

struct MultiCbCtx {
  nrg_multi_cb cb;
  void* user;
  int inst;
};
static void multi_tramp(void* u, const nrg_step* st) {
  MultiCbCtx* m = (MultiCbCtx*)u;
  if (m->cb) m->cb(m->user, m->inst, st);
}

static int run_onechqs_multi(nrg_ctx* c, int ninst, const nrg_onechqs_params* ps, nrg_multi_cb cb, void* user,
                             double* timing_ms) {
  int rc;
  const int nsites = ps[0].nsitesmax;
  for (int i = 0; i < ninst; i++)
    if (ps[i].nsitesmax != nsites) return fail(NRG_ERR_ARG, "nrg_run_onechqs_multi: all instances need the same nsitesmax");
  std::vector<std::unique_ptr<OneChRun>> R(ninst);
  std::vector<MultiCbCtx> cbc(ninst);
  for (int i = 0; i < ninst; i++) {
    cbc[i] = MultiCbCtx{cb, user, i};
    R[i].reset(new OneChRun());
    if ((rc = R[i]->init(c, ps[i], "m" + std::to_string(i) + "_", multi_tramp, &cbc[i]))) return rc;
  }
  bool record_any = false;
  for (int i = 0; i < ninst; i++) record_any = record_any || R[i]->ch.record;
  auto t0 = std::chrono::steady_clock::now();
  int zpar = 0;
  for (int N = 0; N < nsites; N++) {
    int64_t zb0 = 0, wb0 = 0;
    for (int i = 0; i < ninst; i++) {
      R[i]->basis();
      R[i]->ch.layout(zb0, wb0);
      zb0 += R[i]->ch.L_ztot;
      wb0 += R[i]->ch.L_wtot;
    }
    zpar ^= 1;
    DBuf& zb = c->buf(zpar ? "mZ1" : "mZ2");
    DBuf& wb = c->buf("mW");
    if ((rc = zb.reserve(sizeof(double) * (size_t)(zb0 + 1)))) return rc;
    if ((rc = wb.reserve(sizeof(double) * (size_t)(wb0 + 1)))) return rc;
    std::vector<int> n;
    std::vector<int64_t> hoff, woff;
    for (int i = 0; i < ninst; i++) {
      if ((rc = R[i]->assemble((double*)zb.p))) return rc;
      n.insert(n.end(), R[i]->ch.L_bsize.begin(), R[i]->ch.L_bsize.end());
      hoff.insert(hoff.end(), R[i]->ch.L_hoff.begin(), R[i]->ch.L_hoff.end());
      woff.insert(woff.end(), R[i]->ch.L_woff.begin(), R[i]->ch.L_woff.end());
    }
    // every instance's truncation first, then the kept eigenvector rows of all
    // instances in one back-transformation, then the instances' updates
    const bool defer = !record_any;
    if (defer) {
      if ((rc = eig_begin(c, (int)n.size(), n.data(), hoff.data(), (double*)zb.p, woff.data(), (double*)wb.p)))
        return rc;
    } else if ((rc = eig_dev(c, (int)n.size(), n.data(), hoff.data(), (double*)zb.p, woff.data(), (double*)wb.p,
                             nullptr))) {
      return rc;
    }
    std::vector<int> krow(n.size(), 0);
    int base = 0;
    if (defer) {
      // every instance's eigenvalues in one copy; the truncations (host work on
      // each instance's own arrays) on host threads, while nothing else runs
      std::vector<double> wall((size_t)wb0);
      HIPCHK(hipMemcpyAsync(wall.data(), wb.p, sizeof(double) * (size_t)wb0, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      if ((rc = host_par_for(ninst, [&](int i) {
             R[i]->ch.collect_host(wall.data(), (double*)zb.p, (double*)wb.p);
             return R[i]->post_cut(N, nullptr, nullptr);
           })))
        return rc;
      for (int i = 0; i < ninst; i++) {
        R[i]->ch.kept_rows(krow, base);
        base += R[i]->ch.Abasis.NumBlocks();
      }
      if ((rc = eig_finish(c, krow.data()))) return rc;
      // the thermodynamics on host threads while the kept rows are formed
      if ((rc = host_par_for(ninst, [&](int i) { return R[i]->post_thermo(N, nullptr); }))) return rc;
    } else {
      for (int i = 0; i < ninst; i++) {
        if ((rc = R[i]->ch.collect((double*)zb.p, (double*)wb.p))) return rc;
        if ((rc = R[i]->post_cut(N, nullptr, nullptr))) return rc;
        if ((rc = R[i]->post_thermo(N, nullptr))) return rc;
      }
    }
    // every instance's operator updates, then the diagonals for their averages
    // gathered into one buffer and copied back with one wait
    for (int i = 0; i < ninst; i++)
      if ((rc = R[i]->rest_update(N, nullptr, nullptr))) return rc;
    std::vector<int64_t> aoff(ninst + 1, 0);
    for (int i = 0; i < ninst; i++) aoff[i + 1] = aoff[i] + R[i]->ch.op_avgs_len();
    DBuf& ab = c->buf("mOpavg");
    if ((rc = ab.reserve(sizeof(double) * (size_t)(aoff[ninst] + 1)))) return rc;
    for (int i = 0; i < ninst; i++)
      if ((rc = R[i]->ch.op_avgs_issue((double*)ab.p + aoff[i]))) return rc;
    std::vector<double> hall((size_t)aoff[ninst] + 1, 0.0);
    if (aoff[ninst] > 0)
      HIPCHK(hipMemcpyAsync(hall.data(), ab.p, sizeof(double) * (size_t)aoff[ninst], hipMemcpyDeviceToHost,
                            c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < ninst; i++) {
      R[i]->ch.op_avgs_sum(R[i]->bbar, hall.data() + aoff[i], R[i]->ops);
      if ((rc = R[i]->rest_emit(N, multi_tramp, &cbc[i], nullptr))) return rc;
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  auto t1 = std::chrono::steady_clock::now();
  if (timing_ms) *timing_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return 0;
}

// ---------------------------------------------------------------------------
// TwoChQS Anderson: legacy standalone loop (src/TwoChQS/TwoChQS.cpp:204-495);
// qsz: TwoChQSz, the same loop on (Q, Sz) blocks (src/TwoChQSz/TwoChQSz.cpp:103-425)
// ---------------------------------------------------------------------------
//   dm != nullptr (TwoChQSz): also d_up (operator 3), and every shell captured
//   with all its states for the DM-NRG post-processing (dm_capture_all)
static int run_twochqs(nrg_ctx* c, nrg_twochqs_params p, nrg_step_cb cb, void* user, double* timing_ms,
                       bool qsz = false, std::vector<DmShell>* dm = nullptr) {
  const double L = p.Lambda;
  const double bbar = p.betabar > 0 ? p.betabar : 0.727;
  const double HLF = 0.5 * (1.0 + 1.0 / L);
  int rc;
  static const TwoChTables T;
  Chain ch;
  ch.c = c;
  ch.site = qsz ? twochqsz_site() : twochqs_site();
  ch.sz_only = qsz;
  ch.ntypes = 16;
  ch.record = p.record_blocks != 0;
  const SiteDef& st = ch.site;
  // operators: f_ch1, f_ch2 (TwoChQS(z)_UpdateMatrixAfterCutting), plus d for H_0
  ch.specs.resize(3);
  for (int k = 0; k < 2; k++) {
    ch.specs[k].rule = PAIR_QS_UP_UPPER;
    if (qsz)
      ch.specs[k].g = [k](double, double, int ti, int tj) { return qsz_fn_el(k + 1, ti, tj); };
    else
      ch.specs[k].g = [&st, k](double Si, double Sj, int ti, int tj) {
        return twoch_fn_el(st, T, k + 1, Si, Sj, ti, tj);
      };
  }
  ch.specs[2].rule = PAIR_QS_UP_UPPER;
  ch.specs[2].g = [](double, double, int, int) { return 0.0; };
  const bool with_d = qsz && dm;
  if (with_d) {   // d_up: impurity only, past the new site's fermions
    ch.specs.resize(4);
    ch.specs[3].rule = PAIR_QSZ_UP;
    ch.specs[3].need_old = true;
    ch.specs[3].g = [](double, double, int ti, int tj) {
      return ti == tj ? ((__builtin_popcount((unsigned)ti) & 1) ? -1.0 : 1.0) : 0.0;
    };
  }
  ch.init_ops(with_d ? 4 : 3);
  ch.defer = !with_d;
  // src/TwoChQS/TwoChQS.cpp:208-264
  double Ut = 0.5 * p.U / HLF, edt = p.ed / HLF;
  double G1t = p.Gamma1 * (2.0 / M_PI) / (HLF * HLF), G2t = p.Gamma2 * (2.0 / M_PI) / (HLF * HLF);
  double Un = Ut / std::sqrt(L), edn = edt / std::sqrt(L);
  double chi0[2] = {std::sqrt(G1t / L), std::sqrt(G2t / L)};
  // TwoChQS(z)_SetH0Anderson: impurity (N=-1) states and d, basis with the first
  // site of both channels, H_0 assembled with the H_0 fermion-sign rule and no
  // sqrt(Lambda) on the diagonal
  ch.AeigCut = CNRGbasisarray();
  ch.AeigCut.Nshell = -1; ch.AeigCut.NQNumbers = 2;
  if (qsz) {   // src/TwoChQSz/TwoChQSz_SetH0Anderson.cpp:24-135
    ch.AeigCut.QNumbers = {-1, 0, 0, 0.5, 0, -0.5, 1, 0};
    ch.AeigCut.BlockBegEnd = {0, 0, 1, 1, 2, 2, 3, 3};
    ch.AeigCut.BlockBegEndBC = {0, 0, 1, 1, 2, 2, 3, 3};
    ch.AeigCut.dEn = {Un, edn + Un, edn + Un, 2.0 * edn + 3.0 * Un};
    ch.AeigCut.StCameFrom = {0, 1, 2, 3};
    ch.AeigCut.zoff = {0, 1, 2, 3};
    ch.AeigCut.zld = {1, 1, 1, 1};
    if ((rc = ch.set_op(2, {0, 1, 0, 2, 1, 3, 2, 3}, {0, 0, 1, 1, 2, 2, 3, 3}, {1.0, 1.0, -1.0, 1.0}))) return rc;
    // d_up of the impurity (the up entries of the table above)
    if (with_d && (rc = ch.set_op(3, {0, 1, 2, 3}, {0, 0, 1, 1}, {1.0, 1.0}))) return rc;
  } else {
    ch.AeigCut.QNumbers = {-1, 0, 0, 0.5, 1, 0};
    ch.AeigCut.BlockBegEnd = {0, 0, 1, 1, 2, 2};
    ch.AeigCut.BlockBegEndBC = {0, 0, 1, 1, 2, 2};
    ch.AeigCut.dEn = {Un, edn + Un, 2.0 * edn + 3.0 * Un};
    ch.AeigCut.StCameFrom = {0, 1, 2};
    ch.AeigCut.zoff = {0, 1, 2};
    ch.AeigCut.zld = {1, 1, 1};
    if ((rc = ch.set_op(2, {0, 1, 1, 2}, {0, 0, 1, 1}, {1.0, -std::sqrt(2.0)}))) return rc;
  }
  if ((rc = make_basis_dev(c, ch.AeigCut, (ch.tag + "oldb").c_str(), ch.oldb))) return rc;
  auto basis = [&](int befcut) {
    if (qsz) build_basis_qsz(ch.AeigCut, ch.Abasis, st, befcut);
    else build_basis(ch.AeigCut, ch.Abasis, st, 1, befcut);
  };
  basis(1);
  CoefTable h0_tab[2], hn_tab[2];
  for (int k = 0; k < 2; k++) {
    h0_tab[k].ntypes = hn_tab[k].ntypes = 16;
    if (qsz) {
      h0_tab[k].f = [&st, k](double, int a, int b) { return qsz_hn_coef(st, k + 1, a, b, 1); };
      hn_tab[k].f = [&st, k](double, int a, int b) { return qsz_hn_coef(st, k + 1, a, b, 0); };
    } else {
      h0_tab[k].f = [&st, k](double S, int a, int b) { return twoch_hn_coef(st, T, k + 1, S, a, b, 1); };
      hn_tab[k].f = [&st, k](double S, int a, int b) { return twoch_hn_coef(st, T, k + 1, S, a, b, 0); };
    }
  }
  {
    int smax = qsz ? 0 : max_sidx(ch.Abasis);
    std::vector<double> coef[2];
    for (int k = 0; k < 2; k++) coef[k] = h0_tab[k].upto(smax);
    int fops[2] = {2, 2};
    if ((rc = ch.diag_step(2, fops, chi0, coef, 1.0, 0.0))) return rc;
  }
  double DN = HLF * std::pow(L, -(0 - 1) / 2.0);
  double emin = *std::min_element(ch.Aeig.dEn.begin(), ch.Aeig.dEn.end());
  for (auto& x : ch.Aeig.dEn) x -= emin;
  if (p.ncut < ch.Aeig.Nstates())
    return fail(NRG_ERR_ARG, qsz ? "nrg_run_twochqsz: Ncut is below the N=0 basis size; the reference's "
                                   "TwoChQSz_UpdateQm1fQ indexes the uncut basis there"
                                 : "nrg_run_twochqs: Ncut is below the N=0 basis size; the reference's "
                                   "TwoChQS_UpdateQm1fQ indexes the uncut basis there");
  auto finish_step = [&](double emin_) -> int {
    double thT[4] = {DN / bbar, 0, 0, 0};
    double sus[4] = {qsz ? calc_suscep_sz(bbar, ch.Aeig) : calc_suscep(bbar, ch.Aeig), 0, 0, 0};
    double ent[4] = {qsz ? calc_entropy_sz(bbar, ch.Aeig) : calc_entropy(bbar, ch.Aeig), 0, 0, 0};
    double ec = (p.ncut < ch.Aeig.Nstates()) ? ecut_of(ch.Aeig.dEn, p.ncut) : HUGE_VAL;
    const int ns = (int)ch.Aeig.dEn.size();
    ch.Aeig.Kept.assign(ns, 0);
    int nk = 0;
    for (int i = 0; i < ns; i++) { ch.Aeig.Kept[i] = ch.Aeig.dEn[i] <= ec; nk += ch.Aeig.Kept[i]; }
    if (with_d) {
      int r = ch.finish_vectors(true);
      if (r) return r;
      dm->emplace_back();
      r = dm_capture_all(ch, ch.Aeig.Nshell, 3, dm->back());
      if (r) return r;
    }
    // the cut the next iteration starts with (by-value CutStates, no SetDegen)
    ch.cut_states(p.ncut, true, 0.0);
    int r = ch.finish_vectors(false);
    if (r) return r;
    r = ch.update_ops();
    if (r) return r;
    std::vector<double> none;
    ch.emit(cb, user, emin_, ec, nk, thT, sus, ent, DN / bbar, none);
    return 0;
  };
  if ((rc = finish_step(emin))) return rc;
  auto t0 = std::chrono::steady_clock::now();
  for (int N = 1; N <= p.nsitesmax; N++) {
    DN = HLF * std::pow(L, -(N - 1) / 2.0);
    double d0 = 1.0 - std::pow(L, -N), d1 = std::sqrt(1.0 - std::pow(L, -(2 * N - 1)));
    double d2 = std::sqrt(1.0 - std::pow(L, -(2 * N + 1)));
    double chiN[2] = {d0 / (d1 * d2), d0 / (d1 * d2)};
    basis(0);
    int smax = qsz ? 0 : max_sidx(ch.Abasis);
    std::vector<double> coef[2];
    for (int k = 0; k < 2; k++) coef[k] = hn_tab[k].upto(smax);
    int fops[2] = {0, 1};
    if ((rc = ch.diag_step(2, fops, chiN, coef, std::sqrt(L), 0.0))) return rc;
    emin = *std::min_element(ch.Aeig.dEn.begin(), ch.Aeig.dEn.end());
    for (auto& x : ch.Aeig.dEn) x -= emin;
    if ((rc = finish_step(emin))) return rc;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  auto t1 = std::chrono::steady_clock::now();
  if (timing_ms) *timing_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return 0;
}

// The derivative pass: the NRG_main loop with every step's Kept set given (or
// SetKept's when kept is NULL) and ntan tangents along dpar (dU, dGamma, ded).
struct JvpCtx {
  nrg_tangent_cb cb;
  void* user;
  Tangent* tg;
};
static void jvp_tramp(void* u, const nrg_step* st) {
  JvpCtx* j = (JvpCtx*)u;
  nrg_tangent_step dt;
  j->tg->fill(dt, st->nshell, st->nstates, st->nops_avg);
  if (j->cb) j->cb(j->user, st, &dt);
}
static int run_onechqs_jvp(nrg_ctx* c, const nrg_onechqs_params& p, const unsigned char* kept, const int64_t* kept_off,
                           int ntan, const double* dpar, nrg_tangent_cb cb, void* user,
                           std::vector<DmShell>* dm = nullptr) {
  int rc;
  OneChRun R;
  Tangent tg;
  JvpCtx j{cb, user, &tg};
  if ((rc = R.init(c, p, "", nullptr, nullptr))) return rc;
  if ((rc = tg.init(c, ntan, dpar, R.ch, R.Lambda, R.HLF, R.par_chi, p.Gamma))) return rc;
  if ((rc = tg.op_avgs(R.ch, R.bbar))) return rc;
  const double DNm1 = R.HLF * R.Lambda;
  R.ch.emit(jvp_tramp, &j, 0.0, HUGE_VAL, R.ch.Aeig.Nstates(), nullptr, nullptr, nullptr, DNm1 / R.bbar, R.ops);
  for (int N = 0; N < p.nsitesmax; N++) {
    R.basis();
    R.ch.layout(0, 0);
    R.ch.zpar ^= 1;
    DBuf& zb = c->buf(R.ch.zpar ? "Z1" : "Z2");
    DBuf& wb = c->buf("W");
    if ((rc = zb.reserve(sizeof(double) * (size_t)(R.ch.L_ztot + 1)))) return rc;
    if ((rc = wb.reserve(sizeof(double) * (size_t)(R.ch.L_wtot + 1)))) return rc;
    if ((rc = R.assemble((double*)zb.p))) return rc;
    if (ntan > 0) {
      const std::vector<double>& coef = R.hn_tab.upto(max_sidx(R.ch.Abasis));
      if ((rc = tg.assemble(R.ch, coef, R.par_chi, std::sqrt(R.Lambda)))) return rc;
    }
    if ((rc = eig_dev(c, R.ch.Abasis.NumBlocks(), R.ch.L_bsize.data(), R.ch.L_hoff.data(), (double*)zb.p,
                      R.ch.L_woff.data(), (double*)wb.p, nullptr)))
      return rc;
    if ((rc = R.ch.collect((double*)zb.p, (double*)wb.p))) return rc;
    const unsigned char* fixed = nullptr;
    if (kept) {
      if (kept_off[N + 1] - kept_off[N] != R.ch.Aeig.Nstates())
        return fail(NRG_ERR_ARG, "nrg_onechqs_jvp: kept flags of step " + std::to_string(N) + " do not match its " +
                                     std::to_string(R.ch.Aeig.Nstates()) + " states");
      fixed = kept + kept_off[N];
    }
    if ((rc = R.post(N, jvp_tramp, &j, dm, ntan > 0 ? &tg : nullptr, fixed))) return rc;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// the DM-NRG chain with frozen Kept sets and tangents captured per shell
int run_onechqs_dm_jvp(nrg_ctx* c, const nrg_onechqs_params& p, std::vector<DmShell>& dm, const unsigned char* kept,
                       const int64_t* kept_off, int ntan, const double* dpar) {
  return run_onechqs_jvp(c, p, kept, kept_off, ntan, dpar, nullptr, nullptr, &dm);
}

// the TwoChQSz chain with every shell captured for the DM-NRG post-processing
int run_twochqsz_dm(nrg_ctx* c, const nrg_twochqs_params& p, std::vector<DmShell>& dm) {
  return run_twochqs(c, p, nullptr, nullptr, nullptr, true, &dm);
}

int run_onechqs_dm(nrg_ctx* c, const nrg_onechqs_params& p, std::vector<DmShell>& dm, int befcut) {
  return run_onechqs(c, p, nullptr, nullptr, nullptr, &dm, befcut);
}

}  // namespace nrg

extern "C" int nrg_run_onechqs(nrg_ctx* c, const nrg_onechqs_params* p, nrg_step_cb cb, void* user,
                               double* timing_ms) {
  if (!c || !p) return nrg::fail(NRG_ERR_ARG, "nrg_run_onechqs: NULL argument");
  if (p->nsitesmax < 0 || p->ncut < 1 || p->Lambda <= 1.0)
    return nrg::fail(NRG_ERR_ARG, "nrg_run_onechqs: bad parameters");
  HIPCHK(hipSetDevice(c->device));
  return nrg::run_onechqs(c, *p, cb, user, timing_ms);
}

extern "C" int nrg_onechqs_jvp(nrg_ctx* c, const nrg_onechqs_params* p, const unsigned char* kept,
                               const int64_t* kept_off, int ntan, const double* dpar, nrg_tangent_cb cb, void* user) {
  if (!c || !p) return nrg::fail(NRG_ERR_ARG, "nrg_onechqs_jvp: NULL argument");
  if (p->nsitesmax < 0 || p->ncut < 1 || p->Lambda <= 1.0)
    return nrg::fail(NRG_ERR_ARG, "nrg_onechqs_jvp: bad parameters");
  if (ntan < 0 || ntan > NRG_MAX_TAN || (ntan > 0 && !dpar))
    return nrg::fail(NRG_ERR_ARG, "nrg_onechqs_jvp: ntan must be 0..4 with dpar given");
  if (ntan > 0 && !(p->Gamma > 0.0)) return nrg::fail(NRG_ERR_ARG, "nrg_onechqs_jvp: derivatives need Gamma > 0");
  if (kept && !kept_off) return nrg::fail(NRG_ERR_ARG, "nrg_onechqs_jvp: kept without kept_off");
  HIPCHK(hipSetDevice(c->device));
  return nrg::run_onechqs_jvp(c, *p, kept, kept_off, ntan, dpar, cb, user);
}

extern "C" int nrg_run_twochqsz(nrg_ctx* c, const nrg_twochqs_params* p, nrg_step_cb cb, void* user,
                                double* timing_ms) {
  if (!c || !p) return nrg::fail(NRG_ERR_ARG, "nrg_run_twochqsz: NULL argument");
  if (p->nsitesmax < 0 || p->ncut < 1 || p->Lambda <= 1.0)
    return nrg::fail(NRG_ERR_ARG, "nrg_run_twochqsz: bad parameters");
  HIPCHK(hipSetDevice(c->device));
  return nrg::run_twochqs(c, *p, cb, user, timing_ms, true);
}

extern "C" int nrg_onechqs_thermo(nrg_ctx* c, const nrg_onechqs_params* p, double* out, int cap_rows, int* nrows) {
  if (!c || !p || !nrows) return nrg::fail(NRG_ERR_ARG, "nrg_onechqs_thermo: NULL argument");
  if (p->nsitesmax < 1 || p->ncut < 1 || p->Lambda <= 1.0)
    return nrg::fail(NRG_ERR_ARG, "nrg_onechqs_thermo: bad parameters");
  const int rows = 4 * p->nsitesmax;
  *nrows = rows;
  if (!out) return NRG_OK;
  if (cap_rows < rows) return nrg::fail(NRG_ERR_ARG, "nrg_onechqs_thermo: output too small");
  HIPCHK(hipSetDevice(c->device));
  struct Acc {
    std::vector<double> T, chi, S;
  } imp, chain;
  auto collect = [](void* u, const nrg_step* st) {
    if (st->nshell < 0) return;   // CalcThermo runs for N >= Nsites0 = 0
    Acc* a = (Acc*)u;
    for (int nc = 0; nc < 4; nc++) {
      a->T.push_back(st->thermo_T[nc]);
      a->chi.push_back(st->suscep[nc]);
      a->S.push_back(st->entropy[nc]);
    }
  };
  nrg_onechqs_params q = *p;
  q.calcdens = 2;
  q.record_blocks = 0;
  nrg_onechqs_params qc = q;
  qc.U = 0.0; qc.ed = 0.0; qc.Gamma = 0.0;
  int rc;
  if ((rc = nrg::run_onechqs(c, qc, collect, &chain, nullptr))) return rc;
  if ((rc = nrg::run_onechqs(c, q, collect, &imp, nullptr))) return rc;
  if ((int)imp.T.size() != rows || (int)chain.T.size() != rows)
    return nrg::fail(NRG_ERR_STATE, "nrg_onechqs_thermo: step count");
  for (int r = 0; r < rows; r++) {
    const double cchi = chain.chi[r] - 1.0 / 8.0, cS = chain.S[r] - 2.0 * std::log(2.0);
    double* o = out + (size_t)7 * r;
    o[0] = imp.T[r];
    o[1] = imp.chi[r]; o[2] = cchi; o[3] = imp.chi[r] - cchi;
    o[4] = imp.S[r]; o[5] = cS; o[6] = imp.S[r] - cS;
  }
  return NRG_OK;
}

extern "C" int nrg_run_onechqs_multi(nrg_ctx* c, int ninst, const nrg_onechqs_params* p, nrg_multi_cb cb,
                                     void* user, double* timing_ms) {
  if (!c || !p || ninst < 1) return nrg::fail(NRG_ERR_ARG, "nrg_run_onechqs_multi: bad arguments");
  for (int i = 0; i < ninst; i++)
    if (p[i].nsitesmax < 0 || p[i].ncut < 1 || p[i].Lambda <= 1.0)
      return nrg::fail(NRG_ERR_ARG, "nrg_run_onechqs_multi: bad parameters");
  HIPCHK(hipSetDevice(c->device));
  return nrg::run_onechqs_multi(c, ninst, p, cb, user, timing_ms);
}

extern "C" int nrg_run_twochqs(nrg_ctx* c, const nrg_twochqs_params* p, nrg_step_cb cb, void* user,
                               double* timing_ms) {
  if (!c || !p) return nrg::fail(NRG_ERR_ARG, "nrg_run_twochqs: NULL argument");
  if (p->nsitesmax < 0 || p->ncut < 1 || p->Lambda <= 1.0)
    return nrg::fail(NRG_ERR_ARG, "nrg_run_twochqs: bad parameters");
  HIPCHK(hipSetDevice(c->device));
  return nrg::run_twochqs(c, *p, cb, user, timing_ms);
}
