// nrg_driver.cpp -- the NRG_main-style chain driver on the MI355X path.
//
// Host objects keep the reference's data-model semantics
// (include/NRGclasses.hpp:47-759): QNumbers per block, BlockBegEnd inclusive
// pairs, dEn per state, Kept flags, iType/StCameFrom/BlockBegEndBC for basis
// arrays, MatBlockMap/MatBlockBegEnd for operators.  Eigenvectors and
// operator elements stay resident in HBM; the host holds the integer
// bookkeeping (basis build, block maps, truncation decisions) and the O(n)
// FP64 Boltzmann sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <utility>
#include <vector>

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

enum OpKind { OP_FN = 0, OP_IMPONLY = 1, OP_CD = 2 };

// an operator over a kept basis; elements on the device
struct CNRGmatrix {
  OpKind kind = OP_FN;
  bool NeedOld = false, CalcAvg = false;
  std::vector<int> MatBlockMap, MatBlockBegEnd;
  int nblocks_basis = 0;        // blocks of the basis it is synced to
  std::string buf[2];           // device buffer names (double-buffered)
  int cur = 0;                  // which buffer holds the current elements
  int64_t nel = 0;
};

static long key2(double x) { return std::lround(2.0 * x); }

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
// OneChQS model plugin: single site (Q,S,Sz) one state per block
// (src/OneChQS_SetSingleSite.cpp), f^dagger table (src/OpTables.cpp:8-33),
// H_N / f_N / c_d element rules (src/OpMatRules.cpp:1167-1535).
// ---------------------------------------------------------------------------
struct OneChQSSite {
  // type -> (Q, S, Sz); type == site block index
  double Q[4] = {-1, 0, 0, 1}, S[4] = {0, 0.5, 0.5, 0}, Sz[4] = {0, 0.5, -0.5, 0};
  int find(double q, double s, double sz) const {
    for (int t = 0; t < 4; t++)
      if (std::fabs(Q[t] - q) < 1e-20 && std::fabs(S[t] - s) < 1e-20 && std::fabs(Sz[t] - sz) < 1e-20) return t;
    return -1;
  }
  static double fd(int sigma, int ti, int tj) {
    if (sigma == 1) return (ti == 1 && tj == 0) ? 1.0 : ((ti == 3 && tj == 2) ? 1.0 : 0.0);
    if (sigma == -1) return (ti == 2 && tj == 0) ? 1.0 : ((ti == 3 && tj == 1) ? -1.0 : 0.0);
    return 0.0;
  }
  // CoefMatEl * FermiSign of OneChQS_HN_MatEl for a block of spin Si, site types (typep, type)
  double hn_coef(double Si, int tp, int tt) const {
    double Stp = S[tp], Sztp = Sz[tp], Qt = Q[tt], St = S[tt], Szt = Sz[tt];
    double Scfp = Si - Sztp, Scf = Si - Szt;
    double fsign = std::fabs(Qt) < 1e-20 ? -1.0 : 1.0;
    double coef = 0.0;
    for (double sz = -St; sz <= St; sz += 1.0) {
      int ss = find(Qt, St, sz);
      for (int sg = -1; sg <= 1; sg += 2) {
        double ds = 0.5 * sg, szcf = Si - sz, szcfp = szcf - ds, sztp = sz + ds;
        int ssp = find(Q[tp], Stp, sz + ds);
        coef += cgordan(Scf, szcf, St, sz, Si, Si) * cgordan(Scfp, szcfp, Stp, sztp, Si, Si) *
                cgordan(Scfp, szcfp, 0.5, ds, Scf, szcf) * fd(sg, ssp, ss);
      }
    }
    return coef * fsign;
  }
  // OneChQS_fN_MatEl for states of spins (Si,Sj) with the same parent
  double fn_el(double Si, double Sj, int ti, int tj) const {
    double Sti = S[ti], Szti = Sz[ti], Stj = S[tj], Sztj = Sz[tj];
    double Scfi = Si - Szti, Scfj = Sj - Sztj, sigma = Sj - Si, acc = 0.0;
    for (double sz = -Stj; sz <= Stj; sz += 1.0) {
      int ssj = find(Q[tj], Stj, sz), ssi = find(Q[ti], Sti, sz - sigma);
      acc += fd((int)(sigma / 0.5), ssj, ssi) * cgordan(Scfi, Si - sz + sigma, Sti, sz - sigma, Si, Si) *
             cgordan(Scfj, Sj - sz, Stj, sz, Sj, Sj);
    }
    double cgr = cgordan(Si, Si, 0.5, sigma, Sj, Sj);
    return std::fabs(cgr) >= 1e-20 ? acc / cgr : 0.0;
  }
  // OneChQS_cd_MatEl
  double cd_el(double Si, double Sj, int ti, int tj) const {
    double Sti = S[ti], Szti = Sz[ti], Qtj = Q[tj], Stj = S[tj], Sztj = Sz[tj];
    double Scfi = Si - Szti, Scfj = Sj - Sztj, sigma = Sj - Si;
    double cg3 = cgordan(Si, Si, 0.5, sigma, Sj, Sj);
    if (std::fabs(cg3) < 1e-20) {
      sigma = Si - Sj;
      cg3 = cgordan(Si, Si, 0.5, sigma, Sj, Sj);
      if (std::fabs(cg3) < 1e-20) return 0.0;
    }
    double fb = 0.0;
    for (double szcfj = -Scfj; szcfj <= Scfj; szcfj += 1.0) {
      double szcfi = szcfj - sigma, sztj = Sj - szcfj, szti = Si - szcfi;
      int ssj = find(Qtj, Stj, sztj), ssi = find(Q[ti], Sti, szti);
      double fs = std::fabs(Qtj) < 1e-20 ? -1.0 : 1.0;
      if (ssi == ssj)
        fb += cgordan(Scfi, szcfi, Sti, szti, Si, Si) * cgordan(Scfj, szcfj, Stj, sztj, Sj, Sj) *
              cgordan(Scfi, szcfi, 0.5, sigma, Scfj, szcfj) * fs;
    }
    return std::fabs(cg3) > 1e-20 ? fb / cg3 : 0.0;
  }
};

// OneChQS_cd_check (src/OpMatRules.cpp:106-125) / Diag_check (:12-24)
static bool cd_check(const CNRGarray& A, int i, int j) {
  double Qi = A.GetQNumber(i, 0), Si = A.GetQNumber(i, 1), Qj = A.GetQNumber(j, 0), Sj = A.GetQNumber(j, 1);
  return std::fabs(Qj - (Qi + 1.0)) < 1e-20 &&
         (std::fabs(Sj - (Si + 0.5)) < 1e-20 || std::fabs(Sj - (Si - 0.5)) < 1e-20);
}

// ---------------------------------------------------------------------------
// BuildBasis for (Q,S) with SU(2) coupling (src/BuildBasis.cpp:8-432 with
// CommonQNs={2,0,1,0,1}, totSpos={1}).  New blocks in first-appearance order
// of (Q_old+Q_site, S_old+S_site, then S descending to |S_old-S_site|)
// (FilterQNumbers keeps first occurrences); states per new block ordered by
// (old block, site block, old state).
// ---------------------------------------------------------------------------
static void build_basis_qs(const CNRGbasisarray& Ac, CNRGbasisarray& Ab, const OneChQSSite& site) {
  Ab = CNRGbasisarray();
  Ab.Nshell = Ac.Nshell + 1;
  Ab.NQNumbers = 2;
  std::map<std::pair<long, long>, int> seen;
  std::vector<std::pair<double, double>> qns;
  auto add = [&](double q, double s) {
    auto k = std::make_pair(key2(q), key2(s));
    if (seen.emplace(k, (int)qns.size()).second) qns.push_back({q, s});
  };
  for (int b = 0; b < Ac.NumBlocks(); b++) {
    double Qo = Ac.GetQNumber(b, 0), So = Ac.GetQNumber(b, 1);
    for (int s = 0; s < 4; s++) {
      add(Qo + site.Q[s], So + site.S[s]);
      for (double sn = So + site.S[s] - 1.0; sn >= std::fabs(So - site.S[s]); sn -= 1.0) add(Qo + site.Q[s], sn);
    }
  }
  int ist = 0;
  for (auto& qs : qns) {
    Ab.QNumbers.push_back(qs.first);
    Ab.QNumbers.push_back(qs.second);
    Ab.BlockBegEnd.push_back(ist);
    for (int bo = 0; bo < Ac.NumBlocks(); bo++) {
      double Qo = Ac.GetQNumber(bo, 0), So = Ac.GetQNumber(bo, 1);
      for (int s = 0; s < 4; s++) {
        if (key2(qs.first) != key2(site.Q[s] + Qo)) continue;
        if (key2(qs.second) != key2(site.Sz[s] + So) || !triangle(So, site.S[s], qs.second)) continue;
        for (int io = Ac.GetBlockLimit(bo, 0); io <= Ac.GetBlockLimit(bo, 1); io++) {
          Ab.iType.push_back(s);
          Ab.dEn.push_back(Ac.dEn[io]);
          Ab.StCameFrom.push_back(io);
          ist++;
        }
      }
    }
    Ab.BlockBegEnd.push_back(ist - 1);
  }
}

// ---------------------------------------------------------------------------
// device view of an operator over a kept basis
// ---------------------------------------------------------------------------
struct BasisDev {  // per-state block/position, per-block size of a kept basis
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
// thermodynamics (src/CalcThermo.cpp:16-121, src/CalcOpAvg.cpp:13-84)
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

// ---------------------------------------------------------------------------
// the chain driver
// ---------------------------------------------------------------------------
struct Driver {
  nrg_ctx* c;
  nrg_onechqs_params p;
  OneChQSSite site;
  CNRGarray Aeig;
  CNRGbasisarray AeigCut, Abasis;
  std::vector<CNRGmatrix> mats;
  std::vector<int64_t> Aeig_zoff;   // offset of each Aeig block in the Z buffer
  double* dZ = nullptr;
  double* dW = nullptr;
  int parity = 0;                   // Z buffer parity (double-buffered per step)

  int emit(nrg_step_cb cb, void* user, double emin, double ecut, int nkept, const double* th_T, const double* sus,
           const double* ent, double Tavg, const std::vector<double>& ops, const std::vector<double>* hrec,
           const std::vector<double>* zrec) {
    if (!cb) return 0;
    nrg_step st;
    std::memset(&st, 0, sizeof(st));
    st.nshell = Aeig.Nshell; st.nq = Aeig.NQNumbers; st.nblocks = Aeig.NumBlocks();
    st.qn = Aeig.QNumbers.data(); st.bbe = Aeig.BlockBegEnd.data();
    st.nstates = (int)Aeig.dEn.size(); st.en = Aeig.dEn.data();
    std::vector<unsigned char> kp(Aeig.dEn.size(), 1);
    for (size_t i = 0; i < kp.size() && i < Aeig.Kept.size(); i++) kp[i] = Aeig.Kept[i];
    st.kept = kp.data();
    st.emin = emin; st.ecut = ecut; st.nkept = nkept;
    st.hpacked = hrec ? hrec->data() : nullptr;
    st.evec = zrec ? zrec->data() : nullptr;
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
    return 0;
  }

  // CutStates (src/CutStates.cpp:5-96): AeigCut keeps, per surviving block,
  // the energy-ordered prefix up to the first level above Ecut; blocks whose
  // lowest level is above Ecut are dropped.  Eigenvector rows are not moved:
  // AeigCut records where each block's kept rows start in the Z buffer.
  void cut_states(int ncut) {
    CNRGbasisarray A;
    A.Nshell = Aeig.Nshell; A.NQNumbers = Aeig.NQNumbers;
    int nb = Aeig.NumBlocks();
    const bool notrunc = ncut >= Aeig.Nstates();
    std::vector<double> en = Aeig.dEn;
    std::vector<unsigned char> kept = Aeig.Kept;
    double ec = HUGE_VAL;
    if (!notrunc) {
      ec = ecut_of(Aeig.dEn, ncut);
      set_degen(en, p.myeps);   // Acut.SetKept(Ncutoff)
      double ec2 = ecut_of(en, ncut);
      kept.assign(en.size(), 0);
      for (size_t i = 0; i < en.size(); i++) kept[i] = en[i] <= ec2;
    }
    int ist = 0;
    for (int b = 0; b < nb; b++) {
      int b0 = Aeig.GetBlockLimit(b, 0), b1 = Aeig.GetBlockLimit(b, 1);
      int nbc = b1 - b0 + 1;
      int k = nbc;
      if (!notrunc) {
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
        A.Kept.push_back(i + b0 < (int)kept.size() ? kept[b0 + i] : 1);
        A.StCameFrom.push_back(b0 + i);
      }
      A.zoff.push_back(Aeig_zoff[b]);
      A.zld.push_back(nbc);
      ist += k;
    }
    AeigCut = std::move(A);
  }

  int run(nrg_step_cb cb, void* user, double* timing_ms);
};

int Driver::run(nrg_step_cb cb, void* user, double* timing_ms) {
  const double Lambda = p.Lambda;
  const double bbar = p.betabar > 0 ? p.betabar : 0.727;
  if (p.myeps <= 0) p.myeps = 1e-18;
  const double HLF = 0.5 * (1.0 + 1.0 / Lambda);
  int rc;
  // operators: [0] f_N, then n_d, n_d^2, S_d^2 (calcdens 0/2) or c_d (calcdens 1)
  int nm = (p.calcdens == 1) ? 2 : 4;
  mats.assign(nm, CNRGmatrix());
  mats[0].kind = OP_FN;
  for (int m = 1; m < nm; m++) {
    mats[m].kind = (p.calcdens == 1) ? OP_CD : OP_IMPONLY;
    mats[m].NeedOld = true;
    mats[m].CalcAvg = (p.calcdens != 1);
  }
  for (int m = 0; m < nm; m++) {
    mats[m].buf[0] = "op" + std::to_string(m) + "_a";
    mats[m].buf[1] = "op" + std::to_string(m) + "_b";
  }
  // ---- H_{-1}: OneChQS_SetAnderson_Hm1 (src/Main/Main_SetH0.cpp:1794-1969)
  double chi_m1 = std::sqrt(2.0 * p.Gamma / M_PI) / (std::sqrt(Lambda) * HLF);
  if (p.disc_scheme != 1) chi_m1 *= std::sqrt(0.5 * std::log(Lambda) * (Lambda + 1) / (Lambda - 1));
  Aeig = CNRGarray();
  Aeig.Nshell = -1; Aeig.NQNumbers = 2;
  Aeig.QNumbers = {-1, 0, 0, 0.5, 1, 0};
  Aeig.BlockBegEnd = {0, 0, 1, 1, 2, 2};
  Aeig.dEn = {0.5 * p.U / (Lambda * HLF), (p.ed + 0.5 * p.U) / (Lambda * HLF), (2.0 * p.ed + 1.5 * p.U) / (Lambda * HLF)};
  Aeig_zoff = {0, 1, 2};
  {
    DBuf& zb = c->buf("Z0");
    if ((rc = zb.reserve(sizeof(double) * 8))) return rc;
    double ones[3] = {1, 1, 1};
    HIPCHK(hipMemcpyAsync(zb.p, ones, sizeof(ones), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dZ = (double*)zb.p;
  }
  // operators over the 3 H_{-1} states
  {
    std::vector<double> f = {1.0, -std::sqrt(2.0)};
    mats[0].MatBlockMap = {0, 1, 1, 2};
    mats[0].MatBlockBegEnd = {0, 0, 1, 1};
    DBuf& b0 = c->buf(mats[0].buf[0]);
    if ((rc = b0.reserve(sizeof(double) * 2))) return rc;
    HIPCHK(hipMemcpyAsync(b0.p, f.data(), sizeof(double) * 2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    mats[0].cur = 0;
    for (int m = 1; m < nm; m++) {
      std::vector<double> v(3);
      if (p.calcdens == 1) {
        mats[m].MatBlockMap = mats[0].MatBlockMap;
        mats[m].MatBlockBegEnd = mats[0].MatBlockBegEnd;
        v = f;
      } else {
        mats[m].MatBlockMap = {0, 0, 1, 1, 2, 2};
        mats[m].MatBlockBegEnd = {0, 0, 1, 1, 2, 2};
        for (int b = 0; b < 3; b++) {
          double nd = Aeig.GetQNumber(b, 0) + 1.0, Sd = Aeig.GetQNumber(b, 1);
          v[b] = (m == 1) ? nd : (m == 2 ? nd * nd : Sd * (Sd + 1.0));
        }
      }
      DBuf& bm = c->buf(mats[m].buf[0]);
      if ((rc = bm.reserve(sizeof(double) * v.size()))) return rc;
      HIPCHK(hipMemcpyAsync(bm.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      mats[m].cur = 0;
    }
  }
  cut_states(p.ncut);
  HIPCHK(hipStreamSynchronize(c->stream));
  auto op_avgs = [&](std::vector<double>& out) -> int {
    out.clear();
    for (int m = 1; m < nm; m++) {
      if (!mats[m].CalcAvg) continue;
      const CNRGmatrix& M = mats[m];
      int64_t tot = M.MatBlockBegEnd.empty() ? 0 : M.MatBlockBegEnd.back() + 1;
      std::vector<double> h(tot > 0 ? tot : 1);
      if (tot > 0)
        HIPCHK(hipMemcpy(h.data(), c->buf(M.buf[M.cur]).p, sizeof(double) * tot, hipMemcpyDeviceToHost));
      double s = 0, Z = 0;
      for (int b = 0; b < AeigCut.NumBlocks(); b++) {
        double dg = 2.0 * AeigCut.GetQNumber(b, 1) + 1.0;
        int n = AeigCut.GetBlockSize(b);
        int64_t off = -1;
        for (size_t q = 0; q + 1 < M.MatBlockMap.size(); q += 2)
          if (M.MatBlockMap[q] == b && M.MatBlockMap[q + 1] == b) { off = M.MatBlockBegEnd[q]; break; }
        for (int i = 0; i < n; i++) {
          double x = std::exp(-bbar * AeigCut.dEn[AeigCut.GetBlockLimit(b, 0) + i]);
          double el = off >= 0 ? h[off + (int64_t)i * n + i] : 0.0;
          s += dg * el * x; Z += dg * x;
        }
      }
      out.push_back(std::fabs(Z) > 1e-20 ? s / Z : s);
    }
    return 0;
  };
  std::vector<double> ops;
  if ((rc = op_avgs(ops))) return rc;
  double DN = HLF * std::pow(Lambda, -(0 - 2) / 2.0);
  emit(cb, user, 0.0, HUGE_VAL, Aeig.Nstates(), nullptr, nullptr, nullptr, DN / bbar, ops, nullptr, nullptr);

  double par_chi = chi_m1;
  auto t0 = std::chrono::steady_clock::now();
  for (int N = 0; N < p.nsitesmax; N++) {
    DN = HLF * std::pow(Lambda, -(N - 1) / 2.0);
    // ---- BuildBasis (host)
    build_basis_qs(AeigCut, Abasis, site);
    const int nb = Abasis.NumBlocks();
    const int ns = Abasis.Nstates();
    std::vector<int> bbeg(nb), bsize(nb), sidx(nb);
    std::vector<int64_t> h_off(nb), w_off(nb);
    int64_t ztot = 0;
    int nmax = 0, smax = 0;
    for (int b = 0; b < nb; b++) {
      bbeg[b] = Abasis.GetBlockLimit(b, 0);
      bsize[b] = Abasis.GetBlockSize(b);
      sidx[b] = (int)key2(Abasis.GetQNumber(b, 1));
      smax = std::max(smax, sidx[b]);
      h_off[b] = ztot;
      w_off[b] = bbeg[b];
      ztot += (int64_t)bsize[b] * bsize[b];
      nmax = std::max(nmax, bsize[b]);
    }
    // coupling tables C[sidx][tp][t]
    std::vector<double> coef((size_t)(smax + 1) * 16, 0.0);
    for (int si = 0; si <= smax; si++)
      for (int tp = 0; tp < 4; tp++)
        for (int t = 0; t < 4; t++) coef[(size_t)si * 16 + tp * 4 + t] = site.hn_coef(0.5 * si, tp, t);
    // ---- device: upload basis, assemble H_N, eigensolve
    parity ^= 1;
    DBuf& zb = c->buf(parity ? "Z1" : "Z2");
    DBuf& wb = c->buf("W");
    if ((rc = zb.reserve(sizeof(double) * (size_t)(ztot + 1)))) return rc;
    if ((rc = wb.reserve(sizeof(double) * (size_t)(ns + 1)))) return rc;
    double* dZn = (double*)zb.p;
    dW = (double*)wb.p;
    int *d_bbeg, *d_bsize, *d_sidx, *d_stcf, *d_itype;
    int64_t* d_hoff;
    double *d_en, *d_coef, *d_siteq;
    if ((rc = upload(c, "hb_bbeg", bbeg, &d_bbeg))) return rc;
    if ((rc = upload(c, "hb_bsize", bsize, &d_bsize))) return rc;
    if ((rc = upload(c, "hb_sidx", sidx, &d_sidx))) return rc;
    if ((rc = upload(c, "hb_hoff", h_off, &d_hoff))) return rc;
    if ((rc = upload(c, "hb_stcf", Abasis.StCameFrom, &d_stcf))) return rc;
    if ((rc = upload(c, "hb_itype", Abasis.iType, &d_itype))) return rc;
    if ((rc = upload(c, "hb_en", Abasis.dEn, &d_en))) return rc;
    if ((rc = upload(c, "hb_coef", coef, &d_coef))) return rc;
    std::vector<double> siteq(site.Q, site.Q + 4);
    if ((rc = upload(c, "hb_siteq", siteq, &d_siteq))) return rc;
    BasisDev oldb;
    if ((rc = make_basis_dev(c, AeigCut, "oldb", oldb))) return rc;
    HnArgs ha{};
    ha.H = dZn; ha.h_off = d_hoff; ha.bbeg = d_bbeg; ha.bsize = d_bsize; ha.blk_sidx = d_sidx;
    ha.stcf = d_stcf; ha.itype = d_itype; ha.en_par = d_en; ha.site_q = d_siteq;
    ha.coef[0] = d_coef;
    if ((rc = make_oldop(c, mats[0], oldb, "hf0", ha.f[0]))) return rc;
    ha.chi[0] = par_chi;
    ha.sqrtL = std::sqrt(Lambda);
    ha.eN = 0.0;
    ha.nch = 1;
    ha.ntypes = 4;
    {
      double by = 0.0;
      for (int b = 0; b < nb; b++) by += 8.0 * (double)bsize[b] * bsize[b];
      PendingTimer tm;
      if ((rc = c->timer_begin(tm, NRG_K_HN, 0.0, by))) return rc;
      HIPCHK(nrgk_launch_hn(&ha, nb, nmax, c->stream));
      if ((rc = c->timer_end(tm))) return rc;
    }
    std::vector<double> hrec, zrec;
    if (p.record_blocks) {
      std::vector<double> hfull(ztot);
      HIPCHK(hipMemcpyAsync(hfull.data(), dZn, sizeof(double) * ztot, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      for (int b = 0; b < nb; b++) {
        int n = bsize[b];
        for (int i = 0; i < n; i++)
          for (int j = i; j < n; j++) hrec.push_back(hfull[h_off[b] + (int64_t)i * n + j]);
      }
    }
    if ((rc = eig_dev(c, nb, bsize.data(), h_off.data(), dZn, w_off.data(), dW, nullptr))) return rc;
    // ---- eigenvalues to the host: SetE0zero, SetKept, thermo
    Aeig = CNRGarray();
    Aeig.Nshell = Abasis.Nshell; Aeig.NQNumbers = 2;
    Aeig.QNumbers = Abasis.QNumbers;
    Aeig.BlockBegEnd = Abasis.BlockBegEnd;
    Aeig.dEn.resize(ns);
    HIPCHK(hipMemcpyAsync(Aeig.dEn.data(), dW, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
    if (p.record_blocks) {
      zrec.resize(ztot);
      HIPCHK(hipMemcpyAsync(zrec.data(), dZn, sizeof(double) * ztot, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    double emin = *std::min_element(Aeig.dEn.begin(), Aeig.dEn.end());
    for (auto& x : Aeig.dEn) x -= emin;
    set_degen(Aeig.dEn, p.myeps);
    double ecut = ecut_of(Aeig.dEn, p.ncut);
    Aeig.Kept.assign(ns, 0);
    int nkept = 0;
    for (int i = 0; i < ns; i++) { Aeig.Kept[i] = Aeig.dEn[i] <= ecut; nkept += Aeig.Kept[i]; }
    double thT[4], sus[4], ent[4];
    for (int nc = 1; nc <= 4; nc++) {
      double bN = bbar * std::pow(Lambda, (nc - 4) / 8.0);
      thT[nc - 1] = DN / bN;
      sus[nc - 1] = calc_suscep(bN, Aeig);
      ent[nc - 1] = calc_entropy(bN, Aeig);
    }
    Aeig_zoff = h_off;
    dZ = dZn;
    CNRGbasisarray prevCut = std::move(AeigCut);
    cut_states(p.ncut);
    // ---- UpdateMatrices (src/UpdateMatrices.cpp:272-568) on the device
    {
      std::vector<FPair> pairs;
      std::vector<double> G;
      std::vector<RotJob> jobs;
      std::vector<std::vector<int>> nmap(nm), nbe(nm);
      std::vector<int64_t> icount(nm, 0);
      int64_t ftot = 0, auxtot = 0, maxel = 1;
      const int nbc = AeigCut.NumBlocks();
      std::vector<int> job_op;  // operator of each job
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
            bool ok = (mats[m].kind == OP_IMPONLY) ? (ib == jb) : cd_check(AeigCut, ib, jb);
            if (!ok) continue;
            nmap[m].push_back(ib); nmap[m].push_back(jb);
            nbe[m].push_back((int)icount[m]);
            int64_t out_off = icount[m];
            icount[m] += (int64_t)ki * kj;
            nbe[m].push_back((int)icount[m] - 1);
            FPair P{};
            P.ni_bc = ni; P.nj_bc = nj; P.ibeg_bc = ibc_beg; P.jbeg_bc = jbc_beg;
            P.f_off = ftot; P.g_off = (int64_t)G.size(); P.need_old = mats[m].NeedOld ? 1 : 0; P.op = m;
            for (int ti = 0; ti < 4; ti++)
              for (int tj = 0; tj < 4; tj++) {
                double g = 0.0;
                if (mats[m].kind == OP_FN) g = site.fn_el(Si, Sj, ti, tj);
                else if (mats[m].kind == OP_IMPONLY) g = (ti == tj) ? 1.0 : 0.0;
                else g = site.cd_el(Si, Sj, ti, tj);
                G.push_back(g);
              }
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
      // F for every job, then per-operator GEMMs into that operator's new buffer
      DBuf& fb = c->buf("upd_F");
      DBuf& ab = c->buf("upd_aux");
      if ((rc = fb.reserve(sizeof(double) * (size_t)(ftot + 1)))) return rc;
      if ((rc = ab.reserve(sizeof(double) * (size_t)(auxtot + 1)))) return rc;
      FPair* d_pairs;
      double* d_G;
      if ((rc = upload(c, "upd_pairs", pairs, &d_pairs))) return rc;
      if ((rc = upload(c, "upd_G", G, &d_G))) return rc;
      FArgs fa{};
      fa.F = (double*)fb.p; fa.pairs = d_pairs; fa.G = d_G; fa.stcf = d_stcf; fa.itype = d_itype; fa.ntypes = 4;
      for (int m = 0; m < nm; m++)
        if (mats[m].NeedOld)
          if ((rc = make_oldop(c, mats[m], oldb, "uo" + std::to_string(m), fa.old[m]))) return rc;
      {
        PendingTimer tm;
        if ((rc = c->timer_begin(tm, NRG_K_FBASIS, 0.0, 8.0 * (double)ftot))) return rc;
        HIPCHK(nrgk_launch_fbasis(&fa, (int)pairs.size(), maxel, c->stream));
        if ((rc = c->timer_end(tm))) return rc;
      }
      for (int m = 0; m < nm; m++) {
        std::vector<RotJob> jm;
        for (size_t q = 0; q < jobs.size(); q++)
          if (job_op[q] == m) jm.push_back(jobs[q]);
        CNRGmatrix& M = mats[m];
        int nxt = M.cur ^ 1;
        DBuf& ob = c->buf(M.buf[nxt]);
        if ((rc = ob.reserve(sizeof(double) * (size_t)(icount[m] + 1)))) return rc;
        if ((rc = rotate_dev(c, jm, dZ, (double*)fb.p, (double*)ab.p, (double*)ob.p))) return rc;
        M.cur = nxt;
        M.MatBlockMap = nmap[m];
        M.MatBlockBegEnd = nbe[m];
        M.nel = icount[m];
      }
      HIPCHK(hipStreamSynchronize(c->stream));
      if ((rc = c->resolve())) return rc;
    }
    if ((rc = op_avgs(ops))) return rc;
    emit(cb, user, emin, ecut, nkept, thT, sus, ent, DN / bbar, ops, p.record_blocks ? &hrec : nullptr,
         p.record_blocks ? &zrec : nullptr);
    // chi for the next site: SetChainWilson (src/NRGchainClass.cpp:132-155)
    double d0 = 1.0 - std::pow(Lambda, -((double)N + 1.0));
    double d1 = std::sqrt(1.0 - std::pow(Lambda, -(2.0 * (double)N + 1.0)));
    double d2 = std::sqrt(1.0 - std::pow(Lambda, -(2.0 * (double)N + 3.0)));
    par_chi = d0 / (d1 * d2);
    (void)prevCut;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  auto t1 = std::chrono::steady_clock::now();
  if (timing_ms) *timing_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return 0;
}

}  // namespace nrg

extern "C" int nrg_run_onechqs(nrg_ctx* c, const nrg_onechqs_params* p, nrg_step_cb cb, void* user,
                               double* timing_ms) {
  if (!c || !p) return nrg::fail(NRG_ERR_ARG, "nrg_run_onechqs: NULL argument");
  if (p->nsitesmax < 0 || p->ncut < 1 || p->Lambda <= 1.0)
    return nrg::fail(NRG_ERR_ARG, "nrg_run_onechqs: bad parameters");
  HIPCHK(hipSetDevice(c->device));
  nrg::Driver d;
  d.c = c;
  d.p = *p;
  return d.run(cb, user, timing_ms);
}
