// nrg_bin.cpp -- writer and reader of the reference's .bin hand-off files
// (CNRGarray::SaveBin / ReadBinInStream src/NRGarrayClass.cpp:778-945,
// CNRGbasisarray :1947-2075, CNRGmatrix src/NRGmatrixClass.cpp:957-1077):
// every vector is an int32 count followed by its elements, bools are one byte.
#include "nrg_bin.h"

#include <cstdio>
#include <cstring>

namespace nrg {
namespace {

static_assert(sizeof(int) == 4 && sizeof(double) == 8, "the .bin layout has 4-byte ints and 8-byte doubles");

struct Out {
  FILE* f = nullptr;
  bool ok = true;
  void raw(const void* p, size_t n) {
    if (ok && n && std::fwrite(p, 1, n, f) != n) ok = false;
  }
  void i32(int v) { raw(&v, 4); }
  void f64(double v) { raw(&v, 8); }
  void u8(unsigned char v) { raw(&v, 1); }
  void vi(const std::vector<int>& v) { i32((int)v.size()); raw(v.data(), 4 * v.size()); }
  void vd(const std::vector<double>& v) { i32((int)v.size()); raw(v.data(), 8 * v.size()); }
  // complex<double> vectors: count of complex numbers, then (re, im) pairs
  void vc(const std::vector<double>& v) { i32((int)(v.size() / 2)); raw(v.data(), 8 * (v.size() / 2 * 2)); }
  void vb(const std::vector<unsigned char>& v) { i32((int)v.size()); raw(v.data(), v.size()); }
};

struct In {
  FILE* f = nullptr;
  bool ok = true;
  int64_t left = 0;   // bytes not read yet: a count larger than what is left is corrupt
  void raw(void* p, size_t n) {
    if (!ok) return;
    if (n && std::fread(p, 1, n, f) != n) ok = false;
    left -= (int64_t)n;
  }
  int i32() { int v = 0; raw(&v, 4); return v; }
  double f64() { double v = 0; raw(&v, 8); return v; }
  unsigned char u8() { unsigned char v = 0; raw(&v, 1); return v; }
  size_t count(size_t elem) {
    const int n = i32();
    if (n < 0 || (int64_t)n * (int64_t)elem > left) ok = false;
    return ok ? (size_t)n : 0;
  }
  void vi(std::vector<int>& v) { v.resize(count(4)); raw(v.data(), 4 * v.size()); }
  void vd(std::vector<double>& v) { v.resize(count(8)); raw(v.data(), 8 * v.size()); }
  void vc(std::vector<double>& v) { v.resize(2 * count(16)); raw(v.data(), 8 * v.size()); }
  void vb(std::vector<unsigned char>& v) { v.resize(count(1)); raw(v.data(), v.size()); }
};

// CNRGarray::SaveBin field order
void put_array(Out& o, const BinArray& a) {
  o.i32(a.nshell);
  o.i32(a.nq);
  o.vd(a.qn);
  o.vd(a.en);
  o.vd(a.evec);
  o.vc(a.cevec);
  o.vi(a.bbe);
  o.vi(a.idegen);
  o.vb(a.kept);
  o.u8(a.totalS);
  o.vi(a.sqn);
}
void get_array(In& in, BinArray& a) {
  a.nshell = in.i32();
  a.nq = in.i32();
  in.vd(a.qn);
  in.vd(a.en);
  in.vd(a.evec);
  in.vc(a.cevec);
  in.vi(a.bbe);
  in.vi(a.idegen);
  in.vb(a.kept);
  a.totalS = in.u8();
  in.vi(a.sqn);
}

FILE* open(const std::string& path, const char* mode, std::string& err) {
  FILE* f = std::fopen(path.c_str(), mode);
  if (!f) err = "cannot open " + path;
  return f;
}
int close_out(Out& o, const std::string& path, std::string& err) {
  if (std::fclose(o.f) != 0) o.ok = false;
  if (!o.ok) err = "error writing " + path;
  return o.ok ? 0 : -1;
}
int close_in(In& in, const std::string& path, std::string& err) {
  // the reader stops at the last field, as ReadBinInStream does
  std::fclose(in.f);
  if (!in.ok) err = "truncated or corrupt file " + path;
  return in.ok ? 0 : -1;
}

}  // namespace

std::string bin_name(const char* dir, const char* kind, const char* id, int N) {
  std::string s = (dir && *dir) ? std::string(dir) + "/" : std::string();
  return s + kind + "_" + id + "_N" + std::to_string(N) + ".bin";
}

int bin_write(const std::string& path, const BinBasis& a, std::string& err) {
  Out o;
  if (!(o.f = open(path, "wb", err))) return -1;
  put_array(o, a);
  o.vi(a.itype);
  o.vi(a.stcf);
  o.vi(a.bbe_bc);
  o.vi(a.bbe_ev);
  o.i32(a.nq_stcf);
  o.vd(a.stcf_qn);
  return close_out(o, path, err);
}

int bin_write(const std::string& path, const BinMatrix& m, std::string& err) {
  Out o;
  if (!(o.f = open(path, "wb", err))) return -1;
  put_array(o, m);
  o.vd(m.matel);
  o.vc(m.matel_cplx);
  o.vi(m.map);
  o.vi(m.mbe);
  o.u8(m.need_old);
  o.u8(m.upper);
  o.u8(m.is_complex);
  o.f64(m.wigner_eckart_l);
  return close_out(o, path, err);
}

static int open_in(In& in, const std::string& path, std::string& err) {
  if (!(in.f = open(path, "rb", err))) return -1;
  std::fseek(in.f, 0, SEEK_END);
  in.left = (int64_t)std::ftell(in.f);
  std::fseek(in.f, 0, SEEK_SET);
  return 0;
}

int bin_read(const std::string& path, BinBasis& a, std::string& err) {
  In in;
  if (open_in(in, path, err)) return -1;
  get_array(in, a);
  in.vi(a.itype);
  in.vi(a.stcf);
  in.vi(a.bbe_bc);
  in.vi(a.bbe_ev);
  a.nq_stcf = in.i32();
  in.vd(a.stcf_qn);
  return close_in(in, path, err);
}

int bin_read(const std::string& path, BinMatrix& m, std::string& err) {
  In in;
  if (open_in(in, path, err)) return -1;
  get_array(in, m);
  in.vd(m.matel);
  in.vc(m.matel_cplx);
  in.vi(m.map);
  in.vi(m.mbe);
  m.need_old = in.u8();
  m.upper = in.u8();
  m.is_complex = in.u8();
  m.wigner_eckart_l = in.f64();
  return close_in(in, path, err);
}

}  // namespace nrg
