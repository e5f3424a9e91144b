// nrg_bin.h -- the reference's .bin hand-off files: CNRGarray, CNRGbasisarray
// and CNRGmatrix as their SaveBin writes them (int32 counts, doubles, one-byte
// bools, in field order, no padding; src/NRGarrayClass.cpp:778-945,
// :1947-2075, src/NRGmatrixClass.cpp:957-1077).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace nrg {

// CNRGarray's saved fields
struct BinArray {
  int nshell = 0, nq = 0;
  std::vector<double> qn, en, evec;
  std::vector<double> cevec;          // complex eigenvectors, (re, im) pairs
  std::vector<int> bbe, idegen;
  std::vector<unsigned char> kept;    // vector<bool>, one byte per flag
  unsigned char totalS = 0;
  std::vector<int> sqn;               // positions of total-S quantum numbers
};

// CNRGbasisarray = CNRGarray + these
struct BinBasis : BinArray {
  std::vector<int> itype, stcf, bbe_bc, bbe_ev;
  int nq_stcf = 0;
  std::vector<double> stcf_qn;
};

// CNRGmatrix = CNRGarray + these
struct BinMatrix : BinArray {
  std::vector<double> matel;
  std::vector<double> matel_cplx;     // (re, im) pairs
  std::vector<int> map, mbe;
  unsigned char need_old = 0, upper = 0, is_complex = 0;
  double wigner_eckart_l = 0.0;
};

// 0 on success; on failure a message in err
int bin_write(const std::string& path, const BinBasis& a, std::string& err);
int bin_write(const std::string& path, const BinMatrix& m, std::string& err);
int bin_read(const std::string& path, BinBasis& a, std::string& err);
int bin_read(const std::string& path, BinMatrix& m, std::string& err);

// <dir>/<kind>_<id>_N<N>.bin, the names of CNRGCodeHandler::SaveArrays
std::string bin_name(const char* dir, const char* kind, const char* id, int N);

}  // namespace nrg
