// nrg_trunc.h -- host truncation (SetDegen, Ecut) and the error string setter
#pragma once
#include <string>
#include <vector>

namespace nrg {
int fail(int code, const std::string& msg);
void set_degen(std::vector<double>& en, double eps);
double ecut_of(const std::vector<double>& en, int ncut);
}  // namespace nrg
