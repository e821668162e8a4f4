// Fortran unformatted sequential records of doubles (ccq_amd/csrc/restart.cpp).
#pragma once
#include <cstdint>
#include <string>

namespace ccqio {
// one record of n doubles (gfortran layout, subrecords above 2^31 - 9 bytes)
bool write_record(const std::string& path, const double* x, int64_t n);
// the first record of a file, which must hold exactly n doubles; why = reason
bool read_record(const std::string& path, double* x, int64_t n, std::string& why);
}  // namespace ccqio
