// Fortran unformatted records for ccq's amplitude files (host code, part of
// libccq_hip.so; the entry points that use them are in engine.hip).
//
// ccq writes, every iteration, the DIIS vector to t_vec.bin and to the ring
// slot cc_iter_mat-<k>.bin (write_vecs, src/diis.F90:289-349), and keeps T4
// in <bin_file>_4{a,b,c,d,e}.bin (open_t4_files, src/utils/cc_utils.f90:
// 89-145).  All are Fortran unformatted sequential files holding one record
// of doubles.  ccq's own restart keyword does not read t_vec back
// (src/solve_cc.f90:56-60 only recomputes the energy of the freshly zeroed
// vector); here the files are read into the resident amplitudes.
//
// Record layout (gfortran's): a 4-byte signed length marker, the data, the
// same marker.  Records above 2^31 - 9 bytes are split into subrecords whose
// leading markers are negative on every subrecord but the last and whose
// trailing markers are negative on every subrecord but the first.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "restart_io.h"

namespace ccqio {

constexpr int64_t kMaxSub = 2147483639;   // largest subrecord (bytes)

bool write_record(const std::string& path, const double* x, int64_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const char* p = (const char*)x;
  int64_t left = n * (int64_t)sizeof(double);
  bool first = true, ok = true;
  do {
    const int64_t len = left > kMaxSub ? kMaxSub : left;
    const bool last = len == left;
    const int32_t head = (int32_t)(last ? len : -len);
    const int32_t tail = (int32_t)(first ? len : -len);
    ok = ok && std::fwrite(&head, 4, 1, f) == 1;
    ok = ok && (len == 0 || std::fwrite(p, 1, (size_t)len, f) == (size_t)len);
    ok = ok && std::fwrite(&tail, 4, 1, f) == 1;
    p += len;
    left -= len;
    first = false;
  } while (left > 0 && ok);
  return std::fclose(f) == 0 && ok;
}

// the first record of a file, expected to hold exactly n doubles
bool read_record(const std::string& path, double* x, int64_t n, std::string& why) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    why = "cannot open " + path;
    return false;
  }
  char* p = (char*)x;
  int64_t got = 0;
  const int64_t want = n * (int64_t)sizeof(double);
  for (;;) {
    int32_t head = 0, tail = 0;
    if (std::fread(&head, 4, 1, f) != 1) {
      why = path + ": truncated record marker";
      break;
    }
    const int64_t len = head < 0 ? -(int64_t)head : (int64_t)head;
    if (got + len > want) {
      why = path + ": record holds " + std::to_string((got + len) / 8) + "+ values, expected " +
            std::to_string(n);
      break;
    }
    if (len && std::fread(p + got, 1, (size_t)len, f) != (size_t)len) {
      why = path + ": truncated record";
      break;
    }
    got += len;
    if (std::fread(&tail, 4, 1, f) != 1 || (tail < 0 ? -(int64_t)tail : (int64_t)tail) != len) {
      why = path + ": record markers disagree";
      break;
    }
    if (head >= 0) {
      if (got != want) {
        why = path + ": record holds " + std::to_string(got / 8) + " values, expected " +
              std::to_string(n);
        break;
      }
      std::fclose(f);
      return true;
    }
  }
  std::fclose(f);
  return false;
}

}  // namespace ccqio
