// Readers for ccq's integral inputs (host code, part of libccq_hip.so).
//
// ccq reads its integrals with Fortran list-directed READs line by line
// (src/integrals.f90:11-210); at 100 orbitals the two-electron file has
// 1e8 lines and that read takes minutes.  These readers keep ccq's file
// semantics exactly and parse in parallel:
//   * the file is memory-mapped and cut into chunks at line boundaries;
//   * every chunk is parsed by its own thread into (offset, value) records
//     (numbers by std::from_chars, correctly rounded; Fortran 'D' exponents
//     accepted; separators are blanks, tabs and commas as in list-directed
//     input);
//   * the first terminating line (ccq stops at "0 0 0 0 E_nuc" or at the
//     first line it cannot read) is found across chunks; records after it
//     are dropped;
//   * records are bucketed by the slice of the output array they land in
//     and each thread applies its slice's records in file order, so a
//     repeated index keeps the value of its last line, as ccq's sequential
//     loop does.
// Arrays are Fortran column-major with 1-based file indices, exactly as ccq
// stores sys%ints%e1int(orbs,orbs) and e2int(orbs,orbs,orbs,orbs).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ccq_hip.h"

namespace {

thread_local std::string g_err;

struct Mapped {
  const char* p = nullptr;
  size_t n = 0;
  int fd = -1;
  bool open(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    n = (size_t)st.st_size;
    if (n == 0) return true;
    void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) return false;
    p = (const char*)m;
    madvise(m, n, MADV_SEQUENTIAL);
    return true;
  }
  ~Mapped() {
    if (p) munmap((void*)p, n);
    if (fd >= 0) ::close(fd);
  }
};

inline bool is_sep(char c) { return c == ' ' || c == '\t' || c == ',' || c == '\r'; }

// next token on the current line [p, e): false at end of line
inline bool token(const char*& p, const char* e, const char*& t0, const char*& t1) {
  while (p < e && is_sep(*p)) ++p;
  if (p >= e) return false;
  t0 = p;
  while (p < e && !is_sep(*p)) ++p;
  t1 = p;
  return true;
}

inline bool to_int(const char* t0, const char* t1, int64_t& v) {
  if (t0 < t1 && *t0 == '+') ++t0;
  auto r = std::from_chars(t0, t1, v);
  return r.ec == std::errc() && r.ptr == t1;
}

inline bool to_real(const char* t0, const char* t1, double& v) {
  char buf[96];
  size_t len = (size_t)(t1 - t0);
  if (len == 0 || len >= sizeof(buf)) return false;
  size_t k = 0;
  for (const char* q = t0; q < t1; ++q) {
    char c = *q;
    if (c == 'd' || c == 'D') c = 'e';
    if (!(k == 0 && c == '+')) buf[k++] = c;
  }
  auto r = std::from_chars(buf, buf + k, v);
  if (r.ec != std::errc() || r.ptr != buf + k) {
    // a number without exponent digits or with an integer form: from_chars
    // takes both; anything else is not a number
    return false;
  }
  return true;
}

// line boundaries: [b, e) excluding the newline
inline const char* line_end(const char* p, const char* end) {
  const void* q = memchr(p, '\n', (size_t)(end - p));
  return q ? (const char*)q : end;
}

int nthreads_for(size_t bytes) {
  unsigned hw = std::thread::hardware_concurrency();
  int t = (int)std::min<unsigned>(hw ? hw : 1, 64);
  const size_t per = 8u << 20;      // at least 8 MiB per thread
  return std::max(1, std::min<int>(t, (int)(bytes / per) + 1));
}

// chunk starts at line boundaries
std::vector<size_t> chunks(const Mapped& f, int nt) {
  std::vector<size_t> s{0};
  for (int k = 1; k < nt; ++k) {
    size_t pos = f.n * (size_t)k / (size_t)nt;
    if (pos <= s.back()) continue;
    const void* q = memchr(f.p + pos, '\n', f.n - pos);
    if (!q) break;
    size_t b = (size_t)((const char*)q - f.p) + 1;
    if (b > s.back() && b < f.n) s.push_back(b);
  }
  s.push_back(f.n);
  return s;
}

struct Rec {
  int64_t off;
  double val;
};

// Apply per-chunk records (file order) to out[0..total) with last-wins
// semantics, in parallel over slices of the output.
void apply_ordered(std::vector<std::vector<Rec>>& per_chunk, double* out, int64_t total, int nt) {
  const int nc = (int)per_chunk.size();
  nt = std::max(1, nt);
  const int64_t slice = (total + nt - 1) / nt;
  // bucket[c][t]
  std::vector<std::vector<std::vector<Rec>>> bucket(nc, std::vector<std::vector<Rec>>(nt));
  {
    std::vector<std::thread> th;
    for (int c = 0; c < nc; ++c)
      th.emplace_back([&, c] {
        auto& b = bucket[c];
        for (const Rec& r : per_chunk[c]) b[(size_t)(r.off / slice)].push_back(r);
        std::vector<Rec>().swap(per_chunk[c]);
      });
    for (auto& t : th) t.join();
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int c = 0; c < nc; ++c)
        for (const Rec& r : bucket[c][t]) out[r.off] = r.val;
    });
  for (auto& x : th) x.join();
}

int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

}  // namespace

extern "C" {

const char* ccq_hip_read_last_error(void) { return g_err.c_str(); }

// count_file_lines (src/utils/utils.F90:622-649) + the orbital count of
// load_ints (src/integrals.f90:248-251): lines = n(n+1)/2.
int ccq_hip_onebody_orbitals(const char* path, int64_t* orbs) {
  Mapped f;
  if (!f.open(path)) return fail(std::string("cannot open ") + path);
  // every record counts, blank ones included (a list-directed READ with an
  // empty list consumes any record; only the end of file stops the count)
  int64_t lines = 0;
  const char* p = f.p;
  const char* end = f.p + f.n;
  while (p < end) {
    ++lines;
    p = line_end(p, end) + 1;
  }
  *orbs = (int64_t)((-1.0 + std::sqrt(1.0 + 8.0 * (double)lines)) / 2.0);
  return 0;
}

// load_e1int (src/integrals.f90:11-41): n(n+1)/2 lines "value index" in the
// order i = 1..n, j = 1..i; e1int(i,j) = e1int(j,i) = value.
int ccq_hip_read_onebody(const char* path, int64_t orbs, double* e1int) {
  Mapped f;
  if (!f.open(path)) return fail(std::string("cannot open ") + path);
  const char* p = f.p;
  const char* end = f.p + f.n;
  for (int64_t i = 0; i < orbs; ++i)
    for (int64_t j = 0; j <= i; ++j) {
      // list-directed read: skip blank lines
      const char *t0 = nullptr, *t1 = nullptr;
      for (;;) {
        if (p >= end) return fail("load_e1int: end of file before all integrals were read");
        const char* e = line_end(p, end);
        const char* q = p;
        p = e + 1;
        if (token(q, e, t0, t1)) break;
      }
      double v;
      if (!to_real(t0, t1, v)) return fail("load_e1int: unreadable value");
      e1int[i + orbs * j] = v;
      e1int[j + orbs * i] = v;
    }
  return 0;
}

// load_e2int (src/integrals.f90:43-91): lines "i a j b value" (chemists'
// (ia|jb)) stored verbatim as e2int(i,j,a,b); the line with i+a+j+b == 0
// carries the nuclear repulsion and ends the read, so does the first line
// that cannot be read.  Nothing is symmetrised.
int ccq_hip_read_twobody(const char* path, int64_t orbs, double* e2int, double* en_repul) {
  Mapped f;
  if (!f.open(path)) return fail(std::string("cannot open ") + path);
  const int nt = nthreads_for(f.n);
  auto cs = chunks(f, nt);
  const int nc = (int)cs.size() - 1;
  std::vector<std::vector<Rec>> recs(nc);
  // per chunk: index of the first terminating line (sentinel / unreadable /
  // out of range), and the sentinel's value
  std::vector<size_t> stop(nc, (size_t)-1);
  std::vector<int> stop_kind(nc, 0);      // 1 sentinel, 2 unreadable, 3 out of range
  std::vector<double> stop_val(nc, 0.0);
  const int64_t n = orbs, n2 = n * n, n3 = n2 * n;
  {
    std::vector<std::thread> th;
    for (int c = 0; c < nc; ++c)
      th.emplace_back([&, c] {
        const char* p = f.p + cs[c];
        const char* end = f.p + cs[c + 1];
        auto& out = recs[c];
        out.reserve((size_t)((end - p) / 40 + 16));
        while (p < end) {
          const char* e = line_end(p, end);
          const char* q = p;
          const char *t0, *t1;
          int64_t idx[4];
          double v = 0.0;
          bool ok = true;
          for (int k = 0; k < 4 && ok; ++k) ok = token(q, e, t0, t1) && to_int(t0, t1, idx[k]);
          ok = ok && token(q, e, t0, t1) && to_real(t0, t1, v);
          if (ok && idx[0] + idx[1] + idx[2] + idx[3] == 0) {
            stop[c] = out.size();
            stop_kind[c] = 1;
            stop_val[c] = v;
            break;
          }
          if (!ok) {
            stop[c] = out.size();
            stop_kind[c] = 2;
            break;
          }
          const int64_t i = idx[0], a = idx[1], j = idx[2], b = idx[3];
          if (i < 1 || a < 1 || j < 1 || b < 1 || i > n || a > n || j > n || b > n) {
            stop[c] = out.size();
            stop_kind[c] = 3;
            break;
          }
          out.push_back({(i - 1) + n * (j - 1) + n2 * (a - 1) + n3 * (b - 1), v});
          p = e + 1;
        }
      });
    for (auto& t : th) t.join();
  }
  // the first chunk that stopped ends the file for ccq
  int first = -1;
  for (int c = 0; c < nc; ++c)
    if (stop_kind[c]) {
      first = c;
      break;
    }
  if (first >= 0) {
    if (stop_kind[first] == 3) return fail("load_e2int: orbital index out of range");
    recs[first].resize(stop[first]);
    for (int c = first + 1; c < nc; ++c) std::vector<Rec>().swap(recs[c]);
    recs.resize(first + 1);
    if (stop_kind[first] == 1) *en_repul = stop_val[first];
  }
  apply_ordered(recs, e2int, n3 * n, nt);
  return 0;
}

// get_fcidump (src/parser.f90:258-311, parse_key_val src/utils/utils.F90:416-462):
// NORB and NELEC from the header lines before the first line holding "END".
int ccq_hip_fcidump_header(const char* path, int64_t* norb, int64_t* nelec) {
  Mapped f;
  if (!f.open(path)) return fail(std::string("cannot open ") + path);
  *norb = -1;
  *nelec = -1;
  const char* p = f.p;
  const char* end = f.p + f.n;
  auto value_of = [](const std::string& line, const char* key, int64_t& out) {
    size_t k = line.find(key);
    if (k == std::string::npos) return;
    std::string rest = line.substr(k);
    size_t comma = rest.find(',');
    size_t eq = rest.find('=');
    if (eq == std::string::npos || (comma != std::string::npos && eq > comma)) return;
    std::string val = rest.substr(eq + 1, comma == std::string::npos ? std::string::npos
                                                                     : comma - eq - 1);
    const char* q = val.c_str();
    const char* qe = q + val.size();
    const char *t0, *t1;
    int64_t v;
    if (token(q, qe, t0, t1) && to_int(t0, t1, v)) out = v;
  };
  while (p < end) {
    const char* e = line_end(p, end);
    std::string line(p, e);
    p = e + 1;
    if (line.find("END") != std::string::npos) break;
    value_of(line, "NORB", *norb);
    value_of(line, "NELEC", *nelec);
  }
  if (*norb <= 0 || *nelec < 0) return fail("FCIDUMP header without NORB / NELEC");
  return 0;
}

// load_fcidump (src/integrals.f90:93-210): after the header, lines
// "value i a j b".  Orbitals are renumbered in increasing order of the
// orbital energies given on lines "e i 0 0 0" (stable insertion rank,
// src/external/ranking.f90:84-140; energies not given count as 0);
// two-electron lines fill all eight permutations of e2int(i,j,a,b), one-
// electron lines e1int(i,a) = e1int(a,i), the all-zero line is E_nuc.
int ccq_hip_read_fcidump(const char* path, int64_t orbs, double* e1int, double* e2int,
                         double* en_repul) {
  Mapped f;
  if (!f.open(path)) return fail(std::string("cannot open ") + path);
  const char* p = f.p;
  const char* end = f.p + f.n;
  while (p < end) {      // skip the header (through the first line with END)
    const char* e = line_end(p, end);
    bool last = std::string(p, e).find("END") != std::string::npos;
    p = e + 1;
    if (last) break;
  }
  struct Line {
    double v;
    int64_t i, a, j, b;
  };
  std::vector<Line> lines;
  while (p < end) {
    const char* e = line_end(p, end);
    const char* q = p;
    p = e + 1;
    const char *t0, *t1;
    Line L;
    bool ok = token(q, e, t0, t1) && to_real(t0, t1, L.v);
    int64_t* ix[4] = {&L.i, &L.a, &L.j, &L.b};
    for (int k = 0; k < 4 && ok; ++k) ok = token(q, e, t0, t1) && to_int(t0, t1, *ix[k]);
    if (!ok) break;
    lines.push_back(L);
  }
  const int64_t n = orbs;
  std::vector<double> eig((size_t)n, 0.0);
  for (const Line& L : lines)
    if (L.i != 0 && L.a + L.j + L.b == 0) {
      if (L.i < 1 || L.i > n) return fail("load_fcidump: orbital index out of range");
      eig[(size_t)(L.i - 1)] = L.v;
    }
  std::vector<int64_t> rank((size_t)n), inv((size_t)n);
  for (int64_t i = 0; i < n; ++i) rank[(size_t)i] = i;
  for (int64_t i = 1; i < n; ++i) {      // stable insertion rank
    int64_t tmp = rank[(size_t)i], j = i - 1;
    while (j >= 0 && !(eig[(size_t)rank[(size_t)j]] - eig[(size_t)tmp] < 0.0)) {
      rank[(size_t)(j + 1)] = rank[(size_t)j];
      --j;
    }
    rank[(size_t)(j + 1)] = tmp;
  }
  for (int64_t k = 0; k < n; ++k) inv[(size_t)rank[(size_t)k]] = k;
  const int64_t n2 = n * n, n3 = n2 * n;
  auto E2 = [&](int64_t i, int64_t j, int64_t a, int64_t b) -> double& {
    return e2int[i + n * j + n2 * a + n3 * b];
  };
  for (const Line& L : lines) {
    if (L.i != 0 && L.a != 0 && L.j != 0 && L.b != 0) {
      if (std::max({L.i, L.a, L.j, L.b}) > n || std::min({L.i, L.a, L.j, L.b}) < 1)
        return fail("load_fcidump: orbital index out of range");
      const int64_t i = inv[L.i - 1], a = inv[L.a - 1], j = inv[L.j - 1], b = inv[L.b - 1];
      E2(i, j, a, b) = L.v;
      E2(j, i, b, a) = L.v;
      E2(a, b, i, j) = L.v;
      E2(b, a, j, i) = L.v;
      E2(a, j, i, b) = L.v;
      E2(b, i, j, a) = L.v;
      E2(i, b, a, j) = L.v;
      E2(j, a, b, i) = L.v;
    } else if (L.i != 0 && L.a != 0 && L.j + L.b == 0) {
      if (std::max(L.i, L.a) > n || std::min(L.i, L.a) < 1)
        return fail("load_fcidump: orbital index out of range");
      const int64_t i = inv[L.i - 1], a = inv[L.a - 1];
      e1int[i + n * a] = L.v;
      e1int[a + n * i] = L.v;
    } else if (L.i + L.a + L.j + L.b == 0) {
      *en_repul = L.v;
    }
  }
  return 0;
}

// read_sym (src/symmetry.f90:15-119): the first line is the point group
// (1 C1, 2 Cs, 3 C2v, 4 D2h), then one irrep number per orbital.  A missing
// file means C1 with every orbital in irrep 1.
int ccq_hip_read_sym(const char* path, int64_t orbs, int64_t* point_group, int64_t* irreps) {
  Mapped f;
  if (access(path, F_OK) != 0) {
    *point_group = 1;
    for (int64_t i = 0; i < orbs; ++i) irreps[i] = 1;
    return 0;
  }
  if (!f.open(path)) return fail(std::string("cannot open ") + path);
  const char* p = f.p;
  const char* end = f.p + f.n;
  auto next_int = [&](int64_t& v) -> bool {
    while (p < end) {
      const char* e = line_end(p, end);
      const char* q = p;
      p = e + 1;
      const char *t0, *t1;
      if (token(q, e, t0, t1)) return to_int(t0, t1, v);
    }
    return false;
  };
  int64_t pg;
  if (!next_int(pg)) return fail("read_sym: empty symmetry file");
  if (pg < 1 || pg > 4) return fail("read_sym: ERROR: Point-group symmetry not supported");
  *point_group = pg;
  for (int64_t i = 0; i < orbs; ++i) {
    int64_t g;
    if (!next_int(g)) return fail("read_sym: fewer irreps than orbitals");
    irreps[i] = g;
  }
  return 0;
}

}  // extern "C"
