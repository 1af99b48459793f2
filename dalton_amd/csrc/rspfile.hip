// Dalton's coupled-cluster response-vector files (host code only): the
// amplitude ('R0') and multiplier ('L0') files that CC_WRRSP
// (cc/cc_wrrsp.F:20) writes with the records of CC_WVRSP (cc/cc_wrrsp.F:811)
// and CC_RDRSP (cc/cc_rdrsp.F:20) reads with CC_RVRSP (cc/cc_rdrsp.F:1463),
// under the names CC_RWPRE makes (cc/cc_rdrsp.F:354).  They carry restarts
// (.RESTAR reads 'R0' in CCSD_ENERGY, cc/ccsd_energy.F:558) and hand vectors
// between programs.
//
// One file is a Fortran unformatted sequential file of ten records for a
// closed-shell singlet vector of a model without R12 or CC3 parts:
//   1  NWRITE (default INTEGER), MODFIL (CHARACTER*10, blank padded)
//   2  orbital (CPHF) part, 2 NALLAI values
//   3  singles, NT1AM values
//   4  doubles, NT2AM values (packed ai <= bj)
//   5  (empty: triplet doubles)
//   6, 7  (empty: R12 doubles and their triplet part)
//   8  (empty: CC3 effective singles)
//   9, 10  (empty: CC3 effective doubles and their triplet part)
// where NWRITE counts the leading parts that hold data (3 for CCSD, MP2,
// CC2 and CCD; 2 for CCS).
//
// Record framing (the gfortran and LLVM flang runtimes, little endian): each
// record is [n][n bytes][n] with n a 4-byte length.  gfortran splits a record
// longer than 2147483639 bytes into subrecords: the leading length of every
// subrecord but the last, and the trailing length of every subrecord but the
// first, are stored negated.  The writer follows that rule and the reader
// accepts it (and plain records of any size).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

namespace ccgpu {

namespace {

constexpr long kMaxSub = 2147483639L;  // largest data length of one subrecord

// CCGPU_RSP_SUBRECORD (bytes, 8..kMaxSub) lowers the subrecord length so
// that the splitting can be exercised on small files.
long max_sub() {
  const char* e = std::getenv("CCGPU_RSP_SUBRECORD");
  const long v = e ? std::atol(e) : kMaxSub;
  return v >= 8 && v <= kMaxSub ? v : kMaxSub;
}

struct File {
  FILE* f = nullptr;
  std::string path;
  File(const char* p, const char* mode) : path(p) {
    f = std::fopen(p, mode);
    CCGPU_REQUIRE(f != nullptr, std::string("cannot open ") + p);
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void put(const void* b, size_t n) {
    CCGPU_REQUIRE(n == 0 || std::fwrite(b, 1, n, f) == n, "write error on " + path);
  }
  void get(void* b, size_t n) {
    CCGPU_REQUIRE(n == 0 || std::fread(b, 1, n, f) == n, "truncated file " + path);
  }
};

void put_marker(File& F, long v) {
  const int32_t m = (int32_t)v;
  F.put(&m, 4);
}

// One logical record made of the concatenation of parts.
void write_record(File& F, const std::vector<std::pair<const void*, size_t>>& parts) {
  size_t total = 0;
  for (const auto& p : parts) total += p.second;
  size_t part = 0, off = 0;  // position in the parts
  size_t left = total;
  const size_t sub = (size_t)max_sub();
  bool first = true;
  do {
    const size_t n = std::min<size_t>(left, sub);
    const bool last = n == left;
    put_marker(F, last ? (long)n : -(long)n);
    for (size_t k = n; k > 0;) {
      const size_t take = std::min(k, parts[part].second - off);
      F.put((const char*)parts[part].first + off, take);
      off += take;
      k -= take;
      if (off == parts[part].second) ++part, off = 0;
    }
    put_marker(F, first ? (long)n : -(long)n);
    left -= n;
    first = false;
  } while (left > 0);
}

// Reads one logical record; its bytes go to dst (up to cap bytes, the rest
// is skipped).  Returns the record length.
size_t read_record(File& F, void* dst, size_t cap) {
  size_t total = 0;
  bool more = true, first = true;
  while (more) {
    int32_t lead = 0, trail = 0;
    F.get(&lead, 4);
    const long n = lead < 0 ? -(long)lead : (long)lead;
    more = lead < 0;
    const size_t take = dst ? std::min<size_t>((size_t)n, cap > total ? cap - total : 0) : 0;
    if (take) F.get((char*)dst + total, take);
    if ((size_t)n > take) CCGPU_REQUIRE(std::fseek(F.f, (long)n - (long)take, SEEK_CUR) == 0, "seek error on " + F.path);
    F.get(&trail, 4);
    const long t = trail < 0 ? -(long)trail : (long)trail;
    CCGPU_REQUIRE(t == n && (trail < 0) == !first, "record markers do not match in " + F.path);
    total += (size_t)n;
    first = false;
  }
  return total;
}

}  // namespace

std::string rsp_filename(const char* list, int isym, int idx) {
  CCGPU_REQUIRE(list && std::strlen(list) >= 2 && std::strlen(list) <= 3, "rsp_filename: list label of 2 or 3 characters");
  CCGPU_REQUIRE(isym >= 1 && isym <= 8 && idx >= 0 && idx <= 9999, "rsp_filename: symmetry 1..8, index 0..9999");
  // CC_RWPRE: two-character lists (and the listed three-character ones) get
  // a '_' third character; then the symmetry digit; then the index as I3
  // after one blank, blanks becoming '_' (a four-digit index fills both)
  std::string l(list);
  if (l.size() == 2) l += "_";
  l += (char)('0' + isym);
  char num[8];
  if (idx <= 999)
    std::snprintf(num, sizeof num, " %3d", idx);
  else
    std::snprintf(num, sizeof num, "%4d", idx);
  std::string name = "CC" + l + num;
  if (idx <= 999)
    for (char& c : name)
      if (c == ' ') c = '_';
  return name;
}

void rsp_write(const char* path, const char* model, int intsize, long nallai, long nt1, long nt2, const double* vec0,
               const double* vec1, const double* vec2) {
  CCGPU_REQUIRE(intsize == 4 || intsize == 8, "rsp_write: integer size 4 or 8");
  CCGPU_REQUIRE(model && std::strlen(model) <= 10, "rsp_write: model label of at most 10 characters");
  CCGPU_REQUIRE(nallai >= 0 && nt1 >= 0 && nt2 >= 0 && (nt1 == 0 || vec1) && (nt2 == 0 || vec2),
                "rsp_write: bad sizes or null vector");
  char mod[10];
  std::memset(mod, ' ', 10);
  std::memcpy(mod, model, std::strlen(model));
  const int nwrite = nt2 > 0 ? 3 : 2;
  const int32_t n4 = nwrite;
  const int64_t n8 = nwrite;
  File F(path, "wb");
  write_record(F, {{intsize == 4 ? (const void*)&n4 : (const void*)&n8, (size_t)intsize}, {mod, 10}});
  std::vector<double> zeros;
  if (!vec0) zeros.assign((size_t)(2 * nallai), 0.0);
  write_record(F, {{vec0 ? vec0 : zeros.data(), (size_t)(2 * nallai) * 8}});
  write_record(F, {{vec1, (size_t)nt1 * 8}});
  write_record(F, {{vec2, (size_t)nt2 * 8}});
  for (int k = 0; k < 6; ++k) write_record(F, {});
}

int rsp_read(const char* path, long nallai, long nt1, long nt2, double* vec0, double* vec1, double* vec2,
             char* model) {
  File F(path, "rb");
  unsigned char hdr[18];
  const size_t hl = read_record(F, hdr, sizeof hdr);
  CCGPU_REQUIRE(hl == 14 || hl == 18, "rsp_read: header is not (INTEGER, CHARACTER*10) in " + F.path);
  const int isz = (int)hl - 10;
  long nwrite = 0;
  if (isz == 4) {
    int32_t v;
    std::memcpy(&v, hdr, 4);
    nwrite = v;
  } else {
    int64_t v;
    std::memcpy(&v, hdr, 8);
    nwrite = (long)v;
  }
  if (model) {
    std::memcpy(model, hdr + isz, 10);
    model[10] = '\0';
  }
  auto part = [&](double* dst, long n, const char* what) {
    const size_t len = read_record(F, dst, dst ? (size_t)n * 8 : 0);
    CCGPU_REQUIRE(!dst || len == (size_t)n * 8,
                  std::string("rsp_read: ") + what + " record length differs from the expected size in " + F.path);
  };
  part(vec0, 2 * nallai, "orbital");
  if (nwrite >= 2) part(vec1, nt1, "singles");
  if (nwrite >= 3) part(vec2, nt2, "doubles");
  // the remaining records are only checked for intact framing
  for (int c; (c = std::fgetc(F.f)) != EOF;) {
    std::ungetc(c, F.f);
    read_record(F, nullptr, 0);
  }
  return (int)nwrite;
}

// A file of one record of n doubles, as CCRHSN writes its RSPIM
// intermediates (WRITE(LU) (X(I), I = 1, N), e.g. CC_E1IM and CC_E2IM,
// cc/ccsd_rhs.F:1898-1915), with the framing of write_record.
void write_one_record(const char* path, const double* x, long n) {
  CCGPU_REQUIRE(path != nullptr && n >= 0 && (n == 0 || x != nullptr), "write_one_record: bad argument");
  File F(path, "wb");
  write_record(F, {{x, (size_t)n * sizeof(double)}});
}

}  // namespace ccgpu
