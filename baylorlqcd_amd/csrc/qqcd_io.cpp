// qqcd_io.cpp -- the reference's gauge-field and propagator files.
//
// Both are Fortran unformatted sequential files: every record is framed by
// its length in bytes as a little-endian 32-bit integer before and after the
// payload.
//
// Gauge field (GFwrite / GFread, qqcd-test/cfgsprops/glue/gaugetools.f90:
// 1259-1334): after fixsu3, the first 12 reals (rows 1 and 2) of every link,
// one 64-byte record per (real icri, site isite, block ibl) -- icri
// outermost, ibl innermost -- holding u(icri, isite, mu, ieo, ibl) for mu =
// 1..4, ieo = 1..2 (ieo fastest); then a 96-byte trailer: gaction(4) (int32),
// beta, alat(4), tad(4) (float64), icfgsave, myid (int32).  Reading restores
// rows 1 and 2 and rebuilds the links with fixsu3 (:1338-1432): row 1
// normalised, row 2 orthogonalised against it and normalised, row 3 =
// conj(row1 x row2).  The arithmetic here follows that routine operation by
// operation, so written files and read-back links are bit-identical to the
// reference's.
//
// Propagator (FPwrite / FPread, qqcd-test/cfgsprops/quark/quark.f90:
// 3379-3481): one 64-byte record per (site isite, block ibl, real icri) --
// isite outermost -- holding x(icri, isite, id, ieo, ibleo) for id = 1..4,
// ieo = 1..2, from xe for the blocks of global parity 1 and xo otherwise
// (ibleo = the block's index within its parity); then a 40-byte trailer:
// kappa, tmqcd, cSW (float64), icsrc, idsrc, icfg, myid (int32).
//
// Host-only code: the lattice extents are arguments, no device is touched.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qqcd_hip.h"
#include "qqcd_internal.h"

using namespace qqcd;

// fixsu3's sums and products are evaluated exactly as written (no fused
// multiply-adds), as the reference's compiled code evaluates them
#pragma STDC FP_CONTRACT OFF

#define API_BEGIN try {
#define API_END                     \
  }                                 \
  catch (const std::exception &e) { \
    set_error(e.what());            \
    return -1;                      \
  }                                 \
  return 0;

namespace {

struct Dims {
  long nvhalf, ntotal;
};

Dims dims_of(const int n[4]) {
  Geom g{};
  long V = 1;
  for (int m = 0; m < 4; m++) {
    if (n[m] < 2 || n[m] % 2) throw std::runtime_error("lattice extents must be even");
    g.n[m] = n[m];
    V *= n[m];
  }
  if (n[0] % 4) throw std::runtime_error("nx must be divisible by 4 (the reference's site index)");
  g.V = V;
  g.Vh = V / 2;
  return Dims{V / 32, ref_ntotal(g)};
}

// buffered writer / reader of framed records
class RecWriter {
 public:
  RecWriter(const char *path, bool must_be_new) {
    f_ = std::fopen(path, must_be_new ? "wbx" : "wb");
    if (!f_)
      throw std::runtime_error(std::string("cannot create ") + path + (must_be_new ? " (it must not exist)" : ""));
    buf_.reserve(1 << 22);
  }
  ~RecWriter() {
    if (f_) std::fclose(f_);
  }
  void record(const void *p, int32_t len) {
    put(&len, 4);
    put(p, len);
    put(&len, 4);
  }
  void close() {
    flush();
    if (std::fclose(f_) != 0) {
      f_ = nullptr;
      throw std::runtime_error("write error");
    }
    f_ = nullptr;
  }

 private:
  void put(const void *p, size_t n) {
    const char *c = (const char *)p;
    buf_.insert(buf_.end(), c, c + n);
    if (buf_.size() >= (1u << 22)) flush();
  }
  void flush() {
    if (!buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size())
      throw std::runtime_error("write error");
    buf_.clear();
  }
  std::FILE *f_ = nullptr;
  std::vector<char> buf_;
};

class RecReader {
 public:
  explicit RecReader(const char *path) {
    std::FILE *f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    data_.resize(n > 0 ? (size_t)n : 0);
    const size_t got = data_.empty() ? 0 : std::fread(data_.data(), 1, data_.size(), f);
    std::fclose(f);
    if (got != data_.size()) throw std::runtime_error(std::string("read error on ") + path);
  }
  // the payload of the next record, which must be len bytes long
  const char *record(int32_t len) {
    if (pos_ + 8 + (size_t)len > data_.size()) throw std::runtime_error("file ends inside a record");
    int32_t a, b;
    std::memcpy(&a, &data_[pos_], 4);
    std::memcpy(&b, &data_[pos_ + 4 + len], 4);
    if (a != len || b != len) throw std::runtime_error("unexpected record length (not a file of this lattice?)");
    const char *p = &data_[pos_ + 4];
    pos_ += 8 + (size_t)len;
    return p;
  }
  bool at_end() const { return pos_ == data_.size(); }

 private:
  std::vector<char> data_;
  size_t pos_ = 0;
};

inline long uidx(const Dims &d, int c, long i, int mu, int ieo, int ibl) {
  return c + 18 * (i + d.ntotal * (mu + 4 * (ieo + 2L * ibl)));
}
inline long sidx(const Dims &d, int c, long i, int id, int ieo, int ibleo) {
  return c + 6 * (i + d.ntotal * (id + 4 * (ieo + 2L * ibleo)));
}

// fixsu3 on every link of u (rows 1..nvhalf), the reference's operation order
void fixsu3(const Dims &d, double *u) {
  for (int ieo = 0; ieo < 2; ieo++)
    for (int mu = 0; mu < 4; mu++)
      for (long i = 0; i < d.nvhalf; i++)
        for (int ibl = 0; ibl < 16; ibl++) {
          double *a = u + uidx(d, 0, i, mu, ieo, ibl);  // a[0..17] = u(1..18)
          double c1r = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3] + a[4] * a[4] + a[5] * a[5]);
          for (int k = 0; k < 6; k++) a[k] = a[k] / c1r;
          c1r = a[6] * a[0] + a[7] * a[1] + a[8] * a[2] + a[9] * a[3] + a[10] * a[4] + a[11] * a[5];
          const double c1i = a[7] * a[0] - a[6] * a[1] + a[9] * a[2] - a[8] * a[3] + a[11] * a[4] - a[10] * a[5];
          a[6] = a[6] - c1r * a[0] + c1i * a[1];
          a[7] = a[7] - c1r * a[1] - c1i * a[0];
          a[8] = a[8] - c1r * a[2] + c1i * a[3];
          a[9] = a[9] - c1r * a[3] - c1i * a[2];
          a[10] = a[10] - c1r * a[4] + c1i * a[5];
          a[11] = a[11] - c1r * a[5] - c1i * a[4];
          c1r = std::sqrt(a[6] * a[6] + a[7] * a[7] + a[8] * a[8] + a[9] * a[9] + a[10] * a[10] + a[11] * a[11]);
          for (int k = 6; k < 12; k++) a[k] = a[k] / c1r;
          a[16] = a[0] * a[8] - a[1] * a[9] - a[2] * a[6] + a[3] * a[7];
          a[14] = a[4] * a[6] - a[5] * a[7] - a[0] * a[10] + a[1] * a[11];
          a[12] = a[2] * a[10] - a[3] * a[11] - a[4] * a[8] + a[5] * a[9];
          a[17] = -a[0] * a[9] - a[1] * a[8] + a[2] * a[7] + a[3] * a[6];
          a[15] = -a[4] * a[7] - a[5] * a[6] + a[0] * a[11] + a[1] * a[10];
          a[13] = -a[2] * a[11] - a[3] * a[10] + a[4] * a[9] + a[5] * a[8];
        }
}

template <typename T> void put_bytes(char *&p, const T &v) {
  std::memcpy(p, &v, sizeof(T));
  p += sizeof(T);
}
template <typename T> T get_bytes(const char *&p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  p += sizeof(T);
  return v;
}

}  // namespace

namespace qqcd {
// the library's device copy of the links no longer matches the host array u
// (defined with the context, qqcd_capi.cpp)
void forget_host_gauge(const void *u);
}  // namespace qqcd

extern "C" int qqcd_gfwrite(const int n[4], double *u, const char *filename, const int gaction[4], double beta,
                            const double alat[4], const double tad[4], int icfgsave, int myid) {
  API_BEGIN
  const Dims d = dims_of(n);
  qqcd::forget_host_gauge(u);  // fixsu3 re-unitarises u in place
  fixsu3(d, u);
  RecWriter w(filename, true);
  double rec[8];
  for (int c = 0; c < 12; c++)
    for (long i = 0; i < d.nvhalf; i++)
      for (int ibl = 0; ibl < 16; ibl++) {
        for (int mu = 0; mu < 4; mu++) {
          rec[2 * mu] = u[uidx(d, c, i, mu, 0, ibl)];
          rec[2 * mu + 1] = u[uidx(d, c, i, mu, 1, ibl)];
        }
        w.record(rec, sizeof(rec));
      }
  char tr[96], *p = tr;
  for (int k = 0; k < 4; k++) put_bytes<int32_t>(p, gaction[k]);
  put_bytes<double>(p, beta);
  for (int k = 0; k < 4; k++) put_bytes<double>(p, alat[k]);
  for (int k = 0; k < 4; k++) put_bytes<double>(p, tad[k]);
  put_bytes<int32_t>(p, icfgsave);
  put_bytes<int32_t>(p, myid);
  w.record(tr, sizeof(tr));
  w.close();
  API_END
}

extern "C" int qqcd_gfread(const int n[4], double *u, const char *filename, int gaction[4], double *beta,
                           double alat[4], double tad[4], int *icfgsave, int *myidin) {
  API_BEGIN
  const Dims d = dims_of(n);
  RecReader r(filename);
  qqcd::forget_host_gauge(u);  // u is overwritten in place
  for (int c = 0; c < 12; c++)
    for (long i = 0; i < d.nvhalf; i++)
      for (int ibl = 0; ibl < 16; ibl++) {
        const char *p = r.record(64);
        for (int mu = 0; mu < 4; mu++) {
          u[uidx(d, c, i, mu, 0, ibl)] = get_bytes<double>(p);
          u[uidx(d, c, i, mu, 1, ibl)] = get_bytes<double>(p);
        }
      }
  const char *p = r.record(96);
  for (int k = 0; k < 4; k++) gaction[k] = get_bytes<int32_t>(p);
  *beta = get_bytes<double>(p);
  for (int k = 0; k < 4; k++) alat[k] = get_bytes<double>(p);
  for (int k = 0; k < 4; k++) tad[k] = get_bytes<double>(p);
  *icfgsave = get_bytes<int32_t>(p);
  *myidin = get_bytes<int32_t>(p);
  if (!r.at_end()) throw std::runtime_error("gauge file longer than this lattice's");
  fixsu3(d, u);
  API_END
}

extern "C" int qqcd_fpwrite(const int n[4], const double *xe, const double *xo, const char *filename, double kappa,
                            double tmqcd, double cSW, int icsrc, int idsrc, int icfg, int myid) {
  API_BEGIN
  const Dims d = dims_of(n);
  RecWriter w(filename, false);
  double rec[8];
  for (long i = 0; i < d.nvhalf; i++)
    for (int ibl = 0; ibl < 16; ibl++) {
      const bool even = __builtin_popcount(ibl) % 2 == 0;  // global parity of the block (gblclr = 1)
      const double *x = even ? xe : xo;
      const int ibleo = ibl / 2;
      for (int c = 0; c < 6; c++) {
        for (int id = 0; id < 4; id++) {
          rec[2 * id] = x[sidx(d, c, i, id, 0, ibleo)];
          rec[2 * id + 1] = x[sidx(d, c, i, id, 1, ibleo)];
        }
        w.record(rec, sizeof(rec));
      }
    }
  char tr[40], *p = tr;
  put_bytes<double>(p, kappa);
  put_bytes<double>(p, tmqcd);
  put_bytes<double>(p, cSW);
  put_bytes<int32_t>(p, icsrc);
  put_bytes<int32_t>(p, idsrc);
  put_bytes<int32_t>(p, icfg);
  put_bytes<int32_t>(p, myid);
  w.record(tr, sizeof(tr));
  w.close();
  API_END
}

extern "C" int qqcd_fpread(const int n[4], double *xe, double *xo, const char *filename, double *kappa, double *tmqcd,
                           double *cSW, int *icsrc, int *idsrc, int *icfgin, int *myidin) {
  API_BEGIN
  const Dims d = dims_of(n);
  RecReader r(filename);
  for (long i = 0; i < d.nvhalf; i++)
    for (int ibl = 0; ibl < 16; ibl++) {
      const bool even = __builtin_popcount(ibl) % 2 == 0;
      double *x = even ? xe : xo;
      const int ibleo = ibl / 2;
      for (int c = 0; c < 6; c++) {
        const char *p = r.record(64);
        for (int id = 0; id < 4; id++) {
          x[sidx(d, c, i, id, 0, ibleo)] = get_bytes<double>(p);
          x[sidx(d, c, i, id, 1, ibleo)] = get_bytes<double>(p);
        }
      }
    }
  const char *p = r.record(40);
  *kappa = get_bytes<double>(p);
  *tmqcd = get_bytes<double>(p);
  *cSW = get_bytes<double>(p);
  *icsrc = get_bytes<int32_t>(p);
  *idsrc = get_bytes<int32_t>(p);
  *icfgin = get_bytes<int32_t>(p);
  *myidin = get_bytes<int32_t>(p);
  if (!r.at_end()) throw std::runtime_error("propagator file longer than this lattice's");
  API_END
}
