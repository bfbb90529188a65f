// io.hip -- restart primitives and raw (Tecplot) export/import of fields.
// Host code only: fields are copied from HBM once and formatted on the CPU.
//
// Text format.  The reference writes with Fortran list-directed output
// (`write(un,*)`).  Files written here keep that layout and read back, in the
// reference or here, to the same doubles bit for bit; the readers accept any
// list-directed text (flang's, gfortran's, hand-edited files):
//   * a real is printed with the fewest significant digits that read back to
//     the same double, correctly rounded among those (flang's runtime at
//     times prints one digit more or a neighbouring last digit, so a file can
//     differ from the reference's in such digits, never in value); with
//     |x| rounded to one significant digit in [0.1, 1e15), i.e.
//     0.095 <= |x| < 9.5e14, it is written in F form (".1", ".095", "1.",
//     "123.45": no leading zero, a trailing point for whole numbers) and in
//     E form otherwise ("1.E-02", "-2.5E-07",
//     "1.7976931348623157E+308": at least two exponent digits); zero is "0."
//     or "-0.", the non-finite values "NaN", "Inf", "-Inf";
//   * an integer is its decimal digits, a character item its text;
//   * every item is preceded by one blank; an item that would take a record
//     past column 79 starts a new record; every statement ends its record.
#include "common.hpp"

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace moons {
namespace io {

// Significant digits (no trailing zeros) and decimal exponent e with
// ax = 0.digits x 10^e, the shortest string that reads back to ax; ax > 0
// and finite.  For a normal double, up to 15 digits the correctly rounded
// 15-digit form stripped of trailing zeros is the shortest one (any
// <= 15-digit string that reads back lies within one ulp of ax, closer than
// any other 15-digit string);
// at 16 digits the correctly rounded form, else its neighbour across ax
// (the rounding interval is lopsided at powers of two); 17 digits always
// read back.
static int shortest(double ax, char *dig, int *e10) {
  char buf[48];
  // subnormals have fewer significant bits: search every length
  for (int p = ax < DBL_MIN ? 1 : 15; p <= 17; ++p) {
    snprintf(buf, sizeof buf, "%.*e", p - 1, ax);
    bool ok = strtod(buf, nullptr) == ax;
    if (!ok && p == 16) {
      // neighbour of the 16-digit form on the other side of ax
      double d = strtod(buf, nullptr);
      char *ep = strchr(buf, 'e');
      int ex = atoi(ep + 1);
      char m[20];
      int nm = 0;
      for (char *c = buf; c < ep; ++c)
        if (*c >= '0' && *c <= '9') m[nm++] = *c;
      // m holds 16 digits; step the last digit up or down with carries
      int step = d < ax ? 1 : -1;
      int i = nm - 1;
      while (i >= 0) {
        int v = m[i] - '0' + step;
        if (v >= 0 && v <= 9) {
          m[i] = char('0' + v);
          break;
        }
        m[i] = step > 0 ? '0' : '9';
        --i;
      }
      char alt[48] = {0};
      if (i >= 0 && m[0] != '0') {
        int n = snprintf(alt, sizeof alt, "%c.", m[0]);
        memcpy(alt + n, m + 1, nm - 1);
        snprintf(alt + n + nm - 1, sizeof alt - n - nm + 1, "e%d", ex);
      } else if (step < 0) { // 1.000...e(ex) stepped down: sixteen nines one decade lower
        snprintf(alt, sizeof alt, "9.999999999999999e%d", ex - 1);
      }
      if (alt[0]) {
        if (strtod(alt, nullptr) == ax) {
          snprintf(buf, sizeof buf, "%s", alt);
          ok = true;
        }
      }
    }
    if (!ok) continue;
    char *ep = strchr(buf, 'e');
    int n = 0;
    for (char *c = buf; c < ep; ++c)
      if (*c >= '0' && *c <= '9') dig[n++] = *c;
    while (n > 1 && dig[n - 1] == '0') --n;
    dig[n] = 0;
    *e10 = atoi(ep + 1) + 1;
    return n;
  }
  return 0; // not reached for finite ax
}

// x as a list-directed item (without the leading blank); returns its length.
static int format_real(double x, char *out) {
  char *o = out;
  if (std::isnan(x)) return (int)(stpcpy(out, "NaN") - out);
  if (std::signbit(x)) *o++ = '-';
  double ax = std::fabs(x);
  if (std::isinf(ax)) return (int)(stpcpy(o, "Inf") - out);
  if (ax == 0.0) return (int)(stpcpy(o, "0.") - out);
  char dig[24];
  int e = 0;
  int n = shortest(ax, dig, &e);
  // F form when |x| rounded to one significant digit lies in [0.1, 1e15)
  if (ax >= 0.095 && ax < 9.5e14) {
    if (e <= 0) {
      *o++ = '.';
      for (int i = e; i < 0; ++i) *o++ = '0';
      memcpy(o, dig, n);
      o += n;
    } else if (e >= n) {
      memcpy(o, dig, n);
      o += n;
      for (int i = n; i < e; ++i) *o++ = '0';
      *o++ = '.';
    } else {
      memcpy(o, dig, e);
      o += e;
      *o++ = '.';
      memcpy(o, dig + e, n - e);
      o += n - e;
    }
  } else {
    *o++ = dig[0];
    *o++ = '.';
    memcpy(o, dig + 1, n - 1);
    o += n - 1;
    int ex = e - 1;
    o += sprintf(o, "E%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
  }
  *o = 0;
  return (int)(o - out);
}

// Records of list-directed output statements.
struct ListOut {
  std::string s;
  int col = 0;
  void item(const char *t, int n) {
    if (col > 0 && col + 1 + n > 79) {
      s += '\n';
      col = 0;
    }
    s += ' ';
    s.append(t, n);
    col += 1 + n;
  }
  void item(const std::string &t) { item(t.data(), (int)t.size()); }
  void item(long long v) { item(std::to_string(v)); }
  void real(double x) {
    char b[40];
    int n = format_real(x, b);
    item(b, n);
  }
  void end() {
    s += '\n';
    col = 0;
  }
};

static unsigned n_workers(size_t work) {
  unsigned hw = std::thread::hardware_concurrency();
  if (hw == 0) hw = 1;
  if (hw > 16) hw = 16;
  size_t w = work / 65536 + 1;
  return (unsigned)std::min<size_t>(hw, w);
}

// Runs body(t, lo, hi) over [0, n) split into nt contiguous chunks on threads.
template <class F>
static void parallel_chunks(size_t n, unsigned nt, F body) {
  if (nt <= 1) {
    body(0u, (size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t) {
    size_t lo = n * t / nt, hi = n * (t + 1) / nt;
    th.emplace_back([=] { body(t, lo, hi); });
  }
  for (auto &x : th) x.join();
}

// One statement writing every value of f: items formatted in parallel,
// records assembled in order.
static void write_values(ListOut &w, const double *f, size_t n) {
  const int W = 32;
  std::vector<char> buf(n * W);
  std::vector<unsigned char> len(n);
  parallel_chunks(n, n_workers(n), [&](unsigned, size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) len[i] = (unsigned char)format_real(f[i], &buf[i * W]);
  });
  w.s.reserve(w.s.size() + n * 20);
  for (size_t i = 0; i < n; ++i) w.item(&buf[i * W], len[i]);
  w.end();
}

static bool write_file(const std::string &path, const std::string &s) {
  FILE *fp = fopen(path.c_str(), "wb");
  if (!fp) return false;
  bool ok = fwrite(s.data(), 1, s.size(), fp) == s.size();
  ok = (fclose(fp) == 0) && ok;
  return ok;
}

static bool read_file(const std::string &path, std::string &s) {
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) return false;
  fseek(fp, 0, SEEK_END);
  long n = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  s.resize(n > 0 ? (size_t)n : 0);
  bool ok = n <= 0 || fread(&s[0], 1, (size_t)n, fp) == (size_t)n;
  fclose(fp);
  return ok;
}

// List-directed input: each read statement starts at the next record, takes
// its values from as many records as it needs (blank, comma or slash
// separated; r*c repeats; D or E exponents) and leaves the rest of its last
// record unread.
struct ListIn {
  const std::string &s;
  size_t pos = 0;    // start of the next record
  size_t cur = 0;    // scan position within the current statement
  size_t rec_end = 0;
  std::string err;
  explicit ListIn(const std::string &t) : s(t) {}
  bool next_record() {
    if (pos >= s.size()) return false;
    cur = pos;
    size_t nl = s.find('\n', pos);
    rec_end = nl == std::string::npos ? s.size() : nl;
    pos = nl == std::string::npos ? s.size() : nl + 1;
    return true;
  }
  void skip_record() { next_record(); }
  // Reads n reals (or integers, as doubles) for one statement.
  bool values(double *out, size_t n) {
    if (!next_record()) return fail_("end of file");
    size_t got = 0;
    while (got < n) {
      while (cur < rec_end && (s[cur] == ' ' || s[cur] == ',' || s[cur] == '\t' || s[cur] == '\r')) ++cur;
      if (cur >= rec_end) {
        if (!next_record()) return fail_("end of file");
        continue;
      }
      if (s[cur] == '/') return fail_("slash before all values were read");
      size_t t0 = cur;
      while (cur < rec_end && s[cur] != ' ' && s[cur] != ',' && s[cur] != '\t' && s[cur] != '\r' && s[cur] != '/') ++cur;
      std::string tok = s.substr(t0, cur - t0);
      long long rep = 1;
      size_t star = tok.find('*');
      if (star != std::string::npos) {
        rep = atoll(tok.substr(0, star).c_str());
        tok = tok.substr(star + 1);
        if (rep <= 0) return fail_("bad repeat count");
      }
      for (char &c : tok)
        if (c == 'D' || c == 'd' || c == 'Q' || c == 'q') c = 'E';
      char *endp = nullptr;
      errno = 0;
      double v = strtod(tok.c_str(), &endp);
      if (tok.empty() || endp != tok.c_str() + tok.size()) return fail_("not a number: '" + tok + "'");
      for (long long r = 0; r < rep && got < n; ++r) out[got++] = v;
    }
    return true;
  }
  bool fail_(const std::string &m) {
    err = m;
    return false;
  }
};

// Data-location letter and face/edge suffix of construct_suffix_SF
// (code/handwritten/IO/construct_suffix.f90:18-27, get_char_DL
// code/handwritten/fields/data_location_extend.f90:492-502).
static std::string suffix(const std::string &name, int dl) {
  switch (dl) {
  case MOONS_DL_CC: return name + "c";
  case MOONS_DL_NODE: return name + "n";
  case MOONS_DL_FACE_X: return name + "f_x";
  case MOONS_DL_FACE_Y: return name + "f_y";
  case MOONS_DL_FACE_Z: return name + "f_z";
  case MOONS_DL_EDGE_X: return name + "e_x";
  case MOONS_DL_EDGE_Y: return name + "e_y";
  case MOONS_DL_EDGE_Z: return name + "e_z";
  }
  return name;
}

// cp2str (code/pre_generated/datatype_conversion.f90:77-84): F40.20, blanks trimmed.
static std::string time_str(double t) {
  char b[80];
  snprintf(b, sizeof b, "%.20f", t);
  return b;
}

// Coordinates of the data location along d: nodes where the location is
// node-centred along d, cell centres otherwise (get_coordinates_h_g,
// code/handwritten/grid/grid_extend.f90:335-347; hc of init_c,
// code/handwritten/grid/coordinates_extend.f90:78-173).
static std::vector<double> coords(const double *hn, int sn, bool node) {
  if (node) return std::vector<double>(hn, hn + sn);
  std::vector<double> hc(sn - 1);
  for (int i = 0; i < sn - 1; ++i) hc[i] = (hn[i + 1] + hn[i]) / 2.0;
  return hc;
}

static int check_shape(const int s[3]) {
  if (!s || s[0] <= 0 || s[1] <= 0 || s[2] <= 0) return fail(MOONS_ERR_ARG, "shape must be positive");
  return MOONS_OK;
}

// The raw Tecplot file of export_3D_1C_steady/_unsteady
// (code/handwritten/IO/IO_export.f90:265-292) on a one-block mesh: header
// (exp_Header_3D_1C, code/handwritten/IO/exp_Tecplot_Header.f90:106-112),
// zone (exp_Zone_3I, code/handwritten/IO/exp_Tecplot_Zone.f90:20-33) and one
// statement "x y z u" per point, i fastest, pad cells left out at each end
// (exp_3D_1C_GF, code/handwritten/grid_field/GF_export_import/GF_export.f90:70-89).
static int write_raw(const double *hn[3], const int sn[3], int dl, const double *f, const std::string &dir,
                     const std::string &name, int pad, const double *t) {
  std::vector<double> h[3];
  int s[3];
  for (int d = 0; d < 3; ++d) {
    if (!hn[d] || sn[d] < 3) return fail(MOONS_ERR_ARG, "export_raw: a direction has fewer than 3 nodes");
    h[d] = coords(hn[d], sn[d], (dl >> d) & 1);
    s[d] = (int)h[d].size();
    if (s[d] - 2 * pad <= 0) return fail(MOONS_ERR_ARG, "export_raw: pad leaves no points");
  }
  if (pad < 0) return fail(MOONS_ERR_ARG, "export_raw: pad must be >= 0");
  std::string sfx = suffix(name, dl);
  std::string path = dir + sfx + (t ? "_t=" + time_str(*t) : std::string()) + ".dat";
  ListOut w;
  w.item("TITLE = \"3D Scalar Field\"");
  w.end();
  w.item("VARIABLES = \"x\",\"y\",\"z\",\"" + sfx + "\"");
  w.end();
  int e[3] = {s[0] - 2 * pad, s[1] - 2 * pad, s[2] - 2 * pad};
  w.item("ZONE , T =\"1\", I =" + std::to_string(e[0]) + ", J =" + std::to_string(e[1]) + ", K =" +
         std::to_string(e[2]) + " DATAPACKING = POINT");
  w.end();
  // one chunk of k-planes per thread, concatenated in order
  size_t npts = (size_t)e[0] * e[1] * e[2];
  unsigned nt = std::min(n_workers(npts), (unsigned)e[2]);
  std::vector<std::string> part(nt);
  parallel_chunks((size_t)e[2], nt, [&](unsigned idx, size_t lo, size_t hi) {
    ListOut p;
    p.s.reserve((hi - lo) * e[0] * e[1] * 80);
    for (size_t kk = lo; kk < hi; ++kk) {
      int k = (int)kk + pad;
      for (int j = pad; j < s[1] - pad; ++j)
        for (int i = pad; i < s[0] - pad; ++i) {
          p.real(h[0][i]);
          p.real(h[1][j]);
          p.real(h[2][k]);
          p.real(f[i + (size_t)s[0] * (j + (size_t)s[1] * k)]);
          p.end();
        }
    }
    part[idx] = std::move(p.s);
  });
  size_t tot = w.s.size();
  for (auto &p : part) tot += p.size();
  w.s.reserve(tot);
  for (auto &p : part) w.s += p;
  if (!write_file(path, w.s)) return fail(MOONS_ERR_IO, "export_raw: cannot write " + path);
  return MOONS_OK;
}

// import_3D_1C (code/handwritten/IO/IO_import.f90:20-33, imp_3D_1C
// code/handwritten/IO/base_import.f90:142-150, imp_3D_1C_GF
// code/handwritten/grid_field/GF_export_import/GF_import.f90:69-89): two
// header records and the zone record skipped, then one "x y z u" statement
// per point of the padded range; the coordinates are read and dropped, the
// pad cells of f keep their values.
static int read_raw(const int s[3], int dl, double *f, const std::string &dir, const std::string &name, int pad) {
  if (pad < 0 || s[0] - 2 * pad <= 0 || s[1] - 2 * pad <= 0 || s[2] - 2 * pad <= 0)
    return fail(MOONS_ERR_ARG, "import_raw: pad leaves no points");
  std::string path = dir + suffix(name, dl) + ".dat";
  std::string txt;
  if (!read_file(path, txt)) return fail(MOONS_ERR_IO, "import_raw: cannot read " + path);
  ListIn r(txt);
  r.skip_record();
  r.skip_record();
  r.skip_record();
  double v[4];
  for (int k = pad; k < s[2] - pad; ++k)
    for (int j = pad; j < s[1] - pad; ++j)
      for (int i = pad; i < s[0] - pad; ++i) {
        if (!r.values(v, 4)) return fail(MOONS_ERR_IO, "import_raw: " + path + ": " + r.err);
        f[i + (size_t)s[0] * (j + (size_t)s[1] * k)] = v[3];
      }
  return MOONS_OK;
}

} // namespace io
} // namespace moons

using namespace moons;

// ------------------------------------------------------ host-array entry points
extern "C" int moons_gf_write(const char *path, const double *f, const int s[3]) {
  if (!path || !f) return fail(MOONS_ERR_ARG, "moons_gf_write: null argument");
  if (int rc = io::check_shape(s)) return rc;
  io::ListOut w;
  for (int d = 0; d < 3; ++d) w.item((long long)s[d]);
  w.end();
  w.item("f     = ");
  w.end();
  io::write_values(w, f, (size_t)s[0] * s[1] * s[2]);
  w.item("s     = ");
  w.end();
  for (int d = 0; d < 3; ++d) w.item((long long)s[d]);
  w.end();
  w.item("s_1D  = ");
  w.end();
  w.item((long long)s[0] * s[1] * s[2]);
  w.end();
  if (!io::write_file(path, w.s)) return fail(MOONS_ERR_IO, std::string("moons_gf_write: cannot write ") + path);
  return MOONS_OK;
}

extern "C" int moons_gf_read(const char *path, double *f, size_t n, int s_out[3]) {
  if (!path || !s_out) return fail(MOONS_ERR_ARG, "moons_gf_read: null argument");
  std::string txt;
  if (!io::read_file(path, txt)) return fail(MOONS_ERR_IO, std::string("moons_gf_read: cannot read ") + path);
  io::ListIn r(txt);
  double sh[3];
  if (!r.values(sh, 3)) return fail(MOONS_ERR_IO, "moons_gf_read: shape: " + r.err);
  for (int d = 0; d < 3; ++d) s_out[d] = (int)sh[d];
  if (!(sh[0] > 0 && sh[1] > 0 && sh[2] > 0)) return MOONS_OK; // unallocated array on file
  size_t cnt = (size_t)sh[0] * (size_t)sh[1] * (size_t)sh[2];
  if (!f) return MOONS_OK; // shape query
  if (cnt != n) return fail(MOONS_ERR_ARG, "moons_gf_read: array size on file differs from n");
  r.skip_record();
  if (!r.values(f, cnt)) return fail(MOONS_ERR_IO, "moons_gf_read: values: " + r.err);
  return MOONS_OK;
}

extern "C" int moons_raw_write(const double *hn_x, int sn_x, const double *hn_y, int sn_y, const double *hn_z,
                               int sn_z, int dl, const double *f, const char *dir, const char *name, int pad,
                               const double *t) {
  if (!f || !dir || !name) return fail(MOONS_ERR_ARG, "moons_raw_write: null argument");
  if (dl < 0 || dl > 7) return fail(MOONS_ERR_ARG, "moons_raw_write: bad data location");
  const double *hn[3] = {hn_x, hn_y, hn_z};
  int sn[3] = {sn_x, sn_y, sn_z};
  return io::write_raw(hn, sn, dl, f, dir, name, pad, t);
}

extern "C" int moons_raw_read(const int s[3], int dl, double *f, const char *dir, const char *name, int pad) {
  if (!f || !dir || !name) return fail(MOONS_ERR_ARG, "moons_raw_read: null argument");
  if (int rc = io::check_shape(s)) return rc;
  return io::read_raw(s, dl, f, dir, name, pad);
}

// ---------------------------------------------------------- field entry points
extern "C" int moons_field_export(moons_field_t f, const char *path) {
  if (!f || !path) return fail(MOONS_ERR_ARG, "moons_field_export: null argument");
  std::vector<double> h(f->n);
  if (int rc = moons_field_download(f, h.data(), f->n)) return rc;
  return moons_gf_write(path, h.data(), f->s);
}

extern "C" int moons_field_import(moons_field_t f, const char *path) {
  if (!f || !path) return fail(MOONS_ERR_ARG, "moons_field_import: null argument");
  int s[3];
  std::vector<double> h(f->n);
  if (int rc = moons_gf_read(path, nullptr, 0, s)) return rc;
  if (s[0] != f->s[0] || s[1] != f->s[1] || s[2] != f->s[2])
    return fail(MOONS_ERR_ARG, std::string("moons_field_import: shape on file differs from the field: ") + path);
  if (int rc = moons_gf_read(path, h.data(), f->n, s)) return rc;
  return moons_field_upload(f, h.data(), f->n);
}

extern "C" int moons_export_raw(moons_field_t f, const char *dir, const char *name, int pad, const double *t) {
  if (!f || !dir || !name) return fail(MOONS_ERR_ARG, "moons_export_raw: null argument");
  std::vector<double> h(f->n);
  if (int rc = moons_field_download(f, h.data(), f->n)) return rc;
  const HostCoords *c = f->m->hc;
  return moons_raw_write(c[0].hn.data(), c[0].sn, c[1].hn.data(), c[1].sn, c[2].hn.data(), c[2].sn, f->dl, h.data(),
                         dir, name, pad, t);
}

extern "C" int moons_import_raw(moons_field_t f, const char *dir, const char *name, int pad) {
  if (!f || !dir || !name) return fail(MOONS_ERR_ARG, "moons_import_raw: null argument");
  std::vector<double> h(f->n);
  if (int rc = moons_field_download(f, h.data(), f->n)) return rc;
  if (int rc = moons_raw_read(f->s, f->dl, h.data(), dir, name, pad)) return rc;
  return moons_field_upload(f, h.data(), f->n);
}
