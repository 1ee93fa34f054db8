// Matrix files in CONQUEST's store_matrix format for one process: the writer
// follows dump_matrix2 (src/store_matrix_module.f90:340-470, contents from
// set_matrix_store :499-651), the reader maps a file's blocks onto a range of
// this index set the way UpdateMatrix_local does for unmoved atoms
// (src/UpdateInfo_module.f90:1851-2162): block (i, j) of the file goes to the
// block of global atoms i and j whose periodic image sits at R_i - vec_Rij,
// when that block exists in the range; other file blocks are skipped.
//
// File contents, in write order (binary: one Fortran unformatted sequential
// record per write — 4-byte length, payload, 4-byte length; formatted:
// list-directed lines, values as e30.20):
//   inode, nprim
//   nsf_spec_i(1:nprim)        support functions of each primary atom
//   jmax_i(1:nprim)            neighbours of each primary atom in the range
//   jbeta_max_i(1:nprim)       sum of the neighbours' support functions
//   nspin
//   per primary atom: idglob_i; beta_j(1:jmax); idglob_j(1:jmax);
//     vec_Rij(1:3) per neighbour (one record each: (R_i - R_j image) / cell);
//     per spin, the atom's jbeta_max_i * nsf_spec_i values (one record each)
// File name: <root>matrix2.i<index, 2 digits>.p<inode, 6 digits> (get_file_name_2rank,
// src/io_module.f90:3134-3164); one process writes inode 1.
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cqmm.h"
#include "../../include/cqmm_system.h"
#include "cq_common.h"
#include "cq_system.h"

namespace {

using cq::RangeIndex;
using cq::System;

struct Writer {
  FILE *f = nullptr;
  bool binary = true;
  std::vector<char> rec;
  void put(const void *p, size_t n) { rec.insert(rec.end(), (const char *)p, (const char *)p + n); }
  template <class T>
  void put(T v) { put(&v, sizeof(T)); }
  // Close a record: length-prefixed payload (binary) or one text line.
  bool end_record(const std::string &text) {
    if (binary) {
      const uint32_t n = (uint32_t)rec.size();
      const bool ok = fwrite(&n, 4, 1, f) == 1 && (n == 0 || fwrite(rec.data(), n, 1, f) == 1) &&
                      fwrite(&n, 4, 1, f) == 1;
      rec.clear();
      return ok;
    }
    rec.clear();
    return fputs(text.c_str(), f) >= 0;
  }
};

std::string ints_line(const int32_t *v, int64_t n) {
  std::string s;
  char buf[32];
  for (int64_t k = 0; k < n; ++k) {
    snprintf(buf, sizeof buf, " %d", v[k]);
    s += buf;
  }
  return s + "\n";
}

// Fortran Ew.d: [-]0.<d digits>E<sign><2 digits>, right-aligned in w columns; an
// exponent beyond two digits drops the E (0.<d digits><sign><3 digits>), and a
// field that does not fit is w asterisks, as a Fortran write produces.
std::string e_format(double x, int w, int d) {
  char m[64], out[80], pad[128];
  if (x == 0.0 || !std::isfinite(x)) {
    snprintf(out, sizeof out, "%*.*E", w, d, x);
  } else {
    snprintf(m, sizeof m, "%.*E", d - 1, std::fabs(x));  // d.ddd...E+xx: d significant digits
    const char *e = strchr(m, 'E');
    const int ex = atoi(e + 1) + 1;
    std::string digits;
    for (const char *c = m; c < e; ++c)
      if (*c != '.') digits += *c;
    if (std::abs(ex) <= 99)
      snprintf(out, sizeof out, "%s0.%sE%c%02d", x < 0 ? "-" : "", digits.c_str(), ex < 0 ? '-' : '+', std::abs(ex));
    else
      snprintf(out, sizeof out, "%s0.%s%c%03d", x < 0 ? "-" : "", digits.c_str(), ex < 0 ? '-' : '+', std::abs(ex));
  }
  if ((int)strlen(out) > w) return std::string((size_t)w, '*');
  snprintf(pad, sizeof pad, "%*s", w, out);
  return pad;
}

std::string e30_20(double x) { return e_format(x, 30, 20) + "\n"; }

// Fortran Fw.d; a value that does not fit is w asterisks.
std::string f_format(double x, int w, int d) {
  char out[400];
  snprintf(out, sizeof out, "%*.*f", w, d, x);
  if ((int)strlen(out) > w) return std::string((size_t)w, '*');
  return out;
}

// A real as a Fortran read accepts it: D or d exponent letters, or an
// exponent sign straight after the digits (0.123+100). False on anything else.
bool fortran_real(std::string t, double *v) {
  for (char &c : t)
    if (c == 'D' || c == 'd' || c == 'e') c = 'E';
  if (t.find('E') == std::string::npos) {
    const size_t k = t.find_first_of("+-", 1);
    if (k != std::string::npos) t.insert(k, "E");
  }
  if (t.empty() || t.find('*') != std::string::npos) return false;
  char *end = nullptr;
  *v = strtod(t.c_str(), &end);
  return end && *end == '\0';
}

// Tokens of a formatted file or records of a binary one, read in order.
struct Reader {
  FILE *f = nullptr;
  bool binary = true;
  int64_t fsize = 0;  // bytes in the file: no count read from it may ask for more
  std::vector<char> rec;
  size_t pos = 0;
  bool open(const char *path) {
    f = fopen(path, binary ? "rb" : "r");
    if (!f) return false;
    if (fseek(f, 0, SEEK_END) == 0) fsize = ftell(f);
    rewind(f);
    return fsize >= 0;
  }
  int64_t left() const { return fsize - (int64_t)ftell(f); }
  bool next_record() {
    uint32_t n = 0, n2 = 0;
    if (fread(&n, 4, 1, f) != 1) return false;
    if ((int64_t)n + 4 > left()) return false;
    rec.resize(n);
    if (n && fread(rec.data(), n, 1, f) != 1) return false;
    if (fread(&n2, 4, 1, f) != 1 || n2 != n) return false;
    pos = 0;
    return true;
  }
  bool token(std::string &t) {
    t.clear();
    int ch;
    while ((ch = fgetc(f)) != EOF && (ch == ' ' || ch == '\n' || ch == '\r' || ch == '\t' || ch == ','))
      ;
    if (ch == EOF) return false;
    do t += (char)ch;
    while ((ch = fgetc(f)) != EOF && !(ch == ' ' || ch == '\n' || ch == '\r' || ch == '\t' || ch == ','));
    return true;
  }
  // Read n int32 (binary: one record holding exactly n values).
  bool ints(int32_t *v, int64_t n) {
    if (binary) {
      if (!next_record() || rec.size() != (size_t)n * 4) return false;
      if (n) memcpy(v, rec.data(), (size_t)n * 4);
      return true;
    }
    std::string t;
    for (int64_t k = 0; k < n; ++k) {
      if (!token(t)) return false;
      char *end = nullptr;
      errno = 0;
      const long x = strtol(t.c_str(), &end, 10);
      if (errno || *end != '\0' || x < INT32_MIN || x > INT32_MAX) return false;
      v[k] = (int32_t)x;
    }
    return true;
  }
  bool doubles(double *v, int64_t n) {
    if (binary) {
      if (!next_record() || rec.size() != (size_t)n * 8) return false;
      memcpy(v, rec.data(), (size_t)n * 8);
      return true;
    }
    std::string t;
    for (int64_t k = 0; k < n; ++k)
      if (!token(t) || !fortran_real(t, &v[k])) return false;
    return true;
  }
};

// List-directed reads of a text file: each read takes its items from the
// current line on (continuing onto the next lines when one runs short),
// accepts r*c repeat counts, and then moves past the rest of the line, as a
// Fortran list-directed READ does.
struct ListReader {
  std::vector<std::string> lines;
  size_t line = 0;
  bool open(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return false;
    std::string cur;
    int ch;
    while ((ch = fgetc(f)) != EOF) {
      if (ch == '\n') {
        lines.push_back(cur);
        cur.clear();
      } else if (ch != '\r') {
        cur += (char)ch;
      }
    }
    if (!cur.empty()) lines.push_back(cur);
    fclose(f);
    return true;
  }
  // The next n items as strings (repeat counts expanded).
  bool items(std::vector<std::string> &out, size_t n) {
    out.clear();
    while (out.size() < n) {
      if (line >= lines.size()) return false;
      const std::string &l = lines[line++];
      size_t k = 0;
      while (out.size() < n && k < l.size()) {
        while (k < l.size() && (l[k] == ' ' || l[k] == ',' || l[k] == '\t')) ++k;
        if (k >= l.size()) break;
        const size_t b = k;
        while (k < l.size() && !(l[k] == ' ' || l[k] == ',' || l[k] == '\t')) ++k;
        std::string t = l.substr(b, k - b);
        const size_t star = t.find('*');
        long rep = 1;
        if (star != std::string::npos && star > 0) {
          rep = strtol(t.substr(0, star).c_str(), nullptr, 10);
          t = t.substr(star + 1);
          if (rep < 1) return false;
        }
        for (long r = 0; r < rep && out.size() < n; ++r) out.push_back(t);
      }
    }
    return true;
  }
  bool ints(int32_t *v, size_t n) {
    std::vector<std::string> t;
    if (!items(t, n)) return false;
    for (size_t k = 0; k < n; ++k) {
      if (t[k].empty()) return false;  // a null value (r*) sets nothing here
      char *end = nullptr;
      const long x = strtol(t[k].c_str(), &end, 10);
      if (!end || *end) return false;
      v[k] = (int32_t)x;
    }
    return true;
  }
  bool reals(double *v, size_t n) {
    std::vector<std::string> t;
    if (!items(t, n)) return false;
    for (size_t k = 0; k < n; ++k)
      if (!fortran_real(t[k], &v[k])) return false;
    return true;
  }
  // Logicals: T, F, .TRUE., .false. and the like (the letter after an optional dot).
  bool logicals(int32_t *v, size_t n) {
    std::vector<std::string> t;
    if (!items(t, n)) return false;
    for (size_t k = 0; k < n; ++k) {
      const std::string &s = t[k];
      const char c = (char)toupper(s[0] == '.' && s.size() > 1 ? s[1] : s[0]);
      if (c != 'T' && c != 'F') return false;
      v[k] = c == 'T';
    }
    return true;
  }
  // The current line as is (an explicitly formatted record), then move past it.
  bool record(std::string &out) {
    if (line >= lines.size()) return false;
    out = lines[line++];
    return true;
  }
};

// Fixed-column field [b, b + w) of a formatted record (blank past its end).
std::string field(const std::string &rec, size_t b, size_t w) {
  if (b >= rec.size()) return std::string();
  std::string t = rec.substr(b, w);
  const size_t s = t.find_first_not_of(' '), e = t.find_last_not_of(' ');
  return s == std::string::npos ? std::string() : t.substr(s, e - s + 1);
}

struct FileCloser {
  FILE *f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

}  // namespace

extern "C" {

int cqmm_matrix2_file_name(const char *root, int32_t index, int32_t inode, char *out, int32_t out_len) {
  return cq::guarded(__func__, [&]() -> int {
    if (!root || !out || out_len <= 0 || index < 0 || index > 99 || inode < 0 || inode > 999999)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix2_file_name: bad argument");
    const int n = snprintf(out, (size_t)out_len, "%smatrix2.i%02d.p%06d", root, index, inode);
    if (n < 0 || n >= out_len) return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix2_file_name: buffer too short");
    return CQMM_OK;
  });
}

int cqmm_sys_dump_matrix2(const cqmm_system *sys, int32_t range_id, int32_t nspin, const double *const *values,
                          const char *path, int32_t binary) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !values || !path || nspin < 1 || range_id < 0 || range_id >= (int32_t)sys->s.ranges.size())
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_dump_matrix2: bad argument");
    for (int32_t sp = 0; sp < nspin; ++sp)
      if (!values[sp]) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_dump_matrix2: null values");
    const System &s = sys->s;
    const RangeIndex &r = s.ranges[range_id];
    Writer w;
    w.binary = binary != 0;
    w.f = fopen(path, w.binary ? "wb" : "w");
    if (!w.f) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_dump_matrix2: cannot open %s", path);
    FileCloser fc{w.f};
    const int32_t nprim = s.natoms;
    std::vector<int32_t> nsf_i(nprim), jmax(nprim), jbeta(nprim, 0);
    for (int32_t p = 0; p < nprim; ++p) {
      nsf_i[p] = s.pnsf[p];
      jmax[p] = (int32_t)(r.row_beg[p + 1] - r.row_beg[p]);
      for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e) jbeta[p] += r.nb_ndimj[e];
    }
    bool ok = true;
    const int32_t head[2] = {1, nprim};
    w.put(head, 8);
    ok = ok && w.end_record(ints_line(head, 2));
    w.put(nsf_i.data(), nsf_i.size() * 4);
    ok = ok && w.end_record(ints_line(nsf_i.data(), nprim));
    w.put(jmax.data(), jmax.size() * 4);
    ok = ok && w.end_record(ints_line(jmax.data(), nprim));
    w.put(jbeta.data(), jbeta.size() * 4);
    ok = ok && w.end_record(ints_line(jbeta.data(), nprim));
    w.put(nspin);
    ok = ok && w.end_record(ints_line(&nspin, 1));
    std::vector<int32_t> beta, idj;
    for (int32_t p = 0; p < nprim && ok; ++p) {
      const int32_t idi = s.prim_gid[p] + 1;
      w.put(idi);
      ok = ok && w.end_record(ints_line(&idi, 1));
      beta.clear();
      idj.clear();
      for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e) {
        beta.push_back(r.nb_ndimj[e]);
        idj.push_back(s.prim_gid[s.c_prim[r.nb_cover[e]]] + 1);
      }
      w.put(beta.data(), beta.size() * 4);
      ok = ok && w.end_record(ints_line(beta.data(), (int64_t)beta.size()));
      w.put(idj.data(), idj.size() * 4);
      ok = ok && w.end_record(ints_line(idj.data(), (int64_t)idj.size()));
      for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1] && ok; ++e) {
        const int32_t m = r.nb_cover[e];
        const double v[3] = {(s.px[p] - s.cx[m]) / s.box[0], (s.py[p] - s.cy[m]) / s.box[1],
                             (s.pz[p] - s.cz[m]) / s.box[2]};
        w.put(v, 24);
        char line[128];
        snprintf(line, sizeof line, " %.17E %.17E %.17E\n", v[0], v[1], v[2]);
        ok = ok && w.end_record(line);
      }
      const int64_t len = (int64_t)jbeta[p] * nsf_i[p];
      for (int32_t sp = 0; sp < nspin && ok; ++sp)
        for (int64_t k = 0; k < len && ok; ++k) {
          const double x = values[sp][r.row_off[p] + k];
          w.put(x);
          ok = w.end_record(e30_20(x));
        }
    }
    if (!ok) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_dump_matrix2: write to %s failed", path);
    return CQMM_OK;
  });
}

int cqmm_sys_load_matrix2(const cqmm_system *sys, int32_t range_id, const char *path, int32_t binary,
                          int32_t nspin_max, double *const *out, int32_t *nspin, int64_t *blocks_loaded,
                          int64_t *blocks_skipped) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !path || !out || nspin_max < 1 || range_id < 0 || range_id >= (int32_t)sys->s.ranges.size())
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_load_matrix2: bad argument");
    const System &s = sys->s;
    const RangeIndex &r = s.ranges[range_id];
    Reader rd;
    rd.binary = binary != 0;
    const bool opened = rd.open(path);
    FileCloser fc{rd.f};
    if (!opened) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_load_matrix2: cannot open %s", path);
    auto bad = [&](const char *what) {
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_load_matrix2: %s: malformed or truncated (%s)", path, what);
    };
    int32_t head[2];
    if (!rd.ints(head, 2)) return bad("inode, nprim");
    const int32_t nprim = head[1];
    if (nprim < 0 || nprim > s.natoms) return bad("nprim");
    std::vector<int32_t> nsf_i(nprim), jmax(nprim), jbeta(nprim);
    int32_t ns = 0;
    if (!rd.ints(nsf_i.data(), nprim) || !rd.ints(jmax.data(), nprim) || !rd.ints(jbeta.data(), nprim) ||
        !rd.ints(&ns, 1))
      return bad("header");
    // Counts that the file cannot hold (every value takes at least one byte of
    // it) are rejected before anything is sized by them.
    for (int32_t k = 0; k < nprim; ++k)
      if (nsf_i[k] < 1 || nsf_i[k] > 64 || jmax[k] < 0 || jmax[k] > rd.fsize || jbeta[k] < 0 ||
          jbeta[k] > 64 * (int64_t)jmax[k])
        return bad("row sizes");
    if (ns < 1 || ns > nspin_max) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_load_matrix2: %s holds %d spins, room for %d", path, ns, nspin_max);
    for (int32_t sp = 0; sp < ns; ++sp)
      if (!out[sp]) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_load_matrix2: null output");
    // Target blocks by (global i, global j, image shift).
    std::vector<int32_t> prim_of(s.natoms, -1);
    for (int32_t p = 0; p < s.natoms; ++p) prim_of[s.prim_gid[p]] = p;
    auto key = [&](int64_t gi, int64_t gj, int32_t sx, int32_t sy, int32_t sz) {
      return (((gi * s.natoms + gj) * 64 + (sx + 32)) * 64 + (sy + 32)) * 64 + (sz + 32);
    };
    std::unordered_map<int64_t, int64_t> block;
    block.reserve((size_t)(r.row_beg[s.natoms] * 2));
    for (int32_t p = 0; p < s.natoms; ++p)
      for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e) {
        const int32_t m = r.nb_cover[e], ic = s.c_part[m];
        block[key(s.prim_gid[p], s.prim_gid[s.c_prim[m]], s.cs_shift[3 * ic], s.cs_shift[3 * ic + 1],
                  s.cs_shift[3 * ic + 2])] = e;
      }
    int64_t loaded = 0, skipped = 0;
    std::vector<int32_t> beta, idj;
    std::vector<double> vec, vals;
    for (int32_t k = 0; k < nprim; ++k) {
      int32_t idi = 0;
      if (!rd.ints(&idi, 1)) return bad("idglob_i");
      beta.resize(jmax[k]);
      idj.resize(jmax[k]);
      vec.resize(3 * (size_t)jmax[k]);
      if (!rd.ints(beta.data(), jmax[k]) || !rd.ints(idj.data(), jmax[k])) return bad("neighbours");
      for (int32_t jj = 0; jj < jmax[k]; ++jj)
        if (!rd.doubles(&vec[3 * (size_t)jj], 3)) return bad("vec_Rij");
      const int64_t len = (int64_t)jbeta[k] * nsf_i[k];
      if (len * ns > rd.left()) return bad("values");
      vals.resize((size_t)len * ns);
      for (int32_t sp = 0; sp < ns; ++sp)
        for (int64_t q = 0; q < len; ++q)
          if (!rd.doubles(&vals[(size_t)sp * len + q], 1)) return bad("values");
      if (idi < 1 || idi > s.natoms) return bad("idglob_i out of range");
      const int32_t p = prim_of[idi - 1];
      if (nsf_i[k] != s.pnsf[p]) return bad("support functions of atom i differ");
      int64_t voff = 0;
      for (int32_t jj = 0; jj < jmax[k]; ++jj) {
        const int32_t jg = idj[jj] - 1;
        const int64_t nval = (int64_t)beta[jj] * nsf_i[k];
        if (jg < 0 || jg >= s.natoms) return bad("idglob_j out of range");
        if (beta[jj] < 1 || beta[jj] > 64 || voff + nval > len) return bad("neighbour sizes exceed jbeta_max_i");
        const int32_t pj = prim_of[jg];
        // Image of j at R_i - vec * cell: shift = (R_i - R_j) / cell - vec.
        const double fx = (s.px[p] - s.px[pj]) / s.box[0] - vec[3 * (size_t)jj];
        const double fy = (s.py[p] - s.py[pj]) / s.box[1] - vec[3 * (size_t)jj + 1];
        const double fz = (s.pz[p] - s.pz[pj]) / s.box[2] - vec[3 * (size_t)jj + 2];
        const int32_t sx = (int32_t)std::lround(fx), sy = (int32_t)std::lround(fy), sz = (int32_t)std::lround(fz);
        auto it = (std::abs(sx) < 32 && std::abs(sy) < 32 && std::abs(sz) < 32) ? block.find(key(idi - 1, jg, sx, sy, sz))
                                                                                : block.end();
        if (it == block.end() || r.nb_ndimj[it->second] != beta[jj]) {
          ++skipped;
        } else {
          const int64_t off = r.nb_off[it->second];
          for (int32_t sp = 0; sp < ns; ++sp)
            memcpy(out[sp] + off, &vals[(size_t)sp * len + voff], (size_t)nval * 8);
          ++loaded;
        }
        voff += nval;
      }
      if (voff != len) return bad("neighbour sizes do not add up to jbeta_max_i");
    }
    if (nspin) *nspin = ns;
    if (blocks_loaded) *blocks_loaded = loaded;
    if (blocks_skipped) *blocks_skipped = skipped;
    return CQMM_OK;
  });
}

int cqmm_info_global_file_name(int32_t index, char *out, int32_t out_len) {
  return cq::guarded(__func__, [&]() -> int {
    if (!out || out_len <= 0 || index < 0 || index > 99)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_info_global_file_name: bad argument");
    const int n = snprintf(out, (size_t)out_len, "InfoGlobal.i%02d", index);
    if (n < 0 || n >= out_len) return cq::set_error(CQMM_ERR_ARG, "cqmm_info_global_file_name: buffer too short");
    return CQMM_OK;
  });
}

int cqmm_sys_dump_info_global(const cqmm_system *sys, const char *path, int32_t index, int32_t has_md_step,
                              int32_t md_step, const double *velocity) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !path || index < 0 || index > 99)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_dump_info_global: bad argument");
    const System &s = sys->s;
    FILE *f = fopen(path, "w");
    if (!f) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_dump_info_global: cannot open %s", path);
    FileCloser fc{f};
    // Coordinates by global id (atom_coord(1:3, id_glob)); one process owns every atom.
    std::vector<double> xyz(3 * (size_t)s.natoms);
    for (int32_t p = 0; p < s.natoms; ++p) {
      const size_t g = (size_t)s.prim_gid[p];
      xyz[3 * g] = s.px[p];
      xyz[3 * g + 1] = s.py[p];
      xyz[3 * g + 2] = s.pz[p];
    }
    const int32_t step = has_md_step ? md_step : 0;
    bool ok = fprintf(f, " %c %c  = flag_velocity, flag_MDstep\n", velocity ? 'T' : 'F', has_md_step ? 'T' : 'F') > 0;
    ok = ok && fprintf(f, " %11d %11d  # of atoms, # of process \n", s.natoms, 1) > 0;
    ok = ok && fprintf(f, " %11d %11d %11d  npcellx,y,z\n", s.ng[0], s.ng[1], s.ng[2]) > 0;
    ok = ok && fprintf(f, "%s%s%s rcellx,y,z\n", f_format(s.box[0], 25, 15).c_str(), f_format(s.box[1], 25, 15).c_str(),
                       f_format(s.box[2], 25, 15).c_str()) > 0;
    std::string nodes;
    for (int32_t g = 0; g < s.natoms && ok; ++g) nodes += "           1";
    ok = ok && fprintf(f, "%s\n", nodes.c_str()) >= 0;
    ok = ok && fprintf(f, " %11d  MD step\n", step) > 0;
    auto atom_lines = [&](const double *v) {
      for (int32_t g = 0; g < s.natoms && ok; ++g)
        ok = fprintf(f, "     %8d   %s%s%s\n", g + 1, e_format(v[3 * (size_t)g], 22, 13).c_str(),
                     e_format(v[3 * (size_t)g + 1], 22, 13).c_str(), e_format(v[3 * (size_t)g + 2], 22, 13).c_str()) > 0;
    };
    atom_lines(xyz.data());
    if (velocity) {
      ok = ok && fprintf(f, " velocity\n") > 0;
      atom_lines(velocity);
    }
    ok = ok && fprintf(f, " %11d  = index_local in dump_InfoMatGlobal\n", index) > 0;
    if (!ok) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_dump_info_global: write to %s failed", path);
    return CQMM_OK;
  });
}

int cqmm_load_info_global(const char *path, int32_t index, int32_t natoms_max, cqmm_info_global *info,
                          int32_t *glob_to_node, double *coords, double *velocity) {
  return cq::guarded(__func__, [&]() -> int {
    if (!path || !info || natoms_max < 0 || index < 0)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_load_info_global: bad argument");
    ListReader rd;
    if (!rd.open(path)) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_load_info_global: cannot open %s", path);
    auto bad = [&](const char *what) {
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_load_info_global: %s: malformed or truncated (%s)", path, what);
    };
    cqmm_info_global g;
    memset(&g, 0, sizeof g);
    int32_t flags[2], counts[2];
    if (!rd.logicals(flags, 2)) return bad("flag_velocity, flag_MDstep");
    g.has_velocity = flags[0];
    g.has_md_step = flags[1];
    if (!rd.ints(counts, 2)) return bad("ni_in_cell, numprocs");
    g.natoms = counts[0];
    g.numprocs = counts[1];
    // Each atom has a coordinate line, so a count above the file's lines is a
    // corrupt or truncated file (checked before the count sizes anything).
    if (g.natoms < 0 || (size_t)g.natoms > rd.lines.size()) return bad("ni_in_cell");
    if (g.natoms > natoms_max && (glob_to_node || coords || velocity))
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_load_info_global: %s holds %d atoms, room for %d", path, g.natoms,
                            natoms_max);
    if (!rd.ints(g.npcell, 3)) return bad("npcellx,y,z");
    if (!rd.reals(g.rcell, 3)) return bad("rcellx,y,z");
    std::vector<int32_t> nodes((size_t)g.natoms);
    if (g.natoms && !rd.ints(nodes.data(), nodes.size())) return bad("glob_to_node");
    if (!rd.ints(&g.md_step, 1)) return bad("MD step");
    // (5x, i8, 3x, 3e22.13) records: global id in columns 6-13, values from column 17.
    auto atom_lines = [&](double *v) {
      std::string rec;
      for (int32_t a = 0; a < g.natoms; ++a) {
        if (!rd.record(rec)) return false;
        double x[3];
        for (int c = 0; c < 3; ++c)
          if (!fortran_real(field(rec, 16 + 22 * (size_t)c, 22), &x[c])) return false;
        if (v)
          for (int c = 0; c < 3; ++c) v[3 * (size_t)a + c] = x[c];
      }
      return true;
    };
    if (!atom_lines(coords)) return bad("atom_coord");
    if (g.has_velocity) {
      std::string skip;
      if (!rd.record(skip) || !atom_lines(velocity)) return bad("atom_veloc");
    }
    if (!rd.ints(&g.index, 1)) return bad("index");
    if (g.index != index)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_load_info_global: %s holds index %d, expected %d", path, g.index, index);
    if (glob_to_node && g.natoms) memcpy(glob_to_node, nodes.data(), nodes.size() * 4);
    *info = g;
    return CQMM_OK;
  });
}

int cqmm_matrix_file_name(const char *stub, int32_t numprocs, int32_t inode, char *out, int32_t out_len) {
  return cq::guarded(__func__, [&]() -> int {
    if (!stub || !out || out_len <= 0 || numprocs < 1 || inode < 1)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix_file_name: bad argument");
    int digits = 0;
    for (int32_t n = numprocs; n > 0; n /= 10) ++digits;
    const int width = digits > 3 ? digits : 3;
    const int n = snprintf(out, (size_t)out_len, "%smatrix.%0*d", stub, width, inode);
    if (n < 0 || n >= out_len || strlen(stub) + 1 > 80 || n > 80)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix_file_name: name too long");
    return CQMM_OK;
  });
}

int cqmm_dump_matrix_text(const char *path, const double *values, int64_t len) {
  return cq::guarded(__func__, [&]() -> int {
    if (!path || len < 0 || (len && !values)) return cq::set_error(CQMM_ERR_ARG, "cqmm_dump_matrix_text: bad argument");
    FILE *f = fopen(path, "w");
    if (!f) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_dump_matrix_text: cannot open %s", path);
    FileCloser fc{f};
    for (int64_t k = 0; k < len; ++k)
      if (fprintf(f, "%s\n", f_format(values[k], 25, 18).c_str()) < 0)
        return cq::set_errorf(CQMM_ERR_ARG, "cqmm_dump_matrix_text: write to %s failed", path);
    return CQMM_OK;
  });
}

int cqmm_grab_matrix_text(const char *path, double *values, int64_t len) {
  return cq::guarded(__func__, [&]() -> int {
    if (!path || len < 0 || (len && !values)) return cq::set_error(CQMM_ERR_ARG, "cqmm_grab_matrix_text: bad argument");
    ListReader rd;
    if (!rd.open(path)) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_grab_matrix_text: failed to open input file %s", path);
    std::string rec;
    for (int64_t k = 0; k < len; ++k)
      if (!rd.record(rec) || !fortran_real(field(rec, 0, 25), &values[k]))
        return cq::set_errorf(CQMM_ERR_ARG, "cqmm_grab_matrix_text: %s: value %lld missing or unreadable", path,
                              (long long)k + 1);
    return CQMM_OK;
  });
}

}  // extern "C"
