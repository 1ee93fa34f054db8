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
#include <cmath>
#include <cstdint>
#include <cstdio>
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

// Fortran Ew.d with w = 30, d = 20: 0.<20 digits>E<sign><2 digits>, right-aligned.
std::string e30_20(double x) {
  char m[64], out[64];
  if (x == 0.0 || !std::isfinite(x)) {
    snprintf(out, sizeof out, "%30.20E", x);
    return std::string(out) + "\n";
  }
  snprintf(m, sizeof m, "%.19E", std::fabs(x));  // d.ddd...E+xx: 20 significant digits
  const char *e = strchr(m, 'E');
  const int ex = atoi(e + 1) + 1;
  std::string digits;
  for (const char *c = m; c < e; ++c)
    if (*c != '.') digits += *c;
  snprintf(out, sizeof out, "%s0.%sE%c%02d", x < 0 ? "-" : "", digits.c_str(), ex < 0 ? '-' : '+', std::abs(ex));
  char pad[128];
  snprintf(pad, sizeof pad, "%30s\n", out);
  return pad;
}

// Tokens of a formatted file or records of a binary one, read in order.
struct Reader {
  FILE *f = nullptr;
  bool binary = true;
  std::vector<char> rec;
  size_t pos = 0;
  bool next_record() {
    uint32_t n = 0, n2 = 0;
    if (fread(&n, 4, 1, f) != 1) return false;
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
      v[k] = (int32_t)strtol(t.c_str(), nullptr, 10);
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
    for (int64_t k = 0; k < n; ++k) {
      if (!token(t)) return false;
      for (char &c : t)
        if (c == 'D' || c == 'd') c = 'E';
      v[k] = strtod(t.c_str(), nullptr);
    }
    return true;
  }
};

struct FileCloser {
  FILE *f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

}  // namespace

extern "C" {

int cqmm_matrix2_file_name(const char *root, int32_t index, int32_t inode, char *out, int32_t out_len) {
  if (!root || !out || out_len <= 0 || index < 0 || index > 99 || inode < 0 || inode > 999999)
    return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix2_file_name: bad argument");
  const int n = snprintf(out, (size_t)out_len, "%smatrix2.i%02d.p%06d", root, index, inode);
  if (n < 0 || n >= out_len) return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix2_file_name: buffer too short");
  return CQMM_OK;
}

int cqmm_sys_dump_matrix2(const cqmm_system *sys, int32_t range_id, int32_t nspin, const double *const *values,
                          const char *path, int32_t binary) {
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
}

int cqmm_sys_load_matrix2(const cqmm_system *sys, int32_t range_id, const char *path, int32_t binary,
                          int32_t nspin_max, double *const *out, int32_t *nspin, int64_t *blocks_loaded,
                          int64_t *blocks_skipped) {
  if (!sys || !path || !out || nspin_max < 1 || range_id < 0 || range_id >= (int32_t)sys->s.ranges.size())
    return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_load_matrix2: bad argument");
  const System &s = sys->s;
  const RangeIndex &r = s.ranges[range_id];
  Reader rd;
  rd.binary = binary != 0;
  rd.f = fopen(path, rd.binary ? "rb" : "r");
  if (!rd.f) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_load_matrix2: cannot open %s", path);
  FileCloser fc{rd.f};
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
}

}  // extern "C"
