// Host-code memory check (built with -fsanitize=address,undefined by
// tests/test_host_asan.py, CPU only): the index builder on edge-case
// geometries, and the matrix/InfoGlobal file readers on truncated and
// randomly corrupted files, which must return an error status, never read or
// write out of bounds. Links cq_system.cpp and cq_store.cpp alone, so the
// library's error setters are replaced by the two stand-ins below.
#include <cstdarg>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "cqmm.h"
#include "cqmm_matrix.h"
#include "cqmm_system.h"

namespace cq {
int set_error(int status, const std::string &) { return status; }
int set_errorf(int status, const char *, ...) { return status; }
}  // namespace cq

static std::vector<char> slurp(const char *p) {
  std::vector<char> v;
  FILE *f = fopen(p, "rb");
  if (!f) return v;
  int c;
  while ((c = fgetc(f)) != EOF) v.push_back((char)c);
  fclose(f);
  return v;
}
static void spit(const char *p, const std::vector<char> &v, size_t n) {
  FILE *f = fopen(p, "wb");
  fwrite(v.data(), 1, n, f);
  fclose(f);
}

static int geometries() {
  struct G { int n; double dens, dmin; int g[3]; double rc[3]; int mode; };
  const G cases[] = {
      {1, 0.05, 0.0, {1, 1, 1}, {4.0, 5.0, 9.0}, 0},   {1, 0.05, 0.0, {2, 2, 2}, {3.0, 13.0, 16.0}, 1},
      {2, 0.05, 1.5, {1, 2, 3}, {3.0, 3.0, 6.0}, 0},   {7, 0.02, 1.5, {3, 1, 2}, {7.0, 9.0, 16.0}, 2},
      {50, 0.05, 1.5, {2, 3, 4}, {5.0, 6.0, 11.0}, 2}, {30, 0.05, 1.5, {5, 5, 5}, {0.5, 4.0, 4.5}, 0},
  };
  int bad = 0;
  for (const G &c : cases) {
    std::vector<double> xyz(3 * c.n);
    double edge = 0;
    if (cqmm_random_structure(c.n, c.dens, c.dmin, 7, xyz.data(), &edge)) { ++bad; continue; }
    const double box[3] = {edge, edge, edge};
    std::vector<int32_t> nsf(c.n);
    for (int i = 0; i < c.n; i++) nsf[i] = c.mode == 0 ? 4 : c.mode == 1 ? 13 : (i % 2 ? 1 : 22);
    cqmm_system *s = nullptr;
    const double cover = std::max(c.rc[0] + c.rc[1], c.rc[2]);
    if (cqmm_sys_create(c.n, xyz.data(), box, nsf.data(), c.g, cover, &s)) { ++bad; continue; }
    int32_t r[3];
    for (int k = 0; k < 3; k++) bad += cqmm_sys_add_range(s, c.rc[k], &r[k]) != 0;
    for (int k = 0; k < 3; k++) {
      cqmm_range_info ri;
      cqmm_sys_range_info(s, r[k], &ri);
      const size_t nb = (size_t)ri.n_blocks;
      std::vector<int32_t> prim(nb), gi(nb), gj(nb), sh(3 * nb), n1(nb), n2(nb);
      std::vector<int64_t> off(nb);
      bad += cqmm_sys_blocks(s, r[k], prim.data(), gi.data(), gj.data(), sh.data(), off.data(), n1.data(),
                             n2.data()) != 0;
      int64_t np = 0;
      cqmm_sys_transpose_pairs(s, r[k], &np, nullptr, nullptr, nullptr, nullptr);
      std::vector<int64_t> p(np), q(np);
      std::vector<int32_t> a(np), b(np);
      bad += cqmm_sys_transpose_pairs(s, r[k], &np, p.data(), q.data(), a.data(), b.data()) != 0;
      cqmm_matrix_index mi;
      bad += cqmm_sys_matrix_index(s, r[k], &mi) != 0;
      std::vector<double> v(ri.length);
      const double pp[4] = {0.1, 1.5, 0.5, 0.2};
      for (int kind = 0; kind < 4; kind++) bad += cqmm_sys_fill(s, r[k], kind, 3, 1, pp, 4, v.data()) != 0;
    }
    for (int mt = 1; mt <= 2; mt++) {
      cqmm_mult_desc d;
      bad += cqmm_sys_mult_desc(s, mt, r[0], r[1], r[2], &d) != 0;
    }
    int32_t rid;
    cqmm_range_info ri;
    bad += cqmm_sys_add_range(s, -1.0, &rid) == 0;
    bad += cqmm_sys_range_info(s, 99, &ri) == 0;
    cqmm_sys_destroy(s);
  }
  return bad;
}

static int matrix_files(cqmm_system *s, int32_t rid, int32_t rid2, int64_t len) {
  int bad = 0;
  std::vector<double> v(len), w(len);
  const double p[1] = {2.0};
  cqmm_sys_fill(s, rid, 0, 1, 1, p, 1, v.data());
  cqmm_sys_fill(s, rid, 0, 1, 2, p, 1, w.data());
  std::vector<double> a(len), b(len);
  double *outs[2] = {a.data(), b.data()};
  int32_t ns;
  int64_t bl, bs;
  for (int binary = 0; binary < 2; binary++) {
    const double *vals[2] = {v.data(), w.data()};
    const char *path = binary ? "m2.bin" : "m2.txt";
    bad += cqmm_sys_dump_matrix2(s, rid, 2, vals, path, binary) != 0;
    bad += cqmm_sys_load_matrix2(s, rid, path, binary, 2, outs, &ns, &bl, &bs) != 0 || a != v || b != w;
    cqmm_range_info ri;
    cqmm_sys_range_info(s, rid2, &ri);
    std::vector<double> c(ri.length), d(ri.length);
    double *outs2[2] = {c.data(), d.data()};
    bad += cqmm_sys_load_matrix2(s, rid2, path, binary, 2, outs2, &ns, &bl, &bs) != 0;
    bad += cqmm_sys_load_matrix2(s, rid, path, binary, 1, outs, &ns, &bl, &bs) == 0;  // 2 spins, room for 1
    const std::vector<char> orig = slurp(path);
    for (size_t n = 0; n < orig.size(); n += orig.size() / 200 + 1) {
      spit("t.tmp", orig, n);
      bad += cqmm_sys_load_matrix2(s, rid, "t.tmp", binary, 2, outs, &ns, &bl, &bs) == 0;  // truncated
    }
    std::mt19937 g(5 + binary);
    const char cs[] = "0123456789 -+.eE*,/\n";
    for (int t = 0; t < 200; t++) {
      std::vector<char> m = orig;
      for (int k = 1 + g() % 4; k > 0; k--) {
        const size_t at = g() % m.size();
        m[at] = binary ? (char)(g() & 255) : cs[g() % (sizeof(cs) - 1)];
      }
      spit("t.tmp", m, m.size());
      cqmm_sys_load_matrix2(s, rid, "t.tmp", binary, 2, outs, &ns, &bl, &bs);  // any status; no fault
    }
  }
  return bad;
}

static int info_global(cqmm_system *s, int n) {
  int bad = 0;
  std::vector<double> vel(3 * n, 0.25), co(3 * n), ve(3 * n);
  std::vector<int32_t> g2n(n);
  cqmm_info_global info;
  bad += cqmm_sys_dump_info_global(s, "ig.txt", 3, 1, 7, vel.data()) != 0;
  bad += cqmm_load_info_global("ig.txt", 3, n, &info, g2n.data(), co.data(), ve.data()) != 0 || info.natoms != n;
  bad += cqmm_load_info_global("ig.txt", 3, n - 1, &info, g2n.data(), co.data(), ve.data()) == 0;
  const std::vector<char> orig = slurp("ig.txt");
  for (size_t k = 0; k < orig.size(); k += 3) {
    spit("t.tmp", orig, k);
    cqmm_load_info_global("t.tmp", 3, n, &info, g2n.data(), co.data(), ve.data());
  }
  std::mt19937 g(7);
  const char cs[] = "0123456789 -+.eE*,/\nTFtf'\"";
  for (int t = 0; t < 500; t++) {
    std::vector<char> m = orig;
    for (int k = 1 + g() % 4; k > 0; k--) m[g() % m.size()] = cs[g() % (sizeof(cs) - 1)];
    spit("t.tmp", m, m.size());
    cqmm_load_info_global("t.tmp", 3, n, &info, g2n.data(), co.data(), ve.data());
    cqmm_load_info_global("t.tmp", 3, n, &info, nullptr, nullptr, nullptr);
  }
  std::vector<double> x(50), y(50);
  for (int i = 0; i < 50; i++) x[i] = (i - 25) * 1.37e3 + 1e-9 * i;
  bad += cqmm_dump_matrix_text("mt.txt", x.data(), 50) != 0;
  bad += cqmm_grab_matrix_text("mt.txt", y.data(), 50) != 0;
  const std::vector<char> mt = slurp("mt.txt");
  for (size_t k = 0; k < mt.size(); k += 5) {
    spit("t.tmp", mt, k);
    cqmm_grab_matrix_text("t.tmp", y.data(), 50);
  }
  char buf[64];
  for (int L = 0; L < 40; L++) {
    cqmm_matrix2_file_name("root_", 3, 12345, buf, L);
    cqmm_info_global_file_name(7, buf, L);
    cqmm_matrix_file_name("stub_", 2000, 17, buf, L);
  }
  return bad;
}

int main() {
  int bad = geometries();
  printf("geometries bad=%d\n", bad);
  const int n = 12;
  std::vector<double> xyz(3 * n);
  double edge = 0;
  cqmm_random_structure(n, 0.05, 2.0, 3, xyz.data(), &edge);
  const double box[3] = {edge, edge, edge};
  std::vector<int32_t> nsf(n);
  for (int i = 0; i < n; i++) nsf[i] = i % 3 == 0 ? 13 : 4;
  const int32_t ng[3] = {2, 2, 2};
  cqmm_system *s = nullptr;
  if (cqmm_sys_create(n, xyz.data(), box, nsf.data(), ng, 12.0, &s)) return 1;
  int32_t r1, r2;
  cqmm_sys_add_range(s, 5.0, &r1);
  cqmm_sys_add_range(s, 3.5, &r2);
  cqmm_range_info ri;
  cqmm_sys_range_info(s, r1, &ri);
  const int fbad = matrix_files(s, r1, r2, ri.length);
  printf("matrix files bad=%d\n", fbad);
  const int ibad = info_global(s, n);
  printf("info global bad=%d\n", ibad);
  cqmm_sys_destroy(s);
  bad += fbad + ibad;
  printf("done bad=%d\n", bad);
  return bad != 0;
}
