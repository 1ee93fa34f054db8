// Single-process mirror of CONQUEST's sparsity indexing (host side).
//
// For one process every cell partition is in the primary set, so the cover set,
// neighbour lists, halos and local-transpose indices reduce to the loops below.
// Each routine names the reference routine whose output it reproduces; the
// arrays keep that routine's layout and Fortran values because cqmm_plan_create
// consumes exactly what a CONQUEST host would pass.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "cq_common.h"
#include "cq_system.h"
#include "../../include/cqmm_matrix.h"
#include <unordered_map>

using cq::RangeIndex;
using cq::System;

namespace {

constexpr double kRdErr = 1.0e-8;      // RD_ERR, src/numbers_module.f90:96
constexpr double kVerySmall = 1.0e-8;  // very_small, src/numbers_module.f90:93

inline int32_t lab3(int32_t x, int32_t y, int32_t z, int32_t ny, int32_t nz) {
  // 1-based raster label of 1-based (x,y,z); the CC label of group_module.
  return (x - 1) * ny * nz + (y - 1) * nz + z;
}

// Cover set of the whole cell for one process (convert_primary + make_cs,
// src/cover_module.f90:133-429,533-700). Primary-set origin (1,1,1), nleft 0,
// node order = raster order of cell partitions.
void build_cover(System &s) {
  const int32_t P = s.n_parts;
  for (int d = 0; d < 3; ++d) {
    const double dc = s.box[d] / s.ng[d];
    const double lh = -s.rcut_cover;                 // dx*(origin-1-nleft) - rcut
    const double rh = dc * s.ng[d] + s.rcut_cover;   // dx*(origin-1-nleft+nw) + rcut
    const double csl = (0.0 - lh) / dc;
    int32_t nspan = (int32_t)std::ceil(csl);
    int32_t ncov = (int32_t)std::ceil(rh / dc) + nspan;
    const double lhc = dc * (0 - nspan);
    const double rhc = dc * (0 - nspan + ncov);
    if (lh + kVerySmall < lhc) nspan += (int32_t)std::ceil((lhc - lh) / dc);
    if (rh > rhc + kVerySmall) ncov += (int32_t)std::ceil((rh - rhc) / dc);
    s.nspanl[d] = nspan;
    s.ncover[d] = ncov;
  }
  const int32_t ncx = s.ncover[0], ncy = s.ncover[1], ncz = s.ncover[2];
  s.ng_cover = ncx * ncy * ncz;
  // Periodic replicas of each irreducible group along each edge.
  int32_t nrep[3][4096];
  int32_t nmod[3];
  for (int d = 0; d < 3; ++d) {
    const int32_t ng = s.ng[d];
    nmod[d] = ((ng + s.nspanl[d] - 1) / ng) * ng;
    const int32_t nocc = s.ncover[d] / ng;
    const int32_t nrem = s.ncover[d] - nocc * ng;
    for (int32_t g = 1; g <= ng; ++g) nrep[d][g - 1] = (g <= nrem) ? nocc + 1 : nocc;
  }
  // Irreducible set (minx = ngcellx for one process), sorted by node label.
  struct MinEl { int32_t key, gx, gy, gz; };
  std::vector<MinEl> mins;
  mins.reserve(P);
  for (int32_t gx = 1; gx <= s.ng[0]; ++gx)
    for (int32_t gy = 1; gy <= s.ng[1]; ++gy)
      for (int32_t gz = 1; gz <= s.ng[2]; ++gz) {
        const int32_t qx = 1 + (gx - s.nspanl[0] - 1 + nmod[0]) % s.ng[0];
        const int32_t qy = 1 + (gy - s.nspanl[1] - 1 + nmod[1]) % s.ng[1];
        const int32_t qz = 1 + (gz - s.nspanl[2] - 1 + nmod[2]) % s.ng[2];
        mins.push_back({lab3(qx, qy, qz, s.ng[1], s.ng[2]), gx, gy, gz});
      }
  std::sort(mins.begin(), mins.end(), [](const MinEl &a, const MinEl &b) { return a.key < b.key; });
  std::vector<int32_t> nsx, nsy, nsz;
  nsx.reserve(s.ng_cover); nsy.reserve(s.ng_cover); nsz.reserve(s.ng_cover);
  for (const MinEl &m : mins)
    for (int32_t rx = 1; rx <= nrep[0][m.gx - 1]; ++rx)
      for (int32_t ry = 1; ry <= nrep[1][m.gy - 1]; ++ry)
        for (int32_t rz = 1; rz <= nrep[2][m.gz - 1]; ++rz) {
          nsx.push_back(m.gx - 1 - s.nspanl[0] + (rx - 1) * s.ng[0]);
          nsy.push_back(m.gy - 1 - s.nspanl[1] + (ry - 1) * s.ng[1]);
          nsz.push_back(m.gz - 1 - s.nspanl[2] + (rz - 1) * s.ng[2]);
        }
  s.lab_cell.assign(s.ng_cover, 0);
  s.lab_cover.assign(s.ng_cover, 0);
  s.n_ing_cover.assign(s.ng_cover, 0);
  s.icover_ibeg.assign(s.ng_cover, 0);
  s.cs_shift.assign(3 * (size_t)s.ng_cover, 0);
  s.cx.clear(); s.cy.clear(); s.cz.clear(); s.c_prim.clear(); s.c_part.clear();
  int32_t nm = 0;
  for (int32_t ic = 0; ic < s.ng_cover; ++ic) {
    const int32_t qx = 1 + (1 + nsx[ic] + nmod[0] - 1) % s.ng[0];
    const int32_t qy = 1 + (1 + nsy[ic] + nmod[1] - 1) % s.ng[1];
    const int32_t qz = 1 + (1 + nsz[ic] + nmod[2] - 1) % s.ng[2];
    const int32_t mx = 1 + nsx[ic] - qx, my = 1 + nsy[ic] - qy, mz = 1 + nsz[ic] - qz;
    const double xadd = mx * s.dcell[0], yadd = my * s.dcell[1], zadd = mz * s.dcell[2];
    const int32_t q = lab3(qx, qy, qz, s.ng[1], s.ng[2]);
    s.lab_cell[ic] = q;
    s.lab_cover[ic] = (nsx[ic] + s.nspanl[0]) * ncy * ncz + (nsy[ic] + s.nspanl[1]) * ncz +
                      (nsz[ic] + s.nspanl[2]) + 1;
    s.cs_shift[3 * ic + 0] = mx / s.ng[0];
    s.cs_shift[3 * ic + 1] = my / s.ng[1];
    s.cs_shift[3 * ic + 2] = mz / s.ng[2];
    s.icover_ibeg[ic] = nm + 1;
    const int32_t nmg = s.nm_group[q - 1];
    s.n_ing_cover[ic] = nmg;
    for (int32_t a = 0; a < nmg; ++a) {
      const int32_t p = s.icell_beg[q - 1] + a;
      s.cx.push_back(s.px[p] + xadd);
      s.cy.push_back(s.py[p] + yadd);
      s.cz.push_back(s.pz[p] + zadd);
      s.c_prim.push_back(p);
      s.c_part.push_back(ic);
    }
    nm += nmg;
  }
  s.mx_mcover = nm;
}

// Lower bound of the squared distance from a point to cover partition ic.
inline double part_dist2(const System &s, int32_t ic, double x, double y, double z) {
  double d2 = 0.0;
  const double p[3] = {x, y, z};
  for (int d = 0; d < 3; ++d) {
    // Partition covers [ns*dc, (ns+1)*dc) where ns is its cover coordinate.
    const int32_t lab = s.lab_cover[ic] - 1;
    int32_t c;
    if (d == 0) c = lab / (s.ncover[1] * s.ncover[2]);
    else if (d == 1) c = (lab / s.ncover[2]) % s.ncover[1];
    else c = lab % s.ncover[2];
    const int32_t ns = c - s.nspanl[d];
    const double lo = ns * s.dcell[d] - 1e-9, hi = (ns + 1) * s.dcell[d] + 1e-9;
    double e = 0.0;
    if (p[d] < lo) e = lo - p[d];
    else if (p[d] > hi) e = p[d] - hi;
    d2 += e * e;
  }
  return d2;
}

// get_naba (src/matrix_elements_module.f90:257-384) for every primary atom:
// cover partitions in cover-set order, members in partition order, strict
// test d^2 < rcut^2 - RD_ERR.
void build_neighbours(const System &s, RangeIndex &r) {
  const int32_t N = s.natoms;
  const double rc2 = r.rcut * r.rcut;
  std::vector<std::vector<int32_t>> lists(N);
#pragma omp parallel for schedule(dynamic, 8)
  for (int32_t p = 0; p < N; ++p) {
    std::vector<int32_t> &l = lists[p];
    const double x = s.px[p], y = s.py[p], z = s.pz[p];
    for (int32_t ic = 0; ic < s.ng_cover; ++ic) {
      if (s.n_ing_cover[ic] == 0) continue;
      if (part_dist2(s, ic, x, y, z) > rc2 + 1e-6) continue;
      const int32_t m0 = s.icover_ibeg[ic] - 1;
      for (int32_t a = 0; a < s.n_ing_cover[ic]; ++a) {
        const double dx = s.cx[m0 + a] - x;
        const double dy = s.cy[m0 + a] - y;
        const double dz = s.cz[m0 + a] - z;
        if (dx * dx + dy * dy + dz * dz < rc2 - kRdErr) l.push_back(m0 + a);
      }
    }
  }
  r.n_nab.assign(N, 0);
  r.row_beg.assign(N + 1, 0);
  r.row_off.assign(N, 0);
  for (int32_t p = 0; p < N; ++p) {
    r.n_nab[p] = (int32_t)lists[p].size();
    r.row_beg[p + 1] = r.row_beg[p] + r.n_nab[p];
    r.max_nab = std::max(r.max_nab, r.n_nab[p]);
  }
  const int64_t E = r.row_beg[N];
  r.nb_cover.resize(E); r.nb_part.resize(E); r.nb_seq.resize(E);
  r.nb_ndimj.resize(E); r.nb_off.resize(E); r.nb_radius.resize(E);
  int64_t off = 0;
  for (int32_t p = 0; p < N; ++p) {
    r.row_off[p] = off;
    const int32_t nd1 = s.pnsf[p];
    int64_t e = r.row_beg[p];
    for (int32_t m : lists[p]) {
      const int32_t ic = s.c_part[m];
      r.nb_cover[e] = m;
      r.nb_part[e] = ic + 1;
      r.nb_seq[e] = m - (s.icover_ibeg[ic] - 1) + 1;
      r.nb_ndimj[e] = s.pnsf[s.c_prim[m]];
      r.nb_off[e] = off;
      const double dx = s.cx[m] - s.px[p], dy = s.cy[m] - s.py[p], dz = s.cz[m] - s.pz[p];
      r.nb_radius[e] = std::sqrt(dx * dx + dy * dy + dz * dz);
      off += (int64_t)nd1 * r.nb_ndimj[e];
      ++e;
    }
  }
  r.length = off;
}

// Partition-level type(matrix) fields, make_npxyz (src/matrix_elements_module.f90:836-903)
// and the packed index of set_matrix_pointers2 (src/matrix_module.f90:350-421).
void build_bindex(const System &s, RangeIndex &r) {
  const int32_t P = s.n_parts;
  r.part_nabs.assign(P, 0); r.n_atoms_part.assign(P, 0); r.array_posn.assign(P, 0);
  r.part_nd_nabs.assign(P, 0); r.nd_offset64.assign(P, 0); r.nd_offset32.assign(P, 0);
  r.i_acc.assign(s.natoms, 0); r.i_nd_acc.assign(s.natoms, 0);
  const int32_t ncyz = s.ncover[1] * s.ncover[2];
  int64_t nd_off = 0;
  r.bindex.clear();
  for (int32_t nn = 0; nn < P; ++nn) {
    const int32_t na = s.nm_group[nn];
    const int32_t p0 = s.icell_beg[nn];
    int32_t nabs = 0;
    int64_t ndn = 0;
    for (int32_t a = 0; a < na; ++a) {
      const int32_t p = p0 + a;
      r.i_acc[p] = nabs + 1;
      r.i_nd_acc[p] = (int32_t)(ndn + 1);
      nabs += r.n_nab[p];
      for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e) ndn += (int64_t)s.pnsf[p] * r.nb_ndimj[e];
    }
    r.n_atoms_part[nn] = na;
    r.part_nabs[nn] = nabs;
    r.part_nd_nabs[nn] = ndn;
    r.nd_offset64[nn] = nd_off;
    r.nd_offset32[nn] = (int32_t)nd_off;
    if (nd_off > INT32_MAX) r.fits32 = false;
    nd_off += ndn;
    // Packed slice: n_nab, i_acc, i_nd_acc (size1), i_seq (size2), npxyz (3*size2), ndimj (size2).
    r.array_posn[nn] = (int32_t)r.bindex.size() + 1;
    const int32_t size1 = na > 0 ? na : 1, size2 = nabs > 0 ? nabs : 1;
    const size_t base = r.bindex.size();
    r.bindex.resize(base + 3 * (size_t)size1 + 5 * (size_t)size2, 0);
    int32_t *nnab = &r.bindex[base];
    int32_t *iacc = nnab + size1;
    int32_t *indacc = iacc + size1;
    int32_t *iseq = indacc + size1;
    int32_t *npxyz = iseq + size2;
    int32_t *ndimj = npxyz + 3 * size2;
    if (na == 0) { iacc[0] = 1; indacc[0] = 1; }
    int32_t ist = 0;
    for (int32_t a = 0; a < na; ++a) {
      const int32_t p = p0 + a;
      nnab[a] = r.n_nab[p];
      iacc[a] = r.i_acc[p];
      indacc[a] = r.i_nd_acc[p];
      for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e, ++ist) {
        iseq[ist] = r.nb_seq[e];
        const int32_t ibpartp = s.lab_cover[r.nb_part[e] - 1];
        const int32_t nx = 1 + (ibpartp - 1) / ncyz;
        const int32_t ny = 1 + (ibpartp - 1 - (nx - 1) * ncyz) / s.ncover[2];
        const int32_t nz = ibpartp - (nx - 1) * ncyz - (ny - 1) * s.ncover[2];
        npxyz[3 * ist + 0] = nx - 1 - s.nspanl[0] - s.idisp[0][nn];
        npxyz[3 * ist + 1] = ny - 1 - s.nspanl[1] - s.idisp[1][nn];
        npxyz[3 * ist + 2] = nz - 1 - s.nspanl[2] - s.idisp[2][nn];
        ndimj[ist] = r.nb_ndimj[e];
      }
    }
  }
  if (r.bindex.size() > (size_t)INT32_MAX) r.fits32 = false;
}

// make_halo (src/matrix_elements_module.f90:560-660) and get_retr (:517-557).
void build_halo(const System &s, RangeIndex &r) {
  const int32_t N = s.natoms;
  r.i_halo.assign(s.mx_mcover, 0);
  for (int64_t e = 0; e < r.row_beg[N]; ++e) r.i_halo[r.nb_cover[e]] = 1;
  r.i_hbeg.assign(s.ng_cover, 0);
  r.nh_part.clear(); r.j_beg.clear(); r.lab_hcell.clear(); r.lab_hcover.clear();
  r.j_seq.clear(); r.ndimj_h.clear(); r.halo_cover.clear();
  r.np_in_halo = 0; r.ni_in_halo = 0;
  for (int32_t ic = 0; ic < s.ng_cover; ++ic) {
    r.i_hbeg[s.lab_cover[ic] - 1] = s.icover_ibeg[ic];
    int32_t in_part = 0;
    const int32_t m0 = s.icover_ibeg[ic] - 1;
    for (int32_t a = 0; a < s.n_ing_cover[ic]; ++a) {
      if (r.i_halo[m0 + a] == 1) {
        ++in_part;
        ++r.ni_in_halo;
        r.j_seq.push_back(a + 1);
        r.ndimj_h.push_back(s.pnsf[s.c_prim[m0 + a]]);
        r.halo_cover.push_back(m0 + a);
        r.i_halo[m0 + a] = r.ni_in_halo;
      }
    }
    if (in_part > 0) {
      r.j_beg.push_back(r.np_in_halo == 0 ? 1 : r.j_beg.back() + r.nh_part.back());
      ++r.np_in_halo;
      r.nh_part.push_back(in_part);
      r.lab_hcell.push_back(s.lab_cell[ic]);
      r.lab_hcover.push_back(s.lab_cover[ic]);
    }
  }
  r.ndimi.assign(s.pnsf.begin(), s.pnsf.end());
  const int64_t H = r.ni_in_halo;
  r.i_h2d64.assign((size_t)N * H, 0);
  for (int32_t p = 0; p < N; ++p)
    for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e)
      r.i_h2d64[(size_t)p * H + (r.i_halo[r.nb_cover[e]] - 1)] = r.nb_off[e] + 1;
  if (r.length + 1 > INT32_MAX || (int64_t)N * H > INT32_MAX) r.fits32 = false;
  if (r.fits32) {
    r.i_h2d32.resize(r.i_h2d64.size());
    for (size_t t = 0; t < r.i_h2d64.size(); ++t) r.i_h2d32[t] = (int32_t)r.i_h2d64[t];
  } else {
    r.i_h2d32.clear();
  }
  // get_retr: ascending primary atoms per halo atom.
  r.n_hnab.assign(H, 0);
  for (int64_t e = 0; e < r.row_beg[N]; ++e) ++r.n_hnab[r.i_halo[r.nb_cover[e]] - 1];
  r.i_beg.assign(H, 0);
  r.i_nd_beg64.assign(H, 0);
  std::vector<int64_t> cumnd(H, 0);
  for (int32_t p = 0; p < N; ++p)
    for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e) {
      const int32_t h = r.i_halo[r.nb_cover[e]] - 1;
      cumnd[h] += (int64_t)s.pnsf[p] * r.ndimj_h[h];
    }
  int64_t ib = 1, ind = 1;
  for (int64_t h = 0; h < H; ++h) {
    r.i_beg[h] = (int32_t)ib;
    r.i_nd_beg64[h] = ind;
    ib += r.n_hnab[h];
    ind += cumnd[h];
  }
  r.i_prim.assign((size_t)(ib - 1), 0);
  std::vector<int32_t> fill(H, 0);
  for (int32_t p = 0; p < N; ++p)
    for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e) {
      const int32_t h = r.i_halo[r.nb_cover[e]] - 1;
      r.i_prim[r.i_beg[h] - 1 + fill[h]++] = p + 1;
    }
  r.i_nd_beg32.resize(H);
  for (int64_t h = 0; h < H; ++h) r.i_nd_beg32[h] = (int32_t)r.i_nd_beg64[h];
}

double normal01(uint64_t key, uint64_t ctr) {
  const double u1 = cq::u01(key, 2 * ctr), u2 = cq::u01(key, 2 * ctr + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

uint64_t block_key(uint64_t seed, uint64_t tag, int32_t gi, int32_t gj, const int32_t *sh) {
  uint64_t h = cq::mix64(seed);
  h = cq::hash_combine(h, tag);
  h = cq::hash_combine(h, (uint64_t)(uint32_t)gi);
  h = cq::hash_combine(h, (uint64_t)(uint32_t)gj);
  for (int d = 0; d < 3; ++d) h = cq::hash_combine(h, (uint64_t)(uint32_t)(sh[d] + 1024));
  return h;
}

}  // namespace

namespace cq {

// Transpose pairs (p <= q, 1-based offsets) of range rid: block (i,j,s) and
// block (j,i,-s); p == NULL counts them only. Each row's blocks are sorted by
// (column atom, image shift); the partner of block e = (i, j, s) is found by a
// binary search for (i, -s) in the row of j. Pairs come out in storage order
// of their first block.
int transpose_pairs(const System &s, int32_t rid, int64_t *npairs, int64_t *p, int64_t *q, int32_t *n1, int32_t *n2) {
  std::vector<int64_t> vp, vq;
  std::vector<int32_t> va, vb;
  const int st = transpose_pairs(s, rid, vp, vq, va, vb);
  if (st) return st;
  *npairs = (int64_t)vp.size();
  if (p) {
    std::copy(vp.begin(), vp.end(), p);
    std::copy(vq.begin(), vq.end(), q);
    std::copy(va.begin(), va.end(), n1);
    std::copy(vb.begin(), vb.end(), n2);
  }
  return CQMM_OK;
}

int transpose_pairs(const System &s, int32_t rid, std::vector<int64_t> &p, std::vector<int64_t> &q,
                    std::vector<int32_t> &n1, std::vector<int32_t> &n2, std::vector<int64_t> *blocks) {
  const RangeIndex &r = s.ranges[rid];
  const int32_t na = s.natoms;
  const int64_t nb = r.row_beg[na];
  auto key = [](int32_t gj, const int32_t *sh) {
    return (((uint64_t)(uint32_t)gj * 64 + (uint64_t)(sh[0] + 32)) * 64 + (uint64_t)(sh[1] + 32)) * 64 +
           (uint64_t)(sh[2] + 32);
  };
  // Per row: (key of the block's column atom and shift, block), sorted by key.
  std::vector<std::pair<uint64_t, int64_t>> rowkey(nb);
#pragma omp parallel for schedule(dynamic, 16)
  for (int32_t pp = 0; pp < na; ++pp) {
    for (int64_t e = r.row_beg[pp]; e < r.row_beg[pp + 1]; ++e) {
      const int32_t m = r.nb_cover[e];
      rowkey[e] = {key(s.prim_gid[s.c_prim[m]], &s.cs_shift[3 * s.c_part[m]]), e};
    }
    std::sort(rowkey.begin() + r.row_beg[pp], rowkey.begin() + r.row_beg[pp + 1]);
  }
  // Partner of every block (-1: none).
  std::vector<int64_t> partner(nb, -1);
#pragma omp parallel for schedule(dynamic, 16)
  for (int32_t pp = 0; pp < na; ++pp) {
    const int32_t gi = s.prim_gid[pp];
    for (int64_t e = r.row_beg[pp]; e < r.row_beg[pp + 1]; ++e) {
      const int32_t m = r.nb_cover[e];
      const int32_t *sh = &s.cs_shift[3 * s.c_part[m]];
      const int32_t neg[3] = {-sh[0], -sh[1], -sh[2]};
      const int32_t pj = s.c_prim[m];  // the row of the column atom
      const uint64_t k = key(gi, neg);
      auto b = rowkey.begin() + r.row_beg[pj], en = rowkey.begin() + r.row_beg[pj + 1];
      auto it = std::lower_bound(b, en, std::make_pair(k, (int64_t)-1));
      if (it != en && it->first == k) partner[e] = it->second;
    }
  }
  int64_t n = 0;
  for (int64_t e = 0; e < nb; ++e) {
    const int64_t f = partner[e];
    if (f < 0) return cq::set_error(CQMM_ERR_RANGE, "cqmm_sys_transpose_pairs: pattern not symmetric");
    if (f < e) continue;
    ++n;
  }
  p.resize(n);
  q.resize(n);
  n1.resize(n);
  n2.resize(n);
  if (blocks) blocks->resize(2 * n);
  int64_t t = 0;
  for (int32_t pp = 0; pp < na; ++pp)
    for (int64_t e = r.row_beg[pp]; e < r.row_beg[pp + 1]; ++e) {
      const int64_t f = partner[e];
      if (f < e) continue;
      p[t] = r.nb_off[e] + 1;
      q[t] = r.nb_off[f] + 1;
      n1[t] = s.pnsf[pp];
      n2[t] = r.nb_ndimj[e];
      if (blocks) {
        (*blocks)[2 * t] = e;
        (*blocks)[2 * t + 1] = f;
      }
      ++t;
    }
  return CQMM_OK;
}

}  // namespace cq

extern "C" {

int cqmm_random_structure(int32_t natoms, double density, double dmin, uint64_t seed,
                          double *xyz, double *edge) {
  return cq::guarded(__func__, [&]() -> int {
    if (natoms <= 0 || density <= 0.0 || !(dmin >= 0.0) || !xyz)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_random_structure: bad argument");
    const double L = std::cbrt((double)natoms / density);
    const double d2min = dmin * dmin;
    // Cell list over the box for the minimum-image rejection test. It only
    // narrows which placed atoms are compared (every one within dmin is), so
    // its size, capped near the atom count for small dmin, does not change the
    // structure.
    const int32_t nc_cap = 2 * (int32_t)std::ceil(std::cbrt((double)natoms)) + 1;
    const int32_t nc = std::max(1, (int32_t)std::min<double>(nc_cap, std::floor(L / std::max(dmin, 1e-3))));
    const double cs = L / nc;
    std::vector<std::vector<int32_t>> cells((size_t)nc * nc * nc);
    uint64_t key = cq::hash_combine(cq::mix64(seed), 0x5eed5eedULL);
    uint64_t ctr = 0;
    for (int32_t i = 0; i < natoms; ++i) {
      int64_t tries = 0;
      for (;;) {
        if (++tries > 100000000LL) return cq::set_error(CQMM_ERR_ARG, "cqmm_random_structure: cannot place atoms at this density");
        const double x = cq::u01(key, ctr++) * L, y = cq::u01(key, ctr++) * L, z = cq::u01(key, ctr++) * L;
        const int32_t ix = std::min(nc - 1, (int32_t)(x / cs)), iy = std::min(nc - 1, (int32_t)(y / cs)),
                      iz = std::min(nc - 1, (int32_t)(z / cs));
        bool ok = true;
        const int32_t span = (int32_t)std::ceil(dmin / cs);
        for (int32_t ax = -span; ax <= span && ok; ++ax)
          for (int32_t ay = -span; ay <= span && ok; ++ay)
            for (int32_t az = -span; az <= span && ok; ++az) {
              const int32_t cx = ((ix + ax) % nc + nc) % nc, cy = ((iy + ay) % nc + nc) % nc,
                            cz = ((iz + az) % nc + nc) % nc;
              for (int32_t j : cells[((size_t)cx * nc + cy) * nc + cz]) {
                double dx = xyz[3 * j] - x, dy = xyz[3 * j + 1] - y, dz = xyz[3 * j + 2] - z;
                dx -= L * std::nearbyint(dx / L);
                dy -= L * std::nearbyint(dy / L);
                dz -= L * std::nearbyint(dz / L);
                if (dx * dx + dy * dy + dz * dz < d2min) { ok = false; break; }
              }
            }
        if (ok) {
          xyz[3 * i] = x; xyz[3 * i + 1] = y; xyz[3 * i + 2] = z;
          cells[((size_t)ix * nc + iy) * nc + iz].push_back(i);
          break;
        }
      }
    }
    if (edge) *edge = L;
    return CQMM_OK;
  });
}

int cqmm_sys_create(int32_t natoms, const double *xyz, const double box[3], const int32_t *nsf,
                    const int32_t ngrid[3], double rcut_cover, cqmm_system **out) {
  return cq::guarded(__func__, [&]() -> int {
    if (!out || natoms <= 0 || !xyz || !box || !nsf || !ngrid)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_create: null or empty argument");
    for (int d = 0; d < 3; ++d)
      if (ngrid[d] <= 0 || ngrid[d] > 4096 || !(box[d] > 0.0))
        return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_create: bad box or partition grid");
    if (!(rcut_cover > 0.0)) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_create: cover radius must be positive");
    for (int32_t i = 0; i < natoms; ++i)
      if (nsf[i] <= 0 || nsf[i] > 64) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_create: nsf[%d]=%d out of range", i, nsf[i]);
    cqmm_system *h = new cqmm_system();
    System &s = h->s;
    s.natoms = natoms;
    s.rcut_cover = rcut_cover;
    for (int d = 0; d < 3; ++d) {
      s.box[d] = box[d];
      s.ng[d] = ngrid[d];
      s.dcell[d] = box[d] / ngrid[d];
    }
    s.n_parts = s.ng[0] * s.ng[1] * s.ng[2];
    // Partition of each atom and primary order: CC label, then global id.
    std::vector<int32_t> part(natoms);
    for (int32_t i = 0; i < natoms; ++i) {
      int32_t q[3];
      for (int d = 0; d < 3; ++d) {
        double x = xyz[3 * i + d];
        if (!(x >= 0.0 && x < box[d])) {
          delete h;
          return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_create: atom %d outside the cell", i);
        }
        q[d] = std::min(s.ng[d] - 1, (int32_t)(x / s.dcell[d])) + 1;
      }
      part[i] = lab3(q[0], q[1], q[2], s.ng[1], s.ng[2]);
    }
    std::vector<int32_t> order(natoms);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return part[a] < part[b]; });
    s.prim_gid = order;
    s.px.resize(natoms); s.py.resize(natoms); s.pz.resize(natoms);
    s.pnsf.resize(natoms); s.prim_part.resize(natoms);
    for (int32_t p = 0; p < natoms; ++p) {
      const int32_t g = order[p];
      s.px[p] = xyz[3 * g]; s.py[p] = xyz[3 * g + 1]; s.pz[p] = xyz[3 * g + 2];
      s.pnsf[p] = nsf[g];
      s.prim_part[p] = part[g];
    }
    s.nm_group.assign(s.n_parts, 0);
    s.icell_beg.assign(s.n_parts, 0);
    for (int32_t p = 0; p < natoms; ++p) ++s.nm_group[s.prim_part[p] - 1];
    for (int32_t q = 1; q < s.n_parts; ++q) s.icell_beg[q] = s.icell_beg[q - 1] + s.nm_group[q - 1];
    for (int d = 0; d < 3; ++d) s.idisp[d].assign(s.n_parts, 0);
    for (int32_t q = 1; q <= s.n_parts; ++q) {
      const int32_t nx = 1 + (q - 1) / (s.ng[1] * s.ng[2]);
      const int32_t ny = 1 + (q - 1 - (nx - 1) * s.ng[1] * s.ng[2]) / s.ng[2];
      const int32_t nz = q - (nx - 1) * s.ng[1] * s.ng[2] - (ny - 1) * s.ng[2];
      s.idisp[0][q - 1] = nx - 1; s.idisp[1][q - 1] = ny - 1; s.idisp[2][q - 1] = nz - 1;
    }
    s.i_cc2seq.resize(s.n_parts);
    std::iota(s.i_cc2seq.begin(), s.i_cc2seq.end(), 1);
    build_cover(s);
    *out = h;
    return CQMM_OK;
  });
}

int cqmm_sys_destroy(cqmm_system *sys) {
  return cq::guarded(__func__, [&]() -> int {
    delete sys;
    return CQMM_OK;
  });
}

int cqmm_sys_add_range(cqmm_system *sys, double rcut, int32_t *range_id) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !range_id || !(rcut > 0.0)) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_add_range: bad argument");
    System &s = sys->s;
    if (rcut > s.rcut_cover + 1e-12)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_add_range: range %.6g exceeds the cover radius %.6g", rcut, s.rcut_cover);
    RangeIndex r;
    r.rcut = rcut;
    build_neighbours(s, r);
    build_bindex(s, r);
    build_halo(s, r);
    s.ranges.push_back(std::move(r));
    *range_id = (int32_t)s.ranges.size() - 1;
    return CQMM_OK;
  });
}

int cqmm_sys_range_info(const cqmm_system *sys, int32_t rid, cqmm_range_info *info) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !info || rid < 0 || rid >= (int32_t)sys->s.ranges.size())
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_range_info: bad argument");
    const RangeIndex &r = sys->s.ranges[rid];
    info->rcut = r.rcut;
    info->length = r.length;
    info->n_blocks = r.row_beg.back();
    info->ni_in_halo = r.ni_in_halo;
    info->np_in_halo = r.np_in_halo;
    info->max_nab = r.max_nab;
    info->fits_int32 = r.fits32 ? 1 : 0;
    return CQMM_OK;
  });
}

int cqmm_sys_get_info(const cqmm_system *sys, cqmm_sys_info *info) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !info) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_get_info: bad argument");
    const System &s = sys->s;
    info->natoms = s.natoms;
    info->n_parts = s.n_parts;
    info->ng_cover = s.ng_cover;
    info->mx_mcover = s.mx_mcover;
    for (int d = 0; d < 3; ++d) {
      info->ncover[d] = s.ncover[d];
      info->nspanl[d] = s.nspanl[d];
      info->ngrid[d] = s.ng[d];
      info->box[d] = s.box[d];
    }
    return CQMM_OK;
  });
}

static void fill_halo(const RangeIndex &r, cqmm_halo &h) {
  h.np_in_halo = r.np_in_halo;
  h.ni_in_halo = r.ni_in_halo;
  h.nh_part = r.nh_part.data();
  h.j_beg = r.j_beg.data();
  h.lab_hcell = r.lab_hcell.data();
  h.lab_hcover = r.lab_hcover.data();
  h.j_seq = r.j_seq.data();
  h.ndimj = r.ndimj_h.data();
  h.i_h2d = r.i_h2d32.data();
  h.i_halo = r.i_halo.data();
  h.len_i_halo = (int64_t)r.i_halo.size();
  h.i_hbeg = r.i_hbeg.data();
  h.len_i_hbeg = (int64_t)r.i_hbeg.size();
  h.ndimi = r.ndimi.data();
}

int cqmm_sys_mult_desc(const cqmm_system *sys, int32_t mult_type, int32_t ra, int32_t rb, int32_t rc,
                       cqmm_mult_desc *d) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !d) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_mult_desc: null argument");
    const System &s = sys->s;
    const int32_t nr = (int32_t)s.ranges.size();
    if (ra < 0 || ra >= nr || rb < 0 || rb >= nr || rc < 0 || rc >= nr)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_mult_desc: range id out of bounds");
    if (mult_type != 1 && mult_type != 2) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_mult_desc: mult_type must be 1 or 2");
    const RangeIndex &A = s.ranges[ra], &B = s.ranges[rb], &C = s.ranges[rc];
    if (!A.fits32 || !B.fits32 || !C.fits32)
      return cq::set_error(CQMM_ERR_RANGE, "cqmm_sys_mult_desc: a matrix is too long for the reference's int32 index arrays");
    if (A.rcut + B.rcut > s.rcut_cover + 1e-9)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_mult_desc: cover radius %.6g < r(A halo)+r(B) = %.6g",
                            s.rcut_cover, A.rcut + B.rcut);
    std::memset(d, 0, sizeof(*d));
    d->mult_type = mult_type;
    d->n_prim = s.natoms;
    d->ncovery = s.ncover[1];
    d->ncoverz = s.ncover[2];
    fill_halo(A, d->ahalo);
    fill_halo(C, d->chalo);
    d->ltrans.n_hnab = A.n_hnab.data();
    d->ltrans.i_beg = A.i_beg.data();
    d->ltrans.i_nd_beg = A.i_nd_beg32.data();
    d->ltrans.i_prim = A.i_prim.data();
    d->ltrans.len_i_prim = (int64_t)A.i_prim.size();
    d->bmat.n_parts = s.n_parts;
    d->bmat.array_posn = B.array_posn.data();
    d->bmat.n_atoms = B.n_atoms_part.data();
    d->bmat.part_nabs = B.part_nabs.data();
    d->bmat.nd_offset = B.nd_offset32.data();
    d->bmat.bindex = B.bindex.data();
    d->bmat.len_bindex = (int64_t)B.bindex.size();
    d->i_cc2seq = s.i_cc2seq.data();
    d->n_cell_parts = s.n_parts;
    d->lena = A.length;
    d->lenb = B.length;
    d->lenc = C.length;
    return CQMM_OK;
  });
}

int cqmm_sys_blocks(const cqmm_system *sys, int32_t rid, int32_t *prim, int32_t *gid_i, int32_t *gid_j,
                    int32_t *shift, int64_t *offset, int32_t *nd1, int32_t *nd2) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || rid < 0 || rid >= (int32_t)sys->s.ranges.size())
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_blocks: bad argument");
    const System &s = sys->s;
    const RangeIndex &r = s.ranges[rid];
    for (int32_t p = 0; p < s.natoms; ++p)
      for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e) {
        const int32_t m = r.nb_cover[e];
        const int32_t ic = s.c_part[m];
        if (prim) prim[e] = p;
        if (gid_i) gid_i[e] = s.prim_gid[p];
        if (gid_j) gid_j[e] = s.prim_gid[s.c_prim[m]];
        if (shift) for (int d = 0; d < 3; ++d) shift[3 * e + d] = s.cs_shift[3 * ic + d];
        if (offset) offset[e] = r.nb_off[e];
        if (nd1) nd1[e] = s.pnsf[p];
        if (nd2) nd2[e] = r.nb_ndimj[e];
      }
    return CQMM_OK;
  });
}

int cqmm_sys_matrix_index(const cqmm_system *sys, int32_t rid, cqmm_matrix_index *idx) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !idx || rid < 0 || rid >= (int32_t)sys->s.ranges.size())
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_matrix_index: bad argument");
    const System &s = sys->s;
    RangeIndex &r = const_cast<RangeIndex &>(s.ranges[rid]);
    if (r.nab_offset.empty()) {
      r.nab_offset.assign(s.n_parts, 0);
      int64_t acc = 0;
      for (int32_t nn = 0; nn < s.n_parts; ++nn) {
        r.nab_offset[nn] = (int32_t)acc;
        acc += r.part_nabs[nn];
      }
      r.onsite.assign(s.natoms, 0);
      for (int32_t p = 0; p < s.natoms; ++p)
        for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e) {
          const int32_t m = r.nb_cover[e], ic = s.c_part[m];
          if (s.c_prim[m] == p && s.cs_shift[3 * ic] == 0 && s.cs_shift[3 * ic + 1] == 0 && s.cs_shift[3 * ic + 2] == 0)
            r.onsite[p] = r.nb_off[e] + 1;
        }
    }
    idx->n_parts = s.n_parts;
    idx->n_atoms = r.n_atoms_part.data();
    idx->nab_offset = r.nab_offset.data();
    idx->nd_offset = r.nd_offset64.data();
    idx->n_prim = s.natoms;
    idx->n_nab = r.n_nab.data();
    idx->i_acc = r.i_acc.data();
    idx->i_nd_acc = r.i_nd_acc.data();
    idx->ndimi = r.ndimi.data();
    idx->onsite = r.onsite.data();
    idx->i_part = r.nb_part.data();
    idx->i_seq = r.nb_seq.data();
    idx->ndimj = r.nb_ndimj.data();
    idx->n_nabs = (int64_t)r.nb_part.size();
    idx->length = r.length;
    return CQMM_OK;
  });
}

int cqmm_sys_transpose_pairs(const cqmm_system *sys, int32_t rid, int64_t *npairs, int64_t *p, int64_t *q,
                             int32_t *n1, int32_t *n2) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !npairs || rid < 0 || rid >= (int32_t)sys->s.ranges.size())
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_transpose_pairs: bad argument");
    return cq::transpose_pairs(sys->s, rid, npairs, p, q, n1, n2);
  });
}


int cqmm_sys_primary_order(const cqmm_system *sys, int32_t *prim_gid) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !prim_gid) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_primary_order: bad argument");
    std::copy(sys->s.prim_gid.begin(), sys->s.prim_gid.end(), prim_gid);
    return CQMM_OK;
  });
}

int cqmm_sys_fill(const cqmm_system *sys, int32_t rid, int32_t kind, uint64_t seed, uint64_t tag,
                  const double *prm, int32_t np, double *values) {
  return cq::guarded(__func__, [&]() -> int {
    if (!sys || !values || rid < 0 || rid >= (int32_t)sys->s.ranges.size())
      return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_fill: bad argument");
    const int32_t need = kind == 0 ? 1 : kind == 1 ? 2 : kind == 2 ? 4 : 0;
    if (kind < 0 || kind > 3) return cq::set_error(CQMM_ERR_ARG, "cqmm_sys_fill: unknown kind");
    if (np < need || (need > 0 && !prm)) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_sys_fill: kind %d needs %d parameters", kind, need);
    const System &s = sys->s;
    const RangeIndex &r = s.ranges[rid];
    // Onsite energies and the occupied half (kind 2).
    std::vector<double> eps;
    double eps_cut = 0.0;
    std::vector<int64_t> orb0;
    if (kind == 2) {
      orb0.assign(s.natoms + 1, 0);
      for (int32_t p = 0; p < s.natoms; ++p) orb0[p + 1] = orb0[p] + s.pnsf[p];
      eps.resize(orb0[s.natoms]);
      for (int32_t p = 0; p < s.natoms; ++p) {
        const int32_t g = s.prim_gid[p];
        const int32_t z[3] = {0, 0, 0};
        const uint64_t key = block_key(seed, tag ^ 0x0e5e0e5eULL, g, g, z);
        for (int32_t a = 0; a < s.pnsf[p]; ++a) eps[orb0[p] + a] = prm[2] * (2.0 * cq::u01(key, (uint64_t)a) - 1.0);
      }
      std::vector<double> sorted(eps);
      std::sort(sorted.begin(), sorted.end());
      const size_t nocc = sorted.size() / 2;
      eps_cut = nocc < sorted.size() ? sorted[nocc] : 0.0;  // orbitals with eps < eps_cut are occupied
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int32_t p = 0; p < s.natoms; ++p) {
      const int32_t gi = s.prim_gid[p];
      const int32_t nd1 = s.pnsf[p];
      for (int64_t e = r.row_beg[p]; e < r.row_beg[p + 1]; ++e) {
        const int32_t m = r.nb_cover[e];
        const int32_t ic = s.c_part[m];
        const int32_t pj = s.c_prim[m];
        const int32_t gj = s.prim_gid[pj];
        const int32_t *sh = &s.cs_shift[3 * ic];
        const int32_t nsh[3] = {-sh[0], -sh[1], -sh[2]};
        const int32_t nd2 = r.nb_ndimj[e];
        double *blk = values + r.nb_off[e];
        const bool onsite = (gi == gj && sh[0] == 0 && sh[1] == 0 && sh[2] == 0);
        // Canonical separation: antisymmetric under (i,j,s) -> (j,i,-s) bit for bit.
        double rr = 0.0;
        const double pi[3] = {s.px[p], s.py[p], s.pz[p]}, pj3[3] = {s.px[pj], s.py[pj], s.pz[pj]};
        for (int d = 0; d < 3; ++d) {
          const double dd = (pj3[d] - pi[d]) + sh[d] * s.box[d];
          rr += dd * dd;
        }
        rr = std::sqrt(rr);
        const uint64_t kf = block_key(seed, tag, gi, gj, sh);
        const uint64_t kb = block_key(seed, tag, gj, gi, nsh);
        for (int32_t b = 0; b < nd2; ++b)
          for (int32_t a = 0; a < nd1; ++a) {
            double v = 0.0;
            if (kind == 0) {
              v = normal01(kf, (uint64_t)a * 1024 + b) * std::exp(-rr / prm[0]);
            } else if (kind == 3) {
              v = (onsite && a == b) ? 1.0 : 0.0;
            } else if (onsite && kind == 1) {
              v = (a == b) ? 1.0 : 0.0;
            } else if (onsite && kind == 2) {
              if (a == b) {
                const double ea = eps[orb0[p] + a];
                v = ea + (ea < eps_cut ? -prm[3] : prm[3]);
              }
            } else {
              const double g = 0.5 * (normal01(kf, (uint64_t)a * 1024 + b) + normal01(kb, (uint64_t)b * 1024 + a));
              v = prm[0] * g * std::exp(-rr / prm[1]);
            }
            blk[a + (int64_t)nd1 * b] = v;
          }
      }
    }
    return CQMM_OK;
  });
}

}  // extern "C"
