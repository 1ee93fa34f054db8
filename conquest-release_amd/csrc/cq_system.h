// Internal representation of a single-process CONQUEST index set (host side).
// The arrays hold the reference's Fortran values (1-based, 0 = absent) so that
// they can be handed to cqmm_plan_create exactly as a Fortran host would.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/cqmm_system.h"

namespace cq {

// One matrix range: type(matrix) per partition + halo + transpose index.
struct RangeIndex {
  double rcut = 0.0;
  // Rows (primary atoms, primary order).
  std::vector<int32_t> n_nab;       // [n_prim]
  std::vector<int64_t> row_beg;     // [n_prim+1] first neighbour of the row
  std::vector<int64_t> row_off;     // [n_prim] 0-based data offset of the row
  // Neighbours (storage order).
  std::vector<int32_t> nb_cover;    // 0-based cover-set member
  std::vector<int32_t> nb_part;     // cover-set partition np (1-based, CS order)
  std::vector<int32_t> nb_seq;      // sequence number in partition (1-based)
  std::vector<int32_t> nb_ndimj;
  std::vector<int64_t> nb_off;      // 0-based data offset of the block
  std::vector<double> nb_radius;
  int64_t length = 0;
  int32_t max_nab = 0;
  // Partition-level type(matrix) fields and the packed index (set_matrix_pointers2).
  std::vector<int32_t> part_nabs, n_atoms_part, array_posn;
  std::vector<int64_t> part_nd_nabs, nd_offset64;
  std::vector<int32_t> nd_offset32, i_acc, i_nd_acc;  // i_acc/i_nd_acc per primary atom, partition-relative
  std::vector<int32_t> bindex;
  // Halo (make_halo).
  int32_t np_in_halo = 0, ni_in_halo = 0;
  std::vector<int32_t> nh_part, j_beg, lab_hcell, lab_hcover;
  std::vector<int32_t> j_seq, ndimj_h, halo_cover;  // per halo atom
  std::vector<int32_t> i_halo;                      // [mx_mcover]
  std::vector<int32_t> i_hbeg;                      // [ng_cover] by lab_cover
  std::vector<int32_t> ndimi;                       // [n_prim]
  std::vector<int64_t> i_h2d64;                     // [n_prim*ni_in_halo], 1-based offsets
  std::vector<int32_t> i_h2d32;                     // same, when it fits int32
  // Local transpose index (get_retr).
  std::vector<int32_t> n_hnab, i_beg, i_prim;
  std::vector<int64_t> i_nd_beg64;
  std::vector<int32_t> i_nd_beg32;
  // Flattened type(matrix) for cqmm_matrix_index.
  std::vector<int32_t> nab_offset;
  std::vector<int64_t> onsite;
  bool fits32 = true;
};

struct System {
  int32_t natoms = 0;
  double box[3] = {0, 0, 0};
  int32_t ng[3] = {1, 1, 1};
  double dcell[3] = {0, 0, 0};
  double rcut_cover = 0.0;
  // Primary atoms in partition order.
  std::vector<int32_t> prim_gid;     // global id
  std::vector<double> px, py, pz;    // cell coordinates
  std::vector<int32_t> pnsf;
  std::vector<int32_t> prim_part;    // 1-based cell partition (CC label) of each primary atom
  // Cell partitions (CC label order = node order = primary-set order).
  int32_t n_parts = 0;
  std::vector<int32_t> nm_group, icell_beg;  // icell_beg 0-based into primary arrays
  std::vector<int32_t> idisp[3];
  std::vector<int32_t> i_cc2seq;
  // Cover set (make_cs).
  int32_t ncover[3] = {0, 0, 0}, nspanl[3] = {0, 0, 0};
  int32_t ng_cover = 0, mx_mcover = 0;
  std::vector<int32_t> lab_cell, lab_cover, n_ing_cover, icover_ibeg;  // per CS partition (CS order)
  std::vector<int32_t> cs_shift;     // [3*ng_cover] image shift in cells
  std::vector<double> cx, cy, cz;    // member coordinates
  std::vector<int32_t> c_prim;       // member -> primary atom of its cell copy
  std::vector<int32_t> c_part;       // member -> CS partition (0-based)
  // Ranges.
  std::vector<RangeIndex> ranges;
};

// Transpose pairs (p <= q, 1-based offsets) of range rid of the system (the
// work of cqmm_sys_transpose_pairs); p == NULL counts them only.
int transpose_pairs(const System &s, int32_t rid, int64_t *npairs, int64_t *p, int64_t *q, int32_t *n1, int32_t *n2);
// The same into vectors; blocks (if given) receives the two blocks' indices
// in storage order of each pair.
int transpose_pairs(const System &s, int32_t rid, std::vector<int64_t> &p, std::vector<int64_t> &q,
                    std::vector<int32_t> &n1, std::vector<int32_t> &n2, std::vector<int64_t> *blocks = nullptr);

}  // namespace cq

struct cqmm_system {
  cq::System s;
};
