/*
 * cqmm_system.h — single-process mirror of CONQUEST's sparsity indexing, and
 * the synthetic matrices of the benchmark workloads.
 *
 * CONQUEST builds its index arrays in Fortran at start-up: the cover set
 * (make_cs, src/cover_module.f90:133-429), neighbour lists per range
 * (matrix_ini/get_naba, src/matrix_elements_module.f90:67-142,257-384), halos
 * (make_halo, :560-660), local-transpose indices (get_retr, :517-557) and
 * partition offsets (make_npxyz, :836-903), then the mult table (immi,
 * src/mult_module.f90:271-1102). A Fortran host hands those arrays to
 * cqmm_plan_create unchanged. This header lets a host without CONQUEST (tests,
 * benchmarks, the request loop) produce the same arrays for one process from
 * atom positions, and reads them back as a cqmm_mult_desc.
 *
 * Units: lengths in Angstrom; the neighbour test is d^2 < rcut^2 - 1e-8
 * (RD_ERR, src/numbers_module.f90:96) in Angstrom^2.
 */
#ifndef CQMM_SYSTEM_H
#define CQMM_SYSTEM_H

#include <stdint.h>
#include "cqmm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cqmm_system cqmm_system;

/* Atoms: xyz[3*natoms] cell coordinates in [0,box), nsf[natoms] support
 * functions per atom, ngrid[3] partitions per cell edge (regular grid, raster
 * order), rcut_cover the cover-set radius (it must be at least r(A-halo)+r(B)
 * of every product planned on the system). Atom i keeps global id i. */
int cqmm_sys_create(int32_t natoms, const double *xyz, const double box[3],
                    const int32_t *nsf, const int32_t ngrid[3],
                    double rcut_cover, cqmm_system **sys);
int cqmm_sys_destroy(cqmm_system *sys);

/* matrix_ini for one range: neighbour lists, halo and transpose index. */
int cqmm_sys_add_range(cqmm_system *sys, double rcut, int32_t *range_id);

typedef struct cqmm_range_info {
  double rcut;
  int64_t length;              /* doubles in the matrix                       */
  int64_t n_blocks;            /* stored blocks                               */
  int32_t ni_in_halo;          /* atoms in the halo                           */
  int32_t np_in_halo;          /* partitions in the halo                      */
  int32_t max_nab;             /* largest row                                 */
  int32_t fits_int32;          /* 1 when every offset fits the reference's int32 arrays */
} cqmm_range_info;
int cqmm_sys_range_info(const cqmm_system *sys, int32_t range_id,
                        cqmm_range_info *info);

typedef struct cqmm_sys_info {
  int32_t natoms, n_parts, ng_cover, mx_mcover;
  int32_t ncover[3], nspanl[3], ngrid[3];
  double box[3];
} cqmm_sys_info;
int cqmm_sys_get_info(const cqmm_system *sys, cqmm_sys_info *info);

/* Fill a cqmm_mult_desc for product `mult_type` with operand ranges as
 * CONQUEST's immi assigns them (src/mult_module.f90:583-872): type 1 computes
 * C(rc) = A(ra).B(rb); type 2 computes A(ra) = C(rc).B(rb)^T (mat_mult's a is
 * the result). The pointers stay valid while the system lives. */
int cqmm_sys_mult_desc(const cqmm_system *sys, int32_t mult_type, int32_t ra,
                       int32_t rb, int32_t rc, cqmm_mult_desc *desc);

/* Plan of a product straight from the system's indices (64-bit block offsets,
 * for matrices longer than the reference's int32 arrays can address); same
 * products and order as cqmm_plan_create on cqmm_sys_mult_desc. */
int cqmm_sys_plan_create(const cqmm_system *sys, int32_t mult_type, int32_t ra, int32_t rb, int32_t rc,
                         cqmm_plan **plan);
/* The same as cqmm_plan_create_symmetric with the transpose pairs of the
 * result's range (rc for type 1, ra for type 2). */
int cqmm_sys_plan_create_symmetric(const cqmm_system *sys, int32_t mult_type, int32_t ra, int32_t rb, int32_t rc,
                                   cqmm_plan **plan);

/* type(matrix) index of a range as cqmm_ctx_add_range takes it (pointers stay
 * valid while the system lives). Requires cqmm_matrix.h for the struct. */
struct cqmm_matrix_index;
int cqmm_sys_matrix_index(const cqmm_system *sys, int32_t range_id, struct cqmm_matrix_index *idx);
/* Transpose pairs (p <= q, 1-based offsets) of a range: block (i,j,s) and block
 * (j,i,-s). Pass NULL arrays to get the count in *npairs first. */
int cqmm_sys_transpose_pairs(const cqmm_system *sys, int32_t range_id, int64_t *npairs, int64_t *p, int64_t *q,
                             int32_t *n1, int32_t *n2);

/* Block list of a range in storage order: primary atom (0-based, primary
 * order), global ids of the row and column atoms, the column atom's periodic
 * image shift in cells, the 0-based data offset, and block dimensions. Any
 * output pointer may be NULL. */
int cqmm_sys_blocks(const cqmm_system *sys, int32_t range_id, int32_t *prim,
                    int32_t *gid_i, int32_t *gid_j, int32_t *shift,
                    int64_t *offset, int32_t *nd1, int32_t *nd2);
/* Primary order: prim_gid[p] is the global id of primary atom p. */
int cqmm_sys_primary_order(const cqmm_system *sys, int32_t *prim_gid);

/* Synthetic block values (counter-based, independent of enumeration order):
 *  kind 0: entry = N(0,1) * exp(-r_ij / p[0])
 *  kind 1: overlap S = I on the onsite block, symmetrised p[0]*N(0,1)*exp(-r/p[1])
 *          on every other block
 *  kind 2: Hamiltonian H: onsite block diagonal eps_a ~ U(-p[2],p[2]) Ha,
 *          shifted by -p[3] for the half of all orbitals with the lowest eps
 *          and by +p[3] for the rest (opens a gap at half filling); every other
 *          block symmetrised p[0]*N(0,1)*exp(-r/p[1])
 *  kind 3: identity
 * `tag` separates matrices drawn with the same seed. values[length]. */
int cqmm_sys_fill(const cqmm_system *sys, int32_t range_id, int32_t kind,
                  uint64_t seed, uint64_t tag, const double *p, int32_t np,
                  double *values);

/* ---- Matrix files (CONQUEST's store_matrix format, one process) ----
 * File name of dump_matrix2 / grab_matrix2 (get_file_name_2rank,
 * src/io_module.f90:3134-3164): <root>matrix2.i<index:2>.p<inode:6>. */
int cqmm_matrix2_file_name(const char *root, int32_t index, int32_t inode, char *out, int32_t out_len);
/* dump_matrix2 (src/store_matrix_module.f90:340-470) of range `range_id`:
 * header, per primary atom its global id (1-based), neighbours' support
 * functions and global ids, (R_i - R_j image)/cell per neighbour, and the
 * row's values for each of `nspin` arrays. binary: Fortran unformatted
 * sequential records (4-byte markers); else list-directed text, values e30.20. */
int cqmm_sys_dump_matrix2(const cqmm_system *sys, int32_t range_id, int32_t nspin, const double *const *values,
                          const char *path, int32_t binary);
/* Read such a file into range `range_id` (outputs are not cleared first):
 * every file block whose atoms i, j and image (j at R_i - vec*cell) form a
 * block of the range is copied there, as UpdateMatrix_local does
 * (src/UpdateInfo_module.f90:1851-2162) for unmoved atoms; other file blocks
 * are skipped and counted. */
int cqmm_sys_load_matrix2(const cqmm_system *sys, int32_t range_id, const char *path, int32_t binary,
                          int32_t nspin_max, double *const *out, int32_t *nspin, int64_t *blocks_loaded,
                          int64_t *blocks_skipped);

/* Global atom information written beside the matrix files
 * (dump_InfoMatGlobal, src/store_matrix_module.f90:700-781; read back as
 * grab_InfoMatGlobal does, :937-1040). */
typedef struct cqmm_info_global {
  int32_t natoms, numprocs;
  int32_t npcell[3];     /* partitions along a, b, c */
  double rcell[3];       /* cell lengths */
  int32_t md_step;
  int32_t has_velocity, has_md_step;
  int32_t index;         /* the index written last in the file */
} cqmm_info_global;
/* File name InfoGlobal.i<index:2> (get_file_name_2rank without a process
 * number, src/io_module.f90:3134-3164). */
int cqmm_info_global_file_name(int32_t index, char *out, int32_t out_len);
/* Write the file for this system as one process: flags, atom and process
 * counts, partitions, cell (3f25.15), glob_to_node (all 1), MD step,
 * coordinates by global id as (5x,i8,3x,3e22.13) records, velocities (same
 * records, after a "velocity" line) when `velocity` (3*natoms, by global id)
 * is not NULL, and `index`. has_md_step = 0 writes MD step 0 with the flag F. */
int cqmm_sys_dump_info_global(const cqmm_system *sys, const char *path, int32_t index, int32_t has_md_step,
                              int32_t md_step, const double *velocity);
/* Read such a file (list-directed lines as a Fortran READ takes them,
 * including r*c repeats; coordinate records by column). Fails, as
 * grab_InfoMatGlobal aborts, when the index at its end is not `index`, and
 * when it holds more than natoms_max atoms and an output array is given.
 * glob_to_node[natoms], coords[3*natoms] and velocity[3*natoms] may each be
 * NULL (all three NULL: a sizing call that fills `info` only); velocities are
 * written only when the file holds them. */
int cqmm_load_info_global(const char *path, int32_t index, int32_t natoms_max, cqmm_info_global *info,
                          int32_t *glob_to_node, double *coords, double *velocity);

/* The older per-process dump of a matrix's storage array (dump_matrix /
 * grab_matrix, src/io_module.f90:1960-2075): one value per line as f25.18,
 * in storage order. File name <stub>matrix.<inode zero-padded to
 * max(3, digits of numprocs)> (get_file_name, src/io_module.f90:3071-3105). */
int cqmm_matrix_file_name(const char *stub, int32_t numprocs, int32_t inode, char *out, int32_t out_len);
int cqmm_dump_matrix_text(const char *path, const double *values, int64_t len);
/* Reads `len` values; fails when the file holds fewer or a line is not a
 * number (a value too wide for f25.18 is written as asterisks). */
int cqmm_grab_matrix_text(const char *path, double *values, int64_t len);

/* Synthetic structure of the workloads: natoms uniform in a cubic periodic box
 * of edge (natoms/density)^(1/3) with minimum-image separation >= dmin
 * (rejection sampling, seeded). Writes xyz[3*natoms]; returns the edge. */
int cqmm_random_structure(int32_t natoms, double density, double dmin,
                          uint64_t seed, double *xyz, double *edge);

#ifdef __cplusplus
}
#endif
#endif /* CQMM_SYSTEM_H */
