/*
 * cqmm.h — C ABI of the MI355X block-sparse multiply library (libcqmm.so).
 *
 * The library is a drop-in for CONQUEST's sparse product
 *   mat_mult(myid, a, lena, b, lenb, c, lenc, a_b_c)     (src/multiply_module.f90:115)
 * and the kernel seam under it,
 *   m_kern_max / m_kern_min                               (src/multiply_kernel_default.f90:132,325)
 * Every index array below keeps the reference's meaning and its Fortran values
 * (1-based offsets and labels, 0 meaning "absent"); only the storage is a plain
 * C pointer. All integer index arrays are int32, as `integer(integ)` is in the
 * reference (src/datatypes_module.f90).
 *
 * Conventions
 *   - Every function returns CQMM_OK (0) or a negative status; cqmm_last_error()
 *     returns the message of the last failure on the calling thread. The Fortran
 *     shim turns a nonzero status into cq_abort (src/generic_comms.f90:95-100).
 *   - Matrix data pointers are device pointers when CQMM_DEVICE_PTRS is set in
 *     `flags`, host pointers otherwise (the library then stages through HBM).
 *   - A plan is thread-compatible, not thread-safe: one caller at a time per plan.
 */
#ifndef CQMM_H
#define CQMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CQMM_OK 0
#define CQMM_ERR_ARG -1      /* invalid argument or inconsistent descriptor  */
#define CQMM_ERR_HIP -2      /* HIP runtime error                            */
#define CQMM_ERR_NOMEM -3    /* device or host allocation failed             */
#define CQMM_ERR_RANGE -4    /* an index points outside its array            */
#define CQMM_ERR_STATE -5    /* call out of order (e.g. LIFO free violated)  */
#define CQMM_ERR_NOTCONV -6  /* iterative solve stopped without converging   */

#define CQMM_DEVICE_PTRS 1u  /* a, b, c are device pointers                  */
#define CQMM_FP32 2u         /* run the block products in FP32 (MFMA, FP32 accumulate) */
#define CQMM_NO_SYNC 4u      /* return without synchronising the plan stream */
#define CQMM_FP32_ACC64 8u   /* with CQMM_FP32: sum each block's products in FP64 */
#define CQMM_FP32_RUN64 16u  /* with CQMM_FP32: sum a block's products within one chunk of the
                                tile schedule (at most 15) in FP32, the chunks in FP64 */
#define CQMM_BF16X3 32u      /* with CQMM_FP32: split-BF16 x3 products (each operand split into
                                two BF16 halves, hi.hi + hi.lo + lo.hi in BF16 MFMA with FP32
                                accumulation per block product, FP64 block sums); needs a tile
                                schedule of 9x9 to 16x16 blocks or class tiles
                                (CQMM_MIXED_TILES=class), CQMM_ERR_STATE otherwise */

/* type(matrix_halo), src/matrix_module.f90:121-136, built by make_halo
 * (src/matrix_elements_module.f90:560-660). */
typedef struct cqmm_halo {
  int32_t np_in_halo;          /* partitions in halo                          */
  int32_t ni_in_halo;          /* atoms in halo                               */
  const int32_t *nh_part;      /* [np_in_halo] atoms of each halo partition   */
  const int32_t *j_beg;        /* [np_in_halo] first halo atom of partition   */
  const int32_t *lab_hcell;    /* [np_in_halo] simulation-cell partition label */
  const int32_t *lab_hcover;   /* [np_in_halo] cover-set partition label      */
  const int32_t *j_seq;        /* [ni_in_halo] sequence number in partition   */
  const int32_t *ndimj;        /* [ni_in_halo] support functions of halo atom */
  const int32_t *i_h2d;        /* [n_prim*ni_in_halo] block offset of (i,h), 0 if absent */
  const int32_t *i_halo;       /* [len_i_halo] halo number of cover-set member, 0 if none */
  const int32_t *i_hbeg;       /* [len_i_hbeg] start of cover partition (by lab_cover) in i_halo */
  const int32_t *ndimi;        /* [n_prim] support functions of primary atom  */
  int64_t len_i_halo;
  int64_t len_i_hbeg;
} cqmm_halo;

/* type(matrix_trans), src/matrix_module.f90:97-104, built by get_retr
 * (src/matrix_elements_module.f90:517-557). Sized by the A halo. */
typedef struct cqmm_trans {
  const int32_t *n_hnab;       /* [ni_in_halo] primary neighbours of halo atom */
  const int32_t *i_beg;        /* [ni_in_halo] start in i_prim                */
  const int32_t *i_nd_beg;     /* [ni_in_halo] start of the atom's data in the local transpose */
  const int32_t *i_prim;       /* [len_i_prim] primary atom of each pair      */
  int64_t len_i_prim;
} cqmm_trans;

/* The B operand's index: type(matrix) of every primary-set partition
 * (src/matrix_module.f90:63-82) and the packed index array `bindex` laid out
 * by set_matrix_pointers2 (src/matrix_module.f90:350-421):
 *   n_nab(n_atoms) i_acc(n_atoms) i_nd_acc(n_atoms) i_seq(nabs) npxyz(3*nabs) ndimj(nabs)
 * with empty slices counted as length 1. */
typedef struct cqmm_bindex {
  int32_t n_parts;             /* groups_on_node                              */
  const int32_t *array_posn;   /* [n_parts] 1-based start of the partition in bindex */
  const int32_t *n_atoms;      /* [n_parts]                                   */
  const int32_t *part_nabs;    /* [n_parts]                                   */
  const int32_t *nd_offset;    /* [n_parts] data offset of the partition in b */
  const int32_t *bindex;       /* [len_bindex]                                */
  int64_t len_bindex;
} cqmm_bindex;

/* Flattened type(matrix_mult) (src/matrix_module.f90:193-203) for one process. */
typedef struct cqmm_mult_desc {
  int32_t mult_type;           /* 1 = extension (m_kern_max), 2 = reduction (m_kern_min) */
  int32_t n_prim;              /* primary-set atoms                           */
  int32_t ncovery, ncoverz;    /* gcs%ncovery, gcs%ncoverz                     */
  cqmm_halo ahalo;             /* A halo (type 1) / result halo (type 2)      */
  cqmm_halo chalo;             /* C halo (type 1) / left-operand halo (type 2) */
  cqmm_trans ltrans;           /* local transpose index of the A halo         */
  cqmm_bindex bmat;            /* B index                                     */
  const int32_t *i_cc2seq;     /* parts%i_cc2seq: cell partition label -> primary-set sequence */
  int32_t n_cell_parts;        /* length of i_cc2seq                          */
  int64_t lena, lenb, lenc;    /* lengths of a, b, c as mat_mult receives them */
} cqmm_mult_desc;

typedef struct cqmm_plan cqmm_plan;

/* ---- library state ---------------------------------------------------- */
const char *cqmm_last_error(void);
/* Select the HIP device and create the library stream. No reference counterpart. */
int cqmm_init(int device);
int cqmm_finalize(void);
/* Device memory helpers so that Fortran/C hosts need no HIP headers. */
int cqmm_device_alloc(void **ptr, int64_t bytes);
int cqmm_device_free(void *ptr);
int cqmm_memcpy_h2d(void *dst, const void *src, int64_t bytes);
int cqmm_memcpy_d2h(void *dst, const void *src, int64_t bytes);
/* Device-to-device copy of n doubles by a streaming kernel on the library
 * stream (reads 8n bytes, writes 8n bytes; also the reference kernel for
 * scaling HBM traffic counters). Asynchronous. */
int cqmm_device_copy(double *dst, const double *src, int64_t n);
int cqmm_synchronize(void);
/* HIP-event timer on the library stream: the device-side counterpart of the
 * tmr_std_matmult clock around mat_mult (src/multiply_module.f90:171,356,
 * src/timer_stdclocks_module.f90:39-54). stop waits for the stream and returns
 * the elapsed milliseconds between the two events. */
int cqmm_timer_start(void);
int cqmm_timer_stop(float *ms);

/* ---- the drop-in multiply -------------------------------------------- */
/* Validates the descriptor and builds the device (i,k,j) triplet list that
 * m_kern_max / m_kern_min enumerate (src/multiply_kernel_default.f90:169-221,
 * 362-409), with the partition offsets of end_part_comms
 * (src/matrix_comms_module.f90:67-103) and the B-partition slices of
 * prefetch / Mquest_get_local (src/multiply_module.f90:568-623,
 * src/matrix_comms_module.f90:130-169). Reused for every product with the
 * same patterns (CONQUEST's mult(x) table entry, src/mult_module.f90:271-1102).
 * Block dimensions (support functions per atom) 1..64; CQMM_ERR_ARG outside. */
int cqmm_plan_create(const cqmm_mult_desc *desc, cqmm_plan **plan);
/* A plan for a product whose result is symmetric (no reference counterpart:
 * CONQUEST computes every block; the caller asserts the symmetry, e.g. LS.L =
 * LSL and LSL.(LS)^T = LSLSL for symmetric L and S). The npairs transpose
 * pairs of the result's blocks (1-based element offsets p, q of blocks (i,j,s)
 * and (j,i,-s), n1 x n2 the dimensions of block p, as cqmm_ctx_set_transpose
 * takes them) must all be blocks of the result. Of each pair with p != q only
 * block p is computed; every product through the plan then writes block q as
 * its transpose. Half the block products for a result that differs from the
 * full product only by rounding (the mirrored blocks are transposes bit for
 * bit). cqmm_sys_plan_create_symmetric orders each pair so that the computed
 * blocks fill whole tiles of the schedule. */
int cqmm_plan_create_symmetric(const cqmm_mult_desc *desc, int64_t npairs, const int64_t *p, const int64_t *q,
                               const int32_t *n1, const int32_t *n2, cqmm_plan **plan);
int cqmm_plan_destroy(cqmm_plan *plan);

/* mat_mult (src/multiply_module.f90:115-359) for one process.
 * Type 1: c = a . b (c overwritten, a and b read).
 * Type 2: a = c . b^T in the reference's transposed indexing (a overwritten,
 *         b and c read), as matrix_product calls mat_mult(C,B,A)
 *         (src/mult_module.f90:2591-2601). */
int cqmm_mat_mult(cqmm_plan *plan, double *a, int64_t lena, const double *b,
                  int64_t lenb, double *c, int64_t lenc, uint32_t flags);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP stream capture, serving loops that forbid hipMalloc).
 * cqmm_mat_mult with CQMM_FP32 (and CQMM_FP32_ACC64 / CQMM_FP32_RUN64 as
 * given) on device pointers (CQMM_DEVICE_PTRS required), the same kernels and
 * results, but the FP32 operand and result copies live in one static device
 * array of cqmm_noalloc_scratch_bytes() (64 MiB) per device instead of
 * buffers the plan allocates: a product whose FP32 copies (lengths of both
 * operands and the result, each rounded up to 64 elements) exceed it is
 * refused with CQMM_ERR_ARG. Calls no hipMalloc/hipFree. cqmm_init must have
 * run (CQMM_ERR_STATE otherwise). Every call on a device shares that device's
 * array: calls that overlap on one device (other threads or streams) race. */
int cqmm_mat_mult_fp32_noalloc(cqmm_plan *plan, double *a, int64_t lena, const double *b,
                               int64_t lenb, double *c, int64_t lenc, uint32_t flags);
int64_t cqmm_noalloc_scratch_bytes(void);

/* Plan introspection (no reference counterpart; used for checks and reports). */
typedef struct cqmm_plan_info {
  int64_t n_triplets;          /* contributing (i,k,j) block products         */
  int64_t n_out_blocks;        /* output blocks receiving at least one product */
  int64_t n_out_total;         /* all blocks of the output matrix             */
  double flops;                /* sum over triplets of 2*n_i*n_k*n_j          */
  int32_t uniform_nsf;         /* support functions if uniform, else 0        */
  int32_t mult_type;
  int32_t kernel;              /* FP64 kernel cqmm_mat_mult runs (CQMM_KERNEL_*) */
  int32_t n_tiles;             /* output tiles of the tile schedule (0: none)  */
  int64_t n_entries;           /* (tile, contraction atom) entries            */
  int64_t n_loads;             /* operand blocks the tile kernel reads        */
  int32_t tile_family;         /* tile kernel: CQMM_TILES_* below (0: none)   */
  int32_t tile_wgs;            /* class tiles: persistent grid (0: one workgroup per tile) */
  int64_t n_mirror;            /* symmetric-result plans: blocks written as transposes (0: none) */
} cqmm_plan_info;
#define CQMM_TILES_DENSE4 1    /* 4x4 blocks, quadrant MFMA tiles            */
#define CQMM_TILES_DENSE_MIXED 2  /* mixed blocks up to 4x4, quadrant MFMA tiles */
#define CQMM_TILES_MIXED 3     /* mixed blocks up to 16, per-product lists    */
#define CQMM_TILES_CLASS 5     /* mixed blocks up to 16 in class tiles (one row and one
                                  column size per tile), packed MFMA tiles    */
#define CQMM_TILES_UNIFORM 16  /* uniform blocks: 16 + block size             */
int cqmm_plan_get_info(const cqmm_plan *plan, cqmm_plan_info *info);

/* FP64 kernel selection (no reference counterpart). For uniform block sizes
 * the plan also holds an output-tile schedule (8 x 8 output blocks per
 * workgroup, operand blocks staged once per tile in LDS); AUTO runs it when
 * present. TRIPLET runs the per-output-block kernel; TILE fails with
 * CQMM_ERR_STATE when the plan has no tile schedule. Both give the same
 * products; the summation order of type-2 blocks differs (ascending C-halo
 * atom in the tile kernel, B-row order in the triplet kernel). */
#define CQMM_KERNEL_AUTO 0
#define CQMM_KERNEL_TRIPLET 1
#define CQMM_KERNEL_TILE 2
int cqmm_plan_set_kernel(cqmm_plan *plan, int32_t kernel);
/* Grid of the class-tile kernel (no reference counterpart): 0 runs one
 * workgroup per tile; wgs > 0 a persistent grid of wgs workgroups (rounded up
 * to a multiple of 8, at most one per tile), each taking tiles of its XCD's
 * contiguous range in turn. Same products and summation order either way.
 * cqmm_plan_tune picks it for class-tile plans. Ignored by the other kernels. */
int cqmm_plan_set_tile_workgroups(cqmm_plan *plan, int32_t wgs);
/* Kernel autotuning on the caller's operands (no reference counterpart: the
 * reference picks its kernel at compile time, MULT_KERN in src/matrix.obj):
 * runs each kernel the plan has (tile schedule if built, per-output-block
 * always) once, then `reps` more times timed with HIP events on the library
 * stream, and fixes the faster one as the plan's kernel. Same arguments and
 * overwrite semantics as cqmm_mat_mult (the output holds the last run's,
 * correct, result). Average milliseconds per call go to *ms_tile (-1 when
 * the plan has no tile schedule) and *ms_triplet. */
int cqmm_plan_tune(cqmm_plan *plan, double *a, int64_t lena, const double *b, int64_t lenb, double *c,
                   int64_t lenc, uint32_t flags, int32_t reps, float *ms_tile, float *ms_triplet);
/* Copy the triplet list to host memory, in the order the kernel consumes it
 * (grouped by output block, each group in the reference's summation order).
 * Each triplet is (out, p, q): 1-based offsets of the output block, the left
 * operand block and the right operand block in the arrays mat_mult receives
 * (type 1: c, a, b; type 2: a, c, b). Arrays must hold n_triplets entries. */
int cqmm_plan_get_triplets(const cqmm_plan *plan, int64_t *out, int64_t *p,
                           int64_t *q);

#ifdef __cplusplus
}
#endif
#endif /* CQMM_H */
