/*
 * cqmm_matrix.h — device-resident matrix API around the product, mirroring
 * CONQUEST's mult_module (src/mult_module.f90): integer matrix handles, a LIFO
 * temporary-matrix stack, matrix_product / matrix_sum / matrix_scale /
 * matrix_product_trace / matrix_trace / symmetrise_L, all on one GPU, and the
 * density-matrix loops that call them (McWeeny purification, Hotelling S^-1).
 *
 * A context holds, per range id, the reference's type(matrix) index of that
 * range (mat(:,range), src/matrix_data_module.f90:76-133), and per mult id one
 * cqmm_plan (the mult(x) table, src/mult_module.f90:271-1102). Matrix data live
 * in HBM for the life of the handle.
 */
#ifndef CQMM_MATRIX_H
#define CQMM_MATRIX_H

#include <stdint.h>

#include "cqmm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* type(matrix) of all primary-set partitions for one process, flattened
 * (src/matrix_module.f90:63-82; filled by get_naba / make_npxyz,
 * src/matrix_elements_module.f90:257-384,836-903). Per-atom arrays are in
 * primary order; i_acc / i_nd_acc stay partition-relative as in the reference;
 * per-neighbour arrays are the partitions' neighbour lists concatenated. */
typedef struct cqmm_matrix_index {
  int32_t n_parts;
  const int32_t *n_atoms;     /* [n_parts]                                   */
  const int32_t *nab_offset;  /* [n_parts] amat(nn)%offset                   */
  const int64_t *nd_offset;   /* [n_parts] amat(nn)%nd_offset (64-bit here: one GPU holds
                                 matrices longer than 2^31 elements)         */
  int32_t n_prim;
  const int32_t *n_nab;       /* [n_prim]                                    */
  const int32_t *i_acc;       /* [n_prim] partition-relative, 1-based        */
  const int32_t *i_nd_acc;    /* [n_prim] partition-relative, 1-based        */
  const int32_t *ndimi;       /* [n_prim]                                    */
  const int64_t *onsite;      /* [n_prim] 1-based offset of the onsite block, 0 if none */
  const int32_t *i_part;      /* [n_nabs] cover-set partition of neighbour   */
  const int32_t *i_seq;       /* [n_nabs] sequence number in that partition  */
  const int32_t *ndimj;       /* [n_nabs]                                    */
  int64_t n_nabs;
  int64_t length;             /* amat(1)%length                              */
} cqmm_matrix_index;

typedef struct cqmm_ctx cqmm_ctx;

int cqmm_ctx_create(cqmm_ctx **ctx);
int cqmm_ctx_destroy(cqmm_ctx *ctx);
/* Register range `range` (0 <= range < 64) with its index (copied). Replacing a
 * registered range is refused (CQMM_ERR_STATE) while a live matrix or a
 * registered mult uses it; after atoms move, build a new context. */
int cqmm_ctx_add_range(cqmm_ctx *ctx, int32_t range, const cqmm_matrix_index *idx);
/* Transpose pairs of a range for symmetrise_L / matrix_transpose
 * (src/mult_module.f90:1856-1893, 2609-2673): 1-based offsets p, q of blocks
 * (i,j,s) and (j,i,-s) and their dimensions n1 x n2 of block p. Every block
 * appears in exactly one pair (p == q for blocks that are their own transpose). */
int cqmm_ctx_set_transpose(cqmm_ctx *ctx, int32_t range, int64_t npairs, const int64_t *p, const int64_t *q,
                           const int32_t *n1, const int32_t *n2);
/* mult(mult_id) with operand ranges as matrix_product passes them:
 * type 1 C(rc) = A(ra).B(rb); type 2 A(ra) = C(rc).B(rb)^T. Takes ownership of plan. */
int cqmm_ctx_add_mult(cqmm_ctx *ctx, int32_t mult_id, int32_t ra, int32_t rb, int32_t rc, cqmm_plan *plan);

/* allocate_temp_matrix / free_temp_matrix (src/mult_module.f90:3269-3377):
 * zero-filled, handles freed strictly in reverse order of allocation. */
int cqmm_matrix_allocate(cqmm_ctx *ctx, int32_t range, int32_t *handle);
int cqmm_matrix_free(cqmm_ctx *ctx, int32_t handle);
int cqmm_matrix_length(const cqmm_ctx *ctx, int32_t handle, int64_t *length);
int cqmm_matrix_range(const cqmm_ctx *ctx, int32_t handle, int32_t *range);
int cqmm_matrix_device_ptr(const cqmm_ctx *ctx, int32_t handle, double **ptr);
int cqmm_matrix_set(cqmm_ctx *ctx, int32_t handle, const double *host);
int cqmm_matrix_get(cqmm_ctx *ctx, int32_t handle, double *host);

/* matrix_product(A, B, C, mult) (src/mult_module.f90:2556-2606). Runs in the
 * context's product precision (FP64 unless set below). */
int cqmm_matrix_product(cqmm_ctx *ctx, int32_t A, int32_t B, int32_t C, int32_t mult_id);
/* Precision of cqmm_matrix_product on this context (no reference counterpart:
 * the reference multiplies in FP64 only). 0: FP64 (default). 1, 2, 3: FP32
 * copies of both operands and FP32 MFMA, the result widened into the FP64
 * output; 1 sums each output block in FP32, 2 adds each block product's FP32
 * result into FP64, 3 sums a block's products within one chunk of the tile
 * schedule in FP32 and the chunks in FP64; 4 split-BF16 x3 products (as
 * CQMM_BF16X3) where the plan's kernel has them, else mode 2. Sums, traces, transposes and the
 * matrices themselves stay FP64. cqmm_dmm keeps its electron-number
 * correction and the final K, energy and electron count in FP64. */
int cqmm_ctx_set_product_precision(cqmm_ctx *ctx, int32_t precision);
int cqmm_ctx_get_product_precision(const cqmm_ctx *ctx, int32_t *precision);
/* matrix_sum(alpha, A, beta, B) (src/mult_module.f90:2677-2715): A = alpha*A + beta*B
 * on the blocks common to both patterns (matrix_add / addscan,
 * src/multiply_module.f90:807-943); blocks of A absent from B are untouched. */
int cqmm_matrix_sum(cqmm_ctx *ctx, double alpha, int32_t A, double beta, int32_t B);
/* matrix_scale (src/mult_module.f90:2718-2738). */
int cqmm_matrix_scale(cqmm_ctx *ctx, double alpha, int32_t A);
/* matrix_product_trace (src/mult_module.f90:2885-2916): sum_p A_p*B_p, same range,
 * deterministic FP64 reduction. */
int cqmm_matrix_product_trace(cqmm_ctx *ctx, int32_t A, int32_t B, double *result);
/* matrix_trace (src/mult_module.f90:2773-2811): sum of onsite diagonals. */
int cqmm_matrix_trace(cqmm_ctx *ctx, int32_t A, double *result);
/* symmetrise_L (src/mult_module.f90:1856-1893) for any range with transpose pairs:
 * A = A/2 + A^T/2, in place. */
int cqmm_matrix_symmetrise(cqmm_ctx *ctx, int32_t A);
/* matrix_transpose (src/mult_module.f90:2609-2673) on one process: B = A^T, both
 * of one range with transpose pairs. */
int cqmm_matrix_transpose(cqmm_ctx *ctx, int32_t A, int32_t B);

/* ---- McWeeny purification (src/McWeeny.f90) ------------------------------ */
typedef struct cqmm_mcw_setup {
  int32_t range_L, range_S, range_H, range_LS, range_LSL;
  int32_t mult_L_S_LS;     /* type 1: LS  = L.S      (L_S_LS,  src/mult_module.f90:583) */
  int32_t mult_LS_L_LSL;   /* type 1: LSL = LS.L     (LS_L_LSL, :687)                  */
  int32_t mult_LSL_SL_L;   /* type 2: LSLSL(L) = LSL.(LS)^T (LSL_SL_L, :729)           */
  int32_t max_iter;        /* n_L_iterations                                          */
  int32_t precision;       /* 0: all FP64; bulk products with FP32 MFMA on FP32 operands,
                              then FP64 refinement: 1 sums each block's products in FP64
                              registers, 2 sums them in the FP32 MFMA accumulator, 3 sums
                              them in FP32 within each chunk of the tile schedule (at most
                              15 products) and the chunks in FP64, 4 split-BF16 x3 products
                              (operands split into two BF16 halves, hi.hi + hi.lo + lo.hi
                              per block product in FP32, FP64 sums) where all three plans'
                              kernels have them (blocks 9 to 16 or class tiles), else 1;
                              5 as 4, then a second reduced-precision phase as 1 (from the
                              first phase's end to its own stop test), then FP64 */
  int32_t n_final_fp64;    /* FP64 iterations after the FP32 phase (>= 2)              */
  double fp32_tol;         /* FP32 phase ends when |dE| < fp32_tol (Ha, per spin)      */
} cqmm_mcw_setup;

#define CQMM_MCW_TRACE 256
typedef struct cqmm_mcw_result {
  int32_t iterations;      /* McWeeny steps taken (accepted + the final rejected one)   */
  int32_t fp32_iterations; /* of which with FP32 products                              */
  int32_t stop_reason;     /* 1 energy rose, 2 |dE| < 1e-8, 3 c outside [0,1], 4 max_iter */
  int32_t restored;        /* 1 when fewer than 3 steps ran and L was restored         */
  double energy;           /* Tr[rho H] of the last accepted rho (per spin)           */
  double electrons;        /* Tr[rho S] (per spin)                                     */
  double c;                /* last mid-point c                                         */
  double seconds;          /* device wall time of the loop                             */
  double energy_trace[CQMM_MCW_TRACE]; /* Tr[rho H] after step n (n < CQMM_MCW_TRACE), the
                              steps in order, FP32 ones included                        */
} cqmm_mcw_result;

/* McWMin (src/McWeeny.f90:107-272) over handles matH, matS, matL (one spin,
 * fixed population): each step is McW_matrix_multiply_nonds (:620-740). On
 * return matL holds the purified density matrix. */
int cqmm_mcweeny(cqmm_ctx *ctx, const cqmm_mcw_setup *setup, int32_t matH, int32_t matS, int32_t matL,
                 cqmm_mcw_result *result);

/* McWMin for n independent systems at once (no reference counterpart: CONQUEST
 * purifies one system per run). System i is (ctxs[i], setups[i], matH[i],
 * matS[i], matL[i]) exactly as in cqmm_mcweeny; contexts must differ. The
 * systems step in lockstep on the calling thread's stream: each product kind
 * of all systems runs as one grouped tile launch, the step's traces and
 * coefficients stay on the device, one copy per step brings every system's
 * scalars to the host, and a step whose set of systems and precision phases
 * is unchanged replays a captured hipGraph. A system leaves the batch when
 * McWMin stops it; its result and L are what cqmm_mcweeny gives alone (the
 * same kernels and summation order; bitwise). status[i] receives system i's
 * status (a failing system does not stop the others); the return value is
 * nonzero only for argument or device errors of the batch itself. n <= 64. */
#define CQMM_BATCH_NO_GRAPH 1u     /* enqueue every step directly (no hipGraph) */
#define CQMM_BATCH_NO_GROUPING 2u  /* one product launch per system            */
int cqmm_mcweeny_batch(int32_t n, cqmm_ctx *const *ctxs, const cqmm_mcw_setup *setups, const int32_t *matH,
                       const int32_t *matS, const int32_t *matL, uint32_t flags, cqmm_mcw_result *results,
                       int32_t *status);

/* InitMcW (src/McWeeny.f90:340-570): L0 = ((lambda*mubar + Ne)/No) T - (lambda/No) THT
 * from T ~ S^-1, with Gershgorin-type limits of H (GetHLimits, :1020-1109).
 * Products: T_S_TS (type 1), T_H_TH (type 1), TH_T_L (type 1, result at L range). */
typedef struct cqmm_initmcw_setup {
  int32_t range_T, range_S, range_H, range_L, range_TS, range_TH;
  int32_t mult_T_S_TS, mult_T_H_TH, mult_TH_T_L;
  double n_electrons;      /* electrons per spin                                      */
} cqmm_initmcw_setup;
int cqmm_init_mcweeny(cqmm_ctx *ctx, const cqmm_initmcw_setup *setup, int32_t matH, int32_t matS, int32_t matT,
                      int32_t matL, double out_mu_lambda[2]);

/* McWMin with two spin channels (src/McWeeny.f90:107-272, nspin = 2,
 * spin_factor 1). Spin s is (matH[s], matS[s], matL[s]) of one context; pass one
 * S handle twice unless the support functions depend on spin
 * (flag_SpinDependentSF). fixed_populations = 1 (flag_fix_spin_population):
 * each spin steps with McW_matrix_multiply_nonds (:620-740) and its own c, a
 * spin's stop test marks it done (for good), and both spins take every step
 * until both are done; setup->precision must be 0. fixed_populations = 0:
 * McW_matrix_multiply_ds (:772-890), one c = (c1_up + c1_dn)/(c2_up + c2_dn)
 * and one stop test on E_up + E_dn; every precision of cqmm_mcw_setup applies,
 * both spins changing phase together. result[s] receives spin s's energy,
 * electrons and energy_trace; iterations, restored and c are the loop's;
 * stop_reason is the loop's (variable) or why spin s was marked done (fixed).
 * Fewer than 3 steps restore both L. */
int cqmm_mcweeny_spin(cqmm_ctx *ctx, const cqmm_mcw_setup *setup, int32_t fixed_populations, const int32_t matH[2],
                      const int32_t matS[2], const int32_t matL[2], cqmm_mcw_result result[2]);

/* InitMcW with two spin channels (src/McWeeny.f90:340-570, nspin = 2). Fixed
 * populations: each spin as cqmm_init_mcweeny with n_electrons[s]. Variable
 * populations: the direct sum of both channels, n_o_ds = 2 n_o, n_e_ds =
 * n_electrons[0] + n_electrons[1], SXHX summed over spins, the limits of H the
 * extremes over both spins, A_ds = 1 - (SX_up + SX_dn)/n_o_ds, one mubar and
 * lambda, L_s = ((lambda mubar + n_e_ds)/n_o_ds) T_s - (lambda/n_o_ds) T_s H_s T_s.
 * setup->n_electrons is not read. out_mu_lambda (may be NULL) receives
 * {mubar_up, lambda_up, mubar_dn, lambda_dn}. */
int cqmm_init_mcweeny_spin(cqmm_ctx *ctx, const cqmm_initmcw_setup *setup, int32_t fixed_populations,
                           const int32_t matH[2], const int32_t matS[2], const int32_t matT[2], const int32_t matL[2],
                           const double n_electrons[2], double out_mu_lambda[4]);

/* Hotelling iteration for T ~ S^-1 (Iter_Hott_InvS, src/S_matrix_module.f90:746-963):
 * T <- 2T - T.S.T with products T_S_TS (type 1) and TS_T_T (type 2, result at T
 * range); stops when the change of Tr[TS]/n_orb is below tol or after max_iter. */
typedef struct cqmm_hotelling_setup {
  int32_t range_T, range_S, range_TS;
  int32_t mult_T_S_TS, mult_TS_T_T;
  int32_t max_iter;
  double tol;
} cqmm_hotelling_setup;
int cqmm_hotelling(cqmm_ctx *ctx, const cqmm_hotelling_setup *setup, int32_t matS, int32_t matT,
                   int32_t *iterations, double *omega);

/* ---- LNV density-matrix minimisation (src/mult_module.f90, src/DMMinModule.f90) -- */
/* Ranges and mult entries of the LNV products, as matrix_product passes them
 * (mult table, src/mult_module.f90:583-875). Type 1: C(rc) = A(ra).B(rb);
 * type 2: A(ra) = C(rc).B(rb)^T. Ranges LS, LH, LSL and L need transpose pairs. */
typedef struct cqmm_lnv_setup {
  int32_t range_L, range_S, range_H, range_LS, range_LH, range_LSL, range_SLS;
  int32_t mult_L_S_LS;     /* t1 (L, S, LS)      LS = L.S                     */
  int32_t mult_L_H_LH;     /* t1 (L, H, LH)      LH = L.H                     */
  int32_t mult_HL_S_LSL;   /* t1 (LH, S, LSL)    HLS = HL.S                   */
  int32_t mult_HLS_LS_L;   /* t2 (L, LS, LSL)    B = (2HLS + SLH).(SL)^T      */
  int32_t mult_SL_S_SLS;   /* t1 (LS, S, SLS)    SLS = SL.S                   */
  int32_t mult_SLS_LS_L;   /* t2 (L, LS, SLS)    SLSLS = SLS.(SL)^T           */
  int32_t mult_LS_L_LSL;   /* t1 (LS, L, LSL)    LSL = LS.L                   */
  int32_t mult_LSL_SL_H;   /* t2 (H, LS, LSL)    LSLSL = LSL.(LS)^T at H range */
} cqmm_lnv_setup;

#define CQMM_LNV_K 1u      /* K = 3LSL - 2LSLSL (H range)                       */
#define CQMM_LNV_E 2u      /* energy = Tr[H K] (needs a K handle)               */
#define CQMM_LNV_M3 4u     /* M3 = dE/dL (L range)                              */
#define CQMM_LNV_PHI 8u    /* phi = dNe/dL = 6SLS - 6SLSLS and electrons        */

/* LNV_matrix_multiply (src/mult_module.f90:1478-1823), one spin, with doM1 =
 * doM2 = doM4 = false. Per spin: electrons = 3Tr[L.SLS] - 2Tr[L.SLSLS],
 * energy = Tr[H K]. Handles not requested may be -1. */
int cqmm_lnv(cqmm_ctx *ctx, const cqmm_lnv_setup *setup, uint32_t flags, int32_t matH, int32_t matS,
             int32_t matL, int32_t matK, int32_t matM3, int32_t matphi, double *electrons, double *energy);

typedef struct cqmm_dmm_setup {
  cqmm_lnv_setup lnv;
  int32_t range_T, range_TL;
  int32_t mult_T_L_TL;      /* t1 (T, L, TL)                                     */
  int32_t mult_TL_T_L;      /* t2 (L, T, TL): X -> T X T at L range              */
  int32_t n_L_iterations;   /* iterations per stage (n_L_iterations)             */
  int32_t vary_mu;          /* 1: keep the electron number (correct_electron_number
                               and the projected search directions)             */
  int32_t maxpulay;         /* DM.MaxPulay (default 5, at most 16)               */
  int32_t global_tolerance; /* residual not divided by the atom count            */
  int32_t n_atoms;          /* ni_in_cell                                       */
  double tolerance;         /* residual at which lateDM stops                   */
  double min_pulay_step, max_pulay_step;  /* DM.MinPulayStepSize / DM.MaxPulayStepSize */
  double lin_tol;           /* DM.LinTol: earlyDM -> lateDM when |zeta| < lin_tol */
  double electrons_target;  /* electrons per spin (ne_spin_in_cell)            */
  double delta_e;           /* initial delta_e (from McWMin; 0 is allowed)      */
} cqmm_dmm_setup;

typedef struct cqmm_dmm_result {
  int32_t done;             /* residual below tolerance                         */
  int32_t early_iterations, late_iterations, inflexions;
  double energy;            /* 2 Tr[H K] (spin-unpolarised total)               */
  double electrons;         /* 2 (3Tr[LSLS] - 2Tr[LSLSLS])                      */
  double residual;          /* last g                                           */
  double seconds;
} cqmm_dmm_result;

/* correct_electron_number_fixspin (src/DMMinModule.f90:1676-1817): moves L along
 * T phi T^T to the target electron number (cubic model, SolveCubic :2072-2143). */
int cqmm_correct_electrons(cqmm_ctx *ctx, const cqmm_dmm_setup *setup, int32_t matH, int32_t matS, int32_t matT,
                           int32_t matL, double *electrons);
/* FindMinDM (src/DMMinModule.f90:172-312) for one spin, from the current L:
 * earlyDM (:366-723, steepest descent with lineMinL :1395-1614) then lateDM
 * (:809-1339, Pulay mixing with DoPulay, src/Pulay.f90:188-262). K (H range)
 * receives 3LSL - 2LSLSL of the final L. */
int cqmm_dmm(cqmm_ctx *ctx, const cqmm_dmm_setup *setup, int32_t matH, int32_t matS, int32_t matT, int32_t matL,
             int32_t matK, cqmm_dmm_result *result);

#ifdef __cplusplus
}
#endif
#endif /* CQMM_MATRIX_H */
