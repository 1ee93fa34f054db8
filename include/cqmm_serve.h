/*
 * cqmm_serve.h — one-call density-matrix solves for synthetic structures and a
 * request loop that serves many of them on one GPU.
 *
 * No reference counterpart: CONQUEST solves one structure per run. Each solve
 * follows CONQUEST's own sequence for the O(N) path — Hotelling S^-1
 * (src/S_matrix_module.f90:746-1034), InitMcW (src/McWeeny.f90:340-570),
 * McWMin (src/McWeeny.f90:107-272) — on a seeded random structure with the
 * synthetic S and H recipe of include/cqmm_system.h (kinds 1 and 2) plus
 * h_shift*S.
 */
#ifndef CQMM_SERVE_H
#define CQMM_SERVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cqmm_purify_params {
  double r_S, r_H, r_L;     /* ranges (Angstrom); T uses r_S                 */
  double density, dmin;     /* structure: atoms/A^3 and minimum separation    */
  double s_amp, s_decay;    /* S off-diagonal amplitude and decay length      */
  double h_amp, h_decay, h_onsite, h_split, h_shift;  /* H recipe             */
  int32_t precision;        /* cqmm_mcw_setup.precision: 0 FP64; 1 FP32 MFMA + FP64 block
                               sums; 2 FP32 sums; 3 FP32 chunk runs; 4 split-BF16 x3; 5 split-
                               BF16 x3, then FP32 MFMA + FP64 sums; FP64 final steps   */
  int32_t max_iter;         /* McWeeny iterations at most                     */
  double fp32_tol;          /* mixed: FP32 phase also ends when |dE| < fp32_tol*|E|
                               (0: only the McWeeny stop tests end it)        */
  int32_t symmetric;        /* 1: LSL and LSLSL through symmetric-result plans
                               (cqmm_plan_create_symmetric: half the blocks
                               computed, the rest mirrored); 0: every block    */
} cqmm_purify_params;

typedef struct cqmm_purify_result {
  int32_t status;           /* 0, or the failing call's status                 */
  int32_t iterations, fp32_iterations, stop_reason, hotelling_iterations;
  double band_energy;       /* 2 Tr[rho H] (Ha), spin-unpolarised               */
  double n_electrons;       /* 2 Tr[rho S]                                      */
  double ne_error;          /* |n_electrons - 2 floor(n_orb/2)|                 */
  double setup_s, solve_s;  /* host+device set-up; Hotelling+InitMcW+McWeeny    */
  int64_t hbm_estimate;     /* bytes reserved for admission                     */
  int32_t natoms, n_orb;
  int64_t hbm_peak;         /* device bytes the solve held at most (measured)   */
  double mcw_s;             /* McWMin's wall time                               */
  double mcw_flops;         /* useful FLOP of McWMin's products (sum over steps of
                               2 n_i n_k n_j over the triplets of L.S, LS.L, LSL.(LS)^T) */
} cqmm_purify_result;

/* Defaults of the served workload: S/H/L 8/8/12 A, 0.05 atoms/A^3, 2 A. */
void cqmm_purify_default_params(cqmm_purify_params *p);
/* HBM bytes a solve of natoms x nsf would hold (estimate used for admission). */
int64_t cqmm_purify_hbm_estimate(int32_t natoms, int32_t nsf, const cqmm_purify_params *p);
/* One solve on the calling thread (library stream). */
int cqmm_purify(int32_t natoms, int32_t nsf, uint64_t seed, const cqmm_purify_params *p, cqmm_purify_result *r);

typedef struct cqmm_serve_request {
  int32_t natoms, nsf;
  uint64_t seed;
  double arrival_s;         /* arrival time from the start of the run          */
} cqmm_serve_request;

typedef struct cqmm_serve_config {
  int32_t slots;            /* worker threads, one HIP stream each: whole solves (batch 0)
                               or set-ups up to InitMcW (batch 1, 2)              */
  int64_t hbm_budget;       /* bytes; a request is admitted when its estimate fits */
  cqmm_purify_params params;
  int32_t log_progress;     /* nonzero: one line on stderr per finished request */
  int32_t batch;            /* 0: each slot runs its request's whole solve alone;
                               1: McWMin of all prepared requests in lockstep with
                               grouped product launches and hipGraph replay
                               (cqmm_mcweeny_batch's engine); 2: as 1 without graphs */
  int32_t max_batch;        /* batch modes: requests in the McWeeny batch at most (1-64) */
} cqmm_serve_config;

typedef struct cqmm_serve_result {
  cqmm_purify_result r;
  double arrival_s, admit_s, done_s;  /* from the start of the run           */
  int32_t slot;
} cqmm_serve_result;

/* Serve `n` requests: each is released at its arrival time, waits (FIFO) until
 * a slot is free and its HBM estimate fits the budget next to the requests in
 * flight (every admitted request holds its estimate until it finishes), then
 * is set up on that slot's stream. Batch 0: the slot also runs the request's
 * McWMin, overlapping on the device with the other slots. Batch 1/2 (continuous
 * batching): the prepared request joins the McWeeny batch at the next step
 * boundary once fewer than max_batch requests are in it; all requests in the
 * batch advance one step per iteration with one grouped product launch per
 * product kind, and each leaves when McWMin stops it — its result is the one
 * it gets alone (bitwise). Set-ups on the slots overlap with the batch's steps.
 * A failing request records its status and does not stop the others; a
 * request's status stays CQMM_ERR_STATE until it is finished. In batch modes
 * hbm_peak is the set-up's measured peak plus the McWeeny working set.
 * wall_s receives the run's duration. */
int cqmm_serve(const cqmm_serve_request *reqs, int32_t n, const cqmm_serve_config *cfg,
               cqmm_serve_result *results, double *wall_s);

#ifdef __cplusplus
}
#endif
#endif /* CQMM_SERVE_H */
