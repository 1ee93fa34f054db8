// McWeeny purification (McWMin, src/McWeeny.f90:107-272, one step =
// McW_matrix_multiply_nonds :620-740) of one or several independent systems
// in lockstep on one stream.
//
// Every step of every system in the batch is enqueued without a host round
// trip: the three products of each kind run as one grouped tile launch over
// all systems (cq::launch_tile_group; each system's tiles are computed exactly
// as in its own launch, so a result does not depend on what it was batched
// with), the traces c1, c2, E and N_e land in device memory, a one-thread
// kernel turns c1/c2 into the coefficients of rho', and the sums read those
// coefficients on the device. The host reads all systems' scalars with one copy
// per step and runs McWMin's stop tests; accepted steps enqueue L <- rho' and
// symmetrise_L before the next step. Once the set of systems (and their
// precision phase) is unchanged for a step, the step is captured in a hipGraph
// and replayed, one launch per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cqmm.h"
#include "../../include/cqmm_matrix.h"
#include "cq_common.h"
#include "cq_ctx.h"
#include "cq_device.h"
#include "cq_mcw.h"
#include "cq_ops.h"
#include "cq_temps.h"

namespace cq {
namespace {

#define MTRY(x)              \
  do {                       \
    const int st_ = (x);     \
    if (st_) return st_;     \
  } while (0)

// Device scalars of one system, kScal doubles:
// 0 c1 = Tr[S(LSL - LSLSL)], 1 c2 = Tr[S(L - LSL)], 2 E = Tr[rho' H],
// 3 N_e = Tr[rho' S], 4 1 + c, 5 coefficient of L in rho', 6 final scale of
// rho', 7 c as McWMin sees it (-0.01 when |c2| < 1e-6).
constexpr int kScal = 8;

// rho' = ((1+c)LSL - LSLSL + (1-2c)L)/(1-c) for c < 1/2, else ((1+c)LSL - LSLSL)/c
// (McW_matrix_multiply_nonds, src/McWeeny.f90:700-720); `nonds` applies its
// rule that a vanishing c2 reports c = -0.01 (:737), which
// McW_matrix_multiply_ds does not have.
__device__ void mcw_coef(double c1, double c2, bool nonds, double *sc) {
  const double cn = c1 / c2;
  sc[4] = 1.0 + cn;
  if (cn < 0.5) {
    sc[5] = 1.0 - 2.0 * cn;
    sc[6] = 1.0 / (1.0 - cn);
  } else {
    sc[5] = 0.0;
    sc[6] = 1.0 / cn;
  }
  sc[7] = (nonds && fabs(c2) < 1e-6) ? -0.01 : cn;
}

__global__ void k_mcw_coef(double *sc) {
  if (threadIdx.x != 0) return;
  mcw_coef(sc[0], sc[1], true, sc);
}

// Variable spin populations (McW_matrix_multiply_ds, src/McWeeny.f90:806-858):
// one c = (c1_up + c1_dn)/(c2_up + c2_dn) for both spin channels.
__global__ void k_mcw_coef_pair(double *a, double *b) {
  if (threadIdx.x != 0) return;
  const double c1 = a[0] + b[0], c2 = a[1] + b[1];
  mcw_coef(c1, c2, false, a);
  mcw_coef(c1, c2, false, b);
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct McwJob {
  cqmm_ctx *ctx = nullptr;
  cqmm_mcw_setup su{};
  int32_t H = 0, S = 0, L = 0;
  cqmm_mcw_result *res = nullptr;
  std::unique_ptr<Temps> tmp;
  int32_t oldL = 0, rho = 0, LS = 0, LSL = 0, LSLSL = 0, top = 0, bottom = 0, botH = 0;
  float *L32 = nullptr, *S32 = nullptr, *LS32 = nullptr, *LSL32 = nullptr, *LSLSL32 = nullptr;
  bool mixed = false, fp32_phase = false, done = false;
  int mode = 0;  // FP32 product mode of the reduced-precision phase (1-4)
  int mode2 = 0; // a second reduced-precision phase's mode (precision 5: 2), 0 for none
  int fp64_done = 0, n = 0, stop = 4, max_iter = 50, nfinal = 2, status = 0;
  double oldE = 1.0e30, lastE = 0.0, lastNe = 0.0, lastC = 0.0, t0 = 0.0;
  int slot = -1;
  uint64_t serial = 0;
  // Spin pairs: both channels of one McWMin (nspin = 2). `partner` is the other
  // channel, `lead` marks the first (it holds the loop state of variable
  // populations); spin 0 = a spin-unpolarised system, 1 = fixed populations,
  // 2 = variable populations; spin_done = this channel's stop test fired.
  McwJob *partner = nullptr;
  bool lead = true, spin_done = false;
  int spin = 0;
};

McwEngine::McwEngine() = default;

McwEngine::~McwEngine() {
  // Temporaries of systems still admitted (an error ended the loop), newest
  // first: two channels of a spin pair share a context's stack.
  for (auto it = jobs_.rbegin(); it != jobs_.rend(); ++it) (void)(*it)->tmp->release();
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (hsc_) (void)hipHostFree(hsc_);
}

int McwEngine::init(hipStream_t s, bool use_graph) {
  s_ = s;
  use_graph_ = use_graph;
  if (!dsc_.p) MTRY(dsc_.alloc((size_t)kMaxJobs * kScal * sizeof(double)));
  if (!hsc_) MTRY(hip_check(hipHostMalloc((void **)&hsc_, (size_t)kMaxJobs * kScal * sizeof(double), hipHostMallocDefault),
                            "pinned scalars"));
  slot_used_.assign(kMaxJobs, false);
  return CQMM_OK;
}

double *McwEngine::dsc(const McwJob &j) const { return dsc_.as<double>() + (size_t)j.slot * kScal; }
const double *McwEngine::hsc(const McwJob &j) const { return hsc_ + (size_t)j.slot * kScal; }

// McWMin's set-up for one system (src/McWeeny.f90:107-140): temporaries, the
// saved L for the restore rule, FP32 copies for the mixed-precision phase, and
// every map and buffer the step needs, so that no step allocates (graph capture).
int McwEngine::add(cqmm_ctx *ctx, const cqmm_mcw_setup &su, int32_t H, int32_t S, int32_t L, cqmm_mcw_result *res) {
  return add_one(ctx, su, H, S, L, res, nullptr);
}

// Both spin channels of one McWMin (nspin = 2) in one context: two systems
// whose steps the host couples (one c for variable populations, the joint stop
// rule for fixed ones).
int McwEngine::add_spin(cqmm_ctx *ctx, const cqmm_mcw_setup &su, bool fixed, const int32_t H[2], const int32_t S[2],
                        const int32_t L[2], cqmm_mcw_result res[2]) {
  if (fixed && su.precision != 0)
    return set_errorf(CQMM_ERR_ARG, "cqmm_mcweeny_spin: fixed spin populations take precision 0 (not %d)",
                      su.precision);
  if (L[0] == L[1]) return set_error(CQMM_ERR_ARG, "cqmm_mcweeny_spin: one L handle for both spins");
  MTRY(add_one(ctx, su, H[0], S[0], L[0], &res[0], nullptr));
  McwJob *a = jobs_.back().get();
  const int st = add_one(ctx, su, H[1], S[1], L[1], &res[1], ctx);
  if (st) {
    (void)a->tmp->release();
    slot_used_[a->slot] = false;
    jobs_.pop_back();
    return st;
  }
  McwJob *b = jobs_.back().get();
  a->partner = b;
  b->partner = a;
  b->lead = false;
  a->spin = b->spin = fixed ? 1 : 2;
  // Variable populations compare the total energy; McWMin starts it at
  // spin_factor (oldE_up + oldE_dn) = 2e30 (src/McWeeny.f90:145-146).
  if (!fixed) a->oldE = 2.0e30;
  return CQMM_OK;
}

int McwEngine::add_one(cqmm_ctx *ctx, const cqmm_mcw_setup &su, int32_t H, int32_t S, int32_t L, cqmm_mcw_result *res,
                       const cqmm_ctx *shared) {
  if (!ctx || !res) return set_error(CQMM_ERR_ARG, "cqmm_mcweeny: null argument");
  MTRY(ctx_check(ctx, H, "cqmm_mcweeny"));
  MTRY(ctx_check(ctx, S, "cqmm_mcweeny"));
  MTRY(ctx_check(ctx, L, "cqmm_mcweeny"));
  if (hrange(ctx, L) != su.range_L || hrange(ctx, S) != su.range_S || hrange(ctx, H) != su.range_H)
    return set_error(CQMM_ERR_ARG, "cqmm_mcweeny: handle ranges differ from the setup");
  for (int32_t r : {su.range_LS, su.range_LSL})
    if (r < 0 || r >= kMaxRanges || !ctx->ranges[r].set)
      return set_errorf(CQMM_ERR_ARG, "cqmm_mcweeny: range %d not registered", r);
  for (int32_t m : {su.mult_L_S_LS, su.mult_LS_L_LSL, su.mult_LSL_SL_L})
    if (!ctx->mults.count(m)) return set_errorf(CQMM_ERR_ARG, "cqmm_mcweeny: mult %d not registered", m);
  if (!ctx->ranges[su.range_L].has_pairs)
    return set_errorf(CQMM_ERR_STATE, "cqmm_mcweeny: range %d has no transpose pairs", su.range_L);
  if (su.precision < 0 || su.precision > 5) return set_errorf(CQMM_ERR_ARG, "cqmm_mcweeny: precision %d", su.precision);
  // Systems of one batch have their own contexts (their temporaries follow
  // the context's LIFO rule); the second channel of a spin pair shares the first's.
  for (const auto &j : jobs_)
    if (j->ctx == ctx && ctx != shared)
      return set_error(CQMM_ERR_ARG, "cqmm_mcweeny: two systems of one batch share a context");
  int slot = -1;
  for (int k = 0; k < kMaxJobs; ++k)
    if (!slot_used_[k]) {
      slot = k;
      break;
    }
  if (slot < 0) return set_errorf(CQMM_ERR_ARG, "cqmm_mcweeny: more than %d systems in one batch", kMaxJobs);
  std::unique_ptr<McwJob> j(new McwJob());
  j->ctx = ctx;
  j->su = su;
  j->H = H;
  j->S = S;
  j->L = L;
  j->res = res;
  j->t0 = now_s();
  std::memset(res, 0, sizeof(*res));
  j->tmp.reset(new Temps(ctx));
  Temps &t = *j->tmp;
  MTRY(t.alloc(su.range_L, &j->oldL));
  MTRY(t.alloc(su.range_L, &j->rho));
  MTRY(t.alloc(su.range_LS, &j->LS));
  MTRY(t.alloc(su.range_LSL, &j->LSL));
  MTRY(t.alloc(su.range_L, &j->LSLSL));
  MTRY(t.alloc(su.range_S, &j->top));
  MTRY(t.alloc(su.range_S, &j->bottom));
  MTRY(t.alloc(su.range_H, &j->botH));
  j->mixed = su.precision != 0;
  // Product mode of the reduced-precision phase: precision 1 -> FP32 MFMA with
  // FP64 sums, 2 -> FP32 sums, 3 -> FP32 chunk runs, 4 -> split-BF16 x3 where
  // all three products' kernels have it (blocks of 9 to 16 or class tiles),
  // else FP32 MFMA with FP64 sums (e.g. 4x4 blocks); 5 -> split-BF16 x3, then
  // a second phase in FP32 MFMA with FP64 sums before the FP64 steps.
  j->mode = su.precision == 1 ? 2 : su.precision == 2 ? 1 : su.precision == 5 ? 4 : su.precision;
  if (j->mode == 4)
    for (int32_t m : {su.mult_L_S_LS, su.mult_LS_L_LSL, su.mult_LSL_SL_L}) {
      const TripletPlan &tp = ctx->mults.at(m).plan->tp;
      if (!uses_tiles(tp) || !tiles_take_bf16x3(tp.tiles)) j->mode = 2;
    }
  j->mode2 = su.precision == 5 && j->mode == 4 ? 2 : 0;
  j->fp32_phase = j->mixed;
  j->max_iter = su.max_iter > 0 ? su.max_iter : 50;
  j->nfinal = su.n_final_fp64 > 0 ? su.n_final_fp64 : 2;
  if (j->mixed) {
    // FP32 copies are keyed by the L handle: systems never share them.
    const std::string k = "#" + std::to_string(L);
    void *p;
    MTRY(get_scratch(ctx, "L32" + k, hlen(ctx, L) * 4, &p));
    j->L32 = (float *)p;
    MTRY(get_scratch(ctx, "S32" + k, hlen(ctx, S) * 4, &p));
    j->S32 = (float *)p;
    MTRY(get_scratch(ctx, "LS32" + k, hlen(ctx, j->LS) * 4, &p));
    j->LS32 = (float *)p;
    MTRY(get_scratch(ctx, "LSL32" + k, hlen(ctx, j->LSL) * 4, &p));
    j->LSL32 = (float *)p;
    MTRY(get_scratch(ctx, "LSLSL32" + k, hlen(ctx, j->LSLSL) * 4, &p));
    j->LSLSL32 = (float *)p;
    if (j->mode == 4) MTRY(op_to_bf16x2(hlen(ctx, S), dptr(ctx, S), (uint32_t *)j->S32, s_));
    else MTRY(op_to_f32(hlen(ctx, S), dptr(ctx, S), j->S32, s_));
  }
  // Range maps of every sum of the step, built now (not during a capture).
  const int32_t rS = su.range_S, rH = su.range_H, rL = su.range_L, rLSL = su.range_LSL;
  const std::pair<int32_t, int32_t> maps[] = {{rS, rLSL}, {rS, rL}, {rL, rLSL}, {rH, rL}};
  for (const auto &m : maps)
    if (m.first != m.second) {
      const RunMap *rm = nullptr;
      MTRY(summap(ctx, m.first, m.second, &rm));
    }
  MTRY(ensure_red(ctx));
  MTRY(cqmm_matrix_sum(ctx, 0.0, j->oldL, 1.0, L));
  j->slot = slot;
  j->serial = ++serial_;
  slot_used_[slot] = true;
  jobs_.push_back(std::move(j));
  return CQMM_OK;
}

// matrix_sum(alpha, A, beta, B) with optional device coefficients.
int McwEngine::sum(cqmm_ctx *ctx, double a, const double *ap, int32_t A, double b, const double *bp, int32_t B) {
  const int32_t ra = hrange(ctx, A), rb = hrange(ctx, B);
  if (ra == rb) return op_axpby_dev(hlen(ctx, A), a, ap, dptr(ctx, A), b, bp, dptr(ctx, B), s_);
  const RunMap *m = nullptr;
  MTRY(summap(ctx, ra, rb, &m));
  return op_runs_axpby_dev(*m, a, ap, dptr(ctx, A), b, bp, dptr(ctx, B), s_);
}

int McwEngine::trace(cqmm_ctx *ctx, int32_t A, int32_t B, double *d_out) {
  return op_dot(hlen(ctx, A), dptr(ctx, A), dptr(ctx, B), ctx->red_part.as<double>(), d_out, s_);
}

// The three products of one kind for all systems: tile schedules grouped by
// kernel family into launches of up to kMaxGroup products; plans without a
// tile schedule run their own launch.
int McwEngine::products(int which) {
  struct Grp {
    const TilePlan *tp = nullptr;
    bool qt = false;
    int mode = 0;
    std::vector<TileArgs> a;
  };
  std::map<int, Grp> groups;
  // Symmetric-result plans on the tile path: mirrored after the launches.
  struct Mirror {
    const TripletPlan *tp;
    void *O;
    bool f32;
  };
  std::vector<Mirror> mirrors;
  for (const auto &jp : jobs_) {
    McwJob &j = *jp;
    cqmm_ctx *ctx = j.ctx;
    const int32_t mid = which == 0 ? j.su.mult_L_S_LS : which == 1 ? j.su.mult_LS_L_LSL : j.su.mult_LSL_SL_L;
    const TripletPlan &tp = ctx->mults.at(mid).plan->tp;
    // matrix_product(A, B, C): L.S -> LS; LS.L -> LSL; LSL.(LS)^T -> LSLSL.
    const int32_t A = which == 0 ? j.L : which == 1 ? j.LS : j.LSL;
    const int32_t B = which == 0 ? j.S : which == 1 ? j.L : j.LS;
    const int32_t C = which == 0 ? j.LS : which == 1 ? j.LSL : j.LSLSL;
    const void *P, *Q;
    void *O;
    int mode = 0;
    if (j.fp32_phase) {
      float *A32 = which == 0 ? j.L32 : which == 1 ? j.LS32 : j.LSL32;
      float *B32 = which == 0 ? j.S32 : which == 1 ? j.L32 : j.LS32;
      float *C32 = which == 0 ? j.LS32 : which == 1 ? j.LSL32 : j.LSLSL32;
      P = A32;
      Q = B32;
      O = C32;
      mode = j.mode;
    } else {
      P = dptr(ctx, A);
      Q = dptr(ctx, B);
      O = dptr(ctx, C);
    }
    if (!uses_tiles(tp)) {
      if (mode == 0) MTRY(launch_block_products(tp, (const double *)P, (const double *)Q, (double *)O, false, s_));
      else MTRY(launch_block_products_f32(tp, (const float *)P, (const float *)Q, (float *)O, mode, s_));
      continue;
    }
    if (tp.n_mirror > 0) mirrors.push_back({&tp, O, mode != 0});
    if (tp.tiles.n_tiles == 0) continue;
    if (!group_) {
      MTRY(launch_tile_args(tp.tiles, tp.q_trans, mode, tile_launch_args(tp.tiles, P, Q, O), s_));
      continue;
    }
    Grp &g = groups[tile_group_key(tp.tiles, tp.q_trans, mode)];
    g.tp = &tp.tiles;
    g.qt = tp.q_trans;
    g.mode = mode;
    g.a.push_back(tile_launch_args(tp.tiles, P, Q, O));
  }
  for (auto &kv : groups) {
    Grp &g = kv.second;
    for (size_t b = 0; b < g.a.size(); b += kMaxGroup) {
      TileGroup tg;
      std::memset(&tg, 0, sizeof(tg));
      int32_t total = 0;
      const size_t e = std::min(g.a.size(), b + kMaxGroup);
      for (size_t k = b; k < e; ++k) {
        tg.a[k - b] = g.a[k];
        tg.a[k - b].tile_base = total;
        total += g.a[k].n_tiles;
      }
      tg.ngrp = (int32_t)(e - b);
      tg.total = total;
      MTRY(launch_tile_group(*g.tp, g.qt, g.mode, tg, s_));
    }
  }
  for (const Mirror &m : mirrors)
    MTRY(m.f32 ? mirror_result_f32(*m.tp, (float *)m.O, s_) : mirror_result(*m.tp, (double *)m.O, s_));
  return CQMM_OK;
}

// One McW_matrix_multiply_nonds for every system, enqueued on the stream.
int McwEngine::enqueue_step() {
  for (const auto &jp : jobs_) {
    McwJob &j = *jp;
    if (!j.fp32_phase) continue;
    if (j.mode == 4) MTRY(op_to_bf16x2(hlen(j.ctx, j.L), dptr(j.ctx, j.L), (uint32_t *)j.L32, s_));
    else MTRY(op_to_f32(hlen(j.ctx, j.L), dptr(j.ctx, j.L), j.L32, s_));
  }
  // Split-BF16 systems take LS and LSL as operands in split-BF16 words: each
  // FP32 result is widened (LSL, for the traces) and then split in place
  // before the product that reads it.
  auto split_bf16 = [&](int which) -> int {
    for (const auto &jp : jobs_) {
      McwJob &j = *jp;
      if (!j.fp32_phase || j.mode != 4) continue;
      if (which == 0) MTRY(op_f32_to_bf16x2(hlen(j.ctx, j.LS), (uint32_t *)j.LS32, s_));
      if (which == 1) {
        MTRY(op_to_f64(hlen(j.ctx, j.LSL), j.LSL32, dptr(j.ctx, j.LSL), s_));
        MTRY(op_f32_to_bf16x2(hlen(j.ctx, j.LSL), (uint32_t *)j.LSL32, s_));
      }
    }
    return CQMM_OK;
  };
  for (int which = 0; which < 3; ++which) {
    MTRY(products(which));
    if (which < 2) MTRY(split_bf16(which));
  }
  for (const auto &jp : jobs_) {
    McwJob &j = *jp;
    cqmm_ctx *ctx = j.ctx;
    double *sc = dsc(j);
    if (j.fp32_phase) {
      if (j.mode != 4) MTRY(op_to_f64(hlen(ctx, j.LSL), j.LSL32, dptr(ctx, j.LSL), s_));
      MTRY(op_to_f64(hlen(ctx, j.LSLSL), j.LSLSL32, dptr(ctx, j.LSLSL), s_));
    }
    // c1 = Tr[S (LSL - LSLSL)] and c2 = Tr[S (L - LSL)] on the S range.
    MTRY(op_scale(hlen(ctx, j.top), 0.0, dptr(ctx, j.top), s_));
    MTRY(sum(ctx, 0.0, nullptr, j.top, 1.0, nullptr, j.LSL));
    MTRY(sum(ctx, 1.0, nullptr, j.top, -1.0, nullptr, j.LSLSL));
    MTRY(trace(ctx, j.S, j.top, sc + 0));
    MTRY(op_scale(hlen(ctx, j.bottom), 0.0, dptr(ctx, j.bottom), s_));
    MTRY(sum(ctx, 0.0, nullptr, j.bottom, 1.0, nullptr, j.L));
    MTRY(sum(ctx, 1.0, nullptr, j.bottom, -1.0, nullptr, j.LSL));
    MTRY(trace(ctx, j.S, j.bottom, sc + 1));
  }
  for (const auto &jp : jobs_) {
    McwJob &j = *jp;
    if (j.spin == 2) {
      if (j.lead) hipLaunchKernelGGL(k_mcw_coef_pair, dim3(1), dim3(64), 0, s_, dsc(j), dsc(*j.partner));
    } else {
      hipLaunchKernelGGL(k_mcw_coef, dim3(1), dim3(64), 0, s_, dsc(j));
    }
    MTRY(hip_check(hipGetLastError(), "McWeeny coefficient kernel"));
  }
  for (const auto &jp : jobs_) {
    McwJob &j = *jp;
    cqmm_ctx *ctx = j.ctx;
    double *sc = dsc(j);
    // rho' on the L range.
    MTRY(op_scale(hlen(ctx, j.rho), 0.0, dptr(ctx, j.rho), s_));
    MTRY(sum(ctx, 0.0, nullptr, j.rho, -1.0, nullptr, j.LSLSL));
    MTRY(sum(ctx, 1.0, nullptr, j.rho, 0.0, sc + 4, j.LSL));
    MTRY(sum(ctx, 1.0, nullptr, j.rho, 0.0, sc + 5, j.L));
    MTRY(op_scale_dev(hlen(ctx, j.rho), 0.0, sc + 6, dptr(ctx, j.rho), s_));
    // E = Tr[rho' H] (H range), N_e = Tr[rho' S] (S range).
    MTRY(op_scale(hlen(ctx, j.botH), 0.0, dptr(ctx, j.botH), s_));
    MTRY(sum(ctx, 0.0, nullptr, j.botH, 1.0, nullptr, j.rho));
    MTRY(trace(ctx, j.botH, j.H, sc + 2));
    MTRY(sum(ctx, 0.0, nullptr, j.top, 1.0, nullptr, j.rho));
    MTRY(trace(ctx, j.top, j.S, sc + 3));
  }
  return CQMM_OK;
}

// L <- rho' and symmetrise_L (McWMin, src/McWeeny.f90:180-190).
int McwEngine::accept(McwJob &j) {
  MTRY(op_axpby(hlen(j.ctx, j.L), 0.0, dptr(j.ctx, j.L), 1.0, dptr(j.ctx, j.rho), s_));
  return op_symmetrise(j.ctx->ranges[hrange(j.ctx, j.L)].pairs, dptr(j.ctx, j.L), s_);
}

// McWMin's stop tests on step n's scalars (src/McWeeny.f90:166-173), with the
// mixed-precision phase rules: the FP32 phase ends at the first stop test (or
// when |dE| falls below fp32_tol |E|); a rejected FP32 step restarts from the
// last accepted L in FP64; the first n_final_fp64 FP64 steps are kept even if
// FP32-level noise trips the energy tests. `j` is a system or the first
// channel of a variable-population spin pair, whose loop tests the total
// energy and the common c (:207-237) and accepts or rejects both channels.
int McwEngine::after_step(McwJob &j) {
  McwJob *u[2] = {&j, j.partner};
  const int nu = j.partner ? 2 : 1;
  const double *sc[2] = {hsc(j), nu == 2 ? hsc(*j.partner) : nullptr};
  double c2 = sc[0][1], omega = sc[0][2];
  if (nu == 2) {
    c2 = sc[0][1] + sc[1][1];
    omega = sc[0][2] + sc[1][2];
  }
  if (!(std::fabs(c2) > 1e-15)) return set_errorf(CQMM_ERR_NOTCONV, "McWeeny: c2 is near zero (%g)", c2);
  const double cn = sc[0][7];
  const int n = ++j.n;
  for (int k = 0; k < nu; ++k) {
    u[k]->n = n;
    if (n - 1 < CQMM_MCW_TRACE) u[k]->res->energy_trace[n - 1] = sc[k][2];
    if (j.fp32_phase) ++u[k]->res->fp32_iterations;
  }
  if (!j.fp32_phase) ++j.fp64_done;
  int why = 0;
  if (omega > j.oldE) why = 1;
  else if (std::fabs(omega - j.oldE) < 1.0e-8) why = 2;
  else if (cn > 1.0 || cn < 0.0) why = 3;
  auto keep = [&]() -> int {
    for (int k = 0; k < nu; ++k) {
      MTRY(accept(*u[k]));
      u[k]->lastE = sc[k][2];
      u[k]->lastNe = sc[k][3];
      u[k]->lastC = cn;
    }
    return CQMM_OK;
  };
  bool handled = false;
  if (j.fp32_phase) {
    const bool flat = std::fabs(omega - j.oldE) < j.su.fp32_tol * std::max(1.0, std::fabs(omega));
    if (why || flat) {
      if (j.mode2 && j.mode != j.mode2) {
        // The next reduced-precision phase (its S copy in its operand format),
        // from the last accepted L or this step's.
        for (int k = 0; k < nu; ++k) {
          u[k]->mode = u[k]->mode2;
          MTRY(op_to_f32(hlen(u[k]->ctx, u[k]->S), dptr(u[k]->ctx, u[k]->S), u[k]->S32, s_));
        }
      } else {
        for (int k = 0; k < nu; ++k) u[k]->fp32_phase = false;
      }
      if (!(why == 1 || why == 3)) MTRY(keep());
      j.oldE = 1.0e30 * nu;
      handled = true;
    }
  } else if (why && j.fp64_done <= j.nfinal && j.mixed && why != 3) {
    why = 0;
  }
  if (!handled) {
    if (why) {
      for (int k = 0; k < nu; ++k) {
        u[k]->stop = why;
        u[k]->done = true;
      }
      return CQMM_OK;
    }
    MTRY(keep());
    j.oldE = omega;
  }
  if (n >= j.max_iter)
    for (int k = 0; k < nu; ++k) {
      u[k]->n = j.max_iter + 1;
      u[k]->done = true;
    }
  return CQMM_OK;
}

// Fixed spin populations (McWMin, src/McWeeny.f90:156-203): each channel has
// its own c and energy test; a channel whose test fires is marked done for the
// rest of the loop; the loop ends when both are done, and until then both
// channels take every step.
int McwEngine::after_fixed_pair(McwJob &a) {
  McwJob *u[2] = {&a, a.partner};
  const double *sc[2] = {hsc(a), hsc(*a.partner)};
  for (int k = 0; k < 2; ++k)
    if (!(std::fabs(sc[k][1]) > 1e-15))
      return set_errorf(CQMM_ERR_NOTCONV, "McWeeny: c2 of spin %d is near zero (%g)", k + 1, sc[k][1]);
  const int n = ++a.n;
  a.partner->n = n;
  for (int k = 0; k < 2; ++k) {
    if (n - 1 < CQMM_MCW_TRACE) u[k]->res->energy_trace[n - 1] = sc[k][2];
    const double omega = sc[k][2], cn = sc[k][7], oldE = u[k]->oldE;
    int why = 0;
    if (omega > oldE) why = 1;
    else if (std::fabs(omega - oldE) < 1.0e-8) why = 2;
    else if (cn > 1.0 || cn < 0.0) why = 3;
    if (why && !u[k]->spin_done) {
      u[k]->spin_done = true;
      u[k]->stop = why;
    }
  }
  if (u[0]->spin_done && u[1]->spin_done) {
    u[0]->done = u[1]->done = true;
    return CQMM_OK;
  }
  for (int k = 0; k < 2; ++k) {
    MTRY(accept(*u[k]));
    u[k]->oldE = sc[k][2];
    u[k]->lastE = sc[k][2];
    u[k]->lastNe = sc[k][3];
    u[k]->lastC = sc[k][7];
  }
  if (n >= a.max_iter)
    for (int k = 0; k < 2; ++k) {
      u[k]->n = a.max_iter + 1;
      u[k]->done = true;
    }
  return CQMM_OK;
}

// McWMin's end: the restore rule (fewer than 3 steps, src/McWeeny.f90:253-254),
// the results, and the temporaries freed in reverse order.
int McwEngine::finish(McwJob &j) {
  cqmm_mcw_result *res = j.res;
  res->iterations = std::min(j.n, j.max_iter);
  res->stop_reason = j.stop;
  if (j.n < 3) {
    MTRY(op_axpby(hlen(j.ctx, j.L), 0.0, dptr(j.ctx, j.L), 1.0, dptr(j.ctx, j.oldL), s_));
    res->restored = 1;
  }
  res->energy = j.lastE;
  res->electrons = j.lastNe;
  res->c = j.lastC;
  MTRY(j.tmp->release());
  MTRY(hip_check(hipStreamSynchronize(s_), "cqmm_mcweeny"));
  res->seconds = now_s() - j.t0;
  return CQMM_OK;
}

void McwEngine::drop_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  exec_ = nullptr;
  graph_key_.clear();
}

// A captured step holds the device addresses of its systems' matrices, so the
// key names each system by its admission serial (a slot is reused by the next
// system admitted after one finishes) and its precision phase.
std::string McwEngine::key() const {
  std::string k;
  for (const auto &jp : jobs_) k += std::to_string(jp->serial) + (jp->fp32_phase ? (char)('0' + jp->mode) : 'd') + ",";
  return k;
}

// One step for all systems: a graph replay when the batch is unchanged since
// the previous (eager) step, else the step enqueued directly; then the host
// reads every system's scalars and applies the stop tests. Systems that stop
// (or fail) are finished and removed; their results are final.
int McwEngine::step(std::vector<McwJob *> *finished) {
  if (jobs_.empty()) return CQMM_OK;
  const std::string k = key();
  if (use_graph_ && exec_ && k == graph_key_) {
    MTRY(hip_check(hipGraphLaunch(exec_, s_), "hipGraphLaunch"));
    ++graph_steps_;
  } else if (use_graph_ && k == last_eager_key_) {
    drop_graph();
    hipGraph_t g = nullptr;
    MTRY(hip_check(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"));
    const int st = enqueue_step();
    const hipError_t e = hipStreamEndCapture(s_, &g);
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    MTRY(hip_check(e, "hipStreamEndCapture"));
    const hipError_t ei = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MTRY(hip_check(ei, "hipGraphInstantiate"));
    graph_key_ = k;
    ++captures_;
    MTRY(hip_check(hipGraphLaunch(exec_, s_), "hipGraphLaunch"));
    ++graph_steps_;
  } else {
    MTRY(enqueue_step());
    last_eager_key_ = k;
  }
  MTRY(hip_check(hipMemcpyAsync(hsc_, dsc_.p, (size_t)kMaxJobs * kScal * sizeof(double), hipMemcpyDeviceToHost, s_),
                 "D2H McWeeny scalars"));
  MTRY(hip_check(hipStreamSynchronize(s_), "McWeeny step"));
  for (const auto &jp : jobs_) {
    McwJob &j = *jp;
    if (!j.lead) continue;
    const int st = j.spin == 1 ? after_fixed_pair(j) : after_step(j);
    if (st)
      for (McwJob *m : {&j, j.partner})
        if (m) {
          m->status = st;
          m->done = true;
        }
  }
  // Finish what stopped; a spin pair's second channel first, as its
  // temporaries are the newer on the shared context's stack.
  std::vector<McwJob *> order;
  for (const auto &jp : jobs_)
    if (jp->lead && jp->done) {
      if (jp->partner) order.push_back(jp->partner);
      order.push_back(jp.get());
    }
  for (McwJob *j : order) {
    if (!j->status) j->status = finish(*j);
    else (void)j->tmp->release();
    slot_used_[j->slot] = false;
  }
  for (size_t i = 0; i < jobs_.size();) {
    if (!jobs_[i]->done) {
      ++i;
      continue;
    }
    if (finished) finished->push_back(jobs_[i].release());
    jobs_.erase(jobs_.begin() + (long)i);
  }
  return CQMM_OK;
}

int McwEngine::job_status(const McwJob *j) { return j->status; }
cqmm_ctx *McwEngine::job_ctx(const McwJob *j) { return j->ctx; }
void McwEngine::job_delete(McwJob *j) { delete j; }

}  // namespace cq

using namespace cq;

extern "C" {

int cqmm_mcweeny(cqmm_ctx *ctx, const cqmm_mcw_setup *su, int32_t matH, int32_t matS, int32_t matL,
                 cqmm_mcw_result *res) {
  int32_t status = 0;
  const int st = cqmm_mcweeny_batch(1, &ctx, su, &matH, &matS, &matL, 0u, res, &status);
  return st ? st : status;
}

int cqmm_mcweeny_spin(cqmm_ctx *ctx, const cqmm_mcw_setup *su, int32_t fixed_populations, const int32_t matH[2],
                      const int32_t matS[2], const int32_t matL[2], cqmm_mcw_result result[2]) {
  return guarded(__func__, [&]() -> int {
    if (!ctx || !su || !matH || !matS || !matL || !result)
      return set_error(CQMM_ERR_ARG, "cqmm_mcweeny_spin: null argument");
    MTRY(ensure_init());
    McwEngine eng;
    MTRY(eng.init(stream(), true));
    MTRY(eng.add_spin(ctx, *su, fixed_populations != 0, matH, matS, matL, result));
    std::vector<McwJob *> fin;
    int status = CQMM_OK;
    while (eng.active() > 0) {
      MTRY(eng.step(&fin));
      for (McwJob *j : fin) {
        if (!status) status = McwEngine::job_status(j);
        McwEngine::job_delete(j);
      }
      fin.clear();
    }
    return status;
  });
}

int cqmm_mcweeny_batch(int32_t n, cqmm_ctx *const *ctxs, const cqmm_mcw_setup *setups, const int32_t *matH,
                       const int32_t *matS, const int32_t *matL, uint32_t flags, cqmm_mcw_result *results,
                       int32_t *status) {
  return guarded(__func__, [&]() -> int {
    if (n < 0 || (n > 0 && (!ctxs || !setups || !matH || !matS || !matL || !results || !status)))
      return set_error(CQMM_ERR_ARG, "cqmm_mcweeny_batch: bad argument");
    if (n > McwEngine::kMaxJobs)
      return set_errorf(CQMM_ERR_ARG, "cqmm_mcweeny_batch: %d systems (at most %d)", n, McwEngine::kMaxJobs);
    MTRY(ensure_init());
    McwEngine eng;
    MTRY(eng.init(stream(), !(flags & CQMM_BATCH_NO_GRAPH)));
    eng.set_grouping(!(flags & CQMM_BATCH_NO_GROUPING));
    std::map<const cqmm_mcw_result *, int32_t> index;
    for (int32_t i = 0; i < n; ++i) {
      status[i] = CQMM_ERR_STATE;
      index[&results[i]] = i;
    }
    for (int32_t i = 0; i < n; ++i) {
      const int st = eng.add(ctxs[i], setups[i], matH[i], matS[i], matL[i], &results[i]);
      if (st) status[i] = st;
    }
    std::vector<McwJob *> fin;
    while (eng.active() > 0) {
      MTRY(eng.step(&fin));
      for (McwJob *j : fin) {
        status[index[j->res]] = McwEngine::job_status(j);
        McwEngine::job_delete(j);
      }
      fin.clear();
    }
    return CQMM_OK;
  });
}

}  // extern "C"
