// One-call density-matrix solves on synthetic structures and the request loop
// that serves many of them on one GPU (slots = worker threads, each with its
// own default-priority HIP stream; admission by an HBM budget).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/cqmm.h"
#include "../../include/cqmm_matrix.h"
#include "../../include/cqmm_serve.h"
#include "../../include/cqmm_system.h"
#include "cq_common.h"
#include "cq_device.h"
#include "cq_ctx.h"
#include "cq_mcw.h"

namespace {

// mult ids of the purification products (names as in src/mult_module.f90)
enum { L_S_LS = 1, LS_L_LSL, LSL_SL_L, T_S_TS, TS_T_T, T_H_TH, TH_T_L };

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One request's structure, matrices and McWeeny set-up, from the seeded
// structure to the InitMcW guess (owned until the solve ends).
struct Solve {
  cqmm_system *sys = nullptr;
  cqmm_ctx *ctx = nullptr;
  int32_t mS = 0, mH = 0, mT = 0, mL = 0;
  cqmm_mcw_setup ms{};
  cqmm_mcw_result mr{};
  double ne = 0.0, t0 = 0.0, t1 = 0.0;
  int64_t mcw_bytes = 0;  // McWeeny temporaries and FP32 copies
  Solve() = default;
  Solve(const Solve &) = delete;
  Solve &operator=(const Solve &) = delete;
  ~Solve() { release(); }
  void release() {
    if (ctx) cqmm_ctx_destroy(ctx);
    if (sys) cqmm_sys_destroy(sys);
    ctx = nullptr;
    sys = nullptr;
  }
};

#define PTRY(x)                 \
  do {                          \
    int st_ = (x);              \
    if (st_) {                  \
      r->status = st_;          \
      return st_;               \
    }                           \
  } while (0)

double nab_per_row(double density, double r) { return density * 4.0 / 3.0 * M_PI * r * r * r + 1.0; }

}  // namespace

extern "C" {

void cqmm_purify_default_params(cqmm_purify_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->r_S = 8.0;
  p->r_H = 8.0;
  p->r_L = 12.0;
  p->density = 0.05;
  p->dmin = 2.0;
  p->s_amp = 0.1;
  p->s_decay = 1.5;
  p->h_amp = -0.2;
  p->h_decay = 2.0;
  p->h_onsite = 0.5;
  p->h_split = 1.5;
  p->h_shift = 3.0;
  p->precision = 1;
  p->max_iter = 100;
  p->fp32_tol = 0.0;
}

int64_t cqmm_purify_hbm_estimate(int32_t natoms, int32_t nsf, const cqmm_purify_params *p) {
  const double N = natoms, b = (double)nsf * nsf, d = p->density;
  const double rS = p->r_S, rH = p->r_H, rL = p->r_L, rT = p->r_S;
  auto len = [&](double r) { return N * nab_per_row(d, r) * b; };
  const double S = len(rS), H = len(rH), L = len(rL), T = len(rT), LS = len(rL + rS), LSL = len(2 * rL + rS),
               TS = len(rT + rS), TH = len(rT + rH);
  // Matrices: persistent S,H,T,L; Hotelling and McWeeny temporaries (kept in the context's pool);
  // InitMcW's TS, TH, XHX, S-range temporary.
  const double mats = 8.0 * (S + H + T + L + (3 * TS + 3 * T) + (2 * L + LS + LSL + L + 2 * S + H) + (TS + TH + L + S));
  const double f32 = 4.0 * (L + S + LS + LSL + L);
  const double nb = N * nsf;  // products ~ rows x neighbours x neighbours; bytes per product in the plan
  (void)nb;
  auto trip = [&](double ra, double rb) { return N * nab_per_row(d, ra) * nab_per_row(d, rb); };
  const double t_big = std::max(trip(rL + rS, rL), trip(rL, rL + rS));
  const double trips = trip(rL, rS) + 2 * t_big + trip(rT, rS) + trip(rT, rT) + trip(rT, rH) + trip(rT + rH, rT);
  // Resident triplet lists and tile schedules (14 B per product) + the largest
  // build's transient buffers (38 B per product of the biggest product). The
  // coefficients are fitted to the measured per-solve peaks (hbm_peak) of
  // 64-2048 atoms with 4 and 13 support functions, where this sum (x1.1) is
  // 1.18-1.68 times the peak.
  const double plans = 14.0 * trips + 38.0 * t_big;
  return (int64_t)(1.1 * (mats + f32 + plans)) + (64LL << 20);
}

}  // extern "C"

namespace {

// Everything before McWMin for one request, on the calling thread's stream:
// seeded structure, index set and plans of the purification products,
// synthetic S and H, Hotelling S^-1 (src/S_matrix_module.f90:746-963) and
// InitMcW (src/McWeeny.f90:340-570). The stream is synchronised on return.
int prepare(int32_t natoms, int32_t nsf, uint64_t seed, const cqmm_purify_params *p, cqmm_purify_result *r, Solve &sv) {
  std::memset(r, 0, sizeof(*r));
  r->natoms = natoms;
  r->n_orb = natoms * nsf;
  r->hbm_estimate = cqmm_purify_hbm_estimate(natoms, nsf, p);
  sv.t0 = now_s();
  PTRY(cq::ensure_init());
  std::vector<double> xyz(3 * (size_t)natoms);
  double edge = 0.0;
  PTRY(cqmm_random_structure(natoms, p->density, p->dmin, seed, xyz.data(), &edge));
  const double box[3] = {edge, edge, edge};
  const int32_t ng = std::max<int32_t>(1, (int32_t)std::lround(edge / 5.0));
  const int32_t grid[3] = {ng, ng, ng};
  std::vector<int32_t> nsfv(natoms, nsf);
  const double rS = p->r_S, rH = p->r_H, rL = p->r_L, rT = p->r_S;
  const double rLS = rL + rS, rLSL = 2 * rL + rS, rTS = rT + rS, rTH = rT + rH;
  const double rcover = std::max({rLSL, rL + rLS, rTH + rT, 2 * rT, rT + rS, rT + rH});
  PTRY(cqmm_sys_create(natoms, xyz.data(), box, nsfv.data(), grid, rcover, &sv.sys));
  // Ranges in the order S, H, L, T, LS, LSL, TS, TH; equal radii share one range.
  const double radii[8] = {rS, rH, rL, rT, rLS, rLSL, rTS, rTH};
  int32_t id[8];
  std::map<double, int32_t> seen;
  for (int k = 0; k < 8; ++k) {
    auto it = seen.find(radii[k]);
    if (it != seen.end()) {
      id[k] = it->second;
    } else {
      PTRY(cqmm_sys_add_range(sv.sys, radii[k], &id[k]));
      seen[radii[k]] = id[k];
    }
  }
  enum { S = 0, H, Lr, T, LS, LSL, TS, TH };
  PTRY(cqmm_ctx_create(&sv.ctx));
  for (auto &kv : seen) {
    cqmm_matrix_index idx;
    PTRY(cqmm_sys_matrix_index(sv.sys, kv.second, &idx));
    PTRY(cqmm_ctx_add_range(sv.ctx, kv.second, &idx));
  }
  {
    int64_t np = 0;
    PTRY(cqmm_sys_transpose_pairs(sv.sys, id[Lr], &np, nullptr, nullptr, nullptr, nullptr));
    std::vector<int64_t> pp(np), qq(np);
    std::vector<int32_t> a(np), b(np);
    PTRY(cqmm_sys_transpose_pairs(sv.sys, id[Lr], &np, pp.data(), qq.data(), a.data(), b.data()));
    PTRY(cqmm_ctx_set_transpose(sv.ctx, id[Lr], np, pp.data(), qq.data(), a.data(), b.data()));
  }
  struct M { int mid, type, ra, rb, rc; };
  const M mults[7] = {{L_S_LS, 1, id[Lr], id[S], id[LS]},   {LS_L_LSL, 1, id[LS], id[Lr], id[LSL]},
                      {LSL_SL_L, 2, id[Lr], id[LS], id[LSL]}, {T_S_TS, 1, id[T], id[S], id[TS]},
                      {TS_T_T, 2, id[T], id[T], id[TS]},     {T_H_TH, 1, id[T], id[H], id[TH]},
                      {TH_T_L, 1, id[TH], id[T], id[Lr]}};
  for (const M &m : mults) {
    cqmm_plan *plan = nullptr;
    // LSL = LS.L and LSLSL = LSL.(LS)^T are symmetric (L and S are).
    if (p->symmetric && (m.mid == LS_L_LSL || m.mid == LSL_SL_L))
      PTRY(cqmm_sys_plan_create_symmetric(sv.sys, m.type, m.ra, m.rb, m.rc, &plan));
    else
      PTRY(cqmm_sys_plan_create(sv.sys, m.type, m.ra, m.rb, m.rc, &plan));
    PTRY(cqmm_ctx_add_mult(sv.ctx, m.mid, m.ra, m.rb, m.rc, plan));
  }
  cqmm_range_info ri;
  PTRY(cqmm_sys_range_info(sv.sys, id[S], &ri));
  std::vector<double> Sv(ri.length);
  const double ps[2] = {p->s_amp, p->s_decay};
  PTRY(cqmm_sys_fill(sv.sys, id[S], 1, seed, 11, ps, 2, Sv.data()));
  PTRY(cqmm_sys_range_info(sv.sys, id[H], &ri));
  std::vector<double> Hv(ri.length);
  const double ph[4] = {p->h_amp, p->h_decay, p->h_onsite, p->h_split};
  PTRY(cqmm_sys_fill(sv.sys, id[H], 2, seed, 12, ph, 4, Hv.data()));
  if (p->h_shift != 0.0) {
    if (id[S] != id[H]) PTRY(cq::set_error(CQMM_ERR_ARG, "cqmm_purify: h_shift needs r_S == r_H"));
    for (size_t t = 0; t < Hv.size(); ++t) Hv[t] += p->h_shift * Sv[t];
  }
  PTRY(cqmm_matrix_allocate(sv.ctx, id[S], &sv.mS));
  PTRY(cqmm_matrix_allocate(sv.ctx, id[H], &sv.mH));
  PTRY(cqmm_matrix_allocate(sv.ctx, id[T], &sv.mT));
  PTRY(cqmm_matrix_allocate(sv.ctx, id[Lr], &sv.mL));
  PTRY(cqmm_matrix_set(sv.ctx, sv.mS, Sv.data()));
  PTRY(cqmm_matrix_set(sv.ctx, sv.mH, Hv.data()));
  sv.t1 = now_s();
  cqmm_hotelling_setup hs = {id[T], id[S], id[TS], T_S_TS, TS_T_T, 50, 1e-10};
  int32_t hit = 0;
  double om = 0.0;
  PTRY(cqmm_hotelling(sv.ctx, &hs, sv.mS, sv.mT, &hit, &om));
  r->hotelling_iterations = hit;
  sv.ne = (double)(r->n_orb / 2);
  cqmm_initmcw_setup is = {id[T], id[S], id[H], id[Lr], id[TS], id[TH], T_S_TS, T_H_TH, TH_T_L, sv.ne};
  double ml[2];
  PTRY(cqmm_init_mcweeny(sv.ctx, &is, sv.mH, sv.mS, sv.mT, sv.mL, ml));
  sv.ms = cqmm_mcw_setup{id[Lr], id[S], id[H], id[LS], id[LSL], L_S_LS, LS_L_LSL, LSL_SL_L,
                         p->max_iter, p->precision, 2, p->fp32_tol};
  // McWeeny working set: oldL, rho, LSLSL (L), LS, LSL, top, bottom (S), botH (H),
  // and FP32 copies of L, S, LS, LSL, LSLSL in the mixed-precision modes.
  int64_t len[8];
  for (int k = 0; k < 8; ++k) {
    PTRY(cqmm_sys_range_info(sv.sys, id[k], &ri));
    len[k] = ri.length;
  }
  sv.mcw_bytes = 8 * (3 * len[Lr] + len[LS] + len[LSL] + 2 * len[S] + len[H]);
  if (p->precision != 0) sv.mcw_bytes += 4 * (2 * len[Lr] + len[S] + len[LS] + len[LSL]);
  return CQMM_OK;
}

// Useful FLOP of one McWeeny step (the three products).
double step_flops(const Solve &sv) {
  double f = 0.0;
  for (int32_t m : {L_S_LS, LS_L_LSL, LSL_SL_L}) {
    cqmm_plan_info info;
    if (cqmm_plan_get_info(sv.ctx->mults.at(m).plan, &info) == CQMM_OK) f += info.flops;
  }
  return f;
}

// Results of a finished McWMin.
void complete(cqmm_purify_result *r, Solve &sv, double t_end) {
  r->mcw_s = sv.mr.seconds;
  r->mcw_flops = step_flops(sv) * sv.mr.iterations;
  r->iterations = sv.mr.iterations;
  r->fp32_iterations = sv.mr.fp32_iterations;
  r->stop_reason = sv.mr.stop_reason;
  r->band_energy = 2.0 * sv.mr.energy;
  r->n_electrons = 2.0 * sv.mr.electrons;
  r->ne_error = std::fabs(r->n_electrons - 2.0 * sv.ne);
  r->setup_s = sv.t1 - sv.t0;
  r->solve_s = t_end - sv.t1;
}

}  // namespace

extern "C" {

int cqmm_purify(int32_t natoms, int32_t nsf, uint64_t seed, const cqmm_purify_params *p, cqmm_purify_result *r) {
  return cq::guarded(__func__, [&]() -> int {
    if (!p || !r || natoms <= 0 || nsf <= 0 || nsf > 64) return cq::set_error(CQMM_ERR_ARG, "cqmm_purify: bad argument");
    PTRY(cq::ensure_init());
    cq::hbm_peak_reset();
    const int64_t hbm_base = cq::hbm_current();
    Solve sv;
    {
      const int st = prepare(natoms, nsf, seed, p, r, sv);
      if (st) return st;
    }
    PTRY(cqmm_mcweeny(sv.ctx, &sv.ms, sv.mH, sv.mS, sv.mL, &sv.mr));
    complete(r, sv, now_s());
    r->hbm_peak = cq::hbm_peak() - hbm_base;
    return CQMM_OK;
  });
}

int cqmm_serve(const cqmm_serve_request *reqs, int32_t n, const cqmm_serve_config *cfg, cqmm_serve_result *res,
               double *wall_s) {
  return cq::guarded(__func__, [&]() -> int {
    if (!reqs || !cfg || !res || n < 0 || cfg->slots <= 0 || cfg->hbm_budget <= 0 ||
        (cfg->batch && (cfg->max_batch <= 0 || cfg->max_batch > cq::McwEngine::kMaxJobs)))
      return cq::set_error(CQMM_ERR_ARG, "cqmm_serve: bad argument");
    int st = cq::ensure_init();
    if (st) return st;
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    std::vector<int64_t> est(n);
    for (int32_t i = 0; i < n; ++i) {
      std::memset(&res[i], 0, sizeof(res[i]));
      est[i] = cqmm_purify_hbm_estimate(reqs[i].natoms, reqs[i].nsf, &cfg->params);
      res[i].arrival_s = reqs[i].arrival_s;
      res[i].slot = -1;
      res[i].r.status = CQMM_ERR_STATE;  // until a worker finishes the request
    }
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int32_t> pending;          // released, waiting for admission
    std::deque<int32_t> ready;            // prepared, waiting for the batch (batch mode)
    std::vector<std::unique_ptr<Solve>> solves(n);
    bool all_released = false, stop_all = false;
    int32_t n_done = 0;
    int64_t in_use = 0;
    const double t_start = now_s();
    const int device = cq::runtime().device;
    auto log_done = [&](int32_t i, const char *how) {
      if (cfg->log_progress)
        std::fprintf(stderr, "[serve] request %d (%d atoms, %d SF) %s: status %d, %.2f s latency, %d iterations\n", i,
                     reqs[i].natoms, reqs[i].nsf, how, res[i].r.status, res[i].done_s - res[i].arrival_s,
                     res[i].r.iterations);
    };
    // Called with mu held: request i is finished (solved or failed).
    auto finish_locked = [&](int32_t i, const char *how) {
      res[i].done_s = now_s() - t_start;
      in_use -= est[i];
      ++n_done;
      log_done(i, how);
    };
    // Next admissible request: the oldest, if it fits the budget; if not,
    // the first of the next kLookahead that fits (backfill), at most
    // kMaxBypass times while the same request waits at the head, after which
    // admission waits for it. A request that can never fit fails at once.
    constexpr int32_t kLookahead = 64, kMaxBypass = 64;
    int32_t bypassed = 0;
    auto admit_locked = [&]() -> int32_t {
      while (!pending.empty()) {
        const int32_t h = pending.front();
        if (est[h] > cfg->hbm_budget) {
          pending.pop_front();
          res[h].r.status = CQMM_ERR_NOMEM;
          res[h].admit_s = now_s() - t_start;
          in_use += est[h];
          finish_locked(h, "refused");
          cv.notify_all();
          bypassed = 0;
          continue;
        }
        if (in_use + est[h] <= cfg->hbm_budget) {
          pending.pop_front();
          in_use += est[h];
          bypassed = 0;
          return h;
        }
        if (bypassed >= kMaxBypass) return -1;
        const int32_t look = std::min<int32_t>((int32_t)pending.size(), kLookahead + 1);
        for (int32_t k = 1; k < look; ++k) {
          const int32_t r = pending[k];
          if (est[r] <= cfg->hbm_budget && in_use + est[r] <= cfg->hbm_budget) {
            pending.erase(pending.begin() + k);
            in_use += est[r];
            ++bypassed;
            return r;
          }
        }
        return -1;
      }
      return -1;
    };
    auto worker = [&](int slot) {
      (void)hipSetDevice(device);
      hipStream_t s = nullptr;
      const bool have_stream = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
      if (have_stream) cq::set_thread_stream(s);
      for (;;) {
        int32_t i = -1;
        {
          std::unique_lock<std::mutex> lk(mu);
          for (;;) {
            if (stop_all) break;
            i = admit_locked();
            if (i >= 0 || (pending.empty() && all_released)) break;
            cv.wait(lk);
          }
        }
        if (i < 0) break;
        res[i].admit_s = now_s() - t_start;
        res[i].slot = slot;
        if (!have_stream) {
          std::lock_guard<std::mutex> lk(mu);
          res[i].r.status = cq::set_error(CQMM_ERR_HIP, "cqmm_serve: worker stream creation failed");
          finish_locked(i, "failed");
          cv.notify_all();
          continue;
        }
        if (!cfg->batch) {
          // Whole solve on this slot's stream, overlapping with the other slots.
          res[i].r.status = cqmm_purify(reqs[i].natoms, reqs[i].nsf, reqs[i].seed, &cfg->params, &res[i].r);
          (void)hipStreamSynchronize(s);
          std::lock_guard<std::mutex> lk(mu);
          finish_locked(i, "solved");
          cv.notify_all();
          continue;
        }
        // Batch mode: set-up, Hotelling and InitMcW here, McWeeny in the batch.
        std::unique_ptr<Solve> sv(new Solve());
        cq::hbm_peak_reset();
        const int64_t base = cq::hbm_current();
        const int pst = prepare(reqs[i].natoms, reqs[i].nsf, reqs[i].seed, &cfg->params, &res[i].r, *sv);
        res[i].r.hbm_peak = cq::hbm_peak() - base + sv->mcw_bytes;
        std::lock_guard<std::mutex> lk(mu);
        if (pst) {
          res[i].r.status = pst;
          sv.reset();
          finish_locked(i, "failed");
        } else {
          res[i].r.status = CQMM_ERR_STATE;
          solves[i] = std::move(sv);
          ready.push_back(i);
        }
        cv.notify_all();
      }
      cq::set_thread_stream(nullptr);
      if (have_stream) (void)hipStreamDestroy(s);
    };
    // Batch mode: McWMin of every prepared request in lockstep (cq::McwEngine),
    // admitting prepared requests between steps up to max_batch at a time.
    int batch_status = CQMM_OK;
    auto batcher = [&]() {
      (void)hipSetDevice(device);
      hipStream_t s = nullptr;
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        std::lock_guard<std::mutex> lk(mu);
        batch_status = cq::set_error(CQMM_ERR_HIP, "cqmm_serve: batch stream creation failed");
        stop_all = true;
        cv.notify_all();
        return;
      }
      cq::set_thread_stream(s);
      {
        cq::McwEngine eng;
        std::map<cqmm_ctx *, int32_t> of_ctx;
        int est_ = eng.init(s, cfg->batch != 2);
        std::vector<cq::McwJob *> fin;
        while (!est_) {
          std::vector<int32_t> join;
          {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
              while (!ready.empty() && eng.active() + (int)join.size() < cfg->max_batch) {
                join.push_back(ready.front());
                ready.pop_front();
              }
              if (!join.empty() || eng.active() > 0 || n_done == n || stop_all) break;
              cv.wait(lk);
            }
          }
          if (join.empty() && eng.active() == 0) break;
          for (int32_t i : join) {
            Solve &sv = *solves[i];
            const int ast = eng.add(sv.ctx, sv.ms, sv.mH, sv.mS, sv.mL, &sv.mr);
            if (ast) {
              std::lock_guard<std::mutex> lk(mu);
              res[i].r.status = ast;
              solves[i].reset();
              finish_locked(i, "failed");
              cv.notify_all();
            } else {
              of_ctx[sv.ctx] = i;
            }
          }
          if (eng.active() == 0) continue;
          est_ = eng.step(&fin);
          for (cq::McwJob *j : fin) {
            const int32_t i = of_ctx[cq::McwEngine::job_ctx(j)];
            of_ctx.erase(cq::McwEngine::job_ctx(j));
            const int jst = cq::McwEngine::job_status(j);
            cq::McwEngine::job_delete(j);
            if (!jst) complete(&res[i].r, *solves[i], now_s());
            solves[i].reset();
            std::lock_guard<std::mutex> lk(mu);
            res[i].r.status = jst;
            finish_locked(i, jst ? "failed" : "solved");
            cv.notify_all();
          }
          fin.clear();
        }
        if (est_) {
          std::lock_guard<std::mutex> lk(mu);
          batch_status = est_;
          stop_all = true;
          cv.notify_all();
        }
      }
      cq::set_thread_stream(nullptr);
      (void)hipStreamDestroy(s);
    };
    std::vector<std::thread> pool;
    try {
      for (int32_t k = 0; k < cfg->slots; ++k) pool.emplace_back(worker, k);
      if (cfg->batch) pool.emplace_back(batcher);
    } catch (...) {
      {
        std::lock_guard<std::mutex> lk(mu);
        all_released = true;
        stop_all = true;
      }
      cv.notify_all();
      for (auto &th : pool) th.join();
      throw;
    }
    for (int32_t t = 0; t < n; ++t) {
      const int32_t i = order[t];
      const double wait = reqs[i].arrival_s - (now_s() - t_start);
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      {
        std::lock_guard<std::mutex> lk(mu);
        if (stop_all) break;
        pending.push_back(i);
      }
      cv.notify_all();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      all_released = true;
    }
    cv.notify_all();
    for (auto &th : pool) th.join();
    if (wall_s) *wall_s = now_s() - t_start;
    return batch_status;
  });
}

}  // extern "C"
