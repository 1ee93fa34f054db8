// McWeeny purification of independent systems in lockstep on one stream
// (cq_mcw.hip): the batch engine behind cqmm_mcweeny, cqmm_mcweeny_batch and
// the request loop of cqmm_serve.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cqmm_matrix.h"
#include "cq_device.h"

namespace cq {

struct McwJob;

class McwEngine {
 public:
  static constexpr int kMaxJobs = 64;
  McwEngine();
  ~McwEngine();
  McwEngine(const McwEngine &) = delete;
  McwEngine &operator=(const McwEngine &) = delete;
  // Stream for every operation of the engine; graphs on or off.
  int init(hipStream_t s, bool use_graph);
  // Tile products of the systems in one grouped launch per kernel family (on),
  // or one launch per system (off; same results).
  void set_grouping(bool on) { group_ = on; }
  // Admit a system (McWMin's set-up); it joins from the next step.
  int add(cqmm_ctx *ctx, const cqmm_mcw_setup &su, int32_t H, int32_t S, int32_t L, cqmm_mcw_result *res);
  // Admit both spin channels of one McWMin (nspin = 2, one context): fixed or
  // variable spin populations.
  int add_spin(cqmm_ctx *ctx, const cqmm_mcw_setup &su, bool fixed, const int32_t H[2], const int32_t S[2],
               const int32_t L[2], cqmm_mcw_result res[2]);
  // One McWeeny step of every admitted system; systems that stopped are
  // appended to `finished` (ownership passes to the caller: job_delete).
  int step(std::vector<McwJob *> *finished);
  int active() const { return (int)jobs_.size(); }
  int64_t graph_steps() const { return graph_steps_; }
  int64_t captures() const { return captures_; }
  static int job_status(const McwJob *j);
  static cqmm_ctx *job_ctx(const McwJob *j);
  static void job_delete(McwJob *j);

 private:
  int add_one(cqmm_ctx *ctx, const cqmm_mcw_setup &su, int32_t H, int32_t S, int32_t L, cqmm_mcw_result *res,
              const cqmm_ctx *shared);
  int enqueue_step();
  int products(int which);
  int sum(cqmm_ctx *ctx, double a, const double *ap, int32_t A, double b, const double *bp, int32_t B);
  int trace(cqmm_ctx *ctx, int32_t A, int32_t B, double *d_out);
  int accept(McwJob &j);
  int after_step(McwJob &j);
  int after_fixed_pair(McwJob &a);
  int finish(McwJob &j);
  double *dsc(const McwJob &j) const;
  const double *hsc(const McwJob &j) const;
  std::string key() const;
  void drop_graph();

  hipStream_t s_ = nullptr;
  bool use_graph_ = true, group_ = true;
  std::vector<std::unique_ptr<McwJob>> jobs_;
  std::vector<bool> slot_used_;
  DevBuf dsc_;
  double *hsc_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  std::string graph_key_, last_eager_key_;
  int64_t graph_steps_ = 0, captures_ = 0;
  uint64_t serial_ = 0;  // admission count: names a system in the graph key
};

}  // namespace cq
