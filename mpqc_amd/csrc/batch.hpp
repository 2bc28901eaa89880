// Residuals of many independent molecules in grouped launches.
//
// Every molecule is a CCSDContext (integrals, amplitudes and intermediates in
// HBM).  Each molecule's residual is recorded once as an op tape without being
// run; the tapes have the same op sequence (the same code path at different
// sizes), so op k of all molecules becomes one grouped launch: a persistent
// grouped contraction over every molecule's output tiles, or one grouped
// permutation / copy / tau build.  The launch sequence is captured in a
// hipGraph once and replayed for every batched residual.  Each grouped
// contraction's tile edge and split-K target are chosen by an autotuner that
// times the candidates on the batch itself.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "ccsd_ctx.hpp"

namespace mpqc_hip {

class CCSDBatch {
 public:
  CCSDBatch(int n, const int* o, const int* v, int precision);
  ~CCSDBatch();
  int size() const { return (int)mols_.size(); }
  CCSDContext& mol(int k);

  // one residual per molecule at its current amplitudes (graph replay)
  void residual();
  // times every candidate configuration of every grouped contraction and keeps
  // the fastest; returns the batched residual time (ms) after tuning
  double tune(int reps, int verbose);
  // mean time (ms) of `reps` batched residuals
  double time(int reps);

  struct StageInfo {
    std::string tag;
    int kind;      // CCSDContext::TapeOp kind
    int bm;        // contraction tile edge (contractions only)
    long target;   // split-K work-item target (contractions only)
    double ms;     // last timing (tune or time_stages)
    double flops;  // sum of 2 M N K batch over the molecules (contractions)
  };
  std::vector<StageInfo> stages();
  // per-stage timing of one batched residual (each stage launched alone)
  void time_stages(int reps);

 private:
  struct Stage {
    int kind = 0;
    std::string tag;
    std::vector<GemmDesc> gemms;
    GemmPrecision prec = GemmPrecision::F64;
    int bm = 64;
    long target = 2048;
    GroupedGemmPlan gp;
    std::vector<GroupedTAddDesc> tadds;
    std::vector<GroupedEwDesc> ews;
    GroupedPlan plan;
    double ms = 0.0;
    double flops = 0.0;
  };
  std::vector<std::unique_ptr<CCSDContext>> mols_;
  hipStream_t stream_ = nullptr;
  std::vector<Stage> stages_;
  double* ws_ = nullptr;
  size_t ws_n_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  bool planned_ = false;
  GemmPrecision prec_ = GemmPrecision::F64;

  void plan();
  void ensure_planned();
  void build_plan(Stage& s);
  void launch_stage(const Stage& s);
  void capture();
  void release_graph();
  double time_stage(const Stage& s, int reps);
};

// Candidate configurations the autotuner tries for each grouped contraction.
struct GroupedConfig {
  int bm;
  long target;
};
const std::vector<GroupedConfig>& grouped_candidates();

}  // namespace mpqc_hip
