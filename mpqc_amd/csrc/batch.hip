// Grouped (batched) CCSD residuals; see batch.hpp.
#include "batch.hpp"

#include <cstdio>
#include <stdexcept>

namespace mpqc_hip {

namespace {

TTerm term_from(const double* p, const std::array<long, 4>& ext, const std::array<int, 4>& perm,
                double alpha) {
  TTerm t;
  t.p = p;
  long e[4] = {ext[0], ext[1], ext[2], ext[3]};
  int pm[4] = {perm[0], perm[1], perm[2], perm[3]};
  permuted_strides(e, pm, t.s);
  t.alpha = alpha;
  t.vol = ext[0] * ext[1] * ext[2] * ext[3];
  return t;
}

}  // namespace

const std::vector<GroupedConfig>& grouped_candidates() {
  static const std::vector<GroupedConfig> c = {{64, 1024}, {64, 4096}, {64, 16384},
                                               {128, 1024}, {128, 4096}, {128, 16384}};
  return c;
}

CCSDBatch::CCSDBatch(int n, const int* o, const int* v, int precision) {
  if (n <= 0) throw std::invalid_argument("batch needs at least one molecule");
  if (!o || !v) throw std::invalid_argument("null size arrays");
  if (precision != 0 && precision != 1)
    throw std::invalid_argument("batched residuals run in the f64 and f32 forms");
  prec_ = (GemmPrecision)precision;
  try {
    for (int k = 0; k < n; ++k)
      mols_.emplace_back(new CCSDContext(o[k], v[k], nullptr, 0, nullptr, 0, precision, false));
    MPQC_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  } catch (...) {
    mols_.clear();
    throw;
  }
}

CCSDBatch::~CCSDBatch() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  release_graph();
  for (Stage& s : stages_) {
    free_grouped_gemm(s.gp);
    free_grouped(s.plan);
  }
  if (ws_) (void)hipFree(ws_);
  mols_.clear();
  if (stream_) (void)hipStreamDestroy(stream_);
}

CCSDContext& CCSDBatch::mol(int k) {
  if (k < 0 || k >= size()) throw std::invalid_argument("molecule index out of range");
  return *mols_[k];
}

void CCSDBatch::release_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

void CCSDBatch::plan() {
  for (Stage& s : stages_) {
    free_grouped_gemm(s.gp);
    free_grouped(s.plan);
  }
  stages_.clear();
  // derived integral layouts (run per molecule), then the tapes (recorded only)
  std::vector<std::vector<CCSDContext::TapeOp>> tapes(mols_.size());
  for (size_t k = 0; k < mols_.size(); ++k) {
    CCSDContext& c = *mols_[k];
    c.prepare();
    c.tape_ = &tapes[k];
    c.tape_execute_ = false;
    try {
      c.residual();
    } catch (...) {
      c.tape_ = nullptr;
      c.tape_execute_ = true;
      throw;
    }
    c.tape_ = nullptr;
    c.tape_execute_ = true;
  }
  const size_t nops = tapes[0].size();
  for (const auto& t : tapes)
    if (t.size() != nops) throw std::logic_error("batch: residual tapes differ in length");
  size_t ws_need = 0;
  for (size_t i = 0; i < nops; ++i) {
    Stage st;
    const CCSDContext::TapeOp& op0 = tapes[0][i];
    st.kind = op0.kind;
    st.tag = op0.kind == CCSDContext::TapeOp::kGemm ? op0.tag
             : op0.kind == CCSDContext::TapeOp::kTAdd ? op0.tag
             : op0.kind == CCSDContext::TapeOp::kCopy ? "copy"
             : op0.kind == CCSDContext::TapeOp::kTau ? "tau" : "TP";
    for (size_t k = 0; k < tapes.size(); ++k) {
      const CCSDContext::TapeOp& op = tapes[k][i];
      if (op.kind != op0.kind || op.tag != op0.tag) throw std::logic_error("batch: residual tapes differ");
      const CCSDContext& c = *mols_[k];
      switch (op.kind) {
        case CCSDContext::TapeOp::kGemm:
          st.gemms.push_back(op.g);
          st.prec = op.prec;
          st.flops += 2.0 * op.g.M * (double)op.g.N * op.g.K * op.g.batch;
          break;
        case CCSDContext::TapeOp::kTAdd: {
          GroupedTAddDesc d;
          d.out = op.dst;
          for (int q = 0; q < 4; ++q) d.n[q] = op.ext[q];
          d.beta = op.beta;
          d.nterms = op.nterms;
          for (int t = 0; t < op.nterms; ++t)
            d.terms[t] = term_from(op.terms[t].p, op.terms[t].ext, op.terms[t].perm, op.terms[t].alpha);
          st.tadds.push_back(d);
          break;
        }
        case CCSDContext::TapeOp::kCopy:
          st.ews.push_back({0, op.dst, op.src, nullptr, op.n, c.o_, c.v_});
          break;
        case CCSDContext::TapeOp::kTau:
        case CCSDContext::TapeOp::kTP: {
          const long n = (long)c.o_ * c.o_ * c.v_ * c.v_;
          st.ews.push_back({op.kind == CCSDContext::TapeOp::kTau ? 1 : 2, op.dst, op.src, op.src2, n,
                            c.o_, c.v_});
          break;
        }
      }
    }
    if (st.kind == CCSDContext::TapeOp::kGemm)
      for (const GroupedConfig& cf : grouped_candidates())
        ws_need = std::max(ws_need, grouped_gemm_workspace(st.gemms, cf.bm, cf.target));
    stages_.push_back(std::move(st));
  }
  if (ws_need > ws_n_) {
    if (ws_) MPQC_HIP_CHECK(hipFree(ws_));
    ws_ = nullptr;
    MPQC_HIP_CHECK(hipMalloc(&ws_, std::max<size_t>(ws_need, 1) * sizeof(double)));
    ws_n_ = std::max<size_t>(ws_need, 1);
  }
  for (Stage& s : stages_) build_plan(s);
  planned_ = true;
}

void CCSDBatch::build_plan(Stage& s) {
  free_grouped_gemm(s.gp);
  free_grouped(s.plan);
  switch (s.kind) {
    case CCSDContext::TapeOp::kGemm:
      s.gp = plan_grouped_gemm(s.gemms, s.prec, s.bm, s.target, ws_, ws_n_, stream_);
      break;
    case CCSDContext::TapeOp::kTAdd:
      s.plan = plan_grouped_tadd(s.tadds, stream_);
      break;
    default:
      s.plan = plan_grouped_ew(s.ews, stream_);
      break;
  }
}

void CCSDBatch::launch_stage(const Stage& s) {
  if (s.kind == CCSDContext::TapeOp::kGemm) launch_grouped_gemm(s.gp, stream_);
  else launch_grouped(s.plan, stream_);
}

void CCSDBatch::capture() {
  release_graph();
  MPQC_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  try {
    for (const Stage& s : stages_) launch_stage(s);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  MPQC_HIP_CHECK(hipStreamEndCapture(stream_, &graph_));
  MPQC_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
}

void CCSDBatch::ensure_planned() {
  // new integrals on any molecule (set_integral clears its prepared flag): its
  // derived layouts and possibly buffer addresses change, so plan again
  for (auto& m : mols_)
    if (!m->prepared_) planned_ = false;
  if (!planned_) {
    release_graph();
    plan();
  }
}

void CCSDBatch::residual() {
  ensure_planned();
  if (!exec_) capture();
  // the molecules' own streams may still carry uploads
  for (auto& m : mols_) MPQC_HIP_CHECK(hipStreamSynchronize(m->stream_));
  MPQC_HIP_CHECK(hipGraphLaunch(exec_, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
}

double CCSDBatch::time_stage(const Stage& s, int reps) {
  hipEvent_t a, b;
  MPQC_HIP_CHECK(hipEventCreate(&a));
  MPQC_HIP_CHECK(hipEventCreate(&b));
  launch_stage(s);  // warm-up
  MPQC_HIP_CHECK(hipEventRecord(a, stream_));
  for (int r = 0; r < reps; ++r) launch_stage(s);
  MPQC_HIP_CHECK(hipEventRecord(b, stream_));
  MPQC_HIP_CHECK(hipEventSynchronize(b));
  float ms = 0.f;
  MPQC_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

double CCSDBatch::tune(int reps, int verbose) {
  if (reps < 1) throw std::invalid_argument("reps must be positive");
  ensure_planned();
  release_graph();
  for (Stage& s : stages_) {
    if (s.kind != CCSDContext::TapeOp::kGemm) continue;
    double best = 1e300;
    GroupedConfig best_cf{s.bm, s.target};
    for (const GroupedConfig& cf : grouped_candidates()) {
      s.bm = cf.bm;
      s.target = cf.target;
      build_plan(s);
      const double ms = time_stage(s, reps);
      if (verbose) std::printf("  tune %-24s bm=%3d target=%6ld  %.4f ms\n", s.tag.c_str(), cf.bm, cf.target, ms);
      if (ms < best) {
        best = ms;
        best_cf = cf;
      }
    }
    s.bm = best_cf.bm;
    s.target = best_cf.target;
    s.ms = best;
    build_plan(s);
  }
  return time(reps);
}

double CCSDBatch::time(int reps) {
  if (reps < 1) throw std::invalid_argument("reps must be positive");
  residual();  // plans, captures and warms up
  hipEvent_t a, b;
  MPQC_HIP_CHECK(hipEventCreate(&a));
  MPQC_HIP_CHECK(hipEventCreate(&b));
  MPQC_HIP_CHECK(hipEventRecord(a, stream_));
  for (int r = 0; r < reps; ++r) MPQC_HIP_CHECK(hipGraphLaunch(exec_, stream_));
  MPQC_HIP_CHECK(hipEventRecord(b, stream_));
  MPQC_HIP_CHECK(hipEventSynchronize(b));
  float ms = 0.f;
  MPQC_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

void CCSDBatch::time_stages(int reps) {
  ensure_planned();
  for (Stage& s : stages_) s.ms = time_stage(s, reps);
}

std::vector<CCSDBatch::StageInfo> CCSDBatch::stages() {
  ensure_planned();
  std::vector<StageInfo> out;
  for (const Stage& s : stages_) out.push_back({s.tag, s.kind, s.bm, s.target, s.ms, s.flops});
  return out;
}

}  // namespace mpqc_hip
