// Batched, strided, split-K GEMM on CDNA4 matrix cores (FP64 MFMA).
//
// Row-major semantics: C[m,n] = alpha * sum_k op(A)[m,k] op(B)[k,n] + beta * C[m,n]
//   op(A) = A      : A[m*lda + k]       op(A) = A^T : A[k*lda + m]
//   op(B) = B      : B[k*ldb + n]       op(B) = B^T : B[n*ldb + k]
// Batch b uses A + b*sA, B + b*sB, C + b*sC.
//
// This is the contraction engine that replaces TiledArray's SUMMA +
// dgemm for the CCSD residual (reference: src/mpqc/math/external/tiledarray/
// tiledarray_impl.cpp:36-47, which instantiates TA's ContractReduce over
// TA::Tensor<double>).  Every "*" in ccsd_r1_r2.h:70-293 lands here, after the
// operands have been brought into GEMM layout by the permutation kernels.
#pragma once

#include <vector>

#include "common.hpp"

namespace mpqc_hip {

enum class GemmPrecision : int {
  F64 = 0,      // FP64 MFMA (v_mfma_f64_16x16x4_f64), FP64 accumulate
  F32 = 1,      // FP32 MFMA on FP32-rounded operands, FP32 accumulate per K-chunk, FP64 across chunks
  BF16X3 = 2,   // bf16 hi+lo split, 3 products (hh, hl, lh) on bf16 MFMA (~16 bits)
  F16X3 = 3,    // fp16 hi+lo split of power-of-two scaled operands, 3 products (~22 bits)
  BF16X6 = 4,   // bf16 hi+mid+lo split, 6 products (~24 bits)
};

struct GemmDesc {
  bool transA = false, transB = false;
  int M = 0, N = 0, K = 0;
  double alpha = 1.0, beta = 0.0;
  const double* A = nullptr;
  long lda = 0, sA = 0;
  const double* B = nullptr;
  long ldb = 0, sB = 0;
  double* C = nullptr;
  long ldc = 0, sC = 0;
  int batch = 1;
};

// Workspace used for split-K partial sums (owned by the caller).
struct GemmWorkspace {
  double* ptr = nullptr;
  size_t bytes = 0;
};

// Launches on `stream`.  Chooses tile shape and split-K factor.  Deterministic:
// split-K partials are summed in a fixed order by a separate pass.
void gemm(const GemmDesc& d, GemmPrecision prec, const GemmWorkspace& ws, hipStream_t stream);

// C[b] = beta*C[b] + sum_s W[(b - batch0)*nsplit + s]  for batches batch0..batch0+nb-1, with the
// M x N partials W dense; fixed summation order (bitwise reproducible).
void splitk_reduce(const double* W, double* C, long ldc, long sC, int M, int N, int nsplit, int batch0,
                   int nb, double beta, hipStream_t stream);

// ---- grouped contractions: many independent problems with the same transposes
// and precision (F64 or F32) in one persistent launch.  Work items are
// (problem, batch x split slice, output tile); a workgroup walks the item list
// with stride gridDim.x and finds each item's problem by binary search over the
// problems' first items.  Problems whose split-K factor is above one leave
// partial sums in the workspace, summed by a grouped reduce launch.
struct GroupedGemmPlan {
  int bm = 64;             // square tile edge (64 or 128)
  bool transA = false, transB = false;
  GemmPrecision prec = GemmPrecision::F64;
  int nprob = 0;
  long items = 0;          // total work items of the contraction launch
  long reduce_items = 0;   // total 256-element chunks of the reduce launch
  int grid = 0;            // persistent workgroups
  void* dev = nullptr;     // device array of problem records
  size_t ws_doubles = 0;   // split-K workspace used
};
// Workspace (doubles) the plan for (bm, target_items) needs.
size_t grouped_gemm_workspace(const std::vector<GemmDesc>& probs, int bm, long target_items);
// Builds the plan and uploads the problem records (synchronous on `stream`).
GroupedGemmPlan plan_grouped_gemm(const std::vector<GemmDesc>& probs, GemmPrecision prec, int bm,
                                  long target_items, double* ws, size_t ws_doubles,
                                  hipStream_t stream);
void launch_grouped_gemm(const GroupedGemmPlan& plan, hipStream_t stream);
void free_grouped_gemm(GroupedGemmPlan& plan);

// Split-precision paths (BF16X3, F16X3, BF16X6); called by gemm().
void gemm_split(const GemmDesc& d, GemmPrecision prec, const GemmWorkspace& ws, hipStream_t stream);

// Bytes of the workspace that gemm(d, prec, {ptr, ws_bytes}, ...) reads or writes
// (split-K partial sums, and the split forms' reduction head); 0 when it uses none.
size_t gemm_scratch_bytes(const GemmDesc& d, GemmPrecision prec, size_t ws_bytes);
size_t gemm_split_scratch_bytes(const GemmDesc& d, GemmPrecision prec, size_t ws_bytes);

// FP64 contractions with K <= 64, A of at most 256 rows and a long B (no
// transposes) run on a streaming kernel that keeps A in registers; mode 1 sends
// them to the tiled kernels instead (mode 0, the default, restores it).
void set_thin_k_mode(int mode);

// Convenience for plain (non-batched) calls.
inline GemmDesc make_gemm(bool tA, bool tB, int M, int N, int K, double alpha, const double* A,
                          long lda, const double* B, long ldb, double beta, double* C, long ldc) {
  GemmDesc d;
  d.transA = tA; d.transB = tB; d.M = M; d.N = N; d.K = K; d.alpha = alpha; d.beta = beta;
  d.A = A; d.lda = lda; d.B = B; d.ldb = ldb; d.C = C; d.ldc = ldc;
  return d;
}

}  // namespace mpqc_hip
