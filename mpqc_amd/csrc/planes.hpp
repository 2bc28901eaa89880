// Mixed-precision operands kept in HBM as 16-bit "planes".
//
// A tensor x (FP64) is stored as NP half-precision planes p_0..p_{NP-1} with the
// same element layout, such that  x ~= (p_0 + p_1 [+ p_2]) / s  where s is a
// power of two kept in device memory (s = 1 for bf16).  Contractions then read
// 2*NP bytes per element instead of 8 and need no conversion in the GEMM.
#pragma once

#include <cstdint>

#include "common.hpp"
#include "tensor_ops.hpp"

namespace mpqc_hip {

enum class PlaneKind : int {
  BF16x2 = 0,  // hi + lo bf16, products hh + hl + lh
  F16x2 = 1,   // hi + lo fp16 of the power-of-two scaled value, products hh + hl + lh
  BF16x3 = 2,  // hi + mid + lo bf16, six products
};

inline int plane_count(PlaneKind k) { return k == PlaneKind::BF16x3 ? 3 : 2; }

struct Planes {
  uint16_t* p[3] = {nullptr, nullptr, nullptr};
  double* scale = nullptr;  // device scalar (power of two; 1 for bf16)
  long n = 0;               // elements per plane
};

// Writes planes of  out[i0,i1,i2,i3] = sum_t alpha_t in_t[...]  (permuted linear
// combination, as tadd) with extents n; computes the scale on the device first
// (max |out| over the tensor).  `partials` needs kReducePartials doubles.
// Device-resident max |x| of inputs that several splits read (valid while the
// inputs are unchanged): a split finds an input here, reuses its max and
// skips that input's max pass, or records the max it computes.
struct AbsMaxCache {
  static constexpr int kSlots = 8;
  const double* ptr[kSlots] = {};
  double* dev = nullptr;  // kSlots doubles
  bool valid[kSlots] = {};
  void clear() { for (bool& v : valid) v = false; }
};
void split_to_planes(PlaneKind kind, const Planes& out, const long n[4], const TTerm* terms,
                     int nterms, double* partials, hipStream_t stream, const TProd* prod = nullptr,
                     AbsMaxCache* cache = nullptr);

// C[m,n] (FP64) = alpha * sum_k A[m,k] B[n,k] / (sA sB) + beta C, batched, with A and
// B given as planes in k-contiguous layout (A: row m at A + m*lda, B: row n at
// B + n*ldb).  Split-K partial sums go to `ws` when the output is small.
struct PlaneGemmDesc {
  PlaneKind kind;
  int M, N, K;
  double alpha, beta;
  Planes A;
  long lda, sA;
  Planes B;
  long ldb, sB;
  double* C;
  long ldc, sC;
  int batch = 1;
  // optional: max |C| over the M x N result into *cmax (device scalar; cmax_partials: at
  // least 8192 doubles), where the result is written by the split-K pass
  double* cmax = nullptr;
  double* cmax_partials = nullptr;
};
// Returns whether *d.cmax was written.
bool gemm_planes(const PlaneGemmDesc& d, double* ws, size_t ws_bytes, hipStream_t stream);
// Forces the contraction tile (bm x bn; 0, 0 restores the planner's choice).
// Used by measurements that compare tile shapes.
void set_plane_tile_override(int bm, int bn);
// LDS stages of the 128 x 128 two-plane contraction (2: double buffered, two
// workgroups per CU; 1: single buffered, four workgroups per CU).
void set_plane_lds_stages(int stages);
// Selects the contraction variant that loads B fragments straight into registers.
void set_plane_bdirect(bool on);
// Tile order of the contraction grid: 0 XCD ranges of row-major tiles, 1 plain row-major,
// 2 XCD ranges of column-major tiles.
void set_plane_tile_order(int order);
// process-wide: most K splits of a small-grid plane GEMM (default 512)
void set_plane_split_cap(int cap);
// process-wide: 0 lets the planner use the deep-pipeline 256 x 256 two-plane kernel for
// large contractions, 1 keeps every contraction on the 64/128-tile kernel
void set_plane_dl_mode(int mode);
// Workspace (bytes) gemm_planes needs for its split-K partial sums at this shape.
size_t gemm_planes_workspace_bytes(int M, int N, int K, int batch);

}  // namespace mpqc_hip
