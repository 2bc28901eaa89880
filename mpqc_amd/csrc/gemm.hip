// Batched strided split-K GEMM on gfx950 matrix cores.
//
// Tile: BM x BN per workgroup of 256 threads (4 wave64s arranged 2x2), each
// wave owns a (BM/2) x (BN/2) sub-tile made of 16x16 MFMA blocks.  K is staged
// through LDS in BK=16 slices, double buffered: the next slice is fetched into
// registers while the matrix cores work on the current one.
//
// FP64 path: v_mfma_f64_16x16x4_f64 (one double of A and of B per lane per
// instruction, 4 doubles of C per lane).
//
// Operand fragments for a 16x16x4 block (lane l, r = 0..3):
//   A[m = l%16][k = l/16], B[k = l/16][n = l%16], C[m = 4*r + l/16][n = l%16]
// (the accumulator layout was established on the GPU; tests/test_gpu_gemm.py
// checks every tile position through full GEMMs).
#include "gemm.hpp"

#include <algorithm>

namespace mpqc_hip {

namespace {

constexpr int kBK = 16;
constexpr int kThreads = 256;

struct KParams {
  const double* A;
  const double* B;
  double* C;
  double* W;  // split-K workspace (nsplit > 1)
  long lda, ldb, ldc, sA, sB, sC;
  int M, N, K;
  int kchunk;
  int nsplit;
  int nsplit_batch;  // batch entries of a grouped problem (reduce pass)
  int batch0;  // first batch index of this launch
  int tiles_m, tiles_n;
  int per_xcd;  // tiles per XCD (workgroup w runs on XCD w % 8)
  double alpha, beta;
};

constexpr int kXcdCount = 8;
constexpr int kXcdMinTiles = 64;  // fewer output tiles: plain tile order

// XCD-aware tile order: workgroups are dispatched round-robin over the 8 XCDs;
// giving each XCD a contiguous range of row-major tiles keeps the operand rows
// those tiles share inside one XCD's L2.  Returns -1 for padding workgroups.
// With few tiles (per_xcd = 0) the order is plain: a launch of many K splits
// over fewer than eight tiles would otherwise put every split on one XCD.
__device__ inline int xcd_tile(const KParams& p) {
  const int w = blockIdx.x;
  const int t = p.per_xcd == 0 ? w : (w % kXcdCount) * p.per_xcd + w / kXcdCount;
  return t < p.tiles_m * p.tiles_n ? t : -1;
}

// One BM x BN output tile `tile` (row-major tile index) of batch/split slice z
// (z = batch * nsplit + split, batch relative to p.batch0).  Called by all 256
// threads of a workgroup.
template <int BM, int BN, bool TA, bool TB>
__device__ __forceinline__ void dgemm_f64_tile(const KParams& p, int tile, int z) {
  constexpr int PAD = 1;
  constexpr int LDA_S = BM + PAD;
  constexpr int LDB_S = BN + PAD;
  constexpr int A_PER = BM * kBK / kThreads;
  constexpr int B_PER = BN * kBK / kThreads;
  constexpr int WM = BM / 2, WN = BN / 2;  // wave sub-tile
  constexpr int RB = WM / 16, CB = WN / 16;

  __shared__ double As[2][kBK][LDA_S];
  __shared__ double Bs[2][kBK][LDB_S];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  const int tm = tile / p.tiles_n, tn = tile - tm * p.tiles_n;
  const int b = p.batch0 + z / p.nsplit;
  const int s = z - (z / p.nsplit) * p.nsplit;
  const int m0 = tm * BM, n0 = tn * BN;
  const int k_begin = s * p.kchunk;
  const int k_end = min(p.K, k_begin + p.kchunk);

  const double* __restrict__ A = p.A + (long)b * p.sA;
  const double* __restrict__ B = p.B + (long)b * p.sB;

  double ra[A_PER], rb[B_PER];

  auto load_regs = [&](int k0) {
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int e = tid + kThreads * r;
      int mm, kk;
      if (!TA) { mm = e / kBK; kk = e % kBK; } else { kk = e / BM; mm = e % BM; }
      const int gm = m0 + mm, gk = k0 + kk;
      double x = 0.0;
      if (gm < p.M && gk < k_end) x = TA ? A[(long)gk * p.lda + gm] : A[(long)gm * p.lda + gk];
      ra[r] = x;
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int e = tid + kThreads * r;
      int nn, kk;
      if (!TB) { kk = e / BN; nn = e % BN; } else { nn = e / kBK; kk = e % kBK; }
      const int gn = n0 + nn, gk = k0 + kk;
      double x = 0.0;
      if (gn < p.N && gk < k_end) x = TB ? B[(long)gn * p.ldb + gk] : B[(long)gk * p.ldb + gn];
      rb[r] = x;
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int e = tid + kThreads * r;
      int mm, kk;
      if (!TA) { mm = e / kBK; kk = e % kBK; } else { kk = e / BM; mm = e % BM; }
      As[buf][kk][mm] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int e = tid + kThreads * r;
      int nn, kk;
      if (!TB) { kk = e / BN; nn = e % BN; } else { nn = e / kBK; kk = e % kBK; }
      Bs[buf][kk][nn] = rb[r];
    }
  };

  double4_t acc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

  int buf = 0;
  if (k_begin < k_end) {
    load_regs(k_begin);
    store_lds(0);
  }
  __syncthreads();
  for (int kt = k_begin; kt < k_end; kt += kBK) {
    const bool has_next = kt + kBK < k_end;
    if (has_next) load_regs(kt + kBK);
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
      double fa[RB], fb[CB];
#pragma unroll
      for (int i = 0; i < RB; ++i) fa[i] = As[buf][ks * 4 + lk][wm * WM + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < CB; ++j) fb[j] = Bs[buf][ks * 4 + lk][wn * WN + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
    if (has_next) store_lds(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // epilogue
  if (p.nsplit == 1) {
    double* __restrict__ C = p.C + (long)b * p.sC;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int j = 0; j < CB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gm = m0 + wm * WM + i * 16 + r * 4 + lk;
          const int gn = n0 + wn * WN + j * 16 + lr;
          if (gm < p.M && gn < p.N) {
            double* c = C + (long)gm * p.ldc + gn;
            const double v = p.alpha * acc[i][j][r];
            *c = (p.beta == 0.0) ? v : v + p.beta * (*c);
          }
        }
  } else {
    // partial sums to workspace slice z (dense M x N)
    double* __restrict__ Wz = p.W + (long)z * p.M * p.N;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int j = 0; j < CB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gm = m0 + wm * WM + i * 16 + r * 4 + lk;
          const int gn = n0 + wn * WN + j * 16 + lr;
          if (gm < p.M && gn < p.N) Wz[(long)gm * p.N + gn] = p.alpha * acc[i][j][r];
        }
  }
}

template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(kThreads, 2) void dgemm_f64_kernel(KParams p) {
  const int tile = xcd_tile(p);
  if (tile < 0) return;
  dgemm_f64_tile<BM, BN, TA, TB>(p, tile, blockIdx.y);
}

// FP32 path: operands are rounded to FP32 when staged into LDS and multiplied
// on v_mfma_f32_16x16x4_f32 (same operand fragments as the FP64 instruction;
// accumulator layout C[m = 4*(l/16) + r][n = l%16]).  Partial sums are kept in
// FP32 for kFlush consecutive K slices (64 values of k) and then added into
// FP64 accumulators, so long contractions (K = v^2) lose no precision to FP32
// accumulation.
template <int BM, int BN, bool TA, bool TB>
__device__ __forceinline__ void sgemm_f32acc64_tile(const KParams& p, int tile, int z) {
  constexpr int PAD = 1;
  constexpr int LDA_S = BM + PAD;
  constexpr int LDB_S = BN + PAD;
  constexpr int A_PER = BM * kBK / kThreads;
  constexpr int B_PER = BN * kBK / kThreads;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int RB = WM / 16, CB = WN / 16;
  constexpr int kFlush = 4;

  __shared__ float As[2][kBK][LDA_S];
  __shared__ float Bs[2][kBK][LDB_S];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  const int tm = tile / p.tiles_n, tn = tile - tm * p.tiles_n;
  const int b = p.batch0 + z / p.nsplit;
  const int s = z - (z / p.nsplit) * p.nsplit;
  const int m0 = tm * BM, n0 = tn * BN;
  const int k_begin = s * p.kchunk;
  const int k_end = min(p.K, k_begin + p.kchunk);

  const double* __restrict__ A = p.A + (long)b * p.sA;
  const double* __restrict__ B = p.B + (long)b * p.sB;

  float ra[A_PER], rb[B_PER];
  auto load_regs = [&](int k0) {
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int e = tid + kThreads * r;
      int mm, kk;
      if (!TA) { mm = e / kBK; kk = e % kBK; } else { kk = e / BM; mm = e % BM; }
      const int gm = m0 + mm, gk = k0 + kk;
      double x = 0.0;
      if (gm < p.M && gk < k_end) x = TA ? A[(long)gk * p.lda + gm] : A[(long)gm * p.lda + gk];
      ra[r] = (float)x;
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int e = tid + kThreads * r;
      int nn, kk;
      if (!TB) { kk = e / BN; nn = e % BN; } else { nn = e / kBK; kk = e % kBK; }
      const int gn = n0 + nn, gk = k0 + kk;
      double x = 0.0;
      if (gn < p.N && gk < k_end) x = TB ? B[(long)gn * p.ldb + gk] : B[(long)gk * p.ldb + gn];
      rb[r] = (float)x;
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int e = tid + kThreads * r;
      int mm, kk;
      if (!TA) { mm = e / kBK; kk = e % kBK; } else { kk = e / BM; mm = e % BM; }
      As[buf][kk][mm] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int e = tid + kThreads * r;
      int nn, kk;
      if (!TB) { kk = e / BN; nn = e % BN; } else { nn = e / kBK; kk = e % kBK; }
      Bs[buf][kk][nn] = rb[r];
    }
  };

  double4_t acc[RB][CB];
  float4_t facc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
      facc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
    }

  int buf = 0, since_flush = 0;
  if (k_begin < k_end) {
    load_regs(k_begin);
    store_lds(0);
  }
  __syncthreads();
  for (int kt = k_begin; kt < k_end; kt += kBK) {
    const bool has_next = kt + kBK < k_end;
    if (has_next) load_regs(kt + kBK);
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
      float fa[RB], fb[CB];
#pragma unroll
      for (int i = 0; i < RB; ++i) fa[i] = As[buf][ks * 4 + lk][wm * WM + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < CB; ++j) fb[j] = Bs[buf][ks * 4 + lk][wn * WN + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j)
          facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[i], fb[j], facc[i][j], 0, 0, 0);
    }
    if (++since_flush == kFlush || !has_next) {
      since_flush = 0;
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)facc[i][j][r];
          facc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
        }
    }
    if (has_next) store_lds(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  double* __restrict__ C = p.C + (long)b * p.sC;
  double* __restrict__ Wz = p.W + (long)z * p.M * p.N;
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * WM + i * 16 + lk * 4 + r;
        const int gn = n0 + wn * WN + j * 16 + lr;
        if (gm < p.M && gn < p.N) {
          const double v = p.alpha * acc[i][j][r];
          if (p.nsplit == 1) {
            double* c = C + (long)gm * p.ldc + gn;
            *c = (p.beta == 0.0) ? v : v + p.beta * (*c);
          } else {
            Wz[(long)gm * p.N + gn] = v;
          }
        }
      }
}

template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(kThreads, 2) void sgemm_f32acc64_kernel(KParams p) {
  const int tile = xcd_tile(p);
  if (tile < 0) return;
  sgemm_f32acc64_tile<BM, BN, TA, TB>(p, tile, blockIdx.y);
}

// C[b] = beta*C[b] + sum_s W[b*nsplit + s]   (fixed summation order)
__global__ void splitk_reduce_kernel(const double* __restrict__ W, double* __restrict__ C, long ldc,
                                     long sC, int M, int N, int nsplit, int batch0, double beta) {
  const long MN = (long)M * N;
  const int bz = blockIdx.y;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < MN; e += (long)gridDim.x * blockDim.x) {
    const int m = (int)(e / N), n = (int)(e - (long)(e / N) * N);
    double acc = 0.0;
    const double* w = W + (long)bz * nsplit * MN + e;
    for (int s = 0; s < nsplit; ++s) acc += w[(long)s * MN];
    double* c = C + (long)(batch0 + bz) * sC + (long)m * ldc + n;
    *c = (beta == 0.0) ? acc : acc + beta * (*c);
  }
}

// Many partials over a small output (nsplit >= kDeepReduceSplits, M*N*batch below
// kWideReduceMax): a wave owns one output, lane l adds the partials s = l, l + 64, ...
// (a dependent chain of nsplit / 64 loads instead of nsplit / 4), and the lanes
// combine by a fixed butterfly (every lane ends with the same sum).
constexpr int kDeepReduceSplits = 64;
__global__ void splitk_reduce_deep_kernel(const double* __restrict__ W, double* __restrict__ C, long ldc,
                                          long sC, int M, int N, int nsplit, int batch0, int nb, double beta) {
  const long MN = (long)M * N;
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave >= MN * nb) return;
  const int bz = (int)(wave / MN);
  const long e = wave - (long)bz * MN;
  const double* w = W + (long)bz * nsplit * MN + e;
  double acc = 0.0;
  for (int s = lane; s < nsplit; s += 64) acc += w[(long)s * MN];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) {
    const int m = (int)(e / N), n = (int)(e - (long)m * N);
    double* c = C + (long)(batch0 + bz) * sC + (long)m * ldc + n;
    *c = (beta == 0.0) ? acc : acc + beta * (*c);
  }
}

// Same sum for small outputs (M*N*batch below kWideReduceMax): a wave owns 16
// consecutive outputs and its four 16-lane groups add every fourth partial
// (coalesced 128-byte rows), then the groups are combined in a fixed order.
constexpr long kWideReduceMax = 1L << 16;
__global__ void splitk_reduce_wide_kernel(const double* __restrict__ W, double* __restrict__ C,
                                          long ldc, long sC, int M, int N, int nsplit, int batch0,
                                          int nb, double beta) {
  const long MN = (long)M * N;
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long chunks = (MN + 15) / 16;
  if (wave >= chunks * nb) return;
  const int bz = (int)(wave / chunks);
  const long e = (wave - (long)bz * chunks) * 16 + (lane & 15);
  const int g = lane >> 4;
  double acc = 0.0;
  if (e < MN) {
    const double* w = W + (long)bz * nsplit * MN + e;
    for (int s = g; s < nsplit; s += 4) acc += w[(long)s * MN];
  }
  // groups (0,1) and (2,3), then the two pairs
  acc += __shfl_xor(acc, 16);
  acc += __shfl_xor(acc, 32);
  if (g == 0 && e < MN) {
    const int m = (int)(e / N), n = (int)(e - (long)m * N);
    double* c = C + (long)(batch0 + bz) * sC + (long)m * ldc + n;
    *c = (beta == 0.0) ? acc : acc + beta * (*c);
  }
}

// Split-K plan of one GEMM: K slices per output tile, their length, and how many
// batch entries share one launch (bounded by the workspace that holds the
// partial sums).
struct SplitKPlan {
  int nsplit = 1, kchunk = 0, max_batch_per_launch = 65535;
  size_t scratch_bytes = 0;  // workspace the launches touch
};

SplitKPlan plan_splitk(const GemmDesc& d, int BM, int BN, size_t ws_bytes) {
  SplitKPlan pl;
  const int tiles_m = (int)ceil_div(d.M, BM), tiles_n = (int)ceil_div(d.N, BN);
  const long tiles = (long)tiles_m * tiles_n;
  // split-K: aim for ~4 workgroups per CU (256 CUs) when the output is small,
  // with at least 16 K-slices (256 values of k) per split: shorter splits spend
  // their time in the prologue, the epilogue and the partial sums
  int nsplit = 1;
  const long want = 1024;
  if (tiles * d.batch < want && d.K >= 32 * kBK) {
    nsplit = (int)std::min<long>(ceil_div(want, tiles * d.batch), d.K / (16 * kBK));
    nsplit = std::max(1, std::min(nsplit, 512));
  }
  const size_t slice = (size_t)d.M * d.N * sizeof(double);
  // the workspace must hold nsplit partial outputs of at least one batch entry
  if (nsplit > 1 && ws_bytes / slice < (size_t)nsplit)
    nsplit = (int)std::max<size_t>(1, ws_bytes / slice);
  int kchunk = (int)(ceil_div(ceil_div(d.K, nsplit), kBK) * kBK);
  if (nsplit > 1) nsplit = (int)ceil_div(d.K, kchunk);
  if (nsplit == 1) kchunk = (int)(ceil_div(d.K, kBK) * kBK);
  int max_batch_per_launch = 65535 / nsplit;
  if (nsplit > 1)  // workspace capacity limits how many batches go in one launch
    max_batch_per_launch = (int)std::min<size_t>(max_batch_per_launch, ws_bytes / (slice * nsplit));
  if (d.K == 0) kchunk = kBK;
  pl.nsplit = nsplit;
  pl.kchunk = kchunk;
  pl.max_batch_per_launch = max_batch_per_launch;
  pl.scratch_bytes = nsplit > 1 ? slice * nsplit * (size_t)std::min(d.batch, max_batch_per_launch) : 0;
  return pl;
}

// 128-wide tiles unless 64-wide ones pad the output much less, or the
// 128-wide launch (after split-K) has too few workgroups to fill the chip (two
// per CU) and the 64-wide one has more: small outputs with a short K, where
// split-K cannot add parallelism.
bool wide_tiles(const GemmDesc& d, size_t ws_bytes) {
  bool wide = d.M > 64 && d.N > 64;
  if (wide) {
    const double a128 = (double)ceil_div(d.M, 128) * ceil_div(d.N, 128) * 128 * 128;
    const double a64 = (double)ceil_div(d.M, 64) * ceil_div(d.N, 64) * 64 * 64;
    wide = a128 <= 1.25 * a64;
  }
  if (wide) {
    auto wgs = [&](int bm) {
      return ceil_div(d.M, bm) * ceil_div(d.N, bm) * d.batch * (long)plan_splitk(d, bm, bm, ws_bytes).nsplit;
    };
    const long w128 = wgs(128);
    if (w128 < 512 && wgs(64) > w128) wide = false;
  }
  return wide;
}

template <int BM, int BN>
void launch(const GemmDesc& d, GemmPrecision prec, const GemmWorkspace& ws, hipStream_t stream) {
  const int tiles_m = (int)ceil_div(d.M, BM), tiles_n = (int)ceil_div(d.N, BN);
  const long tiles = (long)tiles_m * tiles_n;
  const SplitKPlan plan = plan_splitk(d, BM, BN, ws.bytes);
  const int nsplit = plan.nsplit, kchunk = plan.kchunk, max_batch_per_launch = plan.max_batch_per_launch;
  for (int b0 = 0; b0 < d.batch; b0 += max_batch_per_launch) {
    const int nb = std::min(max_batch_per_launch, d.batch - b0);
    KParams p;
    p.A = d.A; p.B = d.B; p.C = d.C; p.W = ws.ptr;
    p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc; p.sA = d.sA; p.sB = d.sB; p.sC = d.sC;
    p.M = d.M; p.N = d.N; p.K = d.K; p.kchunk = kchunk; p.nsplit = nsplit; p.batch0 = b0;
    p.tiles_m = tiles_m; p.tiles_n = tiles_n; p.alpha = d.alpha; p.beta = d.beta;
    p.per_xcd = tiles >= kXcdMinTiles ? (int)ceil_div(tiles, kXcdCount) : 0;
    dim3 grid((unsigned)(p.per_xcd ? p.per_xcd * kXcdCount : tiles), (unsigned)(nb * nsplit));
    if (prec == GemmPrecision::F32) {
      if (!d.transA && !d.transB) hipLaunchKernelGGL((sgemm_f32acc64_kernel<BM, BN, false, false>), grid, dim3(kThreads), 0, stream, p);
      else if (!d.transA && d.transB) hipLaunchKernelGGL((sgemm_f32acc64_kernel<BM, BN, false, true>), grid, dim3(kThreads), 0, stream, p);
      else if (d.transA && !d.transB) hipLaunchKernelGGL((sgemm_f32acc64_kernel<BM, BN, true, false>), grid, dim3(kThreads), 0, stream, p);
      else hipLaunchKernelGGL((sgemm_f32acc64_kernel<BM, BN, true, true>), grid, dim3(kThreads), 0, stream, p);
      check_launch("sgemm_f32acc64");
    } else {
      if (!d.transA && !d.transB) hipLaunchKernelGGL((dgemm_f64_kernel<BM, BN, false, false>), grid, dim3(kThreads), 0, stream, p);
      else if (!d.transA && d.transB) hipLaunchKernelGGL((dgemm_f64_kernel<BM, BN, false, true>), grid, dim3(kThreads), 0, stream, p);
      else if (d.transA && !d.transB) hipLaunchKernelGGL((dgemm_f64_kernel<BM, BN, true, false>), grid, dim3(kThreads), 0, stream, p);
      else hipLaunchKernelGGL((dgemm_f64_kernel<BM, BN, true, true>), grid, dim3(kThreads), 0, stream, p);
      check_launch("dgemm_f64");
    }
    if (nsplit > 1) splitk_reduce(ws.ptr, d.C, d.ldc, d.sC, d.M, d.N, nsplit, b0, nb, d.beta, stream);
  }
}

// ---------------------------------------------------------------- grouped
struct GProblem {
  KParams p;
  long first;   // first contraction work item
  long rfirst;  // first reduce chunk (problems with nsplit > 1)
};

__device__ inline int find_problem(const GProblem* __restrict__ probs, int nprob, long it, bool reduce) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((reduce ? probs[mid].rfirst : probs[mid].first) <= it) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int BM, int BN, bool TA, bool TB, bool F32>
__global__ __launch_bounds__(kThreads, 2) void grouped_gemm_kernel(const GProblem* __restrict__ probs,
                                                                   int nprob, long items) {
  for (long it = blockIdx.x; it < items; it += gridDim.x) {
    const int q = find_problem(probs, nprob, it, false);
    const KParams& p = probs[q].p;
    const long local = it - probs[q].first;
    const int ntiles = p.tiles_m * p.tiles_n;
    const int z = (int)(local / ntiles);
    const int tile = (int)(local - (long)z * ntiles);
    if (F32) sgemm_f32acc64_tile<BM, BN, TA, TB>(p, tile, z);
    else dgemm_f64_tile<BM, BN, TA, TB>(p, tile, z);
    __syncthreads();  // the next item reuses the LDS tiles
  }
}

// C[b] = beta C[b] + sum_s W[b*nsplit + s] for every problem with nsplit > 1;
// a chunk is 256 consecutive elements of one problem's batch x M x N output.
__global__ __launch_bounds__(256) void grouped_reduce_kernel(const GProblem* __restrict__ probs,
                                                             int nprob, long chunks) {
  for (long ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const int q = find_problem(probs, nprob, ch, true);
    const KParams& p = probs[q].p;
    const long MN = (long)p.M * p.N;
    const long e = (ch - probs[q].rfirst) * 256 + threadIdx.x;
    if (e >= MN * (long)p.nsplit_batch) continue;
    const long b = e / MN, mn = e - b * MN;
    const double* w = p.W + b * p.nsplit * MN + mn;
    double acc = 0.0;
    for (int s = 0; s < p.nsplit; ++s) acc += w[(long)s * MN];
    const long m = mn / p.N, n = mn - m * p.N;
    double* c = p.C + b * p.sC + m * p.ldc + n;
    *c = (p.beta == 0.0) ? acc : acc + p.beta * (*c);
  }
}

template <int BM>
void launch_grouped_bm(const GroupedGemmPlan& pl, hipStream_t stream) {
  const GProblem* d = (const GProblem*)pl.dev;
  const dim3 grid((unsigned)pl.grid);
#define MPQC_GROUPED(TA, TB)                                                                       \
  if (pl.prec == GemmPrecision::F32)                                                              \
    hipLaunchKernelGGL((grouped_gemm_kernel<BM, BM, TA, TB, true>), grid, dim3(kThreads), 0, stream, \
                       d, pl.nprob, pl.items);                                                     \
  else                                                                                            \
    hipLaunchKernelGGL((grouped_gemm_kernel<BM, BM, TA, TB, false>), grid, dim3(kThreads), 0, stream, \
                       d, pl.nprob, pl.items);
  if (!pl.transA && !pl.transB) { MPQC_GROUPED(false, false) }
  else if (!pl.transA && pl.transB) { MPQC_GROUPED(false, true) }
  else if (pl.transA && !pl.transB) { MPQC_GROUPED(true, false) }
  else { MPQC_GROUPED(true, true) }
#undef MPQC_GROUPED
  check_launch("grouped_gemm");
}

struct GroupedSplit {
  int nsplit, kchunk;
};

// split-K per problem: enough items for about target_items in total, at least
// four K slices per split
std::vector<GroupedSplit> grouped_splits(const std::vector<GemmDesc>& probs, int bm, long target) {
  long tiles_total = 0;
  for (const GemmDesc& d : probs)
    tiles_total += (long)ceil_div(d.M, bm) * ceil_div(d.N, bm) * d.batch;
  const long factor = tiles_total >= target ? 1 : ceil_div(target, std::max(1L, tiles_total));
  std::vector<GroupedSplit> out;
  for (const GemmDesc& d : probs) {
    int nsplit = (int)std::max(1L, std::min<long>(std::min<long>(factor, 512), d.K / (4 * kBK)));
    int kchunk = (int)(ceil_div(ceil_div(d.K, nsplit), kBK) * kBK);
    nsplit = (int)ceil_div(d.K, kchunk);
    if (nsplit <= 1) {
      nsplit = 1;
      kchunk = (int)(ceil_div(d.K, kBK) * kBK);
    }
    out.push_back({nsplit, kchunk});
  }
  return out;
}

}  // namespace

size_t grouped_gemm_workspace(const std::vector<GemmDesc>& probs, int bm, long target_items) {
  const auto sp = grouped_splits(probs, bm, target_items);
  size_t n = 0;
  for (size_t q = 0; q < probs.size(); ++q)
    if (sp[q].nsplit > 1) n += (size_t)sp[q].nsplit * probs[q].batch * probs[q].M * probs[q].N;
  return n;
}

GroupedGemmPlan plan_grouped_gemm(const std::vector<GemmDesc>& probs, GemmPrecision prec, int bm,
                                  long target_items, double* ws, size_t ws_doubles,
                                  hipStream_t stream) {
  if (probs.empty()) throw std::invalid_argument("grouped GEMM without problems");
  if (bm != 64 && bm != 128) throw std::invalid_argument("grouped GEMM tile must be 64 or 128");
  if (prec != GemmPrecision::F64 && prec != GemmPrecision::F32)
    throw std::invalid_argument("grouped GEMM runs in F64 or F32");
  GroupedGemmPlan pl;
  pl.bm = bm;
  pl.prec = prec;
  pl.transA = probs[0].transA;
  pl.transB = probs[0].transB;
  pl.nprob = (int)probs.size();
  const auto sp = grouped_splits(probs, bm, target_items);
  std::vector<GProblem> host(probs.size());
  size_t woff = 0;
  for (size_t q = 0; q < probs.size(); ++q) {
    const GemmDesc& d = probs[q];
    if (d.transA != pl.transA || d.transB != pl.transB)
      throw std::invalid_argument("grouped GEMM problems differ in transposes");
    if (d.M <= 0 || d.N <= 0 || d.K <= 0 || d.batch <= 0)
      throw std::invalid_argument("grouped GEMM problem with an empty extent");
    if (d.batch > 65535) throw std::invalid_argument("grouped GEMM batch too large");
    KParams& p = host[q].p;
    p.A = d.A; p.B = d.B; p.C = d.C;
    p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc; p.sA = d.sA; p.sB = d.sB; p.sC = d.sC;
    p.M = d.M; p.N = d.N; p.K = d.K;
    p.kchunk = sp[q].kchunk;
    p.nsplit = sp[q].nsplit;
    p.nsplit_batch = d.batch;
    p.batch0 = 0;
    p.tiles_m = (int)ceil_div(d.M, bm);
    p.tiles_n = (int)ceil_div(d.N, bm);
    p.per_xcd = 0;
    p.alpha = d.alpha;
    p.beta = d.beta;
    p.W = nullptr;
    host[q].first = pl.items;
    host[q].rfirst = pl.reduce_items;
    pl.items += (long)p.tiles_m * p.tiles_n * d.batch * p.nsplit;
    if (p.nsplit > 1) {
      const size_t need = (size_t)p.nsplit * d.batch * d.M * d.N;
      if (woff + need > ws_doubles) throw std::invalid_argument("grouped GEMM workspace too small");
      p.W = ws + woff;
      woff += need;
      pl.reduce_items += ceil_div((long)d.batch * d.M * d.N, 256);
    }
  }
  pl.ws_doubles = woff;
  pl.grid = (int)std::min<long>(pl.items, 4096);
  MPQC_HIP_CHECK(hipMalloc(&pl.dev, host.size() * sizeof(GProblem)));
  MPQC_HIP_CHECK(hipMemcpyAsync(pl.dev, host.data(), host.size() * sizeof(GProblem),
                                hipMemcpyHostToDevice, stream));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream));
  return pl;
}

void launch_grouped_gemm(const GroupedGemmPlan& pl, hipStream_t stream) {
  if (pl.items == 0) return;
  if (pl.bm == 128) launch_grouped_bm<128>(pl, stream);
  else launch_grouped_bm<64>(pl, stream);
  if (pl.reduce_items > 0) {
    const unsigned g = (unsigned)std::min<long>(pl.reduce_items, 65536);
    hipLaunchKernelGGL(grouped_reduce_kernel, dim3(g), dim3(256), 0, stream, (const GProblem*)pl.dev,
                       pl.nprob, pl.reduce_items);
    check_launch("grouped_reduce");
  }
}

void free_grouped_gemm(GroupedGemmPlan& pl) {
  if (pl.dev) (void)hipFree(pl.dev);
  pl.dev = nullptr;
}

void splitk_reduce(const double* W, double* C, long ldc, long sC, int M, int N, int nsplit, int batch0,
                   int nb, double beta, hipStream_t stream) {
  const long MN = (long)M * N;
  if (nb > 65535) throw std::invalid_argument("splitk_reduce: too many batches");
  if (MN * nb <= kWideReduceMax && nsplit >= kDeepReduceSplits) {
    const long waves = MN * nb;
    hipLaunchKernelGGL(splitk_reduce_deep_kernel, dim3((unsigned)ceil_div(waves, 4)), dim3(256), 0, stream, W, C,
                       ldc, sC, M, N, nsplit, batch0, nb, beta);
    check_launch("splitk_reduce_deep");
  } else if (MN * nb <= kWideReduceMax) {
    const long waves = ceil_div(MN, 16) * nb;
    hipLaunchKernelGGL(splitk_reduce_wide_kernel, dim3((unsigned)ceil_div(waves, 4)), dim3(256), 0,
                       stream, W, C, ldc, sC, M, N, nsplit, batch0, nb, beta);
    check_launch("splitk_reduce_wide");
  } else {
    dim3 g2((unsigned)std::min<long>(ceil_div(MN, 256), 4096), (unsigned)nb);
    hipLaunchKernelGGL(splitk_reduce_kernel, g2, dim3(256), 0, stream, W, C, ldc, sC, M, N, nsplit,
                       batch0, beta);
    check_launch("splitk_reduce");
  }
}

// ---------------------------------------------------------------- thin K
// C = alpha A B + beta C for a short contraction (K <= 64) of a small A (M <= 256
// rows) with a long B (K x N), no transposes: t1(a,k) X(k,(b,i,j)) of the CCSD
// residual (ccsd_r1_r2.h:207-211, :238-248, :278-280).  The output streams
// through HBM once (read and written, or only written when beta = 0); the
// tiled kernels spend most of such a launch staging K through LDS.
//
// A lives in registers for the whole launch: one wave per 16-row block of A
// (at least four waves), each holding the 16 x 4 fragments of its block for all K.  The workgroup
// walks steps of 32 columns (two 16-column blocks) with stride gridDim.x: the
// K x 32 tile of B is loaded by the whole workgroup in 16-byte pieces into
// registers during the previous step and stored to LDS (double buffered, one
// barrier per step); each wave reads its B fragments from LDS, reads its rows
// of C before the products, runs 2 x K/4 v_mfma_f64_16x16x4f64 and writes its
// rows of C.  (B fragments loaded by every wave straight from HBM, 8 bytes per
// lane, ran at half this speed: the loads, not the stores or the products,
// bounded that form.)
namespace {

struct ThinKParams {
  const double* A;
  const double* B;
  double* C;
  long lda, ldb, ldc, sA, sB, sC;
  int M, N, K;
  int nsteps;  // 32-column steps
  double alpha, beta;
};

constexpr int kThinMaxWaves = 16;
constexpr int kThinMinWaves = 4;  // waves past the last row block only load B
constexpr int kThinCols = 32;  // columns per step

// KS: steps of 4 along K (K <= 4 KS); VEC: B is 16-byte aligned with an even ldb
template <int KS, bool VEC>
__global__ __launch_bounds__(64 * kThinMaxWaves) void dgemm_thin_k_kernel(ThinKParams p) {
  constexpr int ROWS = 4 * KS, LDS_LD = kThinCols + 2;
  constexpr int CH = ROWS * kThinCols / 2;  // 16-byte pieces of a B tile
  __shared__ double Bs[2][ROWS][LDS_LD];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const long z = blockIdx.y;
  const double* __restrict__ A = p.A + z * p.sA;
  const double* __restrict__ B = p.B + z * p.sB;
  double* __restrict__ C = p.C + z * p.sC;
  // fragments of A: lane holds A(16 wave + lr, 4 s + lk)
  double af[KS];
  {
    const int m = 16 * wave + lr;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int k = 4 * s + lk;
      af[s] = (m < p.M && k < p.K) ? A[(long)m * p.lda + k] : 0.0;
    }
  }
  // the B tile of step st: piece q covers row q / 16, columns 2 (q % 16) + {0, 1}
  constexpr int PER = (CH + 64 * kThinMinWaves - 1) / (64 * kThinMinWaves);  // pieces per thread, at most
  const int per = (CH + nt - 1) / nt;
  double2 rg[PER];
  auto gload = [&](int st) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      if (j >= per) break;
      const int q = tid + j * nt;
      const int row = q / (kThinCols / 2), n = st * kThinCols + 2 * (q % (kThinCols / 2));
      double2 x = {0.0, 0.0};
      if (q < CH && st < p.nsteps && row < p.K) {
        const double* b = B + (long)row * p.ldb + n;
        if (VEC && n + 1 < p.N) {
          x = *reinterpret_cast<const double2*>(b);
        } else {
          if (n < p.N) x.x = b[0];
          if (n + 1 < p.N) x.y = b[1];
        }
      }
      rg[j] = x;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      if (j >= per) break;
      const int q = tid + j * nt;
      if (q < CH) {
        const int row = q / (kThinCols / 2), c = 2 * (q % (kThinCols / 2));
        Bs[buf][row][c] = rg[j].x;
        Bs[buf][row][c + 1] = rg[j].y;
      }
    }
  };
  const bool read_c = p.beta != 0.0;
  const int G = gridDim.x;
  int st = blockIdx.x, buf = 0;
  gload(st);
  sstore(0);
  __syncthreads();
  for (; st < p.nsteps; st += G) {  // uniform over the workgroup
    gload(st + G);
    double cin[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int n = st * kThinCols + 16 * c + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 16 * wave + 4 * r + lk;
        cin[c][r] = (read_c && n < p.N && m < p.M) ? C[(long)m * p.ldc + n] : 0.0;
      }
    }
    double4_t acc[2] = {double4_t{0.0, 0.0, 0.0, 0.0}, double4_t{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int c = 0; c < 2; ++c)
        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], Bs[buf][4 * s + lk][16 * c + lr], acc[c], 0, 0, 0);
    // accumulator: lane holds C(16 wave + 4 r + lk, n)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int n = st * kThinCols + 16 * c + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 16 * wave + 4 * r + lk;
        if (n < p.N && m < p.M)
          C[(long)m * p.ldc + n] = read_c ? fma(p.beta, cin[c][r], p.alpha * acc[c][r]) : p.alpha * acc[c][r];
      }
    }
    sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
}

bool thin_k_shape(const GemmDesc& d) {
  return !d.transA && !d.transB && d.K <= 64 && d.M <= 16 * kThinMaxWaves && d.N >= 2048 && d.batch <= 65535;
}

// Resident workgroups of `waves` waves on the device (a grid-size heuristic
// only: the kernel is correct for any grid).
template <int KS, bool VEC>
int thin_k_resident(int waves) {
  static int cached[kThinMaxWaves + 1] = {0};
  if (cached[waves] == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    MPQC_HIP_CHECK(hipGetDevice(&dev));
    MPQC_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    MPQC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dgemm_thin_k_kernel<KS, VEC>, 64 * waves, 0));
    cached[waves] = std::max(1, cus) * std::max(1, per_cu);
  }
  return cached[waves];
}

template <int KS, bool VEC>
void launch_thin_k_ks(const ThinKParams& p, int waves, int batch, hipStream_t stream) {
  const int grid = (int)std::min<long>(p.nsteps, std::max<long>(1, thin_k_resident<KS, VEC>(waves) / std::max(1, batch)));
  hipLaunchKernelGGL((dgemm_thin_k_kernel<KS, VEC>), dim3((unsigned)grid, (unsigned)batch), dim3(64 * waves), 0,
                     stream, p);
  check_launch("dgemm_thin_k");
}

template <bool VEC>
void launch_thin_k_vec(const ThinKParams& p, int waves, int batch, hipStream_t stream) {
  switch ((p.K + 3) / 4) {
    case 1: case 2: launch_thin_k_ks<2, VEC>(p, waves, batch, stream); break;
    case 3: case 4: launch_thin_k_ks<4, VEC>(p, waves, batch, stream); break;
    case 5: case 6: launch_thin_k_ks<6, VEC>(p, waves, batch, stream); break;
    case 7: case 8: launch_thin_k_ks<8, VEC>(p, waves, batch, stream); break;
    case 9: case 10: launch_thin_k_ks<10, VEC>(p, waves, batch, stream); break;
    case 11: case 12: launch_thin_k_ks<12, VEC>(p, waves, batch, stream); break;
    default: launch_thin_k_ks<16, VEC>(p, waves, batch, stream); break;
  }
}

void launch_thin_k(const GemmDesc& d, hipStream_t stream) {
  ThinKParams p;
  p.A = d.A; p.B = d.B; p.C = d.C;
  p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc; p.sA = d.sA; p.sB = d.sB; p.sC = d.sC;
  p.M = d.M; p.N = d.N; p.K = d.K; p.alpha = d.alpha; p.beta = d.beta;
  p.nsteps = (int)ceil_div(d.N, kThinCols);
  const int waves = std::max(kThinMinWaves, (int)ceil_div(d.M, 16));
  // 16-byte loads of B need every row start 16-byte aligned
  const bool vec = ((uintptr_t)d.B % 16 == 0) && d.ldb % 2 == 0 && (d.batch == 1 || d.sB % 2 == 0);
  if (vec)
    launch_thin_k_vec<true>(p, waves, d.batch, stream);
  else
    launch_thin_k_vec<false>(p, waves, d.batch, stream);
}

// ---------------------------------------------------------------- thin N
// C = alpha A B + beta C for a tall A (M x K, K contiguous, K <= 256) against a
// small B (K x N, N <= 64), no transposes: the T1 products Giabc(k,a,c,d) t1(d,i),
// Giajb(k,b,i,c) t1(c,j), Gijab t1, Gijka t1 (ccsd_r1_r2.h:207-211, :240, :246,
// ccsd_intermediates.h:182-193).  A streams through HBM once; the tiled kernels
// pad N to 64 and stage A through LDS.
//
// B sits in LDS for the whole launch (zero-padded to 16 K x 16 N blocks).  Each
// wave walks 16-row blocks of A: per 16-wide K chunk a lane loads 32 contiguous
// bytes, A(m0 + l%16, 16t + 4g .. 16t + 4g + 3) with g = l/16 (a 16-row block reads
// whole 128-byte row pieces), and issues four v_mfma_f64_16x16x4f64 per 16 columns of
// N, the q-th taking k = 16t + 4g + q from lane group g (any assignment of the 16 k
// of a chunk to the four MFMA k slots works when B uses the same one).
struct ThinNParams {
  const double* A;
  const double* B;
  double* C;
  long lda, ldb, ldc;
  int M, N, K;
  double alpha, beta;
};

constexpr int kThinNThreads = 512;

template <int NB>
constexpr int thin_n_pitch() { return 16 * NB + 4; }  // 4 rows apart land 128 bytes apart in the banks

// NB: 16-column blocks (N <= 16 NB); KC: 16-wide K chunks (K <= 16 KC); VEC: 16-byte
// loads of A (A 16-byte aligned, lda even)
template <int NB, int KC, bool VEC>
__global__ __launch_bounds__(kThinNThreads) void dgemm_thin_n_kernel(ThinNParams p) {
  constexpr int PITCH = thin_n_pitch<NB>();
  __shared__ double Bs[16 * KC * PITCH];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, g = lane >> 4;
  for (int e = tid; e < 16 * KC * 16 * NB; e += kThinNThreads) {
    const int k = e / (16 * NB), n = e - k * (16 * NB);
    Bs[k * PITCH + n] = (k < p.K && n < p.N) ? p.B[(long)k * p.ldb + n] : 0.0;
  }
  __syncthreads();
  const long nblk = ((long)p.M + 15) / 16;
  const long waves = (long)gridDim.x * (kThinNThreads / 64);
  const bool read_c = p.beta != 0.0;
  for (long blk = (long)blockIdx.x * (kThinNThreads / 64) + wave; blk < nblk; blk += waves) {
    const long m = 16 * blk + lr;
    const bool mrow = m < p.M;
    // the B fragments are the same for every block: an opaque column offset keeps the
    // compiler from hoisting all of them out of the loop (they would not fit in VGPRs)
    int col = lr;
    asm volatile("" : "+v"(col));
    const double* __restrict__ a = p.A + (mrow ? m : (long)p.M - 1) * p.lda;
    double av[KC][4];
#pragma unroll
    for (int t = 0; t < KC; ++t) {
      const int k0 = 16 * t + 4 * g;
      if (VEC && k0 + 3 < p.K) {
        const double2 x = *reinterpret_cast<const double2*>(a + k0);
        const double2 y = *reinterpret_cast<const double2*>(a + k0 + 2);
        av[t][0] = x.x; av[t][1] = x.y; av[t][2] = y.x; av[t][3] = y.y;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) av[t][q] = k0 + q < p.K ? a[k0 + q] : 0.0;
      }
    }
    double4_t acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < KC; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double* brow = Bs + (16 * t + 4 * g + q) * PITCH + col;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t][q], brow[16 * nb], acc[nb], 0, 0, 0);
      }
    // accumulator: lane holds C(16 blk + 4 r + g, 16 nb + lr)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int n = 16 * nb + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long mm = 16 * blk + 4 * r + g;
        if (mm < p.M && n < p.N) {
          double* c = p.C + mm * p.ldc + n;
          *c = read_c ? fma(p.beta, *c, p.alpha * acc[nb][r]) : p.alpha * acc[nb][r];
        }
      }
    }
  }
}

bool thin_n_shape(const GemmDesc& d) {
  const long npad = ceil_div(d.N, 16) * 16, kpad = ceil_div(d.K, 16) * 16;
  return !d.transA && !d.transB && d.batch == 1 && d.N <= 64 && d.K <= 256 && d.M >= 4096 &&
         kpad * (npad + 4) <= 10240;  // B in at most 80 KB of LDS
}

template <int NB, int KC, bool VEC>
int thin_n_resident() {
  static int cached = 0;
  if (cached == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    MPQC_HIP_CHECK(hipGetDevice(&dev));
    MPQC_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    MPQC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dgemm_thin_n_kernel<NB, KC, VEC>,
                                                                 kThinNThreads, 0));
    cached = std::max(1, cus) * std::max(1, per_cu);
  }
  return cached;
}

template <int NB, int KC, bool VEC>
void launch_thin_n_t(const ThinNParams& p, hipStream_t stream) {
  const long blocks = ceil_div(p.M, 16), waves_per_wg = kThinNThreads / 64;
  const int grid = (int)std::min<long>(ceil_div(blocks, waves_per_wg), thin_n_resident<NB, KC, VEC>());
  hipLaunchKernelGGL((dgemm_thin_n_kernel<NB, KC, VEC>), dim3((unsigned)grid), dim3(kThinNThreads), 0, stream, p);
  check_launch("dgemm_thin_n");
}

template <int NB, bool VEC>
void launch_thin_n_nb(const ThinNParams& p, hipStream_t stream) {
  switch ((p.K + 15) / 16) {
    case 1: case 2: launch_thin_n_t<NB, 2, VEC>(p, stream); break;
    case 3: case 4: launch_thin_n_t<NB, 4, VEC>(p, stream); break;
    case 5: case 6: launch_thin_n_t<NB, 6, VEC>(p, stream); break;
    case 7: case 8: launch_thin_n_t<NB, 8, VEC>(p, stream); break;
    case 9: case 10: launch_thin_n_t<NB, 10, VEC>(p, stream); break;
    case 11: case 12: launch_thin_n_t<NB, 12, VEC>(p, stream); break;
    default: launch_thin_n_t<NB, 16, VEC>(p, stream); break;
  }
}

template <bool VEC>
void launch_thin_n_vec(const ThinNParams& p, hipStream_t stream) {
  switch ((p.N + 15) / 16) {
    case 1: launch_thin_n_nb<1, VEC>(p, stream); break;
    case 2: launch_thin_n_nb<2, VEC>(p, stream); break;
    case 3: launch_thin_n_nb<3, VEC>(p, stream); break;
    default: launch_thin_n_nb<4, VEC>(p, stream); break;
  }
}

void launch_thin_n(const GemmDesc& d, hipStream_t stream) {
  ThinNParams p;
  p.A = d.A; p.B = d.B; p.C = d.C;
  p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc;
  p.M = d.M; p.N = d.N; p.K = d.K; p.alpha = d.alpha; p.beta = d.beta;
  const bool vec = ((uintptr_t)d.A % 16 == 0) && d.lda % 2 == 0;
  if (vec)
    launch_thin_n_vec<true>(p, stream);
  else
    launch_thin_n_vec<false>(p, stream);
}

// ---------------------------------------------------------------- many small products, one shared A
// C_b = alpha op(A) B_b + beta C_b for b = 0 .. batch-1 with M, N, K <= 64, one A for
// every b (sA = 0), no transposed B: the T2 g_ki term of Rd (ccsd_r1_r2.h:226), a
// product per (a,b) pair of o x o blocks.  One wave per product: the wave holds A's
// v_mfma_f64_16x16x4f64 fragments for the whole launch, loads the product's B
// fragments and C rows (rows of 16 doubles), and walks the products with stride
// (waves in the grid).  The tiled kernels give each such product a 64 x 64 workgroup
// that stages 16-wide K slices through LDS.
struct SmallParams {
  const double* A;
  const double* B;
  double* C;
  long lda, ldb, ldc, sB, sC;
  int M, N, K, batch;
  bool transA;
  double alpha, beta;
};

constexpr int kSmallThreads = 256;

// S: 16-row and 16-column blocks (M, N <= 16 S); KS: steps of 4 along K (K <= 4 KS)
template <int S, int KS>
__global__ __launch_bounds__(kSmallThreads) void dgemm_small_batched_kernel(SmallParams p) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  // A fragments: lane holds op(A)(16 ib + lr, 4 s + lk)
  double af[S][KS];
#pragma unroll
  for (int ib = 0; ib < S; ++ib)
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int i = 16 * ib + lr, k = 4 * s + lk;
      af[ib][s] = (i < p.M && k < p.K) ? (p.transA ? p.A[(long)k * p.lda + i] : p.A[(long)i * p.lda + k]) : 0.0;
    }
  const bool read_c = p.beta != 0.0;
  const long nw = (long)gridDim.x * (kSmallThreads / 64);
  for (long b = (long)blockIdx.x * (kSmallThreads / 64) + wave; b < p.batch; b += nw) {
    const double* __restrict__ B = p.B + b * p.sB;
    double* __restrict__ C = p.C + b * p.sC;
    // one 16-column block of the product at a time (registers: A's fragments stay, the
    // block's B fragments, C rows and accumulators come and go)
#pragma unroll 1
    for (int jb = 0; jb < S; ++jb) {
      const int j = 16 * jb + lr;
      double bf[KS];
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + lk;
        bf[s] = (j < p.N && k < p.K) ? B[(long)k * p.ldb + j] : 0.0;
      }
      double cin[S][4];
#pragma unroll
      for (int ib = 0; ib < S; ++ib)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * ib + 4 * r + lk;
          cin[ib][r] = (read_c && i < p.M && j < p.N) ? C[(long)i * p.ldc + j] : 0.0;
        }
      double4_t acc[S];
#pragma unroll
      for (int ib = 0; ib < S; ++ib) acc[ib] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int ib = 0; ib < S; ++ib)
          acc[ib] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[ib][s], bf[s], acc[ib], 0, 0, 0);
      // accumulator: lane holds C(16 ib + 4 r + lk, 16 jb + lr)
#pragma unroll
      for (int ib = 0; ib < S; ++ib)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * ib + 4 * r + lk;
          if (i < p.M && j < p.N)
            C[(long)i * p.ldc + j] = read_c ? fma(p.beta, cin[ib][r], p.alpha * acc[ib][r]) : p.alpha * acc[ib][r];
        }
    }
  }
}

bool small_batched_shape(const GemmDesc& d) {
  return d.sA == 0 && !d.transB && d.M <= 64 && d.N <= 64 && d.K <= 64 && d.batch >= 256;
}

template <int S, int KS>
void launch_small_batched_t(const SmallParams& p, hipStream_t stream) {
  static int resident = 0;
  if (resident == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    MPQC_HIP_CHECK(hipGetDevice(&dev));
    MPQC_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    MPQC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dgemm_small_batched_kernel<S, KS>,
                                                                 kSmallThreads, 0));
    resident = std::max(1, cus) * std::max(1, per_cu);
  }
  const int grid = (int)std::min<long>(ceil_div(p.batch, kSmallThreads / 64), resident);
  hipLaunchKernelGGL((dgemm_small_batched_kernel<S, KS>), dim3((unsigned)grid), dim3(kSmallThreads), 0, stream, p);
  check_launch("dgemm_small_batched");
}

template <int S>
void launch_small_batched_s(const SmallParams& p, hipStream_t stream) {
  switch ((p.K + 3) / 4) {
    case 1: case 2: launch_small_batched_t<S, 2>(p, stream); break;
    case 3: case 4: launch_small_batched_t<S, 4>(p, stream); break;
    case 5: case 6: launch_small_batched_t<S, 6>(p, stream); break;
    case 7: case 8: launch_small_batched_t<S, 8>(p, stream); break;
    case 9: case 10: launch_small_batched_t<S, 10>(p, stream); break;
    case 11: case 12: launch_small_batched_t<S, 12>(p, stream); break;
    default: launch_small_batched_t<S, 16>(p, stream); break;
  }
}

void launch_small_batched(const GemmDesc& d, hipStream_t stream) {
  SmallParams p;
  p.A = d.A; p.B = d.B; p.C = d.C;
  p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc; p.sB = d.sB; p.sC = d.sC;
  p.M = d.M; p.N = d.N; p.K = d.K; p.batch = d.batch; p.transA = d.transA;
  p.alpha = d.alpha; p.beta = d.beta;
  switch ((std::max(d.M, d.N) + 15) / 16) {
    case 1: launch_small_batched_s<1>(p, stream); break;
    case 2: launch_small_batched_s<2>(p, stream); break;
    case 3: launch_small_batched_s<3>(p, stream); break;
    default: launch_small_batched_s<4>(p, stream); break;
  }
}

// ---------------------------------------------------------------- GEMV
// y = alpha A x + beta y (N = 1, A row-major M x K with long rows): the r1 terms
// Gbar(k,c,l,d) t1(d,l), (2 Gijab - Giajb) t1 and (2 t2 - t2) h (ccsd_r1_r2.h:100-108)
// and the g_ac / g_ki terms (ccsd_intermediates.h:105-148).  One wave per row (rows
// with stride: waves in the grid); lanes read 16 bytes of A and of x per step, eight
// steps in flight, then a shuffle reduction.  HBM-bound: M K 8 bytes per launch.
struct GemvParams {
  const double* A;
  const double* x;
  double* y;
  long lda, incx, incy;
  int M, K;
  double alpha, beta;
};

constexpr int kGemvThreads = 256;

template <bool VEC>
__global__ __launch_bounds__(kGemvThreads) void dgemv_kernel(GemvParams p) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long nw = (long)gridDim.x * (kGemvThreads / 64);
  for (long m = (long)blockIdx.x * (kGemvThreads / 64) + wave; m < p.M; m += nw) {
    const double* __restrict__ a = p.A + m * p.lda;
    double s0 = 0.0, s1 = 0.0;
    if (VEC) {  // 16-byte pieces of the row (A 16-byte aligned, lda even, x contiguous and aligned)
      const int k2 = p.K / 2;  // whole pieces
#pragma unroll 8
      for (int q = lane; q < k2; q += 64) {
        const double2 av = reinterpret_cast<const double2*>(a)[q];
        const double2 xv = reinterpret_cast<const double2*>(p.x)[q];
        s0 = fma(av.x, xv.x, s0);
        s1 = fma(av.y, xv.y, s1);
      }
      if ((p.K & 1) && lane == 0) s0 = fma(a[p.K - 1], p.x[p.K - 1], s0);
    } else {
#pragma unroll 8
      for (int k = lane; k < p.K; k += 64) s0 = fma(a[k], p.x[(long)k * p.incx], s0);
    }
    double sum = s0 + s1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if (lane == 0) {
      double* y = p.y + m * p.incy;
      *y = p.beta == 0.0 ? p.alpha * sum : fma(p.beta, *y, p.alpha * sum);
    }
  }
}

bool gemv_shape(const GemmDesc& d) {
  return d.N == 1 && !d.transA && d.batch == 1 && d.K >= 256 && d.M >= 64;
}

void launch_gemv(const GemmDesc& d, hipStream_t stream) {
  GemvParams p;
  p.A = d.A; p.x = d.B; p.y = d.C;
  p.lda = d.lda; p.incx = d.transB ? 1 : d.ldb; p.incy = d.ldc;
  p.M = d.M; p.K = d.K; p.alpha = d.alpha; p.beta = d.beta;
  const bool vec = ((uintptr_t)d.A % 16 == 0) && d.lda % 2 == 0 && p.incx == 1 && ((uintptr_t)d.B % 16 == 0);
  const int grid = (int)std::min<long>(ceil_div(d.M, kGemvThreads / 64), 4096);
  if (vec)
    hipLaunchKernelGGL(dgemv_kernel<true>, dim3((unsigned)grid), dim3(kGemvThreads), 0, stream, p);
  else
    hipLaunchKernelGGL(dgemv_kernel<false>, dim3((unsigned)grid), dim3(kGemvThreads), 0, stream, p);
  check_launch("dgemv");
}

int g_thin_k_never = 0;  // 1: the tiled kernels take the shapes thin_k_shape() and thin_n_shape() admit

}  // namespace

void set_thin_k_mode(int mode) { g_thin_k_never = mode; }

void gemm(const GemmDesc& d, GemmPrecision prec, const GemmWorkspace& ws, hipStream_t stream) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  if (d.K <= 0) {
    throw std::invalid_argument("gemm: K must be positive");
  }
  if (prec == GemmPrecision::BF16X3 || prec == GemmPrecision::F16X3 || prec == GemmPrecision::BF16X6) {
    gemm_split(d, prec, ws, stream);
    return;
  }
  if (!g_thin_k_never && prec == GemmPrecision::F64 && thin_k_shape(d)) {
    launch_thin_k(d, stream);
    return;
  }
  if (!g_thin_k_never && prec == GemmPrecision::F64 && thin_n_shape(d)) {
    launch_thin_n(d, stream);
    return;
  }
  if (!g_thin_k_never && prec == GemmPrecision::F64 && small_batched_shape(d)) {
    launch_small_batched(d, stream);
    return;
  }
  if (!g_thin_k_never && prec == GemmPrecision::F64 && gemv_shape(d)) {
    launch_gemv(d, stream);
    return;
  }
  if (wide_tiles(d, ws.bytes))
    launch<128, 128>(d, prec, ws, stream);
  else
    launch<64, 64>(d, prec, ws, stream);
}

size_t gemm_scratch_bytes(const GemmDesc& d, GemmPrecision prec, size_t ws_bytes) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0 || d.K <= 0) return 0;
  if (prec == GemmPrecision::BF16X3 || prec == GemmPrecision::F16X3 || prec == GemmPrecision::BF16X6)
    return gemm_split_scratch_bytes(d, prec, ws_bytes);
  if (!g_thin_k_never && prec == GemmPrecision::F64 &&
      (thin_k_shape(d) || thin_n_shape(d) || small_batched_shape(d) || gemv_shape(d)))
    return 0;
  const bool wide = wide_tiles(d, ws_bytes);
  return plan_splitk(d, wide ? 128 : 64, wide ? 128 : 64, ws_bytes).scratch_bytes;
}

}  // namespace mpqc_hip
