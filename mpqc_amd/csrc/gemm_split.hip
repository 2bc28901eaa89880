// Split-precision GEMM on gfx950 16x16x32 half-precision matrix cores.
//
// Each FP64 operand element x is split into half-precision parts whose sum
// reproduces x to 16-24 bits, and the product is formed from the cross terms
// that matter:
//   BF16 2-way (hi + lo, 3 products hh + hl + lh): ~16 bits
//   BF16 3-way (hi + mid + lo, 6 products):         ~24 bits
//   FP16 2-way, power-of-two scaled (3 products):   ~22 bits
// Products run on v_mfma_f32_16x16x32_{bf16,f16}; FP32 partial sums are
// folded into FP64 accumulators every kFlush K-slices.  Operands stay FP64 in
// HBM and are split while being staged into LDS.
//
// Fragment layout of the 16x16x32 instructions (lane l, j = 0..7, r = 0..3):
//   A[m = l%16][k = 8*(l/16) + j], B[k = 8*(l/16) + j][n = l%16],
//   C[m = 4*(l/16) + r][n = l%16]
// LDS holds A as [m][k] and B as [n][k] (k contiguous) so each fragment is one
// 16-byte read.
#include "gemm.hpp"

#include <algorithm>
#include <cmath>

namespace mpqc_hip {

namespace {

constexpr int kThreads = 256;
constexpr int kBKs = 32;        // K per slice = one MFMA K
constexpr int kPadH = 8;        // halves of padding per LDS row (keeps 16-byte alignment)
constexpr int kRowH = kBKs + kPadH;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4_t __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

enum SplitKind { kBF16x2 = 0, kBF16x3 = 1, kF16x2 = 2 };

template <int S>
struct SplitTraits;
template <>
struct SplitTraits<kBF16x2> { static constexpr int parts = 2; static constexpr bool f16 = false; };
template <>
struct SplitTraits<kBF16x3> { static constexpr int parts = 3; static constexpr bool f16 = false; };
template <>
struct SplitTraits<kF16x2> { static constexpr int parts = 2; static constexpr bool f16 = true; };

__device__ inline unsigned short bits_bf16(float f) {
  __bf16 h = (__bf16)f;
  return __builtin_bit_cast(unsigned short, h);
}
__device__ inline float from_bf16(unsigned short u) {
  return (float)__builtin_bit_cast(__bf16, u);
}
__device__ inline unsigned short bits_f16(float f) {
  _Float16 h = (_Float16)f;
  return __builtin_bit_cast(unsigned short, h);
}
__device__ inline float from_f16(unsigned short u) {
  return (float)__builtin_bit_cast(_Float16, u);
}

// split one scaled value into `parts` half-precision words
template <int S>
__device__ inline void split_value(double x, unsigned short* out) {
  if (SplitTraits<S>::f16) {
    const float xf = (float)x;
    const unsigned short h = bits_f16(xf);
    out[0] = h;
    out[1] = bits_f16((float)(x - (double)from_f16(h)));
  } else {
    const float xf = (float)x;
    const unsigned short h = bits_bf16(xf);
    const double r1 = x - (double)from_bf16(h);
    out[0] = h;
    const unsigned short m = bits_bf16((float)r1);
    out[1] = m;
    if (SplitTraits<S>::parts == 3) out[2] = bits_bf16((float)(r1 - (double)from_bf16(m)));
  }
}

struct SParams {
  const double* A;
  const double* B;
  double* C;
  double* W;
  const double* amax;  // device max|A| (F16 only)
  const double* bmax;  // device max|B| (F16 only)
  long lda, ldb, ldc, sA, sB, sC;
  int M, N, K;
  int kchunk;
  int nsplit;
  int batch0;
  int tiles_n;
  double alpha, beta;
};

// power of two s with max|x| * s in [2^13, 2^14)
__device__ inline double f16_scale(const double* amax) {
  const double m = *amax;
  if (!(m > 0.0) || !isfinite(m)) return 1.0;
  int e;
  frexp(m, &e);  // m = f * 2^e, f in [0.5, 1)
  return ldexp(1.0, 14 - e);
}

template <int S, int NT>
__device__ inline void mfma_products(const u16x8_t (&a)[NT], const u16x8_t (&b)[NT], float4_t& c) {
  auto mf = [&](const u16x8_t& x, const u16x8_t& y) {
    if (SplitTraits<S>::f16)
      c = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, x),
                                                 __builtin_bit_cast(f16x8_t, y), c, 0, 0, 0);
    else
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, x),
                                                  __builtin_bit_cast(bf16x8_t, y), c, 0, 0, 0);
  };
  // small terms first
  if (NT == 3) {
    mf(a[1], b[1]);  // mid*mid
    mf(a[0], b[2]);  // hi*lo
    mf(a[2], b[0]);  // lo*hi
  }
  mf(a[0], b[1]);
  mf(a[1], b[0]);
  mf(a[0], b[0]);
}

template <int BM, int BN, bool TA, bool TB, int S>
__global__ __launch_bounds__(kThreads, (BM >= 128 ? 1 : 2)) void gemm_split_kernel(SParams p) {
  constexpr int NP = SplitTraits<S>::parts;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int RB = WM / 16, CB = WN / 16;
  constexpr int AG = BM * (kBKs / 4) / kThreads;  // groups of 4 k-values per thread (A)
  constexpr int BG = BN * (kBKs / 4) / kThreads;
  constexpr int kFlush = 4;  // K-slices (of 32) per FP32 partial sum

  __shared__ __attribute__((aligned(16))) unsigned short As[2][NP][BM][kRowH];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[2][NP][BN][kRowH];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  const int tile = blockIdx.x;
  const int tm = tile / p.tiles_n, tn = tile - tm * p.tiles_n;
  const int z = blockIdx.y;
  const int b = p.batch0 + z / p.nsplit;
  const int s = z - (z / p.nsplit) * p.nsplit;
  const int m0 = tm * BM, n0 = tn * BN;
  const int k_begin = s * p.kchunk;
  const int k_end = min(p.K, k_begin + p.kchunk);

  const double* __restrict__ A = p.A + (long)b * p.sA;
  const double* __restrict__ B = p.B + (long)b * p.sB;
  const double scA = SplitTraits<S>::f16 ? f16_scale(p.amax) : 1.0;
  const double scB = SplitTraits<S>::f16 ? f16_scale(p.bmax) : 1.0;

  double ra[AG][4], rb[BG][4];
  // group g covers rows (m or n) g % ROWS and k = 4*(g / ROWS) .. +3
  auto load_regs = [&](int k0) {
#pragma unroll
    for (int r = 0; r < AG; ++r) {
      const int g = tid + kThreads * r;
      // coalesced: contiguous k within a row for A[m][k], contiguous m for A[k][m]
      const int mm = TA ? g % BM : g / (kBKs / 4);
      const int kq = TA ? (g / BM) * 4 : (g % (kBKs / 4)) * 4;
      const int gm = m0 + mm;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gk = k0 + kq + q;
        double x = 0.0;
        if (gm < p.M && gk < k_end) x = TA ? A[(long)gk * p.lda + gm] : A[(long)gm * p.lda + gk];
        ra[r][q] = x * scA;
      }
    }
#pragma unroll
    for (int r = 0; r < BG; ++r) {
      const int g = tid + kThreads * r;
      const int nn = TB ? g / (kBKs / 4) : g % BN;
      const int kq = TB ? (g % (kBKs / 4)) * 4 : (g / BN) * 4;
      const int gn = n0 + nn;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gk = k0 + kq + q;
        double x = 0.0;
        if (gn < p.N && gk < k_end) x = TB ? B[(long)gn * p.ldb + gk] : B[(long)gk * p.ldb + gn];
        rb[r][q] = x * scB;
      }
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int r = 0; r < AG; ++r) {
      const int g = tid + kThreads * r;
      const int mm = TA ? g % BM : g / (kBKs / 4);
      const int kq = TA ? (g / BM) * 4 : (g % (kBKs / 4)) * 4;
      unsigned short w[4][3];
#pragma unroll
      for (int q = 0; q < 4; ++q) split_value<S>(ra[r][q], w[q]);
#pragma unroll
      for (int part = 0; part < NP; ++part) {
        u16x4_t v = {w[0][part], w[1][part], w[2][part], w[3][part]};
        *reinterpret_cast<u16x4_t*>(&As[buf][part][mm][kq]) = v;
      }
    }
#pragma unroll
    for (int r = 0; r < BG; ++r) {
      const int g = tid + kThreads * r;
      const int nn = TB ? g / (kBKs / 4) : g % BN;
      const int kq = TB ? (g % (kBKs / 4)) * 4 : (g / BN) * 4;
      unsigned short w[4][3];
#pragma unroll
      for (int q = 0; q < 4; ++q) split_value<S>(rb[r][q], w[q]);
#pragma unroll
      for (int part = 0; part < NP; ++part) {
        u16x4_t v = {w[0][part], w[1][part], w[2][part], w[3][part]};
        *reinterpret_cast<u16x4_t*>(&Bs[buf][part][nn][kq]) = v;
      }
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
  for (int kt = k_begin; kt < k_end; kt += kBKs) {
    const bool has_next = kt + kBKs < k_end;
    if (has_next) load_regs(kt + kBKs);
    u16x8_t fb[CB][NP];
#pragma unroll
    for (int j = 0; j < CB; ++j)
#pragma unroll
      for (int part = 0; part < NP; ++part)
        fb[j][part] = *reinterpret_cast<const u16x8_t*>(&Bs[buf][part][wn * WN + j * 16 + lr][lk * 8]);
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      u16x8_t fa[NP];
#pragma unroll
      for (int part = 0; part < NP; ++part)
        fa[part] = *reinterpret_cast<const u16x8_t*>(&As[buf][part][wm * WM + i * 16 + lr][lk * 8]);
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        if constexpr (NP == 3) {
          mfma_products<S, 3>(fa, fb[j], facc[i][j]);
        } else {
          const u16x8_t a2[2] = {fa[0], fa[1]};
          const u16x8_t b2[2] = {fb[j][0], fb[j][1]};
          // hi*lo + lo*hi + hi*hi
          if (SplitTraits<S>::f16) {
            facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, a2[0]), __builtin_bit_cast(f16x8_t, b2[1]), facc[i][j], 0, 0, 0);
            facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, a2[1]), __builtin_bit_cast(f16x8_t, b2[0]), facc[i][j], 0, 0, 0);
            facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, a2[0]), __builtin_bit_cast(f16x8_t, b2[0]), facc[i][j], 0, 0, 0);
          } else {
            facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a2[0]), __builtin_bit_cast(bf16x8_t, b2[1]), facc[i][j], 0, 0, 0);
            facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a2[1]), __builtin_bit_cast(bf16x8_t, b2[0]), facc[i][j], 0, 0, 0);
            facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a2[0]), __builtin_bit_cast(bf16x8_t, b2[0]), facc[i][j], 0, 0, 0);
          }
        }
      }
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

  const double unscale = p.alpha / (scA * scB);
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
          const double v = unscale * acc[i][j][r];
          if (p.nsplit == 1) {
            double* c = C + (long)gm * p.ldc + gn;
            *c = (p.beta == 0.0) ? v : v + p.beta * (*c);
          } else {
            Wz[(long)gm * p.N + gn] = v;
          }
        }
      }
}

// max |x| over an M x K (op) region of a batched operand: stage 1 per block
__global__ __launch_bounds__(256) void absmax_stage1(const double* __restrict__ X, bool trans, int R,
                                                     int Cc, long ld, long sbatch, int batch,
                                                     double* __restrict__ partials) {
  // rows R (m or n), cols Cc (k); element (r, c, b) at X[b*sbatch + (trans ? c*ld + r : r*ld + c)]
  const long total = (long)R * Cc * batch;
  double m = 0.0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long bb = e / ((long)R * Cc);
    const long rem = e - bb * (long)R * Cc;
    long r, c;
    if (trans) { c = rem / R; r = rem - c * R; } else { r = rem / Cc; c = rem - r * Cc; }
    const double x = fabs(X[bb * sbatch + (trans ? c * ld + r : r * ld + c)]);
    m = x > m ? x : m;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void absmax_stage2(const double* __restrict__ partials, int n,
                                                     double* __restrict__ out) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, partials[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

void absmax(const double* X, bool trans, int R, int Cc, long ld, long sbatch, int batch,
            double* partials, double* out, hipStream_t stream) {
  const long total = (long)R * Cc * batch;
  const int g = (int)std::max<long>(1, std::min<long>(1024, ceil_div(total, 256)));
  hipLaunchKernelGGL(absmax_stage1, dim3(g), dim3(256), 0, stream, X, trans, R, Cc, ld, sbatch, batch,
                     partials);
  hipLaunchKernelGGL(absmax_stage2, dim3(1), dim3(256), 0, stream, partials, g, out);
  check_launch("absmax");
}

__global__ void splitk_reduce2(const double* __restrict__ W, double* __restrict__ C, long ldc, long sC,
                               int M, int N, int nsplit, int batch0, double beta) {
  const long MN = (long)M * N;
  const int bz = blockIdx.y;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < MN; e += (long)gridDim.x * blockDim.x) {
    const int m = (int)(e / N), n = (int)(e - (long)(e / N) * N);
    double acc = 0.0;
    const double* w = W + (long)bz * nsplit * MN + e;
    for (int s2 = 0; s2 < nsplit; ++s2) acc += w[(long)s2 * MN];
    double* c = C + (long)(batch0 + bz) * sC + (long)m * ldc + n;
    *c = (beta == 0.0) ? acc : acc + beta * (*c);
  }
}

// workspace layout: [0, 2048) reduction partials, [2048, 2050) max|A|, max|B|,
// [4096, ...) split-K partial sums
constexpr long kHead = 4096;

struct SplitPlan {
  int nsplit = 1, kchunk = 0, max_batch_per_launch = 65535;
  size_t scratch_bytes = 0;  // workspace the launches touch (head included)
};

SplitPlan plan_split(const GemmDesc& d, int BM, int BN, size_t ws_bytes) {
  SplitPlan pl;
  const size_t wbytes = ws_bytes > (size_t)kHead * sizeof(double) ? ws_bytes - kHead * sizeof(double) : 0;
  const int tiles_m = (int)ceil_div(d.M, BM), tiles_n = (int)ceil_div(d.N, BN);
  const long tiles = (long)tiles_m * tiles_n;
  int nsplit = 1;
  const long want = 512;
  if (tiles * d.batch < want && d.K >= 8 * kBKs) {
    nsplit = (int)std::min<long>(ceil_div(want, tiles * d.batch), d.K / (4 * kBKs));
    nsplit = std::max(1, std::min(nsplit, 64));
  }
  int kchunk = (int)(ceil_div(ceil_div(d.K, nsplit), kBKs) * kBKs);
  if (nsplit > 1) nsplit = (int)ceil_div(d.K, kchunk);
  const size_t slice = (size_t)d.M * d.N * sizeof(double);
  int max_batch_per_launch = 65535 / nsplit;
  if (nsplit > 1) {
    const size_t cap = wbytes / (slice * nsplit);
    if (cap == 0) {
      nsplit = 1;
      kchunk = (int)(ceil_div(d.K, kBKs) * kBKs);
      max_batch_per_launch = 65535;
    } else {
      max_batch_per_launch = (int)std::min<size_t>(max_batch_per_launch, cap);
    }
  }
  pl.nsplit = nsplit;
  pl.kchunk = kchunk;
  pl.max_batch_per_launch = max_batch_per_launch;
  pl.scratch_bytes = kHead * sizeof(double) +
                     (nsplit > 1 ? slice * nsplit * (size_t)std::min(d.batch, max_batch_per_launch) : 0);
  return pl;
}

template <int BM, int BN, int S>
void launch_split(const GemmDesc& d, const GemmWorkspace& ws, hipStream_t stream) {
  if (ws.ptr == nullptr || ws.bytes < (size_t)kHead * sizeof(double))
    throw std::invalid_argument("split GEMM needs a workspace");
  double* partials = ws.ptr;
  double* amax = ws.ptr + 2048;
  double* bmax = ws.ptr + 2049;
  double* wsplit = ws.ptr + kHead;
  if (SplitTraits<S>::f16) {
    absmax(d.A, d.transA, d.M, d.K, d.lda, d.sA, d.batch, partials, amax, stream);
    absmax(d.B, !d.transB, d.N, d.K, d.ldb, d.sB, d.batch, partials, bmax, stream);
  }
  const int tiles_n = (int)ceil_div(d.N, BN);
  const long tiles = (long)ceil_div(d.M, BM) * tiles_n;
  const SplitPlan plan = plan_split(d, BM, BN, ws.bytes);
  const int nsplit = plan.nsplit, kchunk = plan.kchunk, max_batch_per_launch = plan.max_batch_per_launch;
  for (int b0 = 0; b0 < d.batch; b0 += max_batch_per_launch) {
    const int nb = std::min(max_batch_per_launch, d.batch - b0);
    SParams p;
    p.A = d.A; p.B = d.B; p.C = d.C; p.W = wsplit; p.amax = amax; p.bmax = bmax;
    p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc; p.sA = d.sA; p.sB = d.sB; p.sC = d.sC;
    p.M = d.M; p.N = d.N; p.K = d.K; p.kchunk = kchunk; p.nsplit = nsplit; p.batch0 = b0;
    p.tiles_n = tiles_n; p.alpha = d.alpha; p.beta = d.beta;
    dim3 grid((unsigned)tiles, (unsigned)(nb * nsplit));
    if (!d.transA && !d.transB) hipLaunchKernelGGL((gemm_split_kernel<BM, BN, false, false, S>), grid, dim3(kThreads), 0, stream, p);
    else if (!d.transA && d.transB) hipLaunchKernelGGL((gemm_split_kernel<BM, BN, false, true, S>), grid, dim3(kThreads), 0, stream, p);
    else if (d.transA && !d.transB) hipLaunchKernelGGL((gemm_split_kernel<BM, BN, true, false, S>), grid, dim3(kThreads), 0, stream, p);
    else hipLaunchKernelGGL((gemm_split_kernel<BM, BN, true, true, S>), grid, dim3(kThreads), 0, stream, p);
    check_launch("gemm_split");
    if (nsplit > 1) {
      const long MN = (long)d.M * d.N;
      dim3 g2((unsigned)std::min<long>(ceil_div(MN, 256), 4096), (unsigned)nb);
      hipLaunchKernelGGL(splitk_reduce2, g2, dim3(256), 0, stream, wsplit, d.C, d.ldc, d.sC, d.M, d.N,
                         nsplit, b0, d.beta);
      check_launch("splitk_reduce");
    }
  }
}

template <int S>
void dispatch_split(const GemmDesc& d, const GemmWorkspace& ws, hipStream_t stream) {
  if (d.M <= 64 || d.N <= 64)
    launch_split<64, 64, S>(d, ws, stream);
  else
    launch_split<128, 128, S>(d, ws, stream);
}

}  // namespace

size_t gemm_split_scratch_bytes(const GemmDesc& d, GemmPrecision prec, size_t ws_bytes) {
  (void)prec;  // every split kind tiles the same way
  const bool small = d.M <= 64 || d.N <= 64;
  return plan_split(d, small ? 64 : 128, small ? 64 : 128, ws_bytes).scratch_bytes;
}

void gemm_split(const GemmDesc& d, GemmPrecision prec, const GemmWorkspace& ws, hipStream_t stream) {
  switch (prec) {
    case GemmPrecision::BF16X3: dispatch_split<kBF16x2>(d, ws, stream); break;
    case GemmPrecision::BF16X6: dispatch_split<kBF16x3>(d, ws, stream); break;
    case GemmPrecision::F16X3: dispatch_split<kF16x2>(d, ws, stream); break;
    default: throw std::invalid_argument("gemm_split: not a split precision");
  }
}

}  // namespace mpqc_hip
