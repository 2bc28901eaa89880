// 16-bit operand planes: the fused permute+split producer and the NT
// contraction kernel on v_mfma_f32_16x16x32_{bf16,f16}.
//
// Kernel layout: 256 threads (2x2 wave64s), BM x BN output tile, K staged
// through LDS in 32-wide slices (one MFMA K), double buffered.  Both operands
// are k-contiguous in HBM and in LDS, so a slice row is copied with 16-byte
// loads and every MFMA fragment (8 consecutive k of one row) is one 16-byte
// LDS read.  The matrix cores accumulate in FP32 over K chunks of at most
// kMaxChunk; chunks are summed in FP64 by the split-K pass, which bounds the
// FP32 accumulation length without FP64 registers in the kernel (two
// workgroups fit per CU).  Workgroups are mapped to tiles XCD by XCD
// (consecutive tiles on the same XCD share operand rows in that XCD's L2).
#include "planes.hpp"

#include <algorithm>
#include <cmath>

namespace mpqc_hip {

namespace {

constexpr int kThr = 256;
constexpr int kBK = 32;
constexpr int kRow = kBK + 8;  // LDS row in halves (16-byte aligned rows)
constexpr int kXcd = 8;
constexpr int kMaxChunk = 2048;  // longest K summed in FP32

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------- producer
struct SplitArgs {
  uint16_t* p[3];
  double* scale;
  long n[4];
  int nterms;
  TTerm t[kMaxTerms];
  int kind;
};

__device__ inline double eval_term_sum(const SplitArgs& a, long e) {
  const long i3 = e % a.n[3];
  long r = e / a.n[3];
  const long i2 = r % a.n[2];
  r /= a.n[2];
  const long i1 = r % a.n[1];
  const long i0 = r / a.n[1];
  double v = 0.0;
  for (int t = 0; t < a.nterms; ++t) {
    const TTerm& tt = a.t[t];
    v += tt.alpha * tt.p[i0 * tt.s[0] + i1 * tt.s[1] + i2 * tt.s[2] + i3 * tt.s[3]];
  }
  return v;
}

__global__ __launch_bounds__(256) void split_max_stage1(SplitArgs a, double* partials) {
  const long total = a.n[0] * a.n[1] * a.n[2] * a.n[3];
  double m = 0.0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256)
    m = fmax(m, fabs(eval_term_sum(a, e)));
  __shared__ double sh[256];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// scale = 2^(14 - e) with max = f 2^e (f16); 1 for bf16
__global__ __launch_bounds__(256) void split_max_stage2(const double* partials, int nb, int kind,
                                                        double* scale) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) m = fmax(m, partials[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double s = 1.0;
    if (kind == (int)PlaneKind::F16x2) {
      const double mx = sh[0];
      if (mx > 0.0 && isfinite(mx)) {
        int e;
        frexp(mx, &e);
        s = ldexp(1.0, 14 - e);
      }
    }
    scale[0] = s;
  }
}

__device__ inline uint16_t f16_bits(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }
__device__ inline double f16_val(uint16_t u) { return (double)(float)__builtin_bit_cast(_Float16, u); }
__device__ inline uint16_t bf16_bits(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ inline double bf16_val(uint16_t u) { return (double)(float)__builtin_bit_cast(__bf16, u); }

__global__ __launch_bounds__(256) void split_write(SplitArgs a) {
  const long total = a.n[0] * a.n[1] * a.n[2] * a.n[3];
  const double s = *a.scale;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const double x = eval_term_sum(a, e) * s;
    if (a.kind == (int)PlaneKind::F16x2) {
      const uint16_t h = f16_bits((float)x);
      a.p[0][e] = h;
      a.p[1][e] = f16_bits((float)(x - f16_val(h)));
    } else {
      const uint16_t h = bf16_bits((float)x);
      const double r1 = x - bf16_val(h);
      const uint16_t m = bf16_bits((float)r1);
      a.p[0][e] = h;
      a.p[1][e] = m;
      if (a.kind == (int)PlaneKind::BF16x3) a.p[2][e] = bf16_bits((float)(r1 - bf16_val(m)));
    }
  }
}

// ---------------------------------------------------------------- contraction
struct PParams {
  const uint16_t* A[3];
  const uint16_t* B[3];
  const double* sa;
  const double* sb;
  double* C;
  double* W;
  long lda, ldb, ldc, sA, sB, sC;
  int M, N, K;
  int kchunk, nsplit, batch0, tiles_m, tiles_n, per_xcd;
  double alpha, beta;
};

template <int NP, bool F16>
__device__ inline void mma3(const u16x8_t* a, const u16x8_t* b, float4_t& c) {
  auto mf = [&](const u16x8_t& x, const u16x8_t& y) {
    if (F16)
      c = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, x),
                                                 __builtin_bit_cast(f16x8_t, y), c, 0, 0, 0);
    else
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, x),
                                                  __builtin_bit_cast(bf16x8_t, y), c, 0, 0, 0);
  };
  if (NP == 3) {  // 6 products, small terms first
    mf(a[1], b[1]);
    mf(a[0], b[2]);
    mf(a[2], b[0]);
  }
  mf(a[0], b[1]);
  mf(a[1], b[0]);
  mf(a[0], b[0]);
}

template <int BM, int BN, int NP, bool F16, bool VEC>
__global__ __launch_bounds__(kThr, 2) void gemm_planes_kernel(PParams p) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int RB = WM / 16, CB = WN / 16;
  constexpr int ACH = BM * (kBK / 8) / kThr;  // 16-byte chunks per thread per plane (A)
  constexpr int BCH = BN * (kBK / 8) / kThr;

  __shared__ __attribute__((aligned(16))) uint16_t As[2][NP][BM][kRow];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][NP][BN][kRow];

  // XCD-aware tile order: workgroup w runs on XCD w % 8; each XCD takes a
  // contiguous range of row-major tiles
  const int w = blockIdx.x;
  const int tile = (w % kXcd) * p.per_xcd + w / kXcd;
  if (tile >= p.tiles_m * p.tiles_n) return;
  const int tm = tile / p.tiles_n, tn = tile - tm * p.tiles_n;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  const int z = blockIdx.y;
  const int b = p.batch0 + z / p.nsplit;
  const int s = z - (z / p.nsplit) * p.nsplit;
  const int m0 = tm * BM, n0 = tn * BN;
  const int k_begin = s * p.kchunk;
  const int k_end = min(p.K, k_begin + p.kchunk);

  u16x8_t ra[NP][ACH], rb[NP][BCH];
  auto load_regs = [&](int k0) {
#pragma unroll
    for (int c = 0; c < ACH; ++c) {
      const int ch = tid + kThr * c;
      const int row = ch >> 2, kc = (ch & 3) * 8;
      const int gm = m0 + row, gk = k0 + kc;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        u16x8_t v = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint16_t* src = p.A[q] + (long)b * p.sA + (long)gm * p.lda + gk;
        if (gm < p.M) {
          if (VEC) {
            if (gk < k_end) v = *reinterpret_cast<const u16x8_t*>(src);
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
              if (gk + j < k_end) v[j] = src[j];
          }
        }
        ra[q][c] = v;
      }
    }
#pragma unroll
    for (int c = 0; c < BCH; ++c) {
      const int ch = tid + kThr * c;
      const int row = ch >> 2, kc = (ch & 3) * 8;
      const int gn = n0 + row, gk = k0 + kc;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        u16x8_t v = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint16_t* src = p.B[q] + (long)b * p.sB + (long)gn * p.ldb + gk;
        if (gn < p.N) {
          if (VEC) {
            if (gk < k_end) v = *reinterpret_cast<const u16x8_t*>(src);
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
              if (gk + j < k_end) v[j] = src[j];
          }
        }
        rb[q][c] = v;
      }
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int c = 0; c < ACH; ++c) {
      const int ch = tid + kThr * c;
      const int row = ch >> 2, kc = (ch & 3) * 8;
#pragma unroll
      for (int q = 0; q < NP; ++q) *reinterpret_cast<u16x8_t*>(&As[buf][q][row][kc]) = ra[q][c];
    }
#pragma unroll
    for (int c = 0; c < BCH; ++c) {
      const int ch = tid + kThr * c;
      const int row = ch >> 2, kc = (ch & 3) * 8;
#pragma unroll
      for (int q = 0; q < NP; ++q) *reinterpret_cast<u16x8_t*>(&Bs[buf][q][row][kc]) = rb[q][c];
    }
  };

  float4_t facc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) facc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

  int buf = 0;
  if (k_begin < k_end) {
    load_regs(k_begin);
    store_lds(0);
  }
  __syncthreads();
  for (int kt = k_begin; kt < k_end; kt += kBK) {
    const bool has_next = kt + kBK < k_end;
    if (has_next) load_regs(kt + kBK);
    u16x8_t fb[CB][NP];
#pragma unroll
    for (int j = 0; j < CB; ++j)
#pragma unroll
      for (int q = 0; q < NP; ++q)
        fb[j][q] = *reinterpret_cast<const u16x8_t*>(&Bs[buf][q][wn * WN + j * 16 + lr][lk * 8]);
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      u16x8_t fa[NP];
#pragma unroll
      for (int q = 0; q < NP; ++q)
        fa[q] = *reinterpret_cast<const u16x8_t*>(&As[buf][q][wm * WM + i * 16 + lr][lk * 8]);
#pragma unroll
      for (int j = 0; j < CB; ++j) mma3<NP, F16>(fa, fb[j], facc[i][j]);
    }
    if (has_next) store_lds(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  const double unscale = p.alpha / ((*p.sa) * (*p.sb));
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
          const double v = unscale * (double)facc[i][j][r];
          if (p.nsplit == 1) {
            double* c = C + (long)gm * p.ldc + gn;
            *c = (p.beta == 0.0) ? v : v + p.beta * (*c);
          } else {
            Wz[(long)gm * p.N + gn] = v;
          }
        }
      }
}

__global__ void planes_splitk_reduce(const double* __restrict__ W, double* __restrict__ C, long ldc,
                                     long sC, int M, int N, int nsplit, int batch0, double beta) {
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

template <int BM, int BN, int NP, bool F16>
void launch_planes(const PlaneGemmDesc& d, double* ws, size_t ws_bytes, hipStream_t stream) {
  const int tiles_m = (int)ceil_div(d.M, BM), tiles_n = (int)ceil_div(d.N, BN);
  const long tiles = (long)tiles_m * tiles_n;
  int nsplit = 1;
  const long want = 512;
  if (tiles * d.batch < want && d.K >= 8 * kBK) {
    nsplit = (int)std::min<long>(ceil_div(want, tiles * d.batch), d.K / (4 * kBK));
    nsplit = std::max(1, std::min(nsplit, 64));
  }
  // FP32 accumulation length bounded by kMaxChunk: more chunks, summed in FP64
  nsplit = std::max<int>(nsplit, (int)ceil_div(d.K, kMaxChunk));
  int kchunk = (int)(ceil_div(ceil_div(d.K, nsplit), kBK) * kBK);
  if (nsplit > 1) nsplit = (int)ceil_div(d.K, kchunk);
  const size_t slice = (size_t)d.M * d.N * sizeof(double);
  int max_batch = 65535 / nsplit;
  if (nsplit > 1) {
    const size_t cap = ws_bytes / (slice * nsplit);
    if (cap == 0) throw std::runtime_error("gemm_planes: split-K workspace too small");
    max_batch = (int)std::min<size_t>(max_batch, cap);
  }
  const bool vec = d.K % 8 == 0 && d.lda % 8 == 0 && d.ldb % 8 == 0 && d.sA % 8 == 0 &&
                   d.sB % 8 == 0;
  const int per_xcd = (int)ceil_div(tiles, kXcd);
  for (int b0 = 0; b0 < d.batch; b0 += max_batch) {
    const int nb = std::min(max_batch, d.batch - b0);
    PParams p;
    for (int q = 0; q < 3; ++q) {
      p.A[q] = d.A.p[q];
      p.B[q] = d.B.p[q];
    }
    p.sa = d.A.scale; p.sb = d.B.scale; p.C = d.C; p.W = ws;
    p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc; p.sA = d.sA; p.sB = d.sB; p.sC = d.sC;
    p.M = d.M; p.N = d.N; p.K = d.K; p.kchunk = kchunk; p.nsplit = nsplit; p.batch0 = b0;
    p.tiles_m = tiles_m; p.tiles_n = tiles_n; p.per_xcd = per_xcd;
    p.alpha = d.alpha; p.beta = d.beta;
    dim3 grid((unsigned)(per_xcd * kXcd), (unsigned)(nb * nsplit));
    if (vec)
      hipLaunchKernelGGL((gemm_planes_kernel<BM, BN, NP, F16, true>), grid, dim3(kThr), 0, stream, p);
    else
      hipLaunchKernelGGL((gemm_planes_kernel<BM, BN, NP, F16, false>), grid, dim3(kThr), 0, stream, p);
    check_launch("gemm_planes");
    if (nsplit > 1) {
      const long MN = (long)d.M * d.N;
      dim3 g2((unsigned)std::min<long>(ceil_div(MN, 256), 4096), (unsigned)nb);
      hipLaunchKernelGGL(planes_splitk_reduce, g2, dim3(256), 0, stream, ws, d.C, d.ldc, d.sC, d.M,
                         d.N, nsplit, b0, d.beta);
      check_launch("planes_splitk_reduce");
    }
  }
}

template <int NP, bool F16>
void dispatch_planes(const PlaneGemmDesc& d, double* ws, size_t ws_bytes, hipStream_t stream) {
  if (d.M <= 64 || d.N <= 64)
    launch_planes<64, 64, NP, F16>(d, ws, ws_bytes, stream);
  else
    launch_planes<128, 128, NP, F16>(d, ws, ws_bytes, stream);
}

}  // namespace

void split_to_planes(PlaneKind kind, const Planes& out, const long n[4], const TTerm* terms,
                     int nterms, double* partials, hipStream_t stream) {
  if (nterms < 1 || nterms > kMaxTerms) throw std::invalid_argument("split_to_planes: bad term count");
  SplitArgs a;
  for (int q = 0; q < 3; ++q) a.p[q] = out.p[q];
  a.scale = out.scale;
  for (int d = 0; d < 4; ++d) a.n[d] = n[d];
  a.nterms = nterms;
  for (int t = 0; t < nterms; ++t) a.t[t] = terms[t];
  a.kind = (int)kind;
  const long total = n[0] * n[1] * n[2] * n[3];
  if (total > out.n) throw std::invalid_argument("split_to_planes: planes too small");
  const int g = (int)std::max<long>(1, std::min<long>(1024, ceil_div(total, 256)));
  hipLaunchKernelGGL(split_max_stage1, dim3(g), dim3(256), 0, stream, a, partials);
  hipLaunchKernelGGL(split_max_stage2, dim3(1), dim3(256), 0, stream, partials, g, (int)kind, out.scale);
  const int g2 = (int)std::max<long>(1, std::min<long>(16384, ceil_div(total, 256)));
  hipLaunchKernelGGL(split_write, dim3(g2), dim3(256), 0, stream, a);
  check_launch("split_to_planes");
}

size_t gemm_planes_workspace_bytes(int M, int N, int K, int batch) {
  const long tiles = ceil_div(M, 128) * ceil_div(N, 128);
  long nsplit = 1;
  if (tiles * batch < 512 && K >= 8 * kBK) nsplit = std::min<long>(64, std::max<long>(1, std::min<long>(ceil_div(512, tiles * batch), K / (4 * kBK))));
  nsplit = std::max<long>(nsplit, ceil_div(K, kMaxChunk)) + 1;
  return nsplit > 2 ? (size_t)nsplit * M * N * batch * sizeof(double) : 0;
}

void gemm_planes(const PlaneGemmDesc& d, double* ws, size_t ws_bytes, hipStream_t stream) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  if (d.K <= 0) throw std::invalid_argument("gemm_planes: K must be positive");
  switch (d.kind) {
    case PlaneKind::BF16x2: dispatch_planes<2, false>(d, ws, ws_bytes, stream); break;
    case PlaneKind::F16x2: dispatch_planes<2, true>(d, ws, ws_bytes, stream); break;
    case PlaneKind::BF16x3: dispatch_planes<3, false>(d, ws, ws_bytes, stream); break;
  }
}

}  // namespace mpqc_hip
