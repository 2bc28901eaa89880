// 16-bit operand planes: the fused permute+split producer and the NT
// contraction kernels on v_mfma_f32_{16x16x32,32x32x16}_{bf16,f16}.
//
// Large two-plane contractions run on the deep-pipeline kernel (256 x 256 tiles,
// eight waves, direct global->LDS loads; described at gemm_planes_dl_kernel).  The
// others run on the 64/128-tile kernel: 256 threads (2x2 wave64s), BM x BN output
// tile, K staged through LDS in 32-wide slices (one MFMA K), double buffered.  Both
// operands are k-contiguous in HBM and in LDS, so a slice row is copied with 16-byte
// loads and every MFMA fragment (8 consecutive k of one row) is one 16-byte LDS read.
// The matrix cores accumulate in FP32 over K chunks of at most kMaxChunk: longer
// contractions are split into chunks (split-K) whose partial sums are added in FP64
// in a fixed order by a separate pass, which bounds the FP32 accumulation length
// without FP64 registers in the main loop.  A workgroup whose K range is longer than
// kMaxChunk (not produced by the planner) would fold its FP32 sums into the FP64
// output after every chunk.  Workgroups are mapped to tiles XCD by XCD (consecutive
// tiles on the same XCD share operand rows in that XCD's L2).
#include "planes.hpp"
#include "gemm.hpp"
#include "permute.hpp"

#include <algorithm>
#include <cmath>

namespace mpqc_hip {

namespace {

constexpr int kThr = 256;
constexpr int kBK = 32;
// LDS rows of one K slice: 32 halves = 64 bytes = four 16-byte chunks, unpadded;
// chunk c of row r is stored at chunk c ^ ((r >> 2) & 3).  A 16-lane phase of the
// fragment reads (16 rows, one chunk) and of the staging writes (4 rows, 4
// chunks) then touches 16 different 16-byte bank groups.
constexpr int kRow = kBK;
__device__ inline int swz(int row, int chunk) { return (chunk ^ ((row >> 2) & 3)) * 8; }
constexpr int kXcd = 8;
constexpr int kXcdMinTiles = 64;  // fewer output tiles: plain tile order (K splits spread over every XCD)
constexpr int kMaxChunk = 2048;  // longest K summed in FP32

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------- producer
__device__ inline uint16_t f16_bits(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }
__device__ inline double f16_val(uint16_t u) { return (double)(float)__builtin_bit_cast(_Float16, u); }
__device__ inline uint16_t bf16_bits(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ inline double bf16_val(uint16_t u) { return (double)(float)__builtin_bit_cast(__bf16, u); }

// max |x| over each distinct input tensor (contiguous reads): partials[src * gridDim.x + block]
struct AbsMaxArgs {
  const double* p[kMaxTerms + 2];
  long n[kMaxTerms + 2];
};

__global__ __launch_bounds__(256) void absmax_stage1(AbsMaxArgs a, double* partials) {
  const int src = blockIdx.y;
  const double* __restrict__ x = a.p[src];
  const long n = a.n[src];
  double m = 0.0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256)
    m = fmax(m, fabs(x[e]));
  __shared__ double sh[256];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[(long)src * gridDim.x + blockIdx.x] = sh[0];
}

// Plane scale from the bound  max|sum_t alpha_t x_t| <= sum_t |alpha_t| max|x_t|  (exact for
// one input): s = 2^(14 - e) with bound = f 2^e, so every scaled value stays below 2^14 and
// far from the fp16 overflow.  bf16 planes need no scale (s = 1).
struct ScaleArgs {
  int nsrc, nterms, nb, kind;
  int src_of[kMaxTerms];
  double absalpha[kMaxTerms];
  int prod_x, prod_y;  // sources of the product term's factors (-1: none)
  double prod_absalpha;
  const double* cached[kMaxTerms + 2];  // max |x| already known (no partials)
  double* record[kMaxTerms + 2];        // where to record a computed max
};

__global__ __launch_bounds__(256) void scale_stage2(const double* partials, ScaleArgs a, double* scale) {
  __shared__ double sh[256];
  __shared__ double srcmax[kMaxTerms + 2];
  for (int src = 0; src < a.nsrc; ++src) {
    if (a.cached[src]) {
      if (threadIdx.x == 0) srcmax[src] = *a.cached[src];
      __syncthreads();
      continue;
    }
    double m = 0.0;
    for (int i = threadIdx.x; i < a.nb; i += 256) m = fmax(m, partials[(long)src * a.nb + i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      srcmax[src] = sh[0];
      if (a.record[src]) *a.record[src] = sh[0];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double s = 1.0;
    if (a.kind == (int)PlaneKind::F16x2) {
      double bound = 0.0;
      for (int t = 0; t < a.nterms; ++t) bound += a.absalpha[t] * srcmax[a.src_of[t]];
      if (a.prod_x >= 0) bound += a.prod_absalpha * srcmax[a.prod_x] * srcmax[a.prod_y];
      if (bound > 0.0 && isfinite(bound)) {
        int e;
        frexp(bound, &e);
        s = ldexp(1.0, 14 - e);
      }
    }
    scale[0] = s;
  }
}

struct SplitOut {
  uint16_t* p[3];
  const double* scale;
  int kind;
};

template <int SHAPE>
__global__ __launch_bounds__(kPThreads) void split_write_kernel(PermGeom g, SplitOut o) {
  const double sc = *o.scale;
  permute_block<SHAPE>(g, blockIdx.x, blockIdx.y, blockIdx.z, [&](long e, double v) {
    const double x = v * sc;
    if (o.kind == (int)PlaneKind::F16x2) {
      const uint16_t h = f16_bits((float)x);
      o.p[0][e] = h;
      o.p[1][e] = f16_bits((float)(x - f16_val(h)));
    } else {
      const uint16_t h = bf16_bits((float)x);
      const double r1 = x - bf16_val(h);
      const uint16_t m = bf16_bits((float)r1);
      o.p[0][e] = h;
      o.p[1][e] = m;
      if (o.kind == (int)PlaneKind::BF16x3) o.p[2][e] = bf16_bits((float)(r1 - bf16_val(m)));
    }
  });
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
  int order;  // tile order: 0 XCD ranges of row-major tiles, 1 plain, 2 XCD ranges column-major
  double alpha, beta;
};

template <bool F16>
__device__ inline void mfma1(const u16x8_t& x, const u16x8_t& y, float4_t& c) {
  if (F16)
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, x),
                                               __builtin_bit_cast(f16x8_t, y), c, 0, 0, 0);
  else
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, x),
                                                __builtin_bit_cast(bf16x8_t, y), c, 0, 0, 0);
}

// Plane products of one K slice in the order of increasing magnitude (small
// terms first): lo*lo-class terms of the 3-plane form, then hi*lo, lo*hi,
// hi*hi.  Product-major order: consecutive MFMAs write different accumulators,
// so each dependent MFMA comes RB * CB instructions after the one it extends.
template <int NP, int RB, int CB, bool F16>
__device__ inline void mma_slice(const u16x8_t (&fa)[RB][NP], const u16x8_t (&fb)[CB][NP],
                                 float4_t (&acc)[RB][CB]) {
  // (plane of A, plane of B) of the six products; the 2-plane form uses the last three
  constexpr int kPA[6] = {1, 0, 2, 0, 1, 0};
  constexpr int kPB[6] = {1, 2, 0, 1, 0, 0};
#pragma unroll
  for (int t = (NP == 3 ? 0 : 3); t < 6; ++t)
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int j = 0; j < CB; ++j) mfma1<F16>(fa[i][kPA[t]], fb[j][kPB[t]], acc[i][j]);
}

// Occupancy per tile size: the 128 x 128 tile fits two workgroups per CU; the
// 256-row tiles use the whole LDS and up to 512 registers per lane (the MFMA
// accumulators of a 128 x 128 wave tile live in AGPRs), one workgroup per CU.
template <int BM, int BN, int NBUF>
constexpr int plane_occupancy() { return NBUF == 1 ? 3 : BM * BN >= 256 * 128 ? 1 : 2; }

// BD (B direct): the B fragments of each wave are loaded from global memory
// (L1/L2) straight into registers one K slice ahead instead of being staged
// through LDS; this halves the LDS traffic per slice (the kernel's limit).
// NBUF: LDS stages of the K slices.  Two overlap the next slice's LDS stores with
// the current products (64 KB at 128 x 128: two workgroups per CU); one halves the
// LDS so that three workgroups share a CU and hide each other's load latency, at the
// price of a second barrier per slice (its A fragments stream as in the 256-row
// tiles, which keeps it within the registers of three waves per SIMD).
template <int BM, int BN, int NP, bool F16, bool VEC, bool BD, int NBUF = 2>
__global__ __launch_bounds__(kThr, (plane_occupancy<BM, BN, NBUF>())) void gemm_planes_kernel(PParams p) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int RB = WM / 16, CB = WN / 16;
  constexpr int ACH = BM * (kBK / 8) / kThr;  // 16-byte chunks per thread per plane (A)
  constexpr int BCH = BD ? 0 : BN * (kBK / 8) / kThr;
  static_assert(!BD || (VEC && NP == 2), "direct B fragments need 16-byte rows and two planes");

  __shared__ __attribute__((aligned(16))) uint16_t As[NBUF][NP][BM][kRow];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[NBUF][NP][BD ? 1 : BN][kRow];

  // XCD-aware tile order: workgroup w runs on XCD w % 8; each XCD takes a
  // contiguous range of row-major tiles
  const int w = blockIdx.x;
  const int tile = (p.order == 1 || p.per_xcd == 0) ? w : (w % kXcd) * p.per_xcd + w / kXcd;
  if (tile >= p.tiles_m * p.tiles_n) return;
  int tm, tn;
  if (p.order == 2) {  // column-major tiles
    tn = tile / p.tiles_m;
    tm = tile - tn * p.tiles_m;
  } else {
    tm = tile / p.tiles_n;
    tn = tile - tm * p.tiles_n;
  }

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

  u16x8_t ra[NP][ACH], rb[NP][BCH > 0 ? BCH : 1];
  // B fragments of this wave for K slice k0 (lane: row n0 + wn*WN + 16j + lr, k = k0 + 8 lk .. +7)
  auto load_bfrag = [&](int k0, u16x8_t (&dst)[CB][NP]) {
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      const int gn = n0 + wn * WN + j * 16 + lr, gk = k0 + lk * 8;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        u16x8_t v = {0, 0, 0, 0, 0, 0, 0, 0};
        if (gn < p.N && gk < k_end)
          v = *reinterpret_cast<const u16x8_t*>(p.B[q] + (long)b * p.sB + (long)gn * p.ldb + gk);
        dst[j][q] = v;
      }
    }
  };
  // VEC loads are unconditional from clamped addresses and the out-of-range
  // chunks are zeroed when they are stored to LDS, so no instruction needs the
  // loaded data before the store (the loads stay in flight across the MFMAs)
  uint32_t amask = ~0u, bmask = ~0u;
  auto load_regs = [&](int k0) {
#pragma unroll
    for (int c = 0; c < ACH; ++c) {
      const int ch = tid + kThr * c;
      const int row = ch >> 2, kc = (ch & 3) * 8;
      const int gm = m0 + row, gk = k0 + kc;
      if constexpr (VEC) {
        const bool ok = gm < p.M && gk < k_end;
        amask = ok ? (amask | (1u << c)) : (amask & ~(1u << c));
        const long off = (long)b * p.sA + (long)min(gm, p.M - 1) * p.lda + min(gk, p.K - 8);
#pragma unroll
        for (int q = 0; q < NP; ++q) ra[q][c] = *reinterpret_cast<const u16x8_t*>(p.A[q] + off);
      } else {
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          u16x8_t v = {0, 0, 0, 0, 0, 0, 0, 0};
          const uint16_t* src = p.A[q] + (long)b * p.sA + (long)gm * p.lda + gk;
          if (gm < p.M) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
              if (gk + j < k_end) v[j] = src[j];
          }
          ra[q][c] = v;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < BCH; ++c) {
      const int ch = tid + kThr * c;
      const int row = ch >> 2, kc = (ch & 3) * 8;
      const int gn = n0 + row, gk = k0 + kc;
      if constexpr (VEC) {
        const bool ok = gn < p.N && gk < k_end;
        bmask = ok ? (bmask | (1u << c)) : (bmask & ~(1u << c));
        const long off = (long)b * p.sB + (long)min(gn, p.N - 1) * p.ldb + min(gk, p.K - 8);
#pragma unroll
        for (int q = 0; q < NP; ++q) rb[q][c] = *reinterpret_cast<const u16x8_t*>(p.B[q] + off);
      } else {
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          u16x8_t v = {0, 0, 0, 0, 0, 0, 0, 0};
          const uint16_t* src = p.B[q] + (long)b * p.sB + (long)gn * p.ldb + gk;
          if (gn < p.N) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
              if (gk + j < k_end) v[j] = src[j];
          }
          rb[q][c] = v;
        }
      }
    }
  };
  auto store_lds = [&](int buf) {
    const u16x8_t zero = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < ACH; ++c) {
      const int ch = tid + kThr * c;
      const int row = ch >> 2, kc = (ch & 3) * 8;
      const bool ok = (amask >> c) & 1u;
#pragma unroll
      for (int q = 0; q < NP; ++q)
        *reinterpret_cast<u16x8_t*>(&As[buf][q][row][swz(row, kc / 8)]) = ok ? ra[q][c] : zero;
    }
#pragma unroll
    for (int c = 0; c < BCH; ++c) {
      const int ch = tid + kThr * c;
      const int row = ch >> 2, kc = (ch & 3) * 8;
      const bool ok = (bmask >> c) & 1u;
#pragma unroll
      for (int q = 0; q < NP; ++q)
        *reinterpret_cast<u16x8_t*>(&Bs[buf][q][row][swz(row, kc / 8)]) = ok ? rb[q][c] : zero;
    }
  };

  float4_t facc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) facc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

  // FP32 sums of at most kMaxChunk values of k are added into the FP64 output
  // (the first flush applies beta to C, or starts the split-K partial)
  const double unscale = p.alpha / ((*p.sa) * (*p.sb));
  double* __restrict__ C = p.C + (long)b * p.sC;
  double* __restrict__ Wz = p.W + (long)z * p.M * p.N;
  bool first_flush = true;
  auto flush = [&]() {
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int j = 0; j < CB; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gm = m0 + wm * WM + i * 16 + lk * 4 + r;
          const int gn = n0 + wn * WN + j * 16 + lr;
          if (gm < p.M && gn < p.N) {
            const double v = unscale * (double)facc[i][j][r];
            double* c = (p.nsplit == 1) ? C + (long)gm * p.ldc + gn : Wz + (long)gm * p.N + gn;
            if (!first_flush) *c += v;
            else if (p.nsplit > 1 || p.beta == 0.0) *c = v;
            else *c = v + p.beta * (*c);
          }
        }
        facc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
      }
    first_flush = false;
  };

  int buf = 0, since_flush = 0;
  u16x8_t fbd[CB][NP];  // direct B fragments of the current slice (BD)
  if (k_begin < k_end) {
    load_regs(k_begin);
    store_lds(0);
    if constexpr (BD) load_bfrag(k_begin, fbd);
  }
  __syncthreads();
  for (int kt = k_begin; kt < k_end; kt += kBK) {
    const bool has_next = kt + kBK < k_end;
    if (has_next) load_regs(kt + kBK);
    u16x8_t fb[CB][NP], fbn[CB][NP];
    if constexpr (BD) {
#pragma unroll
      for (int j = 0; j < CB; ++j)
#pragma unroll
        for (int q = 0; q < NP; ++q) fb[j][q] = fbd[j][q];
      if (has_next) load_bfrag(kt + kBK, fbn);
    } else {
#pragma unroll
      for (int j = 0; j < CB; ++j)
#pragma unroll
        for (int q = 0; q < NP; ++q)
          fb[j][q] = *reinterpret_cast<const u16x8_t*>(&Bs[buf][q][wn * WN + j * 16 + lr][swz(lr, lk)]);
    }
    if constexpr (NP == 2 && (BM >= 256 || NBUF == 1)) {
      // 128-row wave tiles: the B fragments of the slice stay in registers and the A
      // fragments stream one 16-row block at a time, read one block ahead of its products
      u16x8_t fa[2][1][NP];
#pragma unroll
      for (int q = 0; q < NP; ++q)
        fa[0][0][q] = *reinterpret_cast<const u16x8_t*>(&As[buf][q][wm * WM + lr][swz(lr, lk)]);
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        if (i + 1 < RB)
#pragma unroll
          for (int q = 0; q < NP; ++q)
            fa[(i + 1) & 1][0][q] =
                *reinterpret_cast<const u16x8_t*>(&As[buf][q][wm * WM + (i + 1) * 16 + lr][swz(lr, lk)]);
        mma_slice<NP, 1, CB, F16>(fa[i & 1], fb, reinterpret_cast<float4_t(&)[1][CB]>(facc[i]));
      }
    } else if constexpr (NP == 2) {
      // all fragments of the slice first (one LDS latency per slice), then the products
      u16x8_t fa[RB][NP];
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int q = 0; q < NP; ++q)
          fa[i][q] = *reinterpret_cast<const u16x8_t*>(&As[buf][q][wm * WM + i * 16 + lr][swz(lr, lk)]);
      // keep every fragment read of the slice ahead of the first product, so the
      // scheduler issues them together (one LDS latency) and does not re-read
      // fragments between products to save registers
      __builtin_amdgcn_sched_barrier(0);
      mma_slice<NP, RB, CB, F16>(fa, fb, facc);
    } else {
      // three planes: one row block at a time keeps two workgroups per CU
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        u16x8_t fa[1][NP];
#pragma unroll
        for (int q = 0; q < NP; ++q)
          fa[0][q] = *reinterpret_cast<const u16x8_t*>(&As[buf][q][wm * WM + i * 16 + lr][swz(lr, lk)]);
        mma_slice<NP, 1, CB, F16>(fa, fb, reinterpret_cast<float4_t(&)[1][CB]>(facc[i]));
      }
    }
    if constexpr (BD) {
      if (has_next)
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
          for (int q = 0; q < NP; ++q) fbd[j][q] = fbn[j][q];
    }
    if constexpr (BM < 256 && NBUF == 2) {  // (the planner keeps every K range within kMaxChunk)
      if (++since_flush == kMaxChunk / kBK && has_next) {
        flush();
        since_flush = 0;
      }
    }
    if constexpr (NBUF == 2) {
      if (has_next) store_lds(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    } else {
      if (has_next) {
        __syncthreads();  // every wave has read the slice
        store_lds(0);
      }
      __syncthreads();
    }
  }
  flush();  // also covers an empty K range (writes beta C or a zero partial)
}

// ---------------------------------------------------------------- deep-pipeline contraction
// Large two-plane contractions: a 256 x 256 output tile per workgroup of eight waves
// (two per SIMD, a 2 x 4 grid of 128 x 64 wave tiles, each 8 x 4 blocks of
// v_mfma_f32_16x16x32, which keep the matrix cores busier at two waves per SIMD than
// 4 x 2 blocks of 32x32x16), one workgroup per CU.  K advances in 32-wide steps through two
// LDS stages of 64 KB that direct global->LDS loads (global_load_lds_dwordx4, no staging
// registers) fill: a stage row is 64 contiguous bytes of one plane, so every 64-byte
// request of the loads is used whole.  The next step's loads are issued when its stage
// is free and overlap the current step's products; one barrier per step.  The tiles an
// XCD owns are walked in groups of four tile rows, column-major inside a group, so the
// workgroups of one XCD share A rows and B rows in its L2.  The epilogue stages each
// wave's FP32 sums through the (then free) LDS and writes whole rows with 16-byte
// stores: FP32 partials of a K chunk (at most kMaxChunk) for a fixed-order FP64 reduce,
// or C = alpha / (sa sb) sum + beta C in FP64.
constexpr int kDlTile = 256;
constexpr int kDlBK = 32;       // K per stage: one 16x16x32 MFMA step
constexpr int kDlThreads = 512; // eight waves
constexpr int kDlGroup = 4;     // tile rows per group of the tile order

typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));
#define MPQC_LDS(p) ((__attribute__((address_space(3))) void*)(p))

// zero source for the 8-wide k chunks beyond a contraction's K range
__device__ __attribute__((aligned(16))) uint16_t g_dl_zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};

// A stage row is 32 halves (four 16-byte chunks); chunk c of row r sits at chunk
// c ^ dl_sw(r).  A fragment read covers rows r = 0..15 (lane & 15) at logical chunk
// lane >> 4.  dl_sw(r) = r0 | r2 << 1 (bits of r) was chosen by measurement among eight
// maps of bits of r (standalone copy of this kernel at 6080^3, SQ_LDS_BANK_CONFLICT): the
// only one whose fragment reads count no bank-conflict cycles (the others, including maps
// conflict-free for any 8 or 16 consecutive lanes, count about four per read), and the
// fastest by 1 %.
__device__ inline int dl_sw(int r) { return (r & 1) | (((r >> 2) & 1) << 1); }
// The same for fragments of 32 rows of one chunk per 32 lanes (v_mfma_f32_32x32x16).
__device__ inline int dl_sw32(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

template <bool F16>
__device__ inline void mfma32(const u16x8_t& x, const u16x8_t& y, f32x16_t& c) {
  if constexpr (F16)
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_t, x), __builtin_bit_cast(f16x8_t, y), c,
                                               0, 0, 0);
  else
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, x), __builtin_bit_cast(bf16x8_t, y), c,
                                                0, 0, 0);
}

// s_waitcnt with only vmcnt(0) or only lgkmcnt(0) (the immediate holds vmcnt in bits 3:0
// and 15:14, expcnt in 6:4, lgkmcnt in 11:8)
__device__ inline void dl_wait_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }
__device__ inline void dl_wait_lds() { __builtin_amdgcn_s_waitcnt(0xC07F); }
__device__ inline void dl_barrier() {
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_sched_barrier(0);
}

struct DlParams {
  const uint16_t* A[2];
  const uint16_t* B[2];
  const double* sa;
  const double* sb;
  double* C;
  float* W;  // raw FP32 partials (nsplit > 1): [batch in launch][split][M][N]
  long lda, ldb, ldc, sA, sB, sC;
  int M, N, K;
  int kchunk, nsplit, batch0, tiles_m, tiles_n, per_xcd;
  double alpha, beta;
};

template <bool F16>
__global__ __launch_bounds__(kDlThreads, 1) void gemm_planes_dl_kernel(DlParams p) {
  // [stage][operand][plane][row * 32 + physical chunk * 8]
  __shared__ __attribute__((aligned(1024))) uint16_t lds[2][2][2][kDlTile * kDlBK];

  const int w = blockIdx.x;
  const int tile = p.per_xcd == 0 ? w : (w % kXcd) * p.per_xcd + w / kXcd;  // XCD ranges of the tile order
  if (tile >= p.tiles_m * p.tiles_n) return;
  const int grp = tile / (kDlGroup * p.tiles_n), first = grp * kDlGroup;
  const int rows = min(kDlGroup, p.tiles_m - first), in = tile - grp * kDlGroup * p.tiles_n;
  const int tm = first + in % rows, tn = in / rows;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  const int z = blockIdx.y;
  const int b = p.batch0 + z / p.nsplit;
  const int s = z - (z / p.nsplit) * p.nsplit;
  const int m0 = tm * kDlTile, n0 = tn * kDlTile;
  const int k_begin = s * p.kchunk;
  const int k_end = min(p.K, k_begin + p.kchunk);
  const int nsteps = k_end > k_begin ? (k_end - k_begin + kDlBK - 1) / kDlBK : 0;

  // this lane's 16 bytes of each load: tile row wave*32 + hq*16 + lane/4, stored at
  // physical chunk lane&3, i.e. logical chunk (lane&3) ^ dl_sw(row)
  const uint16_t* src[2][2][2];  // [operand][plane][hq]: row start + logical chunk offset
  int kofs[2];                   // [hq]: the chunk's first k relative to the step's k0
#pragma unroll
  for (int hq = 0; hq < 2; ++hq) {
    const int r = wave * 32 + hq * 16 + (lane >> 2);
    kofs[hq] = ((lane & 3) ^ dl_sw(r)) * 8;
    const long ga = (long)b * p.sA + (long)min(m0 + r, p.M - 1) * p.lda + kofs[hq];
    const long gb = (long)b * p.sB + (long)min(n0 + r, p.N - 1) * p.ldb + kofs[hq];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      src[0][q][hq] = p.A[q] + ga;
      src[1][q][hq] = p.B[q] + gb;
    }
  }
  auto issue = [&](int t, int st) {
    const int k0 = k_begin + t * kDlBK;
#pragma unroll
    for (int op = 0; op < 2; ++op)
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int hq = 0; hq < 2; ++hq) {
          const uint16_t* g = k0 + kofs[hq] < k_end ? src[op][q][hq] + k0 : g_dl_zero;
          __builtin_amdgcn_global_load_lds((const void*)g, MPQC_LDS(&lds[st][op][q][(wave * 32 + hq * 16) * kDlBK]),
                                           16, 0, 0);
        }
  };

  // fragments (v_mfma_f32_16x16x32: lane -> row lane & 15, logical chunk lane >> 4): the
  // wave's 128 x 64 tile is 8 x 4 blocks of 16 x 16, one MFMA K step per stage
  const int lr = lane & 15, lk = lane >> 4;
  const int frag = lr * kDlBK + ((lk ^ dl_sw(lr)) * 8);
  u16x8_t fa[8][2], fb[4][2];  // [block][plane]
  float4_t acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.f;

  if (nsteps > 0) {
    issue(0, 0);
    dl_wait_loads();
    dl_barrier();
    for (int t = 0; t < nsteps; ++t) {
      const int st = t & 1;
      // stage st^1 was read in step t-1, which every wave finished before the barrier
      if (t + 1 < nsteps) issue(t + 1, st ^ 1);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
          fa[i][q] = *reinterpret_cast<const u16x8_t*>(&lds[st][0][q][(wm * 128 + i * 16) * kDlBK + frag]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          fb[j][q] = *reinterpret_cast<const u16x8_t*>(&lds[st][1][q][(wn * 64 + j * 16) * kDlBK + frag]);
      }
      // products in increasing magnitude: hi*lo, lo*hi, hi*hi
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) mfma1<F16>(fa[i][0], fb[j][1], acc[i][j]);
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) mfma1<F16>(fa[i][1], fb[j][0], acc[i][j]);
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) mfma1<F16>(fa[i][0], fb[j][0], acc[i][j]);
      // (keeps the waits below every product: hoisted above them, they would leave the
      // next stage's loads less time to land)
      __builtin_amdgcn_sched_barrier(0);
      dl_wait_loads();  // step t+1's stage has landed (this wave's share)
      dl_wait_lds();    // this wave's reads of stage st are done
      dl_barrier();
    }
  }

  // epilogue: each wave stages its 128 x 64 FP32 tile through 16 KB of LDS, 64 rows at a
  // time, and writes rows of 64 values with 16-byte accesses (4 lanes per 64 bytes)
  const double unscale = p.alpha / ((*p.sa) * (*p.sb));
  double* __restrict__ C = p.C + (long)b * p.sC;
  float* __restrict__ Wz = p.W + (long)z * p.M * p.N;
  float* stg = reinterpret_cast<float*>(&lds[0][0][0][0]) + wave * (64 * 64);
  if (nsteps == 0) dl_barrier();  // (the K loop's last barrier otherwise)
#pragma unroll
  for (int i0 = 0; i0 < 8; i0 += 4) {
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)  // (column ^ 16 ((row >> 2) & 3): the four lane groups hit distinct banks)
          stg[(ii * 16 + lk * 4 + r) * 64 + ((j ^ lk) * 16 + lr)] = acc[i0 + ii][j][r];
    dl_wait_lds();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = q * 64 + lane, row = idx >> 4, c4 = idx & 15;
      const f32x4_t v = *reinterpret_cast<const f32x4_t*>(&stg[row * 64 + ((c4 * 4) ^ (((row >> 2) & 3) * 16))]);
      const int gm = m0 + wm * 128 + i0 * 16 + row, gn = n0 + wn * 64 + c4 * 4;
      if (gm >= p.M) continue;
      if (p.nsplit > 1) {
        float* wr = Wz + (long)gm * p.N + gn;
        if (gn + 3 < p.N && ((uintptr_t)wr & 15) == 0) {
          *reinterpret_cast<f32x4_t*>(wr) = v;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (gn + e < p.N) wr[e] = v[e];
        }
      } else {
        double* c = C + (long)gm * p.ldc + gn;
        double o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = unscale * (double)v[e];
        if (gn + 3 < p.N && ((uintptr_t)c & 15) == 0) {
          f64x2_t lo = {o[0], o[1]}, hi = {o[2], o[3]};
          if (p.beta != 0.0) {
            const f64x2_t c0 = *reinterpret_cast<const f64x2_t*>(c), c1 = *reinterpret_cast<const f64x2_t*>(c + 2);
            lo += p.beta * c0;
            hi += p.beta * c1;
          }
          *reinterpret_cast<f64x2_t*>(c) = lo;
          *reinterpret_cast<f64x2_t*>(c + 2) = hi;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (gn + e < p.N) c[e] = p.beta == 0.0 ? o[e] : o[e] + p.beta * c[e];
        }
      }
    }
    dl_wait_lds();
  }
}

// ---------------------------------------------------------------- tall-K contraction
// Two-plane contractions with a small output (M, N <= 160) and a long K (the r1
// intermediates F_oo, F_vv and the Giabc / Gijka tau terms: K = o v^2 or o^2 v): a
// workgroup owns the whole output, padded to MB x NB blocks of 32 x 32,
// for one range of K, so the grid is the K ranges and every operand byte is read once.
// The range's 32-wide steps go through kStages LDS stages filled by direct
// global->LDS loads (rows beyond M or N read the zero page: no HBM traffic), up to
// kStages - 1 of them in flight; wave w of W (four, or eight for large outputs) owns
// output blocks w, w + W, ...  FP32 sums
// over at most kMaxChunk values of k are folded into FP64 registers, and the range's
// FP64 sums (times alpha / (sa sb)) are written as split-K partials for the
// fixed-order reduce.
template <int MB, int NB>
struct TkShape {
  static constexpr int TM = MB * 32, TN = NB * 32;
  static constexpr int kRows = TM + TN;               // stage rows per plane: A rows, then B rows
  static constexpr int kInstr = kRows / 16 * 2;       // 16-row loads per stage (both planes)
  static constexpr int kBlocks = MB * NB;
  // eight waves where four would hold too many accumulators (FP32 and FP64) per lane
  static constexpr int kWaves = kBlocks > 16 && kInstr % 8 == 0 ? 8 : 4;
  static constexpr int kPerWave = kInstr / kWaves;    // equal for every wave
  static constexpr int kStageHalves = 2 * kRows * 32; // [plane][row][32 halves]
  static constexpr int kStages = 4 * kStageHalves * 2 <= 112 * 1024 ? 4 : 3;
  static constexpr int kMaxBlocksPerWave = (kBlocks + kWaves - 1) / kWaves;
  static_assert(kInstr % kWaves == 0, "loads must split evenly over the waves");
};

// s_waitcnt vmcnt(n * L) for n = 0, 1, 2 (expcnt and lgkmcnt not waited on)
template <int L>
__device__ inline void tk_wait_steps(int n) {
  static_assert(2 * L < 64, "vmcnt holds 6 bits");
  if (n <= 0) __builtin_amdgcn_s_waitcnt(0x0F70);
  else if (n == 1) __builtin_amdgcn_s_waitcnt(0x0F70 | (L & 15) | ((L >> 4) << 14));
  else __builtin_amdgcn_s_waitcnt(0x0F70 | ((2 * L) & 15) | (((2 * L) >> 4) << 14));
}

template <int MB, int NB, bool F16>
__global__ __launch_bounds__((64 * TkShape<MB, NB>::kWaves)) void gemm_planes_tallk_kernel(DlParams p) {
  using S = TkShape<MB, NB>;
  constexpr int ST = S::kStages;
  __shared__ __attribute__((aligned(1024))) uint16_t lds[ST][S::kStageHalves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int z = blockIdx.x;
  const int b = p.batch0 + z / p.nsplit;
  const int s = z - (z / p.nsplit) * p.nsplit;
  const int k_begin = s * p.kchunk;
  const int k_end = min(p.K, k_begin + p.kchunk);
  const int nsteps = k_end > k_begin ? (k_end - k_begin + kDlBK - 1) / kDlBK : 0;

  // load q = wave + kWaves i of a stage: plane q / (kRows/16), rows 16 * (q % (kRows/16)) .. +15
  const uint16_t* src[S::kPerWave];
  int kofs[S::kPerWave], dst[S::kPerWave];
  bool valid[S::kPerWave];
#pragma unroll
  for (int i = 0; i < S::kPerWave; ++i) {
    const int q = wave + S::kWaves * i;
    const int plane = q / (S::kRows / 16), rg = q - plane * (S::kRows / 16);
    const int r = rg * 16 + (lane >> 2);
    kofs[i] = ((lane & 3) ^ dl_sw32(r)) * 8;
    dst[i] = (plane * S::kRows + rg * 16) * kDlBK;
    if (r < S::TM) {
      valid[i] = r < p.M;
      src[i] = p.A[plane] + (long)b * p.sA + (long)min(r, p.M - 1) * p.lda + kofs[i];
    } else {
      valid[i] = r - S::TM < p.N;
      src[i] = p.B[plane] + (long)b * p.sB + (long)min(r - S::TM, p.N - 1) * p.ldb + kofs[i];
    }
  }
  auto issue = [&](int t, int st) {
    const int k0 = k_begin + t * kDlBK;
#pragma unroll
    for (int i = 0; i < S::kPerWave; ++i) {
      const uint16_t* g = valid[i] && k0 + kofs[i] < k_end ? src[i] + k0 : g_dl_zero;
      __builtin_amdgcn_global_load_lds((const void*)g, MPQC_LDS(&lds[st][dst[i]]), 16, 0, 0);
    }
  };
  int frag[2];
#pragma unroll
  for (int ss = 0; ss < 2; ++ss)
    frag[ss] = (lane & 31) * kDlBK + (((2 * ss + (lane >> 5)) ^ dl_sw32(lane & 31)) * 8);

  f32x16_t acc[S::kMaxBlocksPerWave];
  double acc64[S::kMaxBlocksPerWave][16];
#pragma unroll
  for (int c = 0; c < S::kMaxBlocksPerWave; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      acc[c][r] = 0.f;
      acc64[c][r] = 0.0;
    }
  auto fold = [&]() {
#pragma unroll
    for (int c = 0; c < S::kMaxBlocksPerWave; ++c)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc64[c][r] += (double)acc[c][r];
        acc[c][r] = 0.f;
      }
  };
  auto products = [&](int st) {
#pragma unroll
    for (int c = 0; c < S::kMaxBlocksPerWave; ++c) {
      const int blk = wave + S::kWaves * c;
      if (blk < S::kBlocks) {  // (wave-uniform)
        const int i = blk / NB, j = blk - (blk / NB) * NB;
#pragma unroll
        for (int ss = 0; ss < 2; ++ss) {
          u16x8_t a[2], bb[2];
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            a[q] = *reinterpret_cast<const u16x8_t*>(&lds[st][(q * S::kRows + i * 32) * kDlBK + frag[ss]]);
            bb[q] = *reinterpret_cast<const u16x8_t*>(&lds[st][(q * S::kRows + S::TM + j * 32) * kDlBK + frag[ss]]);
          }
          mfma32<F16>(a[0], bb[1], acc[c]);
          mfma32<F16>(a[1], bb[0], acc[c]);
          mfma32<F16>(a[0], bb[0], acc[c]);
        }
      }
    }
  };

  if (nsteps > 0) {
#pragma unroll
    for (int t = 0; t < ST - 1; ++t)
      if (t < nsteps) issue(t, t);
    int since_fold = 0;
    for (int t = 0; t < nsteps; ++t) {
      // stage t has landed once at most min(ST - 2, nsteps - 1 - t) later steps' loads are out
      tk_wait_steps<S::kPerWave>(min(ST - 2, nsteps - 1 - t));
      dl_wait_lds();
      dl_barrier();
      // the stage of step t - 1 is free (every wave read it before this barrier)
      if (t + ST - 1 < nsteps) issue(t + ST - 1, (t + ST - 1) % ST);
      products(t % ST);
      if (++since_fold == kMaxChunk / kDlBK) {
        fold();
        since_fold = 0;
      }
    }
    fold();
  }
  const double unscale = p.alpha / ((*p.sa) * (*p.sb));
  double* __restrict__ W = reinterpret_cast<double*>(p.W) + (long)z * p.M * p.N;
#pragma unroll
  for (int c = 0; c < S::kMaxBlocksPerWave; ++c) {
    const int blk = wave + S::kWaves * c;
    if (blk >= S::kBlocks) continue;
    const int i = blk / NB, j = blk - (blk / NB) * NB;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int gm = i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), gn = j * 32 + (lane & 31);
      if (gm < p.M && gn < p.N) W[(long)gm * p.N + gn] = unscale * acc64[c][r];
    }
  }
}

// C[b] = beta C[b] + sum_s alpha / (sa sb) * W[b, s]  (FP32 partials, FP64 sums in split order)
__global__ __launch_bounds__(256) void splitk_reduce_f32_kernel(const float* __restrict__ W, double* __restrict__ C,
                                                                long ldc, long sC, int M, int N, int nsplit,
                                                                int batch0, double alpha, const double* sa,
                                                                const double* sb, double beta) {
  const long MN = (long)M * N;
  const int bz = blockIdx.y;
  const double u = alpha / ((*sa) * (*sb));
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < MN; e += (long)gridDim.x * blockDim.x) {
    const float* wp = W + (long)bz * nsplit * MN + e;
    double acc = 0.0;
    for (int s2 = 0; s2 < nsplit; ++s2) acc += u * (double)wp[(long)s2 * MN];
    const int m = (int)(e / N), n = (int)(e - (long)m * N);
    double* c = C + (long)(batch0 + bz) * sC + (long)m * ldc + n;
    *c = (beta == 0.0) ? acc : acc + beta * (*c);
  }
}

// The same sums four outputs per thread: 16-byte partial loads and FP64 stores, 32-bit
// indices (N % 4 == 0, C rows 16-byte aligned, M N / 4 < 2^31; the host checks).
// With part (one batch), also max |C| of the block's outputs into part[blockIdx.x].
template <bool MAX>
__global__ __launch_bounds__(256) void splitk_reduce_f32_vec_kernel(const float* __restrict__ W,
                                                                    double* __restrict__ C, long ldc, long sC,
                                                                    int M, int N, int nsplit, int batch0,
                                                                    double alpha, const double* sa,
                                                                    const double* sb, double beta,
                                                                    double* __restrict__ part) {
  double cm = 0.0;
  const long MN = (long)M * N;
  const int n4 = N >> 2, total = (int)(MN >> 2);
  const int bz = blockIdx.y;
  const double u = alpha / ((*sa) * (*sb));
  const float* __restrict__ wz = W + (long)bz * nsplit * MN;
  double* __restrict__ cz = C + (long)(batch0 + bz) * sC;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s2 = 0; s2 < nsplit; ++s2) {
      const f32x4_t w = *reinterpret_cast<const f32x4_t*>(wz + (long)s2 * MN + 4 * (long)e);
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[k] += u * (double)w[k];
    }
    const int m = e / n4, c4 = e - m * n4;
    f64x2_t* c = reinterpret_cast<f64x2_t*>(cz + (long)m * ldc + 4 * c4);
    f64x2_t lo = {acc[0], acc[1]}, hi = {acc[2], acc[3]};
    if (beta != 0.0) {
      const f64x2_t c0 = c[0], c1 = c[1];
      lo = f64x2_t{acc[0] + beta * c0[0], acc[1] + beta * c0[1]};
      hi = f64x2_t{acc[2] + beta * c1[0], acc[3] + beta * c1[1]};
    }
    c[0] = lo;
    c[1] = hi;
    if (MAX) cm = fmax(cm, fmax(fmax(fabs(lo[0]), fabs(lo[1])), fmax(fabs(hi[0]), fabs(hi[1]))));
  }
  if (MAX) {
    __shared__ double sh[256];
    sh[threadIdx.x] = cm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
  }
}

// *out = max of part[0..n)
__global__ __launch_bounds__(256) void max_stage2(const double* __restrict__ part, int n, double* __restrict__ out) {
  double m = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) m = fmax(m, part[b]);
  __shared__ double sh[256];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// Launch plan: tile edge, split-K factor and K chunk of one contraction.
struct PlanePlan {
  int bm, bn;  // tile: 64 x 64, 128 x 128, 256 x 128 or 256 x 256
  int nsplit;  // K chunks (each summed in FP32, chunks summed in FP64)
  int kchunk;
};

int g_tile_override_m = 0, g_tile_override_n = 0;
bool g_bdirect = false;  // direct B fragments (selected by measurement)
int g_tile_order = 0;    // see PParams::order
int g_lds_stages = 2;    // LDS stages of the 128 x 128 two-plane kernel
int g_split_cap = 512;   // most K splits of a grid whose tiles alone do not fill the GPU
int g_dl_mode = 0;       // deep-pipeline kernel: 0 where the planner picks it, 1 never

// Plan of the deep-pipeline kernel: used for two-plane contractions whose 256 x 256
// tiles pad the output by at most 50 % (measured: the o(o+-1)/2 x o^2 pair-symmetric
// W_oooo halves at (H2O)8, padded 40-47 %, run in 0.13 ms against 0.18 ms on the 128 x 128
// kernel).  The K split count c (chunks of at most
// kMaxChunk, multiples of 32) minimises a time model: rounds of one workgroup per CU
// (256 CUs) times the chunk's 32-wide steps at ~3 us each (measured), plus the FP32
// partials' write and read (8 bytes per output and split at ~5 TB/s).
struct DlPlan {
  bool use = false;
  int nsplit = 1, kchunk = 0;
};

DlPlan plan_dl(int M, int N, int K, int batch) {
  DlPlan pl;
  if (g_dl_mode == 1 || M < kDlTile || N < kDlTile || K < 2 * kDlBK) return pl;
  const long tiles = (long)ceil_div(M, kDlTile) * ceil_div(N, kDlTile);
  if ((double)tiles * kDlTile * kDlTile > 1.5 * (double)M * N) return pl;
  const long work = tiles * batch;
  const int n_lo = (int)ceil_div(K, kMaxChunk);
  const int n_hi = std::max(n_lo, std::min(std::min(K / 128, n_lo + 64), g_split_cap));
  double best = 0.0;
  int pick = n_lo;
  for (int c = n_lo; c <= n_hi; ++c) {
    const long chunk = ceil_div(ceil_div(K, c), kDlBK) * kDlBK;
    const int splits = (int)ceil_div(K, chunk);
    if (splits != c) continue;
    const double t = (double)ceil_div(work * c, 256) * (double)(chunk / kDlBK) * 3.0 +
                     (c > 1 ? (double)c * M * N * batch * 8.0 / 5.0e6 : 0.0);
    if (c == n_lo || t < best) {
      best = t;
      pick = c;
    }
  }
  pl.kchunk = (int)(ceil_div(ceil_div(K, pick), kDlBK) * kDlBK);
  pl.nsplit = (int)ceil_div(K, pl.kchunk);
  pl.use = true;
  return pl;
}

size_t dl_workspace_bytes(int M, int N, const DlPlan& pl, int batch) {
  return pl.use && pl.nsplit > 1 ? (size_t)pl.nsplit * M * N * batch * sizeof(float) : 0;
}

template <bool F16>
bool launch_planes_dl(const PlaneGemmDesc& d, const DlPlan& pl, double* ws, size_t ws_bytes, hipStream_t stream) {
  bool cmax_done = false;
  const int tiles_m = (int)ceil_div(d.M, kDlTile), tiles_n = (int)ceil_div(d.N, kDlTile);
  const long tiles = (long)tiles_m * tiles_n;
  const int per_xcd = tiles >= kXcdMinTiles ? (int)ceil_div(tiles, kXcd) : 0;  // 0: plain tile order
  const size_t slice = (size_t)d.M * d.N * sizeof(float);
  int max_batch = 65535 / pl.nsplit;
  if (pl.nsplit > 1) max_batch = (int)std::min<size_t>(max_batch, ws_bytes / (slice * pl.nsplit));
  if (max_batch < 1) throw std::runtime_error("gemm_planes: split-K workspace too small");
  for (int b0 = 0; b0 < d.batch; b0 += max_batch) {
    const int nb = std::min(max_batch, d.batch - b0);
    DlParams p;
    for (int q = 0; q < 2; ++q) {
      p.A[q] = d.A.p[q];
      p.B[q] = d.B.p[q];
    }
    p.sa = d.A.scale; p.sb = d.B.scale; p.C = d.C; p.W = reinterpret_cast<float*>(ws);
    p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc; p.sA = d.sA; p.sB = d.sB; p.sC = d.sC;
    p.M = d.M; p.N = d.N; p.K = d.K; p.kchunk = pl.kchunk; p.nsplit = pl.nsplit; p.batch0 = b0;
    p.tiles_m = tiles_m; p.tiles_n = tiles_n; p.per_xcd = per_xcd;
    p.alpha = d.alpha; p.beta = d.beta;
    dim3 grid((unsigned)(per_xcd ? per_xcd * kXcd : tiles), (unsigned)(nb * pl.nsplit));
    hipLaunchKernelGGL((gemm_planes_dl_kernel<F16>), grid, dim3(kDlThreads), 0, stream, p);
    check_launch("gemm_planes_dl");
    if (pl.nsplit > 1) {
      const long MN = (long)d.M * d.N;
      const bool vec = d.N % 4 == 0 && d.ldc % 2 == 0 && (d.batch == 1 || d.sC % 2 == 0) &&
                       reinterpret_cast<uintptr_t>(d.C) % 16 == 0 && MN / 4 < (1L << 31) - 65536L * 256;
      if (vec) {
        dim3 g2((unsigned)std::min<long>(ceil_div(MN / 4, 256), 8192), (unsigned)nb);
        if (d.cmax && d.cmax_partials && d.batch == 1) {
          hipLaunchKernelGGL(splitk_reduce_f32_vec_kernel<true>, g2, dim3(256), 0, stream, p.W, d.C, d.ldc, d.sC,
                             d.M, d.N, pl.nsplit, b0, d.alpha, d.A.scale, d.B.scale, d.beta, d.cmax_partials);
          hipLaunchKernelGGL(max_stage2, dim3(1), dim3(256), 0, stream, d.cmax_partials, (int)g2.x, d.cmax);
          cmax_done = true;
        } else {
          hipLaunchKernelGGL(splitk_reduce_f32_vec_kernel<false>, g2, dim3(256), 0, stream, p.W, d.C, d.ldc, d.sC,
                             d.M, d.N, pl.nsplit, b0, d.alpha, d.A.scale, d.B.scale, d.beta, nullptr);
        }
      } else {
        dim3 g2((unsigned)std::min<long>(ceil_div(MN, 256), 4096), (unsigned)nb);
        hipLaunchKernelGGL(splitk_reduce_f32_kernel, g2, dim3(256), 0, stream, p.W, d.C, d.ldc, d.sC, d.M, d.N,
                           pl.nsplit, b0, d.alpha, d.A.scale, d.B.scale, d.beta);
      }
      check_launch("splitk_reduce_f32");
    }
  }
  return cmax_done;
}

PlanePlan plan_planes(int M, int N, int K, int batch, int force_bm = 0, int force_bn = 0) {
  PlanePlan pl;
  // 128-wide tiles unless 64-wide ones pad the output much less
  pl.bm = (M <= 64 || N <= 64) ? 64 : 128;
  if (pl.bm == 128) {
    const double a128 = (double)ceil_div(M, 128) * ceil_div(N, 128) * 128 * 128;
    const double a64 = (double)ceil_div(M, 64) * ceil_div(N, 64) * 64 * 64;
    if (a128 > 1.25 * a64) pl.bm = 64;
  }
  pl.bn = pl.bm;
  if (g_tile_override_m) {
    pl.bm = g_tile_override_m;
    pl.bn = g_tile_override_n;
  }
  if (force_bm) {
    pl.bm = force_bm;
    pl.bn = force_bn;
  }
  const long tiles = (long)ceil_div(M, pl.bm) * ceil_div(N, pl.bn);
  int nsplit = 1;
  const long want = 1024;  // about 4 workgroups per CU
  if (tiles * batch < want && K >= 8 * kBK) {
    nsplit = (int)std::min<long>(ceil_div(want, tiles * batch), K / (4 * kBK));
    nsplit = std::max(1, std::min(nsplit, g_split_cap));
  }
  // K chunks beyond kMaxChunk go to separate splits as well: streaming partial
  // sums to the workspace measured faster than the in-kernel FP64 flush, whose
  // read-modify-write of the C tile stalls the K loop
  nsplit = std::max<int>(nsplit, (int)ceil_div(K, kMaxChunk));
  pl.kchunk = (int)(ceil_div(ceil_div(K, nsplit), kBK) * kBK);
  pl.nsplit = nsplit > 1 ? (int)ceil_div(K, pl.kchunk) : 1;
  return pl;
}

template <int BM, int BN, int NP, bool F16>
void launch_planes(const PlaneGemmDesc& d, const PlanePlan& pl, double* ws, size_t ws_bytes,
                   hipStream_t stream) {
  const int tiles_m = (int)ceil_div(d.M, BM), tiles_n = (int)ceil_div(d.N, BN);
  const long tiles = (long)tiles_m * tiles_n;
  const int nsplit = pl.nsplit, kchunk = pl.kchunk;
  const size_t slice = (size_t)d.M * d.N * sizeof(double);
  int max_batch = 65535 / nsplit;
  if (nsplit > 1) {
    const size_t cap = ws_bytes / (slice * nsplit);
    if (cap == 0) throw std::runtime_error("gemm_planes: split-K workspace too small");
    max_batch = (int)std::min<size_t>(max_batch, cap);
  }
  const bool vec = d.K % 8 == 0 && d.lda % 8 == 0 && d.ldb % 8 == 0 && d.sA % 8 == 0 &&
                   d.sB % 8 == 0;
  const int per_xcd = tiles >= kXcdMinTiles ? (int)ceil_div(tiles, kXcd) : 0;  // 0: plain tile order
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
    p.tiles_m = tiles_m; p.tiles_n = tiles_n; p.per_xcd = per_xcd; p.order = g_tile_order;
    p.alpha = d.alpha; p.beta = d.beta;
    dim3 grid((unsigned)(per_xcd ? per_xcd * kXcd : tiles), (unsigned)(nb * nsplit));
    if (vec && NP == 2 && BM == 128 && BN == 128 && g_lds_stages == 1 && !g_bdirect)
      hipLaunchKernelGGL((gemm_planes_kernel<BM, BN, NP, F16, true, false, 1>), grid, dim3(kThr), 0, stream, p);
    else if (vec && NP == 2 && g_bdirect)
      hipLaunchKernelGGL((gemm_planes_kernel<BM, BN, NP, F16, true, NP == 2>), grid, dim3(kThr), 0, stream, p);
    else if (vec)
      hipLaunchKernelGGL((gemm_planes_kernel<BM, BN, NP, F16, true, false>), grid, dim3(kThr), 0, stream, p);
    else
      hipLaunchKernelGGL((gemm_planes_kernel<BM, BN, NP, F16, false, false>), grid, dim3(kThr), 0, stream, p);
    check_launch("gemm_planes");
    if (nsplit > 1) splitk_reduce(ws, d.C, d.ldc, d.sC, d.M, d.N, nsplit, b0, nb, d.beta, stream);
  }
}

// Plan of the tall-K kernel: two-plane contractions with M, N <= 160 and K >= 8192;
// output blocks padded to 2 or 5 per dimension; one round of workgroups (as many as
// the LDS lets run at once on 256 CUs), each streaming an equal range of K.
struct TkPlan {
  bool use = false;
  int mb = 0, nb = 0, nsplit = 1, kchunk = 0;
};

template <int MB, int NB>
int tk_occupancy() {
  return (int)std::max<long>(1, (160L * 1024) / ((long)TkShape<MB, NB>::kStages * TkShape<MB, NB>::kStageHalves * 2));
}

TkPlan plan_tallk(int M, int N, int K, int batch) {
  TkPlan pl;
  if (g_dl_mode == 1 || M > 160 || N > 160 || K < 8192) return pl;
  pl.mb = M <= 64 ? 2 : 5;
  pl.nb = N <= 64 ? 2 : 5;
  const int occ = pl.mb == 2 && pl.nb == 2 ? tk_occupancy<2, 2>() : pl.mb == 5 && pl.nb == 5 ? tk_occupancy<5, 5>()
                                                                                              : tk_occupancy<5, 2>();
  const long want = std::max<long>(1, 256L * occ / std::max(1, batch));
  pl.kchunk = (int)(ceil_div(ceil_div(K, want), kDlBK) * kDlBK);
  pl.nsplit = (int)ceil_div(K, pl.kchunk);
  pl.use = true;
  return pl;
}

size_t tk_workspace_bytes(int M, int N, const TkPlan& pl, int batch) {
  return pl.use ? (size_t)pl.nsplit * M * N * batch * sizeof(double) : 0;
}

template <int MB, int NB, bool F16>
void launch_planes_tallk(const PlaneGemmDesc& d, const TkPlan& pl, double* ws, size_t ws_bytes, hipStream_t stream) {
  const size_t slice = (size_t)d.M * d.N * sizeof(double);
  int max_batch = std::min<int>(65535, (int)std::min<size_t>(INT32_MAX / pl.nsplit, ws_bytes / (slice * pl.nsplit)));
  if (max_batch < 1) throw std::runtime_error("gemm_planes: tall-K workspace too small");
  for (int b0 = 0; b0 < d.batch; b0 += max_batch) {
    const int nb = std::min(max_batch, d.batch - b0);
    DlParams p;
    for (int q = 0; q < 2; ++q) {
      p.A[q] = d.A.p[q];
      p.B[q] = d.B.p[q];
    }
    p.sa = d.A.scale; p.sb = d.B.scale; p.C = d.C; p.W = reinterpret_cast<float*>(ws);
    p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc; p.sA = d.sA; p.sB = d.sB; p.sC = d.sC;
    p.M = d.M; p.N = d.N; p.K = d.K; p.kchunk = pl.kchunk; p.nsplit = pl.nsplit; p.batch0 = b0;
    p.tiles_m = 1; p.tiles_n = 1; p.per_xcd = 0;
    p.alpha = d.alpha; p.beta = d.beta;
    hipLaunchKernelGGL((gemm_planes_tallk_kernel<MB, NB, F16>), dim3((unsigned)(nb * pl.nsplit)),
                       dim3(64 * TkShape<MB, NB>::kWaves), 0, stream, p);
    check_launch("gemm_planes_tallk");
    splitk_reduce(ws, d.C, d.ldc, d.sC, d.M, d.N, pl.nsplit, b0, nb, d.beta, stream);
  }
}

template <bool F16>
void dispatch_tallk(const PlaneGemmDesc& d, const TkPlan& pl, double* ws, size_t ws_bytes, hipStream_t stream) {
  if (pl.mb == 2 && pl.nb == 2) launch_planes_tallk<2, 2, F16>(d, pl, ws, ws_bytes, stream);
  else if (pl.mb == 2) launch_planes_tallk<2, 5, F16>(d, pl, ws, ws_bytes, stream);
  else if (pl.nb == 2) launch_planes_tallk<5, 2, F16>(d, pl, ws, ws_bytes, stream);
  else launch_planes_tallk<5, 5, F16>(d, pl, ws, ws_bytes, stream);
}

// Returns whether *d.cmax was written (only the deep-pipeline kernel's split-K pass does).
template <int NP, bool F16>
bool dispatch_planes(const PlaneGemmDesc& d, double* ws, size_t ws_bytes, hipStream_t stream) {
  if constexpr (NP == 2) {
    const bool aligned16 = d.K % 8 == 0 && d.lda % 8 == 0 && d.ldb % 8 == 0 && d.sA % 8 == 0 && d.sB % 8 == 0 &&
                           !g_tile_override_m;
    const DlPlan dl = plan_dl(d.M, d.N, d.K, d.batch);
    if (aligned16 && dl.use && ws_bytes >= dl_workspace_bytes(d.M, d.N, dl, 1))
      return launch_planes_dl<F16>(d, dl, ws, ws_bytes, stream);
    const TkPlan tk = plan_tallk(d.M, d.N, d.K, d.batch);
    if (aligned16 && tk.use && ws_bytes >= tk_workspace_bytes(d.M, d.N, tk, 1)) {
      dispatch_tallk<F16>(d, tk, ws, ws_bytes, stream);
      return false;
    }
  }
  const PlanePlan pl = plan_planes(d.M, d.N, d.K, d.batch);
  if (pl.bm == 64) {
    launch_planes<64, 64, NP, F16>(d, pl, ws, ws_bytes, stream);
  } else if (pl.bm == 256 && pl.bn == 256 && NP == 2) {
    if constexpr (NP == 2) launch_planes<256, 256, NP, F16>(d, pl, ws, ws_bytes, stream);
  } else if (pl.bm == 128 || NP == 3) {  // three planes of a 256-row tile exceed the LDS
    launch_planes<128, 128, NP, F16>(d, pl.bm == 128 ? pl : plan_planes(d.M, d.N, d.K, d.batch, 128, 128),
                                     ws, ws_bytes, stream);
  } else {
    launch_planes<(NP == 3 ? 128 : 256), 128, NP, F16>(d, pl, ws, ws_bytes, stream);
  }
  return false;
}

}  // namespace

void split_to_planes(PlaneKind kind, const Planes& out, const long n[4], const TTerm* terms,
                     int nterms, double* partials, hipStream_t stream, const TProd* prod,
                     AbsMaxCache* cache) {
  if (nterms < 1 || nterms > kMaxTerms) throw std::invalid_argument("split_to_planes: bad term count");
  const long total = n[0] * n[1] * n[2] * n[3];
  if (total > out.n) throw std::invalid_argument("split_to_planes: planes too small");
  if (total == 0) return;
  // scale: per distinct input, max |x| (fp16 planes only)
  ScaleArgs sa{};
  sa.nterms = nterms;
  sa.kind = (int)kind;
  sa.prod_x = sa.prod_y = -1;
  AbsMaxArgs am{};
  auto source = [&](const double* p, long vol) {
    for (int u = 0; u < sa.nsrc; ++u)
      if (am.p[u] == p) return u;
    if (vol <= 0) throw std::invalid_argument("split_to_planes: input volume not set");
    const int u = sa.nsrc++;
    am.p[u] = p;
    am.n[u] = vol;
    return u;
  };
  for (int t = 0; t < nterms; ++t) {
    sa.src_of[t] = source(terms[t].p, terms[t].vol);
    sa.absalpha[t] = std::fabs(terms[t].alpha);
  }
  if (prod) {
    sa.prod_x = source(prod->x, prod->nx);
    sa.prod_y = source(prod->y, prod->ny);
    sa.prod_absalpha = std::fabs(prod->alpha);
  }
  if (kind == PlaneKind::F16x2) {
    // inputs whose max is cached skip the max pass (their stage-1 blocks see n = 0)
    if (cache && cache->dev) {
      for (int u = 0; u < sa.nsrc; ++u)
        for (int k = 0; k < AbsMaxCache::kSlots; ++k)
          if (cache->ptr[k] == am.p[u]) {
            if (cache->valid[k]) {
              sa.cached[u] = cache->dev + k;
              am.n[u] = 0;
            } else {
              sa.record[u] = cache->dev + k;
              cache->valid[k] = true;
            }
          }
    }
    long nmax = 0;
    for (int u = 0; u < sa.nsrc; ++u) nmax = std::max(nmax, am.n[u]);
    sa.nb = (int)std::max<long>(1, std::min<long>(1024, ceil_div(nmax, 256)));
    hipLaunchKernelGGL(absmax_stage1, dim3(sa.nb, sa.nsrc), dim3(256), 0, stream, am, partials);
  } else {
    sa.nsrc = 0;
    sa.nb = 0;
  }
  hipLaunchKernelGGL(scale_stage2, dim3(1), dim3(256), 0, stream, partials, sa, out.scale);
  const PermGeom g = make_perm_geom(n, terms, nterms, prod);
  SplitOut so{{out.p[0], out.p[1], out.p[2]}, out.scale, (int)kind};
  if (g.tiled)
    if (perm_fast_tile(g) == 3)
      hipLaunchKernelGGL(split_write_kernel<3>, perm_grid(g), dim3(kPThreads), perm_lds_bytes(g), stream, g, so);
    else if (perm_fast_tile(g) == 4)
      hipLaunchKernelGGL(split_write_kernel<4>, perm_grid(g), dim3(kPThreads), perm_lds_bytes(g), stream, g, so);
    else
      hipLaunchKernelGGL(split_write_kernel<1>, perm_grid(g), dim3(kPThreads), perm_lds_bytes(g), stream, g, so);
  else
    hipLaunchKernelGGL(split_write_kernel<0>, perm_grid(g), dim3(kPThreads), 0, stream, g, so);
  check_launch("split_to_planes");
}

void set_plane_bdirect(bool on) { g_bdirect = on; }
void set_plane_lds_stages(int stages) {
  if (stages != 1 && stages != 2) throw std::invalid_argument("plane LDS stages must be 1 or 2");
  g_lds_stages = stages;
}
void set_plane_split_cap(int cap) {
  if (cap < 1 || cap > 4096) throw std::invalid_argument("plane_split_cap must be in [1, 4096]");
  g_split_cap = cap;
}

void set_plane_tile_order(int order) {
  if (order < 0 || order > 2) throw std::invalid_argument("unknown tile order");
  g_tile_order = order;
}

void set_plane_tile_override(int bm, int bn) {
  if (!((bm == 0 && bn == 0) || (bm == 64 && bn == 64) || (bm == 128 && bn == 128) ||
        (bm == 256 && bn == 128) || (bm == 256 && bn == 256)))
    throw std::invalid_argument("unsupported plane GEMM tile");
  g_tile_override_m = bm;
  g_tile_override_n = bn;
}

size_t gemm_planes_workspace_bytes(int M, int N, int K, int batch) {
  const PlanePlan pl = plan_planes(M, N, K, batch);
  const size_t old = pl.nsplit > 1 ? (size_t)pl.nsplit * M * N * batch * sizeof(double) : 0;
  return std::max({old, dl_workspace_bytes(M, N, plan_dl(M, N, K, batch), batch),
                   tk_workspace_bytes(M, N, plan_tallk(M, N, K, batch), batch)});
}

void set_plane_dl_mode(int mode) {
  if (mode != 0 && mode != 1) throw std::invalid_argument("plane_dl must be 0 (planner) or 1 (off)");
  g_dl_mode = mode;
}

bool gemm_planes(const PlaneGemmDesc& d, double* ws, size_t ws_bytes, hipStream_t stream) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return false;
  if (d.K <= 0) throw std::invalid_argument("gemm_planes: K must be positive");
  switch (d.kind) {
    case PlaneKind::BF16x2: return dispatch_planes<2, false>(d, ws, ws_bytes, stream);
    case PlaneKind::F16x2: return dispatch_planes<2, true>(d, ws, ws_bytes, stream);
    case PlaneKind::BF16x3: return dispatch_planes<3, false>(d, ws, ws_bytes, stream);
  }
  return false;
}

}  // namespace mpqc_hip
