// HBM-streaming kernels of the CCSD iteration.  All are bandwidth bound; they
// read/write each element once with unit-stride writes and wave-contiguous
// reads along the innermost output index.
#include "tensor_ops.hpp"
#include "permute.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mpqc_hip {

namespace {

constexpr int kBlock = 256;

template <int SHAPE>
__global__ __launch_bounds__(kPThreads) void tadd_kernel(PermGeom g, double* __restrict__ out,
                                                       double beta) {
  permute_block<SHAPE>(g, blockIdx.x, blockIdx.y, blockIdx.z,
                       [&](long o, double v) { out[o] = (beta == 0.0) ? v : beta * out[o] + v; });
}

// Element-wise kernels use 1-D grid-stride loops with 32-bit indices (the
// host checks v^2 o^2 < 2^31).
__global__ __launch_bounds__(kBlock) void build_tau_kernel(double* __restrict__ tau,
                                                           const double* __restrict__ t2,
                                                           const double* __restrict__ t1, int o,
                                                           int v) {
  // tau(a,b,i,j) = t2 + t1(a,i) t1(b,j)
  const int oo = o * o, n = v * v * oo;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) {
    const int ab = e / oo, ij = e - ab * oo;
    const int a = ab / v, b = ab - a * v;
    const int i = ij / o, j = ij - i * o;
    tau[e] = t2[e] + t1[a * o + i] * t1[b * o + j];
  }
}

// tau as above, and per block max |t2| and max |tau| -> part[2 blk], part[2 blk + 1]
__global__ __launch_bounds__(kBlock) void build_tau_absmax_kernel(double* __restrict__ tau,
                                                                  const double* __restrict__ t2,
                                                                  const double* __restrict__ t1, int o, int v,
                                                                  double* __restrict__ part) {
  __shared__ double s2[kBlock], st[kBlock];
  const int oo = o * o, n = v * v * oo;
  double m2 = 0.0, mt = 0.0;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) {
    const int ab = e / oo, ij = e - ab * oo;
    const int a = ab / v, b = ab - a * v;
    const int i = ij / o, j = ij - i * o;
    const double x = t2[e];
    const double y = x + t1[a * o + i] * t1[b * o + j];
    tau[e] = y;
    m2 = fmax(m2, fabs(x));
    mt = fmax(mt, fabs(y));
  }
  s2[threadIdx.x] = m2;
  st[threadIdx.x] = mt;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
      st[threadIdx.x] = fmax(st[threadIdx.x], st[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s2[0];
    part[2 * blockIdx.x + 1] = st[0];
  }
}

__global__ __launch_bounds__(kBlock) void absmax2_stage2(const double* __restrict__ part, int nb,
                                                         double* __restrict__ out) {
  __shared__ double s2[kBlock], st[kBlock];
  double m2 = 0.0, mt = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) {
    m2 = fmax(m2, part[2 * b]);
    mt = fmax(mt, part[2 * b + 1]);
  }
  s2[threadIdx.x] = m2;
  st[threadIdx.x] = mt;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
      st[threadIdx.x] = fmax(st[threadIdx.x], st[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s2[0];
    out[1] = st[0];
  }
}

__global__ __launch_bounds__(kBlock) void build_TP_kernel(double* __restrict__ TP,
                                                          const double* __restrict__ t2,
                                                          const double* __restrict__ t1, int o,
                                                          int v) {
  // TP(a,i,l,d) = 0.5 t2(d,a,i,l) + t1(d,i) t1(a,l)
  const int ov = o * v, n = ov * ov;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) {
    const int ai = e / ov, ld = e - ai * ov;
    const int a = ai / o, i = ai - a * o;
    const int l = ld / v, d = ld - l * v;
    const int src = ((d * v + a) * o + i) * o + l;
    TP[e] = 0.5 * t2[src] + t1[d * o + i] * t1[a * o + l];
  }
}

__global__ __launch_bounds__(kBlock) void jacobi_kernel(double* __restrict__ t1, double* __restrict__ t2,
                                                        const double* __restrict__ r1,
                                                        const double* __restrict__ r2,
                                                        const double* __restrict__ eo,
                                                        const double* __restrict__ ev, int o, int v,
                                                        double* __restrict__ c1, double* __restrict__ c2) {
  const int oo = o * o, n = v * v * oo;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) {
    const int ab = e / oo, ij = e - ab * oo;
    const int a = ab / v, b = ab - a * v;
    const int i = ij / o, j = ij - i * o;
    const double x = t2[e] + r2[e] / (eo[i] + eo[j] - ev[a] - ev[b]);
    t2[e] = x;
    if (c2) c2[e] = x;
  }
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < v * o; e += gridDim.x * kBlock) {
    const int a = e / o, i = e - a * o;
    const double x = t1[e] + r1[e] / (eo[i] - ev[a]);
    t1[e] = x;
    if (c1) c1[e] = x;
  }
}

// ---------------------------------------------------------------- reductions
constexpr int kRedGrid = 1024;

__device__ inline double block_sum(double x) {
  __shared__ double sh[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = x;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < kBlock / 64; ++k) r += sh[k];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

__global__ __launch_bounds__(kBlock) void dot_stage1(const double* __restrict__ x,
                                                     const double* __restrict__ y, long n,
                                                     double* __restrict__ partials) {
  double acc = 0.0;
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock)
    acc += x[e] * y[e];
  const double r = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// out[k] = sum_b partials[k*nb + b]  (one block per k)
__global__ __launch_bounds__(kBlock) void reduce_stage2(const double* __restrict__ partials, int nb,
                                                        double* __restrict__ out, double scale) {
  double acc = 0.0;
  const double* p = partials + (long)blockIdx.x * nb;
  for (int b = threadIdx.x; b < nb; b += kBlock) acc += p[b];
  const double r = block_sum(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = scale * r;
}

__global__ __launch_bounds__(kBlock) void energy_stage1(
    const double* __restrict__ f_ai, const double* __restrict__ r1, const double* __restrict__ t1,
    const double* __restrict__ g, const double* __restrict__ r2, const double* __restrict__ tau,
    int o, int v, double* __restrict__ partials) {
  // sizes checked on the host: v*v*o*o < 2^31
  const int n2 = v * v * o * o;
  const int oo = o * o;
  double acc = 0.0;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < n2; e += gridDim.x * kBlock) {
    const int ab = e / oo, ij = e - ab * oo;
    const int a = ab / v, b = ab - a * v;
    const double tba = tau[(b * v + a) * oo + ij];
    acc += (g[e] + r2[e]) * (2.0 * tau[e] - tba);
  }
  double acc1 = 0.0;
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < (long)v * o; e += (long)gridDim.x * kBlock)
    acc1 += (f_ai[e] + r1[e]) * t1[e];
  const double r = block_sum(acc + 2.0 * acc1);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// The same energy partials (same per-thread expression and grid, so the same bits) with
// |r1|^2 and |r2|^2 from the same reads: partials[k * gridDim.x + block], k = 0 |r1|^2,
// 1 |r2|^2, 2 energy.
__global__ __launch_bounds__(kBlock) void energy_norms_stage1(
    const double* __restrict__ f_ai, const double* __restrict__ r1, const double* __restrict__ t1,
    const double* __restrict__ g, const double* __restrict__ r2, const double* __restrict__ tau,
    int o, int v, double* __restrict__ partials) {
  const int n2 = v * v * o * o;
  const int oo = o * o;
  double acc = 0.0, q2 = 0.0;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < n2; e += gridDim.x * kBlock) {
    const int ab = e / oo, ij = e - ab * oo;
    const int a = ab / v, b = ab - a * v;
    const double tba = tau[(b * v + a) * oo + ij];
    const double r = r2[e];
    acc += (g[e] + r) * (2.0 * tau[e] - tba);
    q2 += r * r;
  }
  double acc1 = 0.0, q1 = 0.0;
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < (long)v * o; e += (long)gridDim.x * kBlock) {
    const double r = r1[e];
    acc1 += (f_ai[e] + r) * t1[e];
    q1 += r * r;
  }
  const double s1 = block_sum(q1), s2 = block_sum(q2), se = block_sum(acc + 2.0 * acc1);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = s1;
    partials[gridDim.x + blockIdx.x] = s2;
    partials[2 * gridDim.x + blockIdx.x] = se;
  }
}

struct MultiDotArgs {
  const double* x1[16];
  const double* x2[16];
  const double* y1;
  const double* y2;
  double* c1;  // copies of y (optional)
  double* c2;
  long n1, n2;
  int nvec;
};

// partials[k*gridDim.x + block] = partial <x_k, y>
__global__ __launch_bounds__(kBlock) void multidot_stage1(MultiDotArgs a, double* __restrict__ partials) {
  double acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.0;
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < a.n2; e += (long)gridDim.x * kBlock) {
    const double y = a.y2[e];
    if (a.c2) a.c2[e] = y;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k < a.nvec) acc[k] += a.x2[k][e] * y;
  }
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < a.n1; e += (long)gridDim.x * kBlock) {
    const double y = a.y1[e];
    if (a.c1) a.c1[e] = y;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k < a.nvec) acc[k] += a.x1[k][e] * y;
  }
  for (int k = 0; k < a.nvec; ++k) {
    const double r = block_sum(acc[k]);
    if (threadIdx.x == 0) partials[(long)k * gridDim.x + blockIdx.x] = r;
  }
}

// The same with the long part (x2, y2) read two elements at a time (16-byte loads;
// every x2[k], y2 and c2 16-byte aligned, n2 even; the host checks).
__global__ __launch_bounds__(kBlock) void multidot_vec_stage1(MultiDotArgs a, double* __restrict__ partials) {
  typedef double d2_t __attribute__((ext_vector_type(2)));
  double acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.0;
  const long n2h = a.n2 >> 1;
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n2h; e += (long)gridDim.x * kBlock) {
    const d2_t y = reinterpret_cast<const d2_t*>(a.y2)[e];
    if (a.c2) reinterpret_cast<d2_t*>(a.c2)[e] = y;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k < a.nvec) {
        const d2_t x = reinterpret_cast<const d2_t*>(a.x2[k])[e];
        acc[k] += x[0] * y[0];
        acc[k] += x[1] * y[1];
      }
  }
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < a.n1; e += (long)gridDim.x * kBlock) {
    const double y = a.y1[e];
    if (a.c1) a.c1[e] = y;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k < a.nvec) acc[k] += a.x1[k][e] * y;
  }
  for (int k = 0; k < a.nvec; ++k) {
    const double r = block_sum(acc[k]);
    if (threadIdx.x == 0) partials[(long)k * gridDim.x + blockIdx.x] = r;
  }
}

struct LincombArgs {
  const double* x[16];
  double c[16];
  int nvec;
};

__global__ __launch_bounds__(kBlock) void scale_matrix_kernel(double* __restrict__ C, long total, int M,
                                                               int N, long sC, long ldc, double beta) {
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < total; e += (long)gridDim.x * kBlock) {
    const long bm = e / N;
    const long n = e - bm * N;
    const long b = bm / M, m = bm - b * M;
    double* c = C + b * sC + m * ldc + n;
    *c = (beta == 0.0) ? 0.0 : beta * *c;
  }
}

__global__ __launch_bounds__(kBlock) void fill_kernel(double* __restrict__ x, long n, double value) {
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock) x[e] = value;
}

__global__ __launch_bounds__(kBlock) void lincomb_kernel(double* __restrict__ y, LincombArgs a, long n) {
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock) {
    double acc = 0.0;
    for (int k = 0; k < a.nvec; ++k) acc += a.c[k] * a.x[k][e];
    y[e] = acc;
  }
}

// t2 = sum_k c_k x_k (lincomb_kernel's sum) and tau = t2 + t1(a,i) t1(b,j) (build_tau's)
// in one pass; with part, also the per-block max |t2| and max |tau| (build_tau_absmax's).
template <bool MAXES>
__global__ __launch_bounds__(kBlock) void lincomb_tau_kernel(double* __restrict__ t2, double* __restrict__ tau,
                                                             const double* __restrict__ t1, LincombArgs a, int o,
                                                             int v, double* __restrict__ part) {
  const int oo = o * o, n = v * v * oo;
  double m2 = 0.0, mt = 0.0;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) {
    double x = 0.0;
    for (int k = 0; k < a.nvec; ++k) x += a.c[k] * a.x[k][e];
    const int ab = e / oo, ij = e - ab * oo;
    const int ai = ab / v, bi = ab - ai * v;
    const int i = ij / o, j = ij - i * o;
    const double y = x + t1[ai * o + i] * t1[bi * o + j];
    t2[e] = x;
    tau[e] = y;
    if (MAXES) {
      m2 = fmax(m2, fabs(x));
      mt = fmax(mt, fabs(y));
    }
  }
  if (MAXES) {
    __shared__ double s2[kBlock], st[kBlock];
    s2[threadIdx.x] = m2;
    st[threadIdx.x] = mt;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
        st[threadIdx.x] = fmax(st[threadIdx.x], st[threadIdx.x + w]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      part[2 * blockIdx.x] = s2[0];
      part[2 * blockIdx.x + 1] = st[0];
    }
  }
}

__global__ __launch_bounds__(kBlock) void nonfinite_stage1(const double* __restrict__ x, long n,
                                                           double* __restrict__ partials) {
  double acc = 0.0;
  for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock)
    acc += isfinite(x[e]) ? 0.0 : 1.0;
  const double r = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__device__ inline long tiled_offset(const long* __restrict__ mi, const long n[4], long i0, long i1,
                                    long i2, long i3) {
  const long* m0 = mi;
  const long* m1 = m0 + 2 * n[0];
  const long* m2 = m1 + 2 * n[1];
  const long* m3 = m2 + 2 * n[2];
  const long lo0 = m0[i0], e0 = m0[n[0] + i0];
  const long lo1 = m1[i1], e1 = m1[n[1] + i1];
  const long lo2 = m2[i2], e2 = m2[n[2] + i2];
  const long lo3 = m3[i3], e3 = m3[n[3] + i3];
  const long base = lo0 * n[1] * n[2] * n[3] + e0 * lo1 * n[2] * n[3] + e0 * e1 * lo2 * n[3] +
                    e0 * e1 * e2 * lo3;
  return base + (((i0 - lo0) * e1 + (i1 - lo1)) * e2 + (i2 - lo2)) * e3 + (i3 - lo3);
}

struct TileArgs {
  long n[4];
};

template <bool TO_DENSE>
__global__ __launch_bounds__(kBlock) void tiles_kernel(double* __restrict__ dst,
                                                       const double* __restrict__ src,
                                                       const long* __restrict__ mi, TileArgs a) {
  const long plane = a.n[2] * a.n[3];
  const long e = (long)blockIdx.x * kBlock + threadIdx.x;
  if (e >= plane) return;
  const long i0 = blockIdx.z, i1 = blockIdx.y;
  const long i2 = e / a.n[3], i3 = e - i2 * a.n[3];
  const long d = ((i0 * a.n[1] + i1) * a.n[2] + i2) * a.n[3] + i3;
  const long t = tiled_offset(mi, a.n, i0, i1, i2, i3);
  if (TO_DENSE) dst[d] = src[t];
  else dst[t] = src[d];
}

int red_grid(long n) { return (int)std::max<long>(1, std::min<long>(kRedGrid, ceil_div(n, kBlock))); }

void check_grid_yz(long y, long z) {
  if (y > 65535 || z > 65535) throw std::invalid_argument("tensor kernel: extent exceeds grid limit");
}

// ---------------------------------------------------------------- grouped
struct GPerm {
  PermGeom g;
  double* out;
  double beta;
  long first;          // first workgroup of this problem
  unsigned gx, gy;     // launch grid of the problem (gz implied)
};

struct GEw {
  int kind;
  double* out;
  const double* a;
  const double* b;
  long n;
  int o, v;
  long first;
};

constexpr int kEwPerBlock = kBlock * 8;  // elements per workgroup of the element-wise kinds

template <class P>
__device__ inline int find_first(const P* __restrict__ probs, int nprob, long it) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].first <= it) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kPThreads) void grouped_tadd_kernel(const GPerm* __restrict__ probs, int nprob) {
  const long it = (long)blockIdx.y * gridDim.x + blockIdx.x;
  const int q = find_first(probs, nprob, it);
  const GPerm& P = probs[q];
  const long local = it - P.first;
  const unsigned bx = (unsigned)(local % P.gx);
  const long rest = local / P.gx;
  const unsigned by = (unsigned)(rest % P.gy), bz = (unsigned)(rest / P.gy);
  if (bz >= (unsigned)P.g.nr0) return;  // grid padding past the last problem
  double* out = P.out;
  const double beta = P.beta;
  permute_block(P.g, bx, by, bz, [&](long o, double v) { out[o] = (beta == 0.0) ? v : beta * out[o] + v; });
}

__global__ __launch_bounds__(kBlock) void grouped_ew_kernel(const GEw* __restrict__ probs, int nprob) {
  const long it = (long)blockIdx.y * gridDim.x + blockIdx.x;
  const int q = find_first(probs, nprob, it);
  const GEw& P = probs[q];
  const long base = (it - P.first) * kEwPerBlock;
  for (int k = 0; k < kEwPerBlock / kBlock; ++k) {
    const long e = base + (long)k * kBlock + threadIdx.x;
    if (e >= P.n) return;
    if (P.kind == 0) {
      P.out[e] = P.a[e];
    } else if (P.kind == 1) {  // tau(a,b,i,j) = t2 + t1(a,i) t1(b,j)
      const int o = P.o, v = P.v, oo = o * o;
      const int ee = (int)e, ab = ee / oo, ij = ee - ab * oo;
      const int a = ab / v, b = ab - a * v, i = ij / o, j = ij - i * o;
      P.out[e] = P.a[e] + P.b[a * o + i] * P.b[b * o + j];
    } else {  // TP(a,i,l,d) = 0.5 t2(d,a,i,l) + t1(d,i) t1(a,l)
      const int o = P.o, v = P.v, ov = o * v;
      const int ee = (int)e, ai = ee / ov, ld = ee - ai * ov;
      const int a = ai / o, i = ai - a * o, l = ld / v, d = ld - l * v;
      P.out[e] = 0.5 * P.a[((d * v + a) * o + i) * o + l] + P.b[d * o + i] * P.b[a * o + l];
    }
  }
}

dim3 grouped_grid(long items) {
  if (items <= 0) throw std::invalid_argument("grouped launch without work");
  const long gx = std::min<long>(items, 1L << 20);
  const long gy = ceil_div(items, gx);
  if (gy > 65535) throw std::invalid_argument("grouped launch too large");
  return dim3((unsigned)gx, (unsigned)gy);
}

}  // namespace

GroupedPlan plan_grouped_tadd(const std::vector<GroupedTAddDesc>& probs, hipStream_t s) {
  if (probs.empty()) throw std::invalid_argument("grouped permutation without problems");
  std::vector<GPerm> host(probs.size());
  GroupedPlan pl;
  pl.kind = 0;
  pl.nprob = (int)probs.size();
  for (size_t q = 0; q < probs.size(); ++q) {
    const GroupedTAddDesc& d = probs[q];
    host[q].g = make_perm_geom(d.n, d.terms, d.nterms);
    const dim3 gr = perm_grid(host[q].g);
    host[q].out = d.out;
    host[q].beta = d.beta;
    host[q].first = pl.items;
    host[q].gx = gr.x;
    host[q].gy = gr.y;
    pl.items += (long)gr.x * gr.y * gr.z;
    pl.lds = std::max(pl.lds, perm_lds_bytes(host[q].g));
  }
  MPQC_HIP_CHECK(hipMalloc(&pl.dev, host.size() * sizeof(GPerm)));
  MPQC_HIP_CHECK(hipMemcpyAsync(pl.dev, host.data(), host.size() * sizeof(GPerm), hipMemcpyHostToDevice, s));
  MPQC_HIP_CHECK(hipStreamSynchronize(s));
  return pl;
}

GroupedPlan plan_grouped_ew(const std::vector<GroupedEwDesc>& probs, hipStream_t s) {
  if (probs.empty()) throw std::invalid_argument("grouped element-wise launch without problems");
  std::vector<GEw> host(probs.size());
  GroupedPlan pl;
  pl.kind = 1;
  pl.nprob = (int)probs.size();
  for (size_t q = 0; q < probs.size(); ++q) {
    const GroupedEwDesc& d = probs[q];
    if (d.kind < 0 || d.kind > 2) throw std::invalid_argument("unknown element-wise kind");
    if (d.kind != 0 && (long)d.v * d.v * d.o * d.o != d.n)
      throw std::invalid_argument("element-wise extents do not match");
    if (d.n >= (1L << 31)) throw std::invalid_argument("element-wise problem too large");
    host[q] = GEw{d.kind, d.out, d.a, d.b, d.n, d.o, d.v, pl.items};
    pl.items += ceil_div(d.n, kEwPerBlock);
  }
  MPQC_HIP_CHECK(hipMalloc(&pl.dev, host.size() * sizeof(GEw)));
  MPQC_HIP_CHECK(hipMemcpyAsync(pl.dev, host.data(), host.size() * sizeof(GEw), hipMemcpyHostToDevice, s));
  MPQC_HIP_CHECK(hipStreamSynchronize(s));
  return pl;
}

void launch_grouped(const GroupedPlan& pl, hipStream_t s) {
  if (pl.items == 0) return;
  const dim3 grid = grouped_grid(pl.items);
  if (pl.kind == 0)
    hipLaunchKernelGGL(grouped_tadd_kernel, grid, dim3(kPThreads), pl.lds, s, (const GPerm*)pl.dev, pl.nprob);
  else
    hipLaunchKernelGGL(grouped_ew_kernel, grid, dim3(kBlock), 0, s, (const GEw*)pl.dev, pl.nprob);
  check_launch("grouped");
}

void free_grouped(GroupedPlan& pl) {
  if (pl.dev) (void)hipFree(pl.dev);
  pl.dev = nullptr;
}

void permuted_strides(const long in_extents[4], const int perm[4], long out_s[4]) {
  long st[4];
  st[3] = 1;
  for (int d = 2; d >= 0; --d) st[d] = st[d + 1] * in_extents[d + 1];
  for (int d = 0; d < 4; ++d) out_s[d] = st[perm[d]];
}

namespace {
PermGeom make_perm_geom_dims(const long n[4], const TTerm* terms, int nterms, const TProd* prod);

// Adjacent output dims that every input also holds adjacent (stride of the outer =
// stride of the inner times the inner extent; dims of extent 1 always qualify) merge
// into one: e.g. tau(c,d,i,j) -> (i,j,c,d) is a plain (cd) x (ij) transpose.  Fewer,
// longer dims give whole 64 x 32 tiles and longer contiguous runs.  Returns false when
// nothing merges.
bool merge_dims(const long n[4], const TTerm* terms, int nterms, const TProd* prod, long nm[4], TTerm* tm,
                TProd* pm) {
  long ext[4], st[kMaxTerms + 2][4];
  const int ns = nterms + (prod ? 2 : 0);
  for (int d = 0; d < 4; ++d) {
    ext[d] = n[d];
    for (int t = 0; t < nterms; ++t) st[t][d] = terms[t].s[d];
    if (prod) {
      st[nterms][d] = prod->sx[d];
      st[nterms + 1][d] = prod->sy[d];
    }
  }
  int nd = 4;
  bool merged = false;
  for (int d = 2; d >= 0; --d) {  // merge d with the (possibly merged) dim right of it
    const int in = d + 1;
    bool ok = true;
    if (ext[d] > 1 && ext[in] > 1)
      for (int t = 0; t < ns && ok; ++t) ok = st[t][d] == st[t][in] * ext[in];
    if (!ok || ext[d] * ext[in] > (1L << 30)) continue;
    const bool keep_inner = ext[in] > 1 || ext[d] == 1;  // stride of the merged dim
    for (int t = 0; t < ns; ++t) st[t][d] = keep_inner ? st[t][in] : st[t][d];
    ext[d] *= ext[in];
    for (int e = in; e < nd - 1; ++e) {  // shift the dims right of `in` one place left
      ext[e] = ext[e + 1];
      for (int t = 0; t < ns; ++t) st[t][e] = st[t][e + 1];
    }
    --nd;
    merged = true;
  }
  if (!merged) return false;
  const int pad = 4 - nd;  // leading dims of extent 1
  for (int d = 0; d < 4; ++d) {
    nm[d] = d < pad ? 1 : ext[d - pad];
    for (int t = 0; t < nterms; ++t) {
      if (d == 0) tm[t] = terms[t];
      tm[t].s[d] = d < pad ? 0 : st[t][d - pad];
    }
    if (prod) {
      if (d == 0) *pm = *prod;
      pm->sx[d] = d < pad ? 0 : st[nterms][d - pad];
      pm->sy[d] = d < pad ? 0 : st[nterms + 1][d - pad];
    }
  }
  return true;
}
}  // namespace

PermGeom make_perm_geom(const long n[4], const TTerm* terms, int nterms, const TProd* prod) {
  if (nterms < 1 || nterms > kMaxTerms) throw std::invalid_argument("permutation: bad term count");
  long nm[4];
  TTerm tm[kMaxTerms];
  TProd pm;
  if (merge_dims(n, terms, nterms, prod, nm, tm, &pm)) {
    try {
      const PermGeom g = make_perm_geom_dims(nm, tm, nterms, prod ? &pm : nullptr);
      (void)perm_grid(g);  // throws for a grid the launch could not take
      return g;
    } catch (const std::invalid_argument&) {
      // (a merged dim too long for the grid or the linear plane: the dims as given)
    }
  }
  return make_perm_geom_dims(n, terms, nterms, prod);
}

namespace {
PermGeom make_perm_geom_dims(const long n[4], const TTerm* terms, int nterms, const TProd* prod) {
  PermGeom g{};
  long os[4];
  os[3] = 1;
  for (int d = 2; d >= 0; --d) os[d] = os[d + 1] * n[d + 1];
  // q: an output dim along which some input that is not contiguous along dim 3 is contiguous
  int q = -1;
  if (n[3] > 1)
    for (int t = 0; t < nterms && q < 0; ++t)
      if (terms[t].s[3] != 1)
        for (int d = 2; d >= 0; --d)
          if (terms[t].s[d] == 1 && n[d] > 1) {
            q = d;
            break;
          }
  g.tiled = q >= 0;
  if (g.tiled) {  // tiles smaller than a workgroup (short dims 3 and q): gathers instead
    const long t3 = ceil_div(n[3], ceil_div(n[3], kPT));
    const long tq = ceil_div(n[q], ceil_div(n[q], std::min<long>(kPT, std::max<long>(1, kPTEl / t3))));
    if (t3 * tq < kPThreads) g.tiled = false;
  }
  if (!g.tiled) q = 2;
  int r[2], nr = 0;
  for (int d = 0; d < 3; ++d)
    if (d != q) r[nr++] = d;
  g.n3 = n[3]; g.nq = n[q]; g.nr0 = n[r[0]]; g.nr1 = n[r[1]];
  g.o3 = 1; g.oq = os[q]; g.or0 = os[r[0]]; g.or1 = os[r[1]];
  // tile extents: equal parts of at most kPT along 3, then along q with t3 * tq <= kPTEl
  g.t3 = (int)ceil_div(n[3], ceil_div(n[3], kPT));
  const long tq_max = std::min<long>(kPT, std::max<long>(1, kPTEl / g.t3));
  g.tq = (int)ceil_div(n[q], ceil_div(n[q], tq_max));
  g.tiles3 = (int)ceil_div(n[3], g.t3);
  g.tilesq = (int)ceil_div(n[q], g.tq);
  g.inv_t3 = 1.0f / (float)g.t3;
  g.inv_tq = 1.0f / (float)g.tq;
  g.nterms = nterms;
  g.lds_mask = 0;
  for (int t = 0; t < nterms; ++t) {
    const TTerm& T = terms[t];
    g.t[t] = PTerm{T.p, T.s[3], T.s[q], T.s[r[0]], T.s[r[1]], T.alpha};
    if (g.tiled && T.s[3] != 1 && T.s[q] == 1) g.lds_mask |= 1u << t;
  }
  if (!g.tiled && g.nq * g.n3 >= (1L << 31)) throw std::invalid_argument("permutation plane too large");
  g.has_prod = prod != nullptr;
  if (prod) {
    long mx = 0, my = 0;
    for (int d = 0; d < 4; ++d) {
      mx += (n[d] - 1) * prod->sx[d];
      my += (n[d] - 1) * prod->sy[d];
    }
    if (!prod->x || !prod->y || mx >= prod->nx || my >= prod->ny)
      throw std::invalid_argument("product term reads past its inputs");
    g.px = prod->x; g.py = prod->y; g.palpha = prod->alpha;
    g.px3 = prod->sx[3]; g.pxq = prod->sx[q]; g.pxr0 = prod->sx[r[0]]; g.pxr1 = prod->sx[r[1]];
    g.py3 = prod->sy[3]; g.pyq = prod->sy[q]; g.pyr0 = prod->sy[r[0]]; g.pyr1 = prod->sy[r[1]];
  }
  return g;
}
}  // namespace

dim3 perm_grid(const PermGeom& g) {
  check_grid_yz(g.nr1, g.nr0);
  const long x = g.tiled ? (long)g.tiles3 * g.tilesq : ceil_div(g.nq * g.n3, (long)kPThreads * kLinPer);
  if (x >= (1L << 31)) throw std::invalid_argument("permutation grid too large");
  return dim3((unsigned)x, (unsigned)g.nr1, (unsigned)g.nr0);
}

void tadd(double* out, const long n[4], double beta, const TTerm* terms, int nterms,
          hipStream_t stream, const TProd* prod) {
  if (nterms > kMaxTerms) throw std::invalid_argument("tadd: too many terms");
  if (n[0] * n[1] * n[2] * n[3] == 0) return;
  const PermGeom g = make_perm_geom(n, terms, nterms, prod);
  if (g.tiled)
    if (perm_fast_tile(g) == 3)
      hipLaunchKernelGGL(tadd_kernel<3>, perm_grid(g), dim3(kPThreads), perm_lds_bytes(g), stream, g, out, beta);
    else if (perm_fast_tile(g) == 4)
      hipLaunchKernelGGL(tadd_kernel<4>, perm_grid(g), dim3(kPThreads), perm_lds_bytes(g), stream, g, out, beta);
    else
      hipLaunchKernelGGL(tadd_kernel<1>, perm_grid(g), dim3(kPThreads), perm_lds_bytes(g), stream, g, out, beta);
  else
    hipLaunchKernelGGL(tadd_kernel<0>, perm_grid(g), dim3(kPThreads), 0, stream, g, out, beta);
  check_launch("tadd");
}

int elem_grid(long n) { return (int)std::max<long>(1, std::min<long>(ceil_div(n, kBlock), 16384)); }

void check_small(int o, int v) {
  if ((long)v * v * o * o >= (1L << 31)) throw std::invalid_argument("v^2 o^2 must be < 2^31");
}

void build_tau(double* tau, const double* t2, const double* t1, int o, int v, hipStream_t s) {
  check_small(o, v);
  hipLaunchKernelGGL(build_tau_kernel, dim3(elem_grid((long)v * v * o * o)), dim3(kBlock), 0, s, tau,
                     t2, t1, o, v);
  check_launch("build_tau");
}

void build_tau_absmax(double* tau, const double* t2, const double* t1, int o, int v, double* partials,
                      double* maxes, hipStream_t s) {
  check_small(o, v);
  const int nb = (int)std::min<long>(elem_grid((long)v * v * o * o), kReducePartials / 2);
  hipLaunchKernelGGL(build_tau_absmax_kernel, dim3(nb), dim3(kBlock), 0, s, tau, t2, t1, o, v, partials);
  hipLaunchKernelGGL(absmax2_stage2, dim3(1), dim3(kBlock), 0, s, partials, nb, maxes);
  check_launch("build_tau_absmax");
}

void build_TP(double* TP, const double* t2, const double* t1, int o, int v, hipStream_t s) {
  check_small(o, v);
  hipLaunchKernelGGL(build_TP_kernel, dim3(elem_grid((long)v * v * o * o)), dim3(kBlock), 0, s, TP,
                     t2, t1, o, v);
  check_launch("build_TP");
}

void jacobi_update(double* t1, double* t2, const double* r1, const double* r2, const double* eps_o,
                   const double* eps_v, int o, int v, hipStream_t s, double* copy1, double* copy2) {
  check_small(o, v);
  hipLaunchKernelGGL(jacobi_kernel, dim3(elem_grid((long)v * v * o * o)), dim3(kBlock), 0, s, t1, t2,
                     r1, r2, eps_o, eps_v, o, v, copy1, copy2);
  check_launch("jacobi");
}

void dot(const double* x, const double* y, long n, double* out, double* partials, hipStream_t s) {
  const int g = red_grid(n);
  hipLaunchKernelGGL(dot_stage1, dim3(g), dim3(kBlock), 0, s, x, y, n, partials);
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(kBlock), 0, s, partials, g, out, 1.0);
  check_launch("dot");
}

// r1(a,i) += sum_{c,k} (2 t2(c,a,k,i) - t2(c,a,i,k)) h(k,c) in one pass over t2:
// workgroup (a, range of c) runs over the o x o blocks t2(c,a,:,:): each block is read
// whole (coalesced) into LDS (rows padded by one double), then thread (i, kq) sums
// over k = kq (mod Q) from LDS; its sums over kq go through LDS in a fixed order into
// part[(chunk v + a) o + i], and a second kernel adds the chunks in order.  The next
// block's loads are issued before the current block's sums.
constexpr int kT2barMaxO = 64;  // the LDS tile: o x (o + 1) doubles
__global__ __launch_bounds__(kBlock) void t2bar_h_stage1(double* __restrict__ part, const double* __restrict__ t2,
                                                         const double* __restrict__ h, int o, int v, int cchunk) {
  __shared__ double tile[kT2barMaxO * (kT2barMaxO + 1)];
  __shared__ double s[kBlock];
  constexpr int kPer = kT2barMaxO * kT2barMaxO / kBlock;  // elements of a block per thread, at most
  const int a = blockIdx.x, chunk = blockIdx.y, t = threadIdx.x;
  const int Q = kBlock / o, i = t % o, kq = t / o;
  const int oo = o * o;
  const int c0 = chunk * cchunk, c1 = min(v, (chunk + 1) * cchunk);
  double reg[kPer];
  auto load = [&](int c) {
    const double* __restrict__ T = t2 + ((long)c * v + a) * oo;  // t2(c,a,:,:)
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = t + q * kBlock;
      reg[q] = (c < c1 && e < oo) ? T[e] : 0.0;
    }
  };
  double acc = 0.0;
  load(c0);
  for (int c = c0; c < c1; ++c) {
    __syncthreads();  // the previous block's sums are done with the tile
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = t + q * kBlock;
      if (e < oo) tile[(e / o) * (o + 1) + e % o] = reg[q];
    }
    __syncthreads();
    load(c + 1);
    if (kq < Q)
      for (int k = kq; k < o; k += Q)
        acc = fma(2.0 * tile[k * (o + 1) + i] - tile[i * (o + 1) + k], h[(long)k * v + c], acc);
  }
  s[t] = kq < Q ? acc : 0.0;
  __syncthreads();
  if (t < o) {
    double sum = 0.0;
    for (int q = 0; q < Q; ++q) sum += s[q * o + t];
    part[((long)chunk * v + a) * o + t] = sum;
  }
}

__global__ __launch_bounds__(kBlock) void t2bar_h_stage2(double* __restrict__ r1, const double* __restrict__ part,
                                                         long ov, int nchunks) {
  const long e = (long)blockIdx.x * kBlock + threadIdx.x;
  if (e >= ov) return;
  double sum = 0.0;
  for (int q = 0; q < nchunks; ++q) sum += part[(long)q * ov + e];
  r1[e] += sum;
}

bool t2bar_h_supported(int o, int v) { return o >= 1 && o <= kT2barMaxO && (long)o * v <= kReducePartials; }

void t2bar_h(double* r1, const double* t2, const double* h, int o, int v, double* partials, hipStream_t s) {
  if (!t2bar_h_supported(o, v)) throw std::invalid_argument("t2bar_h: o must be <= 64 and o v <= partials");
  const long ov = (long)o * v;
  const int nchunks = (int)std::max<long>(1, std::min<long>({(long)v, kReducePartials / ov, 16}));
  const int cchunk = (int)ceil_div(v, nchunks);
  const int used = (int)ceil_div(v, cchunk);
  hipLaunchKernelGGL(t2bar_h_stage1, dim3((unsigned)v, (unsigned)used), dim3(kBlock), 0, s, partials, t2, h, o, v,
                     cchunk);
  hipLaunchKernelGGL(t2bar_h_stage2, dim3((unsigned)ceil_div(ov, kBlock)), dim3(kBlock), 0, s, r1, partials, ov,
                     used);
  check_launch("t2bar_h");
}

void ccsd_energy(const double* f_ai, const double* r1, const double* t1, const double* g_abij,
                 const double* r2, const double* tau, int o, int v, double* out, double* partials,
                 hipStream_t s) {
  if ((long)v * v * o * o >= (1L << 31)) throw std::invalid_argument("ccsd_energy: v^2 o^2 too large");
  const int g = red_grid((long)v * v * o * o);
  hipLaunchKernelGGL(energy_stage1, dim3(g), dim3(kBlock), 0, s, f_ai, r1, t1, g_abij, r2, tau, o, v,
                     partials);
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(kBlock), 0, s, partials, g, out, 1.0);
  check_launch("ccsd_energy");
}

void ccsd_energy_norms(const double* f_ai, const double* r1, const double* t1, const double* g_abij,
                       const double* r2, const double* tau, int o, int v, double* out3, double* partials,
                       hipStream_t s) {
  if ((long)v * v * o * o >= (1L << 31)) throw std::invalid_argument("ccsd_energy: v^2 o^2 too large");
  const int g = red_grid((long)v * v * o * o);
  hipLaunchKernelGGL(energy_norms_stage1, dim3(g), dim3(kBlock), 0, s, f_ai, r1, t1, g_abij, r2, tau, o, v,
                     partials);
  hipLaunchKernelGGL(reduce_stage2, dim3(3), dim3(kBlock), 0, s, partials, g, out3, 1.0);
  check_launch("ccsd_energy_norms");
}

void multi_dot2(const double* const* xs1, const double* const* xs2, int nvec, const double* y1,
                const double* y2, long n1, long n2, double* out, double* partials, hipStream_t s,
                double* copy1, double* copy2) {
  if (nvec > 16) throw std::invalid_argument("multi_dot2: at most 16 vectors");
  MultiDotArgs a;
  for (int k = 0; k < 16; ++k) {
    a.x1[k] = k < nvec ? xs1[k] : nullptr;
    a.x2[k] = k < nvec ? xs2[k] : nullptr;
  }
  a.y1 = y1; a.y2 = y2; a.n1 = n1; a.n2 = n2; a.nvec = nvec;
  a.c1 = copy1; a.c2 = copy2;
  auto al16 = [](const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
  bool vec = n2 % 2 == 0 && n2 >= 2 * (long)kBlock * kRedGrid && al16(y2) && al16(copy2);
  for (int k = 0; k < nvec; ++k) vec = vec && al16(xs2[k]);
  const int g = red_grid(std::max(n1, vec ? n2 / 2 : n2));
  if (vec)
    hipLaunchKernelGGL(multidot_vec_stage1, dim3(g), dim3(kBlock), 0, s, a, partials);
  else
    hipLaunchKernelGGL(multidot_stage1, dim3(g), dim3(kBlock), 0, s, a, partials);
  hipLaunchKernelGGL(reduce_stage2, dim3(nvec), dim3(kBlock), 0, s, partials, g, out, 1.0);
  check_launch("multi_dot2");
}

void scale_matrix(double* C, int batch, int M, int N, long sC, long ldc, double beta, hipStream_t s) {
  const long total = (long)batch * M * N;
  if (total <= 0 || beta == 1.0) return;
  hipLaunchKernelGGL(scale_matrix_kernel, dim3(elem_grid(total)), dim3(kBlock), 0, s, C, total, M, N, sC,
                     ldc, beta);
  check_launch("scale_matrix");
}

void fill(double* x, long n, double value, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(elem_grid(n)), dim3(kBlock), 0, s, x, n, value);
  check_launch("fill");
}

void lincomb(double* y, const double* const* xs, const double* c, int nvec, long n, hipStream_t s) {
  if (nvec > 16) throw std::invalid_argument("lincomb: at most 16 vectors");
  LincombArgs a;
  for (int k = 0; k < 16; ++k) {
    a.x[k] = k < nvec ? xs[k] : nullptr;
    a.c[k] = k < nvec ? c[k] : 0.0;
  }
  a.nvec = nvec;
  const int g = (int)std::min<long>(ceil_div(n, kBlock), 8192);
  hipLaunchKernelGGL(lincomb_kernel, dim3(g), dim3(kBlock), 0, s, y, a, n);
  check_launch("lincomb");
}

void lincomb_tau(double* t2, double* tau, const double* t1, const double* const* xs, const double* c, int nvec,
                 int o, int v, double* partials, double* maxes, hipStream_t s) {
  check_small(o, v);
  if (nvec > 16) throw std::invalid_argument("lincomb: at most 16 vectors");
  LincombArgs a;
  for (int k = 0; k < 16; ++k) {
    a.x[k] = k < nvec ? xs[k] : nullptr;
    a.c[k] = k < nvec ? c[k] : 0.0;
  }
  a.nvec = nvec;
  const int nb = (int)std::min<long>(elem_grid((long)v * v * o * o), kReducePartials / 2);
  if (maxes) {
    hipLaunchKernelGGL(lincomb_tau_kernel<true>, dim3(nb), dim3(kBlock), 0, s, t2, tau, t1, a, o, v, partials);
    hipLaunchKernelGGL(absmax2_stage2, dim3(1), dim3(kBlock), 0, s, partials, nb, maxes);
  } else {
    hipLaunchKernelGGL(lincomb_tau_kernel<false>, dim3(nb), dim3(kBlock), 0, s, t2, tau, t1, a, o, v, partials);
  }
  check_launch("lincomb_tau");
}

void count_nonfinite(const double* x, long n, double* out, double* partials, hipStream_t s) {
  const int g = red_grid(n);
  hipLaunchKernelGGL(nonfinite_stage1, dim3(g), dim3(kBlock), 0, s, x, n, partials);
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(kBlock), 0, s, partials, g, out, 1.0);
  check_launch("count_nonfinite");
}

void make_tile_modeinfo(long* host_out, const long n[4], const long* const hashmarks[4],
                        const int ntiles[4]) {
  long* p = host_out;
  for (int d = 0; d < 4; ++d) {
    const long nd = n[d];
    if (hashmarks[d] == nullptr) {  // one tile covering the mode
      for (long i = 0; i < nd; ++i) { p[i] = 0; p[nd + i] = nd; }
    } else {
      const long* hm = hashmarks[d];
      if (hm[0] != 0 || hm[ntiles[d]] != nd)
        throw std::invalid_argument("tile hashmarks must start at 0 and end at the extent");
      for (int t = 0; t < ntiles[d]; ++t) {
        if (hm[t + 1] <= hm[t]) throw std::invalid_argument("tile hashmarks must increase");
        for (long i = hm[t]; i < hm[t + 1]; ++i) { p[i] = hm[t]; p[nd + i] = hm[t + 1] - hm[t]; }
      }
    }
    p += 2 * nd;
  }
}

void tiles_to_dense(double* dense, const double* tiled, const long n[4], const long* modeinfo,
                    hipStream_t s) {
  TileArgs a;
  for (int d = 0; d < 4; ++d) a.n[d] = n[d];
  check_grid_yz(n[1], n[0]);
  dim3 grid((unsigned)ceil_div(n[2] * n[3], kBlock), (unsigned)n[1], (unsigned)n[0]);
  hipLaunchKernelGGL(tiles_kernel<true>, grid, dim3(kBlock), 0, s, dense, tiled, modeinfo, a);
  check_launch("tiles_to_dense");
}

void dense_to_tiles(double* tiled, const double* dense, const long n[4], const long* modeinfo,
                    hipStream_t s) {
  TileArgs a;
  for (int d = 0; d < 4; ++d) a.n[d] = n[d];
  check_grid_yz(n[1], n[0]);
  dim3 grid((unsigned)ceil_div(n[2] * n[3], kBlock), (unsigned)n[1], (unsigned)n[0]);
  hipLaunchKernelGGL(tiles_kernel<false>, grid, dim3(kBlock), 0, s, tiled, dense, modeinfo, a);
  check_launch("dense_to_tiles");
}

}  // namespace mpqc_hip
