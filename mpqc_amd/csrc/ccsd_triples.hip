// Perturbative triples correction (T) on the GPU from the converged CCSD
// amplitudes held by the context.
//
// Reference: CCSD_T::compute_ccsd_t_straight (src/mpqc/chemistry/qc/lcao/cc/
// ccsd_t.h:1127-1167) with CCSD_T_Reduce (:2274-2335):
//   t3(a,b,c,i,j,k) = <id|ab> t2(d,c,j,k) - <jk|lc> t2(a,b,i,l)        (raw W)
//   t3s = sum of the six simultaneous permutations of (a,i), (b,j), (c,k)
//   v3s = <ij|ab> t1(c,k) + <jk|bc> t1(a,i) + <ik|ac> t1(b,j)
//   E(T) = 1/3 sum (t3s + v3s)(4 t3s + t3s(abckij) + t3s(abcjki)
//                   - 2 (t3s(abckji) + t3s(abcikj) + t3s(abcjik))) / D
// The reference stores t3 over all (a,b,c,i,j,k); here the occupied triples
// are visited as multisets {i <= j <= k}: for a pair (j,k) and a block of i,
// 12 batched GEMMs build the raw W of all six orderings of each triple
// (particle term M = v^2, N = v, K = v; hole term K = o), and one kernel sums
// the energy.  A thread of that kernel owns one virtual multiset
// {a >= b >= c}: it loads the 36 values W[ordering t](permutation q of a,b,c)
// once, forms t3s for every (permutation, ordering) pair from them and adds
// the 36 energy terms of the orbit, weighted by 1/(|stabiliser of abc| *
// |stabiliser of ijk|) so that every ordered (a,b,c,i,j,k) counts once.  No
// t3 element is stored beyond the current block.
#include <cmath>
#include <stdexcept>

#include "ccsd_ctx.hpp"

namespace mpqc_hip {

namespace {

// S3 as position maps: 0 (0,1,2) 1 (0,2,1) 2 (1,0,2) 3 (1,2,0) 4 (2,0,1) 5 (2,1,0)
__constant__ int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
// kComp[p][r] = index of the map n -> p[r[n]]
__constant__ int kComp[6][6];
// energy terms: t3s at occupied order (y'[m[0]], y'[m[1]], y'[m[2]]) with coefficient
// (i,j,k):4  (k,i,j):1  (j,k,i):1  (k,j,i):-2  (i,k,j):-2  (j,i,k):-2
__constant__ int kTermPerm[6] = {0, 4, 3, 5, 1, 2};
__constant__ double kTermCoef[6] = {4.0, 1.0, 1.0, -2.0, -2.0, -2.0};

constexpr int kTThreads = 256;

void upload_s3_tables(hipStream_t s) {
  static const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int comp[6][6];
  for (int p = 0; p < 6; ++p)
    for (int r = 0; r < 6; ++r) {
      const int m[3] = {P[p][P[r][0]], P[p][P[r][1]], P[p][P[r][2]]};
      int idx = -1;
      for (int q = 0; q < 6; ++q)
        if (P[q][0] == m[0] && P[q][1] == m[1] && P[q][2] == m[2]) idx = q;
      comp[p][r] = idx;
    }
  MPQC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(kComp), comp, sizeof(comp), 0, hipMemcpyHostToDevice, s));
}

__device__ inline long tet(long a) { return a * (a + 1) * (a + 2) / 6; }

// virtual multiset number e -> (a >= b >= c)
__device__ inline void unrank_abc(long e, int& a, int& b, int& c) {
  long x = (long)cbrt(6.0 * (double)e);
  while (x > 0 && tet(x) > e) --x;
  while (tet(x + 1) <= e) ++x;
  const long rem = e - tet(x);
  long y = (long)((sqrt(8.0 * (double)rem + 1.0) - 1.0) * 0.5);
  while (y > 0 && y * (y + 1) / 2 > rem) --y;
  while ((y + 1) * (y + 2) / 2 <= rem) ++y;
  a = (int)x;
  b = (int)y;
  c = (int)(rem - y * (y + 1) / 2);
}

struct TriplesArgs {
  const double* W;  // [6 orderings][nbmax][v^3]
  long wsig;        // stride between orderings (nbmax v^3)
  const double* G;  // Gijab (i,j,a,b)
  const double* t1; // (a,i)
  const double* eo;
  const double* ev;
  int o, v, i0, j, k;
  long nabc;        // v (v+1) (v+2) / 6
};

__global__ __launch_bounds__(kTThreads) void triples_energy_kernel(TriplesArgs g, double* partials) {
  const long e = (long)blockIdx.x * kTThreads + threadIdx.x;
  const int z = blockIdx.y;
  double acc = 0.0;
  if (e < g.nabc) {
    int x[3];
    unrank_abc(e, x[0], x[1], x[2]);
    const int y[3] = {g.i0 + z, g.j, g.k};
    const int v = g.v, o = g.o;
    const long v3 = (long)v * v * v;
    double L[6][6];  // L[q][t] = W_t(x_q)
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      const double* Wt = g.W + t * g.wsig + (long)z * v3;
#pragma unroll
      for (int q = 0; q < 6; ++q)
        L[q][t] = Wt[((long)x[kPerm[q][0]] * v + x[kPerm[q][1]]) * v + x[kPerm[q][2]]];
    }
    double TS[6][6];  // t3s(x_p, y_s) = sum_r L[p o r][s o r]
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        double t = 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r) t += L[kComp[p][r]][kComp[s][r]];
        TS[p][s] = t;
      }
    double sum = 0.0;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      const int a0 = x[kPerm[p][0]], a1 = x[kPerm[p][1]], a2 = x[kPerm[p][2]];
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int i0 = y[kPerm[s][0]], i1 = y[kPerm[s][1]], i2 = y[kPerm[s][2]];
        // v3s(x', y') with g_abij(a,b,i,j) = Gijab(i,j,a,b)
        const double vs = g.G[(((long)i0 * o + i1) * v + a0) * v + a1] * g.t1[(long)a2 * o + i2] +
                          g.G[(((long)i1 * o + i2) * v + a1) * v + a2] * g.t1[(long)a0 * o + i0] +
                          g.G[(((long)i0 * o + i2) * v + a0) * v + a2] * g.t1[(long)a1 * o + i1];
        double X = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) X += kTermCoef[m] * TS[p][kComp[s][kTermPerm[m]]];
        sum += (TS[p][s] + vs) * X;
      }
    }
    const int sx = (x[0] == x[1] && x[1] == x[2]) ? 6 : (x[0] == x[1] || x[1] == x[2]) ? 2 : 1;
    const int sy = (y[0] == y[1] && y[1] == y[2]) ? 6 : (y[0] == y[1] || y[1] == y[2]) ? 2 : 1;
    const double D = g.eo[y[0]] + g.eo[y[1]] + g.eo[y[2]] - g.ev[x[0]] - g.ev[x[1]] - g.ev[x[2]];
    acc = sum / (3.0 * D * (double)(sx * sy));
  }
  __shared__ double sh[kTThreads];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kTThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[(long)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

// acc[slot] = sum of partials[0..n) in a fixed order (one workgroup)
__global__ __launch_bounds__(256) void sum_to_slot(const double* partials, long n, double* acc, long slot) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long e = threadIdx.x; e < n; e += 256) s += partials[e];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[slot] = sh[0];
}

struct DevBuf {
  double* p = nullptr;
  explicit DevBuf(size_t n) {
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw std::bad_alloc();
    }
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace

// Giabc(i,a,b,c) = <ia|bc> = (ib|ac) = sum_K Xai(K,b,i) Xab(K,a,c) for the df
// method without a stored Giabc (the reference's get_abci with "[df]",
// ccsd_t.h:2224-2234, through compute4's density-fitted branch)
void CCSDContext::ensure_iabc_for_triples() {
  if ((have_ints_ >> 6) & 1u) return;
  if (!df_) throw std::logic_error("(T) needs Giabc");
  const long Q = naux_, O = o_, V = v_;
  DevBuf T((size_t)(O * V * V * V));
  // T[(i,b)][(a,c)] = Xia(K,(i,b))^T Xab(K,(a,c)) ; Giabc(i,a,b,c) = T(i,b,a,c)
  mpqc_hip::gemm(make_gemm(true, false, (int)(O * V), (int)(V * V), (int)Q, 1.0, Xia_, O * V, Xab_, V * V, 0.0,
                           T.p, V * V), GemmPrecision::F64, ws_, stream_);
  TTerm t;
  t.p = T.p;
  const long in_ext[4] = {O, V, V, V};
  const int pm[4] = {0, 2, 1, 3};
  permuted_strides(in_ext, pm, t.s);
  t.alpha = 1.0;
  t.vol = O * V * V * V;
  const long out_ext[4] = {O, V, V, V};
  tadd(Giabc_, out_ext, 0.0, &t, 1, stream_);
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  have_ints_ |= 1u << 6;
}

double CCSDContext::triples(double block_bytes, double* seconds) {
  if (!have_eps_) throw std::logic_error("(T) needs the orbital energies");
  if (!prepared_) prepare();
  ensure_iabc_for_triples();
  const int o = o_, v = v_;
  const long O = o, V = v, v3 = V * V * V, ov = O * V;
  if (V * (V + 1) * (V + 2) / 6 >= (1L << 31) || V * V >= (1L << 31))
    throw std::invalid_argument("v too large for (T)");
  long nbmax = (long)(block_bytes / (6.0 * sizeof(double) * (double)v3));
  nbmax = std::max(1L, std::min(nbmax, O));
  hipEvent_t ev0, ev1;
  MPQC_HIP_CHECK(hipEventCreate(&ev0));
  MPQC_HIP_CHECK(hipEventCreate(&ev1));
  MPQC_HIP_CHECK(hipEventRecord(ev0, stream_));
  upload_s3_tables(stream_);
  // t2P(j,k,d,c) = t2(d,c,j,k): the B operand of the particle term for a pair (j,k)
  DevBuf t2P((size_t)(ov * ov)), W((size_t)(6 * nbmax * v3));
  {
    TTerm t;
    t.p = t2_;
    const long in_ext[4] = {V, V, O, O};
    const int pm[4] = {2, 3, 0, 1};
    permuted_strides(in_ext, pm, t.s);
    t.vol = ov * ov;
    const long out_ext[4] = {O, O, V, V};
    tadd(t2P.p, out_ext, 0.0, &t, 1, stream_);
  }
  const long nabc = V * (V + 1) * (V + 2) / 6;
  const int gx = (int)ceil_div(nabc, kTThreads);
  long nchunks = 0;
  for (int k = 0; k < o; ++k)
    for (int j = 0; j <= k; ++j) nchunks += ceil_div(j + 1, nbmax);
  DevBuf partials((size_t)gx * nbmax), acc((size_t)nchunks);
  static const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  long slot = 0;
  for (int k = 0; k < o; ++k)
    for (int j = 0; j <= k; ++j)
      for (int i0 = 0; i0 <= j; i0 += (int)nbmax) {
        const int nb = (int)std::min<long>(nbmax, j + 1 - i0);
        const int y0[3] = {i0, j, k};
        for (int s = 0; s < 6; ++s) {
          // ordering s: (o1, o2, o3) = (y[P0], y[P1], y[P2]); position of i among them
          const int o1 = y0[P[s][0]], o2 = y0[P[s][1]], o3 = y0[P[s][2]];
          const int ipos = P[s][0] == 0 ? 0 : P[s][1] == 0 ? 1 : 2;
          double* Ws = W.p + (long)s * nbmax * v3;
          // particle: W(a,b,c) = sum_d Giabc(o1,d,a,b) t2P(o2,o3,d,c)
          GemmDesc dp = make_gemm(true, false, v * v, v, v, 1.0, Giabc_ + (long)o1 * v3, V * V,
                                  t2P.p + ((long)o2 * O + o3) * V * V, V, 0.0, Ws, V);
          dp.sA = ipos == 0 ? v3 : 0;
          dp.sB = ipos == 1 ? O * V * V : ipos == 2 ? V * V : 0;
          dp.sC = v3;
          dp.batch = nb;
          mpqc_hip::gemm(dp, GemmPrecision::F64, ws_, stream_);
          // hole: W(a,b,c) -= sum_l t2(a,b,o1,l) Gijka(o2,o3,l,c)
          GemmDesc dh = make_gemm(false, false, v * v, v, o, -1.0, t2_ + (long)o1 * O, O * O,
                                  Gijka_ + ((long)o2 * O + o3) * O * V, V, 1.0, Ws, V);
          dh.sA = ipos == 0 ? O : 0;
          dh.sB = ipos == 1 ? O * O * V : ipos == 2 ? O * V : 0;
          dh.sC = v3;
          dh.batch = nb;
          mpqc_hip::gemm(dh, GemmPrecision::F64, ws_, stream_);
        }
        TriplesArgs g;
        g.W = W.p; g.wsig = nbmax * v3; g.G = Gijab_; g.t1 = t1_; g.eo = eo_; g.ev = ev_;
        g.o = o; g.v = v; g.i0 = i0; g.j = j; g.k = k; g.nabc = nabc;
        hipLaunchKernelGGL(triples_energy_kernel, dim3(gx, nb), dim3(kTThreads), 0, stream_, g, partials.p);
        check_launch("triples_energy");
        hipLaunchKernelGGL(sum_to_slot, dim3(1), dim3(256), 0, stream_, partials.p, (long)gx * nb, acc.p, slot);
        check_launch("triples_sum");
        ++slot;
      }
  MPQC_HIP_CHECK(hipEventRecord(ev1, stream_));
  std::vector<double> h(nchunks);
  MPQC_HIP_CHECK(hipMemcpyAsync(h.data(), acc.p, nchunks * sizeof(double), hipMemcpyDeviceToHost, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  float ms = 0.f;
  MPQC_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (seconds) *seconds = ms * 1e-3;
  double e = 0.0;
  for (double x : h) e += x;
  if (!std::isfinite(e)) throw std::runtime_error("non-finite (T) energy");
  return e;
}

}  // namespace mpqc_hip
