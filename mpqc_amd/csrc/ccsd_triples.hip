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
// The reference stores t3 over all (a,b,c,i,j,k).  Here the occupied triples
// are visited as multisets {i <= j <= k}: for a pair (j,k) and a block of i,
// five GEMMs build the raw W of the six orderings of each triple.  The
// particle and hole terms are one GEMM with K = v + o: A_i = [<id|ab>^T | t2(a,b,i,:)]
// ((a,b) x (v+o), built once) and B_jk = [t2(:,c,j,k) ; -<jk|:c>] ((v+o) x c).
// The orderings (i,j,k), (i,k,j) stack the block's i along M and the two B's
// along N; the four orderings led by j or k share A and put the block's i
// along N (N = nb v), so every GEMM has long M and N.
//
// Energy: t3s and v3s are invariant under the simultaneous permutation of
// (a,b,c) and (i,j,k), and the coefficients (4, 1, 1, -2, -2, -2) depend only
// on the conjugacy class of the occupied permutation, so the 36 energy terms
// of the orbit pair {a,b,c} x {i,j,k} collapse to 6 x sum_u F(u) with
//   T_u = t3s(x, y_u) = sum_r W_{u o r}(x_r),  F(u) = (T_u + V_u) sum_m c_m T_{u o m},
// x = (a >= b >= c), y = (i <= j <= k), weighted by 1/(|stab x| |stab y|) so
// that every ordered (a,b,c,i,j,k) counts once.  In the split-precision forms
// the W GEMMs read 16-bit planes of A and B (split once per call) and write
// FP64 W; the energy is FP64 throughout.  A workgroup owns a multiset
// of 8^3 cubes of (a,b,c) space and stages, ordering by ordering, the six
// permuted cubes of W_s through LDS, so W is read with coalesced rows once.
#include <cmath>
#include <memory>
#include <stdexcept>

#include "ccsd_ctx.hpp"

namespace mpqc_hip {

namespace {

// S3 as position maps: 0 (0,1,2) 1 (0,2,1) 2 (1,0,2) 3 (1,2,0) 4 (2,0,1) 5 (2,1,0)
__constant__ int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
// kComp[p][r] = index of the map n -> p[r[n]];  kInv[r] = index of r^-1
__constant__ int kComp[6][6];
__constant__ int kInv[6];
// energy terms: t3s at occupied order (y'[m[0]], y'[m[1]], y'[m[2]]) with coefficient
// (i,j,k):4  (k,i,j):1  (j,k,i):1  (k,j,i):-2  (i,k,j):-2  (j,i,k):-2
__constant__ int kTermPerm[6] = {0, 4, 3, 5, 1, 2};
__constant__ double kTermCoef[6] = {4.0, 1.0, 1.0, -2.0, -2.0, -2.0};

constexpr int kT = 8;                        // cube edge
constexpr int kS2 = kT + 1, kS1 = kT * kS2 + 1;  // padded LDS strides (odd: no bank conflicts)
constexpr int kCube = kT * kS1;
constexpr int kTThreads = 256;
static_assert(6 * kT * kT * kT % kTThreads == 0 && 6 * kT * kT * kT / kTThreads <= 32, "cube staging layout");

void upload_s3_tables(hipStream_t s) {
  static const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  auto index_of = [&](const int m[3]) {
    for (int q = 0; q < 6; ++q)
      if (P[q][0] == m[0] && P[q][1] == m[1] && P[q][2] == m[2]) return q;
    return -1;
  };
  int comp[6][6], inv[6];
  for (int p = 0; p < 6; ++p) {
    for (int r = 0; r < 6; ++r) {
      const int m[3] = {P[p][P[r][0]], P[p][P[r][1]], P[p][P[r][2]]};
      comp[p][r] = index_of(m);
    }
    int m[3];
    for (int n = 0; n < 3; ++n) m[P[p][n]] = n;
    inv[p] = index_of(m);
  }
  MPQC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(kComp), comp, sizeof(comp), 0, hipMemcpyHostToDevice, s));
  MPQC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(kInv), inv, sizeof(inv), 0, hipMemcpyHostToDevice, s));
}

__device__ inline long tet(long a) { return a * (a + 1) * (a + 2) / 6; }

// multiset number e -> (a >= b >= c)
__device__ inline void unrank3(long e, int& a, int& b, int& c) {
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
  // W of ordering s for triple z of the block at (a,b,c):
  //   Wb[s] + z * Wz[s] + (a v + b) * Wr[s] + c
  const double* Wb[6];
  long Wz[6], Wr[6];
  const double* G;  // Gijab (i,j,a,b)
  const double* t1; // (a,i)
  const double* eo;
  const double* ev;
  int o, v, i0, j, k;
};

__global__ __launch_bounds__(kTThreads) void triples_energy_kernel(TriplesArgs g, double* partials) {
  __shared__ double tile[6 * kCube];
  __shared__ double red[kTThreads];
  const int z = blockIdx.y;
  const int v = g.v, o = g.o;
  int X[3];
  unrank3(blockIdx.x, X[0], X[1], X[2]);  // cube multiset A >= B >= C
  const int y[3] = {g.i0 + z, g.j, g.k};
  // this thread's two points of cube (A,B,C): local coordinates and validity
  int l[2][3];
  bool ok[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int q = threadIdx.x + h * kTThreads;
    l[h][0] = q / (kT * kT);
    l[h][1] = (q / kT) % kT;
    l[h][2] = q % kT;
    const int a = X[0] * kT + l[h][0], b = X[1] * kT + l[h][1], c = X[2] * kT + l[h][2];
    ok[h] = a < v && b < v && c < v && a >= b && b >= c;
  }
  double T[2][6];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int u = 0; u < 6; ++u) T[h][u] = 0.0;
  // W_s cubes: every thread owns kPerThr elements of the six permuted cubes; the
  // loads of ordering s + 1 are issued before the products of ordering s, from
  // clamped addresses (out-of-range elements are zeroed at the LDS store), so they
  // stay in flight across the compute
  constexpr int kCubeEl = kT * kT * kT, kPerThr = 6 * kCubeEl / kTThreads;
  double pre[kPerThr];
  uint32_t okmask = 0;
  auto gload = [&](int s) {
    const double* Ws = g.Wb[s] + (long)z * g.Wz[s];
    const long rs = g.Wr[s];
#pragma unroll
    for (int k = 0; k < kPerThr; ++k) {
      const int e = threadIdx.x + k * kTThreads;
      const int r = e / kCubeEl, rem = e - r * kCubeEl;
      const int u = rem / (kT * kT), w = (rem / kT) % kT, zz = rem % kT;
      const int ga = X[kPerm[r][0]] * kT + u, gb = X[kPerm[r][1]] * kT + w, gc = X[kPerm[r][2]] * kT + zz;
      const bool in = ga < v && gb < v && gc < v;
      okmask = in ? (okmask | (1u << k)) : (okmask & ~(1u << k));
      pre[k] = Ws[((long)min(ga, v - 1) * v + min(gb, v - 1)) * rs + min(gc, v - 1)];
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int k = 0; k < kPerThr; ++k) {
      const int e = threadIdx.x + k * kTThreads;
      const int r = e / kCubeEl, rem = e - r * kCubeEl;
      const int u = rem / (kT * kT), w = (rem / kT) % kT, zz = rem % kT;
      tile[r * kCube + u * kS1 + w * kS2 + zz] = ((okmask >> k) & 1u) ? pre[k] : 0.0;
    }
  };
  gload(0);
  for (int s = 0; s < 6; ++s) {
    lstore();
    __syncthreads();
    if (s + 1 < 6) gload(s + 1);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (!ok[h]) continue;
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        // x_r lies in cube r at local (l[r0], l[r1], l[r2]); W_s(x_r) adds to T_{s o r^-1}
        const double val = tile[r * kCube + l[h][kPerm[r][0]] * kS1 + l[h][kPerm[r][1]] * kS2 + l[h][kPerm[r][2]]];
        T[h][kComp[s][kInv[r]]] += val;
      }
    }
    __syncthreads();
  }
  double acc = 0.0;
  const int sy = (y[0] == y[1] && y[1] == y[2]) ? 6 : (y[0] == y[1] || y[1] == y[2]) ? 2 : 1;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (!ok[h]) continue;
    const int x[3] = {X[0] * kT + l[h][0], X[1] * kT + l[h][1], X[2] * kT + l[h][2]};
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int i0 = y[kPerm[u][0]], i1 = y[kPerm[u][1]], i2 = y[kPerm[u][2]];
      // v3s(x, y_u) with g_abij(a,b,i,j) = Gijab(i,j,a,b)
      const double vs = g.G[(((long)i0 * o + i1) * v + x[0]) * v + x[1]] * g.t1[(long)x[2] * o + i2] +
                        g.G[(((long)i1 * o + i2) * v + x[1]) * v + x[2]] * g.t1[(long)x[0] * o + i0] +
                        g.G[(((long)i0 * o + i2) * v + x[0]) * v + x[2]] * g.t1[(long)x[1] * o + i1];
      double Xm = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) Xm += kTermCoef[m] * T[h][kComp[u][kTermPerm[m]]];
      sum += (T[h][u] + vs) * Xm;
    }
    const int sx = (x[0] == x[1] && x[1] == x[2]) ? 6 : (x[0] == x[1] || x[1] == x[2]) ? 2 : 1;
    const double D = g.eo[y[0]] + g.eo[y[1]] + g.eo[y[2]] - g.ev[x[0]] - g.ev[x[1]] - g.ev[x[2]];
    acc += 2.0 * sum / (D * (double)(sx * sy));  // 6 sum_u F(u) / (3 D |stab x| |stab y|)
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kTThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[(long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
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

// planes of one (T) operand, owned for the duration of the call
struct TPlanes {
  Planes P;
  std::vector<void*> owned;
  TPlanes(PlaneKind kind, long n) {
    P.n = n;
    for (int q = 0; q < plane_count(kind); ++q) {
      void* ptr = nullptr;
      if (hipMalloc(&ptr, std::max<long>(n, 8) * sizeof(uint16_t)) != hipSuccess) {
        (void)hipGetLastError();
        release();
        throw std::bad_alloc();
      }
      owned.push_back(ptr);
      P.p[q] = (uint16_t*)ptr;
    }
    void* sc = nullptr;
    if (hipMalloc(&sc, sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      release();
      throw std::bad_alloc();
    }
    owned.push_back(sc);
    P.scale = (double*)sc;
  }
  void release() {
    for (void* x : owned) (void)hipFree(x);
    owned.clear();
  }
  ~TPlanes() { release(); }
  TPlanes(const TPlanes&) = delete;
  TPlanes& operator=(const TPlanes&) = delete;
  // the same planes starting at element `off`
  Planes at(long off) const {
    Planes q = P;
    for (int k = 0; k < 3; ++k)
      if (q.p[k]) q.p[k] += off;
    q.n = P.n - off;
    return q;
  }
};

// out (dense, extents out_ext) = alpha * input permuted (out axis d reads input axis perm[d])
void permute_into(double* out, const long out_ext[4], const double* in, const long in_ext[4], const int perm[4],
                  double alpha, hipStream_t s) {
  TTerm t;
  t.p = in;
  permuted_strides(in_ext, perm, t.s);
  t.alpha = alpha;
  t.vol = in_ext[0] * in_ext[1] * in_ext[2] * in_ext[3];
  tadd(out, out_ext, 0.0, &t, 1, s);
}

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
  const long ext[4] = {O, V, V, V};
  const int pm[4] = {0, 2, 1, 3};
  permute_into(Giabc_, ext, T.p, ext, pm, 1.0, stream_);
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  have_ints_ |= 1u << 6;
}

double CCSDContext::triples(double block_bytes, double* seconds) {
  if (!have_eps_) throw std::logic_error("(T) needs the orbital energies");
  if (!prepared_) prepare();
  ensure_iabc_for_triples();
  const int o = o_, v = v_;
  const long O = o, V = v, v3 = V * V * V, VK = V + O;
  const long nB = ceil_div(V, kT);
  const long ncube = nB * (nB + 1) * (nB + 2) / 6;
  if (ncube >= (1L << 31) || V * V >= (1L << 31)) throw std::invalid_argument("v too large for (T)");
  long nbmax = (long)(block_bytes / (6.0 * sizeof(double) * (double)v3));
  nbmax = std::max(1L, std::min(nbmax, O));
  hipEvent_t ev0, ev1;
  MPQC_HIP_CHECK(hipEventCreate(&ev0));
  MPQC_HIP_CHECK(hipEventCreate(&ev1));
  MPQC_HIP_CHECK(hipEventRecord(ev0, stream_));
  upload_s3_tables(stream_);
  // GEMM operands with K = v + o:
  //   Acat(i, (a,b), K)   = [<id|ab> for K = d < v | t2(a,b,i,l) for K = v + l]
  //   Bcat(j, k, K, c)    = [t2(d,c,j,k)           | -<jk|lc>]
  // and three layouts of Bcat that make the N dimension of each GEMM long:
  //   Bt (K, j, k, c), Bt2 (K, k, j, c) = Bcat(j, k, K, c), Bp (j, k, K, 2, c) = [Bcat(j,k) | Bcat(k,j)]
  const size_t d = sizeof(double);
  DevBuf Acat((size_t)(O * V * V * VK)), Bt((size_t)(VK * O * O * V)), Bt2((size_t)(VK * O * O * V)),
      Bp((size_t)(2 * O * O * VK * V)), W((size_t)(6 * nbmax * v3));
  {
    DevBuf tmp((size_t)(O * V * V * std::max(V, O)));
    // Acat[i][(a,b)][d] = Giabc(i,d,a,b); Acat[i][(a,b)][v + l] = t2(a,b,i,l)
    const long e_iabc[4] = {O, V, V, V};
    const int p1[4] = {0, 2, 3, 1};
    permute_into(tmp.p, e_iabc, Giabc_, e_iabc, p1, 1.0, stream_);
    MPQC_HIP_CHECK(hipMemcpy2DAsync(Acat.p, VK * d, tmp.p, V * d, V * d, O * V * V, hipMemcpyDeviceToDevice,
                                    stream_));
    const long e_t2[4] = {V, V, O, O};
    const int p2[4] = {2, 0, 1, 3};
    const long ext2[4] = {O, V, V, O};
    permute_into(tmp.p, ext2, t2_, e_t2, p2, 1.0, stream_);
    MPQC_HIP_CHECK(hipMemcpy2DAsync(Acat.p + V, VK * d, tmp.p, O * d, O * d, O * V * V, hipMemcpyDeviceToDevice,
                                    stream_));
    // Bcat[j][k][d][c] = t2(d,c,j,k); Bcat[j][k][v + l][c] = -Gijka(j,k,l,c)
    DevBuf Bcat((size_t)(O * O * VK * V)), tmp2((size_t)(O * O * std::max(V, O) * V));
    const int p3[4] = {2, 3, 0, 1};
    const long ext3[4] = {O, O, V, V};
    permute_into(tmp2.p, ext3, t2_, e_t2, p3, 1.0, stream_);
    MPQC_HIP_CHECK(hipMemcpy2DAsync(Bcat.p, VK * V * d, tmp2.p, V * V * d, V * V * d, O * O,
                                    hipMemcpyDeviceToDevice, stream_));
    const long e_ijka[4] = {O, O, O, V};
    const int pid[4] = {0, 1, 2, 3};
    permute_into(tmp2.p, e_ijka, Gijka_, e_ijka, pid, -1.0, stream_);
    MPQC_HIP_CHECK(hipMemcpy2DAsync(Bcat.p + V * V, VK * V * d, tmp2.p, O * V * d, O * V * d, O * O,
                                    hipMemcpyDeviceToDevice, stream_));
    const long e_b[4] = {O, O, VK, V};
    const int pt[4] = {2, 0, 1, 3}, pt2[4] = {2, 1, 0, 3}, psw[4] = {1, 0, 2, 3};
    const long e_bt[4] = {VK, O, O, V};
    permute_into(Bt.p, e_bt, Bcat.p, e_b, pt, 1.0, stream_);
    permute_into(Bt2.p, e_bt, Bcat.p, e_b, pt2, 1.0, stream_);
    MPQC_HIP_CHECK(hipMemcpy2DAsync(Bp.p, 2 * V * d, Bcat.p, V * d, V * d, O * O * VK, hipMemcpyDeviceToDevice,
                                    stream_));
    DevBuf BcatT((size_t)(O * O * VK * V));
    permute_into(BcatT.p, e_b, Bcat.p, e_b, psw, 1.0, stream_);  // BcatT(j,k) = Bcat(k,j)
    MPQC_HIP_CHECK(hipMemcpy2DAsync(Bp.p + V, 2 * V * d, BcatT.p, V * d, V * d, O * O * VK,
                                    hipMemcpyDeviceToDevice, stream_));
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));  // temporaries are released on return
  }
  // split-precision forms: the W GEMMs read 16-bit planes of the operands in
  // k-contiguous layout (A as is; B transposed to [c][K] rows), split once here
  const bool sm = split_mode();
  const PlaneKind kind = plane_kind();
  std::unique_ptr<TPlanes> Ap, BTp, BT2p, BPp;
  std::unique_ptr<DevBuf> wsP;
  size_t wsP_bytes = 0;
  if (sm) {
    Ap.reset(new TPlanes(kind, O * V * V * VK));
    BTp.reset(new TPlanes(kind, O * O * V * VK));
    BT2p.reset(new TPlanes(kind, O * O * V * VK));
    BPp.reset(new TPlanes(kind, O * O * 2 * V * VK));
    auto split1 = [&](const Planes& out, const long out_ext[4], const double* src, const long in_ext[4],
                      const int perm[4]) {
      TTerm t;
      t.p = src;
      permuted_strides(in_ext, perm, t.s);
      t.alpha = 1.0;
      t.vol = in_ext[0] * in_ext[1] * in_ext[2] * in_ext[3];
      split_to_planes(kind, out, out_ext, &t, 1, partials_, stream_);
    };
    const long e_a[4] = {1, O, V * V, VK};
    const int pid[4] = {0, 1, 2, 3};
    split1(Ap->P, e_a, Acat.p, e_a, pid);
    const long e_bt[4] = {VK, O, O, V}, e_bto[4] = {O, O, V, VK};
    const int p_bt[4] = {1, 2, 3, 0}, p_bt2[4] = {2, 1, 3, 0};
    split1(BTp->P, e_bto, Bt.p, e_bt, p_bt);    // [o2][o3][c][K]
    split1(BT2p->P, e_bto, Bt.p, e_bt, p_bt2);  // [o3][o2][c][K]
    const long e_bp[4] = {O, O, VK, 2 * V}, e_bpo[4] = {O, O, 2 * V, VK};
    const int p_bp[4] = {0, 1, 3, 2};
    split1(BPp->P, e_bpo, Bp.p, e_bp, p_bp);    // [j][k][(t,c)][K]
    for (long nb = 1; nb <= nbmax; ++nb) {
      wsP_bytes = std::max(wsP_bytes, gemm_planes_workspace_bytes((int)(nb * V * V), (int)(2 * V), (int)VK, 1));
      wsP_bytes = std::max(wsP_bytes, gemm_planes_workspace_bytes((int)(V * V), (int)(nb * V), (int)VK, 1));
    }
    wsP.reset(new DevBuf((wsP_bytes + sizeof(double) - 1) / sizeof(double)));
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  }
  auto plane_gemm = [&](int M, int N, const Planes& A, const Planes& B, double* C, long ldc) {
    PlaneGemmDesc dg;
    dg.kind = kind;
    dg.M = M; dg.N = N; dg.K = (int)VK; dg.alpha = 1.0; dg.beta = 0.0;
    dg.A = A; dg.lda = VK; dg.sA = 0; dg.B = B; dg.ldb = VK; dg.sB = 0;
    dg.C = C; dg.ldc = ldc; dg.sC = 0; dg.batch = 1;
    gemm_planes(dg, wsP->p, wsP_bytes, stream_);
  };
  // W buffers: orderings (i,j,k), (i,k,j) interleaved as [(i,a,b)][2][c]; the other four as
  // [(a,b)][i][c] with row stride nbmax v
  double* W1 = W.p;                       // 2 nbmax v^3
  double* Wg[4];                          // (j,k,i), (j,i,k), (k,j,i), (k,i,j)
  for (int q = 0; q < 4; ++q) Wg[q] = W.p + (2 + q) * nbmax * v3;
  const long ldw = nbmax * V;
  long nchunks = 0;
  for (int k = 0; k < o; ++k)
    for (int j = 0; j <= k; ++j) nchunks += ceil_div(j + 1, nbmax);
  DevBuf partials((size_t)(ncube * nbmax)), acc((size_t)nchunks);
  long slot = 0;
  for (int k = 0; k < o; ++k)
    for (int j = 0; j <= k; ++j)
      for (int i0 = 0; i0 <= j; i0 += (int)nbmax) {
        const int nb = (int)std::min<long>(nbmax, j + 1 - i0);
        // (i,j,k) and (i,k,j): C[(i,a,b)][(t,c)] = Acat[(i,a,b)][K] Bp[j][k][K][(t,c)]
        if (sm)
          plane_gemm((int)(nb * V * V), (int)(2 * V), Ap->at((long)i0 * V * V * VK),
                     BPp->at(((long)j * O + k) * 2 * V * VK), W1, 2 * V);
        else
          mpqc_hip::gemm(make_gemm(false, false, (int)(nb * V * V), (int)(2 * V), (int)VK, 1.0,
                                   Acat.p + (long)i0 * V * V * VK, VK, Bp.p + ((long)j * O + k) * VK * 2 * V,
                                   2 * V, 0.0, W1, 2 * V), GemmPrecision::F64, ws_, stream_);
        // fixed first index o1 in {j, k}: C[(a,b)][(i,c)] = Acat[o1][(a,b)][K] B[K][(i,c)]
        struct G4 { int o1; const double* B; };
        const G4 g4[4] = {{j, Bt.p + ((long)k * O + i0) * V},    // (j,k,i): Bcat(k, i)
                          {j, Bt2.p + ((long)k * O + i0) * V},   // (j,i,k): Bcat(i, k)
                          {k, Bt.p + ((long)j * O + i0) * V},    // (k,j,i): Bcat(j, i)
                          {k, Bt2.p + ((long)j * O + i0) * V}};  // (k,i,j): Bcat(i, j)
        // the same four as planes: B rows (i, c) of BTp[o2][.] / BT2p[o3][.]
        const long pb[4] = {((long)k * O + i0) * V * VK, ((long)k * O + i0) * V * VK,
                            ((long)j * O + i0) * V * VK, ((long)j * O + i0) * V * VK};
        for (int q = 0; q < 4; ++q) {
          if (sm)
            plane_gemm((int)(V * V), (int)(nb * V), Ap->at((long)g4[q].o1 * V * V * VK),
                       (q % 2 == 0 ? BTp : BT2p)->at(pb[q]), Wg[q], ldw);
          else
            mpqc_hip::gemm(make_gemm(false, false, (int)(V * V), (int)(nb * V), (int)VK, 1.0,
                                     Acat.p + (long)g4[q].o1 * V * V * VK, VK, g4[q].B, O * O * V, 0.0, Wg[q],
                                     ldw), GemmPrecision::F64, ws_, stream_);
        }
        TriplesArgs g;
        // orderings by S3 index: 0 (i,j,k) 1 (i,k,j) 2 (j,i,k) 3 (j,k,i) 4 (k,i,j) 5 (k,j,i)
        g.Wb[0] = W1;     g.Wz[0] = 2 * v3; g.Wr[0] = 2 * V;
        g.Wb[1] = W1 + V; g.Wz[1] = 2 * v3; g.Wr[1] = 2 * V;
        g.Wb[2] = Wg[1];  g.Wz[2] = V;      g.Wr[2] = ldw;
        g.Wb[3] = Wg[0];  g.Wz[3] = V;      g.Wr[3] = ldw;
        g.Wb[4] = Wg[3];  g.Wz[4] = V;      g.Wr[4] = ldw;
        g.Wb[5] = Wg[2];  g.Wz[5] = V;      g.Wr[5] = ldw;
        g.G = Gijab_; g.t1 = t1_; g.eo = eo_; g.ev = ev_;
        g.o = o; g.v = v; g.i0 = i0; g.j = j; g.k = k;
        hipLaunchKernelGGL(triples_energy_kernel, dim3((unsigned)ncube, nb), dim3(kTThreads), 0, stream_, g,
                           partials.p);
        check_launch("triples_energy");
        hipLaunchKernelGGL(sum_to_slot, dim3(1), dim3(256), 0, stream_, partials.p, ncube * nb, acc.p, slot);
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
