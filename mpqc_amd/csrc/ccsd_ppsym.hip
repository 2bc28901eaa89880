// Particle-particle ladder (and the hole-hole intermediate's G tau term) on
// pair-symmetric blocks, in the FP64 and split forms.
//
// b(a,b,i,j) = sum_cd Gabcd(a,b,c,d) tau(c,d,i,j) (ccsd_r1_r2.h:275).  With the
// integrals' symmetry Gabcd(b,a,d,c) = Gabcd(a,b,c,d) (both are (ac|bd)) and no
// assumption on tau:
//   b(a,b,i,j) = 1/2 [P(ab,ij) + s(a,b) M(ab,ij)]
//   P(ab,ij) = sum_{c<=d} V+(ab,cd) S(ij,cd),   V+ = G(ab,cd) + G(ab,dc),
//              S(ij,cd) = f_cd [tau(c,d,i,j) + tau(d,c,i,j)],  f_cc = 1/2, else 1
//   M(ab,ij) = sum_{c<d} V-(ab,cd) A(ij,cd),    V- = G(ab,cd) - G(ab,dc),
//              A(ij,cd) = tau(c,d,i,j) - tau(d,c,i,j)
// with V+ symmetric and V- antisymmetric under a<->b, so rows a<=b (P) and a<b (M)
// carry everything; s(a,b) = +1 for a<b, -1 for a>b.  The two GEMMs
// (v(v+1)/2 x o^2 x v(v+1)/2 and v(v-1)/2 x o^2 x v(v-1)/2) hold half the FLOPs of
// the full contraction (2 o^2 v^4).  W(k,l,i,j) += Gijab(k,l,c,d) tau(c,d,i,j)
// (ccsd_intermediates.h:182-193) has the same form with Gijab(l,k,d,c) = Gijab(k,l,c,d),
// rows k<=l, on the same S and A.  The context checks each symmetry once per set of
// integrals and keeps the full contraction where it does not hold.
//
// Pairs are numbered row-major over the upper triangle: (p,q), p<=q < n has index
// p n - p(p-1)/2 + q - p; (p,q), p<q < n has index p (n-1) - p(p-1)/2 + q - p - 1.
// The packed operands' rows are padded with zeros to a multiple of 64 columns (the
// plane GEMM's vector path wants K and the row pitch in multiples of 8).
#include <algorithm>
#include <stdexcept>
#include <vector>

#include "ccsd_ctx.hpp"

namespace mpqc_hip {

namespace {

constexpr int kSymBlock = 256;

__device__ __forceinline__ long sym_pair(long p, long q, long n) { return p * n - p * (p - 1) / 2 + (q - p); }
__device__ __forceinline__ long anti_pair(long p, long q, long n) {
  return p * (n - 1) - p * (p - 1) / 2 + (q - p - 1);
}

// out(R, C) = G(a,b,c,d) + sign G(a,b,d,c) for row pair R = (a,b) (a, b < p) and
// column pair C = (c,d) (c, d < q) of G with extents (p, p, q, q) (tables of packed
// (x, y) pairs), zero in the padding columns
__global__ __launch_bounds__(kSymBlock) void pack_pairs_kernel(double* __restrict__ out,
                                                               const double* __restrict__ G, long p, long q,
                                                               const int2* __restrict__ rows, long nrows,
                                                               const int2* __restrict__ cols, long ncols, long ld,
                                                               double sign) {
  const long total = nrows * ld;
  for (long e = (long)blockIdx.x * kSymBlock + threadIdx.x; e < total; e += (long)gridDim.x * kSymBlock) {
    const long R = e / ld, C = e - R * ld;
    double x = 0.0;
    if (C < ncols) {
      const int2 ab = rows[R], cd = cols[C];
      const long row = ((long)ab.x * p + ab.y) * q;
      x = G[(row + cd.x) * q + cd.y] + sign * G[(row + cd.y) * q + cd.x];
    }
    out[e] = x;
  }
}

// out(C, ij) = f [tau(c,d,i,j) + sign tau(d,c,i,j)] for column pair C = (c,d), read
// straight from tau(a,b,i,j) (coalesced along ij); rows npairs..nrows-1 are zero (the
// padding of the split forms, whose planes take the transpose).  The K x N operand of
// the FP64 form's NN GEMM.
__global__ __launch_bounds__(kSymBlock) void pack_tau_rows_kernel(double* __restrict__ out,
                                                                  const double* __restrict__ tau, long oo, long v,
                                                                  const int2* __restrict__ pairs, long npairs,
                                                                  long nrows, double sign) {
  const long total = nrows * oo;
  for (long e = (long)blockIdx.x * kSymBlock + threadIdx.x; e < total; e += (long)gridDim.x * kSymBlock) {
    const long C = e / oo, ij = e - C * oo;
    double x = 0.0;
    if (C < npairs) {
      const int2 cd = pairs[C];
      x = tau[((long)cd.x * v + cd.y) * oo + ij] + sign * tau[((long)cd.y * v + cd.x) * oo + ij];
      if (cd.x == cd.y) x *= 0.5;
    }
    out[e] = x;
  }
}

// pack_tau_rows_kernel that also writes the block's max |out| into part[blockIdx.x]
__global__ __launch_bounds__(kSymBlock) void pack_tau_rows_max_kernel(double* __restrict__ out,
                                                                      const double* __restrict__ tau, long oo,
                                                                      long v, const int2* __restrict__ pairs,
                                                                      long npairs, long nrows, double sign,
                                                                      double* __restrict__ part) {
  const long total = nrows * oo;
  double m = 0.0;
  for (long e = (long)blockIdx.x * kSymBlock + threadIdx.x; e < total; e += (long)gridDim.x * kSymBlock) {
    const long C = e / oo, ij = e - C * oo;
    double x = 0.0;
    if (C < npairs) {
      const int2 cd = pairs[C];
      x = tau[((long)cd.x * v + cd.y) * oo + ij] + sign * tau[((long)cd.y * v + cd.x) * oo + ij];
      if (cd.x == cd.y) x *= 0.5;
    }
    out[e] = x;
    m = fmax(m, fabs(x));
  }
  __shared__ double sh[kSymBlock];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = kSymBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// *out = max of part[0..n)
__global__ __launch_bounds__(kSymBlock) void sym_max_stage2(const double* __restrict__ part, int n,
                                                            double* __restrict__ out) {
  double m = 0.0;
  for (int b = threadIdx.x; b < n; b += kSymBlock) m = fmax(m, part[b]);
  __shared__ double sh[kSymBlock];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = kSymBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// r2(a,b,i,j) += 1/2 [P(ab,ij) + s(a,b) M(ab,ij)]
__global__ __launch_bounds__(kSymBlock) void unpack_ladder_kernel(double* __restrict__ r2,
                                                                  const double* __restrict__ P,
                                                                  const double* __restrict__ M, long oo,
                                                                  long v) {
  const long total = v * v * oo;
  for (long e = (long)blockIdx.x * kSymBlock + threadIdx.x; e < total; e += (long)gridDim.x * kSymBlock) {
    const long ij = e % oo, ab = e / oo, b = ab % v, a = ab / v;
    const long p = a < b ? a : b, q = a < b ? b : a;
    double x = P[sym_pair(p, q, v) * oo + ij];
    if (a != b) {
      const double m = M[anti_pair(p, q, v) * oo + ij];
      x += a < b ? m : -m;
    }
    r2[e] += 0.5 * x;
  }
}

// per block, over G with extents (p, p, q, q): max |G(a,b,c,d) - G(b,a,d,c)| and
// max |G(a,b,c,d)| -> part[2 blk], part[2 blk + 1]
__global__ __launch_bounds__(kSymBlock) void pair_asymmetry_kernel(double* __restrict__ part,
                                                                   const double* __restrict__ G, long p, long q) {
  __shared__ double sd[kSymBlock], sg[kSymBlock];
  const long total = p * p * q * q;
  double md = 0.0, mg = 0.0;
  for (long e = (long)blockIdx.x * kSymBlock + threadIdx.x; e < total; e += (long)gridDim.x * kSymBlock) {
    const long d = e % q, t0 = e / q, c = t0 % q, t1 = t0 / q, b = t1 % p, a = t1 / p;
    const double x = G[e];
    md = fmax(md, fabs(x - G[((b * p + a) * q + d) * q + c]));
    mg = fmax(mg, fabs(x));
  }
  sd[threadIdx.x] = md;
  sg[threadIdx.x] = mg;
  __syncthreads();
  for (int w = kSymBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + w]);
      sg[threadIdx.x] = fmax(sg[threadIdx.x], sg[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd[0];
    part[2 * blockIdx.x + 1] = sg[0];
  }
}

// per block, over G with extents (n0, n1, n2, n1): max |G(i,x,a,y) - G(i,y,a,x)| and
// max |G(i,x,a,y)| -> part[2 blk], part[2 blk + 1]
__global__ __launch_bounds__(kSymBlock) void swap13_asymmetry_kernel(double* __restrict__ part,
                                                                     const double* __restrict__ G, long n0,
                                                                     long n1, long n2) {
  __shared__ double sd[kSymBlock], sg[kSymBlock];
  const long total = n0 * n1 * n2 * n1;
  double md = 0.0, mg = 0.0;
  for (long e = (long)blockIdx.x * kSymBlock + threadIdx.x; e < total; e += (long)gridDim.x * kSymBlock) {
    const long y = e % n1, t0 = e / n1, a = t0 % n2, t1 = t0 / n2, x = t1 % n1, i = t1 / n1;
    const double g = G[e];
    md = fmax(md, fabs(g - G[((i * n1 + y) * n2 + a) * n1 + x]));
    mg = fmax(mg, fabs(g));
  }
  sd[threadIdx.x] = md;
  sg[threadIdx.x] = mg;
  __syncthreads();
  for (int w = kSymBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + w]);
      sg[threadIdx.x] = fmax(sg[threadIdx.x], sg[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd[0];
    part[2 * blockIdx.x + 1] = sg[0];
  }
}

int sym_grid(long n) { return (int)std::max<long>(1, std::min<long>(ceil_div(n, kSymBlock), 65536)); }

std::vector<int2> pair_table(long n, bool strict) {
  std::vector<int2> t;
  for (long p = 0; p < n; ++p)
    for (long q = strict ? p + 1 : p; q < n; ++q) t.push_back(make_int2((int)p, (int)q));
  return t;
}

long pad64(long n) { return std::max<long>(64, (n + 63) / 64 * 64); }

}  // namespace

// whether Gabcd has the (ab)(cd) pair symmetry the form relies on, to 1e-12 of its
// largest element (integrals from a transform hold it to round-off; a perturbation
// that breaks it sends the ladder back to the full contraction)
bool CCSDContext::pp_symmetric_usable() {
  if (!pp_symmetric_ || !Gabcd_) return false;
  if (pp_checked_) return pp_sym_ok_;
  // whether G with extents (p, p, q, q) holds G(b,a,d,c) = G(a,b,c,d) to 1e-12 of its largest element
  auto symmetric = [&](const double* G, long p, long q) {
    const int blocks = (int)std::max<long>(1, std::min<long>(ceil_div(p * p * q * q, kSymBlock), kReducePartials / 2));
    hipLaunchKernelGGL(pair_asymmetry_kernel, dim3(blocks), dim3(kSymBlock), 0, stream_, partials_, G, p, q);
    check_launch("pair_asymmetry");
    std::vector<double> h(2 * (size_t)blocks);
    MPQC_HIP_CHECK(hipMemcpyAsync(h.data(), partials_, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
    double md = 0.0, mg = 0.0;
    for (int b = 0; b < blocks; ++b) {
      md = std::max(md, h[2 * b]);
      mg = std::max(mg, h[2 * b + 1]);
    }
    return md <= 1e-12 * mg;
  };
  pp_sym_ok_ = symmetric(Gabcd_, v_, v_);
  // Gijab(l,k,d,c) = Gijab(k,l,c,d): the hole-hole intermediate's G tau term in the same form
  w_sym_ok_ = pp_sym_ok_ && symmetric(Gijab_, o_, v_);
  pp_checked_ = true;
  return pp_sym_ok_;
}

// whether Giabc(i,c,a,b) = Giabc(i,b,a,c) (<ic|ab> = (ia|cb) = (ia|bc) = <ib|ac> for real
// orbitals) to 1e-12 of its largest element: then the T1 dressing term
// Giabc(i,c,a,b) t1(c,j) of Rd (ccsd_r1_r2.h:207) is the j term's Zj(i,b,a,j) =
// Giabc(i,b,a,d) t1(d,j) (:240), and Rd takes it from there instead of a contraction
// that reads Giabc again
bool CCSDContext::giabc_symmetric_usable() {
  if (iabc_checked_) return iabc_sym_ok_;
  const long total = (long)o_ * v_ * v_ * v_;
  const int blocks = (int)std::max<long>(1, std::min<long>(ceil_div(total, kSymBlock), kReducePartials / 2));
  hipLaunchKernelGGL(swap13_asymmetry_kernel, dim3(blocks), dim3(kSymBlock), 0, stream_, partials_, Giabc_,
                     (long)o_, (long)v_, (long)v_);
  check_launch("swap13_asymmetry");
  std::vector<double> h(2 * (size_t)blocks);
  MPQC_HIP_CHECK(hipMemcpyAsync(h.data(), partials_, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  double md = 0.0, mg = 0.0;
  for (int b = 0; b < blocks; ++b) {
    md = std::max(md, h[2 * b]);
    mg = std::max(mg, h[2 * b + 1]);
  }
  iabc_sym_ok_ = md <= 1e-12 * mg;
  iabc_checked_ = true;
  return iabc_sym_ok_;
}

// pair tables, packed V+ / V- (FP64) and the per-iteration buffers
void CCSDContext::pack_pp_symmetric() {
  const long O = o_, V = v_, oo = O * O;
  nvs_ = V * (V + 1) / 2;
  nva_ = V * (V - 1) / 2;
  nos_ = O * (O + 1) / 2;
  noa_ = O * (O - 1) / 2;
  ldvs_ = pad64(nvs_);
  ldva_ = pad64(nva_);
  if (!pv_sym_) {
    auto upload_pairs = [&](long n, bool strict) {
      const std::vector<int2> t = pair_table(n, strict);
      int2* d = reinterpret_cast<int2*>(alloc(std::max<size_t>(t.size(), 1)));
      if (!t.empty())
        MPQC_HIP_CHECK(hipMemcpyAsync(d, t.data(), t.size() * sizeof(int2), hipMemcpyHostToDevice, stream_));
      return d;
    };
    pv_sym_ = upload_pairs(V, false);
    pv_anti_ = upload_pairs(V, true);
    po_sym_ = upload_pairs(O, false);
    po_anti_ = upload_pairs(O, true);
    Vpk_ = alloc((size_t)nvs_ * ldvs_);
    Vmk_ = alloc((size_t)std::max<long>(nva_, 1) * ldva_);
    Spk_ = alloc((size_t)oo * ldvs_);
    Apk_ = alloc((size_t)oo * ldva_);
    Ppk_ = alloc((size_t)nvs_ * oo);
    Mpk_ = alloc((size_t)std::max<long>(nva_, 1) * oo);
    Gpk_ = alloc((size_t)nos_ * ldvs_);
    Gmk_ = alloc((size_t)std::max<long>(noa_, 1) * ldva_);
    WPk_ = alloc((size_t)nos_ * oo);
    WMk_ = alloc((size_t)std::max<long>(noa_, 1) * oo);
  }
  hipLaunchKernelGGL(pack_pairs_kernel, dim3(sym_grid(nvs_ * ldvs_)), dim3(kSymBlock), 0, stream_, Vpk_, Gabcd_, V, V,
                     pv_sym_, nvs_, pv_sym_, nvs_, ldvs_, 1.0);
  check_launch("pack_pairs");
  if (nva_ > 0) {
    hipLaunchKernelGGL(pack_pairs_kernel, dim3(sym_grid(nva_ * ldva_)), dim3(kSymBlock), 0, stream_, Vmk_, Gabcd_, V,
                       V, pv_anti_, nva_, pv_anti_, nva_, ldva_, -1.0);
    check_launch("pack_pairs");
  }
  if (w_sym_ok_) {  // G+/-(kl, cd) of Gijab, rows k<=l (k<l)
    hipLaunchKernelGGL(pack_pairs_kernel, dim3(sym_grid(nos_ * ldvs_)), dim3(kSymBlock), 0, stream_, Gpk_, Gijab_, O,
                       V, po_sym_, nos_, pv_sym_, nvs_, ldvs_, 1.0);
    check_launch("pack_pairs");
    if (noa_ > 0 && nva_ > 0) {
      hipLaunchKernelGGL(pack_pairs_kernel, dim3(sym_grid(noa_ * ldva_)), dim3(kSymBlock), 0, stream_, Gmk_, Gijab_,
                         O, V, po_anti_, noa_, pv_anti_, nva_, ldva_, -1.0);
      check_launch("pack_pairs");
    }
  }
  pp_packed_ = true;
}

// the split forms' V+ / V- planes
void CCSDContext::prepare_pp_symmetric() {
  pack_pp_symmetric();
  const long oo = (long)o_ * o_;
  if (!Vp_P_.p[0]) {  // (re-allocated after a change of the plane kind)
    Vp_P_ = alloc_planes(nvs_ * ldvs_);
    Vm_P_ = alloc_planes(std::max<long>(nva_, 1) * ldva_);
    Sp_P_ = alloc_planes(oo * ldvs_);
    Ap_P_ = alloc_planes(oo * ldva_);
    Gp_P_sym_ = alloc_planes(nos_ * ldvs_);
    Gm_P_sym_ = alloc_planes(std::max<long>(noa_, 1) * ldva_);
  }
  const std::array<int, 4> I{0, 1, 2, 3};
  split(Vp_P_, {1, 1, nvs_, ldvs_}, {{Vpk_, {1, 1, nvs_, ldvs_}, I, 1.0}});
  if (nva_ > 0) split(Vm_P_, {1, 1, nva_, ldva_}, {{Vmk_, {1, 1, nva_, ldva_}, I, 1.0}});
  if (w_sym_ok_) {
    split(Gp_P_sym_, {1, 1, nos_, ldvs_}, {{Gpk_, {1, 1, nos_, ldvs_}, I, 1.0}});
    if (noa_ > 0 && nva_ > 0) split(Gm_P_sym_, {1, 1, noa_, ldva_}, {{Gmk_, {1, 1, noa_, ldva_}, I, 1.0}});
  }
}

// S(pair, ij) and A(pair, ij) from tau (FP64 form: the NN GEMMs' K x N operands; split
// forms: padded to a multiple of 64 pairs and transposed into (ij, pair) planes)
void CCSDContext::pack_tau_pairs() {
  const long O = o_, V = v_, oo = O * O;
  const bool sm = split_mode();
  if (!sm && !pp_packed_) pack_pp_symmetric();
  const long rs = sm ? ldvs_ : nvs_, ra = sm ? ldva_ : nva_;
  // fp16 planes: the pack also yields the exact max |S| and |A| their scales need (cache
  // slots 5 and 6), in place of a max pass of each split
  const bool rec = sm && prec_ == GemmPrecision::F16X3 && amax_.dev;
  timed("r2:pp_sym_pack", 0.0, [&] {
    if (rec) {
      const int gs = (int)std::min<long>(ceil_div(rs * oo, kSymBlock), 4096);
      hipLaunchKernelGGL(pack_tau_rows_max_kernel, dim3(gs), dim3(kSymBlock), 0, stream_, Spk_, tau_, oo, V, pv_sym_,
                         nvs_, rs, 1.0, partials_);
      hipLaunchKernelGGL(sym_max_stage2, dim3(1), dim3(kSymBlock), 0, stream_, partials_, gs, amax_.dev + 5);
    } else {
      hipLaunchKernelGGL(pack_tau_rows_kernel, dim3(sym_grid(rs * oo)), dim3(kSymBlock), 0, stream_, Spk_, tau_, oo,
                         V, pv_sym_, nvs_, rs, 1.0);
    }
    check_launch("pack_tau_rows");
    if (nva_ > 0) {
      if (rec) {
        const int ga = (int)std::min<long>(ceil_div(ra * oo, kSymBlock), 4096);
        hipLaunchKernelGGL(pack_tau_rows_max_kernel, dim3(ga), dim3(kSymBlock), 0, stream_, Apk_, tau_, oo, V,
                           pv_anti_, nva_, ra, -1.0, partials_);
        hipLaunchKernelGGL(sym_max_stage2, dim3(1), dim3(kSymBlock), 0, stream_, partials_, ga, amax_.dev + 6);
      } else {
        hipLaunchKernelGGL(pack_tau_rows_kernel, dim3(sym_grid(ra * oo)), dim3(kSymBlock), 0, stream_, Apk_, tau_,
                           oo, V, pv_anti_, nva_, ra, -1.0);
      }
      check_launch("pack_tau_rows");
    }
  });
  if (rec) {
    amax_.ptr[5] = Spk_;
    amax_.valid[5] = true;
    amax_.ptr[6] = nva_ > 0 ? Apk_ : nullptr;
    amax_.valid[6] = nva_ > 0;
  }
  if (sm) {
    const std::array<int, 4> T{0, 1, 3, 2};
    split(Sp_P_, {1, 1, oo, ldvs_}, {{Spk_, {1, 1, ldvs_, oo}, T, 1.0}});
    if (nva_ > 0) split(Ap_P_, {1, 1, oo, ldva_}, {{Apk_, {1, 1, ldva_, oo}, T, 1.0}});
  }
}

// out(p,q,i,j) += 1/2 [P(pq,ij) + s(p,q) M(pq,ij)] from the two halves' products
void CCSDContext::pair_products(const double* Pk, long lp, const Planes& Pp, const double* Mk, long lm,
                                const Planes& Mp, long nrs, long nra, long n, double* out, double* P, double* M,
                                const char* tag_s, const char* tag_a, const char* tag_u) {
  const long oo = (long)o_ * o_;
  if (split_mode()) {
    gemmP((int)nrs, (int)oo, (int)ldvs_, 1.0, Pp, ldvs_, Sp_P_, ldvs_, 0.0, P, oo, tag_s);
    if (nra > 0 && nva_ > 0)
      gemmP((int)nra, (int)oo, (int)ldva_, 1.0, Mp, ldva_, Ap_P_, ldva_, 0.0, M, oo, tag_a);
  } else {
    gemm(false, false, (int)nrs, (int)oo, (int)nvs_, 1.0, Pk, lp, Spk_, oo, 0.0, P, oo, 0, 0, 0, 1, tag_s);
    if (nra > 0 && nva_ > 0)
      gemm(false, false, (int)nra, (int)oo, (int)nva_, 1.0, Mk, lm, Apk_, oo, 0.0, M, oo, 0, 0, 0, 1, tag_a);
  }
  if (nra > 0 && nva_ == 0) fill(M, nra * oo, 0.0, stream_);  // v = 1: no (c<d) pairs, M = 0
  timed(tag_u, 0.0, [&] {
    hipLaunchKernelGGL(unpack_ladder_kernel, dim3(sym_grid(n * n * oo)), dim3(kSymBlock), 0, stream_, out, P, M, oo,
                       n);
    check_launch("unpack_ladder");
  });
}

// :275 b(a,b,i,j) = tau(c,d,i,j) Gabcd(a,b,c,d), added to out (after pack_tau_pairs)
void CCSDContext::pp_ladder_symmetric(double* out) {
  pair_products(Vpk_, ldvs_, Vp_P_, Vmk_, ldva_, Vm_P_, nvs_, nva_, v_, out, Ppk_, Mpk_, "r2:pp_ladder_sym",
                "r2:pp_ladder_anti", "r2:pp_sym_unpack");
}

// ccsd_intermediates.h:182-193: W(k,l,i,j) += Gijab(k,l,c,d) tau(c,d,i,j) in the same
// pair form (rows k<=l; after pack_tau_pairs)
void CCSDContext::w_oooo_symmetric() {
  pair_products(Gpk_, ldvs_, Gp_P_sym_, Gmk_, ldva_, Gm_P_sym_, nos_, noa_, o_, W_, WPk_, WMk_, "r2:W_oooo_sym",
                "r2:W_oooo_anti", "r2:W_sym_unpack");
}

}  // namespace mpqc_hip
