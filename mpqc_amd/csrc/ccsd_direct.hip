// AO integral-direct CCSD residual (method "direct", ccsd.h:73).
//
// The reference keeps the MO blocks with at most three virtual indices and
// evaluates the AO integrals on the fly (AOFactory::compute_direct, ccsd.h:246-247);
// every iteration forms (compute_u2_u11, ccsd.h:573-603)
//   U(p,r,i,j) = [(t2(a,b,i,j) Ca(q,a)) Ca(s,b) + tc(i,q) tc(j,s)] (p q|r s),  tc(i,q) = Ca(q,c) t1(c,i)
// which replaces the g_iabc_bar tau term of r1 (ccsd_r1_r2.h:116-117) and the
// particle-particle ladder with its Giabc terms in r2 (:282-286).  Here the AO
// integrals stay resident in HBM, laid out once as X((p,r),(q,s)) = (pq|rs), so U
// is one GEMM with M = K = nbf^2, N = o^2 over the AO-basis tau
// (Ca tau Ca^T = the bracket above, since tau = t2 + t1 t1), and the MO transforms
// of U are batched GEMMs.
#include <stdexcept>

#include "ccsd_ctx.hpp"

namespace mpqc_hip {

void CCSDContext::set_direct_ao(int nbf, const double* ao_eri, const double* c_occ, const double* c_vir,
                                int memspace) {
  if (nbf <= 0 || !ao_eri || !c_occ || !c_vir) throw std::invalid_argument("set_direct_ao: null or empty input");
  if (memspace != 0 && memspace != 1) throw std::invalid_argument("memspace must be host or device");
  const long n = nbf, O = o_, V = v_, oo = O * O;
  if (O + V > n) throw std::invalid_argument("more active orbitals than basis functions");
  if (n * n * n * n >= (1L << 31)) throw std::invalid_argument("nbf^4 must be < 2^31 for the direct method");
  if (Xao_ && nbf != nbf_) throw std::invalid_argument("the AO basis size is fixed for a context");
  if (!Xao_) {
    nbf_ = nbf;
    Xao_ = alloc((size_t)n * n * n * n);
    Ci_ = alloc((size_t)n * O);
    Ca_ = alloc((size_t)n * V);
    Uao_ = alloc((size_t)n * n * oo);
    TauAO_ = alloc((size_t)n * n * oo);
    T1ao_ = alloc((size_t)n * V * oo);
    Vr_ = alloc((size_t)n * n * oo);
    Yr_ = alloc((size_t)n * O);
    Dct_ = alloc((size_t)n * V);
    CaD_ = alloc((size_t)n * V);
    G1ao_ = alloc((size_t)n * V * oo);
    G2ao_ = alloc((size_t)n * V * oo);
  }
  const hipMemcpyKind kind = memspace ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const long n4 = n * n * n * n;
  double* src = nullptr;
  if (!memspace) {
    if (hipMalloc(&src, (size_t)n4 * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      throw std::bad_alloc();
    }
  }
  try {
    if (!memspace)
      MPQC_HIP_CHECK(hipMemcpyAsync(src, ao_eri, (size_t)n4 * sizeof(double), hipMemcpyHostToDevice, stream_));
    // X(p,r,q,s) = (pq|rs)
    TTerm t;
    t.p = memspace ? ao_eri : src;
    const long ext[4] = {n, n, n, n};
    const int pm[4] = {0, 2, 1, 3};
    permuted_strides(ext, pm, t.s);
    t.vol = n4;
    tadd(Xao_, ext, 0.0, &t, 1, stream_);
    MPQC_HIP_CHECK(hipMemcpyAsync(Ci_, c_occ, (size_t)n * O * sizeof(double), kind, stream_));
    MPQC_HIP_CHECK(hipMemcpyAsync(Ca_, c_vir, (size_t)n * V * sizeof(double), kind, stream_));
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  } catch (...) {
    if (src) (void)hipFree(src);
    throw;
  }
  if (src) MPQC_HIP_CHECK(hipFree(src));
  direct_ = true;
  df_ = false;
  prepared_ = false;
}

// U((p,r),(i,j)) = X((p,r),(q,s)) [Ca tau Ca^T]((q,s),(i,j))
void CCSDContext::direct_u() {
  const int n = nbf_, o = o_, v = v_;
  const long V = v, oo = (long)o * o, N = n;
  // T1ao(q,b,i,j) = Ca(q,a) tau(a,(b,i,j))
  gemm(false, false, n, (int)(V * oo), v, 1.0, Ca_, V, tau_, V * oo, 0.0, T1ao_, V * oo, 0, 0, 0, 1, "u:ca_tau");
  // TauAO(q,s,i,j) = Ca(s,b) T1ao(q,b,(i,j)), one product per q
  gemm(false, false, n, (int)oo, v, 1.0, Ca_, V, T1ao_, oo, 0.0, TauAO_, oo, 0, V * oo, N * oo, n, "u:ca_tau_ca");
  // ccsd.h:591-593
  gemm(false, false, n * n, (int)oo, n * n, 1.0, Xao_, N * N, TauAO_, oo, 0.0, Uao_, oo, 0, 0, 0, 1, "u:ao_eri_tau");
}

// ccsd_r1_r2.h:116-117  r1(a,i) += (2 U(p,r,k,i) - U(p,r,i,k)) Ci(p,k) Ca(r,a)
void CCSDContext::direct_r1() {
  const int n = nbf_, o = o_, v = v_;
  const long O = o, N = n;
  // Vr(r,i,p,k) = 2 U(p,r,k,i) - U(p,r,i,k)
  perm_add(Vr_, {N, O, N, O}, 0.0,
           {{Uao_, {N, N, O, O}, {1, 3, 0, 2}, 2.0}, {Uao_, {N, N, O, O}, {1, 2, 0, 3}, -1.0}});
  // Yr(r,i) = Vr((r,i),(p,k)) Ci((p,k))
  gemm(false, false, n * o, 1, n * o, 1.0, Vr_, N * O, Ci_, 1, 0.0, Yr_, 1, 0, 0, 0, 1, "r1:u_ci");
  // r1(a,i) += Ca(r,a) Yr(r,i)
  gemm(true, false, v, o, n, 1.0, Ca_, v, Yr_, O, 1.0, r1_, O, 0, 0, 0, 1, "r1:u_ca");
}

// ccsd_r1_r2.h:282-286
//   b(a,b,i,j) = (U(p,r,i,j) Ca(r,b) - Ci(r,k) t1(b,k) U(p,r,i,j)) Ca(p,a) - U(p,r,i,j) Ci(p,k) Ca(r,b) t1(a,k)
//              = Ca(p,a) [U (Ca - D)](p,b,i,j) - D(p,a) [U Ca](p,b,i,j),   D(r,b) = Ci(r,k) t1(b,k)
// added to r2 after the symmetrisation, as the ladder it replaces
void CCSDContext::direct_r2() {
  const int n = nbf_, o = o_, v = v_;
  const long V = v, O = o, oo = O * O, N = n;
  const std::array<int, 4> I{0, 1, 2, 3};
  gemm(false, true, n, v, o, 1.0, Ci_, O, t1_, O, 0.0, Dct_, V, 0, 0, 0, 1, "r2:u_ci_t1");
  perm_add(CaD_, {1, 1, N, V}, 0.0, {{Ca_, {1, 1, N, V}, I, 1.0}, {Dct_, {1, 1, N, V}, I, -1.0}});
  // G1(p,b,(i,j)) = (Ca - D)(r,b) U(p,r,(i,j)), G2 with Ca; one product per p
  gemm(true, false, v, (int)oo, n, 1.0, CaD_, V, Uao_, oo, 0.0, G1ao_, oo, 0, N * oo, V * oo, n, "r2:u_cad");
  gemm(true, false, v, (int)oo, n, 1.0, Ca_, V, Uao_, oo, 0.0, G2ao_, oo, 0, N * oo, V * oo, n, "r2:u_ca");
  gemm(true, false, v, (int)(V * oo), n, 1.0, Ca_, V, G1ao_, V * oo, 1.0, r2_, V * oo, 0, 0, 0, 1, "r2:u_b1");
  gemm(true, false, v, (int)(V * oo), n, -1.0, Dct_, V, G2ao_, V * oo, 1.0, r2_, V * oo, 0, 0, 0, 1, "r2:u_b2");
}

}  // namespace mpqc_hip
