// Integral preparation on the GPU: the AO -> MO transform that produces the
// cc::Integrals blocks of the CCSD residual.
//
// Reference: LCAOFactory::compute4 (src/mpqc/chemistry/qc/lcao/factory/
// lcao_factory.h:378-497).  Without density fitting it transforms the AO
// integrals one index at a time (quarter transforms, :421-493) and reads
// physical-notation blocks <pq|rs> = (pr|qs); with density fitting (:386-406)
// a four-index block is the product of two three-index factors over the fitting
// index, result("i,k,j,l") = left("q,i,j") * right("q,k,l").  The Fock blocks
// are <p|F|q> = C_p^T F C_q.
//
// Both entry points write the context's integral buffers exactly as
// set_integral would; temporaries are allocated for the call and released.
#include <stdexcept>

#include "ccsd_ctx.hpp"

namespace mpqc_hip {

namespace {

struct DevTmp {
  double* p = nullptr;
  explicit DevTmp(size_t n) {
    if (n == 0) return;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw std::bad_alloc();
    }
  }
  ~DevTmp() {
    if (p) (void)hipFree(p);
  }
  DevTmp(const DevTmp&) = delete;
  DevTmp& operator=(const DevTmp&) = delete;
};

// out[a,b,c,d] (extents n) = src[off + a*s0 + b*s1 + c*s2 + d*s3]
void strided_copy(double* out, const long n[4], const double* src, long off, const long s[4], long src_vol,
                  hipStream_t stream) {
  TTerm t;
  t.p = src + off;
  for (int d = 0; d < 4; ++d) t.s[d] = s[d];
  t.alpha = 1.0;
  t.vol = src_vol - off;
  // every read stays inside the source
  long last = off;
  for (int d = 0; d < 4; ++d) last += (n[d] - 1) * s[d];
  if (off < 0 || last >= src_vol) throw std::logic_error("strided_copy reads past its source");
  tadd(out, n, 0.0, &t, 1, stream);
}

}  // namespace

// C_act = [C_occ | C_vir] (nbf x (o+v), row-major) on the device.
static void gather_coefficients(double* Cact, int nbf, int o, int v, const double* c_occ, const double* c_vir,
                                int memspace, hipStream_t s) {
  const hipMemcpyKind kind = memspace ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const size_t pitch = (size_t)(o + v) * sizeof(double);
  MPQC_HIP_CHECK(hipMemcpy2DAsync(Cact, pitch, c_occ, (size_t)o * sizeof(double), (size_t)o * sizeof(double), nbf,
                                  kind, s));
  MPQC_HIP_CHECK(hipMemcpy2DAsync(Cact + o, pitch, c_vir, (size_t)v * sizeof(double), (size_t)v * sizeof(double),
                                  nbf, kind, s));
}

// Fia, Fij, Fab = blocks of C_act^T F C_act  (<i|F|a>, <i|F|j>, <a|F|b>)
void CCSDContext::transform_fock(int nbf, const double* fock, const double* Cact, int memspace) {
  const int o = o_, v = v_, nmo = o_ + v_;
  DevTmp Fdev(memspace ? 0 : (size_t)nbf * nbf), T((size_t)nbf * nmo), Fmo((size_t)nmo * nmo);
  const double* F = fock;
  if (!memspace) {
    MPQC_HIP_CHECK(hipMemcpyAsync(Fdev.p, fock, (size_t)nbf * nbf * sizeof(double), hipMemcpyHostToDevice, stream_));
    F = Fdev.p;
  }
  mpqc_hip::gemm(make_gemm(false, false, nbf, nmo, nbf, 1.0, F, nbf, Cact, nmo, 0.0, T.p, nmo),
                 GemmPrecision::F64, ws_, stream_);
  mpqc_hip::gemm(make_gemm(true, false, nmo, nmo, nbf, 1.0, Cact, nmo, T.p, nmo, 0.0, Fmo.p, nmo),
                 GemmPrecision::F64, ws_, stream_);
  const size_t d = sizeof(double), pitch = (size_t)nmo * d;
  MPQC_HIP_CHECK(hipMemcpy2DAsync(Fia_, v * d, Fmo.p + o, pitch, v * d, o, hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipMemcpy2DAsync(Fij_, o * d, Fmo.p, pitch, o * d, o, hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipMemcpy2DAsync(Fab_, v * d, Fmo.p + (size_t)o * nmo + o, pitch, v * d, v,
                                  hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  have_ints_ |= 0x7u;
}

// compute4 without density fitting (lcao_factory.h:421-493): four quarter
// transforms of (mu nu|la si) by C_act, each one GEMM that contracts the
// leading index and appends the new one, (mu nu la si) -> (nu la si p) -> ...
// -> (p q r s); then the physical blocks <PQ|RS>[a,b,c,d] = (p=P0+a, q=R0+c |
// r=Q0+b, s=S0+d) by strided copies.
void CCSDContext::transform_ao(int nbf, const double* eri, const double* fock, const double* c_occ,
                               const double* c_vir, int memspace) {
  if (nbf <= 0 || !eri || !fock || !c_occ || !c_vir) throw std::invalid_argument("transform_ao: null or empty input");
  if (memspace != 0 && memspace != 1) throw std::invalid_argument("memspace must be host or device");
  const long n = nbf, nmo = (long)o_ + v_, O = o_, V = v_;
  if (nmo > n) throw std::invalid_argument("more active orbitals than basis functions");
  if (n * n * n >= (1L << 31) || n * n * n * n >= (1L << 40))
    throw std::invalid_argument("nbf too large for the dense AO transform");
  const long n4 = n * n * n * n;
  DevTmp Cact((size_t)n * nmo);
  gather_coefficients(Cact.p, nbf, o_, v_, c_occ, c_vir, memspace, stream_);
  // buffers: x0 = AO (caller's device buffer or an upload), bufA for (nu la si p) and
  // (si p q r), bufB for (la si p q) and (p q r s)
  DevTmp up(memspace ? 0 : (size_t)n4);
  const double* x0 = eri;
  if (!memspace) {
    MPQC_HIP_CHECK(hipMemcpyAsync(up.p, eri, n4 * sizeof(double), hipMemcpyHostToDevice, stream_));
    x0 = up.p;
  }
  DevTmp bufA((size_t)std::max(n * n * n * nmo, n * nmo * nmo * nmo));
  DevTmp bufB(memspace ? (size_t)std::max(n * n * nmo * nmo, nmo * nmo * nmo * nmo) : 0);
  double* pB = memspace ? bufB.p : up.p;  // the uploaded AO buffer is free after the first quarter
  auto quarter = [&](const double* x, double* y, long rest) {
    // y[(rest), p] = sum_mu x[mu, (rest)] C[mu, p]
    mpqc_hip::gemm(make_gemm(true, false, (int)rest, (int)nmo, nbf, 1.0, x, rest, Cact.p, nmo, 0.0, y, nmo),
                   GemmPrecision::F64, ws_, stream_);
  };
  quarter(x0, bufA.p, n * n * n);              // (nu la si p)
  quarter(bufA.p, pB, n * n * nmo);            // (la si p q)
  quarter(pB, bufA.p, n * nmo * nmo);          // (si p q r)
  quarter(bufA.p, pB, nmo * nmo * nmo);        // (p q r s)
  const double* mo = pB;
  const long mo_vol = nmo * nmo * nmo * nmo;
  if (!Gabcd_) Gabcd_ = alloc((size_t)V * V * V * V);
  struct Blk { double* dst; long P0, Q0, R0, S0, nP, nQ, nR, nS; };
  const Blk blocks[] = {{Gijab_, 0, 0, O, O, O, O, V, V}, {Gijkl_, 0, 0, 0, 0, O, O, O, O},
                        {Giajb_, 0, O, 0, O, O, V, O, V}, {Giabc_, 0, O, O, O, O, V, V, V},
                        {Gijka_, 0, 0, 0, O, O, O, O, V}, {Gabcd_, O, O, O, O, V, V, V, V}};
  for (const Blk& b : blocks) {
    // <PQ|RS>[a,b,c,d] = mo[P0+a, R0+c, Q0+b, S0+d]
    const long ext[4] = {b.nP, b.nQ, b.nR, b.nS};
    const long s[4] = {nmo * nmo * nmo, nmo, nmo * nmo, 1};
    const long off = b.P0 * nmo * nmo * nmo + b.R0 * nmo * nmo + b.Q0 * nmo + b.S0;
    strided_copy(b.dst, ext, mo, off, s, mo_vol, stream_);
  }
  check_launch("transform_ao");
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  have_ints_ |= 0x1f8u;
  transform_fock(nbf, fock, Cact.p, memspace);
  prepared_ = false;
}

// compute4 with density fitting (lcao_factory.h:386-406) for the "df" method:
// X(K,p,q) = sum_L metric(L,K) (L|mu nu) C(mu,p) C(nu,q) gives Xai, Xij, Xab
// (ccsd.h:481-493), then every four-index block the df residual reads is a
// product of two factors over K, e.g. <ij|ab> = (ia|jb) = X(K,i,a) X(K,j,b).
void CCSDContext::transform_df(int nbf, const double* B3, const double* metric, const double* fock,
                               const double* c_occ, const double* c_vir, int memspace) {
  if (naux_ == 0) throw std::logic_error("transform_df needs set_df_basis first");
  if (nbf <= 0 || !B3 || !fock || !c_occ || !c_vir) throw std::invalid_argument("transform_df: null or empty input");
  if (memspace != 0 && memspace != 1) throw std::invalid_argument("memspace must be host or device");
  const long n = nbf, Q = naux_, nmo = (long)o_ + v_, O = o_, V = v_;
  if (nmo > n) throw std::invalid_argument("more active orbitals than basis functions");
  if (Q * n * n >= (1L << 31)) throw std::invalid_argument("naux * nbf^2 too large");
  const hipMemcpyKind kind = memspace ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  DevTmp Cact((size_t)n * nmo);
  gather_coefficients(Cact.p, nbf, o_, v_, c_occ, c_vir, memspace, stream_);
  // fitted AO factors X(K, mu nu) = sum_L metric(L,K) B(L, mu nu)
  DevTmp Xao((size_t)Q * n * n), Bup(metric ? (size_t)Q * n * n : 0), Mup(metric ? (size_t)Q * Q : 0);
  if (metric) {
    MPQC_HIP_CHECK(hipMemcpyAsync(Bup.p, B3, Q * n * n * sizeof(double), kind, stream_));
    MPQC_HIP_CHECK(hipMemcpyAsync(Mup.p, metric, Q * Q * sizeof(double), kind, stream_));
    mpqc_hip::gemm(make_gemm(true, false, (int)Q, (int)(n * n), (int)Q, 1.0, Mup.p, Q, Bup.p, n * n, 0.0, Xao.p,
                             n * n), GemmPrecision::F64, ws_, stream_);
  } else {
    MPQC_HIP_CHECK(hipMemcpyAsync(Xao.p, B3, Q * n * n * sizeof(double), kind, stream_));
  }
  // Y(K,mu,q) = X(K,mu,nu) C(nu,q);  Z(K,p,q) = C(mu,p) Y(K,mu,q)
  DevTmp Y((size_t)Q * n * nmo), Z((size_t)Q * nmo * nmo);
  mpqc_hip::gemm(make_gemm(false, false, (int)(Q * n), (int)nmo, nbf, 1.0, Xao.p, n, Cact.p, nmo, 0.0, Y.p, nmo),
                 GemmPrecision::F64, ws_, stream_);
  {
    GemmDesc d = make_gemm(true, false, (int)nmo, (int)nmo, nbf, 1.0, Cact.p, nmo, Y.p, nmo, 0.0, Z.p, nmo);
    d.sA = 0; d.sB = n * nmo; d.sC = nmo * nmo; d.batch = (int)Q;
    mpqc_hip::gemm(d, GemmPrecision::F64, ws_, stream_);
  }
  const long zvol = Q * nmo * nmo;
  const long zs[4] = {0, nmo * nmo, nmo, 1};
  {
    const long e_ai[4] = {1, Q, V, O}, e_ij[4] = {1, Q, O, O}, e_ab[4] = {1, Q, V, V};
    strided_copy(Xai_, e_ai, Z.p, O * nmo, zs, zvol, stream_);
    strided_copy(Xij_, e_ij, Z.p, 0, zs, zvol, stream_);
    strided_copy(Xab_, e_ab, Z.p, O * nmo + O, zs, zvol, stream_);
  }
  // Xia(K,i,a) contiguous, then the four-index blocks as X^T X GEMMs over K and one
  // permutation each
  DevTmp Xia((size_t)Q * O * V);
  {
    const long e_ia[4] = {1, Q, O, V};
    strided_copy(Xia.p, e_ia, Z.p, O, zs, zvol, stream_);
  }
  const bool full = !reduced_abcd_;
  const long tmp_n = full ? std::max(V * V * V * V, O * V * V * V) : std::max(O * O * V * V, O * O * O * V);
  DevTmp T((size_t)std::max(tmp_n, O * O * O * O));
  // dst[perm(x,y,z,w)] = sum_K A(K,x,y) B(K,z,w): T[(x,y)][(z,w)] = A^T B, then dst = permute(T)
  auto block = [&](double* dst, const double* A, long nx, long ny, const double* Bm, long nz, long nw,
                   std::array<int, 4> perm) {
    mpqc_hip::gemm(make_gemm(true, false, (int)(nx * ny), (int)(nz * nw), (int)Q, 1.0, A, nx * ny, Bm, nz * nw,
                             0.0, T.p, nz * nw), GemmPrecision::F64, ws_, stream_);
    const long in_ext[4] = {nx, ny, nz, nw};
    long out_ext[4];
    const int pm[4] = {perm[0], perm[1], perm[2], perm[3]};
    for (int d = 0; d < 4; ++d) out_ext[d] = in_ext[pm[d]];
    TTerm t;
    t.p = T.p;
    permuted_strides(in_ext, pm, t.s);
    t.alpha = 1.0;
    t.vol = nx * ny * nz * nw;
    tadd(dst, out_ext, 0.0, &t, 1, stream_);
  };
  block(Gijab_, Xia.p, O, V, Xia.p, O, V, {0, 2, 1, 3});  // (ia|jb) -> <ij|ab>
  block(Gijkl_, Xij_, O, O, Xij_, O, O, {0, 2, 1, 3});    // (ik|jl) -> <ij|kl>
  block(Giajb_, Xij_, O, O, Xab_, V, V, {0, 2, 1, 3});    // (ij|ab) -> <ia|jb>
  block(Gijka_, Xij_, O, O, Xia.p, O, V, {0, 2, 1, 3});   // (ik|ja) -> <ij|ka>
  unsigned bits = 0x0b8u | 0xe00u;                        // Gijab Gijkl Giajb Gijka, Xai Xij Xab
  if (full) {
    if (!Gabcd_) Gabcd_ = alloc((size_t)V * V * V * V);
    block(Giabc_, Xia.p, O, V, Xab_, V, V, {0, 2, 1, 3});  // (ib|ac) -> <ia|bc>
    block(Gabcd_, Xab_, V, V, Xab_, V, V, {0, 2, 1, 3});   // (ac|bd) -> <ab|cd>
    bits |= 0x140u;
  }
  check_launch("transform_df");
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  have_ints_ |= bits;
  transform_fock(nbf, fock, Cact.p, memspace);
  prepared_ = false;
}

}  // namespace mpqc_hip
