// EOM-EE-CCSD singlet excitation energies on the device.
//
// At converged amplitudes the Jacobian J = dR/dt of the CCSD residual
// (ccsd_r1_r2.h) is the similarity-transformed Hamiltonian minus E_CC on the
// space of singles and doubles, so its eigenvalues are the EOM-CCSD excitation
// energies that EOM_CCSD::eom_ccsd_davidson_solver (eom_ccsd_impl.h:236-390)
// finds with H-bar built from the intermediates of ccsd_hbar.h.  The
// closed-shell residual is equivariant under the exchange (a,i) <-> (b,j) of
// the doubles, so J commutes with it and the singlets are the eigenvalues of J
// on (ab)(ij)-symmetric doubles.  J^T has the same eigenvalues, and its
// products are exact reverse sweeps of the residual's op tape, so the Davidson
// iterations run on J^T (math::DavidsonDiag::solve / extrapolate,
// davidson_diag.h:128-425, non-symmetric form):
//   - vectors are contiguous (t1(a,i), t2(a,b,i,j)) slots in HBM
//   - the subspace matrix G(i,j) = <b_i, J^T b_j> grows by its new rows and columns
//   - the lowest n_roots eigenvalues of G by real part, with the real parts of
//     their eigenvectors
//   - residuals preconditioned by 1 / (theta - D), D from the diagonals of the
//     dressed Fock matrices FIJ and FAB (EEPred, eom_preconditioner.h:18-74)
//   - once the subspace holds more than n_roots (max_vector - 1) vectors it is
//     rebuilt from the residuals and the Ritz vectors of this and the previous
//     iteration; new vectors are orthonormalised by two Gram-Schmidt passes that
//     drop vectors whose norm falls below the vector threshold (gram_schmidt.h)
//   - converged when max |delta theta| <= convergence and the largest residual
//     norm <= 10 convergence, or when no new vector passes the threshold
// Initial vectors: unit singles on the n_guess smallest e_a - e_i (the
// reference starts from CIS vectors; the converged roots do not depend on it).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <stdexcept>

#include "ccsd_ctx.hpp"
#include "small_eig.hpp"

namespace mpqc_hip {

namespace {

constexpr int kEomBlock = 256;

// out = P r / (theta - D): D = e_a - e_i (singles), e_a + e_b - e_i - e_j
// (doubles, summed in an order symmetric under (a,i) <-> (b,j)), P the
// projection onto t2(a,b,i,j) = t2(b,a,j,i).  Out of place: the doubles read
// their exchanged element.  Without the projection the doubles would keep a
// round-off antisymmetric part, on which P J^T vanishes exactly: Gram-Schmidt
// would amplify it into subspace vectors with a spurious zero root.
// d1 (optional, v x o): the singles' diagonal in place of e_a - e_i
__global__ __launch_bounds__(kEomBlock) void ee_precondition_kernel(double* __restrict__ out,
                                                                    const double* __restrict__ r,
                                                                    const double* __restrict__ eo,
                                                                    const double* __restrict__ ev,
                                                                    const double* __restrict__ d1, int o,
                                                                    int v, double theta, long n1, long n) {
  for (long e = (long)blockIdx.x * kEomBlock + threadIdx.x; e < n; e += (long)gridDim.x * kEomBlock) {
    double d, x;
    if (e < n1) {
      const long a = e / o, i = e - a * o;
      d = d1 ? theta - d1[e] : theta + eo[i] - ev[a];
      x = r[e];
    } else {
      long k = e - n1;
      const long j = k % o;
      k /= o;
      const long i = k % o;
      k /= o;
      const long b = k % v, a = k / v;
      d = theta - (ev[a] + ev[b]) + (eo[i] + eo[j]);
      const long sw = n1 + ((b * v + a) * o + j) * o + i;
      x = 0.5 * (r[e] + r[sw]);
    }
    if (fabs(d) < 1e-8) d = d < 0.0 ? -1e-8 : 1e-8;
    out[e] = x / d;
  }
}

// CIS singlet matrix A((a,i),(b,j)) (extents v, o, v, o): the Fock terms
// A(ai,bi) += F(a,b) and A(ai,aj) -= F(i,j) (separate launches: the diagonal gets both)
__global__ __launch_bounds__(kEomBlock) void cis_fock_vv_kernel(double* __restrict__ A, const double* __restrict__ Fab,
                                                                int o, int v) {
  const long total = (long)v * o * v;
  for (long e = (long)blockIdx.x * kEomBlock + threadIdx.x; e < total; e += (long)gridDim.x * kEomBlock) {
    const long b = e % v, ai = e / v, i = ai % o, a = ai / o;
    A[((a * o + i) * v + b) * o + i] += Fab[a * v + b];
  }
}
__global__ __launch_bounds__(kEomBlock) void cis_fock_oo_kernel(double* __restrict__ A, const double* __restrict__ Fij,
                                                                int o, int v) {
  const long total = (long)v * o * o;
  for (long e = (long)blockIdx.x * kEomBlock + threadIdx.x; e < total; e += (long)gridDim.x * kEomBlock) {
    const long j = e % o, ai = e / o, i = ai % o, a = ai / o;
    A[((a * o + i) * v + a) * o + j] -= Fij[i * o + j];
  }
}

// unit vector e_k
__global__ void unit_vector_kernel(double* __restrict__ x, long k) { x[k] = 1.0; }

// Fixed set of equal-sized device vectors handed out and returned by pointer.
struct SlotPool {
  double* base = nullptr;
  std::vector<double*> free_;
  SlotPool(int nslots, long n) {
    if (hipMalloc(&base, (size_t)nslots * n * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      base = nullptr;
      throw std::bad_alloc();
    }
    for (int k = nslots - 1; k >= 0; --k) free_.push_back(base + (size_t)k * n);
  }
  ~SlotPool() {
    if (base) (void)hipFree(base);
  }
  SlotPool(const SlotPool&) = delete;
  SlotPool& operator=(const SlotPool&) = delete;
  double* get() {
    if (free_.empty()) throw std::logic_error("EOM: vector pool exhausted");
    double* p = free_.back();
    free_.pop_back();
    return p;
  }
  void put(double* p) { free_.push_back(p); }
};

}  // namespace

// Gabcd(a,b,c,d) = (ac|bd) = sum_K Xab(K,a,c) Xab(K,b,d)
void CCSDContext::ensure_abcd_from_factors() {
  if ((have_ints_ >> 8) & 1u) return;
  if (!df_) throw std::logic_error("Gabcd is not set");
  const long Q = naux_, V = v_, vv = V * V;
  if (!Gabcd_) Gabcd_ = alloc((size_t)vv * vv);
  double* T = nullptr;
  if (hipMalloc(&T, (size_t)vv * vv * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    throw std::bad_alloc();
  }
  try {
    mpqc_hip::gemm(make_gemm(true, false, (int)vv, (int)vv, (int)Q, 1.0, Xab_, vv, Xab_, vv, 0.0, T, vv),
                   GemmPrecision::F64, ws_, stream_);
    const long ext[4] = {V, V, V, V};
    const int pm[4] = {0, 2, 1, 3};
    TTerm t;
    t.p = T;
    permuted_strides(ext, pm, t.s);
    t.vol = vv * vv;
    tadd(Gabcd_, ext, 0.0, &t, 1, stream_);
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  } catch (...) {
    (void)hipFree(T);
    throw;
  }
  MPQC_HIP_CHECK(hipFree(T));
  have_ints_ |= 1u << 8;
}

void CCSDContext::eom_apply_jt(const double* y, double* x) {
  const long O = o_, V = v_, n1 = O * V, n2 = n1 * n1;
  zero_adjoints();
  MPQC_HIP_CHECK(hipMemcpyAsync(adj(r1_), y, n1 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipMemcpyAsync(adj(r2_), y + n1, n2 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  reverse_sweep(res_tape_, false);
  MPQC_HIP_CHECK(hipMemcpyAsync(x, adj(t1_), n1 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  // x2 = (t2^ + t2^(b,a,j,i)) / 2
  const long e[4] = {V, V, O, O};
  const int I[4] = {0, 1, 2, 3}, sw[4] = {1, 0, 3, 2};
  TTerm ts[2];
  for (int k = 0; k < 2; ++k) {
    ts[k].p = adj(t2_);
    permuted_strides(e, k ? sw : I, ts[k].s);
    ts[k].alpha = 0.5;
    ts[k].vol = n2;
  }
  tadd(x + n1, e, 0.0, ts, 2, stream_);
}

int CCSDContext::eom_ee(int n_roots, int max_iter, double convergence, int verbose, double* energies) {
  if (n_roots < 1) throw std::invalid_argument("n_roots must be positive");
  if (max_iter < 1) throw std::invalid_argument("max_iter must be positive");
  if (!(convergence > 0.0)) throw std::invalid_argument("convergence must be positive");
  if (!energies) throw std::invalid_argument("null energies buffer");
  TapePrecision tp(*this);  // the Jacobian products are reverse sweeps of the (FP64 in split forms) tape
  const long O = o_, V = v_, n1 = O * V, n2 = n1 * n1, n = n1 + n2;
  if (n_roots > n1) throw std::invalid_argument("n_roots exceeds the number of single excitations");
  const int max_vector = std::max(eom_max_vector_, 2);
  const double threshold = eom_vector_threshold_ > 0.0 ? eom_vector_threshold_ : 10.0 * convergence;
  // tracked roots: the requested ones plus eom_extra_roots above them, which keep
  // near-degenerate states from flipping in and out of the requested set
  const int nt = (int)std::min<long>(n1, (long)n_roots + eom_extra_roots_);
  const int n_guess = (int)std::min<long>(
      n1, std::max<long>(nt, eom_n_guess_ > 0 ? eom_n_guess_ : std::max(2L * n_roots, (long)nt)));

  // The df residual is the standard residual with -f_ov t1 in g_ac on the
  // four-index blocks the factors assemble (ccsd_r1_r2.h:136-184 against
  // :70-293): the same function, so the same Jacobian.
  if (!prepared_) prepare();
  const bool was_df = df_;
  const double was_fov = fov_sign_;
  if (df_) {  // (direct_df: the direct residual on the blocks the factors assemble)
    ensure_iabc_for_triples();
    if (!direct_) ensure_abcd_from_factors();
    df_ = false;
    fov_sign_ = -1.0;
    prepared_ = false;
  }
  struct Restore {
    CCSDContext* c;
    bool df;
    double fov;
    ~Restore() {
      if (c->df_ != df || c->fov_sign_ != fov) {
        c->df_ = df;
        c->fov_sign_ = fov;
        c->prepared_ = false;
      }
    }
  } restore{this, was_df, was_fov};
  record_residual();  // the residual at the current amplitudes, recorded for the sweeps

  // preconditioner diagonals: FIJ(i,i), FAB(a,a) of the recorded residual
  double* dg = adj_hist(110, O + V);
  MPQC_HIP_CHECK(hipMemcpy2DAsync(dg, sizeof(double), FIJ_, (O + 1) * sizeof(double), sizeof(double), O,
                                  hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipMemcpy2DAsync(dg + O, sizeof(double), FAB_, (V + 1) * sizeof(double), sizeof(double), V,
                                  hipMemcpyDeviceToDevice, stream_));
  std::vector<double> hdg(O + V);
  MPQC_HIP_CHECK(hipMemcpyAsync(hdg.data(), dg, (O + V) * sizeof(double), hipMemcpyDeviceToHost, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  // singles' diagonal: FAB(a,a) - FIJ(i,i), plus with eom_cis_diagonal the two-electron
  // part of the CIS singlet diagonal, 2 (ia|ia) - (ii|aa) = 2 Gijab(i,i,a,a) - Giajb(i,a,i,a)
  std::vector<double> hd1((size_t)n1);
  for (long a = 0; a < V; ++a)
    for (long i = 0; i < O; ++i) hd1[a * O + i] = hdg[O + a] - hdg[i];
  double* d1 = nullptr;
  if (eom_cis_diagonal_) {
    std::vector<double> g1((size_t)n1), g2((size_t)n1);
    for (long i = 0; i < O; ++i) {
      MPQC_HIP_CHECK(hipMemcpy2DAsync(g1.data() + i * V, sizeof(double), Gijab_ + (i * O + i) * V * V,
                                      (V + 1) * sizeof(double), sizeof(double), V, hipMemcpyDeviceToHost, stream_));
      MPQC_HIP_CHECK(hipMemcpy2DAsync(g2.data() + i * V, sizeof(double), Giajb_ + i * V * O * V + i * V,
                                      (O * V + 1) * sizeof(double), sizeof(double), V, hipMemcpyDeviceToHost,
                                      stream_));
    }
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
    for (long a = 0; a < V; ++a)
      for (long i = 0; i < O; ++i) hd1[a * O + i] += 2.0 * g1[i * V + a] - g2[i * V + a];
    d1 = adj_hist(111, n1);
    MPQC_HIP_CHECK(hipMemcpyAsync(d1, hd1.data(), n1 * sizeof(double), hipMemcpyHostToDevice, stream_));
  }

  eom_products_ = 0;
  eom_product_seconds_ = 0.0;
  hipEvent_t ev0, ev1;
  MPQC_HIP_CHECK(hipEventCreate(&ev0));
  MPQC_HIP_CHECK(hipEventCreate(&ev1));
  struct Events {
    hipEvent_t a, b;
    ~Events() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } events{ev0, ev1};
  // h = P J^T g, timed
  auto product = [&](const double* g, double* h) {
    MPQC_HIP_CHECK(hipEventRecord(ev0, stream_));
    eom_apply_jt(g, h);
    MPQC_HIP_CHECK(hipEventRecord(ev1, stream_));
    MPQC_HIP_CHECK(hipEventSynchronize(ev1));
    float ms = 0.0f;
    MPQC_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    eom_product_seconds_ += 1e-3 * ms;
    ++eom_products_;
  };

  const int cap = std::max({nt * max_vector, n_guess, 3 * nt});
  SlotPool pool(2 * cap + 3 * nt + 1, n);
  double* scal = scalars_ + 16;

  // out[k] = <xs[k], y> over len elements
  auto dots = [&](const std::vector<const double*>& xs, const double* y, std::vector<double>& out, long len = 0) {
    if (len == 0) len = n;
    out.assign(xs.size(), 0.0);
    for (size_t c0 = 0; c0 < xs.size(); c0 += 16) {
      const int m = (int)std::min<size_t>(16, xs.size() - c0);
      multi_dot2(xs.data() + c0, xs.data() + c0, m, y, y, len, 0, scal, partials_, stream_);
      MPQC_HIP_CHECK(hipMemcpyAsync(out.data() + c0, scal, m * sizeof(double), hipMemcpyDeviceToHost, stream_));
    }
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  };
  // y = sum_k c_k xs[k] (y may be one of xs: every element is read before it is written)
  auto combine = [&](double* y, const std::vector<const double*>& xs, const std::vector<double>& c, long len = 0) {
    if (len == 0) len = n;
    size_t k = 0;
    bool first = true;
    while (first || k < xs.size()) {
      std::vector<const double*> px;
      std::vector<double> pc;
      if (!first) {
        px.push_back(y);
        pc.push_back(1.0);
      }
      while (px.size() < 16 && k < xs.size()) {
        px.push_back(xs[k]);
        pc.push_back(c[k]);
        ++k;
      }
      lincomb(y, px.data(), pc.data(), (int)px.size(), len, stream_);
      first = false;
    }
  };
  auto norm = [&](const double* x, long len = 0) {
    std::vector<double> d;
    dots({x}, x, d, len);
    return std::sqrt(std::max(d[0], 0.0));
  };
  // orthonormalise cand against basis and against each other, twice; vectors
  // whose norm after projection is below the threshold are dropped
  auto gram_schmidt = [&](const std::vector<double*>& basis, std::vector<double*>& cand, long len = 0,
                          SlotPool* pl = nullptr, double thresh = 0.0) {
    if (len == 0) len = n;
    if (!pl) pl = &pool;
    if (thresh == 0.0) thresh = threshold;
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double*> kept;
      for (double* x : cand) {
        std::vector<const double*> against(basis.begin(), basis.end());
        against.insert(against.end(), kept.begin(), kept.end());
        if (!against.empty()) {
          std::vector<double> c;
          dots(against, x, c, len);
          for (double& e : c) e = -e;
          against.insert(against.begin(), x);
          c.insert(c.begin(), 1.0);
          combine(x, against, c, len);
        }
        const double nr = norm(x, len);
        if (!std::isfinite(nr)) throw std::range_error("non-finite EOM vector");
        if (nr < thresh) {
          pl->put(x);
          if (verbose) std::printf("  EOM: Gram-Schmidt drops a vector of norm %.3e\n", nr);
          continue;
        }
        const double* xs[1] = {x};
        const double s = 1.0 / nr;
        lincomb(x, xs, &s, 1, len, stream_);
        kept.push_back(x);
      }
      cand = kept;
    }
  };

  // initial vectors: with eom_cis_guess the lowest CIS singlet vectors (as the reference,
  // eom_ccsd_impl.h:431-450), else unit singles on the smallest singles' diagonal
  std::vector<double*> guess;
  if (eom_cis_guess_) {
    const long m = n1;
    // A(ai,bj) = delta_ij F(a,b) - delta_ab F(i,j) + 2 <ij|ab> - <ia|jb>: the singles block of the
    // Jacobian at t = 0 (symmetric)
    double* A = nullptr;
    if (hipMalloc(&A, (size_t)m * m * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      throw std::bad_alloc();
    }
    struct FreeA {
      double* p;
      ~FreeA() { (void)hipFree(p); }
    } free_a{A};
    {
      const long e4[4] = {V, O, V, O}, eg[4] = {O, O, V, V}, ej[4] = {O, V, O, V};
      const int pg[4] = {2, 0, 3, 1}, pj[4] = {1, 0, 3, 2};
      TTerm ts[2];
      ts[0].p = Gijab_;
      permuted_strides(eg, pg, ts[0].s);
      ts[0].alpha = 2.0;
      ts[0].vol = n1 * n1;
      ts[1].p = Giajb_;
      permuted_strides(ej, pj, ts[1].s);
      ts[1].alpha = -1.0;
      ts[1].vol = n1 * n1;
      tadd(A, e4, 0.0, ts, 2, stream_);
      const int g1 = (int)std::max<long>(1, std::min<long>(ceil_div(V * O * V, kEomBlock), 16384));
      hipLaunchKernelGGL(cis_fock_vv_kernel, dim3(g1), dim3(kEomBlock), 0, stream_, A, Fab_, (int)O, (int)V);
      const int g2 = (int)std::max<long>(1, std::min<long>(ceil_div(V * O * O, kEomBlock), 16384));
      hipLaunchKernelGGL(cis_fock_oo_kernel, dim3(g2), dim3(kEomBlock), 0, stream_, A, Fij_, (int)O, (int)V);
      check_launch("cis_fock");
    }
    double* dA = adj_hist(112, m);
    MPQC_HIP_CHECK(hipMemcpy2DAsync(dA, sizeof(double), A, (m + 1) * sizeof(double), sizeof(double), m,
                                    hipMemcpyDeviceToDevice, stream_));
    std::vector<double> hdA(m);
    MPQC_HIP_CHECK(hipMemcpyAsync(hdA.data(), dA, m * sizeof(double), hipMemcpyDeviceToHost, stream_));
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
    // symmetric Davidson on A for the n_guess lowest roots, to residual norms of 1e-5
    const int ng = n_guess, ccap = std::max(12 * ng, 2 * ng + 2);
    SlotPool cpool(2 * ccap + 2 * ng + 1, m);
    std::vector<double*> cb, ch, cnew;
    {
      std::vector<std::pair<double, long>> ord;
      for (long e = 0; e < m; ++e) ord.push_back({hdA[e], e});
      std::stable_sort(ord.begin(), ord.end(),
                       [](const std::pair<double, long>& x, const std::pair<double, long>& y) { return x.first < y.first; });
      for (int g = 0; g < ng; ++g) {
        double* x = cpool.get();
        fill(x, m, 0.0, stream_);
        hipLaunchKernelGGL(unit_vector_kernel, dim3(1), dim3(1), 0, stream_, x, ord[g].second);
        check_launch("unit_vector");
        cnew.push_back(x);
      }
    }
    std::vector<double> cG((size_t)ccap * ccap, 0.0);
    std::vector<double*> cx;
    const int cgrid = (int)std::max<long>(1, std::min<long>(ceil_div(m, kEomBlock), 16384));
    for (int it = 0; it < 100 && !cnew.empty(); ++it) {
      const int ns = (int)cb.size();
      for (double* b : cnew) {
        double* h = cpool.get();
        mpqc_hip::gemm(make_gemm(false, false, (int)m, 1, (int)m, 1.0, A, m, b, 1, 0.0, h, 1), GemmPrecision::F64,
                       ws_, stream_);
        cb.push_back(b);
        ch.push_back(h);
      }
      cnew.clear();
      const int nv = (int)cb.size();
      for (int ii = ns; ii < nv; ++ii) {
        std::vector<double> row;
        dots(std::vector<const double*>(ch.begin(), ch.begin() + ii + 1), cb[ii], row, m);
        for (int j = 0; j <= ii; ++j) cG[(size_t)ii * ccap + j] = cG[(size_t)j * ccap + ii] = row[j];
      }
      std::vector<double> Gs((size_t)nv * nv);
      for (int i = 0; i < nv; ++i)
        for (int j = 0; j < nv; ++j) Gs[(size_t)i * nv + j] = cG[(size_t)i * ccap + j];
      const std::vector<std::complex<double>> w = small_eigenvalues(nv, Gs);
      std::vector<int> order(nv);
      for (int k = 0; k < nv; ++k) order[k] = k;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a].real() < w[b].real(); });
      for (double* x : cx) cpool.put(x);
      cx.clear();
      std::vector<double*> cres;
      double worst = 0.0;
      for (int r = 0; r < ng && r < nv; ++r) {
        const double theta = w[order[r]].real();
        const std::vector<std::complex<double>> cv = small_eigenvector(nv, Gs, w[order[r]]);
        std::vector<double> c(nv);
        double cn = 0.0;
        for (int j = 0; j < nv; ++j) {
          c[j] = cv[j].real();
          cn += c[j] * c[j];
        }
        cn = std::sqrt(cn);
        for (double& e : c) e /= cn;
        double* x = cpool.get();
        combine(x, std::vector<const double*>(cb.begin(), cb.end()), c, m);
        cx.push_back(x);
        double* rr = cpool.get();
        std::vector<const double*> hx(ch.begin(), ch.end());
        std::vector<double> hc(c);
        hx.push_back(x);
        hc.push_back(-theta);
        combine(rr, hx, hc, m);
        const double nr = norm(rr, m);
        worst = std::max(worst, nr);
        double* p = cpool.get();
        hipLaunchKernelGGL(ee_precondition_kernel, dim3(cgrid), dim3(kEomBlock), 0, stream_, p, rr, dg, dg + O, dA,
                           (int)O, (int)V, theta, m, m);
        check_launch("cis_precondition");
        cpool.put(rr);
        cres.push_back(p);
      }
      if (verbose) std::printf("  CIS %3d  subspace %3d  max|r| %.3e\n", it + 1, nv, worst);
      if (worst < 1e-5) {
        for (double* p : cres) cpool.put(p);
        break;
      }
      if (nv + (int)cres.size() > ccap) {  // restart from the Ritz vectors and their corrections
        for (double* b : cb) cpool.put(b);
        for (double* h : ch) cpool.put(h);
        cb.clear();
        ch.clear();
        std::fill(cG.begin(), cG.end(), 0.0);
        for (double* x : cx) {
          double* g = cpool.get();
          MPQC_HIP_CHECK(hipMemcpyAsync(g, x, m * sizeof(double), hipMemcpyDeviceToDevice, stream_));
          cnew.push_back(g);
        }
        cnew.insert(cnew.end(), cres.begin(), cres.end());
        gram_schmidt({}, cnew, m, &cpool, 1e-10);
      } else {
        cnew = cres;
        gram_schmidt(cb, cnew, m, &cpool, 1e-10);
      }
    }
    // the Ritz vectors as the singles of the first EOM vectors
    for (double* x : cx) {
      double* g = pool.get();
      fill(g, n, 0.0, stream_);
      MPQC_HIP_CHECK(hipMemcpyAsync(g, x, m * sizeof(double), hipMemcpyDeviceToDevice, stream_));
      guess.push_back(g);
    }
    gram_schmidt({}, guess);
  } else {
    std::vector<std::pair<double, long>> order1;
    for (long e = 0; e < n1; ++e) order1.push_back({hd1[e], e});
    std::stable_sort(order1.begin(), order1.end(),
                     [](const std::pair<double, long>& x, const std::pair<double, long>& y) { return x.first < y.first; });
    for (int g = 0; g < n_guess; ++g) {
      double* x = pool.get();
      fill(x, n, 0.0, stream_);
      hipLaunchKernelGGL(unit_vector_kernel, dim3(1), dim3(1), 0, stream_, x, order1[g].second);
      check_launch("unit_vector");
      guess.push_back(x);
    }
  }

  std::vector<double*> B, HB, Xcur, Xprev, R;
  std::vector<double> G;  // row-major, leading dimension cap
  std::vector<double> eig(nt, 0.0), E(nt, 0.0), norms(nt, 0.0);
  double norm_e = 1.0, norm_r = 1.0;
  int iter = 0;
  const int grid = (int)std::max<long>(1, std::min<long>(ceil_div(n, kEomBlock), 16384));
  while (iter < max_iter && (norm_r > 10.0 * convergence || norm_e > convergence) && !guess.empty()) {
    // products with the new vectors
    const int ns = (int)B.size();
    for (double* g : guess) {
      double* h = pool.get();
      product(g, h);
      B.push_back(g);
      HB.push_back(h);
    }
    guess.clear();
    const int nv = (int)B.size();
    if (nv > cap) throw std::logic_error("EOM: subspace exceeds its capacity");
    // G(ii, j) = <b_ii, Hb_j>, G(j, ii) = <b_j, Hb_ii> for the new ii
    G.resize((size_t)cap * cap, 0.0);
    for (int ii = ns; ii < nv; ++ii) {
      std::vector<const double*> hb(HB.begin(), HB.begin() + ii + 1), bb(B.begin(), B.begin() + ii + 1);
      std::vector<double> row, col;
      dots(hb, B[ii], row);
      dots(bb, HB[ii], col);
      for (int j = 0; j <= ii; ++j) {
        G[(size_t)ii * cap + j] = row[j];
        G[(size_t)j * cap + ii] = col[j];
      }
    }
    std::vector<double> Gs((size_t)nv * nv);
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < nv; ++j) Gs[(size_t)i * nv + j] = G[(size_t)i * cap + j];
    const std::vector<std::complex<double>> w = small_eigenvalues(nv, Gs);
    std::vector<int> order(nv);
    for (int k = 0; k < nv; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a].real() < w[b].real(); });
    // Ritz vectors X_r = B C_r and residuals R_r = HB C_r - theta_r X_r
    std::swap(Xprev, Xcur);
    for (double* x : Xcur) pool.put(x);
    Xcur.clear();
    for (double* r : R) pool.put(r);
    R.clear();
    for (int r = 0; r < nt; ++r) {
      const std::complex<double> lam = w[order[r]];
      E[r] = lam.real();
      const std::vector<std::complex<double>> cv = small_eigenvector(nv, Gs, lam);
      std::vector<double> c(nv);
      double cn = 0.0;
      for (int j = 0; j < nv; ++j) {
        c[j] = cv[j].real();
        cn += c[j] * c[j];
      }
      cn = std::sqrt(cn);
      for (double& e : c) e /= cn;
      double* x = pool.get();
      double* rr = pool.get();
      combine(x, std::vector<const double*>(B.begin(), B.end()), c);
      std::vector<const double*> hx(HB.begin(), HB.end());
      std::vector<double> hc(c);
      hx.push_back(x);
      hc.push_back(-E[r]);
      combine(rr, hx, hc);
      Xcur.push_back(x);
      R.push_back(rr);
      norms[r] = norm(rr);
      if (!std::isfinite(norms[r])) throw std::range_error("non-finite EOM residual");
    }
    // a root whose residual and change are both within the criteria gets no new vector
    std::vector<bool> done(nt);
    for (int r = 0; r < nt; ++r)
      done[r] = eom_skip_converged_ && norms[r] <= 10.0 * convergence && std::fabs(eig[r] - E[r]) <= convergence;
    for (int r = 0; r < nt; ++r) {
      double* p = pool.get();
      hipLaunchKernelGGL(ee_precondition_kernel, dim3(grid), dim3(kEomBlock), 0, stream_, p, R[r], dg, dg + O, d1,
                         (int)O, (int)V, E[r], n1, n);
      check_launch("ee_precondition");
      pool.put(R[r]);
      R[r] = p;
    }
    if (nv > nt * (max_vector - 1)) {
      // collapse: the residuals, then the Ritz vectors of the previous and this iteration
      for (double* b : B) pool.put(b);
      for (double* h : HB) pool.put(h);
      B.clear();
      HB.clear();
      std::fill(G.begin(), G.end(), 0.0);
      guess.clear();
      for (int r = 0; r < nt; ++r) {
        if (done[r]) pool.put(R[r]);
        else guess.push_back(R[r]);
      }
      R.clear();
      for (const std::vector<double*>* xs : {&Xprev, &Xcur})
        for (double* x : *xs) {
          double* g = pool.get();
          MPQC_HIP_CHECK(hipMemcpyAsync(g, x, n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
          guess.push_back(g);
        }
      gram_schmidt({}, guess);
    } else {
      guess.clear();
      for (int r = 0; r < nt; ++r) {
        if (done[r]) pool.put(R[r]);
        else guess.push_back(R[r]);
      }
      R.clear();
      gram_schmidt(B, guess);
    }
    norm_e = 0.0;
    norm_r = 0.0;
    for (int r = 0; r < n_roots; ++r) {
      norm_e = std::max(norm_e, std::fabs(eig[r] - E[r]));
      norm_r = std::max(norm_r, norms[r]);
    }
    eig = E;
    ++iter;
    if (verbose) {
      std::printf("  EOM %3d  subspace %3d  max|dE| %.3e  max|r| %.3e ", iter, nv, norm_e, norm_r);
      for (int r = 0; r < n_roots; ++r) std::printf(" %.12f", E[r]);
      std::printf("\n");
      std::fflush(stdout);
    }
  }
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  // As DavidsonDiag::solve: an exhausted subspace (every new vector below the
  // threshold) ends the iterations with the current roots; max_iter without
  // convergence is an error.
  const bool converged = norm_r <= 10.0 * convergence && norm_e <= convergence;
  if (!converged && guess.empty() && verbose)
    std::printf("  EOM: no new vector above the threshold %.3e; stopping\n", threshold);
  if (!converged && !guess.empty()) throw std::runtime_error("EOM Davidson exceeded max_iter");
  for (int r = 0; r < n_roots; ++r) energies[r] = eig[r];
  return iter;
}

}  // namespace mpqc_hip
