// Adjoint of the converged CCSD energy: Lambda equations and integral gradients.
//
// The residual (and the integral preparation) is recorded as a tape of
// primitive ops; the reverse sweep applies each op's adjoint, newest first:
//   copy      dst = src                     src^ += dst^,  dst^ = 0
//   tadd      out = b out + sum_t a_t P_t(x_t)  x_t^ += a_t P_t^-1(out^), out^ *= b
//   gemm      C = a op(A) op(B) + b C        A^ += a C^ op(B)^T (as stored), B^ += a op(A)^T C^,
//                                            C^ *= b; operands shared by a batch are summed
//                                            over the batch through a partial-sum workspace
//   tau       tau = t2 + t1(a,i) t1(b,j)     t2^ += tau^, t1^ += (tau^ + tau^(b,a,j,i)) . t1
//   TP        TP(a,i,l,d) = t2(d,a,i,l)/2 + t1(d,i) t1(a,l)   (same pattern)
// Every adjoint reuses the forward kernels (contractions on the matrix cores,
// permutations through the LDS-tiled permute).  The forward intermediates of
// the recorded residual stay in HBM and are read by the sweep; the tape is
// checked so that no buffer is overwritten after an op has read it.
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "ccsd_ctx.hpp"

namespace mpqc_hip {

namespace {

constexpr int kOcc = 0, kVir = 1, kNone = -1;

TTerm term_of(const double* p, const std::array<long, 4>& ext, const int perm[4], double alpha) {
  TTerm t;
  t.p = p;
  long e[4] = {ext[0], ext[1], ext[2], ext[3]};
  permuted_strides(e, perm, t.s);
  t.alpha = alpha;
  t.vol = ext[0] * ext[1] * ext[2] * ext[3];
  return t;
}

long vol4(const std::array<long, 4>& e) { return e[0] * e[1] * e[2] * e[3]; }

}  // namespace

int CCSDContext::alloc_index(const double* p) const {
  for (size_t k = 0; k < allocations_.size(); ++k)
    if (p >= allocations_[k] && p < allocations_[k] + alloc_sizes_[k]) return (int)k;
  return -1;
}

double* CCSDContext::adj(const double* p) {
  const int k = alloc_index(p);
  if (k < 0) throw std::logic_error("adjoint of a buffer the context does not own");
  if (adj_.size() < allocations_.size()) adj_.resize(allocations_.size(), nullptr);
  if (!adj_[k]) {
    void* q = nullptr;
    if (hipMalloc(&q, alloc_sizes_[k] * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      throw std::bad_alloc();
    }
    adj_[k] = (double*)q;
    fill(adj_[k], (long)alloc_sizes_[k], 0.0, stream_);
  }
  return adj_[k] + (p - allocations_[k]);
}

bool CCSDContext::is_integral(const double* p) const {
  for (const double* q : {Fia_, Fij_, Fab_, Gijab_, Gijkl_, Giajb_, Giabc_, Gijka_, Gabcd_, Xao_, Ci_, Ca_})
    if (q && alloc_index(q) == alloc_index(p)) return true;
  for (const TapeOp& op : prepare_tape_)
    if (alloc_index(op.dst) == alloc_index(p)) return true;
  return false;
}

double* CCSDContext::adj_scratch(long n) {
  if (adj_scratch_n_ < n) {
    if (adj_scratch_) MPQC_HIP_CHECK(hipFree(adj_scratch_));
    adj_scratch_ = nullptr;
    MPQC_HIP_CHECK(hipMalloc(&adj_scratch_, n * sizeof(double)));
    adj_scratch_n_ = n;
  }
  return adj_scratch_;
}

double* CCSDContext::adj_hist(int k, long n) {
  if ((int)lam_hist_.size() <= k) lam_hist_.resize(k + 1, nullptr);
  if (!lam_hist_[k]) lam_hist_[k] = alloc(n);
  return lam_hist_[k];
}

void CCSDContext::zero_adjoints() {
  for (size_t k = 0; k < adj_.size(); ++k)
    if (adj_[k]) fill(adj_[k], (long)alloc_sizes_[k], 0.0, stream_);
}

void CCSDContext::record_residual() {
  if (split_mode()) throw std::logic_error("record_residual: the tape records FP64 or FP32 contractions");
  if (df_) throw std::invalid_argument("the adjoint runs on the standard method");
  res_tape_.clear();
  tape_ = &res_tape_;
  tape_execute_ = true;
  try {
    residual();
  } catch (...) {
    tape_ = nullptr;
    throw;
  }
  tape_ = nullptr;
  check_tape(res_tape_);
}

// No op may write a buffer that an earlier op of the tape has read: the
// reverse sweep reads the forward operands as they were when the op ran.
void CCSDContext::check_tape(const std::vector<TapeOp>& tape) const {
  std::vector<int> first_read(allocations_.size(), -1);
  auto read = [&](const double* p, int i) {
    const int k = alloc_index(p);
    if (k < 0) throw std::logic_error("tape operand not owned by the context");
    if (first_read[k] < 0) first_read[k] = i;
  };
  for (size_t i = 0; i < tape.size(); ++i) {
    const TapeOp& op = tape[i];
    const double* written = nullptr;
    switch (op.kind) {
      case TapeOp::kCopy: read(op.src, (int)i); written = op.dst; break;
      case TapeOp::kTau:
      case TapeOp::kTP: read(op.src, (int)i); read(op.src2, (int)i); written = op.dst; break;
      case TapeOp::kTAdd:
        for (int t = 0; t < op.nterms; ++t) {
          if (alloc_index(op.terms[t].p) == alloc_index(op.dst))
            throw std::logic_error("tape: permutation reads its own output");
          read(op.terms[t].p, (int)i);
        }
        written = op.dst;
        break;
      case TapeOp::kGemm: read(op.g.A, (int)i); read(op.g.B, (int)i); written = op.g.C; break;
    }
    const int k = alloc_index(written);
    if (k < 0) throw std::logic_error("tape output not owned by the context");
    if (first_read[k] >= 0 && first_read[k] < (int)i)
      throw std::logic_error("tape: buffer written after it was read (" + op.tag + ")");
  }
}

void CCSDContext::reverse_sweep(const std::vector<TapeOp>& tape, bool integrals) {
  for (size_t i = tape.size(); i-- > 0;) reverse_op(tape[i], integrals);
}

void CCSDContext::reverse_op(const TapeOp& op, bool integrals) {
  const long O = o_, V = v_;
  const long ov = O * V, oovv = ov * ov;
  auto skip = [&](const double* p) { return !integrals && is_integral(p); };
  switch (op.kind) {
    case TapeOp::kCopy: {
      double* dbar = adj(op.dst);
      if (!skip(op.src)) {
        double* sbar = adj(op.src);
        const double* xs[2] = {sbar, dbar};
        const double c[2] = {1.0, 1.0};
        lincomb(sbar, xs, c, 2, op.n, stream_);
      }
      fill(dbar, op.n, 0.0, stream_);
      break;
    }
    case TapeOp::kTAdd: {
      double* obar = adj(op.dst);
      for (int t = 0; t < op.nterms; ++t) {
        const Src& s = op.terms[t];
        if (skip(s.p)) continue;
        int inv[4];
        for (int d = 0; d < 4; ++d) inv[s.perm[d]] = d;
        const TTerm tt = term_of(obar, op.ext, inv, s.alpha);
        long n[4] = {s.ext[0], s.ext[1], s.ext[2], s.ext[3]};
        tadd(adj(s.p), n, 1.0, &tt, 1, stream_);
      }
      if (op.beta == 0.0) {
        fill(obar, vol4(op.ext), 0.0, stream_);
      } else if (op.beta != 1.0) {
        const double* xs[1] = {obar};
        lincomb(obar, xs, &op.beta, 1, vol4(op.ext), stream_);
      }
      break;
    }
    case TapeOp::kGemm: {
      const GemmDesc& g = op.g;
      double* cbar = adj(g.C);
      // d: adjoint contraction into `target` (stride st); a zero stride with batch > 1
      // means the operand is shared by the batch, so the batch is summed.
      auto run = [&](GemmDesc d, double* target, long ld_target, long st) {
        d.C = target;
        d.ldc = ld_target;
        d.beta = 1.0;
        d.batch = g.batch;
        if (g.batch > 1 && st == 0) {
          const size_t need = (size_t)g.batch * d.M * d.N;
          if (adj_ws_n_ < need) {
            if (adj_ws_) MPQC_HIP_CHECK(hipFree(adj_ws_));
            adj_ws_ = nullptr;
            MPQC_HIP_CHECK(hipMalloc(&adj_ws_, need * sizeof(double)));
            adj_ws_n_ = need;
          }
          d.C = adj_ws_;
          d.ldc = d.N;
          d.sC = (long)d.M * d.N;
          d.beta = 0.0;
          mpqc_hip::gemm(d, op.prec, ws_, stream_);
          splitk_reduce(adj_ws_, target, ld_target, 0, d.M, d.N, g.batch, 0, 1, 1.0, stream_);
        } else {
          d.sC = st;
          mpqc_hip::gemm(d, op.prec, ws_, stream_);
        }
      };
      if (!skip(g.A)) {
        double* abar = adj(g.A);
        GemmDesc d;
        if (!g.transA) {  // A^(M x K) += a C^(M x N) op(B)^T(N x K)
          d = make_gemm(false, !g.transB, g.M, g.K, g.N, g.alpha, cbar, g.ldc, g.B, g.ldb, 1.0, nullptr, 0);
          d.sA = g.sC; d.sB = g.sB;
        } else {          // A^(K x M) += a op(B)(K x N) C^^T(N x M)
          d = make_gemm(g.transB, true, g.K, g.M, g.N, g.alpha, g.B, g.ldb, cbar, g.ldc, 1.0, nullptr, 0);
          d.sA = g.sB; d.sB = g.sC;
        }
        run(d, abar, g.lda, g.sA);
      }
      if (!skip(g.B)) {
        double* bbar = adj(g.B);
        GemmDesc d;
        if (!g.transB) {  // B^(K x N) += a op(A)^T(K x M) C^(M x N)
          d = make_gemm(!g.transA, false, g.K, g.N, g.M, g.alpha, g.A, g.lda, cbar, g.ldc, 1.0, nullptr, 0);
          d.sA = g.sA; d.sB = g.sC;
        } else {          // B^(N x K) += a C^^T(N x M) op(A)(M x K)
          d = make_gemm(true, g.transA, g.N, g.K, g.M, g.alpha, cbar, g.ldc, g.A, g.lda, 1.0, nullptr, 0);
          d.sA = g.sC; d.sB = g.sA;
        }
        run(d, bbar, g.ldb, g.sB);
      }
      if (g.beta != 1.0) scale_matrix(cbar, g.batch, g.M, g.N, g.sC, g.ldc, g.beta, stream_);
      break;
    }
    case TapeOp::kTau:
    case TapeOp::kTP: {
      // out^ -> t2^ (a permutation) and t1^ (two contractions with t1)
      double* xbar = adj(op.dst);
      double* t2bar = adj(op.src);
      double* t1bar = adj(op.src2);
      const std::array<long, 4> e_abij{V, V, O, O}, e_aild{V, O, O, V};
      double* M = adj_scratch(oovv);
      if (op.kind == TapeOp::kTau) {
        const int I[4] = {0, 1, 2, 3};
        const TTerm t = term_of(xbar, e_abij, I, 1.0);
        long n[4] = {V, V, O, O};
        tadd(t2bar, n, 1.0, &t, 1, stream_);
        // M(a,i,b,j) = tau^(a,b,i,j) + tau^(b,a,j,i);  t1^(a,i) += M(ai,bj) t1(bj)
        const int p1[4] = {0, 2, 1, 3}, p2[4] = {1, 3, 0, 2};
        const TTerm ts[2] = {term_of(xbar, e_abij, p1, 1.0), term_of(xbar, e_abij, p2, 1.0)};
        long nm[4] = {V, O, V, O};
        tadd(M, nm, 0.0, ts, 2, stream_);
        mpqc_hip::gemm(make_gemm(false, false, (int)ov, 1, (int)ov, 1.0, M, ov, op.src2, 1, 1.0, t1bar, 1),
             prec_, ws_, stream_);
        fill(xbar, oovv, 0.0, stream_);
      } else {
        // t2^(d,a,i,l) += TP^(a,i,l,d) / 2
        const int p0[4] = {3, 0, 1, 2};
        const TTerm t = term_of(xbar, e_aild, p0, 0.5);
        long n[4] = {V, V, O, O};
        tadd(t2bar, n, 1.0, &t, 1, stream_);
        // t1^(d,i) += sum_al TP^(a,i,l,d) t1(a,l):  M1(d,i,a,l) = TP^(a,i,l,d)
        const int p1[4] = {3, 1, 0, 2};
        const TTerm t1t = term_of(xbar, e_aild, p1, 1.0);
        long nm[4] = {V, O, V, O};
        tadd(M, nm, 0.0, &t1t, 1, stream_);
        mpqc_hip::gemm(make_gemm(false, false, (int)ov, 1, (int)ov, 1.0, M, ov, op.src2, 1, 1.0, t1bar, 1),
             prec_, ws_, stream_);
        // t1^(a,l) += sum_id TP^(a,i,l,d) t1(d,i):  M2(a,l,d,i) = TP^(a,i,l,d)
        const int p2[4] = {0, 2, 3, 1};
        const TTerm t2t = term_of(xbar, e_aild, p2, 1.0);
        tadd(M, nm, 0.0, &t2t, 1, stream_);
        mpqc_hip::gemm(make_gemm(false, false, (int)ov, 1, (int)ov, 1.0, M, ov, op.src2, 1, 1.0, t1bar, 1),
             prec_, ws_, stream_);
        fill(xbar, oovv, 0.0, stream_);
      }
      break;
    }
  }
}

// dE/dt for E = 2 <f_ai, t1> + <g_abij, 2 tau(a,b,i,j) - tau(b,a,i,j)>  (ccsd.h:377-379 at R = 0)
void CCSDContext::energy_t_gradient(double* g1, double* g2) {
  const long O = o_, V = v_, ov = O * V;
  const std::array<long, 4> e_abij{V, V, O, O};
  // dE/dtau = 2 g(a,b,i,j) - g(b,a,i,j) -> g2;  g1 = 2 f_ai + (g2 + g2(b,a,j,i)) . t1
  const int I[4] = {0, 1, 2, 3}, sw[4] = {1, 0, 2, 3};
  const TTerm tg[2] = {term_of(g_abij_, e_abij, I, 2.0), term_of(g_abij_, e_abij, sw, -1.0)};
  long n[4] = {V, V, O, O};
  tadd(g2, n, 0.0, tg, 2, stream_);
  const int p1[4] = {0, 2, 1, 3}, p2[4] = {1, 3, 0, 2};
  const TTerm ts[2] = {term_of(g2, e_abij, p1, 1.0), term_of(g2, e_abij, p2, 1.0)};
  long nm[4] = {V, O, V, O};
  double* M = adj_scratch(ov * ov);
  tadd(M, nm, 0.0, ts, 2, stream_);
  const TTerm tf = term_of(f_ai_, {1, 1, V, O}, I, 2.0);
  long n1[4] = {1, 1, V, O};
  tadd(g1, n1, 0.0, &tf, 1, stream_);
  mpqc_hip::gemm(make_gemm(false, false, (int)ov, 1, (int)ov, 1.0, M, ov, t1_, 1, 1.0, g1, 1), prec_,
                 ws_, stream_);
}

void CCSDContext::jacobian_transpose(const double* y1, const double* y2, double* x1, double* x2) {
  const long n1 = (long)o_ * v_, n2 = n1 * n1;
  TapePrecision tp(*this);
  record_residual();
  zero_adjoints();
  MPQC_HIP_CHECK(hipMemcpyAsync(adj(r1_), y1, n1 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipMemcpyAsync(adj(r2_), y2, n2 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  reverse_sweep(res_tape_, false);
  MPQC_HIP_CHECK(hipMemcpyAsync(x1, adj(t1_), n1 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipMemcpyAsync(x2, adj(t2_), n2 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
}

void CCSDContext::jacobian_transpose_io(const double* y1, const double* y2, double* x1, double* x2,
                                        int layout, int memspace) {
  const long n1 = (long)o_ * v_, n2 = n1 * n1;
  double* dy1 = adj_hist(100, n1);
  double* dy2 = adj_hist(101, n2);
  double* dx1 = adj_hist(102, n1);
  double* dx2 = adj_hist(103, n2);
  upload(dy1, y1, {kNone, kNone, kVir, kOcc}, layout, memspace);
  upload(dy2, y2, {kVir, kVir, kOcc, kOcc}, layout, memspace);
  jacobian_transpose(dy1, dy2, dx1, dx2);
  download(x1, dx1, {kNone, kNone, kVir, kOcc}, layout, memspace);
  download(x2, dx2, {kVir, kVir, kOcc, kOcc}, layout, memspace);
}

// Lambda equations (dR/dt)^T lambda = -dE/dt, solved with the amplitude
// equations' preconditioner: lambda += (J^T lambda + dE/dt) / D, D = e_i - e_a
// (e_i + e_j - e_a - e_b), accelerated by DIIS on the same residual.
int CCSDContext::lambda_solve(int max_iter, double target_precision, int verbose) {
  if (!have_eps_) throw std::logic_error("orbital energies not set");
  if (max_iter < 1) throw std::invalid_argument("max_iter must be positive");
  const long n1 = (long)o_ * v_, n2 = n1 * n1;
  lambda_ready_ = gradient_ready_ = false;
  TapePrecision tp(*this);
  record_residual();
  if (!lam1_) {
    lam1_ = alloc(n1); lam2_ = alloc(n2); gt1_ = alloc(n1); gt2_ = alloc(n2);
  }
  energy_t_gradient(gt1_, gt2_);
  fill(lam1_, n1, 0.0, stream_);
  fill(lam2_, n2, 0.0, stream_);
  const int ndiis = 8;
  std::vector<double*> hx1, hx2, he1, he2;
  for (int k = 0; k < ndiis; ++k) {
    hx1.push_back(adj_hist(4 * k + 0, n1)); hx2.push_back(adj_hist(4 * k + 1, n2));
    he1.push_back(adj_hist(4 * k + 2, n1)); he2.push_back(adj_hist(4 * k + 3, n2));
  }
  double* res1 = adj_hist(4 * ndiis + 0, n1);
  double* res2 = adj_hist(4 * ndiis + 1, n2);
  std::vector<double> B((size_t)ndiis * ndiis, 0.0);
  int nvec = 0, head = 0, iter = 0;
  bool converged = false;
  for (iter = 0; iter < max_iter; ++iter) {
    // J^T lambda
    zero_adjoints();
    MPQC_HIP_CHECK(hipMemcpyAsync(adj(r1_), lam1_, n1 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    MPQC_HIP_CHECK(hipMemcpyAsync(adj(r2_), lam2_, n2 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    reverse_sweep(res_tape_, false);
    {
      const double* x1[2] = {adj(t1_), gt1_};
      const double c[2] = {1.0, 1.0};
      lincomb(res1, x1, c, 2, n1, stream_);
      // By default the t2 part is projected onto amplitudes with t2(a,b,i,j) = t2(b,a,j,i).
      // For integrals with the permutational symmetry of real orbitals the residual keeps
      // t2 in that subspace (ccsd_r1_r2.h:260 symmetrises, and the ladders added after it
      // are symmetric), so the Lambda equations only need to hold there; the
      // antisymmetric block of the Jacobian lacks the orbital-energy diagonal, which makes
      // the full-space solve converge slowly.  With "lambda_symmetric" = 0 the full
      // space is solved: exact also for perturbations that break the integrals'
      // symmetry (which take t2 out of the subspace).
      const long O = o_, V = v_;
      const int I[4] = {0, 1, 2, 3}, sw[4] = {1, 0, 3, 2};
      TTerm ts[3];
      long e[4] = {V, V, O, O};
      const int nt = lambda_symmetric_ ? 3 : 2;
      for (int k = 0; k < nt; ++k) {
        const bool g = lambda_symmetric_ ? k == 2 : k == 1;
        ts[k].p = g ? gt2_ : adj(t2_);
        permuted_strides(e, (lambda_symmetric_ && k == 1) ? sw : I, ts[k].s);
        ts[k].alpha = (lambda_symmetric_ && !g) ? 0.5 : 1.0;
        ts[k].vol = n2;
      }
      tadd(res2, e, 0.0, ts, nt, stream_);
    }
    dot(res1, res1, n1, scalars_ + 0, partials_, stream_);
    dot(res2, res2, n2, scalars_ + 1, partials_, stream_);
    double h[2];
    MPQC_HIP_CHECK(hipMemcpyAsync(h, scalars_, 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
    const double err = std::sqrt(h[0] + h[1]) / (double)(n1 + n2);
    if (!std::isfinite(err)) throw std::range_error("non-finite Lambda residual");
    if (verbose) std::printf("  lambda %3d  residual %.6e\n", iter + 1, err);
    if (err <= target_precision) {
      converged = true;
      break;
    }
    // DIIS history: error = res, x = lambda + res / D
    const int slot = head;
    head = (head + 1) % ndiis;
    if (nvec < ndiis) ++nvec;
    MPQC_HIP_CHECK(hipMemcpyAsync(he1[slot], res1, n1 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    MPQC_HIP_CHECK(hipMemcpyAsync(he2[slot], res2, n2 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    jacobi_update(lam1_, lam2_, res1, res2, eo_, ev_, o_, v_, stream_);
    MPQC_HIP_CHECK(hipMemcpyAsync(hx1[slot], lam1_, n1 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    MPQC_HIP_CHECK(hipMemcpyAsync(hx2[slot], lam2_, n2 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    {
      std::vector<const double*> e1, e2;
      for (int k = 0; k < ndiis; ++k) { e1.push_back(he1[k]); e2.push_back(he2[k]); }
      multi_dot2(e1.data(), e2.data(), nvec, he1[slot], he2[slot], n1, n2, scalars_ + 16, partials_,
                 stream_);
      double d[16];
      MPQC_HIP_CHECK(hipMemcpyAsync(d, scalars_ + 16, nvec * sizeof(double), hipMemcpyDeviceToHost, stream_));
      MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
      for (int k = 0; k < nvec; ++k) B[(size_t)k * ndiis + slot] = B[(size_t)slot * ndiis + k] = d[k];
    }
    if (nvec >= 2) {
      // min |sum c_k e_k| subject to sum c_k = 1 over the vectors in age order; as in the
      // amplitude DIIS (TA::DIIS), the oldest vectors are dropped while the bordered
      // system (normalised by its first diagonal element) is singular to 1e-15
      std::vector<int> age(nvec);
      for (int k = 0; k < nvec; ++k) age[k] = (slot - (nvec - 1) + k + ndiis) % ndiis;
      const double zero_determinant = 1.0e-15, zero_norm = 1.0e-10;
      int nskip = 0;
      double absdet = 0.0;
      std::vector<double> c;
      while (true) {
        const int r = nvec - nskip + 1;
        std::vector<double> A((size_t)r * r, 0.0), rhs(r, 0.0);
        for (int i = 1; i < r; ++i) { A[i] = -1.0; A[(size_t)i * r] = -1.0; }
        rhs[0] = -1.0;
        const double b00 = B[(size_t)age[nskip] * ndiis + age[nskip]];
        const double norm = std::fabs(b00) > zero_norm ? 1.0 / std::fabs(b00) : 1.0;
        for (int i = 1; i < r; ++i)
          for (int j = 1; j < r; ++j)
            A[(size_t)i * r + j] = B[(size_t)age[i - 1 + nskip] * ndiis + age[j - 1 + nskip]] * norm;
        absdet = colpiv_qr_solve(r, A, rhs, c);
        if (!(absdet < zero_determinant && nskip + 1 < nvec)) break;
        ++nskip;
      }
      if (absdet >= zero_determinant) {
        std::vector<const double*> x1, x2;
        std::vector<double> cc;
        for (int k = nskip; k < nvec; ++k) {
          x1.push_back(hx1[age[k]]);
          x2.push_back(hx2[age[k]]);
          cc.push_back(c[k - nskip + 1]);
        }
        lincomb(lam1_, x1.data(), cc.data(), (int)cc.size(), n1, stream_);
        lincomb(lam2_, x2.data(), cc.data(), (int)cc.size(), n2, stream_);
      }
    }
  }
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  if (!converged) throw std::runtime_error("Lambda equations did not converge");
  lambda_ready_ = true;
  return iter + 1;
}

void CCSDContext::energy_gradient() {
  if (!lambda_ready_) throw std::logic_error("solve the Lambda equations first");
  const long O = o_, V = v_, n1 = O * V, n2 = n1 * n1;
  TapePrecision tp(*this);
  zero_adjoints();
  MPQC_HIP_CHECK(hipMemcpyAsync(adj(r1_), lam1_, n1 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipMemcpyAsync(adj(r2_), lam2_, n2 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  reverse_sweep(res_tape_, true);
  // explicit dependence of E on f_ai and g_abij
  const int I[4] = {0, 1, 2, 3}, sw[4] = {1, 0, 2, 3};
  const TTerm tf = term_of(t1_, {1, 1, V, O}, I, 2.0);
  long nf[4] = {1, 1, V, O};
  tadd(adj(f_ai_), nf, 1.0, &tf, 1, stream_);
  const TTerm tg[2] = {term_of(tau_, {V, V, O, O}, I, 2.0), term_of(tau_, {V, V, O, O}, sw, -1.0)};
  long ng[4] = {V, V, O, O};
  tadd(adj(g_abij_), ng, 1.0, tg, 2, stream_);
  reverse_sweep(prepare_tape_, true);
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  gradient_ready_ = true;
}

void CCSDContext::get_lambda(double* l1, double* l2, int layout, int memspace) {
  if (!lambda_ready_) throw std::logic_error("solve the Lambda equations first");
  if (l1) download(l1, lam1_, {kNone, kNone, kVir, kOcc}, layout, memspace);
  if (l2) download(l2, lam2_, {kVir, kVir, kOcc, kOcc}, layout, memspace);
}

void CCSDContext::get_integral_gradient(int which, double* out, int layout, int memspace) {
  if (!gradient_ready_) throw std::logic_error("compute the energy gradient first");
  if (!out) throw std::invalid_argument("null output buffer");
  const int O = kOcc, V = kVir, X = kNone;
  const double* src = nullptr;
  std::array<int, 4> modes{};
  switch (which) {
    case 0: src = Fia_; modes = {X, X, O, V}; break;
    case 1: src = Fij_; modes = {X, X, O, O}; break;
    case 2: src = Fab_; modes = {X, X, V, V}; break;
    case 3: src = Gijab_; modes = {O, O, V, V}; break;
    case 4: src = Gijkl_; modes = {O, O, O, O}; break;
    case 5: src = Giajb_; modes = {O, V, O, V}; break;
    case 6: src = Giabc_; modes = {O, V, V, V}; break;
    case 7: src = Gijka_; modes = {O, O, O, V}; break;
    case 8: src = Gabcd_; modes = {V, V, V, V}; break;
    default: throw std::invalid_argument("unknown integral block");
  }
  download(out, adj(src), modes, layout, memspace);
}

}  // namespace mpqc_hip
