// Device-resident closed-shell CCSD iteration.
//
// One residual evaluation is a fixed sequence of GEMMs on the matrix cores and
// permuted linear combinations streaming through HBM.  The order and the
// coefficients follow src/mpqc/chemistry/qc/lcao/cc/ccsd_r1_r2.h:70-293 term by
// term; the comments give the reference line each step implements.  Operands
// are brought into GEMM layout explicitly (integral layouts once per solve in
// prepare(), amplitude layouts once per iteration), so every contraction runs as
// a dense row-major GEMM with a contiguous K dimension.
#include "ccsd_ctx.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

namespace mpqc_hip {

namespace {
constexpr int kOcc = 0, kVir = 1, kAux = 2, kNone = -1;
}  // namespace

// Column-pivoted Householder QR solve of a small dense system (row-major A).
// Returns |det A|; x gets the rank-truncated solution (components of zero
// pivots are zero).  Mirrors the solve used for the DIIS system.
double colpiv_qr_solve(int n, std::vector<double> A, std::vector<double> b, std::vector<double>& x) {
  auto at = [&](int r, int c) -> double& { return A[(size_t)r * n + c]; };
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<double> diag(n, 0.0);
  for (int k = 0; k < n; ++k) {
    // pivot: column with the largest remaining norm
    int p = k;
    double best = -1.0;
    for (int j = k; j < n; ++j) {
      double s = 0.0;
      for (int r = k; r < n; ++r) s += at(r, j) * at(r, j);
      if (s > best) { best = s; p = j; }
    }
    if (p != k) {
      for (int r = 0; r < n; ++r) std::swap(at(r, k), at(r, p));
      std::swap(perm[k], perm[p]);
    }
    // Householder reflector for column k, rows k..n-1
    double norm = 0.0;
    for (int r = k; r < n; ++r) norm += at(r, k) * at(r, k);
    norm = std::sqrt(norm);
    if (norm == 0.0) { diag[k] = 0.0; continue; }
    const double alpha = at(k, k) > 0 ? -norm : norm;
    std::vector<double> vh(n, 0.0);
    for (int r = k; r < n; ++r) vh[r] = at(r, k);
    vh[k] -= alpha;
    double vn = 0.0;
    for (int r = k; r < n; ++r) vn += vh[r] * vh[r];
    if (vn > 0.0) {
      for (int j = k; j < n; ++j) {
        double s = 0.0;
        for (int r = k; r < n; ++r) s += vh[r] * at(r, j);
        s = 2.0 * s / vn;
        for (int r = k; r < n; ++r) at(r, j) -= s * vh[r];
      }
      double s = 0.0;
      for (int r = k; r < n; ++r) s += vh[r] * b[r];
      s = 2.0 * s / vn;
      for (int r = k; r < n; ++r) b[r] -= s * vh[r];
    }
    diag[k] = at(k, k);
  }
  double absdet = 1.0, maxpiv = 0.0;
  for (int k = 0; k < n; ++k) {
    absdet *= std::fabs(diag[k]);
    maxpiv = std::max(maxpiv, std::fabs(diag[k]));
  }
  const double thresh = maxpiv * 2.220446049250313e-16 * n;
  int rank = 0;
  for (int k = 0; k < n; ++k)
    if (std::fabs(diag[k]) > thresh) ++rank;
  std::vector<double> z(n, 0.0);
  for (int k = rank - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < rank; ++j) s -= at(k, j) * z[j];
    z[k] = s / at(k, k);
  }
  x.assign(n, 0.0);
  for (int k = 0; k < n; ++k) x[perm[k]] = z[k];
  return absdet;
}

// ---------------------------------------------------------------- lifecycle
CCSDContext::CCSDContext(int o, int v, const int64_t* occ_hm, int n_occ_tiles, const int64_t* vir_hm,
                         int n_vir_tiles, int precision, bool with_solver)
    : o_(o), v_(v) {
  try {
    construct(occ_hm, n_occ_tiles, vir_hm, n_vir_tiles, precision, with_solver);
  } catch (...) {
    release();
    throw;
  }
}

void CCSDContext::release() {
  if (stream_) hipStreamSynchronize(stream_);
  drop_residual_graph();
  for (auto& pe : pending_) {
    hipEventDestroy(pe.a);
    hipEventDestroy(pe.b);
  }
  pending_.clear();
  for (auto e : event_pool_) hipEventDestroy(e);
  event_pool_.clear();
  for (double* p : adj_) if (p) hipFree(p);
  adj_.clear();
  if (adj_ws_) hipFree(adj_ws_);
  adj_ws_ = nullptr;
  adj_ws_n_ = 0;
  if (adj_scratch_) hipFree(adj_scratch_);
  adj_scratch_ = nullptr;
  adj_scratch_n_ = 0;
  for (double* p : allocations_) hipFree(p);
  allocations_.clear();
  alloc_sizes_.clear();
  for (auto& pa : plane_allocs_) hipFree(pa.first);
  plane_allocs_.clear();
  if (modeinfo_) hipFree(modeinfo_);
  modeinfo_ = nullptr;
  if (staging_) hipFree(staging_);
  staging_ = nullptr;
  if (stream_) hipStreamDestroy(stream_);
  stream_ = nullptr;
}

void CCSDContext::construct(const int64_t* occ_hm, int n_occ_tiles, const int64_t* vir_hm,
                            int n_vir_tiles, int precision, bool with_solver) {
  const int o = o_, v = v_;
  if (o <= 0 || v <= 0) throw std::invalid_argument("o and v must be positive");
  if ((long)v * v * o * o >= (1L << 31)) throw std::invalid_argument("v^2 o^2 must be < 2^31");
  if (precision < 0 || precision > 4) throw std::invalid_argument("unknown precision");
  prec_ = (GemmPrecision)precision;
  auto take_hm = [](const int64_t* hm, int nt, int n, std::vector<int64_t>& out) {
    if (hm == nullptr) { out = {0, n}; return; }
    if (nt <= 0) throw std::invalid_argument("tile count must be positive");
    out.assign(hm, hm + nt + 1);
    if (out.front() != 0 || out.back() != n) throw std::invalid_argument("hashmarks must span [0, n]");
    for (int t = 0; t < nt; ++t)
      if (out[t + 1] <= out[t]) throw std::invalid_argument("hashmarks must increase");
  };
  take_hm(occ_hm, n_occ_tiles, o, occ_hm_);
  take_hm(vir_hm, n_vir_tiles, v, vir_hm_);
  MPQC_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));

  const long O = o, V = v;
  const long ov = O * V, oo = O * O, vv = V * V;
  const long oovv = oo * vv, ooov = oo * ov, ovvv = ov * vv, oooo = oo * oo;
  Fia_ = alloc(ov); Fij_ = alloc(oo); Fab_ = alloc(vv);
  Gijab_ = alloc(oovv); Gijkl_ = alloc(oooo); Giajb_ = alloc(oovv); Giabc_ = alloc(ovvv);
  Gijka_ = alloc(ooov);
  // Gabcd allocated on upload (largest tensor)
  f_ai_ = alloc(ov); g_abij_ = alloc(oovv); Gbar_ = alloc(oovv); Gbar_cKLD_ = alloc(oovv);
  Gbar_kcld_ = alloc(oovv); P6_ = alloc(oovv); Gbar_akcd_ = alloc(ovvv); P8_ = alloc(ooov);
  S11_ = alloc(ooov); P13_ = alloc(ooov); P14_ = alloc(ovvv); J0_ = alloc(oovv); K0_ = alloc(oovv);
  X18_ = alloc(ooov); X19_ = alloc(ooov); GA_ = alloc(oovv); GB_ = alloc(oovv); GC_ = alloc(oovv);
  eo_ = alloc(O); ev_ = alloc(V);
  t1_ = alloc(ov); t2_ = alloc(oovv); tau_ = alloc(oovv); r1_ = alloc(ov); r2_ = alloc(oovv);
  FIJ_ = alloc(oo); FAB_ = alloc(vv);
  Xoo_ = alloc(oo); Zoo_ = alloc(oo); hkc_ = alloc(ov); hT_ = alloc(ov); t1T_ = alloc(ov);
  tau_akld_ = alloc(oovv); tau_ilab_ = alloc(oovv); Y5_ = alloc(oovv); tau_kcdi_ = alloc(oovv);
  tau_ackl_ = alloc(oovv); Rd_ = alloc(oovv); Q_ = alloc(ooov); M12_ = alloc(ooov);
  S_ = alloc(ooov); gki_ = alloc(oo); gac_ = alloc(vv); t2_abjk_ = alloc(oovv); X16_ = alloc(oovv);
  TP_ = alloc(oovv); t2P_ = alloc(oovv); J_ = alloc(oovv); K_ = alloc(oovv); Zj_ = alloc(oovv);
  Zk_ = alloc(oovv); Xjk_ = alloc(oovv); Y1_ = alloc(oovv); Y2_ = alloc(oovv); RR_ = alloc(oovv);
  tmpR_ = alloc(oovv); Y23_ = alloc(oooo); W_ = alloc(oooo); tmp24_ = alloc(ooov);
  // split-K workspace: enough for 64 slices of the largest small-output GEMM, and
  // for 16 slices of an (ov) x (ov) or v^2 x o^2 output up to 512 MiB (small
  // systems, where those outputs are too few tiles to fill the chip)
  const long ws_elems = std::max<long>({64L * std::max({vv, ov, oo * oo, 4096L}), 1L << 20,
                                        std::min<long>(16L * std::max(ov * ov, vv * oo), 64L << 20)});
  ws_.ptr = alloc(ws_elems);
  ws_.bytes = ws_elems * sizeof(double);
  partials_ = alloc(kReducePartials);
  scalars_ = alloc(64);
  if (with_solver) solver_init(1, 8, 0.0, 1, 1);
  MPQC_HIP_CHECK(hipMemsetAsync(t1_, 0, ov * sizeof(double), stream_));
  MPQC_HIP_CHECK(hipMemsetAsync(t2_, 0, oovv * sizeof(double), stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
}

CCSDContext::~CCSDContext() { release(); }

double* CCSDContext::alloc(size_t n) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw std::bad_alloc();
  }
  allocations_.push_back((double*)p);
  alloc_sizes_.push_back(std::max<size_t>(n, 1));
  return (double*)p;
}

void CCSDContext::check_access(const double* p, long max_off, const char* what) const {
  for (size_t k = 0; k < allocations_.size(); ++k) {
    const double* base = allocations_[k];
    if (p >= base && p < base + alloc_sizes_[k]) {
      if (max_off < 0 || (size_t)(p - base) + (size_t)max_off >= alloc_sizes_[k])
        throw std::logic_error(std::string("out-of-bounds operand in ") + what);
      return;
    }
  }
  throw std::logic_error(std::string("operand not owned by the context in ") + what);
}

void CCSDContext::set_option(const std::string& key, double value) {
  ++gen_;  // a captured residual reflects the options it was captured under
  if (key == "graph") {  // 1: replay the FP64/FP32 residual from a HIP graph
    if (value != 0.0 && value != 1.0) throw std::invalid_argument("graph must be 0 or 1");
    graph_mode_ = (int)value;
  } else if (key == "fov_sign") {
    if (value != 1.0 && value != -1.0) throw std::invalid_argument("fov_sign must be +1 or -1");
    fov_sign_ = value;
  } else if (key == "lambda_symmetric") {
    lambda_symmetric_ = value != 0.0;
    lambda_ready_ = false;
  } else if (key == "kernel_timing") {
    kernel_timing_ = value != 0.0;
  } else if (key == "method") {  // 0 standard, 1 df, 2 direct, 3 direct_df (ccsd.h:73, 94-99)
    if (value != 0.0 && value != 1.0 && value != 2.0 && value != 3.0)
      throw std::invalid_argument("method must be 0 (standard), 1 (df), 2 (direct) or 3 (direct_df)");
    if ((value == 1.0 || value == 3.0) && naux_ == 0)
      throw std::invalid_argument("methods df and direct_df need set_df_basis first");
    if ((value == 2.0 || value == 3.0) && !Xao_)
      throw std::invalid_argument("methods direct and direct_df need set_direct_ao first");
    df_ = value == 1.0 || value == 3.0;
    direct_ = value == 2.0 || value == 3.0;
    prepared_ = false;
  } else if (key == "reduced_abcd_memory") {  // ccsd.h:102, used at :292
    reduced_abcd_ = value != 0.0;
    prepared_ = false;
  } else if (key == "plane_tile") {
    // process-wide tile of the plane contractions: 0 automatic, else rows * 1000 + columns
    const int t = (int)value;
    if (t != value || t < 0) throw std::invalid_argument("plane_tile must be 0 or rows * 1000 + columns");
    set_plane_tile_override(t / 1000, t % 1000);
  } else if (key == "plane_split_cap") {  // process-wide: most K splits of a small-grid plane GEMM
    if ((int)value != value) throw std::invalid_argument("plane_split_cap must be an integer");
    set_plane_split_cap((int)value);
  } else if (key == "plane_tile_order") {  // process-wide: 0 XCD ranges row-major, 1 plain, 2 XCD ranges column-major
    set_plane_tile_order((int)value);
  } else if (key == "plane_lds_stages") {  // process-wide
    set_plane_lds_stages((int)value);
  } else if (key == "solve_restart") {
    // 1: solve() starts from the initial guess (ccsd.h:331-352); 0: from the context's
    // amplitudes (e.g. those of a reduced-precision solve, refined in FP64); 2: from them
    // after projecting t2 onto t2(a,b,i,j) = t2(b,a,j,i), the symmetry of the exact
    // amplitudes for real orbitals (a reduced-precision solve leaves a small
    // antisymmetric part, which FP64 iterations remove only slowly)
    const int r = (int)value;
    if (r != value || r < 0 || r > 2) throw std::invalid_argument("solve_restart must be 0, 1 or 2");
    solve_restart_ = r;
  } else if (key == "plane_dl") {  // process-wide: 0 deep-pipeline kernel where planned, 1 never
    if ((int)value != value) throw std::invalid_argument("plane_dl must be an integer");
    set_plane_dl_mode((int)value);
  } else if (key == "thin_k") {  // process-wide: 0 streaming kernel for FP64 K <= 64 contractions where planned, 1 never
    if ((int)value != value) throw std::invalid_argument("thin_k must be an integer");
    set_thin_k_mode((int)value);
  } else if (key == "eom_max_vector") {  // eom_ccsd.h:33
    if (value < 2 || value > 64 || value != (int)value)
      throw std::invalid_argument("eom_max_vector must be an integer in [2, 64]");
    eom_max_vector_ = (int)value;
  } else if (key == "eom_vector_threshold") {  // eom_ccsd.h:34
    if (!(value >= 0.0)) throw std::invalid_argument("eom_vector_threshold must be >= 0");
    eom_vector_threshold_ = value;
  } else if (key == "eom_extra_roots") {
    if (value < 0 || value > 64 || value != (int)value)
      throw std::invalid_argument("eom_extra_roots must be an integer in [0, 64]");
    eom_extra_roots_ = (int)value;
  } else if (key == "pp_symmetric") {
    pp_symmetric_ = value != 0.0;
    planes_ready_ = false;
  } else if (key == "eom_cis_guess") {
    eom_cis_guess_ = value != 0.0;
  } else if (key == "eom_cis_diagonal") {
    eom_cis_diagonal_ = value != 0.0;
  } else if (key == "eom_skip_converged") {
    eom_skip_converged_ = value != 0.0;
  } else if (key == "eom_n_guess") {
    if (value < 0 || value != (int)value) throw std::invalid_argument("eom_n_guess must be a non-negative integer");
    eom_n_guess_ = (int)value;
  } else if (key == "triples_block_bytes") {
    if (!(value >= 1 << 20)) throw std::invalid_argument("triples_block_bytes must be at least 1 MiB");
    triples_block_bytes_ = value;
  } else if (key == "df_block_bytes") {
    if (!(value >= 1 << 20)) throw std::invalid_argument("df_block_bytes must be at least 1 MiB");
    df_block_bytes_ = value;
    df_block_ = 0;
    release_planes();  // the df ladder's block planes are sized by the block
  } else if (key == "precision") {
    const int p = (int)value;
    if (p < 0 || p > 4 || p != value) throw std::invalid_argument("unknown precision");
    if ((GemmPrecision)p != prec_) {
      prec_ = (GemmPrecision)p;
      planes_ready_ = false;
      release_planes();  // the plane kind may change: every plane buffer is re-allocated
    }
  } else {
    throw std::invalid_argument("unknown option: " + key);
  }
}

// ---------------------------------------------------------------- transfers
void CCSDContext::upload(double* dst, const double* src, const std::array<int, 4>& modes, int layout,
                         int memspace) {
  long n[4];
  long total = 1;
  for (int d = 0; d < 4; ++d) {
    n[d] = modes[d] == kOcc ? o_ : modes[d] == kVir ? v_ : modes[d] == kAux ? naux_ : 1;
    total *= n[d];
  }
  if (memspace != 0 && memspace != 1) throw std::invalid_argument("memspace must be host or device");
  // a caller's device buffer may still be written by work it queued on other streams
  // (the context's stream does not wait for the null stream): let that finish first
  if (memspace) MPQC_HIP_CHECK(hipDeviceSynchronize());
  if (layout == 0) {
    MPQC_HIP_CHECK(hipMemcpyAsync(dst, src, total * sizeof(double),
                                  memspace ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
    return;
  }
  if (layout != 1) throw std::invalid_argument("layout must be dense or TA tiles");
  // tiled: stage raw tiles on the device, then gather into the dense layout
  const double* dsrc = src;
  if (!memspace) {
    if (staging_n_ < (size_t)total) {
      if (staging_) MPQC_HIP_CHECK(hipFree(staging_));
      staging_ = nullptr;
      MPQC_HIP_CHECK(hipMalloc(&staging_, total * sizeof(double)));
      staging_n_ = total;
    }
    MPQC_HIP_CHECK(hipMemcpyAsync(staging_, src, total * sizeof(double), hipMemcpyHostToDevice, stream_));
    dsrc = staging_;
  }
  // mode info
  const long* hms[4];
  int nts[4];
  for (int d = 0; d < 4; ++d) {
    if (modes[d] == kOcc) { hms[d] = occ_hm_.data(); nts[d] = (int)occ_hm_.size() - 1; }
    else if (modes[d] == kVir) { hms[d] = vir_hm_.data(); nts[d] = (int)vir_hm_.size() - 1; }
    else if (modes[d] == kAux) { hms[d] = aux_hm_.data(); nts[d] = (int)aux_hm_.size() - 1; }
    else { hms[d] = nullptr; nts[d] = 1; }
  }
  std::vector<long> mi(2 * (n[0] + n[1] + n[2] + n[3]));
  static_assert(sizeof(long) == sizeof(int64_t), "long must be 64-bit");
  make_tile_modeinfo(mi.data(), n, (const long* const*)hms, nts);
  if (modeinfo_n_ < mi.size()) {
    if (modeinfo_) MPQC_HIP_CHECK(hipFree(modeinfo_));
    modeinfo_ = nullptr;
    MPQC_HIP_CHECK(hipMalloc(&modeinfo_, mi.size() * sizeof(long)));
    modeinfo_n_ = mi.size();
  }
  MPQC_HIP_CHECK(hipMemcpyAsync(modeinfo_, mi.data(), mi.size() * sizeof(long), hipMemcpyHostToDevice, stream_));
  tiles_to_dense(dst, dsrc, n, modeinfo_, stream_);
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
}

void CCSDContext::download(double* dst, const double* src, const std::array<int, 4>& modes, int layout,
                           int memspace) {
  long n[4];
  long total = 1;
  for (int d = 0; d < 4; ++d) {
    n[d] = modes[d] == kOcc ? o_ : modes[d] == kVir ? v_ : modes[d] == kAux ? naux_ : 1;
    total *= n[d];
  }
  if (memspace != 0 && memspace != 1) throw std::invalid_argument("memspace must be host or device");
  if (memspace) MPQC_HIP_CHECK(hipDeviceSynchronize());  // the caller's pending work on its buffer
  if (layout == 0) {
    MPQC_HIP_CHECK(hipMemcpyAsync(dst, src, total * sizeof(double),
                                  memspace ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream_));
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
    return;
  }
  if (layout != 1) throw std::invalid_argument("layout must be dense or TA tiles");
  const long* hms[4];
  int nts[4];
  for (int d = 0; d < 4; ++d) {
    if (modes[d] == kOcc) { hms[d] = occ_hm_.data(); nts[d] = (int)occ_hm_.size() - 1; }
    else if (modes[d] == kVir) { hms[d] = vir_hm_.data(); nts[d] = (int)vir_hm_.size() - 1; }
    else if (modes[d] == kAux) { hms[d] = aux_hm_.data(); nts[d] = (int)aux_hm_.size() - 1; }
    else { hms[d] = nullptr; nts[d] = 1; }
  }
  std::vector<long> mi(2 * (n[0] + n[1] + n[2] + n[3]));
  make_tile_modeinfo(mi.data(), n, (const long* const*)hms, nts);
  if (modeinfo_n_ < mi.size()) {
    if (modeinfo_) MPQC_HIP_CHECK(hipFree(modeinfo_));
    modeinfo_ = nullptr;
    MPQC_HIP_CHECK(hipMalloc(&modeinfo_, mi.size() * sizeof(long)));
    modeinfo_n_ = mi.size();
  }
  MPQC_HIP_CHECK(hipMemcpyAsync(modeinfo_, mi.data(), mi.size() * sizeof(long), hipMemcpyHostToDevice, stream_));
  double* ddst = dst;
  if (!memspace) {
    if (staging_n_ < (size_t)total) {
      if (staging_) MPQC_HIP_CHECK(hipFree(staging_));
      staging_ = nullptr;
      MPQC_HIP_CHECK(hipMalloc(&staging_, total * sizeof(double)));
      staging_n_ = total;
    }
    ddst = staging_;
  }
  dense_to_tiles(ddst, src, n, modeinfo_, stream_);
  if (!memspace)
    MPQC_HIP_CHECK(hipMemcpyAsync(dst, ddst, total * sizeof(double), hipMemcpyDeviceToHost, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
}

void CCSDContext::set_integral(int which, const double* data, int layout, int memspace) {
  if (data == nullptr) throw std::invalid_argument("null integral buffer");
  const int O = kOcc, V = kVir, X = kNone;
  switch (which) {
    case 0: upload(Fia_, data, {X, X, O, V}, layout, memspace); break;
    case 1: upload(Fij_, data, {X, X, O, O}, layout, memspace); break;
    case 2: upload(Fab_, data, {X, X, V, V}, layout, memspace); break;
    case 3: upload(Gijab_, data, {O, O, V, V}, layout, memspace); break;
    case 4: upload(Gijkl_, data, {O, O, O, O}, layout, memspace); break;
    case 5: upload(Giajb_, data, {O, V, O, V}, layout, memspace); break;
    case 6: upload(Giabc_, data, {O, V, V, V}, layout, memspace); break;
    case 7: upload(Gijka_, data, {O, O, O, V}, layout, memspace); break;
    case 8:
      if (!Gabcd_) Gabcd_ = alloc((size_t)v_ * v_ * v_ * v_);
      upload(Gabcd_, data, {V, V, V, V}, layout, memspace);
      break;
    case 9: case 10: case 11: {  // X(K,a,i), X(K,i,j), X(K,a,b)  (ccsd.h:481-493)
      if (naux_ == 0) throw std::logic_error("three-index integrals need set_df_basis first");
      const int A = kAux;
      if (which == 9) upload(Xai_, data, {X, A, V, O}, layout, memspace);
      else if (which == 10) upload(Xij_, data, {X, A, O, O}, layout, memspace);
      else upload(Xab_, data, {X, A, V, V}, layout, memspace);
      break;
    }
    default: throw std::invalid_argument("unknown integral id");
  }
  have_ints_ |= 1u << which;
  prepared_ = false;
}

void CCSDContext::get_integral(int which, double* out, int layout, int memspace) {
  if (out == nullptr) throw std::invalid_argument("null output buffer");
  if (which < 0 || which > 11 || !((have_ints_ >> which) & 1u)) throw std::logic_error("integral block not set");
  const int O = kOcc, V = kVir, X = kNone, A = kAux;
  switch (which) {
    case 0: download(out, Fia_, {X, X, O, V}, layout, memspace); break;
    case 1: download(out, Fij_, {X, X, O, O}, layout, memspace); break;
    case 2: download(out, Fab_, {X, X, V, V}, layout, memspace); break;
    case 3: download(out, Gijab_, {O, O, V, V}, layout, memspace); break;
    case 4: download(out, Gijkl_, {O, O, O, O}, layout, memspace); break;
    case 5: download(out, Giajb_, {O, V, O, V}, layout, memspace); break;
    case 6: download(out, Giabc_, {O, V, V, V}, layout, memspace); break;
    case 7: download(out, Gijka_, {O, O, O, V}, layout, memspace); break;
    case 8: download(out, Gabcd_, {V, V, V, V}, layout, memspace); break;
    case 9: download(out, Xai_, {X, A, V, O}, layout, memspace); break;
    case 10: download(out, Xij_, {X, A, O, O}, layout, memspace); break;
    default: download(out, Xab_, {X, A, V, V}, layout, memspace); break;
  }
}

void CCSDContext::set_orbital_energies(const double* eo, const double* ev) {
  if (!eo || !ev) throw std::invalid_argument("null orbital energies");
  MPQC_HIP_CHECK(hipMemcpyAsync(eo_, eo, o_ * sizeof(double), hipMemcpyHostToDevice, stream_));
  MPQC_HIP_CHECK(hipMemcpyAsync(ev_, ev, v_ * sizeof(double), hipMemcpyHostToDevice, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  // zero denominators would raise SIGFPE in the reference (mpqc_init.cpp FPE traps)
  for (int i = 0; i < o_; ++i)
    for (int a = 0; a < v_; ++a)
      if (eo[i] - ev[a] == 0.0) throw std::range_error("zero Jacobi denominator e_i - e_a");
  have_eps_ = true;
}

void CCSDContext::solver_init(int diis_start, int n_diis, double damp, int ngroup, int ngroup_nstart) {
  if (n_diis < 1 || n_diis > 16) throw std::invalid_argument("n_diis must be in [1,16]");
  if (ngroup < 1 || diis_start < 0 || damp < 0.0) throw std::invalid_argument("invalid DIIS parameters");
  diis_start_ = diis_start;
  n_diis_ = n_diis;
  diis_damp_ = damp;
  diis_ngroup_ = ngroup;
  diis_group_nstart_ = ngroup_nstart;
  diis_iter_ = 0;
  err_slots_.clear();
  x_slots_.clear();
  B_.assign((size_t)n_diis * n_diis, 0.0);
  const long n1 = (long)o_ * v_, n2 = (long)o_ * o_ * v_ * v_;
  while ((int)hist_e1_.size() < n_diis) {
    hist_e1_.push_back(alloc(n1));
    hist_e2_.push_back(alloc(n2));
    hist_x1_.push_back(alloc(n1));
    hist_x2_.push_back(alloc(n2));
  }
  solver_ready_ = true;
}

void CCSDContext::set_amplitudes(const double* t1, const double* t2, int layout, int memspace) {
  upload(t1_, t1, {kNone, kNone, kVir, kOcc}, layout, memspace);
  upload(t2_, t2, {kVir, kVir, kOcc, kOcc}, layout, memspace);
  build_tau(tau_, t2_, t1_, o_, v_, stream_);
  tau_fresh_ = true;
  tau_amax_fresh_ = false;
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
}

void CCSDContext::get_amplitudes(double* t1, double* t2, int layout, int memspace) {
  if (t1) download(t1, t1_, {kNone, kNone, kVir, kOcc}, layout, memspace);
  if (t2) download(t2, t2_, {kVir, kVir, kOcc, kOcc}, layout, memspace);
}

void CCSDContext::get_residuals(double* r1, double* r2, int layout, int memspace) {
  if (r1) download(r1, r1_, {kNone, kNone, kVir, kOcc}, layout, memspace);
  if (r2) download(r2, r2_, {kVir, kVir, kOcc, kOcc}, layout, memspace);
}

void CCSDContext::get_dressed_fock(double* FIJ, double* FAB) {
  if (FIJ) MPQC_HIP_CHECK(hipMemcpyAsync(FIJ, FIJ_, (size_t)o_ * o_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
  if (FAB) MPQC_HIP_CHECK(hipMemcpyAsync(FAB, FAB_, (size_t)v_ * v_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------- helpers
void CCSDContext::gemm(bool tA, bool tB, int M, int N, int K, double alpha, const double* A, long lda,
                       const double* B, long ldb, double beta, double* C, long ldc, long sA, long sB,
                       long sC, int batch, const char* tag) {
  GemmDesc d = make_gemm(tA, tB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  d.sA = sA; d.sB = sB; d.sC = sC; d.batch = batch;
  // in the split-precision forms only the large contractions use 16-bit planes;
  // everything issued through this path runs in FP64
  const GemmPrecision prec = split_mode() ? GemmPrecision::F64 : prec_;
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) throw std::logic_error("empty GEMM");
  if (lda < (tA ? M : K) || ldb < (tB ? K : N) || ldc < N) throw std::logic_error("GEMM leading dimension too small");
  const long bm = (long)(batch - 1);
  check_access(A, (tA ? (long)(K - 1) * lda + (M - 1) : (long)(M - 1) * lda + (K - 1)) + bm * sA, "gemm A");
  check_access(B, (tB ? (long)(N - 1) * ldb + (K - 1) : (long)(K - 1) * ldb + (N - 1)) + bm * sB, "gemm B");
  check_access(C, (long)(M - 1) * ldc + (N - 1) + bm * sC, "gemm C");
  if (tape_) {
    TapeOp op;
    op.kind = TapeOp::kGemm;
    op.g = d;
    op.prec = prec;
    op.tag = tag ? tag : "gemm";
    tape_->push_back(op);
    if (!tape_execute_) return;
  }
  timed(tag ? tag : "gemm", 2.0 * M * (double)N * K * batch,
        [&] { mpqc_hip::gemm(d, prec, ws_, stream_); });
}

void CCSDContext::copy_d2d(double* dst, const double* src, long n) {
  check_access(dst, n - 1, "copy destination");
  check_access(src, n - 1, "copy source");
  if (tape_) {
    TapeOp op;
    op.kind = TapeOp::kCopy;
    op.dst = dst; op.src = src; op.n = n;
    tape_->push_back(op);
    if (!tape_execute_) return;
  }
  MPQC_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
}

void CCSDContext::op_tau() {
  if (tape_) {
    TapeOp op;
    op.kind = TapeOp::kTau;
    op.dst = tau_; op.src = t2_; op.src2 = t1_;
    tape_->push_back(op);
    if (!tape_execute_) return;
  }
  timed("r:tau", 0.0, [&] { build_tau(tau_, t2_, t1_, o_, v_, stream_); });
}

TProd CCSDContext::tp_product() const {
  TProd pr;
  const long O = o_, V = v_;
  pr.x = t1_;  // t1(d,i): d*o + i over (a,i,l,d)
  pr.sx[0] = 0; pr.sx[1] = 1; pr.sx[2] = 0; pr.sx[3] = O;
  pr.nx = O * V;
  pr.y = t1_;  // t1(a,l): a*o + l
  pr.sy[0] = O; pr.sy[1] = 0; pr.sy[2] = 1; pr.sy[3] = 0;
  pr.ny = O * V;
  pr.alpha = 1.0;
  return pr;
}

void CCSDContext::op_TP() {
  if (tape_) {
    TapeOp op;
    op.kind = TapeOp::kTP;
    op.dst = TP_; op.src = t2_; op.src2 = t1_;
    tape_->push_back(op);
    if (!tape_execute_) return;
  }
  // TP(a,i,l,d) = 0.5 t2(d,a,i,l) + t1(d,i) t1(a,l) through the tiled permute
  const long O = o_, V = v_;
  TTerm t;
  t.p = t2_;
  long ext[4] = {V, V, O, O};
  int perm[4] = {1, 2, 3, 0};
  permuted_strides(ext, perm, t.s);
  t.alpha = 0.5;
  t.vol = V * V * O * O;
  const TProd pr = tp_product();
  long n[4] = {V, O, O, V};
  timed("r:TP", 0.0, [&] { tadd(TP_, n, 0.0, &t, 1, stream_, &pr); });
}

int CCSDContext::timing_slot(const char* tag) {
  const std::string t = tag;
  for (size_t k = 0; k < ktimes_.size(); ++k)
    if (ktimes_[k].tag == t) return (int)k;
  ktimes_.push_back({t, 0, 0.0, 0.0});
  return (int)ktimes_.size() - 1;
}

hipEvent_t CCSDContext::take_event() {
  hipEvent_t e;
  if (!event_pool_.empty()) {
    e = event_pool_.back();
    event_pool_.pop_back();
  } else {
    MPQC_HIP_CHECK(hipEventCreate(&e));
  }
  return e;
}

void CCSDContext::resolve_events() {
  if (pending_.empty()) return;
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  for (auto& pe : pending_) {
    float ms = 0.f;
    MPQC_HIP_CHECK(hipEventElapsedTime(&ms, pe.a, pe.b));
    ktimes_[pe.slot].calls += 1;
    ktimes_[pe.slot].ms += ms;
    ktimes_[pe.slot].flops += pe.flops;
    event_pool_.push_back(pe.a);
    event_pool_.push_back(pe.b);
  }
  pending_.clear();
}

std::vector<CCSDContext::KernelTime> CCSDContext::kernel_times() {
  resolve_events();
  return ktimes_;
}

void CCSDContext::reset_kernel_times() {
  resolve_events();
  ktimes_.clear();
}

int CCSDContext::make_terms(std::array<long, 4> n, std::initializer_list<Src> srcs, TTerm* terms) {
  int nt = 0;
  if (srcs.size() > (size_t)kMaxTerms) throw std::logic_error("too many permutation terms");
  for (const Src& s : srcs) {
    long ext[4] = {s.ext[0], s.ext[1], s.ext[2], s.ext[3]};
    int perm[4] = {s.perm[0], s.perm[1], s.perm[2], s.perm[3]};
    terms[nt].p = s.p;
    permuted_strides(ext, perm, terms[nt].s);
    {
      long maxoff = 0, vol = 1;
      bool seen[4] = {false, false, false, false};
      for (int d = 0; d < 4; ++d) {
        if (perm[d] < 0 || perm[d] > 3 || seen[perm[d]]) throw std::logic_error("bad permutation");
        seen[perm[d]] = true;
        if (ext[perm[d]] != n[d]) throw std::logic_error("permutation extent mismatch");
        maxoff += (n[d] - 1) * terms[nt].s[d];
        vol *= ext[d];
      }
      if (maxoff >= vol) throw std::logic_error("permutation reads past its input");
      check_access(s.p, maxoff, "permutation input");
    }
    terms[nt].alpha = s.alpha;
    terms[nt].vol = s.ext[0] * s.ext[1] * s.ext[2] * s.ext[3];
    ++nt;
  }
  return nt;
}

void CCSDContext::perm_add(double* out, std::array<long, 4> n, double beta,
                           std::initializer_list<Src> srcs, const char* tag, int line) {
  TTerm terms[kMaxTerms];
  check_access(out, n[0] * n[1] * n[2] * n[3] - 1, "perm_add out");
  const int nt = make_terms(n, srcs, terms);
  long nn[4] = {n[0], n[1], n[2], n[3]};
  if (tape_) {
    TapeOp op;
    op.kind = TapeOp::kTAdd;
    op.dst = out; op.ext = n; op.beta = beta; op.nterms = nt;
    int k = 0;
    for (const Src& sc : srcs) op.terms[k++] = sc;
    op.tag = tag ? std::string(tag) : "perm@" + std::to_string(line);
    tape_->push_back(op);
    if (!tape_execute_) return;
  }
  std::string t;
  if (kernel_timing_) t = tag ? std::string(tag) : "perm@" + std::to_string(line);
  timed(t.c_str(), 0.0, [&] { tadd(out, nn, beta, terms, nt, stream_); });
}

// ---------------------------------------------------------------- planes
bool CCSDContext::split_mode() const {
  return prec_ == GemmPrecision::BF16X3 || prec_ == GemmPrecision::F16X3 ||
         prec_ == GemmPrecision::BF16X6;
}

PlaneKind CCSDContext::plane_kind() const {
  switch (prec_) {
    case GemmPrecision::F16X3: return PlaneKind::F16x2;
    case GemmPrecision::BF16X6: return PlaneKind::BF16x3;
    default: return PlaneKind::BF16x2;
  }
}

void CCSDContext::ensure_planes(Planes& P, long n) {
  if (P.p[0] && P.n >= n) return;
  if (P.p[0]) throw std::logic_error("plane buffer re-sized without release_planes()");
  P = alloc_planes(n);
}

// frees every 16-bit plane buffer and forgets every Planes handle that pointed into them
void CCSDContext::release_planes() {
  if (!stream_) return;
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  for (auto& pa : plane_allocs_) hipFree(pa.first);
  plane_allocs_.clear();
  for (Planes* P : {&Gabcd_P_, &Giabc_P_, &Gijab_P_, &GAt_P_, &GBt_P_, &GCt_P_, &Gp_P_sym_, &Gm_P_sym_, &Vp_P_,
                    &Vm_P_, &Sp_P_, &Ap_P_, &GbarT_P_, &Gbar_P_, &Gbar_akcd_P_, &P8T_P_, &GiabcT_P_, &P13_P_,
                    &P14_P_, &tau_P_, &tauT_P_, &TP_P_, &t2P_P_, &Xjk_P_, &K_P_, &Y1t_P_, &Y2t_P_, &WT_P_,
                    &t1T_P_, &gac_P_, &tau_ikcd_P_, &tau_ackl_P_, &Gp_P_})
    *P = Planes();
  planes_ready_ = false;
}

Planes CCSDContext::alloc_planes(long n) {
  Planes P;
  P.n = n;
  for (int q = 0; q < plane_count(plane_kind()); ++q) {
    void* ptr = nullptr;
    MPQC_HIP_CHECK(hipMalloc(&ptr, std::max<long>(n, 8) * sizeof(uint16_t)));
    P.p[q] = (uint16_t*)ptr;
    plane_allocs_.push_back({P.p[q], (size_t)std::max<long>(n, 8)});
  }
  P.scale = alloc(1);
  return P;
}

void CCSDContext::split(const Planes& out, std::array<long, 4> n, std::initializer_list<Src> srcs,
                        const TProd* prod, const char* tag, int line) {
  TTerm terms[kMaxTerms];
  const int nt = make_terms(n, srcs, terms);
  long nn[4] = {n[0], n[1], n[2], n[3]};
  std::string t;
  if (kernel_timing_) t = tag ? std::string(tag) : "split@" + std::to_string(line);
  timed(t.c_str(), 0.0,
        [&] { split_to_planes(plane_kind(), out, nn, terms, nt, partials_, stream_, prod, &amax_); });
}

void CCSDContext::gemmP(int M, int N, int K, double alpha, const Planes& A, long lda, const Planes& B,
                        long ldb, double beta, double* C, long ldc, const char* tag, int amax_slot) {
  if (M <= 0 || N <= 0 || K <= 0) throw std::logic_error("empty GEMM");
  if (lda < K || ldb < K || ldc < N) throw std::logic_error("GEMM leading dimension too small");
  if ((long)(M - 1) * lda + K - 1 >= A.n || (long)(N - 1) * ldb + K - 1 >= B.n)
    throw std::logic_error(std::string("out-of-bounds plane operand in ") + tag);
  check_access(C, (long)(M - 1) * ldc + (N - 1), "gemmP C");
  PlaneGemmDesc d;
  d.kind = plane_kind();
  d.M = M; d.N = N; d.K = K; d.alpha = alpha; d.beta = beta;
  d.A = A; d.lda = lda; d.sA = 0; d.B = B; d.ldb = ldb; d.sB = 0;
  d.C = C; d.ldc = ldc; d.sC = 0; d.batch = 1;
  const bool rec = amax_slot >= 0 && prec_ == GemmPrecision::F16X3 && amax_.dev && amax_.ptr[amax_slot] == C &&
                   ldc == N;
  if (rec) {
    d.cmax = amax_.dev + amax_slot;
    d.cmax_partials = partials_;
  }
  bool done = false;
  timed(tag, 2.0 * M * (double)N * K, [&] { done = gemm_planes(d, pws_, pws_bytes_, stream_); });
  if (rec) amax_.valid[amax_slot] = done;
}

// constant operands of the large contractions, split once per solve
void CCSDContext::prepare_planes() {
  const long O = o_, V = v_;
  const long oovv = O * O * V * V, ovvv = O * V * V * V, vvvv = V * V * V * V, oooo = O * O * O * O;
  // o^2 v^2 and smaller planes: every split form uses them
  for (Planes* P : {&Gijab_P_, &GAt_P_, &GBt_P_, &GCt_P_, &tau_P_, &tauT_P_, &TP_P_, &t2P_P_, &Xjk_P_, &K_P_,
                    &Y1t_P_, &Y2t_P_, &GbarT_P_, &Gbar_P_, &tau_ikcd_P_, &tau_ackl_P_})
    ensure_planes(*P, oovv);
  ensure_planes(WT_P_, oooo);
  ensure_planes(P8T_P_, O * O * O * V);
  ensure_planes(P13_P_, O * O * O * V);
  ensure_planes(t1T_P_, O * V);
  ensure_planes(gac_P_, V * V);
  // o v^3 planes: only where Giabc is an input (the df method builds those terms from the factors)
  if (!df_)
    for (Planes* P : {&Giabc_P_, &GiabcT_P_, &Gbar_akcd_P_, &P14_P_}) ensure_planes(*P, ovvv);
  {  // split-K workspace for the largest plane contraction shapes
    const int o = o_, v = v_;
    size_t need = 0;
    const int shapes[][3] = {{v * v, o * o, v * v}, {o * v, o * o, v * v}, {o * o, o * o, v * v},
                             {v * v, o * o, o * o}, {o * v, o * v, o * v}, {v, v, o * o * v},
                             {o, o, o * v * v},     {v, o, o * v * v},     {v, o, v * o * o},
                             {o * v * v, o, v},     {v, v * o * o, v},     {v * v, 1, o * v},
                             {o * o, 1, o * v}};
    for (auto& sh : shapes) need = std::max(need, gemm_planes_workspace_bytes(sh[0], sh[1], sh[2], 1));
    // the pair-symmetric ladder: v(v+1)/2 x o^2 x (v(v+1)/2 padded to 64)
    const int nvs = v * (v + 1) / 2;
    need = std::max(need, gemm_planes_workspace_bytes(nvs, o * o, (nvs + 63) / 64 * 64, 1));
    if (need > pws_bytes_) {
      const size_t n = (need + sizeof(double) - 1) / sizeof(double);
      pws_ = alloc(n);
      pws_bytes_ = n * sizeof(double);
    }
  }
  const std::array<long, 4> e_ijab{O, O, V, V};
  const std::array<int, 4> I{0, 1, 2, 3};
  // (the df method has no Gabcd or Giabc: its ladder and Giabc-type terms come from the factors)
  pp_sym_planes_ = false;
  if (!df_) {
    // (the packed operands' plane split indexes a (pairs x pitch) plane with 32 bits)
    const long nvs = V * (V + 1) / 2, ldvs = (nvs + 63) / 64 * 64;
    pp_sym_planes_ = nvs * ldvs < (1L << 31) && pp_symmetric_usable();
    if (pp_sym_planes_) {
      prepare_pp_symmetric();
    } else {
      // the full contraction's v^4 planes, only where the pair-symmetric form is not usable
      ensure_planes(Gabcd_P_, vvvv);
      split(Gabcd_P_, {V, V, V, V}, {{Gabcd_, {V, V, V, V}, I, 1.0}});
    }
    split(Giabc_P_, {O, V, V, V}, {{Giabc_, {O, V, V, V}, I, 1.0}});
  }
  split(Gijab_P_, {O, O, V, V}, {{Gijab_, e_ijab, I, 1.0}});
  // B^T operands of the J/K GEMMs: GA^T(k,c,l,d) = Gijab(k,l,c,d); GB^T = 2 Gijab(k,l,c,d) -
  // Gijab(k,l,d,c); GC^T(k,c,l,d) = Gijab(k,l,d,c)
  split(GAt_P_, {O, V, O, V}, {{Gijab_, e_ijab, {0, 2, 1, 3}, 1.0}});
  split(GBt_P_, {O, V, O, V}, {{Gijab_, e_ijab, {0, 2, 1, 3}, 2.0}, {Gijab_, e_ijab, {0, 3, 1, 2}, -1.0}});
  split(GCt_P_, {O, V, O, V}, {{Gijab_, e_ijab, {0, 3, 1, 2}, 1.0}});
  // r1 operands: GbarT(c,d,k,l) = gbar(k,l,c,d); gbar(k,l,a,b) and g_iabc_bar(a,k,c,d)
  // themselves; P8T(i,c,k,l) = P8(c,k,l,i)
  split(GbarT_P_, {V, V, O, O}, {{Gbar_, e_ijab, {2, 3, 0, 1}, 1.0}});
  split(Gbar_P_, e_ijab, {{Gbar_, e_ijab, I, 1.0}});
  if (!df_) split(Gbar_akcd_P_, {V, O, V, V}, {{Gbar_akcd_, {V, O, V, V}, I, 1.0}});
  split(P8T_P_, {O, V, O, O}, {{P8_, {V, O, O, O}, {3, 0, 1, 2}, 1.0}});
  // GiabcT(k,a,c,d) = Giabc(k,a,d,c), the A operand of the k-intermediate's T1 term
  if (!df_) split(GiabcT_P_, {O, V, V, V}, {{Giabc_, {O, V, V, V}, {0, 1, 3, 2}, 1.0}});
  // g_ki / g_ac operands (ccsd_r1_r2.h:219,223)
  split(P13_P_, {O, O, O, V}, {{P13_, {O, O, O, V}, I, 1.0}});
  if (!df_) split(P14_P_, {V, V, O, V}, {{P14_, {V, V, O, V}, I, 1.0}});
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  planes_ready_ = true;
}

double CCSDContext::scalar_from_device(const double* d) {
  double h = 0.0;
  MPQC_HIP_CHECK(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  return h;
}

// ---------------------------------------------------------------- set-up
// Integral layouts consumed by the GEMMs (constant for a solve).
void CCSDContext::prepare() {
  // standard: all nine blocks; df (ccsd.h:281-305): no Giabc / Gabcd under
  // reduced_abcd_memory, plus Xai, Xij, Xab
  const unsigned need = (direct_ && !df_) ? 0x0ffu : !df_ ? 0x1ffu : (reduced_abcd_ || direct_) ? 0xebfu : 0xfffu;
  if ((have_ints_ & need) != need)
    throw std::logic_error((direct_ && !df_) ? "method direct needs Fia, Fij, Fab, Gijab, Gijkl, Giajb, Giabc, Gijka"
                           : !df_ ? "all of Fia, Fij, Fab, Gijab, Gijkl, Giajb, Giabc, Gijka, Gabcd must be set"
                           : reduced_abcd_ ? "method df needs Fia, Fij, Fab, Gijab, Gijkl, Giajb, Gijka, Xai, Xij, Xab"
                                           : "method df without reduced_abcd_memory also needs Giabc and Gabcd");
  const bool have_iabc = !df_ || !reduced_abcd_;
  const long O = o_, V = v_;
  const std::array<long, 4> ovv_ext_ijab{O, O, V, V}, e_iajb{O, V, O, V}, e_iabc{O, V, V, V},
      e_ijka{O, O, O, V};
  const std::array<int, 4> I{0, 1, 2, 3};
  ++gen_;
  prepare_tape_.clear();
  std::vector<TapeOp>* saved_tape = tape_;
  tape_ = &prepare_tape_;
  // f_ai("a,i") = Fia("i,a")                                     ccsd.h:331-332
  perm_add(f_ai_, {1, 1, V, O}, 0.0, {{Fia_, {1, 1, O, V}, {0, 1, 3, 2}, 1.0}});
  // g_abij("a,b,i,j") = Gijab("i,j,a,b")                          ccsd.h:336-337
  perm_add(g_abij_, {V, V, O, O}, 0.0, {{Gijab_, ovv_ext_ijab, {2, 3, 0, 1}, 1.0}});
  // g_ijab_bar(k,l,c,d) = 2 G(k,l,c,d) - G(l,k,c,d)               ccsd_r1_r2.h:87
  perm_add(Gbar_, {O, O, V, V}, 0.0,
           {{Gijab_, ovv_ext_ijab, I, 2.0}, {Gijab_, ovv_ext_ijab, {1, 0, 2, 3}, -1.0}});
  perm_add(Gbar_cKLD_, {V, O, O, V}, 0.0, {{Gbar_, ovv_ext_ijab, {2, 0, 1, 3}, 1.0}});
  perm_add(Gbar_kcld_, {O, V, O, V}, 0.0, {{Gbar_, ovv_ext_ijab, {0, 2, 1, 3}, 1.0}});
  // P6(a,i,c,k) = 2 Gijab(k,i,c,a) - Giajb(k,a,i,c)                ccsd_r1_r2.h:107-108
  perm_add(P6_, {V, O, V, O}, 0.0,
           {{Gijab_, ovv_ext_ijab, {3, 1, 2, 0}, 2.0}, {Giajb_, e_iajb, {1, 2, 3, 0}, -1.0}});
  // g_iabc_bar(k,a,c,d) = 2 Giabc(k,a,c,d) - Giabc(k,a,d,c) as (a,k,c,d)   ccsd_r1_r2.h:112-113
  if (have_iabc && !df_)
    perm_add(Gbar_akcd_, {V, O, V, V}, 0.0,
             {{Giabc_, e_iabc, {1, 0, 2, 3}, 2.0}, {Giabc_, e_iabc, {1, 0, 3, 2}, -1.0}});
  // GiabcT(k,a,c,d) = Giabc(k,a,d,c): the k term's Giabc(k,a,d,c) t1(d,i) (:246) as one
  // contraction over d in the FP64 forms
  if (have_iabc && !df_) {
    if (!GiabcT_) GiabcT_ = alloc(O * V * V * V);
    perm_add(GiabcT_, e_iabc, 0.0, {{Giabc_, e_iabc, {0, 1, 3, 2}, 1.0}});
  }
  // P8(c,k,l,i) = 2 Gijka(l,k,i,c) - Gijka(k,l,i,c)                 ccsd_r1_r2.h:120-121
  perm_add(P8_, {V, O, O, O}, 0.0,
           {{Gijka_, e_ijka, {3, 1, 0, 2}, 2.0}, {Gijka_, e_ijka, {3, 0, 1, 2}, -1.0}});
  // S11(a,i,j,k) = Gijka(j,i,k,a)                                   ccsd_r1_r2.h:210-211
  perm_add(S11_, {V, O, O, O}, 0.0, {{Gijka_, e_ijka, {3, 1, 0, 2}, 1.0}});
  // P13(k,i,l,c) = g_ijka_bar(k,l,i,c) = 2 Gijka(k,l,i,c) - Gijka(l,k,i,c)   ccsd_r1_r2.h:219
  perm_add(P13_, {O, O, O, V}, 0.0,
           {{Gijka_, e_ijka, {0, 2, 1, 3}, 2.0}, {Gijka_, e_ijka, {1, 2, 0, 3}, -1.0}});
  // P14(a,c,k,d) = g_iabc_bar(k,a,d,c) = 2 Giabc(k,a,d,c) - Giabc(k,a,c,d)   ccsd_r1_r2.h:223
  if (have_iabc && !df_)
    perm_add(P14_, {V, V, O, V}, 0.0,
             {{Giabc_, e_iabc, {1, 3, 0, 2}, 2.0}, {Giabc_, e_iabc, {1, 2, 0, 3}, -1.0}});
  // J0(a,i,k,c) = Gijab(i,k,a,c);  K0(a,i,k,c) = Giajb(k,a,i,c)     ccsd_r1_r2.h:239,245
  perm_add(J0_, {V, O, O, V}, 0.0, {{Gijab_, ovv_ext_ijab, {2, 0, 1, 3}, 1.0}});
  perm_add(K0_, {V, O, O, V}, 0.0, {{Giajb_, e_iajb, {1, 2, 0, 3}, 1.0}});
  // X18(l,i,k,c) = Gijka(l,k,i,c);  X19(l,i,k,c) = Gijka(k,l,i,c)   ccsd_r1_r2.h:239,246
  perm_add(X18_, {O, O, O, V}, 0.0, {{Gijka_, e_ijka, {0, 2, 1, 3}, 1.0}});
  perm_add(X19_, {O, O, O, V}, 0.0, {{Gijka_, e_ijka, {1, 2, 0, 3}, 1.0}});
  // GA(l,d,k,c) = Gijab(k,l,c,d); GB = 2 Gijab(k,l,c,d) - Gijab(k,l,d,c); GC = Gijab(k,l,d,c)
  //                                                                 ccsd_r1_r2.h:241-243,248
  perm_add(GA_, {O, V, O, V}, 0.0, {{Gijab_, ovv_ext_ijab, {1, 3, 0, 2}, 1.0}});
  perm_add(GB_, {O, V, O, V}, 0.0,
           {{Gijab_, ovv_ext_ijab, {1, 3, 0, 2}, 2.0}, {Gijab_, ovv_ext_ijab, {1, 2, 0, 3}, -1.0}});
  perm_add(GC_, {O, V, O, V}, 0.0, {{Gijab_, ovv_ext_ijab, {1, 2, 0, 3}, 1.0}});
  if (df_) prepare_df();
  tape_ = saved_tape;
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  prepared_ = true;
  planes_ready_ = false;
  pp_packed_ = false;
  pp_checked_ = false;
  iabc_checked_ = false;
}

void CCSDContext::init_guess() {
  if (!prepared_) prepare();
  const long ov = (long)o_ * v_, oovv = ov * ov;
  // t1 = 0, t2 = 0, tau = 0; r1 = f_ai, r2 = g_abij                 ccsd.h:331-352
  MPQC_HIP_CHECK(hipMemsetAsync(t1_, 0, ov * sizeof(double), stream_));
  MPQC_HIP_CHECK(hipMemsetAsync(t2_, 0, oovv * sizeof(double), stream_));
  MPQC_HIP_CHECK(hipMemsetAsync(tau_, 0, oovv * sizeof(double), stream_));
  MPQC_HIP_CHECK(hipMemcpyAsync(r1_, f_ai_, ov * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipMemcpyAsync(r2_, g_abij_, oovv * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  tau_fresh_ = true;
  tau_amax_fresh_ = false;
}

// ---------------------------------------------------------------- residual
// The FP64 and FP32 residuals run the same launches on the same buffers every
// call once the first call has made its checks and allocations, so from the
// second call of a configuration on they are captured once and replayed (no
// host launch gaps).  The split forms choose plane scales on the host and the
// tape records ops on the host, so they stay eager, as do the df and direct
// methods.
bool CCSDContext::graph_eligible() const {
  return graph_mode_ && !split_mode() && !tape_ && !kernel_timing_ && !df_ && !direct_ &&
         res_failed_gen_ != gen_;
}

void CCSDContext::drop_residual_graph() {
  if (res_exec_) (void)hipGraphExecDestroy(res_exec_);
  if (res_graph_) (void)hipGraphDestroy(res_graph_);
  res_exec_ = nullptr;
  res_graph_ = nullptr;
  res_graph_gen_ = -1;
}

bool CCSDContext::capture_residual() {
  drop_residual_graph();
  MPQC_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  bool ok = true;
  try {
    residual_ops();
  } catch (...) {
    ok = false;
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(stream_, &g);
  if (!ok || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    res_failed_gen_ = gen_;  // eager from now on in this configuration
    return false;
  }
  res_graph_ = g;
  MPQC_HIP_CHECK(hipGraphInstantiate(&res_exec_, res_graph_, nullptr, nullptr, 0));
  res_graph_gen_ = gen_;
  return true;
}

void CCSDContext::residual() {
  if (!prepared_) prepare();
  // tau("a,b,i,j") = t2("a,b,i,j") + t1("a,i") t1("b,j") (ccsd.h:389): the update and
  // set_amplitudes leave it current; rebuilt here otherwise (and always on the tape)
  if (!tau_fresh_ || tape_) {
    op_tau();
    tau_amax_fresh_ = false;
  }
  if (!tape_ || tape_execute_) tau_fresh_ = true;  // (a tape recorded without executing runs nothing)
  if (graph_eligible()) {
    if (res_exec_ && res_graph_gen_ == gen_) {
      MPQC_HIP_CHECK(hipGraphLaunch(res_exec_, stream_));
      return;
    }
    if (res_seen_gen_ == gen_) {
      if (capture_residual()) {
        MPQC_HIP_CHECK(hipGraphLaunch(res_exec_, stream_));
        return;
      }
    } else {
      res_seen_gen_ = gen_;
    }
  }
  residual_ops();
}

void CCSDContext::residual_ops() {
  const int o = o_, v = v_;
  const long O = o, V = v;
  const long ov = O * V, oo = O * O, vv = V * V;
  const std::array<long, 4> e_abij{V, V, O, O};
  const std::array<int, 4> I{0, 1, 2, 3};

  // amplitude layouts used by several GEMMs (tau is current: residual())
  perm_add(t1T_, {1, 1, O, V}, 0.0, {{t1_, {1, 1, V, O}, {0, 1, 3, 2}, 1.0}});  // t1T(k,d) = t1(d,k)
  const bool sm = split_mode();
  if (direct_ && sm) throw std::logic_error("the direct residual runs in the f64 and f32 forms");
  const long Q = naux_;
  if (df_) df_r1_intermediates();
  if (direct_) direct_u();  // ccsd.h:423-426: U before r1
  if (sm) {
    // 16-bit planes of the amplitude operands (the integral planes are split once per solve)
    if (!amax_.dev) amax_.dev = alloc(AbsMaxCache::kSlots);
    amax_.clear();
    if (!planes_ready_) prepare_planes();
    amax_.ptr[0] = t1_;
    amax_.ptr[1] = t2_;
    amax_.ptr[2] = tau_;
    amax_.ptr[3] = J_;  // final when the ring operands are split (max |K| then serves two splits)
    amax_.ptr[4] = K_;
    if (tau_amax_fresh_) amax_.valid[1] = amax_.valid[2] = true;  // max |t2|, |tau| from the update
    split(tau_P_, e_abij, {{tau_, e_abij, I, 1.0}});
    // tau^T(i,j,c,d) = tau(c,d,i,j): the B^T operand of every tau contraction
    split(tauT_P_, {O, O, V, V}, {{tau_, e_abij, {2, 3, 0, 1}, 1.0}});
    split(t1T_P_, {1, 1, O, V}, {{t1_, {1, 1, V, O}, {0, 1, 3, 2}, 1.0}});
    // Y2^T(b,j,k,c) = t2(c,b,j,k)
    split(Y2t_P_, {V, O, O, V}, {{t2_, e_abij, {1, 2, 3, 0}, 1.0}});
  }

  // ===================== r1: ccsd_r1_r2.h:70-124 =====================
  // :83-84  r1(a,i) = Fia(i,a) - 2 (Fia(k,c) t1(c,i)) t1(a,k)
  gemm(false, false, o, o, v, 1.0, Fia_, v, t1_, o, 0.0, Xoo_, o, 0, 0, 0, 1, "r1:fia_t1");
  copy_d2d(r1_, f_ai_, ov);
  gemm(false, false, v, o, o, -2.0, t1_, o, Xoo_, o, 1.0, r1_, o, 0, 0, 0, 1, "r1:t1_fia_t1");
  // :89-91  h_ac = F_vv = Fab(a,c) - gbar(k,l,c,d) tau(a,d,k,l); FAB = h_ac; r1 += h_ac t1
  //         (ccsd_intermediates.h:71-77)
  copy_d2d(FAB_, Fab_, vv);
  if (df_) {  // df :155: h_ac = Fab(a,c) - X_ai_tau(K,a,l) Xai(K,c,l):  XtP(a,(K,l)) . XaiP(c,(K,l))^T
    gemm(false, true, v, v, (int)(Q * O), -1.0, XtP_, Q * O, XaiP_, Q * O, 1.0, FAB_, v, 0, 0, 0, 1, "r1:df_F_vv");
  } else if (sm) {  // tau(a,(d,k,l)) . gbar^T(c,(d,k,l))
    gemmP(v, v, o * o * v, -1.0, tau_P_, oo * V, GbarT_P_, oo * V, 1.0, FAB_, v, "r1:F_vv");
  } else {
    perm_add(tau_akld_, {V, O, O, V}, 0.0, {{tau_, e_abij, {0, 2, 3, 1}, 1.0}});
    gemm(false, true, v, v, o * o * v, -1.0, tau_akld_, oo * V, Gbar_cKLD_, oo * V, 1.0, FAB_, v, 0, 0, 0, 1, "r1:F_vv");
  }
  gemm(false, false, v, o, v, 1.0, FAB_, v, t1_, o, 1.0, r1_, o, 0, 0, 0, 1, "r1:FAB_t1");
  // :95-97  h_ki = F_oo = Fij(k,i) + gbar(k,l,a,b) tau(a,b,i,l); FIJ = h_ki; r1 -= t1 h_ki
  //         (ccsd_intermediates.h:57-63)
  copy_d2d(FIJ_, Fij_, oo);
  if (df_) {  // df :161: h_ki = Fij(k,i) + X_ai_tau(K,c,i) Xai(K,c,k):  Xai((K,c),k)^T . Xt((K,c),i)
    gemm(true, false, o, o, (int)(Q * V), 1.0, Xai_, o, Xt_, o, 1.0, FIJ_, o, 0, 0, 0, 1, "r1:df_F_oo");
  } else if (sm) {  // gbar(k,(l,a,b)) . tau^T(i,(l,a,b))
    gemmP(o, o, o * v * v, 1.0, Gbar_P_, O * vv, tauT_P_, O * vv, 1.0, FIJ_, o, "r1:F_oo");
  } else {
    perm_add(tau_ilab_, {O, O, V, V}, 0.0, {{tau_, e_abij, {2, 3, 0, 1}, 1.0}});
    gemm(false, true, o, o, o * v * v, 1.0, Gbar_, O * vv, tau_ilab_, O * vv, 1.0, FIJ_, o, 0, 0, 0, 1, "r1:F_oo");
  }
  gemm(false, false, v, o, o, -1.0, t1_, o, FIJ_, o, 1.0, r1_, o, 0, 0, 0, 1, "r1:t1_FIJ");
  // :101-104 h_kc = F_ov = Fia(k,c) + gbar(k,l,c,d) t1(d,l)   (ccsd_intermediates.h:89-95)
  //          r1 += h_kc (2 t2(c,a,k,i) - t2(c,a,i,k) + t1(c,i) t1(a,k))
  copy_d2d(hkc_, Fia_, ov);
  gemm(false, false, o * v, 1, o * v, 1.0, Gbar_kcld_, ov, t1T_, 1, 1.0, hkc_, 1, 0, 0, 0, 1, "r1:F_ov");
  if (!tape_ && t2bar_h_supported(o, v)) {  // one pass over t2 (the tape keeps the contraction)
    timed("r1:hkc_t2bar", 2.0 * ov * ov, [&] { t2bar_h(r1_, t2_, hkc_, o, v, partials_, stream_); });
  } else {
    perm_add(hT_, {1, 1, V, O}, 0.0, {{hkc_, {1, 1, O, V}, {0, 1, 3, 2}, 1.0}});  // hT(c,k) = h(k,c)
    perm_add(Y5_, {V, O, V, O}, 0.0,
             {{t2_, e_abij, {1, 3, 0, 2}, 2.0}, {t2_, e_abij, {1, 2, 0, 3}, -1.0}});
    gemm(false, false, v * o, 1, v * o, 1.0, Y5_, ov, hT_, 1, 1.0, r1_, 1, 0, 0, 0, 1, "r1:hkc_t2bar");
  }
  gemm(false, false, o, o, v, 1.0, hkc_, v, t1_, o, 0.0, Zoo_, o, 0, 0, 0, 1, "r1:hkc_t1");
  gemm(false, false, v, o, o, 1.0, t1_, o, Zoo_, o, 1.0, r1_, o, 0, 0, 0, 1, "r1:t1_hkc_t1");
  // :107-108 r1 += (2 Gijab(k,i,c,a) - Giajb(k,a,i,c)) t1(c,k)
  gemm(false, false, v * o, 1, v * o, 1.0, P6_, ov, t1_, 1, 1.0, r1_, 1, 0, 0, 0, 1, "r1:g_t1");
  // :110-114 r1 += g_iabc_bar(k,a,c,d) tau(c,d,k,i)
  if (df_) {
    if (direct_)  // direct_df: the standard r1 with U (ccsd.h:436, ccsd_r1_r2.h:116-117)
      direct_r1();
    else  // df :177 r1 += X_ai_tau(K,c,i) Xab(K,a,c):  XabP(a,(K,c)) . Xt((K,c),i)
      gemm(false, false, v, o, (int)(Q * V), 1.0, XabP_, Q * V, Xt_, o, 1.0, r1_, o, 0, 0, 0, 1, "r1:df_Xt_Xab");
    // df :179 r1 -= X_ai_tau(K,a,l) Xij(K,l,i):  XtP(a,(K,l)) . Xij((K,l),i)
    gemm(false, false, v, o, (int)(Q * O), -1.0, XtP_, Q * O, Xij_, o, 1.0, r1_, o, 0, 0, 0, 1, "r1:df_Xt_Xij");
  } else if (direct_) {  // :116-117
    direct_r1();
  } else if (sm) {  // g_iabc_bar(a,(k,c,d)) . tau_ikcd(i,(k,c,d)),  tau_ikcd(i,k,c,d) = tau(c,d,k,i)
    split(tau_ikcd_P_, {O, O, V, V}, {{tau_, e_abij, {3, 2, 0, 1}, 1.0}});
    gemmP(v, o, o * v * v, 1.0, Gbar_akcd_P_, O * vv, tau_ikcd_P_, O * vv, 1.0, r1_, o, "r1:giabc_tau");
  } else {
    perm_add(tau_kcdi_, {O, V, V, O}, 0.0, {{tau_, e_abij, {2, 0, 1, 3}, 1.0}});
    gemm(false, false, v, o, o * v * v, 1.0, Gbar_akcd_, O * vv, tau_kcdi_, o, 1.0, r1_, o, 0, 0, 0, 1, "r1:giabc_tau");
  }
  // :120-121 r1 -= (2 Gijka(l,k,i,c) - Gijka(k,l,i,c)) tau(c,a,k,l)
  if (df_) {
    // (replaced by the X_ai_tau Xij term above)
  } else if (sm) {  // tau_ackl(a,(c,k,l)) . P8T(i,(c,k,l)),  tau_ackl(a,c,k,l) = tau(c,a,k,l)
    split(tau_ackl_P_, e_abij, {{tau_, e_abij, {1, 0, 2, 3}, 1.0}});
    gemmP(v, o, o * o * v, -1.0, tau_ackl_P_, oo * V, P8T_P_, oo * V, 1.0, r1_, o, "r1:gijka_tau");
  } else {
    perm_add(tau_ackl_, {V, V, O, O}, 0.0, {{tau_, e_abij, {1, 0, 2, 3}, 1.0}});
    gemm(false, false, v, o, o * o * v, -1.0, tau_ackl_, oo * V, P8_, o, 1.0, r1_, o, 0, 0, 0, 1, "r1:gijka_tau");
  }

  // ===================== r2: ccsd_r1_r2.h:198-293 =====================
  // Giabc(i,c,a,b) = Giabc(i,b,a,c): Rd's Giabc t1 term is a permutation of Zj (not on the
  // adjoint's tape, which keeps the reference's contractions)
  const bool iabc_sym = !df_ && !tape_ && giabc_symmetric_usable();
  // Zj(k,a,c,i) = Giabc(k,a,c,d) t1(d,i), the j term's Giabc t1 part (:240); in FP64 when it
  // also stands for Rd's term (a T1 dressing term in 16-bit planes slows the split forms'
  // convergence: (H2O)8 bf16x3 36 iterations to 1e-8 instead of 18)
  auto j_giabc_t1 = [&] {
    if (sm && !iabc_sym)
      gemmP(o * v * v, o, v, 1.0, Giabc_P_, v, t1T_P_, v, 0.0, Zj_, o, "r2:j_giabc_t1");
    else
      gemm(false, false, o * v * v, o, v, 1.0, Giabc_, v, t1_, o, 0.0, Zj_, o, 0, 0, 0, 1, "r2:j_giabc_t1");
  };
  if (df_) {
    // df :317 X_ab_t1(K,a,i) = Xab(K,a,b) t1(b,i)
    gemm(false, false, (int)(Q * V), o, v, 1.0, Xab_, v, t1_, o, 0.0, Xabt1_, o, 0, 0, 0, 1, "r2:df_Xab_t1");
    perm_add(Xabt1P_, {1, V, Q, O}, 0.0, {{Xabt1_, {1, Q, V, O}, {0, 2, 1, 3}, 1.0}});
    // df :321-322 Rd(a,b,i,j) = X_ab_t1(K,b,j) (Xai(K,a,i) - Xij(K,k,i) t1(a,k))
    copy_d2d(Zdf_, Xai_, Q * ov);
    gemm(false, false, v, o, o, -1.0, t1_, o, Xij_, o, 1.0, Zdf_, o, 0, oo, ov, (int)Q, "r2:df_Xij_t1");
    gemm(true, false, v * o, v * o, (int)Q, 1.0, Zdf_, ov, Xabt1_, ov, 0.0, Rdf_, ov, 0, 0, 0, 1, "r2:df_Z_Xabt1");
    perm_add(Rd_, {V, V, O, O}, 0.0, {{Rdf_, {V, O, V, O}, {0, 2, 1, 3}, 1.0}});
  } else {
  // :207-208 Rd = (Giabc(i,c,a,b) - Giajb(k,b,i,c) t1(a,k)) t1(c,j)
  if (iabc_sym) {
    //   first term: Zj(i,b,a,j), added in the half-sum pass below
    j_giabc_t1();
  } else {
    //   first term, batched over i: Rd[(a,b), i, j] = sum_c Giabc[i][c,(a,b)] t1[c,j]
    gemm(true, false, v * v, o, v, 1.0, Giabc_, vv, t1_, o, 0.0, Rd_, oo, vv * V, 0, O, o, "r2:giabc_t1");
  }
  //   second term as t1(a,k) * (Giajb(k,b,i,c) t1(c,j))
  gemm(false, false, o * o * v, o, v, 1.0, Giajb_, v, t1_, o, 0.0, Q_, o, 0, 0, 0, 1, "r2:giajb_t1");
  }
  // :210-211 Rd -= (Gijka(j,i,k,a) + Gijab(i,k,a,c) t1(c,j)) t1(b,k)
  gemm(false, false, o * o * v, o, v, 1.0, Gijab_, v, t1_, o, 0.0, M12_, o, 0, 0, 0, 1, "r2:gijab_t1");
  if (df_) {
    perm_add(S_, {V, O, O, O}, 0.0,
             {{S11_, {V, O, O, O}, I, 1.0}, {M12_, {O, O, V, O}, {2, 0, 3, 1}, 1.0}});
    gemm(false, true, v, o * o, o, -1.0, t1_, o, S_, o, 1.0, Rd_, oo, 0, O * oo, V * oo, v, "r2:t1_S");
  } else {
    // Rd enters r2 only as Rd(a,b,i,j) + Rd(b,a,j,i) (:260), so the term
    // -S(a,i,j,k) t1(b,k), S(a,i,j,k) = Gijka(j,i,k,a) + M12(i,k,a,j), may go in as its
    // mirror -t1(a,k) S(b,j,i,k): both t1(a,k) terms in one pass over Rd, with
    // Q(k,b,i,j) += S(b,j,i,k); then Rd -= t1(a,k) Q(k,(b,i,j))
    perm_add(Q_, {O, V, O, O}, 1.0,
             {{S11_, {V, O, O, O}, {3, 0, 2, 1}, 1.0}, {M12_, {O, O, V, O}, {1, 2, 3, 0}, 1.0}});
    gemm(false, false, v, v * o * o, o, -1.0, t1_, o, Q_, V * oo, iabc_sym ? 0.0 : 1.0, Rd_, V * oo, 0, 0, 0, 1,
         "r2:t1_giajb_S");
  }
  // :218-220 g_ki = F_OO(FIJ, Fia, g_ijka_bar, t1) = FIJ + Fia t1 + gbar(k,l,i,c) t1(c,l)
  //          (ccsd_intermediates.h:105-112)
  copy_d2d(gki_, FIJ_, oo);
  gemm(false, false, o, o, v, 1.0, Fia_, v, t1_, o, 1.0, gki_, o, 0, 0, 0, 1, "r2:gki_fia_t1");
  if (sm)
    gemmP(o * o, 1, o * v, 1.0, P13_P_, ov, t1T_P_, ov, 1.0, gki_, 1, "r2:gki_g_t1");
  else
    gemm(false, false, o * o, 1, o * v, 1.0, P13_, ov, t1T_, 1, 1.0, gki_, 1, 0, 0, 0, 1, "r2:gki_g_t1");
  // :222-224 g_ac = F_VV(FAB, Fia, g_iabc_bar, t1) = FAB + f_ov(k,c) t1(a,k) + gbar(k,a,d,c) t1(d,k)
  //          (ccsd_intermediates.h:141-148; fov_sign = +1 is the reference as written)
  copy_d2d(gac_, FAB_, vv);
  // (the df form, :333, has - f(k,c) t1(a,k))
  gemm(false, false, v, v, o, df_ ? -1.0 : fov_sign_, t1_, o, Fia_, v, 1.0, gac_, v, 0, 0, 0, 1, "r2:gac_t1_fia");
  if (df_) {
    // df :334 + 2 (Xai(K,d,k) t1(d,k)) Xab(K,a,c)
    gemm(false, false, (int)Q, 1, v * o, 1.0, Xai_, ov, t1_, 1, 0.0, sK_, 1, 0, 0, 0, 1, "r2:df_Xai_t1");
    gemm(true, false, v * v, 1, (int)Q, 2.0, Xab_, vv, sK_, 1, 1.0, gac_, 1, 0, 0, 0, 1, "r2:df_gac_sXab");
    // df :335 - X_ab_t1(K,a,k) Xai(K,c,k):  Xabt1P(a,(K,k)) . XaiP(c,(K,k))^T
    gemm(false, true, v, v, (int)(Q * O), -1.0, Xabt1P_, Q * O, XaiP_, Q * O, 1.0, gac_, v, 0, 0, 0, 1, "r2:df_gac_Xabt1_Xai");
  } else if (sm)
    gemmP(v * v, 1, o * v, 1.0, P14_P_, ov, t1T_P_, ov, 1.0, gac_, 1, "r2:gac_g_t1");
  else
    gemm(false, false, v * v, 1, o * v, 1.0, P14_, ov, t1T_, 1, 1.0, gac_, 1, 0, 0, 0, 1, "r2:gac_g_t1");
  // :226 Rd += g_ac(a,c) t2(c,b,i,j) - g_ki(k,i) t2(a,b,k,j)
  if (sm) {  // g_ac(a,c) . Y2^T((b,i,j),c)
    split(gac_P_, {1, 1, V, V}, {{gac_, {1, 1, V, V}, I, 1.0}});
    gemmP(v, v * o * o, v, 1.0, gac_P_, v, Y2t_P_, v, 1.0, Rd_, V * oo, "r2:g_ac_t2");
  } else {
    gemm(false, false, v, v * o * o, v, 1.0, gac_, v, t2_, V * oo, 1.0, Rd_, V * oo, 0, 0, 0, 1, "r2:g_ac_t2");
  }
  //   - g_ki(k,i) t2(a,b,k,j): per (a,b), Rd(a,b,:,:) -= g^T t2(a,b,:,:), a batch of v^2
  //   o x o x o products reading t2 and Rd in place (no transposed copies)
  gemm(true, false, o, o, o, -1.0, gki_, o, t2_, o, 1.0, Rd_, o, 0, oo, oo, v * v, "r2:t2_gki");
  // :236 T(d,b,i,l) = 0.5 t2(d,b,i,l) + t1(d,i) t1(b,l), as TP(a,i,l,d) = T(d,a,i,l)
  if (!sm || df_) op_TP();  // (the df j term reads TP in FP64)
  if (!sm) perm_add(t2P_, {V, O, O, V}, 0.0, {{t2_, e_abij, {0, 2, 3, 1}, 1.0}});  // t2P(a,i,l,d) = t2(a,d,i,l)
  // :238-243 j(a,k,i,c) stored as J(a,i,k,c):
  //   Gijab(i,k,a,c) - Gijka(l,k,i,c) t1(a,l) + Giabc(k,a,c,d) t1(d,i)
  //   - Gijab(k,l,c,d) T(d,a,i,l) + 0.5 (2 Gijab(k,l,c,d) - Gijab(k,l,d,c)) t2(a,d,i,l)
  // :245-248 k(k,a,i,c) stored as K(a,i,k,c):
  //   Giajb(k,a,i,c) - Gijka(k,l,i,c) t1(a,l) + Giabc(k,a,d,c) t1(d,i) - Gijab(k,l,d,c) T(d,a,i,l)
  // (the integral terms J0, K0 are added with the Giabc t1 terms below; df: here)
  if (df_) {
    copy_d2d(J_, J0_, ov * ov);
    copy_d2d(K_, K0_, ov * ov);
  }
  const double jk_beta = df_ ? 1.0 : 0.0;
  gemm(false, false, v, o * o * v, o, -1.0, t1_, o, X18_, oo * V, jk_beta, J_, oo * V, 0, 0, 0, 1, "r2:j_gijka_t1");
  gemm(false, false, v, o * o * v, o, -1.0, t1_, o, X19_, oo * V, jk_beta, K_, oo * V, 0, 0, 0, 1, "r2:k_gijka_t1");
  if (df_) {
    // df :353-354 j += X_ab_t1(K,a,i) Xai(K,c,k) - Xai(x,d,l) T(d,a,i,l) Xai(x,c,k)
    //   = U(K,(a,i)) Xia(K,(k,c)),  U = X_ab_t1 - Xia(x,(l,d)) TP((a,i),(l,d))^T
    copy_d2d(Udf_, Xabt1_, Q * ov);
    gemm(false, true, (int)Q, v * o, o * v, -1.0, Xia_, ov, TP_, ov, 1.0, Udf_, ov, 0, 0, 0, 1, "r2:df_j_XT");
    gemm(true, false, v * o, o * v, (int)Q, 1.0, Udf_, ov, Xia_, ov, 1.0, J_, ov, 0, 0, 0, 1, "r2:df_j_U_Xia");
    // df :360 k += (Xai(K,d,k) t1(d,i)) Xab(K,a,c):  V1(K,k,i), then Zk(a,c,k,i) = Xab^T . V1
    gemm(true, false, o, o, v, 1.0, Xai_, o, t1_, o, 0.0, V1_, o, ov, 0, oo, (int)Q, "r2:df_k_Xai_t1");
    gemm(true, false, v * v, o * o, (int)Q, 1.0, Xab_, vv, V1_, oo, 0.0, Zk_, oo, 0, 0, 0, 1, "r2:df_k_V1_Xab");
    perm_add(K_, {V, O, O, V}, 1.0, {{Zk_, {V, V, O, O}, {0, 3, 2, 1}, 1.0}});
  } else {
  if (!iabc_sym) j_giabc_t1();  // (computed with Rd's T1 dressing otherwise)
  if (sm)
    gemmP(o * v * v, o, v, 1.0, GiabcT_P_, v, t1T_P_, v, 0.0, Zk_, o, "r2:k_giabc_t1");
  else
    gemm(false, false, o * v * v, o, v, 1.0, GiabcT_, v, t1_, o, 0.0, Zk_, o, 0, 0, 0, 1, "r2:k_giabc_t1");  // Zk(k,a,c,i)
  perm_add(J_, {V, O, O, V}, 1.0, {{Zj_, {O, V, V, O}, {1, 3, 0, 2}, 1.0}, {J0_, {V, O, O, V}, I, 1.0}});
  perm_add(K_, {V, O, O, V}, 1.0, {{Zk_, {O, V, V, O}, {1, 3, 0, 2}, 1.0}, {K0_, {V, O, O, V}, I, 1.0}});
  }
  if (sm) {
    const std::array<long, 4> e_aild{V, O, O, V};
    // planes of TP and t2P straight from t2 (and t1)
    const TProd pr = tp_product();
    split(TP_P_, e_aild, {{t2_, e_abij, {1, 2, 3, 0}, 0.5}}, &pr);
    split(t2P_P_, e_aild, {{t2_, e_abij, {0, 2, 3, 1}, 1.0}});
    if (!df_)  // (df: the X T X form above)
      gemmP(o * v, o * v, o * v, -1.0, TP_P_, ov, GAt_P_, ov, 1.0, J_, ov, "r2:j_gT");
    gemmP(o * v, o * v, o * v, 0.5, t2P_P_, ov, GBt_P_, ov, 1.0, J_, ov, "r2:j_gbar_t2", 3);  // (final J)
    gemmP(o * v, o * v, o * v, -1.0, TP_P_, ov, GCt_P_, ov, 1.0, K_, ov, "r2:k_gT", 4);        // (final K)
  } else {
    if (!df_)  // (df: the X T X form above)
      gemm(false, false, o * v, o * v, o * v, -1.0, TP_, ov, GA_, ov, 1.0, J_, ov, 0, 0, 0, 1, "r2:j_gT");
    gemm(false, false, o * v, o * v, o * v, 0.5, t2P_, ov, GB_, ov, 1.0, J_, ov, 0, 0, 0, 1, "r2:j_gbar_t2");
    gemm(false, false, o * v, o * v, o * v, -1.0, TP_, ov, GC_, ov, 1.0, K_, ov, 0, 0, 0, 1, "r2:k_gT");
  }
  // :251-252 r2 += 0.5 (2 j(a,k,i,c) - k(k,a,i,c)) (2 t2(c,b,k,j) - t2(b,c,k,j))  -> RR(a,i,b,j)
  if (sm) {
    const std::array<long, 4> e_aikc{V, O, O, V};
    // NT operands: Xjk(a,i,k,c) = 2J - K; Y1^T(b,j,k,c) = 2 t2(c,b,k,j) - t2(b,c,k,j);
    // Y2^T(b,j,k,c) = t2(c,b,j,k)
    split(Xjk_P_, e_aikc, {{J_, e_aikc, I, 2.0}, {K_, e_aikc, I, -1.0}});
    split(Y1t_P_, {V, O, O, V}, {{t2_, e_abij, {1, 3, 2, 0}, 2.0}, {t2_, e_abij, {0, 3, 2, 1}, -1.0}});
    gemmP(o * v, o * v, o * v, 0.5, Xjk_P_, ov, Y1t_P_, ov, 0.0, RR_, ov, "r2:ring");
    split(K_P_, e_aikc, {{K_, e_aikc, I, 1.0}});
    gemmP(o * v, o * v, o * v, 1.0, K_P_, ov, Y2t_P_, ov, 0.0, tmpR_, ov, "r2:crossed_ring");
  } else {
    perm_add(Xjk_, {V, O, O, V}, 0.0, {{J_, {V, O, O, V}, I, 2.0}, {K_, {V, O, O, V}, I, -1.0}});
    perm_add(Y1_, {O, V, V, O}, 0.0,
             {{t2_, e_abij, {2, 0, 1, 3}, 2.0}, {t2_, e_abij, {2, 1, 0, 3}, -1.0}});
    gemm(false, false, o * v, o * v, o * v, 0.5, Xjk_, ov, Y1_, ov, 0.0, RR_, ov, 0, 0, 0, 1, "r2:ring");
    // :254-256 tmp(a,b,i,j) = k(k,a,i,c) t2(c,b,j,k) -> tmpR(a,i,b,j); r2 -= 0.5 tmp + tmp(b,a,i,j)
    perm_add(Y2_, {O, V, V, O}, 0.0, {{t2_, e_abij, {3, 0, 1, 2}, 1.0}});
    gemm(false, false, o * v, o * v, o * v, 1.0, K_, ov, Y2_, ov, 0.0, tmpR_, ov, 0, 0, 0, 1, "r2:crossed_ring");
  }
  // :277-280 tmp(k,a,i,j) = Giabc(k,a,c,d) tau(c,d,i,j); b -= tmp(k,a,j,i) t1(b,k) + tmp(k,b,i,j) t1(a,k).
  // The two terms are mirror images under (a,i) <-> (b,j), and Rd enters r2 as
  // Rd(a,b,i,j) + Rd(b,a,j,i): Rd takes the second (the direct and reduced-memory df
  // ladders form these terms themselves)
  if (!direct_ && !(df_ && reduced_abcd_)) {
    if (sm)
      gemmP(o * v, o * o, v * v, 1.0, Giabc_P_, vv, tauT_P_, vv, 0.0, tmp24_, oo, "r2:pp_t1_giabc_tau");
    else
      gemm(false, false, o * v, o * o, v * v, 1.0, Giabc_, vv, tau_, oo, 0.0, tmp24_, oo, 0, 0, 0, 1,
           "r2:pp_t1_giabc_tau");
    gemm(false, false, v, v * o * o, o, -1.0, t1_, o, tmp24_, V * oo, 1.0, Rd_, V * oo, 0, 0, 0, 1, "r2:pp_t1_a");
  }
  // :260 r2(a,b,i,j) = r2(a,b,i,j) + r2(b,a,j,i); :262 r2 += Gijab(i,j,a,b).  The ring
  // terms come in pairs under (a,i) <-> (b,j), so in two passes: the half-sum
  // H = Rd + RR(a,i,b,j) - 0.5 tmpR(a,i,b,j) - tmpR(b,i,a,j) (in Rd), then
  // r2 = H + H(b,a,j,i) + Gijab: 9 tensor reads and writes in place of 12
  {
    const std::array<long, 4> e_aibj{V, O, V, O};
    if (iabc_sym)  // with Rd's first T1 dressing term, Giabc(i,c,a,b) t1(c,j) = Zj(i,b,a,j)
      perm_add(Rd_, {V, V, O, O}, 1.0,
               {{RR_, e_aibj, {0, 2, 1, 3}, 1.0},
                {tmpR_, e_aibj, {0, 2, 1, 3}, -0.5},
                {tmpR_, e_aibj, {2, 0, 1, 3}, -1.0},
                {Zj_, {O, V, V, O}, {2, 1, 0, 3}, 1.0}});
    else
      perm_add(Rd_, {V, V, O, O}, 1.0,
               {{RR_, e_aibj, {0, 2, 1, 3}, 1.0},
                {tmpR_, e_aibj, {0, 2, 1, 3}, -0.5},
                {tmpR_, e_aibj, {2, 0, 1, 3}, -1.0}});
    perm_add(r2_, {V, V, O, O}, 0.0,
             {{Rd_, e_abij, I, 1.0},
              {Rd_, e_abij, {1, 0, 3, 2}, 1.0},
              {g_abij_, e_abij, I, 1.0}});  // Gijab(i,j,a,b) in the output's order (prepare())
  }
  // :266-269 a_klij = W_oooo (ccsd_intermediates.h:182-193); r2 += a_klij(k,l,i,j) tau(a,b,k,l)
  gemm(false, false, o * o * o, o, v, 1.0, Gijka_, v, t1_, o, 0.0, Y23_, o, 0, 0, 0, 1, "r2:W_gijka_t1");
  perm_add(W_, {O, O, O, O}, 0.0,
           {{Gijkl_, {O, O, O, O}, I, 1.0}, {Y23_, {O, O, O, O}, I, 1.0},
            {Y23_, {O, O, O, O}, {1, 0, 3, 2}, 1.0}});
  // the pair-symmetric forms of W_oooo's G tau term and of the pp ladder share S and A
  const bool psym = sm ? (pp_sym_planes_ && !df_)
                       : (!tape_ && !direct_ && !(df_ && reduced_abcd_) && pp_symmetric_usable());
  if (psym) pack_tau_pairs();
  if (sm) {
    if (psym && w_sym_ok_)
      w_oooo_symmetric();
    else
      gemmP(o * o, o * o, v * v, 1.0, Gijab_P_, vv, tauT_P_, vv, 1.0, W_, oo, "r2:W_oooo");
    // W^T(i,j,k,l) = W(k,l,i,j)
    split(WT_P_, {O, O, O, O}, {{W_, {O, O, O, O}, {2, 3, 0, 1}, 1.0}});
    gemmP(v * v, o * o, o * o, 1.0, tau_P_, oo, WT_P_, oo, 1.0, r2_, oo, "r2:hh_ladder");
    if (df_ && reduced_abcd_) {
      df_pp_ladder();  // df :497-508, the (ab|cd) blocks on 16-bit planes
      return;
    }
    if (df_) throw std::logic_error("the split forms of the df method need reduced_abcd_memory");
    // :275 b(a,b,i,j) = tau(c,d,i,j) Gabcd(a,b,c,d)  -- the particle-particle ladder
    if (psym)
      pp_ladder_symmetric(r2_);
    else
      gemmP(v * v, o * o, v * v, 1.0, Gabcd_P_, vv, tauT_P_, vv, 1.0, r2_, oo, "r2:pp_ladder");
  } else {
    if (psym && w_sym_ok_)
      w_oooo_symmetric();
    else
      gemm(false, false, o * o, o * o, v * v, 1.0, Gijab_, vv, tau_, oo, 1.0, W_, oo, 0, 0, 0, 1, "r2:W_oooo");
    gemm(false, false, v * v, o * o, o * o, 1.0, tau_, oo, W_, oo, 1.0, r2_, oo, 0, 0, 0, 1, "r2:hh_ladder");
    if (direct_) {
      direct_r2();  // :282-286 (direct_df: :509-514)
      return;
    }
    if (df_ && reduced_abcd_) {
      df_pp_ladder();  // df :497-508
      return;
    }
    // :275 b(a,b,i,j) = tau(c,d,i,j) Gabcd(a,b,c,d)  -- the particle-particle ladder (the
    // recorded residual of the adjoint keeps the full contraction)
    if (psym)
      pp_ladder_symmetric(r2_);
    else
      gemm(false, false, v * v, o * o, v * v, 1.0, Gabcd_, vv, tau_, oo, 1.0, r2_, oo, 0, 0, 0, 1, "r2:pp_ladder");
  }
  // (the ladder's T1 terms, :277-280, went into Rd above)
}

// ---------------------------------------------------------------- energy / error / update
void CCSDContext::energy_error(double* E, double* err) {
  if (!prepared_) prepare();
  const long n1 = (long)o_ * v_, n2 = n1 * n1;
  // error = sqrt(|r1|^2 + |r2|^2) / (size(r1) + size(r2))   qc/cc/solvers.h:60-63
  // E = 2 <f_ai + r1, t1> + <g_abij + r2, 2 tau(a,b,i,j) - tau(b,a,i,j)>   ccsd.h:377-379
  // (with the two norms, in one pass over r2)
  timed("en:energy", 0.0, [&] {
    ccsd_energy_norms(f_ai_, r1_, t1_, g_abij_, r2_, tau_, o_, v_, scalars_, partials_, stream_);
  });
  double h[3];
  MPQC_HIP_CHECK(hipMemcpyAsync(h, scalars_, 3 * sizeof(double), hipMemcpyDeviceToHost, stream_));
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
  const double e = std::sqrt(h[0] + h[1]) / (double)(n1 + n2);
  if (!std::isfinite(e) || !std::isfinite(h[2]))
    throw std::range_error("non-finite residual or energy");
  if (err) *err = e;
  if (E) *E = h[2];
}

void CCSDContext::check_finite(const double* x, long n, const char* what) {
  count_nonfinite(x, n, scalars_ + 8, partials_, stream_);
  if (scalar_from_device(scalars_ + 8) != 0.0) throw std::range_error(std::string("non-finite ") + what);
}

// DIISSolver::update (qc/cc/solvers.h:112-121): Jacobi update_only
// (lcao/cc/solvers.h:1153-1158), then TA::DIIS::extrapolate(t, r) with the
// pre-update residuals as error vectors; then tau (ccsd.h:389).
void CCSDContext::update() {
  if (!have_eps_) throw std::logic_error("orbital energies not set");
  if (!solver_ready_) throw std::logic_error("solver not initialised");
  const long n1 = (long)o_ * v_, n2 = n1 * n1;
  // --- DIIS: push the error (copy of r) into the history
  ++diis_iter_;
  const double scale = 1.0 + diis_damp_;
  if ((int)err_slots_.size() == n_diis_) {
    err_slots_.pop_front();
    // drop first row/column of B
    const int m = n_diis_;
    std::vector<double> Bn((size_t)m * m, 0.0);
    for (int i = 1; i < m; ++i)
      for (int j = 1; j < m; ++j) Bn[(size_t)(i - 1) * m + (j - 1)] = B_[(size_t)i * m + j];
    B_ = Bn;
  }
  int eslot = 0;
  {
    std::vector<bool> used(n_diis_, false);
    for (int s : err_slots_) used[s] = true;
    while (used[eslot]) ++eslot;
  }
  err_slots_.push_back(eslot);
  const int nvec = (int)err_slots_.size();
  {
    // the new error vector r goes into its slot in the pass that forms its overlaps with
    // the history (its own overlap read from r itself)
    std::vector<const double*> x1, x2;
    for (int s : err_slots_) {
      x1.push_back(s == eslot ? r1_ : hist_e1_[s]);
      x2.push_back(s == eslot ? r2_ : hist_e2_[s]);
    }
    timed("up:diis_dots", 0.0, [&] {
      multi_dot2(x1.data(), x2.data(), nvec, r1_, r2_, n1, n2, scalars_ + 16, partials_, stream_, hist_e1_[eslot],
                 hist_e2_[eslot]);
    });
    double h[16];
    MPQC_HIP_CHECK(hipMemcpyAsync(h, scalars_ + 16, nvec * sizeof(double), hipMemcpyDeviceToHost, stream_));
    MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
    const int m = n_diis_;
    for (int i = 0; i < nvec; ++i) B_[(size_t)i * m + (nvec - 1)] = B_[(size_t)(nvec - 1) * m + i] = h[i];
  }
  // --- Jacobi: t += r / D, the new t also written into its history slot (x = t)
  if ((int)x_slots_.size() == n_diis_) x_slots_.pop_front();
  int xslot = 0;
  {
    std::vector<bool> used(n_diis_, false);
    for (int s : x_slots_) used[s] = true;
    while (used[xslot]) ++xslot;
  }
  timed("up:jacobi", 0.0, [&] {
    jacobi_update(t1_, t2_, r1_, r2_, eo_, ev_, o_, v_, stream_, hist_x1_[xslot], hist_x2_[xslot]);
  });
  x_slots_.push_back(xslot);
  const bool do_extrap =
      diis_iter_ > diis_start_ && ((diis_iter_ - diis_start_) % diis_ngroup_) < diis_group_nstart_;
  std::vector<const double*> x1, x2;  // extrapolation: history amplitudes and their coefficients
  std::vector<double> c;
  if (do_extrap) {
    const double zero_determinant = 1.0e-15, zero_norm = 1.0e-10;
    const int m = n_diis_;
    int nskip = 0;
    double absdet = 0.0;
    std::vector<double> C;
    while (true) {
      const int rank = nvec - nskip + 1;
      std::vector<double> A((size_t)rank * rank, 0.0), rhs(rank, 0.0);
      for (int i = 0; i < rank; ++i) { A[i] = -1.0; A[(size_t)i * rank] = -1.0; }
      A[0] = 0.0;
      rhs[0] = -1.0;
      double norm = 1.0;
      const double b00 = B_[(size_t)nskip * m + nskip];
      if (std::fabs(b00) > zero_norm) norm = 1.0 / std::fabs(b00);
      for (int i = 1; i < rank; ++i)
        for (int j = 1; j < rank; ++j)
          A[(size_t)i * rank + j] = B_[(size_t)(i - 1 + nskip) * m + (j - 1 + nskip)] * norm;
      for (int i = 0; i < rank; ++i) A[(size_t)i * rank + i] *= scale;
      absdet = colpiv_qr_solve(rank, A, rhs, C);
      ++nskip;
      if (!(absdet < zero_determinant && nskip < nvec)) break;
    }
    if (absdet < zero_determinant) throw std::domain_error("DIIS: poorly-conditioned system");
    --nskip;
    int k = 0;
    for (int s : x_slots_) {
      if (k >= nskip) {
        x1.push_back(hist_x1_[s]);
        x2.push_back(hist_x2_[s]);
        c.push_back(C[k - nskip + 1]);
      }
      ++k;
    }
  }
  // tau("a,b,i,j") = t2 + t1 t1                                     ccsd.h:389
  // (with the maxima the fp16 splits need; after an extrapolation, in the pass that writes t2)
  double* maxes = (prec_ == GemmPrecision::F16X3 && amax_.dev) ? amax_.dev + 1 : nullptr;
  if (do_extrap) {
    timed("up:extrapolate", 0.0, [&] {
      lincomb(t1_, x1.data(), c.data(), (int)c.size(), n1, stream_);
      lincomb_tau(t2_, tau_, t1_, x2.data(), c.data(), (int)c.size(), o_, v_, partials_, maxes, stream_);
    });
  } else if (maxes) {
    timed("up:tau", 0.0, [&] { build_tau_absmax(tau_, t2_, t1_, o_, v_, partials_, maxes, stream_); });
  } else {
    timed("up:tau", 0.0, [&] { build_tau(tau_, t2_, t1_, o_, v_, stream_); });
  }
  tau_amax_fresh_ = maxes != nullptr;
  tau_fresh_ = true;
  MPQC_HIP_CHECK(hipStreamSynchronize(stream_));
}

// CCSD::compute_ccsd solver loop (ccsd.h:344-477) with method = "standard".
double CCSDContext::solve(int max_iter, double target_precision, int verbose, int* n_iter) {
  if (!have_eps_) throw std::logic_error("orbital energies not set");
  if (solve_restart_ == 1) {
    init_guess();
  } else {
    if (solve_restart_ == 2) {
      const long V = v_, O = o_;
      perm_add(Rd_, {V, V, O, O}, 0.0, {{t2_, {V, V, O, O}, {0, 1, 2, 3}, 0.5}, {t2_, {V, V, O, O}, {1, 0, 3, 2}, 0.5}});
      copy_d2d(t2_, Rd_, V * V * O * O);
      tau_fresh_ = false;
      tau_amax_fresh_ = false;
    }
    residual();  // the residual of the current amplitudes in the current form
  }
  solver_init(diis_start_, n_diis_, diis_damp_, diis_ngroup_, diis_group_nstart_);
  trace_.clear();
  double E0 = 0.0, E1 = 0.0, dE = 0.0, err = 1.0;
  int iter = 0;
  auto t0 = std::chrono::steady_clock::now();
  while (iter < max_iter) {
    energy_error(&E1, &err);  // error (:373) and energy (:377-379); E1 uses t of this pass
    dE = std::fabs(E0 - E1);
    E0 = E1;
    const bool converged = dE <= target_precision * 0.1 && err <= target_precision;  // :141-152
    if (iter == 0 || !converged) {
      update();
      auto t1 = std::chrono::steady_clock::now();
      if (iter > 0) {
        const double sec = std::chrono::duration<double>(t1 - t0).count();
        trace_.push_back({iter, dE, err, E1, sec});
        if (verbose) printf("%3i \t %10.5e \t %10.5e \t %15.12f \t %10.4f \n", iter, dE, err, E1, sec);
      }
    } else {
      auto t1 = std::chrono::steady_clock::now();
      const double sec = std::chrono::duration<double>(t1 - t0).count();
      trace_.push_back({iter, dE, err, E1, sec});
      if (verbose) printf("%3i \t %10.5e \t %10.5e \t %15.12f \t %10.4f \n", iter, dE, err, E1, sec);
      break;
    }
    t0 = std::chrono::steady_clock::now();
    residual();
    ++iter;
  }
  if (iter >= max_iter && verbose) printf("\n Warning!! Exceed Max Iteration! \n");
  if (n_iter) *n_iter = iter;
  return E1;
}

}  // namespace mpqc_hip
