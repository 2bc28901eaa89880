// extern "C" boundary (include/mpqc_ccsd.h).  Every exception thrown below the
// boundary becomes a status code plus a message for mpqc_ccsd_last_error().
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/mpqc_ccsd.h"
#include "batch.hpp"
#include "ccsd_ctx.hpp"

struct mpqc_ccsd_batch {
  mpqc_hip::CCSDBatch* impl = nullptr;
  std::string err;
};

struct mpqc_ccsd_ctx {
  mpqc_hip::CCSDContext* impl = nullptr;
  std::string err;
};

namespace {

thread_local std::string g_last_error;  // for failures without a context

int status_of(const std::exception& e) {
  if (dynamic_cast<const mpqc_hip::HipError*>(&e)) return MPQC_CCSD_ERR_HIP;
  if (dynamic_cast<const std::bad_alloc*>(&e)) return MPQC_CCSD_ERR_NOMEM;
  if (dynamic_cast<const std::domain_error*>(&e)) {
    const std::string w = e.what();
    return w.find("DIIS") != std::string::npos ? MPQC_CCSD_ERR_DIIS : MPQC_CCSD_ERR_ARG;
  }
  if (dynamic_cast<const std::range_error*>(&e)) return MPQC_CCSD_ERR_NONFINITE;
  if (dynamic_cast<const std::runtime_error*>(&e) &&
      std::string(e.what()).find("did not converge") != std::string::npos)
    return MPQC_CCSD_ERR_NOCONV;
  if (dynamic_cast<const std::logic_error*>(&e) && !dynamic_cast<const std::invalid_argument*>(&e))
    return MPQC_CCSD_ERR_STATE;
  return MPQC_CCSD_ERR_ARG;
}

template <typename F>
int guarded(mpqc_ccsd_ctx* ctx, F&& f) {
  if (ctx == nullptr || ctx->impl == nullptr) {
    g_last_error = "null context";
    return MPQC_CCSD_ERR_ARG;
  }
  try {
    f(*ctx->impl);
    return MPQC_CCSD_OK;
  } catch (const std::bad_alloc&) {
    ctx->err = "device or host allocation failed";
    return MPQC_CCSD_ERR_NOMEM;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return status_of(e);
  }
}

template <typename F>
int guarded_batch(mpqc_ccsd_batch* b, F&& f) {
  if (b == nullptr || b->impl == nullptr) {
    g_last_error = "null batch";
    return MPQC_CCSD_ERR_ARG;
  }
  try {
    f(*b->impl);
    return MPQC_CCSD_OK;
  } catch (const std::bad_alloc&) {
    b->err = "device or host allocation failed";
    return MPQC_CCSD_ERR_NOMEM;
  } catch (const std::exception& e) {
    b->err = e.what();
    return status_of(e);
  }
}

template <typename F>
int guarded_free(F&& f) {
  try {
    f();
    return MPQC_CCSD_OK;
  } catch (const std::bad_alloc&) {
    g_last_error = "device or host allocation failed";
    return MPQC_CCSD_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return status_of(e);
  }
}

// Split-K workspace for the context-free GEMM entry points: one per device,
// grown on demand, kept until process exit.
struct FreeWorkspace {
  std::mutex m;
  double* ptr[64] = {nullptr};
  size_t bytes[64] = {0};
};
FreeWorkspace g_ws;

mpqc_hip::GemmWorkspace workspace_for_current_device(size_t want) {
  int dev = 0;
  MPQC_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lk(g_ws.m);
  if (g_ws.bytes[dev] < want) {
    if (g_ws.ptr[dev]) MPQC_HIP_CHECK(hipFree(g_ws.ptr[dev]));
    g_ws.ptr[dev] = nullptr;
    g_ws.bytes[dev] = 0;
    void* p = nullptr;
    MPQC_HIP_CHECK(hipMalloc(&p, want));
    g_ws.ptr[dev] = (double*)p;
    g_ws.bytes[dev] = want;
  }
  return {g_ws.ptr[dev], g_ws.bytes[dev]};
}

// Scratch of the allocation-free GEMM entry: a fixed device array (one copy per
// GPU, loaded with the library's code object and held until the process ends).
// Host code never uses the array's name as an address: the device address of each
// GPU's copy comes from hipGetSymbolAddress on that entry's first call on the
// device and is kept per device.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

struct NoAllocScratch {
  std::mutex m;
  double* ptr[64] = {nullptr};
};
NoAllocScratch g_scratch;

mpqc_hip::GemmWorkspace noalloc_scratch_for_current_device() {
  int dev = 0;
  MPQC_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch.m);
  if (g_scratch.ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw mpqc_hip::HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch.ptr[dev] = (double*)p;
  }
  return {g_scratch.ptr[dev], kNoAllocScratchBytes};
}

}  // namespace

extern "C" {

const char* mpqc_ccsd_version(void) { return "mpqc_amd-ccsd 0.1 (gfx950)"; }

int mpqc_ccsd_ctx_create(int o, int v, const int64_t* occ_hashmarks, int n_occ_tiles,
                         const int64_t* vir_hashmarks, int n_vir_tiles, int precision,
                         mpqc_ccsd_ctx** out) {
  if (out == nullptr) {
    g_last_error = "null output pointer";
    return MPQC_CCSD_ERR_ARG;
  }
  *out = nullptr;
  return guarded_free([&] {
    auto* c = new mpqc_ccsd_ctx;
    try {
      c->impl = new mpqc_hip::CCSDContext(o, v, occ_hashmarks, n_occ_tiles, vir_hashmarks,
                                          n_vir_tiles, precision);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int mpqc_ccsd_ctx_destroy(mpqc_ccsd_ctx* ctx) {
  if (ctx == nullptr) return MPQC_CCSD_OK;
  delete ctx->impl;
  delete ctx;
  return MPQC_CCSD_OK;
}

const char* mpqc_ccsd_last_error(const mpqc_ccsd_ctx* ctx) {
  if (ctx == nullptr) return g_last_error.c_str();
  return ctx->err.c_str();
}

int mpqc_ccsd_set_option(mpqc_ccsd_ctx* ctx, const char* key, double value) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (!key) throw std::invalid_argument("null option key");
    c.set_option(key, value);
  });
}

int mpqc_ccsd_set_integral(mpqc_ccsd_ctx* ctx, int which, const double* data, int layout,
                           int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.set_integral(which, data, layout, memspace); });
}

int mpqc_ccsd_set_direct_ao(mpqc_ccsd_ctx* ctx, int nbf, const double* ao_eri, const double* c_occ,
                            const double* c_vir, int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (memspace == 1) MPQC_HIP_CHECK(hipDeviceSynchronize());  // the caller's producers of its buffers
    c.set_direct_ao(nbf, ao_eri, c_occ, c_vir, memspace);
  });
}

int mpqc_ccsd_set_df_basis(mpqc_ccsd_ctx* ctx, int naux, const int64_t* aux_hashmarks,
                           int n_aux_tiles) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.set_df_basis(naux, aux_hashmarks, n_aux_tiles); });
}

int mpqc_ccsd_triples(mpqc_ccsd_ctx* ctx, double* e_t, double* seconds) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (!e_t) throw std::invalid_argument("null (T) output");
    *e_t = c.triples(c.triples_block_bytes_, seconds);
  });
}

int mpqc_ccsd_eom_ee(mpqc_ccsd_ctx* ctx, int n_roots, int max_iter, double convergence, int verbose,
                     double* energies, int* n_iter) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    const int it = c.eom_ee(n_roots, max_iter, convergence, verbose, energies);
    if (n_iter) *n_iter = it;
  });
}

int mpqc_ccsd_eom_stats(mpqc_ccsd_ctx* ctx, int* n_products, double* product_seconds) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (n_products) *n_products = c.eom_products_;
    if (product_seconds) *product_seconds = c.eom_product_seconds_;
  });
}

int mpqc_ccsd_get_integral(mpqc_ccsd_ctx* ctx, int which, double* out, int layout, int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.get_integral(which, out, layout, memspace); });
}

int mpqc_ccsd_transform_ao(mpqc_ccsd_ctx* ctx, int nbf, const double* ao_eri, const double* fock_ao,
                           const double* c_occ, const double* c_vir, int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (memspace == 1) MPQC_HIP_CHECK(hipDeviceSynchronize());  // the caller's producers of its buffers
    c.transform_ao(nbf, ao_eri, fock_ao, c_occ, c_vir, memspace);
  });
}

int mpqc_ccsd_transform_df(mpqc_ccsd_ctx* ctx, int nbf, const double* ao_three_center, const double* metric,
                           const double* fock_ao, const double* c_occ, const double* c_vir, int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (memspace == 1) MPQC_HIP_CHECK(hipDeviceSynchronize());  // the caller's producers of its buffers
    c.transform_df(nbf, ao_three_center, metric, fock_ao, c_occ, c_vir, memspace);
  });
}

int mpqc_ccsd_set_orbital_energies(mpqc_ccsd_ctx* ctx, const double* eps_occ_act,
                                   const double* eps_vir) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.set_orbital_energies(eps_occ_act, eps_vir); });
}

int mpqc_ccsd_solver_init(mpqc_ccsd_ctx* ctx, int diis_start, int n_diis, double diis_damp,
                          int diis_ngroup, int diis_group_nstart) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    c.solver_init(diis_start, n_diis, diis_damp, diis_ngroup, diis_group_nstart);
  });
}

int mpqc_ccsd_set_amplitudes(mpqc_ccsd_ctx* ctx, const double* t1, const double* t2, int layout,
                             int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (!t1 || !t2) throw std::invalid_argument("null amplitude buffer");
    c.set_amplitudes(t1, t2, layout, memspace);
  });
}

int mpqc_ccsd_get_amplitudes(mpqc_ccsd_ctx* ctx, double* t1, double* t2, int layout, int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.get_amplitudes(t1, t2, layout, memspace); });
}

int mpqc_ccsd_get_residuals(mpqc_ccsd_ctx* ctx, double* r1, double* r2, int layout, int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.get_residuals(r1, r2, layout, memspace); });
}

int mpqc_ccsd_get_dressed_fock(mpqc_ccsd_ctx* ctx, double* FIJ, double* FAB) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.get_dressed_fock(FIJ, FAB); });
}

int mpqc_ccsd_init_guess(mpqc_ccsd_ctx* ctx) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.init_guess(); });
}

int mpqc_ccsd_residual(mpqc_ccsd_ctx* ctx) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.residual(); });
}

int mpqc_ccsd_energy_error(mpqc_ccsd_ctx* ctx, double* energy, double* error) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.energy_error(energy, error); });
}

int mpqc_ccsd_update(mpqc_ccsd_ctx* ctx) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.update(); });
}

int mpqc_ccsd_solve(mpqc_ccsd_ctx* ctx, int max_iter, double target_precision, int verbose,
                    double* e_corr, int* n_iter) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (max_iter < 1) throw std::invalid_argument("max_iter must be >= 1");
    if (!(target_precision > 0.0)) throw std::invalid_argument("target_precision must be > 0");
    const double e = c.solve(max_iter, target_precision, verbose, n_iter);
    if (e_corr) *e_corr = e;
  });
}

int mpqc_ccsd_get_trace(mpqc_ccsd_ctx* ctx, int max_rows, int* n_rows, int* iters, double* dE,
                        double* error, double* energy, double* seconds) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    const auto& t = c.trace();
    const int n = (int)std::min<size_t>(t.size(), (size_t)std::max(0, max_rows));
    for (int k = 0; k < n; ++k) {
      if (iters) iters[k] = t[k].iter;
      if (dE) dE[k] = t[k].dE;
      if (error) error[k] = t[k].error;
      if (energy) energy[k] = t[k].energy;
      if (seconds) seconds[k] = t[k].seconds;
    }
    if (n_rows) *n_rows = (int)t.size();
  });
}

int mpqc_ccsd_jacobian_transpose(mpqc_ccsd_ctx* ctx, const double* y1, const double* y2, double* x1,
                                 double* x2, int layout, int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (!y1 || !y2 || !x1 || !x2) throw std::invalid_argument("null buffer");
    c.jacobian_transpose_io(y1, y2, x1, x2, layout, memspace);
  });
}

int mpqc_ccsd_lambda_solve(mpqc_ccsd_ctx* ctx, int max_iter, double target_precision, int verbose,
                           int* n_iter) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    if (!(target_precision > 0.0)) throw std::invalid_argument("target_precision must be > 0");
    const int n = c.lambda_solve(max_iter, target_precision, verbose);
    if (n_iter) *n_iter = n;
  });
}

int mpqc_ccsd_get_lambda(mpqc_ccsd_ctx* ctx, double* l1, double* l2, int layout, int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.get_lambda(l1, l2, layout, memspace); });
}

int mpqc_ccsd_energy_gradient(mpqc_ccsd_ctx* ctx) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) { c.energy_gradient(); });
}

int mpqc_ccsd_get_integral_gradient(mpqc_ccsd_ctx* ctx, int which, double* out, int layout,
                                    int memspace) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    c.get_integral_gradient(which, out, layout, memspace);
  });
}

int mpqc_ccsd_batch_create(int n_mol, const int* o, const int* v, int precision,
                           mpqc_ccsd_batch** out) {
  if (out == nullptr) {
    g_last_error = "null output pointer";
    return MPQC_CCSD_ERR_ARG;
  }
  *out = nullptr;
  return guarded_free([&] {
    auto* b = new mpqc_ccsd_batch;
    try {
      b->impl = new mpqc_hip::CCSDBatch(n_mol, o, v, precision);
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

int mpqc_ccsd_batch_destroy(mpqc_ccsd_batch* b) {
  if (b == nullptr) return MPQC_CCSD_OK;
  delete b->impl;
  delete b;
  return MPQC_CCSD_OK;
}

const char* mpqc_ccsd_batch_last_error(const mpqc_ccsd_batch* b) {
  if (b == nullptr) return g_last_error.c_str();
  return b->err.c_str();
}

int mpqc_ccsd_batch_set_integral(mpqc_ccsd_batch* b, int mol, int which, const double* data,
                                 int layout, int memspace) {
  return guarded_batch(b, [&](mpqc_hip::CCSDBatch& x) {
    x.mol(mol).set_integral(which, data, layout, memspace);
  });
}

int mpqc_ccsd_batch_set_amplitudes(mpqc_ccsd_batch* b, int mol, const double* t1, const double* t2,
                                   int layout, int memspace) {
  return guarded_batch(b, [&](mpqc_hip::CCSDBatch& x) {
    if (!t1 || !t2) throw std::invalid_argument("null amplitude buffer");
    x.mol(mol).set_amplitudes(t1, t2, layout, memspace);
  });
}

int mpqc_ccsd_batch_residual(mpqc_ccsd_batch* b) {
  return guarded_batch(b, [&](mpqc_hip::CCSDBatch& x) { x.residual(); });
}

int mpqc_ccsd_batch_get_residuals(mpqc_ccsd_batch* b, int mol, double* r1, double* r2, int layout,
                                  int memspace) {
  return guarded_batch(b, [&](mpqc_hip::CCSDBatch& x) {
    x.mol(mol).get_residuals(r1, r2, layout, memspace);
  });
}

int mpqc_ccsd_batch_tune(mpqc_ccsd_batch* b, int reps, int verbose, double* ms) {
  return guarded_batch(b, [&](mpqc_hip::CCSDBatch& x) {
    const double t = x.tune(reps, verbose);
    if (ms) *ms = t;
  });
}

int mpqc_ccsd_batch_time(mpqc_ccsd_batch* b, int reps, double* ms) {
  return guarded_batch(b, [&](mpqc_hip::CCSDBatch& x) {
    const double t = x.time(reps);
    if (ms) *ms = t;
  });
}

int mpqc_ccsd_batch_time_stages(mpqc_ccsd_batch* b, int reps) {
  return guarded_batch(b, [&](mpqc_hip::CCSDBatch& x) {
    if (reps < 1) throw std::invalid_argument("reps must be positive");
    x.time_stages(reps);
  });
}

int mpqc_ccsd_batch_get_stages(mpqc_ccsd_batch* b, int max_rows, int* n_rows, char* tags, int* kinds,
                               int* bm, int64_t* target, double* ms, double* flops) {
  return guarded_batch(b, [&](mpqc_hip::CCSDBatch& x) {
    const auto st = x.stages();
    const int n = (int)std::min<size_t>(st.size(), (size_t)std::max(0, max_rows));
    for (int k = 0; k < n; ++k) {
      if (tags) {
        std::memset(tags + 64 * k, 0, 64);
        std::strncpy(tags + 64 * k, st[k].tag.c_str(), 63);
      }
      if (kinds) kinds[k] = st[k].kind;
      if (bm) bm[k] = st[k].bm;
      if (target) target[k] = st[k].target;
      if (ms) ms[k] = st[k].ms;
      if (flops) flops[k] = st[k].flops;
    }
    if (n_rows) *n_rows = (int)st.size();
  });
}

int mpqc_ccsd_get_kernel_times(mpqc_ccsd_ctx* ctx, int max_rows, int* n_rows, char* names,
                               int64_t* calls, double* ms, double* flops) {
  return guarded(ctx, [&](mpqc_hip::CCSDContext& c) {
    const auto kt = c.kernel_times();
    const int n = (int)std::min<size_t>(kt.size(), (size_t)std::max(0, max_rows));
    for (int k = 0; k < n; ++k) {
      if (names) {
        std::memset(names + 64 * k, 0, 64);
        std::strncpy(names + 64 * k, kt[k].tag.c_str(), 63);
      }
      if (calls) calls[k] = kt[k].calls;
      if (ms) ms[k] = kt[k].ms;
      if (flops) flops[k] = kt[k].flops;
    }
    if (n_rows) *n_rows = (int)kt.size();
  });
}

int mpqc_ccsd_gemm(int transa, int transb, int m, int n, int k, double alpha, const double* a,
                   int64_t lda, int64_t stride_a, const double* b, int64_t ldb, int64_t stride_b,
                   double beta, double* c, int64_t ldc, int64_t stride_c, int batch, int precision,
                   void* stream) {
  return guarded_free([&] {
    if (m < 0 || n < 0 || k <= 0 || batch < 1) throw std::invalid_argument("invalid GEMM sizes");
    if (!a || !b || !c) throw std::invalid_argument("null GEMM operand");
    if (precision < 0 || precision > 4) throw std::invalid_argument("unknown precision");
    mpqc_hip::GemmDesc d = mpqc_hip::make_gemm(transa != 0, transb != 0, m, n, k, alpha, a, lda, b,
                                               ldb, beta, c, ldc);
    d.sA = stride_a; d.sB = stride_b; d.sC = stride_c; d.batch = batch;
    const size_t want = std::max<size_t>((size_t)m * n * batch * 64 * sizeof(double), (size_t)64 << 20);
    auto ws = workspace_for_current_device(std::min<size_t>(want, (size_t)1 << 30));
    mpqc_hip::gemm(d, (mpqc_hip::GemmPrecision)precision, ws, (hipStream_t)stream);
  });
}

int mpqc_ccsd_gemm_noalloc(int transa, int transb, int m, int n, int k, double alpha, const double* a,
                           int64_t lda, int64_t stride_a, const double* b, int64_t ldb, int64_t stride_b,
                           double beta, double* c, int64_t ldc, int64_t stride_c, int batch,
                           int precision, void* stream) {
  return guarded_free([&] {
    if (m < 0 || n < 0 || k <= 0 || batch < 1) throw std::invalid_argument("invalid GEMM sizes");
    if (!a || !b || !c) throw std::invalid_argument("null GEMM operand");
    if (precision < 0 || precision > 4) throw std::invalid_argument("unknown precision");
    mpqc_hip::GemmDesc d = mpqc_hip::make_gemm(transa != 0, transb != 0, m, n, k, alpha, a, lda, b,
                                               ldb, beta, c, ldc);
    d.sA = stride_a; d.sB = stride_b; d.sC = stride_c; d.batch = batch;
    // the split-K plan is made for the fixed scratch (fewer K splits where the
    // partial sums of more would not fit); a plan that still needs more is refused
    const size_t need = mpqc_hip::gemm_scratch_bytes(d, (mpqc_hip::GemmPrecision)precision,
                                                     kNoAllocScratchBytes);
    if (need > kNoAllocScratchBytes)
      throw std::invalid_argument("GEMM needs " + std::to_string(need) +
                                  " bytes of scratch; the allocation-free entry holds at most " +
                                  std::to_string(kNoAllocScratchBytes));
    mpqc_hip::gemm(d, (mpqc_hip::GemmPrecision)precision, noalloc_scratch_for_current_device(),
                   (hipStream_t)stream);
  });
}

int mpqc_ccsd_pp_ladder(int o, int v, const double* gabcd_dev, const double* tau_dev,
                        double* r2_dev, double alpha, double beta, int precision, void* stream) {
  return guarded_free([&] {
    if (o <= 0 || v <= 0) throw std::invalid_argument("o and v must be positive");
    if (!gabcd_dev || !tau_dev || !r2_dev) throw std::invalid_argument("null operand");
    if (precision < 0 || precision > 4) throw std::invalid_argument("unknown precision");
    const long vv = (long)v * v, oo = (long)o * o;
    if (vv > 0x7fffffffL) throw std::invalid_argument("v^2 too large");
    // b("a,b,i,j") = tau("c,d,i,j") * Gabcd("a,b,c,d"): M = v^2, N = o^2, K = v^2
    mpqc_hip::GemmDesc d = mpqc_hip::make_gemm(false, false, (int)vv, (int)oo, (int)vv, alpha,
                                               gabcd_dev, vv, tau_dev, oo, beta, r2_dev, oo);
    auto ws = workspace_for_current_device((size_t)1 << 20);
    mpqc_hip::gemm(d, (mpqc_hip::GemmPrecision)precision, ws, (hipStream_t)stream);
  });
}

}  // extern "C"
