// A C++ host of the C-ABI (include/mpqc_ccsd.h), the way MPQC's
// CCSD::compute_ccsd (ccsd.h:268-477) would call it: TiledArray tile layout
// (TiledRange1 hashmarks of trange1_engine.cpp:10-20), host buffers, status
// codes turned into exceptions by the caller.
//
//   abi_c_check cpu            error paths that need no GPU (argument checks
//                              happen before any HIP call)
//   abi_c_check gpu FIXTURE    one residual (cc::compute_cs_ccsd_r1 + r2) and
//                              one converged solve on a small Hamiltonian,
//                              against the oracle's values in FIXTURE
//                              (tests/golden/abi_residual_o2_v5.txt)
//
// Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpqc_ccsd.h"

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  if (!ok) ++failures;
}

// what a host wrapper does with a status: throw with the library's message
void check(int status, const mpqc_ccsd_ctx* ctx, const char* call) {
  if (status != MPQC_CCSD_OK)
    throw std::runtime_error(std::string(call) + ": status " + std::to_string(status) + ": " +
                             mpqc_ccsd_last_error(ctx));
}

int run_cpu() {
  const char* ver = mpqc_ccsd_version();
  expect(ver && std::strstr(ver, "gfx950"), "version names gfx950");

  mpqc_ccsd_ctx* ctx = reinterpret_cast<mpqc_ccsd_ctx*>(0x1);
  int st = mpqc_ccsd_ctx_create(0, 5, nullptr, 0, nullptr, 0, MPQC_CCSD_PREC_F64, &ctx);
  expect(st == MPQC_CCSD_ERR_ARG && ctx == nullptr, "o = 0 rejected, output cleared");
  expect(std::strstr(mpqc_ccsd_last_error(nullptr), "positive") != nullptr, "message says why");

  st = mpqc_ccsd_ctx_create(2, 5, nullptr, 0, nullptr, 0, MPQC_CCSD_PREC_F64, nullptr);
  expect(st == MPQC_CCSD_ERR_ARG, "null output pointer rejected");

  st = mpqc_ccsd_ctx_create(2, 5, nullptr, 0, nullptr, 0, 9, &ctx);
  expect(st == MPQC_CCSD_ERR_ARG && std::strstr(mpqc_ccsd_last_error(nullptr), "precision"),
         "unknown precision rejected");

  const int64_t bad_occ[] = {0, 2, 1, 2};
  st = mpqc_ccsd_ctx_create(2, 5, bad_occ, 3, nullptr, 0, MPQC_CCSD_PREC_F64, &ctx);
  expect(st == MPQC_CCSD_ERR_ARG && std::strstr(mpqc_ccsd_last_error(nullptr), "increase"),
         "non-increasing hashmarks rejected");
  const int64_t short_vir[] = {0, 2, 4};
  st = mpqc_ccsd_ctx_create(2, 5, nullptr, 0, short_vir, 2, MPQC_CCSD_PREC_F64, &ctx);
  expect(st == MPQC_CCSD_ERR_ARG && std::strstr(mpqc_ccsd_last_error(nullptr), "span"),
         "hashmarks not spanning [0, v] rejected");

  expect(mpqc_ccsd_set_option(nullptr, "fov_sign", 1.0) == MPQC_CCSD_ERR_ARG &&
             std::strstr(mpqc_ccsd_last_error(nullptr), "null context"),
         "null context rejected");
  expect(mpqc_ccsd_residual(nullptr) == MPQC_CCSD_ERR_ARG, "residual on a null context rejected");
  expect(mpqc_ccsd_ctx_destroy(nullptr) == MPQC_CCSD_OK, "destroying a null context is a no-op");

  double dummy = 0.0;
  st = mpqc_ccsd_gemm(0, 0, -1, 4, 4, 1.0, &dummy, 4, 0, &dummy, 4, 0, 0.0, &dummy, 4, 0, 1,
                      MPQC_CCSD_PREC_F64, nullptr);
  expect(st == MPQC_CCSD_ERR_ARG, "negative GEMM size rejected");
  st = mpqc_ccsd_pp_ladder(0, 4, &dummy, &dummy, &dummy, 1.0, 0.0, MPQC_CCSD_PREC_F64, nullptr);
  expect(st == MPQC_CCSD_ERR_ARG, "empty ladder rejected");
  st = mpqc_ccsd_pp_ladder(2, 4, nullptr, &dummy, &dummy, 1.0, 0.0, MPQC_CCSD_PREC_F64, nullptr);
  expect(st == MPQC_CCSD_ERR_ARG, "null ladder operand rejected");
  return failures == 0 ? 0 : 1;
}

struct Fixture {
  int o = 0, v = 0;
  std::vector<int64_t> occ_hm, vir_hm;
  std::map<std::string, std::vector<double>> rec;
};

Fixture read_fixture(const char* path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  Fixture f;
  in >> f.o >> f.v;
  for (auto* hm : {&f.occ_hm, &f.vir_hm}) {
    int nt = 0;
    in >> nt;
    hm->resize(nt + 1);
    for (auto& x : *hm) in >> x;
  }
  std::string name;
  long n = 0;
  while (in >> name >> n) {
    std::vector<double> a(n);
    for (auto& x : a) in >> x;
    f.rec[name] = std::move(a);
  }
  if (!in.eof()) throw std::runtime_error("malformed fixture");
  return f;
}

// Dense row-major tensor -> concatenation of row-major tiles in row-major tile
// order (TA::DistArray's layout over the given hashmarks per mode).
std::vector<double> to_tiles(const std::vector<double>& x, const std::vector<const std::vector<int64_t>*>& hm) {
  const int d = (int)hm.size();
  std::vector<long> ext(d), stride(d);
  for (int k = 0; k < d; ++k) ext[k] = hm[k]->back();
  stride[d - 1] = 1;
  for (int k = d - 2; k >= 0; --k) stride[k] = stride[k + 1] * ext[k + 1];
  std::vector<double> out;
  out.reserve(x.size());
  std::vector<int> t(d, 0);
  std::function<void(int, long)> tile_loop;  // elements of one tile, row-major
  std::vector<long> lo(d), hi(d);
  tile_loop = [&](int k, long off) {
    if (k == d) {
      out.push_back(x[off]);
      return;
    }
    for (long i = lo[k]; i < hi[k]; ++i) tile_loop(k + 1, off + i * stride[k]);
  };
  while (true) {
    for (int k = 0; k < d; ++k) {
      lo[k] = (*hm[k])[t[k]];
      hi[k] = (*hm[k])[t[k] + 1];
    }
    tile_loop(0, 0);
    int k = d - 1;
    while (k >= 0 && ++t[k] == (int)hm[k]->size() - 1) t[k--] = 0;
    if (k < 0) break;
  }
  return out;
}

double max_abs(const std::vector<double>& a) {
  double m = 0.0;
  for (double x : a) m = std::fmax(m, std::fabs(x));
  return m;
}

double max_diff(const std::vector<double>& a, const std::vector<double>& b) {
  double m = 0.0;
  for (size_t i = 0; i < a.size(); ++i) m = std::fmax(m, std::fabs(a[i] - b[i]));
  return m;
}

int run_gpu(const char* path) {
  const Fixture f = read_fixture(path);
  const int o = f.o, v = f.v;
  const auto* O = &f.occ_hm;
  const auto* V = &f.vir_hm;
  const std::map<std::string, std::pair<int, std::vector<const std::vector<int64_t>*>>> blocks = {
      {"Fia", {MPQC_INT_FIA, {O, V}}},           {"Fij", {MPQC_INT_FIJ, {O, O}}},
      {"Fab", {MPQC_INT_FAB, {V, V}}},           {"Gijab", {MPQC_INT_GIJAB, {O, O, V, V}}},
      {"Gijkl", {MPQC_INT_GIJKL, {O, O, O, O}}}, {"Giajb", {MPQC_INT_GIAJB, {O, V, O, V}}},
      {"Giabc", {MPQC_INT_GIABC, {O, V, V, V}}}, {"Gijka", {MPQC_INT_GIJKA, {O, O, O, V}}},
      {"Gabcd", {MPQC_INT_GABCD, {V, V, V, V}}}};
  auto setup = [&](mpqc_ccsd_ctx** ctx) {
    check(mpqc_ccsd_ctx_create(o, v, f.occ_hm.data(), (int)f.occ_hm.size() - 1, f.vir_hm.data(),
                               (int)f.vir_hm.size() - 1, MPQC_CCSD_PREC_F64, ctx),
          nullptr, "ctx_create");
    for (const auto& kv : blocks) {
      const auto tiles = to_tiles(f.rec.at(kv.first), kv.second.second);
      check(mpqc_ccsd_set_integral(*ctx, kv.second.first, tiles.data(), MPQC_LAYOUT_TA_TILES, MPQC_MEM_HOST),
            *ctx, "set_integral");
    }
    check(mpqc_ccsd_set_orbital_energies(*ctx, f.rec.at("eps_occ").data(), f.rec.at("eps_vir").data()), *ctx,
          "set_orbital_energies");
    check(mpqc_ccsd_solver_init(*ctx, 1, 8, 0.0, 1, 1), *ctx, "solver_init");
  };

  mpqc_ccsd_ctx* ctx = nullptr;
  setup(&ctx);
  // residual at the fixture's amplitudes, amplitudes handed over as tiles
  const auto t1 = to_tiles(f.rec.at("t1"), {V, O});
  const auto t2 = to_tiles(f.rec.at("t2"), {V, V, O, O});
  check(mpqc_ccsd_set_amplitudes(ctx, t1.data(), t2.data(), MPQC_LAYOUT_TA_TILES, MPQC_MEM_HOST), ctx,
        "set_amplitudes");
  check(mpqc_ccsd_residual(ctx), ctx, "residual");
  std::vector<double> r1((size_t)v * o), r2((size_t)v * v * o * o);
  check(mpqc_ccsd_get_residuals(ctx, r1.data(), r2.data(), MPQC_LAYOUT_DENSE, MPQC_MEM_HOST), ctx,
        "get_residuals");
  const auto& r1x = f.rec.at("r1");
  const auto& r2x = f.rec.at("r2");
  const double e1 = max_diff(r1, r1x) / max_abs(r1x), e2 = max_diff(r2, r2x) / max_abs(r2x);
  std::printf("residual vs oracle: r1 %.2e r2 %.2e (relative)\n", e1, e2);
  expect(e1 < 1e-12 && e2 < 1e-12, "residual matches the oracle to 1e-12");
  // the same residual read back in tile layout
  std::vector<double> r2t(r2.size());
  check(mpqc_ccsd_get_residuals(ctx, nullptr, r2t.data(), MPQC_LAYOUT_TA_TILES, MPQC_MEM_HOST), ctx,
        "get_residuals (tiles)");
  expect(max_diff(r2t, to_tiles(r2, {V, V, O, O})) == 0.0, "tile-layout read-back equals the dense one");
  // a state error comes back as a status with a message, and the context stays usable
  expect(mpqc_ccsd_get_lambda(ctx, r1.data(), nullptr, MPQC_LAYOUT_DENSE, MPQC_MEM_HOST) == MPQC_CCSD_ERR_STATE &&
             std::strlen(mpqc_ccsd_last_error(ctx)) > 0,
         "Lambda read before its solve is a state error");
  double E = 0.0;
  int n_iter = 0;
  check(mpqc_ccsd_solve(ctx, 100, 1e-12, 0, &E, &n_iter), ctx, "solve");
  const double Ex = f.rec.at("E_ccsd")[0];
  std::printf("solve: E %.14f oracle %.14f in %d iterations\n", E, Ex, n_iter);
  expect(std::fabs(E - Ex) < 1e-10, "converged energy matches the oracle to 1e-10 Eh");
  check(mpqc_ccsd_ctx_destroy(ctx), nullptr, "ctx_destroy");
  return failures == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc >= 2 && std::strcmp(argv[1], "cpu") == 0) return run_cpu();
    if (argc >= 3 && std::strcmp(argv[1], "gpu") == 0) return run_gpu(argv[2]);
    std::fprintf(stderr, "usage: %s cpu | gpu FIXTURE\n", argv[0]);
    return 2;
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 1;
  }
}
