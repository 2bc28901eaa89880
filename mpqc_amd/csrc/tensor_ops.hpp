// HBM-streaming tensor kernels for the CCSD iteration: permuted linear
// combinations (the operand-layout changes TiledArray performs implicitly when
// an expression's index order differs from storage order), amplitude updates
// and deterministic reductions.
#pragma once

#include <vector>

#include "common.hpp"

namespace mpqc_hip {

constexpr int kMaxTerms = 10;

// One input of a permuted linear combination: element (i0,i1,i2,i3) of the
// output reads p[i0*s[0] + i1*s[1] + i2*s[2] + i3*s[3]].
struct TTerm {
  const double* p = nullptr;
  long s[4] = {0, 0, 0, 0};
  double alpha = 1.0;
  long vol = 0;  // element count of the input tensor (used for range bounds)
};

// Optional product term of a permuted combination: element (i0,i1,i2,i3) adds
// alpha * x[sum_d i_d sx[d]] * y[sum_d i_d sy[d]] (e.g. t1(d,i) t1(a,l)).
struct TProd {
  const double* x = nullptr;
  long sx[4] = {0, 0, 0, 0};
  long nx = 0;  // element count of x
  const double* y = nullptr;
  long sy[4] = {0, 0, 0, 0};
  long ny = 0;
  double alpha = 1.0;
};

// out[i0,i1,i2,i3] (dense row-major, extents n) = beta*out + sum_t alpha_t * in_t[...]
// (+ the product term when prod is given)
void tadd(double* out, const long n[4], double beta, const TTerm* terms, int nterms,
          hipStream_t stream, const TProd* prod = nullptr);

// Strides of a dense row-major tensor with extents e[0..3] (missing trailing
// extents = 1), permuted so that output axis d reads input axis perm[d].
// perm lists, for each output axis, which input axis it is.
void permuted_strides(const long in_extents[4], const int perm[4], long out_s[4]);

// tau(a,b,i,j) = t2(a,b,i,j) + t1(a,i) t1(b,j)                 (ccsd.h:389)
void build_tau(double* tau, const double* t2, const double* t1, int o, int v, hipStream_t s);
// build_tau, and maxes[0] = max |t2|, maxes[1] = max |tau| (the fp16 plane scales of the
// next residual's t2 and tau operands)
void build_tau_absmax(double* tau, const double* t2, const double* t1, int o, int v, double* partials,
                      double* maxes, hipStream_t s);

// TP(a,i,l,d) = 0.5 t2(d,a,i,l) + t1(d,i) t1(a,l)  -- the "T" intermediate of
// ccsd_r1_r2.h:236 laid out as the (ai),(ld) GEMM operand.
void build_TP(double* TP, const double* t2, const double* t1, int o, int v, hipStream_t s);

// Jacobi update (lcao/cc/solvers.h:165-258):
//   t1 += r1 / (e_i - e_a);  t2 += r2 / (e_i + e_j - e_a - e_b)
// (copy1, copy2: optional second destinations of the updated t1, t2)
void jacobi_update(double* t1, double* t2, const double* r1, const double* r2, const double* eps_o,
                   const double* eps_v, int o, int v, hipStream_t s, double* copy1 = nullptr,
                   double* copy2 = nullptr);

// Deterministic reductions; results land in device memory `out` (double).
// sum_k x[k]*y[k] over n elements, written to out[0].
void dot(const double* x, const double* y, long n, double* out, double* partials, hipStream_t s);

// r1(a,i) += sum_{c,k} (2 t2(c,a,k,i) - t2(c,a,i,k)) h(k,c) (ccsd_r1_r2.h:100-101, h as
// (k,c) with row length v), one pass over t2; needs o <= 64 and o v <= kReducePartials.
bool t2bar_h_supported(int o, int v);
void t2bar_h(double* r1, const double* t2, const double* h, int o, int v, double* partials, hipStream_t s);

// Energy (ccsd.h:377-379):
//   2 <f_ai + r1, t1> + <g_abij + r2, 2 tau(a,b,i,j) - tau(b,a,i,j)>
void ccsd_energy(const double* f_ai, const double* r1, const double* t1, const double* g_abij,
                 const double* r2, const double* tau, int o, int v, double* out, double* partials,
                 hipStream_t s);
// The same energy into out3[2] with |r1|^2 into out3[0] and |r2|^2 into out3[1], from one
// pass over the o^2 v^2 tensors (partials: at least 3 * 1024 doubles).
void ccsd_energy_norms(const double* f_ai, const double* r1, const double* t1, const double* g_abij,
                       const double* r2, const double* tau, int o, int v, double* out3, double* partials,
                       hipStream_t s);

// out[k] = <x_k, y> for k < nvec, x_k = X + k*stride (two members: 1- and 2-body)
// (copy1, copy2: optional destinations that receive y1, y2 in the same pass)
void multi_dot2(const double* const* xs1, const double* const* xs2, int nvec, const double* y1,
                const double* y2, long n1, long n2, double* out, double* partials, hipStream_t s,
                double* copy1 = nullptr, double* copy2 = nullptr);

// C_b[m][n] *= beta for b < batch, m < M, n < N (C_b = C + b*sC, row stride ldc);
// beta = 0 stores zeros
void scale_matrix(double* C, int batch, int M, int N, long sC, long ldc, double beta, hipStream_t s);

// x[0..n) = value
void fill(double* x, long n, double value, hipStream_t s);

// y = sum_k c_k x_k, k < nvec (fixed summation order)
void lincomb(double* y, const double* const* xs, const double* c, int nvec, long n, hipStream_t s);
// t2 = sum_k c[k] xs[k] (as lincomb) and tau = t2 + t1(a,i) t1(b,j) (as build_tau) in one
// pass over the o^2 v^2 elements; with maxes, also max |t2| and max |tau| into maxes[0..1]
// (as build_tau_absmax).
void lincomb_tau(double* t2, double* tau, const double* t1, const double* const* xs, const double* c, int nvec,
                 int o, int v, double* partials, double* maxes, hipStream_t s);

// Number of doubles of `partials` scratch the reductions above need.
constexpr long kReducePartials = 4096 * 16;

// TiledArray tile layout <-> dense row-major (4-D; lower rank uses extent-1 modes).
// Each mode d is split by hashmarks hm_d[0..nt_d] (hm_d[0]=0, hm_d[nt_d]=n_d); tiles are
// stored in row-major tile-ordinal order, each tile row-major over its own extents.
// `modeinfo` is a device array holding, for each of the 4 modes d in turn,
// lo[i] (first index of the tile containing i) and ext[i] (that tile's extent)
// for i < n[d]; build it with make_tile_modeinfo().
void make_tile_modeinfo(long* host_out, const long n[4], const long* const hashmarks[4],
                        const int ntiles[4]);
void tiles_to_dense(double* dense, const double* tiled, const long n[4], const long* modeinfo,
                    hipStream_t s);
void dense_to_tiles(double* tiled, const double* dense, const long n[4], const long* modeinfo,
                    hipStream_t s);

// ---- grouped forms (one launch over many independent problems)
// Permuted linear combinations: problem q writes out_q (extents n_q) =
// beta_q out_q + sum_t terms_q[t].
struct GroupedTAddDesc {
  double* out;
  long n[4];
  double beta;
  int nterms;
  TTerm terms[kMaxTerms];
};
// Element-wise builds: kind 0 copy (out = a, n elements), 1 tau (out = tau from
// t2 = a, t1 = b), 2 TP (out = TP from t2 = a, t1 = b) with extents o, v.
struct GroupedEwDesc {
  int kind;
  double* out;
  const double* a;
  const double* b;
  long n;
  int o, v;
};
struct GroupedPlan {
  void* dev = nullptr;  // device array of problem records
  int nprob = 0;
  long items = 0;       // workgroups of the launch
  int kind = 0;         // 0 permutation, 1 element-wise
  size_t lds = 0;       // dynamic LDS of the launch (largest tile of the permutations)
};
GroupedPlan plan_grouped_tadd(const std::vector<GroupedTAddDesc>& probs, hipStream_t s);
GroupedPlan plan_grouped_ew(const std::vector<GroupedEwDesc>& probs, hipStream_t s);
void launch_grouped(const GroupedPlan& plan, hipStream_t s);
void free_grouped(GroupedPlan& plan);

// Count of non-finite values in x[0..n) -> out[0] (as double).
void count_nonfinite(const double* x, long n, double* out, double* partials, hipStream_t s);

}  // namespace mpqc_hip
