// Density-fitted closed-shell CCSD residual (method "df").
//
// The reference's compute_cs_ccsd_r1_df / compute_cs_ccsd_r2_df
// (src/mpqc/chemistry/qc/lcao/cc/ccsd_r1_r2.h:136-184, 309-520) replace every
// term that needs a four-index integral with three or more virtual indices by
// contractions with the three-index factors X(K,p,q) = (K|G|p q)[inv_sqr]
// (ccsd.h:481-493).  The residual itself lives in CCSDContext::residual(),
// which branches on df_; this file holds the set-up, the X_ai_tau
// intermediate and the lazily built (ab|cd) ladder of reduced_abcd_memory.
//
// Layouts in HBM (row-major, K = fitting index, naux_ of them):
//   Xai (K,a,i)  Xij (K,i,j)  Xab (K,a,b)           as uploaded
//   XaiP (a,K,i) XabP (a,K,b) Xia (K,i,a)           built once in prepare_df()
//   Xt (K,a,i) = X_ai_tau, XtP (a,K,i)              per residual
//   Gblk (a,c,b,d) -> Gp (a,b,c,d) for a block of a  the lazy ladder
#include <stdexcept>

#include "ccsd_ctx.hpp"

namespace mpqc_hip {

void CCSDContext::set_df_basis(int naux, const int64_t* aux_hm, int n_aux_tiles) {
  if (naux <= 0) throw std::invalid_argument("naux must be positive");
  if (naux_ != 0) {
    if (naux != naux_) throw std::logic_error("the fitting basis size is fixed once set");
  } else {
    const long Q = naux, O = o_, V = v_;
    if (Q * V * V >= (1L << 31) || Q * V * O >= (1L << 31))
      throw std::invalid_argument("naux * v^2 must be < 2^31");
    naux_ = naux;
    const long ov = O * V;
    Xai_ = alloc(Q * ov); Xij_ = alloc(Q * O * O); Xab_ = alloc(Q * V * V);
    XaiP_ = alloc(Q * ov); XabP_ = alloc(Q * V * V); Xia_ = alloc(Q * ov);
    Tbar_ = alloc(ov * ov); Xt_ = alloc(Q * ov); XtP_ = alloc(Q * ov);
    Xabt1_ = alloc(Q * ov); Xabt1P_ = alloc(Q * ov); Zdf_ = alloc(Q * ov); Udf_ = alloc(Q * ov);
    sK_ = alloc(Q); V1_ = alloc(Q * O * O); Ydf_ = alloc(Q * V * V);
    Rdf_ = alloc(ov * ov); bdf_ = alloc(ov * ov);
  }
  if (aux_hm == nullptr) {
    aux_hm_ = {0, naux};
  } else {
    if (n_aux_tiles <= 0) throw std::invalid_argument("tile count must be positive");
    aux_hm_.assign(aux_hm, aux_hm + n_aux_tiles + 1);
    if (aux_hm_.front() != 0 || aux_hm_.back() != naux)
      throw std::invalid_argument("hashmarks must span [0, naux]");
    for (int t = 0; t < n_aux_tiles; ++t)
      if (aux_hm_[t + 1] <= aux_hm_[t]) throw std::invalid_argument("hashmarks must increase");
  }
  df_ = true;
  direct_ = false;
  prepared_ = false;
}

// Layouts of the three-index factors used as GEMM operands.
void CCSDContext::prepare_df() {
  const long Q = naux_, O = o_, V = v_;
  // XaiP(a,K,i) = Xai(K,a,i);  XabP(a,K,b) = Xab(K,a,b);  Xia(K,i,a) = Xai(K,a,i)
  perm_add(XaiP_, {1, V, Q, O}, 0.0, {{Xai_, {1, Q, V, O}, {0, 2, 1, 3}, 1.0}});
  perm_add(XabP_, {1, V, Q, V}, 0.0, {{Xab_, {1, Q, V, V}, {0, 2, 1, 3}, 1.0}});
  perm_add(Xia_, {1, Q, O, V}, 0.0, {{Xai_, {1, Q, V, O}, {0, 1, 3, 2}, 1.0}});
}

// ccsd_r1_r2.h:147-148  X_ai_tau(K,a,i) = 2 Xai(K,b,j) tau(a,b,i,j) - Xai(K,b,j) tau(a,b,j,i)
// as Xai(K,(b,j)) . Tbar((a,i),(b,j))^T with Tbar(a,i,b,j) = 2 tau(a,b,i,j) - tau(a,b,j,i)
void CCSDContext::df_r1_intermediates() {
  const long Q = naux_, O = o_, V = v_, ov = O * V;
  const std::array<long, 4> e_abij{V, V, O, O};
  perm_add(Tbar_, {V, O, V, O}, 0.0,
           {{tau_, e_abij, {0, 2, 1, 3}, 2.0}, {tau_, e_abij, {0, 3, 1, 2}, -1.0}});
  gemm(false, true, (int)Q, (int)ov, (int)ov, 1.0, Xai_, ov, Tbar_, ov, 0.0, Xt_, ov, 0, 0, 0, 1,
       "r1:df_X_ai_tau");
  perm_add(XtP_, {1, V, Q, O}, 0.0, {{Xt_, {1, Q, V, O}, {0, 2, 1, 3}, 1.0}});
}

void CCSDContext::ensure_df_blocks() {
  if (Gblk_) return;
  const long V = v_;
  const long v3 = V * V * V;
  long nb = (long)(df_block_bytes_ / (2.0 * sizeof(double) * (double)v3));
  nb = std::max(1L, std::min(nb, V));
  nb = std::min(nb, std::max(1L, ((1L << 31) - 1) / v3));
  if (nb * v3 >= (1L << 31)) throw std::invalid_argument("v^3 must be < 2^31 for the df ladder");
  df_block_ = (int)nb;
  Gblk_ = alloc(nb * v3);
  Gp_ = alloc(nb * v3);
}

// ccsd_r1_r2.h:497-508 (reduced_abcd_memory):
//   Y(K,a,b) = 0.5 Xab(K,a,b) - Xai(K,b,i) t1(a,i)
//   g(a,b,c,d) = Xab(K,a,c) Y(K,b,d)        (df_direct_integrals, Physical notation)
//   b(a,b,i,j) = tau(c,d,i,j) g(a,b,c,d);  b += b(b,a,j,i);  r2 += b
// g is built block by block over a (df_block_ rows), so only two blocks of
// nb v^3 values are ever stored: the GEMM writes (a,c,b,d), a permutation makes
// it (a,b,c,d), and the ladder GEMM consumes it at once.
void CCSDContext::df_pp_ladder() {
  const long Q = naux_, O = o_, V = v_, ov = O * V, oo = O * O, vv = V * V;
  const std::array<long, 4> e_abij{V, V, O, O};
  const int v = v_, o = o_;
  perm_add(Ydf_, {1, Q, V, V}, 0.0, {{Xab_, {1, Q, V, V}, {0, 1, 2, 3}, 0.5}});
  gemm(false, true, v, v, o, -1.0, t1_, o, Xai_, o, 1.0, Ydf_, v, 0, ov, vv, (int)Q, "r2:df_pp_Y");
  ensure_df_blocks();
  const long nbmax = df_block_;
  const bool sm = split_mode();
  if (sm) ensure_planes(Gp_P_, nbmax * V * vv);
  for (long a0 = 0; a0 < V; a0 += nbmax) {
    const long nb = std::min(nbmax, V - a0);
    gemm(true, false, (int)(nb * V), (int)vv, (int)Q, 1.0, Xab_ + a0 * V, vv, Ydf_, vv, 0.0, Gblk_, vv,
         0, 0, 0, 1, "r2:df_pp_abcd");
    if (sm) {
      // the block's (a,b,c,d) layout straight into planes, contracted with the tau^T planes
      split(Gp_P_, {nb, V, V, V}, {{Gblk_, {nb, V, V, V}, {0, 2, 1, 3}, 1.0}});
      gemmP((int)(nb * V), (int)oo, (int)vv, 1.0, Gp_P_, vv, tauT_P_, vv, 0.0, bdf_ + a0 * V * oo, oo,
            "r2:df_pp_ladder");
    } else {
      perm_add(Gp_, {nb, V, V, V}, 0.0, {{Gblk_, {nb, V, V, V}, {0, 2, 1, 3}, 1.0}});
      gemm(false, false, (int)(nb * V), (int)oo, (int)vv, 1.0, Gp_, vv, tau_, oo, 0.0, bdf_ + a0 * V * oo, oo,
           0, 0, 0, 1, "r2:df_pp_ladder");
    }
  }
  perm_add(r2_, e_abij, 1.0, {{bdf_, e_abij, {0, 1, 2, 3}, 1.0}, {bdf_, e_abij, {1, 0, 3, 2}, 1.0}});
}

}  // namespace mpqc_hip
