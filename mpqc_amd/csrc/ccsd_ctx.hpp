// Device-resident CCSD solver context: integrals, amplitudes, residuals,
// intermediates and DIIS history all live in HBM for the whole solve.
#pragma once

#include <array>
#include <deque>
#include <string>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "planes.hpp"
#include "tensor_ops.hpp"

namespace mpqc_hip {

struct DeviceBuffer {
  double* p = nullptr;
  size_t n = 0;  // elements
};

struct TraceRow {
  int iter;
  double dE, error, energy, seconds;
};

// Column-pivoted Householder QR solve of a small dense system (row-major A);
// returns |det A|, x gets the rank-truncated solution.
double colpiv_qr_solve(int n, std::vector<double> A, std::vector<double> b, std::vector<double>& x);

class CCSDContext {
 public:
  // with_solver = false skips the DIIS history (contexts that only evaluate residuals)
  CCSDContext(int o, int v, const int64_t* occ_hm, int n_occ_tiles, const int64_t* vir_hm,
              int n_vir_tiles, int precision, bool with_solver = true);
  ~CCSDContext();

  int o() const { return o_; }
  int v() const { return v_; }

  void set_option(const std::string& key, double value);
  void set_integral(int which, const double* data, int layout, int memspace);
  void get_integral(int which, double* out, int layout, int memspace);
  // density fitting: naux fitting functions tiled by aux_hm (nullptr: one tile);
  // selects the "df" residual (ccsd.h:93-99: method defaults to df when a
  // fitting basis is given)
  void set_df_basis(int naux, const int64_t* aux_hm, int n_aux_tiles);
  bool df() const { return df_; }
  // AO -> MO on the device (LCAOFactory::compute4, lcao_factory.h:378-497): the
  // integral blocks and Fock blocks from AO integrals (mu nu|la si) (nbf^4,
  // chemist notation), the AO Fock matrix and the active occupied / virtual MO
  // coefficients (nbf x o, nbf x v, row-major)
  void transform_ao(int nbf, const double* eri, const double* fock, const double* c_occ, const double* c_vir,
                    int memspace);
  // the density-fitted set: three-centre (K|mu nu) (naux x nbf^2), a metric factor
  // L (naux x naux, L L^T = V^-1; X = L^T B) or nullptr when B is already fitted
  void transform_df(int nbf, const double* B3, const double* metric, const double* fock, const double* c_occ,
                    const double* c_vir, int memspace);
  // AO integral-direct method ("direct", ccsd.h:73, 246-247): the AO integrals
  // (mu nu|la si) (nbf^4, chemist notation) and the active occupied / virtual MO
  // coefficients (nbf x o, nbf x v, row-major) stay resident; every residual forms
  // U(p,r,i,j) = (pq|rs) [Ca tau Ca^T](q,s,i,j) (compute_u2_u11, ccsd.h:573-603) in
  // place of the (ab|cd) ladder and the Giabc tau terms.  Selects the method.
  void set_direct_ao(int nbf, const double* ao_eri, const double* c_occ, const double* c_vir, int memspace);
  bool direct() const { return direct_; }
  void set_orbital_energies(const double* eo, const double* ev);
  void solver_init(int diis_start, int n_diis, double damp, int ngroup, int ngroup_nstart);
  void set_amplitudes(const double* t1, const double* t2, int layout, int memspace);
  void get_amplitudes(double* t1, double* t2, int layout, int memspace);
  void get_residuals(double* r1, double* r2, int layout, int memspace);
  void get_dressed_fock(double* FIJ, double* FAB);

  void init_guess();
  void residual();
  void energy_error(double* E, double* err);
  void update();
  double solve(int max_iter, double target_precision, int verbose, int* n_iter);
  const std::vector<TraceRow>& trace() const { return trace_; }
  // (T) correction from the current (converged) amplitudes (CCSD_T::compute_ccsd_t,
  // ccsd_t.h:144-176, "straight" algorithm :1127-1167); seconds (optional) gets the
  // device time
  double triples(double block_bytes, double* seconds);
  double triples_block_bytes_ = 8.0 * (1 << 30);
  // EOM-EE-CCSD singlet excitation energies (EOM_CCSD, eom_ccsd.h:17 with
  // eom_ccsd_impl.h:236-390): the lowest n_roots eigenvalues of the CCSD
  // Jacobian on (ab)(ij)-symmetric doubles at the current (converged)
  // amplitudes, by Davidson iterations on its transpose; returns the number of
  // iterations
  int eom_ee(int n_roots, int max_iter, double convergence, int verbose, double* energies);
  int eom_max_vector_ = 8;             // subspace vectors per root (eom_ccsd.h:33)
  double eom_vector_threshold_ = 0.0;  // new-vector norm threshold; 0: 10 * convergence (:34)
  int eom_n_guess_ = 0;                // initial singles guesses; 0: 2 n_roots
  int eom_extra_roots_ = 0;            // roots tracked above the requested ones
  bool eom_skip_converged_ = true;     // no new vectors for converged roots
  bool eom_cis_diagonal_ = false;      // singles preconditioned with the CIS singlet diagonal
  bool eom_cis_guess_ = true;          // initial vectors from CIS (else unit singles)
  bool pp_symmetric_ = true;           // split forms: pp ladder on (ab)/(cd) pair-symmetric blocks
  int eom_products_ = 0;               // J^T products of the last eom_ee
  double eom_product_seconds_ = 0.0;   // their device time (HIP events on the context stream)

  // ---- adjoint of the converged energy (no counterpart in the reference).
  // With the amplitudes t converged (R(t) = 0), lambda solves (dR/dt)^T lambda = -dE/dt
  // and dE/d(theta) = dE/d(theta)|_t + lambda^T dR/d(theta) for every integral block.
  // Energy functional: E = 2 <f_ai, t1> + <g_abij, 2 tau(a,b,i,j) - tau(b,a,i,j)>.
  // Jacobian-transpose product at the current amplitudes (device buffers, dense layout).
  void jacobian_transpose(const double* y1, const double* y2, double* x1, double* x2);
  // the same with host or device buffers in dense or tile layout
  void jacobian_transpose_io(const double* y1, const double* y2, double* x1, double* x2, int layout,
                             int memspace);
  int lambda_solve(int max_iter, double target_precision, int verbose);
  void energy_gradient();
  void get_lambda(double* l1, double* l2, int layout, int memspace);
  void get_integral_gradient(int which, double* out, int layout, int memspace);

  // Per-contraction timing with HIP events on the context stream (enabled by
  // the "kernel_timing" option).  Tags name the reference expression.
  struct KernelTime {
    std::string tag;
    long calls = 0;
    double ms = 0.0;
    double flops = 0.0;  // 2*M*N*K*batch summed over calls
  };
  std::vector<KernelTime> kernel_times();
  void reset_kernel_times();

  std::string last_error;

 private:
  // tensor upload / download in dense or TA tile layout; `modes` gives for each
  // of the 4 modes whether it is occupied (0), virtual (1) or absent (-1).
  void upload(double* dst, const double* src, const std::array<int, 4>& modes, int layout,
              int memspace);
  void download(double* dst, const double* src, const std::array<int, 4>& modes, int layout,
                int memspace);
  double* alloc(size_t n);
  // Host-side guard: throws unless [p, p + max_off] lies inside one of this
  // context's device allocations.  Every GEMM and permutation the residual
  // issues is checked before launch.
  void check_access(const double* p, long max_off, const char* what) const;
  void construct(const int64_t* occ_hm, int n_occ_tiles, const int64_t* vir_hm, int n_vir_tiles,
                 int precision, bool with_solver);
  void release();
  void prepare();
  void gemm(bool tA, bool tB, int M, int N, int K, double alpha, const double* A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc, long sA = 0, long sB = 0,
            long sC = 0, int batch = 1, const char* tag = nullptr);
  struct Src {
    const double* p;
    std::array<long, 4> ext;
    std::array<int, 4> perm;
    double alpha;
  };
  // untagged calls are timed under "perm@<source line of the call>"
  void perm_add(double* out, std::array<long, 4> n, double beta, std::initializer_list<Src> srcs,
                const char* tag = nullptr, int line = __builtin_LINE());
  // validated terms of a permuted linear combination
  int make_terms(std::array<long, 4> n, std::initializer_list<Src> srcs, TTerm* terms);
  void copy_d2d(double* dst, const double* src, long n);
  void op_tau();  // tau = t2 + t1 t1  (ccsd.h:389)
  // ---- density-fitted residual pieces (ccsd_r1_r2.h:136-184, 309-520)
  void prepare_df();                    // layouts of the three-index factors
  void ensure_iabc_for_triples();       // <ia|bc> from the factors when the df set lacks it
  void transform_fock(int nbf, const double* fock, const double* Cact, int memspace);
  void df_r1_intermediates();           // X_ai_tau and its (a,K,l) layout
  void df_pp_ladder();                  // lazy (ab|cd) ladder of reduced_abcd_memory
  void ensure_df_blocks();
  void op_TP();   // the T intermediate of ccsd_r1_r2.h:236
  // ---- AO integral-direct residual pieces
  void direct_u();   // U at the current tau (ccsd.h:581-596)
  void direct_r1();  // ccsd_r1_r2.h:116-117
  void direct_r2();  // ccsd_r1_r2.h:282-286
  void pack_pp_symmetric();     // packed V+ / V- (FP64) of the pair-symmetric ladder
  void prepare_pp_symmetric();  // and their planes (split forms)
  void pack_tau_pairs();        // S, A from tau for the pair-symmetric products
  void pair_products(const double* Pk, long lp, const Planes& Pp, const double* Mk, long lm, const Planes& Mp,
                     long nrs, long nra, long n, double* out, double* P, double* M, const char* tag_s,
                     const char* tag_a, const char* tag_u);
  void pp_ladder_symmetric(double* out);  // ccsd_r1_r2.h:275 on pair-symmetric blocks, added to out
  void w_oooo_symmetric();      // ccsd_intermediates.h:182-193, its Gijab tau term likewise

  // ---- op tape: integral preparation and the residual as a list of primitive
  // ops (copies, permuted linear combinations, contractions, tau/T builds).
  // Replayed in reverse by the adjoint; grouped across contexts by the batch.
  struct TapeOp {
    enum Kind : int { kCopy = 0, kTAdd = 1, kGemm = 2, kTau = 3, kTP = 4 };
    Kind kind = kCopy;
    double* dst = nullptr;         // kCopy / kTau / kTP output, kTAdd output
    const double* src = nullptr;   // kCopy source; kTau / kTP: t2
    const double* src2 = nullptr;  // kTau / kTP: t1
    long n = 0;                    // kCopy element count
    std::array<long, 4> ext{};     // kTAdd output extents
    double beta = 0.0;             // kTAdd
    int nterms = 0;                // kTAdd
    Src terms[kMaxTerms];          // kTAdd
    GemmDesc g;                    // kGemm
    GemmPrecision prec = GemmPrecision::F64;
    std::string tag;
  };
  // adjoint state
  std::vector<double*> adj_;      // adjoint buffer of allocation k (nullptr until used)
  std::vector<TapeOp> res_tape_;  // residual recorded at the amplitudes of the last sweep
  double *lam1_ = nullptr, *lam2_ = nullptr, *gt1_ = nullptr, *gt2_ = nullptr;
  double* adj_ws_ = nullptr;      // partial sums of shared-operand contraction adjoints
  size_t adj_ws_n_ = 0;
  double* adj_scratch_ = nullptr;  // permuted copies inside tau / TP adjoints
  long adj_scratch_n_ = 0;
  std::vector<double*> lam_hist_;  // Lambda DIIS history (owned allocations)
  double* adj_scratch(long n);
  double* adj_hist(int k, long n);
  bool lambda_ready_ = false, gradient_ready_ = false;
  bool lambda_symmetric_ = true;  // solve in the (ab)(ij)-symmetric t2 subspace
  int alloc_index(const double* p) const;
  double* adj(const double* p);
  bool is_integral(const double* p) const;
  void zero_adjoints();
  void record_residual();
  void check_tape(const std::vector<TapeOp>& tape) const;
  void reverse_sweep(const std::vector<TapeOp>& tape, bool integrals);
  void reverse_op(const TapeOp& op, bool integrals);
  void energy_t_gradient(double* g1, double* g2);
  // x = P J^T y for contiguous (t1, t2) vectors, P the projection onto
  // t2(a,b,i,j) = t2(b,a,j,i); the residual tape must be recorded
  void eom_apply_jt(const double* y, double* x);
  void ensure_abcd_from_factors();  // Gabcd from the three-index factors
  std::vector<TapeOp>* tape_ = nullptr;  // ops are appended here when set
  bool tape_execute_ = true;             // false: record without launching
  std::vector<TapeOp> prepare_tape_;     // recorded by prepare()
  friend class CCSDBatch;
  friend struct Adjoint;

  // ---- mixed precision: 16-bit planes of the operands of the large contractions
  bool split_mode() const;
  // The op tape (adjoint, EOM) records and sweeps FP64 contractions in the
  // split forms: the 16-bit plane contractions are not on the tape.  The
  // guard switches the contraction precision for one public call.
  struct TapePrecision {
    CCSDContext& c;
    GemmPrecision saved;
    explicit TapePrecision(CCSDContext& x) : c(x), saved(x.prec_) {
      if (x.split_mode()) x.prec_ = GemmPrecision::F64;
    }
    ~TapePrecision() { c.prec_ = saved; }
    TapePrecision(const TapePrecision&) = delete;
    TapePrecision& operator=(const TapePrecision&) = delete;
  };
  PlaneKind plane_kind() const;
  Planes alloc_planes(long n);
  void ensure_planes(Planes& P, long n);  // allocates P (n values per plane) unless it already is
  void release_planes();
  void split(const Planes& out, std::array<long, 4> n, std::initializer_list<Src> srcs,
             const TProd* prod = nullptr, const char* tag = nullptr, int line = __builtin_LINE());
  TProd tp_product() const;  // t1(d,i) t1(a,l) over TP's (a,i,l,d) index order
  AbsMaxCache amax_;         // max |t1|, |t2|, |tau| shared by one residual's splits
  // (amax_slot >= 0: the result is the whole tensor amax_.ptr[amax_slot]; its max |C| is
  // recorded in that slot of the cache where the contraction's split-K pass can produce it)
  void gemmP(int M, int N, int K, double alpha, const Planes& A, long lda, const Planes& B, long ldb,
             double beta, double* C, long ldc, const char* tag, int amax_slot = -1);
  void prepare_planes();
  bool planes_ready_ = false;
  double* pws_ = nullptr;  // split-K workspace of the plane GEMMs
  size_t pws_bytes_ = 0;
  std::vector<std::pair<uint16_t*, size_t>> plane_allocs_;
  Planes Gabcd_P_, Giabc_P_, Gijab_P_, GAt_P_, GBt_P_, GCt_P_;
  // pair-symmetric pp ladder: pair counts, padded row pitches, pair tables, packed operands
  long nvs_ = 0, nva_ = 0, nos_ = 0, noa_ = 0, ldvs_ = 0, ldva_ = 0;
  bool pp_packed_ = false;  // V+ / V- packed from the current Gabcd
  bool pp_checked_ = false, pp_sym_ok_ = false;  // Gabcd's pair symmetry checked / holds
  bool pp_sym_planes_ = false;                   // the split forms' planes are the packed ones
  bool pp_symmetric_usable();
  int2 *pv_sym_ = nullptr, *pv_anti_ = nullptr, *po_sym_ = nullptr, *po_anti_ = nullptr;
  double *Gpk_ = nullptr, *Gmk_ = nullptr, *WPk_ = nullptr, *WMk_ = nullptr;  // W_oooo's G+/- and products
  Planes Gp_P_sym_, Gm_P_sym_;
  bool w_sym_ok_ = false;  // Gijab(l,k,d,c) = Gijab(k,l,c,d) holds (checked with Gabcd's symmetry)
  bool iabc_checked_ = false, iabc_sym_ok_ = false;  // Giabc(i,c,a,b) = Giabc(i,b,a,c) checked / holds
  bool giabc_symmetric_usable();
  double *Vpk_ = nullptr, *Vmk_ = nullptr, *Spk_ = nullptr, *Apk_ = nullptr, *Ppk_ = nullptr, *Mpk_ = nullptr;
  Planes Vp_P_, Vm_P_, Sp_P_, Ap_P_;
  Planes GbarT_P_, Gbar_P_, Gbar_akcd_P_, P8T_P_, GiabcT_P_, P13_P_, P14_P_;
  Planes tau_P_, tauT_P_, TP_P_, t2P_P_, Xjk_P_, K_P_, Y1t_P_, Y2t_P_, WT_P_, t1T_P_, gac_P_;
  Planes tau_ikcd_P_, tau_ackl_P_;
  Planes Gp_P_;  // one (ab|cd) block of the df ladder
  void check_finite(const double* x, long n, const char* what);
  double scalar_from_device(const double* d);

  int o_, v_;
  GemmPrecision prec_;
  double fov_sign_ = 1.0;
  hipStream_t stream_ = nullptr;
  std::vector<int64_t> occ_hm_, vir_hm_;
  std::vector<double*> allocations_;
  std::vector<size_t> alloc_sizes_;

  // density fitting (method "df"): three-index factors X(K,p,q) and their layouts
  bool df_ = false, reduced_abcd_ = true;
  int naux_ = 0;
  std::vector<int64_t> aux_hm_;
  double *Xai_ = nullptr, *Xij_ = nullptr, *Xab_ = nullptr;
  double *XaiP_ = nullptr, *XabP_ = nullptr, *Xia_ = nullptr;
  double *Tbar_ = nullptr, *Xt_ = nullptr, *XtP_ = nullptr, *Xabt1_ = nullptr, *Xabt1P_ = nullptr,
         *Zdf_ = nullptr, *Udf_ = nullptr, *sK_ = nullptr, *V1_ = nullptr, *Ydf_ = nullptr,
         *Rdf_ = nullptr, *bdf_ = nullptr, *Gblk_ = nullptr, *Gp_ = nullptr;
  // AO integral-direct method: X((p,r),(q,s)) = (pq|rs), coefficients, U and its transforms
  bool direct_ = false;
  int nbf_ = 0;
  double *Xao_ = nullptr, *Ci_ = nullptr, *Ca_ = nullptr, *Uao_ = nullptr, *TauAO_ = nullptr, *T1ao_ = nullptr,
         *Vr_ = nullptr, *Yr_ = nullptr, *Dct_ = nullptr, *CaD_ = nullptr, *G1ao_ = nullptr, *G2ao_ = nullptr;
  int df_block_ = 0;            // virtual rows a per (ab|cd) block
  double df_block_bytes_ = 4.0 * (1 << 30);  // budget of the two (ab|cd) block buffers

  // integrals
  double *Fia_ = nullptr, *Fij_ = nullptr, *Fab_ = nullptr, *Gijab_ = nullptr, *Gijkl_ = nullptr,
         *Giajb_ = nullptr, *Giabc_ = nullptr, *Gijka_ = nullptr, *Gabcd_ = nullptr;
  unsigned have_ints_ = 0;
  bool prepared_ = false;
  // derived integral layouts (built once in prepare())
  double *f_ai_, *g_abij_, *Gbar_, *Gbar_cKLD_, *Gbar_kcld_, *P6_, *Gbar_akcd_, *P8_, *S11_,
      *P13_, *P14_, *J0_, *K0_, *X18_, *X19_, *GA_, *GB_, *GC_;
  double* GiabcT_ = nullptr;  // Giabc(k,a,d,c) as (k,a,c,d), the FP64 forms' k term (allocated by prepare())
  // orbital energies
  double *eo_ = nullptr, *ev_ = nullptr;
  bool have_eps_ = false;
  // amplitudes / residuals
  double *t1_, *t2_, *tau_, *r1_, *r2_, *FIJ_, *FAB_;
  // per-iteration scratch
  double *Xoo_, *Zoo_, *hkc_, *hT_, *t1T_, *tau_akld_, *tau_ilab_, *Y5_, *tau_kcdi_, *tau_ackl_,
      *Rd_, *Q_, *M12_, *S_, *gki_, *gac_, *t2_abjk_, *X16_, *TP_, *t2P_, *J_, *K_, *Zj_, *Zk_,
      *Xjk_, *Y1_, *Y2_, *RR_, *tmpR_, *Y23_, *W_, *tmp24_;
  GemmWorkspace ws_;
  double* partials_ = nullptr;
  double* scalars_ = nullptr;  // device scalars
  long* modeinfo_ = nullptr;   // device tile maps (scratch, sized per call)
  size_t modeinfo_n_ = 0;
  double* staging_ = nullptr;  // device staging for tiled / host uploads
  size_t staging_n_ = 0;

  // DIIS (TA::DIIS semantics, see DIISState in ccsd_ctx.hip)
  int diis_start_ = 1, n_diis_ = 8, diis_ngroup_ = 1, diis_group_nstart_ = 1;
  int solve_restart_ = 1;  // solve() from the initial guess (0, 2: from the current amplitudes)
  double diis_damp_ = 0.0;
  int diis_iter_ = 0;
  std::deque<int> err_slots_, x_slots_;
  std::vector<double> B_;  // n_diis x n_diis, row-major (only top-left nvec used)
  std::vector<double*> hist_e1_, hist_e2_, hist_x1_, hist_x2_;
  bool solver_ready_ = false;
  std::vector<TraceRow> trace_;

  bool kernel_timing_ = false;
  // residual replayed from a HIP graph (option "graph"): captured on the second
  // residual() of a configuration, replayed until an option, the integrals or
  // the method change (gen_ counts those changes)
  int graph_mode_ = 1;
  bool tau_fresh_ = false;  // tau_ holds t2 + t1 t1 of the current amplitudes
  bool tau_amax_fresh_ = false;  // amax_.dev[1], [2] hold max |t2|, |tau| of the current amplitudes
  long gen_ = 0, res_seen_gen_ = -1, res_graph_gen_ = -1, res_failed_gen_ = -1;
  hipGraph_t res_graph_ = nullptr;
  hipGraphExec_t res_exec_ = nullptr;
  bool graph_eligible() const;
  void residual_ops();
  bool capture_residual();
  void drop_residual_graph();
  struct PendingEvent {
    int slot;
    hipEvent_t a, b;
    double flops;
  };
  std::vector<PendingEvent> pending_;
  std::vector<KernelTime> ktimes_;
  std::vector<hipEvent_t> event_pool_;
  void resolve_events();
  int timing_slot(const char* tag);
  hipEvent_t take_event();
  // runs f() between two HIP events on the context stream when kernel timing is on
  template <class F>
  void timed(const char* tag, double flops, F&& f) {
    if (!kernel_timing_) {
      f();
      return;
    }
    const int slot = timing_slot(tag);
    hipEvent_t a = take_event(), b = take_event();
    MPQC_HIP_CHECK(hipEventRecord(a, stream_));
    f();
    MPQC_HIP_CHECK(hipEventRecord(b, stream_));
    pending_.push_back({slot, a, b, flops});
  }
};

}  // namespace mpqc_hip
