/*
 * mpqc_ccsd.h -- C-ABI of the MI355X-native closed-shell CCSD residual library.
 *
 * Drop-in boundary for MPQC's CCSD solver loop.  Each entry point replaces one
 * reference interface (paths relative to the MPQC source tree, src/mpqc/...):
 *
 *   mpqc_ccsd_ctx_create          CCSD<Tile,Policy>::compute_ccsd set-up
 *                                 (chemistry/qc/lcao/cc/ccsd.h:268-342) and the
 *                                 TiledRange1 of the amplitudes
 *                                 (chemistry/qc/lcao/expression/trange1_engine.cpp:10-20)
 *   mpqc_ccsd_set_integral        cc::Integrals<Array> members Fia..Gabcd
 *                                 (chemistry/qc/lcao/cc/ccsd_r1_r2.h:24-56), filled at
 *                                 ccsd.h:281-305
 *   mpqc_ccsd_set_direct_ao       the AO integrals and MO coefficients of the "direct" method
 *                                 (ccsd.h:73, 246-247, 313-322; compute_u2_u11 :573-603)
 *   mpqc_ccsd_set_df_basis        the density-fitting basis of the "df" method: its size
 *                                 and the TiledRange1 of the fitting index of
 *                                 (K|G|p q)[inv_sqr] (ccsd.h:93-99, 481-493); selects
 *                                 compute_cs_ccsd_r1_df / compute_cs_ccsd_r2_df
 *                                 (ccsd_r1_r2.h:136-184, 309-520)
 *   mpqc_ccsd_triples             the (T) correction CCSD_T::compute_ccsd_t
 *                                 (chemistry/qc/lcao/cc/ccsd_t.h:144-176; algorithm of
 *                                 compute_ccsd_t_straight :1127-1167 and CCSD_T_Reduce
 *                                 :2274-2335) from the context's amplitudes
 *   mpqc_ccsd_eom_ee              EOM-EE-CCSD singlet excitation energies EOM_CCSD::evaluate
 *                                 (chemistry/qc/lcao/cc/eom/eom_ccsd.h:17, eom_ccsd_impl.h:395-478;
 *                                 Davidson of eom_ccsd_davidson_solver :236-390 with
 *                                 math/linalg/davidson_diag.h:128-425 and EEPred,
 *                                 eom/eom_preconditioner.h:18-74) from the context's amplitudes
 *   mpqc_ccsd_get_integral        reads a cc::Integrals member back (ccsd_r1_r2.h:24-56)
 *   mpqc_ccsd_transform_ao        LCAOFactory::compute4 without density fitting
 *                                 (chemistry/qc/lcao/factory/lcao_factory.h:409-493) for the
 *                                 nine blocks of ccsd.h:281-305, and <p|F|q> for Fia/Fij/Fab
 *   mpqc_ccsd_transform_df        the "[df]" blocks: (K|G|p q)[inv_sqr] (ccsd.h:481-493) and
 *                                 compute4 with density fitting (lcao_factory.h:386-406)
 *   mpqc_ccsd_set_orbital_energies JacobiDIISSolver(kv, f_ii, f_aa)
 *                                 (chemistry/qc/lcao/cc/solvers.h:1140-1146; values from
 *                                 make_diagonal_fpq, chemistry/qc/lcao/scf/mo_build.h:20-26)
 *   mpqc_ccsd_solver_init         DIISSolver(const KeyVal&) keywords diis_start, n_diis,
 *                                 diis_damp, diis_ngroup, diis_group_nstart
 *                                 (chemistry/qc/cc/solvers.h:93-95)
 *   mpqc_ccsd_residual            cc::compute_cs_ccsd_r1 then cc::compute_cs_ccsd_r2
 *                                 (ccsd_r1_r2.h:70-124, 198-293), incl. the FIJ/FAB
 *                                 hand-off (ccsd_r1_r2.h:90,96) and the tau rebuild
 *                                 (ccsd.h:389)
 *   mpqc_ccsd_energy_error        Solver::error (chemistry/qc/cc/solvers.h:60-63) and the
 *                                 energy expression at ccsd.h:376-380
 *   mpqc_ccsd_update              DIISSolver::update = JacobiDIISSolver::update_only
 *                                 (lcao/cc/solvers.h:1153-1158) + TA::DIIS::extrapolate
 *                                 (chemistry/qc/cc/solvers.h:112-121), then tau (ccsd.h:389)
 *   mpqc_ccsd_solve               the whole solver loop CCSD::compute_ccsd (ccsd.h:344-477)
 *   mpqc_ccsd_get_trace           the iteration table printed by detail::print_ccsd
 *                                 (ccsd.h:28-36)
 *   mpqc_ccsd_set/get_amplitudes  CCSD::T1_/T2_ (ccsd.h:121-122, 252-253)
 *   mpqc_ccsd_pp_ladder           the particle-particle ladder contraction alone,
 *                                 b("a,b,i,j") = tau("c,d,i,j") * Gabcd("a,b,c,d")
 *                                 (ccsd_r1_r2.h:275)
 *   mpqc_ccsd_gemm                one TiledArray contraction "C = alpha A*B + beta C"
 *                                 (math/external/tiledarray/tiledarray_impl.cpp:36-47,
 *                                 TA's ContractReduce over TA::Tensor<double>)
 *
 * Adjoint entry points (no counterpart in the reference, which has no Lambda
 * equations and no CCSD gradients): mpqc_ccsd_jacobian_transpose,
 * mpqc_ccsd_lambda_solve, mpqc_ccsd_energy_gradient, mpqc_ccsd_get_lambda,
 * mpqc_ccsd_get_integral_gradient.  They differentiate the converged energy of
 * ccsd.h:376-380 (at R = 0: E = 2 <f_ai, t1> + <g_abij, 2 tau - tau(b,a,i,j)>)
 * through the residual of ccsd_r1_r2.h:70-293.
 *
 * Conventions: row-major FP64 buffers in the reference's index orders: t1 "a,i",
 * t2/r2 "a,b,i,j", Fia "i,a", Fij "i,j", Fab "a,b", Gijab "i,j,a,b" = <ij|ab>,
 * Gijkl "i,j,k,l", Giajb "i,a,j,b", Giabc "i,a,b,c", Gijka "i,j,k,a",
 * Gabcd "a,b,c,d" (physicist notation).  MPQC_LAYOUT_TA_TILES means the buffer is
 * the concatenation, in row-major tile-ordinal order, of row-major tiles over the
 * TiledRange1 hashmarks given at create time (occupied modes use the occupied
 * hashmarks, virtual modes the virtual ones).  MPQC_MEM_HOST / MPQC_MEM_DEVICE
 * says where a buffer lives; a device buffer is read or written only after the
 * work already queued on the device (any stream) has finished, as with hipMemcpy,
 * and the call returns once the copy is complete.  Every function returns an MPQC_CCSD_* status; on
 * failure mpqc_ccsd_last_error() describes it.  Calls on one context must be
 * serialised by the caller (one HIP stream per context).
 */
#ifndef MPQC_CCSD_H
#define MPQC_CCSD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpqc_ccsd_ctx mpqc_ccsd_ctx;

enum {
  MPQC_CCSD_OK = 0,
  MPQC_CCSD_ERR_ARG = 1,       /* invalid argument (InputError in the reference) */
  MPQC_CCSD_ERR_HIP = 2,       /* HIP runtime error */
  MPQC_CCSD_ERR_STATE = 3,     /* call out of order (ProgrammingError in the reference) */
  MPQC_CCSD_ERR_NONFINITE = 4, /* non-finite residual/amplitudes (SIGFPE in the reference) */
  MPQC_CCSD_ERR_DIIS = 5,      /* DIIS system poorly conditioned (TA::DIIS domain_error) */
  MPQC_CCSD_ERR_NOMEM = 6,     /* device allocation failed */
  MPQC_CCSD_ERR_NOCONV = 7     /* an iterative solve did not reach its precision */
};

/* Contraction precisions (the Jacobi update, DIIS and energy are always FP64):
 * F64    FP64 MFMA
 * F32    FP32 MFMA on FP32-rounded operands, FP64 accumulation across K chunks
 * BF16X3 bf16 hi+lo split, 3 products on bf16 MFMA (~16 significant bits)
 * F16X3  fp16 hi+lo split of power-of-two scaled operands, 3 products (~22 bits)
 * BF16X6 bf16 hi+mid+lo split, 6 products (~24 bits) */
enum {
  MPQC_CCSD_PREC_F64 = 0,
  MPQC_CCSD_PREC_F32 = 1,
  MPQC_CCSD_PREC_BF16X3 = 2,
  MPQC_CCSD_PREC_F16X3 = 3,
  MPQC_CCSD_PREC_BF16X6 = 4
};
enum { MPQC_LAYOUT_DENSE = 0, MPQC_LAYOUT_TA_TILES = 1 };
enum { MPQC_MEM_HOST = 0, MPQC_MEM_DEVICE = 1 };
enum {
  MPQC_INT_FIA = 0,
  MPQC_INT_FIJ = 1,
  MPQC_INT_FAB = 2,
  MPQC_INT_GIJAB = 3,
  MPQC_INT_GIJKL = 4,
  MPQC_INT_GIAJB = 5,
  MPQC_INT_GIABC = 6,
  MPQC_INT_GIJKA = 7,
  MPQC_INT_GABCD = 8,
  /* three-index density-fitting factors X(K,p,q) = (K|G|p q)[inv_sqr]
   * (ccsd.h:481-493), index orders "K,a,i", "K,i,j", "K,a,b" */
  MPQC_INT_XAI = 9,
  MPQC_INT_XIJ = 10,
  MPQC_INT_XAB = 11
};

/* Library version string. */
const char* mpqc_ccsd_version(void);

/* Creates a context for o active occupied and v virtual orbitals on the current
 * HIP device.  Hashmarks (n_tiles+1 entries, first 0, last o or v) describe the
 * TiledArray tiling used by MPQC_LAYOUT_TA_TILES buffers; pass NULL for one tile.
 * precision: MPQC_CCSD_PREC_* for the contractions (the update, DIIS and energy
 * are always FP64). */
int mpqc_ccsd_ctx_create(int o, int v, const int64_t* occ_hashmarks, int n_occ_tiles,
                         const int64_t* vir_hashmarks, int n_vir_tiles, int precision,
                         mpqc_ccsd_ctx** out);
int mpqc_ccsd_ctx_destroy(mpqc_ccsd_ctx* ctx);
const char* mpqc_ccsd_last_error(const mpqc_ccsd_ctx* ctx);

/* Numeric options: "fov_sign" (+1 reference F_VV as written, -1 exact CCSD for
 * non-Brillouin references), "precision" (MPQC_CCSD_PREC_*), "kernel_timing"
 * (1 = time every contraction with HIP events on the context stream),
 * "method" (0 standard, 1 df, 2 direct, 3 direct_df; ccsd.h:73, 94-99), "reduced_abcd_memory" (1 default:
 * the df method builds (ab|cd) lazily from Xab instead of reading Gabcd and
 * Giabc, ccsd.h:102,292), "df_block_bytes" (device memory for the lazily built
 * (ab|cd) blocks, default 4 GiB; set before the first residual), "pp_symmetric" (1 default:
 * the particle-particle ladder of ccsd_r1_r2.h:275 as two GEMMs over the symmetric and
 * antisymmetric (c,d) pair combinations, rows a<=b, with half the FLOPs, exact for any
 * amplitudes given Gabcd(b,a,d,c) = Gabcd(a,b,c,d), and W_oooo's Gijab tau term likewise given
 * Gijab(l,k,d,c) = Gijab(k,l,c,d); blocks without the symmetry and the adjoint's recorded
 * residual keep the full contractions; 0: the full contractions), "graph" (1 default: the FP64
 * and FP32 residual is captured into a HIP graph on its second call and replayed; any option
 * change or new integrals re-capture; 0: eager launches), "solve_restart" (1 default: solve()
 * starts from the initial guess; 0: from the context's amplitudes; 2: from them projected onto
 * t2(a,b,i,j) = t2(b,a,j,i)).  Tuning keys that act on
 * every context of the process (measurement aids; the defaults are the measured best):
 * "plane_tile" (tile of the 16-bit plane contractions, rows * 1000 + columns, 0 automatic),
 * "plane_lds_stages" (1 or 2), "plane_split_cap" (most K splits of a plane contraction whose
 * tiles do not fill the GPU, default 512; set before the context's first residual), "plane_tile_order" (0 XCD ranges of row-major tiles, 1 plain,
 * 2 XCD ranges of column-major tiles), "plane_dl" (0 the deep-pipeline plane kernel where planned,
 * 1 never), "thin_k" (0 the streaming kernel for FP64 contractions with K <= 64 and at most 256
 * rows of A, 1 never). */
int mpqc_ccsd_set_option(mpqc_ccsd_ctx* ctx, const char* key, double value);

int mpqc_ccsd_set_integral(mpqc_ccsd_ctx* ctx, int which, const double* data, int layout,
                           int memspace);
int mpqc_ccsd_set_orbital_energies(mpqc_ccsd_ctx* ctx, const double* eps_occ_act,
                                   const double* eps_vir);
/* (T) correction from the amplitudes held by the context (normally after
 * mpqc_ccsd_solve) and its integrals Gijab, Gijka and Giabc (built from Xai and Xab
 * for the df method when not set) and orbital energies.  *e_t gets E(T);
 * *seconds (may be NULL) the device time.  Option "triples_block_bytes" (default
 * 8 GiB) bounds the device memory of the raw triples of one block. */
int mpqc_ccsd_triples(mpqc_ccsd_ctx* ctx, double* e_t, double* seconds);
/* EOM-EE-CCSD: the n_roots lowest singlet excitation energies (Eh, ascending) at
 * the amplitudes held by the context (normally after mpqc_ccsd_solve), written to
 * energies[0..n_roots).  They are the eigenvalues of the Jacobian of the residual
 * on (ab)(ij)-symmetric doubles, found by Davidson iterations (on its transpose,
 * whose products are exact reverse sweeps of the residual).  Converged when the
 * largest change of a root <= convergence and the largest residual norm <=
 * 10 convergence (davidson_diag.h:139-141), or when no new vector passes the
 * vector threshold (the roots are then returned as they stand, as the reference
 * does); an error after max_iter iterations otherwise.  *n_iter (may be NULL) gets the
 * iteration count; verbose prints one line per iteration.  Options
 * "eom_max_vector" (subspace vectors per root, default 8), "eom_vector_threshold"
 * (norm below which a new vector is dropped, default 10 convergence) and
 * "eom_n_guess" (initial vectors, default 2 n_roots), "eom_extra_roots" (roots tracked above the requested ones, default 0),
 * "eom_cis_diagonal" (1: singles preconditioned with the CIS singlet diagonal instead of
 * FAB(a,a) - FIJ(i,i), which also orders the unit guesses; default 0), "eom_cis_guess" (1: the initial vectors are the
 * eom_n_guess lowest CIS singlet eigenvectors, from a Davidson run on the CIS matrix built on the
 * device, as the reference's CIS guess, eom_ccsd_impl.h:431-450; default 1; 0: unit singles).  Runs in the f64 and f32 precision forms; the df method
 * evaluates the same Jacobian on the four-index blocks its factors assemble
 * (Giabc and Gabcd are built on the device when not set). */
int mpqc_ccsd_eom_ee(mpqc_ccsd_ctx* ctx, int n_roots, int max_iter, double convergence, int verbose,
                     double* energies, int* n_iter);
/* Products with the transposed Jacobian made by the last mpqc_ccsd_eom_ee and
 * their device time (HIP events on the context stream).  No counterpart in the
 * reference (its Davidson prints per-iteration times instead). */
int mpqc_ccsd_eom_stats(mpqc_ccsd_ctx* ctx, int* n_products, double* product_seconds);
/* Reads back an integral block held by the context (as set, or as built by the
 * transforms below), in the layout and index order of mpqc_ccsd_set_integral. */
int mpqc_ccsd_get_integral(mpqc_ccsd_ctx* ctx, int which, double* out, int layout, int memspace);
/* AO -> MO on the device: fills every integral block (Fia..Gabcd) from the AO
 * integrals (mu nu|la si) (nbf^4, chemist notation, row-major), the AO Fock matrix
 * (nbf x nbf) and the active occupied / virtual MO coefficients (nbf x o and
 * nbf x v, row-major: C[mu][p]).  All four in host or device memory (memspace). */
int mpqc_ccsd_transform_ao(mpqc_ccsd_ctx* ctx, int nbf, const double* ao_eri, const double* fock_ao,
                           const double* c_occ, const double* c_vir, int memspace);
/* The density-fitted integral set of the "df" method on the device (after
 * mpqc_ccsd_set_df_basis): from the three-centre integrals (K|mu nu) (naux x nbf x nbf)
 * and a metric factor L (naux x naux, L L^T = V^-1 for the Coulomb metric V; the
 * fitted factors are X = L^T (K|mu nu); pass NULL when the buffer is already fitted),
 * builds Xai, Xij, Xab, the four-index blocks Gijab, Gijkl, Giajb, Gijka (and Giabc,
 * Gabcd with reduced_abcd_memory = 0) and Fia, Fij, Fab from the AO Fock matrix. */
int mpqc_ccsd_transform_df(mpqc_ccsd_ctx* ctx, int nbf, const double* ao_three_center, const double* metric,
                           const double* fock_ao, const double* c_occ, const double* c_vir, int memspace);

/* Declares naux fitting functions tiled by aux_hashmarks (NULL: one tile) and
 * switches the context to the "df" residual.  Then MPQC_INT_XAI/XIJ/XAB are set
 * with mpqc_ccsd_set_integral (the fitting index uses these hashmarks in
 * MPQC_LAYOUT_TA_TILES buffers); Giabc and Gabcd are needed only with
 * reduced_abcd_memory = 0.  The df residual runs in every precision form; the 16-bit
 * split forms need reduced_abcd_memory (the (ab|cd) blocks go through planes). */
int mpqc_ccsd_set_df_basis(mpqc_ccsd_ctx* ctx, int naux, const int64_t* aux_hashmarks,
                           int n_aux_tiles);
/* The AO integral-direct method ("direct", ccsd.h:73): the AO integrals (mu nu|la si)
 * (nbf^4, chemist notation, row-major) and the active occupied / virtual MO coefficients
 * (nbf x o, nbf x v, row-major C[mu][p]), host or device memory.  Replaces
 * AOFactory::compute_direct("(mu nu|G|ka la)") and ints.Ci / ints.Ca (ccsd.h:246-247,
 * 313-322); switches the context to the direct residual, which needs every block but
 * Gabcd and forms U = (pq|rs) [Ca tau Ca^T] each iteration (compute_u2_u11,
 * ccsd.h:573-603; used at ccsd_r1_r2.h:116-117, 282-286).  The AO integrals stay
 * resident in HBM instead of being recomputed.  With a fitting basis as well
 * (mpqc_ccsd_set_df_basis), "method" = 3 selects direct_df: the df residual with U
 * (ccsd.h:436, 457; ccsd_r1_r2.h:509-514).  Runs in MPQC_CCSD_PREC_F64 and F32. */
int mpqc_ccsd_set_direct_ao(mpqc_ccsd_ctx* ctx, int nbf, const double* ao_eri, const double* c_occ,
                            const double* c_vir, int memspace);
int mpqc_ccsd_solver_init(mpqc_ccsd_ctx* ctx, int diis_start, int n_diis, double diis_damp,
                          int diis_ngroup, int diis_group_nstart);

int mpqc_ccsd_set_amplitudes(mpqc_ccsd_ctx* ctx, const double* t1, const double* t2, int layout,
                             int memspace);
int mpqc_ccsd_get_amplitudes(mpqc_ccsd_ctx* ctx, double* t1, double* t2, int layout,
                             int memspace);
int mpqc_ccsd_get_residuals(mpqc_ccsd_ctx* ctx, double* r1, double* r2, int layout,
                            int memspace);
/* The effective one-body Hamiltonians FIJ ("k,i") and FAB ("a,c") that
 * compute_cs_ccsd_r1 writes into cc::Integrals (host buffers). */
int mpqc_ccsd_get_dressed_fock(mpqc_ccsd_ctx* ctx, double* FIJ, double* FAB);

/* Sets r1 = f_ai, r2 = g_abij, t = 0 (ccsd.h:331-352). */
int mpqc_ccsd_init_guess(mpqc_ccsd_ctx* ctx);
int mpqc_ccsd_residual(mpqc_ccsd_ctx* ctx);
int mpqc_ccsd_energy_error(mpqc_ccsd_ctx* ctx, double* energy, double* error);
int mpqc_ccsd_update(mpqc_ccsd_ctx* ctx);

/* Full solver loop.  Returns the correlation energy and the number of loop
 * passes; target_precision as Energy::target_precision (default 1e-9). */
int mpqc_ccsd_solve(mpqc_ccsd_ctx* ctx, int max_iter, double target_precision, int verbose,
                    double* e_corr, int* n_iter);
/* Rows of the iteration table of the last solve: iteration, dE, residual error,
 * energy, wall seconds of that iteration.  Returns the row count in *n_rows. */
int mpqc_ccsd_get_trace(mpqc_ccsd_ctx* ctx, int max_rows, int* n_rows, int* iters, double* dE,
                        double* error, double* energy, double* seconds);

/* Accumulated contraction timings (when "kernel_timing" is on): for each tag
 * (the reference expression, e.g. "r2:pp_ladder"), the number of launches,
 * total milliseconds and total FLOPs (2*M*N*K).  names receives max_rows
 * NUL-terminated strings of 64 bytes each.  This is the measurement the
 * reference gets from its per-phase timers (ccsd.h:392-394,442-444,463-465). */
int mpqc_ccsd_get_kernel_times(mpqc_ccsd_ctx* ctx, int max_rows, int* n_rows, char* names,
                               int64_t* calls, double* ms, double* flops);

/* ---- adjoint (no counterpart in the reference) ----
 * x = (dR/dt)^T y at the context's current amplitudes: y1 "a,i", y2 "a,b,i,j" in,
 * x1, x2 out (same layouts).  Runs one forward residual first. */
int mpqc_ccsd_jacobian_transpose(mpqc_ccsd_ctx* ctx, const double* y1, const double* y2, double* x1,
                                 double* x2, int layout, int memspace);
/* Solves (dR/dt)^T lambda = -dE/dt at the current (converged) amplitudes, to
 * sqrt(|res1|^2 + |res2|^2) / (size1 + size2) <= target_precision (the amplitude
 * equations' error measure, chemistry/qc/cc/solvers.h:60-63). */
int mpqc_ccsd_lambda_solve(mpqc_ccsd_ctx* ctx, int max_iter, double target_precision, int verbose,
                           int* n_iter);
int mpqc_ccsd_get_lambda(mpqc_ccsd_ctx* ctx, double* l1, double* l2, int layout, int memspace);
/* dE/d(integral block) for all nine blocks of cc::Integrals (after lambda_solve);
 * read each with mpqc_ccsd_get_integral_gradient (same index order and layout
 * as mpqc_ccsd_set_integral). */
int mpqc_ccsd_energy_gradient(mpqc_ccsd_ctx* ctx);
int mpqc_ccsd_get_integral_gradient(mpqc_ccsd_ctx* ctx, int which, double* out, int layout,
                                    int memspace);

/* ---- batched residuals of many independent molecules ----
 * One context per molecule (o[k], v[k]; one tile per mode, so the tile layout is
 * the dense layout), residuals cc::compute_cs_ccsd_r1 + r2 (ccsd_r1_r2.h:70-293)
 * of all molecules in grouped launches captured in one hipGraph.  The reference
 * evaluates molecules one at a time; there is no batched counterpart. */
typedef struct mpqc_ccsd_batch mpqc_ccsd_batch;
int mpqc_ccsd_batch_create(int n_mol, const int* o, const int* v, int precision,
                           mpqc_ccsd_batch** out);
int mpqc_ccsd_batch_destroy(mpqc_ccsd_batch* b);
const char* mpqc_ccsd_batch_last_error(const mpqc_ccsd_batch* b);
int mpqc_ccsd_batch_set_integral(mpqc_ccsd_batch* b, int mol, int which, const double* data,
                                 int layout, int memspace);
int mpqc_ccsd_batch_set_amplitudes(mpqc_ccsd_batch* b, int mol, const double* t1, const double* t2,
                                   int layout, int memspace);
int mpqc_ccsd_batch_residual(mpqc_ccsd_batch* b);
int mpqc_ccsd_batch_get_residuals(mpqc_ccsd_batch* b, int mol, double* r1, double* r2, int layout,
                                  int memspace);
/* Autotuner: times each candidate tile edge / split-K target of every grouped
 * contraction on this batch, keeps the fastest, re-captures the graph; *ms gets
 * the batched residual time afterwards. */
int mpqc_ccsd_batch_tune(mpqc_ccsd_batch* b, int reps, int verbose, double* ms);
int mpqc_ccsd_batch_time(mpqc_ccsd_batch* b, int reps, double* ms);
int mpqc_ccsd_batch_time_stages(mpqc_ccsd_batch* b, int reps);
/* Per grouped launch: tag (64 bytes each), op kind, tile edge, split-K target,
 * last measured ms, contraction FLOPs. */
int mpqc_ccsd_batch_get_stages(mpqc_ccsd_batch* b, int max_rows, int* n_rows, char* tags, int* kinds,
                               int* bm, int64_t* target, double* ms, double* flops);

/* Device-pointer kernel entry points (stream may be NULL for the null stream). */
int mpqc_ccsd_pp_ladder(int o, int v, const double* gabcd_dev, const double* tau_dev,
                        double* r2_dev, double alpha, double beta, int precision, void* stream);
int mpqc_ccsd_gemm(int transa, int transb, int m, int n, int k, double alpha, const double* a,
                   int64_t lda, int64_t stride_a, const double* b, int64_t ldb, int64_t stride_b,
                   double beta, double* c, int64_t ldc, int64_t stride_c, int batch, int precision,
                   void* stream);

/* mpqc_ccsd_gemm without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as mpqc_ccsd_gemm; its split-K partial
 * sums live in a fixed 64 MiB device array of the library (one per GPU, held
 * until the process ends) instead of a workspace the library allocates.  The
 * split-K plan is made to fit that array; a GEMM whose plan would need more is
 * refused with MPQC_CCSD_ERR_ARG.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race and
 * must be serialised by the caller. */
int mpqc_ccsd_gemm_noalloc(int transa, int transb, int m, int n, int k, double alpha, const double* a,
                           int64_t lda, int64_t stride_a, const double* b, int64_t ldb, int64_t stride_b,
                           double beta, double* c, int64_t ldc, int64_t stride_c, int batch,
                           int precision, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MPQC_CCSD_H */
