/*
 * tincr_tda -- MI355X (gfx950) LC-TDA excited-state hot path, C ABI.
 *
 * Drop-in for tincr's src/excited_states path (Rust).  Each entry point names
 * the reference interface it replaces.  Conventions:
 *   - all arrays are caller-owned, host memory, row-major FP64;
 *   - layouts are the reference's: q_ov [n_atoms, n_occ*n_virt] with column
 *     i*n_virt+a, q_oo [n_atoms, n_occ^2], q_vv [n_atoms, n_virt^2], trial /
 *     CI vectors [n_occ*n_virt, k] (one column per vector, row i*n_virt+a);
 *   - the library copies inputs to the GPU and owns all device memory;
 *   - every call returns a status (TDA_OK == 0); tda_last_error() explains;
 *   - distinct contexts are independent (one HIP stream each); one context
 *     must not be used from two threads at once.
 */
#ifndef TINCR_TDA_H
#define TINCR_TDA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TDA_OK 0
#define TDA_ERR_NOT_CONVERGED 1 /* outputs hold the last Ritz pairs (davidson.rs:80,184 DavidsonError) */
#define TDA_ERR_BAD_ARG 2
#define TDA_ERR_OOM 3
#define TDA_ERR_HIP 4
#define TDA_ERR_STATE 5 /* call order violated (e.g. run before set_system) */

/* precision of the bulk products (q_ov / gamma / exchange contractions) */
#define TDA_PREC_FP64 0 /* reference arithmetic */
#define TDA_PREC_FP32 1 /* FP32 products, FP64 subspace + final FP64 correction */
#define TDA_PREC_BF16 2 /* BF16 exchange (per-atom scaled, FP32 accumulate) */
#define TDA_PREC_FP16 3 /* FP16 exchange (per-atom scaled, FP32 accumulate) */

/* form of the A.x operator */
#define TDA_OP_AO 0      /* exchange contracted in the AO basis from C, SC and gamma_lr (default) */
#define TDA_OP_CHARGES 1 /* exchange from the per-atom charges q_oo / q_vv (three-stage, as solvers/mod.rs:328-360) */

typedef struct tda_ctx tda_ctx;

typedef struct tda_stats {
  int32_t n_iter;           /* Davidson iterations, all phases */
  int32_t n_iter_fp64;      /* iterations of the FP64 correction phase */
  int32_t n_products;       /* single-vector A.x products, all precisions */
  int32_t n_products_fp64;  /* of which in FP64 */
  int32_t converged;        /* 1 if every root met the tolerance in FP64 */
  int32_t n_atom_blocks;    /* atom blocks per exchange (HBM budget) */
  int32_t n_iter_fp32;      /* iterations of the FP32 phase (mixed precision) */
  int32_t n_products_fp32;  /* products of the FP32-accuracy phase */
  int32_t split_fp32;       /* 1: that phase used FP16 hi / lo split products (AO operator) */
  double max_residual;      /* max FP64 residual norm of the returned roots */
  double t_total_ms;        /* tda_run wall time */
  double t_qbuild_ms;       /* q build + gamma_lr.q_oo (inside tda_run if not prepared) */
  double t_products_ms;     /* sum of A.x product time (device events) */
  /* exchange contraction kernels, [stage][class].  TDA_OP_CHARGES: stage 0 =
   * X.q_vv, 1 = G^T.T; TDA_OP_AO: stage 0 = transition densities L X R^T
   * (gamma_lr Hadamard epilogue), 1 = their back-contraction with R'.
   * class 0 = FP64 products, 1 = FP32-accuracy products (FP32, or FP16 hi / lo
   * split when split_fp32: flops then count the three MFMA products), 2 =
   * FP16 / BF16 products.
   * Time of the main MFMA kernel only (device events on the context stream),
   * its launch count and its algorithmic flops. */
  double exch_ms[2][3];
  int32_t exch_launches[2][3];
  double exch_flops[2][3];
} tda_stats;

/* Create a context on HIP device `device`.  hbm_budget_bytes bounds the
 * exchange intermediates (0 = 16 GiB).  precision: TDA_PREC_*. */
int tda_ctx_create(int device, size_t hbm_budget_bytes, int precision, tda_ctx** out);
void tda_ctx_destroy(tda_ctx* ctx);
const char* tda_last_error(const tda_ctx* ctx);
int tda_set_precision(tda_ctx* ctx, int precision);
/* Selects the operator form (TDA_OP_*); both give the reference's A.x. */
int tda_set_operator(tda_ctx* ctx, int op);

/* Replaces trans_charges (src/excited_states/transition_charges.rs:8-15).
 * orbs, s: [n_ao, n_ao]; orbs_per_atom: [n_atoms] (AOs of each atom are
 * contiguous); occupied MOs occ_begin..occ_begin+n_occ-1 and virtual MOs
 * virt_begin..virt_begin+n_virt-1 (contiguous ranges, as :21-27 uses them).
 * Outputs q_ov [n_atoms, n_occ*n_virt], q_oo [n_atoms, n_occ^2],
 * q_vv [n_atoms, n_virt^2] (any may be NULL). */
int tda_trans_charges(tda_ctx* ctx, int n_atoms, int n_ao, const int32_t* orbs_per_atom, const double* orbs,
                      const double* s, int occ_begin, int n_occ, int virt_begin, int n_virt, double* q_ov,
                      double* q_oo, double* q_vv);

/* No counterpart in the reference: tda_trans_charges for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and outputs, but every
 * array is a DEVICE pointer owned by the caller (orbs_per_atom int32
 * [n_atoms], orbs and s [n_ao, n_ao], outputs in the reference layouts; NULL
 * skips an output), and the call allocates nothing: S C and the AO offsets
 * live in a fixed per-GPU scratch array of TDA_NOALLOC_SCRATCH_BYTES (held by
 * each GPU the library runs on until the process ends), which limits
 * n_ao^2 * 8 + 4 (n_atoms + 2) to that size (n_ao <= 3547): larger systems
 * return TDA_ERR_BAD_ARG before any work.  Runs on the context's stream and
 * returns when the outputs are written.  All calls on one GPU share its
 * scratch array: the library serialises its own calls within a process;
 * calls from different processes on one GPU must not overlap.  Counts that
 * are negative or do not sum to n_ao give TDA_ERR_BAD_ARG (outputs then
 * undefined; reads stay inside the n_ao rows). */
#define TDA_NOALLOC_SCRATCH_BYTES ((size_t)96 << 20)
int tda_trans_charges_noalloc(tda_ctx* ctx, int n_atoms, int n_ao, const int32_t* orbs_per_atom, const double* orbs,
                              const double* s, int occ_begin, int n_occ, int virt_begin, int n_virt, double* q_ov,
                              double* q_oo, double* q_vv);

/* Replaces orbe_differences (src/excited_states/utils.rs:8-24):
 * omega[i*n_virt+a] = virt[a] - occ[i]. */
int tda_orbe_differences(const double* occ, int n_occ, const double* virt, int n_virt, double* omega);

/* Loads the inputs System/Monomer::prepare_tda reads from its Properties
 * (src/excited_states/tda/system.rs:10-52, tda/monomer.rs:8-53): orbs, S,
 * orbe, gamma [n_atoms^2], gamma_lr [n_atoms^2] (NULL => no long-range
 * correction, as gammafunction_lc == None, solvers/mod.rs:314), xyz
 * [n_atoms, 3] in bohr (used for transition dipoles, tda/moments.rs:7-14). */
int tda_set_system(tda_ctx* ctx, int n_atoms, int n_ao, const int32_t* orbs_per_atom, const double* orbs,
                   const double* s, const double* orbe, int occ_begin, int n_occ, int virt_begin, int n_virt,
                   const double* gamma, const double* gamma_lr, const double* xyz);

/* Loads a charge-transfer pair H -> L, replacing
 * PairChargeTransfer::prepare_ct_tda (src/excited_states/tda/ct_pair.rs:12-76)
 * and selecting its DavidsonEngine (solvers/mod.rs:165-274): holes in the
 * occupied MOs of monomer H, particles in the virtual MOs of monomer L.
 *   H: n_atoms_h, n_ao_h, orbs_per_atom_h, orbs_h, s_h [n_ao_h^2], orbe_h,
 *      occupied range occ_begin_h..+n_occ_h;
 *   L: n_atoms_l, n_ao_l, orbs_per_atom_l, orbs_l, s_l [n_ao_l^2], orbe_l,
 *      virtual range virt_begin_l..+n_virt_l;
 *   s_hl [n_ao_h, n_ao_l]: AO overlap block (rows H, columns L);
 *   gamma_h [n_atoms_h^2], gamma_l [n_atoms_l^2], gamma_hl [n_atoms_h,
 *      n_atoms_l]: blocks of gamma the pair gamma is assembled from (:20-27);
 *   gamma_lr_hl [n_atoms_h, n_atoms_l] (NULL => no exchange term);
 *   xyz [(n_atoms_h + n_atoms_l), 3], H atoms first (pair atom order, :148).
 * q_ov spans the H and L atoms (:78-123), q_oo is H's, q_vv is L's, and the
 * exchange contracts gamma_lr_hl^T with q_oo (solvers/mod.rs:219-222).  The
 * pair always uses the per-atom-charges operator (TDA_OP_CHARGES); the
 * context's operator setting applies again after the next tda_set_system.
 * tda_get_charges then returns q_oo over the H atoms and q_vv over the L atoms. */
int tda_set_ct_pair(tda_ctx* ctx, int n_atoms_h, int n_ao_h, const int32_t* orbs_per_atom_h, const double* orbs_h,
                    const double* s_h, const double* orbe_h, int occ_begin_h, int n_occ_h, int n_atoms_l, int n_ao_l,
                    const int32_t* orbs_per_atom_l, const double* orbs_l, const double* s_l, const double* orbe_l,
                    int virt_begin_l, int n_virt_l, const double* s_hl, const double* gamma_h,
                    const double* gamma_l, const double* gamma_hl, const double* gamma_lr_hl, const double* xyz);

/* Selects the dense-matrix engine (DavidsonEngine for Array2,
 * src/excited_states/solvers/mod.rs:36-51), as the FMO exciton Hamiltonian
 * uses it (src/fmo/lcmo/basis.rs:497-503): h [dim, dim] row-major,
 * products h x, preconditioner r / (w - diag(h)) with no floor, guess at the
 * smallest diagonal elements.  tda_run then solves in FP64 and accepts only
 * omega and X (q_trans, tr_dipoles, f must be NULL); X is [dim, n_roots]. */
int tda_set_dense(tda_ctx* ctx, int dim, const double* h);

/* prepare_tda: builds q_ov/q_oo/q_vv and omega on the GPU (and
 * gamma_lr.q_oo, once per system instead of once per product call as
 * solvers/mod.rs:328-331 does). */
int tda_prepare(tda_ctx* ctx);

/* Drops the prepared charges, as Properties::reset does for q_ov/q_oo/q_vv/
 * omega (src/properties/mod.rs:23-63): the next tda_prepare / tda_run
 * rebuilds them from the loaded orbitals. */
int tda_clear_charges(tda_ctx* ctx);

/* Copies the prepared charges / omega back (reference layouts; NULL skips). */
int tda_get_charges(tda_ctx* ctx, double* q_ov, double* q_oo, double* q_vv, double* omega);

/* DavidsonEngine::compute_products without the product cache
 * (src/excited_states/solvers/mod.rs:20-34, :280-369): ax = A x for the k
 * columns of x [ov, k], in the context's precision. */
int tda_compute_products(tda_ctx* ctx, const double* x, int k, double* ax);

/* Z-vector solve of the excited-state gradient, tda_zvector_lc /
 * tda_zvector_no_lc (src/gradients/helpers.rs:1130-1223, :1053-1128):
 * conjugate gradients on A z = rhs for each of the k columns of rhs [ov, k]
 * (layout of X), with A the operator of mult_av_lc (helpers.rs:1384-1437:
 * the forward singlet A.x, omega + 2 J - K) or, without gamma_lr,
 * mult_av_nolc (:1351-1382: omega + 4 J), in FP64.  Start z = rhs / omega;
 * a column stops when r.r <= conv (1e-19 in the reference) or after max_iter
 * iterations (1e4 in the reference).  z [ov, k]; n_iter [k] (optional):
 * iterations per column (max_iter if it did not converge). */
int tda_zvector(tda_ctx* ctx, const double* rhs, int k, int max_iter, double conv, double* z, int32_t* n_iter);

/* DavidsonEngine::precondition (solvers/mod.rs:373-380):
 * out = r / (w - omega), entries with |w - omega| < 1e-4 divided by 1. */
int tda_precondition(tda_ctx* ctx, const double* r, double w, double* out);

/* System::run_tda / Monomer::run_tda (src/excited_states/tda/mod.rs:19-131):
 * guess (solvers/utils.rs:17-26), Davidson (solvers/davidson.rs:53-199),
 * q_trans = q_ov X, Mulliken transition dipoles and oscillator strengths
 * (tda/moments.rs:7-24).  Outputs: omega [n_roots] (ci_eigenvalues), X
 * [ov, n_roots] (ci_coefficients), q_trans [n_atoms, n_roots], tr_dipoles
 * [3, n_roots], f [n_roots]; any output may be NULL.  Calls tda_prepare
 * first if needed.  In a mixed precision the Davidson runs on low-precision
 * products and then continues with FP64 products from the Ritz vectors until
 * the FP64 residuals meet `tolerance`. */
int tda_run(tda_ctx* ctx, int n_roots, int max_iter, double tolerance, double* omega, double* X, double* q_trans,
            double* tr_dipoles, double* f, tda_stats* stats);

/* ------------------------------------------------------------------------
 * Adjoint (Hellmann-Feynman) derivatives of the excitation energies w_k =
 * x_k^T A x_k of a System/Monomer with respect to the operator's inputs, for
 * the converged CI vectors.  No reference function computes these directly;
 * they are the gamma / charge terms of the LC-TDA gradient
 * (src/gradients/excited_state.rs:259-487) at frozen orbitals.
 * X: [ov, n_states] (ci_coefficients layout) or NULL for the last tda_run's
 * vectors.  Outputs (any may be NULL):
 *   d_gamma    [n_states, n_atoms, n_atoms]  dw/dgamma_AB    = 2 u_A u_B, u = q_ov x
 *   d_gamma_lr [n_states, n_atoms, n_atoms]  dw/dgamma_lr,AB = -sum_ij q_oo^A_ij (x q_vv^B x^T)_ij
 *   c_qov      [n_states, n_atoms]           dw/dq_ov[A, ia] = c_qov[k][A] * x_ia  (= 4 (gamma u)_A)
 *   d_qoo      [n_states, n_atoms, n_occ^2]  dw/dq_oo^A_ij   = -sum_B gamma_lr,AB (x q_vv^B x^T)_ij
 *   d_qvv      [n_states, n_atoms, n_virt^2] dw/dq_vv^B_ab   = -(x^T G^B x)_ab, G^B = sum_A gamma_lr,BA q_oo^A
 * Entries of gamma / gamma_lr are treated as independent (not symmetrised).
 * ------------------------------------------------------------------------ */
int tda_adjoint(tda_ctx* ctx, int n_states, const double* X, double* d_gamma, double* d_gamma_lr, double* c_qov,
                double* d_qoo, double* d_qvv);

/* Chains the gamma terms to nuclear coordinates at frozen charges.
 * dgamma_dr / dgamma_lr_dr [3 n_atoms, n_atoms, n_atoms] in the layout of
 * gamma_gradients_atomwise (src/scc/gamma_approximation.rs:256-282): entry
 * [3i + d][i][j] = d gamma_ij / d R_i,d, the other rows zero.  Since gamma is
 * symmetric, gamma_ji moves with R_i too:
 *   grad[k][c] = sum_AB g1[c][A][B] (dw_k/dgamma_AB + dw_k/dgamma_BA)
 *              + the same with gamma_lr (dgamma_lr_dr may be NULL).
 * grad: [n_states, 3 n_atoms]. */
int tda_adjoint_nuclear(tda_ctx* ctx, int n_states, const double* X, const double* dgamma_dr,
                        const double* dgamma_lr_dr, double* grad);

/* ------------------------------------------------------------------------
 * Excited-state nuclear gradient d omega / dR of one state, replacing
 * System::tda_gradient_lc (src/gradients/excited_state.rs:259-487) and
 * tda_gradient_nolc (:56-257, the context has no gamma_lr), with the inputs
 * that ground_state_gradient(true) + prepare_excited_grad store in Properties
 * (src/gradients/ground_state.rs:12-131, excited_state.rs:11-54).  Needs a
 * System/Monomer context whose occupied MOs are 0..n_occ-1 and virtual MOs
 * n_occ..n_ao-1 (as System::occ_indices / virt_indices).  FP64 throughout.
 *   form: TDA_GRAD_REFERENCE follows the reference; TDA_GRAD_EXACT replaces
 *     the two LC response terms that finite differences reject (the Z-vector
 *     operator becomes the ground-state orbital Hessian A + B = de + 4 J - K
 *     - K', K'_ia,jb = (ib|ja)_lr, instead of mult_av_lc's TDA A; H+[z] in
 *     W_ij carries (kj|ib) + (kb|ij) instead of twice (kj|ib)).  Without
 *     gamma_lr both forms are the reference's.
 *   omega, x [n_occ * n_virt]: the state's excitation energy and CI vector
 *     (ci_eigenvalues[state], the state's column of ci_coefficients).
 *   dp [n_ao^2]: P - P_ref of the converged SCC (properties p, p_ref).
 *   ds_rows, dh_rows [3][n_ao][n_ao]: dS / dH0 in row form, [k][mu][nu] =
 *     d S[mu][nu] / d R[atom(mu)][k] (zero when mu, nu share an atom); the
 *     reference's grad_s[3A + k][mu][nu] (h0_and_s.rs:196-300) is this entry
 *     for atom(mu) == A and [k][nu][mu] for atom(nu) == A.
 *   dgamma, dgamma_lr [3][n_atoms][n_atoms]: [k][A][B] = d gamma_AB / d R[A][k],
 *     i.e. gamma_gradients_atomwise's entry [3A + k][A][B]
 *     (gamma_approximation.rs:256-282); dgamma_lr NULL without gamma_lr.
 *   max_iter, conv: Z-vector conjugate-gradient limits (1e4, 1e-19 on r.r in
 *     tda_zvector_lc, helpers.rs:1146-1147).
 * Outputs: grad [3 n_atoms] (gradExc); z [n_occ * n_virt] (optional): the
 * Z-vector; z_iter (optional): its CG iterations.
 * ------------------------------------------------------------------------ */
#define TDA_GRAD_REFERENCE 0
#define TDA_GRAD_EXACT 1
int tda_gradient(tda_ctx* ctx, int form, double omega, const double* x, const double* dp, const double* ds_rows,
                 const double* dh_rows, const double* dgamma, const double* dgamma_lr, int max_iter, double conv,
                 double* grad, double* z, int32_t* z_iter);

/* ------------------------------------------------------------------------
 * Batched fragments: the FMO monomer loop (src/fmo/lcmo/basis.rs:462-465,
 * prepare_tda + run_tda per monomer) and the CT-pair loop (basis.rs:172-238,
 * run_ct_tda per pair) in one call; each fragment converges on its own.
 * TDA_BATCH_GROUPED (default): the fragments' operators (per-atom charges
 * form) are prepared by worker threads, then one Davidson runs for all of
 * them together: every step is one grouped launch per operation over the
 * fragments still active (products, projections, subspace eigensolves,
 * residuals, Gram-Schmidt), converged fragments drop out.  FP64 batches use
 * FP64 products; the other precisions FP32 products with an FP64
 * verification of each fragment's roots (subspace algebra always FP64).
 * TDA_BATCH_POOL: worker threads, one context and HIP stream each, solve
 * whole fragments (tda_run), largest estimated cost first.
 * Input arrays are read during tda_batch_run only and must stay valid until
 * it returns; outputs are written there.
 * ------------------------------------------------------------------------ */
typedef struct tda_batch tda_batch;
#define TDA_BATCH_GROUPED 0
#define TDA_BATCH_POOL 1

int tda_batch_create(int device, int n_workers, size_t hbm_budget_bytes, int precision, tda_batch** out);
void tda_batch_destroy(tda_batch* batch);
const char* tda_batch_last_error(const tda_batch* batch);
/* operator form of the monomer solves (TDA_OP_*); CT pairs use the charges
 * form, and so does the grouped driver for every fragment */
int tda_batch_set_operator(tda_batch* batch, int op);
/* TDA_BATCH_GROUPED or TDA_BATCH_POOL (no counterpart in the reference, whose
 * loops are rayon iterators over the fragments) */
int tda_batch_set_mode(tda_batch* batch, int mode);
/* A monomer (tda_set_system arguments) solved as run_tda(n_roots, max_iter,
 * tol) with tda_run's outputs.  Returns the item index (>= 0) or -status. */
int tda_batch_add_system(tda_batch* batch, int n_atoms, int n_ao, const int32_t* orbs_per_atom, const double* orbs,
                         const double* s, const double* orbe, int occ_begin, int n_occ, int virt_begin, int n_virt,
                         const double* gamma, const double* gamma_lr, const double* xyz, int n_roots, int max_iter,
                         double tolerance, double* omega, double* X, double* q_trans, double* tr_dipoles, double* f);
/* A CT pair (tda_set_ct_pair arguments) solved as run_ct_tda(n_roots,
 * max_iter, tol).  Returns the item index (>= 0) or -status. */
int tda_batch_add_ct_pair(tda_batch* batch, int n_atoms_h, int n_ao_h, const int32_t* orbs_per_atom_h,
                          const double* orbs_h, const double* s_h, const double* orbe_h, int occ_begin_h, int n_occ_h,
                          int n_atoms_l, int n_ao_l, const int32_t* orbs_per_atom_l, const double* orbs_l,
                          const double* s_l, const double* orbe_l, int virt_begin_l, int n_virt_l, const double* s_hl,
                          const double* gamma_h, const double* gamma_l, const double* gamma_hl,
                          const double* gamma_lr_hl, const double* xyz, int n_roots, int max_iter, double tolerance,
                          double* omega, double* X, double* q_trans, double* tr_dipoles, double* f);
/* Forget the added items. */
int tda_batch_clear(tda_batch* batch);
/* Totals of the last grouped tda_batch_run (no counterpart in the reference):
 * exch_ms / exch_launches / exch_flops of the grouped exchange launches,
 * n_iter = the largest fragment iteration count, n_products summed,
 * t_total_ms the run's wall time; zero after a pool run. */
int tda_batch_last_totals(const tda_batch* batch, tda_stats* out);
/* Solves every item.  status [n_items] (optional) receives each item's
 * tda_run status, stats [n_items] (optional) its statistics.  Returns TDA_OK,
 * TDA_ERR_NOT_CONVERGED if some item did not converge (outputs hold its last
 * Ritz pairs), or the first other error. */
int tda_batch_run(tda_batch* batch, int32_t* status, tda_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* TINCR_TDA_H */
