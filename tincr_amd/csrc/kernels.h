// Launchers of the non-GEMM kernels (stream.hip, eigh.hip, basis.hip).
#pragma once
#include "common.h"
#include "gemm.h"

namespace tda {

// ---- transition charges (transition_charges.rs:8-98) ----
// out[A][p][q] = 1/2 sum_{mu in A} (C[mu][pb+p] SC[mu][qb+q] + C[mu][qb+q] SC[mu][pb+p])
// for p < P, q < Q; rows p in [P, P_rows) and columns q in [Q, Q_pad) are zero.
// C, SC: [n_ao][ld] row-major FP64.  ao_off: [n_atoms+1] (device).
void launch_qblock(const double* C, const double* SC, int64_t ld, const int* ao_off, int n_atoms,
                   int pb, int P, int P_rows, int qb, int Q, int Q_pad, double* out, hipStream_t s);
// CT pair variant (ct_pair.rs:78-123): rows of atoms in [atom_lo, atom_hi) of a
// combined index space:  out[A][i][a] = 1/2 sum_mu L[mu][i] R[mu][a]
void launch_qprod(const double* Lm, int64_t ldl, const double* Rm, int64_t ldr, const int* ao_off,
                  int atom_lo, int n_atoms, int P, int P_rows, int Q, int Q_pad, double* out,
                  hipStream_t s);

// per-atom max |x| over blocks of `len` elements -> amax[n] (float)
void launch_block_amax(const double* x, int64_t len, int n, float* amax, hipStream_t s);
void launch_row_norms(const double* x, int n, int64_t len, int64_t ld, double* out, hipStream_t s);
// dst[b][e] = (T) (src[b][e] * inv_scale[b])   (inv_scale may be null => 1)
void launch_convert(const double* src, int64_t len, int n, const float* inv_scale, void* dst, int dtype,
                    hipStream_t s);
// FP16 hi / lo split copy: dst[r][c] = hi(src[r][c]), dst[r][lo_off + c] = lo
void launch_convert_split(const double* src, int64_t rows, int64_t cols, int64_t ld_src, void* dst, int64_t ld_dst,
                          int64_t lo_off, hipStream_t s);
void launch_convert_f32_to(const float* src, int64_t n, void* dst, int dtype, hipStream_t s);

// ---- layout conversions ----
// host order x[(i*v+a)*k + c]  <->  device xd[c*ovp + i*vp + a] (pads zero)
void launch_pack_x(const double* src, int o, int v, int vp, int k, double* dst, hipStream_t s);
void launch_unpack_x(const double* src, int o, int v, int vp, int k, double* dst, hipStream_t s);
// delta eps:  de[i*vp + a] = orbe[vb+a] - orbe[ob+i]  (a < v), 0 in pads
void launch_delta_eps(const double* orbe, int ob, int o, int vb, int v, int vp, double* de, hipStream_t s);
// dst[k][L] (double) -> (float/half/bf16)
void launch_cast_vectors(const double* src, int64_t n, void* dst, int dtype, hipStream_t s);
// FP32 -> FP64 element copy
void launch_widen(const float* src, int64_t n, double* dst, hipStream_t s);

// ---- dot products of the Davidson basis (basis.hip) ----
// C[i][j] (ldc) = alpha sum_{l<K} A[i][l] B[j][l] + beta C[i][j]; A [M][lda],
// B [N][ldb], each FP32 (a32 / b32) or FP64, rows 16-byte aligned; FP64
// products.  ws: rowdot_ws_bytes(M, N, K) bytes (may be NULL when 0).
size_t rowdot_ws_bytes(int M, int N, int64_t K);
void launch_rowdot(const void* A, bool a32, int64_t lda, const void* B, bool b32, int64_t ldb, int M, int N,
                   int64_t K, double alpha, double beta, double* C, int64_t ldc, double* ws, hipStream_t s);

// ---- Davidson helpers (davidson.rs:53-185, solvers/mod.rs:373-380) ----
// out = r / d,  d = w - de,  |d| < 1e-4 -> 1.0
// out = r / (w - de), |w - de| < thr replaced by 1 (thr = 0: no floor)
void launch_precondition(const double* r, double w, const double* de, int64_t L, double thr, double* out,
                         hipStream_t s);
// P[c] = precondition(R[idx[c]], u[idx[c]])   idx, u on device
void launch_precondition_many(const double* R, const double* u, const int* idx, int nc, const double* de,
                              int64_t L, double thr, double* P, hipStream_t s);
// R[r][e] = AVY[r][e] - u[r] * RZ[r][e]  (in place on AVY)
void launch_residual(double* AVY, const double* RZ, const double* u, int n, int64_t L, hipStream_t s);
// norms[r] = || X[r] ||  (deterministic two-pass), partial: >= n*256 doubles
void launch_norms(const double* X, int n, int64_t L, double* partial, double* norms, hipStream_t s);
// out[c] = X[c] . Y[c] for n vectors of length L (partial: n * 256 doubles)
void launch_dots(const double* X, const double* Y, int n, int64_t L, double* partial, double* out, hipStream_t s);
// conjugate gradients of the Z-vector solve (helpers.rs:1130-1223), per column c:
// Z = R / de (de == 0 -> 0); Z += a_c P, R -= a_c AP (a_c == 0: skipped);
// P = b_c P + R (b_c < 0: skipped)
void launch_cg_init(const double* R, const double* de, int n, int64_t L, double* Z, hipStream_t s);
void launch_cg_step(double* Z, double* R, const double* P, const double* AP, const double* a, int n, int64_t L,
                    hipStream_t s);
void launch_cg_dir(double* P, const double* R, const double* b, int n, int64_t L, hipStream_t s);
// Gram-Schmidt of nc new vectors among themselves (davidson.rs:162-168),
// after their projection against the old basis: W[j] (= P[j] - V V^T P[j])
// becomes u_j = W[j] - sum_{s<j accepted} q_s (q_s . P[j]); u_j is accepted
// when |u_j| > 1e-8; the accepted q_j = u_j / |u_j| go to Q[0..*count) in
// order.  scratch: gs_block_scratch_doubles(nc) doubles.
void launch_gs_block(const double* P, double* W, double* Q, int64_t L, int nc, double* scratch, int* count,
                     hipStream_t s);
size_t gs_block_scratch_doubles(int nc);
// The same block Gram-Schmidt in coefficient form, for nc <= 64.  With
// u_j = sum_r T[j][r] w_r, the dots it needs are u_s . p_j = (T Gwp)[s][j]
// and |u_j|^2 = T_j Gww T_j^T, so two Gram matrices replace the nc passes
// over the vectors:  Gt [2 nc][nc] with Gt[l][r] = p_l . w_r (l < nc) and
// Gt[nc + l][r] = w_l . w_r.  One wave runs the recurrence of
// launch_gs_block on them and writes Tq [*count][nc], the rows T_j / |u_j|
// of the accepted j in order (then Q = Tq W), wn[j] = |u_j| for every j,
// count[0] = accepted, count[1] = 1 when some |u_j|^2 < 1e-4 |w_j|^2 (the
// estimate has lost more than 4 of its 16 digits, the orthogonality of the
// result would lose as many: the caller falls back to launch_gs_block).
void launch_gs_gram(const double* Gt, int nc, double* Tq, double* wn, int* count, hipStream_t s);
void launch_copy(const double* src, double* dst, int64_t n, hipStream_t s);
void launch_fill(double* dst, double v, int64_t n, hipStream_t s);
// AO offsets [n_atoms + 1] from device per-atom AO counts, clamped to n_ao;
// *bad = 1 when the counts are negative or do not sum to n_ao
void launch_ao_offsets(const int32_t* opa, int n_atoms, int n_ao, int* off, int* bad, hipStream_t s);
// out[mu][nu] = (float) g[atom[mu]][atom[nu]] (mu, nu < n_ao; g [N][N], out row stride ld)
void launch_expand_pairs(const double* g, const int* atom, int n_ao, int N, int64_t ld, float* out, hipStream_t s);
// X[r][pos[r]] = 1 for r < n (rows of length L; pos on the device)
void launch_set_ones(double* X, int64_t L, const int64_t* pos, int n, hipStream_t s);
// zero columns [c0, c0 + cw) of `rows` rows of length ld (elements of `esz`
// bytes, esz even): the never-written pad rows of the exchange intermediate
void launch_zero_runs(void* base, int64_t rows, int64_t ld, int64_t c0, int64_t cw, int esz, hipStream_t s);
// byte-wise zero fill by a kernel (graph-capture safe, unlike hipMemsetAsync)
void launch_zero_bytes(void* dst, size_t bytes, hipStream_t s);
// tiny dense: C[M][N] = A[M][K] . B[K][N]   (FP64, row-major, small sizes)
void launch_small_gemm(const double* A, const double* B, double* C, int M, int N, int K, hipStream_t s);

// ---- AO-factored operator ----
// out[r][col] (r = s*n_ao + mu, s in {0,1}) = (s == swap ? SC : C)[mu][base + col]
// for col < P, 0 for P <= col < P_pad.  transpose=1 writes out[col][r] with row
// stride ld >= 2 n_ao (zeros in columns r >= 2 n_ao); ld is ignored otherwise.
void launch_ao_operand(const double* C, const double* SC, int n_ao, int base, int P, int P_pad, int swap,
                       int transpose, int64_t ld, double* out, hipStream_t s);
// Coulomb from the D-blocks of M (M = -1/4 gamma~ o D, 2n_ao x 2n_ao per vector
// with row stride ld, element type dtype):
// u[A*k + c] = 1/2 sum_{mu in A} (D[(0,mu),(1,mu)] + D[(1,mu),(0,mu)])
void launch_ao_charges(const void* M, int dtype, int k, int n_ao, int64_t ld, const int* ao_off, int n_atoms,
                       const double* gamma_lr, double* u, hipStream_t s,
                       int64_t lo_off = 0);
// transition charges without q_ov: XR [k][2n_ao][ld] (XR_c = R x_c^T), L
// [2n_ao][ld]; u[A*ldu + c] = 1/2 sum_{mu in A} sum_{i<o} (L[(0,mu)][i]
// XR_c[(1,mu)][i] + L[(1,mu)][i] XR_c[(0,mu)][i])
void launch_ao_dens_charges(const double* L, const double* XR, int k, int n_ao, int o, int64_t ld,
                            const int* ao_off, int n_atoms, double* u, int64_t ldu, hipStream_t s);
// out[A*N + B] = -1/2 sum_{mu in A, nu in B} (D[mu][nu] D[n_ao+mu][n_ao+nu] +
// D[mu][n_ao+nu] D[n_ao+mu][nu]) for D [2n_ao][ld] = L x R^T of one vector
void launch_ao_dglr(const double* D, int64_t ld, int n_ao, const int* ao_off, int n_atoms, double* out,
                    hipStream_t s);
// M[c][(0,mu),(1,mu)] += w[atom(mu)*k + c] and M[c][(1,mu),(0,mu)] += same
void launch_ao_add_coulomb(void* M, int dtype, int k, int n_ao, int64_t ld, const int* atom, const double* w,
                           hipStream_t s,
                           int64_t lo_off = 0);

// ---- subspace eigensolver (eigh, davidson.rs:93) ----
// H: m x m symmetric (upper triangle read), leading dim ldh; work >= eigh_work_doubles(m)
// Outputs the n lowest eigenvalues (ascending) and eigenvectors Y[r][0..m) (row per root).
size_t eigh_work_doubles(int m);

// Hellmann-Feynman Coulomb terms: dgam[k][A][B] = 2 u[A][k] u[B][k],
// cq[k][A] = 4 w[A][k] (u, w [N][n]; cq may be NULL)
void launch_hf_coulomb(const double* u, const double* w, int N, int n, double* dgam, double* cq, hipStream_t s);

// S[k] = D[k] + D[k]^T for n matrices N x N
void launch_sym_add(const double* D, int n, int N, double* S, hipStream_t s);

// Mulliken transition dipoles mu [3][n] = xyz^T qt and oscillator strengths
// f [n] = 2/3 w |mu|^2 (xyz [N][3], qt [N][n], w [n]; all device)
void launch_dipoles(const double* xyz, const double* qt, const double* w, int N, int n, double* mu, double* f,
                    hipStream_t s);
// phase_clk (optional, device): wall-clock stamps at the phase boundaries
// (copy, tridiagonalisation, bisection, inverse iteration, back-transform).
void launch_eigh(const double* H, int ldh, int m, int n, double* work, double* evals, double* Y, int ldy,
                 hipStream_t s, unsigned long long* phase_clk = nullptr);
// Batched form for the fragments of a batch: EighDesc b of slot->descs (m <=
// 256, n <= min(m, 64), work m*m + m doubles) solved by workgroup b <
// slot->count.
struct EighDesc {
  const double* H;
  double* work;
  double* evals;
  double* Y;
  int ldh, m, n, ldy;
};
void launch_eigh_batched(const GroupSlot* slot, int grid, hipStream_t s);

}  // namespace tda
