/*
 * Kernel-level test hooks of libtincr_tda.so.  No counterpart in the
 * reference: they expose single kernels (GEMM, subspace eigensolver) so the
 * parity tests can check each one against numpy on its own.
 */
#ifndef TINCR_TDA_DEBUG_H
#define TINCR_TDA_DEBUG_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* C[M,N] = alpha * A.B + beta * C, A and B converted on the device to
 * `in_prec` (TDA_PREC_*), C accumulated and stored in FP64.
 * a_kc: A stored [M][K] (else [K][M]); b_kc: B stored [N][K] (else [K][N]).
 * ksplit: 0 = automatic. */
int tda_debug_gemm(int in_prec, int a_kc, int b_kc, int64_t M, int64_t N, int64_t K, const double* A,
                   const double* B, double* C, double alpha, double beta, int ksplit);

/* As tda_debug_gemm with FP64 products, the operands flagged a32 / b32
 * stored on the device as FP32 (rounded from the given doubles) and widened
 * to FP64 inside the kernel: C = alpha * fl32(A) . fl32(B) + beta * C. */
int tda_debug_gemm_f32store(int a_kc, int b_kc, int64_t M, int64_t N, int64_t K, const double* A, const double* B,
                            double* C, double alpha, double beta, int ksplit, int a32, int b32);

/* Dot-product stream of the Davidson basis (FP64 arithmetic; a32 / b32: that
 * operand stored on the device as FP32, rounded from the doubles):
 * C[M][N] = alpha * A[M][K] . B[N][K]^T + beta * C */
int tda_debug_rowdot(int M, int N, int64_t K, const double* A, int a32, const double* B, int b32, double* C,
                     double alpha, double beta);

/* The K-contiguous 16-bit GEMM path for batch entries z < batch:
 *   C[z][m][n] = epi(alpha * sum_k A[m][k] B[z][n][k])
 * A [M][lda], B [batch][N][ldb] (row strides >= K; elements past K are never
 * read), inputs converted on the device to in_prec (TDA_PREC_BF16 / FP16).
 * epi 0: C in in_prec; 1: C FP64 = alpha acc + diag[m][n] X[z][m][n];
 * 2: C in in_prec = alpha acc * H[r(m)][r(n)], H [nh][nh], r(x) = hidx[x % hmod]
 * (hidx NULL: r(x) = x).  split (FP16 only): A, B and a 16-bit C held as
 * hi / lo pairs (hi = fp16(v), lo = fp16((v - hi) 2^11)), the product
 * evaluated as 2^-11 (A_hi B_lo + A_lo B_hi) + A_hi B_hi.  C is returned as
 * FP64 [batch][M][N] (hi + lo 2^-11 when split); *used = 1 when the
 * dedicated kernel ran (0: the generic kernel).  split is a flag word:
 * bit 0 the split above; bit 1 (epi 2 with hidx) hands the factors over as
 * the precomputed FP32 table hfull[(m % hmod)][(n % hmod)] = H[r(m)][r(n)]
 * the AO operator uses, instead of the gathered H. */
int tda_debug_gemm16(int in_prec, int epi, int64_t M, int64_t N, int64_t K, int batch, int64_t lda, int64_t ldb,
                     const double* A, const double* B, double alpha, const double* H, int nh, const int* hidx,
                     int hmod, const double* diag, const double* X, double* C, int split, int* used);

/* Lowest n eigenpairs of the symmetric m x m matrix H (upper triangle read):
 * evals [n] ascending, Y [n][m] (row r = eigenvector r). */
int tda_debug_eigh(int m, int n, const double* H, double* evals, double* Y);
/* As tda_debug_eigh, and phase_us[5] receives the device time (us) of the
 * phases: copy, tridiagonalisation, bisection, inverse iteration,
 * back-transformation. */
int tda_debug_eigh_timed(int m, int n, const double* H, double* evals, double* Y, double* phase_us);

/* `reps` launches of calib_copy_kernel reading and writing `bytes` each (a
 * known-traffic reference for hardware-counter runs). */
int tda_debug_copy(int64_t bytes, int reps);

#ifdef __cplusplus
}
#endif
#endif
