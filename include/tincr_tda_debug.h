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
