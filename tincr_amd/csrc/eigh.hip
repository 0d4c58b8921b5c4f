// FP64 symmetric eigensolver for the Davidson subspace (davidson.rs:93,
// `a_proj.eigh(UPLO::Upper)` followed by taking the n lowest pairs, :96-97).
//
// One workgroup of 512 threads per problem (m <= 1024):
//   1. Householder reduction to tridiagonal form T = Q^T H Q (matrix in a
//      global-memory work copy, which stays resident in L2);
//   2. the n lowest eigenvalues of T by Sturm-count multisection: a wave
//      evaluates 64 shifts at once, so each round narrows the bracket 65x;
//   3. eigenvectors of T by inverse iteration with a partially pivoted
//      tridiagonal LU (as LAPACK dstein), modified Gram-Schmidt inside
//      clusters of close eigenvalues;
//   4. back-transformation y = Q z, one wave per root.
// Published methods: Householder tridiagonalisation and Sturm bisection as in
// Golub & Van Loan, "Matrix Computations" 8.3 / 8.4; inverse iteration with
// cluster reorthogonalisation as in Dhillon, "Current inverse iteration
// software can fail" (LAWN 120) / LAPACK dstein.
#include "kernels.h"

namespace tda {

constexpr int EIGH_THREADS = 512;
constexpr int EIGH_MAXM = 1024;

size_t eigh_work_doubles(int m) {
  // A copy (m*m) + tau (m) + LU scratch (5*m per root, up to 64 roots)
  return (size_t)m * m + (size_t)m + (size_t)5 * m * 64 + 64;
}

__device__ inline double sturm_count(const double* d, const double* e2, int m, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += (q < 0.0);
  for (int i = 1; i < m; ++i) {
    q = d[i] - x - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
  }
  return (double)cnt;
}

// deterministic start vector entries in (-1, 1)
__device__ inline double start_entry(int i, int r) {
  uint32_t h = (uint32_t)(i * 2654435761u) ^ (uint32_t)(r * 40503u + 0x9e3779b9u);
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return ((double)(h & 0xffffff) / 16777216.0) * 2.0 - 1.0 + 1e-3;
}

// Solve (T - lam I) x = b in place (b -> x) using a partially pivoted LU
// factorisation stored in lu[5*m]: u0 (diag), u1, u2 (super-diagonals), l
// (multiplier), piv (row swap flag).
__device__ void tri_lu(const double* d, const double* e, int m, double lam, double pivmin, double* lu) {
  double* u0 = lu;
  double* u1 = lu + m;
  double* u2 = lu + 2 * m;
  double* l = lu + 3 * m;
  double* piv = lu + 4 * m;
  double cd = d[0] - lam;                 // current row: diag
  double cs1 = (m > 1) ? e[0] : 0.0;      // current row: super 1
  double cs2 = 0.0;                       // current row: super 2
  for (int i = 0; i < m - 1; ++i) {
    const double sub = e[i];
    const double nd = d[i + 1] - lam;
    const double ns1 = (i + 1 < m - 1) ? e[i + 1] : 0.0;
    if (fabs(cd) >= fabs(sub)) {
      double p = cd;
      if (fabs(p) < pivmin) p = (p < 0.0) ? -pivmin : pivmin;
      const double mult = sub / p;
      u0[i] = p; u1[i] = cs1; u2[i] = cs2; l[i] = mult; piv[i] = 0.0;
      cd = nd - mult * cs1;
      cs1 = ns1 - mult * cs2;
      cs2 = 0.0;
    } else {
      const double mult = cd / sub;
      u0[i] = sub; u1[i] = nd; u2[i] = ns1; l[i] = mult; piv[i] = 1.0;
      cd = cs1 - mult * nd;
      cs1 = cs2 - mult * ns1;
      cs2 = 0.0;
    }
  }
  double p = cd;
  if (fabs(p) < pivmin) p = (p < 0.0) ? -pivmin : pivmin;
  u0[m - 1] = p; u1[m - 1] = 0.0; u2[m - 1] = 0.0;
}

__device__ void tri_solve(const double* lu, int m, double* x) {
  const double* u0 = lu;
  const double* u1 = lu + m;
  const double* u2 = lu + 2 * m;
  const double* l = lu + 3 * m;
  const double* piv = lu + 4 * m;
  // forward: apply row swaps and L
  for (int i = 0; i < m - 1; ++i) {
    if (piv[i] != 0.0) {
      const double t = x[i];
      x[i] = x[i + 1];
      x[i + 1] = t - l[i] * x[i];
    } else {
      x[i + 1] -= l[i] * x[i];
    }
  }
  // backward with U (three diagonals)
  for (int i = m - 1; i >= 0; --i) {
    double s = x[i];
    if (i + 1 < m) s -= u1[i] * x[i + 1];
    if (i + 2 < m) s -= u2[i] * x[i + 2];
    x[i] = s / u0[i];
  }
}

__global__ void __launch_bounds__(EIGH_THREADS) eigh_kernel(const double* __restrict__ H, int ldh, int m, int n,
                                                            double* __restrict__ work, double* __restrict__ evals,
                                                            double* __restrict__ Y, int ldy) {
  __shared__ double sv[EIGH_MAXM];   // Householder vector
  __shared__ double sw[EIGH_MAXM];   // p / w
  __shared__ double sd[EIGH_MAXM];   // diagonal of T
  __shared__ double se[EIGH_MAXM];   // off-diagonal of T
  __shared__ double se2[EIGH_MAXM];  // squares of off-diagonal
  __shared__ double red[EIGH_THREADS / 64];
  __shared__ double sval[4];
  __shared__ double slam[64];
  __shared__ int cl_start[65];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  constexpr int NW = EIGH_THREADS / 64;
  double* A = work;
  double* tau = work + (size_t)m * m;
  double* lus = tau + m;

  auto block_sum = [&](double v) -> double {
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += red[w];
    return s;
  };

  // 1. symmetric working copy from the upper triangle
  for (int64_t t = tid; t < (int64_t)m * m; t += EIGH_THREADS) {
    const int i = (int)(t / m), j = (int)(t % m);
    A[t] = (i <= j) ? H[(int64_t)i * ldh + j] : H[(int64_t)j * ldh + i];
  }
  __syncthreads();

  // 2. Householder tridiagonalisation (lower), A[k][i] == A[i][k]
  for (int k = 0; k < m - 2; ++k) {
    const double* rowk = A + (int64_t)k * m;
    double part = 0.0;
    for (int i = k + 2 + tid; i < m; i += EIGH_THREADS) part += rowk[i] * rowk[i];
    const double sigma = block_sum(part);
    const double alpha = rowk[k + 1];
    if (tid == 0) sd[k] = rowk[k];
    if (sigma == 0.0) {
      if (tid == 0) { se[k] = alpha; tau[k] = 0.0; }
      __syncthreads();
      continue;
    }
    const double beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
    const double tk = (beta - alpha) / beta;
    const double sc = 1.0 / (alpha - beta);
    for (int i = k + 1 + tid; i < m; i += EIGH_THREADS) sv[i] = (i == k + 1) ? 1.0 : rowk[i] * sc;
    __syncthreads();
    // p = tk * A22 v
    double pv = 0.0;
    for (int i = k + 1 + tid; i < m; i += EIGH_THREADS) {
      double acc = 0.0;
      for (int j = k + 1; j < m; ++j) acc += A[(int64_t)j * m + i] * sv[j];
      acc *= tk;
      sw[i] = acc;
      pv += acc * sv[i];
    }
    const double ptv = block_sum(pv);
    const double K = 0.5 * tk * ptv;
    for (int i = k + 1 + tid; i < m; i += EIGH_THREADS) sw[i] -= K * sv[i];
    __syncthreads();
    // A22 -= v w^T + w v^T
    const int nt = m - k - 1;
    for (int64_t t = tid; t < (int64_t)nt * nt; t += EIGH_THREADS) {
      const int i = k + 1 + (int)(t / nt), j = k + 1 + (int)(t % nt);
      A[(int64_t)i * m + j] -= sv[i] * sw[j] + sw[i] * sv[j];
    }
    __syncthreads();
    // keep v (below k+1) in row k for the back-transformation
    for (int i = k + 2 + tid; i < m; i += EIGH_THREADS) A[(int64_t)k * m + i] = sv[i];
    if (tid == 0) { se[k] = beta; tau[k] = tk; }
    __syncthreads();
  }
  if (tid == 0) {
    if (m >= 2) {
      sd[m - 2] = A[(int64_t)(m - 2) * m + (m - 2)];
      se[m - 2] = A[(int64_t)(m - 1) * m + (m - 2)];
    }
    sd[m - 1] = A[(int64_t)(m - 1) * m + (m - 1)];
    for (int k = m - 2; k < m; ++k)
      if (k >= 0) tau[k] = 0.0;
  }
  __syncthreads();
  for (int i = tid; i < m - 1; i += EIGH_THREADS) se2[i] = se[i] * se[i];
  __syncthreads();

  // Gershgorin bounds, norm, pivmin
  if (tid == 0) {
    double lo = 1e300, hi = -1e300, tn = 0.0, emax = 0.0;
    for (int i = 0; i < m; ++i) {
      const double r = (i > 0 ? fabs(se[i - 1]) : 0.0) + (i < m - 1 ? fabs(se[i]) : 0.0);
      lo = fmin(lo, sd[i] - r);
      hi = fmax(hi, sd[i] + r);
      tn = fmax(tn, fabs(sd[i]) + r);
      if (i < m - 1) emax = fmax(emax, se2[i]);
    }
    const double pad = 2.0 * 2.2e-16 * fmax(tn, 1e-300) + 1e-300;
    sval[0] = lo - pad;
    sval[1] = hi + pad;
    sval[2] = tn;
    sval[3] = fmax(2.2e-300, 2.2e-16 * 2.2e-16 * fmax(emax, 1e-300));  // pivmin
  }
  __syncthreads();
  const double glo = sval[0], ghi = sval[1], tnorm = sval[2], pivmin = sval[3];

  // 3. eigenvalues: one wave per root, 64-way multisection
  for (int r = wave; r < n; r += NW) {
    double a = glo, b = ghi;
    for (int round = 0; round < 40; ++round) {
      const double x = a + (b - a) * (double)(lane + 1) / 65.0;
      const double c = sturm_count(sd, se2, m, x, pivmin);
      // new a: largest x with count <= r ; new b: smallest x with count >= r+1
      double na = (c <= (double)r) ? x : -1e300;
      double nb = (c >= (double)(r + 1)) ? x : 1e300;
      for (int off = 32; off > 0; off >>= 1) {
        na = fmax(na, __shfl_xor(na, off, 64));
        nb = fmin(nb, __shfl_xor(nb, off, 64));
      }
      const double a2 = fmax(a, na), b2 = fmin(b, nb);
      const bool done = (b2 - a2) <= 2.2e-16 * fmax(fabs(a2), fabs(b2)) + pivmin || (a2 == a && b2 == b);
      a = a2;
      b = b2;
      if (done) break;
    }
    if (lane == 0) evals[r] = 0.5 * (a + b);
  }
  __syncthreads();

  // 4. eigenvectors of T by inverse iteration; cluster leaders run serially
  //    over their cluster (ortol = 1e-3 * ||T||, LAPACK dstein's choice).
  if (tid == 0) {
    int nc = 0;
    for (int r = 0; r < n; ++r) {
      if (r == 0 || evals[r] - evals[r - 1] > 1e-3 * tnorm) cl_start[nc++] = r;
    }
    cl_start[nc] = n;
    sval[0] = (double)nc;
  }
  __syncthreads();
  const int nclus = (int)sval[0];
  for (int c = tid; c < nclus; c += EIGH_THREADS) {
    double* lu = lus + (size_t)5 * m * (c % 64);
    for (int r = cl_start[c]; r < cl_start[c + 1]; ++r) {
      // perturb coincident eigenvalues within the cluster slightly
      double lam = evals[r];
      if (r > cl_start[c] && lam - evals[r - 1] < 10.0 * 2.2e-16 * tnorm)
        lam = evals[r - 1] + 10.0 * 2.2e-16 * tnorm;
      tri_lu(sd, se, m, lam, pivmin, lu);
      double* z = Y + (size_t)r * ldy;
      double nrm = 0.0;
      for (int i = 0; i < m; ++i) { z[i] = start_entry(i, r); nrm += z[i] * z[i]; }
      nrm = 1.0 / sqrt(nrm);
      for (int i = 0; i < m; ++i) z[i] *= nrm;
      for (int it = 0; it < 4; ++it) {
        tri_solve(lu, m, z);
        for (int s = cl_start[c]; s < r; ++s) {
          const double* zs = Y + (size_t)s * ldy;
          double dt = 0.0;
          for (int i = 0; i < m; ++i) dt += zs[i] * z[i];
          for (int i = 0; i < m; ++i) z[i] -= dt * zs[i];
        }
        double mx = 0.0;
        for (int i = 0; i < m; ++i) mx = fmax(mx, fabs(z[i]));
        if (mx == 0.0) { z[r % m] = 1.0; mx = 1.0; }
        double s2 = 0.0;
        for (int i = 0; i < m; ++i) { z[i] /= mx; s2 += z[i] * z[i]; }
        const double inv = 1.0 / sqrt(s2);
        for (int i = 0; i < m; ++i) z[i] *= inv;
      }
    }
  }
  __syncthreads();

  // 5. back-transformation y = H_0 H_1 ... H_{m-3} z, one wave per root
  for (int r = wave; r < n; r += NW) {
    double* y = Y + (size_t)r * ldy;
    for (int k = m - 3; k >= 0; --k) {
      const double tk = tau[k];
      if (tk == 0.0) continue;
      const double* vk = A + (int64_t)k * m;  // v[k+1] = 1 implicit, v[i>k+1] stored
      double part = 0.0;
      for (int i = k + 1 + lane; i < m; i += 64) part += ((i == k + 1) ? 1.0 : vk[i]) * y[i];
      const double dt = wave_sum(part) * tk;
      for (int i = k + 1 + lane; i < m; i += 64) y[i] -= dt * ((i == k + 1) ? 1.0 : vk[i]);
    }
  }
}

void launch_eigh(const double* H, int ldh, int m, int n, double* work, double* evals, double* Y, int ldy,
                 hipStream_t s) {
  TDA_CHECK(m >= 1 && m <= EIGH_MAXM, 2, "eigh: subspace dimension out of range (1..1024)");
  TDA_CHECK(n >= 1 && n <= m && n <= 64, 2, "eigh: number of roots out of range");
  TDA_CHECK(ldy >= m, 2, "eigh: ldy < m");
  eigh_kernel<<<1, EIGH_THREADS, 0, s>>>(H, ldh, m, n, work, evals, Y, ldy);
  TDA_LAUNCH_CHECK();
}

}  // namespace tda
