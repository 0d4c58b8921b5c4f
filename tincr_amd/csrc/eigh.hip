// FP64 symmetric eigensolver for the Davidson subspace (davidson.rs:93,
// `a_proj.eigh(UPLO::Upper)` followed by taking the n lowest pairs, :96-97).
//
// One workgroup of 512 threads per problem (m <= 1024):
//   1. Householder reduction to tridiagonal form T = Q^T H Q.  Each step does
//      a symmetric mat-vec and a rank-2 update of the trailing block, one wave
//      per row (coalesced along the row).  The trailing block lives in global
//      memory (L2-resident) until it fits the LDS pool, then in LDS.
//   2. The n lowest eigenvalues of T by Sturm-count multisection: a wave
//      evaluates 64 shifts at once, so each round narrows the bracket 65x.
//   3. Eigenvectors of T by inverse iteration with a partially pivoted
//      tridiagonal LU (as LAPACK dstein), modified Gram-Schmidt inside
//      clusters of close eigenvalues; factors and iterate live in LDS slots.
//   4. Back-transformation y = H_0 ... H_{m-3} z with the vectors in LDS.
// Published methods: Householder tridiagonalisation and Sturm bisection as in
// Golub & Van Loan, "Matrix Computations" 8.3 / 8.4; inverse iteration with
// cluster reorthogonalisation as in LAPACK dstein (Dhillon, LAWN 120).
#include "kernels.h"

namespace tda {

constexpr int EIGH_THREADS = 512;
constexpr int EIGH_NW = EIGH_THREADS / 64;
constexpr int EIGH_MAXM = 1024;
constexpr int EIGH_POOL = 14336;  // doubles of LDS pool (112 KiB)

size_t eigh_work_doubles(int m) {
  // A copy (m*m) + tau (m) + spill area for vectors (64*m) + slots (6*m*64)
  return (size_t)m * m + (size_t)m + (size_t)64 * m + (size_t)6 * m * 64 + 64;
}

__device__ inline int sturm_count(const double* d, const double* e2, int m, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += (q < 0.0);
  for (int i = 1; i < m; ++i) {
    q = d[i] - x - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
  }
  return cnt;
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

// Partially pivoted LU of (T - lam I): ri = reciprocal of the U diagonal,
// u1, u2 = U super-diagonals, l = multipliers, piv = +1 (no swap) / -1 (swap).
__device__ void tri_lu(const double* d, const double* e, int m, double lam, double pivmin, double* ri, double* u1,
                       double* u2, double* l, double* piv) {
  double cd = d[0] - lam;
  double cs1 = (m > 1) ? e[0] : 0.0;
  double cs2 = 0.0;
  for (int i = 0; i < m - 1; ++i) {
    const double sub = e[i];
    const double nd = d[i + 1] - lam;
    const double ns1 = (i + 1 < m - 1) ? e[i + 1] : 0.0;
    if (fabs(cd) >= fabs(sub)) {
      double p = cd;
      if (fabs(p) < pivmin) p = (p < 0.0) ? -pivmin : pivmin;
      const double rp = 1.0 / p;
      const double mult = sub * rp;
      ri[i] = rp; u1[i] = cs1; u2[i] = cs2; l[i] = mult; piv[i] = 1.0;
      cd = nd - mult * cs1;
      cs1 = ns1 - mult * cs2;
      cs2 = 0.0;
    } else {
      const double rp = 1.0 / sub;
      const double mult = cd * rp;
      ri[i] = rp; u1[i] = nd; u2[i] = ns1; l[i] = mult; piv[i] = -1.0;
      cd = cs1 - mult * nd;
      cs1 = cs2 - mult * ns1;
      cs2 = 0.0;
    }
  }
  double p = cd;
  if (fabs(p) < pivmin) p = (p < 0.0) ? -pivmin : pivmin;
  ri[m - 1] = 1.0 / p; u1[m - 1] = 0.0; u2[m - 1] = 0.0;
}

__device__ void tri_solve(const double* ri, const double* u1, const double* u2, const double* l, const double* piv,
                          int m, double* x) {
  for (int i = 0; i < m - 1; ++i) {
    if (piv[i] < 0.0) {
      const double t = x[i];
      x[i] = x[i + 1];
      x[i + 1] = t - l[i] * x[i];
    } else {
      x[i + 1] -= l[i] * x[i];
    }
  }
  double x1 = 0.0, x2 = 0.0;  // x[i+1], x[i+2]
  for (int i = m - 1; i >= 0; --i) {
    const double v = (x[i] - u1[i] * x1 - u2[i] * x2) * ri[i];
    x[i] = v;
    x2 = x1;
    x1 = v;
  }
}

__global__ void __launch_bounds__(EIGH_THREADS) eigh_kernel(const double* __restrict__ H, int ldh, int m, int n,
                                                            double* __restrict__ work, double* __restrict__ evals,
                                                            double* __restrict__ Y, int ldy) {
  __shared__ double pool[EIGH_POOL];
  __shared__ double sv[EIGH_MAXM];   // Householder vector
  __shared__ double sw[EIGH_MAXM];   // p, then w
  __shared__ double sd[EIGH_MAXM];   // diagonal of T
  __shared__ double se[EIGH_MAXM];   // off-diagonal of T
  __shared__ double se2[EIGH_MAXM];  // squares of off-diagonal
  __shared__ double red[EIGH_NW];
  __shared__ double sval[4];
  __shared__ int cl_start[65];
  __shared__ int snclus;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  double* A = work;
  double* tau = work + (size_t)m * m;

  auto block_sum = [&](double v) -> double {
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < EIGH_NW; ++w) s += red[w];
    return s;
  };

  // 1. symmetric working copy from the upper triangle
  for (int64_t t = tid; t < (int64_t)m * m; t += EIGH_THREADS) {
    const int i = (int)(t / m), j = (int)(t % m);
    A[t] = (i <= j) ? H[(int64_t)i * ldh + j] : H[(int64_t)j * ldh + i];
  }
  __syncthreads();

  // 2. Householder tridiagonalisation; element (i, j) of the trailing matrix
  //    is Am[(i - k0) * lda + (j - k0)].
  double* Am = A;
  int k0 = 0, lda = m;
  bool in_lds = false;
  for (int k = 0; k < m - 2; ++k) {
    if (!in_lds && (m - k) * (m - k) <= EIGH_POOL) {
      // move the trailing block rows/cols k..m-1 into LDS
      const int nb = m - k;
      for (int t = tid; t < nb * nb; t += EIGH_THREADS) {
        const int i = t / nb, j = t % nb;
        pool[t] = Am[(int64_t)(k + i - k0) * lda + (k + j - k0)];
      }
      __syncthreads();
      Am = pool;
      k0 = k;
      lda = nb;
      in_lds = true;
    }
    const double* rowk = Am + (int64_t)(k - k0) * lda - k0;  // rowk[j] = A(k, j)
    double part = 0.0;
    for (int j = k + 2 + tid; j < m; j += EIGH_THREADS) part += rowk[j] * rowk[j];
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
    for (int j = k + 1 + tid; j < m; j += EIGH_THREADS) {
      const double vj = (j == k + 1) ? 1.0 : rowk[j] * sc;
      sv[j] = vj;
      if (j >= k + 2) A[(int64_t)k * m + j] = vj;  // kept for the back-transformation
    }
    __syncthreads();
    // p = tk * A22 v by columns: p_i = sum_j A(j, i) v_j, thread per i, the
    // j loop unrolled so that loads (coalesced over i) overlap
    double pv = 0.0;
    for (int i = k + 1 + tid; i < m; i += EIGH_THREADS) {
      const double* ci = Am + (i - k0) - (int64_t)k0 * lda;  // ci[j * lda] = A(j, i)
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int j = k + 1;
      for (; j + 3 < m; j += 4) {
        a0 += ci[(int64_t)j * lda] * sv[j];
        a1 += ci[(int64_t)(j + 1) * lda] * sv[j + 1];
        a2 += ci[(int64_t)(j + 2) * lda] * sv[j + 2];
        a3 += ci[(int64_t)(j + 3) * lda] * sv[j + 3];
      }
      for (; j < m; ++j) a0 += ci[(int64_t)j * lda] * sv[j];
      const double pi = ((a0 + a1) + (a2 + a3)) * tk;
      sw[i] = pi;
      pv += pi * sv[i];
    }
    const double ptv = block_sum(pv);
    const double K = 0.5 * tk * ptv;
    for (int j = k + 1 + tid; j < m; j += EIGH_THREADS) sw[j] -= K * sv[j];
    __syncthreads();
    // A22 -= v w^T + w v^T, one wave per row
    for (int i = k + 1 + wave; i < m; i += EIGH_NW) {
      double* ri = Am + (int64_t)(i - k0) * lda - k0;
      const double vi = sv[i], wi = sw[i];
      for (int j = k + 1 + lane; j < m; j += 64) ri[j] -= vi * sw[j] + wi * sv[j];
    }
    if (tid == 0) { se[k] = beta; tau[k] = tk; }
    __syncthreads();
  }
  if (tid == 0) {
    auto at = [&](int i, int j) { return Am[(int64_t)(i - k0) * lda + (j - k0)]; };
    if (m >= 2) {
      sd[m - 2] = at(m - 2, m - 2);
      se[m - 2] = at(m - 1, m - 2);
    }
    sd[m - 1] = at(m - 1, m - 1);
    for (int k = (m >= 2 ? m - 2 : 0); k < m; ++k) tau[k] = 0.0;
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
  for (int r = wave; r < n; r += EIGH_NW) {
    double a = glo, b = ghi;
    for (int round = 0; round < 40; ++round) {
      const double x = a + (b - a) * (double)(lane + 1) / 65.0;
      const int c = sturm_count(sd, se2, m, x, pivmin);
      double na = (c <= r) ? x : -1e300;
      double nb = (c >= r + 1) ? x : 1e300;
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

  // 4. eigenvectors of T by inverse iteration.  With eigenvalues bisected to
  //    full precision, vectors of roots separated by gap come out orthogonal
  //    to ~eps*||T||/gap; roots closer than 1e-9*||T|| form clusters that one
  //    thread reorthogonalises serially (modified Gram-Schmidt).
  if (tid == 0) {
    int nc = 0;
    for (int r = 0; r < n; ++r)
      if (r == 0 || evals[r] - evals[r - 1] > 1e-9 * tnorm) cl_start[nc++] = r;
    cl_start[nc] = n;
    snclus = nc;
  }
  __syncthreads();
  const int nclus = snclus;
  // LDS slots of 6*m doubles (u0, u1, u2, l, piv, z); spill to global if needed
  const int slot_sz = 6 * m;
  const int nslots_lds = EIGH_POOL / slot_sz;
  double* gslots = tau + m + (size_t)64 * m;
  for (int c = tid; c < nclus; c += EIGH_THREADS) {
    double* sl = (c < nslots_lds) ? pool + (size_t)c * slot_sz : gslots + (size_t)c * slot_sz;
    double *u0 = sl, *u1 = sl + m, *u2 = sl + 2 * m, *l = sl + 3 * m, *piv = sl + 4 * m, *z = sl + 5 * m;
    for (int r = cl_start[c]; r < cl_start[c + 1]; ++r) {
      double lam = evals[r];
      if (r > cl_start[c] && lam - evals[r - 1] < 10.0 * 2.2e-16 * tnorm) lam = evals[r - 1] + 10.0 * 2.2e-16 * tnorm;
      tri_lu(sd, se, m, lam, pivmin, u0, u1, u2, l, piv);
      double nrm = 0.0;
      for (int i = 0; i < m; ++i) { z[i] = start_entry(i, r); nrm += z[i] * z[i]; }
      nrm = 1.0 / sqrt(nrm);
      for (int i = 0; i < m; ++i) z[i] *= nrm;
      for (int it = 0; it < 2; ++it) {
        tri_solve(u0, u1, u2, l, piv, m, z);
        for (int s = cl_start[c]; s < r; ++s) {
          const double* zs = Y + (size_t)s * ldy;
          double dt = 0.0;
          for (int i = 0; i < m; ++i) dt += zs[i] * z[i];
          for (int i = 0; i < m; ++i) z[i] -= dt * zs[i];
        }
        double s2 = 0.0;
        for (int i = 0; i < m; ++i) s2 += z[i] * z[i];
        if (!(s2 > 0.0) || !isfinite(s2)) {  // rescale by the max entry first
          double mx = 0.0;
          for (int i = 0; i < m; ++i) mx = fmax(mx, fabs(z[i]));
          if (!(mx > 0.0) || !isfinite(mx)) {
            for (int i = 0; i < m; ++i) z[i] = 0.0;
            z[r % m] = 1.0;
            mx = 1.0;
          }
          const double im = 1.0 / mx;
          s2 = 0.0;
          for (int i = 0; i < m; ++i) { z[i] *= im; s2 += z[i] * z[i]; }
        }
        const double inv = 1.0 / sqrt(s2);
        for (int i = 0; i < m; ++i) z[i] *= inv;
      }
      double* y = Y + (size_t)r * ldy;
      for (int i = 0; i < m; ++i) y[i] = z[i];
    }
  }
  __syncthreads();

  // 5. back-transformation y = H_0 H_1 ... H_{m-3} z.  Each wave applies the
  //    reflectors to its roots (r = wave + 8t) held in LDS when they fit.
  const bool y_lds = (size_t)n * m <= (size_t)EIGH_POOL;
  double* yb = y_lds ? pool : Y;
  const int ldb = y_lds ? m : ldy;
  if (y_lds) {
    for (int t = tid; t < n * m; t += EIGH_THREADS) pool[t] = Y[(size_t)(t / m) * ldy + (t % m)];
    __syncthreads();
  }
  constexpr int RPW = 8;  // roots per wave per pass
  for (int r0 = wave; r0 < n; r0 += EIGH_NW * RPW) {
    int nr = 0;
    int rs[RPW];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int r = r0 + q * EIGH_NW;
      rs[q] = r;
      if (r < n) nr = q + 1;
    }
    for (int k = m - 3; k >= 0; --k) {
      const double tk = tau[k];
      if (tk == 0.0) continue;
      const double* vk = A + (int64_t)k * m;  // v[k+1] = 1 implicit, v[i>k+1] stored
      double dts[RPW];
#pragma unroll
      for (int q = 0; q < RPW; ++q) dts[q] = 0.0;
      for (int i = k + 1 + lane; i < m; i += 64) {
        const double vi = (i == k + 1) ? 1.0 : vk[i];
#pragma unroll
        for (int q = 0; q < RPW; ++q)
          if (q < nr) dts[q] += vi * yb[(size_t)rs[q] * ldb + i];
      }
#pragma unroll
      for (int q = 0; q < RPW; ++q)
        if (q < nr) dts[q] = wave_sum(dts[q]) * tk;
      for (int i = k + 1 + lane; i < m; i += 64) {
        const double vi = (i == k + 1) ? 1.0 : vk[i];
#pragma unroll
        for (int q = 0; q < RPW; ++q)
          if (q < nr) yb[(size_t)rs[q] * ldb + i] -= dts[q] * vi;
      }
    }
  }
  if (y_lds) {
    __syncthreads();
    for (int t = tid; t < n * m; t += EIGH_THREADS) Y[(size_t)(t / m) * ldy + (t % m)] = pool[t];
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
