// FP64 symmetric eigensolver for the Davidson subspace (davidson.rs:93,
// `a_proj.eigh(UPLO::Upper)` followed by taking the n lowest pairs, :96-97).
//
// One workgroup of 1024 threads per problem (m <= 1024):
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

constexpr int EIGH_THREADS = 1024;
constexpr int EIGH_NW = EIGH_THREADS / 64;
constexpr int EIGH_MAXM = 1024;
constexpr int EIGH_POOL = 12288;  // doubles of LDS pool (96 KiB)
// Subspaces up to 256 (15 n for n <= 17 roots): the per-index LDS arrays
// shrink to 256 entries and the pool grows to 144 KiB (trailing block in LDS
// from 128 rows instead of 101, more inverse-iteration slots, the
// back-transformation in LDS for more (m, n)).  Measured at m = 240, 16
// roots: 2.82 ms instead of 2.85; a tridiagonalisation step costs about the
// same from LDS as from L2 (both are one CU's bandwidth plus 3-4 barriers).
constexpr int EIGH_MAXM_SMALL = 256;
constexpr int EIGH_POOL_SMALL = 18432;

size_t eigh_work_doubles(int m) {
  // A copy (m*m) + tau (m) + spill area for vectors (64*m) + slots (6*m*64)
  return (size_t)m * m + (size_t)m + (size_t)64 * m + (size_t)6 * m * 64 + 64;
}

// 1 / q from the hardware reciprocal refined by two Newton steps
// (r <- r + r (1 - q r)): a relative error of a few ulp, against the ~10
// dependent instructions of the correctly rounded division.  A Sturm count
// only needs each q_i to be that of slightly perturbed data.
__device__ inline double sturm_recip(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  return fma(fma(-q, r, 1.0), r, r);
}

__device__ inline int sturm_count(const double* d, const double* e2, int m, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += (q < 0.0);
  int i = 1;
  // four steps per trip with their LDS loads issued ahead of the divisions
  for (; i + 4 <= m; i += 4) {
    const double d0 = d[i] - x, d1 = d[i + 1] - x, d2 = d[i + 2] - x, d3 = d[i + 3] - x;
    const double f0 = e2[i - 1], f1 = e2[i], f2 = e2[i + 1], f3 = e2[i + 2];
    q = d0 - f0 * sturm_recip(q);
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
    q = d1 - f1 * sturm_recip(q);
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
    q = d2 - f2 * sturm_recip(q);
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
    q = d3 - f3 * sturm_recip(q);
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
  }
  for (; i < m; ++i) {
    q = d[i] - x - e2[i - 1] * sturm_recip(q);
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

// Partially pivoted LU of (T - lam I) with the forward substitution (row
// swaps and L) applied to x on the fly: ri = reciprocal of the U diagonal,
// u1, u2 = U super-diagonals; then back substitution with U.  x[i] of the
// forward pass is carried in a register.
__device__ void tri_lu_solve(const double* d, const double* e, int m, double lam, double pivmin, double* ri,
                             double* u1, double* u2, double* x) {
  double cd = d[0] - lam;
  double cs1 = (m > 1) ? e[0] : 0.0;
  double cs2 = 0.0;
  double xi = x[0];
  for (int i = 0; i < m - 1; ++i) {
    const double sub = e[i];
    const double nd = d[i + 1] - lam;
    const double ns1 = (i + 1 < m - 1) ? e[i + 1] : 0.0;
    const double xn = x[i + 1];
    if (fabs(cd) >= fabs(sub)) {
      double p = cd;
      if (fabs(p) < pivmin) p = (p < 0.0) ? -pivmin : pivmin;
      const double rp = 1.0 / p;
      const double mult = sub * rp;
      ri[i] = rp; u1[i] = cs1; u2[i] = cs2;
      cd = nd - mult * cs1;
      cs1 = ns1 - mult * cs2;
      cs2 = 0.0;
      x[i] = xi;
      xi = xn - mult * xi;
    } else {
      const double rp = 1.0 / sub;
      const double mult = cd * rp;
      ri[i] = rp; u1[i] = nd; u2[i] = ns1;
      cd = cs1 - mult * nd;
      cs1 = cs2 - mult * ns1;
      cs2 = 0.0;
      x[i] = xn;
      xi = xi - mult * xn;
    }
  }
  double p = cd;
  if (fabs(p) < pivmin) p = (p < 0.0) ? -pivmin : pivmin;
  ri[m - 1] = 1.0 / p; u1[m - 1] = 0.0; u2[m - 1] = 0.0;
  x[m - 1] = xi;
  double x1 = 0.0, x2 = 0.0;  // x[i+1], x[i+2]
  for (int i = m - 1; i >= 0; --i) {
    const double v = (x[i] - u1[i] * x1 - u2[i] * x2) * ri[i];
    x[i] = v;
    x2 = x1;
    x1 = v;
  }
}

// orders the LDS / global accesses of one wave (lanes exchange data)
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

template <int MAXM, int POOL>
__device__ __forceinline__ void eigh_body(const double* __restrict__ H, int ldh, int m, int n,
                                          double* __restrict__ work, double* __restrict__ evals,
                                          double* __restrict__ Y, int ldy, unsigned long long* __restrict__ clk) {
#define EIGH_STAMP(i) \
  if (clk && threadIdx.x == 0) clk[i] = wall_clock64();
  __shared__ double pool[POOL];
  __shared__ double sv[MAXM];   // pending Householder vector v_{k-1}
  __shared__ double sw[MAXM];   // pending w_{k-1}
  __shared__ double snv[MAXM];  // current Householder vector
  __shared__ double snp[MAXM];  // current p
  __shared__ double sd[MAXM];   // diagonal of T
  __shared__ double se[MAXM];   // off-diagonal of T
  __shared__ double se2[MAXM];  // squares of off-diagonal
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

  EIGH_STAMP(0)
  // 1. symmetric working copy from the upper triangle
  for (int64_t t = tid; t < (int64_t)m * m; t += EIGH_THREADS) {
    const int i = (int)(t / m), j = (int)(t % m);
    A[t] = (i <= j) ? H[(int64_t)i * ldh + j] : H[(int64_t)j * ldh + i];
  }
  __syncthreads();

  EIGH_STAMP(1)
  // 2. Householder tridiagonalisation (lower; A(k, j) == A(j, k)).  Element
  //    (i, j) of the working matrix is Am[(i - k0) * lda + (j - k0)].  The
  //    rank-2 update A -= v w^T + w v^T of step k is applied in two places:
  //    to row k+1 right away (together with the norm of its tail, which the
  //    next reflector needs), and to the rest of the trailing block inside
  //    step k+1's mat-vec pass, so each step reads and writes the trailing
  //    block once.  Mat-vec in column form: thread (i, seg) sums A(j, i) v_j
  //    over its segment of j (coalesced over i), partial sums combined
  //    through LDS.  Three or four barriers per step.
  double* Am = A;
  int k0 = 0, lda = m;
  bool in_lds = false;
  double* part = pool + POOL - 2048;  // segment partial sums (global phase only)
  __shared__ double red2[2][EIGH_NW];
  int rb = 0;
  // one-barrier block sum (alternating buffers)
  auto block_sum1 = [&](double v) -> double {
    v = wave_sum(v);
    if (lane == 0) red2[rb][wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < EIGH_NW; ++w) s += red2[rb][w];
    rb ^= 1;
    return s;
  };
  for (int j = tid; j < m; j += EIGH_THREADS) {
    sv[j] = 0.0;
    sw[j] = 0.0;
  }
  double sigma = 0.0;  // |A(k, k+2:)|^2 of the current row k
  {
    double ps = 0.0;
    for (int j = 2 + tid; j < m; j += EIGH_THREADS) ps += A[j] * A[j];
    sigma = block_sum1(ps);  // also orders the sv / sw clearing
  }
  for (int k = 0; k < m - 2; ++k) {
    const int nt = m - k - 1;  // trailing size (rows/cols k+1..m-1)
    if (!in_lds && (m - k) * (m - k) <= POOL - 2048) {
      // move rows/cols k..m-1 into LDS (the pending update travels along)
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
    const double* rowk = Am + (int64_t)(k - k0) * lda - k0;  // rowk[j] = A(k, j), up to date
    const double alpha = rowk[k + 1];
    if (tid == 0) sd[k] = rowk[k];
    double tk = 0.0, beta = alpha, sc = 0.0;
    if (sigma != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
      tk = (beta - alpha) / beta;
      sc = 1.0 / (alpha - beta);
    }
    // Householder vector into snv (kept in global row k for the back-transformation)
    for (int j = k + 1 + tid; j < m; j += EIGH_THREADS) {
      const double vj = (j == k + 1) ? 1.0 : rowk[j] * sc;
      snv[j] = vj;
      if (j >= k + 2) A[(int64_t)k * m + j] = vj;
    }
    __syncthreads();  // B1: v complete
    // fused pass over the trailing block: apply the pending (v_{k-1}, w_{k-1}), then p = A v
    const int segs = max(1, min(8, EIGH_THREADS / nt));
    const int i = k + 1 + tid % nt;
    double acc = 0.0;
    {
      const int seg = tid / nt;
      // The pass is written once and instantiated for the LDS block and for
      // the global one: through the one generic pointer it compiles to flat
      // accesses, whose LDS waits also wait for the global loads in flight.
      auto pass = [&](double* __restrict__ base) {
        double* ci = base + (i - k0) - (int64_t)k0 * lda;  // ci[j * lda] = A(j, i)
        const double vpi = sv[i], wpi = sw[i];
        const int len = (nt + segs - 1) / segs;
        const int j0 = k + 1 + seg * len, j1 = min(m, j0 + len);
        // 16 independent loads in flight per thread: the global phase is
        // latency-bound (one CU streams the whole trailing block)
        int j = j0;
        for (; j + 16 <= j1; j += 16) {
          double a[16];
#pragma unroll
          for (int q = 0; q < 16; ++q) a[q] = ci[(int64_t)(j + q) * lda];
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            a[q] -= sv[j + q] * wpi + sw[j + q] * vpi;
            ci[(int64_t)(j + q) * lda] = a[q];
            acc += a[q] * snv[j + q];
          }
        }
        for (; j < j1; ++j) {
          double a = ci[(int64_t)j * lda] - (sv[j] * wpi + sw[j] * vpi);
          ci[(int64_t)j * lda] = a;
          acc += a * snv[j];
        }
        if (segs > 1) part[seg * nt + (i - k - 1)] = acc;
      };
      if (seg < segs) {
        if (in_lds) pass(pool);
        else pass(A);
      }
      if (segs > 1) {
        __syncthreads();  // B2: partial sums complete
        if (tid < nt) {
          double s2 = 0.0;
          for (int q = 0; q < segs; ++q) s2 += part[q * nt + tid];
          acc = s2;
        }
      }
    }
    // p_i = tk * (A v)_i for i = k+1+tid (tid < nt); p . v by block sum
    double pi = 0.0, vi = 0.0;
    if (tid < nt) {
      pi = acc * tk;
      vi = snv[i];
      if (tid == 0) snp[k + 1] = pi;
    }
    const double ptv = block_sum1(pi * vi);  // B3: also publishes snp[k+1]; the pass is over
    const double K = 0.5 * tk * ptv;
    if (tid == 0) {
      se[k] = beta;
      tau[k] = tk;
    }
    // new pending (v_k, w_k); row k+1 gets it now, with the norm of its tail
    double ps = 0.0;
    if (tid < nt) {
      const double wi = pi - K * vi;
      const double w1 = snp[k + 1] - K;  // w_{k+1} (v_{k+1} = 1)
      sv[i] = vi;
      sw[i] = wi;
      double* rowk1 = Am + (int64_t)(k + 1 - k0) * lda - k0;
      const double a = rowk1[i] - (wi + w1 * vi);
      rowk1[i] = a;
      if (i >= k + 3) ps = a * a;
    }
    sigma = block_sum1(ps);  // B4: row k+1, sv, sw complete
  }
  if (tid == 0) {
    // rows m-2 and m-1: row m-2 is up to date; A(m-1, m-1) still lacks (v_{m-3}, w_{m-3})
    auto at = [&](int i2, int j2) { return Am[(int64_t)(i2 - k0) * lda + (j2 - k0)]; };
    if (m >= 2) {
      sd[m - 2] = at(m - 2, m - 2);
      se[m - 2] = at(m - 2, m - 1);
    }
    sd[m - 1] = at(m - 1, m - 1) - 2.0 * sv[m - 1] * sw[m - 1];
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

  EIGH_STAMP(2)
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

  EIGH_STAMP(3)
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
  // one wave per cluster with an LDS slot of 4*m doubles (ri, u1, u2, z):
  // lane 0 runs the LU recurrences, the wave does the vector operations
  const int slot_sz = 4 * m;
  const int nslots = max(1, min(EIGH_NW, POOL / slot_sz));
  if (wave < nslots) {
    double* sl = pool + (size_t)wave * slot_sz;
    double *ri = sl, *u1 = sl + m, *u2 = sl + 2 * m, *z = sl + 3 * m;
    for (int c = wave; c < nclus; c += nslots) {
      for (int r = cl_start[c]; r < cl_start[c + 1]; ++r) {
        double lam = evals[r];
        if (r > cl_start[c] && lam - evals[r - 1] < 10.0 * 2.2e-16 * tnorm)
          lam = evals[r - 1] + 10.0 * 2.2e-16 * tnorm;
        double nrm = 0.0;
        for (int i = lane; i < m; i += 64) {
          const double zi = start_entry(i, r);
          z[i] = zi;
          nrm += zi * zi;
        }
        nrm = 1.0 / sqrt(wave_sum(nrm));
        for (int i = lane; i < m; i += 64) z[i] *= nrm;
        wave_sync();
        if (lane == 0) tri_lu_solve(sd, se, m, lam, pivmin, ri, u1, u2, z);
        wave_sync();
        for (int s2 = cl_start[c]; s2 < r; ++s2) {  // modified Gram-Schmidt inside the cluster
          const double* zs = Y + (size_t)s2 * ldy;
          double dt = 0.0;
          for (int i = lane; i < m; i += 64) dt += zs[i] * z[i];
          dt = wave_sum(dt);
          for (int i = lane; i < m; i += 64) z[i] -= dt * zs[i];
        }
        double s2 = 0.0;
        for (int i = lane; i < m; i += 64) s2 += z[i] * z[i];
        s2 = wave_sum(s2);
        if (!(s2 > 0.0) || !isfinite(s2)) {  // rescale by the max entry first
          double mx = 0.0;
          for (int i = lane; i < m; i += 64) mx = fmax(mx, fabs(z[i]));
          mx = wave_max(mx);
          if (!(mx > 0.0) || !isfinite(mx)) {
            for (int i = lane; i < m; i += 64) z[i] = (i == r % m) ? 1.0 : 0.0;
            mx = 1.0;
          }
          const double im = 1.0 / mx;
          s2 = 0.0;
          for (int i = lane; i < m; i += 64) {
            z[i] *= im;
            s2 += z[i] * z[i];
          }
          s2 = wave_sum(s2);
        }
        const double inv = 1.0 / sqrt(s2);
        double* y = Y + (size_t)r * ldy;
        for (int i = lane; i < m; i += 64) y[i] = z[i] * inv;
        __threadfence_block();  // later roots of the cluster read y
        wave_sync();
      }
    }
  }
  __syncthreads();

  EIGH_STAMP(4)
  // 5. back-transformation Y = H_0 H_1 ... H_{m-3} Z in blocks of BW
  //    reflectors: H_s ... H_{s+BW-1} = I - V T V^T (T upper triangular,
  //    forward accumulation as LAPACK dlarft), applied last block first:
  //    Z -= V (T (V^T Z)).  Z and the V panel live in LDS when they fit.
  constexpr int BW = 16;
  const int nref = max(0, m - 2);  // reflectors 0..m-3
  const bool fits = (size_t)n * m + (size_t)BW * m + 2 * BW * BW + (size_t)BW * n <= (size_t)POOL;
  if (fits && nref > 0) {
    double* Zs = pool;                       // [n][m]
    double* Vp = pool + (size_t)n * m;       // [BW][m]  (row q = reflector s+q)
    double* Tm = Vp + (size_t)BW * m;        // [BW][BW]
    double* Wm = Tm + BW * BW;               // [BW][n], then Gv [BW][BW]
    for (int t = tid; t < n * m; t += EIGH_THREADS) Zs[t] = Y[(size_t)(t / m) * ldy + (t % m)];
    const int nblk = (nref + BW - 1) / BW;
    for (int blk = nblk - 1; blk >= 0; --blk) {
      const int s0 = blk * BW;
      const int bw = min(BW, nref - s0);
      // panel: Vp[q][i] = v_{s0+q}[i]  (0 for i <= s0+q, 1 at s0+q+1)
      for (int t = tid; t < bw * m; t += EIGH_THREADS) {
        const int q = t / m, i = t % m, kk = s0 + q;
        Vp[t] = (i <= kk) ? 0.0 : (i == kk + 1) ? 1.0 : A[(int64_t)kk * m + i];
      }
      __syncthreads();
      // Gram matrix Gv[p][q] = v_p . v_q (p < q), one wave per pair
      double* Gv = Wm + BW * n;  // [BW][BW]
      for (int pq = wave; pq < bw * bw; pq += EIGH_NW) {
        const int p = pq / bw, q = pq % bw;
        if (p >= q) continue;
        double g = 0.0;
        for (int i = s0 + q + 1 + lane; i < m; i += 64) g += Vp[p * m + i] * Vp[q * m + i];
        g = wave_sum(g);
        if (lane == 0) Gv[p * BW + q] = g;
      }
      __syncthreads();
      // T (dlarft, forward): T[0:q, q] = -tau_q T[0:q, 0:q] Gv[0:q, q]; T[q][q] = tau_q
      if (wave == 0) {
        for (int q = 0; q < bw; ++q) {
          double tq = 0.0;
          if (lane < q)
            for (int r2 = lane; r2 < q; ++r2) tq += Tm[lane * BW + r2] * Gv[r2 * BW + q];
          if (lane < q) Tm[lane * BW + q] = -tau[s0 + q] * tq;
          if (lane == q) Tm[q * BW + q] = tau[s0 + q];
          if (lane > q && lane < BW) Tm[lane * BW + q] = 0.0;
        }
      }
      __syncthreads();
      // W = V^T Z  (bw x n)
      for (int t = tid; t < bw * n; t += EIGH_THREADS) {
        const int q = t / n, r = t % n;
        double acc = 0.0;
        for (int i = s0 + q + 1; i < m; ++i) acc += Vp[q * m + i] * Zs[(size_t)r * m + i];
        Wm[q * n + r] = acc;
      }
      __syncthreads();
      // W = T W  (upper triangular, in place via registers)
      double tw = 0.0;
      int tq2 = -1, tr2 = -1;
      if (tid < bw * n) {
        tq2 = tid / n;
        tr2 = tid % n;
        for (int p = tq2; p < bw; ++p) tw += Tm[tq2 * BW + p] * Wm[p * n + tr2];
      }
      __syncthreads();
      if (tid < bw * n) Wm[tq2 * n + tr2] = tw;
      __syncthreads();
      // Z -= V W
      for (int t = tid; t < n * m; t += EIGH_THREADS) {
        const int r = t / m, i = t % m;
        double acc = 0.0;
        for (int q = 0; q < bw; ++q) acc += Vp[q * m + i] * Wm[q * n + r];
        Zs[t] -= acc;
      }
      __syncthreads();
    }
    for (int t = tid; t < n * m; t += EIGH_THREADS) Y[(size_t)(t / m) * ldy + (t % m)] = Zs[t];
  } else if (nref > 0) {
    // large problems: one wave per root, reflectors one at a time (global memory)
    for (int r = wave; r < n; r += EIGH_NW) {
      double* y = Y + (size_t)r * ldy;
      for (int k = m - 3; k >= 0; --k) {
        const double tk = tau[k];
        if (tk == 0.0) continue;
        const double* vk = A + (int64_t)k * m;
        double part2 = 0.0;
        for (int i = k + 1 + lane; i < m; i += 64) part2 += ((i == k + 1) ? 1.0 : vk[i]) * y[i];
        const double dt = wave_sum(part2) * tk;
        for (int i = k + 1 + lane; i < m; i += 64) y[i] -= dt * ((i == k + 1) ? 1.0 : vk[i]);
      }
    }
  }
  __syncthreads();
  EIGH_STAMP(5)
#undef EIGH_STAMP
}

template <int MAXM, int POOL>
__global__ void __launch_bounds__(EIGH_THREADS) eigh_kernel(const double* __restrict__ H, int ldh, int m, int n,
                                                            double* __restrict__ work, double* __restrict__ evals,
                                                            double* __restrict__ Y, int ldy,
                                                            unsigned long long* __restrict__ clk) {
  eigh_body<MAXM, POOL>(H, ldh, m, n, work, evals, Y, ldy, clk);
}

// Many independent subspace problems (the fragments of a batch): workgroup b
// solves descriptor b < *count; the grid is fixed (a captured launch replays
// with new descriptors).
__global__ void __launch_bounds__(EIGH_THREADS) eigh_batched_kernel(const GroupSlot* __restrict__ slot) {
  if ((int)blockIdx.x >= slot->count) return;
  const EighDesc e = ((const EighDesc*)slot->descs)[blockIdx.x];
  eigh_body<EIGH_MAXM_SMALL, EIGH_POOL_SMALL>(e.H, e.ldh, e.m, e.n, e.work, e.evals, e.Y, e.ldy, nullptr);
}

void launch_eigh_batched(const GroupSlot* slot, int grid, hipStream_t s) {
  if (grid <= 0) return;
  eigh_batched_kernel<<<grid, EIGH_THREADS, 0, s>>>(slot);
  TDA_LAUNCH_CHECK();
}

void launch_eigh(const double* H, int ldh, int m, int n, double* work, double* evals, double* Y, int ldy,
                 hipStream_t s, unsigned long long* phase_clk) {
  TDA_CHECK(m >= 1 && m <= EIGH_MAXM, 2, "eigh: subspace dimension out of range (1..1024)");
  TDA_CHECK(n >= 1 && n <= m && n <= 64, 2, "eigh: number of roots out of range");
  TDA_CHECK(ldy >= m, 2, "eigh: ldy < m");
  if (m <= EIGH_MAXM_SMALL)
    eigh_kernel<EIGH_MAXM_SMALL, EIGH_POOL_SMALL><<<1, EIGH_THREADS, 0, s>>>(H, ldh, m, n, work, evals, Y, ldy,
                                                                         phase_clk);
  else
    eigh_kernel<EIGH_MAXM, EIGH_POOL><<<1, EIGH_THREADS, 0, s>>>(H, ldh, m, n, work, evals, Y, ldy, phase_clk);
  TDA_LAUNCH_CHECK();
}

}  // namespace tda
