// Dot products of the Davidson basis with a few long rows (davidson.rs:89,
// 162-168 in block form): C[i][j] = alpha sum_l A[i][l] B[j][l] + beta C[i][j]
// (projections V^T AV, overlaps V^T P).  The basis V [m][L] is long
// (L = n_occ * v_p, 2.1 million at (H2O)512) and short (m <= ~128), so the
// product is an HBM stream with little arithmetic per byte.  The kernel reads
// A and B straight from HBM into registers (no LDS staging, no barrier inside
// the stream); A may be stored as FP32 (the basis of the FP16 / BF16 phase);
// every product and sum is FP64.  Measured on (H2O)512 rows (L = 2^21, 16
// columns of FP64 B) against the LDS-tiled MFMA GEMM: FP32 A 0.15 / 0.29 /
// 0.40 ms at m = 32 / 80 / 128 (GEMM 0.22 / 0.42 / 0.41), FP64 A 0.20 / 0.42 /
// 0.59 ms (GEMM 0.22 / 0.47 / 0.49); the Ritz-vector products Y V stay on the
// GEMM, which streamed them faster than a register-FMA kernel.
//
// rowdot: each wave takes a 64 x 16 block of C over a slice of l and feeds
// v_mfma_f64_16x16x4 directly from global loads.  Lane (r = lane % 16,
// g = lane / 16) loads 4 consecutive elements l0 + 4g .. l0 + 4g + 3 of its
// A rows and of its B row (one 16-byte load for FP32); MFMA step e of the
// 16-element slice then multiplies element 4g + e in k-slot g.  Any
// assignment of l to (step, slot) shared by A and B sums the same terms, and
// this one makes each lane's loads contiguous.  The four waves of a
// workgroup split the workgroup's slice of l and add their blocks in LDS;
// the workgroups' partial blocks are summed in a fixed order by a second
// kernel (results do not depend on timing).

#include <algorithm>

#include "common.h"
#include "kernels.h"

namespace tda {
namespace {

constexpr int RD_WAVES = 4;

// raw 4-element loads (kept in the storage type until the MFMA needs them)
template <typename T>
struct Raw4 {
  T v[4];
};
template <typename T>
__device__ inline Raw4<T> ld4(const T* p) {
  Raw4<T> r;
  if constexpr (sizeof(T) == 4) {
    const float4 x = *(const float4*)p;
    r.v[0] = x.x; r.v[1] = x.y; r.v[2] = x.z; r.v[3] = x.w;
  } else {
    const double2 x = *(const double2*)p;
    const double2 y = *(const double2*)(p + 2);
    r.v[0] = x.x; r.v[1] = x.y; r.v[2] = y.x; r.v[3] = y.y;
  }
  return r;
}
template <typename T>
__device__ inline Raw4<T> ld4_tail(const T* p, int64_t l, int64_t lend) {
  Raw4<T> r;
#pragma unroll
  for (int e = 0; e < 4; ++e) r.v[e] = (l + e < lend) ? p[e] : (T)0;
  return r;
}

// grid: x = slices of l (ks), y = (16 RT)-row blocks, z = 16-column blocks.
// Partial block of slice x goes to ws[x][M][N] (ks > 1) or, with alpha /
// beta applied, to C (ks == 1).  Each wave takes 32-element steps of the
// slice (every RD_WAVES-th), issuing the loads of both 16-element halves
// before their MFMAs.
template <typename TA, typename TB, int RT, int US = 2>
__global__ void __launch_bounds__(256) rowdot_kernel(const TA* __restrict__ A, int64_t lda, const TB* __restrict__ B,
                                                     int64_t ldb, int M, int N, int64_t K, int64_t slice,
                                                     double alpha, double beta, double* __restrict__ C, int64_t ldc,
                                                     double* __restrict__ ws) {
  constexpr int BR = 16 * RT;  // rows per block
  __shared__ double red[RD_WAVES - 1][BR * 16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int row0 = blockIdx.y * BR, col0 = blockIdx.z * 16;
  const int64_t kb = (int64_t)blockIdx.x * slice;
  const int64_t ke = min(K, kb + slice);
  const TA* ap[RT];
  bool arow[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int row = row0 + 16 * t + r;
    arow[t] = row < M;
    ap[t] = A + (int64_t)(arow[t] ? row : 0) * lda;
  }
  const int col = col0 + r;
  const bool bcol = col < N;
  const TB* bp = B + (int64_t)(bcol ? col : 0) * ldb;
  v4d acc[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  const Raw4<TA> za{};
  const Raw4<TB> zb{};
  for (int64_t k = kb + 16 * US * wave; k < ke; k += 16 * US * RD_WAVES) {
    Raw4<TA> ra[US][RT];
    Raw4<TB> rb[US];
    if (k + 16 * US <= ke) {
#pragma unroll
      for (int h = 0; h < US; ++h) {
        const int64_t kl = k + 16 * h + 4 * g;
#pragma unroll
        for (int t = 0; t < RT; ++t) ra[h][t] = arow[t] ? ld4<TA>(ap[t] + kl) : za;
        rb[h] = bcol ? ld4<TB>(bp + kl) : zb;
      }
    } else {
#pragma unroll
      for (int h = 0; h < US; ++h) {
        const int64_t kl = k + 16 * h + 4 * g;
#pragma unroll
        for (int t = 0; t < RT; ++t) ra[h][t] = arow[t] ? ld4_tail<TA>(ap[t] + kl, kl, ke) : za;
        rb[h] = bcol ? ld4_tail<TB>(bp + kl, kl, ke) : zb;
      }
    }
#pragma unroll
    for (int h = 0; h < US; ++h)
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int t = 0; t < RT; ++t)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)ra[h][t].v[e], (double)rb[h].v[e], acc[t], 0, 0, 0);
  }
  // result element (t, q) of the lane: row 16 t + g + 4 q, column r
  if (wave > 0) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) red[wave - 1][(16 * t + g + 4 * q) * 16 + r] = acc[t][q];
  }
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int lr = 16 * t + g + 4 * q;
      double s = acc[t][q];
#pragma unroll
      for (int w = 0; w < RD_WAVES - 1; ++w) s += red[w][lr * 16 + r];
      const int row = row0 + lr;
      if (row >= M || !bcol) continue;
      if (ws) {
        ws[((int64_t)blockIdx.x * M + row) * N + col] = s;
      } else {
        double* p = C + (int64_t)row * ldc + col;
        *p = alpha * s + (beta != 0.0 ? beta * *p : 0.0);
      }
    }
}

// C[i][j] = alpha sum_x ws[x][i][j] + beta C[i][j], one wave per element,
// partials summed in a fixed order (lane-strided, then the wave tree)
__global__ void __launch_bounds__(256) rowdot_reduce_kernel(const double* __restrict__ ws, int ks, int M, int N,
                                                            double alpha, double beta, double* __restrict__ C,
                                                            int64_t ldc) {
  const int64_t out = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (out >= (int64_t)M * N) return;
  const int i = (int)(out / N), j = (int)(out % N);
  const int64_t MN = (int64_t)M * N;
  double s = 0.0;
  for (int x = lane; x < ks; x += 64) s += ws[x * MN + out];
  s = wave_sum(s);
  if (lane == 0) {
    double* p = C + (int64_t)i * ldc + j;
    *p = alpha * s + (beta != 0.0 ? beta * *p : 0.0);
  }
}

}  // namespace

// 16-row tiles per wave: 8 for FP32 A with M > 64 (the B rows are then read
// once for M <= 128), else 4 (FP64 A at 8 tiles runs one wave per SIMD)
static int rowdot_rt(int M, bool a32) { return a32 && M > 64 ? 8 : 4; }
static int64_t rowdot_slices(int M, int N, int64_t K, bool a32) {
  const int br = 16 * rowdot_rt(M, a32);
  const int64_t tiles = (int64_t)((M + br - 1) / br) * ((N + 15) / 16);
  // about 2048 workgroups, slices of at least 1024 elements
  return std::max<int64_t>(1, std::min<int64_t>((2048 + tiles - 1) / tiles, (K + 1023) / 1024));
}

size_t rowdot_ws_bytes(int M, int N, int64_t K) {
  const int64_t ks = std::max(rowdot_slices(M, N, K, false), rowdot_slices(M, N, K, true));
  return ks > 1 ? (size_t)ks * M * N * 8 : 0;
}

template <typename TA, typename TB>
static void rowdot_t(const void* A, int64_t lda, const void* B, int64_t ldb, int M, int N, int64_t K, int64_t ks,
                     int64_t slice, double alpha, double beta, double* C, int64_t ldc, double* w, hipStream_t s) {
  const int rt = rowdot_rt(M, sizeof(TA) == 4);
  dim3 grid((unsigned)ks, (unsigned)((M + 16 * rt - 1) / (16 * rt)), (unsigned)((N + 15) / 16));
  if (rt == 8)
    rowdot_kernel<TA, TB, 8><<<grid, 256, 0, s>>>((const TA*)A, lda, (const TB*)B, ldb, M, N, K, slice, alpha, beta,
                                                  C, ldc, w);
  else
    rowdot_kernel<TA, TB, 4><<<grid, 256, 0, s>>>((const TA*)A, lda, (const TB*)B, ldb, M, N, K, slice, alpha, beta,
                                                  C, ldc, w);
  TDA_LAUNCH_CHECK();
}

void launch_rowdot(const void* A, bool a32, int64_t lda, const void* B, bool b32, int64_t ldb, int M, int N,
                   int64_t K, double alpha, double beta, double* C, int64_t ldc, double* ws, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  TDA_CHECK(((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0 && lda % 4 == 0 && ldb % 4 == 0, 2,
            "rowdot: rows must be 16-byte aligned");
  int64_t ks = rowdot_slices(M, N, K, a32);
  // slices: multiples of 128 elements (a workgroup's four waves take 32 each)
  const int64_t slice = ((K + ks - 1) / ks + 127) / 128 * 128;
  ks = std::max<int64_t>(1, (K + slice - 1) / slice);
  TDA_CHECK(ks == 1 || ws, 2, "rowdot: workspace needed");
  double* w = ks > 1 ? ws : nullptr;
  if (a32 && b32) rowdot_t<float, float>(A, lda, B, ldb, M, N, K, ks, slice, alpha, beta, C, ldc, w, s);
  else if (a32) rowdot_t<float, double>(A, lda, B, ldb, M, N, K, ks, slice, alpha, beta, C, ldc, w, s);
  else if (b32) rowdot_t<double, float>(A, lda, B, ldb, M, N, K, ks, slice, alpha, beta, C, ldc, w, s);
  else rowdot_t<double, double>(A, lda, B, ldb, M, N, K, ks, slice, alpha, beta, C, ldc, w, s);
  if (ks > 1) {
    const int64_t outs = (int64_t)M * N;
    rowdot_reduce_kernel<<<(unsigned)((outs + 3) / 4), 256, 0, s>>>(ws, (int)ks, M, N, alpha, beta, C, ldc);
    TDA_LAUNCH_CHECK();
  }
}

}  // namespace tda
