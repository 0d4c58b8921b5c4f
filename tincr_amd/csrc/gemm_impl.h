// MFMA GEMM kernel templates (gfx950).  See gemm.h for the contract.
//
// Tiling: a 256-thread workgroup (4 waves of 64) owns a BM x BN tile of C;
// each wave owns a (BM/WGM) x (BN/WGN) sub-tile made of 16x16 MFMA blocks.
// K is walked in BK slices staged through LDS; the next slice is fetched into
// registers (16-byte vector loads along the operand's contiguous dimension)
// while the current one feeds the matrix cores.
//
// MFMA operand layouts used (lane l of 64, 16x16 output block):
//   f64 16x16x4, f32 16x16x4 : A[l%16][l/16], B[l/16][l%16]
//   f16/bf16 16x16x32        : A[l%16][8*(l/16)+e], B[8*(l/16)+e][l%16], e<8
//   result f32 accumulators   : D[4*(l/16)+r][l%16], r<4
//   result f64 16x16x4        : D[(l/16)+4*r][l%16], r<4   (checked on gfx950)
// LDS holds A as [k][m] (E=1 types) or [m][k] (E=8 types) so that fragment
// reads are single ds_read instructions.
#pragma once
#include <type_traits>

#include "gemm.h"

// TDA_XCD_REMAP: XCD-grouped tile order (below).  Measured on the C200H202
// AO products it did not beat the plain order, so it is off by default.
#ifndef TDA_XCD_REMAP
#define TDA_XCD_REMAP 0
#endif
// Two workgroups per CU: caps the registers so that the FP64 128x128 tile
// (64 FP64 accumulators per lane) runs 2 waves per SIMD instead of 1.
#ifndef TDA_GEMM_MIN_BLOCKS
#define TDA_GEMM_MIN_BLOCKS 2
#endif

namespace tda {

template <typename TI>
struct Mfma;
template <>
struct Mfma<double> {
  using TA = double;
  using Acc = v4d;
  using Frag = double;
  static constexpr int KSTEP = 4, E = 1;
  static __device__ inline Acc op(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // result row of register r in lane l: rows interleave with stride 4
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  using TA = float;
  using Acc = v4f;
  using Frag = float;
  static constexpr int KSTEP = 4, E = 1;
  static __device__ inline Acc op(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct Mfma<_Float16> {
  using TA = float;
  using Acc = v4f;
  using Frag = v8h;
  static constexpr int KSTEP = 32, E = 8;
  static __device__ inline Acc op(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct Mfma<__bf16> {
  using TA = float;
  using Acc = v4f;
  using Frag = v8bf;
  static constexpr int KSTEP = 32, E = 8;
  static __device__ inline Acc op(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename T, int N>
struct alignas(16) VecN {
  T v[N];
};

// Tile configurations.
struct CfgBig {  // 128x128, waves 2x2 (64x64 each)
  static constexpr int BM = 128, BN = 128, WGM = 2, WGN = 2, BK16 = 64;
};
struct CfgBigK {  // 128x128 with 128-deep K slices (16-bit inputs): twice the MFMA work per barrier pair
  static constexpr int BM = 128, BN = 128, WGM = 2, WGN = 2, BK16 = 128;
};
struct CfgSkinnyN {  // 64x32, waves 2x2 (32x16 each): N <= 32-ish
  static constexpr int BM = 64, BN = 32, WGM = 2, WGN = 2, BK16 = 64;
};
struct CfgSkinnyM {  // 32x128, waves 1x4 (32x32 each): M <= 32-ish
  static constexpr int BM = 32, BN = 128, WGM = 1, WGN = 4, BK16 = 64;
};
// FP64 products with 16 or fewer columns and a long K (Davidson projections
// V^T AV, DGKS overlaps V^T P): 64-deep K slices, a quarter of the barriers
// per byte of the 16-deep ones (measured 1.46x faster on a 192 x 16 x 2^21
// projection, 85 % of a plain streaming read of the same bytes).
struct CfgSkinnyN16 {  // 64x16, waves 4x1 (16x16 each): N <= 16
  static constexpr int BM = 64, BN = 16, WGM = 4, WGN = 1, BK16 = 64, BK1 = 64;
};
struct CfgSmall {  // 64x64, waves 2x2 (32x32 each)
  static constexpr int BM = 64, BN = 64, WGM = 2, WGN = 2, BK16 = 64;
};

// K slice of the 1-element-fragment types (FP64 / FP32): Cfg::BK1 when the
// configuration names one, else 16.
template <typename Cfg, typename = void>
struct BK1Of {
  static constexpr int value = 16;
};
template <typename Cfg>
struct BK1Of<Cfg, std::void_t<decltype(Cfg::BK1)>> {
  static constexpr int value = Cfg::BK1;
};

template <typename TI, typename Cfg>
struct BKOf {
  static constexpr int value = Mfma<TI>::E == 1 ? BK1Of<Cfg>::value : Cfg::BK16;
};

template <typename TO>
__device__ inline void store_out(TO* p, double v) {
  *p = to_type<TO>(v);
}

// Final value of C(m, n) in batch zb from the accumulated sum.
template <int EPI, typename TO>
__device__ inline double epi_value(const GemmArgs& g, int zb, int64_t m, int64_t n, double acc, double alpha,
                                   const TO* p) {
  double v = alpha * acc;
  if constexpr (EPI == EPI_DIAG) {
    v += g.diag[g.dm(m) + g.dmap(n)] * g.xsrc[zb * g.x_batch + g.xm(m) + g.xn(n)];
  } else if constexpr (EPI == EPI_HADAMARD) {
    v *= g.hsrc[g.hm(hada_index(g, m)) + g.hn(hada_index(g, n))];
  } else {
    if (g.beta != 0.0) v += g.beta * to_double<TO>(*p);
  }
  return v;
}

// LA / LB: element types of A and B in memory when they differ from the MFMA
// type TI (FP32-stored operands of an FP64 product: each element is widened
// when the slice is written to LDS, so the arithmetic is TI's).
template <typename TI, typename TO, typename Cfg, bool A_KC, bool B_KC, int EPI, typename LA = TI, typename LB = TI>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, const int zb, const int zs, const int64_t m0,
                                          const int64_t n0) {
  using MF = Mfma<TI>;
  using TA = typename MF::TA;
  constexpr int BM = Cfg::BM, BN = Cfg::BN, WGM = Cfg::WGM, WGN = Cfg::WGN;
  constexpr int BK = BKOf<TI, Cfg>::value;
  constexpr int E = MF::E, KSTEP = MF::KSTEP;
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int MI = WTM / 16, NI = WTN / 16;
  // 16-byte global loads: VA / VB elements of each operand's memory type
  constexpr int VA = 16 / sizeof(LA), VB = 16 / sizeof(LB);
  static_assert(E == 1 || (std::is_same<LA, TI>::value && std::is_same<LB, TI>::value),
                "widened operands exist for the 1-element-fragment types only");
  constexpr int PAD = (E == 1) ? (sizeof(TI) == 8 ? 2 : 4) : 8;
  // LDS geometry: E==1 -> [k][m] ; E==8 -> [m][k]
  constexpr int A_OUT = (E == 1) ? BK : BM, A_IN = (E == 1) ? BM + PAD : BK + PAD;
  constexpr int B_OUT = (E == 1) ? BK : BN, B_IN = (E == 1) ? BN + PAD : BK + PAD;
  __shared__ TI As[A_OUT * A_IN];
  __shared__ TI Bs[B_OUT * B_IN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / WGN) * WTM;
  const int wn0 = (wave % WGN) * WTN;


  const LA* A = (const LA*)g.A + zb * g.a_batch;
  const LB* B = (const LB*)g.B + zb * g.b_batch;

  const int64_t kbeg = (int64_t)zs * g.k_chunk;
  const int64_t kend = min(g.K, kbeg + g.k_chunk);
  const int nkt = kend > kbeg ? (int)((kend - kbeg + BK - 1) / BK) : 0;

  // ---- loader geometry (compile-time counts) ----
  constexpr int A_VECS = BM * BK / VA;  // vectors per A tile
  constexpr int B_VECS = BN * BK / VB;
  constexpr int A_PER = (A_VECS + 255) / 256;
  constexpr int B_PER = (B_VECS + 255) / 256;
  VecN<LA, VA> ra[A_PER];
  VecN<LB, VB> rb[B_PER];

  // Precompute the (tile-invariant) row/col parts of addresses.
  // For A_KC: vector idx -> (m = idx / (BK/VA), kv = idx % (BK/VA))
  // else      : vector idx -> (k = idx / (BM/VA), mv = idx % (BM/VA))  (B likewise with VB)
  // Row / column parts of the addresses and the per-vector eligibility for
  // the 16-byte load path (tile-invariant; the K part is checked per tile).
  const bool a_base_al = (((uintptr_t)A) & 15) == 0;
  const bool b_base_al = (((uintptr_t)B) & 15) == 0;
  int64_t a_row_off[A_PER];
  bool a_row_ok[A_PER], a_vec_ok[A_PER];
#pragma unroll
  for (int i = 0; i < A_PER; ++i) {
    int idx = tid + i * 256;
    int64_t m;
    if (A_KC) m = m0 + idx / (BK / VA);
    else m = m0 + (idx % (BM / VA)) * VA;
    a_row_ok[i] = (idx < A_VECS) && (m < g.M);
    a_row_off[i] = a_row_ok[i] ? g.am(m) : 0;
    if (A_KC) a_vec_ok[i] = a_row_ok[i] && a_base_al && (a_row_off[i] % VA == 0);
    else a_vec_ok[i] = a_row_ok[i] && a_base_al && (m + VA <= g.M) && (g.ak.s0 % VA == 0);
  }
  int64_t b_col_off[B_PER];
  bool b_col_ok[B_PER], b_vec_ok[B_PER];
#pragma unroll
  for (int i = 0; i < B_PER; ++i) {
    int idx = tid + i * 256;
    int64_t n;
    if (B_KC) n = n0 + idx / (BK / VB);
    else n = n0 + (idx % (BN / VB)) * VB;
    b_col_ok[i] = (idx < B_VECS) && (n < g.N);
    b_col_off[i] = b_col_ok[i] ? g.bn(n) : 0;
    if (B_KC) b_vec_ok[i] = b_col_ok[i] && b_base_al && (b_col_off[i] % VB == 0);
    else b_vec_ok[i] = b_col_ok[i] && b_base_al && (n + VB <= g.N) && (g.bk.s0 % VB == 0);
  }

  auto load_tile = [&](int kt) {
    const int64_t k0 = kbeg + (int64_t)kt * BK;
    // K-map evaluated once per tile (tiles never straddle a K-map boundary).
    const int64_t ak0 = g.ak(k0), bk0 = g.bk(k0);
    const bool full_k = k0 + BK <= kend;
    const bool a_k_al = (ak0 % VA) == 0, b_k_al = (bk0 % VB) == 0;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      int idx = tid + i * 256;
      if (A_KC) {
        const int kk = (idx % (BK / VA)) * VA;
        const LA* p = A + a_row_off[i] + ak0 + kk;  // ak.s0 == 1 (host check)
        if (a_vec_ok[i] && full_k && a_k_al) {
          ra[i] = *(const VecN<LA, VA>*)p;
        } else {
          VecN<LA, VA> v;
#pragma unroll
          for (int e = 0; e < VA; ++e) v.v[e] = (a_row_ok[i] && k0 + kk + e < kend) ? p[e] : (LA)0;
          ra[i] = v;
        }
      } else {
        const int kk = idx / (BM / VA);
        const int64_t m = m0 + (idx % (BM / VA)) * VA;
        const LA* p = A + a_row_off[i] + ak0 + g.ak.s0 * kk;
        const bool krow = (idx < A_VECS) && (k0 + kk < kend);
        if (a_vec_ok[i] && krow && a_k_al) {
          ra[i] = *(const VecN<LA, VA>*)p;
        } else {
          VecN<LA, VA> v;
#pragma unroll
          for (int e = 0; e < VA; ++e) v.v[e] = (krow && a_row_ok[i] && m + e < g.M) ? p[e] : (LA)0;
          ra[i] = v;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int idx = tid + i * 256;
      if (B_KC) {
        const int kk = (idx % (BK / VB)) * VB;
        const LB* p = B + b_col_off[i] + bk0 + kk;  // bk.s0 == 1 (host check)
        if (b_vec_ok[i] && full_k && b_k_al) {
          rb[i] = *(const VecN<LB, VB>*)p;
        } else {
          VecN<LB, VB> v;
#pragma unroll
          for (int e = 0; e < VB; ++e) v.v[e] = (b_col_ok[i] && k0 + kk + e < kend) ? p[e] : (LB)0;
          rb[i] = v;
        }
      } else {
        const int kk = idx / (BN / VB);
        const int64_t n = n0 + (idx % (BN / VB)) * VB;
        const LB* p = B + b_col_off[i] + bk0 + g.bk.s0 * kk;
        const bool krow = (idx < B_VECS) && (k0 + kk < kend);
        if (b_vec_ok[i] && krow && b_k_al) {
          rb[i] = *(const VecN<LB, VB>*)p;
        } else {
          VecN<LB, VB> v;
#pragma unroll
          for (int e = 0; e < VB; ++e) v.v[e] = (krow && b_col_ok[i] && n + e < g.N) ? p[e] : (LB)0;
          rb[i] = v;
        }
      }
    }
  };

  auto store_tile = [&]() {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      int idx = tid + i * 256;
      if (idx >= A_VECS) continue;
      int m, k;
      if (A_KC) { m = idx / (BK / VA); k = (idx % (BK / VA)) * VA; }
      else { k = idx / (BM / VA); m = (idx % (BM / VA)) * VA; }
      if constexpr (E == 1) {  // LDS [k][m]
        if (A_KC) {
#pragma unroll
          for (int e = 0; e < VA; ++e) As[(k + e) * A_IN + m] = (TI)ra[i].v[e];
        } else {
          if constexpr (std::is_same<LA, TI>::value) {
            *(VecN<TI, VA>*)&As[k * A_IN + m] = ra[i];
          } else {
#pragma unroll
            for (int e = 0; e < VA; ++e) As[k * A_IN + m + e] = (TI)ra[i].v[e];
          }
        }
      } else {  // LDS [m][k]
        if (A_KC) {
          *(VecN<TI, VA>*)&As[m * A_IN + k] = ra[i];
        } else {
#pragma unroll
          for (int e = 0; e < VA; ++e) As[(m + e) * A_IN + k] = ra[i].v[e];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int idx = tid + i * 256;
      if (idx >= B_VECS) continue;
      int n, k;
      if (B_KC) { n = idx / (BK / VB); k = (idx % (BK / VB)) * VB; }
      else { k = idx / (BN / VB); n = (idx % (BN / VB)) * VB; }
      if constexpr (E == 1) {  // LDS [k][n]
        if (B_KC) {
#pragma unroll
          for (int e = 0; e < VB; ++e) Bs[(k + e) * B_IN + n] = (TI)rb[i].v[e];
        } else {
          if constexpr (std::is_same<LB, TI>::value) {
            *(VecN<TI, VB>*)&Bs[k * B_IN + n] = rb[i];
          } else {
#pragma unroll
            for (int e = 0; e < VB; ++e) Bs[k * B_IN + n + e] = (TI)rb[i].v[e];
          }
        }
      } else {  // LDS [n][k]
        if (B_KC) {
          *(VecN<TI, VB>*)&Bs[n * B_IN + k] = rb[i];
        } else {
#pragma unroll
          for (int e = 0; e < VB; ++e) Bs[(n + e) * B_IN + k] = rb[i].v[e];
        }
      }
    }
  };

  typename MF::Acc acc[MI][NI];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0;

  if (nkt > 0) load_tile(0);
  for (int kt = 0; kt < nkt; ++kt) {
    __syncthreads();
    store_tile();
    __syncthreads();
    if (kt + 1 < nkt) load_tile(kt + 1);
#pragma unroll
    for (int ks = 0; ks < BK / KSTEP; ++ks) {
      typename MF::Frag fa[MI], fb[NI];
#pragma unroll
      for (int i = 0; i < MI; ++i) {
        const int m = wm0 + i * 16 + (lane & 15);
        if constexpr (E == 1) {
          fa[i] = As[(ks * KSTEP + (lane >> 4)) * A_IN + m];
        } else {
          fa[i] = *(const typename MF::Frag*)&As[m * A_IN + ks * KSTEP + 8 * (lane >> 4)];
        }
      }
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        const int n = wn0 + j * 16 + (lane & 15);
        if constexpr (E == 1) {
          fb[j] = Bs[(ks * KSTEP + (lane >> 4)) * B_IN + n];
        } else {
          fb[j] = *(const typename MF::Frag*)&Bs[n * B_IN + ks * KSTEP + 8 * (lane >> 4)];
        }
      }
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = MF::op(fa[i], fb[j], acc[i][j]);
    }
  }

  // ---- epilogue ----
  if (g.ksplit > 1) {
    TA* ws = (TA*)g.ws + ((int64_t)zb * g.ksplit + zs) * g.M * g.N;
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int j = 0; j < NI; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t m = m0 + wm0 + i * 16 + MF::row(lane, r);
          const int64_t n = n0 + wn0 + j * 16 + (lane & 15);
          if (m < g.M && n < g.N) ws[m * g.N + n] = acc[i][j][r];
        }
    return;
  }
  TO* C = (TO*)g.C + zb * g.c_batch;
  double alpha = g.alpha;
  if (g.batch_scale) alpha *= (double)g.batch_scale[zb];
  const float alpha_f = (float)alpha;
  // Column parts of the output / epilogue addresses, once per column.
  int64_t cn_off[NI], en_off[NI], xn_off[NI];
  bool n_ok[NI];
#pragma unroll
  for (int j = 0; j < NI; ++j) {
    const int64_t n = n0 + wn0 + j * 16 + (lane & 15);
    n_ok[j] = n < g.N;
    cn_off[j] = n_ok[j] ? g.cn(n) : 0;
    en_off[j] = 0;
    xn_off[j] = 0;
    if constexpr (EPI == EPI_DIAG) {
      if (n_ok[j]) {
        en_off[j] = g.dmap(n);
        xn_off[j] = g.xn(n);
      }
    } else if constexpr (EPI == EPI_HADAMARD) {
      if (n_ok[j]) en_off[j] = g.hn(hada_index(g, n));
    }
  }
  const double* xs = nullptr;
  if constexpr (EPI == EPI_DIAG) xs = g.xsrc + zb * g.x_batch;
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t m = m0 + wm0 + i * 16 + MF::row(lane, r);
      if (m >= g.M) continue;
      const int64_t cmo = g.cm(m);
      int64_t emo = 0, xmo = 0;
      if constexpr (EPI == EPI_DIAG) {
        emo = g.dm(m);
        xmo = g.xm(m);
      } else if constexpr (EPI == EPI_HADAMARD) {
        emo = g.hm(hada_index(g, m));
      }
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        if (!n_ok[j]) continue;
        TO* p = C + cmo + cn_off[j];
        if constexpr (std::is_same<TO, double>::value) {
          double v = alpha * (double)acc[i][j][r];
          if constexpr (EPI == EPI_DIAG) {
            v += g.diag[emo + en_off[j]] * xs[xmo + xn_off[j]];
          } else if constexpr (EPI == EPI_HADAMARD) {
            v *= g.hsrc[emo + en_off[j]];
          } else {
            if (g.beta != 0.0) v += g.beta * to_double<TO>(*p);
          }
          *p = v;
        } else {
          // FP32 / FP16 / BF16 outputs: the epilogue in FP32 (the accumulator's precision)
          float v = alpha_f * (float)acc[i][j][r];
          if constexpr (EPI == EPI_HADAMARD) {
            v *= (float)g.hsrc[emo + en_off[j]];
          } else if constexpr (EPI == EPI_STORE) {
            if (g.beta != 0.0) v += (float)g.beta * (float)to_double<TO>(*p);
          } else {
            v += (float)(g.diag[emo + en_off[j]] * xs[xmo + xn_off[j]]);
          }
          *p = (TO)v;
        }
      }
    }
}

template <typename TI, typename TO, typename Cfg, bool A_KC, bool B_KC, int EPI, typename LA = TI, typename LB = TI>
__global__ void __launch_bounds__(256, TDA_GEMM_MIN_BLOCKS) gemm_kernel(GemmArgs g) {
  constexpr int BM = Cfg::BM, BN = Cfg::BN;
#if TDA_XCD_REMAP
  // Workgroup w of the launch is dispatched to XCD w % 8, each XCD with its
  // own L2.  Give every XCD a contiguous run of the tile sequence, and walk
  // the tiles in groups of 8 tile rows, so that the tiles sharing A rows / B
  // columns run on the same L2 at about the same time.
  int64_t w = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
  {
    const int64_t nwg = (int64_t)gridDim.x * gridDim.y * gridDim.z;
    const int64_t q = nwg / 8, r = nwg % 8, x = w % 8, j = w / 8;
    w = (x < r) ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
  }
  const int64_t per_z = (int64_t)gridDim.x * gridDim.y;
  const int zz = (int)(w / per_z);
  const int64_t t = w - (int64_t)zz * per_z;
  constexpr int GM = 8;
  const int64_t gsize = (int64_t)GM * gridDim.x;
  const int64_t first_m = (t / gsize) * GM;
  const int64_t gm = min((int64_t)gridDim.y - first_m, (int64_t)GM);
  const int64_t tile_m = first_m + (t % gsize) % gm;
  const int64_t tile_n = (t % gsize) / gm;
  const int zb = zz / g.ksplit;
  const int zs = zz % g.ksplit;
  const int64_t m0 = tile_m * BM;
  const int64_t n0 = tile_n * BN;
#else
  const int zb = blockIdx.z / g.ksplit;
  const int zs = blockIdx.z % g.ksplit;
  const int64_t m0 = (int64_t)blockIdx.y * BM;
  const int64_t n0 = (int64_t)blockIdx.x * BN;
#endif
  gemm_tile<TI, TO, Cfg, A_KC, B_KC, EPI, LA, LB>(g, zb, zs, m0, n0);
}

// Grouped launch: many independent GEMMs (one GemmArgs each, e.g. one per
// fragment of a batch) in one persistent launch.  slot->tiles[t] =
// {descriptor, batch index z, tile row, tile column}; the workgroups stride
// over the slot->count tiles (no split-K).  The grid is fixed and the work
// list is read through the slot, so a captured launch replays with new
// descriptors and tile lists.
template <typename TI, typename TO, typename Cfg, bool A_KC, bool B_KC, int EPI, typename LA = TI, typename LB = TI>
__global__ void __launch_bounds__(256, TDA_GEMM_MIN_BLOCKS) gemm_grouped_kernel(const GroupSlot* __restrict__ slot) {
  const int nt = slot->count;
  const GemmArgs* __restrict__ descs = (const GemmArgs*)slot->descs;
  const int4* __restrict__ tiles = slot->tiles;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const int4 tl = tiles[t];
    gemm_tile<TI, TO, Cfg, A_KC, B_KC, EPI, LA, LB>(descs[tl.x], tl.y, 0, (int64_t)tl.z * Cfg::BM,
                                                     (int64_t)tl.w * Cfg::BN);
    __syncthreads();  // LDS reuse by the next tile
  }
}


// Sum split-K partials and apply the epilogue.
// Many partials per output (ksplit >= 16): one wave per output element, the
// 64 lanes split the partials and combine them with a wave sum.
template <typename TA, typename TO, int EPI>
__global__ void __launch_bounds__(256) splitk_reduce_wave_kernel(GemmArgs g) {
  const int64_t MN = g.M * g.N;
  const int64_t out = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (out >= MN * g.batch) return;
  const int zb = (int)(out / MN);
  const int64_t mn = out - (int64_t)zb * MN;
  const int64_t m = mn / g.N, n = mn - m * g.N;
  const TA* ws = (const TA*)g.ws + (int64_t)zb * g.ksplit * MN + mn;
  double s = 0.0;
  for (int z = lane; z < g.ksplit; z += 64) s += (double)ws[z * MN];
  s = wave_sum(s);
  if (lane != 0) return;
  double alpha = g.alpha;
  if (g.batch_scale) alpha *= (double)g.batch_scale[zb];
  TO* p = (TO*)g.C + zb * g.c_batch + g.cm(m) + g.cn(n);
  const double v = epi_value<EPI, TO>(g, zb, m, n, s, alpha, p);
  *p = to_type<TO>(v);
}

// Few partials per output: one thread per output element.  Grid: x over
// column chunks of 256, y over rows (strided), z over batches.
template <typename TA, typename TO, int EPI>
__global__ void __launch_bounds__(256) splitk_reduce_kernel(GemmArgs g) {
  const int64_t MN = g.M * g.N;
  const int zb = blockIdx.z;
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= g.N) return;
  double alpha = g.alpha;
  if (g.batch_scale) alpha *= (double)g.batch_scale[zb];
  const int64_t cno = g.cn(n);
  for (int64_t m = blockIdx.y; m < g.M; m += gridDim.y) {
    const int64_t mn = m * g.N + n;
    const TA* ws = (const TA*)g.ws + (int64_t)zb * g.ksplit * MN + mn;
    double s = 0.0;
    for (int z = 0; z < g.ksplit; ++z) s += (double)ws[z * MN];
    TO* p = (TO*)g.C + zb * g.c_batch + g.cm(m) + cno;
    const double v = epi_value<EPI, TO>(g, zb, m, n, s, alpha, p);
    *p = to_type<TO>(v);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

inline int cfg_bm(int c) {
  switch (c) {
    case CFG_BIG: case CFG_BIG_K: return 128;
    case CFG_SKINNY_N: case CFG_SKINNY_N16: return 64;
    case CFG_SKINNY_M: return 32;
    default: return 64;
  }
}
inline int cfg_bn(int c) {
  switch (c) {
    case CFG_BIG: case CFG_BIG_K: return 128;
    case CFG_SKINNY_N: return 32;
    case CFG_SKINNY_N16: return 16;
    case CFG_SKINNY_M: return 128;
    default: return 64;
  }
}
inline int bk_of(DType in, int cfg = CFG_BIG) {
  if (in == DT_F64 && cfg == CFG_SKINNY_N16) return CfgSkinnyN16::BK1;
  if (in == DT_F64 || in == DT_F32) return 16;
  return cfg == CFG_BIG_K ? 128 : 64;
}
inline size_t acc_bytes(DType in) { return in == DT_F64 ? 8 : 4; }

template <typename TI, typename TO, typename Cfg, int EPI, typename LA = TI, typename LB = TI>
void launch_cfg(bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(g.N, Cfg::BN), (unsigned)ceil_div(g.M, Cfg::BM),
            (unsigned)(g.batch * g.ksplit));
  if (a_kc && b_kc) gemm_kernel<TI, TO, Cfg, true, true, EPI, LA, LB><<<grid, 256, 0, s>>>(g);
  else if (a_kc && !b_kc) gemm_kernel<TI, TO, Cfg, true, false, EPI, LA, LB><<<grid, 256, 0, s>>>(g);
  else if (!a_kc && b_kc) gemm_kernel<TI, TO, Cfg, false, true, EPI, LA, LB><<<grid, 256, 0, s>>>(g);
  else gemm_kernel<TI, TO, Cfg, false, false, EPI, LA, LB><<<grid, 256, 0, s>>>(g);
  TDA_LAUNCH_CHECK();
}

template <typename TI, typename TO, int EPI, typename LA = TI, typename LB = TI>
void launch_t(int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s) {
  switch (cfg) {
    case CFG_BIG: launch_cfg<TI, TO, CfgBig, EPI, LA, LB>(a_kc, b_kc, g, s); break;
    case CFG_BIG_K:
      if constexpr (Mfma<TI>::E == 8) launch_cfg<TI, TO, CfgBigK, EPI, LA, LB>(a_kc, b_kc, g, s);
      else launch_cfg<TI, TO, CfgBig, EPI, LA, LB>(a_kc, b_kc, g, s);
      break;
    case CFG_SKINNY_N: launch_cfg<TI, TO, CfgSkinnyN, EPI, LA, LB>(a_kc, b_kc, g, s); break;
    case CFG_SKINNY_M: launch_cfg<TI, TO, CfgSkinnyM, EPI, LA, LB>(a_kc, b_kc, g, s); break;
    case CFG_SKINNY_N16:
      if constexpr (std::is_same<TI, double>::value)
        launch_cfg<TI, TO, CfgSkinnyN16, EPI, LA, LB>(a_kc, b_kc, g, s);
      else launch_cfg<TI, TO, CfgSkinnyN, EPI, LA, LB>(a_kc, b_kc, g, s);
      break;
    default: launch_cfg<TI, TO, CfgSmall, EPI, LA, LB>(a_kc, b_kc, g, s); break;
  }
}

template <typename TA, typename TO, int EPI>
void launch_reduce(GemmArgs& g, hipStream_t s) {
  if (g.ksplit >= 16) {
    const int64_t outs = g.M * g.N * g.batch;
    splitk_reduce_wave_kernel<TA, TO, EPI><<<(unsigned)ceil_div(outs, 4), 256, 0, s>>>(g);
  } else {
    const int64_t nx = ceil_div(g.N, 256);
    // enough row blocks for ~4096 workgroups in total, at most one per row
    const int64_t ny =
        std::max<int64_t>(1, std::min<int64_t>(g.M, std::min<int64_t>(65535, ceil_div(4096, nx * g.batch))));
    dim3 grid((unsigned)nx, (unsigned)ny, (unsigned)g.batch);
    splitk_reduce_kernel<TA, TO, EPI><<<grid, 256, 0, s>>>(g);
  }
  TDA_LAUNCH_CHECK();
}


// Per-type entry points, each defined in its own translation unit
// (gemm_f64.hip, gemm_f32.hip, gemm_f16.hip, gemm_bf16.hip).
void gemm_dispatch_f64(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s);
// FP64 products with one or both operands stored as FP32 (gemm_f64mix.hip)
void gemm_dispatch_f64mix(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s);
void gemm_dispatch_f32(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s);
void gemm_dispatch_f16(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s);
void gemm_dispatch_bf16(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s);
// The K-contiguous 16-bit kernel (gemm16_impl.h), for arguments that pass
// gemm16_ok; tile: 1 = 256x256, 0 = 128x128, -1 = by tile count.
void gemm16_dispatch_f16(Epi epi, GemmArgs& g, hipStream_t s, int tile);
void gemm16_dispatch_bf16(Epi epi, GemmArgs& g, hipStream_t s, int tile);

}  // namespace tda
