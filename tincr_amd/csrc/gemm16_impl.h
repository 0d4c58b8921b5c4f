// 16-bit-input GEMM for operands stored with K contiguous ("NT"), gfx950.
//
//   C[z][cm(m) + cn(n)] = epilogue( alpha * sum_k A[z][m*lda + k] * B[z][n*ldb + k] )
//
// This is the shape of every product of the AO-factored exchange (L XR^T,
// R' M^T, R X^T, L'^T F^T): both operands row-major along K.  Compared with
// the generic index-map kernel (gemm_impl.h) it
//   * stages K slices of 64 in two LDS buffers (one barrier per slice; the
//     next slice is fetched into registers while the current one feeds the
//     matrix cores),
//   * stores each operand row in LDS as eight 16-byte chunks permuted by
//     chunk ^ (row & 7), so the ds_read_b128 fragment reads of 16 rows hit 16
//     different bank groups,
//   * uses 256x256 workgroup tiles (8 waves of 128x64) for large products and
//     128x128 tiles (4 waves of 64x64) for small ones,
//   * computes the transposed 16x16 block (B fragment as the MFMA's first
//     operand), so a lane ends with 4 consecutive columns of one row and
//     stores them as one 8-byte (16-bit types) or two 16-byte (FP64) vectors.
// Split operands (a_lo / b_lo, FP16): the K loop runs three segments,
// A_hi B_lo, A_lo B_hi, then -- after the accumulators are scaled by 2^-11 --
// A_hi B_hi, which recovers about FP32 accuracy from FP16 MFMA (the hi / lo
// splitting of Ootomo and Yokota, "Recovering single precision accuracy from
// Tensor Cores while surpassing the FP32 theoretical peak performance", 2022).
// MFMA: v_mfma_f32_16x16x32_{f16,bf16}, FP32 accumulation.  Lane l supplies
// rows l%16, K elements 8*(l/16)..+7 of both operands; the result block
// holds D[4*(l/16)+r][l%16], r < 4.
#pragma once
#include <type_traits>

#include "gemm_impl.h"

namespace tda {

template <int BM_, int BN_, int WGM_, int WGN_, int MINB_>
struct Cfg16 {
  static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_, NT = WGM_ * WGN_ * 64, MINB = MINB_;
};
using Cfg16Big = Cfg16<256, 256, 2, 4, 1>;  // 8 waves, 128x64 each; 128 KiB LDS
using Cfg16Mid = Cfg16<128, 128, 2, 2, 2>;  // 4 waves, 64x64 each; 64 KiB LDS

template <typename TI>
struct Mfma16;
template <>
struct Mfma16<_Float16> {
  using Frag = v8h;
  static __device__ inline v4f op(Frag a, Frag b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma16<__bf16> {
  using Frag = v8bf;
  static __device__ inline v4f op(Frag a, Frag b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

template <typename T>
struct alignas(8) Vec4 {
  T v[4];
};

// DMA: the K slices go from global memory straight into LDS
// (global_load_lds_dwordx4: each wave instruction fills 1 KiB of LDS, lane l
// its 16-byte chunk l), with no round trip through VGPRs and no LDS stores;
// the two LDS buffers are separate arrays, so the reads of one are not held
// up by the transfers into the other.  Needs K a multiple of 64.
template <typename TI, typename TO, typename Cfg, int EPI, bool SPLIT, bool DMA = false>
__global__ void __launch_bounds__(Cfg::NT, Cfg::MINB) gemm16_nt_kernel(GemmArgs g) {
  using MF = Mfma16<TI>;
  using Frag = typename MF::Frag;
  constexpr int BM = Cfg::BM, BN = Cfg::BN, NT = Cfg::NT;
  constexpr int WTM = BM / Cfg::WGM, WTN = BN / Cfg::WGN;
  constexpr int MI = WTM / 16, NI = WTN / 16;
  constexpr int BK = 64;                  // K elements per slice = 8 chunks of 16 bytes
  constexpr int A_CH = BM * 8, B_CH = BN * 8;  // 16-byte chunks per slice
  constexpr int A_PER = A_CH / NT, B_PER = B_CH / NT;
  static_assert(A_CH % NT == 0 && B_CH % NT == 0, "tile / thread count mismatch");
  constexpr int STAGE = (BM + BN) * BK;   // elements per LDS buffer
  __shared__ __align__(16) TI lds[DMA ? STAGE : 2 * STAGE];
  __shared__ __align__(16) TI lds_b[DMA ? STAGE : 8];  // DMA: the second buffer

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / Cfg::WGN) * WTM;
  const int wn0 = (wave % Cfg::WGN) * WTN;
  const int zb = blockIdx.z;
  const int64_t m0 = (int64_t)blockIdx.y * BM;
  const int64_t n0 = (int64_t)blockIdx.x * BN;
  const int64_t lda = g.am.s0, ldb = g.bn.s0;
  const TI* A = (const TI*)g.A + zb * g.a_batch;
  const TI* B = (const TI*)g.B + zb * g.b_batch;
  const int64_t K = g.K;
  const int nks = (int)((K + BK - 1) / BK);  // K slices per segment
  constexpr bool split_in = SPLIT;  // a_lo, b_lo != 0 (checked on the host)
  const int nkt = split_in ? 3 * nks : nks;
  const uint32_t a_lo = (uint32_t)g.a_lo, b_lo = (uint32_t)g.b_lo;

  // Loader: chunk c = tid + i*NT of a slice is row c/8 (= tid/8 + i*NT/8),
  // K chunk c%8 (= tid%8).  Rows past M / N read row M-1 / N-1 (results
  // discarded at the store).  32-bit element offsets from the batch base.
  const int lr = tid >> 3, lkc = tid & 7;
  uint32_t a_off[A_PER], b_off[B_PER];
#pragma unroll
  for (int i = 0; i < A_PER; ++i)
    a_off[i] = (uint32_t)(min(m0 + lr + i * (NT / 8), g.M - 1) * lda) + lkc * 8;
#pragma unroll
  for (int i = 0; i < B_PER; ++i)
    b_off[i] = (uint32_t)(min(n0 + lr + i * (NT / 8), g.N - 1) * ldb) + lkc * 8;
  VecN<TI, 8> ra[1][A_PER], rb[1][B_PER];  // the next K slice, in flight
  v4f acc[MI][NI];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};


  auto load = [&](int kt, auto set_c) {
    constexpr int P = decltype(set_c)::value;
    // segment (split operands): 0 = A_hi B_lo, 1 = A_lo B_hi, 2 = A_hi B_hi
    int ks = kt;
    uint32_t ao = 0, bo = 0;
    if constexpr (split_in) {
      const int seg = kt / nks;
      ks = kt - seg * nks;
      ao = seg == 1 ? a_lo : 0u;
      bo = seg == 0 ? b_lo : 0u;
    }
    const uint32_t k0 = (uint32_t)ks * BK;
    const uint32_t ka = k0 + ao, kb = k0 + bo;
    if ((int64_t)k0 + BK <= K) {
#pragma unroll
      for (int i = 0; i < A_PER; ++i) ra[P][i] = *(const VecN<TI, 8>*)(A + (a_off[i] + ka));
#pragma unroll
      for (int i = 0; i < B_PER; ++i) rb[P][i] = *(const VecN<TI, 8>*)(B + (b_off[i] + kb));
    } else {
      // last slice of a segment: elements past K read as zero
      const int64_t kk = (int64_t)k0 + lkc * 8;
#pragma unroll
      for (int i = 0; i < A_PER; ++i) {
        if (kk + 8 <= K) {
          ra[P][i] = *(const VecN<TI, 8>*)(A + (a_off[i] + ka));
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) ra[P][i].v[e] = (kk + e < K) ? A[a_off[i] + ka + e] : (TI)0;
        }
      }
#pragma unroll
      for (int i = 0; i < B_PER; ++i) {
        if (kk + 8 <= K) {
          rb[P][i] = *(const VecN<TI, 8>*)(B + (b_off[i] + kb));
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) rb[P][i].v[e] = (kk + e < K) ? B[b_off[i] + kb + e] : (TI)0;
        }
      }
    }
  };
  // split operands: the correction segments are done once 2 nks slices are in
  auto rescale = [&](int kt) {
    if (split_in && kt == 2 * nks) {  // compile-time false without SPLIT
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] *= 4.8828125e-4f;  // 2^-11
    }
  };
  // LDS position of this thread's chunks: row lr + i*NT/8, chunk lkc ^ (row & 7)
  // (NT/8 is a multiple of 8, so row & 7 == lr & 7)
  const int lds_a0 = lr * BK + ((lkc ^ (lr & 7)) << 3);
  auto store = [&](int buf, auto set_c) {
    constexpr int P = decltype(set_c)::value;
    TI* s = lds + buf * STAGE;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) *(VecN<TI, 8>*)(s + lds_a0 + i * (NT / 8) * BK) = ra[P][i];
#pragma unroll
    for (int i = 0; i < B_PER; ++i) *(VecN<TI, 8>*)(s + BM * BK + lds_a0 + i * (NT / 8) * BK) = rb[P][i];
  };

  // fragment row (within the wave tile) and its swizzle key
  const int fr = lane & 15, fq = lane >> 4;
  // One K slice: 2 steps of 32 x (MI x NI) MFMAs.  Fragment reads run one
  // group ahead of the MFMAs that consume them (the next pair of A fragments,
  // and the next step's B fragments, are read while the current pair is
  // multiplied), so the LDS latency overlaps matrix-core work.
  auto frag_a = [&](const TI* s, int ks, int i) {
    const int r = wm0 + i * 16 + fr, kc = ks * 4 + fq;
    return *(const Frag*)(s + r * BK + ((kc ^ (r & 7)) << 3));
  };
  auto frag_b = [&](const TI* s, int ks, int j) {
    const int r = wn0 + j * 16 + fr, kc = ks * 4 + fq;
    return *(const Frag*)(s + BM * BK + r * BK + ((kc ^ (r & 7)) << 3));
  };
  auto compute_at = [&](const TI* s) {
    static_assert(MI % 2 == 0, "A fragments go in pairs");
    Frag fb[NI], fa0 = frag_a(s, 0, 0), fa1 = frag_a(s, 0, 1);
#pragma unroll
    for (int j = 0; j < NI; ++j) fb[j] = frag_b(s, 0, j);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      Frag nb[NI];
#pragma unroll
      for (int ip = 0; ip < MI / 2; ++ip) {
        Frag na0, na1;
        if (ip + 1 < MI / 2) {
          na0 = frag_a(s, ks, 2 * ip + 2);
          na1 = frag_a(s, ks, 2 * ip + 3);
        } else if (ks == 0) {
          na0 = frag_a(s, 1, 0);
          na1 = frag_a(s, 1, 1);
#pragma unroll
          for (int j = 0; j < NI; ++j) nb[j] = frag_b(s, 1, j);
        }
        // keep the reads above this group's MFMAs (the scheduler would
        // otherwise sink them next to their uses)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[2 * ip][j] = MF::op(fb[j], fa0, acc[2 * ip][j]);
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[2 * ip + 1][j] = MF::op(fb[j], fa1, acc[2 * ip + 1][j]);
        __builtin_amdgcn_sched_barrier(0);
        fa0 = na0;
        fa1 = na1;
      }
      if (ks == 0) {
#pragma unroll
        for (int j = 0; j < NI; ++j) fb[j] = nb[j];
      }
    }
  };
  auto compute = [&](int buf) { compute_at(lds + buf * STAGE); };
  using I0 = std::integral_constant<int, 0>;
  if constexpr (DMA) {
    // wave w fills rows 8 (w A_INS + i) .. +7 of the A part and likewise of
    // the B part; lane l: row +l/8, LDS chunk l%8, which holds K chunk
    // (l%8) ^ (row & 7) (the swizzle the fragment reads undo)
    constexpr int NW = NT / 64, A_INS = BM / 8 / NW, B_INS = BN / 8 / NW;
    static_assert(A_INS * 8 * NW == BM && B_INS * 8 * NW == BN, "DMA row mapping");
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int lrow = lane >> 3, kch = (lane & 7) ^ lrow;
    uint32_t da[A_INS], db[B_INS];
#pragma unroll
    for (int i = 0; i < A_INS; ++i)
      da[i] = (uint32_t)(min(m0 + (wv * A_INS + i) * 8 + lrow, g.M - 1) * lda) + kch * 8;
#pragma unroll
    for (int i = 0; i < B_INS; ++i)
      db[i] = (uint32_t)(min(n0 + (wv * B_INS + i) * 8 + lrow, g.N - 1) * ldb) + kch * 8;
    auto dma = [&](int kt, TI* dst) {
      int ks = kt;
      uint32_t ao = 0, bo = 0;
      if constexpr (split_in) {
        const int seg = kt / nks;
        ks = kt - seg * nks;
        ao = seg == 1 ? a_lo : 0u;
        bo = seg == 0 ? b_lo : 0u;
      }
      const uint32_t ka = (uint32_t)ks * BK + ao, kb = (uint32_t)ks * BK + bo;
#pragma unroll
      for (int i = 0; i < A_INS; ++i)
        __builtin_amdgcn_global_load_lds((const void*)(A + (da[i] + ka)),
                                         (__attribute__((address_space(3))) void*)(dst + (wv * A_INS + i) * 8 * BK),
                                         16, 0, 0);
#pragma unroll
      for (int i = 0; i < B_INS; ++i)
        __builtin_amdgcn_global_load_lds((const void*)(B + (db[i] + kb)),
                                         (__attribute__((address_space(3))) void*)(dst + BM * BK +
                                                                                   (wv * B_INS + i) * 8 * BK),
                                         16, 0, 0);
    };
    // vmcnt(0): this wave's transfers into LDS are complete (the barrier
    // after it makes every wave's visible)
    auto dma_wait = [] { __builtin_amdgcn_s_waitcnt(0x0F70); };
    if (nkt > 0) {
      dma(0, lds);
      dma_wait();
      __syncthreads();
    }
    for (int kt = 0; kt < nkt; ++kt) {
      const bool odd = kt & 1;
      if (kt + 1 < nkt) dma(kt + 1, odd ? lds : lds_b);
      rescale(kt);
      compute_at(odd ? lds_b : lds);
      dma_wait();
      __syncthreads();
    }
  } else {
    if (nkt > 0) {
      load(0, I0{});
      store(0, I0{});
      __syncthreads();
    }
    for (int kt = 0; kt < nkt; ++kt) {
      if (kt + 1 < nkt) load(kt + 1, I0{});
      rescale(kt);
      compute(kt & 1);
      if (kt + 1 < nkt) store((kt + 1) & 1, I0{});
      __syncthreads();
    }
  }

  // ---- epilogue: lane holds C[m][n..n+3], m = row l%16 of block i, n = 4*(l/16) of block j ----
  TO* C = (TO*)g.C + zb * g.c_batch;
  const float alpha = (float)(g.batch_scale ? g.alpha * (double)g.batch_scale[zb] : g.alpha);
  // column offsets per block j (4 consecutive columns)
  int64_t cn_off[NI];
  bool cn_vec[NI];
  int64_t n_first[NI];
#pragma unroll
  for (int j = 0; j < NI; ++j) {
    const int64_t n = n0 + wn0 + j * 16 + 4 * fq;
    n_first[j] = n;
    const bool all_in = n + 4 <= g.N;
    cn_off[j] = n < g.N ? g.cn(n) : 0;
    // the 4 columns map to 4 consecutive, vector-aligned elements of C
    bool contig = all_in && g.cn.s0 == 1 && (g.cn.div >= (INT64_C(1) << 62) || (n % g.cn.div) + 4 <= g.cn.div);
    cn_vec[j] = contig && (cn_off[j] % 4 == 0);
  }
  // EPI_HADAMARD: the factors H[hm(r(m)) + hn(r(n))] of the tile.  Runs of
  // equal keys r(.) along the tile's rows / columns share a slot; when the
  // slot table (rows x columns, FP32) fits the LDS it is gathered once per
  // tile, else every factor is read from H.  hrow / hcol: per-lane table
  // offsets (or H offsets when not staged).
  int hrow[MI], hcol[NI][4];
  bool staged = false;
  float* ht = nullptr;
  if (EPI == EPI_HADAMARD && g.hfull) {
    // precomputed factor table: the row / column offsets of each lane's
    // elements; no staging, no barrier
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      const int64_t m = min(m0 + wm0 + i * 16 + fr, g.M - 1);
      hrow[i] = (int)(((uint32_t)m % (uint32_t)g.hmod) * (uint32_t)g.hld);
    }
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) hcol[j][r] = (int)((uint32_t)min(n_first[j] + r, g.N - 1) % (uint32_t)g.hmod);
  } else if constexpr (EPI == EPI_HADAMARD) {
    constexpr int NW = NT / 64;
    int* rslot = (int*)lds;   // [BM] slot of each tile row
    int* cslot = rslot + BM;  // [BN]
    int* rkey = cslot + BN;   // [BM] key of each row slot
    int* ckey = rkey + BM;    // [BN]
    int* tot = ckey + BN;     // [NW] slots opened per wave
    ht = (float*)(tot + 16);
    constexpr int CAP = (int)((sizeof(TI) * (DMA ? 1 : 2) * STAGE) / 4) - 2 * (BM + BN) - 16;
    auto key_of = [&](int64_t x) -> int {
      return g.hidx ? g.hidx[(uint32_t)x % (uint32_t)g.hmod] : (int)x;
    };
    // thread t < BM: row t; BM <= t < BM + BN: column t - BM (BM is a multiple of 64)
    const bool is_row = tid < BM, is_col = !is_row && tid < BM + BN;
    int key = 0;
    bool open = false;
    if (is_row || is_col) {
      const int t = is_row ? tid : tid - BM;
      const int64_t base = is_row ? m0 : n0, last = is_row ? g.M - 1 : g.N - 1;
      key = key_of(min(base + t, last));
      open = t == 0 || key != key_of(min(base + t - 1, last));
    }
    const uint64_t bal = __ballot(open);
    const int below = __popcll(bal & ((UINT64_C(1) << lane) - 1));
    if (lane == 0) tot[wave] = __popcll(bal);
    __syncthreads();
    int R = 0, Cn = 0, off = 0;
    for (int w = 0; w < NW; ++w) {
      if (w < BM / 64) R += tot[w];
      else if (w < (BM + BN) / 64) Cn += tot[w];
      if (w < wave && (is_row ? true : w >= BM / 64)) off += tot[w];
    }
    const int slot = off + below + (open ? 1 : 0) - 1;
    if (is_row) {
      rslot[tid] = slot;
      if (open) rkey[slot] = key;
    } else if (is_col) {
      cslot[tid - BM] = slot;
      if (open) ckey[slot] = key;
    }
    __syncthreads();
    staged = R * Cn <= CAP;
    if (staged) {
      // one wave per table row (its H row offset computed once), lanes along
      // the columns: no index division, coalesced gathers along an H row
      for (int r = wave; r < R; r += NW) {
        const int64_t hr = g.hm(rkey[r]);
        for (int c = lane; c < Cn; c += 64) ht[r * Cn + c] = (float)g.hsrc[hr + g.hn(ckey[c])];
      }
    }
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      const int rs = rslot[wm0 + i * 16 + fr];
      hrow[i] = staged ? rs * Cn : (int)g.hm(rkey[rs]);
    }
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int cs = cslot[wn0 + j * 16 + 4 * fq + r];
        hcol[j][r] = staged ? cs : (int)g.hn(ckey[cs]);
      }
    __syncthreads();
  }
  // 16-bit outputs: 4 consecutive columns of row offset cmo, as one 8-byte
  // vector when contiguous and aligned; with c_lo also the lo part
  // fp16((v - hi) * 2^11) at offset c_lo.
  auto put4 = [&](int64_t cmo, bool c_al, int j, const float* v) {
    TO hi[4], lo[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      hi[r] = (TO)v[r];
      lo[r] = (TO)((v[r] - (float)hi[r]) * 2048.f);
    }
    if (cn_vec[j] && c_al && (g.c_lo % 4 == 0)) {
      Vec4<TO> w;
#pragma unroll
      for (int r = 0; r < 4; ++r) w.v[r] = hi[r];
      *(Vec4<TO>*)(C + cmo + cn_off[j]) = w;
      if (g.c_lo) {
#pragma unroll
        for (int r = 0; r < 4; ++r) w.v[r] = lo[r];
        *(Vec4<TO>*)(C + cmo + cn_off[j] + g.c_lo) = w;
      }
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (n_first[j] + r >= g.N) continue;
        const int64_t o = cmo + g.cn(n_first[j] + r);
        C[o] = hi[r];
        if (g.c_lo) C[o + g.c_lo] = lo[r];
      }
    }
  };
#pragma unroll
  for (int i = 0; i < MI; ++i) {
    const int64_t m = m0 + wm0 + i * 16 + fr;
    if (m >= g.M) continue;
    const int64_t cmo = g.cm(m);
    const bool c_al = (cmo % 4 == 0) && ((g.c_batch * zb) % 4 == 0);
    if constexpr (EPI == EPI_HADAMARD) {
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        if (n_first[j] >= g.N) continue;
        float v[4];
        if (g.hfull) {
          const float* hf = g.hfull + hrow[i];
          float f4[4];
          if (hcol[j][0] % 4 == 0 && hcol[j][3] == hcol[j][0] + 3) {
            const v4f t = *(const v4f*)(hf + hcol[j][0]);
#pragma unroll
            for (int r = 0; r < 4; ++r) f4[r] = t[r];
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) f4[r] = hf[hcol[j][r]];
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = alpha * acc[i][j][r] * f4[r];
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int h = hrow[i] + hcol[j][r];
            v[r] = alpha * acc[i][j][r] * (staged ? ht[h] : (float)g.hsrc[h]);
          }
        }
        put4(cmo, c_al, j, v);
      }
    } else if constexpr (EPI == EPI_DIAG) {
      // FP64 (or FP32) output: C = alpha acc + diag[dm(m) + dmap(n)] * X[z][xm(m) + xn(n)],
      // X FP64 (xsrc) or FP32 (xsrc32), the sum formed in FP64
      const double* xs = g.xsrc32 ? nullptr : g.xsrc + zb * g.x_batch;
      const float* xs32 = g.xsrc32 ? g.xsrc32 + zb * g.x_batch : nullptr;
      const int64_t dmo = g.dm(m), xmo = g.xm(m);
#pragma unroll
      for (int j = 0; j < NI; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t n = n_first[j] + r;
          if (n >= g.N) continue;
          double v = (double)alpha * (double)acc[i][j][r];
          const int64_t xo = xmo + g.xn(n);
          v += g.diag[dmo + g.dmap(n)] * (xs32 ? (double)xs32[xo] : xs[xo]);
          C[cmo + g.cn(n)] = to_type<TO>(v);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        if (n_first[j] >= g.N) continue;
        float v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = alpha * acc[i][j][r];
        put4(cmo, c_al, j, v);
      }
    }
  }
}

// Whether the fast kernel applies: 16-bit inputs, both operands K-contiguous
// with plain row maps, 16-byte aligned rows, no split-K, and an epilogue it
// implements (no beta / batch scale).
inline bool gemm16_ok(DType in, DType out, Epi epi, bool a_kc, bool b_kc, const GemmArgs& g, int ksplit) {
  constexpr int64_t PLAIN = INT64_C(1) << 62;
  if (!(in == DT_F16 || in == DT_BF16) || !a_kc || !b_kc || ksplit > 1) return false;
  if (g.am.div < PLAIN || g.bn.div < PLAIN || g.ak.div < PLAIN || g.bk.div < PLAIN) return false;
  if (g.ak.s0 != 1 || g.bk.s0 != 1 || g.am.s0 % 8 || g.bn.s0 % 8 || g.a_batch % 8 || g.b_batch % 8) return false;
  if ((((uintptr_t)g.A) & 15) || (((uintptr_t)g.B) & 15)) return false;
  if (g.beta != 0.0) return false;
  // 32-bit element offsets within one batch entry of each operand
  if (g.M < 1 || g.N < 1 || (g.M - 1) * g.am.s0 + g.a_lo + g.K >= (INT64_C(1) << 32) ||
      (g.N - 1) * g.bn.s0 + g.b_lo + g.K >= (INT64_C(1) << 32))
    return false;
  // split operands: FP16, both split (lo rows 16-byte aligned), lo past hi
  if ((g.a_lo != 0) != (g.b_lo != 0) || g.a_lo < 0 || g.b_lo < 0 || g.c_lo < 0) return false;
  if ((g.a_lo || g.c_lo) && in != DT_F16) return false;
  if (g.a_lo && (g.a_lo % 8 || g.b_lo % 8 || g.a_lo < g.K || g.b_lo < g.K)) return false;
  if (g.c_lo && epi == EPI_DIAG) return false;
  if (epi == EPI_DIAG) return out == DT_F64 || (out == DT_F32 && g.xsrc32);
  // EPI_HADAMARD: a full factor table with 32-bit offsets, or plain H maps
  // with 32-bit offsets
  if (epi == EPI_HADAMARD && g.hfull &&
      (g.hmod < 1 || g.hld < g.hmod || g.hld % 4 || (((uintptr_t)g.hfull) & 15) ||
       (g.hmod - 1) * g.hld + g.hmod >= (INT64_C(1) << 31)))
    return false;
  if (epi == EPI_HADAMARD && !g.hfull &&
      (g.hm.div < PLAIN || g.hn.div < PLAIN || g.hn.s0 != 1 || g.hm.s0 < 0 || g.hm.s0 > 46340 ||
       (g.hidx && (g.hmod < 1 || g.hmod > INT32_MAX))))
    return false;
  return (in == DT_F16 && out == DT_F16) || (in == DT_BF16 && out == DT_BF16);
}

// Tile choice: the big tile when it gives at least ~2 workgroups per CU
// (256 CUs), else the 128x128 tile.
inline bool gemm16_big(const GemmArgs& g) {
  const int64_t tiles = ceil_div(g.M, 256) * ceil_div(g.N, 256) * (int64_t)g.batch;
  return tiles >= 512;
}

bool lds_dma_on();  // TINCR_TDA_LDS_DMA (default on)

template <typename TI, typename TO, int EPI, typename Cfg>
void launch16_cfg(GemmArgs& g, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(g.N, Cfg::BN), (unsigned)ceil_div(g.M, Cfg::BM), (unsigned)g.batch);
  // with whole K slices the tiles load them straight into LDS
  constexpr bool dma_cfg = true;
  const bool dma = dma_cfg && g.K % 64 == 0 && lds_dma_on();
  if constexpr (std::is_same<TI, _Float16>::value) {
    if (g.a_lo) {
      if constexpr (dma_cfg) {
        if (dma) {
          gemm16_nt_kernel<TI, TO, Cfg, EPI, true, true><<<grid, Cfg::NT, 0, s>>>(g);
          TDA_LAUNCH_CHECK();
          return;
        }
      }
      gemm16_nt_kernel<TI, TO, Cfg, EPI, true><<<grid, Cfg::NT, 0, s>>>(g);
      TDA_LAUNCH_CHECK();
      return;
    }
  }
  if constexpr (dma_cfg) {
    if (dma) {
      gemm16_nt_kernel<TI, TO, Cfg, EPI, false, true><<<grid, Cfg::NT, 0, s>>>(g);
      TDA_LAUNCH_CHECK();
      return;
    }
  }
  gemm16_nt_kernel<TI, TO, Cfg, EPI, false><<<grid, Cfg::NT, 0, s>>>(g);
  TDA_LAUNCH_CHECK();
}

// tile: 1 = 256x256, 0 = 128x128, -1 = by tile count (gemm16_big).  The
// diagonal epilogue (FP64 or FP32 output) always takes the 128x128 tile (the
// 256x256 one spills: with FP32 output its tuning trials wrote 12 GB of
// scratch per launch; a 256x128 tile measured 7 % faster on one 2048 x 1024 x
// 6144 x 8 product but no faster over a solve's mix of batch sizes).
template <typename TI, typename TO, int EPI>
void launch16_t(GemmArgs& g, hipStream_t s, int tile) {
  if constexpr (EPI != EPI_DIAG) {
    if (tile > 0 || (tile < 0 && gemm16_big(g))) {
      launch16_cfg<TI, TO, EPI, Cfg16Big>(g, s);
      return;
    }
  }
  launch16_cfg<TI, TO, EPI, Cfg16Mid>(g, s);
}

// Epilogue / output dispatch for one input type (gemm16_ok has been checked).
template <typename TI>
void launch16_in(Epi epi, GemmArgs& g, hipStream_t s, int tile) {
  if (epi == EPI_HADAMARD) launch16_t<TI, TI, EPI_HADAMARD>(g, s, tile);
  else if (epi == EPI_DIAG && g.xsrc32) launch16_t<TI, float, EPI_DIAG>(g, s, tile);
  else if (epi == EPI_DIAG) launch16_t<TI, double, EPI_DIAG>(g, s, tile);
  else launch16_t<TI, TI, EPI_STORE>(g, s, tile);
}

}  // namespace tda
