// MFMA GEMM: planning, dispatch and the split-K reduction.
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "gemm16_impl.h"

#ifndef TDA_BIG_K
#define TDA_BIG_K 0
#endif

namespace tda {

GemmPlan gemm_plan(DType in, int64_t M, int64_t N, int64_t K, int batch, bool, bool, int ksplit_hint) {
  GemmPlan p;
  if (in == DT_F64 && N <= 16 && K >= 1024) p.cfg = CFG_SKINNY_N16;
  else if (N <= 32 && M > 32) p.cfg = CFG_SKINNY_N;
  else if (M <= 32) p.cfg = CFG_SKINNY_M;
  else if (M * N <= (int64_t)256 * 256 * 4) p.cfg = CFG_SMALL;
  else p.cfg = CFG_BIG;
  const int bk = bk_of(in, p.cfg);
  const int64_t tiles = ceil_div(M, cfg_bm(p.cfg)) * ceil_div(N, cfg_bn(p.cfg)) * batch;
  int ks = 1;
  if (ksplit_hint > 0) {
    ks = ksplit_hint;
  } else {
    // split K only when the tiles alone cannot fill the 256 CUs; then aim
    // for >= ~1024 workgroups while keeping >= 8 K-tiles per split
    const int64_t kt = ceil_div(K, bk);
    if (tiles < 512)
      while (tiles * ks < 1024 && kt / (ks * 2) >= 8) ks *= 2;
  }
  const int64_t kt = ceil_div(K, bk);
  p.k_chunk = ceil_div(kt, ks) * bk;
  p.ksplit = (int)ceil_div(K, p.k_chunk);
  if (p.ksplit < 1) p.ksplit = 1;
  return p;
}

size_t gemm_workspace_bytes(DType in, const GemmPlan& p, int64_t M, int64_t N, int batch) {
  if (p.ksplit <= 1) return 0;
  return (size_t)p.ksplit * batch * M * N * acc_bytes(in);
}


template <int EPI>
static void dispatch_reduce(DType in, DType out, GemmArgs& g, hipStream_t s) {
  const bool acc64 = (in == DT_F64);
  if (acc64) {
    if (out != DT_F64) throw Error{2, "gemm reduce: f64 input needs f64 output"};
    launch_reduce<double, double, EPI>(g, s);
    return;
  }
  if (out == DT_F64) {
    if constexpr (EPI == EPI_HADAMARD) throw Error{2, "gemm reduce: unsupported epilogue"};
    else launch_reduce<float, double, EPI>(g, s);
    return;
  }
  if constexpr (EPI == EPI_DIAG) {
    throw Error{2, "gemm reduce: unsupported epilogue"};
  } else {
    if (out == DT_F32) launch_reduce<float, float, EPI>(g, s);
    else if (out == DT_F16) launch_reduce<float, _Float16, EPI>(g, s);
    else launch_reduce<float, __bf16, EPI>(g, s);
  }
}

// Tile autotuner of the 16-bit kernel.  The first launch of each problem
// (input type, epilogue, split, M, N, K, batch, output strides) times the
// 256x256 and the 128x128 tile on the caller's stream (one warm-up and one
// timed launch each; the launches only write C, so repeating them is
// harmless, and every tile sums each output element in the same K order,
// so the choice does not change results) and keeps the faster one for the
// process.  Problems under 1e10 MFMA flops (a few microseconds; the many
// distinct shapes of small fragments) keep the tile-count rule, as does
// TINCR_TDA_AUTOTUNE=0.
// TINCR_TDA_LDS_DMA=0: the 16-bit big tile stages its K slices through
// VGPRs instead of loading them straight into LDS (read per call)
bool lds_dma_on() {
  const char* e = getenv("TINCR_TDA_LDS_DMA");
  return !(e && e[0] == '0');
}

namespace {
using Tune16Key = std::tuple<int, int, bool, int64_t, int64_t, int64_t, int, int64_t, int64_t>;
std::mutex tune16_mu;
std::map<Tune16Key, int> tune16_cache;

bool autotune16_on() {
  static const bool on = [] {
    const char* e = getenv("TINCR_TDA_AUTOTUNE");
    return !(e && e[0] == '0');
  }();
  return on;
}

void dispatch16(DType in, Epi epi, GemmArgs& g, hipStream_t s, int tile) {
  if (in == DT_F16) gemm16_dispatch_f16(epi, g, s, tile);
  else gemm16_dispatch_bf16(epi, g, s, tile);
}

int tune16(DType in, Epi epi, GemmArgs& g, hipStream_t s) {
  // TINCR_TDA_TILE16=0 / 1 forces the 128x128 / 256x256 tile (tests of the
  // tiles on small problems; the FP64-output epilogue stays 128x128); read
  // per call so that a process can switch
  const char* ft = getenv("TINCR_TDA_TILE16");
  if (ft && (ft[0] == '0' || ft[0] == '1') && ft[1] == 0) return ft[0] - '0';
  // the diagonal epilogue runs on the 128x128 tile (the 256x256 one spills)
  if (epi == EPI_DIAG || !autotune16_on()) return -1;
  const double flops = 2.0 * (double)g.M * (double)g.N * (double)g.K * g.batch * (g.a_lo ? 3.0 : 1.0);
  if (flops < 1e10) return -1;
  const Tune16Key key{(int)in, (int)epi, g.a_lo != 0, g.M, g.N, g.K, g.batch, g.cm.s0, g.c_lo};
  {
    std::lock_guard<std::mutex> lk(tune16_mu);
    auto it = tune16_cache.find(key);
    if (it != tune16_cache.end()) return it->second;
  }
  hipEvent_t ev[4];
  for (auto& e : ev) TDA_HIP(hipEventCreate(&e));
  float ms[2] = {0.f, 0.f};
  try {
    for (int tile = 0; tile < 2; ++tile) {
      dispatch16(in, epi, g, s, tile);  // warm-up (code object load, caches)
      TDA_HIP(hipEventRecord(ev[2 * tile], s));
      dispatch16(in, epi, g, s, tile);
      TDA_HIP(hipEventRecord(ev[2 * tile + 1], s));
    }
    TDA_HIP(hipEventSynchronize(ev[3]));
    for (int tile = 0; tile < 2; ++tile) TDA_HIP(hipEventElapsedTime(&ms[tile], ev[2 * tile], ev[2 * tile + 1]));
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  const int choice = ms[1] < ms[0] ? 1 : 0;
  std::lock_guard<std::mutex> lk(tune16_mu);
  tune16_cache[key] = choice;
  return choice;
}
}  // namespace

void gemm_launch(DType in, DType out, Epi epi, bool a_kc, bool b_kc, GemmArgs g, const GemmPlan& plan,
                 hipStream_t stream, hipEvent_t after_main) {
  TDA_CHECK(g.M >= 0 && g.N >= 0 && g.K >= 0, 2, "gemm: negative extent");
  if (g.M == 0 || g.N == 0) return;
  if (g.batch < 1) g.batch = 1;
  // 16-bit inputs with both operands K-contiguous: the dedicated kernel (no split-K)
  if (gemm16_ok(in, out, epi, a_kc, b_kc, g, 1)) {
    g.ksplit = 1;
    g.k_chunk = g.K;
    dispatch16(in, epi, g, stream, tune16(in, epi, g, stream));
    if (after_main) TDA_HIP(hipEventRecord(after_main, stream));
    return;
  }
  TDA_CHECK(!g.a_lo && !g.b_lo && !g.c_lo, 2, "gemm: split operands need the 16-bit K-contiguous kernel");
  TDA_CHECK(!(g.a_f32 || g.b_f32) || in == DT_F64, 2, "gemm: FP32-stored operands need FP64 products");
  GemmPlan pl = plan;
  // 16-bit inputs on the big tile: 128-deep K slices when the K maps allow
  // (TDA_BIG_K; off by default: it spills registers at two workgroups per CU)
  if (TDA_BIG_K && pl.cfg == CFG_BIG && (in == DT_F16 || in == DT_BF16) && g.K >= 256 &&
      (g.ak.div >= g.K || g.ak.div % 128 == 0) && (g.bk.div >= g.K || g.bk.div % 128 == 0)) {
    pl.cfg = CFG_BIG_K;
    pl.k_chunk = round_up(pl.k_chunk, 128);
    pl.ksplit = (int)ceil_div(g.K, pl.k_chunk);  // <= the planned split: the workspace suffices
  }
  // 64-deep K slices need 64-aligned K-map boundaries; otherwise the 16-deep
  // skinny tile (the planned k_chunk, a multiple of 64, stays valid)
  if (pl.cfg == CFG_SKINNY_N16 &&
      !((g.ak.div >= g.K || g.ak.div % CfgSkinnyN16::BK1 == 0) && (g.bk.div >= g.K || g.bk.div % CfgSkinnyN16::BK1 == 0)))
    pl.cfg = CFG_SKINNY_N;
  g.ksplit = pl.ksplit;
  g.k_chunk = pl.k_chunk;
  const int bk = bk_of(in, pl.cfg);
  TDA_CHECK(g.k_chunk % bk == 0, 2, "gemm: k_chunk must be a multiple of BK");
  // K tiles must not straddle a K-map boundary
  TDA_CHECK(g.ak.div >= g.K || g.ak.div % bk == 0, 2, "gemm: A K-map boundary not tile aligned");
  TDA_CHECK(g.bk.div >= g.K || g.bk.div % bk == 0, 2, "gemm: B K-map boundary not tile aligned");
  // vector loads need unit stride along the contiguous dimension
  if (a_kc) TDA_CHECK(g.ak.s0 == 1, 2, "gemm: A K-contiguous needs ak.s0 == 1");
  else TDA_CHECK(g.am.s0 == 1 && g.am.div >= g.M, 2, "gemm: A M-contiguous needs plain unit am");
  if (b_kc) TDA_CHECK(g.bk.s0 == 1, 2, "gemm: B K-contiguous needs bk.s0 == 1");
  else TDA_CHECK(g.bn.s0 == 1 && g.bn.div >= g.N, 2, "gemm: B N-contiguous needs plain unit bn");
  if (pl.ksplit > 1) TDA_CHECK(g.ws != nullptr, 2, "gemm: split-K needs a workspace");
  switch (in) {
    case DT_F64:
      if (g.a_f32 || g.b_f32) gemm_dispatch_f64mix(epi, out, pl.cfg, a_kc, b_kc, g, stream);
      else gemm_dispatch_f64(epi, out, pl.cfg, a_kc, b_kc, g, stream);
      break;
    case DT_F32: gemm_dispatch_f32(epi, out, pl.cfg, a_kc, b_kc, g, stream); break;
    case DT_F16: gemm_dispatch_f16(epi, out, pl.cfg, a_kc, b_kc, g, stream); break;
    default: gemm_dispatch_bf16(epi, out, pl.cfg, a_kc, b_kc, g, stream); break;
  }
  if (after_main) TDA_HIP(hipEventRecord(after_main, stream));
  if (pl.ksplit > 1) {
    if (epi == EPI_DIAG) dispatch_reduce<EPI_DIAG>(in, out, g, stream);
    else if (epi == EPI_HADAMARD) dispatch_reduce<EPI_HADAMARD>(in, out, g, stream);
    else dispatch_reduce<EPI_STORE>(in, out, g, stream);
  }
}

}  // namespace tda
