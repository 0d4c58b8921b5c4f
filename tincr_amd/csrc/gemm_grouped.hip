// Grouped instantiations of the MFMA GEMM (gemm_impl.h): the products of a
// batch of fragments, one persistent launch per operation (batch.hip).
#include "gemm_impl.h"
#include "internal.h"

namespace tda {

namespace {
template <typename TI, typename TO, int EPI, typename Cfg, bool AK, bool BK>
void gl(const GroupSlot* slot, int grid, hipStream_t s) {
  gemm_grouped_kernel<TI, TO, Cfg, AK, BK, EPI><<<grid, 256, 0, s>>>(slot);
  TDA_LAUNCH_CHECK();
}
template <typename TI, typename TO, int EPI, bool AK, bool BK>
void gl_cfg(int cfg, const GroupSlot* slot, int grid, hipStream_t s) {
  switch (cfg) {
    case CFG_SKINNY_N: gl<TI, TO, EPI, CfgSkinnyN, AK, BK>(slot, grid, s); break;
    case CFG_SKINNY_M: gl<TI, TO, EPI, CfgSkinnyM, AK, BK>(slot, grid, s); break;
    case CFG_SMALL: gl<TI, TO, EPI, CfgSmall, AK, BK>(slot, grid, s); break;
    default: throw Error{2, "grouped gemm: unsupported tile configuration"};
  }
}
}  // namespace

int grouped_cfg_bm(int cfg) { return cfg_bm(cfg); }
int grouped_cfg_bn(int cfg) { return cfg_bn(cfg); }

// The combinations the batched Davidson uses (batch.hip):
//   FP64 in / FP64 out: store with (A, B) K-contiguous (tt), (t, f), (f, f);
//     the diagonal epilogue with (f, f)
//   FP32 in: store to FP64 (tt), store to FP32 (tt), diagonal to FP64 (ff)
void grouped_gemm_launch(DType in, DType out, Epi epi, bool a_kc, bool b_kc, int cfg, const GroupSlot* slot, int grid,
                         hipStream_t s) {
  const int kc = (a_kc ? 2 : 0) + (b_kc ? 1 : 0);
  if (in == DT_F64 && out == DT_F64 && epi == EPI_STORE) {
    if (kc == 3) gl_cfg<double, double, EPI_STORE, true, true>(cfg, slot, grid, s);
    else if (kc == 2) gl_cfg<double, double, EPI_STORE, true, false>(cfg, slot, grid, s);
    else if (kc == 0) gl_cfg<double, double, EPI_STORE, false, false>(cfg, slot, grid, s);
    else throw Error{2, "grouped gemm: unsupported operand layout"};
  } else if (in == DT_F64 && out == DT_F64 && epi == EPI_DIAG && kc == 0) {
    gl_cfg<double, double, EPI_DIAG, false, false>(cfg, slot, grid, s);
  } else if (in == DT_F32 && out == DT_F64 && epi == EPI_STORE && kc == 3) {
    gl_cfg<float, double, EPI_STORE, true, true>(cfg, slot, grid, s);
  } else if (in == DT_F32 && out == DT_F32 && epi == EPI_STORE && kc == 3) {
    gl_cfg<float, float, EPI_STORE, true, true>(cfg, slot, grid, s);
  } else if (in == DT_F32 && out == DT_F64 && epi == EPI_DIAG && kc == 0) {
    gl_cfg<float, double, EPI_DIAG, false, false>(cfg, slot, grid, s);
  } else {
    throw Error{2, "grouped gemm: unsupported type / epilogue combination"};
  }
}

}  // namespace tda
