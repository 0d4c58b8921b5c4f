// FP16-input instantiations of the MFMA GEMM.
#include "gemm16_impl.h"

namespace tda {

void gemm_dispatch_f16(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s) {
  if (out == DT_F16) {
    if (epi == EPI_HADAMARD) launch_t<_Float16, _Float16, EPI_HADAMARD>(cfg, a_kc, b_kc, g, s);
    else if (epi == EPI_STORE) launch_t<_Float16, _Float16, EPI_STORE>(cfg, a_kc, b_kc, g, s);
    else throw Error{2, "gemm f16: unsupported epilogue"};
  } else if (out == DT_F64) {
    if (epi == EPI_DIAG) launch_t<_Float16, double, EPI_DIAG>(cfg, a_kc, b_kc, g, s);
    else if (epi == EPI_STORE) launch_t<_Float16, double, EPI_STORE>(cfg, a_kc, b_kc, g, s);
    else throw Error{2, "gemm f16: unsupported epilogue"};
  } else {
    throw Error{2, "gemm f16: unsupported output"};
  }
}

void gemm16_dispatch_f16(Epi epi, GemmArgs& g, hipStream_t s, int tile) { launch16_in<_Float16>(epi, g, s, tile); }

}  // namespace tda
