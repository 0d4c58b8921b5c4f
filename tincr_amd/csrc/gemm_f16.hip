// FP16-input instantiations of the MFMA GEMM.
#include "gemm_impl.h"

namespace tda {

void gemm_dispatch_f16(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s) {
  if (epi != EPI_STORE) throw Error{2, "gemm f16: only the store epilogue"};
  if (out == DT_F16) launch_t<_Float16, _Float16, EPI_STORE>(cfg, a_kc, b_kc, g, s);
  else if (out == DT_F64) launch_t<_Float16, double, EPI_STORE>(cfg, a_kc, b_kc, g, s);
  else throw Error{2, "gemm f16: unsupported output"};
}

}  // namespace tda
