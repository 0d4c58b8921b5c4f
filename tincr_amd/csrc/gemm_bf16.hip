// BF16-input instantiations of the MFMA GEMM.
#include "gemm_impl.h"

namespace tda {

void gemm_dispatch_bf16(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s) {
  if (epi != EPI_STORE) throw Error{2, "gemm bf16: only the store epilogue"};
  if (out == DT_BF16) launch_t<__bf16, __bf16, EPI_STORE>(cfg, a_kc, b_kc, g, s);
  else if (out == DT_F64) launch_t<__bf16, double, EPI_STORE>(cfg, a_kc, b_kc, g, s);
  else throw Error{2, "gemm bf16: unsupported output"};
}

}  // namespace tda
