// FP64-input instantiations of the MFMA GEMM.
#include "gemm_impl.h"

namespace tda {

void gemm_dispatch_f64(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s) {
  if (out != DT_F64) throw Error{2, "gemm f64: output must be f64"};
  if (epi == EPI_DIAG) launch_t<double, double, EPI_DIAG>(cfg, a_kc, b_kc, g, s);
  else if (epi == EPI_HADAMARD) launch_t<double, double, EPI_HADAMARD>(cfg, a_kc, b_kc, g, s);
  else launch_t<double, double, EPI_STORE>(cfg, a_kc, b_kc, g, s);
}

}  // namespace tda
