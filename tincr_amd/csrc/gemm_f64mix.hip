// FP64 MFMA products with FP32-stored operands (Davidson basis kept as FP32 in
// the FP16 phase): A and / or B are floats in memory, widened to FP64 when
// each K slice is written to LDS; accumulation and the result are FP64.
#include "gemm_impl.h"

namespace tda {

void gemm_dispatch_f64mix(int epi, DType out, int cfg, bool a_kc, bool b_kc, GemmArgs& g, hipStream_t s) {
  if (out != DT_F64 || epi != EPI_STORE) throw Error{2, "gemm f64 (FP32-stored operands): FP64 store only"};
  if (g.a_f32 && g.b_f32) launch_t<double, double, EPI_STORE, float, float>(cfg, a_kc, b_kc, g, s);
  else if (g.a_f32) launch_t<double, double, EPI_STORE, float, double>(cfg, a_kc, b_kc, g, s);
  else launch_t<double, double, EPI_STORE, double, float>(cfg, a_kc, b_kc, g, s);
}

}  // namespace tda
