// Library-internal interfaces between translation units (not part of the C ABI).
#pragma once
#include "../../include/tincr_tda.h"
#include <functional>
#include <utility>
#include <vector>

#include "gemm.h"

namespace tda {

// Dimensions of the prepared per-atom-charges operator of a context (after
// tda_prepare on the charges form, or any CT pair): q_ov [N][o][v_p],
// gamma [N][N], G = gamma_lr q_oo [NL][o_p][o_p], q_vv [NL][v_p][v_p],
// de [o][v_p]; lc = false: no exchange (NL = 0).
struct ChargesDims {
  int N, NL, o, v, op, vp;
  bool lc;
};
int ctx_charges_dims(tda_ctx* c, ChargesDims* d);
// Device-to-device copies of the prepared operator into caller arrays (any
// may be NULL), finished before returning.
int ctx_export_charges(tda_ctx* c, double* qov, double* gamma, double* G, double* qvv, double* de);
const char* ctx_error(tda_ctx* c);

// Grouped GEMM (gemm_grouped.hip): slot->descs are independent problems
// (GemmArgs), slot->tiles[t] = {descriptor, batch index, tile row, tile
// column} for the tile shape of `cfg` (CfgId), slot->count tiles; `grid`
// persistent workgroups.  No split-K.
void grouped_gemm_launch(DType in, DType out, Epi epi, bool a_kc, bool b_kc, int cfg, const GroupSlot* slot, int grid,
                         hipStream_t s);
// tile shape of a CfgId (rows, columns)
int grouped_cfg_bm(int cfg);
int grouped_cfg_bn(int cfg);

}  // namespace tda

namespace tda {

// Prepared system of a context as the excited-state gradient reads it
// (gradient.hip): dimensions, the FP64 charges q_ov [N][o][v_p],
// q_oo [N][o_p][o_p], q_vv [N][v_p][v_p], gamma / gamma_lr [N][N]
// (gamma_lr NULL: no long-range correction), MO coefficients C and overlap S
// [n][n], orbital energies [n], de [o][v_p], AO offsets per atom (device and
// host, N + 1 entries) and the context's stream.  Occupied MOs 0..o-1 and
// virtual MOs o..n-1.
struct GradView {
  int N, n, o, v, op, vp, ob, vb;
  const double *qov, *qoo, *qvv, *gamma, *gamma_lr, *C, *S, *orbe, *de;
  const int* off;
  std::vector<int> off_h;
  hipStream_t stream;
  // device buffers kept between calls (owned by the context, freed with it)
  std::vector<std::pair<void*, size_t>>* cache;
};
struct GradIn {
  int exact;  // 0: the reference's LC response terms, 1: the exact ones
  double omega;
  const double *x, *dp, *ds_rows, *dh_rows, *dgamma, *dgamma_lr;  // host arrays
  int max_iter;
  double conv;
};
struct GradOut {
  double* grad;  // [3N] host
  double* z;     // [o*v] host or NULL
  int z_iter;
};
// apply_a(X, AX): the context's FP64 forward product A X of one vector in
// the device layout [o][v_p].
void gradient_run(const GradView& v, const std::function<void(const double*, double*)>& apply_a, const GradIn& in,
                  GradOut& out);

}  // namespace tda
