// Library-internal interfaces between translation units (not part of the C ABI).
#pragma once
#include "../../include/tincr_tda.h"
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
