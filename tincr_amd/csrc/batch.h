// Batch of fragments shared by the worker-pool driver (batch.hip) and the
// grouped driver (batch_grouped.hip).
#pragma once
#include <string>
#include <vector>

#include "../../include/tincr_tda.h"

namespace tda {

struct BatchItem {
  bool ct = false;
  // monomer / system (tda_set_system arguments)
  int n_atoms = 0, n_ao = 0, ob = 0, o = 0, vb = 0, v = 0;
  const int32_t* opa = nullptr;
  const double *orbs = nullptr, *s = nullptr, *orbe = nullptr, *gamma = nullptr, *gamma_lr = nullptr;
  // CT pair (tda_set_ct_pair arguments)
  int n_atoms_l = 0, n_ao_l = 0;
  const int32_t* opa_l = nullptr;
  const double *orbs_l = nullptr, *s_l = nullptr, *orbe_l = nullptr, *s_hl = nullptr, *gamma_l = nullptr,
               *gamma_hl = nullptr;
  const double* xyz = nullptr;
  // solve and outputs
  int n_roots = 0, max_iter = 0;
  double tol = 0;
  double *omega = nullptr, *X = nullptr, *q_trans = nullptr, *tr_dipoles = nullptr, *f = nullptr;
  double cost = 0;
};

// Solves every item with the grouped kernel set (batch_grouped.hip); status
// and stats as tda_batch_run.  Returns TDA_OK or the first error other than
// TDA_ERR_NOT_CONVERGED (per-item statuses carry that one); err receives a
// message.
// totals (optional): the batch's exchange-kernel times / launches / flops,
// the largest iteration count, summed products, wall time.
int batch_run_grouped(int device, int precision, size_t budget, std::vector<BatchItem>& items, int32_t* status,
                      tda_stats* stats, tda_stats* totals, std::string& err);

}  // namespace tda

struct tda_batch {
  int device = 0;
  int n_workers = 1;
  size_t budget = 0;
  int precision = TDA_PREC_FP64;
  int op = TDA_OP_AO;
  int mode = TDA_BATCH_GROUPED;
  std::vector<tda::BatchItem> items;
  std::string err;
  tda_stats last{};  // totals of the last grouped run
};
