// Batched excited-state solves of many fragments: the FMO monomer loop
// (src/fmo/lcmo/basis.rs:462-465, prepare_tda + run_tda per monomer on a rayon
// pool) and the charge-transfer pair loop (basis.rs:172-238, run_ct_tda per
// ordered pair).  The C entry points and the worker-pool driver
// (TDA_BATCH_POOL): each worker thread owns one context (one HIP stream) and
// takes the next unsolved fragment, largest estimated cost first, so that the
// fragments' kernels overlap on the GPU and the long solves start early.
// The default driver is batch_grouped.hip.  Every fragment converges (or
// fails) on its own; the others are unaffected.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tincr_tda.h"
#include "batch.h"

using tda::BatchItem;


namespace {

// A.x flops of one vector, used only to order the fragments
double estimate_cost(int n_atoms, int o, int v) {
  return 2.0 * n_atoms * (double)o * v * (o + v) + 4.0 * n_atoms * (double)o * v;
}

int add_item(tda_batch* b, BatchItem&& it, int n_roots, int max_iter, double tol) {
  if (n_roots < 1 || n_roots > 64 || (int64_t)n_roots > (int64_t)it.o * it.v || max_iter < 1 || tol <= 0.0) {
    b->err = "bad n_roots / max_iter / tolerance";
    return -TDA_ERR_BAD_ARG;
  }
  it.n_roots = n_roots;
  it.max_iter = max_iter;
  it.tol = tol;
  b->items.push_back(std::move(it));
  return (int)b->items.size() - 1;
}

}  // namespace

extern "C" {

int tda_batch_create(int device, int n_workers, size_t hbm_budget_bytes, int precision, tda_batch** out) {
  if (!out || n_workers < 1 || n_workers > 64) return TDA_ERR_BAD_ARG;
  if (precision < TDA_PREC_FP64 || precision > TDA_PREC_FP16) return TDA_ERR_BAD_ARG;
  tda_batch* b = new (std::nothrow) tda_batch();
  if (!b) return TDA_ERR_OOM;
  b->device = device;
  b->n_workers = n_workers;
  b->budget = hbm_budget_bytes;
  b->precision = precision;
  *out = b;
  return TDA_OK;
}

void tda_batch_destroy(tda_batch* b) { delete b; }

const char* tda_batch_last_error(const tda_batch* b) { return b ? b->err.c_str() : "NULL batch"; }

int tda_batch_set_operator(tda_batch* b, int op) {
  if (!b || (op != TDA_OP_AO && op != TDA_OP_CHARGES)) return TDA_ERR_BAD_ARG;
  b->op = op;
  return TDA_OK;
}

int tda_batch_last_totals(const tda_batch* b, tda_stats* out) {
  if (!b || !out) return TDA_ERR_BAD_ARG;
  *out = b->last;
  return TDA_OK;
}

int tda_batch_set_mode(tda_batch* b, int mode) {
  if (!b || (mode != TDA_BATCH_GROUPED && mode != TDA_BATCH_POOL)) return TDA_ERR_BAD_ARG;
  b->mode = mode;
  return TDA_OK;
}

int tda_batch_add_system(tda_batch* b, int n_atoms, int n_ao, const int32_t* opa, const double* orbs,
                         const double* s, const double* orbe, int occ_begin, int n_occ, int virt_begin, int n_virt,
                         const double* gamma, const double* gamma_lr, const double* xyz, int n_roots, int max_iter,
                         double tol, double* omega, double* X, double* q_trans, double* tr_dipoles, double* f) {
  if (!b) return -TDA_ERR_BAD_ARG;
  BatchItem it;
  it.n_atoms = n_atoms;
  it.n_ao = n_ao;
  it.opa = opa;
  it.orbs = orbs;
  it.s = s;
  it.orbe = orbe;
  it.ob = occ_begin;
  it.o = n_occ;
  it.vb = virt_begin;
  it.v = n_virt;
  it.gamma = gamma;
  it.gamma_lr = gamma_lr;
  it.xyz = xyz;
  it.omega = omega;
  it.X = X;
  it.q_trans = q_trans;
  it.tr_dipoles = tr_dipoles;
  it.f = f;
  it.cost = estimate_cost(n_atoms, n_occ, n_virt);
  return add_item(b, std::move(it), n_roots, max_iter, tol);
}

int tda_batch_add_ct_pair(tda_batch* b, int n_atoms_h, int n_ao_h, const int32_t* opa_h, const double* orbs_h,
                          const double* s_h, const double* orbe_h, int occ_begin_h, int n_occ_h, int n_atoms_l,
                          int n_ao_l, const int32_t* opa_l, const double* orbs_l, const double* s_l,
                          const double* orbe_l, int virt_begin_l, int n_virt_l, const double* s_hl,
                          const double* gamma_h, const double* gamma_l, const double* gamma_hl,
                          const double* gamma_lr_hl, const double* xyz, int n_roots, int max_iter, double tol,
                          double* omega, double* X, double* q_trans, double* tr_dipoles, double* f) {
  if (!b) return -TDA_ERR_BAD_ARG;
  BatchItem it;
  it.ct = true;
  it.n_atoms = n_atoms_h;
  it.n_ao = n_ao_h;
  it.opa = opa_h;
  it.orbs = orbs_h;
  it.s = s_h;
  it.orbe = orbe_h;
  it.ob = occ_begin_h;
  it.o = n_occ_h;
  it.n_atoms_l = n_atoms_l;
  it.n_ao_l = n_ao_l;
  it.opa_l = opa_l;
  it.orbs_l = orbs_l;
  it.s_l = s_l;
  it.orbe_l = orbe_l;
  it.vb = virt_begin_l;
  it.v = n_virt_l;
  it.s_hl = s_hl;
  it.gamma = gamma_h;
  it.gamma_l = gamma_l;
  it.gamma_hl = gamma_hl;
  it.gamma_lr = gamma_lr_hl;
  it.xyz = xyz;
  it.omega = omega;
  it.X = X;
  it.q_trans = q_trans;
  it.tr_dipoles = tr_dipoles;
  it.f = f;
  it.cost = estimate_cost(n_atoms_h + n_atoms_l, n_occ_h, n_virt_l);
  return add_item(b, std::move(it), n_roots, max_iter, tol);
}

int tda_batch_clear(tda_batch* b) {
  if (!b) return TDA_ERR_BAD_ARG;
  b->items.clear();
  return TDA_OK;
}

int tda_batch_run(tda_batch* b, int32_t* status, tda_stats* stats) {
  if (!b) return TDA_ERR_BAD_ARG;
  const int n = (int)b->items.size();
  if (n == 0) return TDA_OK;
  if (b->mode == TDA_BATCH_GROUPED) {
    std::vector<int32_t> st(n, TDA_ERR_STATE);
    const int rc =
        tda::batch_run_grouped(b->device, b->precision, b->budget, b->items, st.data(), stats, &b->last, b->err);
    if (status) std::memcpy(status, st.data(), (size_t)n * sizeof(int32_t));
    if (rc != TDA_OK) return rc;
    for (int i = 0; i < n; ++i)
      if (st[i] != TDA_OK) {
        b->err = "fragment " + std::to_string(i) + ": Davidson Algorithm did not converge!";
        return st[i];
      }
    return TDA_OK;
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->items[a].cost > b->items[c].cost; });
  std::vector<int32_t> rcs(n, TDA_ERR_STATE);
  std::atomic<int> next{0};
  std::mutex mu;
  std::string first_err;
  int first_code = TDA_OK;
  auto note = [&](int code, const std::string& msg) {
    std::lock_guard<std::mutex> lock(mu);
    if (first_code == TDA_OK) {
      first_code = code;
      first_err = msg;
    }
  };
  auto worker = [&]() {
    tda_ctx* ctx = nullptr;
    int rc = tda_ctx_create(b->device, b->budget, b->precision, &ctx);
    if (rc == TDA_OK) rc = tda_set_operator(ctx, b->op);
    if (rc != TDA_OK) {
      note(rc, "tda_batch_run: context creation failed");
      if (ctx) tda_ctx_destroy(ctx);
      return;
    }
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= n) break;
      const int idx = order[k];
      const BatchItem& it = b->items[idx];
      if (it.ct)
        rc = tda_set_ct_pair(ctx, it.n_atoms, it.n_ao, it.opa, it.orbs, it.s, it.orbe, it.ob, it.o, it.n_atoms_l,
                             it.n_ao_l, it.opa_l, it.orbs_l, it.s_l, it.orbe_l, it.vb, it.v, it.s_hl, it.gamma,
                             it.gamma_l, it.gamma_hl, it.gamma_lr, it.xyz);
      else
        rc = tda_set_system(ctx, it.n_atoms, it.n_ao, it.opa, it.orbs, it.s, it.orbe, it.ob, it.o, it.vb, it.v,
                            it.gamma, it.gamma_lr, it.xyz);
      if (rc == TDA_OK)
        rc = tda_run(ctx, it.n_roots, it.max_iter, it.tol, it.omega, it.X, it.q_trans, it.tr_dipoles, it.f,
                     stats ? stats + idx : nullptr);
      rcs[idx] = rc;
      if (rc != TDA_OK && rc != TDA_ERR_NOT_CONVERGED)
        note(rc, "fragment " + std::to_string(idx) + ": " + tda_last_error(ctx));
    }
    tda_ctx_destroy(ctx);
  };
  const int nw = std::min(b->n_workers, n);
  std::vector<std::thread> pool;
  pool.reserve(nw);
  for (int w = 0; w < nw; ++w) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  if (status) std::memcpy(status, rcs.data(), (size_t)n * sizeof(int32_t));
  if (first_code != TDA_OK) {
    b->err = first_err;
    return first_code;
  }
  for (int i = 0; i < n; ++i)
    if (rcs[i] != TDA_OK) {
      b->err = "fragment " + std::to_string(i) + ": Davidson Algorithm did not converge!";
      return rcs[i];
    }
  return TDA_OK;
}

}  // extern "C"
