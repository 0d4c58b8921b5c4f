// C ABI, device context and the Davidson driver of the LC-TDA path.
//
// Device layouts (all row-major; o_p, v_p = n_occ, n_virt rounded up to 64):
//   trial / product vectors  X[k][o][v_p]           (ovp = o * v_p per vector)
//   q_ov                     [N][o][v_p]
//   q_oo, G = gamma_lr q_oo  [N][o_p][o_p]
//   q_vv                     [N][v_p][v_p]
//   exchange intermediate T  [k][v_p][NB*o_p]  per atom block of NB atoms
// Pads are zero, so every dot product over ovp equals the one over o*v.
#include <algorithm>
#include <array>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "../../include/tincr_tda.h"
#include "gemm.h"
#include "internal.h"
#include "kernels.h"

using namespace tda;

namespace {

// Stream of the context whose C-ABI call runs on this thread (set by
// guarded()): the only stream that can still read a buffer being regrown.
thread_local hipStream_t tl_stream = nullptr;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  // Grows the buffer; returns true when (re)allocated.  Work already queued on
  // the owning context's stream may still read the old allocation, so that
  // stream (not the whole device: other contexts' workers keep running) is
  // drained before it is freed.
  bool ensure(size_t n, hipStream_t stream = nullptr) {
    if (n <= bytes) return false;
    if (!stream) stream = tl_stream;
    if (p) TDA_HIP(stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize());
    release();
    if (n == 0) return false;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      bytes = 0;
      throw Error{TDA_ERR_OOM, "hipMalloc of " + std::to_string(n) + " bytes failed"};
    }
    bytes = n;
    return true;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return (T*)p; }
};

DType dtype_of(int prec) {
  switch (prec) {
    case PREC_FP64: return DT_F64;
    case PREC_FP32: return DT_F32;
    case PREC_BF16: return DT_BF16;
    default: return DT_F16;
  }
}

struct Timer {
  hipEvent_t a, b;
  Timer() {
    TDA_HIP(hipEventCreate(&a));
    TDA_HIP(hipEventCreate(&b));
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
};

}  // namespace

struct tda_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int prec = PREC_FP64;
  size_t budget = (size_t)16 << 30;
  std::string err;

  // system
  bool has_system = false, has_lc = false, prepared = false;
  int N = 0, nao = 0, ob = 0, o = 0, vb = 0, v = 0, op = 0, vp = 0;
  // atoms carrying q_oo (Nq) and q_vv (Nx); both N for a System/Monomer.  A
  // charge-transfer pair (ct_pair.rs) has q_oo on the hole fragment H (Nq =
  // N_H) and q_vv on the particle fragment L (Nx = N_L).
  int Nq = 0, Nx = 0;
  bool ct = false;
  int nao_l = 0, lumo_l = 0;
  DevBuf Cl, Sl, SCl, Shl, aooff_l, ctW, ctZ;
  // dense matrix engine (solvers/mod.rs:36-51): Hd [vp][vp] zero-padded,
  // stored as o = 1, v = dim; no preconditioner floor
  bool dense = false;
  DevBuf Hd, dxyz, post;
  DevBuf adjP, adjQ, adjX, adjD;  // adjoint work buffers
  DevBuf one;                     // device 1.0 values
  DevBuf zvw;                     // Z-vector work (R, Z, P, AP)
  std::vector<std::pair<void*, size_t>> grad_cache;  // tda_gradient's work buffers (gradient.hip)
  int last_roots = 0;             // CI vectors of the last tda_run held in RZ
  // thick restart of the FP16 / BF16 Davidson phase, in multiples of n_roots
  // (environment TINCR_TDA_LP_CAP / TINCR_TDA_LP_KEEP); the FP32 and FP64
  // phases keep the reference rule (cap 15 n, collapse to n)
  // FP16 / BF16 phase: basis cap 8 n (measured on SCC (H2O)512, 16 states:
  // caps 6 / 7 / 8 / 10 / 12 / 15 / 20 n gave 720 / 652 / 608 / 663 / 811 /
  // 842 / 873 ms per solve: the subspace work of a large basis costs more
  // than the extra restarts)
  // Small systems (product below 5e10 flops, e.g. C200H202: 1e10) use cap 6 n:
  // there the subspace eigensolver at m = 8 n (1.4 ms at m = 160) costs as
  // much as an iteration's products; measured on SCC C200H202, 20 states:
  // caps 4 / 5 / 6 / 8 n gave 64.5 / 61.0 / 60.1 / 62.7 ms per solve.
  // 0: by product cost (lp_cap_mult_for), else the TINCR_TDA_LP_CAP value.
  int lp_cap_mult = 0, lp_keep_mult = 2;
  int fp32_cap_mult = 15;  // basis cap of the FP32-accuracy phase, multiples of n
  int lp_keep(int n) const { return std::max(n, std::min(64, lp_keep_mult * n)); }
  // rows of the basis / subspace buffers: the reference's 15 n (davidson.rs:74)
  // or a larger cap of the low-precision phases
  int space(int n) const { return std::max(15, std::max(lp_cap_mult, fp32_cap_mult)) * n; }
  int lp_cap(int n, int mult) const { return std::min(space(n), std::max(mult * n, lp_keep(n) + n)); }
  int fp32_cap(int n) const { return std::min(space(n), std::max(fp32_cap_mult * n, lp_keep(n) + n)); }
  double pre_thr = 1e-4;
  // Coulomb factor of the charges-form product: 2 (solvers/mod.rs:311); the
  // no-LC Z-vector operator uses 4 (gradients/helpers.rs:1377)
  double coul_factor = 2.0;
  // precision ladder: floor factor of the low-precision phases, stall window
  double floor_k = 2.0;
  int stall_window = 0;  // 0: from the product cost (product_flops)
  bool fp32_thick = true;
  bool split32 = true;  // FP32-accuracy phase on FP16 hi / lo split products (AO operator)
  double ladder_min_flops = 1e9;  // per-product flops below which mixed contexts run FP64
  int64_t ovp = 0;
  std::vector<double> xyz, h_de;
  DevBuf C, S, SC, orbe, aooff, gamma, gamma_lr;
  // prepared operators (FP64)
  DevBuf qov, qoo, qvv, G, de;
  // operator form: 0 = AO-factored exchange, 1 = per-atom charges (three-stage)
  int op_kind = 0;
  bool mo_built = false;  // q_oo / q_vv present
  // q_ov present; the AO operator (LC, not a CT pair) does not build it:
  // its products, transition charges and adjoint work from the AO factors
  bool qov_built = false;
  // AO-factored operands: Rt [2n_ao][v_p], Lm [2n_ao][o_p], Rpt [v_p][2n_ao],
  // Lpt [o_p][2n_ao]; the AO-expanded gamma_lr is gathered through ao_atom
  DevBuf ao_atom, aoRt, aoLm, aoRpt, aoLpt, aoXR, aoM, aoF;
  // gamma_lr expanded to AO pairs, FP32 [n_ao][gfull_ld]: the Hadamard
  // factors of the 16-bit density GEMM (coalesced loads, no gather)
  DevBuf gfull;
  int64_t gfull_ld = 0;
  // Low-precision operand copies, one set per precision (index PREC_*): the
  // precision ladder of tda_run uses the context's precision and FP32.
  struct LpSet {
    bool ready = false;
    DevBuf qvv, G, Rt, Lm, Rpt, Lpt, scale;  // scale: [s_B*g_B | 1/s_B | 1/g_B] per atom
    // AO operator, FP16 types: power of two multiplying the first GEMM's
    // output XR = R x^T (and divided out of the last GEMM's)
    double xr_scale = 1.0;
  } lps[5];  // per PREC_* (index PREC_F16X: FP16 hi / lo split)
  DevBuf qov32, qvv_amax, G_amax;
  // workspaces
  DevBuf ws_gemm, T, xlp, u, uu, vv32, ax_tmp, stage, stage2;
  // Davidson state
  DevBuf V, AV, H, Y, evals, RZ, R, PW, norms, npart, idx, count, gs_scratch, eigw;
  DevBuf guess;  // int64 positions of the n guess unit vectors in the padded layout
  // P (preconditioned vectors) and W (their projections; the kept Ritz
  // vectors between phases) share one allocation, W = P + pw_rows L, so that
  // one GEMM reads both as the rows of [P; W]
  int64_t pw_rows = 0;
  double* Pb() { return PW.as<double>(); }
  double* Wb() { return PW.as<double>() + pw_rows * ovp; }
  bool gs_gram = true;  // block Gram-Schmidt in coefficient form (launch_gs_gram)
  bool no_gfull = false;  // TINCR_TDA_GFULL=0: 16-bit density GEMM gathers gamma_lr per tile instead
  bool ao_scale = false;  // TINCR_TDA_AO_SCALE=1: FP16 AO products with the power-of-two XR scale
  int gs_fallbacks = 0;  // blocks that fell back to the vector-by-vector kernel
  int reorth_passes = 0;  // second passes taken (log only)
  int norm_restarts = 0;  // FP64 verifications that found non-unit Ritz vectors
  // Pinned host staging of the Davidson's per-iteration scalars (a copy to
  // pageable memory would wait for the stream).  Gram-Schmidt blocks (at most
  // 64 vectors): [128, 192) |p_j|, [192, 256) |u_j|, [320, 322) as int:
  // counts, [384, 448) norms of the accepted vectors; then per root (pin_n
  // slots each): eigenvalues, residual norms, expansion indices (as int).
  double* hpin = nullptr;
  int pin_n = 0;
  void ensure_pinned(int n) {
    if (hpin && n <= pin_n) return;
    if (hpin) {
      TDA_HIP(hipStreamSynchronize(stream));
      (void)hipHostFree(hpin);
      hpin = nullptr;
    }
    pin_n = std::max(64, n);
    TDA_HIP(hipHostMalloc((void**)&hpin, (512 + 3 * (size_t)pin_n) * sizeof(double), hipHostMallocDefault));
  }
  double* pinned() {
    if (!hpin) ensure_pinned(64);
    return hpin;
  }
  double* pin_evals() { return pinned() + 512; }
  double* pin_res() { return pinned() + 512 + pin_n; }
  int* pin_idx() { return (int*)(pinned() + 512 + 2 * (size_t)pin_n); }
  // second Gram-Schmidt pass: 0 never, 1 every block, 2 when the estimated
  // loss of orthogonality of the basis would exceed eta_tol
  int reorth_mode = 2;
  double eta_tol = 1e-8;
  double eta_tol32 = 1e-5;  // the same for an FP32 basis (whose rounding alone is 6e-8)
  // near the precision floor of a low-precision phase (max residual below
  // near_k * delta) the phase hands over after near_stall iterations without
  // a 20 % improvement: there the products' error, not the subspace, limits
  // the residual
  double near_k = 4.0;
  int near_stall = 2;
  bool lp_phase = true;  // the FP16 / BF16 phase of the ladder (off: start at the FP32-accuracy phase)
  // FP16 / BF16 phase of the AO operator: the basis V and its products AV
  // stored as FP32 (the phase's product error, ~1e-4 per unit vector, lies
  // far above FP32 rounding); every product with them is formed in FP64.
  // Halves the bytes of the subspace GEMMs that stream V and AV.
  // -1: by product cost (FP32 basis for products of 5e10 flops and more: at
  // C200H202 the FP64 basis is 1.4 ms per solve faster), else 0 / 1
  int basis32 = -1;
  DevBuf qstage;  // FP64 staging of the new basis vectors of an FP32 basis
  DevBuf ws_basis;  // partial sums of the basis dot products (launch_rowdot)
  // instrumentation
  bool timing = true;
  bool verbose = false;
  double t_products_ms = 0;
  int n_products = 0, n_products_fp64 = 0, n_products_fp32 = 0, n_blocks = 1;
  // per exchange stage s (0: stage 1, 1: stage 3), per precision class p
  // (0: FP64, 1: low precision): main-kernel time, launches, flops
  double st_ms[2][3] = {}, st_flops[2][3] = {};
  int st_n[2][3] = {};
  struct Rec {
    hipEvent_t a, b;
    int stage, pclass;
    double flops;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> ev_pool;

  // rocBLAS handle of the large-subspace eigensolver (created on first use)
  rocblas_handle blas = nullptr;
  DevBuf ev_e, ev_info;  // off-diagonal / info of rocsolver_dsyevd

  ~tda_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (blas) (void)rocblas_destroy_handle(blas);
    for (auto& b : grad_cache)
      if (b.first) (void)hipFree(b.first);
    if (hpin) (void)hipHostFree(hpin);
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    for (auto& ls : lps)
      for (DevBuf* b : {&ls.qvv, &ls.G, &ls.Rt, &ls.Lm, &ls.Rpt, &ls.Lpt, &ls.scale}) b->release();
    for (DevBuf* b : {&C, &S, &SC, &orbe, &aooff, &gamma, &gamma_lr, &qov, &qoo, &qvv, &G, &de, &qov32,
                      &qvv_amax, &G_amax, &ws_gemm, &T, &xlp, &u, &uu, &vv32, &ax_tmp,
                      &stage, &stage2, &V, &AV, &H, &Y, &evals, &RZ, &R, &PW, &norms, &npart, &idx, &count,
                      &gs_scratch, &eigw, &guess, &ev_e, &ev_info, &gfull, &ao_atom, &aoRt, &aoLm, &aoRpt, &aoLpt, &aoXR, &aoM, &aoF,
                      &Cl, &Sl, &SCl, &Shl, &aooff_l, &ctW, &ctZ, &Hd, &dxyz, &post,
                      &adjP, &adjQ, &adjX, &adjD, &one, &zvw, &qstage, &ws_basis})
      b->release();
    if (stream) (void)hipStreamDestroy(stream);
  }

  hipEvent_t get_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    TDA_HIP(hipEventCreate(&e));
    return e;
  }
  // Fold recorded exchange events into the counters (after a sync).
  void harvest_events() {
    for (const Rec& r : recs) {
      float ms = 0;
      TDA_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      st_ms[r.stage][r.pclass] += ms;
      st_flops[r.stage][r.pclass] += r.flops;
      st_n[r.stage][r.pclass] += 1;
      ev_pool.push_back(r.a);
      ev_pool.push_back(r.b);
    }
    recs.clear();
  }
  void reset_counters() {
    t_products_ms = 0;
    n_products = n_products_fp64 = n_products_fp32 = 0;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j) st_ms[i][j] = st_flops[i][j] = 0, st_n[i][j] = 0;
  }
};

namespace {

void set_device(tda_ctx* c) { TDA_HIP(hipSetDevice(c->device)); }

// ---------------------------------------------------------------------------
// GEMM helper
// ---------------------------------------------------------------------------
GemmArgs gargs(const void* A, const void* B, void* C, int64_t M, int64_t N, int64_t K, IdxMap am, IdxMap ak,
               IdxMap bk, IdxMap bn, IdxMap cm, IdxMap cn) {
  GemmArgs g;
  std::memset(&g, 0, sizeof(g));
  g.A = A; g.B = B; g.C = C;
  g.M = M; g.N = N; g.K = K;
  g.am = am; g.ak = ak; g.bk = bk; g.bn = bn; g.cm = cm; g.cn = cn;
  g.batch = 1;
  g.ksplit = 1;
  g.alpha = 1.0;
  g.beta = 0.0;
  g.dm = plain(0); g.dmap = plain(1); g.xm = plain(0); g.xn = plain(0);
  g.hm = plain(0); g.hn = plain(0);
  g.hidx = nullptr;
  g.hmod = 1;
  return g;
}

void run_gemm(tda_ctx* c, DType in, DType out, Epi epi, bool a_kc, bool b_kc, GemmArgs g, int ksplit_hint = 0,
              hipEvent_t after_main = nullptr) {
  GemmPlan plan = gemm_plan(in, g.M, g.N, g.K, g.batch < 1 ? 1 : g.batch, a_kc, b_kc, ksplit_hint);
  size_t wsb = gemm_workspace_bytes(in, plan, g.M, g.N, g.batch < 1 ? 1 : g.batch);
  if (wsb) {
    c->ws_gemm.ensure(wsb);
    g.ws = c->ws_gemm.p;
  }
  gemm_launch(in, out, epi, a_kc, b_kc, g, plan, c->stream, after_main);
}

// Davidson basis products on rows of length L = ovp (FP64 arithmetic; the
// basis may be stored as FP32).
// C[i][j] (ldc) = alpha A_i . B_j + beta C[i][j] for i < M, j < N: the
// register-streaming dot kernel (basis.hip) for FP32 rows and for up to 96
// FP64 rows, the LDS-tiled GEMM for more FP64 rows (measured faster there)
void bdot(tda_ctx* c, const void* A, bool a32, int M, const void* B, bool b32, int N, double* C, int64_t ldc,
          double alpha = 1.0, double beta = 0.0) {
  const int64_t L = c->ovp;
  if (a32 || b32 || M <= 96) {
    const size_t wsb = rowdot_ws_bytes(M, N, L);
    if (wsb) c->ws_basis.ensure(wsb);
    launch_rowdot(A, a32, L, B, b32, L, M, N, L, alpha, beta, C, ldc, wsb ? c->ws_basis.as<double>() : nullptr,
                  c->stream);
    return;
  }
  GemmArgs g = gargs(A, B, C, M, N, L, plain(L), plain(1), plain(1), plain(L), plain(ldc), plain(1));
  g.alpha = alpha;
  g.beta = beta;
  run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, g);
}
// C_r = alpha sum_j Y[r ysr + j ysj] V_j + beta C_r for r < R (rows of C and
// V of length and stride L), on the LDS-tiled GEMM
void bcomb(tda_ctx* c, const double* Y, int64_t ysr, int64_t ysj, int R, int J, const void* V, bool v32, double* C,
           double alpha = 1.0, double beta = 0.0) {
  const int64_t L = c->ovp;
  GemmArgs g = gargs(Y, V, C, R, L, J, plain(ysr), plain(ysj), plain(L), plain(1), plain(L), plain(1));
  g.b_f32 = v32;
  g.alpha = alpha;
  g.beta = beta;
  run_gemm(c, DT_F64, DT_F64, EPI_STORE, ysj == 1, false, g);
}

// ---------------------------------------------------------------------------
// prepare_tda: SC = S C, q blocks, delta eps, G = gamma_lr q_oo, LP copies
// ---------------------------------------------------------------------------
void build_charges(tda_ctx* c, const double* dC, const double* dSC, const int* dao, int n_atoms, int nao, int ob,
                   int o, int vb, int v, int op, int vp, double* qov, double* qoo, double* qvv) {
  if (qov) launch_qblock(dC, dSC, nao, dao, n_atoms, ob, o, o, vb, v, vp, qov, c->stream);
  if (qoo) launch_qblock(dC, dSC, nao, dao, n_atoms, ob, o, op, ob, o, op, qoo, c->stream);
  if (qvv) launch_qblock(dC, dSC, nao, dao, n_atoms, vb, v, vp, vb, v, vp, qvv, c->stream);
}

// Row stride of the AO arrays indexed by (s, mu) along their contiguous
// dimension (M, F, R'^T, L'^T): 2 n_ao rounded up to 64 elements, so that the
// rows start 128-byte aligned and the GEMM takes the 16-byte load path for
// every element type.
int64_t ao_ld(const tda_ctx* c) { return round_up(2 * (int64_t)c->nao, 64); }

void gemm_sc(tda_ctx* c, const double* dS, const double* dC, double* dSC, int nao) {
  // SC = S C  (transition_charges.rs:35)
  GemmArgs g = gargs(dS, dC, dSC, nao, nao, nao, plain(nao), plain(1), plain(nao), plain(1), plain(nao), plain(1));
  run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, false, g);
}

// uu = gamma u  (gamma [N][N], u / uu [N][k]) on the FP64 MFMA GEMM
void gamma_times(tda_ctx* c, int N, int k) {
  GemmArgs g = gargs(c->gamma.p, c->u.p, c->uu.p, N, k, N, plain(N), plain(1), plain(k), plain(1), plain(k), plain(1));
  run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, false, g);
}

// Low-precision copies of the operator operands for precision `prec`.
void prepare_lowprec(tda_ctx* c, int prec) {
  if (prec == PREC_FP64) return;
  tda_ctx::LpSet& ls = c->lps[prec];
  if (ls.ready) return;
  const int N = c->N;
  const size_t esz = prec_bytes(prec);
  if (c->has_lc && c->op_kind == 0 && !c->ct) {
    const int64_t n2 = 2 * (int64_t)c->nao, ld = ao_ld(c);
    struct Pair { DevBuf* src; DevBuf* dst; int64_t len; };
    Pair pairs[4] = {{&c->aoRt, &ls.Rt, n2 * c->vp}, {&c->aoLm, &ls.Lm, n2 * c->op},
                     {&c->aoRpt, &ls.Rpt, ld * c->vp}, {&c->aoLpt, &ls.Lpt, ld * c->op}};
    // row widths of the four operands (rows x width = len)
    const int64_t width[4] = {c->vp, c->op, ld, ld};
    for (int i = 0; i < 4; ++i) {
      Pair& pr = pairs[i];
      if (prec == PREC_F16X) {
        pr.dst->ensure((size_t)pr.len * 2 * esz);
        launch_convert_split(pr.src->as<double>(), pr.len / width[i], width[i], width[i], pr.dst->p, 2 * width[i],
                             width[i], c->stream);
      } else {
        pr.dst->ensure((size_t)pr.len * esz);
        launch_convert(pr.src->as<double>(), pr.len, 1, nullptr, pr.dst->p, prec, c->stream);
      }
    }
    ls.xr_scale = 1.0;
    if ((prec == PREC_FP16 || prec == PREC_F16X) && c->ao_scale) {
      // FP16 range.  The middle matrix of the product is small (on SCC
      // (H2O)512 orbitals and unit x: |XR| <= 0.3 with a median of 1e-6 -
      // 2e-4, |M| <= 2e-3 with a median of 1e-7 - 1e-6), below FP16's
      // normal range (2^-14), where its relative precision is lost.  XR is
      // therefore stored as s XR with s = 2^floor(log2(4096 / rho)), rho =
      // max_row |R_row|: |XR_(s,mu),j| <= |R_(s,mu)| |x_j| <= rho for a
      // unit x, so s XR <= 4096 and M, F (linear in XR) carry the same
      // factor well inside the FP16 range; the last GEMM multiplies by 1/s.
      // Both are exact powers of two.  Emulated on those orbitals: the
      // product's relative error falls from 2.7e-5 - 6.9e-5 to 1.3e-5 -
      // 1.7e-5 (random, localised and smooth unit vectors).  Off by
      // default: the solves it changes got no faster (the FP16 phase runs
      // longer to its lower floor and the FP32-accuracy phase still takes
      // the same iterations; (H2O)512: 617 vs 558 ms, C200H202: 59.4 vs
      // 58.1 ms per solve), and the FP64 verification fixes the results.
      ls.scale.ensure((size_t)n2 * 8);
      launch_row_norms(c->aoRt.as<double>(), (int)n2, c->vp, c->vp, ls.scale.as<double>(), c->stream);
      std::vector<double> rn((size_t)n2);
      TDA_HIP(hipMemcpyAsync(rn.data(), ls.scale.p, (size_t)n2 * 8, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
      double rho = 0.0;
      for (double r : rn) rho = std::max(rho, r);
      if (rho > 0.0 && std::isfinite(rho)) ls.xr_scale = std::exp2(std::floor(std::log2(4096.0 / rho)));
    }
    TDA_HIP(hipStreamSynchronize(c->stream));
    ls.ready = true;
    return;
  }
  // Coulomb operand in FP32 for every mixed precision
  if (!c->qov32.p || c->qov32.bytes < (size_t)N * c->ovp * 4) {
    c->qov32.ensure((size_t)N * c->ovp * 4);
  }
  launch_convert(c->qov.as<double>(), c->ovp, N, nullptr, c->qov32.p, PREC_FP32, c->stream);
  if (c->has_lc) {
    const int64_t lvv = (int64_t)c->vp * c->vp, loo = (int64_t)c->op * c->op;
    const int NX = c->Nx;
    ls.qvv.ensure((size_t)NX * lvv * esz);
    ls.G.ensure((size_t)NX * loo * esz);
    ls.scale.ensure((size_t)NX * 4 * 3);
    float* inv_vv = ls.scale.as<float>() + NX;
    float* inv_g = ls.scale.as<float>() + 2 * NX;
    if (prec == PREC_FP32) {
      launch_convert(c->qvv.as<double>(), lvv, NX, nullptr, ls.qvv.p, PREC_FP32, c->stream);
      launch_convert(c->G.as<double>(), loo, NX, nullptr, ls.G.p, PREC_FP32, c->stream);
    } else {
      // per-atom scaling: stored q_vv^B / s_B and G^B / g_B; stage 1 output
      // is multiplied by s_B * g_B so that stage 3 needs no rescaling.
      c->qvv_amax.ensure((size_t)NX * 4);
      c->G_amax.ensure((size_t)NX * 4);
      launch_block_amax(c->qvv.as<double>(), lvv, NX, c->qvv_amax.as<float>(), c->stream);
      launch_block_amax(c->G.as<double>(), loo, NX, c->G_amax.as<float>(), c->stream);
      std::vector<float> svv(NX), sg(NX), sc(3 * (size_t)NX);
      TDA_HIP(hipMemcpyAsync(svv.data(), c->qvv_amax.p, NX * 4, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipMemcpyAsync(sg.data(), c->G_amax.p, NX * 4, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
      for (int b = 0; b < NX; ++b) {
        const float s1 = svv[b] > 0.f ? svv[b] : 1.f;
        const float s2 = sg[b] > 0.f ? sg[b] : 1.f;
        sc[b] = s1 * s2;
        sc[NX + b] = 1.f / s1;
        sc[2 * NX + b] = 1.f / s2;
      }
      TDA_HIP(hipMemcpyAsync(ls.scale.p, sc.data(), sc.size() * 4, hipMemcpyHostToDevice, c->stream));
      launch_convert(c->qvv.as<double>(), lvv, NX, inv_vv, ls.qvv.p, prec, c->stream);
      launch_convert(c->G.as<double>(), loo, NX, inv_g, ls.G.p, prec, c->stream);
    }
  }
  TDA_HIP(hipStreamSynchronize(c->stream));
  ls.ready = true;
}

void reset_lowprec(tda_ctx* c) {
  for (auto& ls : c->lps) ls.ready = false;
}

// q_oo, q_vv and G = gamma_lr q_oo (solvers/mod.rs:328-331, once per system;
// a CT pair contracts gamma_lr^T with the hole fragment's q_oo, :219-222)
void build_mo_charges(tda_ctx* c) {
  if (c->mo_built) return;
  const int NQ = c->Nq, NX = c->Nx;
  c->qoo.ensure((size_t)NQ * c->op * c->op * 8);
  c->qvv.ensure((size_t)NX * c->vp * c->vp * 8);
  if (!c->ct) {
    build_charges(c, c->C.as<double>(), c->SC.as<double>(), c->aooff.as<int>(), NQ, c->nao, c->ob, c->o, c->vb,
                  c->v, c->op, c->vp, nullptr, c->qoo.as<double>(), c->qvv.as<double>());
  } else {
    build_charges(c, c->C.as<double>(), c->SC.as<double>(), c->aooff.as<int>(), NQ, c->nao, c->ob, c->o, 0, 0,
                  c->op, c->vp, nullptr, c->qoo.as<double>(), nullptr);
    build_charges(c, c->Cl.as<double>(), c->SCl.as<double>(), c->aooff_l.as<int>(), NX, c->nao_l, 0, 0,
                  c->lumo_l, c->v, c->op, c->vp, nullptr, nullptr, c->qvv.as<double>());
  }
  if (c->has_lc) {
    const int64_t loo = (int64_t)c->op * c->op;
    c->G.ensure((size_t)NX * loo * 8);
    GemmArgs g = gargs(c->gamma_lr.p, c->qoo.p, c->G.p, NX, loo, NQ, plain(NQ), plain(1), plain(loo), plain(1),
                       plain(loo), plain(1));
    run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, false, g);
  }
  TDA_HIP(hipStreamSynchronize(c->stream));
  c->mo_built = true;
}

// CT-pair q_ov (ct_pair.rs:78-123):
//   A in H: q[A][i][a] = 1/2 sum_{mu in A} C^H_{mu,i} (S_HL C^L_v)_{mu,a}
//   A in L: q[A][i][a] = 1/2 sum_{mu in A} (S_HL^T C^H_o)_{mu,i} C^L_{mu,a}
void build_ct_qov(tda_ctx* c) {
  const int nh = c->nao, nl = c->nao_l, o = c->o, v = c->v, op = c->op, vp = c->vp;
  c->ctW.ensure((size_t)nh * vp * 8);
  c->ctZ.ensure((size_t)nl * op * 8);
  // W = S_HL C^L[:, lumo:]   [nh][vp]
  GemmArgs gw = gargs(c->Shl.p, c->Cl.as<double>() + c->lumo_l, c->ctW.p, nh, v, nl, plain(nl), plain(1), plain(nl),
                      plain(1), plain(vp), plain(1));
  run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, false, gw);
  // Z = S_HL^T C^H[:, :o]    [nl][op]
  GemmArgs gz = gargs(c->Shl.p, c->C.as<double>() + c->ob, c->ctZ.p, nl, o, nh, plain(1), plain(nl), plain(nh),
                      plain(1), plain(op), plain(1));
  run_gemm(c, DT_F64, DT_F64, EPI_STORE, false, false, gz);
  const int NH = c->Nq, NL = c->Nx;
  launch_qprod(c->C.as<double>() + c->ob, nh, c->ctW.as<double>(), vp, c->aooff.as<int>(), 0, NH, o, o, v, vp,
               c->qov.as<double>(), c->stream);
  launch_qprod(c->ctZ.as<double>(), op, c->Cl.as<double>() + c->lumo_l, nl, c->aooff_l.as<int>(), NH, NL, o, o, v,
               vp, c->qov.as<double>(), c->stream);
}

void build_ao_operands(tda_ctx* c) {
  const int nao = c->nao;
  const int64_t n2 = 2 * (int64_t)nao, ld = ao_ld(c);
  const double* C = c->C.as<double>();
  const double* SC = c->SC.as<double>();
  c->aoRt.ensure((size_t)n2 * c->vp * 8);
  c->aoLm.ensure((size_t)n2 * c->op * 8);
  c->aoRpt.ensure((size_t)ld * c->vp * 8);
  c->aoLpt.ensure((size_t)ld * c->op * 8);
  // Rt rows (SC_v; C_v), Lm rows (SC_o; C_o): they build D = L X R^T
  launch_ao_operand(C, SC, nao, c->vb, c->v, c->vp, 0, 0, 0, c->aoRt.as<double>(), c->stream);
  launch_ao_operand(C, SC, nao, c->ob, c->o, c->op, 0, 0, 0, c->aoLm.as<double>(), c->stream);
  // the back-contraction pairs each D block with the other operator: (C; SC), stored transposed
  launch_ao_operand(C, SC, nao, c->vb, c->v, c->vp, 1, 1, ld, c->aoRpt.as<double>(), c->stream);
  launch_ao_operand(C, SC, nao, c->ob, c->o, c->op, 1, 1, ld, c->aoLpt.as<double>(), c->stream);
  c->gfull_ld = round_up(nao, 4);
  c->gfull.ensure((size_t)nao * c->gfull_ld * sizeof(float));
  launch_expand_pairs(c->gamma_lr.as<double>(), c->ao_atom.as<int>(), nao, c->N, c->gfull_ld, c->gfull.as<float>(),
                      c->stream);
}

void prepare(tda_ctx* c) {
  TDA_CHECK(c->has_system, TDA_ERR_STATE, "tda_prepare: no system loaded");
  if (c->dense) return;
  if (c->prepared) {
    prepare_lowprec(c, c->prec);
    return;
  }
  const int N = c->N, nao = c->nao;
  c->SC.ensure((size_t)nao * nao * 8);
  gemm_sc(c, c->S.as<double>(), c->C.as<double>(), c->SC.as<double>(), nao);
  const bool ao = c->has_lc && c->op_kind == 0 && !c->ct;
  if (ao) {
    c->qov.release();  // built on request only (tda_get_charges)
  } else if (!c->ct) {
    c->qov.ensure((size_t)N * c->ovp * 8);
    build_charges(c, c->C.as<double>(), c->SC.as<double>(), c->aooff.as<int>(), N, nao, c->ob, c->o, c->vb, c->v,
                  c->op, c->vp, c->qov.as<double>(), nullptr, nullptr);
  } else {
    c->qov.ensure((size_t)N * c->ovp * 8);
    c->SCl.ensure((size_t)c->nao_l * c->nao_l * 8);
    gemm_sc(c, c->Sl.as<double>(), c->Cl.as<double>(), c->SCl.as<double>(), c->nao_l);
    build_ct_qov(c);
  }
  c->qov_built = !ao;
  c->de.ensure((size_t)c->ovp * 8);
  launch_delta_eps(c->orbe.as<double>(), c->ob, c->o, c->vb, c->v, c->vp, c->de.as<double>(), c->stream);
  c->mo_built = false;
  if (c->has_lc) {
    if (c->op_kind == 1 || c->ct) build_mo_charges(c);
    else build_ao_operands(c);
  }
  TDA_HIP(hipStreamSynchronize(c->stream));
  c->prepared = true;
  reset_lowprec(c);
  prepare_lowprec(c, c->prec);
}

// ---------------------------------------------------------------------------
// A.x  (solvers/mod.rs:280-369).  X, AX: device [k][ovp] FP64.
// ---------------------------------------------------------------------------
void apply_A(tda_ctx* c, const double* X, int k, double* AX, int prec) {
  if (k == 0) return;
  const int N = c->N, o = c->o, v = c->v, op = c->op, vp = c->vp;
  const int64_t ovp = c->ovp;
  const bool lp = prec != PREC_FP64;
  // ---- Coulomb: u = q_ov X ; vv = gamma u ; AX = de*X + 2 q_ov^T vv ----
  c->u.ensure((size_t)N * k * 8);
  c->uu.ensure((size_t)N * k * 8);
  if (!lp) {
    GemmArgs g = gargs(c->qov.p, X, c->u.p, N, k, ovp, plain(ovp), plain(1), plain(1), plain(ovp), plain(k), plain(1));
    run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, g);
  } else {
    c->xlp.ensure((size_t)k * ovp * std::max<size_t>(4, prec_bytes(prec)));
    launch_cast_vectors(X, (int64_t)k * ovp, c->xlp.p, PREC_FP32, c->stream);
    GemmArgs g = gargs(c->qov32.p, c->xlp.p, c->u.p, N, k, ovp, plain(ovp), plain(1), plain(1), plain(ovp),
                       plain(k), plain(1));
    run_gemm(c, DT_F32, DT_F64, EPI_STORE, true, true, g);
  }
  gamma_times(c, N, k);
  {
    GemmArgs g;
    if (!lp) {
      g = gargs(c->uu.p, c->qov.p, AX, k, ovp, N, plain(1), plain(k), plain(ovp), plain(1), plain(ovp), plain(1));
    } else {
      c->vv32.ensure((size_t)N * k * 4);
      launch_convert(c->uu.as<double>(), (int64_t)N * k, 1, nullptr, c->vv32.p, PREC_FP32, c->stream);
      g = gargs(c->vv32.p, c->qov32.p, AX, k, ovp, N, plain(1), plain(k), plain(ovp), plain(1), plain(ovp),
                plain(1));
    }
    g.alpha = c->coul_factor;  // singlet Coulomb factor 2 (solvers/mod.rs:311)
    g.diag = c->de.as<double>();
    g.dmap = plain(1);
    g.xsrc = X;
    g.xm = plain(ovp);
    g.xn = plain(1);
    run_gemm(c, lp ? DT_F32 : DT_F64, DT_F64, EPI_DIAG, false, false, g);
  }
  if (!c->has_lc) return;
  // ---- long-range exchange, blocked over atoms ----
  const DType dt = lp ? dtype_of(prec) : DT_F64;
  const size_t esz = lp ? prec_bytes(prec) : 8;
  const void* Xin = X;
  if (lp && prec != PREC_FP32) {
    launch_cast_vectors(X, (int64_t)k * ovp, c->xlp.p, prec, c->stream);
  }
  if (lp) Xin = c->xlp.p;  // FP32 cast above serves PREC_FP32
  if (lp) prepare_lowprec(c, prec);
  const void* qvv = lp ? c->lps[prec].qvv.p : c->qvv.p;
  const void* G = lp ? c->lps[prec].G.p : c->G.p;
  const float* s1scale = (lp && prec != PREC_FP32) ? c->lps[prec].scale.as<float>() : nullptr;
  const int NX = c->Nx;
  size_t per_atom = (size_t)k * vp * op * esz;
  size_t tb = std::max<size_t>(per_atom, std::min<size_t>(c->budget, (size_t)8 << 30));
  int NB = (int)std::max<size_t>(1, std::min<size_t>((size_t)NX, tb / per_atom));
  c->n_blocks = (int)ceil_div(NX, NB);
  c->T.ensure((size_t)NB * per_atom);
  for (int B0 = 0; B0 < NX; B0 += NB) {
    const int nb = std::min(NB, NX - B0);
    const int64_t NBop = (int64_t)nb * op;
    // Stage 3 reads T's pad rows j in [o, o_p) of every atom against the zero
    // pads of G; they are never written by stage 1 and the buffer is shared by
    // all precisions, so clear them (0 * NaN-pattern would poison the sum).
    if (op > o) launch_zero_runs(c->T.p, (int64_t)k * vp * nb, op, o, op - o, (int)esz, c->stream);
    const int pclass = prec == PREC_FP64 ? 0 : prec == PREC_FP32 ? 1 : 2;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    if (c->timing) {
      e0 = c->get_event();
      e1 = c->get_event();
      e2 = c->get_event();
      e3 = c->get_event();
      TDA_HIP(hipEventRecord(e0, c->stream));
    }
    // stage 1: T[c][a][(B,j)] = sum_b q_vv[B][a][b] X[c][j][b]
    {
      GemmArgs g = gargs((const char*)qvv + (size_t)B0 * vp * vp * esz, Xin, c->T.p, v, (int64_t)k * o, vp,
                         plain(vp), plain(1), plain(1), split(o, ovp, vp), plain(NBop), split(o, (int64_t)vp * NBop, 1));
      g.batch = nb;
      g.a_batch = (int64_t)vp * vp;
      g.b_batch = 0;
      g.c_batch = op;
      if (s1scale) g.batch_scale = s1scale + B0;
      run_gemm(c, dt, dt, EPI_STORE, true, true, g, 0, e1);
    }
    if (c->timing) TDA_HIP(hipEventRecord(e2, c->stream));
    // stage 3: AX[c][i][a] -= sum_(B,j) G[B][i][j] T[c][a][(B,j)]
    {
      GemmArgs g = gargs((const char*)G + (size_t)B0 * op * op * esz, c->T.p, AX, o, (int64_t)k * v, NBop,
                         plain(op), split(op, (int64_t)op * op, 1), plain(1), split(v, (int64_t)vp * NBop, NBop),
                         plain(vp), split(v, ovp, 1));
      g.alpha = -1.0;
      g.beta = 1.0;
      run_gemm(c, dt, DT_F64, EPI_STORE, true, true, g, 0, e3);
    }
    if (c->timing) {
      // algorithmic flops: stage 1 2*k*o*v*v*nb, stage 3 2*k*o*o*v*nb
      c->recs.push_back({e0, e1, 0, pclass, 2.0 * k * (double)o * v * v * nb});
      c->recs.push_back({e2, e3, 1, pclass, 2.0 * k * (double)o * o * v * nb});
    }
  }
}

// ---------------------------------------------------------------------------
// A.x with the AO-factored exchange.  With the Mulliken charges
// q^A_pq = 1/2 sum_{mu in A}(C_mu,p SC_mu,q + C_mu,q SC_mu,p) the exchange
// sum_AB q_oo^A gamma_lr,AB q_vv^B x regroups as
//   K x = 1/4 L'^T (gamma~ o (L x R^T)) R',  L = (SC_o; C_o), R = (SC_v; C_v),
//   L' = (C_o; SC_o), R' = (C_v; SC_v), gamma~_{nu,mu} = gamma_lr[A(nu)][B(mu)],
// and the Coulomb charges are q_ov x = 1/2 sum_{mu in A} (D_{(0,mu),(1,mu)} +
// D_{(1,mu),(0,mu)}) with D = L x R^T; 2 q_ov^T gamma (q_ov x) adds gamma u on
// the diagonals of the off-diagonal blocks of the middle matrix.
// Four GEMMs per batch of k vectors:
//   G1  XR_c = R x_c^T           [2n_ao][o]      (stored per vector)
//   G2  M_c  = -1/4 gamma~ o (L XR_c^T)          [2n_ao][2n_ao]  (Hadamard epilogue)
//   G3  F_c  = R' M_c^T ...      [v_p][2n_ao]
//   G4  AX_c = de o x_c + L'^T F_c^T             [o][v_p]
// M, F, R'^T and L'^T store their (s, mu) rows with the padded stride ao_ld().
// ---------------------------------------------------------------------------
// x32: X and AX are FP32 vectors (the FP32 basis of the FP16 / BF16 phase;
// 16-bit products only): X is narrowed from FP32, the diagonal term reads it
// as FP32 and AX is written as FP32.
void apply_A_ao(tda_ctx* c, const void* X, int k, void* AX, int prec, bool x32 = false) {
  const int o = c->o, v = c->v, op = c->op, vp = c->vp, nao = c->nao, N = c->N;
  const int64_t ovp = c->ovp, n2 = 2 * (int64_t)nao, ld = ao_ld(c);
  const bool lp = prec != PREC_FP64;
  // PREC_F16X: every 16-bit operand row holds hi then lo (row width w = 2
  // rows of the plain layout; lo at +vp / +op / +ld)
  const bool sp = prec == PREC_F16X;
  const int64_t w = sp ? 2 : 1;
  const DType dt = lp ? dtype_of(prec) : DT_F64;
  const size_t esz = lp ? prec_bytes(prec) : 8;
  const int mprec = lp ? (sp ? PREC_FP16 : prec) : PREC_FP64;  // element type of M
  TDA_CHECK(!x32 || (lp && !sp), TDA_ERR_BAD_ARG, "apply_A_ao: FP32 vectors need 16-bit products");
  const void* Xin = X;
  if (sp) {
    c->xlp.ensure((size_t)k * ovp * 2 * esz);
    launch_convert_split((const double*)X, (int64_t)k * o, vp, vp, c->xlp.p, 2 * vp, vp, c->stream);
    Xin = c->xlp.p;
  } else if (lp) {
    c->xlp.ensure((size_t)k * ovp * esz);
    if (x32) launch_convert_f32_to((const float*)X, (int64_t)k * ovp, c->xlp.p, prec, c->stream);
    else launch_cast_vectors((const double*)X, (int64_t)k * ovp, c->xlp.p, prec, c->stream);
    Xin = c->xlp.p;
  }
  if (lp) prepare_lowprec(c, prec);
  const double s1 = lp ? c->lps[prec].xr_scale : 1.0;  // XR stored times s1 (FP16 range)
  const void* Rt = lp ? c->lps[prec].Rt.p : c->aoRt.p;
  const void* Lm = lp ? c->lps[prec].Lm.p : c->aoLm.p;
  const void* Rpt = lp ? c->lps[prec].Rpt.p : c->aoRpt.p;
  const void* Lpt = lp ? c->lps[prec].Lpt.p : c->aoLpt.p;
  c->aoXR.ensure((size_t)k * n2 * w * op * esz);
  c->aoM.ensure((size_t)k * n2 * w * ld * esz);
  c->aoF.ensure((size_t)k * vp * w * ld * esz);
  const int pclass = prec == PREC_FP64 ? 0 : (prec == PREC_FP32 || sp) ? 1 : 2;
  hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
  if (c->timing)
    for (auto& ev : e) ev = c->get_event();
  // G1: XR[c][(s,mu)][j] = sum_b Rt[(s,mu)][b] X[c][j][b]
  {
    GemmArgs g = gargs(Rt, Xin, c->aoXR.p, n2, (int64_t)k * o, vp, plain(w * vp), plain(1), plain(1), plain(w * vp),
                       plain(w * op), split(o, n2 * w * op, 1));
    if (sp) g.a_lo = g.b_lo = vp, g.c_lo = op;
    g.alpha = s1;
    run_gemm(c, dt, dt, EPI_STORE, true, true, g);
  }
  // G2: M[c] = -1/4 gamma~ o (Lm XR[c]^T)
  if (c->timing) TDA_HIP(hipEventRecord(e[0], c->stream));
  {
    GemmArgs g = gargs(Lm, c->aoXR.p, c->aoM.p, n2, n2, o, plain(w * op), plain(1), plain(1), plain(w * op),
                       plain(w * ld), plain(1));
    g.batch = k;
    g.b_batch = n2 * w * op;
    g.c_batch = n2 * w * ld;
    g.alpha = -0.25;
    if (sp) g.a_lo = g.b_lo = op, g.c_lo = ld;
    // gamma~_{(s,mu),(t,nu)} = gamma_lr[atom(mu)][atom(nu)], gathered per row / column
    g.hsrc = c->gamma_lr.as<double>();
    g.hidx = c->ao_atom.as<int>();
    g.hmod = nao;
    g.hm = plain(N);
    g.hn = plain(1);
    if (lp && !c->no_gfull) {  // 16-bit kernel: the factors from the AO-pair table
      g.hfull = c->gfull.as<float>();
      g.hld = c->gfull_ld;
    }
    run_gemm(c, dt, dt, EPI_HADAMARD, true, true, g, 0, e[1]);
  }
  // Coulomb: u = q_ov x from the D diagonals, w = gamma u, added to M
  c->u.ensure((size_t)N * k * 8);
  c->uu.ensure((size_t)N * k * 8);
  launch_ao_charges(c->aoM.p, mprec, k, nao, w * ld, c->aooff.as<int>(), N, c->gamma_lr.as<double>(),
                    c->u.as<double>(), c->stream, sp ? ld : 0);
  gamma_times(c, N, k);
  launch_ao_add_coulomb(c->aoM.p, mprec, k, nao, w * ld, c->ao_atom.as<int>(), c->uu.as<double>(), c->stream,
                        sp ? ld : 0);
  // G3: F[c][a][(t,nu)] = sum_(s,mu) Rpt[a][(s,mu)] M[c][(t,nu)][(s,mu)]
  if (c->timing) TDA_HIP(hipEventRecord(e[2], c->stream));
  {
    GemmArgs g = gargs(Rpt, c->aoM.p, c->aoF.p, vp, n2, n2, plain(w * ld), plain(1), plain(1), plain(w * ld),
                       plain(w * ld), plain(1));
    g.batch = k;
    g.b_batch = n2 * w * ld;
    g.c_batch = vp * w * ld;
    if (sp) g.a_lo = g.b_lo = ld, g.c_lo = ld;
    run_gemm(c, dt, dt, EPI_STORE, true, true, g, 0, e[3]);
  }
  // G4: AX[c][i][a] = de[i][a] X[c][i][a] + sum_(t,nu) Lpt[i][(t,nu)] F[c][a][(t,nu)]
  {
    GemmArgs g = gargs(Lpt, c->aoF.p, AX, o, vp, n2, plain(w * ld), plain(1), plain(1), plain(w * ld), plain(vp),
                       plain(1));
    g.batch = k;
    g.b_batch = vp * w * ld;
    g.c_batch = ovp;
    if (sp) g.a_lo = g.b_lo = ld;
    g.alpha = 1.0 / s1;
    g.diag = c->de.as<double>();
    g.dm = plain(vp);
    g.dmap = plain(1);
    if (x32) g.xsrc32 = (const float*)X;
    else g.xsrc = (const double*)X;
    g.xm = plain(vp);
    g.xn = plain(1);
    g.x_batch = ovp;
    run_gemm(c, dt, x32 ? DT_F32 : DT_F64, EPI_DIAG, true, true, g);
  }
  if (c->timing) {
    // MFMA flops of the launches (three products per split GEMM)
    const double f = sp ? 3.0 : 1.0;
    c->recs.push_back({e[0], e[1], 0, pclass, f * 2.0 * k * (double)n2 * n2 * o});
    c->recs.push_back({e[2], e[3], 1, pclass, f * 2.0 * k * (double)n2 * n2 * vp});
  }
}

// Transition charges u[A][ldu] (column c0 + c for vector c) of n vectors X
// (device [n][o][v_p]) on the AO operands, without q_ov:
// q_ov^A x = 1/2 sum_{mu in A} (D[(0,mu),(1,mu)] + D[(1,mu),(0,mu)]) with
// D = L x R^T, from the first GEMM of the AO product (XR = R x^T) and a
// per-atom dot over the occupied index.  FP64.
void ao_trans_charges(tda_ctx* c, const double* X, int n, double* u, int64_t ldu) {
  const int o = c->o, op = c->op, vp = c->vp, nao = c->nao;
  const int64_t n2 = 2 * (int64_t)nao, per = n2 * op * 8;
  const int64_t room = std::min<int64_t>((int64_t)c->budget, 8ll << 30) / per;
  const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(n, std::min<int64_t>(64, room)));
  c->aoXR.ensure((size_t)chunk * n2 * op * 8);
  for (int c0 = 0; c0 < n; c0 += chunk) {
    const int kc = std::min(chunk, n - c0);
    GemmArgs g = gargs(c->aoRt.p, X + (size_t)c0 * c->ovp, c->aoXR.p, n2, (int64_t)kc * o, vp, plain(vp), plain(1),
                       plain(1), plain(vp), plain(op), split(o, n2 * op, 1));
    run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, g);
    launch_ao_dens_charges(c->aoLm.as<double>(), c->aoXR.as<double>(), kc, nao, o, op, c->aooff.as<int>(), c->N,
                           u + c0, ldu, c->stream);
  }
}

// d w_k / d gamma_lr [n][N][N] (device) of the AO operator from the AO
// transition densities D = L x_k R^T (G1 and G2 of the product without the
// Hadamard factor), without q_oo / q_vv / G: w_k holds -x^T K x =
// -1/4 sum_pq D'_pq gamma~_pq D_pq, D' = L' x R'^T is D with both halves of
// its index exchanged (L', R' are L, R with their halves swapped), so
// d w_k / d gamma_lr[A][B] = -1/2 sum_{mu in A, nu in B} (D00 D11 + D01 D10).
void ao_dglr(tda_ctx* c, const double* X, int n, double* dglr) {
  const int o = c->o, op = c->op, vp = c->vp, nao = c->nao, N = c->N;
  const int64_t n2 = 2 * (int64_t)nao, ld = ao_ld(c), per = (n2 * op + n2 * ld) * 8;
  const int64_t room = std::min<int64_t>((int64_t)c->budget, 8ll << 30) / per;
  const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(n, std::min<int64_t>(64, room)));
  c->aoXR.ensure((size_t)chunk * n2 * op * 8);
  c->aoM.ensure((size_t)chunk * n2 * ld * 8);
  for (int c0 = 0; c0 < n; c0 += chunk) {
    const int kc = std::min(chunk, n - c0);
    GemmArgs g1 = gargs(c->aoRt.p, X + (size_t)c0 * c->ovp, c->aoXR.p, n2, (int64_t)kc * o, vp, plain(vp), plain(1),
                        plain(1), plain(vp), plain(op), split(o, n2 * op, 1));
    run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, g1);
    GemmArgs g2 = gargs(c->aoLm.p, c->aoXR.p, c->aoM.p, n2, n2, o, plain(op), plain(1), plain(1), plain(op),
                        plain(ld), plain(1));
    g2.batch = kc;
    g2.b_batch = n2 * op;
    g2.c_batch = n2 * ld;
    run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, g2);
    for (int j = 0; j < kc; ++j)
      launch_ao_dglr(c->aoM.as<double>() + (size_t)j * n2 * ld, ld, nao, c->aooff.as<int>(), N,
                     dglr + (size_t)(c0 + j) * N * N, c->stream);
  }
}

// x32: X and AX are FP32 vectors (AO operator, FP16 / BF16 products only).
void products(tda_ctx* c, const void* Xv, int k, void* AXv, int prec, bool x32 = false) {
  const double* X = (const double*)Xv;
  double* AX = (double*)AXv;
  if (c->dense) {
    TDA_CHECK(!x32, TDA_ERR_BAD_ARG, "products: FP32 vectors need the AO operator");
    // AX[c][r] = sum_s H[r][s] X[c][s]  (solvers/mod.rs:41-43), pads of H zero
    const int64_t L = c->ovp;
    GemmArgs g = gargs(X, c->Hd.p, AX, k, L, L, plain(L), plain(1), plain(1), plain(L), plain(L), plain(1));
    run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, g);
    c->n_products += k;
    c->n_products_fp64 += k;
    return;
  }
  const bool ao = c->has_lc && c->op_kind == 0 && !c->ct;
  if (prec == PREC_F16X && !ao) prec = PREC_FP32;  // split products exist for the AO operator only
  int chunk = 64;
  if (ao) {
    // per-vector AO workspace: XR + M + F
    const double n2 = 2.0 * c->nao, ld = (double)ao_ld(c);
    const double per = (n2 * c->op + n2 * ld + c->vp * ld) * (prec == PREC_FP64 ? 8.0 : prec_bytes(prec)) *
                       (prec == PREC_F16X ? 2.0 : 1.0);
    chunk = (int)std::max(1.0, std::min(64.0, std::min((double)c->budget, 8.0 * (1ull << 30)) / per));
  }
  TDA_CHECK(!x32 || ao, TDA_ERR_BAD_ARG, "products: FP32 vectors need the AO operator");
  for (int c0 = 0; c0 < k; c0 += chunk) {
    const int kc = std::min(chunk, k - c0);
    if (x32) apply_A_ao(c, (const float*)Xv + (size_t)c0 * c->ovp, kc, (float*)AXv + (size_t)c0 * c->ovp, prec, true);
    else if (ao) apply_A_ao(c, X + (size_t)c0 * c->ovp, kc, AX + (size_t)c0 * c->ovp, prec);
    else apply_A(c, X + (size_t)c0 * c->ovp, kc, AX + (size_t)c0 * c->ovp, prec);
  }
  c->n_products += k;
  if (prec == PREC_FP64) c->n_products_fp64 += k;
  if (prec == PREC_FP32 || prec == PREC_F16X) c->n_products_fp32 += k;
}

// ---------------------------------------------------------------------------
// Hellmann-Feynman derivatives of w_k = x_k^T A x_k for normalised x_k (the
// converged CI vectors; A symmetric), with respect to the operator's inputs:
//   d w_k / d gamma_AB      = 2 u_A u_B,              u = q_ov x_k
//   d w_k / d q_ov[A, ia]   = 4 (gamma u)_A x_ia
//   d w_k / d gamma_lr,AB   = -sum_ij q_oo^A_ij P^B_ij, P^B = x q_vv^B x^T  (o x o)
//   d w_k / d q_oo^A_ij     = -sum_B gamma_lr,AB P^B_ij
//   d w_k / d q_vv^B_ab     = -(x^T G^B x)_ab,         G^B = sum_A gamma_lr,BA q_oo^A
// (x as the o x v matrix; the terms of A = diag(de) + 2 q_ov^T gamma q_ov - K
// of solvers/mod.rs:307-360).  Xd: device [n][o][v_p].  Device outputs
// dgam, dglr [n][N][N] and cq [n][N]; host outputs dqoo [n][N][o^2],
// dqvv [n][N][v^2] (any may be NULL).
// ---------------------------------------------------------------------------
void copy_unpad_rows(const double* src, int rows, int cols, int ld, double* dst);

void hf_derivatives(tda_ctx* c, const double* Xd, int n, double* dgam, double* dglr, double* cq, double* dqoo,
                    double* dqvv) {
  const int N = c->N, o = c->o, v = c->v, op = c->op, vp = c->vp;
  const int64_t ovp = c->ovp, loo = (int64_t)op * op;
  c->u.ensure((size_t)N * n * 8);
  c->uu.ensure((size_t)N * n * 8);
  if (c->qov_built) {
    GemmArgs gu = gargs(c->qov.p, Xd, c->u.p, N, n, ovp, plain(ovp), plain(1), plain(1), plain(ovp), plain(n),
                        plain(1));
    run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, gu);
  } else {
    ao_trans_charges(c, Xd, n, c->u.as<double>(), n);
  }
  gamma_times(c, N, n);
  if (dgam) launch_hf_coulomb(c->u.as<double>(), c->uu.as<double>(), N, n, dgam, cq, c->stream);
  if (!c->has_lc || (!dglr && !dqoo && !dqvv)) return;
  if (c->has_lc && c->op_kind == 0 && !c->ct && !dqoo && !dqvv) {
    // AO operator: gamma_lr derivatives from the AO transition densities,
    // no q_oo / q_vv / G (N (o^2 + v^2) words) on the device
    ao_dglr(c, Xd, n, dglr);
    TDA_HIP(hipStreamSynchronize(c->stream));
    return;
  }
  build_mo_charges(c);
  const int64_t NBop = (int64_t)N * op;
  c->T.ensure((size_t)vp * NBop * 8);
  c->adjP.ensure((size_t)N * loo * 8);
  std::vector<double> h;
  for (int k = 0; k < n; ++k) {
    const double* xk = Xd + (size_t)k * ovp;
    // T[a][(B,j)] = sum_b q_vv^B[a][b] x[j][b]
    launch_zero_runs(c->T.p, (int64_t)vp * N, op, o, op - o, 8, c->stream);
    GemmArgs g1 = gargs(c->qvv.p, xk, c->T.p, v, o, vp, plain(vp), plain(1), plain(1), plain(vp), plain(NBop),
                        plain(1));
    g1.batch = N;
    g1.a_batch = (int64_t)vp * vp;
    g1.b_batch = 0;
    g1.c_batch = op;
    run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, g1);
    // P[B][i][j] = sum_a T[a][(B,i)] x[j][a]   (rows i in [o, o_p) are zero)
    launch_fill(c->adjP.as<double>(), 0.0, (int64_t)N * loo, c->stream);
    GemmArgs g2 = gargs(c->T.p, xk, c->adjP.p, NBop, o, v, plain(1), plain(NBop), plain(1), plain(vp), plain(op),
                        plain(1));
    run_gemm(c, DT_F64, DT_F64, EPI_STORE, false, true, g2);
    if (dglr) {
      // dglr[k][A][B] = -sum_(ij) q_oo[A][(ij)] P[B][(ij)]
      GemmArgs g3 = gargs(c->qoo.p, c->adjP.p, dglr + (size_t)k * N * N, N, N, loo, plain(loo), plain(1), plain(1),
                          plain(loo), plain(N), plain(1));
      g3.alpha = -1.0;
      run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, g3);
    }
    if (dqoo) {
      // -gamma_lr P  [N][o_p^2]
      c->adjQ.ensure((size_t)N * loo * 8);
      GemmArgs g4 = gargs(c->gamma_lr.p, c->adjP.p, c->adjQ.p, N, loo, N, plain(N), plain(1), plain(loo), plain(1),
                          plain(loo), plain(1));
      g4.alpha = -1.0;
      run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, false, g4);
      h.resize((size_t)N * loo);
      TDA_HIP(hipMemcpyAsync(h.data(), c->adjQ.p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
      for (int A = 0; A < N; ++A)
        copy_unpad_rows(h.data() + (size_t)A * loo, o, o, op, dqoo + ((size_t)k * N + A) * o * o);
    }
    if (dqvv) {
      // Y[B][i][b] = sum_j G[B][i][j] x[j][b];  Z[B][a][b] = -sum_i x[i][a] Y[B][i][b]
      c->adjQ.ensure((size_t)N * std::max<int64_t>((int64_t)o * vp, (int64_t)v * v) * 8 * 2);
      double* Y = c->adjQ.as<double>();
      double* Z = Y + (size_t)N * o * vp;
      GemmArgs g5 = gargs(c->G.p, xk, Y, o, v, o, plain(op), plain(1), plain(vp), plain(1), plain(vp), plain(1));
      g5.batch = N;
      g5.a_batch = loo;
      g5.b_batch = 0;
      g5.c_batch = (int64_t)o * vp;
      run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, false, g5);
      GemmArgs g6 = gargs(xk, Y, Z, v, v, o, plain(1), plain(vp), plain(vp), plain(1), plain(v), plain(1));
      g6.batch = N;
      g6.a_batch = 0;
      g6.b_batch = (int64_t)o * vp;
      g6.c_batch = (int64_t)v * v;
      g6.alpha = -1.0;
      run_gemm(c, DT_F64, DT_F64, EPI_STORE, false, false, g6);
      TDA_HIP(hipMemcpyAsync(dqvv + (size_t)k * N * v * v, Z, (size_t)N * v * v * 8, hipMemcpyDeviceToHost,
                             c->stream));
    }
  }
  TDA_HIP(hipStreamSynchronize(c->stream));
}

// ---------------------------------------------------------------------------
// Davidson (solvers/davidson.rs:53-185) on device-resident vectors
// ---------------------------------------------------------------------------
struct DavResult {
  bool converged = false;
  int iters = 0;
  double max_err = 0;
  double delta = 0;  // measured product error of the phase's precision (0: not measured)
  // a low-precision phase met a non-finite product (16-bit overflow): it
  // stopped, and V[0, n) -- its last finite start vectors -- hold the restart
  bool nonfinite = false;
};

// Runs the block Davidson with products in `prec` starting from the m0
// vectors already in V.  `first_can_converge` relaxes the i>0 rule of the
// reference (used by the FP64 continuation, which starts from Ritz vectors).
// stop_stall > 0: leave early when the max residual has not dropped below
// stall_factor x its best value for that many iterations (low-precision
// phases: 0.5 for FP16/BF16, which hit their accuracy floor quickly; 0.9 for
// FP32, which can reach the tolerance and only stops on true stagnation).
// Stagnation is only counted while max residual < stall_floor (FP32: 10 tol,
// so a slowly converging spectrum is not mistaken for the precision floor).
// cap: maximum basis size; keep: Ritz vectors retained at a collapse.  The
// reference rule is cap = 15 n (davidson.rs:74), keep = n with all products
// recomputed after the collapse (:172-176, solvers/mod.rs:290-302).  With
// keep > n (low-precision phases: thick restart) the basis and its products
// are rotated onto the keep lowest Ritz vectors instead (V <- Y V,
// AV <- Y AV), so no product is recomputed and the projection restarts as
// diag(theta).
// floor_k > 0 (low-precision phases): after the first iteration the phase
// measures delta = |A_prec z - A_fp64 z| for its lowest Ritz vector z (unit
// norm), the size of the product error every residual of the phase carries,
// and leaves once max residual < floor_k * delta: below that the computed
// residuals no longer say how far the roots are from convergence.
// reorth: a second Gram-Schmidt pass for a block of new basis vectors when
// the estimated loss of orthogonality of the basis would exceed eta_tol.
// Algorithmic flops of one single-vector product (AO form: the four GEMMs;
// charges form: the three-stage exchange).
double product_flops(const tda_ctx* c) {
  const double o = c->o, v = c->v, N = c->N, n2 = 2.0 * c->nao;
  if (c->has_lc && c->op_kind == 0 && !c->ct) return 2.0 * n2 * n2 * (o + v) + 4.0 * n2 * o * v;
  return 2.0 * N * o * v * (o + v) + 4.0 * N * o * v;
}

// Stall window of the FP16 / BF16 phase: cheap products (small systems,
// where the subspace work dominates an iteration) hand over after 4
// iterations without progress; expensive ones ride out longer plateaus,
// since their iterations cost less than those of the next precision.  The
// FP32-accuracy phase, whose accuracy floor lies far below the tolerance,
// gets three times the window: handing a slow plateau to FP64 products costs
// far more than riding it out.
int lp_stall_window(const tda_ctx* c) {
  if (c->stall_window > 0) return c->stall_window;
  return product_flops(c) >= 5e10 ? 10 : 4;
}

// Basis cap multiplier of the FP16 / BF16 phase (see lp_cap_mult).
int lp_cap_mult_for(const tda_ctx* c) {
  if (c->lp_cap_mult > 0) return c->lp_cap_mult;
  return product_flops(c) >= 5e10 ? 8 : 6;
}
bool basis32_for(const tda_ctx* c) {
  if (c->basis32 >= 0) return c->basis32 != 0;
  return product_flops(c) >= 5e10;
}

// In-block Gram-Schmidt of the nc new vectors (davidson.rs:162-168): W[j],
// already projected against the basis, becomes u_j; the accepted q_j go to Q
// in order.  Returns their number and, when wn is given, |u_j| per j.  P and
// W are the context's rows (W = P + pw_rows L).  Coefficient form (one GEMM
// for the Gram matrices of [P; W] against W, one writing Q) unless its norm
// estimates have lost too many digits; then the vector-by-vector kernel.
// Synchronizes the stream.
int gs_block(tda_ctx* c, double* P, double* W, double* Q, int nc, double* wn) {
  const int64_t L = c->ovp;
  TDA_CHECK(P == c->Pb() && W == c->Wb(), 2, "gs_block: P / W must be the context's rows");
  int* d_count = c->count.as<int>();
  double* scr = c->gs_scratch.as<double>();
  if (c->gs_gram && nc <= 64) {
    double* Gt = scr;                         // [2 nc][nc]
    double* Tq = scr + (size_t)2 * nc * nc;   // [nc][nc]
    bdot(c, P, false, nc, W, false, nc, Gt, nc);                     // rows [0, nc): P W^T
    bdot(c, W, false, nc, W, false, nc, Gt + (size_t)nc * nc, nc);   // rows [nc, 2 nc): W W^T
    launch_gs_gram(Gt, nc, Tq, c->norms.as<double>(), d_count, c->stream);
    int* hc = (int*)(c->pinned() + 320);
    double* hwn = c->pinned() + 192;
    double* hq = c->pinned() + 384;
    TDA_HIP(hipMemcpyAsync(hc, d_count, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (wn) TDA_HIP(hipMemcpyAsync(hwn, c->norms.p, (size_t)nc * 8, hipMemcpyDeviceToHost, c->stream));
    // Q from all nc rows of T before the count is known (one synchronization
    // instead of two): row j of Q depends on row j of T only, so the first
    // count rows are those of the accepted vectors whatever T holds below.
    // The Gram estimate of |u_j| is a difference of large terms when the
    // block's vectors are nearly dependent among themselves (large
    // coefficients T): the accepted q_j must come out unit vectors, else the
    // block is redone vector by vector from the intact P and W.
    bcomb(c, Tq, nc, 1, nc, nc, W, false, Q);
    launch_norms(Q, nc, L, c->npart.as<double>(), c->norms.as<double>() + 64, c->stream);
    TDA_HIP(hipMemcpyAsync(hq, c->norms.as<double>() + 64, (size_t)nc * 8, hipMemcpyDeviceToHost, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
    if (wn) std::memcpy(wn, hwn, (size_t)nc * 8);
    if (!hc[1]) {
      const int nq = hc[0];
      if (nq == 0) return 0;
      bool unit = true;
      for (int j = 0; j < nq; ++j) unit = unit && std::fabs(hq[j] - 1.0) <= 1e-10;
      if (unit) return nq;
    }
    ++c->gs_fallbacks;
  }
  launch_gs_block(P, W, Q, L, nc, scr, d_count, c->stream);
  double* hwn = c->pinned() + 192;
  int* hc = (int*)(c->pinned() + 320);
  if (wn) {
    launch_norms(W, nc, L, c->npart.as<double>(), c->norms.as<double>(), c->stream);
    TDA_HIP(hipMemcpyAsync(hwn, c->norms.p, (size_t)nc * 8, hipMemcpyDeviceToHost, c->stream));
  }
  TDA_HIP(hipMemcpyAsync(hc, d_count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  TDA_HIP(hipStreamSynchronize(c->stream));
  if (wn) std::memcpy(wn, hwn, (size_t)nc * 8);
  return hc[0];
}

// Lowest n eigenpairs of the m x m projection H (upper triangle of the
// row-major [m][ldh] array; davidson.rs:93 eigh(UPLO::Upper)): evals [n]
// ascending, Y [n][ldh] (row r = eigenvector r).  Subspaces of up to 1024
// vectors and 64 roots run in the one-workgroup kernel of eigh.hip; larger
// ones (many roots: the basis holds 15 per root) in rocsolver_dsyevd on a
// copy (H's upper triangle is the column-major lower one; the column-major
// eigenvectors are the rows of Y).
void subspace_eigh(tda_ctx* c, const double* H, int ldh, int m, int n, double* evals, double* Y) {
  if (m <= 1024 && n <= 64) {
    launch_eigh(H, ldh, m, n, c->eigw.as<double>(), evals, Y, ldh, c->stream);
    return;
  }
  if (!c->blas) {
    if (rocblas_create_handle(&c->blas) != rocblas_status_success)
      throw Error{TDA_ERR_HIP, "rocblas_create_handle failed"};
  }
  if (rocblas_set_stream(c->blas, c->stream) != rocblas_status_success)
    throw Error{TDA_ERR_HIP, "rocblas_set_stream failed"};
  const size_t mm = (size_t)m * ldh;
  c->eigw.ensure(std::max(c->eigw.bytes, (mm + (size_t)ldh) * 8));
  c->ev_e.ensure((size_t)ldh * 8);
  c->ev_info.ensure(sizeof(rocblas_int));
  double* A = c->eigw.as<double>();
  double* D = A + mm;
  launch_copy(H, A, (int64_t)mm, c->stream);
  if (rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_lower, m, A, ldh, D, c->ev_e.as<double>(),
                       c->ev_info.as<rocblas_int>()) != rocblas_status_success)
    throw Error{TDA_ERR_HIP, "rocsolver_dsyevd failed"};
  launch_copy(A, Y, (int64_t)n * ldh, c->stream);
  launch_copy(D, evals, n, c->stream);
}

// Appends the nc <= 64 vectors in the rows [0, nc) of P to the basis
// V[0, m): classical Gram-Schmidt against the basis (W = P - V (V^T P)), then
// in-block against the accepted vectors in order with the 1e-8 drop
// (davidson.rs:162-168).  Returns the number appended at V[m].  With
// `reorth`, the loss of orthogonality eta of the basis is tracked: one
// classical pass against a basis orthogonal to eta leaves a kept vector
// q_j = u_j / |u_j| with components of about (eta + eps) |p_j| / |u_j| along
// it, so the loss grows by |p_j| / |u_j| with every block whose preconditioned
// residuals lie largely in the basis, and a nearly dependent basis returns
// spurious Ritz pairs near 0 whose residuals pass the test.  The block gets a
// second pass ("twice is enough": Daniel, Gragg, Kaufman and Stewart) once
// the estimate would exceed eta_tol; after it the new vectors are orthogonal
// to the basis to ~eps and eta stays where it was.
// v32: the basis is stored as FP32 (V as floats; the products with it are
// FP64): the new vectors are formed in FP64 staging and rounded to FP32 when
// appended, so the basis is orthogonal to ~FP32 rounding (eps32 = 6e-8) and
// eta is tracked from there.
// prec: the phase's product precision.  A basis needs orthogonality only to
// about its products' error: eta_tol32 in the FP16 / BF16 phase (~1e-4 per
// unit vector), as for an FP32 basis; 1e-7 in the FP32-accuracy phase (FP32
// MFMA or FP16 hi / lo split products, ~6e-8); eta_tol for FP64 products.
// (At C200H202 the FP64 threshold in the low-precision phases took second
// passes worth about 2 ms per solve that changed nothing they hand over.)
int append_block(tda_ctx* c, int m, int nc, bool reorth, double& eta, bool v32 = false, int prec = PREC_FP64) {
  const int64_t L = c->ovp;
  double* V = c->V.as<double>();
  float* V32 = (float*)c->V.p;
  const void* Vb = v32 ? (const void*)V32 : (const void*)V;
  double* P = c->Pb();
  double* W = c->Wb();
  double* C0 = c->stage2.as<double>();
  launch_copy(P, W, (int64_t)nc * L, c->stream);
  if (m > 0) {
    bdot(c, Vb, v32, m, P, false, nc, C0, nc);              // C0 = V P^T  [m][nc]
    bcomb(c, C0, 1, nc, nc, m, Vb, v32, W, -1.0, 1.0);      // W -= C0^T V
  }
  std::vector<double> pn(nc), wn(nc);
  if (reorth) {
    launch_norms(P, nc, L, c->npart.as<double>(), c->norms.as<double>(), c->stream);
    TDA_HIP(hipMemcpyAsync(c->pinned() + 128, c->norms.p, nc * 8, hipMemcpyDeviceToHost, c->stream));
  }
  if (v32) c->qstage.ensure((size_t)nc * L * 8);
  double* Q = v32 ? c->qstage.as<double>() : V + (size_t)m * L;
  int h_count = gs_block(c, P, W, Q, nc, reorth ? wn.data() : nullptr);
  if (reorth) std::memcpy(pn.data(), c->pinned() + 128, nc * 8);  // gs_block synchronized the stream
  bool need = false;
  if (reorth && c->reorth_mode != 0 && m > 0) {
    double grow = 0.0;
    for (int j = 0; j < nc; ++j)
      if (wn[j] > 1e-8) grow = std::max(grow, pn[j] / wn[j]);
    const double eps = v32 ? 6e-8 : 2.3e-16;
    const double eta_new = (eta + eps) * grow;
    // an FP32 basis cannot be orthogonal beyond its rounding: the second
    // pass is for losses well above it
    const double tol_p = (v32 || prec == PREC_FP16 || prec == PREC_BF16) ? c->eta_tol32
                         : prec == PREC_FP64                              ? c->eta_tol
                                                                          : std::max(c->eta_tol, 1e-7);
    need = c->reorth_mode == 1 || eta_new > tol_p;
    if (!need) eta = std::max(eta, eta_new);
  }
  if (need && h_count > 0) {
    ++c->reorth_passes;
    bdot(c, Vb, v32, m, Q, false, h_count, C0, h_count);
    bcomb(c, C0, 1, h_count, h_count, m, Vb, v32, Q, -1.0, 1.0);
    launch_copy(Q, P, (int64_t)h_count * L, c->stream);
    launch_copy(Q, W, (int64_t)h_count * L, c->stream);
    h_count = gs_block(c, P, W, Q, h_count, nullptr);
  }
  if (v32 && h_count > 0) launch_cast_vectors(Q, (int64_t)h_count * L, V32 + (size_t)m * L, PREC_FP32, c->stream);
  return h_count;
}

// Fills rows [0, n) of P with the guess unit vectors (their positions in the
// padded layout are in c->guess, set by tda_run).
void fill_guess(tda_ctx* c, double* P, int n, int first = 0) {
  const int64_t L = c->ovp;
  launch_fill(P, 0.0, (int64_t)n * L, c->stream);
  launch_set_ones(P, L, c->guess.as<int64_t>() + first, n, c->stream);
}

// The start block of the FP64 phase, orthonormalised in FP64: the nc rows at
// src (the previous phase's Ritz vectors) go through the block Gram-Schmidt
// (1e-8 drop) into V[0, m).  Ritz vectors of a nearly dependent
// low-precision basis can have norms far from 1 (and Ritz values near 0 with
// tiny residuals); orthonormalised, every FP64 residual of the next
// Rayleigh-Ritz step is that of a unit vector.  When fewer than n vectors
// survive, the guess unit vectors are projected against them and appended.
// Returns m >= n.
int orthonormal_start(tda_ctx* c, const double* src, int nc, int n) {
  const int64_t L = c->ovp;
  double* P = c->Pb();
  int m = 0;
  double eta = 1e-15;
  for (int b0 = 0; b0 < nc; b0 += 64) {
    const int nb = std::min(64, nc - b0);
    if (src + (size_t)b0 * L != P) launch_copy(src + (size_t)b0 * L, P, (int64_t)nb * L, c->stream);
    m += append_block(c, m, nb, true, eta);
  }
  for (int pass = 0; pass < 2 && m < n; ++pass) {
    for (int b0 = 0; b0 < n && m < n; b0 += 64) {
      const int nb = std::min(64, n - b0);
      fill_guess(c, P, nb, b0);
      m += append_block(c, m, nb, true, eta);
    }
  }
  TDA_CHECK(m >= n, TDA_ERR_NOT_CONVERGED, "Davidson: no orthonormal start block of n vectors");
  return std::min(m, c->space(n));
}

DavResult davidson(tda_ctx* c, int n, int m0, int max_iter, double tol, int prec, bool first_can_converge,
                   int stop_stall, std::vector<double>& h_evals, std::vector<double>& h_err, int cap_in = 0,
                   int keep = 0, int* kk_out = nullptr, double stall_factor = 0.5,
                   double stall_floor = 1e300, bool thick_restart = true, double floor_k = 0.0,
                   bool reorth = false, bool verify = false, bool v32 = false) {
  const int64_t L = c->ovp;
  // v32: V and AV hold FP32 vectors (V32 / AV32 over the same allocations);
  // the start vectors V[0, m0) arrive as FP64 and are rounded here.  RZ, R,
  // P, W and every exported vector stay FP64.
  TDA_CHECK(!v32 || (!verify && prec != PREC_FP64 && prec != PREC_F16X && !c->dense), TDA_ERR_BAD_ARG,
            "Davidson: an FP32 basis is for the FP16 / BF16 phase only");
  float* V32 = (float*)c->V.p;
  float* AV32 = (float*)c->AV.p;
  const int cap = cap_in > 0 ? cap_in : 15 * n;  // max_space (davidson.rs:74)
  const int ldh = c->space(n);                    // leading dimension of H and Y
  TDA_CHECK(cap <= ldh, TDA_ERR_BAD_ARG,
            "Davidson: a basis of " + std::to_string(cap) + " vectors exceeds the subspace buffers");
  if (keep < n) keep = n;
  const bool thick = thick_restart && keep > n;
  double* V = c->V.as<double>();
  double* AV = c->AV.as<double>();
  double* H = c->H.as<double>();
  double* Y = c->Y.as<double>();
  double* RZ = c->RZ.as<double>();
  double* R = c->R.as<double>();
  double* P = c->Pb();
  double* W = c->Wb();
  int m = m0, n_done = 0, dim_sub = m0;
  const int dim_sub_origin = m0;
  double eta = v32 ? 6e-8 : 1e-15;  // estimated loss of orthogonality of the basis
  if (v32) {
    launch_copy(V, W, (int64_t)m0 * L, c->stream);  // FP64 copy first: V32 overlays V
    launch_cast_vectors(W, (int64_t)m0 * L, V32, PREC_FP32, c->stream);
  }
  DavResult res;
  std::vector<int> idx;
  double best = 1e300, prev_ritz_sum = 1e300;
  int stall = 0;
  bool rotated = false;
  int n_eig_last = n, m_last = m;
  // on exit in thick mode: the n_eig_last lowest Ritz vectors into W (a
  // thicker start for the next phase)
  auto finish = [&](bool converged) {
    if (kk_out) *kk_out = n;
    if (!kk_out || keep <= n || converged) return;
    int kk = n_eig_last;
    if (kk < std::min(keep, m_last)) {
      // the phase kept the reference rule (n eigenpairs per iteration): one
      // more eigensolve of the final projection for the export
      kk = std::min(keep, m_last);
      subspace_eigh(c, H, ldh, m_last, kk, c->evals.as<double>(), Y);
    }
    if (kk <= n) return;
    bcomb(c, Y, ldh, 1, kk, m_last, v32 ? (const void*)V32 : (const void*)V, v32, W);
    *kk_out = kk;
  };
  for (int it = 0; it < max_iter; ++it) {
    res.iters = it + 1;
    // 1. products for the new columns (cache semantics solvers/mod.rs:290-302)
    int proj_from = n_done;
    if (rotated) {
      proj_from = 0;
      rotated = false;
    } else if (m <= n_done) {
      n_done = 0;
      proj_from = 0;
    }
    // n eigenpairs per iteration; a thick restart (and the export at the
    // end of a thick phase) solves the projection again for its keep
    // lowest pairs
    const int n_eig = n;
    n_eig_last = n_eig;
    m_last = m;
    {
      Timer t;
      TDA_HIP(hipEventRecord(t.a, c->stream));
      if (v32) products(c, V32 + (size_t)n_done * L, m - n_done, AV32 + (size_t)n_done * L, prec, true);
      else products(c, V + (size_t)n_done * L, m - n_done, AV + (size_t)n_done * L, prec);
      TDA_HIP(hipEventRecord(t.b, c->stream));
      // 1.1 projection H[p][q] = V_p . AV_q for the new q (davidson.rs:89)
      if (v32) bdot(c, V32, true, m, AV32 + (size_t)proj_from * L, true, m - proj_from, H + proj_from, ldh);
      else bdot(c, V, false, m, AV + (size_t)proj_from * L, false, m - proj_from, H + proj_from, ldh);
      n_done = m;
      // 2. subspace eigenproblem (davidson.rs:93-97)
      subspace_eigh(c, H, ldh, m, n_eig, c->evals.as<double>(), Y);
      // 3.1 Ritz vectors RZ = Y V, 3.2 residuals R = Y AV - u RZ, 3.3 norms
      // (davidson.rs:101-107)
      bcomb(c, Y, ldh, 1, n, m, v32 ? (const void*)V32 : (const void*)V, v32, RZ);
      bcomb(c, Y, ldh, 1, n, m, v32 ? (const void*)AV32 : (const void*)AV, v32, R);
      launch_residual(R, RZ, c->evals.as<double>(), n, L, c->stream);
      launch_norms(R, n, L, c->npart.as<double>(), c->norms.as<double>(), c->stream);
      h_evals.resize(n);
      h_err.resize(n);
      TDA_HIP(hipMemcpyAsync(c->pin_evals(), c->evals.p, n * 8, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipMemcpyAsync(c->pin_res(), c->norms.p, n * 8, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
      std::memcpy(h_evals.data(), c->pin_evals(), n * 8);
      std::memcpy(h_err.data(), c->pin_res(), n * 8);
      float ms = 0;
      TDA_HIP(hipEventElapsedTime(&ms, t.a, t.b));
      c->t_products_ms += ms;
      c->harvest_events();
    }
    bool finite = true;
    for (int r = 0; r < n; ++r) finite = finite && std::isfinite(h_err[r]) && std::isfinite(h_evals[r]);
    if (c->verbose || !finite) {
      double mx = 0, sum = 0;
      int cvd = 0;
      for (double e : h_err) { mx = std::max(mx, e); sum += e; cvd += (e < tol); }
      // columns as solvers/utils.rs:82-121 prints them, plus precision and basis size
      fprintf(stderr,
              "tincr_tda davidson prec=%d iter=%4d conv=%3d left=%3d dim_sub=%4d m=%4d total=%.3e max=%.3e "
              "w0=%.10f gs_fallbacks=%d reorth_passes=%d\n",
              prec, it + 1, cvd, n - cvd, dim_sub, m, sum, mx, h_evals[0], c->gs_fallbacks, c->reorth_passes);
    }
    if (!finite && prec != PREC_FP64 && !c->dense) {
      // a low-precision product overflowed: the phase ends here and the next
      // (wider) precision restarts from the basis' first n vectors, which
      // are finite (the phase's start vectors or its last restart's Ritz
      // vectors: only products and what is derived from them went bad)
      res.nonfinite = true;
      if (kk_out) *kk_out = n;
      if (v32) {  // the restart vectors back to FP64 in V[0, n)
        launch_widen(V32, (int64_t)n * L, W, c->stream);
        launch_copy(W, V, (int64_t)n * L, c->stream);
      }
      return res;
    }
    if (!finite)
      throw Error{TDA_ERR_HIP, "Davidson: non-finite residual norm or Ritz value (prec " + std::to_string(prec) +
                                   ", iteration " + std::to_string(it + 1) + ", basis " + std::to_string(m) + ")"};
    int roots_cvd = 0;
    double max_err = 0;
    for (double e : h_err) {
      roots_cvd += (e < tol);  // davidson.rs:115-117
      max_err = std::max(max_err, e);
    }
    res.max_err = max_err;
    const int roots_lft = n - roots_cvd;
    if (roots_lft == 0 && (it > 0 || first_can_converge)) {  // davidson.rs:121
      if (verify) {
        // the returned roots must be unit vectors: a residual below tol says
        // nothing about a Ritz vector of norm << 1 (nearly dependent basis)
        launch_norms(RZ, n, L, c->npart.as<double>(), c->norms.as<double>(), c->stream);
        double* hn = c->pin_res();
        TDA_HIP(hipMemcpyAsync(hn, c->norms.p, n * 8, hipMemcpyDeviceToHost, c->stream));
        TDA_HIP(hipStreamSynchronize(c->stream));
        bool unit = true;
        for (int r = 0; r < n; ++r) unit = unit && std::fabs(hn[r] - 1.0) <= 1e-10;
        if (!unit) {
          ++c->norm_restarts;
          if (c->verbose) fprintf(stderr, "tincr_tda davidson: Ritz vector norms off 1, restarting orthonormalised\n");
          m = dim_sub = orthonormal_start(c, RZ, n, n);
          n_done = 0;
          continue;
        }
      }
      res.converged = true;
      finish(true);
      return res;
    }
    if (floor_k > 0 && prec != PREC_FP64 && (prec == PREC_FP16 || prec == PREC_BF16 || tol < 1e-6) && !c->dense &&
        it == 0) {
      // product error of this precision on the lowest Ritz vector, against
      // FP64 products or, for FP16 / BF16 products of the AO operator, the
      // FP16 hi / lo split ones (error ~1e-7, three orders below theirs;
      // a third of an FP64 product's time).  The FP32-accuracy phases
      // (FP32 MFMA, FP16 hi / lo split) are measured only for tolerances
      // below 1e-6: their error, about 1e-7, otherwise lies far below tol, so
      // every rule it enters (hand-over below floor_k delta, near-floor stall,
      // Ritz-value progress) is decided by tol; a phase that stalls above tol
      // still hands over by the stall window.  That saves an FP64 product per
      // solve.
      const bool ao = c->has_lc && c->op_kind == 0 && !c->ct;
      const int ref_prec = (prec == PREC_FP16 || prec == PREC_BF16) && ao && c->split32 ? PREC_F16X : PREC_FP64;
      // A_prec z is already at hand: R[0] = (AV Y)[0] - theta_0 z, so
      // P = R[0] + theta_0 z (AV holds this phase's products; an FP32-stored
      // AV adds ~6e-8, far below the errors measured here)
      c->one.ensure(8);
      launch_copy(R, P, L, c->stream);
      launch_fill(c->one.as<double>(), -h_evals[0], 1, c->stream);
      launch_residual(P, RZ, c->one.as<double>(), 1, L, c->stream);  // P <- R[0] + theta_0 z
      products(c, RZ, 1, W, ref_prec);
      launch_fill(c->one.as<double>(), 1.0, 1, c->stream);
      launch_residual(W, P, c->one.as<double>(), 1, L, c->stream);  // W <- W - P
      launch_norms(W, 1, L, c->npart.as<double>(), c->norms.as<double>(), c->stream);
      TDA_HIP(hipMemcpyAsync(&res.delta, c->norms.p, 8, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
      if (c->verbose) fprintf(stderr, "tincr_tda davidson prec=%d product error delta=%.3e\n", prec, res.delta);
    }
    if (res.delta > 0 && max_err < floor_k * res.delta) {
      finish(false);
      return res;
    }
    // Progress of the Ritz values: while the sum of the n lowest still falls
    // by more than n times the phase's product error per iteration, the
    // subspace is still finding the low states (a dense spectrum, states
    // below every orbital-energy difference) and a residual that does not
    // shrink is no sign of the precision floor.
    double ritz_sum = 0.0;
    for (double e : h_evals) ritz_sum += e;
    const bool ritz_moving = prev_ritz_sum - ritz_sum > n * std::max(res.delta, tol);
    prev_ritz_sum = ritz_sum;
    if (stop_stall > 0 && max_err < stall_floor) {
      const bool near = res.delta > 0 && max_err < c->near_k * res.delta;
      if (max_err < stall_factor * best || (ritz_moving && !near)) {
        best = std::min(best, max_err);
        stall = 0;
      } else if (++stall >= (near ? std::min(stop_stall, c->near_stall) : stop_stall)) {
        finish(false);
        return res;
      }
    }
    if (dim_sub + roots_lft <= cap) {  // davidson.rs:146
      idx.clear();
      for (int r = 0; r < n; ++r)
        if (h_err[r] > tol) idx.push_back(r);  // davidson.rs:150-153
      const int nc = (int)idx.size();
      dim_sub += roots_lft;  // davidson.rs:159 (counts dropped vectors too)
      if (nc > 0) {
        int* hidx = c->pin_idx();  // not in flight: the last copy from it finished at a sync
        std::memcpy(hidx, idx.data(), nc * sizeof(int));
        TDA_HIP(hipMemcpyAsync(c->idx.p, hidx, nc * sizeof(int), hipMemcpyHostToDevice, c->stream));
        // preconditioned residuals in blocks of at most 64, each projected
        // against the basis as it stands (earlier blocks included): the
        // reference's vector-by-vector classical Gram-Schmidt against the
        // growing basis, in order (davidson.rs:162-168)
        for (int b0 = 0; b0 < nc; b0 += 64) {
          const int nb = std::min(64, nc - b0);
          launch_precondition_many(R, c->evals.as<double>(), c->idx.as<int>() + b0, nb, c->de.as<double>(), L,
                                   c->pre_thr, P, c->stream);
          m += append_block(c, m, nb, reorth, eta, v32, prec);
        }
      }
    } else if (thick) {
      // thick restart: rotate basis and products onto the kk lowest Ritz vectors
      const int kk = std::min(keep, m);
      if (kk > n_eig) subspace_eigh(c, H, ldh, m, kk, c->evals.as<double>(), Y);
      bcomb(c, Y, ldh, 1, kk, m, v32 ? (const void*)V32 : (const void*)V, v32, W);
      bcomb(c, Y, ldh, 1, kk, m, v32 ? (const void*)AV32 : (const void*)AV, v32, P);
      if (v32) {
        launch_cast_vectors(W, (int64_t)kk * L, V32, PREC_FP32, c->stream);
        launch_cast_vectors(P, (int64_t)kk * L, AV32, PREC_FP32, c->stream);
      } else {
        launch_copy(W, V, (int64_t)kk * L, c->stream);
        launch_copy(P, AV, (int64_t)kk * L, c->stream);
      }
      m = n_done = dim_sub = kk;
      rotated = true;
    } else {
      // collapse to the Ritz vectors (davidson.rs:172-176); RZ holds n of
      // them, so a phase started from more than n vectors collapses to n
      dim_sub = std::min(dim_sub_origin, n);
      if (v32) launch_cast_vectors(RZ, (int64_t)dim_sub * L, V32, PREC_FP32, c->stream);
      else launch_copy(RZ, V, (int64_t)dim_sub * L, c->stream);
      m = dim_sub;
      n_done = 0;
    }
  }
  finish(false);
  return res;
}

void alloc_davidson(tda_ctx* c, int n, int keep) {
  const int64_t L = c->ovp;
  const int cap = c->space(n);
  keep = std::max(keep, n);
  c->V.ensure((size_t)cap * L * 8);
  c->AV.ensure((size_t)cap * L * 8);
  c->H.ensure((size_t)cap * cap * 8);
  c->Y.ensure((size_t)keep * cap * 8);
  c->evals.ensure((size_t)keep * 8);
  c->RZ.ensure((size_t)n * L * 8);
  c->R.ensure((size_t)n * L * 8);
  c->PW.ensure((size_t)2 * keep * L * 8);
  c->pw_rows = (int64_t)(c->PW.bytes / ((size_t)2 * L * 8));
  // Gram-Schmidt blocks hold up to n new vectors, or the keep Ritz vectors
  // that orthonormal_start passes through gs_block
  const int nb = std::max(n, keep);
  c->ensure_pinned(nb);
  c->norms.ensure((size_t)(nb + 64) * 8);  // [64, 128): norms of a Gram-Schmidt block's q_j
  c->npart.ensure((size_t)nb * 256 * 8);
  c->idx.ensure((size_t)n * sizeof(int));
  c->count.ensure(sizeof(int) * 4);
  c->gs_scratch.ensure(gs_block_scratch_doubles(nb) * 8);
  c->eigw.ensure(eigh_work_doubles(std::min(cap, 1024)) * 8);
  c->stage2.ensure((size_t)cap * n * 8);
}

void check_sizes(int n_atoms, int n_ao, const int32_t* opa, int occ_begin, int n_occ, int virt_begin, int n_virt) {
  TDA_CHECK(n_atoms > 0 && n_ao > 0, TDA_ERR_BAD_ARG, "n_atoms and n_ao must be positive");
  TDA_CHECK(opa != nullptr, TDA_ERR_BAD_ARG, "orbs_per_atom is NULL");
  int64_t s = 0;
  for (int i = 0; i < n_atoms; ++i) {
    TDA_CHECK(opa[i] >= 0, TDA_ERR_BAD_ARG, "negative orbs_per_atom");
    s += opa[i];
  }
  TDA_CHECK(s == n_ao, TDA_ERR_BAD_ARG, "sum(orbs_per_atom) != n_ao");
  TDA_CHECK(n_occ > 0 && n_virt > 0, TDA_ERR_BAD_ARG, "n_occ and n_virt must be positive");
  TDA_CHECK(occ_begin >= 0 && occ_begin + n_occ <= n_ao, TDA_ERR_BAD_ARG, "occupied range outside the MOs");
  TDA_CHECK(virt_begin >= 0 && virt_begin + n_virt <= n_ao, TDA_ERR_BAD_ARG, "virtual range outside the MOs");
}

std::vector<int> ao_offsets(int n_atoms, const int32_t* opa) {
  std::vector<int> off(n_atoms + 1, 0);
  for (int i = 0; i < n_atoms; ++i) off[i + 1] = off[i] + opa[i];
  return off;
}

template <typename F>
int guarded(tda_ctx* c, F&& f) {
  struct StreamScope {  // DevBuf regrowth drains this context's stream
    hipStream_t keep;
    explicit StreamScope(hipStream_t s) : keep(tl_stream) { tl_stream = s; }
    ~StreamScope() { tl_stream = keep; }
  } scope(c ? c->stream : nullptr);
  try {
    if (c) set_device(c);
    f();
    return TDA_OK;
  } catch (const Error& e) {
    if (c) c->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    if (c) c->err = "host allocation failed";
    return TDA_ERR_OOM;
  } catch (...) {
    if (c) c->err = "unknown error";
    return TDA_ERR_HIP;
  }
}

void copy_unpad_rows(const double* src, int rows, int cols, int ld, double* dst) {
  for (int r = 0; r < rows; ++r) std::memcpy(dst + (size_t)r * cols, src + (size_t)r * ld, (size_t)cols * 8);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int tda_ctx_create(int device, size_t hbm_budget_bytes, int precision, tda_ctx** out) {
  if (!out) return TDA_ERR_BAD_ARG;
  *out = nullptr;
  if (precision < TDA_PREC_FP64 || precision > TDA_PREC_FP16) return TDA_ERR_BAD_ARG;
  tda_ctx* c = nullptr;
  try {
    c = new tda_ctx();
  } catch (...) {
    return TDA_ERR_OOM;
  }
  c->device = device;
  c->prec = precision;
  {
    const char* e = getenv("TINCR_TDA_VERBOSE");
    c->verbose = e && e[0] && e[0] != '0';
    const char* ec = getenv("TINCR_TDA_LP_CAP");
    const char* ek = getenv("TINCR_TDA_LP_KEEP");
    if (ec && atoi(ec) >= 2) c->lp_cap_mult = atoi(ec);
    if (ek && atoi(ek) >= 1) c->lp_keep_mult = atoi(ek);
    const char* ef = getenv("TINCR_TDA_FLOOR_K");
    const char* es = getenv("TINCR_TDA_STALL");
    if (ef && atof(ef) >= 0) c->floor_k = atof(ef);
    if (es && atoi(es) >= 1) c->stall_window = atoi(es);
    const char* et = getenv("TINCR_TDA_FP32_THICK");
    if (et) c->fp32_thick = atoi(et) != 0;
    const char* e3 = getenv("TINCR_TDA_FP32_CAP");
    if (e3 && atoi(e3) >= 2) c->fp32_cap_mult = atoi(e3);
    const char* ep = getenv("TINCR_TDA_SPLIT32");
    if (ep) c->split32 = atoi(ep) != 0;
    const char* el = getenv("TINCR_TDA_LADDER_MIN_FLOPS");
    if (el && atof(el) >= 0) c->ladder_min_flops = atof(el);
    const char* eg = getenv("TINCR_TDA_GS_GRAM");
    if (eg) c->gs_gram = atoi(eg) != 0;
    const char* egf = getenv("TINCR_TDA_GFULL");
    if (egf) c->no_gfull = atoi(egf) == 0;
    const char* eas = getenv("TINCR_TDA_AO_SCALE");
    if (eas) c->ao_scale = atoi(eas) != 0;
    const char* elp = getenv("TINCR_TDA_LP_PHASE");
    if (elp) c->lp_phase = atoi(elp) != 0;
    const char* er = getenv("TINCR_TDA_REORTH");
    if (er && (atoi(er) == 0 || atoi(er) == 1)) c->reorth_mode = atoi(er);
    const char* ee = getenv("TINCR_TDA_ETA_TOL");
    if (ee && atof(ee) > 0) c->eta_tol = atof(ee);
    const char* eb = getenv("TINCR_TDA_BASIS32");
    if (eb) c->basis32 = atoi(eb) != 0 ? 1 : 0;
    const char* e32 = getenv("TINCR_TDA_ETA_TOL32");
    if (e32 && atof(e32) > 0) c->eta_tol32 = atof(e32);
    const char* enk = getenv("TINCR_TDA_NEAR_K");
    if (enk && atof(enk) >= 0) c->near_k = atof(enk);
    const char* ens = getenv("TINCR_TDA_NEAR_STALL");
    if (ens && atoi(ens) >= 1) c->near_stall = atoi(ens);
  }
  if (hbm_budget_bytes) c->budget = hbm_budget_bytes;
  int rc = guarded(c, [&] {
    int ndev = 0;
    TDA_HIP(hipGetDeviceCount(&ndev));
    TDA_CHECK(device >= 0 && device < ndev, TDA_ERR_BAD_ARG, "no such HIP device");
    TDA_HIP(hipSetDevice(device));
    TDA_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  });
  if (rc != TDA_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return TDA_OK;
}

void tda_ctx_destroy(tda_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  delete c;
}

const char* tda_last_error(const tda_ctx* c) { return c ? c->err.c_str() : "null context"; }

int tda_set_precision(tda_ctx* c, int precision) {
  if (!c) return TDA_ERR_BAD_ARG;
  if (precision < TDA_PREC_FP64 || precision > TDA_PREC_FP16) {
    c->err = "bad precision";
    return TDA_ERR_BAD_ARG;
  }
  c->prec = precision;
  return TDA_OK;
}

int tda_set_operator(tda_ctx* c, int op) {
  if (!c) return TDA_ERR_BAD_ARG;
  if (op != TDA_OP_AO && op != TDA_OP_CHARGES) {
    c->err = "bad operator form";
    return TDA_ERR_BAD_ARG;
  }
  if (op != c->op_kind) {
    c->op_kind = op;
    c->prepared = false;
    reset_lowprec(c);
  }
  return TDA_OK;
}

int tda_orbe_differences(const double* occ, int n_occ, const double* virt, int n_virt, double* omega) {
  if (!occ || !virt || !omega || n_occ < 0 || n_virt < 0) return TDA_ERR_BAD_ARG;
  for (int i = 0; i < n_occ; ++i)
    for (int a = 0; a < n_virt; ++a) omega[(size_t)i * n_virt + a] = virt[a] - occ[i];
  return TDA_OK;
}

int tda_trans_charges(tda_ctx* c, int n_atoms, int n_ao, const int32_t* opa, const double* orbs, const double* s,
                      int occ_begin, int n_occ, int virt_begin, int n_virt, double* q_ov, double* q_oo,
                      double* q_vv) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    check_sizes(n_atoms, n_ao, opa, occ_begin, n_occ, virt_begin, n_virt);
    TDA_CHECK(orbs && s, TDA_ERR_BAD_ARG, "orbs or s is NULL");
    DevBuf dC, dS, dSC, dao, dq;
    const size_t nn = (size_t)n_ao * n_ao;
    dC.ensure(nn * 8);
    dS.ensure(nn * 8);
    dSC.ensure(nn * 8);
    std::vector<int> off = ao_offsets(n_atoms, opa);
    dao.ensure(off.size() * sizeof(int));
    TDA_HIP(hipMemcpyAsync(dC.p, orbs, nn * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(dS.p, s, nn * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(dao.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    gemm_sc(c, dS.as<double>(), dC.as<double>(), dSC.as<double>(), n_ao);
    struct Blk { double* out; int pb, P, qb, Q; };
    Blk blks[3] = {{q_ov, occ_begin, n_occ, virt_begin, n_virt},
                   {q_oo, occ_begin, n_occ, occ_begin, n_occ},
                   {q_vv, virt_begin, n_virt, virt_begin, n_virt}};
    for (const Blk& b : blks) {
      if (!b.out) continue;
      const size_t cnt = (size_t)n_atoms * b.P * b.Q;
      dq.ensure(cnt * 8);
      launch_qblock(dC.as<double>(), dSC.as<double>(), n_ao, dao.as<int>(), n_atoms, b.pb, b.P, b.P, b.qb, b.Q,
                    b.Q, dq.as<double>(), c->stream);
      TDA_HIP(hipMemcpyAsync(b.out, dq.p, cnt * 8, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
    }
    TDA_HIP(hipStreamSynchronize(c->stream));
    for (DevBuf* b : {&dC, &dS, &dSC, &dao, &dq}) b->release();
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free transition charges.  The only device memory beyond the
// caller's inputs and outputs is `scratch` below: S C [n_ao][n_ao] FP64, the
// AO offsets [n_atoms + 1] and a flag.  Each GPU the library runs on holds
// its own copy of the array until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns for the current
// device (the array's name is a host-side stand-in, not a device address).
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchBytes = TDA_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;  // guards scratch_ptr and serialises the calls per process
std::array<void*, kMaxDevices> scratch_ptr{};

// Device address of `scratch` on the current device (set up on the first
// call on that device, with it current).
void* scratch_address(int device) {
  TDA_CHECK(device >= 0 && device < kMaxDevices, TDA_ERR_BAD_ARG, "device index out of range");
  if (!scratch_ptr[device]) {
    void* ptr = nullptr;
    TDA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    TDA_CHECK(ptr != nullptr, TDA_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    scratch_ptr[device] = ptr;
  }
  return scratch_ptr[device];
}
}  // namespace

extern "C" {

int tda_trans_charges_noalloc(tda_ctx* c, int n_atoms, int n_ao, const int32_t* d_opa, const double* d_orbs,
                              const double* d_s, int occ_begin, int n_occ, int virt_begin, int n_virt,
                              double* d_q_ov, double* d_q_oo, double* d_q_vv) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(n_atoms > 0 && n_ao > 0, TDA_ERR_BAD_ARG, "n_atoms and n_ao must be positive");
    TDA_CHECK(n_occ > 0 && n_virt > 0, TDA_ERR_BAD_ARG, "n_occ and n_virt must be positive");
    TDA_CHECK(occ_begin >= 0 && occ_begin + n_occ <= n_ao, TDA_ERR_BAD_ARG, "occupied range outside the MOs");
    TDA_CHECK(virt_begin >= 0 && virt_begin + n_virt <= n_ao, TDA_ERR_BAD_ARG, "virtual range outside the MOs");
    TDA_CHECK(d_opa && d_orbs && d_s, TDA_ERR_BAD_ARG, "NULL input array");
    const size_t sc_bytes = (size_t)round_up((int64_t)n_ao * n_ao * 8, 256);
    const size_t need = sc_bytes + ((size_t)n_atoms + 2) * sizeof(int);
    TDA_CHECK(need <= kScratchBytes, TDA_ERR_BAD_ARG,
              "tda_trans_charges_noalloc: n_ao = " + std::to_string(n_ao) + ", n_atoms = " +
                  std::to_string(n_atoms) + " need " + std::to_string(need) + " scratch bytes, more than the " +
                  std::to_string(kScratchBytes) + " available");
    std::lock_guard<std::mutex> lock(scratch_mu);
    unsigned char* base = (unsigned char*)scratch_address(c->device);
    double* dSC = (double*)base;
    int* off = (int*)(base + sc_bytes);
    int* bad = off + n_atoms + 1;
    launch_ao_offsets(d_opa, n_atoms, n_ao, off, bad, c->stream);
    // SC = S C without split-K (a split would need a workspace)
    GemmArgs g = gargs(d_s, d_orbs, dSC, n_ao, n_ao, n_ao, plain(n_ao), plain(1), plain(n_ao), plain(1), plain(n_ao),
                       plain(1));
    GemmPlan plan = gemm_plan(DT_F64, n_ao, n_ao, n_ao, 1, true, false, 1);
    TDA_CHECK(gemm_workspace_bytes(DT_F64, plan, n_ao, n_ao, 1) == 0, TDA_ERR_HIP, "S C plan needs a workspace");
    gemm_launch(DT_F64, DT_F64, EPI_STORE, true, false, g, plan, c->stream);
    struct Blk { double* out; int pb, P, qb, Q; };
    const Blk blks[3] = {{d_q_ov, occ_begin, n_occ, virt_begin, n_virt},
                         {d_q_oo, occ_begin, n_occ, occ_begin, n_occ},
                         {d_q_vv, virt_begin, n_virt, virt_begin, n_virt}};
    for (const Blk& b : blks)
      if (b.out) launch_qblock(d_orbs, dSC, n_ao, off, n_atoms, b.pb, b.P, b.P, b.qb, b.Q, b.Q, b.out, c->stream);
    int h_bad = 0;
    TDA_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
    TDA_CHECK(!h_bad, TDA_ERR_BAD_ARG, "orbs_per_atom: negative counts or sum != n_ao (outputs undefined)");
  });
}

int tda_set_system(tda_ctx* c, int n_atoms, int n_ao, const int32_t* opa, const double* orbs, const double* s,
                   const double* orbe, int occ_begin, int n_occ, int virt_begin, int n_virt, const double* gamma,
                   const double* gamma_lr, const double* xyz) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    check_sizes(n_atoms, n_ao, opa, occ_begin, n_occ, virt_begin, n_virt);
    TDA_CHECK(orbs && s && orbe && gamma && xyz, TDA_ERR_BAD_ARG, "NULL input array");
    c->has_system = false;
    c->prepared = false;
    reset_lowprec(c);
    c->ct = false;
    c->dense = false;
    c->pre_thr = 1e-4;
    c->last_roots = 0;
    c->N = n_atoms;
    c->Nq = c->Nx = n_atoms;
    c->nao = n_ao;
    c->ob = occ_begin;
    c->o = n_occ;
    c->vb = virt_begin;
    c->v = n_virt;
    c->op = (int)round_up(n_occ, 64);
    c->vp = (int)round_up(n_virt, 64);
    c->ovp = (int64_t)c->o * c->vp;
    c->has_lc = gamma_lr != nullptr;
    const size_t nn = (size_t)n_ao * n_ao;
    c->C.ensure(nn * 8);
    c->S.ensure(nn * 8);
    c->orbe.ensure((size_t)n_ao * 8);
    c->gamma.ensure((size_t)n_atoms * n_atoms * 8);
    std::vector<int> off = ao_offsets(n_atoms, opa);
    c->aooff.ensure(off.size() * sizeof(int));
    TDA_HIP(hipMemcpyAsync(c->C.p, orbs, nn * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(c->S.p, s, nn * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(c->orbe.p, orbe, (size_t)n_ao * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(c->gamma.p, gamma, (size_t)n_atoms * n_atoms * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(c->aooff.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (c->has_lc) {
      c->gamma_lr.ensure((size_t)n_atoms * n_atoms * 8);
      TDA_HIP(hipMemcpyAsync(c->gamma_lr.p, gamma_lr, (size_t)n_atoms * n_atoms * 8, hipMemcpyHostToDevice,
                             c->stream));
    }
    {
      std::vector<int> atom_of((size_t)n_ao);
      for (int A = 0; A < n_atoms; ++A)
        for (int mu = off[A]; mu < off[A + 1]; ++mu) atom_of[mu] = A;
      c->ao_atom.ensure((size_t)n_ao * sizeof(int));
      TDA_HIP(hipMemcpyAsync(c->ao_atom.p, atom_of.data(), (size_t)n_ao * sizeof(int), hipMemcpyHostToDevice,
                             c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
    }
    c->mo_built = false;
    c->xyz.assign(xyz, xyz + (size_t)n_atoms * 3);
    c->h_de.resize((size_t)n_occ * n_virt);
    for (int i = 0; i < n_occ; ++i)
      for (int a = 0; a < n_virt; ++a) c->h_de[(size_t)i * n_virt + a] = orbe[virt_begin + a] - orbe[occ_begin + i];
    TDA_HIP(hipStreamSynchronize(c->stream));
    c->has_system = true;
  });
}

int tda_set_ct_pair(tda_ctx* c, int n_atoms_h, int n_ao_h, const int32_t* opa_h, const double* orbs_h,
                    const double* s_h, const double* orbe_h, int occ_begin_h, int n_occ_h, int n_atoms_l, int n_ao_l,
                    const int32_t* opa_l, const double* orbs_l, const double* s_l, const double* orbe_l,
                    int virt_begin_l, int n_virt_l, const double* s_hl, const double* gamma_h,
                    const double* gamma_l, const double* gamma_hl, const double* gamma_lr_hl, const double* xyz) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    check_sizes(n_atoms_h, n_ao_h, opa_h, occ_begin_h, n_occ_h, 0, 1);
    check_sizes(n_atoms_l, n_ao_l, opa_l, 0, 1, virt_begin_l, n_virt_l);
    TDA_CHECK(orbs_h && s_h && orbe_h && orbs_l && s_l && orbe_l && s_hl && gamma_h && gamma_l && gamma_hl && xyz,
              TDA_ERR_BAD_ARG, "NULL input array");
    c->has_system = false;
    c->prepared = false;
    reset_lowprec(c);
    const int NH = n_atoms_h, NL = n_atoms_l, N = NH + NL;
    c->ct = true;
    c->dense = false;
    c->pre_thr = 1e-4;
    c->last_roots = 0;
    c->N = N;
    c->Nq = NH;
    c->Nx = NL;
    c->nao = n_ao_h;
    c->nao_l = n_ao_l;
    c->lumo_l = virt_begin_l;
    c->ob = occ_begin_h;
    c->o = n_occ_h;
    c->vb = n_ao_h + virt_begin_l;  // index into the stacked [orbe_h; orbe_l]
    c->v = n_virt_l;
    c->op = (int)round_up(n_occ_h, 64);
    c->vp = (int)round_up(n_virt_l, 64);
    c->ovp = (int64_t)c->o * c->vp;
    c->has_lc = gamma_lr_hl != nullptr;
    const size_t nhh = (size_t)n_ao_h * n_ao_h, nll = (size_t)n_ao_l * n_ao_l, nhl = (size_t)n_ao_h * n_ao_l;
    c->C.ensure(nhh * 8);
    c->S.ensure(nhh * 8);
    c->Cl.ensure(nll * 8);
    c->Sl.ensure(nll * 8);
    c->Shl.ensure(nhl * 8);
    TDA_HIP(hipMemcpyAsync(c->C.p, orbs_h, nhh * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(c->S.p, s_h, nhh * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(c->Cl.p, orbs_l, nll * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(c->Sl.p, s_l, nll * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(c->Shl.p, s_hl, nhl * 8, hipMemcpyHostToDevice, c->stream));
    // stacked orbital energies: omega[i, a] = orbe_l[lumo + a] - orbe_h[i] (ct_pair.rs:43-58)
    std::vector<double> e((size_t)n_ao_h + n_ao_l);
    std::memcpy(e.data(), orbe_h, (size_t)n_ao_h * 8);
    std::memcpy(e.data() + n_ao_h, orbe_l, (size_t)n_ao_l * 8);
    c->orbe.ensure(e.size() * 8);
    TDA_HIP(hipMemcpyAsync(c->orbe.p, e.data(), e.size() * 8, hipMemcpyHostToDevice, c->stream));
    // pair gamma from the monomer blocks and the H-L block of the full gamma (ct_pair.rs:20-27)
    std::vector<double> g((size_t)N * N);
    for (int A = 0; A < NH; ++A)
      for (int B = 0; B < NH; ++B) g[(size_t)A * N + B] = gamma_h[(size_t)A * NH + B];
    for (int A = 0; A < NL; ++A)
      for (int B = 0; B < NL; ++B) g[(size_t)(NH + A) * N + NH + B] = gamma_l[(size_t)A * NL + B];
    for (int A = 0; A < NH; ++A)
      for (int B = 0; B < NL; ++B) {
        g[(size_t)A * N + NH + B] = gamma_hl[(size_t)A * NL + B];
        g[(size_t)(NH + B) * N + A] = gamma_hl[(size_t)A * NL + B];
      }
    c->gamma.ensure(g.size() * 8);
    TDA_HIP(hipMemcpyAsync(c->gamma.p, g.data(), g.size() * 8, hipMemcpyHostToDevice, c->stream));
    std::vector<double> glr_t;
    if (c->has_lc) {
      // stored transposed, [NL][NH], so that G = gamma_lr^T q_oo is one GEMM
      glr_t.resize((size_t)NL * NH);
      for (int A = 0; A < NH; ++A)
        for (int B = 0; B < NL; ++B) glr_t[(size_t)B * NH + A] = gamma_lr_hl[(size_t)A * NL + B];
      c->gamma_lr.ensure(glr_t.size() * 8);
      TDA_HIP(hipMemcpyAsync(c->gamma_lr.p, glr_t.data(), glr_t.size() * 8, hipMemcpyHostToDevice, c->stream));
    }
    std::vector<int> off_h = ao_offsets(NH, opa_h), off_l = ao_offsets(NL, opa_l);
    c->aooff.ensure(off_h.size() * sizeof(int));
    c->aooff_l.ensure(off_l.size() * sizeof(int));
    TDA_HIP(hipMemcpyAsync(c->aooff.p, off_h.data(), off_h.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipMemcpyAsync(c->aooff_l.p, off_l.data(), off_l.size() * sizeof(int), hipMemcpyHostToDevice,
                           c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));  // host staging vectors go out of scope
    c->mo_built = false;
    c->xyz.assign(xyz, xyz + (size_t)N * 3);
    c->h_de.resize((size_t)n_occ_h * n_virt_l);
    for (int i = 0; i < n_occ_h; ++i)
      for (int a = 0; a < n_virt_l; ++a)
        c->h_de[(size_t)i * n_virt_l + a] = orbe_l[virt_begin_l + a] - orbe_h[occ_begin_h + i];
    c->has_system = true;
  });
}

int tda_set_dense(tda_ctx* c, int dim, const double* h) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(dim > 0 && h, TDA_ERR_BAD_ARG, "dim must be positive and h non-NULL");
    c->has_system = false;
    c->prepared = false;
    reset_lowprec(c);
    c->ct = false;
    c->dense = true;
    c->pre_thr = 0.0;
    c->last_roots = 0;
    c->has_lc = false;
    c->N = c->Nq = c->Nx = 0;
    c->nao = 0;
    c->ob = c->vb = 0;
    c->o = 1;
    c->v = dim;
    c->op = 64;
    c->vp = (int)round_up(dim, 64);
    c->ovp = c->vp;
    const int vp = c->vp;
    std::vector<double> hp((size_t)vp * vp, 0.0), d((size_t)vp, 1e300);
    for (int r = 0; r < dim; ++r) {
      std::memcpy(hp.data() + (size_t)r * vp, h + (size_t)r * dim, (size_t)dim * 8);
      d[r] = h[(size_t)r * dim + r];
    }
    c->Hd.ensure(hp.size() * 8);
    c->de.ensure(d.size() * 8);
    TDA_HIP(hipMemcpyAsync(c->Hd.p, hp.data(), hp.size() * 8, hipMemcpyHostToDevice, c->stream));
    // diagonal preconditioner; pad entries 1e300 so that pads stay 0
    TDA_HIP(hipMemcpyAsync(c->de.p, d.data(), d.size() * 8, hipMemcpyHostToDevice, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
    c->h_de.assign(d.begin(), d.begin() + dim);  // guess from diag(H) (basis.rs:500)
    c->xyz.clear();
    c->mo_built = false;
    c->has_system = true;
  });
}

int tda_prepare(tda_ctx* c) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] { prepare(c); });
}

int tda_clear_charges(tda_ctx* c) {
  if (!c) return TDA_ERR_BAD_ARG;
  c->prepared = false;
  reset_lowprec(c);
  return TDA_OK;
}

int tda_get_charges(tda_ctx* c, double* q_ov, double* q_oo, double* q_vv, double* omega) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(!c->dense, TDA_ERR_STATE, "a dense engine has no transition charges");
    prepare(c);
    const int N = c->N, o = c->o, v = c->v, op = c->op, vp = c->vp;
    if (q_ov) {
      if (!c->qov_built) {  // the AO operator does not keep q_ov: build it for the caller
        c->qov.ensure((size_t)N * c->ovp * 8);
        build_charges(c, c->C.as<double>(), c->SC.as<double>(), c->aooff.as<int>(), N, c->nao, c->ob, c->o, c->vb,
                      c->v, c->op, c->vp, c->qov.as<double>(), nullptr, nullptr);
        TDA_HIP(hipStreamSynchronize(c->stream));
        c->qov_built = true;
      }
      std::vector<double> h((size_t)N * c->ovp);
      TDA_HIP(hipMemcpy(h.data(), c->qov.p, h.size() * 8, hipMemcpyDeviceToHost));
      copy_unpad_rows(h.data(), N * o, v, vp, q_ov);
    }
    if ((q_oo || q_vv) && !c->has_lc) throw Error{TDA_ERR_STATE, "q_oo/q_vv are built only with gamma_lr"};
    if (q_oo || q_vv) build_mo_charges(c);
    if (q_oo) {
      std::vector<double> h((size_t)c->Nq * op * op);
      TDA_HIP(hipMemcpy(h.data(), c->qoo.p, h.size() * 8, hipMemcpyDeviceToHost));
      for (int A = 0; A < c->Nq; ++A) copy_unpad_rows(h.data() + (size_t)A * op * op, o, o, op, q_oo + (size_t)A * o * o);
    }
    if (q_vv) {
      std::vector<double> h((size_t)c->Nx * vp * vp);
      TDA_HIP(hipMemcpy(h.data(), c->qvv.p, h.size() * 8, hipMemcpyDeviceToHost));
      for (int A = 0; A < c->Nx; ++A) copy_unpad_rows(h.data() + (size_t)A * vp * vp, v, v, vp, q_vv + (size_t)A * v * v);
    }
    if (omega) std::memcpy(omega, c->h_de.data(), c->h_de.size() * 8);
  });
}

int tda_compute_products(tda_ctx* c, const double* x, int k, double* ax) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(k >= 0 && (k == 0 || (x && ax)), TDA_ERR_BAD_ARG, "bad x/ax/k");
    prepare(c);
    if (k == 0) return;
    const int64_t ov = (int64_t)c->o * c->v;
    c->stage.ensure((size_t)ov * k * 8);
    c->ax_tmp.ensure((size_t)c->ovp * k * 8 * 2);
    double* dX = c->ax_tmp.as<double>();
    double* dAX = dX + (size_t)c->ovp * k;
    TDA_HIP(hipMemcpyAsync(c->stage.p, x, (size_t)ov * k * 8, hipMemcpyHostToDevice, c->stream));
    launch_pack_x(c->stage.as<double>(), c->o, c->v, c->vp, k, dX, c->stream);
    products(c, dX, k, dAX, c->prec);
    launch_unpack_x(dAX, c->o, c->v, c->vp, k, c->stage.as<double>(), c->stream);
    TDA_HIP(hipMemcpyAsync(ax, c->stage.p, (size_t)ov * k * 8, hipMemcpyDeviceToHost, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
    c->harvest_events();
  });
}

// Z-vector solve of the excited-state gradient (gradients/helpers.rs:
// 1130-1223 with LC, :1053-1128 without): conjugate gradients on
// A z = rhs for each column, A the singlet product of mult_av_lc
// (helpers.rs:1384-1437, the forward A.x: de + 2 J - K) or, without LC,
// mult_av_nolc (:1351-1382: de + 4 J); start z = rhs / de, stop when
// r.r <= conv (1e-19 there) or after max_iter iterations (1e4 there).
int tda_zvector(tda_ctx* c, const double* rhs, int k, int max_iter, double conv, double* z, int32_t* n_iter) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(c->has_system && !c->dense, TDA_ERR_STATE, "no TDA system loaded");
    TDA_CHECK(k >= 0 && (k == 0 || (rhs && z)) && max_iter >= 0, TDA_ERR_BAD_ARG, "bad rhs/z/k/max_iter");
    prepare(c);
    if (k == 0) return;
    const int64_t L = c->ovp, ov = (int64_t)c->o * c->v;
    c->stage.ensure((size_t)ov * k * 8);
    c->zvw.ensure((size_t)4 * k * L * 8);
    c->npart.ensure((size_t)k * 256 * 8);
    c->one.ensure((size_t)3 * k * 8);
    double* dR = c->zvw.as<double>();
    double* dZ = dR + (size_t)k * L;
    double* dP = dZ + (size_t)k * L;
    double* dAP = dP + (size_t)k * L;
    double* dots = c->one.as<double>();  // [k] scalars of one step
    double* a_d = dots + k;
    double* b_d = a_d + k;
    struct Coul {  // the no-LC Z-vector operator has Coulomb factor 4
      tda_ctx* c;
      double keep;
      ~Coul() { c->coul_factor = keep; }
    } coul{c, c->coul_factor};
    if (!c->has_lc) c->coul_factor = 4.0;
    TDA_HIP(hipMemcpyAsync(c->stage.p, rhs, (size_t)ov * k * 8, hipMemcpyHostToDevice, c->stream));
    launch_pack_x(c->stage.as<double>(), c->o, c->v, c->vp, k, dR, c->stream);
    // z0 = rhs / de (:1154-1155); r0 = rhs - A z0; p0 = r0 (:1168-1186)
    launch_cg_init(dR, c->de.as<double>(), k, L, dZ, c->stream);
    products(c, dZ, k, dAP, PREC_FP64);
    launch_fill(a_d, 1.0, k, c->stream);
    launch_residual(dR, dAP, a_d, k, L, c->stream);
    launch_copy(dR, dP, (int64_t)k * L, c->stream);
    std::vector<double> rr(k), pap(k), rr_new(k), a(k), b(k);
    std::vector<int32_t> its(k, max_iter);
    std::vector<bool> active(k, true);
    launch_dots(dR, dR, k, L, c->npart.as<double>(), dots, c->stream);
    TDA_HIP(hipMemcpyAsync(rr.data(), dots, k * 8, hipMemcpyDeviceToHost, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
    int left = k;
    for (int it = 0; it < max_iter && left > 0; ++it) {
      products(c, dP, k, dAP, PREC_FP64);  // :1189-1203
      launch_dots(dP, dAP, k, L, c->npart.as<double>(), dots, c->stream);
      TDA_HIP(hipMemcpyAsync(pap.data(), dots, k * 8, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
      for (int j = 0; j < k; ++j) a[j] = active[j] ? rr[j] / pap[j] : 0.0;  // :1206-1210
      TDA_HIP(hipMemcpyAsync(a_d, a.data(), k * 8, hipMemcpyHostToDevice, c->stream));
      launch_cg_step(dZ, dR, dP, dAP, a_d, k, L, c->stream);
      launch_dots(dR, dR, k, L, c->npart.as<double>(), dots, c->stream);
      TDA_HIP(hipMemcpyAsync(rr_new.data(), dots, k * 8, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
      for (int j = 0; j < k; ++j) {
        b[j] = -1.0;
        if (!active[j]) continue;
        if (rr_new[j] <= conv) {  // :1212-1216
          active[j] = false;
          its[j] = it + 1;
          --left;
          continue;
        }
        b[j] = rr_new[j] / rr[j];  // :1217
        rr[j] = rr_new[j];
      }
      TDA_HIP(hipMemcpyAsync(b_d, b.data(), k * 8, hipMemcpyHostToDevice, c->stream));
      launch_cg_dir(dP, dR, b_d, k, L, c->stream);
    }
    launch_unpack_x(dZ, c->o, c->v, c->vp, k, c->stage.as<double>(), c->stream);
    TDA_HIP(hipMemcpyAsync(z, c->stage.p, (size_t)ov * k * 8, hipMemcpyDeviceToHost, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
    c->harvest_events();
    if (n_iter) std::memcpy(n_iter, its.data(), (size_t)k * sizeof(int32_t));
  });
}

int tda_precondition(tda_ctx* c, const double* r, double w, double* out) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(c->has_system, TDA_ERR_STATE, "no system loaded");
    TDA_CHECK(r && out, TDA_ERR_BAD_ARG, "NULL r/out");
    prepare(c);
    const int64_t ov = (int64_t)c->o * c->v;
    c->stage.ensure((size_t)ov * 8);
    c->ax_tmp.ensure((size_t)c->ovp * 8 * 2);
    double* dR = c->ax_tmp.as<double>();
    double* dP = dR + c->ovp;
    TDA_HIP(hipMemcpyAsync(c->stage.p, r, (size_t)ov * 8, hipMemcpyHostToDevice, c->stream));
    launch_pack_x(c->stage.as<double>(), c->o, c->v, c->vp, 1, dR, c->stream);
    launch_precondition(dR, w, c->de.as<double>(), c->ovp, c->pre_thr, dP, c->stream);
    launch_unpack_x(dP, c->o, c->v, c->vp, 1, c->stage.as<double>(), c->stream);
    TDA_HIP(hipMemcpyAsync(out, c->stage.p, (size_t)ov * 8, hipMemcpyDeviceToHost, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
  });
}

namespace {
// CI vectors for the adjoint: host X [ov, n] packed to the device, or the
// last tda_run's vectors
const double* adjoint_vectors(tda_ctx* c, int n, const double* X) {
  TDA_CHECK(c->has_system, TDA_ERR_STATE, "no system loaded");
  TDA_CHECK(!c->ct && !c->dense, TDA_ERR_STATE, "the adjoint is defined for a System/Monomer");
  TDA_CHECK(n >= 1, TDA_ERR_BAD_ARG, "n_states must be positive");
  prepare(c);
  if (!X) {
    TDA_CHECK(n <= c->last_roots, TDA_ERR_STATE, "no CI vectors: run tda_run first or pass X");
    return c->RZ.as<double>();
  }
  const int64_t ov = (int64_t)c->o * c->v;
  c->stage.ensure((size_t)ov * n * 8);
  c->adjX.ensure((size_t)c->ovp * n * 8);
  TDA_HIP(hipMemcpyAsync(c->stage.p, X, (size_t)ov * n * 8, hipMemcpyHostToDevice, c->stream));
  launch_pack_x(c->stage.as<double>(), c->o, c->v, c->vp, n, c->adjX.as<double>(), c->stream);
  return c->adjX.as<double>();
}
}  // namespace

int tda_adjoint(tda_ctx* c, int n_states, const double* X, double* d_gamma, double* d_gamma_lr, double* c_qov,
                double* d_qoo, double* d_qvv) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    const double* Xd = adjoint_vectors(c, n_states, X);
    TDA_CHECK(c->has_lc || (!d_gamma_lr && !d_qoo && !d_qvv), TDA_ERR_STATE,
              "gamma_lr / q_oo / q_vv derivatives need gamma_lr");
    const int N = c->N, n = n_states;
    const size_t nn = (size_t)N * N * n;
    c->adjD.ensure((nn * 2 + (size_t)N * n) * 8);
    double* dg = c->adjD.as<double>();
    double* dglr = dg + nn;
    double* cq = dglr + nn;
    hf_derivatives(c, Xd, n, dg, d_gamma_lr ? dglr : nullptr, cq, d_qoo, d_qvv);
    if (d_gamma) TDA_HIP(hipMemcpyAsync(d_gamma, dg, nn * 8, hipMemcpyDeviceToHost, c->stream));
    if (d_gamma_lr) TDA_HIP(hipMemcpyAsync(d_gamma_lr, dglr, nn * 8, hipMemcpyDeviceToHost, c->stream));
    if (c_qov) TDA_HIP(hipMemcpyAsync(c_qov, cq, (size_t)N * n * 8, hipMemcpyDeviceToHost, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
  });
}

int tda_adjoint_nuclear(tda_ctx* c, int n_states, const double* X, const double* dgamma_dr,
                        const double* dgamma_lr_dr, double* grad) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(dgamma_dr && grad, TDA_ERR_BAD_ARG, "NULL dgamma_dr / grad");
    const double* Xd = adjoint_vectors(c, n_states, X);
    const bool lr = c->has_lc && dgamma_lr_dr;
    const int N = c->N, n = n_states;
    const int64_t nsq = (int64_t)N * N, n3 = 3 * (int64_t)N;
    const size_t nn = (size_t)nsq * n;
    c->adjD.ensure((nn * 3 + (size_t)N * n + (size_t)n3 * n) * 8);
    double* dg = c->adjD.as<double>();
    double* dglr = dg + nn;
    double* sym = dglr + nn;
    double* cq = sym + nn;
    double* g = cq + (size_t)N * n;
    hf_derivatives(c, Xd, n, dg, lr ? dglr : nullptr, cq, nullptr, nullptr);
    // The gradient array holds row i of dgamma/dR_i only (gamma_approximation.rs:
    // 263-279: entry [3i + d][i][j]); gamma_ji moves with R_i as well, so
    // grad[k][c] = sum_AB g1[c][A][B] (dw_k/dgamma_AB + dw_k/dgamma_BA)
    // (+ the gamma_lr term).
    c->adjQ.ensure((size_t)n3 * nsq * 8);
    const double* srcs[2] = {dgamma_dr, lr ? dgamma_lr_dr : nullptr};
    const double* dws[2] = {dg, dglr};
    for (int t = 0; t < 2; ++t) {
      if (!srcs[t]) continue;
      launch_sym_add(dws[t], n, N, sym, c->stream);
      TDA_HIP(hipMemcpyAsync(c->adjQ.p, srcs[t], (size_t)n3 * nsq * 8, hipMemcpyHostToDevice, c->stream));
      GemmArgs gg = gargs(sym, c->adjQ.p, g, n, n3, nsq, plain(nsq), plain(1), plain(1), plain(nsq), plain(n3),
                          plain(1));
      gg.beta = t == 0 ? 0.0 : 1.0;
      run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, gg);
      TDA_HIP(hipStreamSynchronize(c->stream));  // adjQ is reloaded next
    }
    TDA_HIP(hipMemcpyAsync(grad, g, (size_t)n3 * n * 8, hipMemcpyDeviceToHost, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
  });
}

// Excited-state nuclear gradient of one state (tda_gradient_lc,
// src/gradients/excited_state.rs:259-487; tda_gradient_nolc :56-257); the
// work is in gradient.hip.
int tda_gradient(tda_ctx* c, int form, double omega, const double* x, const double* dp, const double* ds_rows,
                 const double* dh_rows, const double* dgamma, const double* dgamma_lr, int max_iter, double conv,
                 double* grad, double* z, int32_t* z_iter) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(c->has_system && !c->dense && !c->ct, TDA_ERR_STATE, "tda_gradient: no System/Monomer loaded");
    TDA_CHECK(form == TDA_GRAD_REFERENCE || form == TDA_GRAD_EXACT, TDA_ERR_BAD_ARG, "tda_gradient: unknown form");
    TDA_CHECK(x && dp && ds_rows && dh_rows && dgamma && grad, TDA_ERR_BAD_ARG, "tda_gradient: NULL input / output");
    TDA_CHECK(!c->has_lc || dgamma_lr, TDA_ERR_BAD_ARG, "tda_gradient: dgamma_lr needed with gamma_lr");
    TDA_CHECK(max_iter >= 0 && conv >= 0.0 && std::isfinite(omega), TDA_ERR_BAD_ARG, "tda_gradient: bad max_iter / conv / omega");
    TDA_CHECK(c->ob == 0 && c->vb == c->o && c->o + c->v == c->nao, TDA_ERR_BAD_ARG,
              "tda_gradient: the occupied and virtual ranges must cover all MOs (occupied first)");
    prepare(c);
    const int N = c->N, o = c->o, op = c->op, vp = c->vp, n = c->nao;
    // device memory of the call (gradient.hip): T / U exchange intermediates,
    // the padded q_ov, about 30 n x n matrices, the derivative inputs
    const int64_t tmax = std::max({(int64_t)o * vp, (int64_t)op * op, (int64_t)vp * vp, (int64_t)vp * op});
    const double need = 8.0 * (2.0 * N * tmax * (c->has_lc ? 1 : 0) + (double)N * op * vp + 40.0 * n * (double)n +
                               6.0 * (double)N * N);
    TDA_CHECK(need <= (double)c->budget, TDA_ERR_OOM,
              "tda_gradient: " + std::to_string((int64_t)(need / 1e6)) + " MB of work buffers exceed the HBM budget");
    if (!c->qov_built) {
      c->qov.ensure((size_t)N * c->ovp * 8);
      build_charges(c, c->C.as<double>(), c->SC.as<double>(), c->aooff.as<int>(), N, c->nao, c->ob, c->o, c->vb,
                    c->v, c->op, c->vp, c->qov.as<double>(), nullptr, nullptr);
      c->qov_built = true;
    }
    build_mo_charges(c);
    GradView gv;
    gv.N = N; gv.n = n; gv.o = o; gv.v = c->v; gv.op = op; gv.vp = vp; gv.ob = c->ob; gv.vb = c->vb;
    gv.qov = c->qov.as<double>();
    gv.qoo = c->qoo.as<double>();
    gv.qvv = c->qvv.as<double>();
    gv.gamma = c->gamma.as<double>();
    gv.gamma_lr = c->has_lc ? c->gamma_lr.as<double>() : nullptr;
    gv.C = c->C.as<double>();
    gv.S = c->S.as<double>();
    gv.orbe = c->orbe.as<double>();
    gv.de = c->de.as<double>();
    gv.off = c->aooff.as<int>();
    gv.off_h.resize(N + 1);
    TDA_HIP(hipMemcpyAsync(gv.off_h.data(), gv.off, (size_t)(N + 1) * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
    gv.stream = c->stream;
    gv.cache = &c->grad_cache;
    GradIn in{form == TDA_GRAD_EXACT ? 1 : 0, omega, x, dp, ds_rows, dh_rows, dgamma, c->has_lc ? dgamma_lr : nullptr,
              max_iter, conv};
    GradOut out{grad, z, 0};
    gradient_run(gv, [&](const double* X, double* AX) { products(c, X, 1, AX, PREC_FP64); }, in, out);
    c->harvest_events();
    if (z_iter) *z_iter = out.z_iter;
  });
}

int tda_run(tda_ctx* c, int n_roots, int max_iter, double tol, double* omega, double* X, double* q_trans,
            double* tr_dipoles, double* f, tda_stats* stats) {
  if (!c) return TDA_ERR_BAD_ARG;
  const auto t0 = std::chrono::steady_clock::now();
  DavResult res, res64;
  int rc = guarded(c, [&] {
    TDA_CHECK(c->has_system, TDA_ERR_STATE, "tda_run: no system loaded");
    const int64_t ov = (int64_t)c->o * c->v;
    TDA_CHECK(n_roots >= 1 && n_roots <= ov, TDA_ERR_BAD_ARG, "n_roots must be in 1..n_occ*n_virt");
    TDA_CHECK(max_iter >= 1 && tol > 0.0, TDA_ERR_BAD_ARG, "bad max_iter / tolerance");
    TDA_CHECK(!c->dense || (!q_trans && !tr_dipoles && !f), TDA_ERR_BAD_ARG,
              "a dense engine has no q_trans / tr_dipoles / f");
    c->reset_counters();
    c->last_roots = 0;  // RZ is about to be overwritten: no stale CI vectors for the adjoint
    const auto tq0 = std::chrono::steady_clock::now();
    prepare(c);
    const double tq = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tq0).count();
    const int n = n_roots;
    const int64_t L = c->ovp;
    // the reference FP64 path: FP64 contexts, dense engines, and systems whose
    // products are too cheap for a precision ladder to pay off (the subspace
    // work of the extra iterations costs more than the products save)
    const bool fp64_only = c->prec == PREC_FP64 || c->dense || product_flops(c) < c->ladder_min_flops;
    alloc_davidson(c, n, fp64_only ? n : c->lp_keep(n));
    // initial subspace: unit vectors at the n smallest omega (solvers/utils.rs:17-26)
    std::vector<int64_t> order(ov);
    std::iota(order.begin(), order.end(), 0);
    std::partial_sort(order.begin(), order.begin() + n, order.end(), [&](int64_t a, int64_t b) {
      return c->h_de[a] < c->h_de[b] || (c->h_de[a] == c->h_de[b] && a < b);
    });
    {
      std::vector<int64_t> pos(n);
      for (int r = 0; r < n; ++r) pos[r] = (order[r] / c->v) * c->vp + (order[r] % c->v);
      c->guess.ensure((size_t)n * sizeof(int64_t));
      TDA_HIP(hipMemcpyAsync(c->guess.p, pos.data(), (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
      fill_guess(c, c->V.as<double>(), n);
      TDA_HIP(hipStreamSynchronize(c->stream));  // pos goes out of scope
    }
    std::vector<double> h_evals, h_err;
    int fp32_iters = 0, p32 = PREC_FP32;
    if (fp64_only) {
      // the reference rules, plus two guards that change nothing where one
      // classical Gram-Schmidt pass keeps the basis orthonormal: a second
      // pass for a block once the tracked loss of orthogonality of the basis
      // would exceed eta_tol (append_block: preconditioned residuals lying
      // largely in the basis compound it, and the Rayleigh-Ritz step of a
      // nearly dependent basis returns spurious Ritz pairs near 0 whose
      // residuals pass the test), and converged roots must be unit vectors
      res = davidson(c, n, n, max_iter, tol, PREC_FP64, false, 0, h_evals, h_err, 0, 0, nullptr, 0.5, 1e300, false,
                     0.0, true, true);
      res64 = res;
    } else {
      // Precision ladder.  Each low-precision phase measures the error of
      // its products after its first iteration and runs until the residuals
      // reach floor_k times that error (or stop improving); FP16/BF16 first,
      // then FP32 (whose floor lies far below `tol`, so it normally converges),
      // then FP64 products of the n Ritz vectors verify (and, if needed,
      // continue).  Each phase restarts from the previous phase's Ritz vectors.
      // The FP16 / BF16 phase uses a thick restart (basis cap lp_cap, keep
      // lp_keep Ritz vectors) and hands its lp_keep lowest Ritz vectors to the
      // FP32 phase, which follows the reference rule.
      // A phase that met a non-finite product (16-bit overflow) hands V[0, n)
      // -- finite, orthonormal -- to the next one, and the 16-bit types are
      // not used again in this solve (the FP16 hi / lo split has the range of
      // FP16): FP16 / BF16 -> FP32 MFMA products -> FP64.
      int used = 0;
      bool done = false, lp_overflow = false;
      int start = n, kk = n;
      if ((c->prec == PREC_FP16 || c->prec == PREC_BF16) && c->lp_phase) {
        const bool v32 = basis32_for(c) && c->has_lc && c->op_kind == 0 && !c->ct;
        res = davidson(c, n, n, max_iter, tol, c->prec, false, lp_stall_window(c), h_evals, h_err,
                       c->lp_cap(n, lp_cap_mult_for(c)),
                       c->lp_keep(n), &kk, 0.8, 1e300, true, c->floor_k, true, false, v32);
        used += res.iters;
        done = res.converged;
        lp_overflow = res.nonfinite;
        if (res.nonfinite) kk = n;  // V[0, n) already holds the restart
        else if (kk > n) launch_copy(c->Wb(), c->V.as<double>(), (int64_t)kk * L, c->stream);
        else launch_copy(c->RZ.as<double>(), c->V.as<double>(), (int64_t)n * L, c->stream);
        start = kk;
      }
      // The FP32 phase keeps the reference rule; if it stops unconverged it
      // exports its lp_keep lowest Ritz vectors so that the FP64 phase does
      // not restart from a thin basis.
      int start64 = n;
      const double* src64 = c->RZ.as<double>();
      if (!done) {
        int kk32 = n;
        // the FP32-accuracy phase: FP16 hi / lo split products when the
        // context runs FP16 / BF16 on the AO operator, else FP32
        const bool ao_op = c->has_lc && c->op_kind == 0 && !c->ct;
        p32 = (c->split32 && ao_op && !lp_overflow && (c->prec == PREC_FP16 || c->prec == PREC_BF16)) ? PREC_F16X
                                                                                                        : PREC_FP32;
        DavResult r32 = davidson(c, n, start, std::max(1, max_iter - used), tol, p32, used > 0,
                                 3 * lp_stall_window(c), h_evals, h_err, c->fp32_cap(n), c->lp_keep(n), &kk32, 0.8, 1e300,
                                 c->fp32_thick, c->floor_k, true);
        used += r32.iters;
        fp32_iters = r32.iters;
        res.iters = used;
        if (r32.nonfinite) src64 = c->V.as<double>();  // its start vectors, V[0, n)
        else if (!r32.converged && kk32 > n) start64 = kk32, src64 = c->Wb();
      }
      // FP64 verification / continuation from the low-precision Ritz vectors,
      // orthonormalised in FP64 first; converged roots must be unit vectors
      if (src64 == c->V.as<double>()) {
        launch_copy(src64, c->Pb(), (int64_t)n * L, c->stream);
        src64 = c->Pb();
      }
      start64 = orthonormal_start(c, src64, start64, n);
      res64 = davidson(c, n, start64, std::max(1, max_iter - used), tol, PREC_FP64, true, 0, h_evals, h_err, 0, 0,
                       nullptr, 0.5, 1e300, false, 0.0, true, true);
      res.iters = used;
    }
    // outputs (tda/mod.rs:88-126)
    const int N = c->N;
    c->last_roots = n;
    if (omega) std::memcpy(omega, h_evals.data(), (size_t)n * 8);
    if (X) {
      c->stage.ensure((size_t)ov * n * 8);
      launch_unpack_x(c->RZ.as<double>(), c->o, c->v, c->vp, n, c->stage.as<double>(), c->stream);
      TDA_HIP(hipMemcpyAsync(X, c->stage.p, (size_t)ov * n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    if (c->dense) TDA_HIP(hipStreamSynchronize(c->stream));
    if (!c->dense) {
      // q_trans = q_ov X (tda/mod.rs:91), mu = xyz^T q_trans, f = 2/3 w |mu|^2 (moments.rs:7-24)
      c->u.ensure((size_t)N * n * 8);
      if (c->qov_built) {
        GemmArgs g = gargs(c->qov.p, c->RZ.p, c->u.p, N, n, L, plain(L), plain(1), plain(1), plain(L), plain(n),
                           plain(1));
        run_gemm(c, DT_F64, DT_F64, EPI_STORE, true, true, g);
      } else {
        ao_trans_charges(c, c->RZ.as<double>(), n, c->u.as<double>(), n);  // AO operator: no q_ov
      }
      c->dxyz.ensure((size_t)N * 3 * 8);
      c->post.ensure((size_t)5 * n * 8);
      double* d_w = c->post.as<double>();
      double* d_mu = d_w + n;
      double* d_f = d_mu + 3 * n;
      TDA_HIP(hipMemcpyAsync(c->dxyz.p, c->xyz.data(), (size_t)N * 3 * 8, hipMemcpyHostToDevice, c->stream));
      TDA_HIP(hipMemcpyAsync(d_w, h_evals.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
      launch_dipoles(c->dxyz.as<double>(), c->u.as<double>(), d_w, N, n, d_mu, d_f, c->stream);
      if (q_trans) TDA_HIP(hipMemcpyAsync(q_trans, c->u.p, (size_t)N * n * 8, hipMemcpyDeviceToHost, c->stream));
      if (tr_dipoles) TDA_HIP(hipMemcpyAsync(tr_dipoles, d_mu, (size_t)3 * n * 8, hipMemcpyDeviceToHost, c->stream));
      if (f) TDA_HIP(hipMemcpyAsync(f, d_f, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
      TDA_HIP(hipStreamSynchronize(c->stream));
    }
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->n_iter = res.iters + (fp64_only ? 0 : res64.iters);
      stats->n_iter_fp64 = fp64_only ? res.iters : res64.iters;
      stats->n_products = c->n_products;
      stats->n_products_fp64 = c->n_products_fp64;
      stats->n_iter_fp32 = fp32_iters;
      stats->n_products_fp32 = c->n_products_fp32;
      stats->split_fp32 = (!fp64_only && fp32_iters > 0 && p32 == PREC_F16X) ? 1 : 0;
      stats->converged = res64.converged ? 1 : 0;
      stats->n_atom_blocks = c->n_blocks;
      stats->max_residual = res64.max_err;
      stats->t_qbuild_ms = tq;
      stats->t_products_ms = c->t_products_ms;
      for (int st = 0; st < 2; ++st)
        for (int p = 0; p < 3; ++p) {
          stats->exch_ms[st][p] = c->st_ms[st][p];
          stats->exch_launches[st][p] = c->st_n[st][p];
          stats->exch_flops[st][p] = c->st_flops[st][p];
        }
    }
    if (!res64.converged) throw Error{TDA_ERR_NOT_CONVERGED, "Davidson Algorithm did not converge!"};
  });
  if (stats)
    stats->t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// library-internal (internal.h): the prepared charges-form operator of a
// context, for the grouped batch driver
// ---------------------------------------------------------------------------
namespace tda {

int ctx_charges_dims(tda_ctx* c, ChargesDims* d) {
  if (!c || !d) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(c->has_system && !c->dense, TDA_ERR_STATE, "no system loaded");
    TDA_CHECK(c->ct || c->op_kind == 1 || !c->has_lc, TDA_ERR_STATE, "charges form expected");
    d->N = c->N;
    d->NL = c->has_lc ? c->Nx : 0;
    d->o = c->o;
    d->v = c->v;
    d->op = c->op;
    d->vp = c->vp;
    d->lc = c->has_lc;
  });
}

int ctx_export_charges(tda_ctx* c, double* qov, double* gamma, double* G, double* qvv, double* de) {
  if (!c) return TDA_ERR_BAD_ARG;
  return guarded(c, [&] {
    TDA_CHECK(c->ct || c->op_kind == 1 || !c->has_lc, TDA_ERR_STATE, "charges form expected");
    prepare(c);
    TDA_CHECK(c->qov_built && (!c->has_lc || c->mo_built), TDA_ERR_STATE, "operator not prepared");
    const size_t N = c->N;
    if (qov) TDA_HIP(hipMemcpyAsync(qov, c->qov.p, N * c->ovp * 8, hipMemcpyDeviceToDevice, c->stream));
    if (gamma) TDA_HIP(hipMemcpyAsync(gamma, c->gamma.p, N * N * 8, hipMemcpyDeviceToDevice, c->stream));
    if (c->has_lc && G)
      TDA_HIP(hipMemcpyAsync(G, c->G.p, (size_t)c->Nx * c->op * c->op * 8, hipMemcpyDeviceToDevice, c->stream));
    if (c->has_lc && qvv)
      TDA_HIP(hipMemcpyAsync(qvv, c->qvv.p, (size_t)c->Nx * c->vp * c->vp * 8, hipMemcpyDeviceToDevice, c->stream));
    if (de) TDA_HIP(hipMemcpyAsync(de, c->de.p, (size_t)c->ovp * 8, hipMemcpyDeviceToDevice, c->stream));
    TDA_HIP(hipStreamSynchronize(c->stream));
  });
}

const char* ctx_error(tda_ctx* c) { return c ? c->err.c_str() : "null context"; }

}  // namespace tda
