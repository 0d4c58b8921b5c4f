// Nuclear gradient of one TDA excitation energy: tda_gradient_lc
// (src/gradients/excited_state.rs:259-487) and tda_gradient_nolc (:56-257),
// with the ground-state pieces they read (h0_and_s_gradients, f_v / f_lr of
// P - P0, src/gradients/ground_state.rs:12-131), on the GPU in FP64.
//
// The reference builds [3N, n, n] arrays (dH0/dR, dS/dR, dgamma_AO/dR, f_v,
// f_lr) and contracts each with an AO matrix.  Every one of those terms is
// linear in a single nuclear derivative (of H0, S, gamma or gamma_lr), so the
// contraction is moved in front: the gradient is
//   g[3A+k] = sum_{mu in A, nu} (dH0_k[mu][nu] (Hm + Hm^T)[mu][nu]
//                               + dS_k[mu][nu] (Sm + Sm^T)[mu][nu])
//           + sum_B (dg_k[A][B] Gat[A][B] + dglr_k[A][B] Lat[A][B])
// with dH0_k / dS_k the row form (derivative w.r.t. the atom of the row AO)
// and dg_k[A][B] = dgamma_AB / dR_A,k.  Hm, Sm [n, n] and Gat, Lat [N, N]
// (atom-block sums) are built once from the state's MO-space response:
//   Hm = D_h = C [[-t_ij, z], [0, t_ab]] C^T
//   Sm = -W_AO + 1/4 Es_v(dP, D_h) + 1/2 Es_v(x, x) - 1/8 Es_lr(dP, D_h) - 1/4 Es_lr(x^T, x)
// where f_v(v) : W = 1/4 (dS : Es_v + dg : Eg_v) and f_lr(v) : W = 1/4 (dS : Es_lr
// + dglr : Eg_lr) are the adjoints of helpers.rs:47-79 and :130-186:
//   Es_v = W o (gsv 1^T + 1 gsv^T) + (v + v^T) o (1 h^T),  Eg_v = sw sv^T,
//   sv = rowsum(S o (v + v^T)), sw = rowsum(S o (W + W^T)), gsv = g_AO sv, h = g_AO sw;
//   Es_lr = (g o W) S v + (g o WS) v + W (Sv o g) + WS (g o v) + vS (g o W)
//         + (vS o g) W + v (g o SW) + (g o v) SW,
//   Eg_lr = (S v^T S) o W + (S v^T) o (WS) + (v^T S) o (SW) + v^T o (SWS)
// (g = gamma_lr expanded to AO pairs).  The n^3 work is about thirty AO GEMMs;
// the MO part (Hav contractions of helpers.rs:621-843, the Z-vector) runs on
// the prepared transition charges.
//
// Z-vector (tda_zvector_lc / _no_lc, helpers.rs:1130-1223, :1053-1128):
// conjugate gradients from z = r / de, stop at r.r <= conv.  Operator: the
// forward A (de + 2 J - K) for the reference LC form; de + 4 J - K - K' (the
// ground-state orbital Hessian A + B, K'_ia,jb = (ib|ja)_lr) for the exact
// form and without LC (where it is mult_av_nolc's de + 4 J).  The exact form
// also carries both exchange patterns in W_ij's H+[z] (e1 + e1^T instead of
// the reference's 2 e1).
#include <algorithm>
#include <cstring>
#include <functional>
#include <vector>

#include "internal.h"
#include "kernels.h"

namespace tda {
namespace {

constexpr int TPB = 256;
inline unsigned grid1(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, TPB), 16384));
}

// Zeroed device buffers of one call, taken in call order from a cache that
// outlives the call (the context's): repeated gradients of one system
// allocate nothing.  A slot too small for its request is reallocated.
struct Pool {
  hipStream_t s;
  std::vector<std::pair<void*, size_t>>* cache;
  size_t next = 0;
  Pool(hipStream_t st, std::vector<std::pair<void*, size_t>>* c) : s(st), cache(c) {}
  double* get(size_t n) {
    const size_t bytes = std::max<size_t>(n, 1) * 8;
    if (next == cache->size()) cache->push_back({nullptr, 0});
    std::pair<void*, size_t>& e = (*cache)[next++];
    if (e.second < bytes) {
      if (e.first) {
        TDA_HIP(hipStreamSynchronize(s));
        (void)hipFree(e.first);
        e = {nullptr, 0};
      }
      void* p = nullptr;
      if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        throw Error{TDA_ERR_OOM, "gradient: hipMalloc of " + std::to_string(bytes) + " bytes failed"};
      }
      e = {p, bytes};
    }
    TDA_HIP(hipMemsetAsync(e.first, 0, bytes, s));
    return (double*)e.first;
  }
  int* get_int(size_t n) { return (int*)get((n + 1) / 2); }
};

// FP64 GEMM on the generic MFMA kernel: C[cm(m)+cn(n)] = alpha sum_k A B + beta C,
// batch z offsets A + z sa, B + z sb, C + z sc.  An operand is read K-contiguous
// when its K map has unit stride, else along M (A) / N (B).
struct Mm {
  hipStream_t s;
  Pool& pool;
  double* ws = nullptr;
  size_t wsb = 0;
  void operator()(int64_t M, int64_t N, int64_t K, const double* A, IdxMap am, IdxMap ak, const double* B, IdxMap bk,
                  IdxMap bn, double* C, IdxMap cm, IdxMap cn, double alpha = 1.0, double beta = 0.0, int batch = 1,
                  int64_t sa = 0, int64_t sb = 0, int64_t sc = 0) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    TDA_CHECK(K > 0, TDA_ERR_BAD_ARG, "gradient: empty contraction");
    GemmArgs g;
    std::memset(&g, 0, sizeof(g));
    g.A = A; g.B = B; g.C = C;
    g.M = M; g.N = N; g.K = K;
    g.am = am; g.ak = ak; g.bk = bk; g.bn = bn; g.cm = cm; g.cn = cn;
    g.batch = batch;
    g.a_batch = sa; g.b_batch = sb; g.c_batch = sc;
    g.ksplit = 1;
    g.alpha = alpha;
    g.beta = beta;
    g.dm = plain(0); g.dmap = plain(1); g.xm = plain(0); g.xn = plain(0);
    g.hm = plain(0); g.hn = plain(0);
    g.hmod = 1;
    const bool a_kc = ak.s0 == 1, b_kc = bk.s0 == 1;
    const GemmPlan p = gemm_plan(DT_F64, M, N, K, batch, a_kc, b_kc);
    const size_t need = gemm_workspace_bytes(DT_F64, p, M, N, batch);
    if (need > wsb) {
      ws = pool.get(need / 8 + 1);
      wsb = need;
    }
    g.ws = ws;
    gemm_launch(DT_F64, DT_F64, EPI_STORE, a_kc, b_kc, g, p, s);
  }
};

// ---------------------------------------------------------------------------
// element-wise kernels
// ---------------------------------------------------------------------------
// out = num / den; then (optional) dst = src
__global__ void cg_ratio_kernel(const double* num, const double* den, double* out, double* dst, const double* src) {
  if (threadIdx.x == 0) {
    const double q = num[0] / den[0];
    if (dst) dst[0] = src[0];
    out[0] = q;
  }
}

// ex[i][a] = e_o[i] x[i][a], xe[i][a] = x[i][a] e_v[a]   (a < v; pads stay 0)
__global__ void scale_x_kernel(const double* __restrict__ x, const double* __restrict__ orbe, int ob, int vb, int o,
                               int v, int vp, double* __restrict__ ex, double* __restrict__ xe) {
  const int64_t n = (int64_t)o * v;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < n; t += (int64_t)gridDim.x * TPB) {
    const int i = (int)(t / v), a = (int)(t % v);
    const int64_t e = (int64_t)i * vp + a;
    ex[e] = orbe[ob + i] * x[e];
    xe[e] = x[e] * orbe[vb + a];
  }
}

// y = a x + b y
__global__ void axpby_kernel(const double* __restrict__ x, double* __restrict__ y, int64_t n, double a, double b) {
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < n; t += (int64_t)gridDim.x * TPB)
    y[t] = a * x[t] + (b != 0.0 ? b * y[t] : 0.0);
}

// out [N][op][vp] from q [N][o][vp] (rows o..op-1 left zero)
__global__ void pad_rows_kernel(const double* __restrict__ q, int N, int o, int op, int vp, double* __restrict__ out) {
  const int64_t n = (int64_t)N * o * vp;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < n; t += (int64_t)gridDim.x * TPB) {
    const int64_t A = t / ((int64_t)o * vp), r = t % ((int64_t)o * vp);
    out[A * op * vp + r] = q[t];
  }
}

// MO matrix [n][n] (occupied block first) from its blocks:
//   oo = s_oo * Boo[p][q] (ld op), ov = s_ov * Bov[p][q-o] (ld vp),
//   vv = s_vv * Bvv[p-o][q-o] (ld vp), vo = 0;
// tri: only q >= p in the oo / vv blocks, their diagonals times dsc.
__global__ void mo_assemble_kernel(const double* __restrict__ Boo, double s_oo, const double* __restrict__ Bov,
                                   double s_ov, const double* __restrict__ Bvv, double s_vv, int o, int n, int op,
                                   int vp, int tri, double dsc, double* __restrict__ M) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < nn; t += (int64_t)gridDim.x * TPB) {
    const int p = (int)(t / n), q = (int)(t % n);
    double val = 0.0;
    if (p < o && q < o) {
      if (Boo && (!tri || q >= p)) val = s_oo * Boo[(int64_t)p * op + q] * (tri && p == q ? dsc : 1.0);
    } else if (p < o) {
      if (Bov) val = s_ov * Bov[(int64_t)p * vp + (q - o)];
    } else if (q >= o) {
      if (Bvv && (!tri || q >= p)) val = s_vv * Bvv[(int64_t)(p - o) * vp + (q - o)] * (tri && p == q ? dsc : 1.0);
    }
    M[t] = val;
  }
}

// out[mu] = sum_nu S[mu][nu] (W[mu][nu] + W[nu][mu]); one wave per row
__global__ void rowsum_sym_kernel(const double* __restrict__ S, const double* __restrict__ W, int n,
                                  double* __restrict__ out) {
  const int row = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double acc = 0.0;
  for (int c = lane; c < n; c += 64) acc += S[(int64_t)row * n + c] * (W[(int64_t)row * n + c] + W[(int64_t)c * n + row]);
  acc = wave_sum(acc);
  if (lane == 0) out[row] = acc;
}

// at[A] = sum_{mu in A} vec[mu]
__global__ void atom_sum_kernel(const double* __restrict__ vec, const int* __restrict__ off, int N,
                                double* __restrict__ at) {
  const int A = blockIdx.x * TPB + threadIdx.x;
  if (A >= N) return;
  double s = 0.0;
  for (int mu = off[A]; mu < off[A + 1]; ++mu) s += vec[mu];
  at[A] = s;
}

// out[mu] = sum_B gamma[atom(mu)][B] at[B]
__global__ void gamma_expand_kernel(const double* __restrict__ gamma, const int* __restrict__ atom, int N, int n,
                                    const double* __restrict__ at, double* __restrict__ out) {
  const int mu = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (mu >= n) return;
  const double* g = gamma + (int64_t)atom[mu] * N;
  double acc = 0.0;
  for (int B = lane; B < N; B += 64) acc += g[B] * at[B];
  acc = wave_sum(acc);
  if (lane == 0) out[mu] = acc;
}

// Es[mu][nu] += alpha (W[mu][nu] (gsv[mu] + gsv[nu]) + (v[mu][nu] + v[nu][mu]) h[nu])
__global__ void es_v_kernel(const double* __restrict__ W, const double* __restrict__ v, const double* __restrict__ gsv,
                            const double* __restrict__ h, double alpha, int n, double* __restrict__ Es) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < nn; t += (int64_t)gridDim.x * TPB) {
    const int mu = (int)(t / n), nu = (int)(t % n);
    Es[t] += alpha * (W[t] * (gsv[mu] + gsv[nu]) + (v[t] + v[(int64_t)nu * n + mu]) * h[nu]);
  }
}

// Gat[A][B] += alpha (sw[A] sv[B] + sw[B] sv[A])
__global__ void gat_kernel(const double* __restrict__ sw, const double* __restrict__ sv, double alpha, int N,
                           double* __restrict__ Gat) {
  const int64_t nn = (int64_t)N * N;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < nn; t += (int64_t)gridDim.x * TPB) {
    const int A = (int)(t / N), B = (int)(t % N);
    Gat[t] += alpha * (sw[A] * sv[B] + sw[B] * sv[A]);
  }
}

// out[mu][nu] = g[atom(mu)][atom(nu)] * M[mu][nu]  (M^T when trans)
__global__ void had_g_kernel(const double* __restrict__ M, int trans, const double* __restrict__ g,
                             const int* __restrict__ atom, int N, int n, double* __restrict__ out) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < nn; t += (int64_t)gridDim.x * TPB) {
    const int mu = (int)(t / n), nu = (int)(t % n);
    const double m = trans ? M[(int64_t)nu * n + mu] : M[t];
    out[t] = g[(int64_t)atom[mu] * N + atom[nu]] * m;
  }
}

// out[mu][nu] = M[nu][mu]
__global__ void transpose_kernel(const double* __restrict__ M, int n, double* __restrict__ out) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < nn; t += (int64_t)gridDim.x * TPB) {
    const int mu = (int)(t / n), nu = (int)(t % n);
    out[t] = M[(int64_t)nu * n + mu];
  }
}

// Eg[mu][nu] += alpha (SVS[nu][mu] W[mu][nu] + VS[nu][mu] WS[mu][nu] + SV[nu][mu] SW[mu][nu] + v[nu][mu] SWS[mu][nu])
__global__ void eg_lr_kernel(const double* __restrict__ SVS, const double* __restrict__ VS,
                             const double* __restrict__ SV, const double* __restrict__ v, const double* __restrict__ W,
                             const double* __restrict__ WS, const double* __restrict__ SW,
                             const double* __restrict__ SWS, double alpha, int n, double* __restrict__ Eg) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < nn; t += (int64_t)gridDim.x * TPB) {
    const int mu = (int)(t / n), nu = (int)(t % n);
    const int64_t tt = (int64_t)nu * n + mu;
    Eg[t] += alpha * (SVS[tt] * W[t] + VS[tt] * WS[t] + SV[tt] * SW[t] + v[tt] * SWS[t]);
  }
}

// out[A][B] = sum_{mu in A, nu in B} (L[mu][nu] + L[nu][mu]); one wave per (A, B)
__global__ void block_sums_kernel(const double* __restrict__ L, int n, const int* __restrict__ off, int N,
                                  double* __restrict__ out) {
  const int64_t pair = (int64_t)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (pair >= (int64_t)N * N) return;
  const int A = (int)(pair / N), B = (int)(pair % N);
  const int a0 = off[A], na = off[A + 1] - a0, b0 = off[B], nb = off[B + 1] - b0;
  double acc = 0.0;
  for (int e = lane; e < na * nb; e += 64) {
    const int mu = a0 + e / nb, nu = b0 + e % nb;
    acc += L[(int64_t)mu * n + nu] + L[(int64_t)nu * n + mu];
  }
  acc = wave_sum(acc);
  if (lane == 0) out[pair] = acc;
}

// grad[3A+k]: one workgroup per (A, k); fixed-order reduction
__global__ void __launch_bounds__(TPB) nuclear_kernel(const double* __restrict__ dh, const double* __restrict__ ds,
                                                      const double* __restrict__ Hm, const double* __restrict__ Sm,
                                                      int n, const int* __restrict__ off, int N,
                                                      const double* __restrict__ dg, const double* __restrict__ Gat,
                                                      const double* __restrict__ dglr, const double* __restrict__ Lat,
                                                      double* __restrict__ grad) {
  __shared__ double red[TPB / 64];
  const int A = blockIdx.x / 3, k = blockIdx.x % 3;
  const int a0 = off[A], na = off[A + 1] - a0;
  const int64_t nn = (int64_t)n * n;
  const double* dhk = dh + k * nn;
  const double* dsk = ds + k * nn;
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)na * n; e += TPB) {
    const int mu = a0 + (int)(e / n), nu = (int)(e % n);
    const int64_t t = (int64_t)mu * n + nu, tt = (int64_t)nu * n + mu;
    acc += dhk[t] * (Hm[t] + Hm[tt]) + dsk[t] * (Sm[t] + Sm[tt]);
  }
  const int64_t row = ((int64_t)k * N + A) * N;
  for (int B = threadIdx.x; B < N; B += TPB) {
    acc += dg[row + B] * Gat[(int64_t)A * N + B];
    if (dglr) acc += dglr[row + B] * Lat[(int64_t)A * N + B];
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < TPB / 64; ++w) s += red[w];
    grad[blockIdx.x] = s;
  }
}

}  // namespace

void gradient_run(const GradView& V, const std::function<void(const double*, double*)>& apply_a, const GradIn& in,
                  GradOut& out) {
  const hipStream_t s = V.stream;
  const int N = V.N, n = V.n, o = V.o, v = V.v, op = V.op, vp = V.vp;
  const bool lc = V.gamma_lr != nullptr;
  const bool exact = in.exact != 0;
  const int64_t ovp = (int64_t)o * vp, oo = (int64_t)op * op, vv = (int64_t)vp * vp, vo = (int64_t)vp * op;
  const int64_t nn = (int64_t)n * n;
  Pool pool(s, V.cache);
  Mm mm{s, pool};
  // ---- inputs ----
  double* stage = pool.get(std::max<int64_t>({(int64_t)o * v, 3 * nn, 3 * (int64_t)N * N, nn}));
  double* X = pool.get(ovp);
  TDA_HIP(hipMemcpyAsync(stage, in.x, (size_t)o * v * 8, hipMemcpyHostToDevice, s));
  launch_pack_x(stage, o, v, vp, 1, X, s);
  double* dP = pool.get(nn);
  double* dS = pool.get(3 * nn);
  double* dH = pool.get(3 * nn);
  double* dg = pool.get(3 * (int64_t)N * N);
  double* dglr = lc ? pool.get(3 * (int64_t)N * N) : nullptr;
  TDA_HIP(hipMemcpyAsync(dP, in.dp, (size_t)nn * 8, hipMemcpyHostToDevice, s));
  TDA_HIP(hipMemcpyAsync(dS, in.ds_rows, (size_t)3 * nn * 8, hipMemcpyHostToDevice, s));
  TDA_HIP(hipMemcpyAsync(dH, in.dh_rows, (size_t)3 * nn * 8, hipMemcpyHostToDevice, s));
  TDA_HIP(hipMemcpyAsync(dg, in.dgamma, (size_t)3 * N * N * 8, hipMemcpyHostToDevice, s));
  if (lc) TDA_HIP(hipMemcpyAsync(dglr, in.dgamma_lr, (size_t)3 * N * N * 8, hipMemcpyHostToDevice, s));
  int* atom = pool.get_int(n);
  {
    std::vector<int> h(n);
    for (int A = 0; A < N; ++A)
      for (int mu = V.off_h[A]; mu < V.off_h[A + 1]; ++mu) h[mu] = A;
    TDA_HIP(hipMemcpyAsync(atom, h.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
    TDA_HIP(hipStreamSynchronize(s));  // h leaves scope
  }
  const double* qov = V.qov;
  const double* qoo = V.qoo;
  const double* qvv = V.qvv;
  const double* G = V.gamma;
  const double* L = V.gamma_lr;
  // ---- U, V, T (excited_state.rs:280-286) ----
  double* ex = pool.get(ovp);
  double* xe = pool.get(ovp);
  scale_x_kernel<<<grid1((int64_t)o * v), TPB, 0, s>>>(X, V.orbe, V.ob, V.vb, o, v, vp, ex, xe);
  TDA_LAUNCH_CHECK();
  double* tab = pool.get(vv);   // x^T x
  double* tij = pool.get(oo);   // x x^T
  double* qab = pool.get(vv);   // omega u_ab + v_ab
  double* qij = pool.get(oo);   // omega u_ij - v_ij + g_ij
  mm(v, v, o, X, plain(1), plain(vp), X, plain(vp), plain(1), tab, plain(vp), plain(1));
  mm(o, o, v, X, plain(vp), plain(1), X, plain(1), plain(vp), tij, plain(op), plain(1));
  mm(v, v, o, ex, plain(1), plain(vp), X, plain(vp), plain(1), qab, plain(vp), plain(1), 2.0);
  axpby_kernel<<<grid1(vv), TPB, 0, s>>>(tab, qab, vv, 2.0 * in.omega, 1.0);
  TDA_LAUNCH_CHECK();
  mm(o, o, v, xe, plain(vp), plain(1), X, plain(1), plain(vp), qij, plain(op), plain(1), -2.0);
  axpby_kernel<<<grid1(oo), TPB, 0, s>>>(tij, qij, oo, 2.0 * in.omega, 1.0);
  TDA_LAUNCH_CHECK();
  // ---- Coulomb parts of the Hav terms: w = gamma (q : V), expanded over q ----
  double* uvec = pool.get(N);
  auto qdot_gamma = [&](const double* Q, int64_t blk, const double* Vm, double* w) {
    mm(N, 1, blk, Q, plain(blk), plain(1), Vm, plain(1), plain(0), uvec, plain(1), plain(1));
    mm(N, 1, N, G, plain(N), plain(1), uvec, plain(1), plain(0), w, plain(1), plain(1));
  };
  // R[e] = alpha sum_A w[A] Q[A][e] + beta R[e]
  auto qexpand = [&](const double* w, const double* Q, int64_t blk, double* R, double alpha, double beta) {
    mm(1, blk, N, w, plain(0), plain(1), Q, plain(blk), plain(1), R, plain(0), plain(1), alpha, beta);
  };
  double* wvv = pool.get(N);
  double* woo = pool.get(N);
  double* wx = pool.get(N);
  double* wd = pool.get(N);
  qdot_gamma(qvv, vv, tab, wvv);
  qdot_gamma(qoo, oo, tij, woo);
  qdot_gamma(qov, ovp, X, wx);
  launch_copy(wvv, wd, N, s);
  axpby_kernel<<<grid1(N), TPB, 0, s>>>(woo, wd, N, -1.0, 1.0);  // wd = wvv - woo
  TDA_LAUNCH_CHECK();
  // ---- exchange parts: T[B] (per-atom product), U = gamma_lr T, contraction ----
  const int64_t tmax = std::max({ovp, oo, vv, vo});
  double* T = lc ? pool.get((int64_t)N * tmax) : nullptr;
  double* U = lc ? pool.get((int64_t)N * tmax) : nullptr;
  auto mix = [&](int64_t blk) {  // U[A] = sum_B gamma_lr[A][B] T[B]
    mm(N, blk, N, L, plain(N), plain(1), T, plain(blk), plain(1), U, plain(blk), plain(1));
  };
  // Each stage writes only the valid rows / columns of its T layout; clear
  // the rest so that U's pads are zero as well.
  auto tzero = [&]() { TDA_HIP(hipMemsetAsync(T, 0, (size_t)N * tmax * 8, s)); };
  double* qovP = nullptr;
  if (lc) {
    qovP = pool.get((int64_t)N * op * vp);
    pad_rows_kernel<<<grid1((int64_t)N * ovp), TPB, 0, s>>>(qov, N, o, op, vp, qovP);
    TDA_LAUNCH_CHECK();
  }
  // e1[k][i] = sum_{A,j} q_oo[A][k][j] (gamma_lr T)[A][i][j], T[B][i][j] = sum_a q_ov[B][i][a] Vm[j][a]
  // (the exchange part of hav_qai_or_wij, helpers.rs:797-818), into R with alpha / beta
  auto exch_qai = [&](const double* Vm, double* R, double alpha, double beta) {
    tzero();
    mm(o, o, v, qov, plain(vp), plain(1), Vm, plain(1), plain(vp), T, plain(op), plain(1), 1.0, 0.0, N, ovp, 0, oo);
    mix(oo);
    mm(o, o, (int64_t)N * op, qoo, plain(op), split(op, oo, 1), U, split(op, oo, 1), plain(op), R, plain(op), plain(1),
       alpha, beta);
  };
  // g_ij = Hav_Tab(t_ab) - Hav_Tij(t_ij) (helpers.rs:683-726): Coulomb 4 (wvv - woo) q_oo
  qexpand(wd, qoo, oo, qij, 4.0, 1.0);
  if (lc) {
    // - 2 E_a, E_a[j][i] = sum_{A,a} q_ov[A][j][a] U[A][i][a], T[B][i][a] = sum_b q_ov[B][i][b] t_ab[a][b]
    tzero();
    mm(o, v, v, qov, plain(vp), plain(1), tab, plain(1), plain(vp), T, plain(vp), plain(1), 1.0, 0.0, N, ovp, 0, ovp);
    mix(ovp);
    mm(o, o, (int64_t)N * vp, qov, plain(vp), split(vp, ovp, 1), U, split(vp, ovp, 1), plain(vp), qij, plain(op),
       plain(1), -2.0, 1.0);
    // + 2 E_b, E_b[j][i] = sum_{A,k} q_oo[A][j][k] U[A][i][k], T[B][i][k] = sum_l q_oo[B][i][l] t_ij[k][l]
    tzero();
    mm(o, o, o, qoo, plain(op), plain(1), tij, plain(1), plain(op), T, plain(op), plain(1), 1.0, 0.0, N, oo, 0, oo);
    mix(oo);
    mm(o, o, (int64_t)N * op, qoo, plain(op), split(op, oo, 1), U, split(op, oo, 1), plain(op), qij, plain(op),
       plain(1), 2.0, 1.0);
  }
  // ---- q_ia = x Hav_Qia_Xpy(x)^T + Hav_Qia_Tab(t_ab) - Hav_Qia_Tij(t_ij) (:318-323) ----
  double* qia = pool.get(ovp);
  {
    double* Rc = pool.get(vv);  // Hav_Qia_Xpy(x) [v][v]: 4 wx q_vv - 2 E_c
    qexpand(wx, qvv, vv, Rc, 4.0, 0.0);
    if (lc) {
      // E_c[b][a] = sum_{A,i} q_ov[A][i][b] U[A][a][i], T[B][a][i] = sum_c q_vv[B][a][c] x[i][c]
      tzero();
      mm(v, o, v, qvv, plain(vp), plain(1), X, plain(1), plain(vp), T, plain(op), plain(1), 1.0, 0.0, N, vv, 0, vo);
      mix(vo);
      mm(v, v, (int64_t)N * op, qovP, plain(1), split(op, (int64_t)op * vp, vp), U, split(op, vo, 1), plain(op), Rc,
         plain(vp), plain(1), -2.0, 1.0);
    }
    // q_ia[i][a] = sum_b x[i][b] Rc[a][b]
    mm(o, v, v, X, plain(vp), plain(1), Rc, plain(1), plain(vp), qia, plain(vp), plain(1));
  }
  qexpand(wd, qov, ovp, qia, 4.0, 1.0);
  if (lc) {
    // - 2 E_d, E_d[i][a] = sum_{A,b} q_ov[A][i][b] U[A][a][b], T[B][a][b] = sum_c q_vv[B][a][c] t_ab[b][c]
    tzero();
    mm(v, v, v, qvv, plain(vp), plain(1), tab, plain(1), plain(vp), T, plain(vp), plain(1), 1.0, 0.0, N, vv, 0, vv);
    mix(vv);
    mm(o, v, (int64_t)N * vp, qov, plain(vp), split(vp, ovp, 1), U, split(vp, vv, 1), plain(vp), qia, plain(vp),
       plain(1), -2.0, 1.0);
    // + 2 E_e, E_e[i][a] = sum_{A,j} q_oo[A][i][j] U[A][a][j], T[B][a][j] = sum_k q_ov[B][k][a] t_ij[j][k]
    tzero();
    mm(v, o, o, qov, plain(1), plain(vp), tij, plain(1), plain(op), T, plain(op), plain(1), 1.0, 0.0, N, ovp, 0, vo);
    mix(vo);
    mm(o, v, (int64_t)N * op, qoo, plain(op), split(op, oo, 1), U, split(op, vo, 1), plain(op), qia, plain(vp),
       plain(1), 2.0, 1.0);
  }
  // ---- q_ai^T = Hav_Qai(x)^T x (:326-329); Rf = 4 wx q_oo - 2 e1(x) ----
  double* Rf = pool.get(oo);
  qexpand(wx, qoo, oo, Rf, 4.0, 0.0);
  if (lc) exch_qai(X, Rf, -2.0, 1.0);
  double* qaiT = pool.get(ovp);
  mm(o, v, o, Rf, plain(1), plain(op), X, plain(vp), plain(1), qaiT, plain(vp), plain(1));
  // r_ia = q_ai^T - q_ia (:332)
  double* rhs = pool.get(ovp);
  launch_copy(qaiT, rhs, ovp, s);
  axpby_kernel<<<grid1(ovp), TPB, 0, s>>>(qia, rhs, ovp, -1.0, 1.0);
  TDA_LAUNCH_CHECK();
  // ---- Z-vector (:335-351) ----
  double* wtmp = pool.get(N);
  auto zop = [&](const double* P, double* AP) {
    apply_a(P, AP);  // de + 2 J (- K)
    if (exact || !lc) {
      qdot_gamma(qov, ovp, P, wtmp);
      qexpand(wtmp, qov, ovp, AP, 2.0, 1.0);  // + 2 J: 4 J in all
    }
    if (exact && lc) {
      // - K' P: T[B][a][b] = sum_j q_ov[B][j][a] P[j][b]; (K'P)[i][a] = sum_{A,b} q_ov[A][i][b] U[A][a][b]
      tzero();
      mm(v, v, o, qov, plain(1), plain(vp), P, plain(vp), plain(1), T, plain(vp), plain(1), 1.0, 0.0, N, ovp, 0, vv);
      mix(vv);
      mm(o, v, (int64_t)N * vp, qov, plain(vp), split(vp, ovp, 1), U, split(vp, vv, 1), plain(vp), AP, plain(vp),
         plain(1), -1.0, 1.0);
    }
  };
  double* Z = pool.get(ovp);
  double* R = pool.get(ovp);
  double* P = pool.get(ovp);
  double* AP = pool.get(ovp);
  double* part = pool.get(256);
  double* dots = pool.get(8);
  launch_cg_init(rhs, V.de, 1, ovp, Z, s);  // z0 = r / de (:1154-1155)
  zop(Z, AP);
  launch_copy(rhs, R, ovp, s);
  axpby_kernel<<<grid1(ovp), TPB, 0, s>>>(AP, R, ovp, -1.0, 1.0);
  TDA_LAUNCH_CHECK();
  launch_copy(R, P, ovp, s);
  // CG scalars stay on the device (rr, p.Ap, alpha, beta); the host reads
  // r.r once per iteration for the stopping test
  double* rr_d = dots;
  double* pap_d = dots + 1;
  double* a_d = dots + 2;
  double* b_d = dots + 3;
  double* rrn_d = dots + 4;
  double rr_new = 0.0;
  launch_dots(R, R, 1, ovp, part, rr_d, s);
  int its = in.max_iter;
  for (int it = 0; it < in.max_iter; ++it) {
    zop(P, AP);
    launch_dots(P, AP, 1, ovp, part, pap_d, s);
    cg_ratio_kernel<<<1, 64, 0, s>>>(rr_d, pap_d, a_d, nullptr, nullptr);  // a = r.r / p.Ap
    TDA_LAUNCH_CHECK();
    launch_cg_step(Z, R, P, AP, a_d, 1, ovp, s);
    launch_dots(R, R, 1, ovp, part, rrn_d, s);
    TDA_HIP(hipMemcpyAsync(&rr_new, rrn_d, 8, hipMemcpyDeviceToHost, s));
    TDA_HIP(hipStreamSynchronize(s));
    if (rr_new <= in.conv) {
      its = it + 1;
      break;
    }
    cg_ratio_kernel<<<1, 64, 0, s>>>(rrn_d, rr_d, b_d, rr_d, rrn_d);  // b = r.r new / old; r.r = new
    TDA_LAUNCH_CHECK();
    launch_cg_dir(P, R, b_d, 1, ovp, s);
  }
  out.z_iter = its;
  if (out.z) {
    launch_unpack_x(Z, o, v, vp, 1, stage, s);
    TDA_HIP(hipMemcpyAsync(out.z, stage, (size_t)o * v * 8, hipMemcpyDeviceToHost, s));
    TDA_HIP(hipStreamSynchronize(s));
  }
  // ---- W (:354-392) ----
  // w_ij = q_ij + Hav_Wij(z): 4 wz q_oo - 2 e1(z) (reference) or - (e1 + e1^T) (exact)
  qdot_gamma(qov, ovp, Z, wtmp);
  qexpand(wtmp, qoo, oo, qij, 4.0, 1.0);
  if (lc) {
    if (!exact) {
      exch_qai(Z, qij, -2.0, 1.0);
    } else {
      double* e1 = pool.get(oo);
      exch_qai(Z, e1, 1.0, 0.0);
      // qij -= e1 + e1^T (o x o block; pads of e1 are zero)
      double* e1t = pool.get(oo);
      transpose_kernel<<<grid1(oo), TPB, 0, s>>>(e1, op, e1t);
      TDA_LAUNCH_CHECK();
      axpby_kernel<<<grid1(oo), TPB, 0, s>>>(e1, qij, oo, -1.0, 1.0);
      axpby_kernel<<<grid1(oo), TPB, 0, s>>>(e1t, qij, oo, -1.0, 1.0);
      TDA_LAUNCH_CHECK();
    }
  }
  // w_ia = q_ai^T + e_o z
  double* wia = qaiT;
  {
    double* ez = pool.get(ovp);
    double* zej = pool.get(ovp);
    scale_x_kernel<<<grid1((int64_t)o * v), TPB, 0, s>>>(Z, V.orbe, V.ob, V.vb, o, v, vp, ez, zej);
    TDA_LAUNCH_CHECK();
    axpby_kernel<<<grid1(ovp), TPB, 0, s>>>(ez, wia, ovp, 1.0, 1.0);
    TDA_LAUNCH_CHECK();
  }
  // ---- AO matrices ----
  const double* C = V.C;
  const double* S = V.S;
  double* Mt = pool.get(nn);
  double* tmp = pool.get(nn);
  auto to_ao = [&](double* D) {  // D = C Mt C^T
    mm(n, n, n, C, plain(n), plain(1), Mt, plain(n), plain(1), tmp, plain(n), plain(1));
    mm(n, n, n, tmp, plain(n), plain(1), C, plain(1), plain(n), D, plain(n), plain(1));
  };
  double* Sm = pool.get(nn);  // S coefficient
  double* Hm = pool.get(nn);  // H0 coefficient: D_h = t_vv - t_oo + z_ao
  double* xao = pool.get(nn);
  // Sm = -W_AO = -C triu(W) C^T with the halved diagonals of w_ij / w_ab
  mo_assemble_kernel<<<grid1(nn), TPB, 0, s>>>(qij, -1.0, wia, -1.0, qab, -1.0, o, n, op, vp, 1, 0.5, Mt);
  TDA_LAUNCH_CHECK();
  to_ao(Sm);
  mo_assemble_kernel<<<grid1(nn), TPB, 0, s>>>(tij, -1.0, Z, 1.0, tab, 1.0, o, n, op, vp, 0, 1.0, Mt);
  TDA_LAUNCH_CHECK();
  to_ao(Hm);
  mo_assemble_kernel<<<grid1(nn), TPB, 0, s>>>(nullptr, 0.0, X, 1.0, nullptr, 0.0, o, n, op, vp, 0, 1.0, Mt);
  TDA_LAUNCH_CHECK();
  to_ao(xao);
  // ---- f_v terms: f_v(dP) : D_h and 2 f_v(x_ao) : x_ao ----
  double* Gat = pool.get((int64_t)N * N);
  double* Lat = lc ? pool.get((int64_t)N * N) : nullptr;
  {
    double* sv = pool.get(n);
    double* sw = pool.get(n);
    double* svat = pool.get(N);
    double* swat = pool.get(N);
    double* gsv = pool.get(n);
    double* hv = pool.get(n);
    auto f_v = [&](const double* vm, const double* Wm, double c) {
      rowsum_sym_kernel<<<(unsigned)ceil_div(n, TPB / 64), TPB, 0, s>>>(S, vm, n, sv);
      rowsum_sym_kernel<<<(unsigned)ceil_div(n, TPB / 64), TPB, 0, s>>>(S, Wm, n, sw);
      atom_sum_kernel<<<(unsigned)ceil_div(N, TPB), TPB, 0, s>>>(sv, V.off, N, svat);
      atom_sum_kernel<<<(unsigned)ceil_div(N, TPB), TPB, 0, s>>>(sw, V.off, N, swat);
      gamma_expand_kernel<<<(unsigned)ceil_div(n, TPB / 64), TPB, 0, s>>>(G, atom, N, n, svat, gsv);
      gamma_expand_kernel<<<(unsigned)ceil_div(n, TPB / 64), TPB, 0, s>>>(G, atom, N, n, swat, hv);
      es_v_kernel<<<grid1(nn), TPB, 0, s>>>(Wm, vm, gsv, hv, 0.25 * c, n, Sm);
      gat_kernel<<<grid1((int64_t)N * N), TPB, 0, s>>>(swat, svat, 0.25 * c, N, Gat);
      TDA_LAUNCH_CHECK();
    };
    f_v(dP, Hm, 1.0);
    f_v(xao, xao, 2.0);
  }
  // ---- f_lr terms: -1/2 f_lr(dP) : D_h and - f_lr(x_ao^T) : x_ao ----
  if (lc) {
    double* Ltot = pool.get(nn);
    double* vt = pool.get(nn);
    double* SV = pool.get(nn);
    double* WS = pool.get(nn);
    double* SW = pool.get(nn);
    double* VS = pool.get(nn);
    double* SWS = pool.get(nn);
    double* SVS = pool.get(nn);
    double* h1 = pool.get(nn);
    double* h2 = pool.get(nn);
    auto nnmm = [&](const double* A, const double* B, double* Cm, double alpha, double beta) {
      mm(n, n, n, A, plain(n), plain(1), B, plain(n), plain(1), Cm, plain(n), plain(1), alpha, beta);
    };
    auto had = [&](const double* M, double* outm) {
      had_g_kernel<<<grid1(nn), TPB, 0, s>>>(M, 0, L, atom, N, n, outm);
      TDA_LAUNCH_CHECK();
    };
    auto f_lr = [&](const double* vm, const double* Wm, double c) {
      const double a = 0.25 * c;
      nnmm(S, vm, SV, 1.0, 0.0);
      nnmm(Wm, S, WS, 1.0, 0.0);
      nnmm(S, Wm, SW, 1.0, 0.0);
      nnmm(vm, S, VS, 1.0, 0.0);
      nnmm(SW, S, SWS, 1.0, 0.0);
      nnmm(SV, S, SVS, 1.0, 0.0);
      had(Wm, h1);  // (g o W) S v + v S (g o W)
      nnmm(h1, SV, Sm, a, 1.0);
      nnmm(VS, h1, Sm, a, 1.0);
      had(WS, h1);  // (g o WS) v
      nnmm(h1, vm, Sm, a, 1.0);
      had(SV, h1);  // W (Sv o g)
      nnmm(Wm, h1, Sm, a, 1.0);
      had(vm, h2);  // WS (g o v) + (g o v) SW
      nnmm(WS, h2, Sm, a, 1.0);
      nnmm(h2, SW, Sm, a, 1.0);
      had(VS, h1);  // (vS o g) W
      nnmm(h1, Wm, Sm, a, 1.0);
      had(SW, h1);  // v (g o SW)
      nnmm(vm, h1, Sm, a, 1.0);
      eg_lr_kernel<<<grid1(nn), TPB, 0, s>>>(SVS, VS, SV, vm, Wm, WS, SW, SWS, a, n, Ltot);
      TDA_LAUNCH_CHECK();
    };
    f_lr(dP, Hm, -0.5);
    transpose_kernel<<<grid1(nn), TPB, 0, s>>>(xao, n, vt);
    TDA_LAUNCH_CHECK();
    f_lr(vt, xao, -1.0);
    block_sums_kernel<<<(unsigned)ceil_div((int64_t)N * N, TPB / 64), TPB, 0, s>>>(Ltot, n, V.off, N, Lat);
    TDA_LAUNCH_CHECK();
  }
  // ---- nuclear contraction ----
  double* gd = pool.get(3 * (int64_t)N);
  nuclear_kernel<<<(unsigned)(3 * N), TPB, 0, s>>>(dH, dS, Hm, Sm, n, V.off, N, dg, Gat, dglr, Lat, gd);
  TDA_LAUNCH_CHECK();
  TDA_HIP(hipMemcpyAsync(out.grad, gd, (size_t)3 * N * 8, hipMemcpyDeviceToHost, s));
  TDA_HIP(hipStreamSynchronize(s));
}

}  // namespace tda
