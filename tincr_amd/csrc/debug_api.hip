// Kernel-level test hooks (include/tincr_tda_debug.h).
#include <cstring>
#include <vector>

#include "../../include/tincr_tda.h"
#include "../../include/tincr_tda_debug.h"
#include "gemm.h"
#include "gemm16_impl.h"
#include "kernels.h"

using namespace tda;

namespace {
struct Buf {
  void* p = nullptr;
  explicit Buf(size_t n) { TDA_HIP(hipMalloc(&p, n ? n : 16)); }
  ~Buf() { (void)hipFree(p); }
};

template <typename F>
int guard(F&& f) {
  try {
    f();
    return TDA_OK;
  } catch (const Error& e) {
    fprintf(stderr, "tincr_tda debug: %s\n", e.msg.c_str());
    return e.code;
  }
}
}  // namespace

extern "C" int tda_debug_gemm(int in_prec, int a_kc, int b_kc, int64_t M, int64_t N, int64_t K, const double* A,
                              const double* B, double* C, double alpha, double beta, int ksplit) {
  return guard([&] {
    TDA_CHECK(in_prec >= 0 && in_prec <= 3, TDA_ERR_BAD_ARG, "bad precision");
    const size_t esz = prec_bytes(in_prec);
    Buf a64(M * K * 8), b64(K * N * 8), a(M * K * esz), b(K * N * esz), c(M * N * 8);
    TDA_HIP(hipMemcpy(a64.p, A, M * K * 8, hipMemcpyHostToDevice));
    TDA_HIP(hipMemcpy(b64.p, B, K * N * 8, hipMemcpyHostToDevice));
    TDA_HIP(hipMemcpy(c.p, C, M * N * 8, hipMemcpyHostToDevice));
    launch_convert((const double*)a64.p, M * K, 1, nullptr, a.p, in_prec, 0);
    launch_convert((const double*)b64.p, K * N, 1, nullptr, b.p, in_prec, 0);
    GemmArgs g;
    std::memset(&g, 0, sizeof(g));
    g.A = a.p; g.B = b.p; g.C = c.p;
    g.M = M; g.N = N; g.K = K;
    g.am = a_kc ? plain(K) : plain(1);
    g.ak = a_kc ? plain(1) : plain(M);
    g.bk = b_kc ? plain(1) : plain(N);
    g.bn = b_kc ? plain(K) : plain(1);
    g.cm = plain(N); g.cn = plain(1);
    g.batch = 1; g.alpha = alpha; g.beta = beta;
    g.dm = plain(0); g.dmap = plain(1); g.xm = plain(0); g.xn = plain(0);
    g.hm = plain(0); g.hn = plain(0);
    const DType dt = in_prec == 0 ? DT_F64 : in_prec == 1 ? DT_F32 : in_prec == 2 ? DT_BF16 : DT_F16;
    GemmPlan plan = gemm_plan(dt, M, N, K, 1, a_kc, b_kc, ksplit);
    size_t wsb = gemm_workspace_bytes(dt, plan, M, N, 1);
    Buf ws(wsb);
    g.ws = ws.p;
    gemm_launch(dt, DT_F64, EPI_STORE, a_kc, b_kc, g, plan, 0);
    TDA_HIP(hipDeviceSynchronize());
    TDA_HIP(hipMemcpy(C, c.p, M * N * 8, hipMemcpyDeviceToHost));
  });
}

extern "C" int tda_debug_gemm_f32store(int a_kc, int b_kc, int64_t M, int64_t N, int64_t K, const double* A,
                                       const double* B, double* C, double alpha, double beta, int ksplit, int a32,
                                       int b32) {
  return guard([&] {
    TDA_CHECK(M > 0 && N > 0 && K > 0 && (a32 || b32), TDA_ERR_BAD_ARG, "bad arguments");
    Buf a64(M * K * 8), b64(K * N * 8), a(M * K * 4), b(K * N * 4), c(M * N * 8);
    TDA_HIP(hipMemcpy(a64.p, A, M * K * 8, hipMemcpyHostToDevice));
    TDA_HIP(hipMemcpy(b64.p, B, K * N * 8, hipMemcpyHostToDevice));
    TDA_HIP(hipMemcpy(c.p, C, M * N * 8, hipMemcpyHostToDevice));
    if (a32) launch_cast_vectors((const double*)a64.p, M * K, a.p, TDA_PREC_FP32, 0);
    if (b32) launch_cast_vectors((const double*)b64.p, K * N, b.p, TDA_PREC_FP32, 0);
    GemmArgs g;
    std::memset(&g, 0, sizeof(g));
    g.A = a32 ? a.p : a64.p;
    g.B = b32 ? b.p : b64.p;
    g.C = c.p;
    g.M = M; g.N = N; g.K = K;
    g.am = a_kc ? plain(K) : plain(1);
    g.ak = a_kc ? plain(1) : plain(M);
    g.bk = b_kc ? plain(1) : plain(N);
    g.bn = b_kc ? plain(K) : plain(1);
    g.cm = plain(N); g.cn = plain(1);
    g.batch = 1; g.alpha = alpha; g.beta = beta;
    g.dm = plain(0); g.dmap = plain(1); g.xm = plain(0); g.xn = plain(0);
    g.hm = plain(0); g.hn = plain(0);
    g.a_f32 = a32 != 0;
    g.b_f32 = b32 != 0;
    GemmPlan plan = gemm_plan(DT_F64, M, N, K, 1, a_kc, b_kc, ksplit);
    Buf ws(gemm_workspace_bytes(DT_F64, plan, M, N, 1));
    g.ws = ws.p;
    gemm_launch(DT_F64, DT_F64, EPI_STORE, a_kc, b_kc, g, plan, 0);
    TDA_HIP(hipDeviceSynchronize());
    TDA_HIP(hipMemcpy(C, c.p, M * N * 8, hipMemcpyDeviceToHost));
  });
}

extern "C" int tda_debug_rowdot(int M, int N, int64_t K, const double* A, int a32, const double* B, int b32,
                                double* C, double alpha, double beta) {
  return guard([&] {
    TDA_CHECK(M > 0 && N > 0 && K > 0, TDA_ERR_BAD_ARG, "bad arguments");
    const int64_t ld = (K + 3) / 4 * 4;  // 16-byte aligned rows
    Buf a64(M * K * 8), b64(N * K * 8), a((size_t)M * ld * 8), b((size_t)N * ld * 8), c((size_t)M * N * 8);
    Buf ws(rowdot_ws_bytes(M, N, K));
    TDA_HIP(hipMemcpy(a64.p, A, M * K * 8, hipMemcpyHostToDevice));
    TDA_HIP(hipMemcpy(b64.p, B, N * K * 8, hipMemcpyHostToDevice));
    TDA_HIP(hipMemcpy(c.p, C, (size_t)M * N * 8, hipMemcpyHostToDevice));
    // rows of K elements at stride ld (FP32 or FP64)
    for (int i = 0; i < M; ++i) {
      if (a32) launch_cast_vectors((const double*)a64.p + i * K, K, (float*)a.p + i * ld, TDA_PREC_FP32, 0);
      else TDA_HIP(hipMemcpy((double*)a.p + i * ld, (const double*)a64.p + i * K, K * 8, hipMemcpyDeviceToDevice));
    }
    for (int j = 0; j < N; ++j) {
      if (b32) launch_cast_vectors((const double*)b64.p + j * K, K, (float*)b.p + j * ld, TDA_PREC_FP32, 0);
      else TDA_HIP(hipMemcpy((double*)b.p + j * ld, (const double*)b64.p + j * K, K * 8, hipMemcpyDeviceToDevice));
    }
    launch_rowdot(a.p, a32 != 0, ld, b.p, b32 != 0, ld, M, N, K, alpha, beta, (double*)c.p, N, (double*)ws.p, 0);
    TDA_HIP(hipDeviceSynchronize());
    TDA_HIP(hipMemcpy(C, c.p, (size_t)M * N * 8, hipMemcpyDeviceToHost));
  });
}

extern "C" int tda_debug_gemm16(int in_prec, int epi, int64_t M, int64_t N, int64_t K, int batch, int64_t lda,
                                int64_t ldb, const double* A, const double* B, double alpha, const double* H, int nh,
                                const int* hidx, int hmod, const double* diag, const double* X, double* C,
                                int flags, int* used) {
  const int split = flags & 1;
  const bool full = (flags & 2) != 0;
  return guard([&] {
    TDA_CHECK(in_prec == 2 || in_prec == 3, TDA_ERR_BAD_ARG, "16-bit precision expected");
    TDA_CHECK(epi >= 0 && epi <= 2 && M > 0 && N > 0 && K > 0 && batch > 0 && lda >= K && ldb >= K,
              TDA_ERR_BAD_ARG, "bad gemm16 arguments");
    TDA_CHECK(!split || in_prec == 3, TDA_ERR_BAD_ARG, "split operands are FP16");
    const DType dt = in_prec == 2 ? DT_BF16 : DT_F16;
    const int64_t w = split ? 2 : 1;  // split: rows [hi | lo] of width 2 ld
    const size_t na = (size_t)M * lda, nb = (size_t)batch * N * ldb, nc = (size_t)batch * M * N;
    Buf a64(na * 8), b64(nb * 8), a(na * 2 * w), b(nb * 2 * w), c(nc * 8 * w);
    TDA_HIP(hipMemcpy(a64.p, A, na * 8, hipMemcpyHostToDevice));
    TDA_HIP(hipMemcpy(b64.p, B, nb * 8, hipMemcpyHostToDevice));
    if (split) {
      launch_convert_split((const double*)a64.p, M, lda, lda, a.p, 2 * lda, lda, 0);
      launch_convert_split((const double*)b64.p, (int64_t)batch * N, ldb, ldb, b.p, 2 * ldb, ldb, 0);
    } else {
      launch_convert((const double*)a64.p, na, 1, nullptr, a.p, in_prec, 0);
      launch_convert((const double*)b64.p, nb, 1, nullptr, b.p, in_prec, 0);
    }
    Buf h(epi == EPI_HADAMARD ? (size_t)nh * nh * 8 : 8), hi(epi == EPI_HADAMARD && hidx ? (size_t)hmod * 4 : 8);
    Buf dg(epi == EPI_DIAG ? (size_t)M * N * 8 : 8), xs(epi == EPI_DIAG ? nc * 8 : 8);
    Buf hf(full && epi == EPI_HADAMARD && hmod > 0 ? (size_t)hmod * ((hmod + 3) / 4 * 4) * 4 : 16);
    GemmArgs g;
    std::memset(&g, 0, sizeof(g));
    g.A = a.p; g.B = b.p; g.C = c.p;
    g.M = M; g.N = N; g.K = K;
    g.am = plain(w * lda); g.ak = plain(1); g.bk = plain(1); g.bn = plain(w * ldb);
    g.cm = plain(w * N); g.cn = plain(1);
    g.batch = batch; g.a_batch = 0; g.b_batch = N * w * ldb; g.c_batch = M * w * N;
    if (split) {
      g.a_lo = lda;
      g.b_lo = ldb;
    }
    g.alpha = alpha;
    g.dm = plain(0); g.dmap = plain(1); g.xm = plain(0); g.xn = plain(0);
    g.hm = plain(0); g.hn = plain(0);
    DType out = dt;
    if (epi == EPI_HADAMARD) {
      TDA_HIP(hipMemcpy(h.p, H, (size_t)nh * nh * 8, hipMemcpyHostToDevice));
      if (hidx) TDA_HIP(hipMemcpy(hi.p, hidx, (size_t)hmod * 4, hipMemcpyHostToDevice));
      g.hsrc = (const double*)h.p;
      g.hidx = hidx ? (const int*)hi.p : nullptr;
      g.hmod = hmod;
      g.hm = plain(nh); g.hn = plain(1);
      if (full) {
        TDA_CHECK(hidx && hmod > 0, TDA_ERR_BAD_ARG, "the factor table needs hidx");
        const int64_t hld = (hmod + 3) / 4 * 4;
        std::vector<float> tab((size_t)hmod * hld, 0.f);
        for (int a = 0; a < hmod; ++a)
          for (int b2 = 0; b2 < hmod; ++b2) tab[(size_t)a * hld + b2] = (float)H[(size_t)hidx[a] * nh + hidx[b2]];
        TDA_HIP(hipMemcpy(hf.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
        g.hfull = (const float*)hf.p;
        g.hld = hld;
      }
    } else if (epi == EPI_DIAG) {
      TDA_HIP(hipMemcpy(dg.p, diag, (size_t)M * N * 8, hipMemcpyHostToDevice));
      TDA_HIP(hipMemcpy(xs.p, X, nc * 8, hipMemcpyHostToDevice));
      g.diag = (const double*)dg.p; g.dm = plain(N); g.dmap = plain(1);
      g.xsrc = (const double*)xs.p; g.xm = plain(N); g.xn = plain(1); g.x_batch = M * N;
      out = DT_F64;
      g.cm = plain(N); g.c_batch = M * N;
    }
    if (split && epi != EPI_DIAG) g.c_lo = N;
    const Epi e = (Epi)epi;
    *used = gemm16_ok(dt, out, e, true, true, g, 1) ? 1 : 0;
    GemmPlan plan = gemm_plan(dt, M, N, K, batch, true, true, 0);
    Buf ws(gemm_workspace_bytes(dt, plan, M, N, batch));
    g.ws = ws.p;
    gemm_launch(dt, out, e, true, true, g, plan, 0);
    TDA_HIP(hipDeviceSynchronize());
    if (out == DT_F64) {
      TDA_HIP(hipMemcpy(C, c.p, nc * 8, hipMemcpyDeviceToHost));
    } else if (split) {
      // rows [hi | lo]: value = hi + lo 2^-11
      std::vector<_Float16> raw(nc * 2);
      TDA_HIP(hipMemcpy(raw.data(), c.p, nc * 4, hipMemcpyDeviceToHost));
      for (size_t r = 0; r < (size_t)batch * M; ++r)
        for (int64_t n = 0; n < N; ++n)
          C[r * N + n] = (double)(float)raw[r * 2 * N + n] + (double)(float)raw[r * 2 * N + N + n] / 2048.0;
    } else {
      std::vector<uint16_t> raw(nc);
      TDA_HIP(hipMemcpy(raw.data(), c.p, nc * 2, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < nc; ++i) {
        if (dt == DT_F16) {
          _Float16 v;
          std::memcpy(&v, &raw[i], 2);
          C[i] = (double)(float)v;
        } else {
          const uint32_t bits = (uint32_t)raw[i] << 16;
          float v;
          std::memcpy(&v, &bits, 4);
          C[i] = (double)v;
        }
      }
    }
  });
}

extern "C" int tda_debug_eigh(int m, int n, const double* H, double* evals, double* Y) {
  return tda_debug_eigh_timed(m, n, H, evals, Y, nullptr);
}

extern "C" int tda_debug_eigh_timed(int m, int n, const double* H, double* evals, double* Y, double* phase_us) {
  return guard([&] {
    Buf h((size_t)m * m * 8), w(eigh_work_doubles(m) * 8), e((size_t)n * 8), y((size_t)n * m * 8), clk(64);
    TDA_HIP(hipMemcpy(h.p, H, (size_t)m * m * 8, hipMemcpyHostToDevice));
    launch_eigh((const double*)h.p, m, m, n, (double*)w.p, (double*)e.p, (double*)y.p, m, 0,
                phase_us ? (unsigned long long*)clk.p : nullptr);
    TDA_HIP(hipDeviceSynchronize());
    if (phase_us) {
      unsigned long long st[6];
      TDA_HIP(hipMemcpy(st, clk.p, sizeof(st), hipMemcpyDeviceToHost));
      int rate_khz = 0;
      TDA_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0));
      for (int i = 0; i < 5; ++i) phase_us[i] = (double)(st[i + 1] - st[i]) * 1e3 / (double)rate_khz;
    }
    TDA_HIP(hipMemcpy(evals, e.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    TDA_HIP(hipMemcpy(Y, y.p, (size_t)n * m * 8, hipMemcpyDeviceToHost));
  });
}

namespace {
// Reads n doubles and writes n doubles: the HBM-traffic calibration of a
// counter run (FETCH_SIZE / WRITE_SIZE per dispatch for known bytes).
__global__ void __launch_bounds__(256) calib_copy_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                         int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) dst[t] = src[t];
}
}  // namespace

extern "C" int tda_debug_copy(int64_t bytes, int reps) {
  return guard([&] {
    TDA_CHECK(bytes >= 8 && reps >= 1, TDA_ERR_BAD_ARG, "bad copy size");
    const int64_t n = bytes / 8;
    Buf a(n * 8), b(n * 8);
    TDA_HIP(hipMemset(a.p, 0, n * 8));
    for (int r = 0; r < reps; ++r) {
      calib_copy_kernel<<<4096, 256, 0, 0>>>((const double*)a.p, (double*)b.p, n);
      TDA_LAUNCH_CHECK();
    }
    TDA_HIP(hipDeviceSynchronize());
  });
}
