// LDS-tiled MFMA GEMM for gfx950 with generic (two-level) index maps.
//
//   C[cm(m) + cn(n)]  =  epilogue( alpha * sum_k A[am(m) + ak(k)] * B[bk(k) + bn(n)] )
//
// The contraction layouts of the TDA path (transition-charge build, Coulomb
// projections, the three-stage exchange and the Davidson subspace algebra)
// are all expressed through these maps, so one kernel family serves them.
#pragma once
#include "common.h"

namespace tda {

enum Epi : int {
  EPI_STORE = 0,     // C = alpha*acc*scale[z] + beta*C
  EPI_DIAG = 1,      // C = alpha*acc + diag[dm(m) + dmap(n)] * X[z*x_batch + xm(m) + xn(n)]
  EPI_HADAMARD = 2,  // C = alpha*acc * H[hm(r(m)) + hn(r(n))], r(x) = hidx ? hidx[x % hmod] : x
};

struct GemmArgs {
  const void* A;
  const void* B;
  void* C;
  int64_t M, N, K;
  IdxMap am, ak, bk, bn, cm, cn;
  // batching over blockIdx.z (outer) and split-K (inner)
  int batch;
  int64_t a_batch, b_batch, c_batch;
  int ksplit;
  int64_t k_chunk;  // multiple of BK
  // split-K partials: ws[(z * M + m) * N + n], accumulator type
  void* ws;
  double alpha, beta;
  const float* batch_scale;  // optional: multiplies alpha per batch index
  // EPI_DIAG
  const double* diag;
  IdxMap dm, dmap;
  const double* xsrc;
  IdxMap xm, xn;
  int64_t x_batch;
  // EPI_HADAMARD
  const double* hsrc;
  IdxMap hm, hn;
  const int* hidx;  // optional row / column gather table (e.g. AO -> atom)
  int64_t hmod;     // its length; indices are reduced modulo hmod first
  // Split 16-bit operands (16-bit kernel only): a value v is stored as
  // hi = fp16(v) and lo = fp16((v - hi) * 2^11), lo at element offset a_lo /
  // b_lo from hi in the same row.  Then
  //   sum_k A B = 2^-11 sum_k (A_hi B_lo + A_lo B_hi) + sum_k A_hi B_hi
  // (the lo x lo term, 2^-22 of the value, is dropped).  0 = not split.
  // c_lo != 0: the result is written split the same way, lo at offset c_lo.
  int64_t a_lo, b_lo, c_lo;
  // EPI_HADAMARD in the 16-bit kernel: when set, the factor of element (m, n)
  // is hfull[(m % hmod) * hld + (n % hmod)] (a precomputed FP32 table, e.g.
  // gamma_lr expanded to AO pairs) instead of H gathered through hidx.
  const float* hfull;
  int64_t hld;
  // FP64 products of FP32-stored operands (a_f32 / b_f32 = 1: A / B hold
  // floats, widened to FP64 on their way into LDS; generic kernel only).
  int a_f32, b_f32;
  // EPI_DIAG of the 16-bit kernel: the x of the diagonal term read as FP32
  // (xsrc32) instead of FP64 (xsrc); the result may then be FP32 as well.
  const float* xsrc32;
};

// Row / column index into H of the Hadamard epilogue.
__host__ __device__ inline int64_t hada_index(const GemmArgs& g, int64_t x) {
  if (!g.hidx) return x;
  const int64_t r = (uint64_t)x < (UINT64_C(1) << 32) && (uint64_t)g.hmod < (UINT64_C(1) << 32)
                        ? (int64_t)((uint32_t)x % (uint32_t)g.hmod)
                        : x % g.hmod;
  return g.hidx[r];
}

// Tile configurations of the generic kernel (gemm_impl.h): 128x128, 64x32,
// 32x128, 64x64, 128x128 with 128-deep K slices (16-bit inputs), 64x16 with
// 64-deep K slices.  CFG_SKINNY_N16 is instantiated for FP64 inputs only
// (other inputs fall back to CFG_SKINNY_N).
enum CfgId : int { CFG_BIG = 0, CFG_SKINNY_N = 1, CFG_SKINNY_M = 2, CFG_SMALL = 3, CFG_BIG_K = 4, CFG_SKINNY_N16 = 5 };

// Work list of one grouped launch (a batch's fragments): descriptors, tiles
// (grouped GEMMs) and their count.  Grouped kernels read it through a pointer
// to a slot at a fixed device address, so a captured launch replays with the
// work list of the current step.
struct GroupSlot {
  const void* descs;
  const int4* tiles;
  int count;
  int pad;
};

// Element type of an operand / result buffer.
enum DType : int { DT_F64 = 0, DT_F32 = 1, DT_BF16 = 2, DT_F16 = 3 };

// Launch C = op(A*B).  `in` is the type of A and B (also selects the MFMA),
// `out` the type of C.  Chooses tile shape and split-K from the problem size
// unless `ksplit_hint` > 0.  When ksplit > 1 a workspace of
// ksplit*batch*M*N accumulators is taken from `ws` (caller sized, see
// gemm_workspace_bytes) and a reduction kernel applies the epilogue.
struct GemmPlan {
  int cfg;
  int ksplit;
  int64_t k_chunk;
};

GemmPlan gemm_plan(DType in, int64_t M, int64_t N, int64_t K, int batch, bool a_kc, bool b_kc,
                   int ksplit_hint = 0);
size_t gemm_workspace_bytes(DType in, const GemmPlan& p, int64_t M, int64_t N, int batch);
// `after_main` (optional) is recorded right after the main GEMM kernel, before
// any split-K reduction, so callers can time the MFMA kernel alone.
void gemm_launch(DType in, DType out, Epi epi, bool a_kc, bool b_kc, GemmArgs args,
                 const GemmPlan& plan, hipStream_t stream, hipEvent_t after_main = nullptr);

}  // namespace tda
