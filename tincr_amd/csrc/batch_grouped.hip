// Grouped Davidson for a batch of fragments (FMO monomers and CT pairs:
// src/fmo/lcmo/basis.rs:462-465, 172-238; each item follows
// src/excited_states/solvers/davidson.rs:53-185 with the engine of
// solvers/mod.rs:57-274).
//
// Small fragments (a few to a few hundred AOs) run one context's Davidson
// far below the GPU's size: every step is a dozen launches of microseconds
// and a host synchronisation.  Here all fragments advance together: each
// step of the Davidson is ONE launch per operation over every fragment still
// active, driven by per-fragment descriptors:
//
//   grouped GEMMs (gemm_grouped_kernel): persistent workgroups walk a tile
//     list {fragment problem, batch index, tile}; the products A.x (Coulomb
//     q_ov^T gamma q_ov x and the two-stage exchange over atoms, as in the
//     single-fragment charges form), the projections V^T AV, the Ritz and
//     residual recombinations Y V / Y AV, the Gram-Schmidt overlaps
//   batched subspace eigensolver (one workgroup per fragment)
//   grouped vector kernels (residual + norm, preconditioner, copies, casts,
//     in-block Gram-Schmidt with the 1e-8 drop, one workgroup per unit)
//
// and the host makes the per-fragment decisions of davidson.rs between two
// synchronisations per step (convergence, expansion by the unconverged roots,
// collapse at 15 n; converged or failed fragments simply get no more work:
// the convergence mask).  The descriptors and tile lists of a step are
// uploaded in one copy.  FP64 batches use FP64 products (the reference
// arithmetic); the others FP32 products (FP32 MFMA, FP32 copies of q_ov,
// q_vv, G) until a fragment converges, then FP64 products verify (and if
// needed continue) from its Ritz vectors, orthonormalised in FP64.
// Gram-Schmidt runs two classical passes against the basis ("twice is
// enough"), then the in-block loop; the subspace algebra is FP64 throughout.
//
// Tile shapes: the heavy product GEMMs pick their tile configuration with a
// timed trial of the candidates on the first step that runs them (cached for
// the process).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "batch.h"
#include "internal.h"
#include "kernels.h"

namespace tda {
namespace {

constexpr int GRID_GEMM = 1024;  // persistent workgroups of a grouped GEMM
constexpr int GRID_VEC = 2048;   // of a grouped vector kernel

// ---------------------------------------------------------------------------
// grouped vector kernels: one workgroup per unit, persistent
// ---------------------------------------------------------------------------
struct VecUnit {
  double* dst;
  const double* src;
  const double* aux;   // residual: Ritz vector; precondition: de
  const double* wptr;  // residual / precondition: the root's Ritz value
  double* out;         // residual: |R| ; guess: unused
  float* dst32;        // cast target
  int64_t L;
  int64_t pos;         // guess: position of the 1
};
enum VecOp : int { VO_RESIDUAL = 0, VO_PRECOND = 1, VO_COPY = 2, VO_CAST = 3, VO_GUESS = 4 };

__device__ inline double block_sum256(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

template <int OP>
__global__ void __launch_bounds__(256) vec_kernel(const GroupSlot* __restrict__ slot) {
  __shared__ double red[4];
  const int nu = slot->count;
  const VecUnit* __restrict__ units = (const VecUnit*)slot->descs;
  for (int u = blockIdx.x; u < nu; u += gridDim.x) {
    const VecUnit e = units[u];
    if constexpr (OP == VO_RESIDUAL) {
      // R -= theta RZ ; |R|   (davidson.rs:104-107)
      const double w = *e.wptr;
      double ss = 0.0;
      for (int64_t t = threadIdx.x; t < e.L; t += 256) {
        const double r = e.dst[t] - w * e.aux[t];
        e.dst[t] = r;
        ss += r * r;
      }
      ss = block_sum256(ss, red);
      if (threadIdx.x == 0) *e.out = sqrt(ss);
    } else if constexpr (OP == VO_PRECOND) {
      // r / (w - de), |w - de| < 1e-4 -> 1   (solvers/mod.rs:373-384)
      const double w = *e.wptr;
      for (int64_t t = threadIdx.x; t < e.L; t += 256) {
        const double d = w - e.aux[t];
        e.dst[t] = e.src[t] / (fabs(d) < 1e-4 ? 1.0 : d);
      }
    } else if constexpr (OP == VO_COPY) {
      for (int64_t t = threadIdx.x; t < e.L; t += 256) e.dst[t] = e.src[t];
    } else if constexpr (OP == VO_CAST) {
      for (int64_t t = threadIdx.x; t < e.L; t += 256) e.dst32[t] = (float)e.src[t];
    } else {  // VO_GUESS: a unit vector (solvers/utils.rs:17-26)
      for (int64_t t = threadIdx.x; t < e.L; t += 256) e.dst[t] = t == e.pos ? 1.0 : 0.0;
    }
    __syncthreads();
  }
}

// In-block Gram-Schmidt of nc vectors W (already projected twice against the
// basis) appended to V from row m (davidson.rs:162-168): each vector is
// projected (two classical passes) against the ones accepted before it and
// kept, normalised, when its norm exceeds 1e-8.  One workgroup per fragment.
struct GsDesc {
  double* W;   // [nc][L], overwritten
  double* Vm;  // V + m L: accepted vectors go here
  int* count;  // accepted
  int64_t L;
  int nc;
};

__global__ void __launch_bounds__(256) gs_inblock_kernel(const GroupSlot* __restrict__ slot) {
  __shared__ double red[4][8];
  __shared__ double coef[8];
  const int nd = slot->count;
  const GsDesc* __restrict__ d = (const GsDesc*)slot->descs;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int b = blockIdx.x; b < nd; b += gridDim.x) {
    const GsDesc e = d[b];
    int acc = 0;
    for (int j = 0; j < e.nc; ++j) {
      double* u = e.W + (int64_t)j * e.L;
      for (int pass = 0; pass < 2; ++pass) {
        for (int s0 = 0; s0 < acc; s0 += 8) {
          const int ns = min(8, acc - s0);
          double part[8];
#pragma unroll
          for (int s = 0; s < 8; ++s) part[s] = 0.0;
          for (int64_t t = threadIdx.x; t < e.L; t += 256) {
            const double ut = u[t];
#pragma unroll
            for (int s = 0; s < 8; ++s)
              if (s < ns) part[s] += e.Vm[(int64_t)(s0 + s) * e.L + t] * ut;
          }
#pragma unroll
          for (int s = 0; s < 8; ++s) {
            const double v = wave_sum(part[s]);
            if (lane == 0) red[wave][s] = v;
          }
          __syncthreads();
          if (threadIdx.x < 8) coef[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                                   red[3][threadIdx.x];
          __syncthreads();
          for (int64_t t = threadIdx.x; t < e.L; t += 256) {
            double ut = u[t];
            for (int s = 0; s < ns; ++s) ut -= coef[s] * e.Vm[(int64_t)(s0 + s) * e.L + t];
            u[t] = ut;
          }
          __syncthreads();
        }
      }
      double ss = 0.0;
      for (int64_t t = threadIdx.x; t < e.L; t += 256) ss += u[t] * u[t];
      __shared__ double rs[4];
      const double nrm = sqrt(block_sum256(ss, rs));
      if (nrm > 1e-8) {
        double* q = e.Vm + (int64_t)acc * e.L;
        const double inv = 1.0 / nrm;
        for (int64_t t = threadIdx.x; t < e.L; t += 256) q[t] = u[t] * inv;
        ++acc;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) *e.count = acc;
    __syncthreads();
  }
}

// Mulliken transition dipoles and oscillator strengths of one fragment per
// workgroup (tda/mod.rs:91-126, moments.rs:7-24): mu[d][c] = sum_A xyz[A][d]
// qt[A][c], f = 2/3 w |mu|^2.
struct DipDesc {
  const double* xyz;  // [N][3]
  const double* qt;   // [N][n]
  const double* w;    // [n]
  double* mu;         // [3][n]
  double* f;          // [n]
  int N, n;
};
__global__ void __launch_bounds__(256) dipoles_grouped_kernel(const GroupSlot* __restrict__ slot) {
  const int nd = slot->count;
  const DipDesc* __restrict__ d = (const DipDesc*)slot->descs;
  for (int b = blockIdx.x; b < nd; b += gridDim.x) {
    const DipDesc e = d[b];
    for (int c = threadIdx.x; c < e.n; c += 256) {
      double m[3] = {0.0, 0.0, 0.0};
      for (int A = 0; A < e.N; ++A)
        for (int k = 0; k < 3; ++k) m[k] += e.xyz[A * 3 + k] * e.qt[(int64_t)A * e.n + c];
      for (int k = 0; k < 3; ++k) e.mu[k * e.n + c] = m[k];
      e.f[c] = 2.0 / 3.0 * e.w[c] * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    }
  }
}

// CI vectors to the reference layout X[(i v + a) n + c] (device [n][o][v_p]).
struct UnpackDesc {
  const double* src;
  double* dst;
  int o, v, vp, n;
};
__global__ void __launch_bounds__(256) unpack_grouped_kernel(const GroupSlot* __restrict__ slot) {
  const int nd = slot->count;
  const UnpackDesc* __restrict__ d = (const UnpackDesc*)slot->descs;
  for (int b = blockIdx.x; b < nd; b += gridDim.x) {
    const UnpackDesc e = d[b];
    const int64_t total = (int64_t)e.o * e.v * e.n;
    for (int64_t t = threadIdx.x; t < total; t += 256) {
      const int64_t ia = t / e.n;
      const int c = (int)(t % e.n);
      const int i = (int)(ia / e.v), a = (int)(ia % e.v);
      e.dst[t] = e.src[(int64_t)c * e.o * e.vp + (int64_t)i * e.vp + a];
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int64_t rup(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

GemmArgs gemm_args(const void* A, const void* B, void* C, int64_t M, int64_t N, int64_t K, IdxMap am, IdxMap ak,
                   IdxMap bk, IdxMap bn, IdxMap cm, IdxMap cn) {
  GemmArgs g;
  std::memset(&g, 0, sizeof(g));
  g.A = A; g.B = B; g.C = C;
  g.M = M; g.N = N; g.K = K;
  g.am = am; g.ak = ak; g.bk = bk; g.bn = bn; g.cm = cm; g.cn = cn;
  g.batch = 1;
  g.ksplit = 1;
  g.k_chunk = K;  // one K range per tile (no split-K in grouped launches)
  g.alpha = 1.0;
  g.dm = plain(0); g.dmap = plain(1); g.xm = plain(0); g.xn = plain(0);
  g.hm = plain(0); g.hn = plain(0);
  g.hmod = 1;
  return g;
}

struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  // grows by at least a quarter, so that slowly growing requests do not
  // reallocate every time
  void ensure(size_t b) {
    if (b <= bytes) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    const size_t want = std::max(b, bytes + bytes / 4);
    bytes = 0;
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      TDA_HIP(hipMalloc(&p, b));
      bytes = b;
      return;
    }
    bytes = want;
  }
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};
struct PinMem {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b) {
    if (b <= bytes) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    TDA_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
    bytes = b;
  }
  ~PinMem() {
    if (p) (void)hipHostFree(p);
  }
};

// The work lists of one segment of a step (descriptors, tile lists),
// packed in one pinned buffer and uploaded with one copy.  The device buffer
// starts with MAX_SLOTS slot headers (GroupSlot: absolute pointers + count),
// which the grouped kernels read; the buffer is reallocated only when it must
// grow (gen counts reallocations: captured graphs hold its address).
struct Plan {
  static constexpr int MAX_SLOTS = 32;
  std::vector<char> host;
  PinMem pin;
  DevMem dev;
  uint64_t gen = 0;
  struct Op {
    size_t desc_off, tile_off;
    int count;
  };
  std::vector<Op> ops;
  void reset() {
    host.assign(MAX_SLOTS * sizeof(GroupSlot), 0);
    ops.clear();
  }
  size_t put(const void* p, size_t bytes) {
    const size_t off = rup((int64_t)host.size(), 256);
    host.resize(off + bytes);
    if (bytes) std::memcpy(host.data() + off, p, bytes);
    return off;
  }
  int add(size_t desc_off, size_t tile_off, int count) {
    TDA_CHECK((int)ops.size() < MAX_SLOTS, 2, "batch plan: too many operations in one segment");
    ops.push_back(Op{desc_off, tile_off, count});
    return (int)ops.size() - 1;
  }
  // a grouped GEMM: descriptors + tiles of configuration cfg
  int gemm(const std::vector<GemmArgs>& d, int cfg) {
    std::vector<int4> tiles;
    const int bm = grouped_cfg_bm(cfg), bn = grouped_cfg_bn(cfg);
    for (int i = 0; i < (int)d.size(); ++i) {
      const int tm = (int)((d[i].M + bm - 1) / bm), tn = (int)((d[i].N + bn - 1) / bn);
      for (int z = 0; z < std::max(1, d[i].batch); ++z)
        for (int a = 0; a < tm; ++a)
          for (int c = 0; c < tn; ++c) tiles.push_back(int4{i, z, a, c});
    }
    const size_t doff = put(d.data(), d.size() * sizeof(GemmArgs));
    const size_t toff = put(tiles.data(), tiles.size() * sizeof(int4));
    return add(doff, toff, (int)tiles.size());
  }
  template <typename T>
  int units(const std::vector<T>& u) {
    return add(put(u.data(), u.size() * sizeof(T)), 0, (int)u.size());
  }
  void upload(hipStream_t s) {
    if (host.empty()) reset();
    // the previous upload's copy and the kernels reading its device copy
    // must be done before either buffer is reused
    TDA_HIP(hipStreamSynchronize(s));
    if (dev.bytes < host.size()) {
      dev.ensure(std::max(host.size(), (size_t)64 << 20));
      ++gen;
    }
    GroupSlot* hs = (GroupSlot*)host.data();
    for (size_t i = 0; i < ops.size(); ++i) {
      hs[i].descs = (const char*)dev.p + ops[i].desc_off;
      hs[i].tiles = (const int4*)((const char*)dev.p + ops[i].tile_off);
      hs[i].count = ops[i].count;
    }
    pin.ensure(host.size());
    std::memcpy(pin.p, host.data(), host.size());
    TDA_HIP(hipMemcpyAsync(dev.p, pin.p, host.size(), hipMemcpyHostToDevice, s));
  }
  const GroupSlot* slot(int op) const { return (const GroupSlot*)dev.p + op; }
  int count(int op) const { return ops[op].count; }
};

// A segment of launches (between two host decisions) replayed from a
// captured graph once its launch sequence (the key: tile configurations,
// grids) and the plan buffer are unchanged; the work lists come from the
// uploaded plan.  TINCR_TDA_BATCH_GRAPH=0: direct launches.
struct SegGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<int> key;
  uint64_t gen = 0;
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  ~SegGraph() { drop(); }
};
bool graphs_on() {
  const char* e = getenv("TINCR_TDA_BATCH_GRAPH");
  return !(e && e[0] == '0');
}
template <typename F>
void run_segment(hipStream_t s, SegGraph& sg, const std::vector<int>& key, uint64_t gen, bool use_graph, F&& launches) {
  if (!use_graph) {
    launches();
    return;
  }
  if (!(sg.exec && sg.key == key && sg.gen == gen)) {
    sg.drop();
    hipGraph_t graph = nullptr;
    TDA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      launches();
    } catch (...) {
      (void)hipStreamEndCapture(s, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    TDA_HIP(hipStreamEndCapture(s, &graph));
    const hipError_t e = hipGraphInstantiate(&sg.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    TDA_HIP(e);
    sg.key = key;
    sg.gen = gen;
  }
  TDA_HIP(hipGraphLaunch(sg.exec, s));
}

// operation classes of the grouped GEMMs (tile autotuner key)
enum OpClass : int {
  OC_U = 0, OC_UU, OC_COUL, OC_ST1, OC_ST3, OC_PROJ, OC_RITZ, OC_GS_DOT, OC_GS_SUB, OC_QT, OC_N
};

struct Tuner {
  std::mutex mu;
  std::map<std::pair<int, int>, int> best;  // (op class, precision class) -> cfg
};
Tuner& tuner() {
  static Tuner t;
  return t;
}

int default_cfg(int oc) {
  switch (oc) {
    case OC_U: case OC_UU: case OC_PROJ: case OC_GS_DOT: case OC_QT: return CFG_SKINNY_N;
    case OC_COUL: case OC_RITZ: case OC_GS_SUB: return CFG_SKINNY_M;
    default: return CFG_SMALL;
  }
}

struct GItem {
  BatchItem* in = nullptr;
  int index = 0;
  ChargesDims d{};
  int n = 0, cap = 0;
  int64_t L = 0;
  // device arrays (inside the group's arena)
  double *qov = nullptr, *gamma = nullptr, *G = nullptr, *qvv = nullptr, *de = nullptr;
  float *qov32 = nullptr, *G32 = nullptr, *qvv32 = nullptr;
  double *V = nullptr, *AV = nullptr, *H = nullptr, *Y = nullptr, *work = nullptr, *RZ = nullptr, *R = nullptr,
         *P = nullptr, *W = nullptr, *C0 = nullptr, *u = nullptr, *uu = nullptr, *T = nullptr, *qt = nullptr,
         *xyz = nullptr, *mu = nullptr, *f = nullptr, *xout = nullptr;
  float *x32 = nullptr, *vv32 = nullptr, *T32 = nullptr;
  double *evals = nullptr, *norms = nullptr;  // in the scalars block
  int* count = nullptr;
  size_t out_off = 0;  // of qt | mu | f in the output block (doubles)
  size_t x_off = 0;    // of X in the output block
  // host state (davidson.rs:53-185)
  int m = 0, n_done = 0, dim_sub = 0, dim_sub_origin = 0, it = 0, it_phase = 0, phase = 0;
  int status = 0;  // 0 active, 1 converged, 2 not converged, 3 failed
  bool first_can_converge = false;
  int action = 0, nc = 0, pf = 0, k = 0, pc = 0;  // pc: next column of the product chunks
  std::vector<int> idx;
  std::vector<double> ev, err;
  int n_products = 0, n_products_fp64 = 0, it64 = 0;
};
enum Action : int { A_NONE = 0, A_EXPAND = 1, A_COLLAPSE = 2, A_SWITCH = 3 };

// Device and page-locked buffers of the grouped driver, kept per device for
// the process and reused by later batches (allocating tens of GB of device
// memory or hundreds of MB of page-locked memory costs more than a batch's
// solves).  One batch at a time per device uses them.
struct BatchMem {
  std::mutex mu;
  DevMem arena, scal, outb;
  // exchange intermediates T of one product chunk, per precision (zeroed
  // when allocated: stage 3 reads T's pad rows against zero pads of G, and
  // every later write is a finite value of the same type)
  DevMem t64, t32;
  PinMem scal_h, out_h;
  Plan plan;
};
BatchMem& batch_mem(int device) {
  static std::mutex m;
  static std::map<int, BatchMem*> mems;
  std::lock_guard<std::mutex> lk(m);
  BatchMem*& p = mems[device];
  if (!p) p = new BatchMem();  // lives until the process ends
  return *p;
}

struct Group {
  int device = 0;
  bool mixed = false;
  hipStream_t s = nullptr;
  std::vector<GItem*> items;
  DevMem& arena;
  DevMem& scal;
  DevMem& outb;
  PinMem& scal_h;
  PinMem& out_h;
  Plan& plan;
  DevMem& t64;
  DevMem& t32;
  size_t scal_stride = 0;
  // exchange kernels, [stage][class] as in tda_stats: events of the grouped
  // launches (read after the group's last synchronisation), flops
  struct Rec {
    hipEvent_t a, b;
    int stage, pclass;
    double flops;
  };
  std::vector<Rec> recs;
  tda_stats* totals = nullptr;
  bool graphs = false;
  SegGraph g_prod[2], g_sub, g_exp;
  explicit Group(BatchMem& m)
      : arena(m.arena), scal(m.scal), outb(m.outb), scal_h(m.scal_h), out_h(m.out_h), plan(m.plan), t64(m.t64),
        t32(m.t32) {}
};

// bytes of T for k vectors of a fragment
size_t t_bytes(const GItem& it, int k, size_t esz) {
  return rup((int64_t)k * it.d.vp * it.d.NL * it.d.op * esz, 256);
}
// pool of T: the largest fragment's n vectors, at least min(budget, all of
// them); budget TINCR_TDA_BATCH_TPOOL_GB (default 16)
void ensure_t_pool(DevMem& pool, const std::vector<GItem*>& items, size_t esz, hipStream_t s) {
  size_t mx = 0, all = 0;
  for (GItem* it : items) {
    mx = std::max(mx, t_bytes(*it, it->n, esz));
    all += t_bytes(*it, it->n, esz);
  }
  const char* e = getenv("TINCR_TDA_BATCH_TPOOL_GB");
  const double gb = e && atof(e) > 0 ? atof(e) : 16.0;
  const size_t want = std::max(mx, std::min(all, (size_t)(gb * (1 << 30))));
  if (want <= pool.bytes) return;
  pool.ensure(want);
  TDA_HIP(hipMemsetAsync(pool.p, 0, pool.bytes, s));
}

// per-fragment device footprint (doubles, 256-byte aligned pieces)
struct Footprint {
  size_t bytes = 0;
  size_t take(size_t b) {
    const size_t off = rup((int64_t)bytes, 256);
    bytes = off + b;
    return off;
  }
};

void layout_item(GItem& it, bool mixed, Footprint& fp, std::vector<size_t>& offs) {
  const ChargesDims& d = it.d;
  const size_t L = it.L, N = d.N, NL = d.NL, op = d.op, vp = d.vp, n = it.n, cap = it.cap;
  offs.clear();
  offs.push_back(fp.take(N * L * 8));          // 0 qov
  offs.push_back(fp.take(N * N * 8));          // 1 gamma
  offs.push_back(fp.take(NL * op * op * 8));   // 2 G
  offs.push_back(fp.take(NL * vp * vp * 8));   // 3 qvv
  offs.push_back(fp.take(L * 8));              // 4 de
  offs.push_back(fp.take(cap * L * 8));        // 5 V
  offs.push_back(fp.take(cap * L * 8));        // 6 AV
  offs.push_back(fp.take(cap * cap * 8));      // 7 H
  offs.push_back(fp.take(n * cap * 8));        // 8 Y
  offs.push_back(fp.take((cap * cap + cap) * 8));  // 9 eigensolver work
  offs.push_back(fp.take(n * L * 8));          // 10 RZ
  offs.push_back(fp.take(n * L * 8));          // 11 R
  offs.push_back(fp.take(n * L * 8));          // 12 P
  offs.push_back(fp.take(n * L * 8));          // 13 W
  offs.push_back(fp.take(cap * n * 8));        // 14 C0
  offs.push_back(fp.take(N * n * 8));          // 15 u
  offs.push_back(fp.take(N * n * 8));          // 16 uu
  offs.push_back(fp.take(8));                  // 17 (unused: T lives in the step's pool)
  offs.push_back(fp.take(N * 3 * 8));          // 18 xyz
  if (mixed) {
    offs.push_back(fp.take(N * L * 4));        // 19 qov32
    offs.push_back(fp.take(NL * op * op * 4)); // 20 G32
    offs.push_back(fp.take(NL * vp * vp * 4)); // 21 qvv32
    offs.push_back(fp.take(n * L * 4));        // 22 x32
    offs.push_back(fp.take(N * n * 4));        // 23 vv32
  }
}

void bind_item(GItem& it, bool mixed, char* base, const std::vector<size_t>& o) {
  it.qov = (double*)(base + o[0]);
  it.gamma = (double*)(base + o[1]);
  it.G = (double*)(base + o[2]);
  it.qvv = (double*)(base + o[3]);
  it.de = (double*)(base + o[4]);
  it.V = (double*)(base + o[5]);
  it.AV = (double*)(base + o[6]);
  it.H = (double*)(base + o[7]);
  it.Y = (double*)(base + o[8]);
  it.work = (double*)(base + o[9]);
  it.RZ = (double*)(base + o[10]);
  it.R = (double*)(base + o[11]);
  it.P = (double*)(base + o[12]);
  it.W = (double*)(base + o[13]);
  it.C0 = (double*)(base + o[14]);
  it.u = (double*)(base + o[15]);
  it.uu = (double*)(base + o[16]);
  it.xyz = (double*)(base + o[18]);
  if (mixed) {
    it.qov32 = (float*)(base + o[19]);
    it.G32 = (float*)(base + o[20]);
    it.qvv32 = (float*)(base + o[21]);
    it.x32 = (float*)(base + o[22]);
    it.vv32 = (float*)(base + o[23]);
  }
}

// The n lowest orbital-energy differences (solvers/utils.rs:17-26, ties by
// index), as positions in the padded layout.
std::vector<int64_t> guess_positions(const GItem& it) {
  const BatchItem& b = *it.in;
  const int o = b.o, v = b.v, vp = it.d.vp;
  const double* eo = b.orbe + b.ob;
  const double* ev = (b.ct ? b.orbe_l : b.orbe) + b.vb;
  std::vector<int64_t> order((size_t)o * v);
  std::iota(order.begin(), order.end(), 0);
  auto de = [&](int64_t x) { return ev[x % v] - eo[x / v]; };
  std::partial_sort(order.begin(), order.begin() + it.n, order.end(), [&](int64_t a, int64_t c) {
    const double da = de(a), dc = de(c);
    return da < dc || (da == dc && a < c);
  });
  std::vector<int64_t> pos(it.n);
  for (int r = 0; r < it.n; ++r) pos[r] = (order[r] / v) * vp + order[r] % v;
  return pos;
}

// Prepares every fragment's charges-form operator with worker contexts and
// copies it into the arena.
void prepare_operators(Group& g, int n_workers, size_t budget) {
  std::atomic<int> next{0};
  std::mutex mu;
  std::string first_err;
  auto worker = [&]() {
    tda_ctx* c = nullptr;
    int rc = tda_ctx_create(g.device, budget, TDA_PREC_FP64, &c);
    if (rc == TDA_OK) rc = tda_set_operator(c, TDA_OP_CHARGES);
    if (rc != TDA_OK) {
      std::lock_guard<std::mutex> lk(mu);
      if (first_err.empty()) first_err = "batch: context creation failed";
      if (c) tda_ctx_destroy(c);
      return;
    }
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= (int)g.items.size()) break;
      GItem& it = *g.items[k];
      const BatchItem& b = *it.in;
      if (b.ct)
        rc = tda_set_ct_pair(c, b.n_atoms, b.n_ao, b.opa, b.orbs, b.s, b.orbe, b.ob, b.o, b.n_atoms_l, b.n_ao_l,
                             b.opa_l, b.orbs_l, b.s_l, b.orbe_l, b.vb, b.v, b.s_hl, b.gamma, b.gamma_l, b.gamma_hl,
                             b.gamma_lr, b.xyz);
      else
        rc = tda_set_system(c, b.n_atoms, b.n_ao, b.opa, b.orbs, b.s, b.orbe, b.ob, b.o, b.vb, b.v, b.gamma,
                            b.gamma_lr, b.xyz);
      ChargesDims d{};
      if (rc == TDA_OK) rc = ctx_charges_dims(c, &d);
      if (rc == TDA_OK && (d.N != it.d.N || d.NL != it.d.NL || d.op != it.d.op || d.vp != it.d.vp))
        rc = TDA_ERR_STATE;
      if (rc == TDA_OK) rc = ctx_export_charges(c, it.qov, it.gamma, it.G, it.qvv, it.de);
      if (rc != TDA_OK) {
        it.status = 3;
        std::lock_guard<std::mutex> lk(mu);
        if (first_err.empty()) first_err = "fragment " + std::to_string(it.index) + ": " + ctx_error(c);
      }
    }
    tda_ctx_destroy(c);
  };
  const int nw = std::max(1, std::min(n_workers, (int)g.items.size()));
  std::vector<std::thread> pool;
  for (int w = 0; w < nw; ++w) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  if (!first_err.empty()) throw Error{TDA_ERR_HIP, first_err};
}

// One grouped launch of a plan op.  With graphs every launch is made (empty
// work lists included) on a fixed grid, so a captured segment fits any step;
// otherwise empty ops are skipped and small ones get a smaller grid.
int vec_grid(const Group& g, int op) {
  return g.graphs ? GRID_VEC : std::max(1, std::min(GRID_VEC, g.plan.count(op)));
}
void run_gemm_op(Group& g, int op, DType in, DType out, Epi epi, bool a_kc, bool b_kc, int cfg) {
  if (!g.graphs && g.plan.count(op) == 0) return;
  grouped_gemm_launch(in, out, epi, a_kc, b_kc, cfg, g.plan.slot(op), GRID_GEMM, g.s);
}
template <int OP>
void run_vec_op(Group& g, int op) {
  if (!g.graphs && g.plan.count(op) == 0) return;
  vec_kernel<OP><<<vec_grid(g, op), 256, 0, g.s>>>(g.plan.slot(op));
  TDA_LAUNCH_CHECK();
}

// Grouped GEMM with the tile configuration of its operation class: on the
// first launch of a class (per precision class) with enough work, each
// candidate runs once on a copy of the descriptors writing a scratch output
// (beta = 0), timed, and the fastest is kept for the process.
struct GemmSpec {
  DType in, out;
  Epi epi;
  bool a_kc, b_kc;
  int oc, pclass;
};

// the tuned configuration of an operation class, or -1
int cached_cfg(const GemmSpec& sp) {
  std::lock_guard<std::mutex> lk(tuner().mu);
  auto f = tuner().best.find({sp.oc, sp.pclass});
  return f == tuner().best.end() ? -1 : f->second;
}

int pick_cfg(Group& g, const GemmSpec& sp, const std::vector<GemmArgs>& d, double flops) {
  const std::pair<int, int> key{sp.oc, sp.pclass};
  {
    std::lock_guard<std::mutex> lk(tuner().mu);
    auto f = tuner().best.find(key);
    if (f != tuner().best.end()) return f->second;
  }
  const int def = default_cfg(sp.oc);
  const char* e = getenv("TINCR_TDA_AUTOTUNE");
  if ((e && e[0] == '0') || flops < 2e9 || d.empty()) return def;
  // scratch outputs: one region per descriptor, sized by its C extent
  std::vector<GemmArgs> t = d;
  size_t total = 0;
  std::vector<size_t> off(t.size());
  const size_t esz = sp.out == DT_F64 ? 8 : 4;
  for (size_t i = 0; i < t.size(); ++i) {
    // extent of C: the largest offset the maps reach, plus one
    const GemmArgs& a = t[i];
    const int64_t ext = a.cm(a.M - 1) + a.cn(a.N - 1) + (int64_t)(std::max(1, a.batch) - 1) * a.c_batch + 1;
    off[i] = total;
    total += rup(ext * esz, 256);
  }
  DevMem scratch;
  scratch.ensure(total);
  for (size_t i = 0; i < t.size(); ++i) {
    t[i].C = (char*)scratch.p + off[i];
    t[i].beta = 0.0;
  }
  const int cands[3] = {CFG_SMALL, CFG_SKINNY_N, CFG_SKINNY_M};
  hipEvent_t ev[6];
  for (auto& x : ev) TDA_HIP(hipEventCreate(&x));
  float ms[3];
  Plan tp;
  tp.reset();
  std::vector<int> ops(3);
  for (int c = 0; c < 3; ++c) ops[c] = tp.gemm(t, cands[c]);
  tp.upload(g.s);
  for (int c = 0; c < 3; ++c) {
    auto launch = [&]() {
      grouped_gemm_launch(sp.in, sp.out, sp.epi, sp.a_kc, sp.b_kc, cands[c], tp.slot(ops[c]), GRID_GEMM, g.s);
    };
    launch();  // warm-up
    TDA_HIP(hipEventRecord(ev[2 * c], g.s));
    launch();
    TDA_HIP(hipEventRecord(ev[2 * c + 1], g.s));
  }
  TDA_HIP(hipEventSynchronize(ev[5]));
  int best = 0;
  for (int c = 0; c < 3; ++c) {
    TDA_HIP(hipEventElapsedTime(&ms[c], ev[2 * c], ev[2 * c + 1]));
    if (ms[c] < ms[best]) best = c;
  }
  for (auto& x : ev) (void)hipEventDestroy(x);
  std::lock_guard<std::mutex> lk(tuner().mu);
  tuner().best[key] = cands[best];
  return cands[best];
}

// A.x for the columns [pc, pc + k) (k <= n: the intermediates hold n
// vectors) of the fragments in `its`, products in FP64 (pclass 0) or FP32
// (pclass 1) (solvers/mod.rs:280-369, charges form).
void products_chunk(Group& g, const std::vector<GItem*>& its, bool fp32);
void products(Group& g, const std::vector<GItem*>& its, bool fp32) {
  if (its.empty()) return;
  // fragments in chunks whose intermediates T fit the pool
  DevMem& pool = fp32 ? g.t32 : g.t64;
  const size_t esz = fp32 ? 4 : 8;
  std::vector<GItem*> chunk;
  size_t used = 0;
  for (size_t i = 0; i <= its.size(); ++i) {
    const size_t tb = i < its.size() ? t_bytes(*its[i], its[i]->k, esz) : 0;
    if (i == its.size() || (!chunk.empty() && used + tb > pool.bytes)) {
      products_chunk(g, chunk, fp32);
      chunk.clear();
      used = 0;
      if (i == its.size()) break;
    }
    GItem* it = its[i];
    if (fp32) it->T32 = (float*)((char*)pool.p + used);
    else it->T = (double*)((char*)pool.p + used);
    used += tb;
    chunk.push_back(it);
  }
}

void products_chunk(Group& g, const std::vector<GItem*>& its, bool fp32) {
  if (its.empty()) return;
  const DType dt = fp32 ? DT_F32 : DT_F64;
  const int pc = fp32 ? 1 : 0;
  std::vector<VecUnit> cast_x;
  std::vector<GemmArgs> du, duu, dcoul, ds1, ds3;
  std::vector<VecUnit> cast_uu;
  double f1 = 0, f3 = 0;
  for (GItem* it : its) {
    const ChargesDims& d = it->d;
    const int k = it->k;
    const int64_t L = it->L, N = d.N;
    double* X = it->V + (size_t)it->pc * L;
    double* AX = it->AV + (size_t)it->pc * L;
    const void* Xin = X;
    if (fp32) {
      VecUnit c{};
      c.src = X;
      c.dst32 = it->x32;
      c.L = (int64_t)k * L;
      cast_x.push_back(c);
      Xin = it->x32;
    }
    // u = q_ov X  [N][k]
    du.push_back(gemm_args(fp32 ? (const void*)it->qov32 : it->qov, Xin, it->u, N, k, L, plain(L), plain(1), plain(1),
                           plain(L), plain(k), plain(1)));
    // uu = gamma u
    duu.push_back(gemm_args(it->gamma, it->u, it->uu, N, k, N, plain(N), plain(1), plain(k), plain(1), plain(k),
                            plain(1)));
    if (fp32) {
      VecUnit c{};
      c.src = it->uu;
      c.dst32 = it->vv32;
      c.L = N * k;
      cast_uu.push_back(c);
    }
    // AX = de o X + 2 q_ov^T uu   (solvers/mod.rs:307-311)
    GemmArgs gc = gemm_args(fp32 ? (const void*)it->vv32 : it->uu, fp32 ? (const void*)it->qov32 : it->qov, AX, k, L,
                            N, plain(1), plain(k), plain(L), plain(1), plain(L), plain(1));
    gc.alpha = 2.0;
    gc.diag = it->de;
    gc.dmap = plain(1);
    gc.xsrc = X;
    gc.xm = plain(L);
    gc.xn = plain(1);
    dcoul.push_back(gc);
    if (d.lc && d.NL > 0) {
      const int o = d.o, v = d.v, op = d.op, vp = d.vp, NL = d.NL;
      const int64_t NBop = (int64_t)NL * op;
      // stage 1: T[c][a][(B,j)] = sum_b q_vv[B][a][b] X[c][j][b]   (solvers/mod.rs:335-350)
      GemmArgs g1 = gemm_args(fp32 ? (const void*)it->qvv32 : it->qvv, Xin, fp32 ? (void*)it->T32 : it->T, v,
                              (int64_t)k * o, vp, plain(vp), plain(1), plain(1), split(o, L, vp), plain(NBop),
                              split(o, (int64_t)vp * NBop, 1));
      g1.batch = NL;
      g1.a_batch = (int64_t)vp * vp;
      g1.b_batch = 0;
      g1.c_batch = op;
      ds1.push_back(g1);
      // stage 3: AX[c][i][a] -= sum_(B,j) G[B][i][j] T[c][a][(B,j)]   (:351-360)
      GemmArgs g3 = gemm_args(fp32 ? (const void*)it->G32 : it->G, fp32 ? (const void*)it->T32 : it->T, AX, o,
                              (int64_t)k * v, NBop, plain(op), split(op, (int64_t)op * op, 1), plain(1),
                              split(v, (int64_t)vp * NBop, NBop), plain(vp), split(v, L, 1));
      g3.alpha = -1.0;
      g3.beta = 1.0;
      ds3.push_back(g3);
      f1 += 2.0 * k * (double)o * v * v * NL;
      f3 += 2.0 * k * (double)o * o * v * NL;
    }
    it->n_products += k;
    if (!fp32) it->n_products_fp64 += k;
  }
  const GemmSpec s_u{dt, DT_F64, EPI_STORE, true, true, OC_U, pc};
  const GemmSpec s_uu{DT_F64, DT_F64, EPI_STORE, true, false, OC_UU, 0};
  const GemmSpec s_c{dt, DT_F64, EPI_DIAG, false, false, OC_COUL, pc};
  const GemmSpec s_1{dt, dt, EPI_STORE, true, true, OC_ST1, pc};
  const GemmSpec s_3{dt, DT_F64, EPI_STORE, true, true, OC_ST3, pc};
  // tile configurations (the exchange stages tuned on their first use)
  const int c1 = ds1.empty() ? CFG_SMALL : pick_cfg(g, s_1, ds1, f1);
  // stage 3 accumulates into AX: its tuning trial (scratch outputs) needs
  // the T of this chunk, so an untuned stage 3 runs from a second upload
  const int c3 = ds3.empty() ? CFG_SMALL : cached_cfg(s_3);
  g.plan.reset();
  const int o_cx = g.plan.units(cast_x);
  const int o_u = g.plan.gemm(du, default_cfg(OC_U));
  const int o_uu = g.plan.gemm(duu, default_cfg(OC_UU));
  const int o_cu = g.plan.units(cast_uu);
  const int o_c = g.plan.gemm(dcoul, default_cfg(OC_COUL));
  const int o_1 = g.plan.gemm(ds1, c1);
  const int o_3 = c3 >= 0 ? g.plan.gemm(ds3, c3) : -1;
  g.plan.upload(g.s);
  // the whole product step is timed (device events around it): its
  // algorithmic flops are the Coulomb and exchange contractions
  double fc = 0;
  for (GItem* it : its) fc += 4.0 * it->k * (double)it->d.N * it->d.o * it->d.v + 2.0 * it->k * (double)it->d.N * it->d.N;
  Group::Rec r{nullptr, nullptr, 0, pc, fc + f1 + f3};
  TDA_HIP(hipEventCreate(&r.a));
  TDA_HIP(hipEventCreate(&r.b));
  TDA_HIP(hipEventRecord(r.a, g.s));
  auto launches = [&](bool with3) {
    run_vec_op<VO_CAST>(g, o_cx);
    run_gemm_op(g, o_u, s_u.in, s_u.out, s_u.epi, s_u.a_kc, s_u.b_kc, default_cfg(OC_U));
    run_gemm_op(g, o_uu, s_uu.in, s_uu.out, s_uu.epi, s_uu.a_kc, s_uu.b_kc, default_cfg(OC_UU));
    run_vec_op<VO_CAST>(g, o_cu);
    run_gemm_op(g, o_c, s_c.in, s_c.out, s_c.epi, s_c.a_kc, s_c.b_kc, default_cfg(OC_COUL));
    run_gemm_op(g, o_1, s_1.in, s_1.out, s_1.epi, s_1.a_kc, s_1.b_kc, c1);
    if (with3) run_gemm_op(g, o_3, s_3.in, s_3.out, s_3.epi, s_3.a_kc, s_3.b_kc, c3);
  };
  if (o_3 >= 0 || ds3.empty()) {
    // tuned (or no exchange): the segment is replayed from a captured graph
    const bool with3 = o_3 >= 0;
    run_segment(g.s, g.g_prod[pc], {c1, c3, with3 ? 1 : 0}, g.plan.gen, g.graphs, [&] { launches(with3); });
  } else {
    const bool gr = g.graphs;
    g.graphs = false;
    launches(false);
    const int c3t = pick_cfg(g, s_3, ds3, f3);
    Plan& p = g.plan;
    p.reset();
    const int o_3t = p.gemm(ds3, c3t);
    p.upload(g.s);
    run_gemm_op(g, o_3t, s_3.in, s_3.out, s_3.epi, s_3.a_kc, s_3.b_kc, c3t);
    g.graphs = gr;
  }
  TDA_HIP(hipEventRecord(r.b, g.s));
  g.recs.push_back(r);
}

void set_status(const GItem& it, int32_t* status, tda_stats* stats, bool mixed) {
  const int idx = it.index;
  status[idx] = it.status == 1 ? TDA_OK : it.status == 2 ? TDA_ERR_NOT_CONVERGED : TDA_ERR_HIP;
  if (stats) {
    tda_stats& st = stats[idx];
    std::memset(&st, 0, sizeof(st));
    st.n_iter = it.it;
    st.n_iter_fp64 = mixed ? it.it64 : it.it;
    st.n_iter_fp32 = mixed ? it.it - it.it64 : 0;
    st.n_products = it.n_products;
    st.n_products_fp64 = it.n_products_fp64;
    st.n_products_fp32 = it.n_products - it.n_products_fp64;
    st.converged = it.status == 1;
    st.n_atom_blocks = 1;
    double mx = 0.0;
    for (double e : it.err) mx = std::max(mx, e);
    st.max_residual = mx;
  }
}

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

// The grouped Davidson of one group of fragments (all fit in one arena).
void run_group(Group& g, int n_workers, size_t budget, int32_t* status, tda_stats* stats) {
  const bool mixed = g.mixed;
  const char* ev = getenv("TINCR_TDA_VERBOSE");
  const bool verbose = ev && ev[0] && ev[0] != '0';
  double t_prep = 0, t_prod = 0, t_sub = 0, t_exp = 0, t_out = 0;
  int steps = 0;
  auto t_start = Clock::now();
  // ---- arena ----
  Footprint fp;
  std::vector<std::vector<size_t>> offs(g.items.size());
  for (size_t i = 0; i < g.items.size(); ++i) layout_item(*g.items[i], mixed, fp, offs[i]);
  g.arena.ensure(fp.bytes);
  // no clearing: every array is written in full before it is read (the
  // exported operators carry their zero pads, products and Gram-Schmidt keep
  // the pads of V, AV, RZ, R zero)
  for (size_t i = 0; i < g.items.size(); ++i) bind_item(*g.items[i], mixed, (char*)g.arena.p, offs[i]);
  ensure_t_pool(g.t64, g.items, 8, g.s);
  if (mixed) ensure_t_pool(g.t32, g.items, 4, g.s);
  // scalars: per fragment [cap] Ritz values, [n] residual norms, 1 count
  int maxcap = 0;
  for (GItem* it : g.items) maxcap = std::max(maxcap, it->cap + it->n + 1);
  g.scal_stride = (size_t)maxcap;
  g.scal.ensure(g.items.size() * g.scal_stride * 8);
  g.scal_h.ensure(g.items.size() * g.scal_stride * 8);
  for (size_t i = 0; i < g.items.size(); ++i) {
    GItem& it = *g.items[i];
    double* sb = (double*)g.scal.p + i * g.scal_stride;
    it.evals = sb;
    it.norms = sb + it.cap;
    it.count = (int*)(sb + it.cap + it.n);
  }
  // xyz: packed on the host, copied per fragment, one synchronisation
  {
    size_t tot = 0;
    for (GItem* it : g.items) tot += (size_t)it->d.N * 3;
    std::vector<double> h(tot, 0.0);
    size_t off = 0;
    for (GItem* it : g.items) {
      const size_t N = it->d.N;
      if (it->in->xyz) std::memcpy(h.data() + off, it->in->xyz, N * 3 * 8);
      TDA_HIP(hipMemcpyAsync(it->xyz, h.data() + off, N * 3 * 8, hipMemcpyHostToDevice, g.s));
      off += N * 3;
    }
    TDA_HIP(hipStreamSynchronize(g.s));
  }
  // ---- operators ----
  prepare_operators(g, n_workers, budget);
  t_prep = ms_since(t_start);
  if (mixed) {  // FP32 copies of q_ov, G, q_vv (one grouped launch)
    std::vector<VecUnit> cu;
    for (GItem* it : g.items) {
      if (it->status != 0) continue;
      const ChargesDims& d = it->d;
      VecUnit u{};
      u.src = it->qov;
      u.dst32 = it->qov32;
      u.L = (int64_t)d.N * it->L;
      cu.push_back(u);
      if (d.NL > 0) {
        u.src = it->G;
        u.dst32 = it->G32;
        u.L = (int64_t)d.NL * d.op * d.op;
        cu.push_back(u);
        u.src = it->qvv;
        u.dst32 = it->qvv32;
        u.L = (int64_t)d.NL * d.vp * d.vp;
        cu.push_back(u);
      }
    }
    g.plan.reset();
    const int o = g.plan.units(cu);
    g.plan.upload(g.s);
    run_vec_op<VO_CAST>(g, o);
  }
  // ---- initial subspace: unit vectors at the n lowest de ----
  {
    std::vector<VecUnit> gu;
    for (GItem* it : g.items) {
      if (it->status != 0) continue;
      const std::vector<int64_t> pos = guess_positions(*it);
      for (int r = 0; r < it->n; ++r) {
        VecUnit u{};
        u.dst = it->V + (size_t)r * it->L;
        u.L = it->L;
        u.pos = pos[r];
        gu.push_back(u);
      }
      it->m = it->dim_sub = it->dim_sub_origin = it->n;
      it->phase = mixed ? 0 : 1;
    }
    g.plan.reset();
    const int o = g.plan.units(gu);
    g.plan.upload(g.s);
    run_vec_op<VO_GUESS>(g, o);
  }
  // ---- Davidson steps ----
  for (;;) {
    std::vector<GItem*> act, a32, a64;
    for (GItem* it : g.items)
      if (it->status == 0) act.push_back(it);
    if (act.empty()) break;
    ++steps;
    auto t0 = Clock::now();
    // products for the new columns (cache rule solvers/mod.rs:290-302: all of
    // them when the basis did not grow), in chunks of at most n vectors (the
    // intermediates u, T hold n)
    for (GItem* it : act) {
      if (it->m <= it->n_done) it->n_done = 0;
      it->pf = it->n_done;
      it->pc = it->n_done;
    }
    for (;;) {
      a32.clear();
      a64.clear();
      for (GItem* it : act)
        if (it->pc < it->m) {
          it->k = std::min(it->n, it->m - it->pc);
          (it->phase == 0 ? a32 : a64).push_back(it);
        }
      if (a32.empty() && a64.empty()) break;
      products(g, a32, true);
      products(g, a64, false);
      for (GItem* it : a32) it->pc += it->k;
      for (GItem* it : a64) it->pc += it->k;
    }
    if (verbose) {
      TDA_HIP(hipStreamSynchronize(g.s));
      t_prod += ms_since(t0);
      t0 = Clock::now();
    }
    // projection, eigenproblem, Ritz vectors, residuals, norms (davidson.rs:89-107)
    {
      std::vector<GemmArgs> dp, dz, dr;
      std::vector<EighDesc> de;
      std::vector<VecUnit> dres;
      for (GItem* it : act) {
        const int64_t L = it->L;
        const int m = it->m, cap = it->cap, n = it->n;
        dp.push_back(gemm_args(it->V, it->AV + (size_t)it->pf * L, it->H + it->pf, m, m - it->pf, L, plain(L),
                               plain(1), plain(1), plain(L), plain(cap), plain(1)));
        EighDesc e{it->H, it->work, it->evals, it->Y, cap, m, n, cap};
        de.push_back(e);
        dz.push_back(gemm_args(it->Y, it->V, it->RZ, n, L, m, plain(cap), plain(1), plain(L), plain(1), plain(L),
                               plain(1)));
        dr.push_back(gemm_args(it->Y, it->AV, it->R, n, L, m, plain(cap), plain(1), plain(L), plain(1), plain(L),
                               plain(1)));
        for (int r = 0; r < n; ++r) {
          VecUnit u{};
          u.dst = it->R + (size_t)r * L;
          u.aux = it->RZ + (size_t)r * L;
          u.wptr = it->evals + r;
          u.out = it->norms + r;
          u.L = L;
          dres.push_back(u);
        }
        it->n_done = m;
      }
      Plan& p = g.plan;
      p.reset();
      const int o_p = p.gemm(dp, default_cfg(OC_PROJ));
      const int o_e = p.units(de);
      const int o_z = p.gemm(dz, default_cfg(OC_RITZ));
      const int o_r = p.gemm(dr, default_cfg(OC_RITZ));
      const int o_res = p.units(dres);
      p.upload(g.s);
      // the eigensolver's grid: every fragment of the group (graphs), or the active ones
      const int egrid = g.graphs ? (int)g.items.size() : (int)de.size();
      run_segment(g.s, g.g_sub, {egrid}, p.gen, g.graphs, [&] {
        run_gemm_op(g, o_p, DT_F64, DT_F64, EPI_STORE, true, true, default_cfg(OC_PROJ));
        launch_eigh_batched(p.slot(o_e), egrid, g.s);
        run_gemm_op(g, o_z, DT_F64, DT_F64, EPI_STORE, true, false, default_cfg(OC_RITZ));
        run_gemm_op(g, o_r, DT_F64, DT_F64, EPI_STORE, true, false, default_cfg(OC_RITZ));
        run_vec_op<VO_RESIDUAL>(g, o_res);
      });
      TDA_HIP(hipMemcpyAsync(g.scal_h.p, g.scal.p, g.items.size() * g.scal_stride * 8, hipMemcpyDeviceToHost, g.s));
      TDA_HIP(hipStreamSynchronize(g.s));
    }
    t_sub += ms_since(t0);
    t0 = Clock::now();
    // per-fragment decisions (davidson.rs:111-176)
    for (size_t i = 0; i < g.items.size(); ++i) {
      GItem& it = *g.items[i];
      if (it.status != 0) continue;
      const double* sb = (const double*)g.scal_h.p + i * g.scal_stride;
      it.ev.assign(sb, sb + it.n);
      it.err.assign(sb + it.cap, sb + it.cap + it.n);
      ++it.it;
      ++it.it_phase;
      if (it.phase == 1 && mixed) ++it.it64;
      bool finite = true;
      int cvd = 0;
      for (int r = 0; r < it.n; ++r) {
        finite = finite && std::isfinite(it.ev[r]) && std::isfinite(it.err[r]);
        cvd += it.err[r] < it.in->tol;
      }
      it.action = A_NONE;
      if (!finite) {
        it.status = 3;
        continue;
      }
      const int lft = it.n - cvd;
      if (lft == 0 && (it.it_phase > 1 || it.first_can_converge)) {
        if (it.phase == 0) it.action = A_SWITCH;  // FP32 phase converged: FP64 verification
        else it.status = 1;
        continue;
      }
      if (it.it >= it.in->max_iter) {
        it.status = 2;
        continue;
      }
      if (it.dim_sub + lft <= it.cap) {
        it.idx.clear();
        for (int r = 0; r < it.n; ++r)
          if (it.err[r] > it.in->tol) it.idx.push_back(r);
        it.dim_sub += lft;
        it.nc = (int)it.idx.size();
        it.action = it.nc > 0 ? A_EXPAND : A_NONE;
      } else {
        it.action = A_COLLAPSE;
      }
    }
    // expansion / collapse / phase switch
    {
      std::vector<VecUnit> dpre, dcp;
      std::vector<GemmArgs> dc0, dsub, dc0b, dsubb;
      std::vector<GsDesc> dgs;
      std::vector<GItem*> grown;
      for (GItem* it : g.items) {
        if (it->status != 0 || it->action == A_NONE) continue;
        const int64_t L = it->L;
        if (it->action == A_COLLAPSE) {
          // guess = the dim_sub_origin lowest Ritz vectors (davidson.rs:172-176)
          VecUnit u{};
          u.dst = it->V;
          u.src = it->RZ;
          u.L = (int64_t)it->dim_sub_origin * L;
          dcp.push_back(u);
          it->m = it->dim_sub = it->dim_sub_origin;
          it->n_done = 0;
          continue;
        }
        if (it->action == A_EXPAND) {
          for (int j = 0; j < it->nc; ++j) {
            VecUnit u{};
            u.dst = it->P + (size_t)j * L;
            u.src = it->R + (size_t)it->idx[j] * L;
            u.aux = it->de;
            u.wptr = it->evals + it->idx[j];
            u.L = L;
            dpre.push_back(u);
          }
          VecUnit c{};
          c.dst = it->W;
          c.src = it->P;
          c.L = (int64_t)it->nc * L;
          dcp.push_back(c);
          const int m = it->m, nc = it->nc;
          dc0.push_back(gemm_args(it->V, it->P, it->C0, m, nc, L, plain(L), plain(1), plain(1), plain(L), plain(nc),
                                  plain(1)));
          GemmArgs s1 = gemm_args(it->C0, it->V, it->W, nc, L, m, plain(1), plain(nc), plain(L), plain(1), plain(L),
                                  plain(1));
          s1.alpha = -1.0;
          s1.beta = 1.0;
          dsub.push_back(s1);
          dc0b.push_back(gemm_args(it->V, it->W, it->C0, m, nc, L, plain(L), plain(1), plain(1), plain(L),
                                   plain(nc), plain(1)));
          dsubb.push_back(s1);
          GsDesc gd{it->W, it->V + (size_t)m * L, it->count, L, nc};
          dgs.push_back(gd);
        } else {  // A_SWITCH: the n Ritz vectors, orthonormalised in FP64, start the FP64 phase
          VecUnit c{};
          c.dst = it->W;
          c.src = it->RZ;
          c.L = (int64_t)it->n * L;
          dcp.push_back(c);
          GsDesc gd{it->W, it->V, it->count, L, it->n};
          dgs.push_back(gd);
        }
        grown.push_back(it);
      }
      Plan& p = g.plan;
      p.reset();
      const int o_pre = p.units(dpre);
      const int o_cp = p.units(dcp);
      const int o_c0 = p.gemm(dc0, default_cfg(OC_GS_DOT));
      const int o_s = p.gemm(dsub, default_cfg(OC_GS_SUB));
      const int o_c0b = p.gemm(dc0b, default_cfg(OC_GS_DOT));
      const int o_sb = p.gemm(dsubb, default_cfg(OC_GS_SUB));
      const int o_gs = p.units(dgs);
      p.upload(g.s);
      run_segment(g.s, g.g_exp, {0}, p.gen, g.graphs, [&] {
        run_vec_op<VO_PRECOND>(g, o_pre);
        run_vec_op<VO_COPY>(g, o_cp);
        // two classical passes against the basis: W = P - V (V^T P), then again on W
        run_gemm_op(g, o_c0, DT_F64, DT_F64, EPI_STORE, true, true, default_cfg(OC_GS_DOT));
        run_gemm_op(g, o_s, DT_F64, DT_F64, EPI_STORE, false, false, default_cfg(OC_GS_SUB));
        run_gemm_op(g, o_c0b, DT_F64, DT_F64, EPI_STORE, true, true, default_cfg(OC_GS_DOT));
        run_gemm_op(g, o_sb, DT_F64, DT_F64, EPI_STORE, false, false, default_cfg(OC_GS_SUB));
        if (g.graphs || p.count(o_gs) > 0) {
          gs_inblock_kernel<<<vec_grid(g, o_gs), 256, 0, g.s>>>(p.slot(o_gs));
          TDA_LAUNCH_CHECK();
        }
      });
      if (!grown.empty()) {
        TDA_HIP(hipMemcpyAsync(g.scal_h.p, g.scal.p, g.items.size() * g.scal_stride * 8, hipMemcpyDeviceToHost,
                               g.s));
      }
      TDA_HIP(hipStreamSynchronize(g.s));
      for (size_t i = 0; i < g.items.size(); ++i) {
        GItem& it = *g.items[i];
        if (it.status != 0 || (it.action != A_EXPAND && it.action != A_SWITCH)) continue;
        const double* sb = (const double*)g.scal_h.p + i * g.scal_stride;
        const int cnt = *(const int*)(sb + it.cap + it.n);
        if (it.action == A_EXPAND) {
          it.m += cnt;
        } else {
          if (cnt < it.n) {
            it.status = 3;  // dependent Ritz vectors: no FP64 start block
            continue;
          }
          it.m = it.dim_sub = it.dim_sub_origin = cnt;
          it.n_done = 0;
          it.phase = 1;
          it.it_phase = 0;
          it.first_can_converge = true;
        }
      }
    }
    t_exp += ms_since(t0);
    if (verbose) {
      int na = 0;
      for (GItem* it : g.items) na += it->status == 0;
      fprintf(stderr, "tincr_tda batch step %3d: %4zu active -> %4d\n", steps, act.size(), na);
    }
  }
  auto t_o = Clock::now();
  // ---- outputs (tda/mod.rs:88-126) ----
  {
    size_t out_doubles = 0;
    for (GItem* it : g.items) {
      it->out_off = out_doubles;
      out_doubles += (size_t)it->d.N * it->n + 4 * (size_t)it->n;
      it->x_off = out_doubles;
      out_doubles += (size_t)it->in->o * it->in->v * it->n;
    }
    g.outb.ensure(out_doubles * 8);
    g.out_h.ensure(out_doubles * 8);
    std::vector<GemmArgs> dq;
    std::vector<DipDesc> dd;
    std::vector<UnpackDesc> du;
    for (GItem* it : g.items) {
      if (it->status == 3) continue;
      double* ob = (double*)g.outb.p + it->out_off;
      it->qt = ob;
      it->mu = ob + (size_t)it->d.N * it->n;
      it->f = it->mu + 3 * (size_t)it->n;
      it->xout = (double*)g.outb.p + it->x_off;
      dq.push_back(gemm_args(it->qov, it->RZ, it->qt, it->d.N, it->n, it->L, plain(it->L), plain(1), plain(1),
                             plain(it->L), plain(it->n), plain(1)));
      dd.push_back(DipDesc{it->xyz, it->qt, it->evals, it->mu, it->f, it->d.N, it->n});
      du.push_back(UnpackDesc{it->RZ, it->xout, it->in->o, it->in->v, it->d.vp, it->n});
    }
    Plan& p = g.plan;
    p.reset();
    const int o_q = p.gemm(dq, default_cfg(OC_QT));
    const int o_d = p.units(dd);
    const int o_u = p.units(du);
    const int nd = (int)dd.size();
    p.upload(g.s);
    const bool gr = g.graphs;
    g.graphs = false;  // one-off launches
    run_gemm_op(g, o_q, DT_F64, DT_F64, EPI_STORE, true, true, default_cfg(OC_QT));
    g.graphs = gr;
    if (nd > 0) {
      dipoles_grouped_kernel<<<std::min(nd, GRID_VEC), 256, 0, g.s>>>(p.slot(o_d));
      TDA_LAUNCH_CHECK();
      unpack_grouped_kernel<<<std::min(nd, GRID_VEC), 256, 0, g.s>>>(p.slot(o_u));
      TDA_LAUNCH_CHECK();
    }
    TDA_HIP(hipMemcpyAsync(g.out_h.p, g.outb.p, out_doubles * 8, hipMemcpyDeviceToHost, g.s));
    TDA_HIP(hipStreamSynchronize(g.s));
    const double* oh = (const double*)g.out_h.p;
    for (GItem* it : g.items) {
      const BatchItem& b = *it->in;
      if (it->status != 3) {
        const double* qh = oh + it->out_off;
        if (b.omega) std::memcpy(b.omega, it->ev.data(), (size_t)it->n * 8);
        if (b.q_trans) std::memcpy(b.q_trans, qh, (size_t)it->d.N * it->n * 8);
        if (b.tr_dipoles) std::memcpy(b.tr_dipoles, qh + (size_t)it->d.N * it->n, 3 * (size_t)it->n * 8);
        if (b.f) std::memcpy(b.f, qh + (size_t)it->d.N * it->n + 3 * it->n, (size_t)it->n * 8);
        if (b.X) std::memcpy(b.X, oh + it->x_off, (size_t)b.o * b.v * it->n * 8);
      }
      set_status(*it, status, stats, mixed);
    }
  }
  t_out = ms_since(t_o);
  if (g.totals) {
    tda_stats& tt = *g.totals;
    for (const Group::Rec& r : g.recs) {
      float ms = 0;
      TDA_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      tt.exch_ms[r.stage][r.pclass] += ms;
      tt.exch_flops[r.stage][r.pclass] += r.flops;
      tt.exch_launches[r.stage][r.pclass] += 1;
    }
    for (GItem* it : g.items) {
      tt.n_iter = std::max(tt.n_iter, it->it);
      tt.n_products += it->n_products;
      tt.n_products_fp64 += it->n_products_fp64;
      tt.converged = tt.converged && it->status == 1;
    }
  }
  for (const Group::Rec& r : g.recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g.recs.clear();
  if (verbose)
    fprintf(stderr,
            "tincr_tda batch: %zu fragments, %d steps; ms: prepare %.1f, products %.1f, subspace %.1f, "
            "expansion %.1f, outputs %.1f, total %.1f\n",
            g.items.size(), steps, t_prep, t_prod, t_sub, t_exp, t_out, ms_since(t_start));
}

}  // namespace

int batch_run_grouped(int device, int precision, size_t budget, std::vector<BatchItem>& items, int32_t* status,
                      tda_stats* stats, tda_stats* totals, std::string& err) {
  const auto t0 = std::chrono::steady_clock::now();
  if (totals) {
    std::memset(totals, 0, sizeof(*totals));
    totals->converged = 1;
  }
  try {
    TDA_HIP(hipSetDevice(device));
    const bool mixed = precision != TDA_PREC_FP64;
    // fragment dimensions (the padded layouts of the single-fragment contexts)
    std::vector<GItem> all(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
      GItem& it = all[i];
      BatchItem& b = items[i];
      it.in = &b;
      it.index = (int)i;
      const bool lc = b.gamma_lr != nullptr;
      it.d.N = b.ct ? b.n_atoms + b.n_atoms_l : b.n_atoms;
      it.d.NL = lc ? (b.ct ? b.n_atoms_l : b.n_atoms) : 0;
      it.d.o = b.o;
      it.d.v = b.v;
      it.d.op = (int)rup(b.o, 64);
      it.d.vp = (int)rup(b.v, 64);
      it.d.lc = lc;
      it.n = b.n_roots;
      it.cap = 15 * b.n_roots;  // max_space (davidson.rs:74)
      it.L = (int64_t)b.o * it.d.vp;
      TDA_CHECK(it.cap <= 256, TDA_ERR_BAD_ARG, "batch: at most 17 roots per fragment in the grouped driver");
    }
    BatchMem& mem = batch_mem(device);
    std::lock_guard<std::mutex> mlk(mem.mu);
    // groups of fragments whose arenas fit the budget (largest first); no
    // budget: 80 % of the device memory that is free or already held
    size_t cap_bytes = budget;
    if (!cap_bytes) {
      size_t fr = 0, tot = 0;
      TDA_HIP(hipMemGetInfo(&fr, &tot));
      cap_bytes = (size_t)(0.8 * (double)(fr + mem.arena.bytes));
    }
    std::vector<int> order(all.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return items[a].cost > items[c].cost; });
    hipStream_t s = nullptr;
    TDA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() {
        if (s) {
          (void)hipStreamSynchronize(s);
          (void)hipStreamDestroy(s);
        }
      }
    } sg{s};
    size_t k = 0;
    const int n_workers = 8;
    while (k < order.size()) {
      Group g(mem);
      g.totals = totals;
      g.graphs = graphs_on();
      g.device = device;
      g.mixed = mixed;
      g.s = s;
      Footprint fp;
      std::vector<size_t> tmp;
      while (k < order.size()) {
        Footprint probe = fp;
        layout_item(all[order[k]], mixed, probe, tmp);
        if (!g.items.empty() && probe.bytes > cap_bytes) break;
        fp = probe;
        g.items.push_back(&all[order[k]]);
        ++k;
      }
      run_group(g, n_workers, budget, status, stats);
    }
  } catch (const Error& e) {
    err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    err = "out of host memory";
    return TDA_ERR_OOM;
  }
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats)
    for (size_t i = 0; i < items.size(); ++i) stats[i].t_total_ms = ms;
  if (totals) totals->t_total_ms = ms;
  return TDA_OK;
}

}  // namespace tda
