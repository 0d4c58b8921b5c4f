// Streaming (HBM-bound) kernels of the TDA path: transition-charge build,
// layout conversions and the Davidson vector updates.
#include <type_traits>

#include "kernels.h"

namespace tda {

static inline int grid_for(int64_t n, int per_block = 256, int cap = 16384) {
  int64_t g = ceil_div(n, per_block);
  if (g < 1) g = 1;
  return (int)std::min<int64_t>(g, cap);
}

// ---------------------------------------------------------------------------
// transition charges  q^A_pq = 1/2 sum_{mu in A}(C_mu,p SC_mu,q + C_mu,q SC_mu,p)
// (transition_charges.rs:38-95).  Block = 256 threads along q, QB_ROWS rows p.
// ---------------------------------------------------------------------------
constexpr int QB_ROWS = 8;

__global__ void __launch_bounds__(256) qblock_kernel(const double* __restrict__ C, const double* __restrict__ SC,
                                                     int64_t ld, const int* __restrict__ ao_off, int pb, int P,
                                                     int P_rows, int qb, int Q, int Q_pad, double* __restrict__ out) {
  const int A = blockIdx.z;
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int p0 = blockIdx.y * QB_ROWS;
  if (q >= Q_pad) return;
  const int mu0 = ao_off[A], mu1 = ao_off[A + 1];
  double acc[QB_ROWS];
#pragma unroll
  for (int r = 0; r < QB_ROWS; ++r) acc[r] = 0.0;
  if (q < Q) {
    for (int mu = mu0; mu < mu1; ++mu) {
      const double cq = C[mu * ld + qb + q];
      const double scq = SC[mu * ld + qb + q];
#pragma unroll
      for (int r = 0; r < QB_ROWS; ++r) {
        const int p = p0 + r;
        if (p < P) {
          const double cp = C[mu * ld + pb + p];
          const double scp = SC[mu * ld + pb + p];
          acc[r] += 0.5 * (cp * scq + cq * scp);
        }
      }
    }
  }
  double* o = out + (int64_t)A * P_rows * Q_pad;
#pragma unroll
  for (int r = 0; r < QB_ROWS; ++r) {
    const int p = p0 + r;
    if (p < P_rows) o[(int64_t)p * Q_pad + q] = (p < P) ? acc[r] : 0.0;
  }
}

void launch_qblock(const double* C, const double* SC, int64_t ld, const int* ao_off, int n_atoms, int pb, int P,
                   int P_rows, int qb, int Q, int Q_pad, double* out, hipStream_t s) {
  if (n_atoms == 0 || P_rows == 0 || Q_pad == 0) return;
  dim3 grid((unsigned)ceil_div(Q_pad, 256), (unsigned)ceil_div(P_rows, QB_ROWS), (unsigned)n_atoms);
  qblock_kernel<<<grid, 256, 0, s>>>(C, SC, ld, ao_off, pb, P, P_rows, qb, Q, Q_pad, out);
  TDA_LAUNCH_CHECK();
}

// out[A][i][a] = 1/2 sum_{mu in A} L[mu][i] R[mu][a]  (ct_pair.rs:99-121)
__global__ void __launch_bounds__(256) qprod_kernel(const double* __restrict__ Lm, int64_t ldl,
                                                    const double* __restrict__ Rm, int64_t ldr,
                                                    const int* __restrict__ ao_off, int atom_lo, int P, int P_rows,
                                                    int Q, int Q_pad, double* __restrict__ out) {
  const int A = blockIdx.z;
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int p0 = blockIdx.y * QB_ROWS;
  if (q >= Q_pad) return;
  const int mu0 = ao_off[A], mu1 = ao_off[A + 1];
  double acc[QB_ROWS];
#pragma unroll
  for (int r = 0; r < QB_ROWS; ++r) acc[r] = 0.0;
  if (q < Q) {
    for (int mu = mu0; mu < mu1; ++mu) {
      const double rq = Rm[mu * ldr + q];
#pragma unroll
      for (int r = 0; r < QB_ROWS; ++r) {
        const int p = p0 + r;
        if (p < P) acc[r] += Lm[mu * ldl + p] * rq;
      }
    }
  }
  double* o = out + (int64_t)(atom_lo + A) * P_rows * Q_pad;
#pragma unroll
  for (int r = 0; r < QB_ROWS; ++r) {
    const int p = p0 + r;
    if (p < P_rows) o[(int64_t)p * Q_pad + q] = (p < P) ? 0.5 * acc[r] : 0.0;
  }
}

void launch_qprod(const double* Lm, int64_t ldl, const double* Rm, int64_t ldr, const int* ao_off, int atom_lo,
                  int n_atoms, int P, int P_rows, int Q, int Q_pad, double* out, hipStream_t s) {
  if (n_atoms == 0 || P_rows == 0 || Q_pad == 0) return;
  dim3 grid((unsigned)ceil_div(Q_pad, 256), (unsigned)ceil_div(P_rows, QB_ROWS), (unsigned)n_atoms);
  qprod_kernel<<<grid, 256, 0, s>>>(Lm, ldl, Rm, ldr, ao_off, atom_lo, P, P_rows, Q, Q_pad, out);
  TDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// per-block max |x| and scaled conversions (per-atom scaling of low-precision
// operands)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) block_amax_kernel(const double* __restrict__ x, int64_t len,
                                                         float* __restrict__ amax) {
  const double* p = x + (int64_t)blockIdx.x * len;
  double m = 0.0;
  for (int64_t e = threadIdx.x; e < len; e += 256) m = fmax(m, fabs(p[e]));
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    amax[blockIdx.x] = (float)r;
  }
}

void launch_block_amax(const double* x, int64_t len, int n, float* amax, hipStream_t s) {
  if (n == 0) return;
  block_amax_kernel<<<n, 256, 0, s>>>(x, len, amax);
  TDA_LAUNCH_CHECK();
}

// Euclidean norm of each of n rows of length len (row stride ld): one wave
// per row, the sum in a fixed order.
__global__ void __launch_bounds__(256) row_norms_kernel(const double* __restrict__ x, int n, int64_t len,
                                                        int64_t ld, double* __restrict__ out) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const double* p = x + (int64_t)row * ld;
  double acc = 0.0;
  for (int64_t e = threadIdx.x & 63; e < len; e += 64) acc += p[e] * p[e];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) out[row] = sqrt(acc);
}

void launch_row_norms(const double* x, int n, int64_t len, int64_t ld, double* out, hipStream_t s) {
  if (n == 0) return;
  row_norms_kernel<<<(n + 3) / 4, 256, 0, s>>>(x, n, len, ld, out);
  TDA_LAUNCH_CHECK();
}

template <typename T>
__global__ void __launch_bounds__(256) convert_kernel(const double* __restrict__ src, int64_t len, int n,
                                                      const float* __restrict__ inv_scale, T* __restrict__ dst) {
  const int64_t total = len * n;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    double v = src[t];
    if (inv_scale) v *= (double)inv_scale[t / len];
    dst[t] = to_type<T>(v);
  }
}

void launch_convert(const double* src, int64_t len, int n, const float* inv_scale, void* dst, int dtype,
                    hipStream_t s) {
  const int64_t total = len * n;
  if (total == 0) return;
  const int g = grid_for(total);
  switch (dtype) {
    case PREC_FP32: convert_kernel<float><<<g, 256, 0, s>>>(src, len, n, inv_scale, (float*)dst); break;
    case PREC_BF16: convert_kernel<__bf16><<<g, 256, 0, s>>>(src, len, n, inv_scale, (__bf16*)dst); break;
    case PREC_FP16: convert_kernel<_Float16><<<g, 256, 0, s>>>(src, len, n, inv_scale, (_Float16*)dst); break;
    default: convert_kernel<double><<<g, 256, 0, s>>>(src, len, n, inv_scale, (double*)dst); break;
  }
  TDA_LAUNCH_CHECK();
}

// dst[r][c] = hi, dst[r][lo_off + c] = lo of src[r][c] (rows of `cols`, row strides ld_src / ld_dst)
__global__ void __launch_bounds__(256) convert_split_kernel(const double* __restrict__ src, int64_t rows,
                                                            int64_t cols, int64_t ld_src, _Float16* __restrict__ dst,
                                                            int64_t ld_dst, int64_t lo_off) {
  const int64_t total = rows * cols;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int64_t r = t / cols, c = t - r * cols;
    _Float16 hi, lo;
    split_f16((float)src[r * ld_src + c], hi, lo);
    dst[r * ld_dst + c] = hi;
    dst[r * ld_dst + lo_off + c] = lo;
  }
}

void launch_convert_split(const double* src, int64_t rows, int64_t cols, int64_t ld_src, void* dst, int64_t ld_dst,
                          int64_t lo_off, hipStream_t s) {
  const int64_t total = rows * cols;
  if (total == 0) return;
  convert_split_kernel<<<grid_for(total), 256, 0, s>>>(src, rows, cols, ld_src, (_Float16*)dst, ld_dst, lo_off);
  TDA_LAUNCH_CHECK();
}

template <typename T>
__global__ void __launch_bounds__(256) convert_f32_kernel(const float* __restrict__ src, int64_t n, T* __restrict__ dst) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
    dst[t] = to_type<T>((double)src[t]);
}

void launch_convert_f32_to(const float* src, int64_t n, void* dst, int dtype, hipStream_t s) {
  if (n == 0) return;
  const int g = grid_for(n);
  switch (dtype) {
    case PREC_BF16: convert_f32_kernel<__bf16><<<g, 256, 0, s>>>(src, n, (__bf16*)dst); break;
    case PREC_FP16: convert_f32_kernel<_Float16><<<g, 256, 0, s>>>(src, n, (_Float16*)dst); break;
    default: throw Error{2, "convert_f32_to: bad dtype"};
  }
  TDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// layout conversions
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) pack_x_kernel(const double* __restrict__ src, int o, int v, int vp, int k,
                                                     double* __restrict__ dst) {
  const int64_t ovp = (int64_t)o * vp;
  const int64_t total = ovp * k;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int c = (int)(t / ovp);
    const int64_t e = t % ovp;
    const int i = (int)(e / vp), a = (int)(e % vp);
    dst[t] = (a < v) ? src[((int64_t)i * v + a) * k + c] : 0.0;
  }
}

__global__ void __launch_bounds__(256) unpack_x_kernel(const double* __restrict__ src, int o, int v, int vp, int k,
                                                       double* __restrict__ dst) {
  const int64_t ov = (int64_t)o * v;
  const int64_t total = ov * k;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int64_t ia = t / k;
    const int c = (int)(t % k);
    const int i = (int)(ia / v), a = (int)(ia % v);
    dst[t] = src[(int64_t)c * o * vp + (int64_t)i * vp + a];
  }
}

void launch_pack_x(const double* src, int o, int v, int vp, int k, double* dst, hipStream_t s) {
  const int64_t total = (int64_t)o * vp * k;
  if (total == 0) return;
  pack_x_kernel<<<grid_for(total), 256, 0, s>>>(src, o, v, vp, k, dst);
  TDA_LAUNCH_CHECK();
}

void launch_unpack_x(const double* src, int o, int v, int vp, int k, double* dst, hipStream_t s) {
  const int64_t total = (int64_t)o * v * k;
  if (total == 0) return;
  unpack_x_kernel<<<grid_for(total), 256, 0, s>>>(src, o, v, vp, k, dst);
  TDA_LAUNCH_CHECK();
}

__global__ void delta_eps_kernel(const double* __restrict__ orbe, int ob, int o, int vb, int v, int vp,
                                 double* __restrict__ de) {
  const int64_t total = (int64_t)o * vp;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int i = (int)(t / vp), a = (int)(t % vp);
    de[t] = (a < v) ? orbe[vb + a] - orbe[ob + i] : 0.0;
  }
}

void launch_delta_eps(const double* orbe, int ob, int o, int vb, int v, int vp, double* de, hipStream_t s) {
  const int64_t total = (int64_t)o * vp;
  if (total == 0) return;
  delta_eps_kernel<<<grid_for(total), 256, 0, s>>>(orbe, ob, o, vb, v, vp, de);
  TDA_LAUNCH_CHECK();
}

template <typename T>
__global__ void __launch_bounds__(256) cast_kernel(const double* __restrict__ src, int64_t n, T* __restrict__ dst) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
    dst[t] = to_type<T>(src[t]);
}

__global__ void __launch_bounds__(256) widen_kernel(const float* __restrict__ src, int64_t n, double* __restrict__ dst) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) dst[t] = src[t];
}
void launch_widen(const float* src, int64_t n, double* dst, hipStream_t s) {
  if (n == 0) return;
  widen_kernel<<<grid_for(n), 256, 0, s>>>(src, n, dst);
  TDA_LAUNCH_CHECK();
}

void launch_cast_vectors(const double* src, int64_t n, void* dst, int dtype, hipStream_t s) {
  if (n == 0) return;
  const int g = grid_for(n);
  switch (dtype) {
    case PREC_FP32: cast_kernel<float><<<g, 256, 0, s>>>(src, n, (float*)dst); break;
    case PREC_BF16: cast_kernel<__bf16><<<g, 256, 0, s>>>(src, n, (__bf16*)dst); break;
    case PREC_FP16: cast_kernel<_Float16><<<g, 256, 0, s>>>(src, n, (_Float16*)dst); break;
    default: throw Error{2, "cast_vectors: bad dtype"};
  }
  TDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Davidson helpers
// ---------------------------------------------------------------------------
// thr = 1e-4 for the TDA engines (solvers/mod.rs:378); 0 (no floor) for a
// dense matrix engine (solvers/mod.rs:44-46)
__device__ inline double precond_denom(double w, double de, double thr) {
  const double d = w - de;
  return fabs(d) < thr ? 1.0 : d;
}

__global__ void __launch_bounds__(256) precondition_kernel(const double* __restrict__ r, double w,
                                                           const double* __restrict__ de, int64_t L, double thr,
                                                           double* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256)
    out[t] = r[t] / precond_denom(w, de[t], thr);
}

void launch_precondition(const double* r, double w, const double* de, int64_t L, double thr, double* out,
                         hipStream_t s) {
  if (L == 0) return;
  precondition_kernel<<<grid_for(L), 256, 0, s>>>(r, w, de, L, thr, out);
  TDA_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(256) precondition_many_kernel(const double* __restrict__ R,
                                                                const double* __restrict__ u,
                                                                const int* __restrict__ idx, int nc,
                                                                const double* __restrict__ de, int64_t L,
                                                                double thr, double* __restrict__ P) {
  const int c = blockIdx.y;
  const int r = idx[c];
  const double w = u[r];
  const double* src = R + (int64_t)r * L;
  double* dst = P + (int64_t)c * L;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256)
    dst[t] = src[t] / precond_denom(w, de[t], thr);
}

void launch_precondition_many(const double* R, const double* u, const int* idx, int nc, const double* de, int64_t L,
                              double thr, double* P, hipStream_t s) {
  if (nc == 0 || L == 0) return;
  dim3 grid((unsigned)grid_for(L, 256, 4096), (unsigned)nc);
  precondition_many_kernel<<<grid, 256, 0, s>>>(R, u, idx, nc, de, L, thr, P);
  TDA_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(256) residual_kernel(double* __restrict__ AVY, const double* __restrict__ RZ,
                                                       const double* __restrict__ u, int64_t L) {
  const int r = blockIdx.y;
  const double w = u[r];
  double* a = AVY + (int64_t)r * L;
  const double* z = RZ + (int64_t)r * L;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256)
    a[t] = a[t] - z[t] * w;
}

void launch_residual(double* AVY, const double* RZ, const double* u, int n, int64_t L, hipStream_t s) {
  if (n == 0 || L == 0) return;
  dim3 grid((unsigned)grid_for(L, 256, 4096), (unsigned)n);
  residual_kernel<<<grid, 256, 0, s>>>(AVY, RZ, u, L);
  TDA_LAUNCH_CHECK();
}

// Deterministic two-pass vector norms: NB_NORM partial sums per vector.
constexpr int NB_NORM = 256;

__global__ void __launch_bounds__(256) norm_partial_kernel(const double* __restrict__ X, int64_t L,
                                                           double* __restrict__ partial) {
  const int r = blockIdx.y;
  const double* x = X + (int64_t)r * L;
  double acc = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256) {
    const double v = x[t];
    acc += v * v;
  }
  acc = wave_sum(acc);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[(int64_t)r * NB_NORM + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void __launch_bounds__(256) norm_final_kernel(const double* __restrict__ partial, int n,
                                                         double* __restrict__ norms) {
  const int r = blockIdx.x;
  double acc = partial[(int64_t)r * NB_NORM + threadIdx.x];
  acc = wave_sum(acc);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) norms[r] = sqrt((red[0] + red[1]) + (red[2] + red[3]));
}

void launch_norms(const double* X, int n, int64_t L, double* partial, double* norms, hipStream_t s) {
  if (n == 0) return;
  norm_partial_kernel<<<dim3(NB_NORM, n), 256, 0, s>>>(X, L, partial);
  TDA_LAUNCH_CHECK();
  norm_final_kernel<<<n, 256, 0, s>>>(partial, n, norms);
  TDA_LAUNCH_CHECK();
}

// Column dot products out[c] = X[c] . Y[c] (same two-pass order as the norms).
__global__ void __launch_bounds__(256) dot_partial_kernel(const double* __restrict__ X, const double* __restrict__ Y,
                                                          int64_t L, double* __restrict__ partial) {
  const int r = blockIdx.y;
  const double* x = X + (int64_t)r * L;
  const double* y = Y + (int64_t)r * L;
  double acc = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256) acc += x[t] * y[t];
  acc = wave_sum(acc);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[(int64_t)r * NB_NORM + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void __launch_bounds__(256) dot_final_kernel(const double* __restrict__ partial, double* __restrict__ out) {
  const int r = blockIdx.x;
  double acc = partial[(int64_t)r * NB_NORM + threadIdx.x];
  acc = wave_sum(acc);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[r] = (red[0] + red[1]) + (red[2] + red[3]);
}

void launch_dots(const double* X, const double* Y, int n, int64_t L, double* partial, double* out, hipStream_t s) {
  if (n == 0) return;
  dot_partial_kernel<<<dim3(NB_NORM, n), 256, 0, s>>>(X, Y, L, partial);
  TDA_LAUNCH_CHECK();
  dot_final_kernel<<<n, 256, 0, s>>>(partial, out);
  TDA_LAUNCH_CHECK();
}

// Conjugate-gradient pieces of the Z-vector solve (per column c):
//   init:  Z = R / de (entries with de == 0, the row pads, stay 0)
//   step:  Z += a_c P, R -= a_c AP
//   dir:   P = b_c P + R          (columns with b_c < 0 are left alone)
__global__ void __launch_bounds__(256) cg_init_kernel(const double* __restrict__ R, const double* __restrict__ de,
                                                      int64_t L, double* __restrict__ Z) {
  const int c = blockIdx.y;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256)
    Z[(int64_t)c * L + t] = de[t] != 0.0 ? R[(int64_t)c * L + t] / de[t] : 0.0;
}
__global__ void __launch_bounds__(256) cg_step_kernel(double* __restrict__ Z, double* __restrict__ R,
                                                      const double* __restrict__ P, const double* __restrict__ AP,
                                                      const double* __restrict__ a, int64_t L) {
  const int c = blockIdx.y;
  const double ac = a[c];
  if (ac == 0.0) return;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256) {
    const int64_t i = (int64_t)c * L + t;
    Z[i] += ac * P[i];
    R[i] -= ac * AP[i];
  }
}
__global__ void __launch_bounds__(256) cg_dir_kernel(double* __restrict__ P, const double* __restrict__ R,
                                                     const double* __restrict__ b, int64_t L) {
  const int c = blockIdx.y;
  const double bc = b[c];
  if (bc < 0.0) return;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256) {
    const int64_t i = (int64_t)c * L + t;
    P[i] = bc * P[i] + R[i];
  }
}

void launch_cg_init(const double* R, const double* de, int n, int64_t L, double* Z, hipStream_t s) {
  if (n == 0) return;
  cg_init_kernel<<<dim3((unsigned)grid_for(L, 256, 4096), (unsigned)n), 256, 0, s>>>(R, de, L, Z);
  TDA_LAUNCH_CHECK();
}
void launch_cg_step(double* Z, double* R, const double* P, const double* AP, const double* a, int n, int64_t L,
                    hipStream_t s) {
  if (n == 0) return;
  cg_step_kernel<<<dim3((unsigned)grid_for(L, 256, 4096), (unsigned)n), 256, 0, s>>>(Z, R, P, AP, a, L);
  TDA_LAUNCH_CHECK();
}
void launch_cg_dir(double* P, const double* R, const double* b, int n, int64_t L, hipStream_t s) {
  if (n == 0) return;
  cg_dir_kernel<<<dim3((unsigned)grid_for(L, 256, 4096), (unsigned)n), 256, 0, s>>>(P, R, b, L);
  TDA_LAUNCH_CHECK();
}

// ---- Gram-Schmidt of a block of new vectors among themselves ----
// After the projection against the old basis (W = P - V V^T P), vector j is
//   u_j = w_j - sum_{s < j accepted} q_s (q_s . p_j),  q_s = u_s / |u_s|,
// accepted when |u_j| > 1e-8 (davidson.rs:162-168).  With q_s . p_j =
// (u_s . p_j) / |u_s| the update reads u_j = w_j - sum_s (u_s . p_j) / |u_s|^2 u_s.
// Step j updates w_j in place and, right-looking, forms its partial dots
// u_j . p_l (l > j) and |u_j|^2 per block; a one-block kernel sums them.
// scr layout: partial [nc][nc][NB_GS] (slot l == j holds |u_j|^2), then
// D [nc][nc] (D[s][l] = u_s . p_l, D[j][j] = |u_j|^2), then acc [nc].
constexpr int NB_GS = 1024;
__global__ void __launch_bounds__(256) gs_block_update_kernel(const double* __restrict__ P, double* __restrict__ W,
                                                              int64_t L, int j, int nc, double* __restrict__ scr) {
  const double* D = scr + (int64_t)nc * nc * NB_GS;
  const double* acc = D + (int64_t)nc * nc;
  __shared__ double coef[64];
  __shared__ double red[4];
  if (threadIdx.x < j) {
    const int s = threadIdx.x;
    coef[s] = acc[s] != 0.0 ? D[(int64_t)s * nc + j] / D[(int64_t)s * nc + s] : 0.0;
  }
  __syncthreads();
  double* wj = W + (int64_t)j * L;
  double nn = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256) {
    double u = wj[t];
    for (int s2 = 0; s2 < j; ++s2) u -= coef[s2] * W[(int64_t)s2 * L + t];
    wj[t] = u;
    nn += u * u;
  }
  double* part = scr + (int64_t)j * nc * NB_GS;
  for (int l = j; l < nc; ++l) {
    double a = 0.0;
    if (l == j) {
      a = nn;
    } else {
      const double* pl = P + (int64_t)l * L;
      for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256)
        a += wj[t] * pl[t];
    }
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) part[(int64_t)l * NB_GS + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
  }
}

// one block per l in [j, nc)
__global__ void __launch_bounds__(256) gs_block_final_kernel(int j, int nc, double* __restrict__ scr) {
  double* D = scr + (int64_t)nc * nc * NB_GS;
  double* acc = D + (int64_t)nc * nc;
  const int l = j + blockIdx.x;
  const double* part = scr + ((int64_t)j * nc + l) * NB_GS;
  __shared__ double red[4];
  double a = 0.0;
  for (int b = threadIdx.x; b < NB_GS; b += 256) a += part[b];
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double v = (red[0] + red[1]) + (red[2] + red[3]);
    D[(int64_t)j * nc + l] = v;
    if (l == j) acc[j] = (sqrt(v) > 1.0e-8) ? 1.0 : 0.0;  // davidson.rs:165
  }
}

// Q[rank(j)] = u_j / |u_j| for the accepted j, in order; *count = accepted
__global__ void __launch_bounds__(256) gs_block_append_kernel(const double* __restrict__ W, double* __restrict__ Q,
                                                              int64_t L, int nc, const double* __restrict__ scr,
                                                              int* __restrict__ count) {
  const double* D = scr + (int64_t)nc * nc * NB_GS;
  const double* acc = D + (int64_t)nc * nc;
  const int j = blockIdx.y;
  int rank = 0;
  for (int s = 0; s < j; ++s) rank += acc[s] != 0.0;
  if (blockIdx.x == 0 && j == 0 && threadIdx.x == 0) {
    int tot = 0;
    for (int s = 0; s < nc; ++s) tot += acc[s] != 0.0;
    *count = tot;
  }
  if (acc[j] == 0.0) return;
  const double inv = 1.0 / sqrt(D[(int64_t)j * nc + j]);
  const double* u = W + (int64_t)j * L;
  double* q = Q + (int64_t)rank * L;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < L; t += (int64_t)gridDim.x * 256) q[t] = u[t] * inv;
}

void launch_gs_block(const double* P, double* W, double* Q, int64_t L, int nc, double* scratch, int* count,
                     hipStream_t s) {
  if (nc == 0) return;
  for (int j = 0; j < nc; ++j) {
    gs_block_update_kernel<<<NB_GS, 256, 0, s>>>(P, W, L, j, nc, scratch);
    TDA_LAUNCH_CHECK();
    gs_block_final_kernel<<<nc - j, 256, 0, s>>>(j, nc, scratch);
    TDA_LAUNCH_CHECK();
  }
  gs_block_append_kernel<<<dim3((unsigned)grid_for(L, 256, 1024), (unsigned)nc), 256, 0, s>>>(W, Q, L, nc, scratch,
                                                                                              count);
  TDA_LAUNCH_CHECK();
}

size_t gs_block_scratch_doubles(int nc) { return (size_t)nc * nc * NB_GS + (size_t)nc * nc + nc; }

// One wave; lane t owns index t (nc <= 64).  LDS accesses between the steps
// are ordered by the workgroup barrier (a single wave: no waiting on others).
__global__ void __launch_bounds__(64) gs_gram_kernel(const double* __restrict__ Gt, int nc, double* __restrict__ Tq,
                                                     double* __restrict__ wn, int* __restrict__ count) {
  __shared__ double T[64][65];  // row j: coefficients of u_j over w_0..w_j
  __shared__ double nrm2[64];
  __shared__ double cs[64];
  __shared__ int acc[64];
  const int t = threadIdx.x;
  const double* Gww = Gt + (int64_t)nc * nc;  // Gww[l * nc + r] = w_l . w_r
  bool ill = false;
  for (int j = 0; j < nc; ++j) {
    // c_s = (u_s . p_j) / |u_s|^2 for the accepted s < j; u_s . p_j = sum_r T[s][r] (p_j . w_r)
    if (t < j) {
      double v = 0.0;
      if (acc[t]) {
        for (int r = 0; r <= t; ++r) v += T[t][r] * Gt[(int64_t)j * nc + r];
        v /= nrm2[t];
      }
      cs[t] = v;
    }
    __syncthreads();
    // u_j = w_j - sum_s c_s u_s  (T[s][t] = 0 for t > s)
    if (t < nc) {
      double v = (t == j) ? 1.0 : 0.0;
      for (int s2 = t; s2 < j; ++s2) v -= cs[s2] * T[s2][t];
      T[j][t] = v;
    }
    __syncthreads();
    double part = 0.0;
    if (t <= j) {
      double g = 0.0;
      for (int r = 0; r <= j; ++r) g += Gww[(int64_t)t * nc + r] * T[j][r];
      part = T[j][t] * g;
    }
    const double nn = wave_sum(part);
    if (t == 0) {
      nrm2[j] = nn;
      acc[j] = sqrt(fmax(nn, 0.0)) > 1.0e-8;  // davidson.rs:165
    }
    ill = ill || !(nn >= 1.0e-4 * Gww[(int64_t)j * nc + j]);
    __syncthreads();
  }
  if (t < nc) {
    wn[t] = sqrt(fmax(nrm2[t], 0.0));
    if (acc[t]) {
      int rank = 0;
      for (int s2 = 0; s2 < t; ++s2) rank += acc[s2];
      const double inv = 1.0 / sqrt(nrm2[t]);
      for (int r = 0; r < nc; ++r) Tq[(int64_t)rank * nc + r] = r <= t ? T[t][r] * inv : 0.0;
    }
  }
  if (t == 0) {
    int tot = 0;
    for (int s2 = 0; s2 < nc; ++s2) tot += acc[s2];
    count[0] = tot;
    count[1] = ill ? 1 : 0;
  }
}

void launch_gs_gram(const double* Gt, int nc, double* Tq, double* wn, int* count, hipStream_t s) {
  TDA_CHECK(nc >= 1 && nc <= 64, 2, "gs_gram: nc must be in 1..64");
  gs_gram_kernel<<<1, 64, 0, s>>>(Gt, nc, Tq, wn, count);
  TDA_LAUNCH_CHECK();
}

__global__ void copy_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) dst[t] = src[t];
}
__global__ void fill_kernel(double* __restrict__ dst, double v, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) dst[t] = v;
}
void launch_copy(const double* src, double* dst, int64_t n, hipStream_t s) {
  if (n == 0) return;
  copy_kernel<<<grid_for(n), 256, 0, s>>>(src, dst, n);
  TDA_LAUNCH_CHECK();
}
void launch_fill(double* dst, double v, int64_t n, hipStream_t s) {
  if (n == 0) return;
  fill_kernel<<<grid_for(n), 256, 0, s>>>(dst, v, n);
  TDA_LAUNCH_CHECK();
}

// AO offsets from per-atom AO counts on the device (one thread; n_atoms is
// small): off[A + 1] = off[A] + opa[A], each clamped to [off[A], n_ao] so that
// no later kernel can index past the n_ao rows; *bad = 1 when some count was
// negative or the counts do not sum to n_ao.
__global__ void ao_offsets_kernel(const int32_t* __restrict__ opa, int n_atoms, int n_ao, int* __restrict__ off,
                                  int* __restrict__ bad) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t sum = 0;
  int b = 0, o = 0;
  off[0] = 0;
  for (int A = 0; A < n_atoms; ++A) {
    const int c = opa[A];
    if (c < 0) b = 1;
    sum += c;
    const int64_t next = (int64_t)o + (c > 0 ? c : 0);
    o = (int)(next < n_ao ? next : n_ao);
    off[A + 1] = o;
  }
  *bad = (b || sum != n_ao) ? 1 : 0;
}
void launch_ao_offsets(const int32_t* opa, int n_atoms, int n_ao, int* off, int* bad, hipStream_t s) {
  ao_offsets_kernel<<<1, 64, 0, s>>>(opa, n_atoms, n_ao, off, bad);
  TDA_LAUNCH_CHECK();
}

// out[mu][nu] = (float) g[atom[mu]][atom[nu]] for mu, nu < n_ao (row stride
// ld; g is [N][N]): an atom-pair table expanded to AO pairs
__global__ void __launch_bounds__(256) expand_pairs_kernel(const double* __restrict__ g, const int* __restrict__ atom,
                                                           int n_ao, int N, int64_t ld, float* __restrict__ out) {
  const int mu = blockIdx.y;
  const int64_t gr = (int64_t)atom[mu] * N;
  for (int nu = blockIdx.x * 256 + threadIdx.x; nu < n_ao; nu += gridDim.x * 256)
    out[(int64_t)mu * ld + nu] = (float)g[gr + atom[nu]];
}
void launch_expand_pairs(const double* g, const int* atom, int n_ao, int N, int64_t ld, float* out, hipStream_t s) {
  if (n_ao == 0) return;
  expand_pairs_kernel<<<dim3((unsigned)ceil_div(n_ao, 256), (unsigned)n_ao), 256, 0, s>>>(g, atom, n_ao, N, ld, out);
  TDA_LAUNCH_CHECK();
}

// X[r][pos[r]] = 1 for r < n (unit guess vectors; rows of length L)
__global__ void set_ones_kernel(double* __restrict__ X, int64_t L, const int64_t* __restrict__ pos, int n) {
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r < n) X[(int64_t)r * L + pos[r]] = 1.0;
}
void launch_set_ones(double* X, int64_t L, const int64_t* pos, int n, hipStream_t s) {
  if (n == 0) return;
  set_ones_kernel<<<(unsigned)ceil_div(n, 64), 64, 0, s>>>(X, L, pos, n);
  TDA_LAUNCH_CHECK();
}

__global__ void zero_words_kernel(uint32_t* __restrict__ dst, int64_t nw) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nw; t += (int64_t)gridDim.x * 256) dst[t] = 0u;
}
__global__ void zero_tail_kernel(unsigned char* __restrict__ dst, int64_t nb) {
  if (threadIdx.x < nb) dst[threadIdx.x] = 0;
}
void launch_zero_bytes(void* dst, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  const int64_t nw = (int64_t)(bytes / 4);
  if (nw) {
    zero_words_kernel<<<grid_for(nw), 256, 0, s>>>((uint32_t*)dst, nw);
    TDA_LAUNCH_CHECK();
  }
  const int64_t tail = (int64_t)(bytes % 4);
  if (tail) {
    zero_tail_kernel<<<1, 64, 0, s>>>((unsigned char*)dst + nw * 4, tail);
    TDA_LAUNCH_CHECK();
  }
}

__global__ void zero_runs_kernel(uint16_t* __restrict__ base, int64_t rows, int64_t ld, int64_t c0, int64_t cw) {
  const int64_t total = rows * cw;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256)
    base[(t / cw) * ld + c0 + t % cw] = 0;
}

void launch_zero_runs(void* base, int64_t rows, int64_t ld, int64_t c0, int64_t cw, int esz, hipStream_t s) {
  if (rows == 0 || cw == 0) return;
  const int64_t u = esz / 2;  // in 16-bit units
  zero_runs_kernel<<<grid_for(rows * cw * u), 256, 0, s>>>((uint16_t*)base, rows, ld * u, c0 * u, cw * u);
  TDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// AO-factored exchange / Coulomb operands
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) ao_operand_kernel(const double* __restrict__ C, const double* __restrict__ SC,
                                                         int n_ao, int base, int P, int P_pad, int swap, int transpose,
                                                         int64_t ld, double* __restrict__ out) {
  const int64_t n2 = 2 * (int64_t)n_ao;
  // transposed: [P_pad][ld] with the ld - n2 pad columns zeroed; else [n2][P_pad]
  const int64_t rows = transpose ? ld : n2;
  const int64_t total = rows * P_pad;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int64_t r = t / P_pad;
    const int col = (int)(t % P_pad);
    double v = 0.0;
    if (r < n2 && col < P) {
      const int sidx = (int)(r / n_ao);
      const int64_t mu = r % n_ao;
      const double* src = (sidx == swap) ? SC : C;
      v = src[mu * n_ao + base + col];
    }
    if (transpose) out[(int64_t)col * ld + r] = v;
    else out[t] = v;
  }
}

void launch_ao_operand(const double* C, const double* SC, int n_ao, int base, int P, int P_pad, int swap,
                       int transpose, int64_t ld, double* out, hipStream_t s) {
  const int64_t total = (transpose ? ld : 2 * (int64_t)n_ao) * P_pad;
  if (total == 0) return;
  ao_operand_kernel<<<grid_for(total), 256, 0, s>>>(C, SC, n_ao, base, P, P_pad, swap, transpose, ld, out);
  TDA_LAUNCH_CHECK();
}

template <typename T>
__global__ void __launch_bounds__(64) ao_charges_kernel(const T* __restrict__ M, int k, int n_ao, int64_t ld,
                                                        const int* __restrict__ ao_off,
                                                        const double* __restrict__ gamma_lr, int n_atoms,
                                                        double* __restrict__ u, int64_t lo_off) {
  const int A = blockIdx.x;
  const int c = blockIdx.y;
  const int64_t n2 = 2 * (int64_t)n_ao;
  const T* Mc = M + (int64_t)c * n2 * ld;
  // M = -1/4 gamma~ o D, gamma~_mu,mu = gamma_lr[A][A] for mu on atom A  =>  D = M / (-1/4 gamma_lr[A][A])
  const double inv = -4.0 / gamma_lr[(int64_t)A * n_atoms + A];
  double acc = 0.0;
  for (int mu = ao_off[A] + threadIdx.x; mu < ao_off[A + 1]; mu += 64) {
    const int64_t i2 = (int64_t)mu * ld + n_ao + mu, i3 = (int64_t)(n_ao + mu) * ld + mu;
    double d2 = to_double<T>(Mc[i2]), d3 = to_double<T>(Mc[i3]);
    if (lo_off) {  // FP16 hi / lo split storage
      d2 += to_double<T>(Mc[i2 + lo_off]) * 4.8828125e-4;
      d3 += to_double<T>(Mc[i3 + lo_off]) * 4.8828125e-4;
    }
    acc += 0.5 * (d2 + d3) * inv;
  }
  acc = wave_sum(acc);
  if (threadIdx.x == 0) u[(int64_t)A * k + c] = acc;
}

void launch_ao_charges(const void* M, int dtype, int k, int n_ao, int64_t ld, const int* ao_off, int n_atoms,
                       const double* gamma_lr, double* u, hipStream_t s, int64_t lo_off) {
  dim3 grid(n_atoms, k);
  switch (dtype) {
    case PREC_FP64:
      ao_charges_kernel<double><<<grid, 64, 0, s>>>((const double*)M, k, n_ao, ld, ao_off, gamma_lr, n_atoms, u,
                                                    lo_off);
      break;
    case PREC_FP32:
      ao_charges_kernel<float><<<grid, 64, 0, s>>>((const float*)M, k, n_ao, ld, ao_off, gamma_lr, n_atoms, u,
                                                    lo_off);
      break;
    case PREC_BF16:
      ao_charges_kernel<__bf16><<<grid, 64, 0, s>>>((const __bf16*)M, k, n_ao, ld, ao_off, gamma_lr, n_atoms, u,
                                                    lo_off);
      break;
    default:
      ao_charges_kernel<_Float16><<<grid, 64, 0, s>>>((const _Float16*)M, k, n_ao, ld, ao_off, gamma_lr, n_atoms, u,
                                                    lo_off);
      break;
  }
  TDA_LAUNCH_CHECK();
}

// Transition charges of k vectors from the AO factors, without the MO
// charges q_ov: with D = L x R^T and XR_c = R x_c^T (the first GEMM of the
// AO-factored product), q_ov^A x_c = 1/2 sum_{mu in A} (D[(0,mu),(1,mu)] +
// D[(1,mu),(0,mu)]) = 1/2 sum_{mu in A} sum_i (L[(0,mu)][i] XR_c[(1,mu)][i] +
// L[(1,mu)][i] XR_c[(0,mu)][i]).  One workgroup per (atom, vector); the sum
// runs in a fixed order (deterministic).
__global__ void __launch_bounds__(256) ao_dens_charges_kernel(const double* __restrict__ L,
                                                              const double* __restrict__ XR, int n_ao, int o,
                                                              int64_t ld, const int* __restrict__ ao_off,
                                                              double* __restrict__ u, int64_t ldu) {
  const int A = blockIdx.x;
  const int c = blockIdx.y;
  const double* X = XR + (int64_t)c * 2 * n_ao * ld;
  double acc = 0.0;
  for (int mu = ao_off[A]; mu < ao_off[A + 1]; ++mu) {
    const double* l0 = L + (int64_t)mu * ld;
    const double* l1 = L + (int64_t)(n_ao + mu) * ld;
    const double* x0 = X + (int64_t)mu * ld;
    const double* x1 = X + (int64_t)(n_ao + mu) * ld;
    for (int i = threadIdx.x; i < o; i += 256) acc += l0[i] * x1[i] + l1[i] * x0[i];
  }
  __shared__ double red[4];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) u[(int64_t)A * ldu + c] = 0.5 * (red[0] + red[1] + red[2] + red[3]);
}

void launch_ao_dens_charges(const double* L, const double* XR, int k, int n_ao, int o, int64_t ld,
                            const int* ao_off, int n_atoms, double* u, int64_t ldu, hipStream_t s) {
  if (k == 0 || n_atoms == 0) return;
  dim3 grid(n_atoms, k);
  ao_dens_charges_kernel<<<grid, 256, 0, s>>>(L, XR, n_ao, o, ld, ao_off, u, ldu);
  TDA_LAUNCH_CHECK();
}

// Derivative of w = x^T A x with respect to gamma_lr through the AO-factored
// exchange: w holds -1/4 sum_pq D'_pq gamma~_pq D_pq with D = L x R^T and
// D' = L' x R'^T = D with both index halves exchanged, so
// dw / dgamma_lr[A][B] = -1/2 sum_{mu in A, nu in B} (D[mu][nu] D[n+mu][n+nu] +
// D[mu][n+nu] D[n+mu][nu])   (n = n_ao; D rows with stride ld).
__global__ void __launch_bounds__(256) ao_dglr_kernel(const double* __restrict__ D, int64_t ld, int n_ao,
                                                      const int* __restrict__ ao_off, int n_atoms,
                                                      double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)n_atoms * n_atoms) return;
  const int A = (int)(t / n_atoms), B = (int)(t % n_atoms);
  double acc = 0.0;
  for (int mu = ao_off[A]; mu < ao_off[A + 1]; ++mu) {
    const double* d0 = D + (int64_t)mu * ld;
    const double* d1 = D + (int64_t)(n_ao + mu) * ld;
    for (int nu = ao_off[B]; nu < ao_off[B + 1]; ++nu)
      acc += d0[nu] * d1[n_ao + nu] + d0[n_ao + nu] * d1[nu];
  }
  out[t] = -0.5 * acc;
}

void launch_ao_dglr(const double* D, int64_t ld, int n_ao, const int* ao_off, int n_atoms, double* out,
                    hipStream_t s) {
  const int64_t total = (int64_t)n_atoms * n_atoms;
  if (total == 0) return;
  ao_dglr_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(D, ld, n_ao, ao_off, n_atoms, out);
  TDA_LAUNCH_CHECK();
}

template <typename T>
__global__ void __launch_bounds__(256) ao_add_coulomb_kernel(T* __restrict__ M, int k, int n_ao, int64_t ld,
                                                             const int* __restrict__ atom,
                                                             const double* __restrict__ w, int64_t lo_off) {
  const int64_t n2 = 2 * (int64_t)n_ao;
  const int64_t total = (int64_t)k * n_ao;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int c = (int)(t / n_ao);
    const int64_t mu = t % n_ao;
    const double add = w[(int64_t)atom[mu] * k + c];
    T* Mc = M + (int64_t)c * n2 * ld;
    T* p1 = Mc + mu * ld + n_ao + mu;
    T* p2 = Mc + (n_ao + mu) * ld + mu;
    if constexpr (std::is_same<T, _Float16>::value) {
      if (lo_off) {  // FP16 hi / lo split storage: add to the joined value, split again
        const float v1 = (float)((double)join_f16(p1[0], p1[lo_off]) + add);
        split_f16(v1, p1[0], p1[lo_off]);
        const float v2 = (float)((double)join_f16(p2[0], p2[lo_off]) + add);
        split_f16(v2, p2[0], p2[lo_off]);
        continue;
      }
    }
    *p1 = to_type<T>(to_double<T>(*p1) + add);
    *p2 = to_type<T>(to_double<T>(*p2) + add);
  }
}

void launch_ao_add_coulomb(void* M, int dtype, int k, int n_ao, int64_t ld, const int* atom, const double* w,
                           hipStream_t s, int64_t lo_off) {
  const int g = grid_for((int64_t)k * n_ao);
  switch (dtype) {
    case PREC_FP64: ao_add_coulomb_kernel<double><<<g, 256, 0, s>>>((double*)M, k, n_ao, ld, atom, w, lo_off); break;
    case PREC_FP32: ao_add_coulomb_kernel<float><<<g, 256, 0, s>>>((float*)M, k, n_ao, ld, atom, w, lo_off); break;
    case PREC_BF16: ao_add_coulomb_kernel<__bf16><<<g, 256, 0, s>>>((__bf16*)M, k, n_ao, ld, atom, w, lo_off); break;
    default: ao_add_coulomb_kernel<_Float16><<<g, 256, 0, s>>>((_Float16*)M, k, n_ao, ld, atom, w, lo_off); break;
  }
  TDA_LAUNCH_CHECK();
}

__global__ void small_gemm_kernel(const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ C,
                                  int M, int N, int K) {
  const int64_t total = (int64_t)M * N;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int m = (int)(t / N), n = (int)(t % N);
    double acc = 0.0;
    for (int k = 0; k < K; ++k) acc += A[(int64_t)m * K + k] * B[(int64_t)k * N + n];
    C[t] = acc;
  }
}
void launch_small_gemm(const double* A, const double* B, double* C, int M, int N, int K, hipStream_t s) {
  const int64_t total = (int64_t)M * N;
  if (total == 0) return;
  small_gemm_kernel<<<grid_for(total), 256, 0, s>>>(A, B, C, M, N, K);
  TDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Mulliken transition dipoles and oscillator strengths (tda/moments.rs:7-24):
// mu[d][r] = sum_A xyz[A][d] q_trans[A][r],  f[r] = 2/3 w[r] |mu[:, r]|^2
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) dipoles_kernel(const double* __restrict__ xyz, const double* __restrict__ qt,
                                                      const double* __restrict__ w, int N, int n,
                                                      double* __restrict__ mu, double* __restrict__ f) {
  __shared__ double red[3][256];
  const int r = blockIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  for (int A = threadIdx.x; A < N; A += 256) {
    const double q = qt[(int64_t)A * n + r];
    for (int d = 0; d < 3; ++d) s[d] += xyz[(int64_t)A * 3 + d] * q;
  }
  for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = s[d];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h)
      for (int d = 0; d < 3; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double m2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      mu[(int64_t)d * n + r] = red[d][0];
      m2 += red[d][0] * red[d][0];
    }
    f[r] = 2.0 / 3.0 * w[r] * m2;
  }
}

void launch_dipoles(const double* xyz, const double* qt, const double* w, int N, int n, double* mu, double* f,
                    hipStream_t s) {
  if (n == 0) return;
  dipoles_kernel<<<n, 256, 0, s>>>(xyz, qt, w, N, n, mu, f);
  TDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Hellmann-Feynman Coulomb derivatives of w_k = x_k^T A x_k (A.x with the
// singlet Coulomb 2 q_ov^T gamma q_ov):  u = q_ov x_k [N][n], w = gamma u,
//   dgam[k][A][B] = d w_k / d gamma_AB = 2 u[A][k] u[B][k]
//   cq[k][A]: d w_k / d q_ov[A, ia] = cq[k][A] x_k[ia],  cq = 4 w[A][k]
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) hf_coulomb_kernel(const double* __restrict__ u, const double* __restrict__ w,
                                                         int N, int n, double* __restrict__ dgam,
                                                         double* __restrict__ cq) {
  const int A = blockIdx.x, k = blockIdx.y;
  const double uA = u[(int64_t)A * n + k];
  double* row = dgam + ((int64_t)k * N + A) * N;
  for (int B = threadIdx.x; B < N; B += 256) row[B] = 2.0 * uA * u[(int64_t)B * n + k];
  if (threadIdx.x == 0 && cq) cq[(int64_t)k * N + A] = 4.0 * w[(int64_t)A * n + k];
}

void launch_hf_coulomb(const double* u, const double* w, int N, int n, double* dgam, double* cq, hipStream_t s) {
  if (N == 0 || n == 0) return;
  hf_coulomb_kernel<<<dim3((unsigned)N, (unsigned)n), 256, 0, s>>>(u, w, N, n, dgam, cq);
  TDA_LAUNCH_CHECK();
}

// S[k][A][B] = D[k][A][B] + D[k][B][A]  (k < n, A, B < N)
__global__ void __launch_bounds__(256) sym_add_kernel(const double* __restrict__ D, int N,
                                                      double* __restrict__ S) {
  const int A = blockIdx.x, k = blockIdx.y;
  const double* d = D + (int64_t)k * N * N;
  double* o = S + ((int64_t)k * N + A) * N;
  for (int B = threadIdx.x; B < N; B += 256) o[B] = d[(int64_t)A * N + B] + d[(int64_t)B * N + A];
}

void launch_sym_add(const double* D, int n, int N, double* S, hipStream_t s) {
  if (n == 0 || N == 0) return;
  sym_add_kernel<<<dim3((unsigned)N, (unsigned)n), 256, 0, s>>>(D, N, S);
  TDA_LAUNCH_CHECK();
}

}  // namespace tda
