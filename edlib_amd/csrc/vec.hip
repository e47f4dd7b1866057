// Vector kernels (FP64): dot products, fused axpy + norm, the Lanczos
// recurrence bookkeeping, and the Gram-Schmidt kernels of the eigensolver.
// Reductions are two-stage and deterministic: each block reduces with
// wavefront shuffles into one partial, one block then sums the partials in a
// fixed order.
#include "edh_internal.h"
#include "edh_kernels.h"

namespace edh {

__device__ __forceinline__ double block_sum(double x, double* s_red) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_red[wid] = x;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += s_red[q];
  return t;  // valid on thread 0
}

__global__ void dot_partial_kernel(const double* a, const double* b, int64_t n, double* partial) {
  __shared__ double s_red[16];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc = fma(a[i], b[i], acc);
  double t = block_sum(acc, s_red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Sums `n` partials in a fixed order into *out (optionally its square root).
__global__ void reduce_partials_kernel(const double* partial, int n, double* out, int take_sqrt) {
  __shared__ double s_red[16];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  double t = block_sum(acc, s_red);
  if (threadIdx.x == 0) *out = take_sqrt ? sqrt(t) : t;
}

// w += (*coef) * x ; partial of |w|^2.  Skipped once the recurrence stopped.
__global__ void axpy_norm_kernel(double* w, const double* x, const double* coef, int64_t n, double* partial,
                                 const int* done) {
  __shared__ double s_red[16];
  if (done && *done) return;
  const double c = *coef;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double y = fma(c, x[i], w[i]);
    w[i] = y;
    acc = fma(y, y, acc);
  }
  double t = block_sum(acc, s_red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Lanczos bookkeeping (Lanczos.h:55-65).  sc[0] = scale of the current
// operand (v_k = s_k W_k), sc[1] = additive scale for the next step,
// sc[2] = pass-B coefficient (-alpha_k s_k).  st[0] = steps taken, st[1] = stop.
// k <= 0: the step number is read on the device (st[0] + 1), so one launch
// sequence serves every step (hipGraph replay).
__global__ void lanczos_alpha_kernel(const double* partial, int n, double* alpha, int k, double* sc, const int* st) {
  __shared__ double s_red[16];
  if (st[1]) return;
  if (k <= 0) k = st[0] + 1;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  double t = block_sum(acc, s_red);
  if (threadIdx.x == 0) {
    alpha[k - 1] = t;
    sc[2] = -t * sc[0];
  }
}

__global__ void lanczos_beta_kernel(const double* partial, int n, double* beta, int k, int nl, double* sc, int* st) {
  __shared__ double s_red[16];
  if (st[1]) return;
  if (k <= 0) k = st[0] + 1;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  double t = block_sum(acc, s_red);
  if (threadIdx.x == 0) {
    const double bet = sqrt(t);
    if (k != nl) beta[k] = bet;
    st[0] = k;
    if (fabs(bet) < 1e-10) {
      st[1] = 1;
    } else {
      sc[1] = -bet * sc[0];
      sc[0] = 1.0 / bet;
    }
  }
}

__global__ void lanczos_init_kernel(double* sc, int* st) {
  sc[0] = 1.0;
  sc[1] = 0.0;
  sc[2] = 0.0;
  st[0] = 0;
  st[1] = 0;
}

// ---- Gram-Schmidt against a block of basis vectors (eigensolver) ----------
// A pass is skipped when a gate is given and reads 0 (ARPACK dsaitr's DGKS
// rule: the second pass runs only when the first one removed more than
// 1 - 0.717 of the norm).
__device__ __forceinline__ bool gated_off(const double* gate) { return gate && *gate == 0.0; }

// h_partial[blk * ld + i] = sum_e V[i][e] w[e] for i < m (m <= kMaxBasis);
// with_nb: slot m of each block holds sum_e w[e]^2 (ld = m + 1, else m)
__global__ void multi_dot_kernel(const double* V, int64_t ldv, int m, const double* w, int64_t n, double* partial,
                                 int with_nb, const double* gate) {
  __shared__ double s_acc[kMaxBasis + 1][16];
  if (gated_off(gate)) return;
  double acc[kMaxBasis];
  double nb = 0.0;
#pragma unroll
  for (int i = 0; i < kMaxBasis; ++i) acc[i] = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double we = w[e];
#pragma unroll
    for (int i = 0; i < kMaxBasis; ++i)
      if (i < m) acc[i] = fma(V[i * ldv + e], we, acc[i]);
    if (with_nb) nb = fma(we, we, nb);
  }
  const int wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < kMaxBasis; ++i) {
    if (i >= m) break;
    double x = acc[i];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if ((threadIdx.x & 63) == 0) s_acc[i][wid] = x;
  }
  if (with_nb) {
    double x = nb;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if ((threadIdx.x & 63) == 0) s_acc[m][wid] = x;
  }
  __syncthreads();
  const int ld = m + (with_nb ? 1 : 0);
  for (int i = threadIdx.x; i < ld; i += blockDim.x) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += s_acc[i][q];
    partial[blockIdx.x * (int64_t)ld + i] = t;
  }
}

// h[i] = sum over blocks of partial[blk*ld + i]; also accumulates into hsum
// if given; with_nb: *nb = sqrt of the sum of slot m.
__global__ void multi_reduce_kernel(const double* partial, int nblk, int m, double* h, double* hsum, int with_nb,
                                    double* nb, const double* gate) {
  __shared__ double s_red[16];
  if (gated_off(gate)) return;
  const int ld = m + (with_nb ? 1 : 0);
  for (int i = 0; i < ld; ++i) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) acc += partial[b * (int64_t)ld + i];
    double t = block_sum(acc, s_red);
    if (threadIdx.x == 0) {
      if (i < m) {
        h[i] = t;
        if (hsum) hsum[i] += t;
      } else {
        *nb = sqrt(t);
      }
    }
    __syncthreads();
  }
}

// w -= sum_i h[i] V[i]; optional partial of |w|^2
__global__ void multi_axpy_kernel(const double* V, int64_t ldv, int m, const double* h, double* w, int64_t n,
                                  double* partial, const double* gate) {
  __shared__ double s_h[kMaxBasis];
  __shared__ double s_red[16];
  if (gated_off(gate)) return;
  for (int i = threadIdx.x; i < m; i += blockDim.x) s_h[i] = h[i];
  __syncthreads();
  double acc = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double x = w[e];
    for (int i = 0; i < m; ++i) x = fma(-s_h[i], V[i * ldv + e], x);
    w[e] = x;
    acc = fma(x, x, acc);
  }
  if (partial) {
    double t = block_sum(acc, s_red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// *nrm = sqrt(sum of the partials); gate_out: *gate_out = (*nrm < 0.717 *nb)
__global__ void ortho_norm_kernel(const double* partial, int n, double* nrm, const double* nb, double* gate_out,
                                  const double* gate) {
  __shared__ double s_red[16];
  if (gated_off(gate)) return;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  double t = block_sum(acc, s_red);
  if (threadIdx.x == 0) {
    *nrm = sqrt(t);
    if (gate_out) *gate_out = *nrm < 0.717 * *nb ? 1.0 : 0.0;
  }
}

// out = x * (1 / *nrm)
__global__ void scale_inv_kernel(const double* x, const double* nrm, double* out, int64_t n) {
  const double s = 1.0 / *nrm;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = x[e] * s;
}

// Y (k columns) = V (m columns) * C (m x k, column-major, host-uploaded), per element.
__global__ void basis_combine_kernel(const double* V, int64_t ldv, int m, const double* C, int k, double* Y,
                                     int64_t ldy, int64_t n) {
  __shared__ double s_c[kMaxBasis * kMaxBasis];
  for (int i = threadIdx.x; i < m * k; i += blockDim.x) s_c[i] = C[i];
  __syncthreads();
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double v[kMaxBasis];
#pragma unroll
    for (int i = 0; i < kMaxBasis; ++i)
      if (i < m) v[i] = V[i * ldv + e];
    for (int j = 0; j < k; ++j) {
      double y = 0.0;
#pragma unroll
      for (int i = 0; i < kMaxBasis; ++i)
        if (i < m) y = fma(s_c[j * m + i], v[i], y);
      Y[j * ldy + e] = y;
    }
  }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
int reduce_grid(int64_t n) {
  int64_t g = (n + kVecBlock - 1) / kVecBlock;
  if (g > kMaxPartials) g = kMaxPartials;
  if (g < 1) g = 1;
  return (int)g;
}

void dev_dot(const double* a, const double* b, int64_t n, double* partial, double* out, hipStream_t s) {
  int g = reduce_grid(n);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(g), dim3(kVecBlock), 0, s, a, b, n, partial);
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kVecBlock), 0, s, partial, g, out, 0);
  EDH_HIP(hipGetLastError());
}

void dev_norm(const double* a, int64_t n, double* partial, double* out, hipStream_t s) {
  int g = reduce_grid(n);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(g), dim3(kVecBlock), 0, s, a, a, n, partial);
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kVecBlock), 0, s, partial, g, out, 1);
  EDH_HIP(hipGetLastError());
}

__global__ void fill_kernel(double* p, int64_t n, double x) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) p[e] = x;
}

// x[i] = uniform [-1, 1) value number i + 1 of the splitmix64 sequence
// started at st0 (state st0 + (i + 1) * 0x9E3779B97F4A7C15, then the
// splitmix64 output mix), top 53 bits scaled by 2^-52, minus 1: element i of
// the host generator's sequence, produced independently per element.
__global__ void splitmix_uniform_kernel(double* x, int64_t n, uint64_t st0) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = st0 + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

void dev_random_uniform(double* x, int64_t n, uint64_t seed, hipStream_t s) {
  const uint64_t st0 = seed ? seed : 0x243F6A8885A308D3ull;
  const int g = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(splitmix_uniform_kernel, dim3(std::max(g, 1)), dim3(256), 0, s, x, n, st0);
  EDH_HIP(hipGetLastError());
}

void dev_fill(double* p, int64_t n, double x, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(reduce_grid(n)), dim3(kVecBlock), 0, s, p, n, x);
  EDH_HIP(hipGetLastError());
}

void dev_scale_inv(const double* x, const double* nrm, double* out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(scale_inv_kernel, dim3(reduce_grid(n)), dim3(kVecBlock), 0, s, x, nrm, out, n);
  EDH_HIP(hipGetLastError());
}

void dev_orthogonalize(const double* V, int64_t ldv, int m, double* w, int64_t n, double* partial, double* h,
                       double* hsum, double* nrm, hipStream_t s, double* nb, const double* gate, double* gate_out) {
  EDH_CHECK(m >= 1 && m <= kMaxBasis, EDH_ERR_LIMIT, "Krylov basis size outside the device limit");
  EDH_CHECK(!gate_out || nb, EDH_ERR_INTERNAL, "DGKS gate needs the norm before the pass");
  const int g = reduce_grid(n);
  const int with_nb = nb ? 1 : 0;
  hipLaunchKernelGGL(multi_dot_kernel, dim3(g), dim3(kVecBlock), 0, s, V, ldv, m, w, n, partial, with_nb, gate);
  hipLaunchKernelGGL(multi_reduce_kernel, dim3(1), dim3(kVecBlock), 0, s, partial, g, m, h, hsum, with_nb, nb, gate);
  hipLaunchKernelGGL(multi_axpy_kernel, dim3(g), dim3(kVecBlock), 0, s, V, ldv, m, h, w, n, partial, gate);
  hipLaunchKernelGGL(ortho_norm_kernel, dim3(1), dim3(kVecBlock), 0, s, partial, g, nrm, nb, gate_out, gate);
  EDH_HIP(hipGetLastError());
}

void dev_combine(const double* V, int64_t ldv, int m, const double* C, int k, double* Y, int64_t ldy, int64_t n,
                 hipStream_t s) {
  EDH_CHECK(m <= kMaxBasis && k <= kMaxBasis, EDH_ERR_LIMIT, "Krylov basis larger than the device limit");
  hipLaunchKernelGGL(basis_combine_kernel, dim3(reduce_grid(n)), dim3(kVecBlock), 0, s, V, ldv, m, C, k, Y, ldy,
                     n);
  EDH_HIP(hipGetLastError());
}

void dev_lanczos_init(double* sc, int* st, hipStream_t s) {
  hipLaunchKernelGGL(lanczos_init_kernel, dim3(1), dim3(1), 0, s, sc, st);
  EDH_HIP(hipGetLastError());
}

void dev_lanczos_alpha(const double* partial, int n, double* alpha, int k, double* sc, const int* st, hipStream_t s) {
  hipLaunchKernelGGL(lanczos_alpha_kernel, dim3(1), dim3(kVecBlock), 0, s, partial, n, alpha, k, sc, st);
  EDH_HIP(hipGetLastError());
}

void dev_lanczos_passB(double* w, const double* x, const double* coef, int64_t n, double* partial, const int* done,
                       double* beta, int k, int nl, double* sc, int* st, hipStream_t s) {
  int g = reduce_grid(n);
  hipLaunchKernelGGL(axpy_norm_kernel, dim3(g), dim3(kVecBlock), 0, s, w, x, coef, n, partial, done);
  hipLaunchKernelGGL(lanczos_beta_kernel, dim3(1), dim3(kVecBlock), 0, s, partial, g, beta, k, nl, sc, st);
  EDH_HIP(hipGetLastError());
}

}  // namespace edh
