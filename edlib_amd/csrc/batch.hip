// Batched Krylov engine: the batching and serving forms of the solver.
//
// Many independent Krylov processes -- eigensolver runs (the thick-restart
// Lanczos that replaces ARPACK in Storage::diag, Storage.h:43-146) and
// Green's-function Lanczos runs (Lanczos::lanczos, Lanczos.h:36-77) -- on
// sectors of different models and dimensions advance together.  A sweep
// advances every active process by one Krylov step and each phase of a
// sweep is ONE launch over all of them: the grouped H*v over the
// concatenated up-rows of every operand, then grouped projections, updates,
// norms and scalings over fixed chunks of every vector.  All descriptors of
// a sweep go to the device in one copy from pinned memory, the Lanczos
// coefficients and stop flags stay on the device, and the host waits only
// when an eigensolver run ends a restart cycle or a Lanczos run ends.
// Processes join and leave between sweeps (continuous batching).
//
// Re-orthogonalisation follows ARPACK's dsaitr: one classical Gram-Schmidt
// pass, and a second one only when the first removed more than 1 - 0.717
// of the norm (the DGKS test), decided on the device per process.
//
// Determinism: a process's arithmetic depends only on its own operands.  Its
// H*v elements do not depend on the grouping, its reductions use its own
// chunks (a size fixed by its own dimension) reduced in a fixed order, and its host-side
// logic (projected matrix, restarts) sees only its own numbers.  A result
// is therefore bit-identical alone and in any mix.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>

#include "batch.h"

namespace edh {

namespace {

constexpr int kChunkMax = 2048;  // elements per reduction chunk (large operands)
constexpr int kChunkMin = 256;   // ... (small operands: more blocks in flight)
constexpr int kChunkHuge = 65536;  // ... (operands above 4 M elements)
constexpr int kOpBlock = 256;  // threads per chunk kernel block
constexpr int kHGroup = 8;     // projections accumulated per pass over a chunk
constexpr double kDgks = 0.717;

// scalar block layout (doubles)
constexpr int S_NRM = kMaxBasis;         // current norm
constexpr int S_ONE = kMaxBasis + 1;     // 1.0  } H*v scales of a Lanczos step:
constexpr int S_NEGNRM = kMaxBasis + 2;  // -nrm }  v scale, w scale
constexpr int S_NB = kMaxBasis + 3;      // norm before the first projection pass
constexpr int S_GATE = kMaxBasis + 4;    // 1 when the second pass is needed
constexpr int S_LZV = kMaxBasis + 5;     // } Lanczos run: scale of the stored current vector (1 / beta_{k-1})
constexpr int S_LZW = kMaxBasis + 6;     // }   and of the previous one in w = H v_k - beta_{k-1} v_{k-1}
constexpr int S_LZK = kMaxBasis + 7;     // Lanczos run: its next step k (1-based)
constexpr int S_T = 2 * kMaxBasis;       // projected-matrix columns / Lanczos history

// Eigensolver control block (after the projected-matrix block of a run's
// scalars): the fields the device kernels keep, then the Ritz values of a
// finished run, the Ritz values kept by the last restart and the
// combination matrix of the last cycle end.
constexpr int C_K = 3;        // kept Ritz vectors (columns 0..k-1 of the basis)
constexpr int C_ITER = 4;     // restart cycles done
constexpr int C_NCONV = 5;    // converged Ritz pairs at the last cycle end
constexpr int C_NAV = 6;      // H*v products
constexpr int C_PENDING = 7;  // 1: thick restart, 2: Ritz vectors to combine this sweep
constexpr int C_SRC = 8;      // basis buffer the combination reads
constexpr int C_ROWS = 9;     // rows and columns of the combination matrix
constexpr int C_COLS = 10;
constexpr int C_RITZ0 = 11;   // lowest Ritz value at the last cycle end (an upper bound of the lowest eigenvalue)
constexpr int C_HDR = 16;     // header size; Ritz values follow
__host__ __device__ inline int ctl_theta(int m) { return C_HDR + m; }
__host__ __device__ inline int ctl_cmat(int m) { return C_HDR + 2 * m; }
__host__ __device__ inline size_t ctl_size(int m) { return (size_t)C_HDR + 2 * m + (size_t)(m + 1) * (m + 1); }

struct BOp {
  const double* X;
  int64_t ldx;
  int cnt;
  double* w;
  int64_t n;
  int chunk0, nch, chunk;
  double* sc;
  double* hsum;
  double* hist_a;
  double* hist_b;
  int* flags;
  int kstep;
  int mode;
  double scale;
  double* nrm_out;
  const double* ctl;  // eigensolver control block or null (then X1, m as below)
  const double* X1;   // second basis buffer
  int m;              // basis size
  const double* tdot;  // Lanczos step: per-task shares of <v, w>
  int64_t ntdot;
  const double* lzk;  // Lanczos step counter (operands in rotation) or null
};

// An eigensolver operation follows its run's control block: the operand is
// column col + 1 of the current basis buffer, columns 0..col are projected
// out and the projections accumulate into column col of the projected
// matrix.
__device__ __forceinline__ BOp resolve(const BOp& in) {
  BOp o = in;
  if (o.ctl) {
    const int64_t col = (int64_t)o.ctl[kCtlCol];
    const double* X = o.ctl[kCtlPhase] != 0.0 ? o.X1 : o.X;
    o.X = X;
    o.cnt = (int)col + 1;
    o.w = const_cast<double*>(X) + (col + 1) * o.ldx;
    if (o.hsum) o.hsum = o.sc + S_T + col * (o.m + 1);
  } else if (o.lzk) {
    // a Lanczos step from the run's own counter: the current vector and
    // the H*v output in their rotation, the k-th history entries
    // (Lanczos.h:47-64; beta_k is not stored at the last step)
    const int k = (int)*o.lzk;
    const int r0 = lz_buffer(k, 0), r2 = lz_buffer(k, 2);
    const double* b0 = o.X;
    const double* b1 = o.X1;
    const double* b2 = o.w;
    o.X = r0 == 0 ? b0 : (r0 == 1 ? b1 : b2);
    o.w = const_cast<double*>(r2 == 0 ? b0 : (r2 == 1 ? b1 : b2));
    o.hist_a = o.hist_a + (k - 1);
    o.hist_b = k != o.m ? o.hist_b + k : nullptr;
    o.kstep = k;
  }
  return o;
}

__device__ __forceinline__ bool op_skip(const BOp& o) {
  if (o.flags && o.flags[0]) return true;
  if (o.ctl && o.ctl[kCtlState] != 0.0) return true;
  if ((o.mode & OP_GATED) && o.sc[S_GATE] == 0.0) return true;
  return false;
}

__device__ __forceinline__ double block_sum256(double x, double* s_red) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int wid = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_red[wid] = x;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += s_red[q];
  return t;  // valid on thread 0
}

// Reduces `cnt` per-thread values over the block (fixed order: shuffle
// tree per wavefront, then the four wavefronts in order) and writes them to
// out[0..cnt); one barrier for the whole set.
template <int K>
__device__ __forceinline__ void block_store_many(const double (&v)[K], int cnt, double (*s_part)[K], double* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int g = 0; g < K; ++g) {
    if (g >= cnt) break;
    double x = v[g];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) s_part[wave][g] = x;
  }
  __syncthreads();
  if ((int)threadIdx.x < cnt) {
    double t = 0.0;
    for (int q = 0; q < kOpBlock / 64; ++q) t += s_part[q][threadIdx.x];
    out[threadIdx.x] = t;
  }
  __syncthreads();
}

// partial[c * kMaxBasis + i] = chunk c's sum of X_i[e] * w[e], i < cnt, and
// (OP_NB) slot cnt = chunk c's sum of w[e]^2
template <int KG>
__global__ __launch_bounds__(kOpBlock) void bdots_kernel(const BOp* __restrict__ ops, const int* __restrict__ cjob,
                                                         double* __restrict__ partial) {
  __shared__ double s_part[kOpBlock / 64][KG];
  const int c = blockIdx.x;
  const BOp o = resolve(ops[cjob[c]]);
  if (op_skip(o)) return;
  const int lc = c - o.chunk0;
  const int64_t e0 = (int64_t)lc * o.chunk;
  const int64_t e1 = min(o.n, e0 + o.chunk);
  const int tot = o.cnt + ((o.mode & OP_NB) ? 1 : 0);
  for (int i0 = 0; i0 < tot; i0 += KG) {
    double acc[KG];
#pragma unroll
    for (int g = 0; g < KG; ++g) acc[g] = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kOpBlock) {
      // all loads of the element first, then the FMAs
      const double we = o.w[e];
      double xv[KG];
#pragma unroll
      for (int g = 0; g < KG; ++g) {
        const int i = i0 + g;
        xv[g] = i < o.cnt ? o.X[(int64_t)i * o.ldx + e] : (i < tot ? we : 0.0);
      }
#pragma unroll
      for (int g = 0; g < KG; ++g) acc[g] = fma(xv[g], we, acc[g]);
    }
    block_store_many<KG>(acc, min(KG, tot - i0), s_part, partial + (size_t)c * kMaxBasis + i0);
  }
}

// First pass's update fused with the second pass's projections (cnt <=
// K): w -= sum_i h[i] X_i; then, from the same registers, chunk c's sums of
// X_i[e] * w_new[e] in slots i < cnt and of w_new[e]^2 in slot cnt.  K is
// the group's largest basis size rounded up to 12 or kFuse: the registers
// it takes set the occupancy (K = 24 runs the NCV = 11 basis ~10 % slower
// in the whole solve).
constexpr int kFuse = 24;
constexpr int kFuseSmall = 12;
template <int K>
__global__ __launch_bounds__(kOpBlock) void bupdate_dots_kernel(const BOp* __restrict__ ops,
                                                                const int* __restrict__ cjob,
                                                                double* __restrict__ partial) {
  __shared__ double s_part[kOpBlock / 64][K + 1];
  __shared__ double s_h[K];
  const int c = blockIdx.x;
  const BOp o = resolve(ops[cjob[c]]);
  if (op_skip(o)) return;
  if ((int)threadIdx.x < K) s_h[threadIdx.x] = (int)threadIdx.x < o.cnt ? o.sc[threadIdx.x] : 0.0;
  __syncthreads();
  const int lc = c - o.chunk0;
  const int64_t e0 = (int64_t)lc * o.chunk;
  const int64_t e1 = min(o.n, e0 + o.chunk);
  const int cnt = o.cnt;
  double acc[K + 1];
#pragma unroll
  for (int g = 0; g <= K; ++g) acc[g] = 0.0;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kOpBlock) {
    double xv[K];
#pragma unroll
    for (int g = 0; g < K; ++g) xv[g] = g < cnt ? o.X[(int64_t)g * o.ldx + e] : 0.0;
    double x = o.w[e];
#pragma unroll
    for (int g = 0; g < K; ++g) x = fma(-s_h[g], xv[g], x);
    o.w[e] = x;
#pragma unroll
    for (int g = 0; g < K; ++g) acc[g] = fma(xv[g], x, acc[g]);
    acc[K] = fma(x, x, acc[K]);
  }
  // slot layout: projections 0..cnt-1, squared norm at cnt
  double v[K + 1];
#pragma unroll
  for (int g = 0; g <= K; ++g) v[g] = g < cnt ? acc[g] : (g == cnt ? acc[K] : 0.0);
  block_store_many<K + 1>(v, cnt + 1, s_part, partial + (size_t)c * kMaxBasis);
}

// One block (four wavefronts) per process.  mode 0: h = its chunks'
// partials, each coefficient summed by one wavefront (lane l takes chunks l,
// l+64, ...; then a fixed shuffle tree), into hsum too, alpha history, norm
// before.  mode 1: norm of the updated vector, DGKS gate, Lanczos beta
// history and stop test.
// Lane l takes chunks l, l + 64, ...: eight at a time into eight running
// sums (their loads in flight together; an operand of 11.8 M elements has
// 5,751 chunks), the rest into the first, then the eight in a fixed tree
// (below 512 chunks only the first is used: the plain running sum).
__device__ __forceinline__ double wave_chunk_sum(const BOp& o, const double* partial, int i, int lane) {
  const double* __restrict__ p = partial + (size_t)o.chunk0 * kMaxBasis + i;
  double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0, t5 = 0.0, t6 = 0.0, t7 = 0.0;
  int lc = lane;
  for (; lc + 7 * 64 < o.nch; lc += 8 * 64) {
    t0 += p[(size_t)lc * kMaxBasis];
    t1 += p[(size_t)(lc + 64) * kMaxBasis];
    t2 += p[(size_t)(lc + 128) * kMaxBasis];
    t3 += p[(size_t)(lc + 192) * kMaxBasis];
    t4 += p[(size_t)(lc + 256) * kMaxBasis];
    t5 += p[(size_t)(lc + 320) * kMaxBasis];
    t6 += p[(size_t)(lc + 384) * kMaxBasis];
    t7 += p[(size_t)(lc + 448) * kMaxBasis];
  }
  for (; lc < o.nch; lc += 64) t0 += p[(size_t)lc * kMaxBasis];
  double t = ((t0 + t1) + (t2 + t3)) + ((t4 + t5) + (t6 + t7));
  for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
  return t;  // valid on lane 0
}

constexpr int kReduceBlock = 1024;  // breduce_kernel: one wavefront per coefficient of an operation
__global__ __launch_bounds__(kReduceBlock) void breduce_kernel(const BOp* __restrict__ ops,
                                                               const double* __restrict__ partial, int mode,
                                                               int gated) {
  const BOp o = resolve(ops[blockIdx.x]);
  if (op_skip(o) || (gated && o.sc[S_GATE] == 0.0)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (mode == 2) {
    // after the fused update: norm in slot cnt; DGKS gate; when the gate is
    // set, the second pass's projections go to h and into hsum
    // (wave 0 sums the norm while waves 1.. sum the projections, which are
    // stored only when the gate is set)
    __shared__ int s_gate;
    const int nw = (int)(blockDim.x >> 6);
    const int ip = wave - 1;
    double tp = 0.0;
    if (wave == 0) {
      const double t = wave_chunk_sum(o, partial, o.cnt, lane);
      if (lane == 0) {
        const double nrm = sqrt(t);
        o.sc[S_NRM] = nrm;
        o.sc[S_ONE] = 1.0;
        o.sc[S_NEGNRM] = -nrm;
        const int g = nrm < kDgks * o.sc[S_NB] ? 1 : 0;
        o.sc[S_GATE] = g;
        s_gate = g;
      }
    } else if (ip < o.cnt) {
      tp = wave_chunk_sum(o, partial, ip, lane);
    }
    __syncthreads();
    if (!s_gate) return;
    if (wave > 0) {
      for (int i = ip; i < o.cnt; i += nw - 1) {
        const double t = i == ip ? tp : wave_chunk_sum(o, partial, i, lane);
        if (lane == 0) {
          o.sc[i] = t;
          if (o.hsum) o.hsum[i] += t;
        }
      }
    }
    if (o.mode & OP_SCALE_NRM) {
      // the second pass and the normalisation run as one update: its norm
      // is |w'|^2 - |s|^2 (w' - X s with s = X^T w' and orthonormal X),
      // accurate to rounding since |s| << |w'|
      __syncthreads();
      if (threadIdx.x == 0) {
        double ss = 0.0;
        for (int i = 0; i < o.cnt; ++i) ss = fma(o.sc[i], o.sc[i], ss);
        const double n2 = o.sc[S_NRM] * o.sc[S_NRM] - ss;
        const double nrm = sqrt(fmax(n2, 0.0));
        o.sc[S_NRM] = nrm;
        o.sc[S_NEGNRM] = -nrm;
      }
    }
    return;
  }
  if (mode == 0) {
    const int tot = o.cnt + ((o.mode & OP_NB) ? 1 : 0);
    for (int i = wave; i < tot; i += (int)(blockDim.x >> 6)) {
      const double t = wave_chunk_sum(o, partial, i, lane);
      if (lane == 0) {
        if (i < o.cnt) {
          o.sc[i] = t;
          if (o.hsum) o.hsum[i] += t;
          if (i == 0 && o.hist_a) *o.hist_a = t;
        } else {
          o.sc[S_NB] = sqrt(t);
        }
      }
    }
  } else if (wave == 0) {
    const double t = wave_chunk_sum(o, partial, 0, lane);
    if (lane == 0) {
      const double nrm = sqrt(t);
      o.sc[S_NRM] = nrm;
      o.sc[S_ONE] = 1.0;
      o.sc[S_NEGNRM] = -nrm;
      if (o.mode & OP_NB) o.sc[S_GATE] = nrm < kDgks * o.sc[S_NB] ? 1.0 : 0.0;
      if (o.mode & OP_START) {
        // GreensFunction.h:408-418: a start vector of squared norm <= 1e-10
        // has no run; otherwise the run starts from it normalised
        const double n2 = nrm * nrm;
        *o.nrm_out = n2;
        if (fabs(n2) > 1e-10) {
          o.sc[S_LZV] = 1.0 / nrm;
        } else {
          o.sc[S_LZV] = 0.0;
          o.flags[0] = 1;
        }
      } else if (o.nrm_out) {
        *o.nrm_out = nrm;
      }
      if (o.mode & OP_LZ) {  // Lanczos.h:61-65
        if (o.hist_b) *o.hist_b = nrm;
        o.flags[1] = o.kstep;
        if (fabs(nrm) < 1e-10) o.flags[0] = 1;
        if (o.lzk) o.sc[S_LZK] = o.kstep + 1;  // the step is complete (the last read of k in this sweep)
      }
      if (o.mode & OP_LZS) {
        // next step: v_{k+1} = w / beta_k stays stored as w, and the term
        // beta_k v_k = beta_k (scale of v_k) u_k
        o.sc[S_LZW] = -nrm * o.sc[S_LZV];
        o.sc[S_LZV] = 1.0 / nrm;
      }
    }
  }
}

// w[e] -= sum_i h[i] X_i[e]; partial[c * kMaxBasis] = sum of the new w[e]^2.
// The projections are applied KG at a time, so KG loads per element are in
// flight together.
template <int KG>
__global__ __launch_bounds__(kOpBlock) void bupdate_kernel(const BOp* __restrict__ ops, const int* __restrict__ cjob,
                                                           double* __restrict__ partial, int gated) {
  __shared__ double s_red[8];
  __shared__ double s_h[kMaxBasis + KG];
  const int c = blockIdx.x;
  const BOp o = resolve(ops[cjob[c]]);
  if (op_skip(o) || (gated && o.sc[S_GATE] == 0.0)) return;
  for (int i = threadIdx.x; i < kMaxBasis + KG; i += blockDim.x) s_h[i] = i < o.cnt ? o.sc[i] : 0.0;
  __syncthreads();
  const int lc = c - o.chunk0;
  const int64_t e0 = (int64_t)lc * o.chunk;
  const int64_t e1 = min(o.n, e0 + o.chunk);
  double nsq = 0.0;
  if (o.cnt <= 1) {
    // a Lanczos update (one projection) or a norm: four elements per
    // iteration with their loads in flight together, the same operations
    // in the same order as the general loop (whose remaining FMAs with
    // zero coefficients leave x unchanged)
    const double h0 = s_h[0];
    const bool upd = o.cnt == 1;
    int64_t e = e0 + threadIdx.x;
    for (; e + 3 * kOpBlock < e1; e += 4 * kOpBlock) {
      double x[4], xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = o.w[e + u * kOpBlock];
      if (upd) {
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = o.X[e + u * kOpBlock];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (upd) {
          x[u] = fma(-h0, xv[u], x[u]);
          o.w[e + u * kOpBlock] = x[u];
        }
        nsq = fma(x[u], x[u], nsq);
      }
    }
    for (; e < e1; e += kOpBlock) {
      double x = o.w[e];
      if (upd) {
        x = fma(-h0, o.X[e], x);
        o.w[e] = x;
      }
      nsq = fma(x, x, nsq);
    }
  } else {
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kOpBlock) {
      double x = o.w[e];
      for (int i0 = 0; i0 < o.cnt; i0 += KG) {
        double xv[KG];
#pragma unroll
        for (int g = 0; g < KG; ++g) xv[g] = i0 + g < o.cnt ? o.X[(int64_t)(i0 + g) * o.ldx + e] : 0.0;
#pragma unroll
        for (int g = 0; g < KG; ++g) x = fma(-s_h[i0 + g], xv[g], x);
      }
      if (o.cnt) o.w[e] = x;
      nsq = fma(x, x, nsq);
    }
  }
  const double t = block_sum256(nsq, s_red);
  if (threadIdx.x == 0) partial[(size_t)c * kMaxBasis] = t;
}

// Second projection pass and normalisation in one update: w = (w - sum_i
// s_i X_i) / nrm where the DGKS gate is set, else w = w / nrm (s and nrm
// from breduce_kernel mode 2).
template <int KG>
__global__ __launch_bounds__(kOpBlock) void bupdate_scale_kernel(const BOp* __restrict__ ops,
                                                                 const int* __restrict__ cjob) {
  __shared__ double s_h[kMaxBasis + KG];
  const int c = blockIdx.x;
  const BOp o = resolve(ops[cjob[c]]);
  if (op_skip(o)) return;
  const int cnt = o.sc[S_GATE] != 0.0 ? o.cnt : 0;
  for (int i = threadIdx.x; i < kMaxBasis + KG; i += blockDim.x) s_h[i] = i < cnt ? o.sc[i] : 0.0;
  __syncthreads();
  const double f = 1.0 / o.sc[S_NRM];
  const int lc = c - o.chunk0;
  const int64_t e0 = (int64_t)lc * o.chunk;
  const int64_t e1 = min(o.n, e0 + o.chunk);
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kOpBlock) {
    double x = o.w[e];
    for (int i0 = 0; i0 < cnt; i0 += KG) {
      double xv[KG];
#pragma unroll
      for (int g = 0; g < KG; ++g) xv[g] = i0 + g < cnt ? o.X[(int64_t)(i0 + g) * o.ldx + e] : 0.0;
#pragma unroll
      for (int g = 0; g < KG; ++g) x = fma(-s_h[i0 + g], xv[g], x);
    }
    o.w[e] = x * f;
  }
}

// w[e] *= 1 / nrm (OP_SCALE_NRM) or *= scale
__global__ __launch_bounds__(kOpBlock) void bscale_kernel(const BOp* __restrict__ ops, const int* __restrict__ cjob) {
  const int c = blockIdx.x;
  const BOp o = resolve(ops[cjob[c]]);
  if (op_skip(o)) return;
  const double f = (o.mode & OP_SCALE_NRM) ? 1.0 / o.sc[S_NRM] : o.scale;
  const int lc = c - o.chunk0;
  const int64_t e0 = (int64_t)lc * o.chunk;
  const int64_t e1 = min(o.n, e0 + o.chunk);
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kOpBlock) o.w[e] *= f;
}

// Lanczos step (one block per run): alpha = sum of the grouped H*v's
// per-task shares of <v, w> (thread l takes shares l, l + 256, ... into four
// running sums, then a fixed block tree); the update coefficient for the
// stored (unnormalised) current vector is alpha times its scale.
__global__ __launch_bounds__(256) void blz_alpha_kernel(const BOp* __restrict__ ops) {
  __shared__ double s_red[8];
  const BOp o = resolve(ops[blockIdx.x]);
  if (op_skip(o)) return;
  double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
  int64_t q = threadIdx.x;
  for (; q + 768 < o.ntdot; q += 1024) {
    t0 += o.tdot[q];
    t1 += o.tdot[q + 256];
    t2 += o.tdot[q + 512];
    t3 += o.tdot[q + 768];
  }
  for (; q < o.ntdot; q += 256) t0 += o.tdot[q];
  const double t = block_sum256((t0 + t1) + (t2 + t3), s_red);
  if (threadIdx.x == 0) {
    *o.hist_a = t;
    o.sc[0] = t * o.sc[S_LZV];
  }
}

__global__ void bfill_kernel(double* p, int64_t n, double x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = x;
}

// ---------------------------------------------------------------------------
// Restart-cycle ends on the device.  After every sweep one 64-lane block per
// eigensolver run advances its column; at a cycle end it assembles the
// projected matrix, diagonalises it, tests convergence (ARPACK's bound
// |beta_m s_{m,i}| <= tol max(|theta_i|, eps^(2/3))) and either plans the
// thick restart or the Ritz vectors; the grouped combination launched next
// carries that plan out.  The host only reads the run's state back.
// ---------------------------------------------------------------------------

// (the single-sector eigensolver's restart_size: ARPACK dsaup2's rule)
__host__ __device__ inline int restart_keep(int nev, int m, int nconv) {
  int kk = nev + (nconv < (m - nev) / 2 ? nconv : (m - nev) / 2);
  if (kk == 1) kk = m >= 6 ? m / 2 : (m > 2 ? 2 : 1);
  if (kk > m - 2) kk = m - 2;
  return kk < 1 ? 1 : kk;
}

__device__ __forceinline__ double wave_sum(double x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;  // on every lane
}

// Value of x held by lane l (l uniform over the wavefront).
__device__ __forceinline__ double lane_value(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Eigen-decomposition of the symmetric n x n matrix a (row-major, LDS) by
// one 64-lane block: Householder reduction to tridiagonal form accumulated
// in q, implicit QL with shifts, then a stable ascending sort (the host
// sym_eig's algorithm).  The vector operations are spread over the lanes
// (sums by shuffle trees); during QL lane i holds d_i and e_i in registers,
// every lane evaluates the same scalar recurrence from values read across
// lanes, and lane r rotates row r of q (stored transposed, qt[c n + r]).  On
// return w holds the eigenvalues and column j of a the eigenvector of w[j].
// Returns false when QL does not converge.
// 1/sqrt(x) for x > 0 (normal): the hardware estimate refined by two
// Newton-Raphson steps y <- y + y (1 - x y^2) / 2, each of which doubles the
// number of correct bits, to within an ulp or so.
__device__ __forceinline__ double rsqrt_newton(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

__device__ bool wave_sym_eig(int n, double* a, double* qt, double* v, double* qq, double* w, int* order) {
  const int lane = threadIdx.x;
  for (int i = lane; i < n * n; i += 64) qt[i] = (i / n == i % n) ? 1.0 : 0.0;
  __syncthreads();
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    const double xi = lane < m ? a[(k + 1 + lane) * n + k] : 0.0;
    const double nx = sqrt(wave_sum(xi * xi));
    if (nx == 0.0) continue;
    const double x0 = lane_value(xi, 0);
    const double alpha = x0 > 0 ? -nx : nx;
    double vi = lane < m ? xi - (lane == 0 ? alpha : 0.0) : 0.0;
    const double nv = sqrt(wave_sum(vi * vi));
    if (nv == 0.0) continue;
    vi /= nv;
    if (lane < m) v[lane] = vi;
    __syncthreads();
    double pi = 0.0;
    if (lane < m) {
#pragma unroll 4
      for (int j = 0; j < m; ++j) pi += a[(k + 1 + lane) * n + k + 1 + j] * v[j];
    }
    const double K = wave_sum(vi * pi);
    const double qi = pi - K * vi;
    if (lane < m) qq[lane] = qi;
    __syncthreads();
    if (lane < m) {
#pragma unroll 4
      for (int j = 0; j < m; ++j) a[(k + 1 + lane) * n + k + 1 + j] -= 2.0 * (vi * qq[j] + qi * v[j]);
    }
    if (lane == 0) {
      a[(k + 1) * n + k] = alpha;
      a[k * n + k + 1] = alpha;
    }
    if (lane >= 1 && lane < m) {
      a[(k + 1 + lane) * n + k] = 0.0;
      a[k * n + k + 1 + lane] = 0.0;
    }
    if (lane < n) {
      double sum = 0.0;
#pragma unroll 4
      for (int j = 0; j < m; ++j) sum += qt[(k + 1 + j) * n + lane] * v[j];
#pragma unroll 4
      for (int j = 0; j < m; ++j) qt[(k + 1 + j) * n + lane] -= 2.0 * sum * v[j];
    }
    __syncthreads();
  }
  double dr = lane < n ? a[lane * n + lane] : 0.0;
  double er = lane + 1 < n ? a[(lane + 1) * n + lane] : 0.0;
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      // m = the first index >= l whose off-diagonal is negligible (else
      // n - 1): every lane tests its own, one ballot
      const double dnext = __shfl_down(dr, 1, 64);
      const bool split = lane < n - 1 && fabs(er) <= eps * (fabs(dr) + fabs(dnext));
      const unsigned long long hits = __ballot(split) & (~0ull << l);
      const int m = hits ? __ffsll((long long)hits) - 1 : n - 1;
      if (m == l) break;
      if (++iter > 200) return false;
      const double dl = lane_value(dr, l), el = lane_value(er, l);
      double g = (lane_value(dr, l + 1) - dl) / (2.0 * el);
      double r = sqrt(fma(g, g, 1.0));
      g = lane_value(dr, m) - dl + el / (g + copysign(r, g));
      double sn = 1.0, c = 1.0, pp = 0.0;
      bool early = false;
      // the eigenvector rows: rotation i turns rows i and i + 1, and row
      // i + 1 is the row that rotation i + 1 produced, so it is carried in a
      // register instead of being read back
      double carry = lane < n ? qt[m * n + lane] : 0.0;
      for (int i = m - 1; i >= l; --i) {
        const double ei = lane_value(er, i);
        const double f = sn * ei;
        const double b = c * ei;
        // the entries are moderate (Ritz values and couplings of H), so
        // f^2 + g^2 cannot overflow; r and 1/r from one reciprocal square
        // root (the rotation's latency chain is the QL sweep's cost)
        const double x2 = fma(f, f, g * g);
        if (x2 == 0.0) {
          if (lane == i + 1) {
            er = 0.0;
            dr -= pp;
          }
          if (lane == m) er = 0.0;
          if (lane < n) qt[(i + 1) * n + lane] = carry;
          early = true;
          break;
        }
        const double rinv = rsqrt_newton(x2);
        r = x2 * rinv;
        if (lane == i + 1) er = r;
        sn = f * rinv;
        c = g * rinv;
        g = lane_value(dr, i + 1) - pp;
        r = (lane_value(dr, i) - g) * sn + 2.0 * c * b;
        pp = sn * r;
        if (lane == i + 1) dr = g + pp;
        g = c * r - b;
        if (lane < n) {
          const double q0 = qt[i * n + lane];
          qt[(i + 1) * n + lane] = sn * q0 + c * carry;
          carry = c * q0 - sn * carry;
        }
      }
      if (early) continue;
      if (lane < n) qt[l * n + lane] = carry;
      if (lane == l) {
        dr -= pp;
        er = g;
      }
      if (lane == m) er = 0.0;
    }
  }
  // stable ascending order: lane i's rank among (d, index)
  if (lane < n) {
    int rk = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = lane_value(dr, j);
      rk += (dj < dr || (dj == dr && j < lane)) ? 1 : 0;
    }
    order[rk] = lane;
    qq[lane] = dr;
  }
  __syncthreads();
  if (lane < n) {
    w[lane] = qq[order[lane]];
#pragma unroll 4
    for (int j = 0; j < n; ++j) a[lane * n + j] = qt[order[j] * n + lane];
  }
  __syncthreads();
  return true;
}

struct BCyc {
  double* sc;
  double* ctl;
  int m, nev, maxitr;
  double tol;
  double* status;  // where this sweep's status (header and Ritz values) is gathered
};

__device__ void cycle_end(const BCyc& o, double* s_e, int* s_done, int* s_nconv_p) {
  int& s_nconv = *s_nconv_p;
  double* ctl = o.ctl;
  const int lane = threadIdx.x;
  const int m = o.m;
  const double state = ctl[kCtlState];
  const int col = (int)ctl[kCtlCol];
  __syncthreads();  // every lane has read the header before lane 0 writes it
  if (lane == 0) ctl[C_PENDING] = 0.0;
  if (state != 0.0) return;
  if (lane == 0) ctl[C_NAV] += 1.0;
  if (col + 1 < m) {
    if (lane == 0) ctl[kCtlCol] = col + 1;
    return;
  }
  // cycle end: T from the projections (columns k..m-1) and the kept Ritz values
  const int k = (int)ctl[C_K];
  const int phase = (int)ctl[kCtlPhase];
  double* a = s_e;
  double* qt = a + m * m;
  double* v = qt + m * m;
  double* qq = v + m;
  double* w = qq + m;
  int* order = (int*)(w + m);
  const double* Tcol = o.sc + S_T;
  const double* theta_kept = ctl + ctl_theta(m);
  for (int i = lane; i < m * m; i += 64) {
    const int r = i / m, c = i % m;
    const int hi = r > c ? r : c, lo = r > c ? c : r;
    a[i] = (hi < k) ? (r == c ? theta_kept[r] : 0.0) : Tcol[(size_t)hi * (m + 1) + lo];
  }
  __syncthreads();
  if (!wave_sym_eig(m, a, qt, v, qq, w, order)) {
    if (lane == 0) ctl[kCtlState] = 2.0;
    return;
  }
  const double bres = o.sc[S_NRM];
  if (lane == 0) {
    const double eps23 = 3.666852862501036e-11;  // eps^(2/3)
    int nc = 0;
    for (int i = 0; i < o.nev; ++i) {
      const double bound = fabs(bres * a[(m - 1) * m + i]);
      if (bound <= o.tol * fmax(fabs(w[i]), eps23)) s_done[nc++] = i;
    }
    s_nconv = nc;
  }
  __syncthreads();
  const int nc = s_nconv;
  const int iter = (int)ctl[C_ITER] + 1;
  if (lane == 0) ctl[C_RITZ0] = w[0];
  double* C = ctl + ctl_cmat(m);
  if (nc == o.nev || iter >= o.maxitr) {
    // the converged Ritz pairs are the result (after maxitr: dseupd with info = 1)
    if (lane < nc) ctl[C_HDR + lane] = w[s_done[lane]];
    if (lane < m)
      for (int c = 0; c < nc; ++c) C[c * m + lane] = a[lane * m + s_done[c]];
    __syncthreads();
    if (lane == 0) {
      ctl[C_ITER] = iter;
      ctl[C_NCONV] = nc;
      ctl[C_SRC] = phase;
      ctl[C_ROWS] = m;
      ctl[C_COLS] = nc;
      ctl[C_PENDING] = nc > 0 ? 2.0 : 0.0;
      ctl[kCtlState] = 1.0;
    }
    return;
  }
  // thick restart: new basis column c < kk = sum_i A[i][c] V_i, column kk = V_m
  const int kk = restart_keep(o.nev, m, nc);
  if (lane <= m) {
    for (int c = 0; c < kk; ++c) C[c * (m + 1) + lane] = lane < m ? a[lane * m + c] : 0.0;
    C[kk * (m + 1) + lane] = lane == m ? 1.0 : 0.0;
  }
  if (lane < kk) ctl[ctl_theta(m) + lane] = w[lane];
  for (int i = lane; i < (m + 1) * (m + 1); i += 64) o.sc[S_T + i] = 0.0;
  __syncthreads();
  if (lane == 0) {
    ctl[C_ITER] = iter;
    ctl[C_NCONV] = nc;
    ctl[C_K] = kk;
    ctl[kCtlCol] = kk;
    ctl[C_SRC] = phase;
    ctl[kCtlPhase] = phase ^ 1;
    ctl[C_ROWS] = m + 1;
    ctl[C_COLS] = kk + 1;
    ctl[C_PENDING] = 1.0;
  }
}

__global__ __launch_bounds__(64) void bcycle_kernel(const BCyc* __restrict__ ops) {
  extern __shared__ double s_e[];
  __shared__ int s_done[kMaxBasis];
  __shared__ int s_nconv;
  const BCyc o = ops[blockIdx.x];
  cycle_end(o, s_e, s_done, &s_nconv);
  __syncthreads();
  // the status read of this sweep: header and Ritz values
  for (int i = threadIdx.x; i < C_HDR + o.nev; i += 64) o.status[i] = o.ctl[i];
}

// Grouped combination planned by the cycle ends of this sweep (chunks of
// runs with nothing pending return at once): Y_j[e] = sum_i C[j rows + i]
// X_i[e], X = the basis buffer read, Y = the other buffer (restart) or the
// Ritz-vector array (finished run).
struct BComb {
  const double* X0;
  double* X1;
  double* evecs;
  const double* ctl;
  int64_t n;
  int m, chunk0, nch, chunk;
};

template <int K>  // K >= rows of every combination in the launch
__global__ __launch_bounds__(kOpBlock) void bcombine_kernel(const BComb* __restrict__ ops,
                                                            const int* __restrict__ cjob) {
  __shared__ double s_c[kMaxBasis * kMaxBasis];
  const int c = blockIdx.x;
  const BComb& o = ops[cjob[c]];
  const double pend = o.ctl[C_PENDING];
  if (pend == 0.0) return;
  const int src = (int)o.ctl[C_SRC];
  const double* X = src ? o.X1 : o.X0;
  double* Y = pend == 2.0 ? o.evecs : (src ? const_cast<double*>(o.X0) : o.X1);
  const int rows = (int)o.ctl[C_ROWS], cols = (int)o.ctl[C_COLS];
  const double* C = o.ctl + ctl_cmat(o.m);
  const int lc = c - o.chunk0;
  for (int i = threadIdx.x; i < rows * cols; i += kOpBlock) s_c[i] = C[i];
  __syncthreads();
  const int64_t e0 = (int64_t)lc * o.chunk;
  const int64_t e1 = min(o.n, e0 + o.chunk);
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kOpBlock) {
    double xv[K];
#pragma unroll
    for (int i = 0; i < K; ++i) xv[i] = i < rows ? X[(int64_t)i * o.n + e] : 0.0;
    for (int jj = 0; jj < cols; ++jj) {
      double y = 0.0;
#pragma unroll
      for (int i = 0; i < K; ++i)
        if (i < rows) y = fma(s_c[jj * rows + i], xv[i], y);
      Y[(int64_t)jj * o.n + e] = y;
    }
  }
}

// Grouped read: the scalars of every process that reports this sweep,
// packed into one device buffer (one block per process) for one copy.
struct BGather {
  const double* src;
  int64_t off;
  int count;
};

__global__ void bgather_kernel(const BGather* __restrict__ ops, double* __restrict__ dst) {
  const BGather& o = ops[blockIdx.x];
  for (int i = threadIdx.x; i < o.count; i += blockDim.x) dst[o.off + i] = o.src[i];
}


// Chunk size of an operand of n elements: about n / 256, a power of two in
// [kChunkMin, kChunkMax] (up to kChunkHuge above 4 M elements, so that
// there are at most ~2048 chunks) -- it depends on n only, so a process reduces the
// same way whatever else is in the batch.
int chunk_of(int64_t n) {
  int c = kChunkMin;
  while (c < kChunkMax && (int64_t)c * 256 < n) c *= 2;
  // very large operands: longer chunks, at most ~2048 of them (their
  // partial sums are what the reductions read)
  while (c < kChunkHuge && (int64_t)c * 4096 < n) c *= 2;
  return c;
}
int nchunks(int64_t n) { return (int)((n + chunk_of(n) - 1) / chunk_of(n)); }

size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// Pinned status read-back buffers outlive their batch: a released buffer
// is kept for the next batch of the process, since allocating and freeing
// pinned memory costs about a millisecond per call.
std::mutex g_pinned_mu;
std::vector<std::pair<double*, size_t>> g_pinned_free;  // (buffer, capacity in doubles)

double* pinned_take(size_t need, size_t& cap) {
  {
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    size_t best = g_pinned_free.size();
    for (size_t q = 0; q < g_pinned_free.size(); ++q)
      if (g_pinned_free[q].second >= need &&
          (best == g_pinned_free.size() || g_pinned_free[q].second < g_pinned_free[best].second))
        best = q;
    if (best < g_pinned_free.size()) {
      const auto r = g_pinned_free[best];
      g_pinned_free.erase(g_pinned_free.begin() + (std::ptrdiff_t)best);
      cap = r.second;
      return r.first;
    }
  }
  double* h = nullptr;
  EDH_HIP(hipHostMalloc((void**)&h, need * sizeof(double)));
  cap = need;
  return h;
}

// Only once no copy into the buffer is pending.
void pinned_give(double* h, size_t cap) {
  std::lock_guard<std::mutex> lk(g_pinned_mu);
  g_pinned_free.push_back({h, cap});
}

}  // namespace

Batch::~Batch() {
  if (device >= 0) (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& j : jobs)
    if (j) release_job(*j);
  for (ReadSlab& r : slabs) {
    if (r.host) pinned_give(r.host, r.cap);
    if (r.ev) (void)hipEventDestroy(r.ev);
  }
  for (auto& kv : maps) (void)hipFreeAsync(kv.second[4], stream);
  for (auto& s : secs)
    if (s) s->dev.release();
  issue_frees_until(nullptr);
  if (stream) (void)hipStreamSynchronize(stream);
  drop_graph();
  reap_graphs(true);
  if (graph_ev) (void)hipEventDestroy(graph_ev);
  for (hipEvent_t& e : ahead_ev)
    if (e) (void)hipEventDestroy(e);
  uploads.release();
  if (h_obs) (void)hipHostFree(h_obs);
  for (DevBuf* b : {&d_arena, &d_partial, &d_scal, &d_gather, &d_obs}) b->release();
  for (int i = 0; i < 2; ++i) {
    if (h_stage[i]) (void)hipHostFree(h_stage[i]);
    if (ev_stage[i]) (void)hipEventDestroy(ev_stage[i]);
  }
  if (stream) (void)hipStreamDestroy(stream);
}

double* Batch::dalloc(BJob& j, size_t count) {
  void* p = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(double);
  // stream-ordered: every reader and writer of a job's memory runs on the
  // batch stream, so the memory can return to the pool in stream order
  hipError_t e = hipMallocAsync(&p, bytes, stream);
  if (e != hipSuccess) throw Error(EDH_ERR_NOMEM, std::string("batch: hipMallocAsync failed: ") + hipGetErrorString(e));
  j.allocs.push_back(p);
  j.bytes += bytes;
  bytes_in_use += bytes;
  return (double*)p;
}

void Batch::drop_graph() {
  if (!sweep_graph) return;
  if (graph_launched) {
    retired_graphs.push_back({sweep_graph, graph_ev});
    graph_ev = nullptr;
  } else {
    (void)hipGraphExecDestroy(sweep_graph);
  }
  sweep_graph = nullptr;
  graph_launched = false;
}

void Batch::reap_graphs(bool wait) {
  size_t keep = 0;
  for (size_t q = 0; q < retired_graphs.size(); ++q) {
    auto& r = retired_graphs[q];
    const bool done = wait ? hipEventSynchronize(r.second) == hipSuccess : hipEventQuery(r.second) == hipSuccess;
    if (done) {
      (void)hipGraphExecDestroy(r.first);
      (void)hipEventDestroy(r.second);
    } else {
      retired_graphs[keep++] = r;
    }
  }
  retired_graphs.resize(keep);
}

void Batch::issue_frees_until(hipEvent_t ev) {
  reap_graphs(false);
  while (!pending_free.empty()) {
    if (ev && hipEventQuery(ev) == hipSuccess) return;
    (void)hipFreeAsync(pending_free.back(), stream);
    pending_free.pop_back();
  }
}

void Batch::release_job(BJob& j) {
  for (void* p : j.allocs) defer_free(p);
  j.allocs.clear();
  bytes_in_use -= j.bytes;
  j.bytes = 0;
  j.h_read = nullptr;
}

// An eigensolver run's status as read back after a sweep: header of its
// control block, then its Ritz values.
void Batch::eig_status(BJob& j, const double* h) {
  if (j.state != JOB_ACTIVE) return;
  j.n_av = (int)h[C_NAV];
  j.iter = (int)h[C_ITER];
  j.nconv = (int)h[C_NCONV];
  if (j.iter >= 1) j.ritz0 = h[C_RITZ0];
  const double st = h[kCtlState];
  if (st == 0.0) return;
  if (st == 1.0) {
    const int nk = (int)h[C_NCONV];
    j.evals.assign(h + C_HDR, h + C_HDR + nk);
    j.nev = nk;
    j.state = JOB_DONE;
  } else {
    j.state = JOB_FAILED;  // the projected matrix's QL iteration did not converge
  }
  // the bases are no longer read (the Ritz vectors were combined before this
  // status was; sweeps issued since skip the run before touching them)
  for (double* b : {j.V, j.Valt}) {
    auto it = std::find(j.allocs.begin(), j.allocs.end(), (void*)b);
    if (it != j.allocs.end()) {
      const size_t bb = (size_t)(j.m + 1) * j.n * sizeof(double);
      j.allocs.erase(it);
      j.bytes -= bb;
      bytes_in_use -= bb;
      defer_free(b);
    }
  }
  j.V = j.Valt = nullptr;
}

// Handles the oldest grouped read if its copy has completed (waits for it
// when `wait`); returns whether it was handled.
bool Batch::drain_slab(bool wait) {
  if (slab_fifo.empty()) return false;
  ReadSlab& r = slabs[slab_fifo.front()];
  if (wait) {
    issue_frees_until(r.ev);
    EDH_HIP(hipEventSynchronize(r.ev));
  } else if (hipEventQuery(r.ev) != hipSuccess) {
    return false;
  }
  for (size_t q = 0; q < r.entries.size(); ++q) {
    const int id = r.entries[q].job;
    if (id < 0 || id >= (int)jobs.size() || !jobs[id]) continue;
    BJob& j = *jobs[id];
    if (j.kind == JOB_EIG) {
      eig_status(j, r.host + r.offs[q]);
    } else if (j.state == JOB_ACTIVE && j.awaiting) {
      j.awaiting = false;
      j.h_read = r.host + r.offs[q];
      finish_lanczos(j);
    }
  }
  r.busy = false;
  r.entries.clear();
  r.offs.clear();
  slab_fifo.erase(slab_fifo.begin());
  return true;
}

int Batch::add_sector(const Model& m, int nup, int ndn) {
  auto s = std::make_unique<BSector>();
  s->model.reset(new Model(m));
  build_sector(*s->model, nup, ndn, s->tabs, false);
  upload_sector(s->tabs, m.Ns, s->dev, stream, &uploads);
  s->bytes = s->tabs.table_bytes;
  bytes_in_use += s->bytes;
  secs.push_back(std::move(s));
  return (int)secs.size() - 1;
}

std::vector<int> Batch::add_sectors(const Model& m, const std::vector<std::pair<int, int>>& list) {
  const size_t n = list.size();
  std::vector<std::unique_ptr<BSector>> built(n);
  std::vector<std::exception_ptr> err(n);
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t i; (i = next++) < n;) {
      try {
        auto s = std::make_unique<BSector>();
        s->model.reset(new Model(m));
        build_sector(*s->model, list[i].first, list[i].second, s->tabs, false);
        built[i] = std::move(s);
      } catch (...) {
        err[i] = std::current_exception();
      }
    }
  };
  const size_t nt = std::min<size_t>({n, 16, std::max(1u, std::thread::hardware_concurrency())});
  std::vector<std::thread> th;
  for (size_t t = 1; t < nt; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
  std::vector<int> ids;
  for (auto& s : built) {
    upload_sector(s->tabs, m.Ns, s->dev, stream, &uploads);
    s->bytes = s->tabs.table_bytes;
    bytes_in_use += s->bytes;
    secs.push_back(std::move(s));
    ids.push_back((int)secs.size() - 1);
  }
  return ids;
}

BSector& Batch::sector(int id) {
  EDH_CHECK(id >= 0 && id < (int)secs.size() && secs[id], EDH_ERR_ARG, "batch: unknown sector id");
  return *secs[id];
}

BJob& Batch::job(int id) {
  EDH_CHECK(id >= 0 && id < (int)jobs.size() && jobs[id], EDH_ERR_ARG, "batch: unknown job id");
  return *jobs[id];
}

void Batch::remove_sector(int id) {
  BSector& s = sector(id);
  for (auto& j : jobs)
    EDH_CHECK(!j || j->sec != id || j->state != JOB_ACTIVE, EDH_ERR_STATE, "batch: sector has an active job");
  // no wait: the tables and maps return to the pool after the stream's
  // earlier work (every reader of them) has run
  for (auto it = maps.begin(); it != maps.end();) {
    if (std::get<0>(it->first) == id) {
      defer_free(it->second[4]);
      map_blobs.erase(it->first);  // its copy preceded every reader the host has seen finish
      it = maps.erase(it);
    } else {
      ++it;
    }
  }
  bytes_in_use -= s.bytes;
  for (void* p : s.dev.allocs) defer_free(p);
  s.dev.allocs.clear();
  s.dev.release();
  secs[id].reset();
}

// Dense solution of a sector whose Krylov space is the whole sector (the
// single-sector eigensolver's rule: ncv >= n or n <= 64).
void Batch::finish_dense(BJob& j) {
  BSector& s = sector(j.sec);
  const int64_t n = j.n;
  std::vector<double> w;
  std::vector<double> H;
  dense_sector_matrix(*s.model, s.tabs, H);
  std::vector<double> Hs(H.size());
  for (int64_t a = 0; a < n; ++a)
    for (int64_t b = 0; b < n; ++b) Hs[a * n + b] = 0.5 * (H[a * n + b] + H[b * n + a]);
  sym_eig((int)n, Hs, w);
  std::vector<double> X(n * (size_t)j.nev);
  j.evals.assign(j.nev, 0.0);
  for (int q = 0; q < j.nev; ++q) {
    j.evals[q] = w[q];
    for (int64_t i = 0; i < n; ++i) X[q * n + i] = Hs[i * n + q];
  }
  j.evecs = dalloc(j, X.size());
  EDH_HIP(hipMemcpyAsync(j.evecs, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  EDH_HIP(hipStreamSynchronize(stream));
  j.nconv = j.nev;
  j.state = JOB_DONE;
}

int Batch::submit_eig(int sec, int nev_req, int ncv_req, double tol, int maxitr, uint64_t seed) {
  BSector& s = sector(sec);
  EDH_CHECK(nev_req >= 1 && ncv_req >= 2, EDH_ERR_ARG, "batch: bad nev/ncv");
  auto jp = std::make_unique<BJob>();
  BJob& j = *jp;
  j.kind = JOB_EIG;
  j.sec = sec;
  j.n = s.tabs.N;
  j.tol = tol;
  j.maxitr = maxitr;
  jobs.push_back(std::move(jp));
  const int id = (int)jobs.size() - 1;
  const int64_t n = j.n;
  if (n == 0) {
    j.nev = 0;
    j.state = JOB_DONE;
    return id;
  }
  if (n == 1) {  // Storage.h:49-52 and the one-state eigenpair (SpinResolvedStorage.h:52-56)
    j.nev = 1;
    j.evals.assign(1, host_diag(*s.model, s.tabs, 0, 0));
    j.evecs = dalloc(j, 1);
    const double one = 1.0;
    EDH_HIP(hipMemcpyAsync(j.evecs, &one, sizeof(double), hipMemcpyHostToDevice, stream));
    EDH_HIP(hipStreamSynchronize(stream));
    j.nconv = 1;
    j.state = JOB_DONE;
    return id;
  }
  const int ncv = (int)std::min<int64_t>(ncv_req, n);
  j.nev = std::min(nev_req, ncv - 1);
  if (ncv >= n || n <= 64) {
    finish_dense(j);
    return id;
  }
  j.m = std::min(ncv, kMaxBasis - 1);
  EDH_CHECK(j.nev < j.m - 1, EDH_ERR_LIMIT, "batch: NEV too large for the device Krylov basis (max 46)");
  j.V = dalloc(j, (size_t)(j.m + 1) * n);
  j.Valt = dalloc(j, (size_t)(j.m + 1) * n);
  // scalars: projections, norms, projected matrix, then the control block
  // (all zero: column 0, phase 0, active)
  const size_t nsc = S_T + (size_t)(j.m + 1) * (j.m + 1) + ctl_size(j.m);
  j.sc = dalloc(j, nsc);
  fill_dev(j.sc, nsc);
  j.ctl = j.sc + S_T + (size_t)(j.m + 1) * (j.m + 1);
  j.evecs = dalloc(j, (size_t)j.nev * n);
  // random start vector, uniform in [-1, 1) (the single-sector eigensolver's generator)
  dev_random_uniform(j.V, n, seed, stream);
  j.pending_start_norm = true;  // normalised by the next sweep
  j.state = JOB_ACTIVE;
  return id;
}

int Batch::submit_lanczos(int sec, const double* v0_dev, int nl) {
  BSector& s = sector(sec);
  EDH_CHECK(nl >= 1, EDH_ERR_ARG, "batch: nl must be >= 1");
  auto jp = std::make_unique<BJob>();
  BJob& j = *jp;
  j.kind = JOB_LZ;
  j.sec = sec;
  j.n = s.tabs.N;
  j.nl = nl;
  j.alpha.assign(nl, 0.0);
  j.beta.assign(nl + 1, 0.0);
  jobs.push_back(std::move(jp));
  const int id = (int)jobs.size() - 1;
  if (j.n == 0) {
    j.state = JOB_DONE;
    return id;
  }
  // one allocation per run (one stream-ordered free when it ends): the
  // three vectors, the scalars (alpha history (nl), beta history (nl + 1),
  // stop flags, start vector's squared norm) and the per-task shares of
  // <v, w>, each 256-byte aligned
  auto al = [](size_t c) { return (c + 31) & ~size_t(31); };
  const size_t nv = al((size_t)j.n), nsc = S_T + 2 * (size_t)nl + 3;
  j.ntdot = hv_batch_routed_lanczos(s.dev) ? hv_batch_rows_partials(s.tabs.up.D)
                                                       : hv_batch_tasks(s.tabs.up.D, s.tabs.dn.D);
  double* base = dalloc(j, 3 * nv + al(nsc) + (size_t)j.ntdot);
  for (int q = 0; q < 3; ++q) j.lz[q] = j.lzb[q] = base + q * nv;
  j.sc = base + 3 * nv;
  j.tdot = j.sc + al(nsc);
  fill_dev(j.sc, nsc);
  dev_fill(j.sc + S_LZV, 1, 1.0, stream);  // a given start vector is normalised
  dev_fill(j.sc + S_LZK, 1, 1.0, stream);  // the first step
  j.d_alpha = j.sc + S_T;
  j.d_beta = j.d_alpha + nl;
  j.d_flags = (int*)(j.d_beta + nl + 1);
  dev_fill(j.d_alpha + 2 * (size_t)nl + 2, 1, 1.0, stream);  // squared start norm (set again for a gathered start)
  if (v0_dev)
    EDH_HIP(hipMemcpyAsync(j.lz[0], v0_dev, (size_t)j.n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  j.step = 1;
  j.state = JOB_ACTIVE;
  return id;
}

// Device gather maps of c / c^+ on `orb` out of sector src (the single-sector
// a_adag maps), cached per (sector, orbital, operator).
void Batch::gather_maps(int src, int orb, bool annihilate, std::vector<void*>& out) {
  auto key = std::make_tuple(src, orb, annihilate ? 1 : 0);
  auto it = maps.find(key);
  if (it != maps.end()) {
    out = it->second;
    return;
  }
  BSector& a = sector(src);
  const int Ns = a.model->Ns;
  int nup2, ndn2;
  std::vector<int32_t> mu, md;
  std::vector<int8_t> su, sd;
  adag_gather_maps(Ns, a.tabs.nup, a.tabs.ndn, orb, annihilate, nup2, ndn2, mu, md, su, sd);
  // the four maps in one allocation and one copy from a host blob that the
  // batch keeps (so nothing waits for the copy)
  const size_t sz[4] = {mu.size() * 4, md.size() * 4, su.size(), sd.size()};
  const void* src_p[4] = {mu.data(), md.data(), su.data(), sd.data()};
  size_t off[4], tot = 0;
  for (int q = 0; q < 4; ++q) {
    off[q] = tot;
    tot = align16(tot + std::max<size_t>(sz[q], 8));
  }
  std::vector<char>& blob = map_blobs[key];
  blob.assign(tot, 0);
  for (int q = 0; q < 4; ++q)
    if (sz[q]) std::memcpy(blob.data() + off[q], src_p[q], sz[q]);
  void* base = nullptr;
  EDH_HIP(hipMallocAsync(&base, tot, stream));
  EDH_HIP(hipMemcpyAsync(base, blob.data(), tot, hipMemcpyHostToDevice, stream));
  std::vector<void*> p(5, nullptr);
  for (int q = 0; q < 4; ++q) p[q] = static_cast<char*>(base) + off[q];
  p[4] = base;  // the allocation
  maps[key] = p;
  out = p;
}

// Green's-function start vectors c(+)_orb |psi> (GreensFunction.h:395-416,
// 437-458; one orbital, or the sum over the two orbitals of a non-local
// pair) gathered on the device, and a Lanczos run from
// each.  Nothing waits for the device: one grouped reduction gives every
// start vector's norm, which becomes its run's first scale (the stored
// vector stays unnormalised) or, at a squared norm <= 1e-10, stops the run
// before its first step.  The squared norms reach the host with the runs'
// coefficients (norm2 of a finished run).
void Batch::submit_gf(const std::vector<GfRequest>& req, int nl, std::vector<int>& ids) {
  EDH_CHECK(nl >= 1, EDH_ERR_ARG, "batch: nl must be >= 1");
  ids.assign(req.size(), -1);
  OpGroup g{G_NORM, {}};
  for (size_t r = 0; r < req.size(); ++r) {
    const GfRequest& q = req[r];
    BJob& e = job(q.eig_job);
    EDH_CHECK(e.kind == JOB_EIG && e.state == JOB_DONE && e.evecs, EDH_ERR_STATE,
              "batch: eigensolver job has no eigenvectors");
    EDH_CHECK(q.index >= 0 && q.index < e.nev, EDH_ERR_ARG, "batch: eigenpair index out of range");
    BSector& a = sector(e.sec);
    BSector& b = sector(q.dst_sec);
    const int Ns = a.model->Ns;
    EDH_CHECK(q.orb >= 0 && q.orb < 2 * Ns, EDH_ERR_ARG, "batch: orbital out of range");
    const int spin = q.orb >= Ns ? 1 : 0;
    EDH_CHECK(q.orb2 < 0 || (q.orb2 < 2 * Ns && (q.orb2 >= Ns ? 1 : 0) == spin && q.orb2 != q.orb), EDH_ERR_ARG,
              "batch: second orbital out of range, of the other spin or equal to the first");
    const int d = q.annihilate ? -1 : 1;
    EDH_CHECK(b.tabs.nup == a.tabs.nup + (spin == 0 ? d : 0) && b.tabs.ndn == a.tabs.ndn + (spin == 1 ? d : 0),
              EDH_ERR_ARG, "batch: target sector does not match the operator");
    if (b.tabs.N == 0) continue;
    std::vector<void*> mp;
    gather_maps(e.sec, q.orb, q.annihilate != 0, mp);
    const int id = submit_lanczos(q.dst_sec, nullptr, nl);
    BJob& j = job(id);
    dev_separable_gather(e.evecs + (size_t)q.index * e.n, a.tabs.dn.D, (const int32_t*)mp[0], (const int32_t*)mp[1],
                         (const int8_t*)mp[2], (const int8_t*)mp[3], j.lz[0], b.tabs.up.D, b.tabs.dn.D, stream);
    if (q.orb2 >= 0) {
      // the second orbital's vector added to the first (non-local start
      // vectors, GreensFunction.h:290-327)
      std::vector<void*> mp2;
      gather_maps(e.sec, q.orb2, q.annihilate != 0, mp2);
      dev_separable_gather(e.evecs + (size_t)q.index * e.n, a.tabs.dn.D, (const int32_t*)mp2[0],
                           (const int32_t*)mp2[1], (const int8_t*)mp2[2], (const int8_t*)mp2[3], j.lz[0], b.tabs.up.D,
                           b.tabs.dn.D, stream, true);
    }
    BOpHost o;
    o.w = j.lz[0];
    o.n = j.n;
    o.sc = j.sc;
    o.flags = j.d_flags;
    o.mode = OP_START;
    o.nrm_out = j.d_alpha + 2 * (size_t)nl + 2;
    g.ops.push_back(o);
    ids[r] = id;
  }
  std::vector<OpGroup> pre{g}, post;
  issue(pre, nullptr, post);
}

void Batch::free_job(int id) {
  BJob& j = job(id);
  EDH_CHECK(j.state != JOB_ACTIVE, EDH_ERR_STATE, "batch: job still active");
  release_job(j);
  jobs[id].reset();
}

void Batch::fill_dev(double* p, size_t n) {
  if (!n) return;
  hipLaunchKernelGGL(bfill_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 1024)), dim3(256), 0, stream, p,
                     (int64_t)n, 0.0);
  EDH_HIP(hipGetLastError());
}

// Packs every descriptor of one launch sequence (groups before the H*v, the
// grouped H*v, groups after it) into one pinned staging buffer, copies it to
// the device in one transfer and launches the sequence.
void Batch::issue(std::vector<OpGroup>& pre, const std::vector<std::pair<int, HvArgs>>* hv,
                  std::vector<OpGroup>& post, const std::vector<BJob*>* cycles, const std::vector<ReadReq>* reads) {
  struct Placed {
    int kind, nops, nch;
    size_t ops_off, cj_off;
    int maxcnt;  // largest projection count of the group's operations
    int maxnch;  // most chunks of one operation
  };
  std::vector<Placed> pl_pre, pl_post;
  size_t off = 0;
  // sections start 16-byte aligned; the alignment gaps are zeroed below, so
  // that two sweeps with the same descriptors have the same bytes (the
  // staging buffers are reused and hold older contents)
  std::vector<std::pair<size_t, size_t>> gaps;
  auto pad = [&gaps](size_t end) {
    const size_t a = align16(end);
    if (a > end) gaps.push_back({end, a - end});
    return a;
  };
  // eigensolver cycle ends and the combinations they plan: one cycle
  // descriptor per run, one combination descriptor per run with its chunks
  const int ncyc = cycles ? (int)cycles->size() : 0;
  int comb_nch = 0, max_m = 0;
  for (int q = 0; q < ncyc; ++q) {
    comb_nch += nchunks((*cycles)[q]->n);
    max_m = std::max(max_m, (*cycles)[q]->m);
  }
  const size_t cyc_off = off;
  off = pad(off + (size_t)ncyc * sizeof(BCyc));
  const size_t comb_ops_off = off;
  off = pad(off + (size_t)ncyc * sizeof(BComb));
  const size_t comb_cj_off = off;
  off = pad(off + (size_t)comb_nch * sizeof(int));
  auto place = [&](std::vector<OpGroup>& gs, std::vector<Placed>& out) {
    for (OpGroup& g : gs) {
      if (g.ops.empty()) continue;
      int nch = 0, maxcnt = 0, maxnch = 0;
      for (const BOpHost& h : g.ops) {
        nch += nchunks(h.n);
        maxnch = std::max(maxnch, nchunks(h.n));
        maxcnt = std::max(maxcnt, h.ctl ? h.m : h.cnt);
      }
      Placed p{g.kind, (int)g.ops.size(), nch, 0, 0, maxcnt, maxnch};
      p.ops_off = off;
      off = pad(off + g.ops.size() * sizeof(BOp));
      p.cj_off = off;
      off = pad(off + (size_t)nch * sizeof(int));
      out.push_back(p);
    }
  };
  place(pre, pl_pre);
  const size_t hv_off = off;
  const size_t jb = hv_batch_job_bytes();
  const bool has_hv = hv && !hv->empty();
  std::vector<HvBatchUnit> units;
  std::vector<int> routed;  // operands for the rows kernel (launched after the grouped one)
  if (has_hv) {
    // rows per unit: one while the rows of the sweep are few (parallelism),
    // up to ~4K elements per unit when there are many (less scheduling)
    int64_t total_rows = 0;
    // Lanczos operands (device scales and <v, w> shares, no control
    // block) follow hv_batch_routed_lanczos, the others hv_batch_routed
    auto routed_op = [&](size_t q) {
      const DevSector& ds = sector((*hv)[q].first).dev;
      const HvArgs& a = (*hv)[q].second;
      const bool lanczos = !a.ctl && a.dscale && a.tdot;
      return lanczos ? hv_batch_routed_lanczos(ds) : hv_batch_routed(ds.up.D, ds.dn.D);
    };
    for (size_t q = 0; q < hv->size(); ++q) {
      const DevSector& ds = sector((*hv)[q].first).dev;
      if (!routed_op(q)) total_rows += ds.up.D;
    }
    const int64_t spread = std::max<int64_t>(1, total_rows / 8192);
    for (size_t q = 0; q < hv->size(); ++q) {
      const DevSector& ds = sector((*hv)[q].first).dev;
      if (routed_op(q)) {
        routed.push_back((int)q);
        continue;
      }
      const int rg = hv_batch_row_group(ds.dn.D);
      const int ru = (int)((std::min<int64_t>(hv_batch_unit_rows(ds.dn.D), spread) + rg - 1) / rg * rg);
      for (int64_t r = 0; r < ds.up.D; r += ru)
        units.push_back({(int)q, (int)r, (int)std::min<int64_t>(ru, ds.up.D - r), 0});
    }
    off = pad(off + hv->size() * jb);
  }
  int hv_lds = 0;
  if (has_hv) {
    std::vector<int64_t> spans(units.size());
    for (size_t i = 0; i < units.size(); ++i)
      spans[i] = (int64_t)units[i].nrows * sector((*hv)[units[i].job].first).dev.dn.D;
    hv_lds = hv_batch_lds_elems(spans);
  }
  const size_t units_off = off;
  off = pad(off + units.size() * sizeof(HvBatchUnit));
  place(post, pl_post);
  // grouped read of the processes reporting this sweep
  const int nread = reads ? (int)reads->size() : 0;
  const size_t gat_off = off;
  // (the gather kernel's descriptors: the reads after the first ncyc)
  off = pad(off + (size_t)std::max(nread - ncyc, 0) * sizeof(BGather));
  size_t read_total = 0;
  for (int q = 0; q < nread; ++q) read_total += (*reads)[q].count;
  if (off == 0) return;
  // the first ncyc reads are the eigensolver runs' statuses, written into
  // the gather buffer by the cycle-end kernel; the gather kernel copies the rest
  EDH_CHECK(nread >= ncyc, EDH_ERR_INTERNAL, "batch: status reads missing");
  for (int q = 0; q < ncyc; ++q)
    EDH_CHECK((*reads)[q].src == (*cycles)[q]->ctl, EDH_ERR_INTERNAL, "batch: status read out of order");
  if (nread) d_gather.ensure(read_total * sizeof(double));
  // pinned staging, double-buffered: a buffer is rewritten only after the
  // copy that last read it has completed
  const int si = stage_i;
  stage_i ^= 1;
  if (ev_used[si]) {
    issue_frees_until(ev_stage[si]);
    EDH_HIP(hipEventSynchronize(ev_stage[si]));
  }
  if (off > h_stage_bytes[si]) {
    if (h_stage[si]) EDH_HIP(hipHostFree(h_stage[si]));
    h_stage[si] = nullptr;
    const size_t cap = std::max<size_t>(off * 2, 1 << 20);
    EDH_HIP(hipHostMalloc(&h_stage[si], cap));
    h_stage_bytes[si] = cap;
  }
  if (!ev_stage[si]) EDH_HIP(hipEventCreateWithFlags(&ev_stage[si], hipEventDisableTiming));
  char* hb = (char*)h_stage[si];
  auto fill = [&](std::vector<OpGroup>& gs, std::vector<Placed>& pls) {
    size_t q = 0;
    for (OpGroup& g : gs) {
      if (g.ops.empty()) continue;
      const Placed& p = pls[q++];
      int* cj = (int*)(hb + p.cj_off);
      int c0 = 0;
      for (size_t i = 0; i < g.ops.size(); ++i) {
        const BOpHost& h = g.ops[i];
        BOp o;
        std::memset(&o, 0, sizeof(o));  // padding too: steady sweeps compare the bytes
        o.X = h.X;
        o.ldx = h.ldx;
        o.cnt = h.cnt;
        o.w = h.w;
        o.n = h.n;
        o.chunk0 = c0;
        o.nch = nchunks(h.n);
        o.chunk = chunk_of(h.n);
        o.sc = h.sc;
        o.hsum = h.hsum;
        o.hist_a = h.hist_a;
        o.hist_b = h.hist_b;
        o.flags = h.flags;
        o.kstep = h.kstep;
        o.mode = h.mode;
        o.scale = h.scale;
        o.nrm_out = h.nrm_out;
        o.ctl = h.ctl;
        o.X1 = h.X1;
        o.m = h.m;
        o.tdot = h.tdot;
        o.ntdot = h.ntdot;
        o.lzk = h.lzk;
        std::memcpy(hb + p.ops_off + i * sizeof(BOp), &o, sizeof(BOp));
        for (int c = 0; c < o.nch; ++c) cj[c0 + c] = (int)i;
        c0 += o.nch;
      }
    }
  };
  fill(pre, pl_pre);
  if (has_hv) {
    for (size_t q = 0; q < hv->size(); ++q)
      hv_batch_fill(sector((*hv)[q].first).dev, (*hv)[q].second, 0, hb + hv_off + q * jb);
    if (!units.empty()) std::memcpy(hb + units_off, units.data(), units.size() * sizeof(HvBatchUnit));
  }
  fill(post, pl_post);
  {
    int c0 = 0;
    size_t soff = 0;
    int* cj = (int*)(hb + comb_cj_off);
    for (int q = 0; q < ncyc; ++q) {
      const BJob& j = *(*cycles)[q];
      BCyc y;
      std::memset(&y, 0, sizeof(y));
      y.sc = j.sc;
      y.ctl = j.ctl;
      y.m = j.m;
      y.nev = j.nev;
      y.maxitr = j.maxitr;
      y.tol = j.tol;
      y.status = d_gather.as<double>() + soff;
      soff += (*reads)[q].count;
      std::memcpy(hb + cyc_off + q * sizeof(BCyc), &y, sizeof(BCyc));
      BComb b;
      std::memset(&b, 0, sizeof(b));
      b.X0 = j.V;
      b.X1 = j.Valt;
      b.evecs = j.evecs;
      b.ctl = j.ctl;
      b.n = j.n;
      b.m = j.m;
      b.chunk0 = c0;
      b.nch = nchunks(j.n);
      b.chunk = chunk_of(j.n);
      std::memcpy(hb + comb_ops_off + q * sizeof(BComb), &b, sizeof(BComb));
      for (int c = 0; c < b.nch; ++c) cj[c0 + c] = q;
      c0 += b.nch;
    }
    int64_t roff = 0;
    for (int q = 0; q < nread; ++q) {
      if (q >= ncyc) {
        BGather g;
        std::memset(&g, 0, sizeof(g));
        g.src = (*reads)[q].src;
        g.off = roff;
        g.count = (int)(*reads)[q].count;
        std::memcpy(hb + gat_off + (q - ncyc) * sizeof(BGather), &g, sizeof(BGather));
      }
      roff += (int64_t)(*reads)[q].count;
    }
  }
  for (const auto& g : gaps) std::memset(hb + g.first, 0, g.second);
  int max_ch = 1;
  for (auto* v : {&pl_pre, &pl_post})
    for (const Placed& p : *v) max_ch = std::max(max_ch, p.nch);
  d_arena.ensure(off);
  d_partial.ensure((size_t)max_ch * kMaxBasis * sizeof(double));
  // launch configuration beside the descriptor bytes: together they fix
  // every kernel and argument of the sweep
  std::vector<int64_t> cfg{(int64_t)off, hv_block, hv_lds, (int64_t)units.size(), ncyc, comb_nch, max_m, nread,
                           (int64_t)read_total, (int64_t)(uintptr_t)d_partial.p, (int64_t)(uintptr_t)d_gather.p};
  for (auto* v : {&pl_pre, &pl_post})
    for (const Placed& p : *v)
      for (int64_t x : {(int64_t)p.kind, (int64_t)p.nops, (int64_t)p.nch, (int64_t)p.ops_off, (int64_t)p.cj_off,
                        (int64_t)p.maxcnt})
        cfg.push_back(x);
  // (sweeps with rows-kernel operands launch eagerly: their arguments are
  // not in the descriptor bytes)
  const bool steady = use_graphs && routed.empty() && d_arena.p == last_arena && cfg == last_cfg &&
                      last_desc.size() == off &&
                      std::memcmp(last_desc.data(), hb, off) == 0;
  if (!steady) {
    drop_graph();
    EDH_HIP(hipMemcpyAsync(d_arena.p, hb, off, hipMemcpyHostToDevice, stream));
    EDH_HIP(hipEventRecord(ev_stage[si], stream));
    ev_used[si] = true;
    last_desc.assign(hb, hb + off);
    last_cfg = cfg;
    last_arena = d_arena.p;
  }
  // at most kAhead sweeps in flight: wait for the one issued kAhead sweeps ago
  const int ai = ahead_i;
  ahead_i = (ahead_i + 1) % kAhead;
  if (ahead_used[ai]) {
    issue_frees_until(ahead_ev[ai]);
    EDH_HIP(hipEventSynchronize(ahead_ev[ai]));
  }
  if (!ahead_ev[ai]) EDH_HIP(hipEventCreateWithFlags(&ahead_ev[ai], hipEventDisableTiming));
  char* db = d_arena.as<char>();
  double* part = d_partial.as<double>();
  auto launch = [&](const std::vector<Placed>& pls) {
    for (const Placed& p : pls) {
      const BOp* dops = (const BOp*)(db + p.ops_off);
      const int* dcj = (const int*)(db + p.cj_off);
      // projections (and the norm slot) in one pass over a chunk when they fit twelve registers
      const bool small = p.maxcnt + 1 <= kFuseSmall;
      // one wavefront per coefficient; sixteen of them when the operations
      // have many chunks (the reduction's loads are then its cost)
      const int rblock = p.maxnch >= 512 ? kReduceBlock : 256;
      auto dots = small ? bdots_kernel<kFuseSmall> : bdots_kernel<kHGroup>;
      auto update = small ? bupdate_kernel<kFuseSmall> : bupdate_kernel<kHGroup>;
      if (p.kind == G_PROJECT) {
        hipLaunchKernelGGL(dots, dim3(p.nch), dim3(kOpBlock), 0, stream, dops, dcj, part);
        hipLaunchKernelGGL(breduce_kernel, dim3(p.nops), dim3(rblock), 0, stream, dops, part, 0, 0);
        hipLaunchKernelGGL(update, dim3(p.nch), dim3(kOpBlock), 0, stream, dops, dcj, part, 0);
        hipLaunchKernelGGL(breduce_kernel, dim3(p.nops), dim3(rblock), 0, stream, dops, part, 1, 0);
      } else if (p.kind == G_REORTH) {
        // pass 1 projections (with the norm before) -> fused pass 1 update and
        // pass 2 projections (with the DGKS gate) -> pass 2 update where gated
        hipLaunchKernelGGL(dots, dim3(p.nch), dim3(kOpBlock), 0, stream, dops, dcj, part);
        hipLaunchKernelGGL(breduce_kernel, dim3(p.nops), dim3(rblock), 0, stream, dops, part, 0, 0);
        auto fused = p.maxcnt <= kFuseSmall ? bupdate_dots_kernel<kFuseSmall> : bupdate_dots_kernel<kFuse>;
        hipLaunchKernelGGL(fused, dim3(p.nch), dim3(kOpBlock), 0, stream, dops, dcj, part);
        hipLaunchKernelGGL(breduce_kernel, dim3(p.nops), dim3(rblock), 0, stream, dops, part, 2, 0);
        auto finish = small ? bupdate_scale_kernel<kFuseSmall> : bupdate_scale_kernel<kHGroup>;
        hipLaunchKernelGGL(finish, dim3(p.nch), dim3(kOpBlock), 0, stream, dops, dcj);
      } else if (p.kind == G_LZ) {
        // alpha from the H*v's shares, w -= alpha v_k with |w|^2, beta_k and scales
        hipLaunchKernelGGL(blz_alpha_kernel, dim3(p.nops), dim3(256), 0, stream, dops);
        hipLaunchKernelGGL(update, dim3(p.nch), dim3(kOpBlock), 0, stream, dops, dcj, part, 0);
        hipLaunchKernelGGL(breduce_kernel, dim3(p.nops), dim3(rblock), 0, stream, dops, part, 1, 0);
      } else if (p.kind == G_NORM) {
        hipLaunchKernelGGL(update, dim3(p.nch), dim3(kOpBlock), 0, stream, dops, dcj, part, 0);
        hipLaunchKernelGGL(breduce_kernel, dim3(p.nops), dim3(rblock), 0, stream, dops, part, 1, 0);
      } else {
        hipLaunchKernelGGL(bscale_kernel, dim3(p.nch), dim3(kOpBlock), 0, stream, dops, dcj);
      }
    }
  };
  auto kernels = [&]() {
    launch(pl_pre);
    if (has_hv)
      launch_hv_batch(db + hv_off, (const HvBatchUnit*)(db + units_off), (int)units.size(), hv_block, hv_lds,
                      stream);
    {
      // Lanczos operands of one sector share a multi-operand launch (up to
      // kMultiMax at a time, the same arithmetic per operand), the others
      // take one launch each
      std::vector<std::pair<int, std::vector<HvArgs>>> groups;  // (sector, operands)
      for (int q : routed) {
        const int sid = (*hv)[q].first;
        const HvArgs& a = (*hv)[q].second;
        const DevSector& ds = sector(sid).dev;
        if (a.ctl || !a.dscale || !a.tdot || hv_batch_rows_multi_fit(ds) < 2) {
          launch_hv_batch_rows(ds, a, stream);
          continue;
        }
        auto it = std::find_if(groups.begin(), groups.end(), [&](const auto& g) {
          return g.first == sid && (int)g.second.size() < hv_batch_rows_multi_fit(ds);
        });
        if (it == groups.end()) {
          groups.push_back({sid, {}});
          it = groups.end() - 1;
        }
        it->second.push_back(a);
      }
      for (const auto& g : groups) {
        const DevSector& ds = sector(g.first).dev;
        if (g.second.size() == 1) launch_hv_batch_rows(ds, g.second[0], stream);
        else launch_hv_batch_rows_multi(ds, g.second.data(), (int)g.second.size(), stream);
      }
    }
    launch(pl_post);
    if (ncyc) {
      const size_t smem = (size_t)(2 * max_m * max_m + 3 * max_m) * sizeof(double) + (size_t)max_m * sizeof(int);
      hipLaunchKernelGGL(bcycle_kernel, dim3(ncyc), dim3(64), smem, stream, (const BCyc*)(db + cyc_off));
      // the combination reads at most max_m + 1 basis vectors per element:
      // registers for twelve of them when that suffices (occupancy)
      auto comb = max_m + 1 <= kFuseSmall ? bcombine_kernel<kFuseSmall> : bcombine_kernel<kMaxBasis>;
      hipLaunchKernelGGL(comb, dim3(comb_nch), dim3(kOpBlock), 0, stream, (const BComb*)(db + comb_ops_off),
                         (const int*)(db + comb_cj_off));
    }
    if (nread > ncyc)
      hipLaunchKernelGGL(bgather_kernel, dim3(nread - ncyc), dim3(256), 0, stream, (const BGather*)(db + gat_off),
                         d_gather.as<double>());
  };
  if (!steady) {
    kernels();
  } else {
    if (!sweep_graph) {
      // the second identical sweep in a row: capture its kernels once
      hipGraph_t g = nullptr;
      EDH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      kernels();
      EDH_HIP(hipStreamEndCapture(stream, &g));
      const hipError_t e = hipGraphInstantiate(&sweep_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        // this sweep runs eagerly; capture is tried again at the next
        // steady stretch, up to eight failures per batch (the first is
        // reported), then the batch launches eagerly
        sweep_graph = nullptr;
        (void)hipGetLastError();
        if (graph_failures++ == 0)
          fprintf(stderr, "edhip batch: hipGraphInstantiate failed (%s); the sweep runs eagerly\n",
                  hipGetErrorString(e));
        if (graph_failures >= 8) use_graphs = false;
        last_cfg.clear();  // not steady next sweep: the upload and capture start over
        kernels();
      }
    }
    if (sweep_graph) {
      if (!graph_ev) EDH_HIP(hipEventCreateWithFlags(&graph_ev, hipEventDisableTiming));
      EDH_HIP(hipGraphLaunch(sweep_graph, stream));
      EDH_HIP(hipEventRecord(graph_ev, stream));
      graph_launched = true;
      ++graph_sweeps;
    }
  }
  if (nread) {
    // a free slab large enough, else a new one
    int si2 = -1;
    for (size_t q = 0; q < slabs.size(); ++q)
      if (!slabs[q].busy && slabs[q].cap >= read_total) {
        si2 = (int)q;
        break;
      }
    if (si2 < 0) {
      ReadSlab r;
      r.host = pinned_take(std::max<size_t>(read_total, 4096), r.cap);
      EDH_HIP(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
      slabs.push_back(r);
      si2 = (int)slabs.size() - 1;
    }
    ReadSlab& r = slabs[si2];
    EDH_HIP(hipMemcpyAsync(r.host, d_gather.p, read_total * sizeof(double), hipMemcpyDeviceToHost, stream));
    EDH_HIP(hipEventRecord(r.ev, stream));
    r.busy = true;
    r.entries = *reads;
    r.offs.clear();
    size_t roff = 0;
    for (const ReadReq& q : *reads) {
      r.offs.push_back(roff);
      roff += q.count;
    }
    slab_fifo.push_back(si2);
  }
  EDH_HIP(hipEventRecord(ahead_ev[ai], stream));
  ahead_used[ai] = true;
  EDH_HIP(hipGetLastError());
}

// One sweep: every active process advances one Krylov step.
int Batch::step() {
  // statuses and Lanczos coefficients that have arrived, in sweep order;
  // when no process can run, wait for the oldest read
  while (drain_slab(false)) {
  }
  auto runnable = [&]() {
    for (auto& jp : jobs)
      if (jp && jp->state == JOB_ACTIVE && !jp->awaiting) return true;
    return false;
  };
  while (!runnable() && drain_slab(true)) {
  }
  std::vector<BJob*> eig, lz, starting;
  std::vector<int> eig_id, lz_id;
  for (size_t id = 0; id < jobs.size(); ++id) {
    BJob* jp = jobs[id].get();
    if (!jp || jp->state != JOB_ACTIVE || jp->awaiting) continue;
    if (jp->kind == JOB_EIG) {
      eig.push_back(jp);
      eig_id.push_back((int)id);
      if (jp->pending_start_norm) starting.push_back(jp);
    } else {
      lz.push_back(jp);
      lz_id.push_back((int)id);
    }
  }
  if (eig.empty() && lz.empty()) {
    int active = 0;
    for (auto& jp : jobs)
      if (jp && jp->state == JOB_ACTIVE) ++active;
    return active;
  }
  std::vector<OpGroup> pre, post;
  // (0) newly submitted eigensolver runs: normalise the start vector
  OpGroup g_norm{G_NORM, {}}, g_scale{G_SCALE, {}};
  for (BJob* j : starting) {
    BOpHost o;
    o.w = j->V;
    o.n = j->n;
    o.sc = j->sc;
    g_norm.ops.push_back(o);
    o.mode = OP_SCALE_NRM;
    g_scale.ops.push_back(o);
    j->pending_start_norm = false;
  }
  // (1) Lanczos runs past their first step: rotate (prev, cur, w) <- (cur,
  // w, prev) (Lanczos.h:47-51); v_k = w / beta_{k-1} is not formed, the
  // stored vector is used with its scale
  for (BJob* j : lz)
    if (j->step > 1) {
      double* t = j->lz[1];
      j->lz[1] = j->lz[0];
      j->lz[0] = j->lz[2];
      j->lz[2] = t;
    }
  pre.push_back(g_norm);
  pre.push_back(g_scale);
  // (2) grouped H*v: eigensolver V_{col+1} = H V_col (column and buffer
  // from the run's control block); Lanczos w = H v_k - beta_{k-1} v_{k-1}
  // (Lanczos.h:55)
  std::vector<std::pair<int, HvArgs>> hv;
  for (BJob* j : eig) {
    HvArgs a{};
    a.v = j->V;
    a.w = j->V;
    a.wprev = a.w;
    a.vscale = 1.0;
    a.mode = 0;
    a.ctl = j->ctl;
    a.v_alt = j->Valt;
    hv.push_back({j->sec, a});
  }
  for (BJob* j : lz) {
    HvArgs a{};
    const DevSector& ds = sector(j->sec).dev;
    if (hv_batch_routed_lanczos(ds)) {
      // the rows kernel takes this step's buffers
      a.v = j->lz[0];
      a.w = j->lz[2];
      a.wprev = j->lz[1];
      a.mode = j->step > 1 ? 1 : 0;
    } else {
      // the grouped kernel resolves the rotation from the run's step
      // counter, so the descriptor is the same at every step
      a.v = j->lzb[0];
      a.wprev = j->lzb[1];
      a.w = j->lzb[2];
      a.lzk = j->sc + S_LZK;
    }
    a.vscale = 1.0;
    // {1 / beta_{k-1}, -beta_{k-1} / beta_{k-2}}; at the first step {1 / |v_1|, 0}
    a.dscale = j->sc + S_LZV;
    a.done = j->d_flags;
    a.tdot = j->tdot;  // alpha = <v_k, w> from the same pass
    hv.push_back({j->sec, a});
  }
  // (3) first projection pass: eigensolver against V_0..V_col (into the
  // projected-matrix column, with the norm before for the DGKS test);
  // Lanczos alpha = <v, w>, w -= alpha v, beta = |w| (Lanczos.h:56-65)
  OpGroup p1{G_PROJECT, {}}, p2{G_PROJECT, {}}, rf{G_REORTH, {}}, sc{G_SCALE, {}}, glz{G_LZ, {}};
  for (BJob* j : eig) {
    BOpHost o;
    o.X = j->V;
    o.X1 = j->Valt;
    o.ctl = j->ctl;
    o.m = j->m;
    o.ldx = j->n;
    o.w = j->V;
    o.n = j->n;
    o.sc = j->sc;
    o.hsum = j->sc + S_T;
    o.mode = OP_NB;
    if (j->m <= kFuse) {
      // both passes and the normalisation in one group (fused)
      o.mode = OP_NB | OP_SCALE_NRM;
      rf.ops.push_back(o);
    } else {
      p1.ops.push_back(o);
      o.mode = OP_GATED;  // (4) second pass only where the DGKS test asks for it
      p2.ops.push_back(o);
      BOpHost t = o;
      t.hsum = nullptr;
      t.mode = OP_SCALE_NRM;
      sc.ops.push_back(t);
    }
  }
  for (BJob* j : lz) {
    // the step resolved on the device from the run's counter (the same
    // descriptor at every step; the host's step count equals that counter
    // while the run is live)
    BOpHost o;
    o.X = j->lzb[0];
    o.X1 = j->lzb[1];
    o.w = j->lzb[2];
    o.lzk = j->sc + S_LZK;
    o.m = j->nl;
    o.ldx = j->n;
    o.cnt = 1;
    o.n = j->n;
    o.sc = j->sc;
    o.hist_a = j->d_alpha;
    o.hist_b = j->d_beta;  // Lanczos.h:64: not stored at the last step
    o.flags = j->d_flags;
    o.mode = OP_LZ | OP_LZS;
    o.tdot = j->tdot;
    o.ntdot = j->ntdot;
    glz.ops.push_back(o);
  }
  post.push_back(glz);
  post.push_back(p1);
  post.push_back(rf);
  post.push_back(p2);
  post.push_back(sc);
  // (5) cycle ends on the device, then the grouped read: every eigensolver
  // run's status, and the coefficients of Lanczos runs that end here (they
  // sit out until the read has arrived)
  std::vector<ReadReq> reads;
  for (size_t q = 0; q < eig.size(); ++q) reads.push_back({eig_id[q], eig[q]->ctl, (size_t)C_HDR + eig[q]->nev});
  for (size_t q = 0; q < lz.size(); ++q)
    if (lz[q]->step == lz[q]->nl) reads.push_back({lz_id[q], lz[q]->d_alpha, 2 * (size_t)lz[q]->nl + 3});
  issue(pre, &hv, post, &eig, &reads);
  ++sweeps;
  for (BJob* j : lz) {
    if (j->step == j->nl) j->awaiting = true;
    else ++j->step;
  }
  int active = 0;
  for (auto& jp : jobs)
    if (jp && jp->state == JOB_ACTIVE) ++active;
  return active;
}

void Batch::finish_lanczos(BJob& j) {
  const int nl = j.nl;
  const double* hist = j.h_read;
  int fl[2];
  std::memcpy(fl, hist + 2 * nl + 1, sizeof(fl));
  j.nlanc = fl[1];
  for (int q = 0; q < nl; ++q) j.alpha[q] = q < j.nlanc ? hist[q] : 0.0;
  for (int q = 0; q <= nl; ++q) j.beta[q] = (q >= 1 && q <= j.nlanc && q != nl) ? hist[nl + q] : 0.0;
  j.norm2 = hist[2 * nl + 2];
  j.state = JOB_DONE;
  release_job(j);  // the coefficients are on the host now
  j.lz[0] = j.lz[1] = j.lz[2] = nullptr;
  j.sc = nullptr;
}

// Picks the grouped H*v block size for the sectors of the batch by timing
// each candidate on scratch operands (v = 0, so the timing is data-free).
int Batch::autotune(const std::vector<int>& sec_ids) {
  std::vector<int> ids;
  for (int id : sec_ids)
    if (id >= 0 && id < (int)secs.size() && secs[id] && secs[id]->tabs.N > 0) ids.push_back(id);
  if (ids.empty()) return hv_block;
  int64_t tot = 0;
  for (int id : ids) tot += secs[id]->tabs.N;
  double* scratch = nullptr;
  if (hipMalloc(&scratch, (size_t)2 * tot * sizeof(double)) != hipSuccess) return hv_block;
  hipLaunchKernelGGL(bfill_kernel, dim3(1024), dim3(256), 0, stream, scratch, 2 * tot, 0.0);
  std::vector<std::pair<int, HvArgs>> hv;
  int64_t off = 0;
  for (int id : ids) {
    HvArgs a{};
    a.v = scratch + off;
    a.w = scratch + tot + off;
    a.wprev = a.w;
    a.vscale = 1.0;
    a.mode = 0;
    hv.push_back({id, a});
    off += secs[id]->tabs.N;
  }
  std::vector<OpGroup> none;
  hipEvent_t e0, e1;
  EDH_HIP(hipEventCreate(&e0));
  EDH_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int choice = hv_block;
  for (int b : {128, 256, 512}) {
    hv_block = b;
    issue(none, &hv, none);  // warm-up
    EDH_HIP(hipEventRecord(e0, stream));
    for (int r = 0; r < 3; ++r) issue(none, &hv, none);
    EDH_HIP(hipEventRecord(e1, stream));
    EDH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    EDH_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      choice = b;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  EDH_HIP(hipStreamSynchronize(stream));
  (void)hipFree(scratch);
  hv_block = choice;
  return choice;
}

// Times the grouped H*v of an eigensolver sweep, w = H v over one operand
// per listed sector (seeded uniform v in [-1, 1)), as the steady sweeps run
// it (replayed from a captured graph when graphs are on): mean milliseconds
// per launch over `reps` launches between two HIP events on the batch's
// stream, the launch's algorithmic bytes (16 n + the per-spin table bytes,
// summed over the sectors: read v, write w, read the tables once) and its
// element count.
void Batch::time_hv(const std::vector<int>& sec_ids, int reps, uint64_t seed, double* avg_ms, int64_t* bytes,
                    int64_t* elements) {
  EDH_CHECK(reps >= 1, EDH_ERR_ARG, "batch: reps must be >= 1");
  std::vector<int> ids;
  int64_t tot = 0, bsum = 0;
  for (int id : sec_ids) {
    EDH_CHECK(id >= 0 && id < (int)secs.size() && secs[id], EDH_ERR_ARG, "batch: no such sector");
    const BSector& s = *secs[id];
    if (s.tabs.N == 0) continue;
    ids.push_back(id);
    tot += s.tabs.N;
    bsum += 16 * s.tabs.N + s.tabs.table_bytes;
  }
  *avg_ms = 0.0;
  *bytes = bsum;
  *elements = tot;
  if (ids.empty()) return;
  double* buf = nullptr;
  EDH_HIP(hipMalloc(&buf, (size_t)2 * tot * sizeof(double)));
  std::vector<std::pair<int, HvArgs>> hv;
  int64_t off = 0;
  for (int id : ids) {
    HvArgs a{};
    a.v = buf + off;
    a.w = buf + tot + off;
    a.wprev = a.w;
    a.vscale = 1.0;
    a.mode = 0;
    hv.push_back({id, a});
    off += secs[id]->tabs.N;
  }
  dev_random_uniform(buf, tot, seed, stream);
  std::vector<OpGroup> none;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  auto release = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    const hipError_t es = hipStreamSynchronize(stream);
    (void)hipFree(buf);
    return es;
  };
  hipError_t err = hipSuccess;
  try {
    err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) {
      issue(none, &hv, none);  // upload
      issue(none, &hv, none);  // the same sweep again: captured when graphs are on
      err = hipEventRecord(e0, stream);
      for (int r = 0; r < reps && err == hipSuccess; ++r) {
        issue(none, &hv, none);
        err = hipGetLastError();
      }
      if (err == hipSuccess) err = hipEventRecord(e1, stream);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    }
  } catch (...) {
    (void)release();
    throw;
  }
  const hipError_t es = release();
  EDH_HIP(err);
  EDH_HIP(es);
  *avg_ms = ms / reps;
}

}  // namespace edh
