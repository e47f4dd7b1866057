// Matrix-free spin-resolved H*v on CDNA4 (gfx950), FP64.
//
// Replaces SpinResolvedStorage::av (SpinResolvedStorage.h:58-115).  The state
// vector is the D_up x D_dn row-major matrix V[u][d] (down index contiguous).
//   w[u][d] = (diag(u,d) v[u][d] + w_in)            diagonal, on the fly
//           + sum_j Hdn[d][j] v[u][j]               down hopping (row in LDS)
//           + sum_k Hup[u][k] v[k][d]               up hopping (coalesced rows)
//           + sum_V c * v[u'][d']                   impurity interaction, on the fly
// diag(u,d) = Eup[u] + Edn[d] + X(bits_u, bits_d), X from the cross table mode.
//
// One work-group processes one up-row at a time, in the processing order
// row_order (reverse Cuthill-McKee of the up-spin hopping graph, so that the
// neighbour rows read by the up-spin term are reused from the Infinity Cache).
// The row is staged in LDS (scaled by vscale) so the down-spin gathers hit
// LDS; the down-spin table is a sliced ELL with slice height 64, so a
// wavefront's entry loads coalesce.
#include "edh_internal.h"

namespace edh {

struct HvParams {
  int Ns, nb;  // sites per spin, bath bits (Ns - ml) below the impurity bits
  int Dup, Ddn;
  const int32_t* row_order;
  // up spin
  const int32_t* up_bits;
  const double* up_e;
  const int32_t* up_rowp;
  const uint32_t* up_packed;
  const int32_t* up_rank;
  // down spin
  const int32_t* dn_bits;
  const double* dn_e;
  const int32_t* dn_slptr;
  const int32_t* dn_sllen;
  const uint32_t* dn_ell;
  const int32_t* dn_rank;
  const double* values;
  // cross term
  double cross_u;
  const double* cross_imp;
  const double* cross_rowtab;
  // impurity interaction
  int kan_ml;
  const int32_t* kan_ptr;
  const int32_t* kan_a;
  const int32_t* kan_b;
  const double* kan_c;
  // operands
  const double* v;
  double* w;
  const double* wprev;
  double vscale, wscale;
  int accumulate;
  double* dot_partial;
  const double* dscale;
  const int* done;
};

template <int CROSS>
__device__ __forceinline__ double cross_term(const HvParams& p, int u, int bu, int bd, int na) {
  if (CROSS == CROSS_POPC) return p.cross_u * (double)__builtin_popcount(bu & bd);
  if (CROSS == CROSS_IMP) return p.cross_imp[(bu >> p.nb) * na + (bd >> p.nb)];
  if (CROSS == CROSS_GENERAL) {
    double x = 0.0;
    const double* r = p.cross_rowtab + (size_t)u * p.Ns;
    unsigned m = (unsigned)bd;
    while (m) {
      int b = __builtin_ctz(m);
      x += r[b];
      m &= m - 1;
    }
    return x;
  }
  return 0.0;
}

// One block processes one up-row at a time (grid-stride over the processing
// order).  Each thread owns J columns d = c0 + tid + j*T of the row and keeps
// their accumulators in registers, so every neighbour row of the up-spin
// term issues J independent coalesced loads per thread.
template <int J, int CROSS, bool KAN, bool LDS>
__global__ void hv_rows_kernel(HvParams p) {
  extern __shared__ double s_v[];
  __shared__ double s_red[16];
  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int Ddn = p.Ddn;
  const int na = 1 << (p.Ns - p.nb);
  const unsigned lowmask = (1u << p.nb) - 1u;
  if (p.done && *p.done) return;
  const double vscale = p.dscale ? p.dscale[0] : p.vscale;
  const double wscale = p.dscale ? p.dscale[1] : p.wscale;
  const double* __restrict__ V = p.v;
  const double* __restrict__ vals = p.values;
  double dsum = 0.0;

  for (int slot = blockIdx.x; slot < p.Dup; slot += gridDim.x) {
    const int u = __builtin_amdgcn_readfirstlane(p.row_order[slot]);
    const size_t rowoff = (size_t)u * Ddn;
    if (LDS) {
      for (int d = tid; d < Ddn; d += T) s_v[d] = V[rowoff + d] * vscale;
      __syncthreads();
    }
    const int bu = p.up_bits[u];
    const double eu = p.up_e[u];
    const int hb = p.up_rowp[u], he = p.up_rowp[u + 1];
    for (int c0 = 0; c0 < Ddn; c0 += T * J) {
      double acc[J], vv[J];
      int bd[J];
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int d = c0 + tid + j * T;
        acc[j] = 0.0;
        vv[j] = 0.0;
        bd[j] = 0;
        if (d < Ddn) {
          vv[j] = LDS ? s_v[d] : V[rowoff + d] * vscale;
          if (p.accumulate) acc[j] = wscale * p.wprev[rowoff + d];
          bd[j] = p.dn_bits[d];
        }
      }
      // diagonal, generated from the bit patterns
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int d = c0 + tid + j * T;
        if (d < Ddn) acc[j] = fma(eu + p.dn_e[d] + cross_term<CROSS>(p, u, bu, bd[j], na), vv[j], acc[j]);
      }
      // down-spin hopping: sliced ELL, gathers from the staged row
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int d = c0 + tid + j * T;
        if (d < Ddn) {
          const int sl = d >> 6;
          const int len = p.dn_sllen[sl];
          const uint32_t* __restrict__ ep = p.dn_ell + p.dn_slptr[sl] + (d & 63);
          double a0 = 0.0, a1 = 0.0;
          int k = 0;
          for (; k + 1 < len; k += 2) {
            const uint32_t e0 = ep[k * kSlice], e1 = ep[(k + 1) * kSlice];
            const int c0_ = (int)(e0 & kColMask), c1_ = (int)(e1 & kColMask);
            const double x0 = LDS ? s_v[c0_] : V[rowoff + c0_] * vscale;
            const double x1 = LDS ? s_v[c1_] : V[rowoff + c1_] * vscale;
            a0 = fma(vals[e0 >> kColBits], x0, a0);
            a1 = fma(vals[e1 >> kColBits], x1, a1);
          }
          if (k < len) {
            const uint32_t e0 = ep[k * kSlice];
            const int c0_ = (int)(e0 & kColMask);
            a0 = fma(vals[e0 >> kColBits], LDS ? s_v[c0_] : V[rowoff + c0_] * vscale, a0);
          }
          acc[j] += a0 + a1;
        }
      }
      // up-spin hopping: linear combination of whole rows (coalesced)
      for (int q = hb; q < he; ++q) {
        const uint32_t e = p.up_packed[q];
        const double h = vals[e >> kColBits] * vscale;
        const double* __restrict__ src = V + (size_t)(e & kColMask) * Ddn + c0 + tid;
        double x[J];
#pragma unroll
        for (int j = 0; j < J; ++j) x[j] = (c0 + tid + j * T < Ddn) ? src[j * T] : 0.0;
#pragma unroll
        for (int j = 0; j < J; ++j) acc[j] = fma(h, x[j], acc[j]);
      }
      if (KAN) {
        const int a = bu >> p.nb;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int d = c0 + tid + j * T;
          if (d < Ddn) {
            const int b = bd[j] >> p.nb;
            const int kb = p.kan_ptr[a * na + b], ke = p.kan_ptr[a * na + b + 1];
            for (int q = kb; q < ke; ++q) {
              const int up2 = p.up_rank[(bu & lowmask) | (p.kan_a[q] << p.nb)];
              const int dn2 = p.dn_rank[(bd[j] & lowmask) | (p.kan_b[q] << p.nb)];
              acc[j] = fma(p.kan_c[q], V[(size_t)up2 * Ddn + dn2] * vscale, acc[j]);
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int d = c0 + tid + j * T;
        if (d < Ddn) {
          p.w[rowoff + d] = acc[j];
          dsum = fma(vv[j], acc[j], dsum);
        }
      }
    }
    if (LDS) __syncthreads();
  }
  if (p.dot_partial) {
    for (int off = 32; off > 0; off >>= 1) dsum += __shfl_down(dsum, off, 64);
    const int wid = tid >> 6;
    if ((tid & 63) == 0) s_red[wid] = dsum;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int q = 0; q < (T >> 6); ++q) t += s_red[q];
      p.dot_partial[blockIdx.x] = t;
    }
  }
}

namespace {

// Largest dynamic LDS a block may use on the current device (queried once per device).
size_t lds_budget() {
  static int cached[64] = {0};
  int dev = 0;
  EDH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return 64 * 1024;
  if (!cached[dev]) {
    int v = 0;
    EDH_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    cached[dev] = v > 0 ? v : 64 * 1024;
  }
  // leave room for the block's static reduction scratch
  return (size_t)cached[dev] - 1024;
}

struct Launch {
  int J, block, grid;
  size_t smem;
  bool lds;
};

// Columns per thread J and threads per block T: T = 64 * ceil(Ddn / (64 J))
// with J the smallest of {1,2,4,8,16} that keeps T <= 1024 and near 256
// threads for short rows.
Launch plan(const DevSector& ds) {
  Launch L;
  const int64_t Ddn = ds.dn.D;
  const size_t rowb = (size_t)Ddn * sizeof(double);
  L.lds = rowb <= lds_budget();
  int J = 1;
  while (J < 16 && (Ddn + J - 1) / J > 256) J *= 2;
  int64_t T = ((Ddn + J - 1) / J + 63) / 64 * 64;
  if (T > 1024) T = 1024;
  if (T < 64) T = 64;
  L.J = J;
  L.block = (int)T;
  L.grid = (int)std::min<int64_t>(ds.up.D, 1 << 20);
  L.smem = L.lds ? rowb : 0;
  return L;
}

template <int J, int CROSS, bool KAN, bool LDS>
void launch_t(const HvParams& p, const Launch& L, hipStream_t s) {
  auto k = hv_rows_kernel<J, CROSS, KAN, LDS>;
  if (L.smem > 64 * 1024)
    EDH_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.smem));
  hipLaunchKernelGGL(k, dim3(L.grid), dim3(L.block), L.smem, s, p);
  EDH_HIP(hipGetLastError());
}

template <int J, int CROSS>
void launch_c(const HvParams& p, const Launch& L, bool kan, hipStream_t s) {
  if (kan) {
    if (L.lds) launch_t<J, CROSS, true, true>(p, L, s);
    else launch_t<J, CROSS, true, false>(p, L, s);
  } else {
    if (L.lds) launch_t<J, CROSS, false, true>(p, L, s);
    else launch_t<J, CROSS, false, false>(p, L, s);
  }
}

template <int J>
void launch_j(const HvParams& p, const Launch& L, int cross, bool kan, hipStream_t s) {
  switch (cross) {
    case CROSS_NONE: launch_c<J, CROSS_NONE>(p, L, kan, s); break;
    case CROSS_POPC: launch_c<J, CROSS_POPC>(p, L, kan, s); break;
    case CROSS_IMP: launch_c<J, CROSS_IMP>(p, L, kan, s); break;
    default: launch_c<J, CROSS_GENERAL>(p, L, kan, s); break;
  }
}

}  // namespace

int hv_num_blocks(const DevSector& ds) { return plan(ds).grid; }

void launch_hv(const DevSector& ds, const HvArgs& a, hipStream_t s, int variant) {
  (void)variant;
  if (ds.N == 0) return;
  Launch L = plan(ds);
  HvParams p;
  p.Ns = ds.Ns;
  p.nb = ds.Ns - ds.imp_ml;
  p.Dup = (int)ds.up.D;
  p.Ddn = (int)ds.dn.D;
  p.row_order = ds.row_order;
  p.up_bits = ds.up.bits;
  p.up_e = ds.up.ediag;
  p.up_rowp = ds.up.rowp;
  p.up_packed = ds.up.packed;
  p.up_rank = ds.up.rank;
  p.dn_bits = ds.dn.bits;
  p.dn_e = ds.dn.ediag;
  p.dn_slptr = ds.dn.sl_ptr;
  p.dn_sllen = ds.dn.sl_len;
  p.dn_ell = ds.dn.ell;
  p.dn_rank = ds.dn.rank;
  p.values = ds.values;
  p.cross_u = ds.cross_u;
  p.cross_imp = ds.cross_imp;
  p.cross_rowtab = ds.cross_rowtab;
  p.kan_ml = ds.kan_ml;
  p.kan_ptr = ds.kan_ptr;
  p.kan_a = ds.kan_a;
  p.kan_b = ds.kan_b;
  p.kan_c = ds.kan_c;
  p.v = a.v;
  p.w = a.w;
  p.wprev = a.wprev;
  p.vscale = a.vscale;
  p.wscale = a.wscale;
  p.accumulate = a.mode;
  p.dot_partial = a.dot_partial;
  p.dscale = a.dscale;
  p.done = a.done;
  const bool kan = ds.kan_ml > 0;
  switch (L.J) {
    case 1: launch_j<1>(p, L, ds.cross_mode, kan, s); break;
    case 2: launch_j<2>(p, L, ds.cross_mode, kan, s); break;
    case 4: launch_j<4>(p, L, ds.cross_mode, kan, s); break;
    case 8: launch_j<8>(p, L, ds.cross_mode, kan, s); break;
    default: launch_j<16>(p, L, ds.cross_mode, kan, s); break;
  }
}

}  // namespace edh
