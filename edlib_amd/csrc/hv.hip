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
// One work-group processes R up-rows per pass; each row is staged in LDS
// (scaled by vscale) so the down-spin gathers hit LDS.  The down-spin table is
// a sliced ELL with slice height 64, so a wavefront's entry loads coalesce.
#include "edh_internal.h"

namespace edh {

struct HvParams {
  int Ns, nb;  // sites per spin, bath bits (Ns - ml) below the impurity bits
  int Dup, Ddn;
  int R;    // rows per pass
  int tpr;  // threads per row
  int ngroups;
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

template <int CROSS, bool KAN, bool LDS>
__global__ void hv_kernel(HvParams p) {
  extern __shared__ double s_v[];
  __shared__ double s_red[16];
  const int tid = threadIdx.x;
  const int g = tid / p.tpr;
  const int lt = tid - g * p.tpr;
  const int Ddn = p.Ddn;
  const int na = 1 << (p.Ns - p.nb);
  const unsigned lowmask = (1u << p.nb) - 1u;
  if (p.done && *p.done) return;
  const double vscale = p.dscale ? p.dscale[0] : p.vscale;
  const double wscale = p.dscale ? p.dscale[1] : p.wscale;
  double dsum = 0.0;

  for (int grp = blockIdx.x; grp < p.ngroups; grp += gridDim.x) {
    const int slot = grp * p.R + g;
    const bool active = (g < p.R) && (slot < p.Dup);
    // Row index is uniform per row group; readfirstlane lets the compiler use scalar loads.
    int u = active ? p.row_order[slot] : 0;
    u = __builtin_amdgcn_readfirstlane(u);
    const size_t rowoff = (size_t)u * Ddn;
    double* sv = s_v + (size_t)g * Ddn;
    if (LDS) {
      if (active)
        for (int d = lt; d < Ddn; d += p.tpr) sv[d] = p.v[rowoff + d] * vscale;
      __syncthreads();
    }
    if (active) {
      const int bu = p.up_bits[u];
      const double eu = p.up_e[u];
      const int hb = p.up_rowp[u], he = p.up_rowp[u + 1];
      for (int d = lt; d < Ddn; d += p.tpr) {
        const int bd = p.dn_bits[d];
        const double vd = LDS ? sv[d] : p.v[rowoff + d] * vscale;
        const double dg = eu + p.dn_e[d] + cross_term<CROSS>(p, u, bu, bd, na);
        double acc = dg * vd + (p.accumulate ? wscale * p.wprev[rowoff + d] : 0.0);
        // down-spin hopping (sliced ELL, gathers from the staged row)
        const int sl = d >> 6, lane = d & 63;
        const int len = p.dn_sllen[sl];
        const uint32_t* ep = p.dn_ell + p.dn_slptr[sl] + lane;
        for (int k = 0; k < len; ++k) {
          const uint32_t e = ep[k * kSlice];
          const int c = (int)(e & kColMask);
          const double x = LDS ? sv[c] : p.v[rowoff + c] * vscale;
          acc = fma(p.values[e >> kColBits], x, acc);
        }
        // up-spin hopping: linear combination of whole rows, coalesced over d
        for (int q = hb; q < he; ++q) {
          const uint32_t e = p.up_packed[q];
          const size_t src = (size_t)(e & kColMask) * Ddn + d;
          acc = fma(p.values[e >> kColBits], p.v[src] * vscale, acc);
        }
        if (KAN) {
          const int a = bu >> p.nb, b = bd >> p.nb;
          const int kb = p.kan_ptr[a * na + b], ke = p.kan_ptr[a * na + b + 1];
          for (int q = kb; q < ke; ++q) {
            const int up2 = p.up_rank[(bu & lowmask) | (p.kan_a[q] << p.nb)];
            const int dn2 = p.dn_rank[(bd & lowmask) | (p.kan_b[q] << p.nb)];
            acc = fma(p.kan_c[q], p.v[(size_t)up2 * Ddn + dn2] * vscale, acc);
          }
        }
        p.w[rowoff + d] = acc;
        dsum = fma(vd, acc, dsum);
      }
    }
    if (LDS) __syncthreads();
  }
  if (p.dot_partial) {
    // block reduction: wavefront shuffle, then across waves through LDS
    for (int off = 32; off > 0; off >>= 1) dsum += __shfl_down(dsum, off, 64);
    const int wid = tid >> 6;
    if ((tid & 63) == 0) s_red[wid] = dsum;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += s_red[q];
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
  int R, tpr, block, grid, ngroups;
  size_t smem;
  bool lds;
};

Launch plan(const DevSector& ds) {
  Launch L;
  const size_t rowb = (size_t)ds.dn.D * sizeof(double);
  L.lds = rowb <= lds_budget();
  int R = 1;
  if (L.lds) {
    while (R < 8 && (size_t)(R * 2) * rowb <= 64 * 1024 && ds.dn.D * R * 2 <= 8 * 512) R *= 2;
  }
  L.R = R;
  L.block = (R == 1 && ds.dn.D >= 4096) ? 1024 : 512;
  L.tpr = L.block / R;
  L.ngroups = (int)((ds.up.D + R - 1) / R);
  L.grid = L.ngroups;
  L.smem = L.lds ? (size_t)R * rowb : 0;
  return L;
}

template <int CROSS, bool KAN, bool LDS>
void launch_t(const HvParams& p, const Launch& L, hipStream_t s) {
  auto k = hv_kernel<CROSS, KAN, LDS>;
  if (L.smem > 64 * 1024) EDH_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.smem));
  hipLaunchKernelGGL(k, dim3(L.grid), dim3(L.block), L.smem, s, p);
  EDH_HIP(hipGetLastError());
}

template <int CROSS>
void launch_c(const HvParams& p, const Launch& L, bool kan, hipStream_t s) {
  if (kan) {
    if (L.lds) launch_t<CROSS, true, true>(p, L, s);
    else launch_t<CROSS, true, false>(p, L, s);
  } else {
    if (L.lds) launch_t<CROSS, false, true>(p, L, s);
    else launch_t<CROSS, false, false>(p, L, s);
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
  p.R = L.R;
  p.tpr = L.tpr;
  p.ngroups = L.ngroups;
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
  switch (ds.cross_mode) {
    case CROSS_NONE: launch_c<CROSS_NONE>(p, L, kan, s); break;
    case CROSS_POPC: launch_c<CROSS_POPC>(p, L, kan, s); break;
    case CROSS_IMP: launch_c<CROSS_IMP>(p, L, kan, s); break;
    default: launch_c<CROSS_GENERAL>(p, L, kan, s); break;
  }
}

}  // namespace edh
