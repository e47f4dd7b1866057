// Matrix-free spin-resolved H*v on CDNA4 (gfx950), FP64.
//
// Replaces SpinResolvedStorage::av (SpinResolvedStorage.h:58-115).  The state
// vector is the D_up x D_dn row-major matrix V[u][d] (down index contiguous).
//   w[u][d] = diag(u,d) v[u][d] + w_in[u][d]          diagonal, on the fly
//           + sum_j Hdn[d][j] v[u][j]                  down hopping
//           + sum_k Hup[u][k] v[k][d]                  up hopping
//           + sum_V c v[u'][d']                        impurity interaction, on the fly
// with diag(u,d) = Eup[u] + Edn[d] + X(bits_u, bits_d).
//
// Rows kernel: one work-group per up-row at a time (rows taken in a reverse
// Cuthill-McKee order of the up-spin hopping graph).  The row is staged in
// LDS, so the down-spin gathers hit LDS; each thread owns J columns and keeps
// their accumulators in registers, so the up-spin term issues J independent
// coalesced row loads per neighbour.  The down-spin table is a sliced ELL
// (slice height 64 = one wavefront); when the spin basis and its value set
// are small, 16-bit entries packed by slot quads are software pipelined.
//
// Up-spin slab kernel (two-pass variant): the rows kernel runs without the
// up-spin term, then each XCD sweeps all rows over its own column slab of 16
// columns, so the slab (D_up x 128 B) stays in that XCD's L2 while every
// neighbour row segment is re-read from it.
#include <algorithm>
#include <cstdlib>

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
  const int32_t* dn_q4ptr;
  const uint64_t* dn_q4;
  double dn_vt[4];
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
  // slab kernel geometry
  int nslabs, rows_per_wg, nrb;
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

__device__ __forceinline__ void block_reduce_store(double x, double* s_red, double* out) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_red[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += s_red[q];
    *out = t;
  }
}

// Decodes a compact entry's value: EF == 2 uses one class bit (two values),
// EF == 1 two class bits (four values).
template <int EF>
__device__ __forceinline__ double q4_value(const HvParams& p, uint32_t e) {
  if (EF == 2) return (e & 0x4000u) ? p.dn_vt[1] : p.dn_vt[0];
  const double lo = (e & 0x4000u) ? p.dn_vt[1] : p.dn_vt[0];
  const double hi = (e & 0x4000u) ? p.dn_vt[3] : p.dn_vt[2];
  return (e & 0x8000u) ? hi : lo;
}

// EF: down-spin entry format, 0 = 32-bit sliced ELL, 1/2 = 16-bit entries in
// slot quads (four / two distinct values).
template <int J, int CROSS, bool KAN, bool LDS, int EF, bool UP>
__global__ __launch_bounds__(1024) void hv_rows_kernel(HvParams p) {
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
      if (EF && tid == 0) s_v[Ddn] = 0.0;  // target of the padding entries
      __syncthreads();
    }
    const int bu = p.up_bits[u];
    const double eu = p.up_e[u];
    const int hb = p.up_rowp[u], he = p.up_rowp[u + 1];
    for (int c0 = 0; c0 < Ddn; c0 += T * J) {
      // Columns past the row end are clamped to the last column: every load
      // below is unconditional (no per-column branch, so all loads of a step
      // are in flight together) and such columns are never stored.
      int dc[J];
#pragma unroll
      for (int j = 0; j < J; ++j) dc[j] = min(c0 + tid + j * T, Ddn - 1);
      double acc[J];
      int bd[J];
#pragma unroll
      for (int j = 0; j < J; ++j) {
        bd[j] = p.dn_bits[dc[j]];
        acc[j] = p.accumulate ? wscale * __builtin_nontemporal_load(p.wprev + rowoff + dc[j]) : 0.0;
      }
      // diagonal, generated from the bit patterns
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const double vd = LDS ? s_v[dc[j]] : V[rowoff + dc[j]] * vscale;
        acc[j] = fma(eu + p.dn_e[dc[j]] + cross_term<CROSS>(p, u, bu, bd[j], na), vd, acc[j]);
      }
      // down-spin hopping, gathers from the staged row
      if (EF) {
        // Slot quads (four 16-bit entries in one 64-bit word), software
        // pipelined: quad q+1 of every column is in flight while quad q is
        // consumed.  Loads are clamped to valid addresses and replaced by the
        // padding quad past a column's end, so the loop has no branches.
        int nq[J], qb[J];
        int maxq = 0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int sl = dc[j] >> 6;
          nq[j] = (p.dn_sllen[sl] + 3) >> 2;
          // an empty slice has no quads: point at quad 0 (the array is not
          // empty on this path) so the clamped loads below stay in bounds
          qb[j] = nq[j] > 0 ? p.dn_q4ptr[sl] + (dc[j] & 63) : 0;
          maxq = max(maxq, nq[j]);
        }
        const uint64_t pad1 = (uint64_t)(uint32_t)Ddn;
        const uint64_t padq = pad1 | (pad1 << 16) | (pad1 << 32) | (pad1 << 48);
        uint64_t cur[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const uint64_t t = p.dn_q4[qb[j]];
          cur[j] = nq[j] > 0 ? t : padq;
        }
        for (int q = 0; q < maxq; ++q) {
          uint64_t nx[J];
#pragma unroll
          for (int j = 0; j < J; ++j) {
            const int qq = min(q + 1, max(nq[j] - 1, 0));
            const uint64_t t = p.dn_q4[qb[j] + qq * kSlice];
            nx[j] = (q + 1 < nq[j]) ? t : padq;
          }
#pragma unroll
          for (int j = 0; j < J; ++j) {
#pragma unroll
            for (int sub = 0; sub < 4; ++sub) {
              const uint32_t e = (uint32_t)(cur[j] >> (16 * sub)) & 0xFFFFu;
              acc[j] = fma(q4_value<EF>(p, e), s_v[e & 0x3FFFu], acc[j]);
            }
          }
#pragma unroll
          for (int j = 0; j < J; ++j) cur[j] = nx[j];
        }
      } else {
        // 32-bit sliced ELL: slot k of all J columns loaded together
        int len[J], base[J];
        int maxlen = 0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int sl = dc[j] >> 6;
          len[j] = p.dn_sllen[sl];
          base[j] = len[j] > 0 ? p.dn_slptr[sl] + (dc[j] & 63) : 0;
          maxlen = max(maxlen, len[j]);
        }
        for (int k = 0; k < maxlen; ++k) {
          // k < maxlen means some column has entries, so the array is not
          // empty; indices are clamped into each column's own slots
          uint32_t e[J];
#pragma unroll
          for (int j = 0; j < J; ++j) {
            const uint32_t t = p.dn_ell[base[j] + min(k, max(len[j] - 1, 0)) * kSlice];
            e[j] = k < len[j] ? t : 0u;
          }
#pragma unroll
          for (int j = 0; j < J; ++j) {
            const int c = (int)(e[j] & kColMask);
            acc[j] = fma(vals[e[j] >> kColBits], LDS ? s_v[c] : V[rowoff + c] * vscale, acc[j]);
          }
        }
      }
      // up-spin hopping: linear combination of whole rows (coalesced)
      if (UP) {
        for (int q = hb; q < he; ++q) {
          const uint32_t e = p.up_packed[q];
          const double h = vals[e >> kColBits] * vscale;
          const double* __restrict__ src = V + (size_t)(e & kColMask) * Ddn;
          double x[J];
#pragma unroll
          for (int j = 0; j < J; ++j) x[j] = src[dc[j]];
#pragma unroll
          for (int j = 0; j < J; ++j) acc[j] = fma(h, x[j], acc[j]);
        }
      }
      if (KAN) {
        const int a = bu >> p.nb;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          {
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
          __builtin_nontemporal_store(acc[j], p.w + rowoff + d);
          if (UP && p.dot_partial) dsum = fma(LDS ? s_v[d] : V[rowoff + d] * vscale, acc[j], dsum);
        }
      }
    }
    if (LDS) __syncthreads();
  }
  if (UP && p.dot_partial) block_reduce_store(dsum, s_red, p.dot_partial + blockIdx.x);
}

// Up-spin term over column slabs of 16 (one 128-byte line per row segment):
// w[u][c] += sum_k Hup[u][k] v[k][c].  Persistent work-groups read the XCD
// they run on (HW_REG_XCC_ID) and claim (slab, row block) items from that
// XCD's queue: XCD x owns slabs x, x+8, x+16, ... and its work-groups sweep
// all row blocks of one slab before the next, so the slab (D_up x 128 B)
// stays in that XCD's L2 while the ~17 neighbour segments of each row are
// re-read from it.  A work-group that finds its queue empty takes items from
// the other XCDs' queues, so every item is processed whatever XCD ids are
// reported.  A wavefront covers four rows of the slab.
constexpr int kSlabW = 16;
constexpr int kSlabWG = 256;
constexpr int kXcd = 8;

__global__ void zero_counters_kernel(int* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kSlabWG) void hv_upslab_kernel(HvParams p, int* counters) {
  __shared__ double s_red[16];
  __shared__ int s_item;
  if (p.done && *p.done) return;
  const double vscale = p.dscale ? p.dscale[0] : p.vscale;
  unsigned xcc;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
  const int home = (int)(xcc & (kXcd - 1));
  const int Ddn = p.Ddn;
  const int lane = threadIdx.x & 63;
  const int quarter = lane >> 4;
  const int wave = threadIdx.x >> 6;
  const double* __restrict__ V = p.v;
  double dsum = 0.0;
  for (int t = 0; t < kXcd; ++t) {
    const int x = (home + t) & (kXcd - 1);
    const int nsl_x = p.nslabs > x ? (p.nslabs - x + kXcd - 1) / kXcd : 0;
    const int items = nsl_x * p.nrb;
    for (;;) {
      if (threadIdx.x == 0) s_item = atomicAdd(counters + x, 1);
      __syncthreads();
      const int item = s_item;
      __syncthreads();
      if (item >= items) break;
      const int slab = x + kXcd * (item / p.nrb);
      const int rb = item - (item / p.nrb) * p.nrb;
      const int c = slab * kSlabW + (lane & 15);
      const int r0 = rb * p.rows_per_wg;
      const int r1 = min(r0 + p.rows_per_wg, p.Dup);
      for (int r = r0 + wave * 4 + quarter; r < r1; r += (kSlabWG / 64) * 4) {
        if (c >= Ddn) continue;
        const int hb = p.up_rowp[r], he = p.up_rowp[r + 1];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int q = hb;
        for (; q + 3 < he; q += 4) {
          const uint32_t e0 = p.up_packed[q], e1 = p.up_packed[q + 1], e2 = p.up_packed[q + 2],
                         e3 = p.up_packed[q + 3];
          const double x0 = V[(size_t)(e0 & kColMask) * Ddn + c], x1 = V[(size_t)(e1 & kColMask) * Ddn + c];
          const double x2 = V[(size_t)(e2 & kColMask) * Ddn + c], x3 = V[(size_t)(e3 & kColMask) * Ddn + c];
          a0 = fma(p.values[e0 >> kColBits], x0, a0);
          a1 = fma(p.values[e1 >> kColBits], x1, a1);
          a2 = fma(p.values[e2 >> kColBits], x2, a2);
          a3 = fma(p.values[e3 >> kColBits], x3, a3);
        }
        for (; q < he; ++q) {
          const uint32_t e = p.up_packed[q];
          a0 = fma(p.values[e >> kColBits], V[(size_t)(e & kColMask) * Ddn + c], a0);
        }
        const double acc = (a0 + a1) + (a2 + a3);
        const size_t o = (size_t)r * Ddn + c;
        const double wn = __builtin_nontemporal_load(p.w + o) + acc * vscale;
        __builtin_nontemporal_store(wn, p.w + o);
        if (p.dot_partial) dsum = fma(V[o] * vscale, wn, dsum);
      }
    }
  }
  if (p.dot_partial) block_reduce_store(dsum, s_red, p.dot_partial + blockIdx.x);
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
  return (size_t)cached[dev] - 1024;  // room for the block's static reduction scratch
}

struct Launch {
  int J, block, grid;
  size_t smem;
  bool lds;
  int ef;  // down-spin entry format (see hv_rows_kernel)
};

// Columns per thread J (smallest power of two <= 8 keeping rows near 256
// threads) and threads per block T = 64 * ceil(ceil(Ddn / J) / 64) <= 1024;
// longer rows take several column chunks.
Launch plan(const DevSector& ds) {
  Launch L;
  const int64_t Ddn = ds.dn.D;
  const size_t rowb = (size_t)(Ddn + 1) * sizeof(double);
  L.lds = rowb <= lds_budget();
  L.ef = (L.lds && ds.dn.ell_q4 != nullptr) ? (ds.dn.nvt16 <= 2 ? 2 : 1) : 0;
  int J = 1;
  while (J < 8 && (Ddn + J - 1) / J > 256) J *= 2;
  int64_t T = ((Ddn + J - 1) / J + 63) / 64 * 64;
  T = std::min<int64_t>(std::max<int64_t>(T, 64), 1024);
  L.J = J;
  L.block = (int)T;
  L.grid = (int)std::min<int64_t>(ds.up.D, 1 << 20);
  L.smem = L.lds ? rowb : 0;
  return L;
}

template <int J, int CROSS, bool KAN, bool LDS, int EF>
void launch_t(const HvParams& p, const Launch& L, bool up, hipStream_t s) {
  auto k = up ? hv_rows_kernel<J, CROSS, KAN, LDS, EF, true> : hv_rows_kernel<J, CROSS, KAN, LDS, EF, false>;
  if (L.smem > 64 * 1024)
    EDH_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.smem));
  hipLaunchKernelGGL(k, dim3(L.grid), dim3(L.block), L.smem, s, p);
  EDH_HIP(hipGetLastError());
}

template <int J, int CROSS>
void launch_c(const HvParams& p, const Launch& L, bool kan, bool up, hipStream_t s) {
  if (kan) {
    if (L.ef == 2) launch_t<J, CROSS, true, true, 2>(p, L, up, s);
    else if (L.ef == 1) launch_t<J, CROSS, true, true, 1>(p, L, up, s);
    else if (L.lds) launch_t<J, CROSS, true, true, 0>(p, L, up, s);
    else launch_t<J, CROSS, true, false, 0>(p, L, up, s);
  } else {
    if (L.ef == 2) launch_t<J, CROSS, false, true, 2>(p, L, up, s);
    else if (L.ef == 1) launch_t<J, CROSS, false, true, 1>(p, L, up, s);
    else if (L.lds) launch_t<J, CROSS, false, true, 0>(p, L, up, s);
    else launch_t<J, CROSS, false, false, 0>(p, L, up, s);
  }
}

template <int J>
void launch_j(const HvParams& p, const Launch& L, int cross, bool kan, bool up, hipStream_t s) {
  switch (cross) {
    case CROSS_NONE: launch_c<J, CROSS_NONE>(p, L, kan, up, s); break;
    case CROSS_POPC: launch_c<J, CROSS_POPC>(p, L, kan, up, s); break;
    case CROSS_IMP: launch_c<J, CROSS_IMP>(p, L, kan, up, s); break;
    default: launch_c<J, CROSS_GENERAL>(p, L, kan, up, s); break;
  }
}

// Slab pass geometry; returns its (persistent) grid size.
int slab_geometry(const DevSector& ds, int& nslabs, int& rows_per_wg, int& nrb) {
  nslabs = (int)((ds.dn.D + kSlabW - 1) / kSlabW);
  rows_per_wg = 32;
  nrb = (int)((ds.up.D + rows_per_wg - 1) / rows_per_wg);
  const int64_t items = (int64_t)nslabs * nrb;
  return (int)std::min<int64_t>(items, 2048);
}

HvParams make_params(const DevSector& ds, const HvArgs& a) {
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
  p.dn_q4ptr = ds.dn.q4_ptr;
  p.dn_q4 = ds.dn.ell_q4;
  for (int q = 0; q < 4; ++q) p.dn_vt[q] = ds.dn.vt16[q];
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
  slab_geometry(ds, p.nslabs, p.rows_per_wg, p.nrb);
  return p;
}

}  // namespace

int hv_num_blocks(const DevSector& ds) {
  int nslabs, rpw, nrb;
  const int grid = slab_geometry(ds, nslabs, rpw, nrb);
  return ds.variant == HV_TWO_PASS ? grid : plan(ds).grid;
}

void launch_hv(const DevSector& ds, const HvArgs& a, hipStream_t s, int variant) {
  if (ds.N == 0) return;
  if (variant == HV_AUTO) variant = ds.variant;
  Launch L = plan(ds);
  HvParams p = make_params(ds, a);
  const bool kan = ds.kan_ml > 0;
  const bool up = variant != HV_TWO_PASS;
  switch (L.J) {
    case 1: launch_j<1>(p, L, ds.cross_mode, kan, up, s); break;
    case 2: launch_j<2>(p, L, ds.cross_mode, kan, up, s); break;
    case 4: launch_j<4>(p, L, ds.cross_mode, kan, up, s); break;
    default: launch_j<8>(p, L, ds.cross_mode, kan, up, s); break;
  }
  if (!up) {
    // the rows pass wrote the partial result; the slab pass adds the up-spin
    // term and produces the dot partials (the rows pass skips them)
    int nslabs, rpw, nrb;
    const int grid = slab_geometry(ds, nslabs, rpw, nrb);
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, ds.counters, kXcd);
    hipLaunchKernelGGL(hv_upslab_kernel, dim3(grid), dim3(kSlabWG), 0, s, p, ds.counters);
    EDH_HIP(hipGetLastError());
  }
}

// Times both variants on scratch vectors and keeps the faster one (an
// environment override EDH_HV_VARIANT=0/1 fixes the choice).
int hv_autotune(DevSector& ds, hipStream_t s) {
  const char* env = getenv("EDH_HV_VARIANT");
  if (env && (env[0] == '0' || env[0] == '1')) return env[0] - '0';
  if (ds.N < (1 << 16)) return HV_SINGLE;  // launch-bound sizes: one kernel
  double *v = nullptr, *w = nullptr;
  if (hipMalloc(&v, ds.N * sizeof(double)) != hipSuccess) return HV_SINGLE;
  if (hipMalloc(&w, ds.N * sizeof(double)) != hipSuccess) {
    (void)hipFree(v);
    return HV_SINGLE;
  }
  EDH_HIP(hipMemsetAsync(v, 0, ds.N * sizeof(double), s));
  EDH_HIP(hipMemsetAsync(w, 0, ds.N * sizeof(double), s));
  HvArgs a{};
  a.v = v;
  a.w = w;
  a.wprev = w;
  a.vscale = 1.0;
  a.wscale = 1.0;
  a.mode = 1;
  hipEvent_t e0, e1;
  EDH_HIP(hipEventCreate(&e0));
  EDH_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int choice = HV_SINGLE;
  for (int var = HV_SINGLE; var <= HV_TWO_PASS; ++var) {
    launch_hv(ds, a, s, var);  // warm-up
    EDH_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 2; ++r) launch_hv(ds, a, s, var);
    EDH_HIP(hipEventRecord(e1, s));
    EDH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    EDH_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      choice = var;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  EDH_HIP(hipStreamSynchronize(s));
  (void)hipFree(v);
  (void)hipFree(w);
  return choice;
}

}  // namespace edh
