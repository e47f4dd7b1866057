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
// Cuthill-McKee order of the up-spin hopping graph, 32 consecutive rows of
// that order to each XCD at a time).  The row is staged in
// LDS, so the down-spin gathers hit LDS; each thread owns J columns and keeps
// their accumulators in registers, so the up-spin term issues J independent
// coalesced row loads per neighbour.  The down-spin table is a sliced ELL
// (slice height 64 = one wavefront).  When the hopping values are {a} or
// {a, -a} (uniform hopping) each slot is the LDS byte offset of its column
// with the sign in bit 0, read four slots per 16-byte load, so a slot costs
// one mask, one sign insert and one FMA; otherwise, when the spin basis and
// its value set are small, 16-bit entries are packed by slot quads.  Either
// way all quads of a column group are loaded at once, one group ahead.
//
// Up-spin slab kernel (two-pass variant): the rows kernel runs without the
// up-spin term, then each XCD sweeps all rows over its own column slab of 16
// columns, so the slab (D_up x 128 B) stays in that XCD's L2 while every
// neighbour row segment is re-read from it.
#include <algorithm>
#include <cstdlib>
#include <cstring>

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
  // down spin
  const int32_t* dn_bits;
  const double* dn_e;
  const int32_t* dn_slptr;
  const int32_t* dn_sllen;
  const uint32_t* dn_ell;
  const int32_t* dn_q4ptr;
  const uint64_t* dn_q4;
  double dn_vt[4];
  const uint32_t* dn_s4;
  double dn_s4_scale;
  const double* values;
  int nvals;  // entries of values (staged in LDS after the row with EF == 4)
  int nkan;   // impurity terms (staged in LDS by the multi-operand kernel)
  // cross term
  double cross_u;
  const double* cross_imp;
  const double* cross_rowtab;
  // impurity interaction
  int kan_ml;
  const int32_t* kan_ptr;
  const int64_t* kan_off;  // partner state's element offset per term
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
  // slab kernel geometry; off32: every element offset fits 32 bits
  int nslabs, rows_per_wg, nrb, off32;
  int xcd_rows;  // rows kernel: XCD-grouped row assignment (measured 2 % faster on the 16-site sector)
  // rows kernel in a batch: an eigensolver run's control block (v and w are
  // then columns col and col + 1 of the basis in v or v_alt; a run whose
  // state is not 0 is skipped)
  const double* ctl;
  const double* v_alt;
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

// One signed byte-offset slot: a += (+-1) * row[e / 8], the sign in bit 0.
__device__ __forceinline__ double s4_fma(uint32_t e, const char* sv, double a) {
  const double sg = __hiloint2double((int)(0x3FF00000u | (e << 31)), 0);
  return fma(sg, *reinterpret_cast<const double*>(sv + (e & ~1u)), a);
}

// Quads held in registers per column group (slices with more quads finish
// in a plain loop).
constexpr int kQuadRegs = 6;
// Row staging: loads per thread issued before the LDS writes.
constexpr int kStage = 8;

// EF: down-spin entry format: 0 = 32-bit sliced ELL (column | value index),
// 1/2 = 16-bit entries in slot quads (four / two distinct values),
// 3 = signed byte offsets in 16-byte slot quads (values {a} or {a, -a}),
// 4 = the 32-bit sliced ELL with the value table staged in LDS after the row.
//
// Thread tid of a block with T threads owns columns c0 + tid + j*T (j < J)
// of each chunk c0; a wavefront therefore covers exactly one ELL slice per j,
// so slot counts are wavefront-uniform and the slot loops need no padding
// beyond their own slice.
// CTL: the operands are resolved from an eigensolver run's control block
// (batch sweeps only; a separate instantiation keeps the other launches'
// register allocation as it is).
template <int J, int CROSS, bool KAN, bool LDS, int EF, bool UP, bool CTL>
__global__ __launch_bounds__(1024) void hv_rows_kernel(HvParams p) {
  extern __shared__ double s_v[];
  __shared__ double s_red[16];
  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int Ddn = p.Ddn;
  const int na = 1 << (p.Ns - p.nb);
  if (p.done && *p.done) return;
  const double vscale = p.dscale ? p.dscale[0] : p.vscale;
  const double wscale = p.dscale ? p.dscale[1] : p.wscale;
  const double* __restrict__ V = p.v;
  double* __restrict__ W = p.w;
  if (CTL) {
    if (p.ctl[kCtlState] != 0.0) return;  // uniform over the grid
    const int64_t N = (int64_t)p.Dup * p.Ddn;
    const double* base = p.ctl[kCtlPhase] != 0.0 ? p.v_alt : p.v;
    const int64_t col = (int64_t)p.ctl[kCtlCol];
    V = base + col * N;
    W = const_cast<double*>(base) + (col + 1) * N;
  }
  const double* __restrict__ vals = p.values;
  double dsum = 0.0;
  // EF == 4: the value table after the row (the first row's barrier orders
  // these writes before any read)
  double* s_vals = s_v + Ddn + 1;
  if (EF == 4)
    for (int i = tid; i < p.nvals; i += T) s_vals[i] = vals[i];

  for (int slot0 = blockIdx.x; slot0 < p.Dup; slot0 += gridDim.x) {
    // work-groups go to the XCDs in turn (b mod 8): within each window of
    // 256 consecutive work-groups, XCD x takes 32 consecutive rows of the
    // order, so the rows in flight on one XCD share neighbour rows in its L2
    int slot = slot0;
    if (p.xcd_rows && slot0 < (p.Dup & ~255))
      slot = (slot0 & ~255) + ((slot0 & 7) << 5) + ((slot0 >> 3) & 31);
    const int u = __builtin_amdgcn_readfirstlane(p.row_order[slot]);
    const size_t rowoff = (size_t)u * Ddn;
    if (LDS) {
      // kStage loads per thread in flight together, then the LDS writes
      for (int d0 = 0; d0 < Ddn; d0 += kStage * T) {
        double x[kStage];
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
          const int d = d0 + tid + k * T;
          x[k] = d < Ddn ? V[rowoff + d] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
          const int d = d0 + tid + k * T;
          if (d < Ddn) s_v[d] = x[k] * vscale;
        }
      }
      if (EF && tid == 0) s_v[Ddn] = 0.0;  // target of the padding entries
      __syncthreads();
    }
    const int bu = p.up_bits[u];
    const double eu = p.up_e[u];
    const int hb = p.up_rowp[u], he = p.up_rowp[u + 1];
    for (int c0 = 0; c0 < Ddn; c0 += T * J) {
      // Columns past the row end are clamped to the last column: they are
      // loaded like the others but never stored.
#define DCOL(j) min(c0 + tid + (j) * T, Ddn - 1)
      double acc[J];
      // accumulated input (the J loads issued together) and diagonal
      // (generated from the bit patterns)
      double win[J];
      if (p.accumulate) {
#pragma unroll
        for (int j = 0; j < J; ++j) win[j] = __builtin_nontemporal_load(p.wprev + rowoff + DCOL(j));
      }
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int d = DCOL(j);
        const int bd = p.dn_bits[d];
        const double wp = p.accumulate ? wscale * win[j] : 0.0;
        const double vd = LDS ? s_v[d] : V[rowoff + d] * vscale;
        acc[j] = fma(eu + p.dn_e[d] + cross_term<CROSS>(p, u, bu, bd, na), vd, wp);
      }
      auto down_term = [&]() {
        // down-spin hopping, gathers from the staged row
        if (EF == 3) {
          // all quads of column group j are loaded at once and those of group
          // j+1 are in flight while group j is consumed
          const uint4* __restrict__ S4 = reinterpret_cast<const uint4*>(p.dn_s4);
          const char* sv = reinterpret_cast<const char*>(s_v);
          uint4 qa[kQuadRegs], qn[kQuadRegs];
          int nqa, nqn = 0;
          const uint4* qpa;
          {
            const int d0 = DCOL(0);
            const int sl = __builtin_amdgcn_readfirstlane(d0 >> 6);
            nqa = __builtin_amdgcn_readfirstlane((p.dn_sllen[sl] + 3) >> 2);
            qpa = S4 + p.dn_q4ptr[sl] + (d0 & 63);
  #pragma unroll
            for (int q = 0; q < kQuadRegs; ++q)
              if (q < nqa) qa[q] = qpa[q * kSlice];
          }
  #pragma unroll
          for (int j = 0; j < J; ++j) {
            const uint4* qpn = qpa;
            if (j + 1 < J) {
              const int d1 = DCOL(j + 1);
              const int sl = __builtin_amdgcn_readfirstlane(d1 >> 6);
              nqn = __builtin_amdgcn_readfirstlane((p.dn_sllen[sl] + 3) >> 2);
              qpn = S4 + p.dn_q4ptr[sl] + (d1 & 63);
  #pragma unroll
              for (int q = 0; q < kQuadRegs; ++q)
                if (q < nqn) qn[q] = qpn[q * kSlice];
            }
            double a0 = 0.0, a1 = 0.0;
  #pragma unroll
            for (int q = 0; q < kQuadRegs; ++q) {
              if (q < nqa) {
                a0 = s4_fma(qa[q].x, sv, a0);
                a1 = s4_fma(qa[q].y, sv, a1);
                a0 = s4_fma(qa[q].z, sv, a0);
                a1 = s4_fma(qa[q].w, sv, a1);
              }
            }
            for (int q = kQuadRegs; q < nqa; ++q) {
              const uint4 c = qpa[q * kSlice];
              a0 = s4_fma(c.x, sv, a0);
              a1 = s4_fma(c.y, sv, a1);
              a0 = s4_fma(c.z, sv, a0);
              a1 = s4_fma(c.w, sv, a1);
            }
            acc[j] = fma(p.dn_s4_scale, a0 + a1, acc[j]);
            if (j + 1 < J) {
  #pragma unroll
              for (int q = 0; q < kQuadRegs; ++q) qa[q] = qn[q];
              nqa = nqn;
              qpa = qpn;
            }
          }
        } else if (EF == 1 || EF == 2) {
          // 16-bit entries, four per 64-bit quad, same pipelining
          constexpr int kQ = 8;
          uint64_t qa[kQ], qn[kQ];
          int nqa, nqn = 0;
          const uint64_t* qpa;
          {
            const int d0 = DCOL(0);
            const int sl = __builtin_amdgcn_readfirstlane(d0 >> 6);
            nqa = __builtin_amdgcn_readfirstlane((p.dn_sllen[sl] + 3) >> 2);
            qpa = p.dn_q4 + p.dn_q4ptr[sl] + (d0 & 63);
  #pragma unroll
            for (int q = 0; q < kQ; ++q)
              if (q < nqa) qa[q] = qpa[q * kSlice];
          }
  #pragma unroll
          for (int j = 0; j < J; ++j) {
            const uint64_t* qpn = qpa;
            if (j + 1 < J) {
              const int d1 = DCOL(j + 1);
              const int sl = __builtin_amdgcn_readfirstlane(d1 >> 6);
              nqn = __builtin_amdgcn_readfirstlane((p.dn_sllen[sl] + 3) >> 2);
              qpn = p.dn_q4 + p.dn_q4ptr[sl] + (d1 & 63);
  #pragma unroll
              for (int q = 0; q < kQ; ++q)
                if (q < nqn) qn[q] = qpn[q * kSlice];
            }
            double a0 = 0.0, a1 = 0.0;
  #pragma unroll
            for (int q = 0; q < kQ; ++q) {
              if (q < nqa) {
                const uint64_t c = qa[q];
                const uint32_t e0 = (uint32_t)c & 0xFFFFu, e1 = (uint32_t)(c >> 16) & 0xFFFFu;
                const uint32_t e2 = (uint32_t)(c >> 32) & 0xFFFFu, e3 = (uint32_t)(c >> 48);
                a0 = fma(q4_value<EF>(p, e0), s_v[e0 & 0x3FFFu], a0);
                a1 = fma(q4_value<EF>(p, e1), s_v[e1 & 0x3FFFu], a1);
                a0 = fma(q4_value<EF>(p, e2), s_v[e2 & 0x3FFFu], a0);
                a1 = fma(q4_value<EF>(p, e3), s_v[e3 & 0x3FFFu], a1);
              }
            }
            for (int q = kQ; q < nqa; ++q) {
              const uint64_t c = qpa[q * kSlice];
  #pragma unroll
              for (int sub = 0; sub < 4; ++sub) {
                const uint32_t e = (uint32_t)(c >> (16 * sub)) & 0xFFFFu;
                a0 = fma(q4_value<EF>(p, e), s_v[e & 0x3FFFu], a0);
              }
            }
            acc[j] += a0 + a1;
            if (j + 1 < J) {
  #pragma unroll
              for (int q = 0; q < kQ; ++q) qa[q] = qn[q];
              nqa = nqn;
              qpa = qpn;
            }
          }
        } else {
          // 32-bit sliced ELL: four slots of all J columns loaded together,
          // then consumed in slot order
          int len[J], base[J];
          int maxlen = 0;
  #pragma unroll
          for (int j = 0; j < J; ++j) {
            const int sl = DCOL(j) >> 6;
            len[j] = p.dn_sllen[sl];
            base[j] = len[j] > 0 ? p.dn_slptr[sl] + (DCOL(j) & 63) : 0;
            maxlen = max(maxlen, len[j]);
          }
          const double* __restrict__ vt = EF == 4 ? s_vals : vals;
          // k < maxlen means some column has entries, so the array is not
          // empty; indices are clamped into each column's own slots, and a
          // column past its length reads entry 0 (value 0.0)
          constexpr int kSlots = 4;
          for (int k0 = 0; k0 < maxlen; k0 += kSlots) {
            uint32_t e[kSlots][J];
  #pragma unroll
            for (int s = 0; s < kSlots; ++s)
  #pragma unroll
              for (int j = 0; j < J; ++j) e[s][j] = p.dn_ell[base[j] + min(k0 + s, max(len[j] - 1, 0)) * kSlice];
  #pragma unroll
            for (int s = 0; s < kSlots; ++s) {
              if (k0 + s >= maxlen) break;
  #pragma unroll
              for (int j = 0; j < J; ++j) {
                const uint32_t x = k0 + s < len[j] ? e[s][j] : 0u;
                const int c = (int)(x & kColMask);
                acc[j] = fma(vt[x >> kColBits], LDS ? s_v[c] : V[rowoff + c] * vscale, acc[j]);
              }
            }
          }
        }
      };
      auto up_term = [&]() {
        // up-spin hopping: linear combination of whole rows (coalesced); the
        // loads of four (then two) neighbour rows are in flight together,
        // each column's FMAs still in neighbour order.  The quad forms (EF 1-3)
        // take one row at a time: their slot quads already fill the registers
        // (two at a time made the 16-site sector 4.05 -> 4.22 ms).
        constexpr bool kWide = EF == 0 || EF == 4;
        if (UP) {
          int q = hb;
          for (; kWide && q + 3 < he; q += 4) {
            uint32_t e[4];
            double h[4];
            const double* __restrict__ src[4];
  #pragma unroll
            for (int r = 0; r < 4; ++r) {
              e[r] = p.up_packed[q + r];
              h[r] = vals[e[r] >> kColBits] * vscale;
              src[r] = V + (size_t)(e[r] & kColMask) * Ddn;
            }
            double x[4][J];
  #pragma unroll
            for (int r = 0; r < 4; ++r)
  #pragma unroll
              for (int j = 0; j < J; ++j) x[r][j] = src[r][DCOL(j)];
  #pragma unroll
            for (int j = 0; j < J; ++j)
              acc[j] = fma(h[3], x[3][j], fma(h[2], x[2][j], fma(h[1], x[1][j], fma(h[0], x[0][j], acc[j]))));
          }
          for (; kWide && q + 1 < he; q += 2) {
            const uint32_t e0 = p.up_packed[q], e1 = p.up_packed[q + 1];
            const double h0 = vals[e0 >> kColBits] * vscale, h1 = vals[e1 >> kColBits] * vscale;
            const double* __restrict__ src0 = V + (size_t)(e0 & kColMask) * Ddn;
            const double* __restrict__ src1 = V + (size_t)(e1 & kColMask) * Ddn;
            double x0[J], x1[J];
  #pragma unroll
            for (int j = 0; j < J; ++j) {
              x0[j] = src0[DCOL(j)];
              x1[j] = src1[DCOL(j)];
            }
  #pragma unroll
            for (int j = 0; j < J; ++j) acc[j] = fma(h1, x1[j], fma(h0, x0[j], acc[j]));
          }
          for (; q < he; ++q) {
            const uint32_t e = p.up_packed[q];
            const double h = vals[e >> kColBits] * vscale;
            const double* __restrict__ src = V + (size_t)(e & kColMask) * Ddn;
            double x[J];
  #pragma unroll
            for (int j = 0; j < J; ++j) x[j] = src[DCOL(j)];
  #pragma unroll
            for (int j = 0; j < J; ++j) acc[j] = fma(h, x[j], acc[j]);
          }
        }
      };
      // wavefronts of odd index take the up-spin rows first: while half of a
      // work-group's wavefronts gather from the row in LDS, the other half
      // wait on neighbour rows from L2 and beyond (each column's sum then
      // runs diagonal, up-spin, down-spin terms instead of diagonal,
      // down-spin, up-spin)
      if (UP && ((tid >> 6) & 1)) {
        up_term();
        down_term();
      } else {
        down_term();
        up_term();
      }
      if (KAN) {
        const int a = bu >> p.nb;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int bd = p.dn_bits[DCOL(j)];
          const int b = bd >> p.nb;
          const int kb = p.kan_ptr[a * na + b], ke = p.kan_ptr[a * na + b + 1];
          for (int q = kb; q < ke; ++q)
            acc[j] = fma(p.kan_c[q], V[(int64_t)rowoff + DCOL(j) + p.kan_off[q]] * vscale, acc[j]);
        }
      }
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int d = c0 + tid + j * T;
        if (d < Ddn) {
          __builtin_nontemporal_store(acc[j], W + rowoff + d);
          if (UP && p.dot_partial) dsum = fma(LDS ? s_v[d] : V[rowoff + d] * vscale, acc[j], dsum);
        }
      }
#undef DCOL
    }
    if (LDS) __syncthreads();
  }
  if (UP && p.dot_partial) block_reduce_store(dsum, s_red, p.dot_partial + blockIdx.x);
}

// Several Lanczos operands of one sector in one launch (the batch's
// Green's-function runs on sectors routed to the rows kernel).  Each operand
// gets exactly the arithmetic of hv_rows_kernel<4, CROSS, KAN, true, EF,
// true, false> -- the same column ownership, the same order of its terms
// (diagonal, then down- and up-spin hopping in the wavefront's order, then
// the impurity terms) and the same share of <v, w> per work-group -- so its
// result does not depend on which operands share the launch; the sector
// tables, the diagonal and the down-spin entries are read once for all of
// them, and their rows sit side by side in LDS.  Operands nk..K-1 and
// stopped runs are skipped (uniform over the grid).
struct HvMulti {
  const double* v[kMultiMax];
  double* w[kMultiMax];
  const double* wprev[kMultiMax];
  const double* dscale[kMultiMax];  // {vscale, wscale} on the device
  const int* done[kMultiMax];
  double* dot[kMultiMax];
  int accumulate[kMultiMax];
  int nk;
};

template <int CROSS, bool KAN, int EF>
__global__ __launch_bounds__(1024) void hv_rows_multi_kernel(HvParams p, HvMulti m) {
  constexpr int J = 4, K = kMultiMax;
  extern __shared__ double s_v[];  // K rows of Ddn + 1, then (EF 4) the value table
  __shared__ double s_red[16];
  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int Ddn = p.Ddn;
  const int ld = Ddn + 1;
  const int na = 1 << (p.Ns - p.nb);
  bool live[K];
  double vs[K], ws[K], dsum[K];
  bool any = false;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    live[k] = k < m.nk && !(m.done[k] && *m.done[k]);
    vs[k] = live[k] ? m.dscale[k][0] : 0.0;
    ws[k] = live[k] ? m.dscale[k][1] : 0.0;
    dsum[k] = 0.0;
    any = any || live[k];
  }
  if (!any) return;
  const double* __restrict__ vals = p.values;
  double* s_vals = s_v + K * ld;
  if (EF == 4)
    for (int i = tid; i < p.nvals; i += T) s_vals[i] = vals[i];
  // the impurity terms' tables after the values (the first row's barrier
  // orders these writes before any read): coefficients, element offsets,
  // then the (a, b) pointers
  double* s_kc = s_vals + (EF == 4 ? p.nvals : 0);
  int64_t* s_koff = reinterpret_cast<int64_t*>(s_kc + (KAN ? p.nkan : 0));
  int* s_kptr = reinterpret_cast<int*>(s_koff + (KAN ? p.nkan : 0));
  if (KAN) {
    for (int i = tid; i < p.nkan; i += T) {
      s_kc[i] = p.kan_c[i];
      s_koff[i] = p.kan_off[i];
    }
    for (int i = tid; i <= na * na; i += T) s_kptr[i] = p.kan_ptr[i];
  }

  for (int slot0 = blockIdx.x; slot0 < p.Dup; slot0 += gridDim.x) {
    int slot = slot0;
    if (p.xcd_rows && slot0 < (p.Dup & ~255))
      slot = (slot0 & ~255) + ((slot0 & 7) << 5) + ((slot0 >> 3) & 31);
    const int u = __builtin_amdgcn_readfirstlane(p.row_order[slot]);
    const size_t rowoff = (size_t)u * Ddn;
    // the rows of all operands: four elements of each per thread and round,
    // every load of a round in flight together
    constexpr int kStageM = 4;
    for (int d0 = 0; d0 < Ddn; d0 += kStageM * T) {
      double x[K][kStageM];
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int q = 0; q < kStageM; ++q) {
          const int d = d0 + tid + q * T;
          x[k][q] = (live[k] && d < Ddn) ? m.v[k][rowoff + d] : 0.0;
        }
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int q = 0; q < kStageM; ++q) {
          const int d = d0 + tid + q * T;
          if (live[k] && d < Ddn) s_v[k * ld + d] = x[k][q] * vs[k];
        }
    }
    if (EF && tid == 0)
      for (int k = 0; k < K; ++k) s_v[k * ld + Ddn] = 0.0;
    __syncthreads();
    const int bu = p.up_bits[u];
    const double eu = p.up_e[u];
    const int hb = p.up_rowp[u], he = p.up_rowp[u + 1];
    for (int c0 = 0; c0 < Ddn; c0 += T * J) {
#define DCOL(j) min(c0 + tid + (j) * T, Ddn - 1)
      double acc[K][J];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        double win[J];
        if (live[k] && m.accumulate[k]) {
#pragma unroll
          for (int j = 0; j < J; ++j) win[j] = __builtin_nontemporal_load(m.wprev[k] + rowoff + DCOL(j));
        }
#pragma unroll
        for (int j = 0; j < J; ++j) acc[k][j] = (live[k] && m.accumulate[k]) ? ws[k] * win[j] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int d = DCOL(j);
        const int bd = p.dn_bits[d];
        const double dg = eu + p.dn_e[d] + cross_term<CROSS>(p, u, bu, bd, na);
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (live[k]) acc[k][j] = fma(dg, s_v[k * ld + d], acc[k][j]);
      }
      auto down_term = [&]() {
        int len[J], base[J];
        int maxlen = 0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int sl = DCOL(j) >> 6;
          len[j] = p.dn_sllen[sl];
          base[j] = len[j] > 0 ? p.dn_slptr[sl] + (DCOL(j) & 63) : 0;
          maxlen = max(maxlen, len[j]);
        }
        const double* __restrict__ vt = EF == 4 ? s_vals : vals;
        constexpr int kSlots = 4;
        for (int k0 = 0; k0 < maxlen; k0 += kSlots) {
          uint32_t e[kSlots][J];
#pragma unroll
          for (int sq = 0; sq < kSlots; ++sq)
#pragma unroll
            for (int j = 0; j < J; ++j) e[sq][j] = p.dn_ell[base[j] + min(k0 + sq, max(len[j] - 1, 0)) * kSlice];
#pragma unroll
          for (int sq = 0; sq < kSlots; ++sq) {
            if (k0 + sq >= maxlen) break;
#pragma unroll
            for (int j = 0; j < J; ++j) {
              const uint32_t x = k0 + sq < len[j] ? e[sq][j] : 0u;
              const int c = (int)(x & kColMask);
              const double val = vt[x >> kColBits];
#pragma unroll
              for (int k = 0; k < K; ++k)
                if (live[k]) acc[k][j] = fma(val, s_v[k * ld + c], acc[k][j]);
            }
          }
        }
      };
      auto up_term = [&]() {
        // one neighbour row at a time for all operands (each operand's FMAs
        // still in neighbour order)
        for (int q = hb; q < he; ++q) {
          const uint32_t e = p.up_packed[q];
          const double hv = vals[e >> kColBits];
          const size_t nrow = (size_t)(e & kColMask) * Ddn;
          double x[K][J];
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (live[k])
#pragma unroll
              for (int j = 0; j < J; ++j) x[k][j] = m.v[k][nrow + DCOL(j)];
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (live[k]) {
              const double h = hv * vs[k];
#pragma unroll
              for (int j = 0; j < J; ++j) acc[k][j] = fma(h, x[k][j], acc[k][j]);
            }
        }
      };
      if ((tid >> 6) & 1) {
        up_term();
        down_term();
      } else {
        down_term();
        up_term();
      }
      if (KAN) {
        const int a = bu >> p.nb;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int bd = p.dn_bits[DCOL(j)];
          const int b = bd >> p.nb;
          const int kb = s_kptr[a * na + b], ke = s_kptr[a * na + b + 1];
          for (int q = kb; q < ke; ++q) {
            const double c = s_kc[q];
            const int64_t off = (int64_t)rowoff + DCOL(j) + s_koff[q];
#pragma unroll
            for (int k = 0; k < K; ++k)
              if (live[k]) acc[k][j] = fma(c, m.v[k][off] * vs[k], acc[k][j]);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int d = c0 + tid + j * T;
        if (d < Ddn) {
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (live[k]) {
              __builtin_nontemporal_store(acc[k][j], m.w[k] + rowoff + d);
              dsum[k] = fma(s_v[k * ld + d], acc[k][j], dsum[k]);
            }
        }
      }
#undef DCOL
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!live[k]) continue;
    block_reduce_store(dsum[k], s_red, m.dot[k] + blockIdx.x);
    __syncthreads();  // s_red is reused by the next operand's reduction
  }
}

// Up-spin term over column slabs of 16 (one 128-byte line per row segment):
// w[u][c] += sum_k Hup[u][k] v[k][c].  Persistent work-groups read the XCD
// they run on (HW_REG_XCC_ID) and claim (slab, row block) items from that
// XCD's queue: XCD x owns slabs x, x+8, x+16, ... and its work-groups sweep
// all row blocks of one slab before the next, so the slab (D_up x 128 B)
// stays in that XCD's L2 while the ~17 neighbour segments of each row are
// re-read from it.  A work-group that finds its queue empty takes items from
// the other XCDs' queues, so every item is processed whatever XCD ids are
// reported.  The CSR entries of an item's rows are staged in LDS as (row
// offset, value) pairs; a wavefront covers four rows of the slab.
constexpr int kSlabW = 16;
constexpr int kSlabWG = 256;
constexpr int kSlabRows = 64;
constexpr int kSlabEnt = 2048;
constexpr int kXcd = 8;

__global__ void zero_counters_kernel(int* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kSlabWG) void hv_upslab_kernel(HvParams p, int* counters) {
  __shared__ double s_red[16];
  __shared__ int s_item;
  __shared__ uint32_t s_off[kSlabEnt];
  __shared__ double s_val[kSlabEnt];
  if (p.done && *p.done) return;
  const double vscale = p.dscale ? p.dscale[0] : p.vscale;
  unsigned xcc;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
  const int home = (int)(xcc & (kXcd - 1));
  const int Ddn = p.Ddn;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int quarter = lane >> 4;
  const int wave = tid >> 6;
  const double* __restrict__ V = p.v;
  double dsum = 0.0;
  for (int t = 0; t < kXcd; ++t) {
    const int x = (home + t) & (kXcd - 1);
    const int nsl_x = p.nslabs > x ? (p.nslabs - x + kXcd - 1) / kXcd : 0;
    const int items = nsl_x * p.nrb;
    for (;;) {
      if (tid == 0) s_item = atomicAdd(counters + x, 1);
      __syncthreads();  // also: the previous item's readers of s_off/s_val are done
      const int item = s_item;
      if (item >= items) {
        __syncthreads();  // everyone has read s_item before it is rewritten
        break;
      }
      const int slab = x + kXcd * (item / p.nrb);
      const int rb = item - (item / p.nrb) * p.nrb;
      const int c = slab * kSlabW + (lane & 15);
      const int r0 = rb * p.rows_per_wg;
      const int r1 = min(r0 + p.rows_per_wg, p.Dup);
      const int e0 = p.up_rowp[r0], ne = p.up_rowp[r1] - e0;
      const bool staged = p.off32 && ne <= kSlabEnt;
      if (staged)
        for (int i = tid; i < ne; i += kSlabWG) {
          const uint32_t e = p.up_packed[e0 + i];
          s_off[i] = (e & kColMask) * (uint32_t)Ddn;
          s_val[i] = p.values[e >> kColBits];
        }
      __syncthreads();
      for (int r = r0 + wave * 4 + quarter; r < r1; r += (kSlabWG / 64) * 4) {
        if (c >= Ddn) continue;
        const int hb = p.up_rowp[r], he = p.up_rowp[r + 1];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        if (staged) {
          int q = hb - e0;
          const int qe = he - e0;
          for (; q + 3 < qe; q += 4) {
            const double x0 = V[s_off[q] + c], x1 = V[s_off[q + 1] + c];
            const double x2 = V[s_off[q + 2] + c], x3 = V[s_off[q + 3] + c];
            a0 = fma(s_val[q], x0, a0);
            a1 = fma(s_val[q + 1], x1, a1);
            a2 = fma(s_val[q + 2], x2, a2);
            a3 = fma(s_val[q + 3], x3, a3);
          }
          for (; q < qe; ++q) a0 = fma(s_val[q], V[s_off[q] + c], a0);
        } else {
          for (int q = hb; q < he; ++q) {
            const uint32_t e = p.up_packed[q];
            a0 = fma(p.values[e >> kColBits], V[(size_t)(e & kColMask) * Ddn + c], a0);
          }
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


// ---------------------------------------------------------------------------
// Grouped H*v: every operand of a batch in one launch.  The up-rows of all
// operands are cut into work units (operand, first row, row count; about
// 4K elements each) listed on the host, and work-group b takes units b,
// b + grid, ...  The down-spin hopping reads the 32-bit sliced ELL (present
// for every sector), so any model and value set is handled.
struct HvBatchJob {
  HvParams p;
  int64_t row0;
  int cross, kan;
  const double* ctl;    // eigensolver control block or null
  const double* v_alt;  // second basis buffer (with ctl)
  double* tdot;         // per-task shares of <v, w>, or null
  const double* lzk;    // Lanczos run's step counter (buffers in rotation) or null
};

// One element of the grouped H*v: w[u][d] (written to W) from the operand
// V scaled by vscale; returns v[u][d] * w[u][d] (scaled v).  With STAGED,
// R is row u of V copied into LDS, and the element's own value and its
// down-spin gathers are read from there; the arithmetic (operands and
// order) is the same either way, so the result does not depend on it.
template <bool STAGED>
__device__ __forceinline__ double hv_batch_element(const HvBatchJob& J, const HvParams& p, const double* __restrict__ V,
                                                   double* __restrict__ W, const double* __restrict__ Wprev,
                                                   bool accum, int u, int d, double vscale, double wscale, int na,
                                                   const double* R) {
  const int Ddn = p.Ddn;
  const double* __restrict__ vals = p.values;
  const int cross = J.cross;
  const int sl = d >> 6;
  const size_t rowoff = (size_t)u * Ddn;
  if (!STAGED) R = V + rowoff;

  const int bu = p.up_bits[u];
  const int bd = p.dn_bits[d];
  double x;
  if (cross == CROSS_POPC) x = cross_term<CROSS_POPC>(p, u, bu, bd, na);
  else if (cross == CROSS_IMP) x = cross_term<CROSS_IMP>(p, u, bu, bd, na);
  else if (cross == CROSS_GENERAL) x = cross_term<CROSS_GENERAL>(p, u, bu, bd, na);
  else x = 0.0;
  double acc = accum ? wscale * Wprev[rowoff + d] : 0.0;
  const double vself = R[d];
  acc = fma(p.up_e[u] + p.dn_e[d] + x, vself * vscale, acc);
  // down-spin hopping (sliced ELL, slice = d / 64)
  const int len = p.dn_sllen[sl];
  const uint32_t* __restrict__ el = p.dn_ell + p.dn_slptr[sl] + (d & 63);
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    const uint32_t e0 = el[k * kSlice], e1 = el[(k + 1) * kSlice];
    const uint32_t e2 = el[(k + 2) * kSlice], e3 = el[(k + 3) * kSlice];
    const double x0 = R[e0 & kColMask], x1 = R[e1 & kColMask];
    const double x2 = R[e2 & kColMask], x3 = R[e3 & kColMask];
    acc = fma(vals[e0 >> kColBits], x0 * vscale, acc);
    acc = fma(vals[e1 >> kColBits], x1 * vscale, acc);
    acc = fma(vals[e2 >> kColBits], x2 * vscale, acc);
    acc = fma(vals[e3 >> kColBits], x3 * vscale, acc);
  }
  for (; k < len; ++k) {
    const uint32_t e = el[k * kSlice];
    acc = fma(vals[e >> kColBits], R[e & kColMask] * vscale, acc);
  }
  // up-spin hopping: rows of the same columns (coalesced)
  const int hb = p.up_rowp[u], he = p.up_rowp[u + 1];
  int q = hb;
  for (; q + 4 <= he; q += 4) {
    const uint32_t e0 = p.up_packed[q], e1 = p.up_packed[q + 1];
    const uint32_t e2 = p.up_packed[q + 2], e3 = p.up_packed[q + 3];
    const double x0 = V[(size_t)(e0 & kColMask) * Ddn + d], x1 = V[(size_t)(e1 & kColMask) * Ddn + d];
    const double x2 = V[(size_t)(e2 & kColMask) * Ddn + d], x3 = V[(size_t)(e3 & kColMask) * Ddn + d];
    acc = fma(vals[e0 >> kColBits] * vscale, x0, acc);
    acc = fma(vals[e1 >> kColBits] * vscale, x1, acc);
    acc = fma(vals[e2 >> kColBits] * vscale, x2, acc);
    acc = fma(vals[e3 >> kColBits] * vscale, x3, acc);
  }
  for (; q < he; ++q) {
    const uint32_t e = p.up_packed[q];
    acc = fma(vals[e >> kColBits] * vscale, V[(size_t)(e & kColMask) * Ddn + d], acc);
  }
  if (J.kan) {
    const int a = bu >> p.nb, b = bd >> p.nb;
    const int kb = p.kan_ptr[a * na + b], ke = p.kan_ptr[a * na + b + 1];
    for (int r = kb; r < ke; ++r) acc = fma(p.kan_c[r], V[(int64_t)rowoff + d + p.kan_off[r]] * vscale, acc);
  }
  W[rowoff + d] = acc;
  return vself * vscale * acc;
}

__device__ __forceinline__ double wave_sum64(double x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// Wavefront-task form of the grouped H*v: no LDS and no block barrier.  A
// unit's rows are cut into wavefront tasks: 64-column slices of one row when
// the down dimension is 64 or more (a slice is exactly one ELL slice, so the
// slot count is wavefront-uniform), else R = 64 / Dp whole rows per
// wavefront (Dp = the down dimension rounded up to a power of two).  The
// loads of four slots are issued before their FMAs.  Units start at multiples of R
// rows, so a task's (row group, slice) does not depend on the grouping; a
// Lanczos run's share of <v, w> is stored there (tdot), one value per task.
constexpr int kStageMax = 4608;  // elements of the largest staged unit (36 KB)

// A unit of at most lds_elems elements (its rows are contiguous in V) is
// first copied into LDS by the whole work-group, so the down-spin gathers
// and the element's own value come from LDS rather than through L1.
__global__ __launch_bounds__(512) void hv_batch_wave_kernel(const HvBatchJob* __restrict__ jobs,
                                                            const HvBatchUnit* __restrict__ units, int nunits,
                                                            int lds_elems) {
  extern __shared__ double s_unit[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  for (int ui = blockIdx.x; ui < nunits; ui += gridDim.x) {
    const HvBatchUnit U = units[ui];
    const HvBatchJob& J = jobs[U.job];
    const HvParams& p = J.p;
    if (p.done && *p.done) continue;  // a stopped Lanczos run (uniform over the block)
    const double* __restrict__ V = p.v;
    double* __restrict__ W = p.w;
    const double* __restrict__ Wprev = p.wprev;
    bool accum = p.accumulate != 0;
    if (J.ctl) {  // eigensolver run: its basis column and buffer
      if (J.ctl[kCtlState] != 0.0) continue;
      const int64_t N = (int64_t)p.Dup * p.Ddn;
      const double* base = J.ctl[kCtlPhase] != 0.0 ? J.v_alt : p.v;
      const int64_t col = (int64_t)J.ctl[kCtlCol];
      V = base + col * N;
      W = const_cast<double*>(base) + (col + 1) * N;
    } else if (J.lzk) {  // Lanczos run: its buffers in rotation by the step counter
      const int k = (int)*J.lzk;
      const double* b[3] = {p.v, p.wprev, p.w};
      const int r0 = lz_buffer(k, 0), r1 = lz_buffer(k, 1), r2 = lz_buffer(k, 2);
      V = r0 == 0 ? b[0] : (r0 == 1 ? b[1] : b[2]);
      Wprev = r1 == 0 ? b[0] : (r1 == 1 ? b[1] : b[2]);
      W = const_cast<double*>(r2 == 0 ? b[0] : (r2 == 1 ? b[1] : b[2]));
      accum = k > 1;
    }
    const int Ddn = p.Ddn;
    const int lg = Ddn >= 64 ? 6 : (Ddn <= 1 ? 0 : 32 - __clz(Ddn - 1));  // log2 Dp
    const int nsl = (Ddn + 63) >> 6;
    const int ngr = (U.nrows + (64 >> lg) - 1) >> (6 - lg);
    const int ntask = ngr * nsl;
    const double vscale = p.dscale ? p.dscale[0] : p.vscale;
    const double wscale = p.dscale ? p.dscale[1] : p.wscale;
    const int na = 1 << (p.Ns - p.nb);
    // the unit's rows are V[row0 * Ddn, (row0 + nrows) * Ddn): staged when
    // they fit (uniform over the work-group)
    const int span = U.nrows * Ddn;
    const bool staged = span <= lds_elems;
    if (staged) {
      const double* __restrict__ src = V + (size_t)U.row0 * Ddn;
      const int T = blockDim.x;
      int i = threadIdx.x;
      for (; i + 3 * T < span; i += 4 * T) {
        const double x0 = src[i], x1 = src[i + T], x2 = src[i + 2 * T], x3 = src[i + 3 * T];
        s_unit[i] = x0;
        s_unit[i + T] = x1;
        s_unit[i + 2 * T] = x2;
        s_unit[i + 3 * T] = x3;
      }
      for (; i < span; i += T) s_unit[i] = src[i];
      __syncthreads();
    }
    for (int t = wave; t < ntask; t += nw) {
      const int g = t / nsl, sl = t - g * nsl;
      const int u = U.row0 + (g << (6 - lg)) + (lane >> lg);
      const int d = (sl << 6) + (lane & ((1 << lg) - 1));
      const bool live = u < U.row0 + U.nrows && d < Ddn;
      double vw = 0.0;
      if (live) {
        vw = staged ? hv_batch_element<true>(J, p, V, W, Wprev, accum, u, d, vscale, wscale, na,
                                             s_unit + (size_t)(u - U.row0) * Ddn)
                    : hv_batch_element<false>(J, p, V, W, Wprev, accum, u, d, vscale, wscale, na, nullptr);
      }
      if (J.tdot) {
        // Lanczos runs: this task's share of <v, w>, one wavefront tree,
        // at (row group, slice) -- the same place whatever the grouping
        const double t2 = wave_sum64(vw);
        if (lane == 0) J.tdot[(size_t)((U.row0 >> (6 - lg)) + g) * nsl + sl] = t2;
      }
    }
    if (staged) __syncthreads();  // the next unit overwrites the staged rows
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
  return (size_t)cached[dev] - 1024;  // room for the block's static reduction scratch
}

struct Launch {
  int J, block, grid;
  size_t smem;
  bool lds;
  int ef;  // down-spin entry format (see hv_rows_kernel)
};

// Columns per thread J (smallest power of two <= 4 keeping rows near 256
// threads) and threads per block T <= 1024, a multiple of 64; longer rows
// take several column chunks of T * J columns.
Launch plan(const DevSector& ds) {
  Launch L;
  const int64_t Ddn = ds.dn.D;
  const size_t rowb = (size_t)(Ddn + 1) * sizeof(double);
  L.lds = rowb <= lds_budget();
  L.ef = 0;
  size_t valb = 0;
  if (L.lds && ds.dn.ell_s4 != nullptr) L.ef = 3;
  else if (L.lds && ds.dn.ell_q4 != nullptr) L.ef = ds.dn.nvt16 <= 2 ? 2 : 1;
  else if (L.lds && rowb + (size_t)ds.nvalues * sizeof(double) <= lds_budget()) {
    L.ef = 4;  // 32-bit entries with the value table beside the row in LDS
    valb = (size_t)ds.nvalues * sizeof(double);
  }
  int J = 1;
  while (J < 4 && (Ddn + J - 1) / J > 256) J *= 2;
  // (J, largest work-group) = (4, 1024) is within 5 % of the best of
  // (1..4, 256..1024) on 14- and 16-site sectors of all three models
  const int64_t tmax = 1024;
  // as few chunks as tmax threads allow, split evenly so that few lanes idle
  const int64_t nch = (Ddn + (int64_t)J * tmax - 1) / ((int64_t)J * tmax);
  int64_t T = ((Ddn + J * nch - 1) / (J * nch) + 63) / 64 * 64;
  T = std::min<int64_t>(std::max<int64_t>(T, 64), tmax);
  L.J = J;
  L.block = (int)T;
  L.grid = (int)std::min<int64_t>(ds.up.D, 1 << 20);
  L.smem = L.lds ? rowb + valb : 0;
  return L;
}

template <int J, int CROSS, bool KAN, bool LDS, int EF>
void launch_t(const HvParams& p, const Launch& L, bool up, hipStream_t s) {
  auto k = !up     ? hv_rows_kernel<J, CROSS, KAN, LDS, EF, false, false>
           : p.ctl ? hv_rows_kernel<J, CROSS, KAN, LDS, EF, true, true>
                   : hv_rows_kernel<J, CROSS, KAN, LDS, EF, true, false>;
  if (L.smem > 64 * 1024)
    EDH_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.smem));
  hipLaunchKernelGGL(k, dim3(L.grid), dim3(L.block), L.smem, s, p);
  EDH_HIP(hipGetLastError());
}

template <int J, int CROSS>
void launch_c(const HvParams& p, const Launch& L, bool kan, bool up, hipStream_t s) {
  if (kan) {
    if (L.ef == 4) launch_t<J, CROSS, true, true, 4>(p, L, up, s);
    else if (L.ef == 3) launch_t<J, CROSS, true, true, 3>(p, L, up, s);
    else if (L.ef == 2) launch_t<J, CROSS, true, true, 2>(p, L, up, s);
    else if (L.ef == 1) launch_t<J, CROSS, true, true, 1>(p, L, up, s);
    else if (L.lds) launch_t<J, CROSS, true, true, 0>(p, L, up, s);
    else launch_t<J, CROSS, true, false, 0>(p, L, up, s);
  } else {
    if (L.ef == 4) launch_t<J, CROSS, false, true, 4>(p, L, up, s);
    else if (L.ef == 3) launch_t<J, CROSS, false, true, 3>(p, L, up, s);
    else if (L.ef == 2) launch_t<J, CROSS, false, true, 2>(p, L, up, s);
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
  rows_per_wg = kSlabRows;
  nrb = (int)((ds.up.D + rows_per_wg - 1) / rows_per_wg);
  const int64_t items = (int64_t)nslabs * nrb;
  return (int)std::min<int64_t>(items, 2048);
}

HvParams make_params(const DevSector& ds, const HvArgs& a) {
  HvParams p;
  std::memset(&p, 0, sizeof(p));  // padding too: batched sweeps compare descriptor bytes
  p.Ns = ds.Ns;
  p.nb = ds.Ns - ds.imp_ml;
  p.Dup = (int)ds.up.D;
  p.Ddn = (int)ds.dn.D;
  p.row_order = ds.row_order;
  p.up_bits = ds.up.bits;
  p.up_e = ds.up.ediag;
  p.up_rowp = ds.up.rowp;
  p.up_packed = ds.up.packed;
  p.dn_bits = ds.dn.bits;
  p.dn_e = ds.dn.ediag;
  p.dn_slptr = ds.dn.sl_ptr;
  p.dn_sllen = ds.dn.sl_len;
  p.dn_ell = ds.dn.ell;
  p.dn_q4ptr = ds.dn.q4_ptr;
  p.dn_q4 = ds.dn.ell_q4;
  p.dn_s4 = ds.dn.ell_s4;
  p.dn_s4_scale = ds.dn.s4_scale;
  for (int q = 0; q < 4; ++q) p.dn_vt[q] = ds.dn.vt16[q];
  p.values = ds.values;
  p.nvals = ds.nvalues;
  p.nkan = ds.nkan;
  p.cross_u = ds.cross_u;
  p.cross_imp = ds.cross_imp;
  p.cross_rowtab = ds.cross_rowtab;
  p.kan_ml = ds.kan_ml;
  p.kan_ptr = ds.kan_ptr;
  p.kan_off = ds.kan_off;
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
  p.off32 = (uint64_t)ds.N < (uint64_t(1) << 32);
  p.xcd_rows = 1;
  p.ctl = a.ctl;
  p.v_alt = a.v_alt;
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
    default: launch_j<4>(p, L, ds.cross_mode, kan, up, s); break;
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
  if (ds.N < (1 << 22)) return HV_SINGLE;  // short rows: the slab pass never pays
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

namespace edh {

size_t hv_batch_job_bytes() { return sizeof(HvBatchJob); }

void hv_batch_fill(const DevSector& ds, const HvArgs& a, int64_t row0, void* out) {
  HvBatchJob j;
  std::memset(&j, 0, sizeof(j));
  j.p = make_params(ds, a);
  j.row0 = row0;
  j.cross = ds.cross_mode;
  j.kan = ds.kan_ml > 0 ? 1 : 0;
  j.ctl = a.ctl;
  j.v_alt = a.v_alt;
  j.tdot = a.tdot;
  j.lzk = a.lzk;
  std::memcpy(out, &j, sizeof(j));
}

int hv_batch_row_group(int64_t ddn) {
  if (ddn >= 64) return 1;
  int dp = 1;
  while (dp < ddn) dp *= 2;
  return 64 / dp;
}

int hv_batch_unit_rows(int64_t ddn) {
  const int r = (int)std::max<int64_t>(1, std::min<int64_t>(64, 4096 / std::max<int64_t>(ddn, 1)));
  const int g = hv_batch_row_group(ddn);
  return (r + g - 1) / g * g;
}

int64_t hv_batch_tasks(int64_t dup, int64_t ddn) {
  const int g = hv_batch_row_group(ddn);
  return (dup + g - 1) / g * ((ddn + 63) / 64);
}

int hv_batch_lds_elems(const std::vector<int64_t>& spans) {
  // units of up to kStageMax elements (the usual case) are staged, with an
  // allocation that leaves room for four 512-thread work-groups per CU;
  // longer rows are read through the caches (staging one 100-KB row per
  // work-group measured 1.4-2x slower on the 16-site sectors)
  int64_t mx = 0;
  for (int64_t s : spans)
    if (s <= kStageMax) mx = std::max(mx, s);
  return (int)mx;
}

void launch_hv_batch(const void* djobs, const HvBatchUnit* dunits, int nunits, int block, int lds_elems,
                     hipStream_t s) {
  if (nunits <= 0) return;
  EDH_CHECK(block >= 64 && block <= 512 && block % 64 == 0, EDH_ERR_ARG, "grouped H*v: bad block size");
  EDH_CHECK(lds_elems >= 0 && lds_elems <= kStageMax, EDH_ERR_ARG, "grouped H*v: staged unit too large");
  const int grid = std::min(nunits, 1 << 20);
  const size_t smem = (size_t)lds_elems * sizeof(double);
  hipLaunchKernelGGL(hv_batch_wave_kernel, dim3(grid), dim3(block), smem, s, (const HvBatchJob*)djobs, dunits,
                     nunits, lds_elems);
  EDH_HIP(hipGetLastError());
}

// Large operands go to the rows kernel: rows too long to stage in the
// grouped kernel, or at least 1024 rows of at least 1024 elements (its
// row-per-work-group form then beats the grouped per-element form: 14-site
// Kanamori sector (7,7), 11.8 M states, 0.40 against 0.56 ms per H*v).
bool hv_batch_routed(int64_t dup, int64_t ddn) {
  return dup * ddn >= (int64_t(1) << 22) && (ddn > kStageMax || (ddn >= 1024 && dup >= 1024));
}

namespace {
// LDS of the multi-operand kernel: the operands' rows, the value table (EF
// 4) and the impurity terms' tables.
size_t multi_smem(const DevSector& ds, int ef) {
  size_t b = (size_t)kMultiMax * (ds.dn.D + 1) * sizeof(double);
  if (ef == 4) b += (size_t)ds.nvalues * sizeof(double);
  if (ds.kan_ml > 0) {
    const int na = 1 << ds.kan_ml;
    b += (size_t)ds.nkan * (sizeof(double) + sizeof(int64_t)) + (size_t)(na * na + 1) * sizeof(int);
  }
  return b;
}

// The sector's entry format and row length allow the multi-operand kernel.
bool multi_ok(const DevSector& ds) {
  const Launch L = plan(ds);
  if (!L.lds || L.J != 4 || (L.ef != 0 && L.ef != 4)) return false;
  return multi_smem(ds, L.ef) <= lds_budget();
}
}  // namespace

bool hv_batch_routed_lanczos(const DevSector& ds) {
  const int64_t dup = ds.up.D, ddn = ds.dn.D;
  if (hv_batch_routed(dup, ddn)) return true;
  return dup >= 256 && ddn >= 512 && ddn <= kStageMax && multi_ok(ds);
}

int64_t hv_batch_rows_partials(int64_t dup) { return std::min<int64_t>(dup, 1 << 20); }

int hv_batch_rows_multi_fit(const DevSector& ds) {
  return hv_batch_routed_lanczos(ds) && multi_ok(ds) ? kMultiMax : 0;
}

template <int CROSS, bool KAN, int EF>
void launch_multi_t(const HvParams& p, const HvMulti& m, const Launch& L, size_t smem, hipStream_t s) {
  auto k = hv_rows_multi_kernel<CROSS, KAN, EF>;
  if (smem > 64 * 1024)
    EDH_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(k, dim3(L.grid), dim3(L.block), smem, s, p, m);
  EDH_HIP(hipGetLastError());
}

template <int CROSS>
void launch_multi_c(const HvParams& p, const HvMulti& m, const Launch& L, size_t smem, bool kan, hipStream_t s) {
  if (kan) {
    if (L.ef == 4) launch_multi_t<CROSS, true, 4>(p, m, L, smem, s);
    else launch_multi_t<CROSS, true, 0>(p, m, L, smem, s);
  } else {
    if (L.ef == 4) launch_multi_t<CROSS, false, 4>(p, m, L, smem, s);
    else launch_multi_t<CROSS, false, 0>(p, m, L, smem, s);
  }
}

void launch_hv_batch_rows_multi(const DevSector& ds, const HvArgs* a, int count, hipStream_t s) {
  const int fit = hv_batch_rows_multi_fit(ds);
  EDH_CHECK(count >= 2 && count <= fit, EDH_ERR_INTERNAL, "grouped H*v: operand count does not fit one launch");
  const Launch L = plan(ds);
  EDH_CHECK(L.grid == hv_batch_rows_partials(ds.up.D), EDH_ERR_INTERNAL, "grouped H*v: rows grid mismatch");
  HvMulti m;
  std::memset(&m, 0, sizeof(m));
  for (int k = 0; k < count; ++k) {
    EDH_CHECK(a[k].dscale && a[k].tdot && !a[k].lzk && !a[k].ctl, EDH_ERR_INTERNAL,
              "grouped H*v: multi-operand launch takes Lanczos operands with device scales");
    m.v[k] = a[k].v;
    m.w[k] = a[k].w;
    m.wprev[k] = a[k].wprev;
    m.dscale[k] = a[k].dscale;
    m.done[k] = a[k].done;
    m.dot[k] = a[k].tdot;
    m.accumulate[k] = a[k].mode;
  }
  m.nk = count;
  const HvParams p = make_params(ds, a[0]);
  const size_t smem = multi_smem(ds, L.ef);
  EDH_CHECK(smem <= lds_budget(), EDH_ERR_INTERNAL, "grouped H*v: multi-operand rows do not fit LDS");
  const bool kan = ds.kan_ml > 0;
  switch (ds.cross_mode) {
    case CROSS_NONE: launch_multi_c<CROSS_NONE>(p, m, L, smem, kan, s); break;
    case CROSS_POPC: launch_multi_c<CROSS_POPC>(p, m, L, smem, kan, s); break;
    case CROSS_IMP: launch_multi_c<CROSS_IMP>(p, m, L, smem, kan, s); break;
    default: launch_multi_c<CROSS_GENERAL>(p, m, L, smem, kan, s); break;
  }
}

void launch_hv_batch_rows(const DevSector& ds, const HvArgs& a, hipStream_t s) {
  EDH_CHECK(hv_batch_routed_lanczos(ds), EDH_ERR_INTERNAL,
            "grouped H*v: operand not routed to the rows kernel");
  EDH_CHECK(plan(ds).grid == hv_batch_rows_partials(ds.up.D), EDH_ERR_INTERNAL, "grouped H*v: rows grid mismatch");
  EDH_CHECK(!a.lzk, EDH_ERR_INTERNAL, "grouped H*v: rows-kernel operands take resolved buffers");
  HvArgs b = a;
  b.dot_partial = a.tdot;  // one share of <v, w> per work-group (row slot)
  launch_hv(ds, b, s, HV_SINGLE);
}

}  // namespace edh
