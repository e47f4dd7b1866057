// Static observables of one eigenvector on the device.
//
// Replaces the per-eigenvector loop of StaticObservables::
// calculate_static_observables_eigenvector (StaticObservables.h:356-419),
// which walks every basis state on the host.  With w = v[u][d]^2 and the
// occupation bits of the up state (row u) and down state (column d), every
// observable is an entry of the second-moment matrix of the occupation bits:
//   UU[a][b] = sum_u R_u bu_a bu_b          R_u = sum_d w      (row sums)
//   UD[a][b] = sum_u bu_a D_u[b]            D_u[b] = sum_d w bd_b
//   DD[a][b] = sum_d C_d bd_a bd_b          C_d = sum_u w      (column sums)
// and N_eff = 1 / sum w^2.  The device makes one pass over the rows (R_u,
// D_u, sum w^2 per row) and one over the columns (C_d, by row tiles), then
// forms UU, UD, DD and sum w^2 in two fixed-order reduction stages; the host
// reads those 3 N_s^2 + 1 numbers and assembles the observables with the
// reference's orbital conventions.
#include <algorithm>
#include <cstring>

#include "edh_ctx.h"

namespace edh {

namespace {

constexpr int kObsBlock = 256;
constexpr int kObsMaxSites = 32;
constexpr int kColTile = 128;  // rows per column-sum tile
constexpr int kMomBlocks = 512;  // first-stage blocks of the moment reduction

// One block per row u (grid-stride): out[u * (Ns + 2) + b] = sum_d w [bit b
// of bits_d] for b < Ns, then sum_d w, then sum_d w^2.  MS >= Ns bit slots
// per element (16 for up to 16 sites: half the bit tests of 32).
template <int MS>
__global__ __launch_bounds__(kObsBlock) void obs_rows_kernel(const double* __restrict__ v, int64_t Dup, int64_t Ddn,
                                                             const int32_t* __restrict__ dn_bits, int Ns,
                                                             double* __restrict__ out) {
  constexpr int kObsMaxSites = MS;
  __shared__ double s_part[kObsBlock / 64][kObsMaxSites + 2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t u = blockIdx.x; u < Dup; u += gridDim.x) {
    double acc[kObsMaxSites + 2];
#pragma unroll
    for (int b = 0; b < kObsMaxSites + 2; ++b) acc[b] = 0.0;
    const double* row = v + u * Ddn;
    for (int64_t d = threadIdx.x; d < Ddn; d += kObsBlock) {
      const double x = row[d];
      const double w = x * x;
      const unsigned bits = (unsigned)dn_bits[d];
#pragma unroll
      for (int b = 0; b < kObsMaxSites; ++b)
        if ((bits >> b) & 1u) acc[b] += w;
      acc[kObsMaxSites] += w;
      acc[kObsMaxSites + 1] = fma(w, w, acc[kObsMaxSites + 1]);
    }
    // fixed-order reduction: shuffle tree per wavefront, then the
    // wavefronts in order
#pragma unroll
    for (int b = 0; b < kObsMaxSites + 2; ++b) {
      double x = acc[b];
      for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
      if (lane == 0) s_part[wave][b] = x;
    }
    __syncthreads();
    const int nout = Ns + 2;
    if ((int)threadIdx.x < nout) {
      const int b = (int)threadIdx.x < Ns ? (int)threadIdx.x : kObsMaxSites + ((int)threadIdx.x - Ns);
      double t = 0.0;
      for (int q = 0; q < kObsBlock / 64; ++q) t += s_part[q][b];
      out[u * nout + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// Column sums by row tiles: part[tile * Ddn + d] = sum over the tile's rows
// of v[u][d]^2 (coalesced over d).
__global__ __launch_bounds__(kObsBlock) void obs_cols_kernel(const double* __restrict__ v, int64_t Dup, int64_t Ddn,
                                                             double* __restrict__ part) {
  const int64_t d = blockIdx.x * (int64_t)kObsBlock + threadIdx.x;
  const int64_t tile = blockIdx.y;
  if (d >= Ddn) return;
  const int64_t u0 = tile * kColTile, u1 = std::min<int64_t>(Dup, u0 + kColTile);
  double s0 = 0.0, s1 = 0.0;
  int64_t u = u0;
  for (; u + 1 < u1; u += 2) {
    const double x0 = v[u * Ddn + d], x1 = v[(u + 1) * Ddn + d];
    s0 = fma(x0, x0, s0);
    s1 = fma(x1, x1, s1);
  }
  if (u < u1) {
    const double x = v[u * Ddn + d];
    s0 = fma(x, x, s0);
  }
  part[tile * Ddn + d] = s0 + s1;
}

// C_d = sum over the tiles in order.
__global__ __launch_bounds__(kObsBlock) void obs_cols_sum_kernel(const double* __restrict__ part, int64_t ntile,
                                                                 int64_t Ddn, double* __restrict__ col) {
  const int64_t d = blockIdx.x * (int64_t)kObsBlock + threadIdx.x;
  if (d >= Ddn) return;
  double s = 0.0;
  for (int64_t t = 0; t < ntile; ++t) s += part[t * Ddn + d];
  col[d] = s;
}

// First stage of the moments: block k takes rows [k R, (k + 1) R) and
// columns [k C, (k + 1) C); thread t = a Ns + b (a, b < Ns) sums
// UU[a][b], UD[a][b] over its rows and DD[a][b] over its columns; thread 0
// also sum w^2.  part[k * (3 Ns^2 + 1) + ...] = (UU, UD, DD, sumsq).
__global__ __launch_bounds__(1024) void obs_moments_kernel(const double* __restrict__ rows, const double* __restrict__ col,
                                                           const int32_t* __restrict__ up_bits,
                                                           const int32_t* __restrict__ dn_bits, int64_t Dup,
                                                           int64_t Ddn, int Ns, double* __restrict__ part) {
  const int t = threadIdx.x;
  const int nout = Ns + 2, nm = 3 * Ns * Ns + 1;
  const int k = blockIdx.x;
  const int64_t R = (Dup + gridDim.x - 1) / gridDim.x, C = (Ddn + gridDim.x - 1) / gridDim.x;
  double* o = part + (size_t)k * nm;
  if (t < Ns * Ns) {
    const int a = t / Ns, b = t - a * Ns;
    double uu = 0.0, ud = 0.0, dd = 0.0;
    for (int64_t u = k * R; u < min(Dup, (k + 1) * R); ++u) {
      const unsigned bu = (unsigned)up_bits[u];
      if (!((bu >> a) & 1u)) continue;
      const double* r = rows + u * nout;
      if ((bu >> b) & 1u) uu += r[Ns];
      ud += r[b];
    }
    for (int64_t d = k * C; d < min(Ddn, (k + 1) * C); ++d) {
      const unsigned bd = (unsigned)dn_bits[d];
      if (((bd >> a) & 1u) && ((bd >> b) & 1u)) dd += col[d];
    }
    o[t] = uu;
    o[Ns * Ns + t] = ud;
    o[2 * Ns * Ns + t] = dd;
  }
  if (t == 0) {
    double s = 0.0;
    for (int64_t u = k * R; u < min(Dup, (k + 1) * R); ++u) s += rows[u * nout + Ns + 1];
    o[3 * Ns * Ns] = s;
  }
}

// Second stage, one block per moment i: thread t sums the first-stage
// blocks t, t + 256, ... in order, then a fixed shuffle tree and the four
// wavefronts in order give out[i].
__global__ __launch_bounds__(256) void obs_moments_sum_kernel(const double* __restrict__ part, int nblk, int nm,
                                                              double* __restrict__ out) {
  __shared__ double s_w[4];
  const int i = blockIdx.x;
  double s = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 256) s += part[(size_t)k * nm + i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[i] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

}  // namespace

void static_obs_run(Ctx& c, const double* dv, int norb, double* out) {
  static_obs_sector(c.tabs, c.dev, c.model->Ns, dv, norb, c.obs, c.stream, out);
}

size_t static_obs_scratch(const SectorTables& st, int Ns) {
  const int64_t Dup = st.up.D, Ddn = st.dn.D;
  const int64_t ntile = (Dup + kColTile - 1) / kColTile;
  const size_t nm = 3 * (size_t)Ns * Ns + 1;
  return (size_t)Dup * (Ns + 2) + (size_t)ntile * Ddn + (size_t)Ddn + (size_t)kMomBlocks * nm + nm;
}

size_t static_obs_host(const SectorTables&, int Ns) { return 3 * (size_t)Ns * Ns + 1; }

void static_obs_launch(const SectorTables& st, const DevSector& ds, int Ns, const double* dv, double* scratch,
                       hipStream_t stream, double* host) {
  EDH_CHECK(Ns <= kObsMaxSites, EDH_ERR_LIMIT, "static observables: more than 32 sites per spin");
  const int64_t Dup = st.up.D, Ddn = st.dn.D;
  const int nout = Ns + 2;
  const int64_t ntile = (Dup + kColTile - 1) / kColTile;
  const size_t rows_n = (size_t)Dup * nout, part_n = (size_t)ntile * Ddn;
  double* rows = scratch;
  double* part = rows + rows_n;
  double* col = part + part_n;
  const int nm = 3 * Ns * Ns + 1;
  double* mpart = col + Ddn;
  double* mom = mpart + (size_t)kMomBlocks * nm;
  const int grid_rows = (int)std::min<int64_t>(Dup, 4096);
  if (Ns <= 16)
    hipLaunchKernelGGL(obs_rows_kernel<16>, dim3(grid_rows), dim3(kObsBlock), 0, stream, dv, Dup, Ddn, ds.dn.bits, Ns,
                       rows);
  else
    hipLaunchKernelGGL(obs_rows_kernel<kObsMaxSites>, dim3(grid_rows), dim3(kObsBlock), 0, stream, dv, Dup, Ddn,
                       ds.dn.bits, Ns, rows);
  hipLaunchKernelGGL(obs_cols_kernel, dim3((unsigned)((Ddn + kObsBlock - 1) / kObsBlock), (unsigned)ntile),
                     dim3(kObsBlock), 0, stream, dv, Dup, Ddn, part);
  hipLaunchKernelGGL(obs_cols_sum_kernel, dim3((unsigned)((Ddn + kObsBlock - 1) / kObsBlock)), dim3(kObsBlock), 0,
                     stream, part, ntile, Ddn, col);
  hipLaunchKernelGGL(obs_moments_kernel, dim3(kMomBlocks), dim3((unsigned)std::max(64, (Ns * Ns + 63) / 64 * 64)), 0,
                     stream, rows, col, ds.up.bits, ds.dn.bits, Dup, Ddn, Ns, mpart);
  hipLaunchKernelGGL(obs_moments_sum_kernel, dim3(nm), dim3(256), 0, stream, mpart, kMomBlocks, nm, mom);
  EDH_HIP(hipGetLastError());
  EDH_HIP(hipMemcpyAsync(host, mom, (size_t)nm * sizeof(double), hipMemcpyDeviceToHost, stream));
}

void static_obs_sector(const SectorTables& st, const DevSector& ds, int Ns, const double* dv, int norb,
                       DevBuf& scratch, hipStream_t stream, double* out) {
  EDH_CHECK(norb >= 1 && norb <= Ns, EDH_ERR_ARG, "static observables: interacting orbitals out of range");
  scratch.ensure(static_obs_scratch(st, Ns) * sizeof(double));
  std::vector<double> host(static_obs_host(st, Ns));
  static_obs_launch(st, ds, Ns, dv, scratch.as<double>(), stream, host.data());
  EDH_HIP(hipStreamSynchronize(stream));
  static_obs_reduce(st, Ns, norb, host.data(), out);
}

void static_obs_reduce(const SectorTables& st, int Ns, int norb, const double* host, double* out) {
  EDH_CHECK(norb >= 1 && norb <= Ns, EDH_ERR_ARG, "static observables: interacting orbitals out of range");
  // second moments of the occupation bits (bit b of a per-spin pattern is
  // site Ns - 1 - b, FermionicModel.h:35-68), from the device
  const double* UU = host;
  const double* UD = host + (size_t)Ns * Ns;
  const double* DD = host + 2 * (size_t)Ns * Ns;
  const double sumsq = host[3 * (size_t)Ns * Ns];
  // orbital o < Ns is the up spin of site o, o >= Ns the down spin of site
  // o - Ns; E[x y] for two spin-orbitals from the three blocks
  auto moment = [&](int o1, int o2) -> double {
    const bool up1 = o1 < Ns, up2 = o2 < Ns;
    const int b1 = Ns - 1 - (up1 ? o1 : o1 - Ns), b2 = Ns - 1 - (up2 ? o2 : o2 - Ns);
    if (up1 && up2) return UU[(size_t)b1 * Ns + b2];
    if (!up1 && !up2) return DD[(size_t)b1 * Ns + b2];
    if (up1) return UD[(size_t)b1 * Ns + b2];
    return UD[(size_t)b2 * Ns + b1];
  };
  double* N = out;
  double* Nu = N + norb;
  double* Nd = Nu + norb;
  double* M = Nd + norb;
  double* Docc = M + norb;
  double* MM = Docc + norb;
  double* Neff = MM + (size_t)norb * norb;
  for (int o = 0; o < norb; ++o) {
    const double nu = moment(o, o), nd = moment(o + Ns, o + Ns);
    Nu[o] = nu;
    Nd[o] = nd;
    N[o] = nu + nd;
    M[o] = nu - nd;
    Docc[o] = moment(o, o + Ns);
  }
  // (el_up(o) - el_down(o)) (el_up(o2) - el_down2(o2)), where el_down2 reads
  // orbital o2 + interacting_orbitals (StaticObservables.h:378-381): the
  // down spin of site o2 when every orbital interacts, otherwise the up
  // spin of site o2 + norb (a bath site) while that index is below Ns
  for (int o = 0; o < norb; ++o)
    for (int o2 = 0; o2 < norb; ++o2) {
      const int d2 = o2 + norb;
      MM[(size_t)o * norb + o2] =
          moment(o, o2) - moment(o, d2) - moment(o + Ns, o2) + moment(o + Ns, d2);
    }
  Neff[0] = 1.0 / sumsq;
}

}  // namespace edh
