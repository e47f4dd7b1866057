// Batched Krylov engine: the batching and serving forms of the solver.
//
// Many independent Krylov processes -- eigensolver runs (the thick-restart
// Lanczos that replaces ARPACK in Storage::diag, Storage.h:43-146) and
// Green's-function Lanczos runs (Lanczos::lanczos, Lanczos.h:36-77) -- on
// sectors of different models and dimensions advance together.  A sweep
// advances every active process by one Krylov step and each phase of a
// sweep is ONE launch over all of them: the grouped H*v over the
// concatenated up-rows of every operand, then grouped projections, updates,
// norms and scalings over fixed chunks of every vector.  Processes join and
// leave between sweeps (continuous batching).
//
// Determinism: a process's arithmetic depends only on its own operands.  Its
// H*v elements do not depend on the grouping, its reductions use its own
// chunks (kChunk elements each) summed in a fixed order, and its host-side
// logic (projected matrix, restarts, breakdown test) sees only its own
// numbers.  A result is therefore bit-identical alone and in any mix.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

#include "batch.h"

namespace edh {

namespace {

constexpr int kChunk = 16384;  // elements per reduction chunk
constexpr int kOpBlock = 256;  // threads per chunk kernel block
constexpr int kHGroup = 8;     // projections accumulated per pass over a chunk

// One grouped vector operation on one process: project w on the cnt
// vectors X (stride ldx), subtract, take norms, scale.
struct BOp {
  const double* X;
  int64_t ldx;
  int cnt;
  double* w;
  int64_t n;
  int chunk0;   // first global chunk of this process
  int nch;      // chunks of this process
  double* h;    // device: projection coefficients (kMaxBasis)
  double* hsum; // device: accumulated coefficients (projected matrix column) or null
  double* nrm;  // device: norm output
  double scale; // bscale: factor when nrm is null
};

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

// partial[c * kMaxBasis + i] = sum over chunk c of X_i[e] * w[e], i < cnt
__global__ __launch_bounds__(kOpBlock) void bdots_kernel(const BOp* __restrict__ ops, const int* __restrict__ cjob,
                                                         double* __restrict__ partial) {
  __shared__ double s_red[8];
  const int c = blockIdx.x;
  const BOp& o = ops[cjob[c]];
  const int lc = c - o.chunk0;
  const int64_t e0 = (int64_t)lc * kChunk;
  const int64_t e1 = min(o.n, e0 + kChunk);
  for (int i0 = 0; i0 < o.cnt; i0 += kHGroup) {
    double acc[kHGroup];
#pragma unroll
    for (int g = 0; g < kHGroup; ++g) acc[g] = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kOpBlock) {
      const double we = o.w[e];
#pragma unroll
      for (int g = 0; g < kHGroup; ++g)
        if (i0 + g < o.cnt) acc[g] = fma(o.X[(int64_t)(i0 + g) * o.ldx + e], we, acc[g]);
    }
#pragma unroll
    for (int g = 0; g < kHGroup; ++g) {
      if (i0 + g >= o.cnt) break;
      const double t = block_sum256(acc[g], s_red);
      if (threadIdx.x == 0) partial[(size_t)c * kMaxBasis + i0 + g] = t;
    }
  }
}

// One block per process: h[i] = sum of its chunks' partials in chunk order;
// hsum[i] += h[i] when given.  mode 1: nrm = sqrt(sum of partial slot 0).
__global__ __launch_bounds__(64) void breduce_kernel(const BOp* __restrict__ ops, const double* __restrict__ partial,
                                                     int mode) {
  const BOp& o = ops[blockIdx.x];
  if (mode == 0) {
    for (int i = threadIdx.x; i < o.cnt; i += blockDim.x) {
      double t = 0.0;
      for (int lc = 0; lc < o.nch; ++lc) t += partial[(size_t)(o.chunk0 + lc) * kMaxBasis + i];
      o.h[i] = t;
      if (o.hsum) o.hsum[i] += t;
    }
  } else if (threadIdx.x == 0) {
    double t = 0.0;
    for (int lc = 0; lc < o.nch; ++lc) t += partial[(size_t)(o.chunk0 + lc) * kMaxBasis];
    *o.nrm = sqrt(t);
  }
}

// w[e] -= sum_i h[i] X_i[e]; partial[c * kMaxBasis] = sum of the new w[e]^2
__global__ __launch_bounds__(kOpBlock) void bupdate_kernel(const BOp* __restrict__ ops, const int* __restrict__ cjob,
                                                           double* __restrict__ partial) {
  __shared__ double s_red[8];
  __shared__ double s_h[kMaxBasis];
  const int c = blockIdx.x;
  const BOp& o = ops[cjob[c]];
  for (int i = threadIdx.x; i < o.cnt; i += blockDim.x) s_h[i] = o.h[i];
  __syncthreads();
  const int lc = c - o.chunk0;
  const int64_t e0 = (int64_t)lc * kChunk;
  const int64_t e1 = min(o.n, e0 + kChunk);
  double nsq = 0.0;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kOpBlock) {
    double x = o.w[e];
    for (int i = 0; i < o.cnt; ++i) x = fma(-s_h[i], o.X[(int64_t)i * o.ldx + e], x);
    o.w[e] = x;
    nsq = fma(x, x, nsq);
  }
  const double t = block_sum256(nsq, s_red);
  if (threadIdx.x == 0) partial[(size_t)c * kMaxBasis] = t;
}

// w[e] *= 1 / nrm (or *= scale when nrm is null)
__global__ __launch_bounds__(kOpBlock) void bscale_kernel(const BOp* __restrict__ ops, const int* __restrict__ cjob) {
  const int c = blockIdx.x;
  const BOp& o = ops[cjob[c]];
  const double f = o.nrm ? 1.0 / *o.nrm : o.scale;
  const int lc = c - o.chunk0;
  const int64_t e0 = (int64_t)lc * kChunk;
  const int64_t e1 = min(o.n, e0 + kChunk);
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kOpBlock) o.w[e] *= f;
}

__global__ void bfill_kernel(double* p, int64_t n, double x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = x;
}

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int nchunks(int64_t n) { return (int)((n + kChunk - 1) / kChunk); }

}  // namespace

Batch::~Batch() {
  if (device >= 0) (void)hipSetDevice(device);
  for (auto& j : jobs)
    if (j) release_job(*j);
  for (auto& s : secs)
    if (s) s->dev.release();
  for (DevBuf* b : {&d_hv, &d_ops, &d_cjob, &d_partial, &d_scal, &gtab}) b->release();
  if (h_pin) (void)hipHostFree(h_pin);
  if (stream) (void)hipStreamDestroy(stream);
}

double* Batch::dalloc(BJob& j, size_t count) {
  void* p = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(double);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) throw Error(EDH_ERR_NOMEM, std::string("batch: hipMalloc failed: ") + hipGetErrorString(e));
  j.allocs.push_back(p);
  j.bytes += bytes;
  bytes_in_use += bytes;
  return (double*)p;
}

void Batch::release_job(BJob& j) {
  for (void* p : j.allocs) (void)hipFree(p);
  j.allocs.clear();
  bytes_in_use -= j.bytes;
  j.bytes = 0;
}

int Batch::add_sector(const Model& m, int nup, int ndn) {
  auto s = std::make_unique<BSector>();
  s->model.reset(new Model(m));
  build_sector(*s->model, nup, ndn, s->tabs, false);
  upload_sector(s->tabs, m.Ns, s->dev, stream);
  s->bytes = s->tabs.table_bytes;
  bytes_in_use += s->bytes;
  secs.push_back(std::move(s));
  return (int)secs.size() - 1;
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
  bytes_in_use -= s.bytes;
  s.dev.release();
  secs[id].reset();
}

// Dense solution of a sector whose Krylov space is the whole sector (the
// same rule as the single-sector eigensolver: ncv >= n or n <= 64).
void Batch::finish_dense(BJob& j) {
  BSector& s = sector(j.sec);
  const int64_t n = j.n;
  std::vector<double> w;
  std::vector<double> H(n * n, 0.0);
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
  auto jp = std::make_unique<BJob>();
  BJob& j = *jp;
  j.kind = JOB_EIG;
  j.sec = sec;
  j.n = s.tabs.N;
  j.tol = tol;
  j.maxitr = maxitr;
  EDH_CHECK(nev_req >= 1 && ncv_req >= 2, EDH_ERR_ARG, "batch: bad nev/ncv");
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
  j.dsc = dalloc(j, (size_t)(j.m + 1) * (j.m + 1) + 2 * kMaxBasis);  // T columns, norms, h scratch
  j.T.assign((size_t)j.m * j.m, 0.0);
  // random start vector, uniform in [-1, 1) (the single-sector eigensolver's generator)
  std::vector<double> x(n);
  uint64_t st = seed ? seed : 0x243F6A8885A308D3ull;
  for (int64_t i = 0; i < n; ++i) x[i] = (double)(splitmix64(st) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  EDH_HIP(hipMemcpyAsync(j.V, x.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, stream));
  EDH_HIP(hipStreamSynchronize(stream));
  // normalise on the device with the process's own chunked reduction
  j.k = 0;
  j.col = 0;
  j.pending_start_norm = true;
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
  for (int q = 0; q < 3; ++q) j.lz[q] = dalloc(j, (size_t)j.n);
  j.dsc = dalloc(j, 2 * kMaxBasis);
  EDH_HIP(hipMemcpyAsync(j.lz[0], v0_dev, (size_t)j.n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  j.step = 1;
  j.state = JOB_ACTIVE;
  return id;
}

// Green's-function start vector c(+)_orb psi in sector dst, its squared
// norm, and (when the norm exceeds 1e-10) a Lanczos run from the normalised
// vector (GreensFunction.h:395-416,437-458).  Returns the job id, or -1
// when the vector vanishes.
int Batch::submit_gf(int src_sec, const double* psi, int orb, bool annihilate, int dst_sec, int nl, double* norm) {
  BSector& b = sector(dst_sec);
  EDH_CHECK(nl >= 1, EDH_ERR_ARG, "batch: nl must be >= 1");
  *norm = 0.0;
  if (b.tabs.N == 0) return -1;
  auto jp = std::make_unique<BJob>();
  BJob& j = *jp;
  j.kind = JOB_LZ;
  j.sec = dst_sec;
  j.n = b.tabs.N;
  j.nl = nl;
  j.alpha.assign(nl, 0.0);
  j.beta.assign(nl + 1, 0.0);
  for (int q = 0; q < 3; ++q) j.lz[q] = dalloc(j, (size_t)j.n);
  j.dsc = dalloc(j, 2 * kMaxBasis);
  adag(src_sec, psi, orb, annihilate, dst_sec, j.lz[0]);
  std::vector<BOpHost> ops{{nullptr, 0, 0, j.lz[0], j.n, j.dsc, nullptr, j.dsc + kMaxBasis, 1.0}};
  run_ops(ops, PH_NORM);
  double nrm = 0.0;
  EDH_HIP(hipMemcpyAsync(&nrm, j.dsc + kMaxBasis, sizeof(double), hipMemcpyDeviceToHost, stream));
  EDH_HIP(hipStreamSynchronize(stream));
  *norm = nrm * nrm;
  if (!(std::fabs(*norm) > 1e-10)) {
    release_job(j);
    return -1;
  }
  run_ops(ops, PH_SCALE);
  j.step = 1;
  j.state = JOB_ACTIVE;
  jobs.push_back(std::move(jp));
  return (int)jobs.size() - 1;
}

void Batch::free_job(int id) {
  BJob& j = job(id);
  EDH_CHECK(j.state != JOB_ACTIVE, EDH_ERR_STATE, "batch: job still active");
  release_job(j);
  jobs[id].reset();
}

// ---------------------------------------------------------------------------
// Descriptor upload and the grouped launches of one phase.
// ---------------------------------------------------------------------------
void Batch::run_ops(std::vector<BOpHost>& ops, int phase) {
  if (ops.empty()) return;
  std::vector<BOp> d(ops.size());
  std::vector<int> cj;
  int c0 = 0;
  for (size_t q = 0; q < ops.size(); ++q) {
    const BOpHost& h = ops[q];
    BOp& o = d[q];
    o.X = h.X;
    o.ldx = h.ldx;
    o.cnt = h.cnt;
    o.w = h.w;
    o.n = h.n;
    o.chunk0 = c0;
    o.nch = nchunks(h.n);
    o.h = h.h;
    o.hsum = h.hsum;
    o.nrm = h.nrm;
    o.scale = h.scale;
    for (int c = 0; c < o.nch; ++c) cj.push_back((int)q);
    c0 += o.nch;
  }
  const int nch = c0;
  d_ops.ensure(d.size() * sizeof(BOp));
  d_cjob.ensure(cj.size() * sizeof(int));
  d_partial.ensure((size_t)std::max(nch, 1) * kMaxBasis * sizeof(double));
  EDH_HIP(hipMemcpyAsync(d_ops.p, d.data(), d.size() * sizeof(BOp), hipMemcpyHostToDevice, stream));
  EDH_HIP(hipMemcpyAsync(d_cjob.p, cj.data(), cj.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  const BOp* dops = d_ops.as<BOp>();
  const int* dcj = d_cjob.as<int>();
  double* part = d_partial.as<double>();
  const int nops = (int)ops.size();
  if (phase == PH_PROJECT) {
    // h = X^T w (into hsum too); w -= X h; nrm = |w|
    hipLaunchKernelGGL(bdots_kernel, dim3(nch), dim3(kOpBlock), 0, stream, dops, dcj, part);
    hipLaunchKernelGGL(breduce_kernel, dim3(nops), dim3(64), 0, stream, dops, part, 0);
    hipLaunchKernelGGL(bupdate_kernel, dim3(nch), dim3(kOpBlock), 0, stream, dops, dcj, part);
    hipLaunchKernelGGL(breduce_kernel, dim3(nops), dim3(64), 0, stream, dops, part, 1);
  } else if (phase == PH_NORM) {
    // nrm = |w| (cnt = 0: the update only accumulates the squares)
    hipLaunchKernelGGL(bupdate_kernel, dim3(nch), dim3(kOpBlock), 0, stream, dops, dcj, part);
    hipLaunchKernelGGL(breduce_kernel, dim3(nops), dim3(64), 0, stream, dops, part, 1);
  } else {
    hipLaunchKernelGGL(bscale_kernel, dim3(nch), dim3(kOpBlock), 0, stream, dops, dcj);
  }
  EDH_HIP(hipGetLastError());
}

void Batch::run_hv(std::vector<std::pair<int, HvArgs>>& hv) {
  if (hv.empty()) return;
  const size_t jb = hv_batch_job_bytes();
  h_hv.resize(hv.size() * jb);
  int64_t row0 = 0, maxd = 0;
  for (size_t q = 0; q < hv.size(); ++q) {
    const DevSector& ds = sector(hv[q].first).dev;
    hv_batch_fill(ds, hv[q].second, row0, h_hv.data() + q * jb);
    row0 += ds.up.D;
    maxd = std::max<int64_t>(maxd, ds.dn.D);
  }
  d_hv.ensure(h_hv.size());
  EDH_HIP(hipMemcpyAsync(d_hv.p, h_hv.data(), h_hv.size(), hipMemcpyHostToDevice, stream));
  const int grid = (int)std::min<int64_t>(row0, 1 << 20);
  launch_hv_batch(d_hv.p, (int)hv.size(), row0, maxd, hv_block, grid, stream);
}

// One sweep: every active process advances one Krylov step.
int Batch::step() {
  std::vector<BJob*> eig_start, eig, lz;
  for (auto& jp : jobs) {
    if (!jp || jp->state != JOB_ACTIVE) continue;
    BJob& j = *jp;
    if (j.kind == JOB_EIG) (j.pending_start_norm ? eig_start : eig).push_back(&j);
    else lz.push_back(&j);
  }
  if (eig_start.empty() && eig.empty() && lz.empty()) return 0;
  std::vector<BOpHost> ops;
  // (0) newly submitted eigensolver runs: normalise the start vector
  if (!eig_start.empty()) {
    ops.clear();
    for (BJob* j : eig_start) ops.push_back({j->V, j->n, 0, j->V, j->n, j->dsc, nullptr, j->dsc + kMaxBasis, 1.0});
    run_ops(ops, PH_NORM);
    run_ops(ops, PH_SCALE);
    for (BJob* j : eig_start) {
      j->pending_start_norm = false;
      eig.push_back(j);
      fill_dev(j->dsc + 2 * kMaxBasis, (size_t)(j->m + 1) * (j->m + 1));
    }
  }
  // (1) Lanczos runs past their first step: v_k = w / beta_{k-1}, in place,
  // then rotate (prev, cur, w) <- (cur, w, prev)
  ops.clear();
  for (BJob* j : lz)
    if (j->step > 1) ops.push_back({nullptr, 0, 0, j->lz[2], j->n, nullptr, nullptr, nullptr, 1.0 / j->last_beta});
  run_ops(ops, PH_SCALE);
  for (BJob* j : lz)
    if (j->step > 1) {
      double* t = j->lz[1];
      j->lz[1] = j->lz[0];
      j->lz[0] = j->lz[2];
      j->lz[2] = t;
    }
  // (2) grouped H*v: eigensolver w = V_{col+1} = H V_col ; Lanczos
  // w = H v_k - beta_{k-1} v_{k-1} (Lanczos.h:47-55)
  std::vector<std::pair<int, HvArgs>> hv;
  for (BJob* j : eig) {
    HvArgs a{};
    a.v = j->V + (size_t)j->col * j->n;
    a.w = j->V + (size_t)(j->col + 1) * j->n;
    a.wprev = a.w;
    a.vscale = 1.0;
    a.wscale = 0.0;
    a.mode = 0;
    hv.push_back({j->sec, a});
  }
  for (BJob* j : lz) {
    HvArgs a{};
    a.v = j->lz[0];
    a.w = j->lz[2];
    a.wprev = j->lz[1];
    a.vscale = 1.0;
    a.wscale = j->step > 1 ? -j->last_beta : 0.0;
    a.mode = j->step > 1 ? 1 : 0;
    hv.push_back({j->sec, a});
  }
  run_hv(hv);
  // (3) first projection pass: eigensolver against V_0..V_col (into the
  // projected-matrix column); Lanczos alpha = <v, w>, w -= alpha v, beta = |w|
  ops.clear();
  for (BJob* j : eig) {
    double* Tcol = j->dsc + 2 * kMaxBasis + (size_t)j->col * (j->m + 1);
    ops.push_back({j->V, j->n, j->col + 1, j->V + (size_t)(j->col + 1) * j->n, j->n, j->dsc, Tcol,
                   j->dsc + kMaxBasis, 0.0});
  }
  for (BJob* j : lz) ops.push_back({j->lz[0], j->n, 1, j->lz[2], j->n, j->dsc, nullptr, j->dsc + kMaxBasis, 0.0});
  run_ops(ops, PH_PROJECT);
  // (4) second pass (eigensolver only) and normalisation of the new basis vector
  if (!eig.empty()) {
    ops.resize(eig.size());
    run_ops(ops, PH_PROJECT);
    run_ops(ops, PH_SCALE);
  }
  // (5) scalars back to the host: one copy per process into pinned memory
  const size_t per = 2 + kMaxBasis;
  const size_t need = (eig.size() + lz.size()) * per * sizeof(double);
  if (need > h_pin_bytes) {
    if (h_pin) EDH_HIP(hipHostFree(h_pin));
    h_pin = nullptr;
    EDH_HIP(hipHostMalloc(&h_pin, need));
    h_pin_bytes = need;
  }
  double* hp = (double*)h_pin;
  size_t slot = 0;
  for (BJob* j : lz) {
    EDH_HIP(hipMemcpyAsync(hp + slot * per, j->dsc, sizeof(double), hipMemcpyDeviceToHost, stream));
    EDH_HIP(hipMemcpyAsync(hp + slot * per + 1, j->dsc + kMaxBasis, sizeof(double), hipMemcpyDeviceToHost, stream));
    ++slot;
  }
  std::vector<BJob*> cycle_end;
  for (BJob* j : eig)
    if (j->col + 1 == j->m) cycle_end.push_back(j);
  for (BJob* j : cycle_end) {
    j->Tc.resize((size_t)j->m * (j->m + 1));
    j->nr_last = 0.0;
    EDH_HIP(hipMemcpyAsync(j->Tc.data(), j->dsc + 2 * kMaxBasis, j->Tc.size() * sizeof(double),
                           hipMemcpyDeviceToHost, stream));
    EDH_HIP(hipMemcpyAsync(&j->nr_last, j->dsc + kMaxBasis, sizeof(double), hipMemcpyDeviceToHost, stream));
  }
  EDH_HIP(hipStreamSynchronize(stream));
  ++sweeps;
  // (6) host bookkeeping per process
  slot = 0;
  for (BJob* j : lz) {
    const double al = hp[slot * per], be = hp[slot * per + 1];
    ++slot;
    const int k = j->step;
    j->alpha[k - 1] = al;
    if (k != j->nl) j->beta[k] = be;  // Lanczos.h:64
    j->nlanc = k;
    if (std::fabs(be) < 1e-10 || k == j->nl) {  // Lanczos.h:65
      j->state = JOB_DONE;
      release_job(*j);  // the coefficients are on the host
      j->lz[0] = j->lz[1] = j->lz[2] = nullptr;
      j->dsc = nullptr;
      continue;
    }
    j->last_beta = be;
    j->step = k + 1;
  }
  for (BJob* j : eig) {
    ++j->n_av;
    if (j->col + 1 < j->m) {
      ++j->col;
      continue;
    }
    end_cycle(*j);
  }
  int active = 0;
  for (auto& jp : jobs)
    if (jp && jp->state == JOB_ACTIVE) ++active;
  return active;
}

void Batch::fill_dev(double* p, size_t n) {
  if (!n) return;
  hipLaunchKernelGGL(bfill_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 1024)), dim3(256), 0, stream, p,
                     (int64_t)n, 0.0);
  EDH_HIP(hipGetLastError());
}

// End of a restart cycle of one eigensolver run: Ritz values of the
// projected matrix, convergence test, then a thick restart or the result
// (the single-sector eigensolver's rules).
void Batch::end_cycle(BJob& j) {
  const int m = j.m, nev = j.nev;
  const int64_t n = j.n;
  for (int c = j.k; c < m; ++c)
    for (int i = 0; i <= c; ++i) {
      j.T[(size_t)i * m + c] = j.Tc[(size_t)c * (m + 1) + i];
      j.T[(size_t)c * m + i] = j.Tc[(size_t)c * (m + 1) + i];
    }
  const double bres = j.nr_last;
  std::vector<double> A = j.T, theta;
  sym_eig(m, A, theta);
  const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
  int nconv_now = 0;
  for (int i = 0; i < nev; ++i) {
    const double bound = std::fabs(bres * A[(size_t)(m - 1) * m + i]);
    if (bound <= j.tol * std::max(std::fabs(theta[i]), eps23)) ++nconv_now;
  }
  const bool conv = nconv_now == nev;
  ++j.iter;
  if (conv || j.iter >= j.maxitr) {
    std::vector<double> Cm((size_t)m * nev);
    for (int c = 0; c < nev; ++c)
      for (int i = 0; i < m; ++i) Cm[(size_t)c * m + i] = A[(size_t)i * m + c];
    double* Cd = dalloc(j, Cm.size());
    EDH_HIP(hipMemcpyAsync(Cd, Cm.data(), Cm.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    j.evecs = dalloc(j, (size_t)nev * n);
    dev_combine(j.V, n, m, Cd, nev, j.evecs, n, n, stream);
    EDH_HIP(hipStreamSynchronize(stream));
    j.evals.assign(theta.begin(), theta.begin() + nev);
    j.nconv = conv ? nev : 0;
    j.state = conv ? JOB_DONE : JOB_FAILED;
    // the Krylov bases are no longer needed
    for (double* b : {j.V, j.Valt}) {
      auto it = std::find(j.allocs.begin(), j.allocs.end(), (void*)b);
      if (it != j.allocs.end()) {
        (void)hipFree(*it);
        j.allocs.erase(it);
        const size_t bb = (size_t)(m + 1) * n * sizeof(double);
        j.bytes -= bb;
        bytes_in_use -= bb;
      }
    }
    j.V = j.Valt = nullptr;
    return;
  }
  const int kk = restart_size(nev, m, nconv_now);
  std::vector<double> Cm((size_t)m * kk);
  for (int c = 0; c < kk; ++c)
    for (int i = 0; i < m; ++i) Cm[(size_t)c * m + i] = A[(size_t)i * m + c];
  if (!j.Cbuf) j.Cbuf = dalloc(j, (size_t)kMaxBasis * kMaxBasis);
  EDH_HIP(hipMemcpyAsync(j.Cbuf, Cm.data(), Cm.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  dev_combine(j.V, n, m, j.Cbuf, kk, j.Valt, n, n, stream);
  EDH_HIP(hipMemcpyAsync(j.Valt + (size_t)kk * n, j.V + (size_t)m * n, (size_t)n * sizeof(double),
                         hipMemcpyDeviceToDevice, stream));
  std::swap(j.V, j.Valt);
  std::fill(j.T.begin(), j.T.end(), 0.0);
  for (int i = 0; i < kk; ++i) {
    j.T[(size_t)i * m + i] = theta[i];
    const double si = bres * A[(size_t)(m - 1) * m + i];
    j.T[(size_t)i * m + kk] = si;
    j.T[(size_t)kk * m + i] = si;
  }
  fill_dev(j.dsc + 2 * kMaxBasis, (size_t)(m + 1) * (m + 1));
  j.k = kk;
  j.col = kk;
}

// Gather map of c / c^+ between two sectors of the batch (the single-sector
// a_adag maps) applied on the device.
void Batch::adag(int src, const double* in, int orb, bool annihilate, int dst, double* out) {
  BSector& a = sector(src);
  BSector& b = sector(dst);
  const int Ns = a.model->Ns;
  EDH_CHECK(orb >= 0 && orb < 2 * Ns, EDH_ERR_ARG, "batch: orbital out of range");
  int nup2, ndn2;
  std::vector<int32_t> mu, md;
  std::vector<int8_t> su, sd;
  adag_gather_maps(Ns, a.tabs.nup, a.tabs.ndn, orb, annihilate, nup2, ndn2, mu, md, su, sd);
  EDH_CHECK(nup2 == b.tabs.nup && ndn2 == b.tabs.ndn, EDH_ERR_ARG, "batch: target sector does not match the operator");
  const size_t bytes = (mu.size() + md.size()) * 4 + su.size() + sd.size() + 64;
  gtab.ensure(bytes);
  char* g = gtab.as<char>();
  int32_t* dmu = (int32_t*)g;
  int32_t* dmd = dmu + mu.size();
  int8_t* dsu = (int8_t*)(dmd + md.size());
  int8_t* dsd = dsu + su.size();
  EDH_HIP(hipMemcpyAsync(dmu, mu.data(), mu.size() * 4, hipMemcpyHostToDevice, stream));
  EDH_HIP(hipMemcpyAsync(dmd, md.data(), md.size() * 4, hipMemcpyHostToDevice, stream));
  EDH_HIP(hipMemcpyAsync(dsu, su.data(), su.size(), hipMemcpyHostToDevice, stream));
  EDH_HIP(hipMemcpyAsync(dsd, sd.data(), sd.size(), hipMemcpyHostToDevice, stream));
  dev_separable_gather(in, a.tabs.dn.D, dmu, dmd, dsu, dsd, out, b.tabs.up.D, b.tabs.dn.D, stream);
  EDH_HIP(hipStreamSynchronize(stream));
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
  hipEvent_t e0, e1;
  EDH_HIP(hipEventCreate(&e0));
  EDH_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int choice = hv_block;
  for (int b : {128, 256, 512}) {
    hv_block = b;
    run_hv(hv);  // warm-up
    EDH_HIP(hipEventRecord(e0, stream));
    for (int r = 0; r < 3; ++r) run_hv(hv);
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
  (void)hipFree(scratch);
  hv_block = choice;
  return choice;
}

}  // namespace edh
