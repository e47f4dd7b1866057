// C ABI of libedhip (see include/edhip.h).  Every entry catches all C++
// exceptions and turns them into a status code plus a per-thread message.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <limits>

#include "batch.h"
#include "edh_ctx.h"

struct edh_model {
  std::unique_ptr<edh::Model> m;
};
struct edh_ctx {
  edh::Ctx c;
};
struct edh_batch {
  edh::Batch b;
};

namespace {
thread_local std::string g_err;
}

#define API_BEGIN try {
#define API_END                                   \
  return EDH_OK;                                  \
  }                                               \
  catch (const edh::Error& e) {                   \
    g_err = e.what();                             \
    return e.code;                                \
  }                                               \
  catch (const std::bad_alloc&) {                 \
    g_err = "out of host memory";                 \
    return EDH_ERR_NOMEM;                         \
  }                                               \
  catch (const std::exception& e) {               \
    g_err = e.what();                             \
    return EDH_ERR_INTERNAL;                      \
  }                                               \
  catch (...) {                                   \
    g_err = "unknown error";                      \
    return EDH_ERR_INTERNAL;                      \
  }

namespace edh {

Ctx::~Ctx() {
  dev.release();
  stage_a.release();
  stage_b.release();
  partial.release();
  scal.release();
  for (auto& b : lz) b.release();
  for (auto& b : krylov) b.release();
  evecs.release();
  small.release();
  gtab.release();
  obs.release();
  if (stream) {
    (void)hipStreamSynchronize(stream);  // the tables return to the pool in stream order
    (void)hipStreamDestroy(stream);
  }
}

namespace {

void require_ctx(edh_ctx* ctx) { EDH_CHECK(ctx != nullptr, EDH_ERR_ARG, "null context"); }
void use_device(Ctx& c) { EDH_HIP(hipSetDevice(c.device)); }
void require_sector(Ctx& c) { EDH_CHECK(c.has_sector, EDH_ERR_STATE, "no sector set (call edh_set_sector)"); }

// Copies a vector argument to the device when it is a host pointer.
const double* stage_in(Ctx& c, DevBuf& b, const double* p, int64_t n, int mem) {
  if (mem == EDH_MEM_DEVICE) return p;
  EDH_CHECK(mem == EDH_MEM_HOST, EDH_ERR_ARG, "invalid memory kind");
  b.ensure((size_t)n * sizeof(double));
  if (n) EDH_HIP(hipMemcpyAsync(b.p, p, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  return b.as<double>();
}

int popc_parity(int32_t x) { return __builtin_popcount((unsigned)x) & 1; }

// Index/sign maps for c (annihilate) or c^+ on orbital orb of a vector in
// sector (nup, ndn): target index (u', d') reads source (mu[u'], md[d']) times
// su[u'] sd[d'] (-1 index = zero).  The sign counts occupied orbitals with a
// smaller index (FermionicModel.h:45-68); for a down orbital that includes
// every up electron.
void adag_maps(int Ns, int nup, int ndn, int orb, bool annihilate, int& nup2, int& ndn2, std::vector<int32_t>& mu,
               std::vector<int32_t>& md, std::vector<int8_t>& su, std::vector<int8_t>& sd) {
  const int spin = orb >= Ns ? 1 : 0;
  const int site = orb - spin * Ns;
  const int pb = Ns - 1 - site;
  nup2 = nup + (spin == 0 ? (annihilate ? -1 : 1) : 0);
  ndn2 = ndn + (spin == 1 ? (annihilate ? -1 : 1) : 0);
  std::vector<int32_t> tu = spin_basis(Ns, nup2), td = spin_basis(Ns, ndn2);
  const int nsrc = spin == 0 ? nup : ndn;
  std::vector<int32_t> src = spin_basis(Ns, nsrc);
  std::vector<int32_t> rank(1u << Ns, -1);
  for (size_t r = 0; r < src.size(); ++r) rank[src[r]] = (int32_t)r;
  auto build = [&](const std::vector<int32_t>& tgt, std::vector<int32_t>& mp, std::vector<int8_t>& sg, int extra) {
    mp.assign(tgt.size(), -1);
    sg.assign(tgt.size(), 1);
    for (size_t r = 0; r < tgt.size(); ++r) {
      const int32_t b = tgt[r];
      const bool has = (b >> pb) & 1;
      if (annihilate == has) continue;  // c needs the orbital empty afterwards, c^+ occupied
      const int32_t s = b ^ (1 << pb);
      mp[r] = rank[s];
      sg[r] = (int8_t)(((popc_parity(s >> (pb + 1)) + extra) & 1) ? -1 : 1);
    }
  };
  auto identity = [](size_t D, std::vector<int32_t>& mp, std::vector<int8_t>& sg) {
    mp.resize(D);
    sg.assign(D, 1);
    for (size_t r = 0; r < D; ++r) mp[r] = (int32_t)r;
  };
  if (spin == 0) {
    build(tu, mu, su, 0);
    identity(td.size(), md, sd);
  } else {
    identity(tu.size(), mu, su);
    build(td, md, sd, nup & 1);
  }
}

}  // namespace

void adag_gather_maps(int Ns, int nup, int ndn, int orb, bool annihilate, int& nup2, int& ndn2,
                      std::vector<int32_t>& mu, std::vector<int32_t>& md, std::vector<int8_t>& su,
                      std::vector<int8_t>& sd) {
  adag_maps(Ns, nup, ndn, orb, annihilate, nup2, ndn2, mu, md, su, sd);
}

}  // namespace edh

using namespace edh;

extern "C" {

const char* edh_last_error(void) { return g_err.c_str(); }
int edh_version(void) { return 100; }

int edh_model_hubbard(edh_model** out, int Ns, const double* t, const double* U, const double* J, const double* xmu,
                      const double* Hmag) {
  API_BEGIN
  EDH_CHECK(out && t && U && xmu, EDH_ERR_ARG, "null argument");
  edh_model* m = new edh_model();
  m->m.reset(build_hubbard(Ns, t, U, J, xmu, Hmag));
  *out = m;
  API_END
}

int edh_model_siam(edh_model** out, int Ns, int ml, const double* Epsk, const double* Vk, const double* H0,
                   const double* Ut, double mu) {
  API_BEGIN
  EDH_CHECK(out && Vk && H0 && Ut && (Epsk || Ns == ml), EDH_ERR_ARG, "null argument");
  edh_model* m = new edh_model();
  m->m.reset(build_siam(Ns, ml, Epsk, Vk, H0, Ut, mu));
  *out = m;
  API_END
}

int edh_model_from_terms(edh_model** out, const edh_model_desc* d) {
  API_BEGIN
  EDH_CHECK(out && d, EDH_ERR_ARG, "null argument");
  edh_model* m = new edh_model();
  m->m.reset(build_terms(*d));
  *out = m;
  API_END
}

int edh_model_apply(const edh_model* m, long long state, int cap, long long* targets, double* coef, int* count) {
  API_BEGIN
  EDH_CHECK(m && count && cap >= 0 && (cap == 0 || (targets && coef)), EDH_ERR_ARG, "null argument");
  const Model& md = *m->m;
  EDH_CHECK(state >= 0 && (md.Ns >= 31 || state < (1ll << (2 * md.Ns))), EDH_ERR_ARG, "state out of range");
  int c = 0;
  for (const std::vector<Transition>* L : {&md.T, &md.V})
    for (const Transition& tr : *L) {
      if (!md.valid(tr, state)) continue;
      state_t k;
      int sg;
      const double v = md.apply(tr, state, k, sg);
      if (c < cap) {
        targets[c] = k;
        coef[c] = sg * v;
      }
      ++c;
    }
  *count = c;
  API_END
}

int edh_model_free(edh_model* m) {
  delete m;
  return EDH_OK;
}

int edh_model_sites(const edh_model* m, int* Ns, int* ml) {
  API_BEGIN
  EDH_CHECK(m, EDH_ERR_ARG, "null model");
  if (Ns) *Ns = m->m->Ns;
  if (ml) *ml = m->m->ml;
  API_END
}

int edh_model_transitions(const edh_model* m, int which, int* count, int* ints, double* vals) {
  API_BEGIN
  EDH_CHECK(m && count, EDH_ERR_ARG, "null argument");
  const std::vector<Transition>& L = which == 0 ? m->m->T : m->m->V;
  *count = (int)L.size();
  if (ints && vals)
    for (size_t q = 0; q < L.size(); ++q) {
      const Transition& t = L[q];
      int* o = ints + 7 * q;
      o[0] = t.kind;
      o[1] = t.i;
      o[2] = t.j;
      o[3] = t.k;
      o[4] = t.l;
      o[5] = t.s;
      o[6] = t.sp;
      vals[q] = t.value;
    }
  API_END
}

int edh_model_diagonal(const edh_model* m, long long state, double* out) {
  API_BEGIN
  EDH_CHECK(m && out, EDH_ERR_ARG, "null argument");
  *out = m->m->diagonal_reference(state);
  API_END
}

int edh_sector_csr(const edh_model* m, int nup, int ndn, int which, int64_t* nrows, int64_t* nnz, int* rowp, int* col,
                   double* val) {
  API_BEGIN
  EDH_CHECK(m && nrows && nnz, EDH_ERR_ARG, "null argument");
  SectorTables st;
  build_sector(*m->m, nup, ndn, st, false);
  const SpinTables& sp = which == 0 ? st.up : st.dn;
  *nrows = sp.D;
  *nnz = (int64_t)sp.col.size();
  if (rowp) std::copy(sp.rowp.begin(), sp.rowp.end(), rowp);
  if (col) std::copy(sp.col.begin(), sp.col.end(), col);
  if (val) std::copy(sp.val.begin(), sp.val.end(), val);
  API_END
}

int edh_sector_diagonal(const edh_model* m, int nup, int ndn, double* out) {
  API_BEGIN
  EDH_CHECK(m && out, EDH_ERR_ARG, "null argument");
  SectorTables st;
  build_sector(*m->m, nup, ndn, st, false);
  for (int64_t u = 0; u < st.up.D; ++u)
    for (int64_t d = 0; d < st.dn.D; ++d) out[u * st.dn.D + d] = host_diag(*m->m, st, u, d);
  API_END
}

int edh_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
  API_BEGIN
  EDH_CHECK(free_bytes && total_bytes, EDH_ERR_ARG, "null argument");
  int n = 0;
  EDH_HIP(hipGetDeviceCount(&n));
  EDH_CHECK(device >= 0 && device < n, EDH_ERR_ARG, "invalid device index");
  EDH_HIP(hipSetDevice(device));
  EDH_HIP(hipMemGetInfo(free_bytes, total_bytes));
  API_END
}

int edh_create(edh_ctx** out, int device) {
  API_BEGIN
  EDH_CHECK(out, EDH_ERR_ARG, "null argument");
  int n = 0;
  EDH_HIP(hipGetDeviceCount(&n));
  EDH_CHECK(device >= 0 && device < n, EDH_ERR_ARG, "invalid device index");
  std::unique_ptr<edh_ctx> c(new edh_ctx());
  c->c.device = device;
  EDH_HIP(hipSetDevice(device));
  EDH_HIP(hipStreamCreate(&c->c.stream));
  edh::configure_mem_pool(device);
  edh::noalloc_register_device(device);
  *out = c.release();
  API_END
}

int edh_destroy(edh_ctx* ctx) {
  API_BEGIN
  if (ctx) {
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    delete ctx;
  }
  API_END
}

int edh_set_model(edh_ctx* ctx, const edh_model* m) {
  API_BEGIN
  require_ctx(ctx);
  EDH_CHECK(m, EDH_ERR_ARG, "null model");
  Ctx& c = ctx->c;
  c.model.reset(new Model(*m->m));
  c.has_sector = false;
  API_END
}

int edh_set_sector(edh_ctx* ctx, int nup, int ndn, int64_t* dim) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  EDH_CHECK(c.model, EDH_ERR_STATE, "no model set");
  use_device(c);
  c.has_sector = false;
  build_sector(*c.model, nup, ndn, c.tabs, false);
  upload_sector(c.tabs, c.model->Ns, c.dev, c.stream);
  c.has_sector = true;
  if (dim) *dim = c.tabs.N;
  API_END
}

int edh_constant_shift(edh_ctx* ctx, double shift) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  require_sector(c);
  use_device(c);
  for (double& x : c.tabs.up.ediag) x += shift;
  EDH_HIP(hipMemcpyAsync(c.dev.up.ediag, c.tabs.up.ediag.data(), c.tabs.up.ediag.size() * sizeof(double),
                         hipMemcpyHostToDevice, c.stream));
  EDH_HIP(hipStreamSynchronize(c.stream));
  API_END
}

int edh_sector_info(edh_ctx* ctx, int64_t* dup, int64_t* ddn, int64_t* tbytes) {
  API_BEGIN
  require_ctx(ctx);
  require_sector(ctx->c);
  if (dup) *dup = ctx->c.tabs.up.D;
  if (ddn) *ddn = ctx->c.tabs.dn.D;
  if (tbytes) *tbytes = ctx->c.tabs.table_bytes;
  API_END
}

int edh_alloc(edh_ctx* ctx, size_t bytes, void** dptr) {
  API_BEGIN
  require_ctx(ctx);
  EDH_CHECK(dptr, EDH_ERR_ARG, "null argument");
  use_device(ctx->c);
  hipError_t e = hipMalloc(dptr, bytes);
  EDH_CHECK(e == hipSuccess, EDH_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  API_END
}

int edh_free(edh_ctx* ctx, void* dptr) {
  API_BEGIN
  require_ctx(ctx);
  use_device(ctx->c);
  if (dptr) EDH_HIP(hipFree(dptr));
  API_END
}

int edh_memcpy(edh_ctx* ctx, void* dst, const void* src, size_t bytes, int kind) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  use_device(c);
  hipMemcpyKind k = kind == EDH_COPY_H2D ? hipMemcpyHostToDevice
                                         : (kind == EDH_COPY_D2H ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  EDH_HIP(hipMemcpyAsync(dst, src, bytes, k, c.stream));
  EDH_HIP(hipStreamSynchronize(c.stream));
  API_END
}

int edh_synchronize(edh_ctx* ctx) {
  API_BEGIN
  require_ctx(ctx);
  use_device(ctx->c);
  EDH_HIP(hipStreamSynchronize(ctx->c.stream));
  API_END
}

int edh_static_observables(edh_ctx* ctx, const double* v, int64_t n, int mem, int norb, double* out) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  require_sector(c);
  use_device(c);
  EDH_CHECK(n == c.tabs.N, EDH_ERR_ARG, "vector size does not match the sector dimension");
  EDH_CHECK(v && out, EDH_ERR_ARG, "null argument");
  const double* dv = stage_in(c, c.stage_a, v, n, mem);
  static_obs_run(c, dv, norb, out);
  API_END
}

int edh_av(edh_ctx* ctx, const double* v, double* w, int64_t n, int clear, int mem) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  require_sector(c);
  use_device(c);
  EDH_CHECK(n == c.tabs.N, EDH_ERR_ARG, "vector size does not match the sector dimension");
  EDH_CHECK(v && w, EDH_ERR_ARG, "null vector");
  const double* dv = stage_in(c, c.stage_a, v, n, mem);
  double* dw = w;
  if (mem == EDH_MEM_HOST) {
    c.stage_b.ensure((size_t)n * sizeof(double));
    dw = c.stage_b.as<double>();
    if (!clear) EDH_HIP(hipMemcpyAsync(dw, w, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  }
  EDH_CHECK(dv != dw, EDH_ERR_ARG, "v and w must not alias");
  HvArgs a{};
  a.v = dv;
  a.w = dw;
  a.wprev = dw;
  a.vscale = 1.0;
  a.wscale = 1.0;
  a.mode = clear ? 0 : 1;
  launch_hv(c.dev, a, c.stream);
  if (mem == EDH_MEM_HOST) {
    EDH_HIP(hipMemcpyAsync(w, dw, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    EDH_HIP(hipStreamSynchronize(c.stream));
  }
  API_END
}

int edh_get_hv_variant(edh_ctx* ctx, int* variant) {
  API_BEGIN
  require_ctx(ctx);
  require_sector(ctx->c);
  EDH_CHECK(variant, EDH_ERR_ARG, "null argument");
  *variant = ctx->c.dev.variant;
  API_END
}

int edh_set_hv_variant(edh_ctx* ctx, int variant) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  require_sector(c);
  use_device(c);
  EDH_CHECK(variant >= -1 && variant <= 1, EDH_ERR_ARG, "variant must be -1, 0 or 1");
  c.dev.variant = variant < 0 ? hv_autotune(c.dev, c.stream) : variant;
  API_END
}

int edh_av_time(edh_ctx* ctx, const double* v, double* w, int clear, int warmup, int reps, double* avg_ms) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  require_sector(c);
  use_device(c);
  EDH_CHECK(v && w && avg_ms && reps > 0, EDH_ERR_ARG, "bad argument");
  HvArgs a{};
  a.v = v;
  a.w = w;
  a.wprev = w;
  a.vscale = 1.0;
  a.wscale = 1.0;
  a.mode = clear ? 0 : 1;
  for (int i = 0; i < warmup; ++i) launch_hv(c.dev, a, c.stream);
  hipEvent_t e0, e1;
  EDH_HIP(hipEventCreate(&e0));
  EDH_HIP(hipEventCreate(&e1));
  EDH_HIP(hipEventRecord(e0, c.stream));
  for (int i = 0; i < reps; ++i) launch_hv(c.dev, a, c.stream);
  EDH_HIP(hipEventRecord(e1, c.stream));
  EDH_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  EDH_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / reps;
  API_END
}

int edh_vv(edh_ctx* ctx, const double* a, const double* b, int64_t n, double* out, int mem) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  use_device(c);
  EDH_CHECK(out, EDH_ERR_ARG, "null output");
  const double* da = stage_in(c, c.stage_a, a, n, mem);
  const double* db = (a == b) ? da : stage_in(c, c.stage_b, b, n, mem);
  c.partial.ensure((size_t)kMaxPartials * kMaxBasis * sizeof(double));
  c.scal.ensure(64 * sizeof(double));
  double* res = c.scal.as<double>();
  dev_dot(da, db, n, c.partial.as<double>(), res, c.stream);
  EDH_HIP(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  EDH_HIP(hipStreamSynchronize(c.stream));
  API_END
}

int edh_vv_noalloc(edh_ctx* ctx, const double* a, const double* b, int64_t n, double* out) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  use_device(c);
  EDH_CHECK(a && b && out, EDH_ERR_ARG, "null device pointer");
  noalloc_dot(c.device, a, b, n, out, c.stream);
  API_END
}

}  // extern "C"

namespace edh {
namespace {

// Runs the recurrence with v0 already in lz[0]; returns steps; fills host alpha/beta.
int lanczos_device(Ctx& c, int nl, double* alpha, double* beta, double* step_ms) {
  const int64_t n = c.tabs.N;
  hipStream_t s = c.stream;
  const int nblk = std::max(hv_num_blocks(c.dev), kMaxPartials);
  c.partial.ensure((size_t)std::max(nblk, kMaxPartials * kMaxBasis) * sizeof(double));
  c.scal.ensure((size_t)(8 + 2 * nl + 2) * sizeof(double));
  double* sc = c.scal.as<double>();
  int* st = (int*)(sc + 4);
  double* dalpha = sc + 8;
  double* dbeta = dalpha + nl;
  EDH_HIP(hipMemsetAsync(dalpha, 0, (size_t)(2 * nl + 1) * sizeof(double), s));
  dev_lanczos_init(sc, st, s);
  double* cur = c.lz[0].as<double>();
  double* prev = c.lz[1].as<double>();
  double* out = c.lz[2].as<double>();
  double* partial = c.partial.as<double>();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (step_ms) {
    EDH_HIP(hipEventCreate(&e0));
    EDH_HIP(hipEventCreate(&e1));
    EDH_HIP(hipEventRecord(e0, s));
  }
  // One Lanczos step = H*v with the alpha partial sums, the alpha reduction,
  // then axpy + norm and the beta reduction; every scalar and the step number
  // live on the device.  The step is captured once per buffer rotation
  // (period 3) as a hipGraph and replayed; EDH_LANCZOS_GRAPH=0 launches the
  // kernels directly instead.
  auto step = [&](double* cu, double* pr, double* ou, int k) {
    HvArgs a{};
    a.v = cu;
    a.w = ou;
    a.wprev = pr;
    a.mode = 1;  // step 1: the w scale is 0 and the previous vector is zeroed
    a.dscale = sc;
    a.dot_partial = partial;
    a.done = st + 1;
    launch_hv(c.dev, a, s);
    dev_lanczos_alpha(partial, hv_num_blocks(c.dev), dalpha, k, sc, st, s);
    dev_lanczos_passB(ou, cu, sc + 2, n, partial, st + 1, dbeta, k, nl, sc, st, s);
  };
  dev_fill(prev, n, 0.0, s);
  const char* genv = getenv("EDH_LANCZOS_GRAPH");
  const bool use_graph = nl >= 4 && !(genv && genv[0] == '0');
  bool graphed = false;
  // executable graphs are destroyed only after the stream has drained: HIP
  // does not defer the destruction of an exec whose launches still run
  hipGraphExec_t ex[3] = {nullptr, nullptr, nullptr};
  if (use_graph) {
    double* rot[3] = {cur, prev, out};
    bool ok = true;
    for (int ph = 0; ph < 3 && ok; ++ph) {
      hipGraph_t g = nullptr;
      if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
        ok = false;
        break;
      }
      // phase ph: (cur, prev, out) = rotation^ph of the three buffers
      step(rot[(3 - ph) % 3], rot[(4 - ph) % 3], rot[(5 - ph) % 3], 0);
      if (hipStreamEndCapture(s, &g) != hipSuccess || !g) {
        ok = false;
        break;
      }
      if (hipGraphInstantiate(&ex[ph], g, nullptr, nullptr, 0) != hipSuccess) ok = false;
      (void)hipGraphDestroy(g);
    }
    if (ok) {
      for (int k = 1; k <= nl; ++k) EDH_HIP(hipGraphLaunch(ex[(k - 1) % 3], s));
      graphed = true;
    } else {
      for (auto& e : ex) {
        if (e) (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
    }
    (void)hipGetLastError();
  }
  if (!graphed) {
    for (int k = 1; k <= nl; ++k) {
      step(cur, prev, out, k);
      double* t = prev;
      prev = cur;
      cur = out;
      out = t;
    }
  }
  if (step_ms) {
    EDH_HIP(hipEventRecord(e1, s));
    EDH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    EDH_HIP(hipEventElapsedTime(&ms, e0, e1));
    *step_ms = (double)ms / nl;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  int host_st[2];
  EDH_HIP(hipMemcpyAsync(alpha, dalpha, nl * sizeof(double), hipMemcpyDeviceToHost, s));
  EDH_HIP(hipMemcpyAsync(beta, dbeta, (nl + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
  EDH_HIP(hipMemcpyAsync(host_st, st, sizeof(host_st), hipMemcpyDeviceToHost, s));
  const hipError_t se = hipStreamSynchronize(s);
  for (auto& e : ex)
    if (e) (void)hipGraphExecDestroy(e);
  EDH_HIP(se);
  return host_st[0];
}

}  // namespace
}  // namespace edh

extern "C" {

int edh_lanczos(edh_ctx* ctx, const double* v0, int64_t n, int nl, double* alpha, double* beta, int* nlanc, int mem) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  require_sector(c);
  use_device(c);
  EDH_CHECK(n == c.tabs.N, EDH_ERR_ARG, "vector size does not match the sector dimension");
  EDH_CHECK(nl >= 1 && alpha && beta && nlanc && v0, EDH_ERR_ARG, "bad argument");
  for (auto& b : c.lz) b.ensure((size_t)n * sizeof(double));
  EDH_HIP(hipMemcpyAsync(c.lz[0].p, v0, (size_t)n * sizeof(double),
                         mem == EDH_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c.stream));
  *nlanc = lanczos_device(c, nl, alpha, beta, nullptr);
  API_END
}

int edh_lanczos_time(edh_ctx* ctx, const double* v0, int nl, double* alpha, double* beta, int* nlanc,
                     double* avg_step_ms) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  require_sector(c);
  use_device(c);
  EDH_CHECK(nl >= 1 && alpha && beta && nlanc && v0 && avg_step_ms, EDH_ERR_ARG, "bad argument");
  const int64_t n = c.tabs.N;
  for (auto& b : c.lz) b.ensure((size_t)n * sizeof(double));
  EDH_HIP(hipMemcpyAsync(c.lz[0].p, v0, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  *nlanc = lanczos_device(c, nl, alpha, beta, avg_step_ms);
  API_END
}

int edh_adag(edh_ctx* ctx, int nup, int ndn, int orb, int annihilate, const double* in, double* out, int64_t* out_n,
             int mem) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  EDH_CHECK(c.model, EDH_ERR_STATE, "no model set");
  use_device(c);
  const int Ns = c.model->Ns;
  EDH_CHECK(orb >= 0 && orb < 2 * Ns, EDH_ERR_ARG, "orbital index out of range");
  EDH_CHECK(nup >= 0 && nup <= Ns && ndn >= 0 && ndn <= Ns, EDH_ERR_ARG, "sector out of range");
  int nup2, ndn2;
  const int spin = orb >= Ns ? 1 : 0;
  nup2 = nup + (spin == 0 ? (annihilate ? -1 : 1) : 0);
  ndn2 = ndn + (spin == 1 ? (annihilate ? -1 : 1) : 0);
  Binomial C(Ns);
  if (nup2 < 0 || nup2 > Ns || ndn2 < 0 || ndn2 > Ns) {
    if (out_n) *out_n = 0;
    return EDH_OK;
  }
  const int64_t Du2 = C(Ns, nup2), Dd2 = C(Ns, ndn2), Dd = C(Ns, ndn), N = C(Ns, nup) * Dd;
  if (out_n) *out_n = Du2 * Dd2;
  if (!out) return EDH_OK;
  EDH_CHECK(in, EDH_ERR_ARG, "null input");
  std::vector<int32_t> mu, md;
  std::vector<int8_t> su, sd;
  adag_maps(Ns, nup, ndn, orb, annihilate != 0, nup2, ndn2, mu, md, su, sd);
  const size_t bytes = (mu.size() + md.size()) * 4 + su.size() + sd.size() + 64;
  c.gtab.ensure(bytes);
  char* g = c.gtab.as<char>();
  int32_t* dmu = (int32_t*)g;
  int32_t* dmd = dmu + mu.size();
  int8_t* dsu = (int8_t*)(dmd + md.size());
  int8_t* dsd = dsu + su.size();
  EDH_HIP(hipMemcpyAsync(dmu, mu.data(), mu.size() * 4, hipMemcpyHostToDevice, c.stream));
  EDH_HIP(hipMemcpyAsync(dmd, md.data(), md.size() * 4, hipMemcpyHostToDevice, c.stream));
  EDH_HIP(hipMemcpyAsync(dsu, su.data(), su.size(), hipMemcpyHostToDevice, c.stream));
  EDH_HIP(hipMemcpyAsync(dsd, sd.data(), sd.size(), hipMemcpyHostToDevice, c.stream));
  const double* din = stage_in(c, c.stage_a, in, N, mem);
  double* dout = out;
  if (mem == EDH_MEM_HOST) {
    c.stage_b.ensure((size_t)Du2 * Dd2 * sizeof(double));
    dout = c.stage_b.as<double>();
  }
  dev_separable_gather(din, Dd, dmu, dmd, dsu, dsd, dout, Du2, Dd2, c.stream);
  if (mem == EDH_MEM_HOST)
    EDH_HIP(hipMemcpyAsync(out, dout, (size_t)Du2 * Dd2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  EDH_HIP(hipStreamSynchronize(c.stream));
  API_END
}

int edh_local_op(edh_ctx* ctx, int nup, int ndn, int op, int norb, const int* orbs, const double* in, double* out,
                 int mem) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  EDH_CHECK(c.model, EDH_ERR_STATE, "no model set");
  use_device(c);
  const int Ns = c.model->Ns;
  EDH_CHECK(op == EDH_OP_SZ || op == EDH_OP_N, EDH_ERR_ARG, "unknown operator");
  EDH_CHECK(norb >= 0 && (norb == 0 || orbs) && in && out, EDH_ERR_ARG, "bad argument");
  for (int q = 0; q < norb; ++q) EDH_CHECK(orbs[q] >= 0 && orbs[q] < Ns, EDH_ERR_ARG, "orbital out of range");
  std::vector<int32_t> bu = spin_basis(Ns, nup), bd = spin_basis(Ns, ndn);
  std::vector<double> fu(bu.size(), 0.0), fd(bd.size(), 0.0);
  const double cu = op == EDH_OP_SZ ? 0.5 : 1.0, cd = op == EDH_OP_SZ ? -0.5 : 1.0;
  for (size_t r = 0; r < bu.size(); ++r)
    for (int q = 0; q < norb; ++q) fu[r] += cu * ((bu[r] >> (Ns - 1 - orbs[q])) & 1);
  for (size_t r = 0; r < bd.size(); ++r)
    for (int q = 0; q < norb; ++q) fd[r] += cd * ((bd[r] >> (Ns - 1 - orbs[q])) & 1);
  c.gtab.ensure((fu.size() + fd.size()) * sizeof(double));
  double* dfu = c.gtab.as<double>();
  double* dfd = dfu + fu.size();
  EDH_HIP(hipMemcpyAsync(dfu, fu.data(), fu.size() * 8, hipMemcpyHostToDevice, c.stream));
  EDH_HIP(hipMemcpyAsync(dfd, fd.data(), fd.size() * 8, hipMemcpyHostToDevice, c.stream));
  const int64_t N = (int64_t)bu.size() * bd.size();
  const double* din = stage_in(c, c.stage_a, in, N, mem);
  double* dout = out;
  if (mem == EDH_MEM_HOST) {
    c.stage_b.ensure((size_t)N * sizeof(double));
    dout = c.stage_b.as<double>();
  }
  dev_separable_scale(din, dfu, dfd, dout, (int64_t)bu.size(), (int64_t)bd.size(), c.stream);
  if (mem == EDH_MEM_HOST) EDH_HIP(hipMemcpyAsync(out, dout, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  EDH_HIP(hipStreamSynchronize(c.stream));
  API_END
}

int edh_cf_eval(edh_ctx* ctx, int nz, const double* zre, const double* zim, double ev, int nlanc, int isign, double expb,
                const double* alpha, const double* beta, double* out_re, double* out_im) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  use_device(c);
  EDH_CHECK(nz >= 0 && nlanc >= 1 && zre && zim && alpha && beta && out_re && out_im, EDH_ERR_ARG, "bad argument");
  const size_t need = (size_t)(4 * nz + 2 * nlanc + 2) * sizeof(double);
  c.small.ensure(need);
  double* dz = c.small.as<double>();
  double* dzi = dz + nz;
  double* da = dzi + nz;
  double* db = da + nlanc;
  double* dr = db + nlanc + 1;
  double* di = dr + nz;
  EDH_HIP(hipMemcpyAsync(dz, zre, nz * 8, hipMemcpyHostToDevice, c.stream));
  EDH_HIP(hipMemcpyAsync(dzi, zim, nz * 8, hipMemcpyHostToDevice, c.stream));
  EDH_HIP(hipMemcpyAsync(da, alpha, nlanc * 8, hipMemcpyHostToDevice, c.stream));
  EDH_HIP(hipMemcpyAsync(db, beta, (nlanc + 1) * 8, hipMemcpyHostToDevice, c.stream));
  dev_cf_eval(nz, dz, dzi, ev, nlanc, isign, expb, da, db, dr, di, c.stream);
  EDH_HIP(hipMemcpyAsync(out_re, dr, nz * 8, hipMemcpyDeviceToHost, c.stream));
  EDH_HIP(hipMemcpyAsync(out_im, di, nz * 8, hipMemcpyDeviceToHost, c.stream));
  EDH_HIP(hipStreamSynchronize(c.stream));
  API_END
}

int edh_eigsh(edh_ctx* ctx, int nev, int ncv, double tol, int maxitr, uint64_t seed, double* evals, double* evecs,
              int* nconv, int* n_av) {
  API_BEGIN
  require_ctx(ctx);
  Ctx& c = ctx->c;
  require_sector(c);
  use_device(c);
  EDH_CHECK(evals && nconv && n_av && nev >= 1 && ncv >= 2, EDH_ERR_ARG, "bad argument");
  eigsh_run(c, nev, ncv, tol, maxitr, seed, evals, evecs, nconv, n_av);
  API_END
}


// ---------------------------------------------------------------------------
// Batched engine
// ---------------------------------------------------------------------------
int edh_batch_create(edh_batch** out, int device) {
  API_BEGIN
  EDH_CHECK(out, EDH_ERR_ARG, "null argument");
  int ndev = 0;
  EDH_HIP(hipGetDeviceCount(&ndev));
  EDH_CHECK(device >= 0 && device < ndev, EDH_ERR_ARG, "device index out of range");
  EDH_HIP(hipSetDevice(device));
  std::unique_ptr<edh_batch> b(new edh_batch());
  b->b.device = device;
  EDH_HIP(hipStreamCreate(&b->b.stream));
  edh::configure_mem_pool(device);
  *out = b.release();
  API_END
}

int edh_batch_destroy(edh_batch* b) {
  delete b;
  return EDH_OK;
}

namespace {
edh::Batch& use_batch(edh_batch* b) {
  EDH_CHECK(b != nullptr, EDH_ERR_ARG, "null batch");
  EDH_HIP(hipSetDevice(b->b.device));
  return b->b;
}
}  // namespace

int edh_batch_add_sector(edh_batch* b, const edh_model* m, int nup, int ndn, int* sec_id, int64_t* dim) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(m && sec_id, EDH_ERR_ARG, "null argument");
  *sec_id = B.add_sector(*m->m, nup, ndn);
  if (dim) *dim = B.sector(*sec_id).tabs.N;
  API_END
}

int edh_batch_add_sectors(edh_batch* b, const edh_model* m, int count, const int* nups, const int* ndns, int* sec_ids,
                          int64_t* dims) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(m && count >= 0 && (count == 0 || (nups && ndns && sec_ids)), EDH_ERR_ARG, "null argument");
  std::vector<std::pair<int, int>> list(count);
  for (int q = 0; q < count; ++q) list[q] = {nups[q], ndns[q]};
  const std::vector<int> ids = B.add_sectors(*m->m, list);
  for (int q = 0; q < count; ++q) {
    sec_ids[q] = ids[q];
    if (dims) dims[q] = B.sector(ids[q]).tabs.N;
  }
  API_END
}

int edh_batch_remove_sector(edh_batch* b, int sec_id) {
  API_BEGIN
  use_batch(b).remove_sector(sec_id);
  API_END
}

int edh_batch_submit_eig(edh_batch* b, int sec_id, int nev, int ncv, double tol, int maxitr, uint64_t seed,
                         int* job_id) {
  API_BEGIN
  EDH_CHECK(job_id, EDH_ERR_ARG, "null argument");
  *job_id = use_batch(b).submit_eig(sec_id, nev, ncv, tol, maxitr, seed);
  API_END
}

int edh_batch_submit_lanczos(edh_batch* b, int sec_id, const double* v0, int nl, int mem, int* job_id) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(v0 && job_id, EDH_ERR_ARG, "null argument");
  const int64_t n = B.sector(sec_id).tabs.N;
  const double* src = v0;
  if (mem == EDH_MEM_HOST) {
    B.d_scal.ensure((size_t)std::max<int64_t>(n, 1) * sizeof(double));
    EDH_HIP(hipMemcpyAsync(B.d_scal.p, v0, (size_t)n * sizeof(double), hipMemcpyHostToDevice, B.stream));
    src = B.d_scal.as<double>();
  }
  *job_id = B.submit_lanczos(sec_id, src, nl);
  EDH_HIP(hipStreamSynchronize(B.stream));
  API_END
}

int edh_batch_submit_gf(edh_batch* b, int count, const int* eig_jobs, const int* indices, const int* orbs,
                        const int* annihilate, const int* dst_secs, int nl, int* job_ids) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(count >= 0, EDH_ERR_ARG, "negative count");
  EDH_CHECK(count == 0 || (eig_jobs && indices && orbs && annihilate && dst_secs && job_ids), EDH_ERR_ARG,
            "null argument");
  std::vector<edh::GfRequest> req(count);
  for (int r = 0; r < count; ++r) req[r] = {eig_jobs[r], indices[r], orbs[r], annihilate[r] ? 1 : 0, dst_secs[r]};
  std::vector<int> ids;
  B.submit_gf(req, nl, ids);
  for (int r = 0; r < count; ++r) job_ids[r] = ids[r];
  API_END
}

int edh_batch_submit_gf2(edh_batch* b, int count, const int* eig_jobs, const int* indices, const int* orbs,
                         const int* orbs2, const int* annihilate, const int* dst_secs, int nl, int* job_ids) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(count >= 0, EDH_ERR_ARG, "negative count");
  EDH_CHECK(count == 0 || (eig_jobs && indices && orbs && orbs2 && annihilate && dst_secs && job_ids), EDH_ERR_ARG,
            "null argument");
  std::vector<edh::GfRequest> req(count);
  for (int r = 0; r < count; ++r) {
    req[r] = {eig_jobs[r], indices[r], orbs[r], annihilate[r] ? 1 : 0, dst_secs[r]};
    req[r].orb2 = orbs2[r] < 0 ? -1 : orbs2[r];
  }
  std::vector<int> ids;
  B.submit_gf(req, nl, ids);
  for (int r = 0; r < count; ++r) job_ids[r] = ids[r];
  API_END
}

int edh_batch_lanczos_norm2(edh_batch* b, int count, const int* job_ids, double* norm2) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(count >= 0 && (count == 0 || (job_ids && norm2)), EDH_ERR_ARG, "bad argument");
  for (int q = 0; q < count; ++q) {
    const edh::BJob& j = B.job(job_ids[q]);
    EDH_CHECK(j.kind == edh::JOB_LZ, EDH_ERR_ARG, "batch: not a Lanczos job");
    norm2[q] = j.state == edh::JOB_DONE ? j.norm2 : std::numeric_limits<double>::quiet_NaN();
  }
  API_END
}

int edh_batch_step(edh_batch* b, int nsweeps, int* n_active) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  int act = 0;
  for (int i = 0; i < std::max(nsweeps, 1); ++i) {
    act = B.step();
    if (act == 0) break;
  }
  if (n_active) *n_active = act;
  API_END
}

int edh_batch_job_info(edh_batch* b, int job_id, int* state, int* kind, int64_t* n, int* count, int* n_av) {
  API_BEGIN
  edh::BJob& j = use_batch(b).job(job_id);
  if (state) *state = j.state;
  if (kind) *kind = j.kind;
  if (n) *n = j.n;
  if (count) *count = j.kind == edh::JOB_EIG ? j.nev : j.nlanc;
  if (n_av) *n_av = j.kind == edh::JOB_EIG ? j.n_av : j.nlanc;
  API_END
}

int edh_batch_job_states(edh_batch* b, int count, const int* job_ids, int* states) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(count >= 0 && (count == 0 || (job_ids && states)), EDH_ERR_ARG, "bad argument");
  for (int q = 0; q < count; ++q) states[q] = B.job(job_ids[q]).state;
  API_END
}

int edh_batch_eig_bounds(edh_batch* b, int count, const int* job_ids, double* bounds) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(count >= 0 && (count == 0 || (job_ids && bounds)), EDH_ERR_ARG, "bad argument");
  for (int q = 0; q < count; ++q) {
    const edh::BJob& j = B.job(job_ids[q]);
    EDH_CHECK(j.kind == edh::JOB_EIG, EDH_ERR_ARG, "batch: not an eigensolver job");
    if (j.state == edh::JOB_DONE && !j.evals.empty()) bounds[q] = j.evals[0];
    else if (j.state == edh::JOB_ACTIVE && j.iter >= 1) bounds[q] = j.ritz0;
    else bounds[q] = std::numeric_limits<double>::quiet_NaN();
  }
  API_END
}

int edh_batch_eig_result(edh_batch* b, int job_id, double* evals, double* evecs, int* nconv) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  edh::BJob& j = B.job(job_id);
  EDH_CHECK(j.kind == edh::JOB_EIG && j.state != edh::JOB_ACTIVE, EDH_ERR_STATE, "batch: no eigensolver result");
  if (nconv) *nconv = j.nconv;
  if (evals)
    for (int q = 0; q < (int)j.evals.size(); ++q) evals[q] = j.evals[q];
  if (evecs && j.evecs && j.nev > 0) {
    EDH_HIP(hipMemcpyAsync(evecs, j.evecs, (size_t)j.nev * j.n * sizeof(double), hipMemcpyDeviceToHost, B.stream));
    EDH_HIP(hipStreamSynchronize(B.stream));
  }
  EDH_CHECK(j.state == edh::JOB_DONE, EDH_ERR_STATE, "batch: eigensolver job was cancelled");
  API_END
}

int edh_batch_static_observables(edh_batch* b, int job_id, int index, int norb, double* out) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  edh::BJob& j = B.job(job_id);
  EDH_CHECK(j.kind == edh::JOB_EIG && j.state == edh::JOB_DONE && j.evecs, EDH_ERR_STATE,
            "batch: eigensolver job has no eigenvectors");
  EDH_CHECK(index >= 0 && index < j.nev && out, EDH_ERR_ARG, "batch: eigenpair index out of range");
  edh::BSector& s = B.sector(j.sec);
  edh::static_obs_sector(s.tabs, s.dev, s.model->Ns, j.evecs + (size_t)index * j.n, norb, B.d_obs, B.stream, out);
  API_END
}

int edh_batch_static_observables_many(edh_batch* b, int count, const int* job_ids, const int* indices, int norb,
                                      double* out) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(count >= 0 && (count == 0 || (job_ids && indices && out)), EDH_ERR_ARG, "null argument");
  std::vector<size_t> soff(count + 1, 0), hoff(count + 1, 0);
  for (int q = 0; q < count; ++q) {
    edh::BJob& j = B.job(job_ids[q]);
    EDH_CHECK(j.kind == edh::JOB_EIG && j.state == edh::JOB_DONE && j.evecs, EDH_ERR_STATE,
              "batch: eigensolver job has no eigenvectors");
    EDH_CHECK(indices[q] >= 0 && indices[q] < j.nev, EDH_ERR_ARG, "batch: eigenpair index out of range");
    edh::BSector& s = B.sector(j.sec);
    EDH_CHECK(norb >= 1 && norb <= s.model->Ns, EDH_ERR_ARG, "static observables: interacting orbitals out of range");
    soff[q + 1] = soff[q] + ((edh::static_obs_scratch(s.tabs, s.model->Ns) + 31) & ~size_t(31));
    hoff[q + 1] = hoff[q] + edh::static_obs_host(s.tabs, s.model->Ns);
  }
  if (count == 0) return EDH_OK;
  B.d_obs.ensure(soff[count] * sizeof(double));
  if (hoff[count] > B.h_obs_count) {
    EDH_HIP(hipStreamSynchronize(B.stream));  // no copy into the old buffer is pending
    if (B.h_obs) EDH_HIP(hipHostFree(B.h_obs));
    B.h_obs = nullptr;
    B.h_obs_count = 0;
    EDH_HIP(hipHostMalloc((void**)&B.h_obs, hoff[count] * sizeof(double), hipHostMallocDefault));
    B.h_obs_count = hoff[count];
  }
  for (int q = 0; q < count; ++q) {
    edh::BJob& j = B.job(job_ids[q]);
    edh::BSector& s = B.sector(j.sec);
    edh::static_obs_launch(s.tabs, s.dev, s.model->Ns, j.evecs + (size_t)indices[q] * j.n, B.d_obs.as<double>() + soff[q],
                           B.stream, B.h_obs + hoff[q]);
  }
  EDH_HIP(hipStreamSynchronize(B.stream));
  const size_t osz = 5 * (size_t)norb + (size_t)norb * norb + 1;
  for (int q = 0; q < count; ++q) {
    edh::BJob& j = B.job(job_ids[q]);
    edh::BSector& s = B.sector(j.sec);
    edh::static_obs_reduce(s.tabs, s.model->Ns, norb, B.h_obs + hoff[q], out + q * osz);
  }
  API_END
}

int edh_batch_lanczos_result(edh_batch* b, int job_id, double* alpha, double* beta, int* nlanc) {
  API_BEGIN
  edh::BJob& j = use_batch(b).job(job_id);
  EDH_CHECK(j.kind == edh::JOB_LZ && j.state == edh::JOB_DONE, EDH_ERR_STATE, "batch: no Lanczos result");
  if (nlanc) *nlanc = j.nlanc;
  if (alpha) std::copy(j.alpha.begin(), j.alpha.end(), alpha);
  if (beta) std::copy(j.beta.begin(), j.beta.end(), beta);
  API_END
}

int edh_batch_cancel(edh_batch* b, int job_id) {
  API_BEGIN
  edh::BJob& j = use_batch(b).job(job_id);
  if (j.state == edh::JOB_ACTIVE) j.state = edh::JOB_FAILED;
  API_END
}

int edh_batch_free_job(edh_batch* b, int job_id) {
  API_BEGIN
  use_batch(b).free_job(job_id);
  API_END
}

int edh_batch_memory(edh_batch* b, size_t* bytes) {
  API_BEGIN
  EDH_CHECK(bytes, EDH_ERR_ARG, "null argument");
  *bytes = use_batch(b).bytes_in_use;
  API_END
}

int edh_batch_set_graphs(edh_batch* b, int on) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  B.use_graphs = on != 0;
  if (!B.use_graphs) B.drop_graph();
  API_END
}

int edh_batch_sweeps(edh_batch* b, int64_t* sweeps, int64_t* graph_sweeps) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  if (sweeps) *sweeps = (int64_t)B.sweeps;
  if (graph_sweeps) *graph_sweeps = (int64_t)B.graph_sweeps;
  API_END
}

int edh_batch_autotune(edh_batch* b, int nsec, const int* sec_ids, int* block) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(nsec >= 0 && (nsec == 0 || sec_ids), EDH_ERR_ARG, "bad argument");
  std::vector<int> ids(sec_ids, sec_ids + nsec);
  const int r = B.autotune(ids);
  if (block) *block = r;
  API_END
}

int edh_batch_time_hv(edh_batch* b, int nsec, const int* sec_ids, int reps, uint64_t seed, double* avg_ms,
                      int64_t* bytes, int64_t* elements) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(nsec >= 0 && (nsec == 0 || sec_ids) && avg_ms && bytes && elements, EDH_ERR_ARG, "bad argument");
  std::vector<int> ids(sec_ids, sec_ids + nsec);
  B.time_hv(ids, reps, seed, avg_ms, bytes, elements);
  API_END
}

namespace {
void batch_cf_many(edh_batch* b, int njobs, const int* job_ids, const double* ev, const int* isign, const double* expb,
                   const double* excited, int nz, const double* wr, const double* wn, double* out_re, double* out_im) {
  edh::Batch& B = use_batch(b);
  EDH_CHECK(njobs >= 0 && nz >= 0, EDH_ERR_ARG, "bad argument");
  if (njobs == 0 || nz == 0) return;
  EDH_CHECK(job_ids && ev && isign && expb && excited && wn && out_re && out_im, EDH_ERR_ARG, "null argument");
  std::vector<size_t> off(njobs);
  size_t tot = 0;
  for (int j = 0; j < njobs; ++j) {
    edh::BJob& J = B.job(job_ids[j]);
    EDH_CHECK(J.kind == edh::JOB_LZ && J.state == edh::JOB_DONE && J.nlanc >= 1, EDH_ERR_STATE,
              "batch: continued fraction of an unfinished Lanczos run");
    off[j] = tot;
    tot += J.alpha.size() + J.beta.size();
  }
  std::vector<double> coef(tot);
  for (int j = 0; j < njobs; ++j) {
    edh::BJob& J = B.job(job_ids[j]);
    std::copy(J.alpha.begin(), J.alpha.end(), coef.begin() + off[j]);
    std::copy(J.beta.begin(), J.beta.end(), coef.begin() + off[j] + J.alpha.size());
  }
  const size_t need = tot * 8 + 2 * (size_t)nz * 8 + (size_t)njobs * sizeof(edh::CfJob) + 2 * (size_t)njobs * nz * 8 + 64;
  B.d_scal.ensure(need);
  char* base = B.d_scal.as<char>();
  double* dcoef = (double*)base;
  double* dwn = dcoef + tot;
  double* dwr = dwn + nz;
  edh::CfJob* djobs = (edh::CfJob*)(dwr + nz);
  double* dre = (double*)(djobs + njobs);
  double* dim = dre + (size_t)njobs * nz;
  std::vector<edh::CfJob> cj(njobs);
  for (int j = 0; j < njobs; ++j) {
    edh::BJob& J = B.job(job_ids[j]);
    cj[j] = {ev[j], expb[j], excited[j], isign[j], J.nlanc, dcoef + off[j], dcoef + off[j] + J.alpha.size()};
  }
  EDH_HIP(hipMemcpyAsync(dcoef, coef.data(), tot * 8, hipMemcpyHostToDevice, B.stream));
  EDH_HIP(hipMemcpyAsync(dwn, wn, (size_t)nz * 8, hipMemcpyHostToDevice, B.stream));
  if (wr) EDH_HIP(hipMemcpyAsync(dwr, wr, (size_t)nz * 8, hipMemcpyHostToDevice, B.stream));
  EDH_HIP(hipMemcpyAsync(djobs, cj.data(), cj.size() * sizeof(edh::CfJob), hipMemcpyHostToDevice, B.stream));
  edh::dev_cf_batch(njobs, nz, wr ? dwr : nullptr, dwn, djobs, dre, dim, B.stream);
  EDH_HIP(hipMemcpyAsync(out_re, dre, (size_t)njobs * nz * 8, hipMemcpyDeviceToHost, B.stream));
  EDH_HIP(hipMemcpyAsync(out_im, dim, (size_t)njobs * nz * 8, hipMemcpyDeviceToHost, B.stream));
  EDH_HIP(hipStreamSynchronize(B.stream));
}
}  // namespace

int edh_batch_cf_many(edh_batch* b, int njobs, const int* job_ids, const double* ev, const int* isign,
                      const double* expb, const double* excited, int nz, const double* wn, double* out_re,
                      double* out_im) {
  API_BEGIN
  batch_cf_many(b, njobs, job_ids, ev, isign, expb, excited, nz, nullptr, wn, out_re, out_im);
  API_END
}

int edh_batch_cf_many_z(edh_batch* b, int njobs, const int* job_ids, const double* ev, const int* isign,
                        const double* expb, const double* excited, int nz, const double* zre, const double* zim,
                        double* out_re, double* out_im) {
  API_BEGIN
  EDH_CHECK(nz == 0 || zre, EDH_ERR_ARG, "null argument");
  batch_cf_many(b, njobs, job_ids, ev, isign, expb, excited, nz, zre, zim, out_re, out_im);
  API_END
}

int edh_batch_cf_eval(edh_batch* b, int nz, const double* zre, const double* zim, double ev, int nlanc, int isign,
                      double expb, const double* alpha, const double* beta, double* out_re, double* out_im) {
  API_BEGIN
  edh::Batch& B = use_batch(b);
  EDH_CHECK(nz >= 0 && nlanc >= 1 && zre && zim && alpha && beta && out_re && out_im, EDH_ERR_ARG, "bad argument");
  B.d_scal.ensure((size_t)(4 * nz + 2 * nlanc + 2) * sizeof(double));
  double* dz = B.d_scal.as<double>();
  double* dzi = dz + nz;
  double* da = dzi + nz;
  double* db = da + nlanc;
  double* dr = db + nlanc + 1;
  double* di = dr + nz;
  EDH_HIP(hipMemcpyAsync(dz, zre, nz * 8, hipMemcpyHostToDevice, B.stream));
  EDH_HIP(hipMemcpyAsync(dzi, zim, nz * 8, hipMemcpyHostToDevice, B.stream));
  EDH_HIP(hipMemcpyAsync(da, alpha, nlanc * 8, hipMemcpyHostToDevice, B.stream));
  EDH_HIP(hipMemcpyAsync(db, beta, (nlanc + 1) * 8, hipMemcpyHostToDevice, B.stream));
  dev_cf_eval(nz, dz, dzi, ev, nlanc, isign, expb, da, db, dr, di, B.stream);
  EDH_HIP(hipMemcpyAsync(out_re, dr, nz * 8, hipMemcpyDeviceToHost, B.stream));
  EDH_HIP(hipMemcpyAsync(out_im, di, nz * 8, hipMemcpyDeviceToHost, B.stream));
  EDH_HIP(hipStreamSynchronize(B.stream));
  API_END
}

}  // extern "C"
