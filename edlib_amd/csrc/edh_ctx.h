// Context object behind the edh_ctx handle.
#pragma once

#include "edh_internal.h"
#include "edh_kernels.h"

namespace edh {

// Growable device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b) {
    if (b <= bytes) return;
    if (p) EDH_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) throw Error(EDH_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<Model> model;
  SectorTables tabs;
  DevSector dev;
  bool has_sector = false;
  DevBuf stage_a, stage_b;  // host-pointer staging
  DevBuf partial;           // reduction partials
  DevBuf scal;              // device scalars
  DevBuf lz[3];             // Lanczos work vectors
  DevBuf krylov[2];         // thick-restart bases (double-buffered)
  DevBuf evecs;             // eigenvectors of the last eigsh (resident)
  DevBuf small;             // small device arrays (coefficients, tables)
  DevBuf gtab;              // gather tables for a_adag / operators
  DevBuf obs;               // static-observable row sums and column tiles
  int n_evecs = 0;
  ~Ctx();
};

// Ritz vectors kept by a thick restart of an m-vector basis with nev
// wanted pairs of which nconv have converged.
int restart_size(int nev, int m, int nconv);

void eigsh_run(Ctx& c, int nev, int ncv, double tol, int maxitr, uint64_t seed, double* evals, double* evecs,
               int* nconv, int* n_av);

// Static observables of one device eigenvector of the current sector
// (StaticObservables.h:356-419): out = N, N_up, N_dn, M, D_occ (norb each),
// M_i M_j (norb x norb, row-major), N_eff.
void static_obs_run(Ctx& c, const double* dv, int norb, double* out);
// The same for a vector of any uploaded sector (tables st / ds), with its
// scratch and stream.
void static_obs_sector(const SectorTables& st, const DevSector& ds, int Ns, const double* dv, int norb,
                       DevBuf& scratch, hipStream_t stream, double* out);
// The same in two phases, for many vectors with one wait: the device sums
// of one vector (static_obs_scratch(st, Ns) doubles of device scratch,
// copied into static_obs_host(st, Ns) doubles of pinned host memory), then,
// once the copies have arrived, the host reduction into out.
size_t static_obs_scratch(const SectorTables& st, int Ns);
size_t static_obs_host(const SectorTables& st, int Ns);
void static_obs_launch(const SectorTables& st, const DevSector& ds, int Ns, const double* dv, double* scratch,
                       hipStream_t stream, double* host);
void static_obs_reduce(const SectorTables& st, int Ns, int norb, const double* host, double* out);

// Diagonal element of state (u, d) of the current sector from the host tables.
double host_diag(const Model& m, const SectorTables& st, int64_t u, int64_t d);

// Dense sector matrix (row = source state) from the host tables.
void dense_sector_matrix(const Model& m, const SectorTables& st, std::vector<double>& H);

// Gather maps of c (annihilate) or c^+ on orbital orb from sector (nup, ndn)
// into (nup2, ndn2): target (u', d') reads source (mu[u'], md[d']) times
// su[u'] sd[d'] (index -1 = zero).
void adag_gather_maps(int Ns, int nup, int ndn, int orb, bool annihilate, int& nup2, int& ndn2,
                      std::vector<int32_t>& mu, std::vector<int32_t>& md, std::vector<int8_t>& su,
                      std::vector<int8_t>& sd);

// Allocation-free inner product (noalloc.hip): scratch array resolved per
// device at context creation.
void noalloc_register_device(int device);
void noalloc_dot(int device, const double* a, const double* b, int64_t n, double* out, hipStream_t s);

}  // namespace edh
