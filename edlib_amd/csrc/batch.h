// Batched Krylov engine (grouped launches over many sectors and processes).
#pragma once

#include <memory>
#include <utility>
#include <vector>

#include "edh_ctx.h"

namespace edh {

enum JobKind { JOB_EIG = 0, JOB_LZ = 1 };
enum JobState { JOB_ACTIVE = 0, JOB_DONE = 1, JOB_FAILED = 2 };
enum OpPhase { PH_PROJECT = 0, PH_NORM = 1, PH_SCALE = 2 };

// One (model, sector) resident on the device.
struct BSector {
  std::unique_ptr<Model> model;
  SectorTables tabs;
  DevSector dev;
  size_t bytes = 0;
};

// One Krylov process.
struct BJob {
  int kind = JOB_EIG, state = JOB_ACTIVE, sec = -1;
  int64_t n = 0;
  std::vector<void*> allocs;
  size_t bytes = 0;
  double* dsc = nullptr;  // device scalars: h[kMaxBasis], nrm, projected-matrix columns
  // eigensolver run
  int nev = 0, m = 0, k = 0, col = 0, iter = 0, maxitr = 0, nconv = 0, n_av = 0;
  double tol = 0.0, nr_last = 0.0;
  bool pending_start_norm = false;
  double *V = nullptr, *Valt = nullptr, *evecs = nullptr, *Cbuf = nullptr;
  std::vector<double> T, Tc, evals;
  // Lanczos run: lz[0] = v_k, lz[1] = v_{k-1}, lz[2] = w
  int nl = 0, step = 0, nlanc = 0;
  double last_beta = 0.0;
  double* lz[3] = {nullptr, nullptr, nullptr};
  std::vector<double> alpha, beta;
};

// Host description of one grouped vector operation (see BOp in batch.hip).
struct BOpHost {
  const double* X;
  int64_t ldx;
  int cnt;
  double* w;
  int64_t n;
  double* h;
  double* hsum;
  double* nrm;
  double scale;
};

struct Batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<BSector>> secs;
  std::vector<std::unique_ptr<BJob>> jobs;
  DevBuf d_hv, d_ops, d_cjob, d_partial, d_scal, gtab;
  std::vector<char> h_hv;
  void* h_pin = nullptr;
  size_t h_pin_bytes = 0;
  size_t bytes_in_use = 0;
  int hv_block = 256;
  long long sweeps = 0;
  ~Batch();

  int add_sector(const Model& m, int nup, int ndn);
  void remove_sector(int id);
  BSector& sector(int id);
  BJob& job(int id);
  int submit_eig(int sec, int nev, int ncv, double tol, int maxitr, uint64_t seed);
  int submit_lanczos(int sec, const double* v0_dev, int nl);
  int submit_gf(int src_sec, const double* psi, int orb, bool annihilate, int dst_sec, int nl, double* norm);
  void free_job(int id);
  int step();
  void adag(int src, const double* in, int orb, bool annihilate, int dst, double* out);
  int autotune(const std::vector<int>& sec_ids);

 private:
  double* dalloc(BJob& j, size_t count);
  void release_job(BJob& j);
  void finish_dense(BJob& j);
  void end_cycle(BJob& j);
  void run_ops(std::vector<BOpHost>& ops, int phase);
  void run_hv(std::vector<std::pair<int, HvArgs>>& hv);
  void fill_dev(double* p, size_t n);
};

}  // namespace edh
