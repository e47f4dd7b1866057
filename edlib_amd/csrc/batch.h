// Batched Krylov engine (grouped launches over many sectors and processes).
#pragma once

#include <limits>
#include <map>
#include <memory>
#include <tuple>
#include <utility>
#include <vector>

#include "edh_ctx.h"

namespace edh {

enum JobKind { JOB_EIG = 0, JOB_LZ = 1 };
enum JobState { JOB_ACTIVE = 0, JOB_DONE = 1, JOB_FAILED = 2 };

// One (model, sector) resident on the device.
struct BSector {
  std::unique_ptr<Model> model;
  SectorTables tabs;
  DevSector dev;
  size_t bytes = 0;
};

// One Krylov process.  sc is its device scalar block: projection
// coefficients, norms and flags, then (eigensolver) the projected-matrix
// columns or (Lanczos) the coefficient history and stop flags.
struct BJob {
  int kind = JOB_EIG, state = JOB_ACTIVE, sec = -1;
  int64_t n = 0;
  std::vector<void*> allocs;
  size_t bytes = 0;
  double* sc = nullptr;
  // eigensolver run (its column, restarts and convergence are kept on the
  // device in the control block ctl; the host sees them through the status
  // read of every sweep)
  int nev = 0, m = 0, iter = 0, maxitr = 0, nconv = 0, n_av = 0;
  double tol = 0.0;
  double ritz0 = 0.0;  // lowest Ritz value at the last cycle end (valid once iter >= 1)
  bool pending_start_norm = false;
  double *V = nullptr, *Valt = nullptr, *evecs = nullptr, *ctl = nullptr;
  std::vector<double> evals;
  // Lanczos run: lz[0] = v_k, lz[1] = v_{k-1}, lz[2] = w; step = next step;
  // lzb = the three buffers in their order at step 1 (the device resolves
  // the rotation from its own step counter)
  int nl = 0, step = 0, nlanc = 0;
  double* lz[3] = {nullptr, nullptr, nullptr};
  double* lzb[3] = {nullptr, nullptr, nullptr};
  double *d_alpha = nullptr, *d_beta = nullptr, *tdot = nullptr;
  int64_t ntdot = 0;
  int* d_flags = nullptr;  // [0] stopped (|beta| < 1e-10), [1] steps taken
  std::vector<double> alpha, beta;
  double norm2 = std::numeric_limits<double>::quiet_NaN();  // start vector's squared norm (Green's-function runs)
  // scalars of a finished Lanczos run: a place in one of the batch's pinned
  // read slabs (the run sits out the sweeps until that copy has completed)
  const double* h_read = nullptr;
  bool awaiting = false;
};

// Host description of one grouped vector operation on one process.
struct BOpHost {
  const double* X = nullptr;  // vectors to project on (cnt of them, stride ldx)
  int64_t ldx = 0;
  int cnt = 0;
  double* w = nullptr;  // operand
  int64_t n = 0;
  double* sc = nullptr;      // scalar block
  double* hsum = nullptr;    // projected-matrix column (accumulated) or null
  double* hist_a = nullptr;  // Lanczos alpha_k output or null
  double* hist_b = nullptr;  // Lanczos beta_k output or null
  int* flags = nullptr;      // Lanczos stop flags or null
  int kstep = 0;
  // Lanczos step resolved on the device: the run's step counter k; X, X1, w
  // are then its three buffers at step 1, hist_a / hist_b the alpha and
  // beta histories and m the run's length (kstep, the histories' entries
  // and the rotation follow k)
  const double* lzk = nullptr;
  int mode = 0;  // OP_* bits
  double scale = 1.0;
  double* nrm_out = nullptr;  // extra copy of the norm, or null
  // eigensolver operation: X, w, cnt and hsum follow the run's control block
  // (basis buffers X and X1, basis size m)
  const double* ctl = nullptr;
  const double* X1 = nullptr;
  int m = 0;
  // Lanczos step (G_LZ): the grouped H*v's per-task shares of <v, w>
  const double* tdot = nullptr;
  int64_t ntdot = 0;
};

// OP_LZS: the Lanczos vectors are kept unnormalised (the next step's H*v
// and update scale them), so the norm also sets the next step's scales.
// OP_START: the norm of a Lanczos run's unnormalised start vector sets its
// first scale and its squared norm (nrm_out), or stops the run when the
// squared norm is <= 1e-10 (the vector vanishes).
enum OpMode { OP_NB = 1, OP_GATED = 2, OP_LZ = 4, OP_SCALE_NRM = 8, OP_LZS = 16, OP_START = 32 };
enum GroupKind { G_PROJECT = 0, G_NORM = 1, G_SCALE = 2, G_REORTH = 3, G_LZ = 4 };

struct OpGroup {
  int kind;
  std::vector<BOpHost> ops;
};

// Scalars of one process to be copied back with the sweep's grouped read
// (an eigensolver run's status every sweep, a Lanczos run's coefficients at
// its end).
struct ReadReq {
  int job;
  const double* src;
  size_t count;
};

// Pinned host buffer receiving one sweep's grouped read; its entries are
// handled, in sweep order, once the copy has completed.
struct ReadSlab {
  double* host = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool busy = false;
  std::vector<ReadReq> entries;
  std::vector<size_t> offs;  // entry q's place in host
};

struct GfRequest {
  int eig_job, index, orb, annihilate, dst_sec;
  int orb2 = -1;  // >= 0: start vector (c(+)_orb + c(+)_orb2)|psi>
};

struct Batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<BSector>> secs;
  std::vector<std::unique_ptr<BJob>> jobs;
  DevBuf d_arena, d_partial, d_scal;
  UploadStage uploads;  // pinned staging of sector tables (no wait per sector)
  void* h_stage[2] = {nullptr, nullptr};
  size_t h_stage_bytes[2] = {0, 0};
  hipEvent_t ev_stage[2] = {nullptr, nullptr};
  bool ev_used[2] = {false, false};
  int stage_i = 0;
  // (sector, orbital, annihilate) -> device gather maps (four arrays, then
  // their one allocation) and the host blob they were copied from
  std::map<std::tuple<int, int, int>, std::vector<void*>> maps;
  std::map<std::tuple<int, int, int>, std::vector<char>> map_blobs;
  // stream-ordered frees not yet issued: a hipFreeAsync costs tens of
  // microseconds of host time, so they are issued while the host would
  // otherwise wait for the device (the memory is no longer referenced by
  // anything issued after the logical free, so issuing later is safe)
  std::vector<void*> pending_free;
  void defer_free(void* p) {
    if (p) pending_free.push_back(p);
  }
  void issue_frees_until(hipEvent_t ev);  // until ev completes (all when ev is null)
  DevBuf d_gather;
  DevBuf d_obs;  // static-observable scratch
  double* h_obs = nullptr;  // pinned host copies of the static-observable sums
  size_t h_obs_count = 0;
  std::vector<ReadSlab> slabs;
  std::vector<int> slab_fifo;  // slabs in flight, oldest first

  size_t bytes_in_use = 0;
  int hv_block = 256;
  // Steady sweeps (the same descriptors as the sweep before: no process
  // started or finished, no Lanczos run at its last step, no operand on the
  // rows kernel) replay one captured graph of the sweep's kernels and skip
  // the descriptor upload.
  bool use_graphs = true;
  std::vector<char> last_desc;     // descriptor bytes of the last sweep
  std::vector<int64_t> last_cfg;   // and its launch configuration
  const void* last_arena = nullptr;
  // The host runs at most kAhead sweeps ahead of the device (an event per
  // sweep in a ring): it learns of finished processes, and so starts the
  // work that waits on them, within that many sweeps.
  static constexpr int kAhead = 4;
  hipEvent_t ahead_ev[kAhead] = {};
  bool ahead_used[kAhead] = {};
  int ahead_i = 0;
  hipGraphExec_t sweep_graph = nullptr;
  hipEvent_t graph_ev = nullptr;  // recorded after every launch of sweep_graph
  bool graph_launched = false;
  size_t graph_sweeps = 0;  // sweeps run as graph replays
  size_t graph_failures = 0;  // captures whose instantiation failed (those sweeps ran eagerly)
  // Executable graphs that were launched are destroyed only once their last
  // launch has completed (HIP does not defer the destruction of an exec that
  // is still running): retired ones wait here with the event of that launch.
  std::vector<std::pair<hipGraphExec_t, hipEvent_t>> retired_graphs;
  void drop_graph();
  void reap_graphs(bool wait);
  long long sweeps = 0;
  ~Batch();

  int add_sector(const Model& m, int nup, int ndn);
  // many sectors of one model: tables built on up to 16 host threads, then
  // uploaded in order; ids in the order of `list`
  std::vector<int> add_sectors(const Model& m, const std::vector<std::pair<int, int>>& list);
  void remove_sector(int id);
  BSector& sector(int id);
  BJob& job(int id);
  int submit_eig(int sec, int nev, int ncv, double tol, int maxitr, uint64_t seed);
  int submit_lanczos(int sec, const double* v0_dev, int nl);
  // Green's-function start vectors and their Lanczos runs; ids[i] = -1 when
  // vector i vanishes; norms[i] = its squared norm.
  void submit_gf(const std::vector<GfRequest>& req, int nl, std::vector<int>& ids);
  void free_job(int id);
  int step();
  int autotune(const std::vector<int>& sec_ids);
  void time_hv(const std::vector<int>& sec_ids, int reps, uint64_t seed, double* avg_ms, int64_t* bytes,
               int64_t* elements);

 private:
  double* dalloc(BJob& j, size_t count);
  void release_job(BJob& j);
  bool drain_slab(bool wait);
  void eig_status(BJob& j, const double* h);
  void finish_dense(BJob& j);
  void finish_lanczos(BJob& j);
  void issue(std::vector<OpGroup>& pre, const std::vector<std::pair<int, HvArgs>>* hv, std::vector<OpGroup>& post,
             const std::vector<BJob*>* cycles = nullptr, const std::vector<ReadReq>* reads = nullptr);
  void fill_dev(double* p, size_t n);
  void gather_maps(int src, int orb, bool annihilate, std::vector<void*>& out);
};

}  // namespace edh
