/*
 * edhip -- C ABI of the MI355X-native spin-resolved exact-diagonalisation
 * engine (libedhip.so).  Plain pointers and sizes only; every entry returns
 * an int status (EDH_OK == 0, negative on error) and never lets a C++
 * exception cross the boundary.  The message of the last failure is
 * available from edh_last_error() (per calling thread).
 *
 * Each entry names the reference (Q-solvers/EDLib) interface it replaces.
 * Memory kinds: EDH_MEM_HOST means the vector arguments are host pointers
 * (staged through device buffers owned by the context), EDH_MEM_DEVICE means
 * they are device pointers on the context's GPU and the call stays on the
 * context's stream (asynchronous unless stated otherwise).
 */
#ifndef EDHIP_H
#define EDHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EDH_OK 0
#define EDH_ERR_ARG -1
#define EDH_ERR_HIP -2
#define EDH_ERR_STATE -3
#define EDH_ERR_LIMIT -4
#define EDH_ERR_NOMEM -5
#define EDH_ERR_NOCONV -6
#define EDH_ERR_INTERNAL -9

#define EDH_MODEL_HUBBARD 0
#define EDH_MODEL_SIAM 1

#define EDH_MEM_HOST 0
#define EDH_MEM_DEVICE 1

#define EDH_COPY_H2D 0
#define EDH_COPY_D2H 1
#define EDH_COPY_D2D 2

#define EDH_OP_SZ 0 /* SzOperator::action, ChiLoc.h:57-61 */
#define EDH_OP_N 1  /* NOperator::action,  ChiLoc.h:85-89 */

typedef struct edh_ctx edh_ctx;
typedef struct edh_model edh_model;

/* Message of the last failed call on this thread ("" if none). */
const char* edh_last_error(void);
int edh_version(void);

/* ---- Models (host side; no GPU needed) -------------------------------- */

/* HubbardModel<double>(params) constructor, HubbardModel.h:48-77.
 * t: Ns*Ns hopping/values (H_hop = -sum t_ij c^+_j c_i); U, xmu, Hmag: Ns;
 * J: Ns*Ns exchange (may be NULL); Hmag may be NULL. */
int edh_model_hubbard(edh_model** out, int Ns, const double* t, const double* U, const double* J,
                      const double* xmu, const double* Hmag);

/* SingleImpurityAndersonModel<double>(params) constructor,
 * SingleImpurityAndersonModel.h:195-275.  Epsk: (Ns-ml)*2 (Bath/Epsk),
 * Vk: ml*(Ns-ml)*2 (Bath/Vk_m), H0: ml*ml*2 (H0_m), Ut: 2*2*ml^4
 * (interaction/values), mu (mu). */
int edh_model_siam(edh_model** out, int Ns, int ml, const double* Epsk, const double* Vk, const double* H0,
                   const double* Ut, double mu);
/* Any SzSymmetry model given by its terms: the storage-facing interface of
 * the reference models (T_states(), V_states(), valid/set, diagonal() and
 * interacting_orbitals(), as SpinResolvedStorage::fill and fill_spin use
 * them, SpinResolvedStorage.h:120-155,622-638) reduced to plain arrays.
 *   one-body term q: hop_value[q] * c^+_{j s} c_{i s} with (i, j, s) =
 *     hop[3q..3q+2]; the model's sign rule is folded into the value (the
 *     reference's Hubbard terms carry -t, HubbardModel.h:104-105, its SIAM
 *     terms +V, SingleImpurityAndersonModel.h:77-84);
 *   two-body term q: inter_value[q] * c^+_{i s} c^+_{j sp} c_{l sp} c_{k s}
 *     with (i, j, k, l, s, sp) = inter[6q..6q+5], signs as
 *     InnerInteractionState::set (SingleImpurityAndersonModel.h:156-163);
 *     every site index below interacting_orbitals (at most 6 of them);
 *   diagonal: e0 + sum_{o} eps[o] n_o + sum_s sum_{i<j} same[s][i][j] n_is n_js
 *     + sum_{i,j} cross[i][j] n_i(up) n_j(dn), o = site + spin*Ns.
 * Terms are applied in the order given (the order of the CSR merge). */
typedef struct edh_model_desc {
  int Ns;
  int interacting_orbitals;
  int n_hop;
  const int* hop;
  const double* hop_value;
  int n_int;
  const int* inter;
  const double* inter_value;
  double e0;
  const double* eps;   /* 2*Ns */
  const double* same;  /* 2*Ns*Ns, only i < j read */
  const double* cross; /* Ns*Ns */
} edh_model_desc;
#define EDH_MODEL_TERMS 2
int edh_model_from_terms(edh_model** out, const edh_model_desc* d);

int edh_model_free(edh_model* m);
int edh_model_sites(const edh_model* m, int* Ns, int* interacting_orbitals);

/* Off-diagonal action of the model on one basis state (host, no GPU):
 * every valid term in term order (T_states, then V_states) as
 * (target state, sign * value); *count is the number of terms applied
 * (at most cap are written).  Model::valid + Model::set,
 * HubbardModel.h:86-108 / SingleImpurityAndersonModel.h:346-357. */
int edh_model_apply(const edh_model* m, long long state, int cap, long long* targets, double* coef, int* count);

/* T_states() (which=0) / V_states() (which=1) as (kind,i,j,k,l,s,sp) + value;
 * pass NULL arrays to query *count.  HubbardModel.h:141-147. */
int edh_model_transitions(const edh_model* m, int which, int* count, int* ints, double* vals);

/* Model::diagonal(state) evaluated in the reference's term order,
 * HubbardModel.h:116-132 / SingleImpurityAndersonModel.h:282-306. */
int edh_model_diagonal(const edh_model* m, long long state, double* out);

/* Host-side table introspection for one (nup, ndn) sector, no GPU needed:
 * which = 0 up hopping, 1 down hopping (SpinResolvedStorage::fill_spin,
 * SpinResolvedStorage.h:622-638).  rowp has nrows+1 entries; pass NULL
 * arrays to query sizes. */
int edh_sector_csr(const edh_model* m, int nup, int ndn, int which, int64_t* nrows, int64_t* nnz, int* rowp,
                   int* col, double* val);
/* Diagonal of every state of the sector as the device evaluates it
 * (decomposed form), for checks against Model::diagonal. */
int edh_sector_diagonal(const edh_model* m, int nup, int ndn, double* out);

/* ---- Context ------------------------------------------------------------ */

/* Free and total device memory of GPU `device` (hipMemGetInfo), for
 * admission by footprint in a server (no counterpart in the reference). */
int edh_device_memory(int device, size_t* free_bytes, size_t* total_bytes);

int edh_create(edh_ctx** out, int device);
int edh_destroy(edh_ctx* ctx);

/* Copies the model into the context (Hamiltonian(p) constructor, Hamiltonian.h:40-42). */
int edh_set_model(edh_ctx* ctx, const edh_model* m);

/* SpinResolvedStorage::reset + fill for sector (nup, ndn)
 * (SpinResolvedStorage.h:120-155,172-250): builds the per-spin tables and
 * uploads them.  *dim receives the sector dimension (ntot()). */
int edh_set_sector(edh_ctx* ctx, int nup, int ndn, int64_t* dim);
/* SpinResolvedStorage::constant_shift (SpinResolvedStorage.h:464-466): adds
 * `shift` to every diagonal element of the current sector (until the next
 * edh_set_sector, as the reference's refill discards it). */
int edh_constant_shift(edh_ctx* ctx, double shift);
/* Per-spin dimensions and algorithmic table bytes of the current sector. */
int edh_sector_info(edh_ctx* ctx, int64_t* dim_up, int64_t* dim_dn, int64_t* table_bytes);

/* ---- Device memory helpers ------------------------------------------------ */
int edh_alloc(edh_ctx* ctx, size_t bytes, void** dptr);
int edh_free(edh_ctx* ctx, void* dptr);
int edh_memcpy(edh_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);
int edh_synchronize(edh_ctx* ctx);

/* ---- Hot path -------------------------------------------------------------- */

/* SpinResolvedStorage::av(v, w, n, clear) (SpinResolvedStorage.h:58-115):
 * w = H v + (clear ? 0 : w).  With EDH_MEM_HOST the call is synchronous. */
int edh_av(edh_ctx* ctx, const double* v, double* w, int64_t n, int clear, int mem);

/* StaticObservables::calculate_static_observables_eigenvector
 * (StaticObservables.h:356-419) for one eigenvector v (n = sector dimension)
 * of the current sector, with norb interacting orbitals (sites 0..norb-1):
 * out = N[norb], N_up[norb], N_dn[norb], M[norb], D_occ[norb],
 * M_i M_j[norb*norb] (row-major, the reference's orbital-index rule for the
 * second down occupation), N_eff[1] -- 5*norb + norb*norb + 1 doubles.  The
 * caller averages over eigenpairs with Boltzmann weights
 * (StaticObservables.h:106-158).  At most 32 sites per spin. */
int edh_static_observables(edh_ctx* ctx, const double* v, int64_t n, int mem, int norb, double* out);

/* H*v kernel variant of the current sector: 0 = single pass (rows), 1 = two
 * passes (rows, then XCD-aware column slabs for the up-spin term).  The
 * sector set-up times both and keeps the faster; edh_set_hv_variant
 * overrides the choice until the next edh_set_sector (-1 re-runs the timing). */
int edh_get_hv_variant(edh_ctx* ctx, int* variant);
int edh_set_hv_variant(edh_ctx* ctx, int variant);

/* Times `reps` back-to-back av launches on device vectors with HIP events on
 * the context stream after `warmup` untimed launches; *avg_ms = mean time. */
int edh_av_time(edh_ctx* ctx, const double* v, double* w, int clear, int warmup, int reps, double* avg_ms);

/* SpinResolvedStorage::vv (SpinResolvedStorage.h:390-400): *out = <a, b>. */
int edh_vv(edh_ctx* ctx, const double* a, const double* b, int64_t n, double* out, int mem);

/* edh_vv on device vectors without allocating: *out (a device scalar) =
 * <a, b>, queued on the context's stream, the reduction partials in a
 * static device array of 2048 doubles (16 KiB) per GPU instead of context
 * buffers.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops).  Calls that overlap on one GPU share that array and race. */
int edh_vv_noalloc(edh_ctx* ctx, const double* a, const double* b, int64_t n, double* out);

/* Lanczos::lanczos (Lanczos.h:36-77) on the current sector from the
 * normalised start vector v0 (length n): alpha[nl], beta[nl+1] (beta[0]
 * unused), *nlanc = number of steps taken (breaks when |beta| < 1e-10).
 * The whole recurrence stays on the device (fused H*v + dot, fused axpy +
 * norm); alpha/beta are host arrays. */
int edh_lanczos(edh_ctx* ctx, const double* v0, int64_t n, int nl, double* alpha, double* beta, int* nlanc,
                int mem);
/* Same, timing the recurrence only (v0 on the device); *avg_step_ms. */
int edh_lanczos_time(edh_ctx* ctx, const double* v0, int nl, double* alpha, double* beta, int* nlanc,
                     double* avg_step_ms);

/* SpinResolvedStorage::a_adag (SpinResolvedStorage.h:293-381): applies c
 * (annihilate=1) or c^+ to orbital orb_spin = site + spin*Ns of a vector in
 * sector (nup, ndn); out has the size of the resulting sector (*out_n). */
int edh_adag(edh_ctx* ctx, int nup, int ndn, int orb_spin, int annihilate, const double* in, double* out,
             int64_t* out_n, int mem);

/* ChiLoc::operation (ChiLoc.h:474-530): out[s] = (sum_o action_o(s)) in[s]
 * over the orbitals orbs[0..norb) in sector (nup, ndn); op = EDH_OP_SZ/N. */
int edh_local_op(edh_ctx* ctx, int nup, int ndn, int op, int norb, const int* orbs, const double* in,
                 double* out, int mem);

/* Lanczos::get_frac_point (Lanczos.h:201-224) at nz complex energies, on the
 * device: out = ev*expb*det[1]/det[0] with the reference's overflow rescaling. */
int edh_cf_eval(edh_ctx* ctx, int nz, const double* zre, const double* zim, double ev, int nlanc, int isign,
                double expb, const double* alpha, const double* beta, double* out_re, double* out_im);

/* Storage::diag (Storage.h:43-146) replacement: lowest `nev` eigenpairs of
 * the current sector by thick-restart Lanczos with full reorthogonalisation
 * on the device (which="SA"); ncv Krylov vectors, convergence when the Ritz
 * residual <= tol*max(|theta|, eps^(2/3)), at most maxitr restarts.  evals
 * (host, nev) ascending; evecs (host, n*nev column-major) may be NULL.
 * Special cases of Storage::diag: n == 1 -> (diag, [1]).  After maxitr
 * restarts the converged pairs are returned and *nconv counts them (ARPACK
 * info = 1). */
int edh_eigsh(edh_ctx* ctx, int nev, int ncv, double tol, int maxitr, uint64_t seed, double* evals,
              double* evecs, int* nconv, int* n_av);

/* ------------------------------------------------------------------------
 * Batched engine.  A batch holds sectors of any number of models and many
 * Krylov processes on them: eigensolver runs (edh_eigsh's algorithm) and
 * Green's-function Lanczos runs (edh_lanczos's recurrence).  edh_batch_step
 * advances every active process by one Krylov step per sweep, each phase of
 * the sweep one grouped kernel launch over all processes; processes may be
 * submitted between sweeps (continuous batching).  A process's arithmetic
 * depends only on its own operands, so its result is the same alone and in
 * any mix.  These entries stand in for the per-sector loops of
 * Hamiltonian::diag (Hamiltonian.h:54-101) and GreensFunction::compute
 * (GreensFunction.h:109-159), which run one sector / one start vector at a
 * time in the reference.
 * ---------------------------------------------------------------------- */
typedef struct edh_batch edh_batch;

#define EDH_JOB_ACTIVE 0
#define EDH_JOB_DONE 1
#define EDH_JOB_FAILED 2
#define EDH_JOB_EIG 0
#define EDH_JOB_LANCZOS 1

int edh_batch_create(edh_batch** out, int device);
int edh_batch_destroy(edh_batch* b);
/* Builds the tables of sector (nup, ndn) of model m on the device
 * (SpinResolvedStorage::fill for that sector); returns its id and dimension. */
int edh_batch_add_sector(edh_batch* b, const edh_model* m, int nup, int ndn, int* sec_id, int64_t* dim);
/* edh_batch_add_sector for `count` sectors (nups[q], ndns[q]) of one model:
 * the tables are built on several host threads, so admitting an impurity
 * solve's sectors costs less host time. */
int edh_batch_add_sectors(edh_batch* b, const edh_model* m, int count, const int* nups, const int* ndns, int* sec_ids,
                          int64_t* dims);
int edh_batch_remove_sector(edh_batch* b, int sec_id);
/* Eigensolver run on a sector (Storage::diag, Storage.h:43-146, as edh_eigsh);
 * trivial sectors (n <= 64 or ncv >= n) are solved at submission. */
int edh_batch_submit_eig(edh_batch* b, int sec_id, int nev, int ncv, double tol, int maxitr, uint64_t seed,
                         int* job_id);
/* Lanczos run (Lanczos::lanczos, Lanczos.h:36-77) from the normalised v0. */
int edh_batch_submit_lanczos(edh_batch* b, int sec_id, const double* v0, int nl, int mem, int* job_id);
/* Green's-function start vectors, `count` of them: vector r is
 * c(+)_{orbs[r]} |psi> of eigenpair indices[r] of the finished eigensolver
 * job eig_jobs[r], in sector dst_secs[r] (annihilate[r] selects c), and a
 * Lanczos run of nl steps from out/|out| (GreensFunction.h:395-416,437-458
 * for one orbital each); job_ids[r] is its id, or -1 when the target sector
 * is empty.  Nothing waits for the device: a run whose start vector has
 * <out|out> <= 1e-10 stops before its first step (its continued fraction
 * is not used), and <out|out> is known once the run has finished
 * (edh_batch_lanczos_norm2). */
int edh_batch_submit_gf(edh_batch* b, int count, const int* eig_jobs, const int* indices, const int* orbs,
                        const int* annihilate, const int* dst_secs, int nl, int* job_ids);
/* edh_batch_submit_gf with a second orbital per request: orbs2[q] >= 0 (same
 * spin as orbs[q], a different orbital) gives the start vector
 * (c(+)_orbs[q] + c(+)_orbs2[q])|psi> of the non-local Green's function
 * (GreensFunction::nonlocal_contribution, GreensFunction.h:290-327, with
 * create_particles / annihilate_particles over two orbitals, :389-462);
 * orbs2[q] < 0 is the one-orbital request. */
int edh_batch_submit_gf2(edh_batch* b, int count, const int* eig_jobs, const int* indices, const int* orbs,
                         const int* orbs2, const int* annihilate, const int* dst_secs, int nl, int* job_ids);
/* Squared start-vector norms <out|out> of finished Green's-function runs
 * (NaN while a run has not finished; 1 for runs from
 * edh_batch_submit_lanczos, whose start vector is normalised). */
int edh_batch_lanczos_norm2(edh_batch* b, int count, const int* job_ids, double* norm2);
/* Runs up to nsweeps sweeps (stops early when nothing is active). */
int edh_batch_step(edh_batch* b, int nsweeps, int* n_active);
/* state EDH_JOB_*, kind EDH_JOB_EIG/LANCZOS, dimension, count = nev or
 * Lanczos steps, n_av = H*v applications. */
int edh_batch_job_info(edh_batch* b, int job_id, int* state, int* kind, int64_t* n, int* count, int* n_av);
/* states[q] = EDH_JOB_* of job_ids[q], q < count. */
int edh_batch_job_states(edh_batch* b, int count, const int* job_ids, int* states);
/* Upper bounds of the lowest eigenvalue of eigensolver jobs: the lowest
 * eigenvalue of a finished job, the lowest Ritz value of its last restart
 * cycle for a running one, NaN before its first cycle ends.  (No
 * counterpart in the reference; used to start Green's-function runs early
 * only for pairs that can pass the Boltzmann cutoff.) */
int edh_batch_eig_bounds(edh_batch* b, int count, const int* job_ids, double* bounds);
/* evals (count of edh_batch_job_info, ascending); evecs (host, n*count
 * column-major) may be NULL.  After maxitr restarts only the converged pairs
 * are returned (nconv < the requested nev), as Storage::diag does. */
int edh_batch_eig_result(edh_batch* b, int job_id, double* evals, double* evecs, int* nconv);
/* edh_static_observables for eigenpair `index` of a finished eigensolver
 * job (its vector stays on the device). */
int edh_batch_static_observables(edh_batch* b, int job_id, int index, int norb, double* out);
/* edh_batch_static_observables for `count` eigenpairs (job_ids[q],
 * indices[q]) with one wait for the device; out holds count blocks of
 * 5*norb + norb*norb + 1 values in the same layout. */
int edh_batch_static_observables_many(edh_batch* b, int count, const int* job_ids, const int* indices, int norb,
                                      double* out);
/* alpha[nlanc], beta[nl+1] (beta[0] unused), as edh_lanczos. */
int edh_batch_lanczos_result(edh_batch* b, int job_id, double* alpha, double* beta, int* nlanc);
/* Stops a process (it takes no further sweeps; its state becomes
 * EDH_JOB_FAILED) so that it can be freed. */
int edh_batch_cancel(edh_batch* b, int job_id);
int edh_batch_free_job(edh_batch* b, int job_id);
/* Device bytes held by the batch's sectors and processes. */
int edh_batch_memory(edh_batch* b, size_t* bytes);
/* Steady sweeps -- the same descriptors as the sweep before (eigensolver
 * runs only, nothing started or finished) -- replay one hipGraph of the
 * sweep's kernels without re-uploading descriptors (on by default; 0 turns
 * it off).  edh_batch_sweeps counts the sweeps and those run as replays. */
int edh_batch_set_graphs(edh_batch* b, int on);
int edh_batch_sweeps(edh_batch* b, int64_t* sweeps, int64_t* graph_sweeps);
/* Autotuner: times the grouped H*v over the given sectors for each block
 * size and keeps the fastest (returned in *block). */
int edh_batch_autotune(edh_batch* b, int nsec, const int* sec_ids, int* block);
/* Measurement: mean time of the grouped H*v launch of an eigensolver sweep,
 * w = H v over one operand per listed sector (seeded uniform v), over `reps`
 * launches timed by HIP events on the batch's stream (graph-replayed when
 * graphs are on); *bytes = its algorithmic bytes, the sum over the sectors
 * of 16 n + the per-spin table bytes; *elements = the sum of n.  No
 * counterpart in the reference (its sector loop applies av() one sector at a
 * time, Hamiltonian.h:54-101). */
int edh_batch_time_hv(edh_batch* b, int nsec, const int* sec_ids, int reps, uint64_t seed, double* avg_ms,
                      int64_t* bytes, int64_t* elements);
/* Continued fractions of many finished Lanczos runs in one launch
 * (Lanczos::compute_continued_fraction, Lanczos.h:82-100, for each): run j
 * at z = excited[j]*isign[j] + i*wn[q], q < nz, weight ev[j]*expb[j];
 * out_re/out_im are njobs x nz row-major. */
int edh_batch_cf_many(edh_batch* b, int njobs, const int* job_ids, const double* ev, const int* isign,
                      const double* expb, const double* excited, int nz, const double* wn, double* out_re,
                      double* out_im);
/* edh_batch_cf_many at arbitrary points: run j at z = zre[q] + excited[j] *
 * isign[j] + i zim[q] -- the real-frequency mesh's w_q + i pi/beta
 * (CommonUtils.h:18-20, MeshFactory.h:27-34) or any other contour. */
int edh_batch_cf_many_z(edh_batch* b, int njobs, const int* job_ids, const double* ev, const int* isign,
                        const double* expb, const double* excited, int nz, const double* zre, const double* zim,
                        double* out_re, double* out_im);
/* Lanczos::get_frac_point on the batch's stream (as edh_cf_eval). */
int edh_batch_cf_eval(edh_batch* b, int nz, const double* zre, const double* zim, double ev, int nlanc, int isign,
                      double expb, const double* alpha, const double* beta, double* out_re, double* out_im);

#ifdef __cplusplus
}
#endif

#endif /* EDHIP_H */
