// Launchers of the vector, Gram-Schmidt and Green's-function kernels.
#pragma once

#include "edh_internal.h"

namespace edh {

constexpr int kVecBlock = 256;      // threads per block of the vector kernels
constexpr int kMaxPartials = 2048;  // blocks of a grid-stride reduction
constexpr int kMaxBasis = 48;       // largest Krylov basis handled on the device

int reduce_grid(int64_t n);
void dev_dot(const double* a, const double* b, int64_t n, double* partial, double* out, hipStream_t s);
void dev_norm(const double* a, int64_t n, double* partial, double* out, hipStream_t s);
void dev_scale_inv(const double* x, const double* nrm, double* out, int64_t n, hipStream_t s);
void dev_fill(double* p, int64_t n, double x, hipStream_t s);  // a kernel, never a memset (graph-safe)
// Random start vector: uniform [-1, 1) from splitmix64 (seed 0 -> a fixed
// default), the same values as the sequential host generator.
void dev_random_uniform(double* x, int64_t n, uint64_t seed, hipStream_t s);
// w <- w - V V^T w (one classical Gram-Schmidt pass over m columns); h = V^T w
// (device, m), hsum += h when non-null; *nrm = |w| afterwards (device scalar).
// nb: also *nb = |w| before the pass; gate: the pass does nothing when *gate
// is 0; gate_out: *gate_out = 1 when |w| after < 0.717 |w| before (DGKS).
void dev_orthogonalize(const double* V, int64_t ldv, int m, double* w, int64_t n, double* partial, double* h,
                       double* hsum, double* nrm, hipStream_t s, double* nb = nullptr, const double* gate = nullptr,
                       double* gate_out = nullptr);
// Y[:, j] = sum_i C[i + j*m] V[:, i]  (C on the device, column-major m x k)
void dev_combine(const double* V, int64_t ldv, int m, const double* C, int k, double* Y, int64_t ldy, int64_t n,
                 hipStream_t s);
void dev_lanczos_init(double* sc, int* st, hipStream_t s);
void dev_lanczos_alpha(const double* partial, int n, double* alpha, int k, double* sc, const int* st, hipStream_t s);
void dev_lanczos_passB(double* w, const double* x, const double* coef, int64_t n, double* partial, const int* done,
                       double* beta, int k, int nl, double* sc, int* st, hipStream_t s);

// Green's-function helpers (gf.hip)
void dev_separable_gather(const double* in, int64_t ldin, const int32_t* mu, const int32_t* md, const int8_t* su,
                          const int8_t* sd, double* out, int64_t Dup, int64_t Ddn, hipStream_t s,
                          bool accumulate = false);
void dev_separable_scale(const double* in, const double* fu, const double* fd, double* out, int64_t Dup, int64_t Ddn,
                         hipStream_t s);
// One continued fraction of a batch (device pointers into a packed array).
struct CfJob {
  double ev, expb, excited;
  int isign, nlanc;
  const double* alpha;
  const double* beta;
};
// Run j at point q: z = (wr ? wr[q] : 0) + excited_j * isign_j + i wn[q].
void dev_cf_batch(int njobs, int nz, const double* wr, const double* wn, const CfJob* jobs, double* out_re,
                  double* out_im, hipStream_t s);
void dev_cf_eval(int nz, const double* zre, const double* zim, double ev, int nlanc, int isign, double expb,
                 const double* alpha, const double* beta, double* out_re, double* out_im, hipStream_t s);

// Symmetric eigen-decomposition on the host (eig.hip): A (n x n, row-major)
// is overwritten by the eigenvectors (columns), w by ascending eigenvalues.
void sym_eig(int n, std::vector<double>& A, std::vector<double>& w);

}  // namespace edh
