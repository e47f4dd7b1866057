// Eigensolver replacing ARPACK's dsaupd/dseupd in Storage::diag
// (Storage.h:43-146): thick-restart Lanczos with full reorthogonalisation
// (classical Gram-Schmidt, a second pass when the DGKS test of ARPACK's
// dsaitr asks for it), Krylov basis resident on the device, the projected
// matrix diagonalised on the host.  Parameters follow the
// reference: ncv = min(NCV, ntot), nev = min(NEV, ncv-1), which = "SA",
// convergence when the Ritz residual bound <= tol * max(|theta|, eps^(2/3)).
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <numeric>

#include "edh_ctx.h"

namespace edh {

// ---------------------------------------------------------------------------
// Dense symmetric eigen-decomposition: Householder reduction to tridiagonal
// form with accumulated reflections, then implicit QL sweeps with Wilkinson
// shifts (Golub & Van Loan, "Matrix Computations", sections 8.3 and 8.3.5).
// ---------------------------------------------------------------------------
void sym_eig(int n, std::vector<double>& A, std::vector<double>& w) {
  std::vector<double> Q(n * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i) Q[i * (size_t)n + i] = 1.0;
  auto a = [&](int i, int j) -> double& { return A[i * (size_t)n + j]; };
  auto q = [&](int i, int j) -> double& { return Q[i * (size_t)n + j]; };
  std::vector<double> v(n), p(n), qq(n), d(n), e(n, 0.0);
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;  // length of the column below the diagonal
    double nx = 0.0;
    for (int i = 0; i < m; ++i) nx += a(k + 1 + i, k) * a(k + 1 + i, k);
    nx = std::sqrt(nx);
    if (nx == 0.0) continue;
    const double x0 = a(k + 1, k);
    const double alpha = x0 > 0 ? -nx : nx;
    for (int i = 0; i < m; ++i) v[i] = a(k + 1 + i, k);
    v[0] -= alpha;
    double nv = 0.0;
    for (int i = 0; i < m; ++i) nv += v[i] * v[i];
    nv = std::sqrt(nv);
    if (nv == 0.0) continue;
    for (int i = 0; i < m; ++i) v[i] /= nv;
    // p = A_sub v ; K = v^T p ; qq = p - K v ; A_sub -= 2 (v qq^T + qq v^T)
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      for (int j = 0; j < m; ++j) s += a(k + 1 + i, k + 1 + j) * v[j];
      p[i] = s;
    }
    double K = 0.0;
    for (int i = 0; i < m; ++i) K += v[i] * p[i];
    for (int i = 0; i < m; ++i) qq[i] = p[i] - K * v[i];
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) a(k + 1 + i, k + 1 + j) -= 2.0 * (v[i] * qq[j] + qq[i] * v[j]);
    a(k + 1, k) = alpha;
    a(k, k + 1) = alpha;
    for (int i = 1; i < m; ++i) {
      a(k + 1 + i, k) = 0.0;
      a(k, k + 1 + i) = 0.0;
    }
    // Q <- Q H on columns k+1..n-1
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int j = 0; j < m; ++j) s += q(r, k + 1 + j) * v[j];
      for (int j = 0; j < m; ++j) q(r, k + 1 + j) -= 2.0 * s * v[j];
    }
  }
  for (int i = 0; i < n; ++i) d[i] = a(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);  // e[i] couples i and i+1
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      EDH_CHECK(++iter <= 200, EDH_ERR_NOCONV, "tridiagonal QL did not converge");
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i = m - 1;
      bool early = false;
      for (; i >= l; --i) {
        double f = s * e[i];
        const double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= pp;
          e[m] = 0.0;
          early = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        for (int k = 0; k < n; ++k) {
          f = q(k, i + 1);
          q(k, i + 1) = s * q(k, i) + c * f;
          q(k, i) = c * q(k, i) - s * f;
        }
      }
      if (early) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
  w.resize(n);
  for (int j = 0; j < n; ++j) {
    w[j] = d[order[j]];
    for (int i = 0; i < n; ++i) A[i * (size_t)n + j] = Q[i * (size_t)n + order[j]];
  }
}

double host_diag(const Model& m, const SectorTables& st, int64_t u, int64_t d) {
  const int Ns = m.Ns;
  const int32_t bu = st.up.bits[u], bd = st.dn.bits[d];
  double x = st.up.ediag[u] + st.dn.ediag[d];
  for (int i = 0; i < Ns; ++i)
    for (int j = 0; j < Ns; ++j)
      if (((bu >> (Ns - 1 - i)) & 1) && ((bd >> (Ns - 1 - j)) & 1)) x += m.C[i * Ns + j];
  return x;
}

// Dense sector matrix from the host tables, row = source state (the
// storages' convention: w[i] += H[i][j] v[j]).
void dense_sector_matrix(const Model& m, const SectorTables& st, std::vector<double>& H) {
  const int64_t Du = st.up.D, Dd = st.dn.D, n = st.N;
  H.assign(n * n, 0.0);
  for (int64_t u = 0; u < Du; ++u)
    for (int64_t d = 0; d < Dd; ++d) {
      const int64_t i = u * Dd + d;
      H[i * n + i] += host_diag(m, st, u, d);
      for (int q = st.dn.rowp[d]; q < st.dn.rowp[d + 1]; ++q) H[i * n + u * Dd + st.dn.col[q]] += st.dn.val[q];
      for (int q = st.up.rowp[u]; q < st.up.rowp[u + 1]; ++q) H[i * n + st.up.col[q] * Dd + d] += st.up.val[q];
      if (st.kan_ml > 0) {
        const int nb = m.Ns - st.kan_ml, na = 1 << st.kan_ml;
        const int32_t bu = st.up.bits[u], bd = st.dn.bits[d];
        const int a = bu >> nb, b = bd >> nb;
        const int32_t low = (1 << nb) - 1;
        for (int q = st.kan_ptr[a * na + b]; q < st.kan_ptr[a * na + b + 1]; ++q) {
          const int32_t u2 = st.up.rank[(bu & low) | (st.kan[q].new_a << nb)];
          const int32_t d2 = st.dn.rank[(bd & low) | (st.kan[q].new_b << nb)];
          H[i * n + (int64_t)u2 * Dd + d2] += st.kan[q].coef;
        }
      }
    }
}

// ARPACK's dsaup2 rule: keep the nev wanted Ritz vectors plus one more per
// converged pair, at most half of the new space, and when that leaves a
// single vector, half the basis (m >= 6) or two (m > 2) -- dsaup2's guard
// against stagnation for one wanted eigenvalue (this is what its implicit
// restart with exact shifts retains).
int restart_size(int nev, int m, int nconv) {
  int kk = nev + std::min(nconv, (m - nev) / 2);
  if (kk == 1) kk = m >= 6 ? m / 2 : (m > 2 ? 2 : 1);
  return std::max(1, std::min(kk, m - 2));
}

void eigsh_run(Ctx& c, int nev_req, int ncv_req, double tol, int maxitr, uint64_t seed, double* evals,
               double* evecs_host, int* nconv, int* n_av) {
  EDH_CHECK(c.has_sector, EDH_ERR_STATE, "eigsh: no sector set");
  const int64_t n = c.tabs.N;
  hipStream_t s = c.stream;
  *n_av = 0;
  *nconv = 0;
  c.n_evecs = 0;
  if (n == 0) return;
  if (n == 1) {  // Storage.h:49-52, zero_eigenapair (SpinResolvedStorage.h:52-56)
    evals[0] = host_diag(*c.model, c.tabs, 0, 0);
    if (evecs_host) evecs_host[0] = 1.0;
    c.evecs.ensure(sizeof(double));
    double one = 1.0;
    EDH_HIP(hipMemcpyAsync(c.evecs.p, &one, sizeof(double), hipMemcpyHostToDevice, s));
    EDH_HIP(hipStreamSynchronize(s));
    c.n_evecs = 1;
    *nconv = 1;
    return;
  }
  const int ncv = (int)std::min<int64_t>(ncv_req, n);
  const int nev = std::min(nev_req, ncv - 1);
  EDH_CHECK(nev >= 1, EDH_ERR_ARG, "eigsh: nev must be >= 1");
  if (ncv >= n || n <= 64) {
    // The Krylov space is the whole sector: exact dense solution.
    std::vector<double> H, w;
    dense_sector_matrix(*c.model, c.tabs, H);
    std::vector<double> Hs(H.size());
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < n; ++j) Hs[i * n + j] = 0.5 * (H[i * n + j] + H[j * n + i]);
    sym_eig((int)n, Hs, w);
    std::vector<double> X(n * (size_t)nev);
    for (int j = 0; j < nev; ++j) {
      evals[j] = w[j];
      for (int64_t i = 0; i < n; ++i) X[j * n + i] = Hs[i * n + j];
    }
    if (evecs_host) std::copy(X.begin(), X.end(), evecs_host);
    c.evecs.ensure(X.size() * sizeof(double));
    EDH_HIP(hipMemcpyAsync(c.evecs.p, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice, s));
    EDH_HIP(hipStreamSynchronize(s));
    c.n_evecs = nev;
    *nconv = nev;
    return;
  }
  const int m = std::min(ncv, kMaxBasis - 1);
  EDH_CHECK(nev < m - 1, EDH_ERR_LIMIT, "eigsh: NEV too large for the device Krylov basis (max 46)");
  const size_t vbytes = (size_t)n * sizeof(double);
  c.krylov[0].ensure((size_t)(m + 1) * vbytes);
  c.krylov[1].ensure((size_t)(m + 1) * vbytes);
  c.partial.ensure((size_t)kMaxPartials * kMaxBasis * sizeof(double));
  // device scalars: hsum columns (m x (m+1)), norms (m+1), h scratch
  c.small.ensure((size_t)(m * (m + 1) + (m + 1) + kMaxBasis + kMaxBasis * kMaxBasis + 2) * sizeof(double));
  double* Tcol = c.small.as<double>();             // column j at Tcol + j*(m+1)
  double* norms = Tcol + (size_t)m * (m + 1);      // norm after orthogonalising column j
  double* hs = norms + (m + 1);                    // scratch h
  double* Cdev = hs + kMaxBasis;                   // combine coefficients
  double* dgks = Cdev + kMaxBasis * kMaxBasis;     // norm before the first pass, second-pass gate
  double* V = c.krylov[0].as<double>();
  double* Valt = c.krylov[1].as<double>();
  double* partial = c.partial.as<double>();

  // random start vector, uniform in [-1, 1) (splitmix64, generated on the device)
  {
    dev_random_uniform(V, n, seed, s);
    dev_norm(V, n, partial, norms + m, s);
    dev_scale_inv(V, norms + m, V, n, s);
  }
  std::vector<double> T((size_t)m * m, 0.0), Tc((size_t)m * (m + 1)), nr(m + 1), theta, Y;
  int k = 0;  // kept Ritz vectors at the start of the cycle
  const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
  for (int iter = 0;; ++iter) {
    EDH_HIP(hipMemsetAsync(Tcol, 0, (size_t)m * (m + 1) * sizeof(double), s));
    for (int j = k; j < m; ++j) {
      double* vj = V + (size_t)j * n;
      double* w = V + (size_t)(j + 1) * n;
      HvArgs a{};
      a.v = vj;
      a.w = w;
      a.vscale = 1.0;
      a.mode = 0;
      launch_hv(c.dev, a, s);
      ++*n_av;
      double* hcol = Tcol + (size_t)j * (m + 1);
      // classical Gram-Schmidt, repeated only when the DGKS test asks for it
      dev_orthogonalize(V, n, j + 1, w, n, partial, hs, hcol, norms + j, s, dgks, nullptr, dgks + 1);
      dev_orthogonalize(V, n, j + 1, w, n, partial, hs, hcol, norms + j, s, nullptr, dgks + 1, nullptr);
      dev_scale_inv(w, norms + j, w, n, s);
    }
    EDH_HIP(hipMemcpyAsync(Tc.data(), Tcol, Tc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    EDH_HIP(hipMemcpyAsync(nr.data(), norms, (m + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
    EDH_HIP(hipStreamSynchronize(s));
    for (int j = k; j < m; ++j)
      for (int i = 0; i <= j; ++i) {
        T[(size_t)i * m + j] = Tc[(size_t)j * (m + 1) + i];
        T[(size_t)j * m + i] = Tc[(size_t)j * (m + 1) + i];
      }
    const double bres = nr[m - 1];
    std::vector<double> A = T;
    sym_eig(m, A, theta);
    Y = A;  // columns = eigenvectors of T
    std::vector<int> done_i;
    for (int i = 0; i < nev; ++i) {
      const double bound = std::fabs(bres * Y[(size_t)(m - 1) * m + i]);
      if (bound <= tol * std::max(std::fabs(theta[i]), eps23)) done_i.push_back(i);
    }
    const int nconv_now = (int)done_i.size();
    const bool conv = nconv_now == nev;
    if (conv || iter + 1 >= maxitr) {
      // Like dsaupd/dseupd after maxitr (info = 1, Storage.h:94-129): the
      // converged Ritz pairs are returned and nconv says how many there are.
      const int nk = nconv_now;
      std::vector<double> Cm((size_t)m * std::max(nk, 1));
      for (int q = 0; q < nk; ++q)
        for (int i = 0; i < m; ++i) Cm[(size_t)q * m + i] = Y[(size_t)i * m + done_i[q]];
      if (nk > 0) {
        EDH_HIP(hipMemcpyAsync(Cdev, Cm.data(), (size_t)m * nk * sizeof(double), hipMemcpyHostToDevice, s));
        c.evecs.ensure((size_t)nk * vbytes);
        dev_combine(V, n, m, Cdev, nk, c.evecs.as<double>(), n, n, s);
        if (evecs_host)
          EDH_HIP(hipMemcpyAsync(evecs_host, c.evecs.p, (size_t)nk * vbytes, hipMemcpyDeviceToHost, s));
      }
      EDH_HIP(hipStreamSynchronize(s));
      for (int q = 0; q < nk; ++q) evals[q] = theta[done_i[q]];
      c.n_evecs = nk;
      *nconv = nk;
      return;
    }
    // thick restart: keep kk lowest Ritz vectors, continue from the residual
    const int kk = restart_size(nev, m, nconv_now);
    std::vector<double> Cm((size_t)m * kk);
    for (int j = 0; j < kk; ++j)
      for (int i = 0; i < m; ++i) Cm[(size_t)j * m + i] = Y[(size_t)i * m + j];
    EDH_HIP(hipMemcpyAsync(Cdev, Cm.data(), Cm.size() * sizeof(double), hipMemcpyHostToDevice, s));
    dev_combine(V, n, m, Cdev, kk, Valt, n, n, s);
    EDH_HIP(hipMemcpyAsync(Valt + (size_t)kk * n, V + (size_t)m * n, vbytes, hipMemcpyDeviceToDevice, s));
    std::swap(V, Valt);
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < kk; ++i) {
      T[(size_t)i * m + i] = theta[i];
      const double si = bres * Y[(size_t)(m - 1) * m + i];
      T[(size_t)i * m + kk] = si;
      T[(size_t)kk * m + i] = si;
    }
    k = kk;
  }
}

}  // namespace edh
