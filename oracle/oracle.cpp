// CPU oracle for the spin-resolved exact-diagonalisation hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU and in plain C++,
// the algorithm of the reference EDLib library (Q-solvers/EDLib) for the
// spin-resolved H*v path: basis enumeration and ranking, the Hubbard and
// single-impurity Anderson (Kanamori) model terms, the per-spin CSR hopping
// tables with CRSMatrix merge semantics, SpinResolvedStorage::fill/av/a_adag,
// the SOCRSStorage algorithm (OpenMP, timed as the CPU baseline), the plain
// Lanczos recurrence and the continued-fraction evaluation.  Every function
// cites the reference file:line it follows.  Only tests/, the smoke check and
// the bench's cpu_baseline leg may load this library; the product never does.
//
// Pinning (see tests/test_oracle_*.py): golden vectors gom1.dat / xiats.dat /
// xiatd.dat from test/input/GF_Chi, the 4-site ring ground-state energy from
// test/HubbardModel_Test.cpp:54, closed-form ranks, brute-force dense
// Hamiltonians (numpy) and an independent ARPACK run (scipy eigsh).

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef long long state_t;

// ---------------------------------------------------------------------------
// Binomials and the lexicographic combination basis.
// Combination.h:23-50 (c_n_k table, init_state, next_combination),
// SzSymmetry.h:191-216 (basis[k][r] and ninv tables).
// ---------------------------------------------------------------------------
struct Binom {
  int n;
  std::vector<int64_t> t;  // (n+1)x(n+1), C(a,b) with C(a,b)=0 for b>a here
  explicit Binom(int n_) : n(n_), t((n_ + 1) * (n_ + 1), 0) {
    for (int a = 0; a <= n; ++a) {
      t[a * (n + 1)] = 1;
      for (int b = 1; b <= a; ++b)
        t[a * (n + 1) + b] = t[(a - 1) * (n + 1) + b - 1] + (b <= a - 1 ? t[(a - 1) * (n + 1) + b] : 0);
    }
  }
  int64_t operator()(int a, int b) const {
    if (a < 0 || b < 0 || b > a) return 0;
    return t[a * (n + 1) + b];
  }
};

// Lexicographic successor of a sorted position tuple, Combination.h:33-45.
bool next_combination(int n, int k, std::vector<int>& pos) {
  for (int i = k - 1; i >= 0; --i) {
    if (pos[i] < n - k + i) {
      ++pos[i];
      for (int j = i + 1; j < k; ++j) pos[j] = pos[j - 1] + 1;
      return true;
    }
  }
  return false;
}

// All k-subsets of n bit positions in the reference order (SzSymmetry.h:191-216).
std::vector<int> enumerate_basis(int n, int k) {
  std::vector<int> out;
  std::vector<int> pos(k);
  for (int i = 0; i < k; ++i) pos[i] = i;
  while (true) {
    int bits = 0;
    for (int i = 0; i < k; ++i) bits |= 1 << pos[i];
    out.push_back(bits);
    if (!next_combination(n, k, pos)) break;
  }
  return out;
}

// NSymmetry::num / index recursion, NSymmetry.h:94-96,155-162 (N = sites per spin).
int nsym_num(const Binom& C, int N, state_t b, int n, int m) {
  int res = 0;
  if (((b & (1ll << (N - n))) == 0) && ((n - 1) > 0) && (m > 0) && (m < n))
    res = nsym_num(C, N, b, n - 1, m);
  else if (((n - 1) > 0) && (m > 0) && (m < n))
    res = (int)C(n - 1, m) + nsym_num(C, N, b, n - 1, m - 1);
  return res;
}
int nsym_index(const Binom& C, int N, int nel, state_t st) {
  return (int)C(N, nel) - nsym_num(C, N, st, N, nel) - 1;
}

// ---------------------------------------------------------------------------
// Fermionic helpers, FermionicModel.h:35-68.  Orbital o (= site + spin*Ns)
// sits at bit (Ip-1-o); the sign counts occupied orbitals with index < o.
// ---------------------------------------------------------------------------
inline int check_state(state_t nst, int im, int Ip) { return (int)((nst >> (Ip - 1 - im)) & 1ll); }
inline void op_a(int i, state_t jold, state_t& k, int& isign, int Ip) {
  long long cnt = 0;
  for (int ll = 0; ll < i; ++ll) cnt += ((jold >> (Ip - ll - 1)) & 1ll);
  isign = (cnt % 2) == 0 ? 1 : -1;
  k = jold - (1ll << (Ip - i - 1));
}
inline void op_adag(int i, state_t jold, state_t& k, int& isign, int Ip) {
  long long cnt = 0;
  for (int ll = 0; ll < i; ++ll) cnt += ((jold >> (Ip - ll - 1)) & 1ll);
  isign = (cnt % 2) == 0 ? 1 : -1;
  k = jold + (1ll << (Ip - i - 1));
}

// ---------------------------------------------------------------------------
// Models.
// ---------------------------------------------------------------------------
struct Transition {
  // Hopping / hybridisation (kind 0): c^+_{j s} c_{i s}; i=a, j=b.
  // Off-diagonal Coulomb (kind 1): c^+_{i s} c^+_{j s'} c_{l s'} c_{k s}.
  int kind;
  int i, j, k, l, s, sp;
  double value;
};

struct Model {
  int kind;  // 0 Hubbard, 1 SIAM
  int Ns, ms, Ip;
  // Hubbard (HubbardModel.h:48-77)
  std::vector<double> t, U, J, xmu, Hmag, Eps;  // Eps Ns x ms
  // SIAM (SingleImpurityAndersonModel.h:195-275)
  int ml = 0;
  std::vector<double> Epsk;  // (Ns-ml) x 2
  std::vector<double> Vk;    // ml x (Ns-ml) x 2
  std::vector<double> H0;    // ml x ml x 2
  std::vector<double> Ut;    // 2 x 2 x ml^4
  double mu = 0.0;
  std::vector<Transition> T, V;

  double Uten(int s1, int s2, int i, int j, int k, int l) const {
    int m = ml;
    return Ut[((((s1 * 2 + s2) * m + i) * m + j) * m + k) * m + l];
  }

  // HubbardModel::valid / SIAM InnerHybridizationState::valid (HubbardModel.h:86-88,
  // SingleImpurityAndersonModel.h:74-76) and InnerInteractionState::valid (:129-142).
  int valid(const Transition& tr, state_t nst) const {
    if (tr.kind == 0)
      return check_state(nst, tr.i + tr.s * Ns, Ip) * (1 - check_state(nst, tr.j + tr.s * Ns, Ip));
    if (check_state(nst, tr.k + tr.s * Ns, Ip) != 0) {
      state_t k3 = nst - (1ll << (Ip - 1 - tr.k - tr.s * Ns));
      if (check_state(k3, tr.l + tr.sp * Ns, Ip) != 0) {
        state_t k4 = k3 - (1ll << (Ip - 1 - tr.l - tr.sp * Ns));
        if (check_state(k4, tr.j + tr.sp * Ns, Ip) == 0) {
          state_t k2 = k4 | (1ll << (Ip - 1 - tr.j - tr.sp * Ns));
          return 1 - check_state(k2, tr.i + tr.s * Ns, Ip);
        }
      }
    }
    return 0;
  }
  // HubbardModel::set (HubbardModel.h:99-108, sign = -s1 s2),
  // InnerHybridizationState::set (SingleImpurityAndersonModel.h:77-84, sign = +s1 s2),
  // InnerInteractionState::set (:153-162) and value() = U/2 (:168).
  double set(const Transition& tr, state_t nst, state_t& k, int& sign) const {
    if (tr.kind == 0) {
      state_t k1, k2;
      int s1, s2;
      op_a(tr.i + tr.s * Ns, nst, k1, s1, Ip);
      op_adag(tr.j + tr.s * Ns, k1, k2, s2, Ip);
      k = k2;
      sign = (this->kind == 0) ? -s1 * s2 : s1 * s2;
      return tr.value;
    }
    state_t k1, k2, k3, k4;
    int s1, s2, s3, s4;
    op_a(tr.k + tr.s * Ns, nst, k3, s1, Ip);
    op_a(tr.l + tr.sp * Ns, k3, k4, s2, Ip);
    op_adag(tr.j + tr.sp * Ns, k4, k2, s3, Ip);
    op_adag(tr.i + tr.s * Ns, k2, k1, s4, Ip);
    k = k1;
    sign = s1 * s2 * s3 * s4;
    return 0.5 * tr.value;
  }
  double tvalue(const Transition& tr) const { return tr.kind == 0 ? tr.value : 0.5 * tr.value; }

  // HubbardModel::diagonal, HubbardModel.h:116-132 (note _xmu indexed by spin, :120).
  // SingleImpurityAndersonModel::diagonal, SingleImpurityAndersonModel.h:282-306.
  double diagonal(state_t st) const {
    double x = 0.0;
    if (kind == 0) {
      for (int im = 0; im < Ns; ++im) {
        for (int is = 0; is < ms; ++is) x += (Eps[im * ms + is] - xmu[is]) * check_state(st, im + is * Ns, Ip);
        x += U[im] * check_state(st, im, Ip) * check_state(st, im + Ns, Ip);
        x += Hmag[im] * (check_state(st, im + Ns, Ip) - check_state(st, im, Ip));
        for (int im2 = 0; im2 < Ns; ++im2)
          x += J[im * Ns + im2] * (check_state(st, im, Ip) - check_state(st, im + Ns, Ip)) *
               (check_state(st, im2, Ip) - check_state(st, im2 + Ns, Ip));
      }
      return x;
    }
    int nb = Ns - ml;
    for (int is = 0; is < ms; ++is)
      for (int ik = 0; ik < nb; ++ik) x += Epsk[ik * 2 + is] * check_state(st, ik + ml + is * Ns, Ip);
    for (int im = 0; im < ml; ++im) {
      for (int is = 0; is < ms; ++is) x += (H0[(im * ml + im) * 2 + is] - mu) * check_state(st, im + is * Ns, Ip);
      for (int is = 0; is < ms; ++is)
        x += 0.5 * Uten(is, is, im, im, im, im) * check_state(st, im, Ip) * check_state(st, im + Ns, Ip);
      for (int jm = 0; jm < ml; ++jm)
        for (int is = 0; is < ms; ++is)
          if (im != jm) {
            x += 0.5 * (Uten(is, is, im, jm, im, jm) - Uten(is, is, im, jm, jm, im)) * check_state(st, im + is * Ns, Ip) *
                 check_state(st, jm + is * Ns, Ip);
            x += 0.5 * (Uten(is, 1 - is, im, jm, im, jm)) * check_state(st, im + is * Ns, Ip) *
                 check_state(st, jm + (1 - is) * Ns, Ip);
          }
    }
    return x;
  }
};

// T_states construction, HubbardModel.h:68-76.
void build_hubbard_transitions(Model& m) {
  for (int ii = 0; ii < m.Ns; ++ii)
    for (int jj = 0; jj < m.Ns; ++jj)
      if (std::abs(m.t[ii * m.Ns + jj]) > 1e-10)
        for (int is = 0; is < m.ms; ++is) m.T.push_back(Transition{0, ii, jj, 0, 0, is, 0, m.t[ii * m.Ns + jj]});
}

// T_states / V_states construction, SingleImpurityAndersonModel.h:232-274.
void build_siam_transitions(Model& m) {
  int ml = m.ml, nb = m.Ns - ml;
  for (int im = 0; im < ml; ++im)
    for (int jm = 0; jm < im; ++jm)
      for (int is = 0; is < m.ms; ++is) {
        double h = m.H0[(im * ml + jm) * 2 + is];
        if (std::abs(h) > 1e-10) {
          m.T.push_back(Transition{0, im, jm, 0, 0, is, 0, h});
          m.T.push_back(Transition{0, jm, im, 0, 0, is, 0, h});
        }
      }
  for (int im = 0; im < ml; ++im)
    for (int ik = 0; ik < nb; ++ik)
      for (int is = 0; is < m.ms; ++is) {
        double v = m.Vk[(im * nb + ik) * 2 + is];
        if (std::abs(v) > 1e-10) {
          int imk = ik + ml;
          m.T.push_back(Transition{0, im, imk, 0, 0, is, 0, v});
          m.T.push_back(Transition{0, imk, im, 0, 0, is, 0, v});
        }
      }
  for (int s1 = 0; s1 < m.ms; ++s1)
    for (int s2 = 0; s2 < m.ms; ++s2)
      for (int i = 0; i < ml; ++i)
        for (int j = 0; j < ml; ++j)
          for (int k = 0; k < ml; ++k)
            for (int l = 0; l < ml; ++l) {
              if (((i == l) && (j == k) && (s1 == s2)) || ((i == k) && (j == l))) continue;
              double u = m.Uten(s1, s2, i, j, k, l);
              if (std::abs(u) != 0.0) m.V.push_back(Transition{1, i, j, k, l, s1, s2, u});
            }
}

// ---------------------------------------------------------------------------
// Small CSR builder with the reference merge semantics, CRSMatrix.h:30-98.
// ---------------------------------------------------------------------------
struct CSR {
  std::vector<double> val;
  std::vector<int> col;
  std::vector<int> rowp;
  int vind = 0;
  void init(size_t nrows) {
    val.clear();
    col.clear();
    rowp.assign(nrows + 1, 0);
    vind = 0;
  }
  void add(int i, int j, double t, int sign) {
    if (std::abs(t) == 0) return;  // CRSMatrix.h:51-53
    bool has = false;
    size_t found = 0;
    for (int q = rowp[i]; q < vind; ++q)
      if (col[q] == j) {
        has = true;
        found = q;
      }
    if (has) {
      val[found] += sign * t;
    } else {
      if ((int)col.size() <= vind) {
        col.resize(vind + 1);
        val.resize(vind + 1);
      }
      col[vind] = j;
      val[vind] = sign * t;
      ++vind;
    }
  }
  // compress + endLine, CRSMatrix.h:82-98 (including the skip of the element
  // that shifts into a removed slot).
  void end_line(int i) {
    for (int q = rowp[i]; q < vind; ++q) {
      if (std::abs(val[q]) < 1e-15) {
        for (int r = q; r < vind - 1; ++r) {
          val[r] = val[r + 1];
          col[r] = col[r + 1];
        }
        --vind;
      }
    }
    rowp[i + 1] = vind;
  }
};

// ---------------------------------------------------------------------------
// SpinResolvedStorage restatement: fill (SpinResolvedStorage.h:120-155,
// 172-250, 622-638) and av (:58-115), non-MPI path.
// ---------------------------------------------------------------------------
struct SRS {
  const Model* m;
  int nup, ndn;
  int64_t Dup, Ddn, N;
  CSR Hup, Hdn, Hloc;
  bool has_loc = false;
  size_t int_start = 0;
  std::vector<double> diag;
  std::vector<int> bup, bdn;  // per-spin basis bits (low Ns bits)
};

void fill_spin(const Model& m, const Binom& C, int nel, int shift, CSR& H) {
  std::vector<int> basis = enumerate_basis(m.Ns, nel);
  H.init(basis.size());
  for (size_t i = 0; i < basis.size(); ++i) {
    state_t nst = basis[i];
    for (const Transition& tr : m.T) {
      if (m.valid(tr, nst << shift)) {
        state_t k;
        int sg;
        double v = m.set(tr, nst << shift, k, sg);
        int j = nsym_index(C, m.Ns, nel, k >> shift);
        H.add((int)i, j, v, sg);
      }
    }
    H.end_line((int)i);
  }
}

SRS* srs_fill(const Model& m, int nup, int ndn) {
  SRS* s = new SRS();
  s->m = &m;
  s->nup = nup;
  s->ndn = ndn;
  Binom C(m.Ns);
  s->Dup = C(m.Ns, nup);
  s->Ddn = C(m.Ns, ndn);
  s->N = s->Dup * s->Ddn;
  s->bup = enumerate_basis(m.Ns, nup);
  s->bdn = enumerate_basis(m.Ns, ndn);
  fill_spin(m, C, nup, m.Ns, s->Hup);
  fill_spin(m, C, ndn, 0, s->Hdn);
  s->diag.assign(s->N, 0.0);
  const int64_t Ddn = s->Ddn;
  const std::vector<int>& bu = s->bup;
  const std::vector<int>& bd = s->bdn;
  // The diagonal loop is independent per state; it is parallel here only to keep
  // large sectors tractable (the values are identical to the serial loop).
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->N; ++i) {
    state_t nst = ((state_t)bu[i / Ddn] << m.Ns) + bd[i % Ddn];
    s->diag[i] = m.diagonal(nst);
  }
  s->int_start = s->N;
  if (!m.V.empty()) {
    s->has_loc = true;
    s->Hloc.init(s->N);
    // rank lookup for the target state (SzSymmetry::index, SzSymmetry.h:123-130)
    std::vector<int> ninv_u(1 << m.Ns, -1), ninv_d(1 << m.Ns, -1);
    for (size_t r = 0; r < bu.size(); ++r) ninv_u[bu[r]] = (int)r;
    for (size_t r = 0; r < bd.size(); ++r) ninv_d[bd[r]] = (int)r;
    for (int64_t i = 0; i < s->N; ++i) {
      state_t nst = ((state_t)bu[i / Ddn] << m.Ns) + bd[i % Ddn];
      for (const Transition& tr : m.V) {
        if (m.valid(tr, nst)) {
          s->int_start = std::min<size_t>(i, s->int_start);
          state_t k;
          int sg;
          m.set(tr, nst, k, sg);
          int j = ninv_u[(int)(k >> m.Ns)] * (int)Ddn + ninv_d[(int)(k & ((1ll << m.Ns) - 1))];
          s->Hloc.add((int)i, j, m.tvalue(tr), sg);
        }
      }
      s->Hloc.end_line((int)i);
    }
  }
  return s;
}

void srs_av(const SRS* s, const double* v, double* w, bool clear) {
  const int64_t n = s->N, Ddn = s->Ddn, Dup = s->Dup;
  for (int64_t i = 0; i < n; ++i) w[i] = s->diag[i] * v[i] + (clear ? 0.0 : w[i]);
  for (int64_t k = 0; k < Dup; ++k)
    for (int64_t i = 0; i < Ddn; ++i)
      for (int j = s->Hdn.rowp[i]; j < s->Hdn.rowp[i + 1]; ++j)
        w[i + k * Ddn] += s->Hdn.val[j] * v[s->Hdn.col[j] + k * Ddn];
  for (int64_t i = 0; i < Dup; ++i)
    for (int j = s->Hup.rowp[i]; j < s->Hup.rowp[i + 1]; ++j)
      for (int64_t k = 0; k < Ddn; ++k) w[i * Ddn + k] += s->Hup.val[j] * v[(int64_t)s->Hup.col[j] * Ddn + k];
  if (s->has_loc)
    for (int64_t i = (int64_t)s->int_start; i < n; ++i)
      for (int j = s->Hloc.rowp[i]; j < s->Hloc.rowp[i + 1]; ++j) w[i] += s->Hloc.val[j] * v[s->Hloc.col[j]];
}

// Same arithmetic, threaded over up-rows (used only as the timed CPU baseline
// where the sign-only CRS storage does not fit in memory).
void srs_av_omp(const SRS* s, const double* v, double* w, bool clear) {
  const int64_t Ddn = s->Ddn, Dup = s->Dup;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < Dup; ++k) {
    double* wr = w + k * Ddn;
    const double* vr = v + k * Ddn;
    const double* dg = s->diag.data() + k * Ddn;
    for (int64_t i = 0; i < Ddn; ++i) wr[i] = dg[i] * vr[i] + (clear ? 0.0 : wr[i]);
    for (int64_t i = 0; i < Ddn; ++i)
      for (int j = s->Hdn.rowp[i]; j < s->Hdn.rowp[i + 1]; ++j) wr[i] += s->Hdn.val[j] * vr[s->Hdn.col[j]];
    for (int j = s->Hup.rowp[k]; j < s->Hup.rowp[k + 1]; ++j) {
      const double t = s->Hup.val[j];
      const double* vs = v + (int64_t)s->Hup.col[j] * Ddn;
      for (int64_t i = 0; i < Ddn; ++i) wr[i] += t * vs[i];
    }
  }
  if (s->has_loc)
    for (int64_t i = (int64_t)s->int_start; i < s->N; ++i)
      for (int j = s->Hloc.rowp[i]; j < s->Hloc.rowp[i + 1]; ++j) w[i] += s->Hloc.val[j] * v[s->Hloc.col[j]];
}

// ---------------------------------------------------------------------------
// SOCRSStorage restatement (SOCRSStorage.h:46-87 av, :114-160 fill,
// :303-326 addElement), OpenMP with static row chunks.  Signs are kept one
// per byte, which is what the reference's bitmap effectively spends
// (sizeof(char) bits per byte).
// ---------------------------------------------------------------------------
struct SOCRS {
  const Model* m;
  int64_t N, Ddn;
  std::vector<int> bu, bd, ninv_u, ninv_d;
  std::vector<int> col;
  std::vector<int8_t> sgn;
  std::vector<double> dvals;
  std::vector<int64_t> row_off, vind_off;
  int nthreads;
};

SOCRS* socrs_fill(const Model& m, int nup, int ndn) {
  SOCRS* s = new SOCRS();
  s->m = &m;
  s->bu = enumerate_basis(m.Ns, nup);
  s->bd = enumerate_basis(m.Ns, ndn);
  s->Ddn = s->bd.size();
  s->N = (int64_t)s->bu.size() * s->Ddn;
  s->ninv_u.assign(1 << m.Ns, 0);
  s->ninv_d.assign(1 << m.Ns, 0);
  for (size_t r = 0; r < s->bu.size(); ++r) s->ninv_u[s->bu[r]] = (int)r;
  for (size_t r = 0; r < s->bd.size(); ++r) s->ninv_d[s->bd[r]] = (int)r;
#ifdef _OPENMP
  s->nthreads = omp_get_max_threads();
#else
  s->nthreads = 1;
#endif
  const size_t per = m.T.size() + m.V.size();
  s->col.assign((size_t)s->N * per, 0);
  s->sgn.assign((size_t)s->N * per, 0);
  s->dvals.assign(s->N, 0.0);
  int nt = s->nthreads;
  s->row_off.assign(nt + 1, 0);
  int64_t step = s->N / nt;
  for (int i = 0; i <= nt; ++i) s->row_off[i] = step * i;
  int64_t more = s->N - s->row_off[nt];
  for (int i = 0; i < more; ++i) s->row_off[i] += i;
  for (int i = (int)more; i <= nt; ++i) s->row_off[i] += more;
  s->vind_off.assign(nt + 1, 0);
  for (int i = 0; i <= nt; ++i) s->vind_off[i] = (int64_t)per * s->row_off[i];
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int id = omp_get_thread_num();
#else
    int id = 0;
#endif
    int64_t vind = s->vind_off[id];
    for (int64_t i = s->row_off[id]; i < s->row_off[id + 1]; ++i) {
      state_t nst = ((state_t)s->bu[i / s->Ddn] << m.Ns) + s->bd[i % s->Ddn];
      s->dvals[i] = m.diagonal(nst);
      for (int pass = 0; pass < 2; ++pass) {
        const std::vector<Transition>& L = pass == 0 ? m.T : m.V;
        for (const Transition& tr : L) {
          if (m.valid(tr, nst)) {
            state_t k;
            int sg;
            m.set(tr, nst, k, sg);
            s->col[vind] = s->ninv_u[(int)(k >> m.Ns)] * (int)s->Ddn + s->ninv_d[(int)(k & ((1ll << m.Ns) - 1))];
            s->sgn[vind] = sg < 0 ? 1 : 0;
            ++vind;
          }
        }
      }
    }
  }
  return s;
}

void socrs_av(const SOCRS* s, const double* v, double* w, bool clear) {
  const Model& m = *s->m;
#pragma omp parallel num_threads(s->nthreads)
  {
#ifdef _OPENMP
    int id = omp_get_thread_num();
#else
    int id = 0;
#endif
    int64_t vind = s->vind_off[id];
    for (int64_t i = s->row_off[id]; i < s->row_off[id + 1]; ++i) {
      state_t nst = ((state_t)s->bu[i / s->Ddn] << m.Ns) + s->bd[i % s->Ddn];
      double acc = s->dvals[i] * v[i] + (clear ? 0.0 : w[i]);
      for (int pass = 0; pass < 2; ++pass) {
        const std::vector<Transition>& L = pass == 0 ? m.T : m.V;
        for (const Transition& tr : L) {
          int test = m.valid(tr, nst);
          if (test) {
            acc += m.tvalue(tr) * (1 - 2 * s->sgn[vind]) * v[s->col[vind]];
            ++vind;
          }
        }
      }
      w[i] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Sector list and order, SzSymmetry.h:51-54 (comparator) and :88-94 (std::sort
// of the i-major list).  The comparator is not a strict weak ordering; the
// resulting order is libstdc++-introsort-specific, so this must be compiled
// against libstdc++ like the reference.
// ---------------------------------------------------------------------------
struct Sec {
  int nup, ndn;
  size_t size;
  bool operator<(const Sec& s) const {
    return size < s.size || (size == s.size && nup < s.nup && ndn < s.ndn) ||
           (size == s.size && nup == s.nup && ndn < s.ndn);
  }
};

}  // namespace

// ===========================================================================
// C ABI for ctypes.
// ===========================================================================
extern "C" {

long long orc_cnk(int n, int k) {
  Binom C(n);
  return C(n, k);
}

int orc_basis(int Ns, int k, int* out) {
  std::vector<int> b = enumerate_basis(Ns, k);
  std::memcpy(out, b.data(), b.size() * sizeof(int));
  return (int)b.size();
}

int orc_nsym_index(int Ns, int nel, long long st) {
  Binom C(Ns);
  return nsym_index(C, Ns, nel, st);
}

// Writes the sector order; list==NULL means all (i,j) pairs, i-major.
int orc_sector_order(int Ns, int nlist, const int* list, int* out_nup, int* out_ndn) {
  Binom C(Ns);
  std::vector<Sec> secs;
  if (list) {
    for (int q = 0; q < nlist; ++q)
      secs.push_back(Sec{list[2 * q], list[2 * q + 1], (size_t)(C(Ns, list[2 * q]) * C(Ns, list[2 * q + 1]))});
  } else {
    for (int i = 0; i <= Ns; ++i)
      for (int j = 0; j <= Ns; ++j) secs.push_back(Sec{i, j, (size_t)(C(Ns, i) * C(Ns, j))});
  }
  std::sort(secs.begin(), secs.end(), std::less<Sec>());
  for (size_t q = 0; q < secs.size(); ++q) {
    out_nup[q] = secs[q].nup;
    out_ndn[q] = secs[q].ndn;
  }
  return (int)secs.size();
}

// Eigenpair order: std::set with EigenPair::operator< (EigenPair.h:39-41),
// inserted with consecutive ids as in Hamiltonian::diag (Hamiltonian.h:79-83).
void orc_eigenpair_order(int n, const double* evals, int* out_ids) {
  struct EP {
    double e;
    int id;
    bool operator<(const EP& o) const { return (e < o.e) || (id < o.id); }
  };
  std::set<EP> s;
  for (int i = 0; i < n; ++i) s.insert(EP{evals[i], i});
  int q = 0;
  for (const EP& p : s) out_ids[q++] = p.id;
}

void* orc_model_hubbard(int Ns, const double* t, const double* U, const double* J, const double* xmu,
                        const double* Hmag) {
  Model* m = new Model();
  m->kind = 0;
  m->Ns = Ns;
  m->ms = 2;
  m->Ip = 2 * Ns;
  m->t.assign(t, t + Ns * Ns);
  m->U.assign(U, U + Ns);
  m->J.assign(J, J + Ns * Ns);
  m->xmu.assign(xmu, xmu + Ns);
  m->Hmag.assign(Hmag, Hmag + Ns);
  m->Eps.assign(Ns * 2, 0.0);  // never read from input, HubbardModel.h:49
  build_hubbard_transitions(*m);
  return m;
}

void* orc_model_siam(int Ns, int ml, const double* Epsk, const double* Vk, const double* H0, const double* Ut,
                     double mu) {
  Model* m = new Model();
  m->kind = 1;
  m->Ns = Ns;
  m->ms = 2;
  m->Ip = 2 * Ns;
  m->ml = ml;
  int nb = Ns - ml;
  m->Epsk.assign(Epsk, Epsk + nb * 2);
  m->Vk.assign(Vk, Vk + ml * nb * 2);
  m->H0.assign(H0, H0 + ml * ml * 2);
  m->Ut.assign(Ut, Ut + 4 * ml * ml * ml * ml);
  m->mu = mu;
  build_siam_transitions(*m);
  return m;
}

void orc_model_free(void* m) { delete (Model*)m; }

int orc_model_ntrans(void* mp, int which) {
  Model* m = (Model*)mp;
  return (int)(which == 0 ? m->T.size() : m->V.size());
}

// Transition list as (kind,i,j,k,l,s,sp) + value.
void orc_model_trans(void* mp, int which, int* ints, double* vals) {
  Model* m = (Model*)mp;
  const std::vector<Transition>& L = which == 0 ? m->T : m->V;
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
}

double orc_model_diagonal(void* mp, long long st) { return ((Model*)mp)->diagonal(st); }

// Applies the whole Hamiltonian transition list to one full basis state
// (used by the dense brute force): writes target states, signed values.
int orc_model_apply(void* mp, long long st, long long* targets, double* coef, int cap) {
  Model* m = (Model*)mp;
  int n = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const std::vector<Transition>& L = pass == 0 ? m->T : m->V;
    for (const Transition& tr : L)
      if (m->valid(tr, st)) {
        state_t k;
        int sg;
        double v = m->set(tr, st, k, sg);
        if (n < cap) {
          targets[n] = k;
          coef[n] = sg * v;
        }
        ++n;
      }
  }
  return n;
}

void* orc_srs_fill(void* mp, int nup, int ndn) { return srs_fill(*(Model*)mp, nup, ndn); }
void orc_srs_free(void* s) { delete (SRS*)s; }
long long orc_srs_dim(void* sp) { return ((SRS*)sp)->N; }
void orc_srs_av(void* s, const double* v, double* w, int clear) { srs_av((SRS*)s, v, w, clear != 0); }
void orc_srs_av_omp(void* s, const double* v, double* w, int clear) { srs_av_omp((SRS*)s, v, w, clear != 0); }
void orc_srs_diag(void* sp, double* out) {
  SRS* s = (SRS*)sp;
  std::memcpy(out, s->diag.data(), s->diag.size() * sizeof(double));
}
// which: 0 up, 1 down, 2 loc.  Returns nnz; fills rowp (nrows+1), col, val if non-null.
long long orc_srs_csr(void* sp, int which, int* rowp, int* col, double* val) {
  SRS* s = (SRS*)sp;
  const CSR& H = which == 0 ? s->Hup : (which == 1 ? s->Hdn : s->Hloc);
  long long nnz = H.vind;
  if (rowp) std::memcpy(rowp, H.rowp.data(), H.rowp.size() * sizeof(int));
  if (col) std::memcpy(col, H.col.data(), nnz * sizeof(int));
  if (val) std::memcpy(val, H.val.data(), nnz * sizeof(double));
  return nnz;
}
long long orc_srs_int_start(void* sp) { return (long long)((SRS*)sp)->int_start; }

void* orc_socrs_fill(void* mp, int nup, int ndn) { return socrs_fill(*(Model*)mp, nup, ndn); }
void orc_socrs_free(void* s) { delete (SOCRS*)s; }
void orc_socrs_av(void* s, const double* v, double* w, int clear) { socrs_av((SOCRS*)s, v, w, clear != 0); }
int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// SpinResolvedStorage::a_adag (SpinResolvedStorage.h:293-381, non-MPI) for
// orbital index `orb` (= site + spin*Ns); source sector (nup,ndn), target
// sector (nup2,ndn2); annihilate != 0 applies c, else c^+.
void orc_adag(void* mp, int nup, int ndn, int nup2, int ndn2, int orb, int annihilate, const double* in,
              double* out) {
  Model* m = (Model*)mp;
  std::vector<int> bu = enumerate_basis(m->Ns, nup), bd = enumerate_basis(m->Ns, ndn);
  std::vector<int> bu2 = enumerate_basis(m->Ns, nup2), bd2 = enumerate_basis(m->Ns, ndn2);
  std::vector<int> ninv_u(1 << m->Ns, 0), ninv_d(1 << m->Ns, 0);
  for (size_t r = 0; r < bu2.size(); ++r) ninv_u[bu2[r]] = (int)r;
  for (size_t r = 0; r < bd2.size(); ++r) ninv_d[bd2[r]] = (int)r;
  int64_t Ddn = bd.size(), N = bu.size() * Ddn, Ddn2 = bd2.size(), N2 = bu2.size() * Ddn2;
  for (int64_t q = 0; q < N2; ++q) out[q] = 0.0;
  for (int64_t ind = 0; ind < N; ++ind) {
    state_t nst = ((state_t)bu[ind / Ddn] << m->Ns) + bd[ind % Ddn];
    if (check_state(nst, orb, m->Ip) == (annihilate ? 1 : 0)) {
      state_t k;
      int sg;
      if (annihilate)
        op_a(orb, nst, k, sg, m->Ip);
      else
        op_adag(orb, nst, k, sg, m->Ip);
      int64_t i1 = (int64_t)ninv_u[(int)(k >> m->Ns)] * Ddn2 + ninv_d[(int)(k & ((1ll << m->Ns) - 1))];
      out[i1] = sg * in[ind];
    }
  }
}

// Per-state occupations for observables and the susceptibility operators:
// writes n_up[i*norb+o], n_dn[i*norb+o] for the first norb orbitals.
void orc_occupations(void* mp, int nup, int ndn, int norb, int8_t* nu, int8_t* nd) {
  Model* m = (Model*)mp;
  std::vector<int> bu = enumerate_basis(m->Ns, nup), bd = enumerate_basis(m->Ns, ndn);
  int64_t Ddn = bd.size(), N = bu.size() * Ddn;
  for (int64_t i = 0; i < N; ++i) {
    state_t nst = ((state_t)bu[i / Ddn] << m->Ns) + bd[i % Ddn];
    for (int o = 0; o < norb; ++o) {
      nu[i * norb + o] = (int8_t)check_state(nst, o, m->Ip);
      nd[i * norb + o] = (int8_t)check_state(nst, o + m->Ns, m->Ip);
    }
  }
}

// Lanczos::lanczos recurrence (Lanczos.h:36-77) with SRS::av in accumulate
// mode and the serial SRS::vv dot (SpinResolvedStorage.h:390-400).
// v is overwritten (as in the reference).  Returns nlanc.
// Lanczos::lanczos (Lanczos.h:36-77) on the OpenMP av, the vector loops and
// reductions split over threads (same recurrence as orc_lanczos; the sums
// differ from the serial ones by rounding only).
int orc_lanczos_omp(void* sp, double* v, int Nl, double* alfa, double* beta) {
  SRS* s = (SRS*)sp;
  int64_t size = s->N;
  std::vector<double> w(size, 0.0);
  double bet = 0;
  int nlanc = 0;
  for (int q = 0; q < Nl; ++q) alfa[q] = 0.0;
  for (int q = 0; q <= Nl; ++q) beta[q] = 0.0;
  for (int iter = 1; iter <= Nl; ++iter) {
    ++nlanc;
    if (iter != 1) {
#pragma omp parallel for schedule(static)
      for (int64_t j = 0; j < size; ++j) {
        double d = v[j];
        v[j] = w[j] / bet;
        w[j] = -bet * d;
      }
    }
    srs_av_omp(s, v, w.data(), false);
    double alf = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : alf)
    for (int64_t k = 0; k < size; ++k) alf += w[k] * v[k];
    alfa[iter - 1] = alf;
    double b2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : b2)
    for (int64_t j = 0; j < size; ++j) {
      w[j] -= alf * v[j];
      b2 += w[j] * w[j];
    }
    bet = std::sqrt(b2);
    if (iter != Nl) beta[iter] = bet;
    if (std::abs(bet) < 1e-10) break;
  }
  return nlanc;
}

int orc_lanczos(void* sp, double* v, int Nl, double* alfa, double* beta) {
  SRS* s = (SRS*)sp;
  int64_t size = s->N;
  std::vector<double> w(size, 0.0);
  double alf = 0, bet = 0;
  int nlanc = 0;
  for (int q = 0; q < Nl; ++q) alfa[q] = 0.0;
  for (int q = 0; q <= Nl; ++q) beta[q] = 0.0;
  for (int iter = 1; iter <= Nl; ++iter) {
    ++nlanc;
    if (iter != 1) {
      for (int64_t j = 0; j < size; ++j) {
        double d = v[j];
        v[j] = w[j] / bet;
        w[j] = -bet * d;
      }
    }
    srs_av(s, v, w.data(), false);
    alf = 0.0;
    for (int64_t k = 0; k < size; ++k) alf += w[k] * v[k];
    alfa[iter - 1] = alf;
    for (int64_t j = 0; j < size; ++j) w[j] -= alf * v[j];
    bet = 0.0;
    for (int64_t k = 0; k < size; ++k) bet += w[k] * w[k];
    bet = std::sqrt(bet);
    if (iter != Nl) beta[iter] = bet;
    if (std::abs(bet) < 1e-10) break;
  }
  return nlanc;
}

// Lanczos::get_frac_point (Lanczos.h:201-224) for a list of complex energies.
void orc_frac_points(int nz, const double* zre, const double* zim, double ev, int nlanc, int isign, double expb,
                     const double* alfa, const double* beta, double* out_re, double* out_im) {
  typedef std::complex<double> cd;
  std::vector<cd> det(nlanc), dl(nlanc);
  for (int q = 0; q < nz; ++q) {
    cd ener(zre[q], zim[q]);
    cd swp = 0.0;
    for (int i = 0; i < nlanc; ++i) dl[i] = ener - ((double)alfa[i] * isign);
    if (nlanc == 1) {
      swp += ev * expb / dl[0];
    } else {
      det[nlanc - 1] = dl[nlanc - 1];
      det[nlanc - 2] = dl[nlanc - 2] * dl[nlanc - 1] - std::pow(beta[nlanc - 1], 2);
      for (int i = nlanc - 3; i >= 0; --i) {
        det[i] = (dl[i] * det[i + 1] - std::pow(beta[i + 1], 2) * det[i + 2]);
        if (std::abs(det[i]) > (std::numeric_limits<float>::max() / 2.0) && i != 0) {
          double shift = 1.0 / (std::numeric_limits<float>::max() / 1000.0);
          det[i] *= shift;
          det[i + 1] *= shift;
        }
      }
      swp += ev * expb * det[1] / det[0];
    }
    out_re[q] = swp.real();
    out_im[q] = swp.imag();
  }
}

}  // extern "C"
