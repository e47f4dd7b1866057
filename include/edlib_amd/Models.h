// Models with the reference's constructors and accessors, backed by the
// engine's model objects (edh_model_*): HubbardModel.h:41-222,
// SingleImpurityAndersonModel.h:185-417, FermionicModel.h:21-119.
#ifndef EDLIB_AMD_MODELS_H
#define EDLIB_AMD_MODELS_H

#include <complex>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../edhip.h"
#include "Params.h"
#include "Symmetry.h"

namespace EDLib {

inline void edh_throw(int code) {
  if (code != EDH_OK) throw std::runtime_error(std::string("edhip: ") + edh_last_error());
}

// In-place inverse of the n x n complex matrix a (row-major) by Gauss-Jordan
// elimination with partial pivoting; false if a pivot vanishes.
inline bool complex_inverse(int n, std::vector<std::complex<double>>& a) {
  std::vector<std::complex<double>> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::abs(a[(size_t)r * n + c]) > std::abs(a[(size_t)piv * n + c])) piv = r;
    if (std::abs(a[(size_t)piv * n + c]) == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)piv * n + k], a[(size_t)c * n + k]);
        std::swap(inv[(size_t)piv * n + k], inv[(size_t)c * n + k]);
      }
    const std::complex<double> d = 1.0 / a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] *= d;
      inv[(size_t)c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const std::complex<double> f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  a.swap(inv);
  return true;
}

namespace Model {

// FermionicModel.h:21-119 (bit encoding and signs; Dyson/bare GF out of scope).
class FermionicModel {
 public:
  FermionicModel(params& p) : Ns_(p["NSITES"].as<int>()), ms_(p["NSPINS"].as<int>()), Ip_(Ns_ * ms_) {}
  int checkState(long long nst, int im, int Ip) const { return (int)((nst >> (Ip - 1 - im)) & 1ll); }
  void a(int i, long long jold, long long& k, int& isign) const {
    isign = (__builtin_popcountll((unsigned long long)jold >> (Ip_ - i)) & 1) ? -1 : 1;
    k = jold - (1ll << (Ip_ - i - 1));
  }
  void adag(int i, long long jold, long long& k, int& isign) const {
    isign = (__builtin_popcountll((unsigned long long)jold >> (Ip_ - i)) & 1) ? -1 : 1;
    k = jold + (1ll << (Ip_ - i - 1));
  }
  int orbitals() const { return Ns_; }
  int max_total_electrons() const { return Ip_; }
  int spins() const { return ms_; }

 protected:
  int Ns_, ms_, Ip_;
};

// One term of T_states() / V_states(): (kind, i, j, k, l, s, sp) + value
// as the engine lists them, in the reference's order.  kind 0 is the hop
// c^+_{j s} c_{i s} (Hubbard::InnerState, HubbardModel.h:22-37, and
// InnerHybridizationState, SingleImpurityAndersonModel.h:74-103); kind 1 is
// c^+_{i s} c^+_{j sp} c_{l sp} c_{k s} with value U/2
// (InnerInteractionState, SingleImpurityAndersonModel.h:105-184).
template <typename prec>
class InnerState {
 public:
  InnerState(const int* ints, double v) : kind_(ints[0]), i_(ints[1]), j_(ints[2]), k_(ints[3]), l_(ints[4]),
                                          s_(ints[5]), sp_(ints[6]), value_((prec)v) {}
  std::pair<int, int> indicies() const { return std::make_pair(i_, j_); }
  int spin() const { return s_; }
  int i() const { return i_; }
  int j() const { return j_; }
  int k() const { return k_; }
  int l() const { return l_; }
  prec U() const { return value_; }
  prec value() const { return kind_ == 0 ? value_ : prec(0.5) * value_; }
  int kind() const { return kind_; }
  int sigma() const { return s_; }
  int sigmaprime() const { return sp_; }

 private:
  int kind_, i_, j_, k_, l_, s_, sp_;
  prec value_;
};

// Holds the engine's model object and the term lists read from it.
template <typename prec>
class EngineModel {
 public:
  typedef InnerState<prec> St;
  const std::vector<St>& T_states() const { return T_; }
  const std::vector<St>& V_states() const { return V_; }

 protected:
  struct Free {
    void operator()(edh_model* m) const { edh_model_free(m); }
  };
  std::unique_ptr<edh_model, Free> m_;
  std::vector<St> T_, V_;
  int hop_rule_ = -1;  // -s1 s2 (HubbardModel.h:104-105) or +s1 s2 (SingleImpurityAndersonModel.h:83)
  void adopt(edh_model* m, int hop_rule) {
    m_.reset(m);
    hop_rule_ = hop_rule;
    for (int which = 0; which < 2; ++which) {
      int cnt = 0;
      edh_throw(edh_model_transitions(m, which, &cnt, nullptr, nullptr));
      std::vector<int> ints(7 * (size_t)cnt);
      std::vector<double> vals(cnt);
      if (cnt) edh_throw(edh_model_transitions(m, which, &cnt, ints.data(), vals.data()));
      for (int q = 0; q < cnt; ++q) (which ? V_ : T_).push_back(St(ints.data() + 7 * q, vals[q]));
    }
  }
  // valid / set on a full 2Ns-bit state (FermionicModel.h:35-68 signs)
  static int occ(long long st, int o, int Ip) { return (int)((st >> (Ip - 1 - o)) & 1ll); }
  static int sgn(long long st, int o, int Ip) {
    return (__builtin_popcountll((unsigned long long)st >> (Ip - o)) & 1) ? -1 : 1;
  }
  int term_valid(const St& t, long long nst, int Ns) const {
    const int Ip = 2 * Ns;
    if (t.kind() == 0) return occ(nst, t.i() + t.spin() * Ns, Ip) * (1 - occ(nst, t.j() + t.spin() * Ns, Ip));
    const int ok = t.k() + t.sigma() * Ns, ol = t.l() + t.sigmaprime() * Ns, oj = t.j() + t.sigmaprime() * Ns,
              oi = t.i() + t.sigma() * Ns;
    if (!occ(nst, ok, Ip)) return 0;
    const long long a = nst - (1ll << (Ip - 1 - ok));
    if (!occ(a, ol, Ip)) return 0;
    const long long b = a - (1ll << (Ip - 1 - ol));
    if (occ(b, oj, Ip)) return 0;
    return 1 - occ(b | (1ll << (Ip - 1 - oj)), oi, Ip);
  }
  prec term_set(const St& t, long long nst, long long& k, int& sign, int Ns) const {
    const int Ip = 2 * Ns;
    if (t.kind() == 0) {
      const int oi = t.i() + t.spin() * Ns, oj = t.j() + t.spin() * Ns;
      const int s1 = sgn(nst, oi, Ip);
      const long long k1 = nst - (1ll << (Ip - 1 - oi));
      sign = hop_rule_ * s1 * sgn(k1, oj, Ip);
      k = k1 + (1ll << (Ip - 1 - oj));
      return t.value();
    }
    const int ok = t.k() + t.sigma() * Ns, ol = t.l() + t.sigmaprime() * Ns, oj = t.j() + t.sigmaprime() * Ns,
              oi = t.i() + t.sigma() * Ns;
    int g = sgn(nst, ok, Ip);
    long long a = nst - (1ll << (Ip - 1 - ok));
    g *= sgn(a, ol, Ip);
    a -= 1ll << (Ip - 1 - ol);
    g *= sgn(a, oj, Ip);
    a += 1ll << (Ip - 1 - oj);
    g *= sgn(a, oi, Ip);
    k = a + (1ll << (Ip - 1 - oi));
    sign = g;
    return t.value();
  }
};

template <typename prcsn>
class HubbardModel : public FermionicModel, public EngineModel<prcsn> {
 public:
  typedef prcsn precision;
  typedef Symmetry::SzSymmetry SYMMETRY;
  typedef Symmetry::SzSymmetry::Sector Sector;
  typedef InnerState<prcsn> St;
  using EngineModel<prcsn>::T_states;
  using EngineModel<prcsn>::V_states;

  // Reads hopping/values, interaction/values, chemical_potential/values and
  // the optional magnetic_field/values, exchange/values (HubbardModel.h:55-66).
  HubbardModel(params& p) : FermionicModel(p), symmetry_(p) {
    alps_shim::input_file in(p["INPUT_FILE"].as<std::string>());
    std::vector<double> t = vec(in, "hopping/values", Ns_ * Ns_), U = vec(in, "interaction/values", Ns_),
                        xmu = vec(in, "chemical_potential/values", Ns_);
    std::vector<double> H(Ns_, 0.0), J(Ns_ * Ns_, 0.0);
    if (in.is_data("magnetic_field/values")) H = vec(in, "magnetic_field/values", Ns_);
    if (in.is_data("exchange/values")) J = vec(in, "exchange/values", Ns_ * Ns_);
    edh_model* m = nullptr;
    edh_throw(edh_model_hubbard(&m, Ns_, t.data(), U.data(), J.data(), xmu.data(), H.data()));
    this->adopt(m, -1);
    t_ = t;
    xmu_ = xmu;
  }
  // Bare Green's function matrix at frequency z for spin s, g[I * Ns + J]
  // (HubbardModel.h:181-201): [z + mu_I - eps_I + t_IJ]^-1 with the site
  // energies eps_I = 0 (they are never read from input, HubbardModel.h:49)
  // and mu indexed by site.
  void bare_gf_matrix(std::complex<double> z, int /*s*/, std::vector<std::complex<double>>& g) const {
    const int n = Ns_;
    g.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
      g[(size_t)i * n + i] = z + xmu_[i];
      for (int j = 0; j < n; ++j) g[(size_t)i * n + j] += t_[(size_t)i * n + j];
    }
    if (!complex_inverse(n, g)) throw std::runtime_error("bare Green's function: singular matrix");
  }
  precision diagonal(long long state) const {
    double x = 0;
    edh_throw(edh_model_diagonal(this->m_.get(), state, &x));
    return (precision)x;
  }
  // HubbardModel::valid / set (HubbardModel.h:86-108)
  int valid(const St& state, long long nst) { return this->term_valid(state, nst, Ns_); }
  precision set(const St& state, long long nst, long long& k, int& sign) {
    return this->term_set(state, nst, k, sign, Ns_);
  }
  int interacting_orbitals() const { return Ns_; }
  Symmetry::SzSymmetry& symmetry() { return symmetry_; }
  const Symmetry::SzSymmetry& symmetry() const { return symmetry_; }

 private:
  static std::vector<double> vec(const alps_shim::input_file& in, const std::string& k, size_t n) {
    const alps_shim::dataset& d = in.get(k);
    if (d.data.size() != n) throw std::invalid_argument("input dataset " + k + " has the wrong size");
    return d.data;
  }
  Symmetry::SzSymmetry symmetry_;
  std::vector<double> t_, xmu_;
};

template <typename prcsn>
class SingleImpurityAndersonModel : public FermionicModel, public EngineModel<prcsn> {
 public:
  typedef prcsn precision;
  typedef Symmetry::SzSymmetry SYMMETRY;
  typedef Symmetry::SzSymmetry::Sector Sector;
  typedef InnerState<prcsn> St;
  using EngineModel<prcsn>::T_states;
  using EngineModel<prcsn>::V_states;

  // Reads Bath/Epsk, Bath/Vk_m, H0_m, mu and interaction/values, with the
  // reference's argument checks (SingleImpurityAndersonModel.h:195-230).
  SingleImpurityAndersonModel(params& p) : FermionicModel(p), symmetry_(p), ml_(p["siam.NORBITALS"].as<int>()) {
    if (ml_ > Ns_)
      throw std::invalid_argument(
          "Incorrect values for the total number of sites and the number of orbitals. Please check input file.");
    if (ms_ != 2) throw std::invalid_argument("Incorrect values for the number of spins. Please check input file.");
    alps_shim::input_file in(p["INPUT_FILE"].as<std::string>());
    const int nb = Ns_ - ml_;
    const alps_shim::dataset& e = in.get("Bath/Epsk/values");
    if ((int)(e.data.size() / 2) != nb)
      throw std::invalid_argument(
          "Total number of state does not equal to sum of the total number of bath levels and the number of impurity "
          "orbitals");
    std::vector<double> Vk(ml_ * nb * 2, 0.0), H0(ml_ * ml_ * 2, 0.0);
    for (int im = 0; im < ml_; ++im) {
      std::ostringstream s1, s2;
      s1 << "Bath/Vk_" << im << "/values";
      s2 << "H0_" << im << "/values";
      const alps_shim::dataset& v = in.get(s1.str());
      const alps_shim::dataset& h = in.get(s2.str());
      if ((int)(h.data.size() / 2) > ml_)
        throw std::invalid_argument("Inter orbital hoppings array dimension are bigger than number of impurity orbitals");
      for (size_t q = 0; q < v.data.size() && q < (size_t)nb * 2; ++q) Vk[im * nb * 2 + q] = v.data[q];
      for (size_t q = 0; q < h.data.size(); ++q) H0[im * ml_ * 2 + q] = h.data[q];
    }
    const alps_shim::dataset& U = in.get("interaction/values");
    if (U.shape.size() != 6 || (int)U.shape[2] != ml_)
      throw std::invalid_argument("Incorrect number of orbitals. Please check input file.");
    const double mu = in.get("mu").data.at(0);
    edh_model* m = nullptr;
    edh_throw(edh_model_siam(&m, Ns_, ml_, e.data.data(), Vk.data(), H0.data(), U.data.data(), mu));
    this->adopt(m, +1);
    epsk_ = e.data;
    vk_ = Vk;
    h0_ = H0;
  }
  // Bare Green's function matrix of the impurity orbitals at frequency z for
  // spin s, g[I * ml + J] (SingleImpurityAndersonModel.h:378-391):
  // 1 / (z - H0_II - sum_k V_Ik^2 / (z - eps_k)) on the diagonal, as the
  // reference computes it (without the chemical potential and without the
  // off-diagonal H0); the reference defines it for one impurity orbital,
  // the diagonal form is its extension to several.
  void bare_gf_matrix(std::complex<double> z, int s, std::vector<std::complex<double>>& g) const {
    const int nb = Ns_ - ml_;
    g.assign((size_t)ml_ * ml_, 0.0);
    for (int i = 0; i < ml_; ++i) {
      std::complex<double> delta = 0.0;
      for (int k = 0; k < nb; ++k) {
        const double v = vk_[((size_t)i * nb + k) * 2 + s];
        delta += v * v / (z - epsk_[(size_t)k * 2 + s]);
      }
      g[(size_t)i * ml_ + i] = 1.0 / (z - h0_[((size_t)i * ml_ + i) * 2 + s] - delta);
    }
  }
  precision diagonal(long long state) const {
    double x = 0;
    edh_throw(edh_model_diagonal(this->m_.get(), state, &x));
    return (precision)x;
  }
  // SingleImpurityAndersonModel::valid / set (:346-357)
  int valid(const St& state, long long nst) { return this->term_valid(state, nst, Ns_); }
  precision set(const St& state, long long nst, long long& k, int& sign) {
    return this->term_set(state, nst, k, sign, Ns_);
  }
  int interacting_orbitals() const { return ml_; }
  Symmetry::SzSymmetry& symmetry() { return symmetry_; }
  const Symmetry::SzSymmetry& symmetry() const { return symmetry_; }

 private:
  Symmetry::SzSymmetry symmetry_;
  int ml_;
  std::vector<double> epsk_, vk_, h0_;
};

}  // namespace Model
}  // namespace EDLib

#endif
