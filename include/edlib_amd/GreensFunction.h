// Lanczos continued-fraction Green's function, local susceptibility and
// static observables on top of HIPSpinResolvedStorage: Lanczos.h:36-224,
// GreensFunction.h:25-462, ChiLoc.h:21-531, StaticObservables.h:106-421,
// MeshFactory.h:19-34, CommonUtils.h:14-20.
//
// The Lanczos recurrence runs on the device (storage().lanczos), and so does
// the continued-fraction evaluation on the mesh (storage().cf_eval).
#ifndef EDLIB_AMD_GREENSFUNCTION_H
#define EDLIB_AMD_GREENSFUNCTION_H

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "Hamiltonian.h"
#include "Params.h"

namespace EDLib {
namespace gf {

enum statistics { FERMIONIC = 1, BOSONIC = 0 };

// Positive Matsubara mesh: w_n = (2n+1) pi / beta or 2n pi / beta, n < NOMEGA.
struct MatsubaraMesh {
  std::vector<double> pts;
  double beta;
  bool real = false;
  MatsubaraMesh(params& p, statistics s) : beta(p["lanc.BETA"].as<double>()) {
    int n = p["lanc.NOMEGA"].as<int>();
    for (int i = 0; i < n; ++i) pts.push_back((2 * i + (s == FERMIONIC ? 1 : 0)) * M_PI / beta);
  }
  // z = i w_n (Lanczos.h:146-148)
  std::complex<double> freq_point(int i) const { return std::complex<double>(0.0, pts[i]); }
  int zero_freq() const { return 1; }
  int extent() const { return (int)pts.size(); }
};

// Linear real-frequency mesh on [EMIN, EMAX]; z = w + i pi / beta.
struct RealFrequencyMesh {
  std::vector<double> pts;
  double beta;
  bool real = true;
  RealFrequencyMesh(params& p) : beta(p["lanc.BETA"].as<double>()) {
    int n = p["lanc.NOMEGA"].as<int>();
    double a = p["lanc.EMIN"].as<double>(), b = p["lanc.EMAX"].as<double>();
    for (int i = 0; i < n; ++i) pts.push_back(n > 1 ? a + (b - a) * i / (n - 1) : a);
  }
  std::complex<double> freq_point(int i) const { return std::complex<double>(pts[i], M_PI / beta); }
  int zero_freq() const { return 0; }
  int extent() const { return (int)pts.size(); }
};

// G(w, orbital[, spin]) container.
struct GFData {
  int nw = 0, norb = 0, nspin = 1;
  std::vector<std::complex<double>> d;
  void init(int w, int o, int s) {
    nw = w;
    norb = o;
    nspin = s;
    d.assign((size_t)w * o * s, 0.0);
  }
  std::complex<double>& operator()(int w, int o, int s = 0) { return d[((size_t)w * norb + o) * nspin + s]; }
  const std::complex<double>& operator()(int w, int o, int s = 0) const { return d[((size_t)w * norb + o) * nspin + s]; }
  void scale(double x) {
    for (auto& v : d) v *= x;
  }
};

template <class Ham, class Mesh>
class Lanczos {
 public:
  typedef typename Ham::ModelType::precision precision;
  Lanczos(params& p, Ham& h, const Mesh& mesh)
      : ham_(h), mesh_(mesh), nl_(p["lanc.NLANC"].as<int>()), beta_(p["lanc.BETA"].as<double>()) {}
  const Mesh& omega() const { return mesh_; }
  double beta() const { return beta_; }

 protected:
  // Lanczos::lanczos (Lanczos.h:36-77) for the model's current sector.
  int lanczos(const std::vector<precision>& v) {
    ham_.fill();
    if (v.empty()) return 0;
    return ham_.storage().lanczos(v, nl_, alfa_, beta_lanc_);
  }
  static double expb(double beta, double e, double e0) { return beta * (e - e0) > 25 ? 0.0 : std::exp(-beta * (e - e0)); }
  // compute_continued_fraction (Lanczos.h:82-100): adds to gf(:, site, spin).
  void continued_fraction(double ev, double excited, double ground, int nlanc, int isign, GFData& gf, int site,
                          int spin) {
    std::vector<double> zr(mesh_.extent()), zi(mesh_.extent()), re, im;
    for (int w = 0; w < mesh_.extent(); ++w) {
      std::complex<double> z = mesh_.freq_point(w) + excited * isign;
      zr[w] = z.real();
      zi[w] = z.imag();
    }
    ham_.storage().cf_eval(zr, zi, ev, nlanc, isign, expb(beta_, excited, ground), alfa_, beta_lanc_, re, im);
    for (int w = 0; w < mesh_.extent(); ++w) gf(w, site, spin) += std::complex<double>(re[w], im[w]);
  }
  // compute_sym_continued_fraction + update_static (Lanczos.h:105-124,165-169).
  void sym_continued_fraction(double ev, double excited, double ground, int nlanc, GFData& gf, int site) {
    const double eb = expb(beta_, excited, ground);
    if (!mesh_.real) gf(0, site) -= ev * beta_ * eb;
    const int z0 = mesh_.zero_freq(), nw = mesh_.extent() - z0;
    std::vector<double> zr(2 * nw), zi(2 * nw), re, im;
    for (int w = 0; w < nw; ++w) {
      std::complex<double> a = mesh_.freq_point(w + z0) + excited, b = -mesh_.freq_point(w + z0) + excited;
      zr[w] = a.real();
      zi[w] = a.imag();
      zr[nw + w] = b.real();
      zi[nw + w] = b.imag();
    }
    ham_.storage().cf_eval(zr, zi, ev, nlanc, 1, eb, alfa_, beta_lanc_, re, im);
    for (int w = 0; w < nw; ++w)
      gf(w + z0, site) += std::complex<double>(re[w], im[w]) + std::complex<double>(re[nw + w], im[nw + w]);
  }
  Ham& ham_;
  Mesh mesh_;
  int nl_;
  double beta_;
  std::vector<precision> alfa_, beta_lanc_;
};

inline bool passes_cutoff(double bf, double cutoff) {
  return !(std::abs(cutoff - bf) > std::numeric_limits<double>::epsilon() && bf < cutoff);
}

// Orbital lists of GreensFunction.h:65-102 / ChiLoc.h:131-168.
inline void orbital_lists(params& p, const char* key, int norb, std::vector<size_t>& diag,
                          std::vector<std::array<size_t, 2>>& pairs) {
  std::vector<std::array<size_t, 2>> all;
  alps_shim::input_file in(p["INPUT_FILE"].as<std::string>());
  if (in.is_data(key)) {
    const auto& d = in.get(key).data;
    for (size_t q = 0; q + 1 < d.size(); q += 2) all.push_back({(size_t)d[q], (size_t)d[q + 1]});
  } else {
    for (int i = 0; i < norb; ++i) all.push_back({(size_t)i, (size_t)i});
  }
  for (auto& a : all) {
    if (a[0] == a[1])
      diag.push_back(a[0]);
    else
      pairs.push_back(a);
  }
  std::sort(diag.begin(), diag.end());
  diag.erase(std::unique(diag.begin(), diag.end()), diag.end());
  for (auto& pr : pairs)
    for (size_t o : pr)
      if (std::find(diag.begin(), diag.end(), o) == diag.end()) diag.push_back(o);
}

template <class Ham, class Mesh = MatsubaraMesh>
class GreensFunction : public Lanczos<Ham, Mesh> {
  typedef Lanczos<Ham, Mesh> Base;
  typedef typename Base::precision precision;
  typedef typename Ham::ModelType::Sector Sector;
  typedef EigenPair<precision, Sector> Pair;

 public:
  GreensFunction(params& p, Ham& h, const Mesh& mesh)
      : Base(p, h, mesh), model_(h.model()), cutoff_(p["lanc.BOLTZMANN_CUTOFF"].as<double>()) {
    if (p["storage.EIGENVALUES_ONLY"].as<int>() == 1)
      throw std::logic_error("Eigenvectors have not been computed. Green's function can not be evaluated.");
    const int norb = model_.interacting_orbitals();
    orbital_lists(p, "GreensFunction_orbitals/values", norb, orbs_, pairs_);
    const int nw = mesh.extent();
    Gg_.init(nw, norb, 2);
    Gl_.init(nw, norb, 2);
    Ggij_.init(nw, norb * norb, 2);
    Glij_.init(nw, norb * norb, 2);
  }

  // GreensFunction::compute (GreensFunction.h:109-159)
  void compute() {
    const int norb = model_.interacting_orbitals(), nw = this->mesh_.extent();
    Gg_.init(nw, norb, 2);
    Gl_.init(nw, norb, 2);
    Ggij_.init(nw, norb * norb, 2);
    Glij_.init(nw, norb * norb, 2);
    Z_ = 0.0;
    const auto& pairs = this->ham_.eigenpairs();
    if (pairs.empty()) return;
    const Pair& gs = *pairs.begin();
    for (const Pair& e : pairs) Z_ += std::exp(-(e.eigenvalue() - gs.eigenvalue()) * this->beta_);
    for (const Pair& pr : pairs) {
      double bf = std::exp(-(pr.eigenvalue() - gs.eigenvalue()) * this->beta_);
      if (!passes_cutoff(bf, cutoff_)) continue;
      for (size_t o : orbs_)
        for (int s = 0; s < 2; ++s) {
          contribute(pr, gs, {o}, s, false, Gg_, (int)o);
          contribute(pr, gs, {o}, s, true, Gl_, (int)o);
        }
      for (auto& ij : pairs_)
        for (int s = 0; s < 2; ++s) {
          contribute(pr, gs, {ij[0], ij[1]}, s, false, Ggij_, (int)(norb * ij[0] + ij[1]));
          contribute(pr, gs, {ij[0], ij[1]}, s, true, Glij_, (int)(norb * ij[0] + ij[1]));
        }
    }
    Gg_.scale(1.0 / Z_);
    Gl_.scale(1.0 / Z_);
    Ggij_.scale(1.0 / Z_);
    Glij_.scale(1.0 / Z_);
    // non_local_gf (GreensFunction.h:336-359)
    for (int w = 0; w < nw; ++w)
      for (int s = 0; s < 2; ++s) {
        for (auto& ij : pairs_) {
          int k = (int)(norb * ij[0] + ij[1]);
          for (size_t o : ij) {
            Ggij_(w, k, s) -= Gg_(w, (int)o, s);
            Glij_(w, k, s) -= Gl_(w, (int)o, s);
          }
          Ggij_(w, k, s) *= 0.5;
          Glij_(w, k, s) *= 0.5;
        }
        for (size_t o : orbs_) {
          Ggij_(w, (int)(norb * o + o), s) = Gg_(w, (int)o, s);
          Glij_(w, (int)(norb * o + o), s) = Gl_(w, (int)o, s);
        }
      }
    G_ = Gg_;
    for (size_t q = 0; q < G_.d.size(); ++q) G_.d[q] += Gl_.d[q];
    Gij_ = Ggij_;
    for (size_t q = 0; q < Gij_.d.size(); ++q) Gij_.d[q] += Glij_.d[q];
  }

  const GFData& G() const { return G_; }
  const GFData& G_g() const { return Gg_; }
  const GFData& G_l() const { return Gl_; }
  const GFData& G_ij() const { return Gij_; }
  double statsum() const { return Z_; }

  // Bare Green's function of the model on this mesh in G_ij's layout
  // (bare(w, I * norb + J, s)) and the self-energy from the Dyson equation,
  // Sigma = G0^-1 - G_ij^-1 per frequency and spin (FermionicModel.h:84-108,
  // GreensFunction.h:202-223).  G_ij holds the pairs listed in the input
  // (all of them for a full self-energy).
  void selfenergy(GFData& bare, GFData& sigma) const {
    const int norb = model_.interacting_orbitals(), nw = this->mesh_.extent();
    bare.init(nw, norb * norb, 2);
    sigma.init(nw, norb * norb, 2);
    std::vector<std::complex<double>> g0, g;
    for (int w = 0; w < nw; ++w)
      for (int s = 0; s < 2; ++s) {
        model_.bare_gf_matrix(this->mesh_.freq_point(w), s, g0);
        g.assign((size_t)norb * norb, 0.0);
        for (int k = 0; k < norb * norb; ++k) {
          bare(w, k, s) = g0[k];
          g[k] = Gij_(w, k, s);
        }
        if (!complex_inverse(norb, g0) || !complex_inverse(norb, g))
          throw std::runtime_error("Dyson equation: singular Green's function matrix");
        for (int k = 0; k < norb * norb; ++k) sigma(w, k, s) = g0[k] - g[k];
      }
  }
  // GreensFunction::compute_selfenergy (GreensFunction.h:202-223): stores
  // G0_omega and Sigma_omega under path and returns the self-energy.
  template <class Archive>
  GFData compute_selfenergy(Archive& ar, const std::string& path) const {
    GFData bare, sigma;
    selfenergy(bare, sigma);
    ar.write_gf(path + "/G0_omega", bare, this->mesh_, FERMIONIC);
    ar.write_gf(path + "/Sigma_omega", sigma, this->mesh_, FERMIONIC);
    return sigma;
  }

  // GreensFunction::save (GreensFunction.h:167-194): G_omega (G_omega_r on
  // the real axis), the statistical sum at path/@Statsum and G_ij_omega
  // when orbital pairs are listed; Archive = EDLib::archive (HDF5.h).
  template <class Archive>
  void save(Archive& ar, const std::string& path) const {
    const std::string sfx = this->mesh_.real ? "_r" : "";
    if (!orbs_.empty()) ar.write_gf(path + "/G_omega" + sfx, G_, this->mesh_, FERMIONIC);
    ar.write(path + "/@Statsum", Z_);
    if (!pairs_.empty()) ar.write_gf(path + "/G_ij_omega" + sfx, Gij_, this->mesh_, FERMIONIC);
  }

 private:
  // create_particles / annihilate_particles + lanczos + continued fraction
  // (GreensFunction.h:389-462, 241-282).
  void contribute(const Pair& pr, const Pair& gs, const std::vector<size_t>& orbs, int spin, bool annihilate,
                  GFData& target, int idx) {
    auto& sym = model_.symmetry();
    sym.set_sector(pr.sector());
    if (annihilate ? !sym.can_destroy_particle(spin) : !sym.can_create_particle(spin)) return;
    const int Ns = model_.orbitals();
    if (!annihilate && (spin == 0 ? pr.sector().nup() : pr.sector().ndown()) >= Ns) return;  // no empty orbital
    Sector next = annihilate ? sym.destroy_particle(spin) : sym.create_particle(spin);
    std::vector<precision> out(next.size(), 0.0), tmp;
    for (size_t o : orbs) {
      this->ham_.storage().a_adag((int)(o + spin * Ns), pr.eigenvector(), tmp, next, annihilate);
      for (size_t q = 0; q < out.size(); ++q) out[q] += tmp[q];
    }
    double norm = this->ham_.storage().vv(out, out);
    if (!(std::abs(norm) > 1e-10)) return;
    const double sn = std::sqrt(norm);
    for (auto& x : out) x /= sn;
    sym.set_sector(next);
    int nlanc = this->lanczos(out);
    this->continued_fraction(norm, pr.eigenvalue(), gs.eigenvalue(), nlanc, annihilate ? -1 : 1, target, idx, spin);
  }
  typename Ham::ModelType& model_;
  double cutoff_, Z_ = 0.0;
  std::vector<size_t> orbs_;
  std::vector<std::array<size_t, 2>> pairs_;
  GFData Gg_, Gl_, Ggij_, Glij_, G_, Gij_;
};

// Local susceptibility (ChiLoc.h:180-411): Op = EDH_OP_SZ (average 0 by
// default) or EDH_OP_N (average 1 by default).
template <class Ham, class Mesh = MatsubaraMesh>
class ChiLoc : public Lanczos<Ham, Mesh> {
  typedef Lanczos<Ham, Mesh> Base;
  typedef typename Base::precision precision;
  typedef typename Ham::ModelType::Sector Sector;
  typedef EigenPair<precision, Sector> Pair;

 public:
  ChiLoc(params& p, Ham& h, const Mesh& mesh)
      : Base(p, h, mesh), model_(h.model()), cutoff_(p["lanc.BOLTZMANN_CUTOFF"].as<double>()) {
    if (p["storage.EIGENVALUES_ONLY"].as<int>() == 1)
      throw std::logic_error("Eigenvectors have not been computed. Green's function can not be evaluated.");
    orbital_lists(p, "ChiLoc_orbitals/values", model_.interacting_orbitals(), orbs_, pairs_);
  }

  void compute(int op, const double* avg_ptr = nullptr) {
    type_ = op == EDH_OP_SZ ? "Sz" : "N";  // SzOperator::name / NOperator::name (ChiLoc.h:65, 90)
    const int norb = model_.interacting_orbitals(), nw = this->mesh_.extent();
    const double avg = avg_ptr ? *avg_ptr : (op == EDH_OP_SZ ? 0.0 : 1.0);
    gf_.init(nw, norb, 1);
    gfij_.init(nw, norb * norb, 1);
    Z_ = 0.0;
    const auto& pairs = this->ham_.eigenpairs();
    if (pairs.empty()) return;
    const Pair& gs = *pairs.begin();
    for (const Pair& e : pairs) Z_ += std::exp(-(e.eigenvalue() - gs.eigenvalue()) * this->beta_);
    for (const Pair& pr : pairs) {
      double bf = std::exp(-(pr.eigenvalue() - gs.eigenvalue()) * this->beta_);
      if (!passes_cutoff(bf, cutoff_)) continue;
      for (size_t o : orbs_) contribute(pr, gs, op, {(int)o}, true, gf_, (int)o);
      for (auto& ij : pairs_) contribute(pr, gs, op, {(int)ij[0], (int)ij[1]}, false, gfij_, (int)(ij[0] * norb + ij[1]));
    }
    gf_.scale(1.0 / Z_);
    gfij_.scale(1.0 / Z_);
    if (!this->mesh_.real) {
      for (size_t o : orbs_) zero_freq(gf_, (int)o, avg);
      for (auto& ij : pairs_) zero_freq(gfij_, (int)(ij[0] * norb + ij[1]), avg);
    }
    for (int w = 0; w < nw; ++w) {
      for (auto& ij : pairs_) {
        int k = (int)(ij[0] * norb + ij[1]);
        for (size_t o : ij) gfij_(w, k) -= gf_(w, (int)o);
        gfij_(w, k) *= 0.5;
      }
      for (size_t o : orbs_) gfij_(w, (int)(o * norb + o)) = gf_(w, (int)o);
    }
  }
  const GFData& G() const { return gf_; }
  const GFData& G_ij() const { return gfij_; }

  // ChiLoc::save (ChiLoc.h:418-437): Chi<Sz|N>_omega of the last compute,
  // path/@Statsum and Chi<Sz|N>_ij_omega when orbital pairs are listed.
  template <class Archive>
  void save(Archive& ar, const std::string& path) const {
    if (!orbs_.empty()) ar.write_gf(path + "/Chi" + type_ + "_omega", gf_, this->mesh_, BOSONIC);
    ar.write(path + "/@Statsum", Z_);
    if (!pairs_.empty()) ar.write_gf(path + "/Chi" + type_ + "_ij_omega", gfij_, this->mesh_, BOSONIC);
  }

 private:
  void contribute(const Pair& pr, const Pair& gs, int op, const std::vector<int>& orbs, bool need_norm, GFData& g,
                  int idx) {
    auto& sym = model_.symmetry();
    sym.set_sector(pr.sector());
    std::vector<precision> out;
    this->ham_.storage().local_op(op, orbs, pr.eigenvector(), out);
    double norm = this->ham_.storage().vv(out, out);
    const double sn = std::sqrt(norm);
    for (auto& x : out) x /= sn;
    if (need_norm && !(norm > 1e-9)) return;
    int nlanc = this->lanczos(out);
    this->sym_continued_fraction(norm, pr.eigenvalue(), gs.eigenvalue(), nlanc, g, idx);
  }
  // zero_freq_contribution + get_tail (ChiLoc.h:358-411)
  void zero_freq(GFData& G, int i, double avg) {
    const auto& om = this->mesh_.pts;
    const int nw = (int)om.size();
    const double beta = this->beta_;
    auto tail = [&](int f, double& c2, double& c4) {
      double o1 = om[f] * om[f], o2 = om[f - 1] * om[f - 1];
      double g1 = G(f, i).real(), g2 = G(f - 1, i).real();
      c2 = -(g2 * o2 * o2 - g1 * o1 * o1) / (o1 - o2);
      c4 = -(g1 * o1 * o1 * o2 - g2 * o2 * o2 * o1) / (o1 - o2);
      return c2 * beta * beta / 24.0 + c4 * beta * beta * beta * beta / 1440.0;
    };
    double c2, c4, c2b, c4b;
    double t1 = tail(nw - 1, c2, c4), t2 = tail(nw - 2, c2b, c4b);
    if (std::abs(t1 - t2) / std::abs(t1) > 1e-4)
      std::cerr << "Not enough frequencies to compute high frequency tail. Please increase number of frequencies. Diff: "
                << std::abs(t1 - t2) / std::abs(t1) << std::endl;
    double s = 0.0;
    for (int w = 1; w < nw; ++w) s += G(w, i).real() - c2 / (om[w] * om[w]) - c4 / (om[w] * om[w] * om[w] * om[w]);
    G(0, i) -= 2 * s + 2 * t1 - avg * avg * beta;
  }
  typename Ham::ModelType& model_;
  std::string type_ = "Sz";
  double cutoff_, Z_ = 0.0;
  std::vector<size_t> orbs_;
  std::vector<std::array<size_t, 2>> pairs_;
  GFData gf_, gfij_;
};

}  // namespace gf

// StaticObservables::calculate_static_observables (StaticObservables.h:106-158,356-421).
template <class Ham>
class StaticObservables {
 public:
  typedef typename Ham::ModelType::precision precision;
  StaticObservables(params& p) : beta_(p["lanc.BETA"].as<double>()), cutoff_(p["lanc.BOLTZMANN_CUTOFF"].as<double>()) {
    if (p["storage.EIGENVALUES_ONLY"].as<int>() == 1)
      throw std::logic_error("Eigenvectors have not been computed. Static observables can not be evaluated.");
  }
  // StaticObservables::calculate_static_observables (StaticObservables.h:
  // 106-158): Boltzmann-weighted average over the stored eigenpairs of the
  // per-eigenvector sums, which the engine forms on the device
  // (edh_static_observables, StaticObservables.h:356-419).
  std::map<std::string, std::vector<precision>> calculate_static_observables(Ham& ham) {
    const int norb = ham.model().interacting_orbitals();
    std::map<std::string, std::vector<precision>> avg = {
        {"N", std::vector<precision>(norb, 0.0)},     {"N_up", std::vector<precision>(norb, 0.0)},
        {"N_dn", std::vector<precision>(norb, 0.0)},  {"M", std::vector<precision>(norb, 0.0)},
        {"D_occ", std::vector<precision>(norb, 0.0)}, {"M_i M_j", std::vector<precision>(norb * norb, 0.0)},
        {"N_eff", std::vector<precision>(1, 0.0)},    {"E", std::vector<precision>(1, 0.0)}};
    const auto& pairs = ham.eigenpairs();
    if (pairs.empty()) return avg;
    const double e0 = pairs.begin()->eigenvalue();
    double sum = 0.0;
    auto& sym = ham.model().symmetry();
    std::vector<double> out;
    const char* keys[5] = {"N", "N_up", "N_dn", "M", "D_occ"};
    for (const auto& pr : pairs) {
      double bf = std::exp(-(pr.eigenvalue() - e0) * beta_);
      if (!gf::passes_cutoff(bf, cutoff_)) continue;
      sym.set_sector(pr.sector());
      ham.storage().fill();
      ham.storage().static_observables(pr.eigenvector(), norb, out);
      for (int k = 0; k < 5; ++k)
        for (int o = 0; o < norb; ++o) avg[keys[k]][o] += out[(size_t)k * norb + o] * bf;
      for (int q = 0; q < norb * norb; ++q) avg["M_i M_j"][q] += out[5 * (size_t)norb + q] * bf;
      avg["N_eff"][0] += out[5 * (size_t)norb + (size_t)norb * norb] * bf;
      avg["E"][0] += pr.eigenvalue() * bf;
      sum += bf;
    }
    for (auto& kv : avg)
      for (auto& x : kv.second) x /= sum;
    return avg;
  }

 private:
  double beta_, cutoff_;
};

}  // namespace EDLib

#endif
