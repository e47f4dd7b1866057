// HIPSpinResolvedStorage<Model>: the Storage concept of the reference
// (Storage.h:16-248 + SpinResolvedStorage.h:20-639) served by the MI355X
// engine through the C ABI in edhip.h.  Drop-in for
//   Hamiltonian<SpinResolvedStorage<HubbardModel<double>>>  ->
//   Hamiltonian<HIPSpinResolvedStorage<HubbardModel<double>>>.
//
// av() keeps the reference's calling convention (host pointers owned by the
// caller, w = H v + (clear ? 0 : w)); diag() replaces the ARPACK reverse-
// communication loop by the engine's device-resident thick-restart Lanczos
// with the same parameters (nev, ncv, which = "SA", tol = 1e-14, maxitr = 1000).
#ifndef EDLIB_AMD_HIPSPINRESOLVEDSTORAGE_H
#define EDLIB_AMD_HIPSPINRESOLVEDSTORAGE_H

#include <algorithm>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "../edhip.h"
#include "Models.h"
#include "Params.h"
#include "Symmetry.h"

namespace EDLib {
namespace Storage {

template <class ModelType>
class HIPSpinResolvedStorage {
  static_assert(std::is_base_of<Symmetry::SzSymmetry, typename ModelType::SYMMETRY>::value,
                "Model have wrong symmetry.");

 public:
  typedef ModelType Model;
  typedef typename Model::precision prec;
  static_assert(std::is_same<prec, double>::value, "the engine computes in FP64");

  HIPSpinResolvedStorage(params& p, Model& m)
      : model_(m), nev_(p["arpack.NEV"].as<int>()), eval_only_(p["storage.EIGENVALUES_ONLY"].as<int>()),
        Ns_(p["NSITES"].as<int>()) {
    ncv_ = p.exists("arpack.NCV") ? p["arpack.NCV"].as<int>() : 2 * nev_ + 3;  // Storage.h:26-30
    int dev = p.exists("device.ID") ? p["device.ID"].as<int>() : 0;
    edh_ctx* c = nullptr;
    edh_throw(edh_create(&c, dev));
    ctx_.reset(c);
    edh_throw(edh_set_model(ctx_.get(), m.engine_model()));
  }

  // ---- Storage concept ---------------------------------------------------
  void av(prec* v, prec* w, int n, bool clear = true) {
    edh_throw(edh_av(ctx_.get(), v, w, n, clear ? 1 : 0, EDH_MEM_HOST));
  }

  // SpinResolvedStorage::fill: per-spin tables for the model's current sector.
  void fill() {
    reset();
    if (n_ == 0) return;
    int64_t dim = 0;
    const auto& s = model_.symmetry().sector();
    edh_throw(edh_set_sector(ctx_.get(), s.nup(), s.ndown(), &dim));
    filled_ = true;
  }
  void reset(int = 0) {
    model_.symmetry().init();
    const auto& s = model_.symmetry().sector();
    n_ = ntot_ = (int)s.size();
    filled_ = false;
  }
  void init() { model_.symmetry().init(); }
  void print() {}

  size_t vector_size(typename Model::Sector sector) { return sector.size(); }

  // a_adag(i, invec, outvec, next_sec, a): c_i (a = true) or c^+_i on invec
  // (a vector of the model's current sector) into outvec (next_sec).
  void a_adag(int i, const std::vector<prec>& invec, std::vector<prec>& outvec,
              const typename Model::Sector& next_sec, bool a) {
    const auto& s = model_.symmetry().sector();
    int64_t outn = 0;
    outvec.assign(next_sec.size(), prec(0));
    edh_throw(edh_adag(ctx_.get(), s.nup(), s.ndown(), i, a ? 1 : 0, invec.data(), outvec.data(), &outn,
                       EDH_MEM_HOST));
  }

  prec vv(const std::vector<prec>& v, const std::vector<prec>& w) {
    double out = 0;
    edh_throw(edh_vv(ctx_.get(), v.data(), w.data(), (int64_t)v.size(), &out, EDH_MEM_HOST));
    return out;
  }

  void constant_shift(prec shift) { edh_throw(edh_constant_shift(ctx_.get(), shift)); }
  void zero_eigenapair() {}
  void prepare_work_arrays(prec*, size_t = 0) {}
  int finalize(int info, bool = true, bool = false) { return info; }

  // Storage::diag (Storage.h:43-146): lowest eigenpairs of the current sector.
  int diag() {
    evals_.clear();
    evecs_.clear();
    if (n_ == 0) return 0;
    if (!filled_) fill();
    const int ncv = std::min(ncv_, ntot_);
    const int nev = ntot_ == 1 ? 1 : std::min(nev_, ncv - 1);
    std::vector<double> ev(std::max(nev, 1));
    std::vector<double> X;
    if (!eval_only_) X.resize((size_t)n_ * nev);
    int nconv = 0, nav = 0;
    int rc = edh_eigsh(ctx_.get(), nev_, ncv_, 1e-14, 1000, 0x45444c6962ull, ev.data(), eval_only_ ? nullptr : X.data(),
                       &nconv, &nav);
    edh_throw(rc);
    if (nconv < std::min(nev_, (int)std::min<int64_t>(ncv_, n_) - 1) && n_ > 1)
      std::cout << "Maximum number of iterations reached." << std::endl;  // Storage.h:122-124
    evals_.assign(ev.begin(), ev.begin() + nconv);
    if (!eval_only_) {
      evecs_.resize(nconv);
      for (int i = 0; i < nconv; ++i) evecs_[i].assign(X.begin() + (size_t)i * n_, X.begin() + (size_t)(i + 1) * n_);
    } else {
      evecs_.assign(nconv, std::vector<prec>(1, prec(0)));
    }
    return 0;
  }

  // Device-resident Lanczos recurrence (Lanczos.h:36-77) on the filled
  // sector: returns nlanc; alpha (nl) and beta (nl+1) as in the reference.
  int lanczos(const std::vector<prec>& v, int nl, std::vector<prec>& alpha, std::vector<prec>& beta) {
    alpha.assign(nl, 0.0);
    beta.assign(nl + 1, 0.0);
    int nlanc = 0;
    edh_throw(edh_lanczos(ctx_.get(), v.data(), (int64_t)v.size(), nl, alpha.data(), beta.data(), &nlanc, EDH_MEM_HOST));
    return nlanc;
  }
  // Susceptibility operator on a vector of the current sector (ChiLoc.h:474-530).
  void local_op(int op, const std::vector<int>& orbs, const std::vector<prec>& in, std::vector<prec>& out) {
    const auto& s = model_.symmetry().sector();
    out.assign(in.size(), 0.0);
    edh_throw(edh_local_op(ctx_.get(), s.nup(), s.ndown(), op, (int)orbs.size(), orbs.data(), in.data(), out.data(),
                           EDH_MEM_HOST));
  }
  // Per-eigenvector static observables of a vector of the filled sector,
  // on the device (StaticObservables.h:356-419): N, N_up, N_dn, M, D_occ
  // (norb each), M_i M_j (norb x norb), N_eff.
  void static_observables(const std::vector<prec>& v, int norb, std::vector<double>& out) {
    out.assign(5 * (size_t)norb + (size_t)norb * norb + 1, 0.0);
    edh_throw(edh_static_observables(ctx_.get(), v.data(), (int64_t)v.size(), EDH_MEM_HOST, norb, out.data()));
  }
  // Continued fraction on the device (Lanczos.h:201-224).
  void cf_eval(const std::vector<double>& zre, const std::vector<double>& zim, double ev, int nlanc, int isign,
               double expb, const std::vector<prec>& alpha, const std::vector<prec>& beta, std::vector<double>& re,
               std::vector<double>& im) {
    re.assign(zre.size(), 0.0);
    im.assign(zre.size(), 0.0);
    edh_throw(edh_cf_eval(ctx_.get(), (int)zre.size(), zre.data(), zim.data(), ev, nlanc, isign, expb, alpha.data(),
                          beta.data(), re.data(), im.data()));
  }

  const std::vector<prec>& eigenvalues() const { return evals_; }
  std::vector<prec>& eigenvalues() { return evals_; }
  const std::vector<std::vector<prec>>& eigenvectors() const { return evecs_; }
  std::vector<std::vector<prec>>& eigenvectors() { return evecs_; }
  int& n() { return n_; }
  int& ntot() { return ntot_; }
  edh_ctx* context() { return ctx_.get(); }

 private:
  struct Destroy {
    void operator()(edh_ctx* c) const { edh_destroy(c); }
  };
  Model& model_;
  std::unique_ptr<edh_ctx, Destroy> ctx_;
  int nev_, ncv_, eval_only_, Ns_;
  int n_ = 0, ntot_ = 0;
  bool filled_ = false;
  std::vector<prec> evals_;
  std::vector<std::vector<prec>> evecs_;
};

}  // namespace Storage
}  // namespace EDLib

#endif
