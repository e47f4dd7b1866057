// HIPSpinResolvedStorage<Model>: the reference's storage concept
// (Storage.h:15-248 + SpinResolvedStorage.h:20-639) served by the MI355X
// engine through the C ABI in edhip.h.  Drop-in for
//   Hamiltonian<SpinResolvedStorage<HubbardModel<double>>>  ->
//   Hamiltonian<HIPSpinResolvedStorage<HubbardModel<double>>>.
//
// It derives from Storage<prec> (the "Storage.h" beside this file: the
// reference's own inside the reference tree) and reads the model only
// through the interface SpinResolvedStorage::fill uses (T_states(),
// V_states(), valid/set, diagonal(), interacting_orbitals(), symmetry()),
// so the reference's HubbardModel and SingleImpurityAndersonModel plug in
// unchanged.  The constructor takes the caller's parameter object
// (alps::params in the reference) as it is.
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

#include "ModelTerms.h"
#include "Storage.h"
#include "SzSymmetry.h"
#include "edhip.h"

namespace EDLib {
namespace Storage {

template <class ModelType>
class HIPSpinResolvedStorage : public Storage<typename ModelType::precision> {
  static_assert(std::is_base_of<Symmetry::SzSymmetry, typename ModelType::SYMMETRY>::value,
                "Model have wrong symmetry.");

 public:
  typedef ModelType Model;
  typedef typename Model::precision prec;
  static_assert(std::is_same<prec, double>::value, "the engine computes in FP64");
  typedef Storage<prec> Base;
  using Base::n;
  using Base::ntot;
  using Base::eigenvalues;
  using Base::eigenvectors;

  // SpinResolvedStorage(alps::params&, Model&), SpinResolvedStorage.h:47-48.
  // Keys: arpack.NEV, arpack.NCV, storage.EIGENVALUES_ONLY, NSITES; the
  // optional device.ID picks the GPU (default 0).
  template <class Params>
  HIPSpinResolvedStorage(Params& p, Model& m) : Base(p), model_(m) {
    nev_ = p["arpack.NEV"];
    eval_only_ = p["storage.EIGENVALUES_ONLY"];
    Ns_ = p["NSITES"];
    if (p.exists("arpack.NCV")) {
      ncv_ = p["arpack.NCV"];
    } else {
      ncv_ = 2 * nev_ + 3;  // Storage.h:26-30
    }
    int dev = 0;
    if (p.exists("device.ID")) dev = p["device.ID"];
    EngineModelPtr em = engine_model_from_interface(m, Ns_);
    edh_ctx* c = nullptr;
    edh_check(edh_create(&c, dev));
    ctx_.reset(c);
    edh_check(edh_set_model(ctx_.get(), em.get()));
  }

  // ---- Storage concept ---------------------------------------------------
  virtual void av(prec* v, prec* w, int n, bool clear = true) {
    edh_check(edh_av(ctx_.get(), v, w, n, clear ? 1 : 0, EDH_MEM_HOST));
  }

  // SpinResolvedStorage::fill: per-spin tables for the model's current sector.
  void fill() {
    reset();
    if (n() == 0) return;
    int64_t dim = 0;
    const auto& s = model_.symmetry().sector();
    edh_check(edh_set_sector(ctx_.get(), s.nup(), s.ndown(), &dim));
    filled_ = true;
  }
  void reset(int = 0) {
    model_.symmetry().init();
    const auto& s = model_.symmetry().sector();
    n() = ntot() = (int)s.size();
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
    edh_check(edh_adag(ctx_.get(), s.nup(), s.ndown(), i, a ? 1 : 0, invec.data(), outvec.data(), &outn,
                       EDH_MEM_HOST));
  }

  prec vv(const std::vector<prec>& v, const std::vector<prec>& w) {
    double out = 0;
    edh_check(edh_vv(ctx_.get(), v.data(), w.data(), (int64_t)v.size(), &out, EDH_MEM_HOST));
    return out;
  }

  // Adds shift to every diagonal element of the filled sector, until the
  // next fill() (SpinResolvedStorage.h:464-466; its refill discards it too).
  void constant_shift(prec shift) { edh_check(edh_constant_shift(ctx_.get(), shift)); }

  // ntot() == 1: eigenvalue = the sector's one diagonal element, eigenvector
  // [1] (SpinResolvedStorage.h:52-56).  The element is read through av on
  // the unit vector, so a constant_shift is included as in the reference.
  virtual void zero_eigenapair() {
    if (!filled_) fill();
    prec one = 1.0, d = 0.0;
    av(&one, &d, 1, true);
    eigenvalues().assign(1, d);
    eigenvectors().assign(1, std::vector<prec>(1, prec(1.0)));
  }
  virtual void prepare_work_arrays(prec*, size_t = 0) {}
  virtual int finalize(int info, bool = true, bool = false) { return info; }

  // Storage::diag (Storage.h:43-146): lowest eigenpairs of the current
  // sector; n == 0 gives none, ntot == 1 goes through zero_eigenapair.
  int diag() {
    eigenvalues().clear();
    eigenvectors().clear();
    if (n() == 0) return finalize(0, true, true);
    if (!filled_) fill();
    if (ntot() == 1) {
      zero_eigenapair();
      return finalize(0);
    }
    const int nsec = n();
    const int ncv = std::min(ncv_, ntot());
    const int nev = std::min(nev_, ncv - 1);
    std::vector<double> ev(std::max(nev, 1));
    std::vector<double> X;
    if (!eval_only_) X.resize((size_t)nsec * nev);
    int nconv = 0, nav = 0;
    edh_check(edh_eigsh(ctx_.get(), nev_, ncv_, 1e-14, 1000, 0x45444c6962ull, ev.data(),
                        eval_only_ ? nullptr : X.data(), &nconv, &nav));
    if (nconv < nev) std::cout << "Maximum number of iterations reached." << std::endl;  // Storage.h:122-124
    eigenvalues().assign(ev.begin(), ev.begin() + nconv);
    if (!eval_only_) {
      eigenvectors().resize(nconv);
      for (int i = 0; i < nconv; ++i)
        eigenvectors()[i].assign(X.begin() + (size_t)i * nsec, X.begin() + (size_t)(i + 1) * nsec);
    } else {
      eigenvectors().assign(nconv, std::vector<prec>(1, prec(0)));
    }
    return finalize(0);
  }

  // Device-resident Lanczos recurrence (Lanczos.h:36-77) on the filled
  // sector: returns nlanc; alpha (nl) and beta (nl+1) as in the reference.
  int lanczos(const std::vector<prec>& v, int nl, std::vector<prec>& alpha, std::vector<prec>& beta) {
    alpha.assign(nl, 0.0);
    beta.assign(nl + 1, 0.0);
    int nlanc = 0;
    edh_check(edh_lanczos(ctx_.get(), v.data(), (int64_t)v.size(), nl, alpha.data(), beta.data(), &nlanc, EDH_MEM_HOST));
    return nlanc;
  }
  // Susceptibility operator on a vector of the current sector (ChiLoc.h:474-530).
  void local_op(int op, const std::vector<int>& orbs, const std::vector<prec>& in, std::vector<prec>& out) {
    const auto& s = model_.symmetry().sector();
    out.assign(in.size(), 0.0);
    edh_check(edh_local_op(ctx_.get(), s.nup(), s.ndown(), op, (int)orbs.size(), orbs.data(), in.data(), out.data(),
                           EDH_MEM_HOST));
  }
  // Per-eigenvector static observables of a vector of the filled sector,
  // on the device (StaticObservables.h:356-419): N, N_up, N_dn, M, D_occ
  // (norb each), M_i M_j (norb x norb), N_eff.
  void static_observables(const std::vector<prec>& v, int norb, std::vector<double>& out) {
    out.assign(5 * (size_t)norb + (size_t)norb * norb + 1, 0.0);
    edh_check(edh_static_observables(ctx_.get(), v.data(), (int64_t)v.size(), EDH_MEM_HOST, norb, out.data()));
  }
  // Continued fraction on the device (Lanczos.h:201-224).
  void cf_eval(const std::vector<double>& zre, const std::vector<double>& zim, double ev, int nlanc, int isign,
               double expb, const std::vector<prec>& alpha, const std::vector<prec>& beta, std::vector<double>& re,
               std::vector<double>& im) {
    re.assign(zre.size(), 0.0);
    im.assign(zre.size(), 0.0);
    edh_check(edh_cf_eval(ctx_.get(), (int)zre.size(), zre.data(), zim.data(), ev, nlanc, isign, expb, alpha.data(),
                          beta.data(), re.data(), im.data()));
  }

  edh_ctx* context() { return ctx_.get(); }

 private:
  struct Destroy {
    void operator()(edh_ctx* c) const { edh_destroy(c); }
  };
  Model& model_;
  std::unique_ptr<edh_ctx, Destroy> ctx_;
  int nev_ = 0, ncv_ = 0, eval_only_ = 0, Ns_ = 0;
  bool filled_ = false;
};

}  // namespace Storage
}  // namespace EDLib

#endif
