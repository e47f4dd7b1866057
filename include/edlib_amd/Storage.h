// Storage<prec>: the virtual interface of the reference's storage base class
// (Storage.h:15-248) for building the drop-in outside the reference tree.
// Inside the reference tree this file is not used: HIPSpinResolvedStorage.h
// includes "Storage.h" from its own directory, which there is the
// reference's (whose diag() drives ARPACK; the HIP storage hides it with its
// own device eigensolver, as this mirror leaves diag() to the derived class).
#ifndef EDLIB_AMD_STORAGE_H
#define EDLIB_AMD_STORAGE_H

#include <cstddef>
#include <vector>

namespace EDLib {
namespace Storage {

template <typename prec>
class Storage {
 public:
  // Storage(alps::params&) reads arpack.NEV, storage.EIGENVALUES_ONLY and
  // arpack.NCV (default 2 NEV + 3), Storage.h:20-31; any parameter object
  // with operator[] and exists() serves.
  template <class Params>
  explicit Storage(Params& p) : _nev(p["arpack.NEV"]), _eval_only(p["storage.EIGENVALUES_ONLY"]) {
    if (p.exists("arpack.NCV")) {
      _ncv = p["arpack.NCV"];
    } else {
      _ncv = 2 * _nev + 3;
    }
  }
  virtual ~Storage() {}

  // ntot() == 1: (A(0,0), [1]) without a diagonalisation (Storage.h:33-38).
  virtual void zero_eigenapair() = 0;
  // w = H v (+ w when !clear), Storage.h:174.
  virtual void av(prec* v, prec* w, int n, bool clear = true) = 0;
  virtual void prepare_work_arrays(prec* /*w*/, size_t /*shift*/ = 0) {}
  virtual int finalize(int info, bool /*bcast*/ = true, bool /*empty*/ = false) { return info; }

  const std::vector<prec>& eigenvalues() const { return evals; }
  std::vector<prec>& eigenvalues() { return evals; }
  const std::vector<std::vector<prec>>& eigenvectors() const { return evecs; }
  std::vector<std::vector<prec>>& eigenvectors() { return evecs; }

 protected:
  int& n() { return _n; }
  int& ntot() { return _ntot; }

 private:
  int _ntot = 0;
  int _n = 0;
  int _nev;
  int _ncv;
  int _eval_only;
  std::vector<prec> evals;
  std::vector<std::vector<prec>> evecs;
};

}  // namespace Storage
}  // namespace EDLib

#endif
