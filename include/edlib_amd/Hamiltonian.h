// Hamiltonian<Storage> and EigenPair with the reference's semantics:
// Hamiltonian.h:22-157 (sector loop: fill, diag, collect into a std::set) and
// EigenPair.h:14-51 (operator< = (E < E') || (id < id'), which makes the set
// iterate as a stable sort by eigenvalue).
#ifndef EDLIB_AMD_HAMILTONIAN_H
#define EDLIB_AMD_HAMILTONIAN_H

#include <iomanip>
#include <iostream>
#include <set>
#include <vector>

#include "HIPSpinResolvedStorage.h"
#include "Models.h"
#include "Params.h"

namespace EDLib {

template <typename precision, class SectorType>
class EigenPair {
 public:
  EigenPair(const precision& eval, const std::vector<precision>& evec, int id, SectorType sec)
      : eigenvalue_(eval), eigenvector_(evec), id_(id), sector_(sec) {}
  precision eigenvalue() const { return eigenvalue_; }
  const std::vector<precision>& eigenvector() const { return eigenvector_; }
  const SectorType& sector() const { return sector_; }
  int id() const { return id_; }
  bool operator<(const EigenPair& p) const { return (eigenvalue_ < p.eigenvalue_) || (id_ < p.id_); }
  bool operator>(const EigenPair& p) const { return (eigenvalue_ > p.eigenvalue_) || (id_ > p.id_); }

 private:
  precision eigenvalue_;
  std::vector<precision> eigenvector_;
  int id_;
  SectorType sector_;
};

template <class Storage>
class Hamiltonian {
 public:
  typedef typename Storage::Model Model;
  typedef typename Storage::Model ModelType;
  typedef Storage StorageType;
  typedef typename Model::precision prec;

  Hamiltonian(params& p) : model_(p), storage_(p, model_) {}

  void fill() { storage_.fill(); }

  void diag(bool verbose = true) {
    int k = 0;
    while (model_.symmetry().next_sector()) {
      if (verbose) std::cout << "Diagonalize sector " << model_.symmetry().sector() << std::endl;
      fill();
      int info = storage_.diag();
      if (info != 0) {
        std::cerr << "Eigenvalue have not been computed." << std::endl;
      } else {
        const std::vector<prec>& evals = storage_.eigenvalues();
        const std::vector<std::vector<prec>>& evecs = storage_.eigenvectors();
        for (size_t i = 0; i < evals.size(); ++i, ++k)
          eigenpairs_.insert(EigenPair<prec, typename Model::Sector>(evals[i], evecs[i], k, model_.symmetry().sector()));
      }
    }
    if (verbose) {
      std::cout << "Here is the list of eigenvalues:" << std::endl;
      std::streamsize pr = std::cout.precision();
      std::cout << std::setprecision(14);
      for (const auto& e : eigenpairs_) {
        std::cout << e.eigenvalue() << " ";
        e.sector().print();
        std::cout << std::endl;
      }
      std::cout << std::setprecision(pr);
    }
  }

  Storage& storage() { return storage_; }
  const std::set<EigenPair<prec, typename Model::Sector>>& eigenpairs() const { return eigenpairs_; }
  Model& model() { return model_; }
  void constant_shift(prec shift) { storage_.constant_shift(shift); }

 private:
  Model model_;  // constructed before the storage, which keeps a reference to it
  Storage storage_;
  std::set<EigenPair<prec, typename Model::Sector>> eigenpairs_;
};

typedef Hamiltonian<Storage::HIPSpinResolvedStorage<Model::HubbardModel<double>>> HIPHubbardHamiltonian;
typedef Hamiltonian<Storage::HIPSpinResolvedStorage<Model::SingleImpurityAndersonModel<double>>> HIPSIAMHamiltonian;

}  // namespace EDLib

#endif
