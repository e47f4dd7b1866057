// Spin-resolved symmetry (conserved N_up, N_down) with the reference's API
// and ordering: SzSymmetry.h:19-234, Combination.h:13-77.
#ifndef EDLIB_AMD_SYMMETRY_H
#define EDLIB_AMD_SYMMETRY_H

#include <algorithm>
#include <cstdint>
#include <functional>
#include <iostream>
#include <queue>
#include <vector>

#include "Params.h"

namespace EDLib {

// Binomial table; lexicographic successor of a sorted position tuple.
class Combination {
 public:
  explicit Combination(int N) : n_(N), t_((N + 1) * (N + 1), 0) {
    for (int a = 0; a <= N; ++a) {
      t_[a * (N + 1)] = 1;
      for (int b = 1; b <= a; ++b) t_[a * (N + 1) + b] = t_[(a - 1) * (N + 1) + b - 1] + (b < a ? t_[(a - 1) * (N + 1) + b] : 0);
    }
  }
  int c_n_k(int n, int k) const {
    if (k < 0 || n < 0 || k > n) return 0;
    return (int)t_[n * (n_ + 1) + k];
  }
  void init_state(int k, std::vector<int>& pos) const {
    for (int i = 0; i < k; ++i) pos[i] = i;
  }
  bool next_combination(int n, int k, std::vector<int>& pos) const {
    for (int i = k - 1; i >= 0; --i)
      if (pos[i] < n - k + i) {
        ++pos[i];
        for (int j = i + 1; j < k; ++j) pos[j] = pos[j - 1] + 1;
        return true;
      }
    return false;
  }

 private:
  int n_;
  std::vector<int64_t> t_;
};

namespace Symmetry {

class SzSymmetry {
 public:
  class Sector {
   public:
    Sector(int up, int down, size_t size) : nup_(up), ndown_(down), size_(size) {}
    int nup() const { return nup_; }
    int ndown() const { return ndown_; }
    size_t size() const { return size_; }
    void print(std::ostream& o = std::cout) const { o << nup_ << " " << ndown_; }
    // Same comparator as SzSymmetry.h:51-54 (not a strict weak ordering; the
    // resulting std::sort order is reproduced by using it unchanged).
    bool operator<(const Sector& s) const {
      return size_ < s.size_ || (size_ == s.size_ && nup_ < s.nup_ && ndown_ < s.ndown_) ||
             (size_ == s.size_ && nup_ == s.nup_ && ndown_ < s.ndown_);
    }
    bool operator>(const Sector& s) const { return s < *this; }
    bool operator==(const Sector& s) const { return nup_ == s.nup_ && ndown_ == s.ndown_; }
    friend std::ostream& operator<<(std::ostream& o, const Sector& c) {
      return o << " (nup: " << c.nup_ << " ndown: " << c.ndown_ << ") size: " << c.size_;
    }

   private:
    int nup_, ndown_;
    size_t size_;
  };

  explicit SzSymmetry(int N) : Ns_(N), comb_(N), current_(-1, -1, 0) { initial_fill(); }

  SzSymmetry(params& p) : Ns_(p["NSITES"].as<int>()), comb_(Ns_), current_(-1, -1, 0) {
    initial_fill();
    std::vector<Sector> secs;
    if (p.exists("arpack.SECTOR") && p["arpack.SECTOR"].as<bool>()) {
      alps_shim::input_file in(p["INPUT_FILE"].as<std::string>());
      const alps_shim::dataset& d = in.get("sectors/values");
      for (size_t q = 0; q + 1 < d.data.size(); q += 2) {
        int a = (int)d.data[q], b = (int)d.data[q + 1];
        secs.push_back(Sector(a, b, (size_t)comb_.c_n_k(Ns_, a) * comb_.c_n_k(Ns_, b)));
      }
    } else {
      for (int i = 0; i <= Ns_; ++i)
        for (int j = 0; j <= Ns_; ++j) secs.push_back(Sector(i, j, (size_t)comb_.c_n_k(Ns_, i) * comb_.c_n_k(Ns_, j)));
    }
    std::sort(secs.begin(), secs.end(), std::less<Sector>());
    for (const Sector& s : secs) sectors_.push(s);
  }

  bool next_sector() {
    if (sectors_.empty()) return false;
    current_ = sectors_.front();
    sectors_.pop();
    return true;
  }
  void set_sector(const Sector& s) {
    current_ = s;
    init();
  }
  const Sector& sector() const { return current_; }
  std::queue<Sector>& sectors() { return sectors_; }
  const Combination& comb() const { return comb_; }
  int sites() const { return Ns_; }

  void reset() {
    state_ = 0;
    ind_ = 0;
  }
  void init() { reset(); }
  bool next_state() {
    if (ind_ >= (int64_t)current_.size()) return false;
    state_ = state_by_index(ind_);
    ++ind_;
    return true;
  }
  long long state() const { return state_; }

  // SzSymmetry.h:114-121 (up-major packing)
  long long state_by_index(int64_t ind) const {
    const int64_t cdo = comb_.c_n_k(Ns_, current_.ndown());
    long long res = basis_[current_.nup()][ind / cdo];
    res <<= Ns_;
    res += basis_[current_.ndown()][ind % cdo];
    return res;
  }
  // SzSymmetry.h:123-133
  int64_t index(long long st, const Sector& s) const {
    const long long up = st >> Ns_, dn = st & ((1ll << Ns_) - 1);
    return (int64_t)ninv_[up] * comb_.c_n_k(Ns_, s.ndown()) + ninv_[dn];
  }
  int64_t index(long long st) const { return index(st, current_); }

  // SzSymmetry.h:169-185 (can_create_particle keeps the reference's bound).
  bool can_create_particle(int spin) const {
    return spin == 0 ? (current_.nup() - 1) < Ns_ : (current_.ndown() - 1) < Ns_;
  }
  bool can_destroy_particle(int spin) const { return spin == 0 ? current_.nup() > 0 : current_.ndown() > 0; }
  Sector destroy_particle(int spin) const {
    int a = current_.nup() - (1 - spin), b = current_.ndown() - spin;
    return Sector(a, b, (size_t)comb_.c_n_k(Ns_, a) * comb_.c_n_k(Ns_, b));
  }
  Sector create_particle(int spin) const {
    int a = current_.nup() + (1 - spin), b = current_.ndown() + spin;
    return Sector(a, b, (size_t)comb_.c_n_k(Ns_, a) * comb_.c_n_k(Ns_, b));
  }

 private:
  void initial_fill() {
    basis_.assign(Ns_ + 1, std::vector<int>());
    ninv_.assign((size_t)1 << Ns_, 0);
    std::vector<int> pos(Ns_ + 1);
    for (int k = 0; k <= Ns_; ++k) {
      comb_.init_state(k, pos);
      int r = 0;
      do {
        int bits = 0;
        for (int i = 0; i < k; ++i) bits |= 1 << pos[i];
        basis_[k].push_back(bits);
        ninv_[bits] = r++;
      } while (comb_.next_combination(Ns_, k, pos));
    }
  }
  int Ns_;
  Combination comb_;
  Sector current_;
  std::queue<Sector> sectors_;
  std::vector<std::vector<int>> basis_;
  std::vector<int> ninv_;  // a pattern's popcount fixes its per-spin sector, so one table serves all
  long long state_ = 0;
  int64_t ind_ = 0;
};

}  // namespace Symmetry
}  // namespace EDLib

#endif
