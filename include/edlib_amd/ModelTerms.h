// Engine model from any model class with the reference models' storage-facing
// interface -- the calls SpinResolvedStorage::fill / fill_spin make
// (SpinResolvedStorage.h:120-155,622-638):
//   m.T_states(), m.V_states()          term lists (HubbardModel.h:141-147,
//                                        SingleImpurityAndersonModel.h:365-375)
//   m.valid(term, state), m.set(term, state, k, sign) -> value
//   m.diagonal(state), m.interacting_orbitals()
// Nothing else of the model is read, so the reference's own HubbardModel and
// SingleImpurityAndersonModel serve as they are.
//
// The terms are recovered by applying them to few-particle states:
//   * a one-body term acts on exactly one one-particle state |o>, giving
//     |o'>; its coefficient on that state (both fermion signs are +1 there)
//     is sign * value, the model's sign rule included;
//   * a two-body term acts on exactly one two-particle state (the pair it
//     annihilates) or on none (a vanishing term); it is written as
//     c^+_i c^+_j c_l c_k with (i, j) the created pair in ascending orbital
//     order, and its coefficient is fixed so that the engine's signs on that
//     state reproduce the model's;
//   * the diagonal is a quadratic form of the occupations; its constant,
//     linear and pair parts come from diagonal() on the empty, one- and
//     two-particle states.
// Then every recovered term and the diagonal are compared with the model on
// a fixed sample of many-particle states; a model that is not of this form
// (a term that is not a one- or two-body operator, a diagonal with a
// three-body part, Sz not conserved) is rejected with std::invalid_argument.
#ifndef EDLIB_AMD_MODELTERMS_H
#define EDLIB_AMD_MODELTERMS_H

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "edhip.h"

namespace EDLib {
namespace Storage {

inline void edh_check(int code) {
  if (code != EDH_OK) throw std::runtime_error(std::string("edhip: ") + edh_last_error());
}

struct EngineModelFree {
  void operator()(edh_model* m) const { edh_model_free(m); }
};
typedef std::unique_ptr<edh_model, EngineModelFree> EngineModelPtr;

namespace detail {

inline int popcount64(long long x) { return __builtin_popcountll((unsigned long long)x); }

// Sign of c_o or c^+_o on st: parity of the occupied orbitals before o
// (orbital o at bit Ip-1-o; FermionicModel.h:45-68).
inline int op_sign(long long st, int o, int Ip) {
  return (popcount64((long long)((unsigned long long)st >> (Ip - o))) & 1) ? -1 : 1;
}

// Sign of c^+_i c^+_j c_l c_k on st, applied right to left
// (InnerInteractionState::set, SingleImpurityAndersonModel.h:156-163).
inline int four_op_sign(long long st, int oi, int oj, int ok, int ol, int Ip) {
  int g = op_sign(st, ok, Ip);
  long long a = st - (1ll << (Ip - 1 - ok));
  g *= op_sign(a, ol, Ip);
  a -= 1ll << (Ip - 1 - ol);
  g *= op_sign(a, oj, Ip);
  a += 1ll << (Ip - 1 - oj);
  g *= op_sign(a, oi, Ip);
  return g;
}

inline int orbital_of_bit(long long single, int Ip) {
  int o = 0;
  while (o < Ip && !((single >> (Ip - 1 - o)) & 1ll)) ++o;
  return o;
}

}  // namespace detail

// Builds the engine's copy of model m (Ns sites per spin).
template <class Model>
EngineModelPtr engine_model_from_interface(Model& m, int Ns) {
  using detail::orbital_of_bit;
  const int Ip = 2 * Ns;
  if (Ns < 1 || Ns > 24) throw std::invalid_argument("model terms: NSITES must be in [1, 24]");
  auto bit = [Ip](int o) { return 1ll << (Ip - 1 - o); };
  std::vector<int> hop, inter;
  std::vector<double> hop_value, inter_value;

  // one-body terms
  for (const auto& term : m.T_states()) {
    int found = 0;
    for (int o = 0; o < Ip; ++o) {
      const long long nst = bit(o);
      if (!m.valid(term, nst)) continue;
      long long k = 0;
      int sign = 0;
      const double value = m.set(term, nst, k, sign);
      if (detail::popcount64(k) != 1) throw std::invalid_argument("model terms: a T_states term is not a one-body hop");
      const int o2 = orbital_of_bit(k, Ip);
      if (o / Ns != o2 / Ns) throw std::invalid_argument("model terms: a T_states term does not conserve Sz");
      if (o2 == o) throw std::invalid_argument("model terms: a T_states term does not move the electron");
      hop.insert(hop.end(), {o % Ns, o2 % Ns, o / Ns});
      hop_value.push_back(sign * value);
      ++found;
    }
    if (found != 1) throw std::invalid_argument("model terms: a T_states term is not a one-body hop");
  }

  // two-body terms
  for (const auto& term : m.V_states()) {
    int found = 0;
    for (int o1 = 0; o1 < Ip && found == 0; ++o1)
      for (int o2 = o1 + 1; o2 < Ip && found == 0; ++o2) {
        const long long nst = bit(o1) | bit(o2);
        if (!m.valid(term, nst)) continue;
        long long k = 0;
        int sign = 0;
        const double value = m.set(term, nst, k, sign);
        if (detail::popcount64(k) != 2) throw std::invalid_argument("model terms: a V_states term is not two-body");
        const int c1 = orbital_of_bit(k, Ip), c2 = orbital_of_bit(k & ~bit(c1), Ip);
        const int si = c1 / Ns, sj = c2 / Ns;
        // k pairs with i (spin si), l with j (spin sj)
        int ok = o1, ol = o2;
        if (o1 / Ns != si || o2 / Ns != sj) std::swap(ok, ol);
        if (ok / Ns != si || ol / Ns != sj) throw std::invalid_argument("model terms: a V_states term does not conserve Sz");
        const int g = detail::four_op_sign(nst, c1, c2, ok, ol, Ip);
        inter.insert(inter.end(), {c1 % Ns, c2 % Ns, ok % Ns, ol % Ns, si, sj});
        inter_value.push_back(g * sign * value);
        ++found;
      }
  }

  // diagonal as a quadratic form of the occupations
  const double e0 = m.diagonal(0ll);
  std::vector<double> e1(Ip), eps(Ip), same(2 * Ns * Ns, 0.0), cross(Ns * Ns, 0.0);
  for (int o = 0; o < Ip; ++o) e1[o] = m.diagonal(bit(o));
  for (int o = 0; o < Ip; ++o) eps[o] = e1[o] - e0;
  for (int o1 = 0; o1 < Ip; ++o1)
    for (int o2 = o1 + 1; o2 < Ip; ++o2) {
      const double w = m.diagonal(bit(o1) | bit(o2)) - e1[o1] - e1[o2] + e0;
      if (o1 / Ns == o2 / Ns) {
        same[(o1 / Ns * Ns + o1 % Ns) * Ns + o2 % Ns] = w;
      } else {
        cross[(o1 % Ns) * Ns + o2 % Ns] = w;  // o1 spin up, o2 spin down
      }
    }

  edh_model_desc d;
  d.Ns = Ns;
  d.interacting_orbitals = m.interacting_orbitals();
  d.n_hop = (int)hop_value.size();
  d.hop = hop.data();
  d.hop_value = hop_value.data();
  d.n_int = (int)inter_value.size();
  d.inter = inter.data();
  d.inter_value = inter_value.data();
  d.e0 = e0;
  d.eps = eps.data();
  d.same = same.data();
  d.cross = cross.data();
  edh_model* raw = nullptr;
  edh_check(edh_model_from_terms(&raw, &d));
  EngineModelPtr em(raw);

  // check on a fixed sample of states of every particle number
  uint64_t x = 0x9E3779B97F4A7C15ull;
  auto next = [&x]() {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const long long mask = (Ip == 64) ? -1ll : (long long)((1ull << Ip) - 1);
  std::vector<long long> tgt(64);
  std::vector<double> cf(64);
  for (int q = 0; q < 256; ++q) {
    long long st = (long long)next() & mask;
    if (q < 2 * Ns + 1) {  // one state of each particle number first
      st = 0;
      for (int o = 0; o < q; ++o) st |= bit(o);
    }
    double dref = m.diagonal(st), deng = 0.0;
    edh_check(edh_model_diagonal(em.get(), st, &deng));
    if (std::abs(dref - deng) > 1e-11 * (1.0 + std::abs(dref)))
      throw std::invalid_argument("model terms: diagonal() is not a quadratic form of the occupations");
    std::map<long long, double> want;
    for (const auto& term : m.T_states())
      if (m.valid(term, st)) {
        long long k = 0;
        int sign = 0;
        const double v = m.set(term, st, k, sign);
        want[k] += sign * v;
      }
    for (const auto& term : m.V_states())
      if (m.valid(term, st)) {
        long long k = 0;
        int sign = 0;
        const double v = m.set(term, st, k, sign);
        want[k] += sign * v;
      }
    int cnt = 0;
    edh_check(edh_model_apply(em.get(), st, (int)tgt.size(), tgt.data(), cf.data(), &cnt));
    if (cnt > (int)tgt.size()) {
      tgt.resize(cnt);
      cf.resize(cnt);
      edh_check(edh_model_apply(em.get(), st, cnt, tgt.data(), cf.data(), &cnt));
    }
    std::map<long long, double> got;
    for (int c = 0; c < cnt; ++c) got[tgt[c]] += cf[c];
    for (const auto& kv : want) {
      const double g = got.count(kv.first) ? got[kv.first] : 0.0;
      if (std::abs(g - kv.second) > 1e-12 * (1.0 + std::abs(kv.second)))
        throw std::invalid_argument("model terms: an off-diagonal term is not reproduced");
    }
    for (const auto& kv : got)
      if (!want.count(kv.first) && std::abs(kv.second) > 1e-12)
        throw std::invalid_argument("model terms: an off-diagonal term is not reproduced");
  }
  return em;
}

}  // namespace Storage
}  // namespace EDLib

#endif
