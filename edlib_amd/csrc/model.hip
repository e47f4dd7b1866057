// Host-side model builders and per-sector table construction.
//
// The transition lists follow HubbardModel.h:68-76 and
// SingleImpurityAndersonModel.h:232-274; matrix elements follow
// HubbardModel.h:86-108 and SingleImpurityAndersonModel.h:74-84,129-168; the
// per-spin CSR tables follow SpinResolvedStorage::fill_spin
// (SpinResolvedStorage.h:622-638) with CRSMatrix's merge/compress rules
// (CRSMatrix.h:46-98), so the stored tables equal the reference's exactly.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "edh_internal.h"

namespace edh {

Binomial::Binomial(int n_) : n(n_), t((n_ + 1) * (n_ + 1), 0) {
  for (int a = 0; a <= n; ++a) {
    t[a * (n + 1)] = 1;
    for (int b = 1; b <= a; ++b) t[a * (n + 1) + b] = t[(a - 1) * (n + 1) + b - 1] + (b < a ? t[(a - 1) * (n + 1) + b] : 0);
  }
}

// k-subsets of Ns bit positions, ordered by the lexicographic order of their
// ascending position tuples (the order of Combination::next_combination).
std::vector<int32_t> spin_basis(int Ns, int k) {
  std::vector<int32_t> out;
  if (k < 0 || k > Ns) return out;
  std::vector<int> pos(k);
  for (int i = 0; i < k; ++i) pos[i] = i;
  for (;;) {
    int32_t bits = 0;
    for (int i = 0; i < k; ++i) bits |= 1 << pos[i];
    out.push_back(bits);
    int i = k - 1;
    while (i >= 0 && pos[i] >= Ns - k + i) --i;
    if (i < 0) break;
    ++pos[i];
    for (int j = i + 1; j < k; ++j) pos[j] = pos[j - 1] + 1;
  }
  return out;
}

namespace {

inline int occ(state_t st, int orb, int Ip) { return (int)((st >> (Ip - 1 - orb)) & 1ll); }

// Sign of c_o / c^+_o: parity of occupied orbitals with index < o (FermionicModel.h:45-68).
inline int op_sign(state_t st, int orb, int Ip) {
  unsigned long long above = (unsigned long long)st >> (Ip - orb);
  return (__builtin_popcountll(above) & 1) ? -1 : 1;
}

}  // namespace

int Model::valid(const Transition& tr, state_t nst) const {
  const int Ip = 2 * Ns;
  if (tr.kind == 0) return occ(nst, tr.i + tr.s * Ns, Ip) * (1 - occ(nst, tr.j + tr.s * Ns, Ip));
  // c^+_i c^+_j c_l c_k applied right to left
  int ok = tr.k + tr.s * Ns, ol = tr.l + tr.sp * Ns, oj = tr.j + tr.sp * Ns, oi = tr.i + tr.s * Ns;
  if (!occ(nst, ok, Ip)) return 0;
  state_t s1 = nst & ~(1ll << (Ip - 1 - ok));
  if (!occ(s1, ol, Ip)) return 0;
  state_t s2 = s1 & ~(1ll << (Ip - 1 - ol));
  if (occ(s2, oj, Ip)) return 0;
  state_t s3 = s2 | (1ll << (Ip - 1 - oj));
  return 1 - occ(s3, oi, Ip);
}

double Model::apply(const Transition& tr, state_t nst, state_t& k, int& sign) const {
  const int Ip = 2 * Ns;
  if (tr.kind == 0) {
    int oi = tr.i + tr.s * Ns, oj = tr.j + tr.s * Ns;
    int s1 = op_sign(nst, oi, Ip);
    state_t k1 = nst - (1ll << (Ip - 1 - oi));
    int s2 = op_sign(k1, oj, Ip);
    k = k1 + (1ll << (Ip - 1 - oj));
    sign = hop_sign * s1 * s2;
    return tr.value;
  }
  int ok = tr.k + tr.s * Ns, ol = tr.l + tr.sp * Ns, oj = tr.j + tr.sp * Ns, oi = tr.i + tr.s * Ns;
  int g = op_sign(nst, ok, Ip);
  state_t a = nst - (1ll << (Ip - 1 - ok));
  g *= op_sign(a, ol, Ip);
  a -= (1ll << (Ip - 1 - ol));
  g *= op_sign(a, oj, Ip);
  a += (1ll << (Ip - 1 - oj));
  g *= op_sign(a, oi, Ip);
  a += (1ll << (Ip - 1 - oi));
  k = a;
  sign = g;
  return v_scale * tr.value;
}

double Model::diagonal_reference(state_t st) const {
  const int Ip = 2 * Ns;
  double x = 0.0;
  if (kind == EDH_MODEL_TERMS) {
    // the decomposed form the model was given in
    x = e0;
    for (int s = 0; s < 2; ++s)
      for (int i = 0; i < Ns; ++i) {
        if (!occ(st, i + s * Ns, Ip)) continue;
        x += e[s][i];
        for (int j = i + 1; j < Ns; ++j)
          if (occ(st, j + s * Ns, Ip)) x += S[s][i * Ns + j];
      }
    for (int i = 0; i < Ns; ++i)
      for (int j = 0; j < Ns; ++j)
        if (occ(st, i, Ip) && occ(st, j + Ns, Ip)) x += C[i * Ns + j];
    return x;
  }
  if (kind == EDH_MODEL_HUBBARD) {
    for (int im = 0; im < Ns; ++im) {
      for (int is = 0; is < 2; ++is) x += (0.0 - xmu[is]) * occ(st, im + is * Ns, Ip);
      x += U[im] * occ(st, im, Ip) * occ(st, im + Ns, Ip);
      x += Hmag[im] * (occ(st, im + Ns, Ip) - occ(st, im, Ip));
      for (int im2 = 0; im2 < Ns; ++im2)
        x += J[im * Ns + im2] * (occ(st, im, Ip) - occ(st, im + Ns, Ip)) * (occ(st, im2, Ip) - occ(st, im2 + Ns, Ip));
    }
    return x;
  }
  int nb = Ns - ml;
  for (int is = 0; is < 2; ++is)
    for (int ik = 0; ik < nb; ++ik) x += Epsk[ik * 2 + is] * occ(st, ik + ml + is * Ns, Ip);
  for (int im = 0; im < ml; ++im) {
    for (int is = 0; is < 2; ++is) x += (H0[(im * ml + im) * 2 + is] - mu) * occ(st, im + is * Ns, Ip);
    for (int is = 0; is < 2; ++is) x += 0.5 * uten(is, is, im, im, im, im) * occ(st, im, Ip) * occ(st, im + Ns, Ip);
    for (int jm = 0; jm < ml; ++jm)
      for (int is = 0; is < 2; ++is)
        if (im != jm) {
          x += 0.5 * (uten(is, is, im, jm, im, jm) - uten(is, is, im, jm, jm, im)) * occ(st, im + is * Ns, Ip) *
               occ(st, jm + is * Ns, Ip);
          x += 0.5 * uten(is, 1 - is, im, jm, im, jm) * occ(st, im + is * Ns, Ip) * occ(st, jm + (1 - is) * Ns, Ip);
        }
  }
  return x;
}

Model* build_hubbard(int Ns, const double* t, const double* U, const double* J, const double* xmu,
                     const double* Hmag) {
  EDH_CHECK(Ns >= 1 && Ns <= 24, EDH_ERR_ARG, "Hubbard model: NSITES must be in [1, 24]");
  Model* m = new Model();
  m->kind = EDH_MODEL_HUBBARD;
  m->Ns = Ns;
  m->ml = Ns;
  m->hop_sign = -1;  // -t c^+ c (HubbardModel.h:105)
  m->t.assign(t, t + Ns * Ns);
  m->U.assign(U, U + Ns);
  m->J.assign(Ns * Ns, 0.0);
  if (J) m->J.assign(J, J + Ns * Ns);
  m->xmu.assign(xmu, xmu + Ns);
  m->Hmag.assign(Ns, 0.0);
  if (Hmag) m->Hmag.assign(Hmag, Hmag + Ns);
  for (int ii = 0; ii < Ns; ++ii)
    for (int jj = 0; jj < Ns; ++jj)
      if (std::abs(m->t[ii * Ns + jj]) > 1e-10)
        for (int is = 0; is < 2; ++is) m->T.push_back(Transition{0, ii, jj, 0, 0, is, 0, m->t[ii * Ns + jj]});
  // Diagonal decomposition of HubbardModel::diagonal (HubbardModel.h:116-132).
  // The chemical potential is indexed by spin there (_xmu[is], :120): kept.
  for (int s = 0; s < 2; ++s) {
    m->e[s].assign(Ns, 0.0);
    m->S[s].assign(Ns * Ns, 0.0);
  }
  m->C.assign(Ns * Ns, 0.0);
  for (int i = 0; i < Ns; ++i) {
    const double Jii = m->J[i * Ns + i];
    m->e[0][i] += -m->xmu[0] - m->Hmag[i] + Jii;
    m->e[1][i] += -m->xmu[1] + m->Hmag[i] + Jii;
    m->C[i * Ns + i] += m->U[i] - 2.0 * Jii;
    for (int j = 0; j < Ns; ++j) {
      if (i == j) continue;
      const double Jij = m->J[i * Ns + j];
      m->S[0][i * Ns + j] += Jij;
      m->S[1][i * Ns + j] += Jij;
      m->C[i * Ns + j] -= Jij;  // n_i(up) n_j(dn)
      m->C[j * Ns + i] -= Jij;  // n_i(dn) n_j(up)
    }
  }
  return m;
}

Model* build_siam(int Ns, int ml, const double* Epsk, const double* Vk, const double* H0, const double* Ut,
                  double mu) {
  EDH_CHECK(Ns >= 1 && Ns <= 24, EDH_ERR_ARG, "SIAM: NSITES must be in [1, 24]");
  EDH_CHECK(ml >= 1 && ml <= Ns, EDH_ERR_ARG,
            "Incorrect values for the total number of sites and the number of orbitals.");
  EDH_CHECK(ml <= 6, EDH_ERR_LIMIT, "SIAM: at most 6 impurity orbitals are supported on the device");
  Model* m = new Model();
  m->kind = EDH_MODEL_SIAM;
  m->Ns = Ns;
  m->ml = ml;
  m->hop_sign = +1;  // SingleImpurityAndersonModel.h:83
  const int nb = Ns - ml;
  m->Epsk.assign(Epsk, Epsk + nb * 2);
  m->Vk.assign(Vk, Vk + ml * nb * 2);
  m->H0.assign(H0, H0 + ml * ml * 2);
  m->Ut.assign(Ut, Ut + 4 * ml * ml * ml * ml);
  m->mu = mu;
  // T_states: inter-orbital hoppings, then hybridisation (SingleImpurityAndersonModel.h:232-254).
  for (int im = 0; im < ml; ++im)
    for (int jm = 0; jm < im; ++jm)
      for (int is = 0; is < 2; ++is) {
        double h = m->H0[(im * ml + jm) * 2 + is];
        if (std::abs(h) > 1e-10) {
          m->T.push_back(Transition{0, im, jm, 0, 0, is, 0, h});
          m->T.push_back(Transition{0, jm, im, 0, 0, is, 0, h});
        }
      }
  for (int im = 0; im < ml; ++im)
    for (int ik = 0; ik < nb; ++ik)
      for (int is = 0; is < 2; ++is) {
        double v = m->Vk[(im * nb + ik) * 2 + is];
        if (std::abs(v) > 1e-10) {
          m->T.push_back(Transition{0, im, ik + ml, 0, 0, is, 0, v});
          m->T.push_back(Transition{0, ik + ml, im, 0, 0, is, 0, v});
        }
      }
  // V_states, skipping density-density terms (SingleImpurityAndersonModel.h:256-274).
  for (int s1 = 0; s1 < 2; ++s1)
    for (int s2 = 0; s2 < 2; ++s2)
      for (int i = 0; i < ml; ++i)
        for (int j = 0; j < ml; ++j)
          for (int k = 0; k < ml; ++k)
            for (int l = 0; l < ml; ++l) {
              if (((i == l) && (j == k) && (s1 == s2)) || ((i == k) && (j == l))) continue;
              double u = m->uten(s1, s2, i, j, k, l);
              if (std::abs(u) != 0.0) m->V.push_back(Transition{1, i, j, k, l, s1, s2, u});
            }
  // Diagonal decomposition of SingleImpurityAndersonModel::diagonal (:282-306).
  for (int s = 0; s < 2; ++s) {
    m->e[s].assign(Ns, 0.0);
    m->S[s].assign(Ns * Ns, 0.0);
  }
  m->C.assign(Ns * Ns, 0.0);
  for (int s = 0; s < 2; ++s)
    for (int ik = 0; ik < nb; ++ik) m->e[s][ik + ml] += m->Epsk[ik * 2 + s];
  for (int im = 0; im < ml; ++im) {
    for (int s = 0; s < 2; ++s) m->e[s][im] += m->H0[(im * ml + im) * 2 + s] - mu;
    m->C[im * Ns + im] += 0.5 * (m->uten(0, 0, im, im, im, im) + m->uten(1, 1, im, im, im, im));
    for (int jm = 0; jm < ml; ++jm) {
      if (im == jm) continue;
      for (int s = 0; s < 2; ++s)
        m->S[s][im * Ns + jm] += 0.5 * (m->uten(s, s, im, jm, im, jm) - m->uten(s, s, im, jm, jm, im));
      m->C[im * Ns + jm] += 0.5 * m->uten(0, 1, im, jm, im, jm);  // n_im(up) n_jm(dn)
      m->C[jm * Ns + im] += 0.5 * m->uten(1, 0, im, jm, im, jm);  // n_im(dn) n_jm(up)
    }
  }
  return m;
}

Model* build_terms(const edh_model_desc& d) {
  const int Ns = d.Ns;
  EDH_CHECK(Ns >= 1 && Ns <= 24, EDH_ERR_ARG, "model terms: NSITES must be in [1, 24]");
  EDH_CHECK(d.n_hop >= 0 && d.n_int >= 0, EDH_ERR_ARG, "model terms: negative term count");
  EDH_CHECK((d.n_hop == 0 || (d.hop && d.hop_value)) && (d.n_int == 0 || (d.inter && d.inter_value)) && d.eps &&
                d.same && d.cross,
            EDH_ERR_ARG, "model terms: null array");
  const int ml = d.interacting_orbitals;
  EDH_CHECK(ml >= 0 && ml <= Ns, EDH_ERR_ARG, "model terms: interacting_orbitals out of range");
  std::unique_ptr<Model> m(new Model());
  m->kind = EDH_MODEL_TERMS;
  m->Ns = Ns;
  m->ml = ml;
  m->hop_sign = +1;  // the sign rule is folded into the values
  m->v_scale = 1.0;  // two-body values are the coefficients themselves
  for (int q = 0; q < d.n_hop; ++q) {
    const int* h = d.hop + 3 * q;
    EDH_CHECK(h[0] >= 0 && h[0] < Ns && h[1] >= 0 && h[1] < Ns && (h[2] == 0 || h[2] == 1), EDH_ERR_ARG,
              "model terms: one-body term index out of range");
    EDH_CHECK(h[0] != h[1], EDH_ERR_ARG, "model terms: a one-body term must move the electron (i != j)");
    m->T.push_back(Transition{0, h[0], h[1], 0, 0, h[2], 0, d.hop_value[q]});
  }
  for (int q = 0; q < d.n_int; ++q) {
    const int* v = d.inter + 6 * q;
    for (int a = 0; a < 4; ++a)
      EDH_CHECK(v[a] >= 0 && v[a] < ml, EDH_ERR_ARG, "model terms: two-body term outside the interacting orbitals");
    EDH_CHECK((v[4] == 0 || v[4] == 1) && (v[5] == 0 || v[5] == 1), EDH_ERR_ARG, "model terms: spin index");
    m->V.push_back(Transition{1, v[0], v[1], v[2], v[3], v[4], v[5], d.inter_value[q]});
  }
  EDH_CHECK(m->V.empty() || ml <= 6, EDH_ERR_LIMIT, "model terms: at most 6 interacting orbitals with two-body terms");
  m->e0 = d.e0;
  for (int s = 0; s < 2; ++s) {
    m->e[s].assign(d.eps + s * Ns, d.eps + (s + 1) * Ns);
    m->S[s].assign(Ns * Ns, 0.0);
    for (int i = 0; i < Ns; ++i)
      for (int j = i + 1; j < Ns; ++j) m->S[s][i * Ns + j] = d.same[(s * Ns + i) * Ns + j];
  }
  m->C.assign(d.cross, d.cross + Ns * Ns);
  return m.release();
}

// ---------------------------------------------------------------------------
// Per-sector tables.
// ---------------------------------------------------------------------------
namespace {

// CSR row builder with the reference's merge/compress semantics (CRSMatrix.h:46-98).
struct RowBuilder {
  std::vector<int32_t>& col;
  std::vector<double>& val;
  size_t row_start;
  void add(int32_t j, double t, int sign) {
    if (std::abs(t) == 0) return;
    long found = -1;
    for (size_t q = row_start; q < col.size(); ++q)
      if (col[q] == j) found = (long)q;
    if (found >= 0) {
      val[found] += sign * t;
    } else {
      col.push_back(j);
      val.push_back(sign * t);
    }
  }
  void end() {
    // Removing |value| < 1e-15 shifts the rest left and skips the element that
    // lands in the freed slot, exactly as CRSMatrix::compress does.
    size_t q = row_start;
    while (q < col.size()) {
      if (std::abs(val[q]) < 1e-15) {
        col.erase(col.begin() + q);
        val.erase(val.begin() + q);
        ++q;  // the element now at q is not examined
      } else {
        ++q;
      }
    }
  }
};

int value_index(std::vector<double>& values, std::map<uint64_t, int>& idx, double v) {
  uint64_t key;
  std::memcpy(&key, &v, 8);
  auto it = idx.find(key);
  if (it != idx.end()) return it->second;
  EDH_CHECK((int)values.size() < kMaxValues, EDH_ERR_LIMIT, "more than 1024 distinct hopping values");
  int id = (int)values.size();
  values.push_back(v);
  idx[key] = id;
  return id;
}

void build_spin(const Model& m, int spin, int nel, SpinTables& sp, std::vector<double>& values,
                std::map<uint64_t, int>& vidx, bool need_rank) {
  const int Ns = m.Ns;
  sp = SpinTables();  // no table of an earlier sector or model may survive
  sp.nel = nel;
  sp.bits = spin_basis(Ns, nel);
  sp.D = (int64_t)sp.bits.size();
  EDH_CHECK(sp.D <= (int64_t)kColMask, EDH_ERR_LIMIT, "per-spin dimension exceeds 2^22");
  std::vector<int32_t> rank(1u << Ns, -1);
  for (int64_t r = 0; r < sp.D; ++r) rank[sp.bits[r]] = (int32_t)r;
  // single-spin diagonal part
  sp.ediag.assign(sp.D, 0.0);
  for (int64_t r = 0; r < sp.D; ++r) {
    const int32_t b = sp.bits[r];
    double x = 0.0;
    for (int i = 0; i < Ns; ++i) {
      if (!((b >> (Ns - 1 - i)) & 1)) continue;
      x += m.e[spin][i];
      for (int j = 0; j < Ns; ++j)
        if (j != i && ((b >> (Ns - 1 - j)) & 1)) x += m.S[spin][i * Ns + j];
    }
    sp.ediag[r] = spin == 0 ? x + m.e0 : x;
  }
  // hopping CSR: rows are source states, columns the targets (fill_spin)
  const int shift = spin == 0 ? Ns : 0;
  sp.rowp.assign(sp.D + 1, 0);
  sp.col.clear();
  sp.val.clear();
  for (int64_t r = 0; r < sp.D; ++r) {
    RowBuilder rb{sp.col, sp.val, sp.col.size()};
    const state_t nst = (state_t)sp.bits[r] << shift;
    for (const Transition& tr : m.T) {
      if (m.valid(tr, nst)) {
        state_t k;
        int sg;
        double v = m.apply(tr, nst, k, sg);
        int32_t j = rank[(int32_t)(k >> shift)];
        rb.add(j, v, sg);
      }
    }
    rb.end();
    sp.rowp[r + 1] = (int32_t)sp.col.size();
  }
  sp.packed.resize(sp.col.size());
  for (size_t q = 0; q < sp.col.size(); ++q)
    sp.packed[q] = (uint32_t)sp.col[q] | ((uint32_t)value_index(values, vidx, sp.val[q]) << kColBits);
  // Sliced ELL: slice s holds rows [64s, 64s+64), entry k of lane L at
  // ptr + k*64 + L.  Within a row the entries are reordered so that, slot by
  // slot, each 16-lane group gathers from distinct 8-byte LDS banks (column
  // mod 16): greedy over lanes, picking the pending entry whose bank is least
  // used in the slot so far.  Padding entries read one shared zero slot
  // (broadcast), so they never conflict.
  int64_t nsl = (sp.D + kSlice - 1) / kSlice;
  sp.sl_ptr.assign(nsl + 1, 0);
  sp.sl_len.assign(nsl, 0);
  for (int64_t s = 0; s < nsl; ++s) {
    int len = 0;
    for (int64_t r = s * kSlice; r < std::min<int64_t>(sp.D, (s + 1) * kSlice); ++r)
      len = std::max(len, sp.rowp[r + 1] - sp.rowp[r]);
    sp.sl_len[s] = len;
    sp.sl_ptr[s + 1] = sp.sl_ptr[s] + len * kSlice;
  }
  sp.ell.assign(sp.sl_ptr[nsl], 0u);
  sp.ell_pad.assign(sp.sl_ptr[nsl], 1);
  std::vector<std::vector<int>> pending(kSlice);
  for (int64_t s = 0; s < nsl; ++s) {
    for (int L = 0; L < kSlice; ++L) {
      pending[L].clear();
      const int64_t r = s * kSlice + L;
      if (r < sp.D)
        for (int q = sp.rowp[r]; q < sp.rowp[r + 1]; ++q) pending[L].push_back(q);
    }
    for (int k = 0; k < sp.sl_len[s]; ++k) {
      int used[4][16] = {{0}};
      for (int L = 0; L < kSlice; ++L) {
        if (pending[L].empty()) continue;
        int* u = used[L >> 4];
        size_t best = 0;
        for (size_t c = 1; c < pending[L].size(); ++c)
          if (u[sp.col[pending[L][c]] & 15] < u[sp.col[pending[L][best]] & 15]) best = c;
        const int q = pending[L][best];
        pending[L].erase(pending[L].begin() + best);
        ++u[sp.col[q] & 15];
        sp.ell[sp.sl_ptr[s] + k * kSlice + L] = sp.packed[q];
        sp.ell_pad[sp.sl_ptr[s] + k * kSlice + L] = 0;
      }
    }
  }
  // padding entries of the 32-bit ELL: column 0 with value index 0 (= 0.0)
  // Compact ELL when the spin has < 16384 states and <= 4 distinct values:
  // padding entries point at column D, a zero slot the kernel appends to the
  // row it stages in LDS.
  sp.ell16.clear();
  {
    std::vector<double> vt;
    bool ok = sp.D < 16384;
    for (size_t q = 0; q < sp.val.size() && ok; ++q)
      if (std::find(vt.begin(), vt.end(), sp.val[q]) == vt.end()) {
        if (vt.size() == 4) ok = false;
        else vt.push_back(sp.val[q]);
      }
    if (ok) {
      for (int q = 0; q < 4; ++q) sp.vt16[q] = q < (int)vt.size() ? vt[q] : 0.0;
      sp.nvt16 = (int)vt.size();
      sp.ell16.assign(sp.ell.size(), (uint16_t)sp.D);
      for (size_t q = 0; q < sp.ell.size(); ++q) {
        if (sp.ell_pad[q]) continue;
        const uint32_t e = sp.ell[q];
        const double v = values[e >> kColBits];
        const int vi = (int)(std::find(vt.begin(), vt.end(), v) - vt.begin());
        sp.ell16[q] = (uint16_t)((e & kColMask) | ((uint32_t)vi << 14));
      }
      const int64_t nsl = (int64_t)sp.sl_len.size();
      sp.q4_ptr.assign(nsl + 1, 0);
      for (int64_t sl = 0; sl < nsl; ++sl) sp.q4_ptr[sl + 1] = sp.q4_ptr[sl] + (sp.sl_len[sl] + 3) / 4 * kSlice;
      const uint64_t pad = (uint64_t)(uint16_t)sp.D;
      sp.ell_q4.assign(sp.q4_ptr[nsl], pad | (pad << 16) | (pad << 32) | (pad << 48));
      for (int64_t sl = 0; sl < nsl; ++sl)
        for (int k = 0; k < sp.sl_len[sl]; ++k)
          for (int L = 0; L < kSlice; ++L) {
            const uint64_t e = sp.ell16[sp.sl_ptr[sl] + k * kSlice + L];
            uint64_t& w = sp.ell_q4[sp.q4_ptr[sl] + (k / 4) * kSlice + L];
            const int sh = 16 * (k % 4);
            w = (w & ~(0xFFFFull << sh)) | (e << sh);
          }
    }
  }
  {
    std::vector<double> vt;
    for (size_t q = 0; q < sp.val.size() && vt.size() <= 2; ++q)
      if (std::find(vt.begin(), vt.end(), sp.val[q]) == vt.end()) vt.push_back(sp.val[q]);
    const bool ok = !vt.empty() && (vt.size() == 1 || (vt.size() == 2 && vt[1] == -vt[0])) &&
                    sp.D < (int64_t(1) << 28);
    if (ok) {
      const int64_t nsl = (int64_t)sp.sl_len.size();
      std::vector<int32_t> qp(nsl + 1, 0);
      for (int64_t sl = 0; sl < nsl; ++sl) qp[sl + 1] = qp[sl] + (sp.sl_len[sl] + 3) / 4 * kSlice;
      if (sp.q4_ptr.empty()) sp.q4_ptr = qp;
      sp.s4_scale = vt[0];
      sp.ell_s4.assign((size_t)qp[nsl] * 4, (uint32_t)(sp.D * 8));
      for (int64_t sl = 0; sl < nsl; ++sl)
        for (int k = 0; k < sp.sl_len[sl]; ++k)
          for (int L = 0; L < kSlice; ++L) {
            const size_t q = sp.sl_ptr[sl] + (size_t)k * kSlice + L;
            if (sp.ell_pad[q]) continue;
            const uint32_t e = sp.ell[q];
            const double v = values[e >> kColBits];
            sp.ell_s4[((size_t)qp[sl] + (k / 4) * kSlice + L) * 4 + (k % 4)] =
                ((e & kColMask) << 3) | (v == vt[0] ? 0u : 1u);
          }
    }
  }
  if (need_rank) sp.rank = std::move(rank);
}

}  // namespace

void build_sector(const Model& m, int nup, int ndn, SectorTables& st, bool need_rank) {
  const int Ns = m.Ns;
  EDH_CHECK(nup >= 0 && nup <= Ns && ndn >= 0 && ndn <= Ns, EDH_ERR_ARG, "sector out of range");
  st = SectorTables();
  st.nup = nup;
  st.ndn = ndn;
  st.imp_ml = m.ml;
  st.values.assign(1, 0.0);
  std::map<uint64_t, int> vidx;
  {
    double z = 0.0;
    uint64_t key;
    std::memcpy(&key, &z, 8);
    vidx[key] = 0;
  }
  const bool kan = !m.V.empty();
  build_spin(m, 0, nup, st.up, st.values, vidx, need_rank || kan);
  build_spin(m, 1, ndn, st.dn, st.values, vidx, need_rank || kan);
  st.N = st.up.D * st.dn.D;
  st.table_bytes = 12ll * (int64_t)(st.up.col.size() + st.dn.col.size()) + 4ll * (st.up.D + st.dn.D + 2);

  // Cross (up x down) part of the diagonal.
  bool any = false, diag_only = true, uniform = true, imp_only = true;
  const double u0 = m.C[0];
  for (int i = 0; i < Ns; ++i)
    for (int j = 0; j < Ns; ++j) {
      double c = m.C[i * Ns + j];
      if (c != 0.0) any = true;
      if (i != j && c != 0.0) diag_only = false;
      if (i == j && c != u0) uniform = false;
      if ((i >= m.ml || j >= m.ml) && c != 0.0) imp_only = false;
    }
  st.cross_imp.clear();
  st.cross_rowtab.clear();
  if (!any) {
    st.cross_mode = CROSS_NONE;
  } else if (diag_only && uniform) {
    st.cross_mode = CROSS_POPC;
    st.cross_u = u0;
  } else if (imp_only && m.ml <= 6) {
    st.cross_mode = CROSS_IMP;
    const int ml = m.ml, na = 1 << ml;
    st.cross_imp.assign(na * na, 0.0);
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < na; ++b) {
        double x = 0.0;
        for (int i = 0; i < ml; ++i)
          for (int j = 0; j < ml; ++j)
            if (((a >> (ml - 1 - i)) & 1) && ((b >> (ml - 1 - j)) & 1)) x += m.C[i * Ns + j];
        st.cross_imp[a * na + b] = x;
      }
  } else {
    st.cross_mode = CROSS_GENERAL;
    // R_u[p] for down bit position p: sum over occupied up sites i of C[i][site(p)]
    st.cross_rowtab.assign(st.up.D * Ns, 0.0);
    for (int64_t u = 0; u < st.up.D; ++u)
      for (int p = 0; p < Ns; ++p) {
        const int j = Ns - 1 - p;
        double x = 0.0;
        for (int i = 0; i < Ns; ++i)
          if ((st.up.bits[u] >> (Ns - 1 - i)) & 1) x += m.C[i * Ns + j];
        st.cross_rowtab[u * Ns + p] = x;
      }
  }

  // Off-diagonal interaction (V_states) acting on impurity bits only.  The
  // fermion sign of each term depends on the impurity bits and on the parity
  // of the up-electron count (all up orbitals precede the down ones), so one
  // representative state per impurity configuration fixes the whole table.
  st.kan_ml = 0;
  st.kan_ptr.clear();
  st.kan.clear();
  if (kan) {
    const int ml = m.ml, na = 1 << ml, nb = Ns - ml;
    for (const Transition& tr : m.V)
      EDH_CHECK(tr.i < ml && tr.j < ml && tr.k < ml && tr.l < ml, EDH_ERR_ARG, "interaction outside impurity");
    st.kan_ml = ml;
    st.kan_ptr.assign(na * na + 1, 0);
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < na; ++b) {
        const int idx = a * na + b;
        st.kan_ptr[idx + 1] = st.kan_ptr[idx];
        int bath_up = nup - __builtin_popcount(a), bath_dn = ndn - __builtin_popcount(b);
        if (bath_up < 0 || bath_up > nb || bath_dn < 0 || bath_dn > nb) continue;
        // representative: bath electrons in the lowest bath bits
        state_t up = ((state_t)a << nb) | ((1ll << bath_up) - 1);
        state_t dn = ((state_t)b << nb) | ((1ll << bath_dn) - 1);
        state_t nst = (up << Ns) | dn;
        std::vector<KanEntry> row;
        for (const Transition& tr : m.V) {
          if (!m.valid(tr, nst)) continue;
          state_t k;
          int sg;
          double v = m.apply(tr, nst, k, sg);
          if (std::abs(v) == 0) continue;
          int32_t na_ = (int32_t)((k >> Ns) >> nb), nb_ = (int32_t)((k & ((1ll << Ns) - 1)) >> nb);
          bool merged = false;
          for (KanEntry& e : row)
            if (e.new_a == na_ && e.new_b == nb_) {
              e.coef += sg * v;
              merged = true;
            }
          if (!merged) row.push_back(KanEntry{na_, nb_, sg * v});
        }
        for (const KanEntry& e : row)
          if (std::abs(e.coef) >= 1e-15) st.kan.push_back(e);
        st.kan_ptr[idx + 1] = (int32_t)st.kan.size();
      }
  }
}

}  // namespace edh
