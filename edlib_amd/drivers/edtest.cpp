// Drop-in checks written like the reference's own tests, on the engine:
//   gfchi   -- test/Lanczos_Test.cpp:29-121 (G vs gom1.dat, chi_Sz vs xiats.dat,
//              chi_N vs xiatd.dat) with HIPSpinResolvedStorage in place of the
//              CSR storage;
//   4ring   -- test/HubbardModel_Test.cpp:27-71 (E0, ground sector, observables);
//   4ring_nomag -- test/StaticObservables_Test.cpp:27-72;
//   dyson_hubbard0, dyson_hubbard, dyson_siam0 -- the self-energy from the
//              Dyson equation (GreensFunction.h:202-223), see dyson();
//   storage  -- the storage's corners through its public members: every
//              one-state sector's pair is (its diagonal element, [1])
//              (zero_eigenapair, SpinResolvedStorage.h:52-56), and
//              constant_shift adds shift * v to av until the next fill
//              (SpinResolvedStorage.h:464-466).
// Usage: edtest <case> <golden dir>.  Prints one JSON line; exit code 0 on pass.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "edlib_amd/GreensFunction.h"

using namespace EDLib;

static std::vector<std::array<double, 3>> read_dat(const std::string& f, int n) {
  std::ifstream in(f);
  std::vector<std::array<double, 3>> out;
  for (int i = 0; i < n; ++i) {
    std::array<double, 3> r;
    in >> r[0] >> r[1] >> r[2];
    out.push_back(r);
  }
  return out;
}

static int gfchi(const std::string& dir) {
  params p;
  define_parameters(p);
  p["NSITES"] = 4;
  p["NSPINS"] = 2;
  p["INPUT_FILE"] = dir + "/GF_Chi_input.json";
  p["arpack.SECTOR"] = false;
  p["storage.MAX_DIM"] = 36;
  p["storage.EIGENVALUES_ONLY"] = 0;
  p["arpack.NEV"] = 100;
  p["lanc.BETA"] = 20;
  p["lanc.EMIN"] = -4.0;
  p["lanc.EMAX"] = 4.0;
  p["lanc.NOMEGA"] = 1000;
  HIPHubbardHamiltonian ham(p);
  ham.diag(false);
  gf::MatsubaraMesh fmesh(p, gf::FERMIONIC), bmesh(p, gf::BOSONIC);
  gf::GreensFunction<HIPHubbardHamiltonian> greens(p, ham, fmesh);
  greens.compute();
  StaticObservables<HIPHubbardHamiltonian> so(p);
  auto obs = so.calculate_static_observables(ham);
  gf::ChiLoc<HIPHubbardHamiltonian> susc(p, ham, bmesh);
  double avg = 0.0;
  for (double x : obs["M"]) avg += x / (2.0 * obs["M"].size());
  susc.compute(EDH_OP_SZ, &avg);
  gf::GFData chisz = susc.G();
  avg = 0.0;
  for (double x : obs["N"]) avg += x / double(obs["N"].size());
  susc.compute(EDH_OP_N, &avg);
  gf::GFData chin = susc.G();
  auto g = read_dat(dir + "/gom1.dat", 200);
  auto xs = read_dat(dir + "/xiats.dat", 200);
  auto xd = read_dat(dir + "/xiatd.dat", 200);
  double dg = 0, ds = 0, dn = 0;
  for (int i = 0; i < 200; ++i) {
    for (int s = 0; s < 2; ++s) dg = std::max(dg, std::abs(greens.G()(i, 0, s) - std::complex<double>(g[i][1], g[i][2])));
    ds = std::max(ds, std::abs(chisz(i, 0) - std::complex<double>(-0.25 * xs[i][1], 0.0)));
    dn = std::max(dn, std::abs(chin(i, 0) - std::complex<double>(-xd[i][1], 0.0)));
  }
  bool ok = dg <= 1e-10 && ds <= 1e-9 && dn <= 1e-9;
  std::printf("{\"case\": \"gfchi\", \"max_dG\": %.3e, \"max_dChiSz\": %.3e, \"max_dChiN\": %.3e, \"npairs\": %zu, \"ok\": %s}\n",
              dg, ds, dn, ham.eigenpairs().size(), ok ? "true" : "false");
  return ok ? 0 : 1;
}

static int ring4(const std::string& dir, bool mag) {
  params p;
  define_parameters(p);
  p["NSITES"] = 4;
  p["NSPINS"] = 2;
  p["INPUT_FILE"] = dir + (mag ? "/4ring_input.json" : "/4ring_nomagfield_input.json");
  p["arpack.SECTOR"] = false;
  p["storage.MAX_DIM"] = 36;
  p["storage.EIGENVALUES_ONLY"] = 0;
  p["arpack.NEV"] = 1;
  HIPHubbardHamiltonian ham(p);
  ham.diag(false);
  const auto& gs = *ham.eigenpairs().begin();
  StaticObservables<HIPHubbardHamiltonian> so(p);
  auto obs = so.calculate_static_observables(ham);
  bool ok = true;
  if (mag) {
    ok = std::abs(gs.eigenvalue() - (-11.8443)) < 1e-4 && gs.sector().nup() == 2 && gs.sector().ndown() == 2;
    for (int o = 0; o < 4; ++o) ok = ok && std::abs(obs["N"][o] - 1.0) < 1e-8 && obs["N_up"][o] > obs["N_dn"][o] && obs["M"][o] > 0.0;
  } else {
    ok = obs["N"].size() == 4 && obs["M_i M_j"].size() == 16 && obs["N_eff"].size() == 1;
    for (int o = 0; o < 4; ++o)
      ok = ok && std::abs(obs["N"][o] - 1.0) < 1e-8 && std::abs(obs["N_up"][o] - 0.5) < 1e-8 &&
           std::abs(obs["N_dn"][o] - 0.5) < 1e-8 && std::abs(obs["M"][o]) < 1e-7 &&
           std::abs(obs["D_occ"][o] - obs["D_occ"][0]) < 1e-8;
    ok = ok && obs["N_eff"][0] > 0.0;
  }
  std::printf("{\"case\": \"%s\", \"E0\": %.12f, \"nup\": %d, \"ndn\": %d, \"N0\": %.12f, \"M0\": %.3e, \"ok\": %s}\n",
              mag ? "4ring" : "4ring_nomag", gs.eigenvalue(), gs.sector().nup(), gs.sector().ndown(), obs["N"][0],
              obs["M"][0], ok ? "true" : "false");
  return ok ? 0 : 1;
}

static int storage_corners(const std::string& dir) {
  params p;
  define_parameters(p);
  p["NSITES"] = 4;
  p["NSPINS"] = 2;
  p["INPUT_FILE"] = dir + "/4ring_input.json";
  p["arpack.SECTOR"] = false;
  p["storage.MAX_DIM"] = 36;
  p["storage.EIGENVALUES_ONLY"] = 0;
  p["arpack.NEV"] = 1;
  HIPHubbardHamiltonian ham(p);
  ham.diag(false);
  auto& m = ham.model();
  auto& st = ham.storage();
  bool ok = true;
  int ones = 0;
  double worst = 0.0;
  std::string pairs_json;
  // one-state sectors: (0,0), (0,4), (4,0), (4,4); the state's bits are the
  // up half over the down half, all set or all clear
  for (const auto& pr : ham.eigenpairs()) {
    const int a = pr.sector().nup(), b = pr.sector().ndown();
    if (pr.sector().size() != 1) continue;
    ++ones;
    const long long s = ((long long)(a ? 0xF : 0) << 4) | (b ? 0xF : 0);
    worst = std::max(worst, std::abs(pr.eigenvalue() - m.diagonal(s)));
    char buf[96];
    std::snprintf(buf, sizeof(buf), "%s[%d, %d, %.15g]", pairs_json.empty() ? "" : ", ", a, b, pr.eigenvalue());
    pairs_json += buf;
    ok = ok && pr.eigenvector().size() == 1 && pr.eigenvector()[0] == 1.0;
  }
  ok = ok && ones == 4 && worst < 1e-12;
  // constant_shift on sector (2,2): av(v) + shift * v, discarded by the next fill
  m.symmetry().set_sector(Symmetry::SzSymmetry::Sector(2, 2, 36));
  st.fill();
  std::vector<double> v(36), w0(36), w1(36), w2(36);
  for (int i = 0; i < 36; ++i) v[i] = std::sin(1.0 + i);
  st.av(v.data(), w0.data(), 36, true);
  st.constant_shift(0.75);
  st.av(v.data(), w1.data(), 36, true);
  st.fill();
  st.av(v.data(), w2.data(), 36, true);
  double dshift = 0.0, dfill = 0.0;
  for (int i = 0; i < 36; ++i) {
    dshift = std::max(dshift, std::abs(w1[i] - w0[i] - 0.75 * v[i]));
    dfill = std::max(dfill, std::abs(w2[i] - w0[i]));
  }
  ok = ok && dshift < 1e-13 && dfill == 0.0;
  // the shift reaches a one-state sector's pair too
  m.symmetry().set_sector(Symmetry::SzSymmetry::Sector(4, 4, 1));
  st.fill();
  st.constant_shift(0.5);
  st.zero_eigenapair();
  const double zshift = std::abs(st.eigenvalues()[0] - (m.diagonal(0xFF) + 0.5));
  ok = ok && zshift < 1e-12;
  std::printf("{\"case\": \"storage\", \"one_state_pairs\": [%s], \"max_dE\": %.3e, \"shift_err\": %.3e, "
              "\"refill_err\": %.3e, \"zero_pair_shift_err\": %.3e, \"ok\": %s}\n",
              pairs_json.c_str(), worst, dshift, dfill, zshift, ok ? "true" : "false");
  return ok ? 0 : 1;
}

// A JSON input file (the layout of tests/golden/*_input.json) under /tmp.
static std::string write_input(const std::string& tag,
                               const std::vector<std::pair<std::string, std::pair<std::vector<size_t>, std::vector<double>>>>& sets) {
  const std::string path = "/tmp/edtest_" + tag + "_" + std::to_string((long long)getpid()) + ".json";
  std::ofstream f(path);
  f.precision(17);
  f << "{";
  for (size_t q = 0; q < sets.size(); ++q) {
    f << (q ? ",\n" : "\n") << " \"" << sets[q].first << "\": {\"shape\": [";
    for (size_t i = 0; i < sets[q].second.first.size(); ++i) f << (i ? ", " : "") << sets[q].second.first[i];
    f << "], \"data\": [";
    for (size_t i = 0; i < sets[q].second.second.size(); ++i) f << (i ? ", " : "") << sets[q].second.second[i];
    f << "]}";
  }
  f << "\n}\n";
  return path;
}

static std::vector<double> all_pairs(int n) {
  std::vector<double> v;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      v.push_back(i);
      v.push_back(j);
    }
  return v;
}

// Dyson equation (GreensFunction.h:202-223, FermionicModel.h:84-108):
//   hubbard0 -- 4-site ring at U = 0, all orbital pairs: Sigma = 0;
//   hubbard  -- 4-site ring, U = 5, mu = U/2: Sigma_ii(i w_max) -> U <n> = 2.5;
//   siam0    -- Anderson impurity at U = 0: Sigma = -mu, because the
//               reference's bare function 1/(z - H0 - Delta) has no mu
//               (SingleImpurityAndersonModel.h:378-391).
static int dyson(const std::string& which) {
  params p;
  define_parameters(p);
  p["NSPINS"] = 2;
  p["arpack.SECTOR"] = false;
  p["storage.EIGENVALUES_ONLY"] = 0;
  p["arpack.NEV"] = 100;
  p["lanc.BETA"] = 10.0;
  p["lanc.NOMEGA"] = 1000;
  double maxdev = 0.0, tail = 0.0;
  bool ok = true;
  if (which == "siam0") {
    const double mu = 1.3;
    p["NSITES"] = 4;
    p["siam.NORBITALS"] = 1;
    p["INPUT_FILE"] = write_input(which, {{"Bath/Epsk/values", {{3, 2}, {-1.1, -1.1, 0.2, 0.2, 0.9, 0.9}}},
                                          {"Bath/Vk_0/values", {{3, 2}, {0.5, 0.5, 0.3, 0.3, 0.4, 0.4}}},
                                          {"H0_0/values", {{1, 2}, {0.1, 0.1}}},
                                          {"interaction/values", {{2, 2, 1, 1, 1, 1}, {0, 0, 0, 0}}},
                                          {"mu", {{}, {mu}}},
                                          {"GreensFunction_orbitals/values", {{1, 2}, {0, 0}}}});
    HIPSIAMHamiltonian ham(p);
    ham.diag(false);
    gf::MatsubaraMesh mesh(p, gf::FERMIONIC);
    gf::GreensFunction<HIPSIAMHamiltonian> greens(p, ham, mesh);
    greens.compute();
    gf::GFData bare, sigma;
    greens.selfenergy(bare, sigma);
    for (int w = 0; w < sigma.nw; ++w)
      for (int s = 0; s < 2; ++s) maxdev = std::max(maxdev, std::abs(sigma(w, 0, s) - std::complex<double>(-mu, 0.0)));
    ok = maxdev < 1e-7;
  } else {
    const bool interacting = which == "hubbard";
    const double U = interacting ? 5.0 : 0.0, mu = interacting ? 2.5 : 0.7;
    std::vector<double> t(16, 0.0);
    for (int i = 0; i < 4; ++i) t[i * 4 + (i + 1) % 4] = t[((i + 1) % 4) * 4 + i] = -1.0;
    p["NSITES"] = 4;
    p["INPUT_FILE"] = write_input(which, {{"hopping/values", {{4, 4}, t}},
                                          {"interaction/values", {{4}, std::vector<double>(4, U)}},
                                          {"chemical_potential/values", {{4}, std::vector<double>(4, mu)}},
                                          {"GreensFunction_orbitals/values", {{16, 2}, all_pairs(4)}}});
    HIPHubbardHamiltonian ham(p);
    ham.diag(false);
    gf::MatsubaraMesh mesh(p, gf::FERMIONIC);
    gf::GreensFunction<HIPHubbardHamiltonian> greens(p, ham, mesh);
    greens.compute();
    gf::GFData bare, sigma;
    greens.selfenergy(bare, sigma);
    for (int w = 0; w < sigma.nw; ++w)
      for (int k = 0; k < 16; ++k)
        for (int s = 0; s < 2; ++s) maxdev = std::max(maxdev, std::abs(sigma(w, k, s)));
    if (interacting) {
      tail = sigma(sigma.nw - 1, 0, 0).real();
      ok = std::abs(tail - U / 2) < 0.02;
      for (int i = 0; i < 4; ++i)
        for (int s = 0; s < 2; ++s)
          ok = ok && std::abs(sigma(sigma.nw - 1, i * 4 + i, s).real() - tail) < 1e-6;
    } else {
      ok = maxdev < 1e-7;
    }
  }
  std::remove(p["INPUT_FILE"].as<std::string>().c_str());
  std::printf("{\"case\": \"dyson_%s\", \"max_abs_sigma_or_dev\": %.3e, \"sigma_tail\": %.6f, \"ok\": %s}\n",
              which.c_str(), maxdev, tail, ok ? "true" : "false");
  return ok ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: edtest <gfchi|4ring|4ring_nomag> <golden dir>\n");
    return 2;
  }
  std::string c = argv[1], dir = argv[2];
  try {
    if (c == "gfchi") return gfchi(dir);
    if (c == "4ring") return ring4(dir, true);
    if (c == "4ring_nomag") return ring4(dir, false);
    if (c == "storage") return storage_corners(dir);
    if (c.rfind("dyson_", 0) == 0) return dyson(c.substr(6));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 3;
  }
  return 2;
}
