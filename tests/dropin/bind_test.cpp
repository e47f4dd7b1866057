// The drop-in storage bound to models that expose only the reference
// models' public interface (ref_models.h).
//
//   bindtest tables hubbard <input.json> <Ns>
//   bindtest tables siam <input.json> <Ns> <ml>
//       no GPU: the engine model the storage builds from the interface
//       (engine_model_from_interface) against the engine's own builder
//       (edh_model_hubbard / edh_model_siam) on the same arrays: every
//       sector's per-spin CSR bit for bit, every state's diagonal to 1e-12,
//       and the off-diagonal action on every Fock state to 1e-14.
//   bindtest sectors <Ns>
//       no GPU: the drop-in SzSymmetry's sector order.
//   bindtest hamiltonian hubbard <input.json> <Ns> <nev>
//   bindtest hamiltonian siam <input.json> <Ns> <ml> <nev>
//       GPU: Hamiltonian<HIPSpinResolvedStorage<Ref model>> diagonalised
//       over all sectors (Hamiltonian.h:54-101); prints the ordered
//       eigenvalues with their sectors and the Green's function G_00 on the
//       first Matsubara points (GreensFunction.h:109-159), as JSON.
// Prints one JSON line; exit code 0 when the run finished (the Python test
// judges the numbers).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "edlib_amd/GreensFunction.h"
#include "ref_models.h"

using namespace EDLib;

static void setup(params& p, int Ns, const std::string& input, int ml, int nev) {
  define_parameters(p);
  p["NSITES"] = Ns;
  p["NSPINS"] = 2;
  p["INPUT_FILE"] = input;
  p["arpack.SECTOR"] = false;
  p["storage.EIGENVALUES_ONLY"] = 0;
  p["arpack.NEV"] = nev;
  p["lanc.BETA"] = 20;
  p["lanc.NOMEGA"] = 64;
  p["lanc.NLANC"] = 100;
  if (ml > 0) p["siam.NORBITALS"] = ml;
}

template <class RefModel>
static int compare_tables(RefModel& ref, const edh_model* direct, int Ns) {
  Storage::EngineModelPtr em = Storage::engine_model_from_interface(ref, Ns);
  int64_t csr_mismatch = 0, sectors = 0;
  double dmax = 0.0, omax = 0.0;
  for (int nu = 0; nu <= Ns; ++nu)
    for (int nd = 0; nd <= Ns; ++nd) {
      ++sectors;
      for (int which = 0; which < 2; ++which) {
        int64_t nr[2], nnz[2];
        Storage::edh_check(edh_sector_csr(em.get(), nu, nd, which, &nr[0], &nnz[0], nullptr, nullptr, nullptr));
        Storage::edh_check(edh_sector_csr(direct, nu, nd, which, &nr[1], &nnz[1], nullptr, nullptr, nullptr));
        if (nr[0] != nr[1] || nnz[0] != nnz[1]) {
          ++csr_mismatch;
          continue;
        }
        std::vector<int> rp[2], col[2];
        std::vector<double> val[2];
        const edh_model* ms[2] = {em.get(), direct};
        for (int q = 0; q < 2; ++q) {
          rp[q].resize(nr[q] + 1);
          col[q].resize(std::max<int64_t>(nnz[q], 1));
          val[q].resize(std::max<int64_t>(nnz[q], 1));
          Storage::edh_check(
              edh_sector_csr(ms[q], nu, nd, which, &nr[q], &nnz[q], rp[q].data(), col[q].data(), val[q].data()));
        }
        if (rp[0] != rp[1] || col[0] != col[1]) ++csr_mismatch;
        for (int64_t e = 0; e < nnz[0]; ++e)
          if (std::memcmp(&val[0][e], &val[1][e], sizeof(double)) != 0) {
            ++csr_mismatch;
            break;
          }
      }
      const int64_t D = (int64_t)std::llround(std::tgamma(Ns + 1.0) / std::tgamma(nu + 1.0) / std::tgamma(Ns - nu + 1.0)) *
                        (int64_t)std::llround(std::tgamma(Ns + 1.0) / std::tgamma(nd + 1.0) / std::tgamma(Ns - nd + 1.0));
      std::vector<double> d0(D), d1(D);
      Storage::edh_check(edh_sector_diagonal(em.get(), nu, nd, d0.data()));
      Storage::edh_check(edh_sector_diagonal(direct, nu, nd, d1.data()));
      for (int64_t i = 0; i < D; ++i) dmax = std::max(dmax, std::abs(d0[i] - d1[i]));
    }
  // off-diagonal action (one- and two-body terms) on every Fock state
  std::vector<long long> t(4096);
  std::vector<double> c(4096);
  for (long long st = 0; st < (1ll << (2 * Ns)); ++st) {
    std::map<long long, double> acc[2];
    const edh_model* ms[2] = {em.get(), direct};
    for (int q = 0; q < 2; ++q) {
      int cnt = 0;
      Storage::edh_check(edh_model_apply(ms[q], st, (int)t.size(), t.data(), c.data(), &cnt));
      for (int r = 0; r < cnt; ++r) acc[q][t[r]] += c[r];
    }
    for (int q = 0; q < 2; ++q)
      for (const auto& kv : acc[q]) {
        const double other = acc[1 - q].count(kv.first) ? acc[1 - q][kv.first] : 0.0;
        omax = std::max(omax, std::abs(kv.second - other));
      }
  }
  const bool ok = csr_mismatch == 0 && dmax < 1e-12 && omax < 1e-14;
  std::printf("{\"ok\": %s, \"sectors\": %lld, \"csr_mismatch\": %lld, \"diag_max\": %.3e, \"offdiag_max\": %.3e}\n",
              ok ? "true" : "false", (long long)sectors, (long long)csr_mismatch, dmax, omax);
  return ok ? 0 : 1;
}

template <class Model>
static int run_hamiltonian(params& p) {
  typedef Hamiltonian<Storage::HIPSpinResolvedStorage<Model>> Ham;
  Ham ham(p);
  ham.diag(false);
  std::ostringstream ev;
  ev.precision(17);
  int q = 0;
  for (const auto& e : ham.eigenpairs()) {
    if (q) ev << ", ";
    ev << "[" << e.eigenvalue() << ", " << e.sector().nup() << ", " << e.sector().ndown() << "]";
    if (++q == 12) break;
  }
  gf::MatsubaraMesh fmesh(p, gf::FERMIONIC);
  gf::GreensFunction<Ham> greens(p, ham, fmesh);
  greens.compute();
  const gf::GFData& G = greens.G();
  std::ostringstream g;
  g.precision(17);
  for (int w = 0; w < 16; ++w) g << (w ? ", " : "") << "[" << G(w, 0, 0).real() << ", " << G(w, 0, 0).imag() << "]";
  std::printf("{\"ok\": true, \"eigenpairs\": [%s], \"G00\": [%s]}\n", ev.str().c_str(), g.str().c_str());
  return 0;
}

int main(int argc, char** argv) {
  try {
    if (argc == 3 && std::string(argv[1]) == "sectors") {
      // the drop-in SzSymmetry's sector order (std::sort with the
      // reference comparator, SzSymmetry.h:51-54,72-98)
      params p;
      define_parameters(p);
      p["NSITES"] = std::atoi(argv[2]);
      p["arpack.SECTOR"] = false;
      Symmetry::SzSymmetry sym(p);
      std::ostringstream o;
      int q = 0;
      while (sym.next_sector()) o << (q++ ? ", " : "") << "[" << sym.sector().nup() << ", " << sym.sector().ndown() << "]";
      std::printf("{\"ok\": true, \"sectors\": [%s]}\n", o.str().c_str());
      return 0;
    }
    if (argc < 5) throw std::invalid_argument("usage: bindtest tables|hamiltonian hubbard|siam <input> <Ns> [ml] [nev]");
    const std::string what = argv[1], kind = argv[2], input = argv[3];
    const int Ns = std::atoi(argv[4]);
    const int ml = kind == "siam" ? std::atoi(argv[5]) : 0;
    const int nev = argc > (kind == "siam" ? 6 : 5) ? std::atoi(argv[kind == "siam" ? 6 : 5]) : 4;
    params p;
    setup(p, Ns, input, ml, nev);
    if (what == "tables") {
      alps_shim::input_file in(input);
      edh_model* m = nullptr;
      if (kind == "hubbard") {
        RefIface::HubbardModel<double> ref(p);
        const int n = Ns;
        std::vector<double> H(n, 0.0), J(n * n, 0.0);
        if (in.is_data("magnetic_field/values")) H = in.get("magnetic_field/values").data;
        if (in.is_data("exchange/values")) J = in.get("exchange/values").data;
        Storage::edh_check(edh_model_hubbard(&m, Ns, in.get("hopping/values").data.data(),
                                             in.get("interaction/values").data.data(), J.data(),
                                             in.get("chemical_potential/values").data.data(), H.data()));
        Storage::EngineModelPtr direct(m);
        return compare_tables(ref, direct.get(), Ns);
      }
      RefIface::SingleImpurityAndersonModel<double> ref(p);
      const int nb = Ns - ml;
      std::vector<double> Vk(ml * nb * 2, 0.0), H0(ml * ml * 2, 0.0);
      for (int im = 0; im < ml; ++im) {
        std::vector<double> v = in.get("Bath/Vk_" + std::to_string(im) + "/values").data;
        std::vector<double> h = in.get("H0_" + std::to_string(im) + "/values").data;
        std::copy(v.begin(), v.end(), Vk.begin() + im * nb * 2);
        std::copy(h.begin(), h.end(), H0.begin() + im * ml * 2);
      }
      Storage::edh_check(edh_model_siam(&m, Ns, ml, in.get("Bath/Epsk/values").data.data(), Vk.data(), H0.data(),
                                        in.get("interaction/values").data.data(), in.get("mu").data.at(0)));
      Storage::EngineModelPtr direct(m);
      return compare_tables(ref, direct.get(), Ns);
    }
    if (kind == "hubbard") return run_hamiltonian<RefIface::HubbardModel<double>>(p);
    return run_hamiltonian<RefIface::SingleImpurityAndersonModel<double>>(p);
  } catch (const std::exception& e) {
    std::cerr << "bindtest: " << e.what() << std::endl;
    std::printf("{\"ok\": false, \"error\": \"see stderr\"}\n");
    return 2;
  }
}
