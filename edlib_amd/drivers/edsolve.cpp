// Drop-in drivers with the reference's examples' steps and HDF5 files:
//   edsolve hubbard  [file.param] [--KEY=value ...]  -- examples/Hubbard.cpp:28-67
//     (diag, eigenvalues, G(i w_n), chi_Sz and chi_N, all under "results")
//   edsolve anderson [file.param] [--KEY=value ...]  -- examples/Anderson.cpp:28-55
//     (diag, eigenvalues, G(i w_n))
// Both also store the static observables (HDF5Utils.h:85-89).  INPUT_FILE
// is the reference's input.h5; OUTPUT_FILE the HDF5 result file.
//   edsolve --dump file.h5  prints every numeric dataset of an HDF5 file and
//     the Statsum attribute of its results group as one JSON object
//     {"path": {"shape": [...], "data": [...]}, ...} (for tests).
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "edlib_amd/HDF5.h"

using namespace EDLib;

static int dump(const std::string& path) {
  auto sets = alps_shim::read_h5(path);
  const hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f >= 0) {
    if (H5Lexists(f, "results", H5P_DEFAULT) > 0 && H5Aexists_by_name(f, "results", "Statsum", H5P_DEFAULT) > 0) {
      const hid_t a = H5Aopen_by_name(f, "results", "Statsum", H5P_DEFAULT, H5P_DEFAULT);
      double z = 0.0;
      H5Aread(a, H5T_NATIVE_DOUBLE, &z);
      H5Aclose(a);
      sets["results/@Statsum"] = alps_shim::dataset{{}, {z}};
    }
    H5Fclose(f);
  }
  std::printf("{");
  bool first = true;
  for (const auto& kv : sets) {
    std::printf("%s\n \"%s\": {\"shape\": [", first ? "" : ",", kv.first.c_str());
    for (size_t i = 0; i < kv.second.shape.size(); ++i) std::printf("%s%zu", i ? ", " : "", kv.second.shape[i]);
    std::printf("], \"data\": [");
    for (size_t i = 0; i < kv.second.data.size(); ++i) std::printf("%s%.17g", i ? ", " : "", kv.second.data[i]);
    std::printf("]}");
    first = false;
  }
  std::printf("\n}\n");
  return 0;
}

template <class Ham>
static void run(params& p, bool susceptibilities) {
  archive ar(p["OUTPUT_FILE"].as<std::string>(), "w");
  Ham ham(p);
  ham.diag();
  hdf5::save_eigen_pairs(ham, ar, "results");
  gf::MatsubaraMesh fmesh(p, gf::FERMIONIC);
  gf::GreensFunction<Ham> greens(p, ham, fmesh);
  greens.compute();
  greens.save(ar, "results");
  if (susceptibilities) {
    gf::MatsubaraMesh bmesh(p, gf::BOSONIC);
    gf::ChiLoc<Ham> susc(p, ham, bmesh);
    susc.compute(EDH_OP_SZ);
    susc.save(ar, "results");
    susc.compute(EDH_OP_N);
    susc.save(ar, "results");
  }
  StaticObservables<Ham> so(p);
  hdf5::save_static_observables(so.calculate_static_observables(ham), ar, "results");
}

int main(int argc, char** argv) {
  if (argc >= 3 && std::string(argv[1]) == "--dump") {
    try {
      return dump(argv[2]);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "error: %s\n", e.what());
      return 3;
    }
  }
  if (argc < 2) {
    std::fprintf(stderr, "usage: edsolve <hubbard|anderson> [file.param] [--KEY=value ...] | --dump file.h5\n");
    return 2;
  }
  const std::string model = argv[1];
  try {
    params p(argc - 1, argv + 1);
    define_parameters(p);
    if (model == "hubbard")
      run<HIPHubbardHamiltonian>(p, true);
    else if (model == "anderson")
      run<HIPSIAMHamiltonian>(p, false);
    else
      throw std::invalid_argument("unknown model " + model);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 3;
  }
  return 0;
}
