// Results in the reference's HDF5 layout (the paths its drivers write through
// alps::hdf5::archive): results/eigenvalues/{N, data, sectors/<i>/{nup,
// ndown, size}} (HDF5Utils.h:59-76), results/static_observables/<name>
// (HDF5Utils.h:85-89), results/G_omega, results/G_ij_omega and
// results/@Statsum (GreensFunction.h:167-194), results/Chi<Sz|N>_omega
// (ChiLoc.h:418-437).  A Green's function is written in the layout of
// ALPSCore's gf library (the H5GF format its save() writes, which the
// reference calls at GreensFunction.h:167-194): version/{major, minor,
// reference, originator}; data with the complex parts as a trailing
// dimension of 2 and the integer attribute __complex__ = 1
// (data[w][orbital][spin][re/im]; susceptibilities data[w][orbital][re/im]);
// mesh/N, then one group per index: mesh/1 the frequency mesh (kind
// MATSUBARA with N, statistics 1 fermionic / 0 bosonic, beta,
// positive_only = 1 and points; or kind REAL_FREQUENCY with N and points),
// mesh/2 the orbital and mesh/3 the spin index (kind INDEX, N, points).
// ALPSCore itself is not used.
// Needs the HDF5 C library: define EDLIB_AMD_HDF5 and link -lhdf5.
#ifndef EDLIB_AMD_HDF5_H
#define EDLIB_AMD_HDF5_H

#ifndef EDLIB_AMD_HDF5
#define EDLIB_AMD_HDF5
#endif
#include <hdf5.h>

#include <string>
#include <vector>

#include "GreensFunction.h"

namespace EDLib {
namespace alps_shim {

class archive {
 public:
  archive() {}
  archive(const std::string& path, const std::string& mode) { open(path, mode); }
  ~archive() { close(); }
  archive(const archive&) = delete;
  archive& operator=(const archive&) = delete;

  // mode "w": a new file (truncated); "a": an existing file, appended to
  void open(const std::string& path, const std::string& mode) {
    close();
    if (mode == "w")
      f_ = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    else
      f_ = H5Fopen(path.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    if (f_ < 0) throw std::runtime_error("cannot open HDF5 file " + path);
  }
  void close() {
    if (f_ >= 0) H5Fclose(f_);
    f_ = -1;
  }

  // A dataset of doubles at `path` (intermediate groups created; an existing
  // dataset of that name is replaced).  shape {} is a scalar.
  void write(const std::string& path, const std::vector<double>& v, const std::vector<hsize_t>& shape) {
    write_typed(path, v.data(), H5T_NATIVE_DOUBLE, shape);
  }
  void write(const std::string& path, const std::vector<long long>& v, const std::vector<hsize_t>& shape) {
    write_typed(path, v.data(), H5T_NATIVE_LLONG, shape);
  }
  void write(const std::string& path, double x) {
    if (is_attribute(path))
      write_attribute(path, x);
    else
      write(path, std::vector<double>{x}, {});
  }
  void write(const std::string& path, long long x) { write(path, std::vector<long long>{x}, {}); }

  // A variable-length string dataset (as alps::hdf5 writes std::string).
  void write(const std::string& path_in, const std::string& str) {
    const std::string path = strip(path_in);
    parent_group(path);
    if (H5Lexists(f_, path.c_str(), H5P_DEFAULT) > 0) H5Ldelete(f_, path.c_str(), H5P_DEFAULT);
    const hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, H5T_VARIABLE);
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t d = H5Dcreate2(f_, path.c_str(), t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const char* c = str.c_str();
    const herr_t w = d >= 0 ? H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, &c) : -1;
    if (d >= 0) H5Dclose(d);
    H5Sclose(sp);
    H5Tclose(t);
    if (w < 0) throw std::runtime_error("HDF5: cannot write " + path);
  }
  void write(const std::string& path, const char* str) { write(path, std::string(str)); }

  // A Green's function in the H5GF layout (see the header comment).
  template <class Mesh>
  void write_gf(const std::string& path, const gf::GFData& g, const Mesh& mesh, int statistics) {
    std::vector<double> d(g.d.size() * 2);
    for (size_t q = 0; q < g.d.size(); ++q) {
      d[2 * q] = g.d[q].real();
      d[2 * q + 1] = g.d[q].imag();
    }
    write(path + "/version/major", 0ll);
    write(path + "/version/minor", 1ll);
    write(path + "/version/reference", "https://github.com/ALPSCore/H5GF/blob/master/H5GF.rst");
    write(path + "/version/originator", "EDLib on the edhip engine (H5GF layout of the ALPSCore GF library)");
    if (g.nspin > 1)
      write(path + "/data", d, {(hsize_t)g.nw, (hsize_t)g.norb, (hsize_t)g.nspin, 2});
    else
      write(path + "/data", d, {(hsize_t)g.nw, (hsize_t)g.norb, 2});
    dataset_attribute(path + "/data", "__complex__", 1);
    write(path + "/mesh/N", (long long)(g.nspin > 1 ? 3 : 2));
    write(path + "/mesh/1/kind", mesh.real ? "REAL_FREQUENCY" : "MATSUBARA");
    write(path + "/mesh/1/N", (long long)g.nw);
    write(path + "/mesh/1/points", mesh.pts, {(hsize_t)mesh.pts.size()});
    if (!mesh.real) {
      write(path + "/mesh/1/beta", mesh.beta);
      write(path + "/mesh/1/statistics", (long long)statistics);
      write(path + "/mesh/1/positive_only", 1ll);
    }
    index_mesh(path + "/mesh/2", g.norb);
    if (g.nspin > 1) index_mesh(path + "/mesh/3", g.nspin);
  }

 private:
  static bool is_attribute(const std::string& path) {
    const size_t s = path.find_last_of('/');
    return path.compare(s == std::string::npos ? 0 : s + 1, 1, "@") == 0;
  }
  static std::string strip(std::string p) {
    while (!p.empty() && p.back() == '/') p.pop_back();
    while (!p.empty() && p.front() == '/') p.erase(0, 1);
    return p;
  }
  void index_mesh(const std::string& path, int n) {
    write(path + "/kind", "INDEX");
    write(path + "/N", (long long)n);
    std::vector<long long> pts(n);
    for (int i = 0; i < n; ++i) pts[i] = i;
    write(path + "/points", pts, {(hsize_t)n});
  }
  void parent_group(const std::string& path) {
    const size_t s = path.find_last_of('/');
    if (s != std::string::npos) {
      const hid_t g = group(path.substr(0, s));
      if (g < 0) throw std::runtime_error("HDF5: cannot create group for " + path);
      H5Gclose(g);
    }
  }
  // an integer attribute on an existing dataset
  void dataset_attribute(const std::string& path_in, const char* name, int x) {
    const std::string path = strip(path_in);
    const hid_t d = H5Dopen2(f_, path.c_str(), H5P_DEFAULT);
    if (d < 0) throw std::runtime_error("HDF5: no dataset " + path);
    if (H5Aexists(d, name) > 0) H5Adelete(d, name);
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t a = H5Acreate2(d, name, H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
    const herr_t w = a >= 0 ? H5Awrite(a, H5T_NATIVE_INT, &x) : -1;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    H5Dclose(d);
    if (w < 0) throw std::runtime_error("HDF5: cannot write attribute " + path + "/@" + name);
  }
  // opens (creating as needed) the group at `path`
  // whether every component of `path` exists (H5Lexists needs the parents)
  bool exists(const std::string& path) {
    size_t pos = 0;
    for (;;) {
      const size_t s = path.find('/', pos);
      const std::string part = path.substr(0, s);
      if (H5Lexists(f_, part.c_str(), H5P_DEFAULT) <= 0) return false;
      if (s == std::string::npos) return true;
      pos = s + 1;
    }
  }
  hid_t group(const std::string& path) {
    if (path.empty()) return H5Gopen2(f_, "/", H5P_DEFAULT);
    if (exists(path)) return H5Gopen2(f_, path.c_str(), H5P_DEFAULT);
    const hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
    H5Pset_create_intermediate_group(lcpl, 1);
    const hid_t g = H5Gcreate2(f_, path.c_str(), lcpl, H5P_DEFAULT, H5P_DEFAULT);
    H5Pclose(lcpl);
    return g;
  }
  void write_typed(const std::string& path_in, const void* data, hid_t type, const std::vector<hsize_t>& shape) {
    const std::string path = strip(path_in);
    const size_t s = path.find_last_of('/');
    if (s != std::string::npos) {
      const hid_t g = group(path.substr(0, s));
      if (g < 0) throw std::runtime_error("HDF5: cannot create group for " + path);
      H5Gclose(g);
    }
    if (H5Lexists(f_, path.c_str(), H5P_DEFAULT) > 0) H5Ldelete(f_, path.c_str(), H5P_DEFAULT);
    const hid_t sp = shape.empty() ? H5Screate(H5S_SCALAR) : H5Screate_simple((int)shape.size(), shape.data(), nullptr);
    const hid_t d = H5Dcreate2(f_, path.c_str(), type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const herr_t w = d >= 0 ? H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) : -1;
    if (d >= 0) H5Dclose(d);
    H5Sclose(sp);
    if (w < 0) throw std::runtime_error("HDF5: cannot write " + path);
  }
  void write_attribute(const std::string& path_in, double x) {
    const std::string path = strip(path_in);
    const size_t s = path.find_last_of('/');
    const std::string gp = s == std::string::npos ? "" : path.substr(0, s);
    const std::string name = path.substr(s == std::string::npos ? 1 : s + 2);
    const hid_t g = group(gp);
    if (g < 0) throw std::runtime_error("HDF5: cannot create group " + gp);
    if (H5Aexists(g, name.c_str()) > 0) H5Adelete(g, name.c_str());
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t a = H5Acreate2(g, name.c_str(), H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    const herr_t w = a >= 0 ? H5Awrite(a, H5T_NATIVE_DOUBLE, &x) : -1;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    H5Gclose(g);
    if (w < 0) throw std::runtime_error("HDF5: cannot write attribute " + path);
  }
  hid_t f_ = -1;
};

}  // namespace alps_shim

using alps_shim::archive;

namespace hdf5 {

// HDF5Utils.h:59-76: the eigenvalues in their stored order, and each
// pair's sector.
template <class Ham>
void save_eigen_pairs(const Ham& h, archive& ar, const std::string& path) {
  ar.write(path + "/eigenvalues/N", (long long)h.eigenpairs().size());
  std::vector<double> values;
  int i = 0;
  for (const auto& e : h.eigenpairs()) {
    values.push_back(e.eigenvalue());
    const std::string sp = path + "/eigenvalues/sectors/" + std::to_string(i);
    ar.write(sp + "/nup", (long long)e.sector().nup());
    ar.write(sp + "/ndown", (long long)e.sector().ndown());
    ar.write(sp + "/size", (long long)e.sector().size());
    ++i;
  }
  ar.write(path + "/eigenvalues/data", values, {(hsize_t)values.size()});
}

// HDF5Utils.h:85-89
inline void save_static_observables(const std::map<std::string, std::vector<double>>& obs, archive& ar,
                                    const std::string& root) {
  for (const auto& kv : obs) ar.write(root + "/static_observables/" + kv.first, kv.second, {(hsize_t)kv.second.size()});
}

}  // namespace hdf5
}  // namespace EDLib

#endif
