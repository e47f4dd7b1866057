// Parameter object with the key/default semantics the reference takes from
// alps::params (EDParams.h:11-37), plus a reader for the model input files.
//
// Input files hold the datasets of the reference's input.h5 schema
// (HubbardModel.h:55-66, SingleImpurityAndersonModel.h:199-222,
// SzSymmetry.h:77-86): either the HDF5 file itself (a name ending in .h5;
// needs EDLIB_AMD_HDF5 and the HDF5 C library) or the same datasets as JSON:
//   {"hopping/values": {"shape": [Ns, Ns], "data": [...]}, "mu": {"shape": [], "data": [x]}, ...}
#ifndef EDLIB_AMD_PARAMS_H
#define EDLIB_AMD_PARAMS_H

#ifdef EDLIB_AMD_HDF5
#include <hdf5.h>
#endif

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace EDLib {
namespace alps_shim {

class params;

class param_ref {
 public:
  param_ref(params& p, const std::string& k) : p_(p), k_(k) {}
  template <class T>
  T as() const;
  template <class T>
  operator T() const {
    return as<T>();
  }
  operator std::string() const;
  template <class T>
  param_ref& operator=(const T& v);
  param_ref& operator=(const char* v);
  template <class T>
  bool operator==(const T& v) const {
    return as<T>() == v;
  }

 private:
  params& p_;
  std::string k_;
};

class params {
 public:
  params() {}
  // Reads "key=value" lines and [section] headers (section.key) from a .param file,
  // then "--key=value" command-line overrides.
  params(int argc, const char* const* argv) {
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      if (a.rfind("--", 0) == 0) {
        size_t eq = a.find('=');
        if (eq != std::string::npos) vals_[a.substr(2, eq - 2)] = a.substr(eq + 1);
      } else {
        load_ini(a);
      }
    }
  }
  void load_ini(const std::string& file) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("cannot open parameter file " + file);
    std::string line, section;
    while (std::getline(in, line)) {
      line = trim(line);
      if (line.empty() || line[0] == '#') continue;
      if (line[0] == '[') {
        section = trim(line.substr(1, line.find(']') - 1));
        continue;
      }
      size_t eq = line.find('=');
      if (eq == std::string::npos) continue;
      std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
      vals_[section.empty() ? k : section + "." + k] = v;
    }
  }
  template <class T>
  params& define(const std::string& k, const T& def, const std::string& = "") {
    if (!vals_.count(k)) {
      std::ostringstream s;
      s.precision(17);
      s << def;
      vals_[k] = s.str();
    }
    return *this;
  }
  template <class T>
  params& define(const std::string&, const std::string& = "") {
    return *this;  // declared without a default: exists() stays false until set
  }
  bool exists(const std::string& k) const { return vals_.count(k) != 0; }
  param_ref operator[](const std::string& k) { return param_ref(*this, k); }
  const std::string& raw(const std::string& k) const {
    auto it = vals_.find(k);
    if (it == vals_.end()) throw std::runtime_error("parameter " + k + " is not defined");
    return it->second;
  }
  void set(const std::string& k, const std::string& v) { vals_[k] = v; }

 private:
  static std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
  }
  std::map<std::string, std::string> vals_;
};

namespace detail {
template <class T>
inline T parse(const std::string& s) {
  std::istringstream in(s);
  T v{};
  in >> v;
  if (in.fail()) throw std::runtime_error("cannot convert parameter value '" + s + "'");
  return v;
}
template <>
inline bool parse<bool>(const std::string& s) {
  std::string l;
  for (char c : s) l += (char)std::tolower(c);
  if (l == "true" || l == "1" || l == "yes") return true;
  if (l == "false" || l == "0" || l == "no" || l.empty()) return false;
  return std::atof(s.c_str()) != 0.0;
}
template <>
inline std::string parse<std::string>(const std::string& s) {
  return s;
}
}  // namespace detail

template <class T>
T param_ref::as() const {
  return detail::parse<T>(p_.raw(k_));
}
inline param_ref::operator std::string() const { return p_.raw(k_); }
template <class T>
param_ref& param_ref::operator=(const T& v) {
  std::ostringstream s;
  s.precision(17);
  s << std::boolalpha << v;
  p_.set(k_, s.str());
  return *this;
}
inline param_ref& param_ref::operator=(const char* v) {
  p_.set(k_, v);
  return *this;
}

// ---------------------------------------------------------------------------
// Datasets of an input file (see the header comment).
// ---------------------------------------------------------------------------
struct dataset {
  std::vector<size_t> shape;
  std::vector<double> data;
};

#ifdef EDLIB_AMD_HDF5
// Every integer or floating-point dataset of an HDF5 file, read as doubles
// with its shape, under its path relative to the root ("hopping/values").
inline std::map<std::string, dataset> read_h5(const std::string& path) {
  std::map<std::string, dataset> out;
  const hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) throw std::runtime_error("cannot open HDF5 file " + path);
  struct Ctx {
    std::map<std::string, dataset>* out;
    bool ok;
  } ctx{&out, true};
  auto visit = [](hid_t obj, const char* name, const H5O_info_t* info, void* op) -> herr_t {
    Ctx& c = *static_cast<Ctx*>(op);
    if (info->type != H5O_TYPE_DATASET) return 0;
    const hid_t d = H5Dopen2(obj, name, H5P_DEFAULT);
    if (d < 0) return 0;
    const hid_t t = H5Dget_type(d);
    const H5T_class_t cls = H5Tget_class(t);
    if (cls == H5T_INTEGER || cls == H5T_FLOAT) {
      const hid_t sp = H5Dget_space(d);
      const int nd = H5Sget_simple_extent_ndims(sp);
      std::vector<hsize_t> dims(nd > 0 ? nd : 0);
      if (nd > 0) H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
      dataset ds;
      size_t n = 1;
      for (hsize_t x : dims) {
        ds.shape.push_back((size_t)x);
        n *= (size_t)x;
      }
      ds.data.assign(n, 0.0);
      if (n && H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, ds.data.data()) < 0) c.ok = false;
      (*c.out)[name] = ds;
      H5Sclose(sp);
    }
    H5Tclose(t);
    H5Dclose(d);
    return 0;
  };
  const herr_t r = H5Ovisit(f, H5_INDEX_NAME, H5_ITER_NATIVE, visit, &ctx);
  H5Fclose(f);
  if (r < 0 || !ctx.ok) throw std::runtime_error("cannot read the datasets of " + path);
  return out;
}
#endif

class input_file {
 public:
  explicit input_file(const std::string& path) {
    if (path.size() > 3 && path.compare(path.size() - 3, 3, ".h5") == 0) {
#ifdef EDLIB_AMD_HDF5
      sets_ = read_h5(path);
      return;
#else
      throw std::runtime_error("input file " + path + ": built without HDF5 (define EDLIB_AMD_HDF5, link -lhdf5)");
#endif
    }
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open input file " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    txt_ = ss.str();
    pos_ = 0;
    parse_top();
  }
  bool is_data(const std::string& k) const { return sets_.count(k) != 0; }
  const std::map<std::string, dataset>& datasets() const { return sets_; }
  const dataset& get(const std::string& k) const {
    auto it = sets_.find(k);
    if (it == sets_.end()) throw std::runtime_error("input dataset " + k + " not found");
    return it->second;
  }

 private:
  void ws() {
    while (pos_ < txt_.size() && std::isspace((unsigned char)txt_[pos_])) ++pos_;
  }
  void expect(char c) {
    ws();
    if (pos_ >= txt_.size() || txt_[pos_] != c) throw std::runtime_error(std::string("input file: expected ") + c);
    ++pos_;
  }
  std::string str() {
    expect('"');
    std::string s;
    while (pos_ < txt_.size() && txt_[pos_] != '"') s += txt_[pos_++];
    ++pos_;
    return s;
  }
  double num() {
    ws();
    const char* b = txt_.c_str() + pos_;
    char* e = nullptr;
    double v = std::strtod(b, &e);
    if (e == b) throw std::runtime_error("input file: number expected");
    pos_ += (size_t)(e - b);
    return v;
  }
  std::vector<double> numlist() {
    std::vector<double> v;
    expect('[');
    ws();
    if (txt_[pos_] == ']') {
      ++pos_;
      return v;
    }
    for (;;) {
      v.push_back(num());
      ws();
      if (txt_[pos_] == ',') {
        ++pos_;
        continue;
      }
      expect(']');
      return v;
    }
  }
  void parse_top() {
    expect('{');
    ws();
    if (txt_[pos_] == '}') return;
    for (;;) {
      std::string key = str();
      expect(':');
      expect('{');
      dataset d;
      for (;;) {
        std::string f = str();
        expect(':');
        std::vector<double> v = numlist();
        if (f == "shape")
          for (double x : v) d.shape.push_back((size_t)x);
        else
          d.data = v;
        ws();
        if (txt_[pos_] == ',') {
          ++pos_;
          continue;
        }
        expect('}');
        break;
      }
      sets_[key] = d;
      ws();
      if (txt_[pos_] == ',') {
        ++pos_;
        continue;
      }
      expect('}');
      return;
    }
  }
  std::string txt_;
  size_t pos_;
  std::map<std::string, dataset> sets_;
};

}  // namespace alps_shim

typedef alps_shim::params params;

// Parameter keys and defaults of EDParams.h:11-37.
inline void define_parameters(params& p) {
  p.define<int>("NSITES", 4, "Number of sites");
  p.define<int>("NSPINS", 2, "Number of spins");
  p.define<std::string>("INPUT_FILE", "input.json", "File with initial data");
  p.define<std::string>("OUTPUT_FILE", "sim.json", "File with results");
  p.define<bool>("arpack.SECTOR", false, "Read symmetry sectors from file");
  p.define<size_t>("storage.MAX_DIM", 5000, "Maximum dimension of the Hamiltonian matrix.");
  p.define<size_t>("storage.MAX_SIZE", 70000, "Maximum size of the matrix arrays.");
  p.define<int>("storage.EIGENVALUES_ONLY", 0, "Compute only eigenvalues.");
  p.define<int>("spinstorage.ORBITAL_NUMBER", 1, "Number of orbitals with interaction");
  p.define<int>("arpack.NEV", 2, "Number of eigenvalues to find");
  p.define<int>("arpack.NCV", "Number of convergent values");
  p.define<int>("lanc.NOMEGA", 32, "Number of fermionic frequencies");
  p.define<double>("lanc.EMIN", -3, "Lowest real frequency value");
  p.define<double>("lanc.EMAX", 3, "Largest real frequency value");
  p.define<int>("lanc.NLANC", 100, "Number of Lanczos iterations");
  p.define<double>("lanc.BETA", 10.0, "Inverse temperature");
  p.define<double>("lanc.BOLTZMANN_CUTOFF", 1e-12, "Cutoff for Boltzmann factor");
  p.define<int>("siam.NORBITALS", 1, "Number of orbitals in single impurity Anderson Model.");
  p.define<int>("device.ID", 0, "HIP device used by the storage");
}

}  // namespace EDLib

#endif
