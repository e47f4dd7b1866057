// Upload of one sector's tables to the device.
#include <algorithm>

#include "edh_internal.h"

namespace edh {

namespace {

template <class T>
T* upload(DevSector& ds, const std::vector<T>& h, hipStream_t s) {
  if (h.empty()) return nullptr;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, h.size() * sizeof(T));
  if (e != hipSuccess) throw Error(EDH_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  ds.allocs.push_back(p);
  EDH_HIP(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return (T*)p;
}

void upload_spin(DevSector& ds, const SpinTables& sp, DevSpin& d, hipStream_t s) {
  d.D = sp.D;
  d.bits = upload(ds, sp.bits, s);
  d.ediag = upload(ds, sp.ediag, s);
  d.rowp = upload(ds, sp.rowp, s);
  d.packed = upload(ds, sp.packed, s);
  d.sl_ptr = upload(ds, sp.sl_ptr, s);
  d.sl_len = upload(ds, sp.sl_len, s);
  d.ell = upload(ds, sp.ell, s);
  d.ell16 = upload(ds, sp.ell16, s);
  for (int q = 0; q < 4; ++q) d.vt16[q] = sp.vt16[q];
  d.nvt16 = sp.nvt16;
  d.q4_ptr = upload(ds, sp.q4_ptr, s);
  d.ell_q4 = upload(ds, sp.ell_q4, s);
  d.ell_s4 = upload(ds, sp.ell_s4, s);
  d.s4_scale = sp.s4_scale;
  d.rank = upload(ds, sp.rank, s);
}

// Reverse Cuthill-McKee order of the graph whose adjacency is the CSR
// structure (rows and columns are basis states of one spin): breadth-first
// from a minimum-degree vertex of each component, neighbours by increasing
// degree, then reversed.
std::vector<int32_t> rcm_order(const std::vector<int32_t>& rowp, const std::vector<int32_t>& col, int64_t n) {
  std::vector<int32_t> deg(n), order;
  order.reserve(n);
  for (int64_t i = 0; i < n; ++i) deg[i] = rowp[i + 1] - rowp[i];
  std::vector<char> seen(n, 0);
  std::vector<int32_t> byDeg(n);
  for (int64_t i = 0; i < n; ++i) byDeg[i] = (int32_t)i;
  std::stable_sort(byDeg.begin(), byDeg.end(), [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
  std::vector<int32_t> nb;
  for (int32_t start : byDeg) {
    if (seen[start]) continue;
    size_t head = order.size();
    order.push_back(start);
    seen[start] = 1;
    while (head < order.size()) {
      const int32_t x = order[head++];
      nb.clear();
      for (int q = rowp[x]; q < rowp[x + 1]; ++q)
        if (!seen[col[q]]) {
          seen[col[q]] = 1;
          nb.push_back(col[q]);
        }
      std::stable_sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

}  // namespace

void DevSector::release() {
  for (void* p : allocs) (void)hipFree(p);
  allocs.clear();
  *this = DevSector();
}

void upload_sector(const SectorTables& st, int Ns, DevSector& ds, hipStream_t s) {
  ds.release();
  ds.Ns = Ns;
  ds.nup = st.nup;
  ds.ndn = st.ndn;
  ds.N = st.N;
  ds.imp_ml = st.imp_ml;
  ds.table_bytes = st.table_bytes;
  upload_spin(ds, st.up, ds.up, s);
  upload_spin(ds, st.dn, ds.dn, s);
  ds.values = upload(ds, st.values, s);
  ds.cross_mode = st.cross_mode;
  ds.cross_u = st.cross_u;
  ds.cross_imp = upload(ds, st.cross_imp, s);
  ds.cross_rowtab = upload(ds, st.cross_rowtab, s);
  ds.kan_ml = st.kan_ml;
  if (st.kan_ml > 0) {
    std::vector<int32_t> a(st.kan.size()), b(st.kan.size());
    std::vector<double> c(st.kan.size());
    for (size_t q = 0; q < st.kan.size(); ++q) {
      a[q] = st.kan[q].new_a;
      b[q] = st.kan[q].new_b;
      c[q] = st.kan[q].coef;
    }
    ds.kan_ptr = upload(ds, st.kan_ptr, s);
    ds.kan_a = upload(ds, a, s);
    ds.kan_b = upload(ds, b, s);
    ds.kan_c = upload(ds, c, s);
  }
  ds.counters = upload(ds, std::vector<int32_t>(64, 0), s);
  std::vector<int32_t> order = rcm_order(st.up.rowp, st.up.col, st.up.D);
  ds.row_order = upload(ds, order, s);
  EDH_HIP(hipStreamSynchronize(s));
  ds.variant = hv_autotune(ds, s);
}

}  // namespace edh
