// Upload of one sector's tables to the device.
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
  d.rank = upload(ds, sp.rank, s);
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
  std::vector<int32_t> order(st.up.D);
  for (int64_t u = 0; u < st.up.D; ++u) order[u] = (int32_t)u;
  ds.row_order = upload(ds, order, s);
  EDH_HIP(hipStreamSynchronize(s));
}

}  // namespace edh
