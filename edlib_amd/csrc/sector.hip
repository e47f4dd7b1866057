// Upload of one sector's tables to the device.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>

#include "edh_internal.h"

namespace edh {

namespace {

// All tables of a sector go to the device as one allocation and one copy:
// the host arrays are first packed (256-byte aligned) into one blob, and
// the device pointers are set once the blob has its device address.
class Packer {
 public:
  template <class T>
  void add(T*& field, const std::vector<T>& h) {
    field = nullptr;
    if (h.empty()) return;
    const size_t off = (blob_.size() + 255) & ~size_t(255);
    blob_.resize(off + h.size() * sizeof(T));
    std::memcpy(blob_.data() + off, h.data(), h.size() * sizeof(T));
    set_.push_back([&field, off](char* base) { field = reinterpret_cast<T*>(base + off); });
  }
  // Without a stage the blob must stay alive until the copy has completed
  // (the caller synchronises the stream); with one the copy reads pinned
  // staging memory that the stage keeps until the copy is done.
  void commit(DevSector& ds, hipStream_t s, UploadStage* stage) {
    if (blob_.empty()) return;
    void* p = nullptr;
    hipError_t e = hipMallocAsync(&p, blob_.size(), s);
    if (e != hipSuccess) throw Error(EDH_ERR_NOMEM, std::string("hipMallocAsync failed: ") + hipGetErrorString(e));
    ds.allocs.push_back(p);
    if (stage) {
      int c = -1;
      char* h = stage->take(blob_.size(), c);
      std::memcpy(h, blob_.data(), blob_.size());
      EDH_HIP(hipMemcpyAsync(p, h, blob_.size(), hipMemcpyHostToDevice, s));
      stage->mark(c, s);
    } else {
      EDH_HIP(hipMemcpyAsync(p, blob_.data(), blob_.size(), hipMemcpyHostToDevice, s));
    }
    for (auto& f : set_) f(static_cast<char*>(p));
  }

 private:
  std::vector<char> blob_;
  std::vector<std::function<void(char*)>> set_;
};

void upload_spin(Packer& pk, const SpinTables& sp, DevSpin& d) {
  d.D = sp.D;
  pk.add(d.bits, sp.bits);
  pk.add(d.ediag, sp.ediag);
  pk.add(d.rowp, sp.rowp);
  pk.add(d.packed, sp.packed);
  pk.add(d.sl_ptr, sp.sl_ptr);
  pk.add(d.sl_len, sp.sl_len);
  pk.add(d.ell, sp.ell);
  pk.add(d.ell16, sp.ell16);
  for (int q = 0; q < 4; ++q) d.vt16[q] = sp.vt16[q];
  d.nvt16 = sp.nvt16;
  pk.add(d.q4_ptr, sp.q4_ptr);
  pk.add(d.ell_q4, sp.ell_q4);
  pk.add(d.ell_s4, sp.ell_s4);
  d.s4_scale = sp.s4_scale;
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
  for (void* p : allocs) (void)hipFreeAsync(p, astream);
  allocs.clear();
  *this = DevSector();
}

void configure_mem_pool(int device) {
  static bool done[64] = {false};
  if (device < 0 || device >= 64 || done[device]) return;
  hipMemPool_t pool;
  EDH_HIP(hipDeviceGetDefaultMemPool(&pool, device));
  uint64_t keep = UINT64_MAX;
  EDH_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  done[device] = true;
}

char* UploadStage::take(size_t bytes, int& c) {
  // a chunk with room, else one whose copies have all completed, else a new one
  for (size_t q = 0; q < chunks.size(); ++q)
    if (chunks[q].cap - chunks[q].used >= bytes) {
      c = (int)q;
      char* h = chunks[q].host + chunks[q].used;
      chunks[q].used = (chunks[q].used + bytes + 255) & ~size_t(255);
      return h;
    }
  for (size_t q = 0; q < chunks.size(); ++q)
    if (chunks[q].cap >= bytes && hipEventQuery(chunks[q].ev) == hipSuccess) {
      c = (int)q;
      chunks[q].used = (bytes + 255) & ~size_t(255);
      return chunks[q].host;
    }
  Chunk k;
  k.cap = std::max<size_t>(size_t(16) << 20, bytes);
  EDH_HIP(hipHostMalloc((void**)&k.host, k.cap, hipHostMallocDefault));
  EDH_HIP(hipEventCreateWithFlags(&k.ev, hipEventDisableTiming));
  k.used = (bytes + 255) & ~size_t(255);
  chunks.push_back(k);
  c = (int)chunks.size() - 1;
  return k.host;
}

void UploadStage::mark(int c, hipStream_t s) { EDH_HIP(hipEventRecord(chunks[c].ev, s)); }

void UploadStage::release() {
  for (Chunk& k : chunks) {
    if (k.host) (void)hipHostFree(k.host);
    if (k.ev) (void)hipEventDestroy(k.ev);
  }
  chunks.clear();
}

void upload_sector(const SectorTables& st, int Ns, DevSector& ds, hipStream_t s, UploadStage* stage) {
  ds.release();
  ds.astream = s;
  ds.Ns = Ns;
  ds.nup = st.nup;
  ds.ndn = st.ndn;
  ds.N = st.N;
  ds.imp_ml = st.imp_ml;
  ds.table_bytes = st.table_bytes;
  Packer pk;
  upload_spin(pk, st.up, ds.up);
  upload_spin(pk, st.dn, ds.dn);
  pk.add(ds.values, st.values);
  ds.nvalues = (int)st.values.size();
  ds.cross_mode = st.cross_mode;
  ds.cross_u = st.cross_u;
  pk.add(ds.cross_imp, st.cross_imp);
  pk.add(ds.cross_rowtab, st.cross_rowtab);
  ds.kan_ml = st.kan_ml;
  if (st.kan_ml > 0) {
    // Each impurity term moves the impurity bits (a, b) -> (a', b') with the
    // same occupation counts.  In the basis order (positions in ascending
    // lexicographic order, impurity orbitals on the highest positions) the
    // states sharing their bath bits form one contiguous run ordered by the
    // impurity pattern alone, so the partner of (u, d) is (u + du, d + dd)
    // with du, dd fixed by (a, a') and (b, b'): one element offset per term,
    // du * D_dn + dd, read from the representative state of its group.
    const int ml = st.kan_ml, na = 1 << ml, nb = ds.Ns - ml;
    std::vector<int64_t> koff(st.kan.size());
    std::vector<double> kc(st.kan.size());
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < na; ++b) {
        const int idx = a * na + b;
        if (st.kan_ptr[idx + 1] == st.kan_ptr[idx]) continue;
        const int bath_up = st.nup - __builtin_popcount(a), bath_dn = st.ndn - __builtin_popcount(b);
        const int32_t lu = (int32_t)((1ll << bath_up) - 1), ld = (int32_t)((1ll << bath_dn) - 1);
        const int32_t ru = st.up.rank[lu | (a << nb)], rd = st.dn.rank[ld | (b << nb)];
        for (int q = st.kan_ptr[idx]; q < st.kan_ptr[idx + 1]; ++q) {
          const KanEntry& e = st.kan[q];
          EDH_CHECK(__builtin_popcount(e.new_a) == __builtin_popcount(a) &&
                        __builtin_popcount(e.new_b) == __builtin_popcount(b),
                    EDH_ERR_INTERNAL, "impurity term changes an occupation count");
          const int64_t du = st.up.rank[lu | (e.new_a << nb)] - ru, dd = st.dn.rank[ld | (e.new_b << nb)] - rd;
          koff[q] = du * st.dn.D + dd;
          kc[q] = e.coef;
        }
      }
    pk.add(ds.kan_ptr, st.kan_ptr);
    pk.add(ds.kan_off, koff);
    pk.add(ds.kan_c, kc);
    ds.nkan = (int)kc.size();
  }
  pk.add(ds.counters, std::vector<int32_t>(64, 0));
  pk.add(ds.row_order, rcm_order(st.up.rowp, st.up.col, st.up.D));
  pk.commit(ds, s, stage);
  if (stage) return;
  EDH_HIP(hipStreamSynchronize(s));
  ds.variant = hv_autotune(ds, s);
}

}  // namespace edh
