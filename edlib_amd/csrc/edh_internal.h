// Internal declarations shared by the edhip library sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/edhip.h"

namespace edh {

typedef long long state_t;

// Error carrying a status code across the C ABI.
struct Error : public std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define EDH_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw ::edh::Error(EDH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

#define EDH_CHECK(cond, code, msg)                 \
  do {                                             \
    if (!(cond)) throw ::edh::Error((code), (msg)); \
  } while (0)

// ---------------------------------------------------------------------------
// Binomials and the per-spin lexicographic basis (reference Combination.h:23-50,
// SzSymmetry.h:191-216).  Site s of a spin channel lives at bit (Ns-1-s).
// ---------------------------------------------------------------------------
struct Binomial {
  int n = 0;
  std::vector<int64_t> t;
  explicit Binomial(int n_ = 0);
  int64_t operator()(int a, int b) const {
    if (a < 0 || b < 0 || b > a || a > n) return 0;
    return t[a * (n + 1) + b];
  }
};

std::vector<int32_t> spin_basis(int Ns, int k);

// ---------------------------------------------------------------------------
// Model description (host side).
// ---------------------------------------------------------------------------
struct Transition {
  int kind;  // 0: c^+_{j s} c_{i s};  1: c^+_{i s} c^+_{j s'} c_{l s'} c_{k s}
  int i, j, k, l, s, sp;
  double value;
};

struct Model {
  int kind = 0;  // EDH_MODEL_HUBBARD / EDH_MODEL_SIAM
  int Ns = 0;
  int ml = 0;         // interacting (impurity) orbitals: Ns for Hubbard
  int hop_sign = -1;  // sign rule of kind-0 transitions: -s1 s2 (Hubbard) or +s1 s2 (SIAM)
  double v_scale = 0.5;  // kind-1 coefficient = v_scale * value (the SIAM's 1/2 U, :166)
  std::vector<Transition> T, V;
  // Diagonal decomposition: sum_s e[s][i] n_is + sum_s sum_{i!=j} S[s][i][j] n_is n_js
  //                        + sum_{i,j} C[i][j] n_i(up) n_j(dn); site indices.
  std::vector<double> e[2], S[2], C;
  double e0 = 0.0;  // constant part (terms-built models: diagonal of the empty state)
  // Original parameters, kept for reference-order diagonal evaluation.
  std::vector<double> t, U, J, xmu, Hmag;
  std::vector<double> Epsk, Vk, H0, Ut;
  double mu = 0.0;

  double uten(int s1, int s2, int i, int j, int k, int l) const {
    int m = ml;
    return Ut[((((s1 * 2 + s2) * m + i) * m + j) * m + k) * m + l];
  }
  int valid(const Transition& tr, state_t nst) const;
  double apply(const Transition& tr, state_t nst, state_t& k, int& sign) const;
  double diagonal_reference(state_t st) const;  // model.diagonal() in the reference's order
};

Model* build_hubbard(int Ns, const double* t, const double* U, const double* J, const double* xmu,
                     const double* Hmag);
Model* build_siam(int Ns, int ml, const double* Epsk, const double* Vk, const double* H0, const double* Ut,
                  double mu);
Model* build_terms(const edh_model_desc& d);

// ---------------------------------------------------------------------------
// Per-sector tables (host copies) and their device mirror.
// ---------------------------------------------------------------------------
enum CrossMode { CROSS_NONE = 0, CROSS_POPC = 1, CROSS_IMP = 2, CROSS_GENERAL = 3 };

// Packed hopping entry: column in the low 22 bits, value-table index in the high 10.
constexpr int kColBits = 22;
constexpr uint32_t kColMask = (1u << kColBits) - 1u;
constexpr int kMaxValues = 1 << (32 - kColBits);
constexpr int kSlice = 64;  // sliced-ELL slice height (one wavefront)

struct SpinTables {
  int nel = 0;
  int64_t D = 0;
  std::vector<int32_t> bits;    // basis patterns
  std::vector<double> ediag;    // single-spin diagonal part per basis state
  std::vector<int32_t> rowp;    // CSR row pointer (source-state rows)
  std::vector<int32_t> col;     // CSR columns (reference CRSMatrix semantics)
  std::vector<double> val;      // CSR values
  std::vector<uint32_t> packed; // packed CSR entries
  std::vector<int32_t> sl_ptr, sl_len;  // sliced ELL
  std::vector<uint32_t> ell;
  std::vector<uint8_t> ell_pad;  // 1 where the ELL slot is padding
  // Compact ELL (same slices): column in the low 14 bits, index into vt16 in
  // the high 2, when the spin has <= 16384 states and <= 4 distinct values.
  std::vector<uint16_t> ell16;
  double vt16[4] = {0, 0, 0, 0};
  int nvt16 = 0;  // distinct values used by ell16
  // The compact ELL regrouped by slot quads: quad q of lane L of slice s at
  // q4_ptr[s] + q*64 + L packs slots 4q..4q+3 (lowest 16 bits first).
  std::vector<int32_t> q4_ptr;
  std::vector<uint64_t> ell_q4;
  // Signed byte-offset ELL (same quads, 32-bit slots) when the values are
  // {a} or {a, -a}: slot = 8 * column | (value == -a), padding = 8 * D.
  std::vector<uint32_t> ell_s4;
  double s4_scale = 0.0;
  std::vector<int32_t> rank;    // rank by bit pattern (2^Ns), -1 if not in the basis
};

struct KanEntry {
  int32_t new_a, new_b;  // impurity bit patterns after the move
  double coef;           // value * sign
};

struct SectorTables {
  int nup = 0, ndn = 0;
  int64_t N = 0;
  SpinTables up, dn;
  std::vector<double> values;  // distinct hopping values (index 0 = 0.0)
  int imp_ml = 0;  // impurity (interacting) orbitals: their bits sit above nb = Ns - imp_ml bath bits
  int cross_mode = CROSS_NONE;
  double cross_u = 0.0;
  std::vector<double> cross_imp;    // (2^ml)^2 table
  std::vector<double> cross_rowtab; // Dup x Ns (GENERAL: R_u[bitpos])
  // Kanamori / general V terms on impurity bits
  int kan_ml = 0;
  std::vector<int32_t> kan_ptr;  // (2^ml)^2 + 1
  std::vector<KanEntry> kan;
  int64_t table_bytes = 0;  // per-spin hopping table bytes (algorithmic T)
};

void build_sector(const Model& m, int nup, int ndn, SectorTables& st, bool need_rank);

// H*v variants: single pass (rows, up-spin term from the cache hierarchy) or
// two passes (rows without the up-spin term, then XCD-aware column slabs for
// it).  HV_AUTO picks the one timed fastest for the sector.
enum HvVariant { HV_SINGLE = 0, HV_TWO_PASS = 1, HV_AUTO = -1 };

// Device mirror of one sector.
struct DevSpin {
  int64_t D = 0;
  int32_t* bits = nullptr;
  double* ediag = nullptr;
  int32_t* rowp = nullptr;
  uint32_t* packed = nullptr;
  int32_t* sl_ptr = nullptr;
  int32_t* sl_len = nullptr;
  uint32_t* ell = nullptr;
  uint16_t* ell16 = nullptr;
  double vt16[4] = {0, 0, 0, 0};
  int nvt16 = 0;
  int32_t* q4_ptr = nullptr;
  uint64_t* ell_q4 = nullptr;
  uint32_t* ell_s4 = nullptr;
  double s4_scale = 0.0;
};

struct DevSector {
  int Ns = 0, nup = 0, ndn = 0;
  int64_t N = 0;
  DevSpin up, dn;
  double* values = nullptr;
  int nvalues = 0;  // entries of values
  int nkan = 0;     // impurity terms (entries of kan_off and kan_c)
  int imp_ml = 0;
  int cross_mode = 0;
  double cross_u = 0.0;
  double* cross_imp = nullptr;
  double* cross_rowtab = nullptr;
  int kan_ml = 0;
  int32_t* kan_ptr = nullptr;
  int64_t* kan_off = nullptr;  // element offset of each term's partner state
  double* kan_c = nullptr;
  int32_t* row_order = nullptr;  // processing order of up rows
  int variant = HV_SINGLE;       // chosen H*v variant
  int* counters = nullptr;       // work queues of the slab pass (8 ints)
  int64_t table_bytes = 0;
  std::vector<void*> allocs;
  hipStream_t astream = nullptr;  // stream the tables were allocated on (stream-ordered pool)
  void release();
};

// Pinned staging for table uploads that must not wait for the stream:
// chunks of pinned host memory, each reusable once the copies out of it
// (marked by its event) have completed.
struct UploadStage {
  struct Chunk {
    char* host = nullptr;
    size_t cap = 0, used = 0;
    hipEvent_t ev = nullptr;
  };
  std::vector<Chunk> chunks;
  // pinned space for `bytes`; the caller copies out of it on stream s and
  // calls mark(c, s) after that copy was issued
  char* take(size_t bytes, int& c);
  void mark(int c, hipStream_t s);
  void release();  // after the stream has finished every copy
};

// Uploads a sector's tables in one allocation and one copy.  Without a
// stage the call waits for the copy (and then times the H*v variants);
// with one it returns at once and the variants are not timed.
void upload_sector(const SectorTables& st, int Ns, DevSector& ds, hipStream_t s, UploadStage* stage = nullptr);

// Keeps freed stream-ordered allocations in the device's default pool (no
// return to the driver at synchronisations), so that allocation in the
// solve loops costs no driver call and no device-wide wait.
void configure_mem_pool(int device);

// H*v launch parameters.
struct HvArgs {
  const double* v;      // gathered operand (scaled by vscale on load)
  double* w;            // output
  const double* wprev;  // optional additive input (scaled by wscale), may alias w
  double vscale;        // multiplies every gathered v value
  double wscale;        // multiplies wprev
  int mode;             // 0: w = Hv ; 1: w = Hv + wscale*wprev
  double* dot_partial;  // optional: per-block partial sums of v*w (after scale)
  const double* dscale = nullptr;  // optional device scalars {vscale, wscale} overriding the values above
  const int* done = nullptr;       // optional device stop flag: the launch does nothing when set
  // optional eigensolver control block (grouped H*v only): v and w are then
  // columns col and col + 1 of the basis held in v (phase 0) or v_alt
  // (phase 1), and a run whose state is not 0 is skipped
  const double* ctl = nullptr;
  const double* v_alt = nullptr;
  // optional (grouped H*v only): per-task shares of <v, w> (v scaled), at
  // least hv_batch_tasks(D_up, D_dn) of them
  double* tdot = nullptr;
  // optional (grouped H*v only): a Lanczos run's step counter k (a double on
  // the device); v, wprev and w are then given as the run's three buffers
  // b0, b1, b2 and the launch takes v = b[r], wprev = b[(r + 1) % 3],
  // w = b[(r + 2) % 3] with r = (3 - (k - 1) % 3) % 3, accumulating wprev
  // from the second step on (the buffer rotation of Lanczos.h:47-51)
  const double* lzk = nullptr;
};

// Buffer of a Lanczos run's three in rotation at step k >= 1: role 0 = the
// current vector, 1 = the previous one, 2 = the H*v output.
__host__ __device__ inline int lz_buffer(int k, int role) { return ((3 - (k - 1) % 3) % 3 + role) % 3; }

// Control block of a batched eigensolver run (doubles, on the device): the
// basis column being extended, which of the two basis buffers holds the
// basis, and the run state (0 = active).
constexpr int kCtlCol = 0, kCtlPhase = 1, kCtlState = 2;

void launch_hv(const DevSector& ds, const HvArgs& a, hipStream_t s, int variant = HV_AUTO);
int hv_num_blocks(const DevSector& ds);
int hv_autotune(DevSector& ds, hipStream_t s);

// Grouped H*v over many operands (sectors of any models): one launch covers
// the up-rows of all of them, cut into work units.  A job record is filled
// on the host (hv_batch_fill) into a buffer of hv_batch_job_bytes() per job
// and copied to the device with the unit list.
// Work unit of the grouped H*v: rows [row0, row0 + nrows) of job `job`.
struct HvBatchUnit {
  int job, row0, nrows, pad;
};
size_t hv_batch_job_bytes();
void hv_batch_fill(const DevSector& ds, const HvArgs& a, int64_t row0, void* out);
int hv_batch_unit_rows(int64_t ddn);  // rows per unit for a down dimension (a multiple of the row group)
int hv_batch_row_group(int64_t ddn);  // rows per wavefront task
int64_t hv_batch_tasks(int64_t dup, int64_t ddn);  // wavefront tasks of a whole operand
// LDS elements for a launch whose units span `spans` elements each (units
// above the result run unstaged)
int hv_batch_lds_elems(const std::vector<int64_t>& spans);
void launch_hv_batch(const void* djobs, const HvBatchUnit* dunits, int nunits, int block, int lds_elems,
                     hipStream_t s);
// Large operands (N >= 2^22 with D_dn > 4608, or with D_up and D_dn both
// >= 1024) of a batch go to the single-sector rows kernel instead (their
// rows staged whole in LDS, up to 1024-thread work-groups); the choice
// depends on the operand alone.  Its
// Lanczos <v, w> shares are one per work-group (hv_batch_rows_partials).
bool hv_batch_routed(int64_t dup, int64_t ddn);
// The same for Lanczos operands (Green's-function runs): also sectors of at
// least 256 rows of 512..4608 elements whose format has the multi-operand
// kernel, so that their runs share launches (the decision depends on the
// sector only, so a run's arithmetic does not depend on the other runs).
bool hv_batch_routed_lanczos(const DevSector& ds);
int64_t hv_batch_rows_partials(int64_t dup);
void launch_hv_batch_rows(const DevSector& ds, const HvArgs& a, hipStream_t s);
// Several Lanczos operands of one rows-kernel sector in one launch, each with
// the single-operand launch's arithmetic: hv_batch_rows_multi_fit gives how
// many fit one launch (LDS), 0 when the sector's format has no such kernel;
// the operands carry device scales (dscale) and their <v, w> shares (tdot).
constexpr int kMultiMax = 4;
int hv_batch_rows_multi_fit(const DevSector& ds);
void launch_hv_batch_rows_multi(const DevSector& ds, const HvArgs* a, int count, hipStream_t s);

}  // namespace edh
