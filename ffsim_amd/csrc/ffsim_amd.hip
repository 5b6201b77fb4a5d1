// ffsim_amd: C ABI, device context and launch orchestration (gfx950).
//
// Layout in HBM: a state is a row-major dim_a x dim_b complex matrix (alpha
// strings index rows).  An orbital rotation of the alpha sector runs as
// fused window passes in MODE_ROWS over the state; the beta sector either
// runs in MODE_COLS directly (when a whole beta string vector fits in one LDS
// tile) or, for large norb, on the transposed state: the last alpha pass
// stores the transposed state into the caller's work buffer, the beta passes
// run in MODE_ROWS there, and the last beta pass stores back transposed.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ffsim_amd.h"
#include "diag_params.hpp"
#include "host_core.hpp"
#include "kernels.hpp"
#include "status.hpp"

namespace ffsim_amd {

// ------------------------------------------------------------------ errors
namespace {
thread_local std::string g_last_error;
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

// Resolves the allocation-free entry point's device scratch on the current
// device (noalloc.hip).
void noalloc_setup();

// ------------------------------------------------------------- device state
__global__ void fill_zero_kernel(double* __restrict__ x, long n) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
    x[k] = 0.0;
}

__global__ void diag_col_term_kernel(int n_cols, const uint64_t* __restrict__ col_str,
                                     const DiagHeader* __restrict__ hdr,
                                     const double* __restrict__ coef, double* __restrict__ gamma,
                                     int evolve, long coef_stride, long gamma_stride) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  hdr += blockIdx.y;
  coef += blockIdx.y * coef_stride;
  gamma += blockIdx.y * gamma_stride;
  const int n = hdr->norb;
  const int z = hdr->z_rep;
  const int flags = hdr->flags;
  const double* A_col = coef + n * n;
  const double* lin_col = coef + 3 * n * n + n;
  const uint64_t bits = col_str[c];
  double acc = 0.0;
  if (flags & DF_QUAD) {
    for (int j = 0; j < n; ++j) {
      const int oj = (bits >> j) & 1;
      for (int k = j; k < n; ++k) {
        const int ok = (bits >> k) & 1;
        if (z) {
          if (k > j) acc += ((oj ^ ok) ? -1.0 : 1.0) * A_col[j * n + k];
        } else if (oj && ok) {
          acc += A_col[j * n + k];
        }
      }
    }
  }
  if (flags & DF_LIN) {
    for (int j = 0; j < n; ++j)
      if ((bits >> j) & 1) acc += lin_col[j];
  }
  if (evolve) {
    double sn, cs;
    sincos(-hdr->scale * acc, &sn, &cs);
    gamma[2 * c] = cs;
    gamma[2 * c + 1] = sn;
  } else {
    gamma[c] = acc;
  }
}

// Factors of a diagonal Coulomb evolution for the strings of the state's row
// sector (number representation), used by the fused store: for string a,
// out[a][q] = exp(-i s sum_{p in a} X[p][q]) for q < norb and
// out[a][norb] = exp(-i s rho_a) (quadratic and linear row terms).
__global__ void diag_row_factor_kernel(int n_str, const uint64_t* __restrict__ str,
                                       const DiagHeader* __restrict__ hdr,
                                       const double* __restrict__ coef,
                                       Cx<double>* __restrict__ out) {
  const int n = hdr->norb, nf = n + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n_str * nf) return;
  const int a = (int)(t / nf), q = (int)(t % nf);
  const int flags = hdr->flags;
  const double* A_row = coef;
  const double* X = coef + 2 * n * n;
  const double* lin_row = coef + 3 * n * n;
  const uint64_t bits = str[a];
  double e = 0.0;
  if (q < n) {
    if (flags & DF_CROSS)
      for (int p = 0; p < n; ++p)
        if ((bits >> p) & 1) e += X[p * n + q];
  } else {
    if (flags & DF_QUAD)
      for (int j = 0; j < n; ++j)
        if ((bits >> j) & 1)
          for (int k = j; k < n; ++k)
            if ((bits >> k) & 1) e += A_row[j * n + k];
    if (flags & DF_LIN)
      for (int j = 0; j < n; ++j)
        if ((bits >> j) & 1) e += lin_row[j];
  }
  double sn, cs;
  sincos(-hdr->scale * e, &sn, &cs);
  out[t] = Cx<double>{cs, sn};
}

namespace {

struct DevStrings {
  uint64_t* d = nullptr;
  int64_t dim = 0;
};
// One-body link tables of one (norb, nocc): per target string, n_links pairs
// (source string, p | q << 8 | negative sign << 16).
struct DevLinks {
  int32_t* d = nullptr;
  int n_links = 0;
};
struct DevWindow {
  Fiber* fibers = nullptr;
  int32_t* rows = nullptr;
  int n_fibers = 0, n_active = 0, max_rows = 0, max_rows_active = 0, lo = 0, w = 0;
  long rows_all = 0, rows_active = 0;  // strings covered by all / by the active fibers
};
// Pair tables of one rotation (i, j) for (norb, nocc): the strings with i
// occupied / j empty in ascending address order, and their partners.
struct DevPairs {
  int64_t* first = nullptr;
  int64_t* second = nullptr;
  long n = 0;
};
struct DevSchedule {
  uint32_t* chunks = nullptr;
  int32_t* index = nullptr;
};
struct DevGroups {
  uint32_t* desc = nullptr;
  uint16_t* members = nullptr;
  int32_t* index = nullptr;
  std::vector<int32_t> host_index;  // copy of index for LDS sizing on the host
  GroupTables host;                 // host tables (wave schedules are built from them)
  std::map<std::pair<int, int>, DevSchedule> schedules;  // (n_waves, W)
  int w = 0, w2 = 0, n_lo2 = 0;
  GroupArgs args(const DevSchedule& ds) const {
    return GroupArgs{desc, members, index, n_lo2, w2, ds.chunks, ds.index};
  }
};

// Program ring: per-call gate parameters (Givens steps, phases, diagonal
// operator coefficients) are staged on the host, copied once per call through
// a pinned slot to the matching device slot, and a slot is reused only after
// the event recorded behind the call's last kernel has completed.  A slot's
// device buffer holds the parameters followed by the call's device-only
// scratch (written by kernels, never copied).  Slot buffers are allocated on
// first use and grown when a call needs more.
struct ProgramRing {
  static constexpr int kSlots = 64;
  static constexpr size_t kMinBytes = 1u << 20;
  static constexpr size_t kMaxBytes = 256u << 20;  // parameters + scratch of one call
  char* h[kSlots] = {};
  char* d[kSlots] = {};
  size_t cap_h[kSlots] = {}, cap_d[kSlots] = {};
  hipEvent_t ev[kSlots] = {};
  bool armed[kSlots] = {};  // ev[slot] is recorded behind the slot's last use
  bool busy[kSlots] = {};   // a call owns the slot (from Program construction to finish)
  int next = 0;
  std::condition_variable freed;
};

struct Context {
  // rocBLAS handle of the batched small-state path (created on first use;
  // calls that use it hold blas_mu, the handle being bound to one stream at a time)
  rocblas_handle blas = nullptr;
  std::mutex blas_mu;
  int device = -1;
  size_t lds_max = 65536;
  int n_cu = 0;
  std::mutex mu;
  std::map<std::pair<int, int>, DevStrings> strings;
  std::map<std::pair<int, int>, DevLinks> links;
  std::map<std::tuple<int, int, int, int>, DevWindow> windows;
  std::map<std::pair<int, bool>, DevGroups> groups;  // (window width, merged complements)
  std::map<std::tuple<int, int, int, int>, DevPairs> pairs;  // (norb, nocc, i, j)
  // one-body window passes: fibers of (norb, nocc, window mask, between mask)
  // and the link tables of each window width
  std::map<std::tuple<int, int, uint64_t, uint64_t>, DevWindow> mask_windows;
  // streams on which the Givens path of a batch runs several small states at
  // once (created on first use, default priority)
  std::vector<hipStream_t> aux;
  std::map<int, std::pair<uint32_t*, int32_t*>> ob_links;
  ProgramRing ring;
};

// Kernel timing for benchmarks: when enabled, every fiber / diagonal kernel
// launch is bracketed by HIP events on its own stream; totals are read back
// (after synchronizing) by ffsim_amd_profile_end.
struct Profiler {
  std::mutex mu;
  bool on = false;
  std::vector<hipEvent_t> start, stop;
  std::vector<int> kind;
  std::vector<double> bytes;
  size_t used = 0;
};
Profiler g_prof;

enum { PROF_FIBER = 0, PROF_DIAG = 1, PROF_KINDS = 2 };

// Returns the slot index to close with prof_stop, or -1 when profiling is off.
long prof_start(int kind, double bytes, hipStream_t stream) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  if (!g_prof.on) return -1;
  if (g_prof.used == g_prof.start.size()) {
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    g_prof.start.push_back(a);
    g_prof.stop.push_back(b);
    g_prof.kind.push_back(0);
    g_prof.bytes.push_back(0);
  }
  const size_t k = g_prof.used++;
  g_prof.kind[k] = kind;
  g_prof.bytes[k] = bytes;
  HIP_CHECK(hipEventRecord(g_prof.start[k], stream));
  return (long)k;
}
void prof_stop(long k, hipStream_t stream) {
  if (k < 0) return;
  HIP_CHECK(hipEventRecord(g_prof.stop[k], stream));
}

std::mutex g_ctx_mu;
std::map<int, std::unique_ptr<Context>> g_ctx;
int g_n_cu = 256;  // compute units of the current device (persistent grids)

Context& context() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  auto it = g_ctx.find(dev);
  if (it != g_ctx.end()) return *it->second;
  auto ctx = std::make_unique<Context>();
  ctx->device = dev;
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  ctx->lds_max = prop.sharedMemPerBlock;
  ctx->n_cu = prop.multiProcessorCount;
  g_n_cu = std::max(1, prop.multiProcessorCount);
  for (int k = 0; k < ProgramRing::kSlots; ++k)
    HIP_CHECK(hipEventCreateWithFlags(&ctx->ring.ev[k], hipEventDisableTiming));
  Context& ref = *ctx;
  g_ctx[dev] = std::move(ctx);
  return ref;
}

template <typename T>
T* upload_table(const std::vector<T>& v) {
  T* d = nullptr;
  size_t bytes = std::max<size_t>(v.size() * sizeof(T), sizeof(T));
  HIP_CHECK(hipMalloc((void**)&d, bytes));
  if (!v.empty()) HIP_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

const DevStrings& dev_strings(Context& ctx, int norb, int nocc) {
  std::lock_guard<std::mutex> lk(ctx.mu);
  auto key = std::make_pair(norb, nocc);
  auto it = ctx.strings.find(key);
  if (it != ctx.strings.end()) return it->second;
  std::vector<uint64_t> s = make_strings(norb, nocc);
  DevStrings ds;
  ds.dim = (int64_t)s.size();
  ds.d = upload_table(s);
  return ctx.strings[key] = ds;
}

// Links of every string T of (norb, nocc): for p in T and q = p or q not in T,
// the source S = T - p + q and the sign of a+_p a_q |S> (Jordan-Wigner order of
// ascending orbitals).  Addresses are colex ranks.
const DevLinks& dev_links(Context& ctx, int norb, int nocc) {
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    auto it = ctx.links.find({norb, nocc});
    if (it != ctx.links.end()) return it->second;
  }
  std::vector<uint64_t> strs = make_strings(norb, nocc);
  const int n_links = nocc * (norb - nocc + 1);
  std::vector<int32_t> tab((size_t)strs.size() * n_links * 2);
  for (size_t t = 0; t < strs.size(); ++t) {
    const uint64_t T = strs[t];
    int l = 0;
    for (int p = 0; p < norb; ++p) {
      if (!((T >> p) & 1)) continue;
      for (int q = 0; q < norb; ++q) {
        if (q != p && ((T >> q) & 1)) continue;
        const uint64_t S = (T & ~(1ull << p)) | (1ull << q);
        const uint64_t S0 = S & ~(1ull << q);
        const int par = (__builtin_popcountll(S & ((1ull << q) - 1)) +
                         __builtin_popcountll(S0 & ((1ull << p) - 1))) & 1;
        tab[(t * n_links + l) * 2] = (int32_t)string_rank(S);
        tab[(t * n_links + l) * 2 + 1] = p | (q << 8) | (par << 16);
        ++l;
      }
    }
  }
  DevLinks dl;
  dl.n_links = n_links;
  dl.d = upload_table(tab);
  std::lock_guard<std::mutex> lk(ctx.mu);
  return ctx.links.emplace(std::make_pair(norb, nocc), dl).first->second;
}

const DevWindow& dev_mask_window(Context& ctx, int norb, int nocc, uint64_t mask,
                                 uint64_t between) {
  std::lock_guard<std::mutex> lk(ctx.mu);
  auto key = std::make_tuple(norb, nocc, mask, between);
  auto it = ctx.mask_windows.find(key);
  if (it != ctx.mask_windows.end()) return it->second;
  WindowLayout L = make_mask_layout(norb, nocc, mask, between);
  DevWindow dw;
  dw.fibers = upload_table(L.fibers);
  dw.rows = upload_table(L.row_table);
  dw.n_fibers = (int)L.fibers.size();
  dw.n_active = L.n_active;
  dw.max_rows = L.max_rows;
  dw.max_rows_active = L.max_rows_active;
  dw.w = L.w;
  dw.rows_all = (long)L.row_table.size();
  dw.rows_active = dw.rows_all;
  return ctx.mask_windows.emplace(key, dw).first->second;
}

std::pair<uint32_t*, int32_t*> dev_one_body_links(Context& ctx, int w) {
  std::lock_guard<std::mutex> lk(ctx.mu);
  auto it = ctx.ob_links.find(w);
  if (it != ctx.ob_links.end()) return it->second;
  OneBodyLinks B = make_one_body_links(w);
  std::pair<uint32_t*, int32_t*> d(upload_table(B.links), upload_table(B.index));
  ctx.ob_links.emplace(w, d);
  return d;
}

const DevWindow& dev_window(Context& ctx, int norb, int nocc, int lo, int w) {
  std::lock_guard<std::mutex> lk(ctx.mu);
  auto key = std::make_tuple(norb, nocc, lo, w);
  auto it = ctx.windows.find(key);
  if (it != ctx.windows.end()) return it->second;
  WindowLayout L = make_window_layout(norb, nocc, lo, w);
  DevWindow dw;
  dw.fibers = upload_table(L.fibers);
  dw.rows = upload_table(L.row_table);
  dw.n_fibers = (int)L.fibers.size();
  dw.n_active = L.n_active;
  dw.max_rows = L.max_rows;
  dw.max_rows_active = L.max_rows_active;
  dw.lo = lo;
  dw.w = w;
  for (size_t f = 0; f < L.fibers.size(); ++f) {
    dw.rows_all += L.fibers[f].n_rows;
    if ((int)f < L.n_active) dw.rows_active += L.fibers[f].n_rows;
  }
  return ctx.windows[key] = dw;
}

const DevPairs& dev_pairs(Context& ctx, int norb, int nocc, int i, int j) {
  std::lock_guard<std::mutex> lk(ctx.mu);
  auto key = std::make_tuple(norb, nocc, i, j);
  auto it = ctx.pairs.find(key);
  if (it != ctx.pairs.end()) return it->second;
  std::vector<uint64_t> s = make_strings(norb, nocc);
  std::vector<int64_t> first, second;
  for (size_t a = 0; a < s.size(); ++a) {
    const uint64_t b = s[a];
    if (((b >> i) & 1) && !((b >> j) & 1)) {
      first.push_back((int64_t)a);
      second.push_back((int64_t)string_rank(b ^ (1ull << i) ^ (1ull << j)));
    }
  }
  DevPairs dp;
  dp.n = (long)first.size();
  dp.first = upload_table(first);
  dp.second = upload_table(second);
  return ctx.pairs[key] = dp;
}

// Group tables of a window width.  c64 tables merge complemented groups into
// the mirrored class (fewer, fuller chunks: 4.6% faster); the c128 kernel keeps
// Re(sigma) in SGPRs, which complemented lanes cannot share (4% slower merged).
const DevGroups& dev_groups(Context& ctx, int w, bool merge) {
  std::lock_guard<std::mutex> lk(ctx.mu);
  const auto key = std::make_pair(w, merge);
  auto it = ctx.groups.find(key);
  if (it != ctx.groups.end()) return it->second;
  GroupTables T = make_group_tables(w, kMaxSubWindow, merge);
  DevGroups dg;
  dg.desc = upload_table(T.desc);
  dg.members = upload_table(T.members);
  dg.index = upload_table(T.index);
  dg.host_index = T.index;
  dg.host = T;
  dg.w = w;
  dg.w2 = T.w2;
  dg.n_lo2 = T.n_lo2;
  return ctx.groups[key] = dg;
}

const DevSchedule& dev_schedule(Context& ctx, const DevGroups& gr, int n_waves, int W) {
  std::lock_guard<std::mutex> lk(ctx.mu);
  DevGroups& g = const_cast<DevGroups&>(gr);
  auto key = std::make_pair(n_waves, W);
  auto it = g.schedules.find(key);
  if (it != g.schedules.end()) return it->second;
  WaveSchedule ws = make_wave_schedule(g.host, n_waves, std::max(1, 64 / W));
  DevSchedule ds;
  ds.chunks = upload_table(ws.chunks);
  ds.index = upload_table(ws.index);
  return g.schedules[key] = ds;
}

size_t pow2_ceil(size_t x) {
  size_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// A call's program: parameter bytes are appended on the host, device-only
// scratch is reserved, and both get device addresses at commit().  The slot is
// the call's from construction until finish() (or destruction): other threads
// skip it, and waiting for its previous use happens outside the context lock.
struct Program {
  Context* ctx;
  int slot = -1;
  hipStream_t stream;
  std::vector<char> host;      // parameter bytes
  size_t scratch_bytes = 0;    // device-only scratch, placed after the parameters
  size_t scratch_base = 0;     // device offset of the scratch (known at commit)
  bool prepared = false, committed = false, finished = false;

  Program(Context& c, hipStream_t s) : ctx(&c), stream(s) {
    ProgramRing& R = c.ring;
    std::unique_lock<std::mutex> lk(c.mu);
    for (;;) {
      for (int k = 0; k < ProgramRing::kSlots && slot < 0; ++k) {
        const int cand = (R.next + k) % ProgramRing::kSlots;
        if (!R.busy[cand]) slot = cand;
      }
      if (slot >= 0) break;
      R.freed.wait(lk);
    }
    R.next = (slot + 1) % ProgramRing::kSlots;
    R.busy[slot] = true;
    const bool wait = R.armed[slot];
    R.armed[slot] = false;
    lk.unlock();
    if (wait) {
      const hipError_t e = hipEventSynchronize(R.ev[slot]);
      if (e != hipSuccess) {
        release(false);
        throw Failure{FFSIM_ERR_HIP, std::string("program slot wait: ") + hipGetErrorString(e)};
      }
    }
  }
  Program(const Program&) = delete;
  Program& operator=(const Program&) = delete;
  ~Program() {
    if (finished) return;
    // an error left the call unfinished: anything already enqueued keeps the
    // slot until it has run
    bool armed = false;
    if (committed) armed = hipEventRecord(ctx->ring.ev[slot], stream) == hipSuccess;
    release(armed);
  }
  void release(bool armed) {
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      ctx->ring.armed[slot] = armed;
      ctx->ring.busy[slot] = false;
    }
    ctx->ring.freed.notify_one();
  }
  // Appends n bytes (16-byte aligned; zeros when src is null), returns the offset.
  size_t append(const void* src, size_t n) {
    require(!prepared, FFSIM_ERR_VALUE, "program appended after its layout was fixed");
    const size_t off = (host.size() + 15) & ~size_t(15);
    require(off + n + scratch_bytes <= ProgramRing::kMaxBytes, FFSIM_ERR_VALUE,
            "gate program too large");
    host.resize(off + n, 0);
    if (src && n) std::memcpy(host.data() + off, src, n);
    return off;
  }
  // Reserves n bytes of device-only scratch; false when the call would exceed
  // the slot limit.  Addresses via scratch<T>(off) after commit().
  bool reserve(size_t n, size_t* off) {
    require(!prepared, FFSIM_ERR_VALUE, "program scratch reserved after its layout was fixed");
    const size_t o = (scratch_bytes + 255) & ~size_t(255);
    if (o + n + host.size() + 256 > ProgramRing::kMaxBytes) return false;
    scratch_bytes = o + n;
    *off = o;
    return true;
  }
  template <typename T>
  T* dev(size_t off) const {
    return reinterpret_cast<T*>(ctx->ring.d[slot] + off);
  }
  template <typename T>
  T* scratch(size_t off) const {
    return reinterpret_cast<T*>(ctx->ring.d[slot] + scratch_base + off);
  }
  // Fixes the layout and the slot's buffers: device addresses (dev, scratch)
  // are valid from here on; nothing may be appended or reserved afterwards.
  void prepare() {
    if (prepared) return;
    ProgramRing& R = ctx->ring;
    const size_t used = host.size();
    scratch_base = (used + 255) & ~size_t(255);
    const size_t need_d = scratch_base + scratch_bytes;
    // the slot's previous use has completed, so its buffers can be replaced
    if (R.cap_h[slot] < used) {
      if (R.h[slot]) HIP_CHECK(hipHostFree(R.h[slot]));
      R.h[slot] = nullptr;
      R.cap_h[slot] = 0;
      const size_t cap = std::max(ProgramRing::kMinBytes, pow2_ceil(used));
      HIP_CHECK(hipHostMalloc((void**)&R.h[slot], cap, hipHostMallocDefault));
      R.cap_h[slot] = cap;
    }
    if (R.cap_d[slot] < need_d || !R.d[slot]) {
      if (R.d[slot]) HIP_CHECK(hipFree(R.d[slot]));
      R.d[slot] = nullptr;
      R.cap_d[slot] = 0;
      const size_t cap = std::max(ProgramRing::kMinBytes, pow2_ceil(need_d));
      HIP_CHECK(hipMalloc((void**)&R.d[slot], cap));
      R.cap_d[slot] = cap;
    }
    prepared = true;
  }
  // Overwrites n appended bytes at off (before commit).
  void write(size_t off, const void* src, size_t n) {
    require(!committed && off + n <= host.size(), FFSIM_ERR_VALUE, "program write out of range");
    std::memcpy(host.data() + off, src, n);
  }
  void commit() {
    prepare();
    ProgramRing& R = ctx->ring;
    const size_t used = host.size();
    if (used) {
      std::memcpy(R.h[slot], host.data(), used);
      HIP_CHECK(hipMemcpyAsync(R.d[slot], R.h[slot], used, hipMemcpyHostToDevice, stream));
    }
    committed = true;
  }
  void finish() {
    HIP_CHECK(hipEventRecord(ctx->ring.ev[slot], stream));
    finished = true;
    release(true);
  }
};

size_t elem_bytes(int dtype) { return dtype == FFSIM_C64 ? 8 : 16; }

// LDS kept free for the staged program of a pass (sub-passes + steps).
constexpr size_t kProgramLdsReserve = 20480;
// LDS kept free beside the tile, row indices and row strings of a window pass.
constexpr size_t kGroupLdsReserve = 16384;

int tile_width(int dtype) {
  static int env_w = [] {
    const char* e = std::getenv("FFSIM_AMD_TILE_W");
    return e ? std::atoi(e) : 0;
  }();
  if (env_w == 4 || env_w == 8 || env_w == 16) return env_w;
  // c128: 8 columns (128-byte row segments, one tile per CU).  c64: 4 columns,
  // so that several independent tiles share a CU and one tile's barrier waits
  // overlap another's work (measured 8% faster than 8 columns, 20% than 16).
  return dtype == FFSIM_C64 ? 4 : 8;
}

// The beta rotation runs in place with strings on columns when a whole beta
// vector fits one tile and its window (all norb orbitals) is no wider than
// the row-mode window cap, whose group tables the LDS reserve is sized for.
int window_cap();
bool beta_in_columns(int norb, long dim_b, int dtype, const Context& ctx) {
  return norb <= window_cap() &&
         (size_t)dim_b * (tile_width(dtype) + 1) * elem_bytes(dtype) + kProgramLdsReserve <=
             ctx.lds_max &&
         dim_b < 65536;
}

// Largest fiber of a window of width w for nocc electrons on norb orbitals.
int max_fiber(int norb, int nocc, int w) {
  int lo_k = std::max(0, nocc - (norb - w)), hi_k = std::min(nocc, w);
  uint64_t m = 0;
  for (int kp = lo_k; kp <= hi_k; ++kp) m = std::max(m, binom(w, kp));
  return (int)m;
}

int window_cap() {
  static int cap = [] {
    const char* e = std::getenv("FFSIM_AMD_MAX_WINDOW");
    const int v = e ? std::atoi(e) : 0;
    return (v >= 2 && v <= 16) ? v : 12;
  }();
  return cap;
}

// Largest window (up to the cap) whose fiber tiles fit in LDS.
int choose_window(Context& ctx, int norb, int nocc, int dtype) {
  const int W = tile_width(dtype);
  // tile row + row index + row string (phase passes); the reserve holds the group
  // tables, wave schedule and orbital phases
  const size_t per_row = (size_t)(W + 1) * elem_bytes(dtype) + 4 + (norb <= 32 ? 4 : 8);
  int best = std::min(norb, 2);
  for (int w = std::min(norb, window_cap()); w >= 2; --w) {
    if ((size_t)max_fiber(norb, nocc, w) * per_row + kGroupLdsReserve <= ctx.lds_max) {
      best = w;
      break;
    }
  }
  if (norb <= 1) best = std::max(norb, 1);
  return best;
}

// LDS of a phase pass beyond the tile: orbital phases and one string per row,
// and with a fused diagonal the batch columns' factors and the store's tables.
size_t diag_lds_bytes(bool diag, bool fuse, int n_diag, int rows, int W, int n_fac, int win) {
  if (!diag && !fuse) return 0;
  size_t b = 2 * kMaxDiagOrbitals * sizeof(double) + align16((size_t)rows * (n_diag <= 32 ? 4 : 8));
  if (fuse)
    b += ((size_t)W * n_fac + (size_t)fuse_nibble_entries(win, W)) * 2 * sizeof(double);
  return b;
}

// LDS of a fiber launch, following the kernel's layout (the tile stride is
// padded in column mode only); n_waves sizes the wave schedule.
size_t fiber_lds_bytes(size_t elem, int W, bool cols, int n_waves, int rows, int gd_cap,
                       int gm_cap, bool diag, bool fuse, int n_diag, int n_fac, int win) {
  const int S = (cols && W >= 8) ? W + 1 : W;
  return align16((size_t)rows * S * elem) + align16((size_t)rows * 4) +
         kMaxLoSlots * sizeof(GroupSlot) + align16((size_t)gd_cap * 4) +
         align16((size_t)gm_cap * 2) + (size_t)kMaxLoSlots * n_waves * 8 +
         align16((size_t)gd_cap * 4) + diag_lds_bytes(diag, fuse, n_diag, rows, W, n_fac, win);
}

// mt: several states of one shape (in place: src and dst unused; no
// transposing store, no fused diagonal), their tiles in one persistent
// launch; null for one state.
template <typename T, int W, int MODE, int W2, int NTHR, bool FUSE = false, bool MULTI = false>
void launch_fiber_t(const void* src, void* dst, long ld_src, long ld_dst, int n_batch, bool tstore,
                    const DevWindow& win, bool all_fibers, const DevGroups& gr, const LoSlots& los,
                    int gd_cap, int gm_cap, const SubPass* d_subs, int n_sub,
                    const LayerPair* d_layers, int n_layers, const double* d_diag,
                    const uint64_t* d_strings, int n_diag, size_t lds_max, hipStream_t stream,
                    const DiagFuse& fz = DiagFuse{}, const MultiTiles* mt = nullptr) {
  static_assert(!(MULTI && FUSE), "the multi-state pass has no fused diagonal");
  require(MULTI == (mt != nullptr) && (!MULTI || !tstore), FFSIM_ERR_VALUE,
          "fiber launch: inconsistent multi-state arguments");
  const int n_states = MULTI ? mt->n_states : 1;
  const int n_f = all_fibers ? win.n_fibers : win.n_active;
  if (n_f == 0 || n_batch == 0 || n_states == 0) return;
  const int rows = all_fibers ? win.max_rows : win.max_rows_active;
  const size_t lds = fiber_lds_bytes(sizeof(Cx<T>), W, MODE == MODE_COLS, NTHR / 64, rows,
                                     gd_cap, gm_cap, d_diag != nullptr, FUSE, n_diag, fz.n_fac,
                                     fz.win);
  require(lds <= lds_max, FFSIM_ERR_VALUE, "fiber tile exceeds LDS");
  auto kern = fiber_kernel<T, W, MODE, NTHR, W2, FUSE, MULTI>;
  const DevSchedule& sched = dev_schedule(context(), gr, NTHR / 64, W);
  static bool attr_set = false;
  if (!attr_set) {
    HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_max));
    attr_set = true;
  }
  const int n_bt = (n_batch + W - 1) / W;
  const long n_tiles = (long)n_bt * n_f * n_states;
  require(n_tiles < (1l << 40), FFSIM_ERR_VALUE, "grid too large");
  // persistent grid: as many blocks as fit on the device at once
  int occ = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)kern, NTHR, lds));
  const long per_cu = std::max(1, occ);
  const long grid = std::min<long>(n_tiles, per_cu * g_n_cu);
  // algorithmic bytes: every element of every launched fiber read once and written once
  const double moved = 2.0 * (double)sizeof(Cx<T>) * (double)n_batch * (double)n_states *
                       (double)(all_fibers ? win.rows_all : win.rows_active);
  const long pk = prof_start(PROF_FIBER, moved, stream);
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NTHR), lds, stream,
                     reinterpret_cast<const Cx<T>*>(src), reinterpret_cast<Cx<T>*>(dst), ld_src,
                     ld_dst, n_batch, tstore ? 1 : 0, win.fibers, n_f, win.rows, gr.args(sched), los,
                     d_subs, n_sub, d_layers, n_layers, d_diag, d_strings, n_diag, rows, gd_cap,
                     gm_cap, fz, MULTI ? *mt : MultiTiles{nullptr, 1, 0, 0});
  HIP_CHECK(hipGetLastError());
  prof_stop(pk, stream);
}

template <typename T>
void launch_fiber_multi(int mode, const MultiTiles& mt, long ld, int n_batch,
                        const DevWindow& win, bool all_fibers, const DevGroups& gr,
                        const LoSlots& los, int gd_cap, int gm_cap, const SubPass* d_subs,
                        int n_sub, const LayerPair* d_layers, int n_layers, const double* d_diag,
                        const uint64_t* d_strings, int n_diag, size_t lds_max, int W,
                        hipStream_t stream) {
  require(gr.w2 == kMaxSubWindow && W == (std::is_same<T, double>::value ? 8 : 4),
          FFSIM_ERR_VALUE, "multi-state pass: unsupported shape");
  constexpr int WW = std::is_same<T, double>::value ? 8 : 4;
  const int rows_l = all_fibers ? win.max_rows : win.max_rows_active;
  const size_t lds_l = fiber_lds_bytes(sizeof(Cx<T>), WW, mode == MODE_COLS, 16, rows_l, gd_cap,
                                       gm_cap, d_diag != nullptr, false, n_diag, 0, 0);
  const bool wide = 2 * lds_l > lds_max;
#define FFSIM_MULTI_LAUNCH(MM, NT)                                                           \
  launch_fiber_t<T, WW, MM, kMaxSubWindow, NT, false, true>(                                 \
      nullptr, nullptr, ld, ld, n_batch, false, win, all_fibers, gr, los, gd_cap, gm_cap,      \
      d_subs, n_sub, d_layers, n_layers, d_diag, d_strings, n_diag, lds_max, stream,           \
      DiagFuse{}, &mt)
  if (mode == MODE_ROWS) {
    if (wide) FFSIM_MULTI_LAUNCH(MODE_ROWS, 512); else FFSIM_MULTI_LAUNCH(MODE_ROWS, 256);
  } else {
    if (wide) FFSIM_MULTI_LAUNCH(MODE_COLS, 512); else FFSIM_MULTI_LAUNCH(MODE_COLS, 256);
  }
#undef FFSIM_MULTI_LAUNCH
}

template <typename T>
void launch_fiber(int mode, const void* src, void* dst, long ld_src, long ld_dst, int n_batch,
                  bool tstore, const DevWindow& win, bool all_fibers, const DevGroups& gr,
                  const LoSlots& los, int gd_cap, int gm_cap,
                  const SubPass* d_subs, int n_sub, const LayerPair* d_layers, int n_layers,
                  const double* d_diag, const uint64_t* d_strings, int n_diag, size_t lds_max,
                  int W, hipStream_t stream, const DiagFuse* fz) {
  // gr.w2 < kMaxSubWindow only for windows narrower than the sub-window
#define FFSIM_FIBER_LAUNCH_N(WW, MM, W2, NT)                                                    \
  launch_fiber_t<T, WW, MM, W2, NT>(src, dst, ld_src, ld_dst, n_batch, tstore, win, all_fibers,  \
                                    gr, los, gd_cap, gm_cap, d_subs, n_sub, d_layers, n_layers,  \
                                    d_diag, d_strings, n_diag, lds_max, stream)
  // 512 threads when one block fills the CU's LDS, 256 when two or more fit
  const int rows_l = all_fibers ? win.max_rows : win.max_rows_active;
  const size_t lds_l = fiber_lds_bytes(sizeof(Cx<T>), W, mode == MODE_COLS, 16, rows_l, gd_cap,
                                       gm_cap, d_diag != nullptr, fz != nullptr, n_diag,
                                       fz ? fz->n_fac : 0, fz ? fz->win : 0);
  const bool wide = 2 * lds_l > lds_max;
  if (fz) {
    // the fused diagonal store exists for the production tile widths (8
    // columns for complex128, 4 for complex64) of a transposing row pass
    constexpr int WF = std::is_same<T, double>::value ? 8 : 4;
    require(W == WF && mode == MODE_ROWS && gr.w2 == kMaxSubWindow && tstore, FFSIM_ERR_VALUE,
            "fused diagonal: unsupported pass");
    if (wide)
      launch_fiber_t<T, WF, MODE_ROWS, kMaxSubWindow, 512, true>(
          src, dst, ld_src, ld_dst, n_batch, tstore, win, all_fibers, gr, los, gd_cap, gm_cap,
          d_subs, n_sub, d_layers, n_layers, d_diag, d_strings, n_diag, lds_max, stream, *fz);
    else
      launch_fiber_t<T, WF, MODE_ROWS, kMaxSubWindow, 256, true>(
          src, dst, ld_src, ld_dst, n_batch, tstore, win, all_fibers, gr, los, gd_cap, gm_cap,
          d_subs, n_sub, d_layers, n_layers, d_diag, d_strings, n_diag, lds_max, stream, *fz);
    return;
  }
#define FFSIM_FIBER_LAUNCH(WW, MM, W2)               \
  do {                                               \
    if (wide)                                        \
      FFSIM_FIBER_LAUNCH_N(WW, MM, W2, 512);         \
    else                                             \
      FFSIM_FIBER_LAUNCH_N(WW, MM, W2, 256);         \
  } while (0)
#define FFSIM_FIBER_CASE(WW)                                                                    \
  if (W == WW) {                                                                               \
    if (gr.w2 == kMaxSubWindow) {                                                              \
      if (mode == MODE_ROWS)                                                                   \
        FFSIM_FIBER_LAUNCH(WW, MODE_ROWS, kMaxSubWindow);                                      \
      else                                                                                     \
        FFSIM_FIBER_LAUNCH(WW, MODE_COLS, kMaxSubWindow);                                      \
      return;                                                                                  \
    }                                                                                          \
  }
  // production widths with the full sub-window
  FFSIM_FIBER_CASE(4)
  FFSIM_FIBER_CASE(8)
  FFSIM_FIBER_CASE(16)
  // windows narrower than the sub-window (norb < 6): one width suffices
  if (sizeof(T) == 8) W = 8; else W = 16;
#undef FFSIM_FIBER_LAUNCH
#define FFSIM_FIBER_LAUNCH(WW, MM, W2) FFSIM_FIBER_LAUNCH_N(WW, MM, W2, 256)
  if (W == 8) {
    switch (gr.w2) {
      case 2: if (mode == MODE_ROWS) FFSIM_FIBER_LAUNCH(8, MODE_ROWS, 2); else FFSIM_FIBER_LAUNCH(8, MODE_COLS, 2); return;
      case 3: if (mode == MODE_ROWS) FFSIM_FIBER_LAUNCH(8, MODE_ROWS, 3); else FFSIM_FIBER_LAUNCH(8, MODE_COLS, 3); return;
      case 4: if (mode == MODE_ROWS) FFSIM_FIBER_LAUNCH(8, MODE_ROWS, 4); else FFSIM_FIBER_LAUNCH(8, MODE_COLS, 4); return;
      default: if (mode == MODE_ROWS) FFSIM_FIBER_LAUNCH(8, MODE_ROWS, 5); else FFSIM_FIBER_LAUNCH(8, MODE_COLS, 5); return;
    }
  } else {
    switch (gr.w2) {
      case 2: if (mode == MODE_ROWS) FFSIM_FIBER_LAUNCH(16, MODE_ROWS, 2); else FFSIM_FIBER_LAUNCH(16, MODE_COLS, 2); return;
      case 3: if (mode == MODE_ROWS) FFSIM_FIBER_LAUNCH(16, MODE_ROWS, 3); else FFSIM_FIBER_LAUNCH(16, MODE_COLS, 3); return;
      case 4: if (mode == MODE_ROWS) FFSIM_FIBER_LAUNCH(16, MODE_ROWS, 4); else FFSIM_FIBER_LAUNCH(16, MODE_COLS, 4); return;
      default: if (mode == MODE_ROWS) FFSIM_FIBER_LAUNCH(16, MODE_ROWS, 5); else FFSIM_FIBER_LAUNCH(16, MODE_COLS, 5); return;
    }
  }
#undef FFSIM_FIBER_CASE
#undef FFSIM_FIBER_LAUNCH
#undef FFSIM_FIBER_LAUNCH_N
  throw Failure{FFSIM_ERR_VALUE, "unsupported tile width"};
}

void fiber_pass(int dtype, int mode, const void* src, void* dst, long ld_src, long ld_dst,
                int n_batch, bool tstore, const DevWindow& win, bool all_fibers,
                const DevGroups& gr, const LoSlots& los, int gd_cap, int gm_cap,
                const SubPass* d_subs, int n_sub, const LayerPair* d_layers,
                int n_layers, const double* d_diag, const uint64_t* d_strings, int n_diag,
                Context& ctx, hipStream_t stream, const DiagFuse* fz = nullptr) {
  const int W = tile_width(dtype);
  if (dtype == FFSIM_C128)
    launch_fiber<double>(mode, src, dst, ld_src, ld_dst, n_batch, tstore, win, all_fibers, gr,
                         los, gd_cap, gm_cap, d_subs, n_sub, d_layers, n_layers, d_diag, d_strings, n_diag, ctx.lds_max, W,
                         stream, fz);
  else
    launch_fiber<float>(mode, src, dst, ld_src, ld_dst, n_batch, tstore, win, all_fibers, gr,
                        los, gd_cap, gm_cap, d_subs, n_sub, d_layers, n_layers, d_diag, d_strings, n_diag, ctx.lds_max, W,
                        stream, fz);
}

// One spin sector's orbital rotation, prepared on the host.
struct SectorRotation {
  std::vector<KStep> steps;  // global p
  std::vector<cdouble> diag;
  bool trivial_diag = true;
};

// Pair positions of a generic n x n decomposition in each order (no rotation
// skipped), as the two decompositions emit them.
std::vector<KStep> generic_steps(int n, bool triangle) {
  std::vector<int> ps;
  if (triangle) {
    for (int row = n - 1; row >= 1; --row)
      for (int col = 0; col < row; ++col) ps.push_back(col);
  } else {
    std::vector<int> right, left;
    for (int d = 0; d + 1 < n; ++d)
      for (int t = 0; t <= d; ++t) {
        if (d % 2 == 0)
          right.push_back(d - t);
        else
          left.push_back(n - 2 - d + t);
      }
    ps = right;
    ps.insert(ps.end(), left.rbegin(), left.rend());
  }
  std::vector<KStep> steps;
  for (int p : ps) {
    KStep k{};
    k.p = p;
    k.c = 1.0;
    steps.push_back(k);
  }
  return steps;
}

// Whether the passes of a sector rotation at window width w run the triangle
// order: when it needs fewer window passes than ffsim's (Clements) order, or
// as many with fewer layer pairs.  At 12-orbital windows that is norb 15, 17,
// 18 and 20; at norb 18 it makes 30 fiber launches per LUCJ n_reps=4
// application instead of 40 (measured 17% faster), while at norb 16 the
// Clements order's fewer layer pairs measured 2% faster for complex64.
// FFSIM_AMD_GIVENS_ORDER=clements or =triangle forces one order.
bool triangle_order(int norb, int w) {
  const char* e = std::getenv("FFSIM_AMD_GIVENS_ORDER");
  if (e && std::strcmp(e, "clements") == 0) return false;
  if (e && std::strcmp(e, "triangle") == 0) return true;
  static std::mutex mu;
  static std::map<std::pair<int, int>, bool> cache;
  const std::pair<int, int> key(norb, w);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  long cost[2][3];
  for (int t = 0; t < 2; ++t) {
    std::vector<PlannedPass> passes = plan_passes(generic_steps(norb, t == 1), norb, w, true, -1);
    long lp = 0, sub = 0;
    for (const PlannedPass& P : passes) {
      std::vector<SubPass> subs;
      std::vector<LayerPair> layers;
      plan_sub_passes(P.steps, w, kMaxSubWindow, subs, layers);
      lp += (long)layers.size();
      sub += (long)subs.size();
    }
    cost[t][0] = (long)passes.size();
    cost[t][1] = lp;
    cost[t][2] = sub;
  }
  const bool tri = std::lexicographical_compare(cost[1], cost[1] + 3, cost[0], cost[0] + 3);
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = tri;
  return tri;
}

// w: the window width of the sector's passes (norb for a column-mode pass).
SectorRotation prepare_rotation(const cdouble* mat, int norb, int w) {
  SectorRotation r;
  GivensDecomposition g = triangle_order(norb, w) ? triangle_decomposition(mat, norb, 1e-12)
                                                  : givens_decomposition(mat, norb, 1e-12);
  for (const auto& rot : g.rotations) r.steps.push_back(kernel_step_from_rotation(rot));
  r.diag = g.diag;
  for (const auto& d : g.diag)
    if (d != cdouble(1.0, 0.0)) r.trivial_diag = false;
  return r;
}

// Runs the fused passes of one sector rotation.  The state is `src` with
// strings of this sector on rows (MODE_ROWS, ld = n_batch) or on columns
// (MODE_COLS).  If tstore_dst is non-null the final pass writes the
// transposed state there (MODE_ROWS only).
// Diagonal operator helper.  rows/cols strings are for the state's row and
// column sectors; matrices are given for (row sector, col sector, cross).
// Appends a diagonal operator's header and coefficients to a call's program;
// returns their offsets.
std::pair<size_t, size_t> append_diag_params(Program& prog, const DiagSpec& spec) {
  const DiagHeader h = diag_header(spec);
  const size_t hoff = prog.append(&h, sizeof(h));
  const std::vector<double> coef = diag_coefficients(spec);
  const size_t coff = prog.append(coef.data(), coef.size() * sizeof(double));
  return {hoff, coff};
}

// Factor tables of a diagonal Coulomb evolution to fuse into the final
// transposing pass (offsets in the call's program; see DiagFuse).
struct FuseTables {
  size_t bfac_off = 0, rfac_off = 0;
  int n_fac = 0;
};

// The sub-window slots of a pass and the largest group tables that any fiber
// of the pass stages (over the window's electron counts kp).
void pass_slots(const DevGroups& gr, const std::vector<SubPass>& subs, LoSlots& los, int& gd_cap,
                int& gm_cap) {
  los = LoSlots{};
  for (const SubPass& sp : subs) {
    require(sp.slot < kMaxLoSlots, FFSIM_ERR_VALUE, "too many sub-window starts");
    los.lo2[sp.slot] = sp.lo2;
    los.n = std::max(los.n, sp.slot + 1);
  }
  gd_cap = 0;
  gm_cap = 0;
  for (int kp = 0; kp <= gr.w; ++kp) {
    int nd = 0, nm = 0;
    for (int u = 0; u < los.n; ++u) {
      const int* ix = gr.host_index.data() + ((size_t)kp * gr.n_lo2 + los.lo2[u]) * 4;
      nd += ix[1];
      nm += ix[3];
    }
    gd_cap = std::max(gd_cap, nd);
    gm_cap = std::max(gm_cap, nm);
  }
}

void run_sector(Context& ctx, Program& prog, int dtype, int mode, void* state, long ld,
                int n_batch, int norb, int nocc, const SectorRotation& rot, void* tstore_dst,
                long ld_t, hipStream_t stream, std::vector<std::function<void()>>& launches,
                const FuseTables* fuse = nullptr, bool* fused = nullptr) {
  if (fused) *fused = false;
  const int w = (mode == MODE_COLS) ? norb : choose_window(ctx, norb, nocc, dtype);
  const bool tstore = tstore_dst != nullptr;
  std::vector<PlannedPass> passes =
      plan_passes(rot.steps, norb, std::max(w, 1), tstore, -1);
  const DevStrings& ds = dev_strings(ctx, norb, nocc);
  const DevGroups& gr = dev_groups(ctx, std::max(w, 1), dtype == FFSIM_C64);
  const bool need_diag = !rot.trivial_diag;
  require(!need_diag || norb <= kMaxDiagOrbitals, FFSIM_ERR_VALUE, "norb above 64");
  size_t diag_off = 0;
  if (need_diag) {
    std::vector<double> dd(2 * norb);
    for (int k = 0; k < norb; ++k) {
      dd[2 * k] = rot.diag[k].real();
      dd[2 * k + 1] = rot.diag[k].imag();
    }
    diag_off = prog.append(dd.data(), dd.size() * sizeof(double));
  }
  std::vector<SubPass> subs;
  std::vector<LayerPair> sub_layers;
  for (const PlannedPass& P : passes) {
    const bool last = P.last;
    const bool ts = last && tstore;
    const bool dg = last && need_diag;
    if (P.steps.empty() && !ts && !dg) continue;
    plan_sub_passes(P.steps, std::max(w, 1), kMaxSubWindow, subs, sub_layers);
    const size_t soff = prog.append(sub_layers.data(), sub_layers.size() * sizeof(LayerPair));
    const size_t poff = prog.append(subs.data(), subs.size() * sizeof(SubPass));
    const int n_layers = (int)sub_layers.size();
    const int n_sub = (int)subs.size();
    LoSlots los{};
    int gd_cap = 0, gm_cap = 0;
    pass_slots(gr, subs, los, gd_cap, gm_cap);
    const DevWindow& win = dev_window(ctx, norb, nocc, P.lo, std::max(w, 1));
    void* dst = ts ? tstore_dst : state;
    const long ld_dst = ts ? ld_t : ld;
    const bool all = ts || dg;
    // fuse the diagonal into this pass when the kernel's fused layout applies
    bool fz_on = false;
    // (the fused store takes the pass's window to be [0, w), w <= 16)
    const int wf = dtype == FFSIM_C128 ? 8 : 4;
    if (ts && fuse && mode == MODE_ROWS && tile_width(dtype) == wf && gr.w2 == kMaxSubWindow &&
        n_sub > 0 && P.lo == 0 && w <= 16) {
      const int rows = win.max_rows;
      const size_t lds = fiber_lds_bytes(elem_bytes(dtype), wf, false, 16, rows, gd_cap, gm_cap,
                                         dg, true, norb, fuse->n_fac, w);
      const int nthr = 2 * lds > ctx.lds_max ? 512 : 256;
      fz_on = lds <= ctx.lds_max && rows <= 1024 && wf * fuse->n_fac <= nthr;
    }
    if (fz_on && fused) *fused = true;
    const FuseTables ft = fz_on ? *fuse : FuseTables{};
    launches.push_back([=, &ctx, &prog, &win, &gr, &ds]() {
      DiagFuse dfz{};
      if (fz_on)
        dfz = DiagFuse{prog.scratch<Cx<double>>(ft.bfac_off), prog.scratch<Cx<double>>(ft.rfac_off),
                       ft.n_fac, w};
      fiber_pass(dtype, mode, state, dst, ld, ld_dst, n_batch, ts, win, all, gr, los, gd_cap,
                 gm_cap, prog.dev<SubPass>(poff), n_sub, prog.dev<LayerPair>(soff), n_layers,
                 dg ? prog.dev<double>(diag_off) : nullptr, ds.d, norb, ctx, stream,
                 fz_on ? &dfz : nullptr);
    });
  }
}

// Row pitch (elements) of the transposed state in the work buffer: dim_a
// rounded up to 32 bytes, so that every row's tile segments start on the
// 32-byte boundaries of the memory requests.  (At norb=16 a complex64 row of
// 12870 elements is 16 bytes off on every other row, and an unaligned 32-byte
// segment costs two requests.)
long transposed_ld(int dtype, long dim_a) {
  const long q = 32 / (long)elem_bytes(dtype);
  return (dim_a + q - 1) / q * q;
}

// Full spinful orbital rotation (both sectors) on a dim_a x dim_b state.
// With `then_diag`, the diagonal Coulomb evolution that follows the rotation is
// fused into the final pass when that pass is a transposing c128 row pass and
// everything fits; returns true when it was (the caller applies it otherwise).
// ld_w: row pitch of the transposed state in `work` (0: transposed_ld, the
// layout ffsim_orbital_rotation_work_elements sizes the buffer for).
bool orbital_rotation_impl(Context& ctx, void* vec, void* work, int dtype, int norb, int na,
                           int nb, const cdouble* mat_a, const cdouble* mat_b,
                           hipStream_t stream, const DiagSpec* then_diag = nullptr,
                           long ld_w = 0) {
  const long dim_a = (long)binom(norb, na), dim_b = (long)binom(norb, nb);
  if (dim_a == 0 || dim_b == 0 || norb == 0) return false;
  require(dim_a < (1l << 31) && dim_b < (1l << 31), FFSIM_ERR_VALUE, "dimension too large");
  const bool beta_cols = beta_in_columns(norb, dim_b, dtype, ctx);
  const int w_a = choose_window(ctx, norb, na, dtype);
  const int w_b = beta_cols ? norb : choose_window(ctx, norb, nb, dtype);
  if (ld_w == 0) ld_w = transposed_ld(dtype, dim_a);
  require(ld_w >= dim_a, FFSIM_ERR_VALUE, "work row pitch below dim_a");
  SectorRotation ra, rb;
  if (mat_a) ra = prepare_rotation(mat_a, norb, w_a);
  if (mat_b) {
    if (mat_a && std::memcmp(mat_a, mat_b, sizeof(cdouble) * norb * norb) == 0 &&
        triangle_order(norb, w_a) == triangle_order(norb, w_b))
      rb = ra;
    else
      rb = prepare_rotation(mat_b, norb, w_b);
  }
  Program prog(ctx, stream);
  std::vector<std::function<void()>> launches;
  const bool beta_transposed = mat_b && !beta_cols;
  if (beta_transposed) require(work != nullptr, FFSIM_ERR_VALUE, "work buffer required");
  // factor tables of the fused diagonal: alpha strings are the final pass's
  // batch columns, beta strings its tile rows
  FuseTables ft;
  bool want_fuse = then_diag && then_diag->flags && !then_diag->z_rep &&
                   then_diag->norb == norb && beta_transposed &&
                   norb + 1 <= 64;
  if (want_fuse) {
    ft.n_fac = norb + 1;
    want_fuse = prog.reserve((size_t)dim_a * ft.n_fac * sizeof(Cx<double>), &ft.bfac_off) &&
                prog.reserve((size_t)dim_b * sizeof(Cx<double>), &ft.rfac_off);
  }
  if (want_fuse) {
    const std::pair<size_t, size_t> par = append_diag_params(prog, *then_diag);
    const size_t hoff = par.first, coff = par.second;
    const DevStrings& sa = dev_strings(ctx, norb, na);
    const DevStrings& sb = dev_strings(ctx, norb, nb);
    const int nf = ft.n_fac;
    launches.push_back([=, &prog, &sa, &sb]() {
      const long n_a = dim_a * nf;
      hipLaunchKernelGGL(diag_row_factor_kernel, dim3((unsigned)((n_a + 255) / 256)), dim3(256),
                         0, stream, (int)dim_a, sa.d, prog.dev<DiagHeader>(hoff),
                         prog.dev<double>(coff), prog.scratch<Cx<double>>(ft.bfac_off));
      HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(diag_col_term_kernel, dim3((unsigned)((dim_b + 255) / 256)), dim3(256),
                         0, stream, (int)dim_b, sb.d, prog.dev<DiagHeader>(hoff),
                         prog.dev<double>(coff), prog.scratch<double>(ft.rfac_off), 1, 0l, 0l);
      HIP_CHECK(hipGetLastError());
    });
  }
  bool fused = false;
  if (mat_a || beta_transposed) {
    SectorRotation empty;
    const SectorRotation& r = mat_a ? ra : empty;
    run_sector(ctx, prog, dtype, MODE_ROWS, vec, dim_b, (int)dim_b, norb, na, r,
               beta_transposed ? work : nullptr, ld_w, stream, launches);
  }
  if (mat_b) {
    if (beta_cols) {
      run_sector(ctx, prog, dtype, MODE_COLS, vec, dim_b, (int)dim_a, norb, nb, rb, nullptr, 0,
                 stream, launches);
    } else {
      run_sector(ctx, prog, dtype, MODE_ROWS, work, ld_w, (int)dim_a, norb, nb, rb, vec, dim_b,
                 stream, launches, want_fuse ? &ft : nullptr, &fused);
    }
  }
  prog.commit();
  for (auto& f : launches) f();
  prog.finish();
  return fused;
}


void diag_impl(Context& ctx, int op, const void* src, void* dst, int dtype, int norb, int nrow_occ,
               int ncol_occ, const DiagSpec& spec, hipStream_t stream) {
  const long n_rows = (long)binom(norb, nrow_occ), n_cols = (long)binom(norb, ncol_occ);
  if (n_rows == 0 || n_cols == 0) return;
  require(norb <= 63, FFSIM_ERR_VALUE, "norb > 63 unsupported");
  const DevStrings& rs = dev_strings(ctx, norb, nrow_occ);
  const DevStrings& cs = dev_strings(ctx, norb, ncol_occ);
  Program prog(ctx, stream);
  require(spec.norb == norb, FFSIM_ERR_VALUE, "diagonal operator of another norb");
  const auto [hoff, coff] = append_diag_params(prog, spec);
  // per column: gamma (contractions) or its phase factor (cos, sin) (evolution)
  const int evolve = op == DIAG_EVOLVE ? 1 : 0;
  size_t goff = 0;
  require(prog.reserve((size_t)n_cols * (evolve ? 2 : 1) * sizeof(double), &goff),
          FFSIM_ERR_VALUE, "diagonal operator: too many columns");
  prog.commit();
  const DiagHeader* dh = prog.dev<DiagHeader>(hoff);
  const double* dc = prog.dev<double>(coff);
  double* dg = prog.scratch<double>(goff);
  require(diag_grid(n_rows, n_cols).y <= 65535, FFSIM_ERR_VALUE, "too many columns");
  launch_diag_col_term(n_cols, cs.d, dh, dc, dg, evolve, stream);
  HIP_CHECK(hipGetLastError());
  const double moved = (double)elem_bytes(dtype) * (double)n_rows * (double)n_cols *
                       (op == DIAG_CONTRACT_ACC ? 3.0 : 2.0);
  const long pk = prof_start(PROF_DIAG, moved, stream);
  if (dtype == FFSIM_C128)
    launch_diag_pass_t<double>(op, src, dst, n_rows, n_cols, rs.d, cs.d, dg, dh, dc, stream);
  else
    launch_diag_pass_t<float>(op, src, dst, n_rows, n_cols, rs.d, cs.d, dg, dh, dc, stream);
  HIP_CHECK(hipGetLastError());
  prof_stop(pk, stream);
  prog.finish();
}

// Diagonal operators applied to several states of one shape in place (op:
// DIAG_EVOLVE or DIAG_CONTRACT): specs holds one operator per state, or one
// for all of them.  The column terms of each distinct operator, then one
// elementwise launch with grid.z over the states; each state goes through
// the same arithmetic as diag_impl's.
void diag_batched_impl(Context& ctx, int op, int n_states, void* const* vecs, int dtype,
                       int norb, int nrow_occ, int ncol_occ, const std::vector<DiagSpec>& specs,
                       hipStream_t stream) {
  const long n_rows = (long)binom(norb, nrow_occ), n_cols = (long)binom(norb, ncol_occ);
  if (n_rows == 0 || n_cols == 0 || n_states == 0) return;
  const bool shared = specs.size() == 1;
  require(shared || (int)specs.size() == n_states, FFSIM_ERR_VALUE,
          "batched diagonal operator: one operator per state");
  require(op == DIAG_EVOLVE || op == DIAG_CONTRACT, FFSIM_ERR_VALUE, "unsupported operation");
  require(norb <= 63, FFSIM_ERR_VALUE, "norb > 63 unsupported");
  for (const DiagSpec& sp : specs)
    require(sp.norb == norb, FFSIM_ERR_VALUE, "diagonal operator of another norb");
  require(diag_grid(n_rows, n_cols).y <= 65535, FFSIM_ERR_VALUE, "too many columns");
  const DevStrings& rs = dev_strings(ctx, norb, nrow_occ);
  const DevStrings& cs = dev_strings(ctx, norb, ncol_occ);
  const int evolve = op == DIAG_EVOLVE ? 1 : 0;
  const long gamma_len = n_cols * (evolve ? 2 : 1);
  constexpr int kChunk = 256;
  for (int i0 = 0; i0 < n_states; i0 += kChunk) {
    const int m = std::min(kChunk, n_states - i0);
    const int n_ops = shared ? 1 : m;
    Program prog(ctx, stream);
    // one header per state (the kernel indexes them by state), coefficients
    // and column terms per distinct operator
    std::vector<DiagHeader> hs(m);
    std::vector<double> coef;
    for (int i = 0; i < m; ++i) {
      const DiagSpec& sp = specs[shared ? 0 : i0 + i];
      hs[i] = diag_header(sp);
      if (i < n_ops) {
        const std::vector<double> c = diag_coefficients(sp);
        coef.insert(coef.end(), c.begin(), c.end());
      }
    }
    const long coef_len = (long)coef.size() / n_ops;
    const size_t hoff = prog.append(hs.data(), hs.size() * sizeof(DiagHeader));
    const size_t coff = prog.append(coef.data(), coef.size() * sizeof(double));
    std::vector<const void*> ptrs(vecs + i0, vecs + i0 + m);
    const size_t soff = prog.append(ptrs.data(), m * sizeof(void*));
    size_t goff = 0;
    require(prog.reserve((size_t)n_ops * gamma_len * sizeof(double), &goff), FFSIM_ERR_VALUE,
            "diagonal operator: too many columns");
    prog.commit();
    const DiagHeader* dh = prog.dev<DiagHeader>(hoff);
    const double* dc = prog.dev<double>(coff);
    double* dg = prog.scratch<double>(goff);
    const long cstride = shared ? 0 : coef_len, gstride = shared ? 0 : gamma_len;
    launch_diag_col_term(n_cols, cs.d, dh, dc, dg, evolve, stream, n_ops, cstride, gstride);
    HIP_CHECK(hipGetLastError());
    const double moved = 2.0 * (double)elem_bytes(dtype) * (double)n_rows * (double)n_cols * m;
    const long pk = prof_start(PROF_DIAG, moved, stream);
    if (dtype == FFSIM_C128)
      launch_diag_pass_t<double>(op, nullptr, nullptr, n_rows, n_cols, rs.d, cs.d, dg, dh, dc,
                                 stream, prog.dev<void*>(soff), m, cstride, gstride);
    else
      launch_diag_pass_t<float>(op, nullptr, nullptr, n_rows, n_cols, rs.d, cs.d, dg, dh, dc,
                                stream, prog.dev<void*>(soff), m, cstride, gstride);
    HIP_CHECK(hipGetLastError());
    prof_stop(pk, stream);
    prog.finish();
  }
}

std::vector<double> vec_or_empty(const double* m, int n) {
  if (!m) return {};
  return std::vector<double>(m, m + n);
}

void check_dtype(int dtype) {
  require(dtype == FFSIM_C128 || dtype == FFSIM_C64, FFSIM_ERR_VALUE, "unknown dtype");
}
void check_sizes(int norb, int na, int nb) {
  require(norb >= 0 && norb <= 63, FFSIM_ERR_VALUE, "norb must be in [0, 63]");
  require(na >= 0 && nb >= 0, FFSIM_ERR_VALUE, "electron counts must be non-negative");
}

// Sequence of gates  V_{n}^... : for t = 0..n_terms-1: rotate by rot[t], then
// diag-Coulomb evolution by mats[t] for time times[t]; finally rotate by
// rot[n_terms].  rot[t] == nullptr skips that rotation.
void diag_coulomb_sequence_impl(Context& ctx, void* vec, void* work, int dtype, int norb, int na,
                                int nb, int n_terms, const cdouble* const* rots,
                                const double* const* mats /* aa, ab, bb per term (3 ptrs) */,
                                const double* times, int z_rep, hipStream_t st, long ld_w = 0) {
  for (int t = 0; t < n_terms; ++t) {
    DiagSpec s = diag_coulomb_spec(norb, mats[3 * t], mats[3 * t + 1], mats[3 * t + 2], z_rep,
                                   times[t]);
    bool fused = false;
    if (rots[t])
      fused = orbital_rotation_impl(ctx, vec, work, dtype, norb, na, nb, rots[t], rots[t], st,
                                    s.flags ? &s : nullptr, ld_w);
    if (s.flags && !fused) diag_impl(ctx, DIAG_EVOLVE, vec, vec, dtype, norb, na, nb, s, st);
  }
  if (rots[n_terms])
    orbital_rotation_impl(ctx, vec, work, dtype, norb, na, nb, rots[n_terms], rots[n_terms], st,
                          nullptr, ld_w);
}

std::vector<cdouble> matmul(const cdouble* A, const cdouble* B, int n, bool adj_a) {
  std::vector<cdouble> C((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const cdouble a = adj_a ? std::conj(A[(size_t)k * n + i]) : A[(size_t)i * n + k];
      for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)k * n + j];
    }
  return C;
}


// Launches a set of row-pair (or column-pair) rotations; every rotation of the
// set must touch disjoint rows (columns).
void launch_pair_rotations(int dtype, void* vec, long n_rows, long n_cols, bool on_rows,
                           const PairRotationSet& set, int n_rot, hipStream_t stream) {
  long max_pairs = 0;
  for (int r = 0; r < n_rot; ++r) max_pairs = std::max(max_pairs, set.r[r].n_pairs);
  if (n_rot == 0 || max_pairs == 0 || n_rows == 0 || n_cols == 0) return;
  if (on_rows) {
    const dim3 grid((unsigned)std::min<long>(max_pairs, 65535), (unsigned)n_rot);
    if (dtype == FFSIM_C128)
      hipLaunchKernelGGL(givens_rows_kernel<double>, grid, dim3(256), 0, stream,
                         (Cx<double>*)vec, n_cols, set);
    else
      hipLaunchKernelGGL(givens_rows_kernel<float>, grid, dim3(256), 0, stream,
                         (Cx<float>*)vec, n_cols, set);
  } else {
    const dim3 grid((unsigned)((max_pairs + 255) / 256), (unsigned)std::min<long>(n_rows, 1024),
                    (unsigned)n_rot);
    require(grid.x <= 0x7fffffffu, FFSIM_ERR_VALUE, "too many pairs");
    if (dtype == FFSIM_C128)
      hipLaunchKernelGGL(givens_cols_kernel<double>, grid, dim3(256), 0, stream,
                         (Cx<double>*)vec, n_rows, n_cols, set);
    else
      hipLaunchKernelGGL(givens_cols_kernel<float>, grid, dim3(256), 0, stream,
                         (Cx<float>*)vec, n_rows, n_cols, set);
  }
  HIP_CHECK(hipGetLastError());
}

// Copies a host index list (ffsim passes usize arrays) into the call's
// program as int64, after checking every entry against the dimension.
size_t append_indices(Program& prog, const uint64_t* idx, int64_t n, int64_t bound,
                      const char* what) {
  std::vector<int64_t> v((size_t)n);
  for (int64_t k = 0; k < n; ++k) {
    require(idx[k] < (uint64_t)bound, FFSIM_ERR_VALUE,
            std::string(what) + ": index " + std::to_string(idx[k]) + " out of bounds for " +
                std::to_string(bound) + " rows");
    v[(size_t)k] = (int64_t)idx[k];
  }
  return prog.append(v.data(), v.size() * sizeof(int64_t));
}

// The merged rotations of a spin-balanced UCJ operator
// (python/ffsim/variational/ucj_spin_balanced.py:657-696): V_0 = U_0^H,
// V_r = U_r^H U_{r-1}, V_{n_reps} = F U_last (or U_last).
std::vector<std::vector<cdouble>> ucj_merged_rotations(int n, int n_reps, const cdouble* U,
                                                       const cdouble* final_rot) {
  const size_t nn = (size_t)n * n;
  std::vector<std::vector<cdouble>> V(n_reps + 1);
  std::vector<cdouble> eye(nn, 0.0);
  for (int k = 0; k < n; ++k) eye[(size_t)k * n + k] = 1.0;
  for (int r = 0; r < n_reps; ++r)
    V[r] = matmul(U + r * nn, r == 0 ? eye.data() : U + (r - 1) * nn, n, true);
  const cdouble* last = n_reps ? U + (n_reps - 1) * nn : eye.data();
  if (final_rot)
    V[n_reps] = matmul(final_rot, last, n, false);
  else
    V[n_reps] = std::vector<cdouble>(last, last + nn);
  return V;
}

// ld_w: row pitch of the transposed state in `work` (0: transposed_ld).
void ucj_impl(Context& ctx, void* vec, void* work, int dtype, int n, int na, int nb, int n_reps,
              const double* diag_coulomb_mats, const double* orbital_rotations,
              const double* final_orbital_rotation, hipStream_t stream, long ld_w = 0) {
  const size_t nn = (size_t)n * n;
  std::vector<std::vector<cdouble>> V = ucj_merged_rotations(
      n, n_reps, reinterpret_cast<const cdouble*>(orbital_rotations),
      reinterpret_cast<const cdouble*>(final_orbital_rotation));
  std::vector<const cdouble*> rots(n_reps + 1);
  for (int r = 0; r <= n_reps; ++r) rots[r] = V[r].data();
  std::vector<std::vector<double>> halved(n_reps);
  std::vector<const double*> mats(3 * n_reps);
  std::vector<double> times(n_reps, -1.0);
  for (int r = 0; r < n_reps; ++r) {
    const double* J = diag_coulomb_mats + r * 2 * nn;
    halved[r].assign(J, J + nn);
    for (int k = 0; k < n; ++k) halved[r][(size_t)k * n + k] *= 0.5;
    mats[3 * r] = halved[r].data();
    mats[3 * r + 1] = J + nn;
    mats[3 * r + 2] = halved[r].data();
  }
  diag_coulomb_sequence_impl(ctx, vec, work, dtype, n, na, nb, n_reps, rots.data(), mats.data(),
                             times.data(), 0, stream, ld_w);
}

// Where the GEMM path is taken by default.  An orbital rotation there costs
// 8 dim^3 flops per state and sector against one HBM pass of the state on the
// Givens path, and the path's fixed cost (a compound-matrix launch and two
// batched GEMMs per rotation) is shared by the batch, as the Givens path's
// launches are when all states share them.  Measured per state (LUCJ
// n_reps=2, complex128, MI355X; GEMM vs Givens with all states per launch):
// norb=8 (70 strings) 0.26 vs 0.54 ms alone, 0.051 vs 0.067 ms in a batch of
// 8 and 0.017 vs 0.024 in a batch of 32; norb=9 (126 strings) 0.052 vs 0.061
// and 0.021 vs 0.029; norb=10 (4,4) (210 strings) 0.067 vs 0.047 and 0.053 vs
// 0.033; norb=10 (5,5) (252 strings) 0.077 vs 0.051 and 0.069 vs 0.033;
// norb=12 (924 strings) 2.6 vs 0.16 ms in a batch of 32.
constexpr long kGemmAlwaysDim = 100;   // any batch
constexpr long kGemmBatchDim = 150;    // batches of at least kGemmMinBatch states
constexpr int kGemmMinBatch = 8;

// FFSIM_AMD_BATCH_PATH=gemm | sweep | states forces a path (read at every
// call, for measurements that compare them): the compound-matrix GEMMs, the
// Givens passes over all states at once (where they apply), or the Givens
// passes state by state.
int batch_path_override() {
  const char* e = std::getenv("FFSIM_AMD_BATCH_PATH");
  if (!e) return 0;
  if (std::string(e) == "gemm") return 1;
  if (std::string(e) == "sweep") return 2;
  if (std::string(e) == "states") return 3;
  return 0;
}

bool gemm_path(int norb, int na, int nb, int n_states) {
  const long da = (long)binom(norb, na), db = (long)binom(norb, nb);
  const bool ok = na <= kMaxCompoundK && nb <= kMaxCompoundK && da <= 1024 && db <= 1024;
  if (batch_path_override() == 1) return ok;
  if (batch_path_override() >= 2) return false;
  const long d = std::max(da, db);
  return ok && (d <= kGemmAlwaysDim || (d <= kGemmBatchDim && n_states >= kGemmMinBatch));
}

template <typename T>
void launch_compound(const Cx<double>* U, int norb, int k, const uint64_t* strings, long dim,
                     Cx<T>* out, int n_mats, hipStream_t stream) {
  const dim3 grid((unsigned)((dim * dim + 255) / 256), (unsigned)n_mats);
  switch (k) {
#define FFSIM_COMPOUND(KK) \
  case KK:                \
    hipLaunchKernelGGL((compound_kernel<T, KK>), grid, dim3(256), 0, stream, U, norb, strings, dim, out); \
    break;
    FFSIM_COMPOUND(0)
    FFSIM_COMPOUND(1)
    FFSIM_COMPOUND(2)
    FFSIM_COMPOUND(3)
    FFSIM_COMPOUND(4)
    FFSIM_COMPOUND(5)
    FFSIM_COMPOUND(6)
#undef FFSIM_COMPOUND
    default:
      throw Failure{FFSIM_ERR_VALUE, "compound matrix: too many electrons in a sector"};
  }
  HIP_CHECK(hipGetLastError());
}

void blas_check(rocblas_status st, const char* what) {
  require(st == rocblas_status_success, FFSIM_ERR_HIP,
          std::string(what) + ": rocBLAS status " + rocblas_status_to_string(st));
}

// One complex GEMM per state, C_i = op(A_i) B_i, column-major, no conjugation.
template <typename T>
void gemm_batched(rocblas_handle h, bool trans_a, long m, long n, long k, const void* const* A,
                  long lda, const void* const* B, long ldb, void* const* C, long ldc, int count) {
  const rocblas_operation ta = trans_a ? rocblas_operation_transpose : rocblas_operation_none;
  if constexpr (std::is_same<T, double>::value) {
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    blas_check(rocblas_zgemm_batched(h, ta, rocblas_operation_none, (rocblas_int)m,
                                     (rocblas_int)n, (rocblas_int)k, &one,
                                     reinterpret_cast<const rocblas_double_complex* const*>(A),
                                     (rocblas_int)lda,
                                     reinterpret_cast<const rocblas_double_complex* const*>(B),
                                     (rocblas_int)ldb, &zero,
                                     reinterpret_cast<rocblas_double_complex* const*>(C),
                                     (rocblas_int)ldc, count),
               "zgemm_batched");
  } else {
    const rocblas_float_complex one{1.0f, 0.0f}, zero{0.0f, 0.0f};
    blas_check(rocblas_cgemm_batched(h, ta, rocblas_operation_none, (rocblas_int)m,
                                     (rocblas_int)n, (rocblas_int)k, &one,
                                     reinterpret_cast<const rocblas_float_complex* const*>(A),
                                     (rocblas_int)lda,
                                     reinterpret_cast<const rocblas_float_complex* const*>(B),
                                     (rocblas_int)ldb, &zero,
                                     reinterpret_cast<rocblas_float_complex* const*>(C),
                                     (rocblas_int)ldc, count),
               "cgemm_batched");
  }
}

// A batch of spin-balanced UCJ applications of one shape on the GEMM path:
// per merged rotation, the compound matrices of every state (one launch per
// sector), then psi <- L_a psi (into work) and psi <- work L_b^T as batched
// GEMMs; per diagonal term one batched elementwise pass.
template <typename T>
void ucj_batched_gemm(Context& ctx, int n_states, void* const* vecs, void* const* works, int n,
                      int na, int nb, int n_reps, const double* diag_coulomb_mats,
                      const double* orbital_rotations, const double* final_rots,
                      hipStream_t stream) {
  const long da = (long)binom(n, na), db = (long)binom(n, nb);
  const size_t nn = (size_t)n * n;
  const int n_steps = n_reps + 1;
  const DevStrings& sa = dev_strings(ctx, n, na);
  const DevStrings& sb = dev_strings(ctx, n, nb);
  // states per program: compound matrices of every step of a chunk fit the slot's scratch
  const size_t per_state = (size_t)n_steps * (size_t)(da * da + db * db) * sizeof(Cx<T>);
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>(n_states, (96u << 20) / per_state));
  for (int i0 = 0; i0 < n_states; i0 += chunk) {
    const int m = std::min(chunk, n_states - i0);
    Program prog(ctx, stream);
    // merged rotations, step-major: U[(step * m + i) * nn]
    std::vector<cdouble> Us((size_t)n_steps * m * nn);
    std::vector<double> J((size_t)std::max(n_reps, 1) * m * 2 * nn);
    for (int i = 0; i < m; ++i) {
      const int st = i0 + i;
      std::vector<std::vector<cdouble>> V = ucj_merged_rotations(
          n, n_reps, reinterpret_cast<const cdouble*>(orbital_rotations) + (size_t)st * n_reps * nn,
          final_rots ? reinterpret_cast<const cdouble*>(final_rots) + (size_t)st * nn : nullptr);
      for (int r = 0; r < n_steps; ++r)
        std::copy(V[r].begin(), V[r].end(), Us.begin() + ((size_t)r * m + i) * nn);
      for (int r = 0; r < n_reps; ++r) {
        const double* src = diag_coulomb_mats + ((size_t)st * n_reps + r) * 2 * nn;
        double* dst = J.data() + ((size_t)r * m + i) * 2 * nn;
        std::copy(src, src + 2 * nn, dst);
        for (int k = 0; k < n; ++k) dst[(size_t)k * n + k] *= 0.5;  // same-spin: diagonal halved
      }
    }
    const size_t u_off = prog.append(Us.data(), Us.size() * sizeof(cdouble));
    const size_t j_off = prog.append(J.data(), J.size() * sizeof(double));
    size_t la_off = 0, lb_off = 0;
    require(prog.reserve((size_t)n_steps * m * da * da * sizeof(Cx<T>), &la_off) &&
                prog.reserve((size_t)n_steps * m * db * db * sizeof(Cx<T>), &lb_off),
            FFSIM_ERR_VALUE, "batched UCJ: compound matrices exceed the program scratch");
    // pointer arrays: per step, alpha (A: vec, B: L_a, C: work), beta (A: L_b, B: work,
    // C: vec), then the states for the diagonal passes; device addresses of the
    // compound matrices are known after commit, so these are filled below
    const size_t ptr_count = (size_t)n_steps * 6 * m + m;
    const size_t p_off = prog.append(nullptr, ptr_count * sizeof(void*));
    prog.prepare();
    std::vector<void*> ptrs(ptr_count);
    Cx<T>* La = prog.scratch<Cx<T>>(la_off);
    Cx<T>* Lb = prog.scratch<Cx<T>>(lb_off);
    for (int r = 0; r < n_steps; ++r)
      for (int i = 0; i < m; ++i) {
        void** q = ptrs.data() + (size_t)r * 6 * m;
        q[i] = vecs[i0 + i];
        q[m + i] = La + ((size_t)r * m + i) * da * da;
        q[2 * m + i] = works[i0 + i];
        q[3 * m + i] = Lb + ((size_t)r * m + i) * db * db;
        q[4 * m + i] = works[i0 + i];
        q[5 * m + i] = vecs[i0 + i];
      }
    for (int i = 0; i < m; ++i) ptrs[(size_t)n_steps * 6 * m + i] = vecs[i0 + i];
    prog.write(p_off, ptrs.data(), ptr_count * sizeof(void*));
    prog.commit();
    const Cx<double>* dU = prog.dev<Cx<double>>(u_off);
    launch_compound<T>(dU, n, na, sa.d, da, La, n_steps * m, stream);
    // the beta sector's compound matrices follow those of alpha in the same buffer order
    launch_compound<T>(dU, n, nb, sb.d, db, Lb, n_steps * m, stream);
    void** dptr = prog.dev<void*>(p_off);
    std::lock_guard<std::mutex> lk(ctx.blas_mu);
    if (!ctx.blas) blas_check(rocblas_create_handle(&ctx.blas), "create_handle");
    blas_check(rocblas_set_stream(ctx.blas, stream), "set_stream");
    for (int r = 0; r < n_steps; ++r) {
      void** q = dptr + (size_t)r * 6 * m;
      // alpha: X' = X L_a (column-major views of the row-major state and matrix)
      gemm_batched<T>(ctx.blas, false, db, da, da, (const void* const*)q, db,
                      (const void* const*)(q + m), da, (void* const*)(q + 2 * m), db, m);
      // beta: X' = L_b^T X
      gemm_batched<T>(ctx.blas, true, db, da, db, (const void* const*)(q + 3 * m), db,
                      (const void* const*)(q + 4 * m), db, (void* const*)(q + 5 * m), db, m);
      if (r < n_reps) {
        const dim3 grid((unsigned)((da * db + 255) / 256), (unsigned)m);
        hipLaunchKernelGGL(diag_batch_kernel<T>, grid, dim3(256), 0, stream,
                           (Cx<T>* const*)(dptr + (size_t)n_steps * 6 * m),
                           prog.dev<double>(j_off) + (size_t)r * m * 2 * nn, n, sa.d, da, sb.d,
                           db, -1.0);
        HIP_CHECK(hipGetLastError());
      }
    }
    prog.finish();
  }
}

// One sector rotation of every state of a batch, planned for multi-state
// window passes: the pass structure (window starts, sub-passes, layer-pair
// counts) is common to the states, and each state's layer pairs and orbital
// phases follow at fixed strides.
struct MultiSector {
  struct Pass {
    int lo = 0;
    bool dg = false;  // applies the orbital phases (the last pass)
    std::vector<SubPass> subs;
    int n_layers = 0;
    std::vector<LayerPair> layers;  // state-major, n_layers per state
  };
  std::vector<Pass> passes;
  bool need_diag = false;
  std::vector<double> diag;  // state-major, (re, im) of norb phases per state
};

// Plans the passes of the states' rotations (in place, window width w);
// false when the states' pass structures differ (a rotation with vanishing
// Givens angles skips rotations that the others apply).
bool plan_multi_sector(const std::vector<const SectorRotation*>& rots, int norb, int w,
                       MultiSector& ms) {
  const size_t m = rots.size();
  std::vector<std::vector<PlannedPass>> all(m);
  for (size_t i = 0; i < m; ++i) {
    all[i] = plan_passes(rots[i]->steps, norb, w, false, -1);
    if (all[i].size() != all[0].size()) return false;
    if (!rots[i]->trivial_diag) ms.need_diag = true;
  }
  if (ms.need_diag) {
    ms.diag.assign(m * 2 * norb, 0.0);
    for (size_t i = 0; i < m; ++i)
      for (int k = 0; k < norb; ++k) {
        ms.diag[(i * norb + k) * 2] = rots[i]->diag[k].real();
        ms.diag[(i * norb + k) * 2 + 1] = rots[i]->diag[k].imag();
      }
  }
  std::vector<SubPass> subs;
  std::vector<LayerPair> layers;
  for (size_t k = 0; k < all[0].size(); ++k) {
    const PlannedPass& P0 = all[0][k];
    for (size_t i = 1; i < m; ++i) {
      const PlannedPass& Pi = all[i][k];
      if (Pi.lo != P0.lo || Pi.last != P0.last || Pi.steps.empty() != P0.steps.empty())
        return false;
    }
    MultiSector::Pass out;
    out.lo = P0.lo;
    out.dg = P0.last && ms.need_diag;
    if (P0.steps.empty() && !out.dg) continue;
    for (size_t i = 0; i < m; ++i) {
      plan_sub_passes(all[i][k].steps, w, kMaxSubWindow, subs, layers);
      if (i == 0) {
        out.subs = subs;
        out.n_layers = (int)layers.size();
        out.layers.reserve(m * layers.size());
      } else if (subs.size() != out.subs.size() || (int)layers.size() != out.n_layers ||
                 std::memcmp(subs.data(), out.subs.data(), subs.size() * sizeof(SubPass)) != 0) {
        return false;
      }
      out.layers.insert(out.layers.end(), layers.begin(), layers.end());
    }
    ms.passes.push_back(std::move(out));
  }
  return true;
}

// Queues the passes of a planned multi-state sector rotation: one fiber
// launch per window pass over the tiles of every state (`states`, a device
// array of the m state addresses in the call's program).
void queue_multi_sector(Context& ctx, Program& prog, const MultiSector& ms, int dtype, int mode,
                        size_t states_off, int m, long ld, int n_batch, int norb, int nocc, int w,
                        hipStream_t stream, std::vector<std::function<void()>>& launches) {
  const DevStrings& ds = dev_strings(ctx, norb, nocc);
  const DevGroups& gr = dev_groups(ctx, w, dtype == FFSIM_C64);
  size_t diag_off = 0;
  if (ms.need_diag) diag_off = prog.append(ms.diag.data(), ms.diag.size() * sizeof(double));
  for (const MultiSector::Pass& P : ms.passes) {
    const size_t soff = prog.append(P.layers.data(), P.layers.size() * sizeof(LayerPair));
    const size_t poff = prog.append(P.subs.data(), P.subs.size() * sizeof(SubPass));
    LoSlots los{};
    int gd_cap = 0, gm_cap = 0;
    pass_slots(gr, P.subs, los, gd_cap, gm_cap);
    const DevWindow& win = dev_window(ctx, norb, nocc, P.lo, w);
    const int n_sub = (int)P.subs.size(), n_layers = P.n_layers;
    const bool dg = P.dg;
    launches.push_back([=, &ctx, &prog, &win, &gr, &ds]() {
      const MultiTiles mt{prog.dev<void*>(states_off), m, (long)n_layers, 2l * norb};
      const double* d_diag = dg ? prog.dev<double>(diag_off) : nullptr;
      if (dtype == FFSIM_C128)
        launch_fiber_multi<double>(mode, mt, ld, n_batch, win, dg, gr, los, gd_cap, gm_cap,
                                   prog.dev<SubPass>(poff), n_sub, prog.dev<LayerPair>(soff),
                                   n_layers, d_diag, ds.d, norb, ctx.lds_max, tile_width(dtype),
                                   stream);
      else
        launch_fiber_multi<float>(mode, mt, ld, n_batch, win, dg, gr, los, gd_cap, gm_cap,
                                  prog.dev<SubPass>(poff), n_sub, prog.dev<LayerPair>(soff),
                                  n_layers, d_diag, ds.d, norb, ctx.lds_max, tile_width(dtype),
                                  stream);
    });
  }
}

// States per program of the multi-state path.
constexpr int kMultiChunk = 256;

// The window widths of the multi-state path for this shape; false where it
// does not apply: a beta vector must fit one tile (the beta sector runs in
// place with strings on columns, its window all norb orbitals), both windows
// must hold a full register sub-window, and the tiles must have the
// production width.
bool multi_windows(Context& ctx, int dtype, int n, int na, int nb, int& w_a, int& w_b) {
  const int ww = dtype == FFSIM_C128 ? 8 : 4;
  const long db = (long)binom(n, nb);
  if (n < kMaxSubWindow || n > kMaxDiagOrbitals || tile_width(dtype) != ww ||
      !beta_in_columns(n, db, dtype, ctx))
    return false;
  w_a = choose_window(ctx, n, na, dtype);
  w_b = n;
  if (w_a < kMaxSubWindow) return false;
  return dev_groups(ctx, w_a, dtype == FFSIM_C64).w2 == kMaxSubWindow &&
         dev_groups(ctx, w_b, dtype == FFSIM_C64).w2 == kMaxSubWindow;
}

// One orbital rotation per state, planned for multi-state passes.
struct MultiRotation {
  MultiSector a, b;
  bool has_a = false, has_b = false;
};

// mats_a[i], mats_b[i]: state i's matrices; a sector whose first entry is null
// is not rotated (then every entry is null).  The beta sector reuses the
// alpha decomposition where ffsim_apply_orbital_rotation does.  False when
// the states' pass structures differ.
bool plan_multi_rotation(int n, int w_a, int w_b, const std::vector<const cdouble*>& mats_a,
                         const std::vector<const cdouble*>& mats_b, MultiRotation& mr) {
  const size_t m = mats_a.size();
  mr.has_a = mats_a[0] != nullptr;
  mr.has_b = mats_b[0] != nullptr;
  const bool same_order = triangle_order(n, w_a) == triangle_order(n, w_b);
  std::vector<SectorRotation> ra(m), rb(m);
  std::vector<const SectorRotation*> pa(m), pb(m);
  for (size_t i = 0; i < m; ++i) {
    if (mr.has_a) ra[i] = prepare_rotation(mats_a[i], n, w_a);
    if (mr.has_b) {
      if (mr.has_a && same_order &&
          std::memcmp(mats_a[i], mats_b[i], sizeof(cdouble) * n * n) == 0)
        rb[i] = ra[i];
      else
        rb[i] = prepare_rotation(mats_b[i], n, w_b);
    }
    pa[i] = &ra[i];
    pb[i] = &rb[i];
  }
  return (!mr.has_a || plan_multi_sector(pa, n, w_a, mr.a)) &&
         (!mr.has_b || plan_multi_sector(pb, n, w_b, mr.b));
}

// Queues a planned multi-state rotation: the alpha sector's passes (strings on
// rows), then the beta sector's (strings on columns), in place.
void queue_multi_rotation(Context& ctx, Program& prog, const MultiRotation& mr, int dtype,
                          size_t states_off, int m, int n, int na, int nb, int w_a, int w_b,
                          hipStream_t stream, std::vector<std::function<void()>>& launches) {
  const long da = (long)binom(n, na), db = (long)binom(n, nb);
  if (mr.has_a)
    queue_multi_sector(ctx, prog, mr.a, dtype, MODE_ROWS, states_off, m, db, (int)db, n, na, w_a,
                       stream, launches);
  if (mr.has_b)
    queue_multi_sector(ctx, prog, mr.b, dtype, MODE_COLS, states_off, m, db, (int)da, n, nb, w_b,
                       stream, launches);
}

// A batch of spin-balanced UCJ applications of one shape on the Givens path
// with every state's tiles in the same launches: per merged rotation one
// fiber launch per window pass of each sector over all states, per diagonal
// term one column-factor launch and one elementwise launch over all states.
// Each state goes through the same arithmetic as alone (ucj_impl), so the
// results are bitwise those of the single-state entry.  Applies when the
// beta rotation runs in place with strings on columns, the window holds a
// full sub-window, and the states' rotations share one pass structure (as
// rotations without vanishing Givens angles do); returns false otherwise,
// before anything is launched.
template <typename T>
bool ucj_batched_multi(Context& ctx, int n_states, void* const* vecs, int dtype, int n, int na,
                       int nb, int n_reps, const double* diag_coulomb_mats,
                       const double* orbital_rotations, const double* final_rots,
                       hipStream_t stream) {
  const long da = (long)binom(n, na), db = (long)binom(n, nb);
  int w_a = 0, w_b = 0;
  if (!multi_windows(ctx, dtype, n, na, nb, w_a, w_b)) return false;
  require(diag_grid(da, db).y <= 65535, FFSIM_ERR_VALUE, "too many columns");
  const size_t nn = (size_t)n * n;
  const int n_steps = n_reps + 1;
  for (int i0 = 0; i0 < n_states; i0 += kMultiChunk) {
    const int m = std::min(kMultiChunk, n_states - i0);
    // plan every state's rotations first: nothing is launched when they differ
    std::vector<std::vector<std::vector<cdouble>>> V(m);
    for (int i = 0; i < m; ++i) {
      const int st = i0 + i;
      V[i] = ucj_merged_rotations(
          n, n_reps,
          reinterpret_cast<const cdouble*>(orbital_rotations) + (size_t)st * n_reps * nn,
          final_rots ? reinterpret_cast<const cdouble*>(final_rots) + (size_t)st * nn : nullptr);
    }
    std::vector<MultiRotation> mr(n_steps);
    for (int r = 0; r < n_steps; ++r) {
      std::vector<const cdouble*> pm(m);
      for (int i = 0; i < m; ++i) pm[i] = V[i][r].data();
      if (!plan_multi_rotation(n, w_a, w_b, pm, pm, mr[r])) {
        if (i0 == 0) return false;
        // earlier chunks are queued: the rest state by state (the beta sector
        // runs in place, no work buffer)
        for (int st = i0; st < n_states; ++st)
          ucj_impl(ctx, vecs[st], nullptr, dtype, n, na, nb, n_reps,
                   diag_coulomb_mats + (size_t)st * n_reps * 2 * nn,
                   orbital_rotations + (size_t)st * n_reps * 2 * nn,
                   final_rots ? final_rots + (size_t)st * 2 * nn : nullptr, stream);
        return true;
      }
    }
    Program prog(ctx, stream);
    std::vector<const void*> ptrs(vecs + i0, vecs + i0 + m);
    const size_t states_off = prog.append(ptrs.data(), m * sizeof(void*));
    // the diagonal terms: headers, then coefficients, state-major per term
    const DevStrings& sa = dev_strings(ctx, n, na);
    const DevStrings& sb = dev_strings(ctx, n, nb);
    std::vector<size_t> hoffs(n_reps), coffs(n_reps);
    long coef_len = 0;
    for (int r = 0; r < n_reps; ++r) {
      std::vector<DiagHeader> hs(m);
      std::vector<double> cs;
      for (int i = 0; i < m; ++i) {
        const double* J = diag_coulomb_mats + ((size_t)(i0 + i) * n_reps + r) * 2 * nn;
        std::vector<double> halved(J, J + nn);
        for (int k = 0; k < n; ++k) halved[(size_t)k * n + k] *= 0.5;
        const DiagSpec spec = diag_coulomb_spec(n, halved.data(), J + nn, halved.data(), 0, -1.0);
        hs[i] = diag_header(spec);
        const std::vector<double> c = diag_coefficients(spec);
        coef_len = (long)c.size();
        cs.insert(cs.end(), c.begin(), c.end());
      }
      hoffs[r] = prog.append(hs.data(), hs.size() * sizeof(DiagHeader));
      coffs[r] = prog.append(cs.data(), cs.size() * sizeof(double));
    }
    size_t goff = 0;
    require(n_reps == 0 || prog.reserve((size_t)m * 2 * db * sizeof(double), &goff),
            FFSIM_ERR_VALUE, "batched UCJ: column factors exceed the program scratch");
    std::vector<std::function<void()>> launches;
    for (int r = 0; r < n_steps; ++r) {
      queue_multi_rotation(ctx, prog, mr[r], dtype, states_off, m, n, na, nb, w_a, w_b, stream,
                           launches);
      if (r == n_reps) break;
      const size_t hoff = hoffs[r], coff = coffs[r];
      launches.push_back([=, &prog, &sa, &sb]() {
        const DiagHeader* dh = prog.dev<DiagHeader>(hoff);
        const double* dc = prog.dev<double>(coff);
        double* dg = prog.scratch<double>(goff);
        launch_diag_col_term(db, sb.d, dh, dc, dg, 1, stream, m, coef_len, 2 * db);
        HIP_CHECK(hipGetLastError());
        const double moved = 2.0 * sizeof(Cx<T>) * (double)da * (double)db * m;
        const long pk = prof_start(PROF_DIAG, moved, stream);
        launch_diag_pass_t<T>(DIAG_EVOLVE, nullptr, nullptr, da, db, sa.d, sb.d, dg, dh, dc,
                              stream, prog.dev<void*>(states_off), m, coef_len, 2 * db);
        HIP_CHECK(hipGetLastError());
        prof_stop(pk, stream);
      });
    }
    prog.commit();
    for (auto& f : launches) f();
    prog.finish();
  }
  return true;
}

// Orbital rotations of a batch of states of one shape (mats_a / mats_b: null,
// or one norb x norb matrix per state): where the multi-state path applies,
// one launch per window pass and sector over all states; otherwise (or when
// the states' pass structures differ) state by state on the stream, so the
// states' work buffers may be one buffer.  Either way every state's result is
// bitwise what ffsim_apply_orbital_rotation gives.
void orbital_rotation_batched_impl(Context& ctx, int n_states, void* const* vecs,
                                   void* const* works, int dtype, int n, int na, int nb,
                                   const cdouble* mats_a, const cdouble* mats_b,
                                   hipStream_t stream) {
  const long da = (long)binom(n, na), db = (long)binom(n, nb);
  if (da == 0 || db == 0 || n == 0 || (!mats_a && !mats_b)) return;
  const size_t nn = (size_t)n * n;
  int w_a = 0, w_b = 0;
  const bool multi = n_states >= 2 && multi_windows(ctx, dtype, n, na, nb, w_a, w_b);
  for (int i0 = 0; i0 < n_states; i0 += kMultiChunk) {
    const int m = std::min(kMultiChunk, n_states - i0);
    std::vector<const cdouble*> pa(m), pb(m);
    for (int i = 0; i < m; ++i) {
      pa[i] = mats_a ? mats_a + (size_t)(i0 + i) * nn : nullptr;
      pb[i] = mats_b ? mats_b + (size_t)(i0 + i) * nn : nullptr;
    }
    MultiRotation mr;
    if (multi && m >= 2 && plan_multi_rotation(n, w_a, w_b, pa, pb, mr)) {
      Program prog(ctx, stream);
      std::vector<const void*> ptrs(vecs + i0, vecs + i0 + m);
      const size_t states_off = prog.append(ptrs.data(), m * sizeof(void*));
      std::vector<std::function<void()>> launches;
      queue_multi_rotation(ctx, prog, mr, dtype, states_off, m, n, na, nb, w_a, w_b, stream,
                           launches);
      prog.commit();
      for (auto& f : launches) f();
      prog.finish();
    } else {
      for (int i = 0; i < m; ++i)
        orbital_rotation_impl(ctx, vecs[i0 + i], works ? works[i0 + i] : nullptr, dtype, n, na,
                              nb, pa[i], pb[i], stream);
    }
  }
}

constexpr long kMaxAuxStreams = 4;

void ucj_batched_impl(Context& ctx, int n_states, void* const* vecs, void* const* works,
                      int dtype, int n, int na, int nb, int n_reps,
                      const double* diag_coulomb_mats, const double* orbital_rotations,
                      const double* final_rots, hipStream_t stream) {
  const size_t nn = (size_t)n * n;
  const long da = (long)binom(n, na), db = (long)binom(n, nb);
  if (da == 0 || db == 0) return;
  if (gemm_path(n, na, nb, n_states)) {
    for (int i = 0; i < n_states; ++i)
      require(works[i] != nullptr, FFSIM_ERR_VALUE, "batched UCJ: a work buffer per state");
    if (dtype == FFSIM_C128)
      ucj_batched_gemm<double>(ctx, n_states, vecs, works, n, na, nb, n_reps, diag_coulomb_mats,
                               orbital_rotations, final_rots, stream);
    else
      ucj_batched_gemm<float>(ctx, n_states, vecs, works, n, na, nb, n_reps, diag_coulomb_mats,
                              orbital_rotations, final_rots, stream);
    return;
  }
  if (n_states >= 2 && batch_path_override() != 3) {
    const bool done =
        dtype == FFSIM_C128
            ? ucj_batched_multi<double>(ctx, n_states, vecs, dtype, n, na, nb, n_reps,
                                        diag_coulomb_mats, orbital_rotations, final_rots, stream)
            : ucj_batched_multi<float>(ctx, n_states, vecs, dtype, n, na, nb, n_reps,
                                       diag_coulomb_mats, orbital_rotations, final_rots, stream);
    if (done) return;
  }
  // A launch over a small state has fewer tiles than the device has CUs, so
  // several states run at once, one per auxiliary stream, between events on
  // the caller's stream (each state's work stays in order on its stream).
  // Only for calls on the default stream: a caller that runs its own streams
  // (the server) already keeps the device's few hardware queues busy, and
  // more streams than queues only add false dependencies (measured: serving
  // saturation 1.17x instead of 1.25x with them).
  const long tiles = std::max(da, db) / 8;
  const int n_aux = (n_states < 2 || stream != nullptr) ? 0
                                 : (int)std::min<long>({kMaxAuxStreams, n_states,
                                                        std::max<long>(1, g_n_cu / std::max(1l, tiles))});
  if (n_aux < 2) {
    for (int i = 0; i < n_states; ++i)
      ucj_impl(ctx, vecs[i], works[i], dtype, n, na, nb, n_reps,
               diag_coulomb_mats + (size_t)i * n_reps * 2 * nn,
               orbital_rotations + (size_t)i * n_reps * 2 * nn,
               final_rots ? final_rots + (size_t)i * 2 * nn : nullptr, stream, da);
    return;
  }
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    while ((int)ctx.aux.size() < n_aux) {
      hipStream_t s = nullptr;
      HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      ctx.aux.push_back(s);
    }
  }
  auto event = [] {
    hipEvent_t e = nullptr;
    HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
  };
  hipEvent_t start = event();
  HIP_CHECK(hipEventRecord(start, stream));
  for (int k = 0; k < n_aux; ++k) HIP_CHECK(hipStreamWaitEvent(ctx.aux[k], start, 0));
  for (int i = 0; i < n_states; ++i)
    ucj_impl(ctx, vecs[i], works[i], dtype, n, na, nb, n_reps,
             diag_coulomb_mats + (size_t)i * n_reps * 2 * nn,
             orbital_rotations + (size_t)i * n_reps * 2 * nn,
             final_rots ? final_rots + (size_t)i * 2 * nn : nullptr, ctx.aux[i % n_aux], da);
  for (int k = 0; k < n_aux; ++k) {
    hipEvent_t done = event();
    HIP_CHECK(hipEventRecord(done, ctx.aux[k]));
    HIP_CHECK(hipStreamWaitEvent(stream, done, 0));
    HIP_CHECK(hipEventDestroy(done));  // released once it has completed
  }
  HIP_CHECK(hipEventDestroy(start));
}
}  // namespace
}  // namespace ffsim_amd

using namespace ffsim_amd;

// Tile width and workgroup of the one-body window passes: 64-byte batch
// segments, two workgroups per CU.
constexpr int kOneBodyThreads = 512;
int one_body_tile_width(int dtype) { return dtype == FFSIM_C64 ? 8 : 4; }

template <typename T>
void launch_one_body_pass(int mode, const void* src, void* dst, long ld, int n_batch,
                          const DevWindow& win, const uint32_t* links, const int32_t* lidx,
                          const double* mloc, double scale, int acc, size_t lds_max,
                          hipStream_t st) {
  constexpr int W = sizeof(T) == 8 ? 4 : 8;
  const size_t lds = align16((size_t)win.max_rows * W * sizeof(Cx<T>)) +
                     2 * (size_t)win.w * win.w * sizeof(double);
  require(lds <= lds_max && win.max_rows <= 1024, FFSIM_ERR_VALUE, "one-body tile exceeds LDS");
  const long n_tiles = (long)win.n_fibers * ((n_batch + W - 1) / W);
  auto kern = mode == MODE_ROWS ? one_body_fiber_kernel<T, W, MODE_ROWS, kOneBodyThreads>
                                : one_body_fiber_kernel<T, W, MODE_COLS, kOneBodyThreads>;
  static bool attr_set[2] = {false, false};
  if (!attr_set[mode]) {
    HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_max));
    attr_set[mode] = true;
  }
  int occ = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)kern,
                                                         kOneBodyThreads, lds));
  const long grid = std::min<long>(n_tiles, (long)std::max(1, occ) * g_n_cu);
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kOneBodyThreads), lds, st,
                     reinterpret_cast<const Cx<T>*>(src), reinterpret_cast<Cx<T>*>(dst), ld,
                     n_batch, win.fibers, win.n_fibers, win.rows, win.max_rows, links, lidx, mloc,
                     win.w, scale, acc);
  HIP_CHECK(hipGetLastError());
}

extern "C" {

const char* ffsim_amd_last_error(void) { return g_last_error.c_str(); }
int ffsim_amd_version(void) { return 1; }

int ffsim_amd_init(int device) {
  return guarded([&] {
    HIP_CHECK(hipSetDevice(device));
    context();
    noalloc_setup();
  });
}

int ffsim_amd_profile_begin(void) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.on = true;
    g_prof.used = 0;
  });
}

int ffsim_amd_profile_end(double* out) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.on = false;
    for (int k = 0; k < 3 * PROF_KINDS; ++k) out[k] = 0.0;
    for (size_t k = 0; k < g_prof.used; ++k) {
      HIP_CHECK(hipEventSynchronize(g_prof.stop[k]));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, g_prof.start[k], g_prof.stop[k]));
      const int kind = g_prof.kind[k];
      out[3 * kind + 0] += ms;
      out[3 * kind + 1] += 1.0;
      out[3 * kind + 2] += g_prof.bytes[k];
    }
    g_prof.used = 0;
  });
}

int ffsim_make_strings(int norb, int nocc, int64_t* out) {
  return guarded([&] {
    require(norb >= 0 && norb <= 63 && nocc >= 0, FFSIM_ERR_VALUE, "bad norb/nocc");
    std::vector<uint64_t> s = make_strings(norb, nocc);
    for (size_t k = 0; k < s.size(); ++k) out[k] = (int64_t)s[k];
  });
}

int ffsim_gen_occslst(int norb, int nocc, uint64_t* out) {
  return guarded([&] {
    require(norb >= 0 && norb <= 63 && nocc >= 0, FFSIM_ERR_VALUE, "bad norb/nocc");
    std::vector<uint64_t> s = make_strings(norb, nocc);
    size_t k = 0;
    for (uint64_t b : s)
      for (int o = 0; o < norb; ++o)
        if ((b >> o) & 1) out[k++] = (uint64_t)o;
  });
}

int ffsim_zero_one_subspace_indices(int norb, int nocc, int i, int j, uint64_t* out) {
  int written = 0;
  int rc = guarded([&] {
    require(norb >= 2 && norb <= 63 && i >= 0 && j >= 0 && i < norb && j < norb && i != j,
            FFSIM_ERR_VALUE, "bad orbitals");
    std::vector<uint64_t> s = make_strings(norb, nocc);
    std::vector<uint64_t> first, second;
    for (size_t a = 0; a < s.size(); ++a) {
      const uint64_t b = s[a];
      if (((b >> i) & 1) && !((b >> j) & 1)) {
        first.push_back(a);
        second.push_back(string_rank(b ^ (1ull << i) ^ (1ull << j)));
      }
    }
    for (auto v : first) out[written++] = v;
    for (auto v : second) out[written++] = v;
  });
  return rc == FFSIM_OK ? written : rc;
}

int ffsim_givens_decomposition(const double* mat, int n, double tol, double* c, double* s,
                               int32_t* i, int32_t* j, int32_t* n_rot, double* diag) {
  return guarded([&] {
    require(n >= 0, FFSIM_ERR_VALUE, "mat must be a square matrix");
    GivensDecomposition g = givens_decomposition(reinterpret_cast<const cdouble*>(mat), n, tol);
    for (size_t k = 0; k < g.rotations.size(); ++k) {
      c[k] = g.rotations[k].c;
      s[2 * k] = g.rotations[k].s.real();
      s[2 * k + 1] = g.rotations[k].s.imag();
      i[k] = g.rotations[k].i;
      j[k] = g.rotations[k].j;
    }
    *n_rot = (int32_t)g.rotations.size();
    for (int k = 0; k < n; ++k) {
      diag[2 * k] = g.diag[k].real();
      diag[2 * k + 1] = g.diag[k].imag();
    }
  });
}

int64_t ffsim_orbital_rotation_work_elements(int dtype, int norb, int n_alpha, int n_beta) {
  int64_t out = 0;
  int rc = guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    Context& ctx = context();
    const long dim_a = (long)binom(norb, n_alpha), dim_b = (long)binom(norb, n_beta);
    const bool beta_cols = beta_in_columns(norb, dim_b, dtype, ctx);
    out = beta_cols ? 0 : (int64_t)transposed_ld(dtype, dim_a) * dim_b;
  });
  return rc == FFSIM_OK ? out : rc;
}

int ffsim_apply_orbital_rotation(void* vec, void* work, int dtype, int norb, int n_alpha,
                                 int n_beta, const double* mat_a, const double* mat_b,
                                 void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    Context& ctx = context();
    orbital_rotation_impl(ctx, vec, work, dtype, norb, n_alpha, n_beta,
                          reinterpret_cast<const cdouble*>(mat_a),
                          reinterpret_cast<const cdouble*>(mat_b), (hipStream_t)stream);
  });
}

int ffsim_apply_orbital_rotation_batched(int n_states, void* const* vecs, void* const* works,
                                         int dtype, int norb, int n_alpha, int n_beta,
                                         const double* mats_a, const double* mats_b,
                                         void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    require(n_states >= 0, FFSIM_ERR_VALUE, "negative batch size");
    if (n_states == 0) return;
    require(vecs != nullptr, FFSIM_ERR_VALUE, "batched rotation: no state array");
    for (int i = 0; i < n_states; ++i)
      require(vecs[i] != nullptr, FFSIM_ERR_VALUE, "batched rotation: null state");
    Context& ctx = context();
    orbital_rotation_batched_impl(ctx, n_states, vecs, works, dtype, norb, n_alpha, n_beta,
                                  reinterpret_cast<const cdouble*>(mats_a),
                                  reinterpret_cast<const cdouble*>(mats_b), (hipStream_t)stream);
  });
}

int ffsim_apply_diag_coulomb_evolution(void* vec, int dtype, int norb, int n_alpha, int n_beta,
                                       const double* mat_aa, const double* mat_ab,
                                       const double* mat_bb, double time, int z_representation,
                                       void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    Context& ctx = context();
    DiagSpec s = diag_coulomb_spec(norb, mat_aa, mat_ab, mat_bb, z_representation, time);
    if (!s.flags) return;
    diag_impl(ctx, DIAG_EVOLVE, vec, vec, dtype, norb, n_alpha, n_beta, s, (hipStream_t)stream);
  });
}

int ffsim_apply_diag_coulomb_evolution_batched(int n_states, void* const* vecs, int dtype,
                                               int norb, int n_alpha, int n_beta,
                                               const double* mat_aa, const double* mat_ab,
                                               const double* mat_bb, double time,
                                               int z_representation, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    require(n_states >= 0, FFSIM_ERR_VALUE, "negative batch size");
    if (n_states == 0) return;
    require(vecs != nullptr, FFSIM_ERR_VALUE, "batched diagonal evolution: no state array");
    for (int i = 0; i < n_states; ++i)
      require(vecs[i] != nullptr, FFSIM_ERR_VALUE, "batched diagonal evolution: null state");
    Context& ctx = context();
    DiagSpec s = diag_coulomb_spec(norb, mat_aa, mat_ab, mat_bb, z_representation, time);
    if (!s.flags) return;
    diag_batched_impl(ctx, DIAG_EVOLVE, n_states, vecs, dtype, norb, n_alpha, n_beta, {s},
                      (hipStream_t)stream);
  });
}

int ffsim_apply_num_op_sum_evolution(void* vec, int dtype, int norb, int n_alpha, int n_beta,
                                     const double* coeffs_a, const double* coeffs_b, double time,
                                     void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    if (!coeffs_a && !coeffs_b) return;
    Context& ctx = context();
    DiagSpec s;
    s.norb = norb;
    s.scale = time;
    s.flags = DF_LIN;
    s.lin_row = vec_or_empty(coeffs_a, norb);
    s.lin_col = vec_or_empty(coeffs_b, norb);
    diag_impl(ctx, DIAG_EVOLVE, vec, vec, dtype, norb, n_alpha, n_beta, s, (hipStream_t)stream);
  });
}

int ffsim_contract_diag_coulomb(const void* vec, void* out, int dtype, int norb, int n_alpha,
                                int n_beta, const double* mat_aa, const double* mat_ab,
                                const double* mat_bb, int z_representation, double scale,
                                int accumulate, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    Context& ctx = context();
    DiagSpec s = diag_coulomb_spec(norb, mat_aa, mat_ab, mat_bb, z_representation, scale);
    diag_impl(ctx, accumulate ? DIAG_CONTRACT_ACC : DIAG_CONTRACT, vec, out, dtype, norb,
              n_alpha, n_beta, s, (hipStream_t)stream);
  });
}

int ffsim_contract_diag_coulomb_batched(int n_states, void* const* vecs, int dtype, int norb,
                                        int n_alpha, int n_beta, const double* mats_aa,
                                        const double* mats_ab, const double* mats_bb,
                                        int z_representation, double scale, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    require(n_states >= 0, FFSIM_ERR_VALUE, "negative batch size");
    if (n_states == 0) return;
    require(vecs != nullptr, FFSIM_ERR_VALUE, "batched contraction: no state array");
    for (int i = 0; i < n_states; ++i)
      require(vecs[i] != nullptr, FFSIM_ERR_VALUE, "batched contraction: null state");
    Context& ctx = context();
    const size_t nn = (size_t)norb * norb;
    std::vector<DiagSpec> specs;
    specs.reserve(n_states);
    for (int i = 0; i < n_states; ++i)
      specs.push_back(diag_coulomb_spec(norb, mats_aa ? mats_aa + i * nn : nullptr,
                                        mats_ab ? mats_ab + i * nn : nullptr,
                                        mats_bb ? mats_bb + i * nn : nullptr, z_representation,
                                        scale));
    diag_batched_impl(ctx, DIAG_CONTRACT, n_states, vecs, dtype, norb, n_alpha, n_beta, specs,
                      (hipStream_t)stream);
  });
}

int ffsim_contract_num_op_sum(const void* vec, void* out, int dtype, int norb, int n_alpha,
                              int n_beta, const double* coeffs_a, const double* coeffs_b,
                              double scale, int accumulate, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    Context& ctx = context();
    DiagSpec s;
    s.norb = norb;
    s.scale = scale;
    s.flags = DF_LIN;
    s.lin_row = vec_or_empty(coeffs_a, norb);
    s.lin_col = vec_or_empty(coeffs_b, norb);
    diag_impl(ctx, accumulate ? DIAG_CONTRACT_ACC : DIAG_CONTRACT, vec, out, dtype, norb,
              n_alpha, n_beta, s, (hipStream_t)stream);
  });
}

int ffsim_contract_one_body(const void* vec, void* out, int dtype, int norb, int n_alpha,
                            int n_beta, const double* mat, double scale, int accumulate,
                            void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    require(vec != out, FFSIM_ERR_VALUE, "out must not alias vec");
    require(norb <= 64, FFSIM_ERR_VALUE, "norb above 64");
    Context& ctx = context();
    const long da = (long)binom(norb, n_alpha), db = (long)binom(norb, n_beta);
    hipStream_t st = (hipStream_t)stream;
    if (da == 0 || db == 0) return;
    require(da < (1l << 31) && db < (1l << 31), FFSIM_ERR_VALUE, "dimension too large");
    // window passes when every sector's windows fit a tile
    const int elem = (int)elem_bytes(dtype), Wt = one_body_tile_width(dtype);
    auto fits = [&](int nocc, int w) {
      return (size_t)max_fiber(norb, nocc, std::min(w, norb)) * Wt * elem + 2 * w * w * 8 + 64 <=
             ctx.lds_max / 2;
    };
    int w = std::min(norb, 12);
    while (w > 1 && !(fits(n_alpha, w) && fits(n_beta, w))) --w;
    const std::vector<OneBodyPass> passes = one_body_passes(norb, w);
    if (!passes.empty() && norb >= 1 && norb <= 31) {
      Program prog(ctx, st);
      const cdouble* M = reinterpret_cast<const cdouble*>(mat);
      std::vector<size_t> moffs;
      for (const OneBodyPass& P : passes) {
        std::vector<int> loc;
        for (int b = 0; b < norb; ++b)
          if ((P.mask >> b) & 1) loc.push_back(b);
        const int wl = (int)loc.size();
        std::vector<double> m((size_t)2 * wl * wl, 0.0);
        for (int i = 0; i < wl; ++i)
          for (int j = 0; j < wl; ++j)
            if (P.take[(size_t)loc[i] * norb + loc[j]]) {
              m[2 * (i * wl + j)] = M[(size_t)loc[i] * norb + loc[j]].real();
              m[2 * (i * wl + j) + 1] = M[(size_t)loc[i] * norb + loc[j]].imag();
            }
        moffs.push_back(prog.append(m.data(), m.size() * sizeof(double)));
      }
      prog.commit();
      int acc = accumulate;
      for (int sector = 0; sector < 2; ++sector) {
        const int nocc = sector == 0 ? n_alpha : n_beta;
        const int mode = sector == 0 ? MODE_ROWS : MODE_COLS;
        const long n_batch = sector == 0 ? db : da;
        for (size_t k = 0; k < passes.size(); ++k) {
          const DevWindow& win = dev_mask_window(ctx, norb, nocc, passes[k].mask, passes[k].between);
          const auto lk = dev_one_body_links(ctx, win.w);
          if (dtype == FFSIM_C128)
            launch_one_body_pass<double>(mode, vec, out, db, (int)n_batch, win, lk.first,
                                         lk.second, prog.dev<double>(moffs[k]), scale, acc,
                                         ctx.lds_max, st);
          else
            launch_one_body_pass<float>(mode, vec, out, db, (int)n_batch, win, lk.first,
                                        lk.second, prog.dev<double>(moffs[k]), scale, acc,
                                        ctx.lds_max, st);
          acc = 1;
        }
      }
      prog.finish();
      return;
    }
    Program prog(ctx, st);
    const size_t moff = prog.append(mat, sizeof(double) * 2 * norb * norb);
    prog.commit();
    const double* dm = prog.dev<double>(moff);
    bool first = true;
    for (int sector = 0; sector < 2; ++sector) {
      const int nocc = sector == 0 ? n_alpha : n_beta;
      const DevLinks& L = dev_links(ctx, norb, nocc);
      const int acc = first ? accumulate : 1;
      if (L.n_links == 0) {
        if (first && !accumulate) {
          // no electrons in this sector: its term is zero
          const long n = da * db * (dtype == FFSIM_C128 ? 2 : 1);
          hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)std::min<long>((n + 255) / 256, 65536)),
                             dim3(256), 0, st, (double*)out, n);
          first = false;
        }
        continue;
      }
      const dim3 grid = sector == 0 ? dim3((unsigned)da, (unsigned)((db + 255) / 256))
                                    : dim3((unsigned)((db + 255) / 256), (unsigned)da);
      if (dtype == FFSIM_C128) {
        if (sector == 0)
          hipLaunchKernelGGL((one_body_kernel<double, 0>), grid, dim3(256), 0, st,
                             (const Cx<double>*)vec, (Cx<double>*)out, (int)da, (int)db, L.d,
                             L.n_links, dm, norb, scale, acc);
        else
          hipLaunchKernelGGL((one_body_kernel<double, 1>), grid, dim3(256), 0, st,
                             (const Cx<double>*)vec, (Cx<double>*)out, (int)da, (int)db, L.d,
                             L.n_links, dm, norb, scale, acc);
      } else {
        if (sector == 0)
          hipLaunchKernelGGL((one_body_kernel<float, 0>), grid, dim3(256), 0, st,
                             (const Cx<float>*)vec, (Cx<float>*)out, (int)da, (int)db, L.d,
                             L.n_links, dm, norb, scale, acc);
        else
          hipLaunchKernelGGL((one_body_kernel<float, 1>), grid, dim3(256), 0, st,
                             (const Cx<float>*)vec, (Cx<float>*)out, (int)da, (int)db, L.d,
                             L.n_links, dm, norb, scale, acc);
      }
      HIP_CHECK(hipGetLastError());
      first = false;
    }
    prog.finish();
  });
}

int ffsim_axpby(void* out, const void* x, const void* y, int dtype, int64_t n, double alpha_re,
                double alpha_im, double beta_re, double beta_im, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    if (n <= 0) return;
    context();
    const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    if (dtype == FFSIM_C128)
      hipLaunchKernelGGL(axpby_kernel<double>, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                         reinterpret_cast<Cx<double>*>(out),
                         reinterpret_cast<const Cx<double>*>(x),
                         reinterpret_cast<const Cx<double>*>(y), (long)n, alpha_re, alpha_im,
                         beta_re, beta_im);
    else
      hipLaunchKernelGGL(axpby_kernel<float>, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                         reinterpret_cast<Cx<float>*>(out), reinterpret_cast<const Cx<float>*>(x),
                         reinterpret_cast<const Cx<float>*>(y), (long)n, alpha_re, alpha_im,
                         beta_re, beta_im);
    HIP_CHECK(hipGetLastError());
  });
}

int ffsim_apply_ucj_spin_balanced(void* vec, void* work, int dtype, int norb, int n_alpha,
                                  int n_beta, int n_reps, const double* diag_coulomb_mats,
                                  const double* orbital_rotations,
                                  const double* final_orbital_rotation, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    require(n_reps >= 0, FFSIM_ERR_VALUE, "n_reps must be non-negative");
    Context& ctx = context();
    ucj_impl(ctx, vec, work, dtype, norb, n_alpha, n_beta, n_reps, diag_coulomb_mats,
             orbital_rotations, final_orbital_rotation, (hipStream_t)stream);
  });
}

int ffsim_apply_ucj_spin_balanced_batched(int n_states, void* const* vecs, void* const* works,
                                          int dtype, int norb, int n_alpha, int n_beta,
                                          int n_reps, const double* diag_coulomb_mats,
                                          const double* orbital_rotations,
                                          const double* final_orbital_rotations, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    require(n_reps >= 0 && n_states >= 0, FFSIM_ERR_VALUE, "negative count");
    if (n_states == 0) return;
    Context& ctx = context();
    ucj_batched_impl(ctx, n_states, vecs, works, dtype, norb, n_alpha, n_beta, n_reps,
                     diag_coulomb_mats, orbital_rotations, final_orbital_rotations,
                     (hipStream_t)stream);
  });
}

int ffsim_diag_coulomb_grad_reduce(const void* psi, const void* lam, int dtype, int norb,
                                   int n_alpha, int n_beta, double* out, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    require(norb <= 32, FFSIM_ERR_VALUE, "gradient reductions support norb <= 32");
    Context& ctx = context();
    const long da = (long)binom(norb, n_alpha), db = (long)binom(norb, n_beta);
    if (da == 0 || db == 0) return;
    const DevStrings& cs = dev_strings(ctx, norb, n_beta);
    hipStream_t st = (hipStream_t)stream;
    double* rsum = out;
    double* csum = out + da;
    double* M = out + da + db;
    // column sums over row chunks (enough blocks to fill the device), then the
    // chunks of each column added in order
    const long col_blocks = (db + 255) / 256;
    const long n_chunks =
        std::max(1l, std::min<long>({da, 32l, (4l * g_n_cu + col_blocks - 1) / col_blocks}));
    const long rows_per_chunk = (da + n_chunks - 1) / n_chunks;
    Program prog(ctx, st);
    size_t poff = 0;
    require(prog.reserve((size_t)(n_chunks * db) * sizeof(double), &poff), FFSIM_ERR_VALUE,
            "gradient reductions: scratch too large");
    prog.commit();
    double* part = prog.scratch<double>(poff);
    const dim3 cgrid((unsigned)col_blocks, (unsigned)n_chunks);
    if (dtype == FFSIM_C128) {
      hipLaunchKernelGGL(grad_rows_kernel<double>, dim3((unsigned)da), dim3(256), 0, st,
                         (const Cx<double>*)psi, (const Cx<double>*)lam, (int)da, (int)db, norb,
                         cs.d, rsum, M);
      hipLaunchKernelGGL(grad_cols_kernel<double>, cgrid, dim3(256), 0, st,
                         (const Cx<double>*)psi, (const Cx<double>*)lam, (int)da, (int)db,
                         (int)rows_per_chunk, part);
    } else {
      hipLaunchKernelGGL(grad_rows_kernel<float>, dim3((unsigned)da), dim3(256), 0, st,
                         (const Cx<float>*)psi, (const Cx<float>*)lam, (int)da, (int)db, norb,
                         cs.d, rsum, M);
      hipLaunchKernelGGL(grad_cols_kernel<float>, cgrid, dim3(256), 0, st,
                         (const Cx<float>*)psi, (const Cx<float>*)lam, (int)da, (int)db,
                         (int)rows_per_chunk, part);
    }
    hipLaunchKernelGGL(col_chunk_sum_kernel, dim3((unsigned)col_blocks), dim3(256), 0, st, part,
                       (int)n_chunks, (int)db, csum);
    HIP_CHECK(hipGetLastError());
    prog.finish();
  });
}

int ffsim_transition_rdm(const void* lam, const void* psi, int dtype, int norb, int n_alpha,
                         int n_beta, double* out, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    Context& ctx = context();
    const long da = (long)binom(norb, n_alpha), db = (long)binom(norb, n_beta);
    hipStream_t st = (hipStream_t)stream;
    const long n_out = 4l * norb * norb;
    if (n_out == 0) return;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)std::min<long>((n_out + 255) / 256, 1024)),
                       dim3(256), 0, st, out, n_out);
    if (da == 0 || db == 0) return;
    // binomial table C(c, t) for c, t <= norb, uploaded through the program ring
    Program prog(ctx, st);
    std::vector<uint64_t> tab((size_t)(norb + 1) * (norb + 1));
    for (int c = 0; c <= norb; ++c)
      for (int t = 0; t <= norb; ++t) tab[(size_t)c * (norb + 1) + t] = binom(c, t);
    const size_t toff = prog.append(tab.data(), tab.size() * sizeof(uint64_t));
    // per-(pair, string) dot products of a chunk of strings, then the beta
    // sector's lam^T and psi^T when they fit the call's scratch (its strings
    // then index rows, as the alpha sector's do, and every read is coalesced)
    const long n2 = 2l * norb * norb;
    const long n_str_max = std::max(da, db);
    const long chunk = std::min<long>(n_str_max, std::max<long>(4, (64l << 20) / (n2 * 8)));
    size_t poff = 0;
    require(prog.reserve((size_t)(n2 * chunk * 8), &poff), FFSIM_ERR_VALUE,
            "transition RDM: scratch too large");
    const size_t state_bytes = (size_t)(da * db) * elem_bytes(dtype);
    size_t tpoff = 0;
    const bool transpose_b = prog.reserve(2 * state_bytes, &tpoff);
    prog.commit();
    const DevStrings& sa = dev_strings(ctx, norb, n_alpha);
    const DevStrings& sb = dev_strings(ctx, norb, n_beta);
    double* part = prog.scratch<double>(poff);
    const uint64_t* dtab = prog.dev<uint64_t>(toff);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      const Cx<T>* L = (const Cx<T>*)lam;
      const Cx<T>* P = (const Cx<T>*)psi;
      // one sector: strings n_str on rows (cols = false) or on columns
      auto sector = [&](const Cx<T>* l, const Cx<T>* x, bool cols, long n_str, long n_other,
                        const uint64_t* strs, double* rdm) {
        for (long s0 = 0; s0 < n_str; s0 += chunk) {
          const long cnt = std::min(chunk, n_str - s0);
          const long nz = n2 * cnt;
          hipLaunchKernelGGL(fill_zero_kernel,
                             dim3((unsigned)std::min<long>((nz + 255) / 256, 65536)), dim3(256), 0,
                             st, part, nz);
          const dim3 grid((unsigned)cnt);  // one string per block, its links over 4 waves
          if (cols)
            hipLaunchKernelGGL((trdm_links_kernel<T, true>), grid, dim3(256), 0, st, l, x,
                               (int)n_str, (int)n_other, norb, strs, dtab, (int)s0, (int)cnt,
                               part);
          else
            hipLaunchKernelGGL((trdm_links_kernel<T, false>), grid, dim3(256), 0, st, l, x,
                               (int)n_str, (int)n_other, norb, strs, dtab, (int)s0, (int)cnt,
                               part);
          hipLaunchKernelGGL(trdm_reduce_kernel, dim3((unsigned)n2), dim3(256), 0, st, part,
                             (int)cnt, rdm);
        }
      };
      sector(L, P, false, da, db, sa.d, out);
      if (transpose_b) {
        Cx<T>* lt = prog.scratch<Cx<T>>(tpoff);
        Cx<T>* pt = lt + da * db;
        const dim3 tg((unsigned)((db + 31) / 32), (unsigned)((da + 31) / 32));
        hipLaunchKernelGGL(transpose_kernel<T>, tg, dim3(256), 0, st, L, lt, (int)da, (int)db);
        hipLaunchKernelGGL(transpose_kernel<T>, tg, dim3(256), 0, st, P, pt, (int)da, (int)db);
        sector(lt, pt, false, db, da, sb.d, out + n2);
      } else {
        sector(L, P, true, db, da, sb.d, out + n2);
      }
    };
    if (dtype == FFSIM_C128)
      run(double{});
    else
      run(float{});
    HIP_CHECK(hipGetLastError());
    prog.finish();
  });
}

int ffsim_apply_diag_coulomb_sequence(void* vec, void* work, int dtype, int norb, int n_alpha,
                                      int n_beta, int n_terms, const double* rotations,
                                      const double* mats, const double* times,
                                      int z_representation, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    check_sizes(norb, n_alpha, n_beta);
    require(n_terms >= 0, FFSIM_ERR_VALUE, "n_terms must be non-negative");
    Context& ctx = context();
    const size_t nn = (size_t)norb * norb;
    std::vector<const cdouble*> rots(n_terms + 1);
    for (int t = 0; t <= n_terms; ++t)
      rots[t] = reinterpret_cast<const cdouble*>(rotations) + t * nn;
    std::vector<const double*> m(3 * n_terms);
    for (int t = 0; t < n_terms; ++t)
      for (int k = 0; k < 3; ++k) m[3 * t + k] = mats + (3 * t + k) * nn;
    diag_coulomb_sequence_impl(ctx, vec, work, dtype, norb, n_alpha, n_beta, n_terms, rots.data(),
                               m.data(), times, z_representation, (hipStream_t)stream);
  });
}

int ffsim_apply_givens_rotation_in_place(void* vec, int dtype, int64_t n_rows, int64_t n_cols,
                                         double c, double s_re, double s_im,
                                         const uint64_t* slice1, const uint64_t* slice2,
                                         int64_t n_pairs, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n_rows >= 0 && n_cols >= 0 && n_pairs >= 0, FFSIM_ERR_VALUE, "negative size");
    if (n_pairs == 0 || n_cols == 0) return;
    Context& ctx = context();
    hipStream_t st = (hipStream_t)stream;
    Program prog(ctx, st);
    const size_t o1 = append_indices(prog, slice1, n_pairs, n_rows, "slice1");
    const size_t o2 = append_indices(prog, slice2, n_pairs, n_rows, "slice2");
    prog.commit();
    PairRotationSet set{};
    set.r[0] = PairRotation{prog.dev<int64_t>(o1), prog.dev<int64_t>(o2), (long)n_pairs, c, s_re,
                            s_im};
    launch_pair_rotations(dtype, vec, n_rows, n_cols, true, set, 1, st);
    prog.finish();
  });
}

int ffsim_apply_phase_shift_in_place(void* vec, int dtype, int64_t n_rows, int64_t n_cols,
                                     double phase_re, double phase_im, const uint64_t* indices,
                                     int64_t n_indices, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n_rows >= 0 && n_cols >= 0 && n_indices >= 0, FFSIM_ERR_VALUE, "negative size");
    if (n_indices == 0 || n_cols == 0) return;
    Context& ctx = context();
    hipStream_t st = (hipStream_t)stream;
    Program prog(ctx, st);
    const size_t oi = append_indices(prog, indices, n_indices, n_rows, "indices");
    prog.commit();
    const unsigned grid = (unsigned)std::min<int64_t>(n_indices, 65535);
    if (dtype == FFSIM_C128)
      hipLaunchKernelGGL(phase_rows_kernel<double>, dim3(grid), dim3(256), 0, st,
                         (Cx<double>*)vec, (long)n_cols, prog.dev<int64_t>(oi), (long)n_indices,
                         phase_re, phase_im);
    else
      hipLaunchKernelGGL(phase_rows_kernel<float>, dim3(grid), dim3(256), 0, st,
                         (Cx<float>*)vec, (long)n_cols, prog.dev<int64_t>(oi), (long)n_indices,
                         phase_re, phase_im);
    HIP_CHECK(hipGetLastError());
    prog.finish();
  });
}

int ffsim_apply_givens_layer(void* vec, int dtype, int norb, int nocc, int64_t n_other,
                             int strings_on_rows, int n_rot, const double* c, const double* s,
                             const int32_t* i, const int32_t* j, void* stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(norb >= 0 && norb <= 63 && nocc >= 0 && n_other >= 0 && n_rot >= 0,
            FFSIM_ERR_VALUE, "bad sizes");
    require(n_rot <= kMaxLayerRotations, FFSIM_ERR_VALUE, "too many rotations in one layer");
    uint64_t used = 0;
    for (int r = 0; r < n_rot; ++r) {
      require(i[r] >= 0 && j[r] >= 0 && i[r] < norb && j[r] < norb, FFSIM_ERR_VALUE,
              "orbital index out of range");
      require(i[r] == j[r] + 1 || j[r] == i[r] + 1, FFSIM_ERR_ASSERT,
              "Target orbitals must be adjacent.");
      const uint64_t m = (1ull << i[r]) | (1ull << j[r]);
      require(!(used & m), FFSIM_ERR_VALUE, "rotations of a layer must act on disjoint orbitals");
      used |= m;
    }
    const long dim = (long)binom(norb, nocc);
    if (n_rot == 0 || dim == 0 || n_other == 0) return;
    require(dim < (1l << 31) && n_other < (1l << 31), FFSIM_ERR_VALUE, "dimension too large");
    Context& ctx = context();
    hipStream_t st = (hipStream_t)stream;
    // Two rotations of a layer act on disjoint orbitals but share strings (0101
    // pairs with 0110 under (0,1) and with 1001 under (2,3)), so they cannot run
    // as independent launches over the same rows.  The window passes apply them
    // in one fused pass; where those cannot run in place (strings on columns
    // with a string vector too long for one tile) each rotation is one launch.
    const bool fused = strings_on_rows || beta_in_columns(norb, dim, dtype, ctx);
    if (fused) {
      SectorRotation rot;
      for (int r = 0; r < n_rot; ++r) {
        KStep k{};
        const cdouble sv(s[2 * r], s[2 * r + 1]);
        // (c, s) on (first = i occupied / j empty, second = partner) as a
        // kernel step on (p occupied / p+1 empty, partner)
        const cdouble sigma = (j[r] == i[r] + 1) ? sv : -std::conj(sv);
        k.p = std::min(i[r], j[r]);
        k.c = c[r];
        k.sr = sigma.real();
        k.si = sigma.imag();
        rot.steps.push_back(k);
      }
      Program prog(ctx, st);
      std::vector<std::function<void()>> launches;
      if (strings_on_rows)
        run_sector(ctx, prog, dtype, MODE_ROWS, vec, (long)n_other, (int)n_other, norb, nocc, rot,
                   nullptr, 0, st, launches);
      else
        run_sector(ctx, prog, dtype, MODE_COLS, vec, dim, (int)n_other, norb, nocc, rot, nullptr,
                   0, st, launches);
      prog.commit();
      for (auto& f : launches) f();
      prog.finish();
      return;
    }
    for (int r = 0; r < n_rot; ++r) {
      const DevPairs& dp = dev_pairs(ctx, norb, nocc, i[r], j[r]);
      PairRotationSet set{};
      set.r[0] = PairRotation{dp.first, dp.second, dp.n, c[r], s[2 * r], s[2 * r + 1]};
      launch_pair_rotations(dtype, vec, (long)n_other, dim, false, set, 1, st);
    }
  });
}

}  // extern "C"
