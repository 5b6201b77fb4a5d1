// Allocation-free diagonal Coulomb evolution.
//
// ffsim_apply_diag_coulomb_evolution_noalloc computes what
// ffsim_apply_diag_coulomb_evolution computes, with the same kernels, but
// every device buffer beyond the caller's state (the operator's parameters,
// the alpha and beta string tables and the per-column phase factors) lives in
// one file-scope device array, `scratch`, instead of the library's program
// ring and table cache.  The call makes no hipMalloc / hipFree and no other
// allocating call: the parameters reach the device as kernel arguments and the
// string tables are generated on the device.
//
// Every GPU holds its own copy of `scratch` (kNoAllocScratchBytes) for the
// life of the process.  Calls on one GPU share it, so calls that overlap in
// time on one GPU (different streams, or different host threads) race; calls
// on one stream are ordered and safe.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ffsim_amd.h"
#include "diag_params.hpp"
#include "host_core.hpp"
#include "kernels.hpp"
#include "status.hpp"

namespace ffsim_amd {
namespace {

constexpr size_t kNoAllocScratchBytes = 8u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Per device: the device address of `scratch` (host code never uses the
// symbol's name as an address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// scratch layout: header and coefficients | alpha strings | beta strings | gamma
constexpr size_t kParamBytes = 128u << 10;
constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

unsigned char* scratch_for_current_device() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  require(dev >= 0 && dev < kMaxDevices, FFSIM_ERR_VALUE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[dev];
}

// Parameters travel as kernel arguments, in chunks of this many doubles.
constexpr int kChunkDoubles = 448;
struct ParamChunk {
  long off;  // destination offset in doubles
  int n;
  int pad;
  double v[kChunkDoubles];
};

__global__ void write_params_kernel(double* dst, ParamChunk ch) {
  for (int k = threadIdx.x; k < ch.n; k += blockDim.x) dst[ch.off + k] = ch.v[k];
}

// The string of every address of (norb, nocc): the colex unranking of
// addr = sum_t C(c_t, t + 1) over the set bits c_0 < c_1 < ..., built from a
// Pascal table in LDS (exact in 64 bits for norb <= 63).
__global__ void __launch_bounds__(256)
    unrank_strings_kernel(int norb, int nocc, long dim, uint64_t* __restrict__ out) {
  __shared__ uint64_t C[64][64];  // C[n][k], n <= norb, k <= nocc
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) C[e / 64][e % 64] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int n = 0; n <= norb; ++n) {
      C[n][0] = 1;
      for (int k = 1; k <= nocc && k <= n; ++k) C[n][k] = C[n - 1][k - 1] + (k < n ? C[n - 1][k] : 0);
    }
  }
  __syncthreads();
  const long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= dim) return;
  uint64_t r = (uint64_t)a, s = 0;
  int hi = norb - 1;
  for (int t = nocc; t >= 1; --t) {
    int c = t - 1;
    while (c + 1 <= hi && C[c + 1][t] <= r) ++c;
    s |= 1ull << c;
    r -= C[c][t];
    hi = c - 1;
  }
  out[a] = s;
}

void write_params(double* dst, const std::vector<double>& v, hipStream_t stream) {
  for (size_t k0 = 0; k0 < v.size(); k0 += kChunkDoubles) {
    ParamChunk ch{};
    ch.off = (long)k0;
    ch.n = (int)std::min<size_t>(kChunkDoubles, v.size() - k0);
    std::memcpy(ch.v, v.data() + k0, ch.n * sizeof(double));
    hipLaunchKernelGGL(write_params_kernel, dim3(1), dim3(256), 0, stream, dst, ch);
    HIP_CHECK(hipGetLastError());
  }
}

void unrank(int norb, int nocc, long dim, uint64_t* out, hipStream_t stream) {
  hipLaunchKernelGGL(unrank_strings_kernel, dim3((unsigned)((dim + 255) / 256)), dim3(256), 0,
                     stream, norb, nocc, dim, out);
  HIP_CHECK(hipGetLastError());
}

}  // namespace

// Called by ffsim_amd_init with the device current.
void noalloc_setup() { scratch_for_current_device(); }

}  // namespace ffsim_amd

using namespace ffsim_amd;

extern "C" int ffsim_apply_diag_coulomb_evolution_noalloc(
    void* vec, int dtype, int norb, int n_alpha, int n_beta, const double* mat_aa,
    const double* mat_ab, const double* mat_bb, double time, int z_representation, void* stream) {
  return guarded([&] {
    require(dtype == FFSIM_C128 || dtype == FFSIM_C64, FFSIM_ERR_VALUE, "unknown dtype");
    require(norb >= 0 && norb <= 63, FFSIM_ERR_VALUE, "norb must be in [0, 63]");
    require(n_alpha >= 0 && n_beta >= 0, FFSIM_ERR_VALUE, "electron counts must be non-negative");
    const long dim_a = (long)binom(norb, n_alpha), dim_b = (long)binom(norb, n_beta);
    const size_t off_a = kParamBytes;
    const size_t off_b = off_a + align256((size_t)dim_a * 8);
    const size_t off_g = off_b + align256((size_t)dim_b * 8);
    const size_t total = off_g + align256((size_t)dim_b * 16);
    require(total <= kNoAllocScratchBytes, FFSIM_ERR_VALUE,
            "state too large for the allocation-free entry point: its tables need " +
                std::to_string(total) + " bytes of the " + std::to_string(kNoAllocScratchBytes) +
                "-byte device scratch (use ffsim_apply_diag_coulomb_evolution)");
    DiagSpec spec = diag_coulomb_spec(norb, mat_aa, mat_ab, mat_bb, z_representation, time);
    if (!spec.flags || dim_a == 0 || dim_b == 0) return;
    require(diag_grid(dim_a, dim_b).y <= 65535, FFSIM_ERR_VALUE, "too many columns");
    unsigned char* base = scratch_for_current_device();
    hipStream_t st = (hipStream_t)stream;
    // header (4 doubles) then the coefficients, all written by kernels
    const DiagHeader h = diag_header(spec);
    static_assert(sizeof(DiagHeader) == 4 * sizeof(double), "header packs into 4 doubles");
    const std::vector<double> coef = diag_coefficients(spec);
    require((4 + coef.size()) * sizeof(double) <= kParamBytes, FFSIM_ERR_VALUE,
            "operator parameters exceed the scratch");
    std::vector<double> params(4 + coef.size());
    std::memcpy(params.data(), &h, sizeof(h));
    std::copy(coef.begin(), coef.end(), params.begin() + 4);
    double* dparams = reinterpret_cast<double*>(base);
    write_params(dparams, params, st);
    uint64_t* sa = reinterpret_cast<uint64_t*>(base + off_a);
    uint64_t* sb = reinterpret_cast<uint64_t*>(base + off_b);
    double* gamma = reinterpret_cast<double*>(base + off_g);
    unrank(norb, n_alpha, dim_a, sa, st);
    unrank(norb, n_beta, dim_b, sb, st);
    const DiagHeader* dh = reinterpret_cast<const DiagHeader*>(dparams);
    const double* dc = dparams + 4;
    launch_diag_col_term(dim_b, sb, dh, dc, gamma, 1, st);
    HIP_CHECK(hipGetLastError());
    if (dtype == FFSIM_C128)
      launch_diag_pass_t<double>(DIAG_EVOLVE, vec, vec, dim_a, dim_b, sa, sb, gamma, dh, dc, st);
    else
      launch_diag_pass_t<float>(DIAG_EVOLVE, vec, vec, dim_a, dim_b, sa, sb, gamma, dh, dc, st);
    HIP_CHECK(hipGetLastError());
  });
}
