// Device kernels of the ffsim_amd library (gfx950).
//
// fiber_kernel: applies a program of Givens rotations (and optionally the
//   diagonal phases of an orbital rotation) to every fiber of a window.  One
//   workgroup owns one fiber x W batch vectors, staged in LDS; the rotations
//   run as register sub-passes: each lane loads one group of strings that
//   agree outside a 6-orbital sub-window, applies that sub-window's rotations
//   in registers and writes the group back (a rotation pairs local strings
//   x, y that differ by moving one electron between adjacent orbitals).  The
//   last pass of a sector can store the state transposed and, with the
//   fused-diagonal instantiation, multiply by a diagonal Coulomb evolution.
// diag_kernel: one elementwise pass over the state for a diagonal operator
//   E(a, b) = row term + column term + cross term, used for diag-Coulomb and
//   number-operator evolution (psi *= exp(-i t E)) and contraction
//   (out (+)= s E psi).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "host_core.hpp"

namespace ffsim_amd {

template <typename T>
struct alignas(2 * sizeof(T)) Cx {
  T re, im;
};

enum { MODE_ROWS = 0, MODE_COLS = 1 };

// ---------------------------------------------------------------------------
// Compile-time pair lists of a sub-window: among the k2-subsets of w2 bits in
// ascending order (member index = rank), the pairs (x, y) where x has bit p
// set / p+1 clear and y = x with the electron moved from p to p+1.
struct SubPairList {
  int n;
  unsigned char x[16], y[16];
};

constexpr int popc_const(unsigned v) {
  int c = 0;
  while (v) {
    c += v & 1u;
    v >>= 1;
  }
  return c;
}

constexpr int rank_const(unsigned s, int w2, int k2) {
  int r = 0;
  for (unsigned t = 0; t < s; ++t)
    if (popc_const(t) == k2) ++r;
  (void)w2;
  return r;
}

template <int W2, int K2, int P>
constexpr SubPairList make_sub_pairs() {
  SubPairList pl{};
  pl.n = 0;
  int r = 0;
  for (unsigned s = 0; s < (1u << W2); ++s) {
    if (popc_const(s) != K2) continue;
    if (((s >> P) & 1u) && !((s >> (P + 1)) & 1u)) {
      pl.x[pl.n] = (unsigned char)r;
      pl.y[pl.n] = (unsigned char)rank_const(s ^ (3u << P), W2, K2);
      ++pl.n;
    }
    ++r;
  }
  return pl;
}

constexpr int binom_const(int n, int k) {
  if (k < 0 || k > n) return 0;
  int r = 1;
  for (int t = 1; t <= k; ++t) r = r * (n - k + t) / t;
  return r;
}

// x' = c x + s y, y' = c y - conj(s) x.  The terms of y' that need the old x are
// formed first, so that both outputs can overwrite their inputs in place (no
// register copies when the update sits on a conditional path).
template <typename T>
__device__ __forceinline__ void rot_pair(Cx<T>& a, Cx<T>& b, T c, T sr, T si) {
  const T ur = sr * a.re + si * a.im;
  const T ui = sr * a.im - si * a.re;
  a.re = c * a.re + (sr * b.re - si * b.im);
  a.im = c * a.im + (sr * b.im + si * b.re);
  b.re = c * b.re - ur;
  b.im = c * b.im - ui;
}

// FP64: the same twelve operations as AMD instructions whose outputs are tied to
// their inputs, so both elements are updated in their own registers.  Left to
// itself the compiler renames them and, on the conditional slot paths, copies
// them back with v_mov_b64 (2 copies per pair, ~17% of the VALU stream).
template <>
__device__ __forceinline__ void rot_pair<double>(Cx<double>& a, Cx<double>& b, double c, double sr,
                                                 double si) {
  double ur, ui, t1, t2;
  asm("v_mul_f64 %0, %1, %2" : "=v"(ur) : "s"(sr), "v"(a.re));
  asm("v_fma_f64 %0, %1, %2, %0" : "+v"(ur) : "s"(si), "v"(a.im));
  asm("v_mul_f64 %0, %1, %2" : "=v"(ui) : "s"(sr), "v"(a.im));
  asm("v_fma_f64 %0, -%1, %2, %0" : "+v"(ui) : "s"(si), "v"(a.re));
  asm("v_mul_f64 %0, %1, %2" : "=v"(t1) : "s"(si), "v"(b.im));
  asm("v_fma_f64 %0, %1, %2, -%0" : "+v"(t1) : "s"(sr), "v"(b.re));
  asm("v_mul_f64 %0, %1, %2" : "=v"(t2) : "s"(si), "v"(b.re));
  asm("v_fma_f64 %0, %1, %2, %0" : "+v"(t2) : "s"(sr), "v"(b.im));
  asm("v_fma_f64 %0, %1, %0, %2" : "+v"(a.re) : "s"(c), "v"(t1));
  asm("v_fma_f64 %0, %1, %0, %2" : "+v"(a.im) : "s"(c), "v"(t2));
  asm("v_fma_f64 %0, %1, %0, -%2" : "+v"(b.re) : "s"(c), "v"(ur));
  asm("v_fma_f64 %0, %1, %0, -%2" : "+v"(b.im) : "s"(c), "v"(ui));
}

// The same with Re(sigma) in a VGPR: lanes of complemented groups run with the
// negated value (see group_apply).
__device__ __forceinline__ void rot_pair_vsr(Cx<double>& a, Cx<double>& b, double c, double sr,
                                             double si) {
  double ur, ui, t1, t2;
  asm("v_mul_f64 %0, %1, %2" : "=v"(ur) : "v"(sr), "v"(a.re));
  asm("v_fma_f64 %0, %1, %2, %0" : "+v"(ur) : "s"(si), "v"(a.im));
  asm("v_mul_f64 %0, %1, %2" : "=v"(ui) : "v"(sr), "v"(a.im));
  asm("v_fma_f64 %0, -%1, %2, %0" : "+v"(ui) : "s"(si), "v"(a.re));
  asm("v_mul_f64 %0, %1, %2" : "=v"(t1) : "s"(si), "v"(b.im));
  asm("v_fma_f64 %0, %1, %2, -%0" : "+v"(t1) : "v"(sr), "v"(b.re));
  asm("v_mul_f64 %0, %1, %2" : "=v"(t2) : "s"(si), "v"(b.re));
  asm("v_fma_f64 %0, %1, %2, %0" : "+v"(t2) : "v"(sr), "v"(b.im));
  asm("v_fma_f64 %0, %1, %0, %2" : "+v"(a.re) : "s"(c), "v"(t1));
  asm("v_fma_f64 %0, %1, %0, %2" : "+v"(a.im) : "s"(c), "v"(t2));
  asm("v_fma_f64 %0, %1, %0, -%2" : "+v"(b.re) : "s"(c), "v"(ur));
  asm("v_fma_f64 %0, %1, %0, -%2" : "+v"(b.im) : "s"(c), "v"(ui));
}
// FP32 (complex64 states): the same in-place form with v_fma_f32 / v_mul_f32.
template <>
__device__ __forceinline__ void rot_pair<float>(Cx<float>& a, Cx<float>& b, float c, float sr,
                                                float si) {
  float ur, ui, t1, t2;
  asm("v_mul_f32 %0, %1, %2" : "=v"(ur) : "v"(sr), "v"(a.re));
  asm("v_fma_f32 %0, %1, %2, %0" : "+v"(ur) : "v"(si), "v"(a.im));
  asm("v_mul_f32 %0, %1, %2" : "=v"(ui) : "v"(sr), "v"(a.im));
  asm("v_fma_f32 %0, -%1, %2, %0" : "+v"(ui) : "v"(si), "v"(a.re));
  asm("v_mul_f32 %0, %1, %2" : "=v"(t1) : "v"(si), "v"(b.im));
  asm("v_fma_f32 %0, %1, %2, -%0" : "+v"(t1) : "v"(sr), "v"(b.re));
  asm("v_mul_f32 %0, %1, %2" : "=v"(t2) : "v"(si), "v"(b.re));
  asm("v_fma_f32 %0, %1, %2, %0" : "+v"(t2) : "v"(sr), "v"(b.im));
  asm("v_fma_f32 %0, %1, %0, %2" : "+v"(a.re) : "v"(c), "v"(t1));
  asm("v_fma_f32 %0, %1, %0, %2" : "+v"(a.im) : "v"(c), "v"(t2));
  asm("v_fma_f32 %0, %1, %0, -%2" : "+v"(b.re) : "v"(c), "v"(ur));
  asm("v_fma_f32 %0, %1, %0, -%2" : "+v"(b.im) : "v"(c), "v"(ui));
}
// c64 passes every coefficient in a VGPR already.
__device__ __forceinline__ void rot_pair_vsr(Cx<float>& a, Cx<float>& b, float c, float sr,
                                             float si) {
  rot_pair<float>(a, b, c, sr, si);
}

template <typename T, int W2, int K2, int P>
__device__ __forceinline__ void rot_sub(Cx<T>* v, T c, T sr, T si) {
  if constexpr (P + 1 < W2) {
    constexpr SubPairList pl = make_sub_pairs<W2, K2, P>();
#pragma unroll
    for (int q = 0; q < pl.n; ++q) rot_pair(v[pl.x[q]], v[pl.y[q]], c, sr, si);
  }
}

template <typename T, int W2, int K2, int P>
__device__ __forceinline__ void rot_sub_vsr(Cx<T>* v, T c, T sr, T si) {
  if constexpr (P + 1 < W2) {
    constexpr SubPairList pl = make_sub_pairs<W2, K2, P>();
#pragma unroll
    for (int q = 0; q < pl.n; ++q) rot_pair_vsr(v[pl.x[q]], v[pl.y[q]], c, sr, si);
  }
}

// One slot of a layer pair.  A complemented group (flip) applies the rotation
// with Re(sigma) negated: its members pair up with x and y exchanged, and
// x' = c x + s y, y' = c y - conj(s) x with x <-> y is the same update with
// s -> -conj(s).  Self-complementary classes (2 K2 == W2) are never flipped,
// and c128 tables have no complemented groups (Re(sigma) stays in SGPRs).
template <typename T, int W2, int K2, int P>
__device__ __forceinline__ void slot_apply(Cx<T>* v, const double* c, const double* sr,
                                           const double* si, int mask, bool flip) {
  if constexpr (P + 1 < W2) {
    if (mask & (1 << P)) {
      if constexpr (2 * K2 == W2 || sizeof(T) == 8) {
        rot_sub<T, W2, K2, P>(v, (T)c[P], (T)sr[P], (T)si[P]);
      } else {
        const T s = flip ? -(T)sr[P] : (T)sr[P];
        rot_sub_vsr<T, W2, K2, P>(v, (T)c[P], s, (T)si[P]);
      }
    }
  }
}

// Applies the layer pairs of one sub-pass to one register group (one batch column).
template <typename T, int W2, int K2>
__device__ __forceinline__ void group_apply(Cx<T>* tile, int S, int j,
                                            const uint16_t* __restrict__ members,
                                            const LayerPair* lp, int n_lp, bool flip) {
  constexpr int N = binom_const(W2, K2);
  Cx<T> v[N];
  int idx[N];
#pragma unroll
  for (int m = 0; m < N; ++m) idx[m] = (int)members[m] * S + j;
#pragma unroll
  for (int m = 0; m < N; ++m) v[m] = tile[idx[m]];
  for (int l = 0; l < n_lp; ++l) {
    // one batch of scalar loads for the whole layer pair (uniform over the wave)
    const LayerPair& L = lp[l];
    const int mask = __builtin_amdgcn_readfirstlane(L.mask);
    double c[kSlots], sr[kSlots], si[kSlots];
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
      c[q] = L.c[q];
      sr[q] = L.sr[q];
      si[q] = L.si[q];
    }
    // all fifteen coefficients in SGPRs here: otherwise the compiler sinks
    // each slot's loads into that slot's branch, and the wave waits for the
    // scalar cache once per slot instead of once per layer pair (measured
    // 0.7% per complex128 launch and 0.8% per complex64 launch, same bits)
    asm volatile("" ::"s"(c[0]), "s"(sr[0]), "s"(si[0]), "s"(c[1]), "s"(sr[1]), "s"(si[1]),
                 "s"(c[2]), "s"(sr[2]), "s"(si[2]), "s"(c[3]), "s"(sr[3]), "s"(si[3]),
                 "s"(c[4]), "s"(sr[4]), "s"(si[4]));
    slot_apply<T, W2, K2, 0>(v, c, sr, si, mask, flip);
    slot_apply<T, W2, K2, 2>(v, c, sr, si, mask, flip);
    slot_apply<T, W2, K2, 4>(v, c, sr, si, mask, flip);
    slot_apply<T, W2, K2, 1>(v, c, sr, si, mask, flip);
    slot_apply<T, W2, K2, 3>(v, c, sr, si, mask, flip);
  }
#pragma unroll
  for (int m = 0; m < N; ++m) tile[idx[m]] = v[m];
}

template <typename T, int W2>
__device__ __forceinline__ void group_dispatch(int k2, bool flip, Cx<T>* tile, int S, int j,
                                               const uint16_t* members, const LayerPair* lp,
                                               int n_lp) {
  switch (k2) {
    case 1:
      if constexpr (W2 >= 2) group_apply<T, W2, 1>(tile, S, j, members, lp, n_lp, flip);
      break;
    case 2:
      if constexpr (W2 >= 3) group_apply<T, W2, 2>(tile, S, j, members, lp, n_lp, flip);
      break;
    case 3:
      if constexpr (W2 >= 4) group_apply<T, W2, 3>(tile, S, j, members, lp, n_lp, flip);
      break;
    case 4:
      if constexpr (W2 >= 5) group_apply<T, W2, 4>(tile, S, j, members, lp, n_lp, flip);
      break;
    case 5:
      if constexpr (W2 >= 6) group_apply<T, W2, 5>(tile, S, j, members, lp, n_lp, flip);
      break;
    default:
      break;
  }
}

// Program of one window pass, as stored in device memory:
//   SubPass subs[n_sub]; LayerPair layers[n_layers] (slots relative to each lo2).
struct GroupArgs {
  const uint32_t* desc;
  const uint16_t* members;
  const int32_t* index;
  int n_lo2, w2;
  const uint32_t* chunks;   // wave schedule (see WaveSchedule)
  const int32_t* wave_idx;
};

// Distinct sub-window starts of a pass (a SubPass's `slot` indexes these).
constexpr int kMaxLoSlots = 8;
struct LoSlots {
  int n;
  int lo2[kMaxLoSlots];
};

// Per-block view of one sub-window start's group table, staged in LDS.
struct GroupSlot {
  int d_off;  // first descriptor in LDS
  int nd;     // number of groups
  int pad[2];
};

__host__ __device__ constexpr size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// A diagonal Coulomb evolution fused into the store of a transposing pass
// (the last pass of the second sector's rotation).  Element (tile string i,
// batch string j) is multiplied by exp(-i t E) formed from complex FP64
// factors: rfac of string i, the batch term bfac[j][norb], and bfac[j][q] for
// every occupied orbital q of string i.
struct DiagFuse {
  const Cx<double>* bfac;  // [n_batch][n_fac]: per orbital q < norb, then the batch term
  const Cx<double>* rfac;  // [strings of the tile sector]
  int n_fac;               // norb + 1 (at most 64)
  int win;                 // the pass's window is orbitals [0, win), win <= 16
};
// LDS of the fused store's per-tile tables: for every 4-orbital nibble t of the
// window and every occupation pattern p of it, the product of the batch
// columns' factors over p; then one factor per column for the batch term and
// the fiber's orbitals outside the window (the same for every tile string).
__host__ __device__ constexpr int fuse_nibble_entries(int win, int W) {
  return (((win + 3) >> 2) * 16 + 1) * W;
}

// Several states of one shape in one launch (the batched Givens path): tile t
// covers fiber t / (n_states n_bt) of state (t / n_bt) % n_states, so the
// tiles of one fiber stay together across states and share the staged
// tables.  Each state is rotated in place at its own address, with its own
// layer-pair coefficients and orbital phases at the given strides.
struct MultiTiles {
  void* const* states;  // device array of the states' addresses
  int n_states;
  long layer_stride;  // LayerPairs between consecutive states' coefficients
  long diag_stride;   // doubles between consecutive states' phases
};

// Element (string i, batch j) of a fiber tile lives at
//   MODE_ROWS: state[row_table[i] * ld + b0 + j]   (strings index rows)
//   MODE_COLS: state[(b0 + j) * ld + row_table[i]] (strings index columns)
// tstore (MODE_ROWS only): write element (i, j) to dst[(b0 + j) * ld_dst +
// row_table[i]] instead, i.e. store the transposed state.
//
// Persistent: block b handles tiles b, b + gridDim.x, ... in fiber-major
// order (tile = fiber * n_bt + batch tile), so consecutive tiles of a block
// share their fiber and its staged tables.  While a tile is transformed in
// LDS, the next tile's loads are already in flight into registers.
//
// LDS: tile[max_rows * S] | rows[max_rows] | slots | desc | members | wave schedule
template <typename T, int W, int MODE, int NTHR, int W2, bool FUSE = false, bool MULTI = false>
__global__ void __launch_bounds__(NTHR)
    fiber_kernel(const Cx<T>* src, Cx<T>* dst, long ld_src, long ld_dst, int n_batch, int tstore,
                 const Fiber* __restrict__ fibers, int n_fibers,
                 const int32_t* __restrict__ row_table, GroupArgs ga, LoSlots los,
                 const SubPass* __restrict__ subs_g, int n_sub,
                 const LayerPair* __restrict__ layers_g, int n_layers,
                 const double* __restrict__ diag, const uint64_t* __restrict__ strings,
                 int n_diag, int max_rows, int gd_cap, int gm_cap, DiagFuse fz, MultiTiles mt) {
  static_assert(!FUSE || MODE == MODE_ROWS, "the fused diagonal store is a row-mode store");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // LDS stride of one string's batch.  Row mode: unpadded (the register
  // sub-passes read 8 columns of one string per 8 lanes, and a padded stride
  // puts the strings of neighbouring lane groups on the same banks: 4-5%
  // slower measured).  Column mode, whose lanes walk strings: padded.
  constexpr int S = (MODE == MODE_COLS && W >= 8) ? W + 1 : W;
  constexpr int RSTEP = NTHR / W;          // rows per sweep in MODE_ROWS
  constexpr int CH = 16;                   // register chunk of one thread
  size_t off = 0;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(smem_raw);
  off += align16((size_t)max_rows * S * sizeof(Cx<T>));
  int32_t* rows = reinterpret_cast<int32_t*>(smem_raw + off);
  off += align16((size_t)max_rows * sizeof(int32_t));
  GroupSlot* gslot = reinterpret_cast<GroupSlot*>(smem_raw + off);
  off += (size_t)kMaxLoSlots * sizeof(GroupSlot);
  uint32_t* gdesc = reinterpret_cast<uint32_t*>(smem_raw + off);
  off += align16((size_t)gd_cap * sizeof(uint32_t));
  uint16_t* gmem = reinterpret_cast<uint16_t*>(smem_raw + off);
  off += align16((size_t)gm_cap * sizeof(uint16_t));
  // wave schedule of the staged fiber: per slot and wave (chunk offset, count), then chunks
  constexpr int NW = NTHR / 64;
  int32_t* wsched = reinterpret_cast<int32_t*>(smem_raw + off);
  off += (size_t)kMaxLoSlots * NW * 2 * sizeof(int32_t);
  uint32_t* wchunks = reinterpret_cast<uint32_t*>(smem_raw + off);
  off += align16((size_t)gd_cap * sizeof(uint32_t));
  // Passes that apply orbital phases (diag != nullptr) also hold the phases and
  // the fiber's row strings (32-bit when norb <= 32) in LDS, so that the store
  // phase issues no global loads: a global load there would wait for every
  // store issued before it.  Only those launches allocate this region.
  double* dph = reinterpret_cast<double*>(smem_raw + off);
  off += (size_t)2 * kMaxDiagOrbitals * sizeof(double);
  uint32_t* rstr32 = reinterpret_cast<uint32_t*>(smem_raw + off);
  uint64_t* rstr64 = reinterpret_cast<uint64_t*>(smem_raw + off);
  const bool narrow_str = n_diag <= 32;
  // fused diagonal: the batch strings' factors of the current tile, with the
  // orbital phases of the rotation folded into the per-orbital ones
  Cx<double>* ffac = reinterpret_cast<Cx<double>*>(
      smem_raw + off + align16((size_t)max_rows * (narrow_str ? 4 : 8)));
  const bool have_str = diag != nullptr || FUSE;
  Cx<double>* ntab = ffac + (FUSE ? W * fz.n_fac : 0);  // fused store tables

  const int tid = threadIdx.x;
  const int n_bt = (n_batch + W - 1) / W;
  const int NS = MULTI ? mt.n_states : 1;
  const long per_f = (long)n_bt * NS;  // tiles of one fiber (over all states)
  const long n_tiles = per_f * n_fibers;
  // the state of tile t and its address (one state: the src / dst arguments)
  auto state_of = [&](long tt) -> int { return MULTI ? (int)((tt / n_bt) % NS) : 0; };
  auto src_of = [&](int st) -> const Cx<T>* {
    if constexpr (MULTI) return reinterpret_cast<const Cx<T>*>(mt.states[st]);
    else return src;
  };
  auto dst_of = [&](int st) -> Cx<T>* {
    if constexpr (MULTI) return reinterpret_cast<Cx<T>*>(mt.states[st]);
    else return dst;
  };


  // stage the row addresses and group tables of fiber f
  // The group tables and wave schedule depend only on the fiber's electron count
  // in the window (kp); `tables` is false when the previous fiber had the same kp.
  // fused diagonal: the tile strings' factors of this thread's store rows
  // (i = tid + k NTHR; fibers of a fused pass hold at most 1024 strings), fixed
  // while the fiber is
  constexpr int RF = FUSE ? (1024 + NTHR - 1) / NTHR : 1;
  Cx<double> rf[RF];
#pragma unroll
  for (int k = 0; k < RF; ++k) rf[k] = Cx<double>{1.0, 0.0};
  auto stage_fiber = [&](const Fiber& fd, bool tables) {
    for (int i = tid; i < fd.n_rows; i += NTHR) rows[i] = row_table[fd.row_off + i];
    if (FUSE) {
#pragma unroll
      for (int k = 0; k < RF; ++k) {
        const int i = tid + k * NTHR;
        rf[k] = i < fd.n_rows ? fz.rfac[row_table[fd.row_off + i]] : Cx<double>{1.0, 0.0};
      }
    }
    if (have_str) {
      for (int i = tid; i < fd.n_rows; i += NTHR) {
        const uint64_t sv = strings[rows[i]];
        if (narrow_str)
          rstr32[i] = (uint32_t)sv;
        else
          rstr64[i] = sv;
      }
    }
    if (n_sub > 0 && tables) {
      int d_base = 0, m_base = 0;
      for (int u = 0; u < los.n; ++u) {
        const int* ix = ga.index + (fd.kp * ga.n_lo2 + los.lo2[u]) * 4;
        const int d0 = ix[0], nd = ix[1], m0 = ix[2], nm = ix[3];
        if (tid == 0) gslot[u] = GroupSlot{d_base, nd, {0, 0}};
        for (int k = tid; k < nd; k += NTHR) {
          const uint32_t d = ga.desc[d0 + k];
          gdesc[d_base + k] = ((((d >> 8) - (uint32_t)m0) + (uint32_t)m_base) << 8) | (d & 0xff);
        }
        for (int k = tid; k < nm; k += NTHR) gmem[m_base + k] = ga.members[m0 + k];
        d_base += nd;
        m_base += nm;
      }
      // wave schedule: the chunks of one (kp, lo2) are contiguous over its waves
      int c_base = 0;
      for (int u = 0; u < los.n; ++u) {
        const int32_t* wi = ga.wave_idx + ((fd.kp * ga.n_lo2 + los.lo2[u]) * NW) * 2;
        const int first = wi[0];
        int total = 0;
        for (int wv = 0; wv < NW; ++wv) total += wi[2 * wv + 1];
        if (tid < NW) {
          wsched[(u * NW + tid) * 2] = c_base + (wi[2 * tid] - first);
          wsched[(u * NW + tid) * 2 + 1] = wi[2 * tid + 1];
        }
        for (int k = tid; k < total; k += NTHR) wchunks[c_base + k] = ga.chunks[first + k];
        c_base += total;
      }
    }
  };
  // issue every global load of tile (nr rows, batch offset b0, nb columns) into buf
  // Row addresses come from `rt`: the staged table in LDS (next tile in the same
  // fiber) or the global row table of the next fiber.  The two are separate
  // instantiations: a pointer that may be either would compile to flat loads,
  // and waiting for those (vmcnt) would also drain the stores in flight.
  // All row addresses are read first and the element loads issued after them, so
  // the global-table case waits once for its addresses instead of once per row.
  auto load_regs = [&](Cx<T>* buf, int nr, int b0, int nb, const int32_t* rt, const Cx<T>* sp) {
    int ra[CH];
    if (MODE == MODE_ROWS) {
      const int j = tid % W, i0 = tid / W;
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int i = i0 + k * RSTEP;
        ra[k] = (i < nr && j < nb) ? rt[i] : -1;
      }
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        buf[k] = Cx<T>{T(0), T(0)};
        if (ra[k] >= 0) buf[k] = sp[(long)ra[k] * ld_src + b0 + j];
      }
    } else {
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int i = tid + (k / W) * NTHR, j = k % W;
        ra[k] = (i < nr && j < nb) ? rt[i] : -1;
      }
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int j = k % W;
        buf[k] = Cx<T>{T(0), T(0)};
        if (ra[k] >= 0) buf[k] = sp[(long)(b0 + j) * ld_src + ra[k]];
      }
    }
  };
  auto regs_to_lds = [&](const Cx<T>* buf, int nr) {
    if (MODE == MODE_ROWS) {
      const int j = tid % W, i0 = tid / W;
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int i = i0 + k * RSTEP;
        if (i < nr) tile[i * S + j] = buf[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int i = tid + (k / W) * NTHR, j = k % W;
        if (i < nr) tile[i * S + j] = buf[k];
      }
    }
  };
  // rows [first, nr) straight from global memory into LDS
  constexpr int kCovered = (MODE == MODE_ROWS) ? CH * RSTEP : (CH / W) * NTHR;
  auto load_rows = [&](int first, int nr, int b0, int nb, const Cx<T>* sp) {
    if (nr <= first) return;
    if (MODE == MODE_ROWS) {
      const int j = tid % W, i0 = tid / W;
#pragma unroll 4
      for (int i = first + i0; i < nr; i += RSTEP)
        tile[i * S + j] = (j < nb) ? sp[(long)rows[i] * ld_src + b0 + j] : Cx<T>{T(0), T(0)};
    } else {
      for (int i = first + tid; i < nr; i += NTHR)
        for (int j = 0; j < W; ++j)
          tile[i * S + j] = (j < nb) ? sp[(long)(b0 + j) * ld_src + rows[i]] : Cx<T>{T(0), T(0)};
    }
  };
  auto row_phase = [&](int i, T& pr, T& pim) {
    double re = 1.0, im = 0.0;
    uint64_t bits = narrow_str ? (uint64_t)rstr32[i] : rstr64[i];
    while (bits) {
      const int o = __builtin_ctzll(bits);
      bits &= bits - 1;
      const double dr = dph[2 * o], di = dph[2 * o + 1];
      const double t = re * dr - im * di;
      im = re * di + im * dr;
      re = t;
    }
    pr = (T)re;
    pim = (T)im;
  };
  auto store_tile = [&](int nr, int b0, int nb, Cx<T>* dp) {
    if (MODE == MODE_ROWS && !tstore) {
      const int j = tid % W, i0 = tid / W;
      if (j < nb) {
#pragma unroll 4
        for (int i = i0; i < nr; i += RSTEP) {
          Cx<T> v = tile[i * S + j];
          if (diag) {
            T pr, pim;
            row_phase(i, pr, pim);
            const T t = v.re * pr - v.im * pim;
            v.im = v.re * pim + v.im * pr;
            v.re = t;
          }
          dp[(long)rows[i] * ld_dst + b0 + j] = v;
        }
      }
    } else if (FUSE) {
      // orbital phases and the diagonal evolution.  Tables first: products of
      // the batch columns' factors over each window nibble's patterns, and the
      // batch term times the outside orbitals' factors (shared by the fiber).
      auto cmul = [](const Cx<double>& x, const Cx<double>& y) {
        return Cx<double>{x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re};
      };
      const int nt = (fz.win + 3) >> 2;
      const uint64_t outside = (narrow_str ? (uint64_t)rstr32[0] : rstr64[0]) &
                               ~((1ull << fz.win) - 1);
      for (int e = tid; e < (nt * 16 + 1) * W; e += NTHR) {
        const int j = e % W, k = e / W;
        const Cx<double>* F = ffac + j * fz.n_fac;
        Cx<double> P{1.0, 0.0};
        if (k < nt * 16) {
          const int t4 = 4 * (k >> 4), pat = k & 15;
#pragma unroll
          for (int b = 0; b < 4; ++b)
            if (((pat >> b) & 1) && t4 + b < fz.win) P = cmul(P, F[t4 + b]);
        } else {
          P = F[fz.n_fac - 1];
          uint64_t bits = outside;
          while (bits) {
            const int q = __builtin_ctzll(bits);
            bits &= bits - 1;
            P = cmul(P, F[q]);
          }
        }
        ntab[k * W + j] = P;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < RF; ++k) {
        const int i = tid + k * NTHR;
        if (i >= nr) break;
        const Cx<double> r0f = rf[k];
        const uint32_t sv = (uint32_t)(narrow_str ? (uint64_t)rstr32[i] : rstr64[i]);
        Cx<double> P[W];
#pragma unroll
        for (int j = 0; j < W; ++j) P[j] = cmul(r0f, ntab[nt * 16 * W + j]);
        for (int t = 0; t < nt; ++t) {
          const int pat = (sv >> (4 * t)) & 15;
#pragma unroll
          for (int j = 0; j < W; ++j) P[j] = cmul(P[j], ntab[(t * 16 + pat) * W + j]);
        }
        const long r = rows[i];
#pragma unroll
        for (int j = 0; j < W; ++j) {
          if (j < nb) {
            Cx<T> v = tile[i * S + j];
            const T t = v.re * (T)P[j].re - v.im * (T)P[j].im;
            v.im = v.re * (T)P[j].im + v.im * (T)P[j].re;
            v.re = t;
            dp[(long)(b0 + j) * ld_dst + r] = v;
          }
        }
      }
    } else {
      for (int i = tid; i < nr; i += NTHR) {
        T pr = T(1), pim = T(0);
        if (diag) row_phase(i, pr, pim);
        const long r = rows[i];
#pragma unroll
        for (int j = 0; j < W; ++j) {
          if (j < nb) {
            Cx<T> v = tile[i * S + j];
            if (diag) {
              const T t = v.re * pr - v.im * pim;
              v.im = v.re * pim + v.im * pr;
              v.re = t;
            }
            dp[(long)(b0 + j) * ld_dst + r] = v;
          }
        }
      }
    }
  };

  // fused diagonal: thread t < W * n_fac holds factor q = t % n_fac of batch
  // column j = t / n_fac (read from a clamped, always valid address)
  auto load_bfac = [&](int b0x) -> Cx<double> {
    if (!FUSE) return Cx<double>{1.0, 0.0};
    const int t = min(tid, W * fz.n_fac - 1);
    const int j = t / fz.n_fac, q = t % fz.n_fac;
    return fz.bfac[(long)min(b0x + j, n_batch - 1) * fz.n_fac + q];
  };
  auto bfac_to_lds = [&](Cx<double> v, int b0x) {
    if (!FUSE || tid >= W * fz.n_fac) return;
    const int j = tid / fz.n_fac, q = tid % fz.n_fac;
    if (b0x + j >= n_batch) v = Cx<double>{1.0, 0.0};
    if (diag && q < fz.n_fac - 1) {
      const double dr = dph[2 * q], di = dph[2 * q + 1];
      v = Cx<double>{v.re * dr - v.im * di, v.re * di + v.im * dr};
    }
    ffac[tid] = v;
  };

  // XCD-aware order: workgroups are dealt round-robin over the 8 XCDs, so block
  // b runs on XCD b % 8.  Renumbering makes each XCD's blocks take consecutive
  // tiles, i.e. neighbouring batch tiles of one fiber, whose rows share the
  // 128-byte lines that straddle tile borders; those lines then hit in that
  // XCD's L2 instead of being fetched twice from HBM.
  const int G = gridDim.x;
  const int bid = (G & 7) == 0 ? (int)(blockIdx.x & 7) * (G >> 3) + (int)(blockIdx.x >> 3)
                               : (int)blockIdx.x;
  long t = bid;
  if (t >= n_tiles) return;
  if (diag)
    for (int k = tid; k < 2 * n_diag; k += NTHR)
      dph[k] = diag[(MULTI ? (long)state_of(t) * mt.diag_stride : 0) + k];
  int f = (int)(t / per_f);
  Fiber fd = fibers[f];
  stage_fiber(fd, true);
  __syncthreads();
  Cx<T> buf[CH];
  Cx<double> bfn;
  {
    const int b0 = (int)(t % n_bt) * W;
    const int nb = min(W, n_batch - b0);
    load_rows(0, fd.n_rows, b0, nb, src_of(state_of(t)));
    if (FUSE) bfac_to_lds(load_bfac(b0), b0);
  }
  while (true) {
    const int nr = fd.n_rows;
    const int b0 = (int)(t % n_bt) * W;
    const int nb = min(W, n_batch - b0);
    const long tn = t + gridDim.x;
    const int fn = tn < n_tiles ? (int)(tn / per_f) : -1;
    const int st = state_of(t);
    const int stn = tn < n_tiles ? state_of(tn) : st;
    const bool same_fiber = (fn == f);
    // the next tile is prefetched also when it starts a new fiber: its rows then
    // come from the global row table (one more dependent load, hidden by the
    // sub-passes below) instead of the staged table
    const Fiber fnext = fibers[fn >= 0 ? fn : f];  // a uniform (scalar) load
    const Fiber fdn = same_fiber ? fd : fnext;
    const int nrn = fn >= 0 ? fdn.n_rows : 0;
    const int b0n = (int)(tn % n_bt) * W;
    const int nbn = fn >= 0 ? min(W, n_batch - b0n) : 0;
    // unconditional (no loads after the last tile) so that buf is never carried
    // around the loop: a loop-carried buf makes the compiler copy it at the back
    // edge behind a vmcnt(0) that would also drain this tile's stores
    if (same_fiber)
      load_regs(buf, nrn, b0n, nbn, rows, src_of(stn));
    else
      load_regs(buf, nrn, b0n, nbn, row_table + fdn.row_off, src_of(stn));
    if (FUSE) bfn = load_bfac(b0n);

    // ---- register sub-passes on the current tile
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int lane = tid % 64;
    for (int sp = 0; sp < n_sub; ++sp) {
      __syncthreads();
      // the pass program is uniform: scalar loads straight from device memory
      const SubPass P = subs_g[sp];
      const GroupSlot gs = gslot[P.slot];
      const LayerPair* lp = layers_g + (MULTI ? (long)st * mt.layer_stride : 0) + P.lp0;
      const int c0 = wsched[(P.slot * NW + wave) * 2];
      const int nc = wsched[(P.slot * NW + wave) * 2 + 1];
      for (int c = 0; c < nc; ++c) {
        const uint32_t ch = wchunks[c0 + c];
        const int gl = lane / W, j = lane % W;
        if (gl < (int)(ch & 0xff)) {
          const uint32_t d = gdesc[gs.d_off + (int)(ch >> 8) + gl];
          // c64 tables hold complemented groups (bit 7); c128 tables never do
          if constexpr (sizeof(T) == 4)
            group_dispatch<T, W2>(d & 0x7f, (d >> 7) & 1, tile, S, j, gmem + (d >> 8), lp,
                                  P.n_lp);
          else
            group_dispatch<T, W2>(d & 0xff, false, tile, S, j, gmem + (d >> 8), lp, P.n_lp);
        }
      }
    }
    // Only the prefetch loads are outstanding here.  Waiting for them now lets the
    // stores below stay in flight while the prefetched tile is written to LDS
    // (vmcnt counts stores too: a wait placed after them would drain them).
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt/lgkmcnt unchanged
    __syncthreads();
    store_tile(nr, b0, nb, dst_of(st));
    __syncthreads();
    if (tn >= n_tiles) break;
    // the next tile's state: its orbital phases (read by the next store, after
    // several barriers)
    if (MULTI && diag && stn != st)
      for (int k = tid; k < 2 * n_diag; k += NTHR) dph[k] = diag[(long)stn * mt.diag_stride + k];
    if (!same_fiber) {
      f = fn;
      const bool tables = fdn.kp != fd.kp;
      fd = fdn;
      stage_fiber(fd, tables);
      __syncthreads();
    }
    regs_to_lds(buf, nrn);
    if (FUSE) bfac_to_lds(bfn, b0n);
    load_rows(kCovered, nrn, b0n, nbn, src_of(stn));
    t = tn;
  }
}

// Per-call parameters of the diagonal operator, stored in device memory.
// coef layout (doubles, n = norb): A_row[n*n] | A_col[n*n] | X[n*n] |
// lin_row[n] | lin_col[n].
struct DiagHeader {
  int norb;
  int z_rep;     // use +-1 signs (empty: +1, occupied: -1) instead of occupations
  int flags;     // bit0: quadratic row/col terms, bit1: cross term, bit2: linear terms
  int pad;
  double scale;  // evolve: angle = -scale * E; contract: out (+)= scale * E * psi
  double pad2;
};

enum { DIAG_EVOLVE = 0, DIAG_CONTRACT = 1, DIAG_CONTRACT_ACC = 2 };
enum { DF_QUAD = 1, DF_CROSS = 2, DF_LIN = 4 };

// Column term gamma[c] (quadratic + linear) for every column string.  For
// evolution it writes the column's phase factor exp(-i scale gamma[c]) as
// (cos, sin) pairs instead.  grid.y runs over operators (the states of a
// batch): operator y has header hdr[y], coefficients coef + y coef_stride and
// writes gamma + y gamma_stride.
__global__ void diag_col_term_kernel(int n_cols, const uint64_t* __restrict__ col_str,
                                     const DiagHeader* __restrict__ hdr,
                                     const double* __restrict__ coef, double* __restrict__ gamma,
                                     int evolve, long coef_stride, long gamma_stride);

// One pass of a diagonal operator over an R-row x n_cols_blk-column block of
// the state.  E(a, b) = rho_a + gamma_b + cross(a, b), cross = sum over the
// occupied orbitals q of b of w_a[q] (number representation) or
// wsum_a - 2 sum_{q in b} w_a[q] (Z representation).
//
// Evolution multiplies by exp(-i s E) formed as a product of factors: the
// row factor exp(-i s rho_a) (with exp(-i s wsum_a) in Z form), the column
// factor (precomputed per column) and one factor per occupied orbital of b,
// exp(-i s w_a[q]) or exp(2 i s w_a[q]), tabled per row in LDS.  This costs a
// few complex multiplies per element instead of an FP64 sin/cos.
//
// Each thread walks its columns two at a time with the next column's loads
// in flight while the current one is formed and stored; loads read a clamped
// column (always in bounds) and only stores are guarded, so that the memory
// counters never wait on a branch.
//
// grid.z runs over the states of a batch when `states` is given: state z is
// transformed in place at states[z] by operator z (header hdr[z],
// coefficients coef + z coef_stride, column values colv + z colv_stride), the
// same arithmetic as a launch for that state alone.
template <typename T, int R, int MODE>
__global__ void __launch_bounds__(256)
    diag_kernel(const Cx<T>* src, Cx<T>* dst, long ld, int n_rows,
                int n_cols, int n_cols_blk, const uint64_t* __restrict__ row_str,
                const uint64_t* __restrict__ col_str, const double* __restrict__ colv,
                const DiagHeader* __restrict__ hdr, const double* __restrict__ coef,
                Cx<T>* const* __restrict__ states, long coef_stride, long colv_stride) {
  if (states) {
    const int zs = blockIdx.z;
    src = states[zs];
    dst = states[zs];
    hdr += zs;
    coef += zs * coef_stride;
    if (colv) colv += zs * colv_stride;
  }
  constexpr int NT = 256;
  constexpr int LPR = NT / R;  // lanes per row in the row-term reduction (<= 64)
  static_assert(LPR <= 64 && (LPR & (LPR - 1)) == 0, "R must be a power of two >= 4");
  __shared__ double w_row[R][64];
  __shared__ Cx<double> fac[R][64];  // evolution: per-orbital factors
  __shared__ double rho[R];
  __shared__ Cx<double> rowf[R];     // evolution: row factors
  const int n = hdr->norb;
  const int z = hdr->z_rep;
  const int flags = hdr->flags;
  const double scale = hdr->scale;
  const int r0 = blockIdx.x * R;
  const int c_lo = blockIdx.y * n_cols_blk;
  const int c_hi = min(n_cols, c_lo + n_cols_blk);
  const double* A_row = coef;
  const double* X = coef + 2 * n * n;
  const double* lin_row = coef + 3 * n * n;
  const int tid = threadIdx.x;

  // cross vectors w_a[q] (rows beyond n_rows get zeros and are never stored)
  for (int t = tid; t < R * n; t += NT) {
    const int r = t / n, q = t % n;
    double acc = 0.0;
    if (r0 + r < n_rows && (flags & DF_CROSS)) {
      const uint64_t bits = row_str[r0 + r];
      for (int i = 0; i < n; ++i) {
        const int occ = (bits >> i) & 1;
        if (z)
          acc += (occ ? -1.0 : 1.0) * X[i * n + q];
        else if (occ)
          acc += X[i * n + q];
      }
    }
    w_row[r][q] = acc;
  }
  // row terms: LPR lanes per row, each summing a share of the (j, k >= j) pairs
  {
    const int r = tid / LPR, l = tid % LPR;
    double acc = 0.0;
    if (r0 + r < n_rows) {
      const uint64_t bits = row_str[r0 + r];
      if (flags & DF_QUAD) {
        for (int p = l; p < n * n; p += LPR) {
          const int j = p / n, k = p % n;
          if (k < j) continue;
          const int oj = (bits >> j) & 1, ok = (bits >> k) & 1;
          if (z) {
            if (k > j) acc += ((oj ^ ok) ? -1.0 : 1.0) * A_row[j * n + k];
          } else if (oj && ok) {
            acc += A_row[j * n + k];
          }
        }
      }
      if (flags & DF_LIN)
        for (int j = l; j < n; j += LPR)
          if ((bits >> j) & 1) acc += lin_row[j];
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
    if (l == 0) rho[r] = acc;
  }
  __syncthreads();
  if (MODE == DIAG_EVOLVE) {
    // per-orbital factors and row factors, angle = -scale * term
    for (int t = tid; t < R * n; t += NT) {
      const int r = t / n, q = t % n;
      const double a = -scale * (z ? -2.0 * w_row[r][q] : w_row[r][q]);
      double sn, cs;
      sincos(a, &sn, &cs);
      fac[r][q] = Cx<double>{cs, sn};
    }
    if (tid < R) {
      double e = rho[tid];
      if (z && (flags & DF_CROSS)) {
        double ws = 0.0;
        for (int q = 0; q < n; ++q) ws += w_row[tid][q];
        e += ws;
      }
      double sn, cs;
      sincos(-scale * e, &sn, &cs);
      rowf[tid] = Cx<double>{cs, sn};
    }
  } else if (z && (flags & DF_CROSS) && tid < R) {
    double ws = 0.0;
    for (int q = 0; q < n; ++q) ws += w_row[tid][q];
    rho[tid] += ws;
    for (int q = 0; q < n; ++q) w_row[tid][q] *= -2.0;
  }
  __syncthreads();

  struct Col {
    Cx<T> v[R];
    uint64_t bits;
    double g0, g1;  // evolution: column factor (cos, sin); else gamma, unused
  };
  auto load = [&](int c, Col& k) {
    const int cc = min(c, c_hi - 1);
    k.bits = col_str[cc];
    if (MODE == DIAG_EVOLVE) {
      k.g0 = colv ? colv[2 * cc] : 1.0;
      k.g1 = colv ? colv[2 * cc + 1] : 0.0;
    } else {
      k.g0 = colv ? colv[cc] : 0.0;
      k.g1 = 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int rr = min(r0 + r, n_rows - 1);
      k.v[r] = src[(long)rr * ld + cc];
    }
  };
  auto form = [&](int c, const Col& k) {
    if (c >= c_hi) return;
    if (MODE == DIAG_EVOLVE) {
      double pr[R], pi[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const Cx<double> f = rowf[r];
        pr[r] = f.re * k.g0 - f.im * k.g1;
        pi[r] = f.re * k.g1 + f.im * k.g0;
      }
      if (flags & DF_CROSS) {
        uint64_t bits = k.bits;
        while (bits) {
          const int q = __builtin_ctzll(bits);
          bits &= bits - 1;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const Cx<double> f = fac[r][q];
            const double t = pr[r] * f.re - pi[r] * f.im;
            pi[r] = pr[r] * f.im + pi[r] * f.re;
            pr[r] = t;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r0 + r >= n_rows) break;
        const Cx<T> v = k.v[r];
        Cx<T> o;
        o.re = v.re * (T)pr[r] - v.im * (T)pi[r];
        o.im = v.re * (T)pi[r] + v.im * (T)pr[r];
        dst[(long)(r0 + r) * ld + c] = o;
      }
    } else {
      double cross[R];
#pragma unroll
      for (int r = 0; r < R; ++r) cross[r] = 0.0;
      if (flags & DF_CROSS) {
        uint64_t bits = k.bits;
        while (bits) {
          const int q = __builtin_ctzll(bits);
          bits &= bits - 1;
#pragma unroll
          for (int r = 0; r < R; ++r) cross[r] += w_row[r][q];
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r0 + r >= n_rows) break;
        const long idx = (long)(r0 + r) * ld + c;
        const T f = (T)(scale * (rho[r] + k.g0 + cross[r]));
        const Cx<T> v = k.v[r];
        Cx<T> o{f * v.re, f * v.im};
        if (MODE == DIAG_CONTRACT_ACC) {
          const Cx<T> p = dst[idx];
          o.re += p.re;
          o.im += p.im;
        }
        dst[idx] = o;
      }
    }
  };
  // two columns per step: B's loads are in flight while A is formed, and A's
  // next loads while B is formed
  Col A, B;
  int c = c_lo + tid;
  if (c < c_hi) load(c, A);
  for (; c < c_hi; c += 2 * NT) {
    load(c + NT, B);
    form(c, A);
    load(c + 2 * NT, A);
    form(c + NT, B);
  }
}

// ---------------------------------------------------------------------------
// Adjoint (gradient) reductions.
//
// grad_rows_kernel: for one row a per block, with w(a,b) = Im(conj(lam) psi):
//   rsum[a] = sum_b w,   M[a][j] = sum_{b : j in b} w   (j < norb <= 32).
// grad_cols_kernel: per chunk of rows, the column sums of w (one column per
// thread, grid.y = row chunk); col_chunk_sum_kernel adds the chunks.
template <typename T>
__global__ void __launch_bounds__(256)
    grad_rows_kernel(const Cx<T>* __restrict__ psi, const Cx<T>* __restrict__ lam, int n_rows,
                     int n_cols, int norb, const uint64_t* __restrict__ col_str,
                     double* __restrict__ rsum, double* __restrict__ M) {
  __shared__ double acc[32 * 256];
  __shared__ double red[256];
  const int a = blockIdx.x;
  const int tid = threadIdx.x;
  for (int j = 0; j < norb; ++j) acc[j * 256 + tid] = 0.0;
  double rs = 0.0;
  for (int c = tid; c < n_cols; c += 256) {
    const Cx<T> p = psi[(long)a * n_cols + c], l = lam[(long)a * n_cols + c];
    const double w = (double)l.re * (double)p.im - (double)l.im * (double)p.re;
    rs += w;
    uint64_t bits = col_str[c];
    while (bits) {
      const int j = __builtin_ctzll(bits);
      bits &= bits - 1;
      acc[j * 256 + tid] += w;
    }
  }
  red[tid] = rs;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      red[tid] += red[tid + s];
      for (int j = 0; j < norb; ++j) acc[j * 256 + tid] += acc[j * 256 + tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
    rsum[a] = red[0];
    for (int j = 0; j < norb; ++j) M[(long)a * norb + j] = acc[j * 256];
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    grad_cols_kernel(const Cx<T>* __restrict__ psi, const Cx<T>* __restrict__ lam, int n_rows,
                     int n_cols, int rows_per_chunk, double* __restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  const int a0 = blockIdx.y * rows_per_chunk, a1 = min(n_rows, a0 + rows_per_chunk);
  double s = 0.0;
  for (int a = a0; a < a1; ++a) {
    const Cx<T> p = psi[(long)a * n_cols + c], l = lam[(long)a * n_cols + c];
    s += (double)l.re * (double)p.im - (double)l.im * (double)p.re;
  }
  part[(long)blockIdx.y * n_cols + c] = s;
}

// csum[c] = sum over the row chunks k (in order) of part[k][c].
static __global__ void __launch_bounds__(256)
    col_chunk_sum_kernel(const double* __restrict__ part, int n_chunks, int n_cols,
                         double* __restrict__ csum) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  double s = 0.0;
#pragma unroll 8
  for (int k = 0; k < n_chunks; ++k) s += part[(long)k * n_cols + c];
  csum[c] = s;
}

// Transition one-body RDM of one spin sector:
//   rdm[p][q] += sum over strings S with q in S and (p == q or p not in S) of
//   sign(a+_p a_q |S>) * sum_other conj(lam[S']) psi[S],   S' = S - q + p.
// sector 0: strings index rows (dot products along rows);
// sector 1: strings index columns (dot products along columns).
// The sign is (-1)^(occupied below q in S) * (-1)^(occupied below p in S - q).
__device__ __forceinline__ uint64_t colex_rank_dev(uint64_t s, const uint64_t* binom_tab,
                                                   int stride) {
  uint64_t r = 0;
  int t = 0;
  while (s) {
    const int c = __builtin_ctzll(s);
    s &= s - 1;
    ++t;
    if (c >= t) r += binom_tab[c * stride + t];
  }
  return r;
}

// Transition one-body RDM of one spin sector, one 4-wave block per string S
// of the sector (its links dealt round-robin to the waves, so that a small
// sector still puts several waves on every SIMD): for q in S and p = q or p
// not in S, T = S - q + p and
//   part[(p norb + q) 2 + c][S - s_begin] = sign * sum_o conj(lam(T, o)) psi(S, o)
// (c: re, im) for the strings s_begin <= S < s_begin + n_chunk, where (S, o)
// addresses row S of an n_str x n_other row-major matrix (COLS = false) or
// column S of an n_other x n_str one (COLS = true).  The links of S run in a
// wave-uniform loop, each dot product is reduced across the wave, and every
// (pair, string) entry is written by exactly one wave, so trdm_reduce_kernel
// sums them in a fixed order (deterministic results).  part must be zero where
// S has no link (p, q).
template <typename T, bool COLS>
__global__ void __launch_bounds__(256)
    trdm_links_kernel(const Cx<T>* __restrict__ lam, const Cx<T>* __restrict__ psi, int n_str,
                      int n_other, int norb, const uint64_t* __restrict__ strings,
                      const uint64_t* __restrict__ binom_tab, int s_begin, int n_chunk,
                      double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s_loc = blockIdx.x;
  const int s_idx = s_begin + s_loc;
  if (s_loc >= n_chunk || s_idx >= n_str) return;
  const uint64_t S = strings[s_idx];
  const int stride = norb + 1;
  auto at = [&](const Cx<T>* m, long str, int o) -> Cx<T> {
    return COLS ? m[(long)o * n_str + str] : m[str * n_other + o];
  };
  // psi's row in registers when it is at most 16 values per lane (row-major
  // sectors): each link then reads only lam's row
  constexpr int XR = 16;
  const bool xreg = !COLS && n_other <= 64 * XR;
  Cx<T> xr[XR];
#pragma unroll
  for (int k = 0; k < XR; ++k) {
    const int o = lane + 64 * k;
    xr[k] = (xreg && o < n_other) ? psi[(long)s_idx * n_other + o] : Cx<T>{T(0), T(0)};
  }
  uint64_t qbits = S;
  int link = 0;  // the string's links in (q, p) order, dealt round-robin to the 4 waves
  while (qbits) {
    const int q = __builtin_ctzll(qbits);
    qbits &= qbits - 1;
    const uint64_t S0 = S & ~(1ull << q);
    const int sign_q = __popcll(S & ((1ull << q) - 1)) & 1;
    for (int p = 0; p < norb; ++p) {
      if (p != q && ((S >> p) & 1)) continue;
      if ((link++ & 3) != wave) continue;
      const uint64_t S1 = S0 | (1ull << p);
      const int sign_p = __popcll(S0 & ((1ull << p) - 1)) & 1;
      const double sgn = ((sign_q ^ sign_p) & 1) ? -1.0 : 1.0;
      const long t_idx = (long)colex_rank_dev(S1, binom_tab, stride);
      double re = 0.0, im = 0.0;
      if (xreg) {
        // the same terms in the same order as the loop below
#pragma unroll
        for (int k = 0; k < XR; ++k) {
          const int o = lane + 64 * k;
          if (o < n_other) {
            const Cx<T> l = lam[t_idx * n_other + o], x = xr[k];
            re += (double)l.re * (double)x.re + (double)l.im * (double)x.im;
            im += (double)l.re * (double)x.im - (double)l.im * (double)x.re;
          }
        }
      } else {
#pragma unroll 4
        for (int o = lane; o < n_other; o += 64) {
          const Cx<T> l = at(lam, t_idx, o), x = at(psi, s_idx, o);
          // conj(l) * x
          re += (double)l.re * (double)x.re + (double)l.im * (double)x.im;
          im += (double)l.re * (double)x.im - (double)l.im * (double)x.re;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off);
        im += __shfl_xor(im, off);
      }
      if (lane == 0) {
        const long e = 2l * (p * norb + q);
        part[e * n_chunk + s_loc] = sgn * re;
        part[(e + 1) * n_chunk + s_loc] = sgn * im;
      }
    }
  }
}

// rdm[e] += sum over the chunk's strings s of part[e][s], one block per entry
// e, in a fixed order.
static __global__ void __launch_bounds__(256)
    trdm_reduce_kernel(const double* __restrict__ part, int n_chunk, double* __restrict__ rdm) {
  __shared__ double red[256];
  const int e = blockIdx.x, tid = threadIdx.x;
  const double* row = part + (long)e * n_chunk;
  double acc = 0.0;
  for (int s = tid; s < n_chunk; s += 256) acc += row[s];
  red[tid] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) rdm[e] += red[0];
}

// out (cols x rows) = in^T for a row-major rows x cols matrix, 32 x 32 tiles
// through LDS (padded against bank conflicts).
template <typename T>
__global__ void __launch_bounds__(256)
    transpose_kernel(const Cx<T>* __restrict__ in, Cx<T>* __restrict__ out, int rows, int cols) {
  __shared__ Cx<T> tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int r = r0 + ty + k, c = c0 + tx;
    if (r < rows && c < cols) tile[ty + k][tx] = in[(long)r * cols + c];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int c = c0 + ty + k, r = r0 + tx;
    if (r < rows && c < cols) out[(long)c * rows + r] = tile[tx][ty + k];
  }
}

// One-body contraction out (+)= scale * sum_pq M_pq a+_p a_q psi, one spin sector
// per launch, as a gather over link tables: string T's links are the L strings
// S with a+_p a_q |S> = sign |T> (p in T, q = p or q not in T), each stored as
// (S, p | q << 8 | (sign < 0) << 16).  m: norb x norb complex (re, im pairs).
// SECTOR 0: strings index rows, one block row T x 256 columns, the links of T
//   are uniform over the block (scalar loads), reads of psi rows coalesced.
// SECTOR 1: strings index columns, thread = (row a, column T).
// accumulate: 0 -> out = scale acc, 1 -> out += scale acc.
template <typename T, int SECTOR>
__global__ void __launch_bounds__(256)
    one_body_kernel(const Cx<T>* __restrict__ psi, Cx<T>* __restrict__ out, int n_rows,
                    int n_cols, const int32_t* __restrict__ links, int n_links,
                    const double* __restrict__ m, int norb, double scale, int accumulate) {
  int row, col;
  if (SECTOR == 0) {
    row = blockIdx.x;
    col = blockIdx.y * 256 + threadIdx.x;
  } else {
    row = blockIdx.y;
    col = blockIdx.x * 256 + threadIdx.x;
  }
  if (row >= n_rows || col >= n_cols) return;
  const int tstr = SECTOR == 0 ? row : col;
  const int32_t* lk = links + (long)tstr * n_links * 2;
  double ar = 0.0, ai = 0.0;
  for (int l = 0; l < n_links; ++l) {
    const int src = lk[2 * l];
    const int info = lk[2 * l + 1];
    const int p = info & 0xff, q = (info >> 8) & 0xff;
    const double sg = (info >> 16) & 1 ? -1.0 : 1.0;
    const double mr = sg * m[2 * (p * norb + q)], mi = sg * m[2 * (p * norb + q) + 1];
    const Cx<T> x = SECTOR == 0 ? psi[(long)src * n_cols + col] : psi[(long)row * n_cols + src];
    ar += mr * (double)x.re - mi * (double)x.im;
    ai += mr * (double)x.im + mi * (double)x.re;
  }
  const long o = (long)row * n_cols + col;
  double r = scale * ar, i = scale * ai;
  if (accumulate) {
    r += (double)out[o].re;
    i += (double)out[o].im;
  }
  out[o] = Cx<T>{(T)r, (T)i};
}

// One-body contraction by window passes: out (+)= scale * sum_{p,q in the
// window} m[p][q] a+_p a_q psi for one spin sector, where m is the pass's
// window-local coefficient matrix (zero for the pairs another pass applies).
// A workgroup stages one fiber x W batch vectors of psi in LDS; every string T
// of the fiber then gathers its kp (w - kp + 1) sources S (links of T, see
// make_one_body_links) from LDS, so psi is read once per pass.  Fiber::pad
// negates the fiber's contributions: the parity of its fixed electrons between
// the two parts of a split window.
// MODE_ROWS: element (string r, batch b) at src[r * ld + b]; MODE_COLS: at
// src[b * ld + r].  accumulate: 0 -> out = scale acc, 1 -> out += scale acc.
template <typename T, int W, int MODE, int NTHR>
__global__ void __launch_bounds__(NTHR)
    one_body_fiber_kernel(const Cx<T>* __restrict__ src, Cx<T>* __restrict__ dst, long ld,
                          int n_batch, const Fiber* __restrict__ fibers, int n_fibers,
                          const int32_t* __restrict__ row_table, int max_rows,
                          const uint32_t* __restrict__ links,
                          const int32_t* __restrict__ link_index,
                          const double* __restrict__ mloc, int w, double scale, int accumulate) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(smem_raw);
  double* mc = reinterpret_cast<double*>(smem_raw + align16((size_t)max_rows * W * sizeof(Cx<T>)));
  const int tid = threadIdx.x;
  for (int i = tid; i < 2 * w * w; i += NTHR) mc[i] = mloc[i];
  const int n_bt = (n_batch + W - 1) / W;
  const long n_tiles = (long)n_fibers * n_bt;
  // XCD-aware order (as in fiber_kernel): each XCD takes runs of consecutive
  // tiles, so the lines that neighbouring batch tiles share hit in its L2
  const int G = gridDim.x;
  const long bid = (G & 7) == 0 ? (long)(blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3)
                                : (long)blockIdx.x;
  auto addr = [&](long r, long b) { return MODE == MODE_ROWS ? r * ld + b : b * ld + r; };
  for (long t = bid; t < n_tiles; t += G) {
    const Fiber fd = fibers[t / n_bt];
    const int b0 = (int)(t % n_bt) * W;
    const int nb = min(W, n_batch - b0);
    const int nr = fd.n_rows;
    const int32_t* rows = row_table + fd.row_off;
    __syncthreads();  // the previous tile's store is done reading the LDS tile
    for (int e = tid; e < nr * W; e += NTHR) {
      const int i = MODE == MODE_ROWS ? e / W : e % nr;
      const int j = MODE == MODE_ROWS ? e % W : e / nr;
      tile[i * W + j] = j < nb ? src[addr(rows[i], b0 + j)] : Cx<T>{T(0), T(0)};
    }
    __syncthreads();
    // one thread per string T: each link (one table word, one coefficient) is
    // applied to all W batch columns of the tile.  Results go back into the
    // tile once every thread is done reading it, and are stored with the
    // load's access pattern (lanes along strings in MODE_COLS, so that a
    // fiber's contiguous addresses are written together).
    const int off = link_index[2 * fd.kp], nl = link_index[2 * fd.kp + 1];
    const double fs = fd.pad ? -scale : scale;
    constexpr int SPT = (1024 + NTHR - 1) / NTHR;  // strings per thread (fibers <= 1024)
    double ar[SPT][W], ai[SPT][W];
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      const int i = tid + k * NTHR;
#pragma unroll
      for (int j = 0; j < W; ++j) ar[k][j] = ai[k][j] = 0.0;
      if (i >= nr) continue;
      const uint32_t* lk = links + off + i;
#pragma unroll 4
      for (int l = 0; l < nl; ++l) {
        const uint32_t u = lk[(long)l * nr];
        const int S = (int)(u & 0xffff), pq = ((u >> 16) & 31) * w + ((u >> 21) & 31);
        double mr = mc[2 * pq], mi = mc[2 * pq + 1];
        if (u >> 26) {
          mr = -mr;
          mi = -mi;
        }
        const Cx<T>* xs = tile + S * W;
#pragma unroll
        for (int j = 0; j < W; ++j) {
          const Cx<T> x = xs[j];
          ar[k][j] += mr * (double)x.re - mi * (double)x.im;
          ai[k][j] += mr * (double)x.im + mi * (double)x.re;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      const int i = tid + k * NTHR;
      if (i >= nr) continue;
#pragma unroll
      for (int j = 0; j < W; ++j) tile[i * W + j] = Cx<T>{(T)(fs * ar[k][j]), (T)(fs * ai[k][j])};
    }
    __syncthreads();
    for (int e = tid; e < nr * W; e += NTHR) {
      const int i = MODE == MODE_ROWS ? e / W : e % nr;
      const int j = MODE == MODE_ROWS ? e % W : e / nr;
      if (j >= nb) continue;
      const long o = addr(rows[i], b0 + j);
      Cx<T> y = tile[i * W + j];
      if (accumulate) {
        const Cx<T> z = dst[o];
        y.re += z.re;
        y.im += z.im;
      }
      dst[o] = y;
    }
  }
}

// x[0..n) = 0 (defined in ffsim_amd.hip).
__global__ void fill_zero_kernel(double* __restrict__ x, long n);

// out = alpha x + beta y over n complex entries.
template <typename T>
__global__ void axpby_kernel(Cx<T>* out, const Cx<T>* x,
                             const Cx<T>* y, long n, double ar, double ai, double br,
                             double bi) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const Cx<T> a = x[k];
    Cx<T> o;
    o.re = (T)ar * a.re - (T)ai * a.im;
    o.im = (T)ar * a.im + (T)ai * a.re;
    if (y) {
      const Cx<T> b = y[k];
      o.re += (T)br * b.re - (T)bi * b.im;
      o.im += (T)br * b.im + (T)bi * b.re;
    }
    out[k] = o;
  }
}

}  // namespace ffsim_amd

namespace ffsim_amd {

// ---------------------------------------------------------------------------
// Single Givens rotations and phase shifts on explicit index lists, the
// granularity of ffsim's own native calls (one rotation of row pairs per call).
// Rotation k of a launch pairs first[p] with second[p] for p < n_pairs and
// applies x' = c x + s y, y' = c y - conj(s) x to the pair.
struct PairRotation {
  const int64_t* first;
  const int64_t* second;
  long n_pairs;
  double c, sr, si;
};
// The rotations of one launch, passed by value.  Rotations of one launch run
// concurrently, so they must touch disjoint rows (columns): two rotations on
// disjoint orbital pairs generally do not (they share strings), and run as
// separate launches.
constexpr int kMaxLayerRotations = 32;
struct PairRotationSet {
  PairRotation r[kMaxLayerRotations];
};

// Strings index rows of an (n_rows x n_cols) matrix: one block per row pair at a
// time, threads along the row (coalesced).  blockIdx.y selects the rotation.
template <typename T>
__global__ void __launch_bounds__(256)
    givens_rows_kernel(Cx<T>* __restrict__ vec, long n_cols, PairRotationSet rots) {
  const PairRotation R = rots.r[blockIdx.y];
  const T c = (T)R.c, sr = (T)R.sr, si = (T)R.si;
  for (long p = blockIdx.x; p < R.n_pairs; p += gridDim.x) {
    Cx<T>* x = vec + R.first[p] * n_cols;
    Cx<T>* y = vec + R.second[p] * n_cols;
    for (long k = threadIdx.x; k < n_cols; k += 256) {
      Cx<T> a = x[k], b = y[k];
      rot_pair<T>(a, b, c, sr, si);
      x[k] = a;
      y[k] = b;
    }
  }
}

// Strings index columns: one thread per (pair, row), rows strided over gridDim.y.
template <typename T>
__global__ void __launch_bounds__(256)
    givens_cols_kernel(Cx<T>* __restrict__ vec, long n_rows, long n_cols,
                       PairRotationSet rots) {
  const PairRotation R = rots.r[blockIdx.z];
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= R.n_pairs) return;
  const long cx = R.first[p], cy = R.second[p];
  const T c = (T)R.c, sr = (T)R.sr, si = (T)R.si;
  for (long r = blockIdx.y; r < n_rows; r += gridDim.y) {
    Cx<T> a = vec[r * n_cols + cx], b = vec[r * n_cols + cy];
    rot_pair<T>(a, b, c, sr, si);
    vec[r * n_cols + cx] = a;
    vec[r * n_cols + cy] = b;
  }
}

// Rows idx[k] of an (n_rows x n_cols) matrix multiplied by a phase.
template <typename T>
__global__ void __launch_bounds__(256)
    phase_rows_kernel(Cx<T>* __restrict__ vec, long n_cols, const int64_t* __restrict__ idx,
                      long n_idx, double pr, double pi) {
  for (long q = blockIdx.x; q < n_idx; q += gridDim.x) {
    Cx<T>* row = vec + idx[q] * n_cols;
    for (long k = threadIdx.x; k < n_cols; k += 256) {
      const Cx<T> a = row[k];
      row[k] = Cx<T>{(T)(a.re * pr - a.im * pi), (T)(a.re * pi + a.im * pr)};
    }
  }
}

}  // namespace ffsim_amd

namespace ffsim_amd {

// ---------------------------------------------------------------------------
// Batched small-state path (serving): an orbital rotation as the compound
// matrix Lambda^k(U) applied by a GEMM, and a batched diagonal pass.
constexpr int kMaxCompoundK = 6;

// out[m][t][s] = det(U_m[T, S]) with T = strings[t], S = strings[s] (k-subsets of
// norb orbitals, rows and columns in ascending orbital order): one thread per
// entry, FP64 Gaussian elimination with partial pivoting.
template <typename T, int K>
__global__ void __launch_bounds__(256)
    compound_kernel(const Cx<double>* __restrict__ U, int norb,
                    const uint64_t* __restrict__ strings, long dim, Cx<T>* __restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long m = blockIdx.y;
  if (e >= dim * dim) return;
  Cx<T>* o = out + m * dim * dim + e;
  if constexpr (K == 0) {
    *o = Cx<T>{T(1), T(0)};
  } else {
    const long t = e / dim, s = e - t * dim;
    const Cx<double>* u = U + m * norb * norb;
    int r[K], c[K];
    uint64_t bt = strings[t], bs = strings[s];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      r[q] = __builtin_ctzll(bt);
      bt &= bt - 1;
      c[q] = __builtin_ctzll(bs);
      bs &= bs - 1;
    }
    double ar[K][K], ai[K][K];
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const Cx<double> x = u[r[i] * norb + c[j]];
        ar[i][j] = x.re;
        ai[i][j] = x.im;
      }
    double dr = 1.0, di = 0.0;
#pragma unroll
    for (int col = 0; col < K; ++col) {
      // pivot: the largest |a[row][col]| at or below the diagonal (swapped in place,
      // the swap flips the sign of the determinant)
#pragma unroll
      for (int row = col + 1; row < K; ++row) {
        const double mp = ar[col][col] * ar[col][col] + ai[col][col] * ai[col][col];
        const double mr = ar[row][col] * ar[row][col] + ai[row][col] * ai[row][col];
        if (mr > mp) {
#pragma unroll
          for (int j = 0; j < K; ++j) {
            const double tr = ar[col][j], ti = ai[col][j];
            ar[col][j] = ar[row][j];
            ai[col][j] = ai[row][j];
            ar[row][j] = tr;
            ai[row][j] = ti;
          }
          dr = -dr;
          di = -di;
        }
      }
      const double pr = ar[col][col], pi = ai[col][col];
      const double t0 = dr * pr - di * pi;
      di = dr * pi + di * pr;
      dr = t0;
      const double p2 = pr * pr + pi * pi;
      if (p2 == 0.0) {
        dr = di = 0.0;
        break;
      }
      const double ir = pr / p2, ii = -pi / p2;  // 1 / pivot
#pragma unroll
      for (int row = col + 1; row < K; ++row) {
        const double fr = ar[row][col] * ir - ai[row][col] * ii;
        const double fi = ar[row][col] * ii + ai[row][col] * ir;
#pragma unroll
        for (int j = col + 1; j < K; ++j) {
          ar[row][j] -= fr * ar[col][j] - fi * ai[col][j];
          ai[row][j] -= fr * ai[col][j] + fi * ar[col][j];
        }
      }
    }
    *o = Cx<T>{(T)dr, (T)di};
  }
}

// psi_i[a][b] *= exp(-i t E_i(a, b)) for a batch of states of one shape, number
// representation, one same-spin matrix A_i (diagonal halved) for both sectors
// and the opposite-spin matrix X_i: E = sum_{j <= k in a} A[j][k] +
// sum_{j <= k in b} A[j][k] + sum_{p in a, q in b} X[p][q].  mats holds A_i then
// X_i per state; one thread per element, state i = blockIdx.y.
template <typename T>
__global__ void __launch_bounds__(256)
    diag_batch_kernel(Cx<T>* const* __restrict__ states, const double* __restrict__ mats,
                      int norb, const uint64_t* __restrict__ str_a, long dim_a,
                      const uint64_t* __restrict__ str_b, long dim_b, double time) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= dim_a * dim_b) return;
  const long a = e / dim_b, b = e - a * dim_b;
  const double* A = mats + (long)blockIdx.y * 2 * norb * norb;
  const double* X = A + norb * norb;
  const uint64_t sa = str_a[a], sb = str_b[b];
  double E = 0.0;
  for (uint64_t x = sa; x; x &= x - 1) {
    const int j = __builtin_ctzll(x);
    for (uint64_t y = x; y; y &= y - 1) E += A[j * norb + __builtin_ctzll(y)];
    for (uint64_t y = sb; y; y &= y - 1) E += X[j * norb + __builtin_ctzll(y)];
  }
  for (uint64_t x = sb; x; x &= x - 1) {
    const int j = __builtin_ctzll(x);
    for (uint64_t y = x; y; y &= y - 1) E += A[j * norb + __builtin_ctzll(y)];
  }
  double sn, cs;
  sincos(-time * E, &sn, &cs);
  Cx<T>* p = states[blockIdx.y] + e;
  const Cx<T> v = *p;
  *p = Cx<T>{(T)(v.re * cs - v.im * sn), (T)(v.re * sn + v.im * cs)};
}

}  // namespace ffsim_amd
