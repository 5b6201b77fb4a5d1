// Device kernels of the ffsim_amd library (gfx950).
//
// fiber_kernel: applies a program of Givens rotations (and optionally the
//   diagonal phases of an orbital rotation) to every fiber of a window.  One
//   workgroup owns one fiber x W batch vectors, staged in LDS; rotations run
//   one after another on the LDS tile (a rotation pairs local strings x, y that
//   differ by moving one electron between adjacent window orbitals).
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

// Element (string i, batch j) of a fiber tile lives at
//   MODE_ROWS: state[row_table[i] * ld + b0 + j]   (strings index rows)
//   MODE_COLS: state[(b0 + j) * ld + row_table[i]] (strings index columns)
// tstore (MODE_ROWS only): write element (i, j) to dst[(b0 + j) * ld_dst +
// row_table[i]] instead, i.e. store the transposed state.
template <typename T, int W, int MODE>
__global__ void __launch_bounds__(256)
    fiber_kernel(const Cx<T>* src, Cx<T>* dst, long ld_src, long ld_dst,
                 int n_batch, int tstore, const Fiber* __restrict__ fibers, int n_fibers,
                 const int32_t* __restrict__ row_table, const uint32_t* __restrict__ pair_tab,
                 const int32_t* __restrict__ pair_idx, int wm1, const KStep* __restrict__ steps,
                 int n_steps, const double* __restrict__ diag,
                 const uint64_t* __restrict__ strings) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(smem_raw);
  constexpr int S = W + 1;  // padded LDS stride (conflict-free column walks)
  const int f = blockIdx.x % n_fibers;
  const int bt = blockIdx.x / n_fibers;
  const Fiber fd = fibers[f];
  const int nr = fd.n_rows;
  const int b0 = bt * W;
  const int nb = min(W, n_batch - b0);
  const int32_t* rows = row_table + fd.row_off;
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;

  // ---- load the tile
  if (MODE == MODE_ROWS) {
    const int j = tid % W;
    for (int i = tid / W; i < nr; i += nthr / W) {
      Cx<T> v = {T(0), T(0)};
      if (j < nb) v = src[(long)rows[i] * ld_src + b0 + j];
      tile[i * S + j] = v;
    }
  } else {
    for (int j = 0; j < W; ++j) {
      for (int i = tid; i < nr; i += nthr) {
        Cx<T> v = {T(0), T(0)};
        if (j < nb) v = src[(long)(b0 + j) * ld_src + rows[i]];
        tile[i * S + j] = v;
      }
    }
  }
  __syncthreads();

  // ---- Givens program
  for (int s = 0; s < n_steps; ++s) {
    const KStep st = steps[s];
    const int pi = (fd.kp * wm1 + st.p) * 2;
    const int off = pair_idx[pi];
    const int cnt = pair_idx[pi + 1];
    const T c = (T)st.c, sr = (T)st.sr, si = (T)st.si;
    for (int e = tid; e < cnt * W; e += nthr) {
      const int q = e / W, j = e % W;
      const uint32_t pr = pair_tab[off + q];
      const int x = pr & 0xffff, y = pr >> 16;
      const Cx<T> a = tile[x * S + j];
      const Cx<T> b = tile[y * S + j];
      Cx<T> na, nb2;
      na.re = c * a.re + (sr * b.re - si * b.im);
      na.im = c * a.im + (sr * b.im + si * b.re);
      nb2.re = c * b.re - (sr * a.re + si * a.im);
      nb2.im = c * b.im - (sr * a.im - si * a.re);
      tile[x * S + j] = na;
      tile[y * S + j] = nb2;
    }
    __syncthreads();
  }

  // ---- optional diagonal phases D (orbital rotation tail) and store
  auto row_phase = [&](int i, T& pr, T& pim) {
    double re = 1.0, im = 0.0;
    uint64_t bits = strings[rows[i]];
    while (bits) {
      const int o = __builtin_ctzll(bits);
      bits &= bits - 1;
      const double dr = diag[2 * o], di = diag[2 * o + 1];
      const double t = re * dr - im * di;
      im = re * di + im * dr;
      re = t;
    }
    pr = (T)re;
    pim = (T)im;
  };
  if (MODE == MODE_ROWS && !tstore) {
    const int j = tid % W;
    if (j < nb) {
      for (int i = tid / W; i < nr; i += nthr / W) {
        Cx<T> v = tile[i * S + j];
        if (diag) {
          T pr, pim;
          row_phase(i, pr, pim);
          const T t = v.re * pr - v.im * pim;
          v.im = v.re * pim + v.im * pr;
          v.re = t;
        }
        dst[(long)rows[i] * ld_dst + b0 + j] = v;
      }
    }
  } else {
    for (int i = tid; i < nr; i += nthr) {
      T pr = T(1), pim = T(0);
      if (diag) row_phase(i, pr, pim);
      const long r = rows[i];
      for (int j = 0; j < nb; ++j) {
        Cx<T> v = tile[i * S + j];
        if (diag) {
          const T t = v.re * pr - v.im * pim;
          v.im = v.re * pim + v.im * pr;
          v.re = t;
        }
        dst[(long)(b0 + j) * ld_dst + r] = v;
      }
    }
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

// Column term gamma[c] (quadratic + linear) for every column string.
__global__ void diag_col_term_kernel(int n_cols, const uint64_t* __restrict__ col_str,
                                     const DiagHeader* __restrict__ hdr,
                                     const double* __restrict__ coef, double* __restrict__ gamma);

template <typename T, int R, int MODE>
__global__ void __launch_bounds__(256)
    diag_kernel(const Cx<T>* src, Cx<T>* dst, long ld, int n_rows,
                int n_cols, const uint64_t* __restrict__ row_str,
                const uint64_t* __restrict__ col_str, const double* __restrict__ gamma,
                const DiagHeader* __restrict__ hdr, const double* __restrict__ coef) {
  __shared__ double w_row[R][64];
  __shared__ double rho[R];
  __shared__ double wsum[R];
  const int n = hdr->norb;
  const int z = hdr->z_rep;
  const int flags = hdr->flags;
  const double scale = hdr->scale;
  const int r0 = blockIdx.x * R;
  const double* A_row = coef;
  const double* X = coef + 2 * n * n;
  const double* lin_row = coef + 3 * n * n;
  // per-row terms
  for (int t = threadIdx.x; t < R * n; t += blockDim.x) {
    const int r = t / n, jp = t % n;
    double acc = 0.0;
    if (r0 + r < n_rows && (flags & DF_CROSS)) {
      const uint64_t bits = row_str[r0 + r];
      for (int i = 0; i < n; ++i) {
        const int occ = (bits >> i) & 1;
        if (z)
          acc += (occ ? -1.0 : 1.0) * X[i * n + jp];
        else if (occ)
          acc += X[i * n + jp];
      }
    }
    w_row[r][jp] = acc;
  }
  if (threadIdx.x < R) {
    const int r = threadIdx.x;
    double acc = 0.0;
    if (r0 + r < n_rows) {
      const uint64_t bits = row_str[r0 + r];
      if (flags & DF_QUAD) {
        for (int j = 0; j < n; ++j) {
          const int oj = (bits >> j) & 1;
          for (int k = j; k < n; ++k) {
            const int ok = (bits >> k) & 1;
            if (z) {
              if (k > j) acc += ((oj ^ ok) ? -1.0 : 1.0) * A_row[j * n + k];
            } else if (oj && ok) {
              acc += A_row[j * n + k];
            }
          }
        }
      }
      if (flags & DF_LIN) {
        for (int j = 0; j < n; ++j)
          if ((bits >> j) & 1) acc += lin_row[j];
      }
    }
    rho[r] = acc;
  }
  __syncthreads();
  if (threadIdx.x < R) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += w_row[threadIdx.x][j];
    wsum[threadIdx.x] = s;
  }
  __syncthreads();

  for (int c = threadIdx.x; c < n_cols; c += blockDim.x) {
    const uint64_t cb = col_str[c];
    const double g = gamma ? gamma[c] : 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (r0 + r >= n_rows) break;
      double cross = 0.0;
      if (flags & DF_CROSS) {
        uint64_t bits = cb;
        while (bits) {
          const int jp = __builtin_ctzll(bits);
          bits &= bits - 1;
          cross += w_row[r][jp];
        }
        if (z) cross = wsum[r] - 2.0 * cross;
      }
      const double E = rho[r] + g + cross;
      const long idx = (long)(r0 + r) * ld + c;
      const Cx<T> v = src[idx];
      Cx<T> o;
      if (MODE == DIAG_EVOLVE) {
        double sn, cs;
        const double ang = -scale * E;
        if (sizeof(T) == 8) {
          sincos(ang, &sn, &cs);
        } else {
          const double red = ang - 6.283185307179586 * rint(ang * 0.15915494309189535);
          float sf, cf;
          sincosf((float)red, &sf, &cf);
          sn = sf;
          cs = cf;
        }
        o.re = v.re * (T)cs - v.im * (T)sn;
        o.im = v.re * (T)sn + v.im * (T)cs;
      } else {
        const T f = (T)(scale * E);
        o.re = f * v.re;
        o.im = f * v.im;
        if (MODE == DIAG_CONTRACT_ACC) {
          const Cx<T> p = dst[idx];
          o.re += p.re;
          o.im += p.im;
        }
      }
      dst[idx] = o;
    }
  }
}

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
