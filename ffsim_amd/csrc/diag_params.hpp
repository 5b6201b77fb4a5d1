// Host-side description of a diagonal operator (diag-Coulomb / number-operator
// terms) and its packed device parameters, shared by the library's source files.
#pragma once

#include <algorithm>
#include <cstring>
#include <vector>

#include "kernels.hpp"

namespace ffsim_amd {

// Matrices are given for (row sector, column sector, cross); empty = absent.
struct DiagSpec {
  int norb = 0, z_rep = 0, flags = 0;
  double scale = 1.0;
  std::vector<double> A_row, A_col, X, lin_row, lin_col;
};

inline std::vector<double> mat_or_empty(const double* m, int n) {
  if (!m) return {};
  return std::vector<double>(m, m + (size_t)n * n);
}

// A diagonal Coulomb operator from the prepared matrices of the C ABI
// (number representation: aa/bb diagonals already halved; Z representation:
// raw matrices, scaled by 1/4 here).
inline DiagSpec diag_coulomb_spec(int norb, const double* aa, const double* ab, const double* bb,
                                  int z_rep, double scale) {
  DiagSpec s;
  s.norb = norb;
  s.z_rep = z_rep ? 1 : 0;
  s.scale = scale;
  s.A_row = mat_or_empty(aa, norb);
  s.A_col = mat_or_empty(bb, norb);
  s.X = mat_or_empty(ab, norb);
  if (z_rep) {
    for (auto* v : {&s.A_row, &s.A_col, &s.X})
      for (auto& x : *v) x *= 0.25;
  }
  s.flags = ((aa || bb) ? DF_QUAD : 0) | (ab ? DF_CROSS : 0);
  return s;
}

// Device layout of the coefficients: A_row[n*n] | A_col[n*n] | X[n*n] |
// lin_row[n] | lin_col[n].
inline DiagHeader diag_header(const DiagSpec& spec) {
  DiagHeader h{};
  h.norb = spec.norb;
  h.z_rep = spec.z_rep;
  h.flags = spec.flags;
  h.scale = spec.scale;
  return h;
}

inline std::vector<double> diag_coefficients(const DiagSpec& spec) {
  const size_t nn = (size_t)spec.norb * spec.norb;
  std::vector<double> coef(3 * nn + 2 * spec.norb, 0.0);
  auto put = [&](const std::vector<double>& v, size_t off, size_t n) {
    if (!v.empty()) std::copy(v.begin(), v.begin() + n, coef.begin() + off);
  };
  put(spec.A_row, 0, nn);
  put(spec.A_col, nn, nn);
  put(spec.X, 2 * nn, nn);
  put(spec.lin_row, 3 * nn, spec.norb);
  put(spec.lin_col, 3 * nn + spec.norb, spec.norb);
  return coef;
}

// Rows per block and columns per block of the diagonal pass.
constexpr int kDiagRows = 8;
constexpr int kDiagColsPerBlock = 2048;

// The column-term pre-pass: gamma[c] (contractions) or its phase factor
// (evolution), for n_ops operators at the given strides (headers consecutive).
inline void launch_diag_col_term(long n_cols, const uint64_t* col_str, const DiagHeader* hdr,
                                 const double* coef, double* gamma, int evolve,
                                 hipStream_t stream, int n_ops = 1, long coef_stride = 0,
                                 long gamma_stride = 0) {
  hipLaunchKernelGGL(diag_col_term_kernel, dim3((unsigned)((n_cols + 255) / 256), (unsigned)n_ops),
                     dim3(256), 0, stream, (int)n_cols, col_str, hdr, coef, gamma, evolve,
                     coef_stride, gamma_stride);
}

// The elementwise pass over an n_rows x n_cols state (blocks of kDiagRows rows x
// up to kDiagColsPerBlock columns).  The grid limits are checked by the caller
// with diag_grid.
inline dim3 diag_grid(long n_rows, long n_cols) {
  const long cols_blk = std::min<long>(n_cols, kDiagColsPerBlock);
  return dim3((unsigned)((n_rows + kDiagRows - 1) / kDiagRows),
              (unsigned)((n_cols + cols_blk - 1) / cols_blk));
}

// With `states`, n_states states are transformed in place (src and dst
// unused), state z by the operator at hdr + z, coef + z coef_stride and
// gamma + z gamma_stride.
template <typename T>
void launch_diag_pass_t(int op, const void* src, void* dst, long n_rows, long n_cols,
                        const uint64_t* row_str, const uint64_t* col_str, const double* gamma,
                        const DiagHeader* hdr, const double* coef, hipStream_t stream,
                        void* const* states = nullptr, int n_states = 1, long coef_stride = 0,
                        long gamma_stride = 0) {
  const int cols_blk = (int)std::min<long>(n_cols, kDiagColsPerBlock);
  dim3 grid = diag_grid(n_rows, n_cols);
  grid.z = (unsigned)n_states;
#define FFSIM_DIAG_LAUNCH(MM)                                                                \
  hipLaunchKernelGGL((diag_kernel<T, kDiagRows, MM>), grid, dim3(256), 0, stream,            \
                     reinterpret_cast<const Cx<T>*>(src), reinterpret_cast<Cx<T>*>(dst), n_cols, \
                     (int)n_rows, (int)n_cols, cols_blk, row_str, col_str, gamma, hdr, coef,    \
                     reinterpret_cast<Cx<T>* const*>(states), coef_stride, gamma_stride)
  if (op == DIAG_EVOLVE)
    FFSIM_DIAG_LAUNCH(DIAG_EVOLVE);
  else if (op == DIAG_CONTRACT)
    FFSIM_DIAG_LAUNCH(DIAG_CONTRACT);
  else
    FFSIM_DIAG_LAUNCH(DIAG_CONTRACT_ACC);
#undef FFSIM_DIAG_LAUNCH
}

}  // namespace ffsim_amd
