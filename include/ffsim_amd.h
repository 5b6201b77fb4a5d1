/*
 * ffsim_amd: C ABI of the MI355X (gfx950) gate-application library.
 *
 * Every entry point returns 0 on success and a negative status on failure;
 * ffsim_amd_last_error() then describes the failure (thread-local string).
 *
 * State vectors are device pointers to row-major dim_a x dim_b complex
 * matrices (address = addr_a * dim_b + addr_b, python/ffsim/states/bitstring.py:246),
 * interleaved (re, im), in FP64 (dtype FFSIM_C128) or FP32 (dtype FFSIM_C64).
 * Matrices describing gates (unitaries, Coulomb matrices, coefficients) are
 * HOST pointers to FP64 data; complex matrices are interleaved complex128,
 * row-major, exactly the memory of a C-contiguous numpy complex128 array.
 * `stream` is a hipStream_t (NULL = the default stream); all device work of a
 * call is enqueued on it and the call returns without synchronizing.
 * Scratch state buffers (`work`) are owned by the caller and must hold
 * dim_a * dim_b elements of the state dtype.
 */
#ifndef FFSIM_AMD_H
#define FFSIM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFSIM_C128 0
#define FFSIM_C64 1

#define FFSIM_OK 0
#define FFSIM_ERR_VALUE (-1)   /* bad argument (python ValueError)            */
#define FFSIM_ERR_ASSERT (-2)  /* violated precondition (python AssertionError) */
#define FFSIM_ERR_HIP (-3)     /* HIP runtime failure                          */
#define FFSIM_ERR_NOTIMPL (-4) /* unsupported combination (NotImplementedError) */

/* Library status. */
const char* ffsim_amd_last_error(void);
int ffsim_amd_version(void);
/* Selects and initialises `device` (tables, program ring).  Optional: every
 * device entry point initialises the current device on first use. */
int ffsim_amd_init(int device);

/* Kernel timing (benchmark support; no counterpart in the reference).
 * Between begin and end every fiber (Givens) kernel and diagonal kernel launch
 * is bracketed by HIP events on the stream it runs on.  end synchronizes and
 * writes 6 doubles: {fiber ms, fiber launches, fiber algorithmic bytes,
 * diagonal ms, diagonal launches, diagonal algorithmic bytes}. */
int ffsim_amd_profile_begin(void);
int ffsim_amd_profile_end(double* out);

/* ---- host-side tables (bit-exact with ffsim's cached pyscf tables) ----- */

/* Replaces ffsim._cistring.make_strings (python/ffsim/_cistring.py:21-24):
 * the C(norb, nocc) occupation bitmasks in ascending order. */
int ffsim_make_strings(int norb, int nocc, int64_t* out);
/* Replaces ffsim._cistring.gen_occslst (python/ffsim/_cistring.py:27-30):
 * out is C(norb, nocc) x nocc, occupied orbitals ascending per string. */
int ffsim_gen_occslst(int norb, int nocc, uint64_t* out);
/* Replaces gates.orbital_rotation._zero_one_subspace_indices
 * (python/ffsim/gates/orbital_rotation.py:202-212): 2 * C(norb-2, nocc-1)
 * addresses, the strings with i occupied / j empty followed by their partners
 * with j occupied / i empty.  Returns the number of entries written. */
int ffsim_zero_one_subspace_indices(int norb, int nocc, int i, int j, uint64_t* out);

/* Replaces ffsim._lib.givens_decomposition (src/linalg/givens.rs:69-149,
 * python/ffsim/_lib.pyi:42-44).  mat: n x n complex128.  Writes up to
 * n(n-1)/2 rotations (c[k], s[2k..2k+1], i[k], j[k]) with U = D G_L^* ... G_1^*
 * and the n diagonal phases (2n doubles).  *n_rot receives L. */
int ffsim_givens_decomposition(const double* mat, int n, double tol, double* c, double* s,
                               int32_t* i, int32_t* j, int32_t* n_rot, double* diag);

/* ---- device gate application ------------------------------------------ */

/* Number of state elements the `work` buffer of ffsim_apply_orbital_rotation,
 * ffsim_apply_ucj_spin_balanced and ffsim_apply_diag_coulomb_sequence needs
 * for this shape (0: none, NULL may be passed).  Negative on error.  No
 * reference counterpart: it sizes the transposition buffer that replaces the
 * reference's transpose copies (python/ffsim/gates/orbital_rotation.py:143,
 * :152); the transposed state's rows are padded to 32 bytes, so this can
 * exceed dim_a * dim_b by a few elements per beta string. */
int64_t ffsim_orbital_rotation_work_elements(int dtype, int norb, int n_alpha, int n_beta);

/* Replaces the spinful path of ffsim.apply_orbital_rotation
 * (python/ffsim/gates/orbital_rotation.py:117-154 -> _lib.apply_givens_rotation_in_place,
 * src/gates/orbital_rotation.rs:19-104, and _lib.apply_phase_shift_in_place,
 * src/gates/phase_shift.rs:16-30).  mat_a / mat_b: norb x norb complex128 or
 * NULL (no rotation of that spin sector).  In place on vec; work may be NULL
 * when norb is small enough that a whole string vector fits on chip. */
int ffsim_apply_orbital_rotation(void* vec, void* work, int dtype, int norb, int n_alpha,
                                 int n_beta, const double* mat_a, const double* mat_b,
                                 void* stream);

/* ffsim_apply_orbital_rotation for n_states states of one shape, each with its
 * own matrices: batching has no counterpart in the reference, whose callers
 * loop over apply_orbital_rotation (the factors of
 * DoubleFactorizedHamiltonian._linear_operator_,
 * python/ffsim/hamiltonians/double_factorized_hamiltonian.py:244-275).
 * vecs / works: HOST arrays of n_states DEVICE pointers; works may be NULL
 * (or hold NULLs) where ffsim_orbital_rotation_work_elements is 0, and holds
 * buffers of that many elements otherwise (states that use one run one after
 * another on the stream, so the entries may all name the same buffer).
 * mats_a / mats_b: NULL (that sector is not rotated) or n_states x norb x
 * norb complex128.  Where a beta vector fits one tile (6 <= norb <= 12),
 * each window pass of each sector is one launch over all states; otherwise
 * the states run one by one.  Each result is bitwise equal to
 * ffsim_apply_orbital_rotation's. */
int ffsim_apply_orbital_rotation_batched(int n_states, void* const* vecs, void* const* works,
                                         int dtype, int norb, int n_alpha, int n_beta,
                                         const double* mats_a, const double* mats_b,
                                         void* stream);

/* Replaces _lib.apply_diag_coulomb_evolution_in_place_num_rep / _z_rep
 * (src/gates/diag_coulomb.rs:19-190).  The matrices are real norb x norb,
 * already in the form ffsim's _get_mat_exp exponentiates
 * (python/ffsim/gates/diag_coulomb.py:223-275) but before exponentiation:
 * number representation -> diagonals of aa/bb halved; Z representation ->
 * raw matrices (the 1/4 factor is applied here).  NULL = no interaction of
 * that type.  Applies vec[a,b] *= exp(-i time E(a,b)). */
int ffsim_apply_diag_coulomb_evolution(void* vec, int dtype, int norb, int n_alpha, int n_beta,
                                       const double* mat_aa, const double* mat_ab,
                                       const double* mat_bb, double time, int z_representation,
                                       void* stream);

/* ffsim_apply_diag_coulomb_evolution with one operator on n_states states of
 * one shape (vecs: a HOST array of DEVICE pointers), in one elementwise
 * launch over all of them; batching has no counterpart in the reference,
 * whose callers loop (the adjoint sweep here evolves psi and lambda alike).
 * Each result is bitwise equal to ffsim_apply_diag_coulomb_evolution's. */
int ffsim_apply_diag_coulomb_evolution_batched(int n_states, void* const* vecs, int dtype,
                                               int norb, int n_alpha, int n_beta,
                                               const double* mat_aa, const double* mat_ab,
                                               const double* mat_bb, double time,
                                               int z_representation, void* stream);

/* Replaces _lib.apply_num_op_sum_evolution_in_place on both spin sectors
 * (src/gates/num_op_sum.rs:18-36, python/ffsim/gates/num_op_sum.py:175-219).
 * coeffs_a / coeffs_b: norb real coefficients or NULL. */
int ffsim_apply_num_op_sum_evolution(void* vec, int dtype, int norb, int n_alpha, int n_beta,
                                     const double* coeffs_a, const double* coeffs_b, double time,
                                     void* stream);

/* Replaces _lib.contract_diag_coulomb_into_buffer_num_rep / _z_rep
 * (src/contract/diag_coulomb.rs:19-174): out (=|+=) scale * E(a,b) vec[a,b],
 * with the matrices prepared as python/ffsim/contract/diag_coulomb.py:102-130
 * does (number representation: aa/bb diagonals halved).  accumulate = 0
 * overwrites out, 1 adds into it.  out may alias vec when accumulate = 0. */
int ffsim_contract_diag_coulomb(const void* vec, void* out, int dtype, int norb, int n_alpha,
                                int n_beta, const double* mat_aa, const double* mat_ab,
                                const double* mat_bb, int z_representation, double scale,
                                int accumulate, void* stream);

/* ffsim_contract_diag_coulomb in place on n_states states of one shape, each
 * with its own matrices (mats_*: NULL or n_states x norb x norb, in the
 * convention of ffsim_contract_diag_coulomb), in one elementwise launch:
 * vec_i = scale * E_i(a, b) * vec_i.  Batching has no counterpart in the
 * reference, whose DoubleFactorizedHamiltonian._linear_operator_ contracts
 * its terms one by one (python/ffsim/hamiltonians/
 * double_factorized_hamiltonian.py:244-275).  Each result is bitwise equal to
 * ffsim_contract_diag_coulomb's (with vec == out). */
int ffsim_contract_diag_coulomb_batched(int n_states, void* const* vecs, int dtype, int norb,
                                        int n_alpha, int n_beta, const double* mats_aa,
                                        const double* mats_ab, const double* mats_bb,
                                        int z_representation, double scale, void* stream);

/* Replaces _lib.contract_num_op_sum_spin_into_buffer applied to both spin
 * sectors (src/contract/num_op_sum.rs:18-39, python/ffsim/contract/num_op_sum.py:27-72):
 * out (=|+=) scale * (sum_{i in a} ca[i] + sum_{j in b} cb[j]) vec[a,b]. */
int ffsim_contract_num_op_sum(const void* vec, void* out, int dtype, int norb, int n_alpha,
                              int n_beta, const double* coeffs_a, const double* coeffs_b,
                              double scale, int accumulate, void* stream);

/* out = alpha * x + beta * y elementwise over n complex entries (device
 * helper for linear-operator sums; out may alias x or y). */
int ffsim_axpby(void* out, const void* x, const void* y, int dtype, int64_t n, double alpha_re,
                double alpha_im, double beta_re, double beta_im, void* stream);

/* Replaces UCJOpSpinBalanced._apply_unitary_
 * (python/ffsim/variational/ucj_spin_balanced.py:657-696).
 * diag_coulomb_mats: n_reps x 2 x norb x norb real (aa, ab); orbital_rotations:
 * n_reps x norb x norb complex128; final_orbital_rotation: NULL or norb x norb. */
int ffsim_apply_ucj_spin_balanced(void* vec, void* work, int dtype, int norb, int n_alpha,
                                  int n_beta, int n_reps, const double* diag_coulomb_mats,
                                  const double* orbital_rotations,
                                  const double* final_orbital_rotation, void* stream);

/* UCJOpSpinBalanced._apply_unitary_ (python/ffsim/variational/ucj_spin_balanced.py:
 * 657-696) for a batch of n_states states of one shape (norb, n_alpha, n_beta,
 * dtype), each with its own operator; batching has no counterpart in the
 * reference (its serving unit is one call per state).  vecs / works: HOST
 * arrays of n_states DEVICE pointers (states and their work buffers, each
 * dim_a * dim_b elements; works may hold NULLs where
 * ffsim_orbital_rotation_work_elements is 0 and the GEMM path below is not
 * taken).  diag_coulomb_mats: n_states x n_reps x 2 x norb x norb;
 * orbital_rotations: n_states x n_reps x norb x norb complex128;
 * final_orbital_rotations: NULL or n_states x norb x norb.  Small states
 * (at most 6 electrons per sector and at most 100 strings, or 150 strings in
 * batches of 8 or more) take a GEMM path: each merged orbital rotation is the
 * compound matrix Lambda^k(U) (built on the device) applied to all states by
 * batched complex GEMMs, and each diagonal term one batched elementwise pass.
 * Other states take the Givens path.  Where a beta vector fits one tile
 * (6 <= norb <= 12), every window pass and diagonal term is one launch over
 * all states, with results bitwise equal to the single-state entry's;
 * otherwise the states run one by one on
 * the stream (on the default stream, small states run on up to 4
 * library-owned streams at once, between events on the default stream). */
int ffsim_apply_ucj_spin_balanced_batched(int n_states, void* const* vecs, void* const* works,
                                          int dtype, int norb, int n_alpha, int n_beta,
                                          int n_reps, const double* diag_coulomb_mats,
                                          const double* orbital_rotations,
                                          const double* final_orbital_rotations, void* stream);

/* One-body contraction: out (+)= scale * sum_{pq, sigma} mat[p][q] a+_{p sigma} a_{q sigma} vec,
 * mat a HOST norb x norb complex128 matrix.  Replaces contract_one_body
 * (python/ffsim/contract/one_body.py:24-56, i.e. pyscf's contract_1e over the link
 * tables of _cistring.py:33-42).  Up to norb = 18 it runs as window passes
 * (one per sector up to 12 orbitals, three above); wider states use a gather
 * over link tables.  out must not alias vec.  norb <= 63. */
int ffsim_contract_one_body(const void* vec, void* out, int dtype, int norb, int n_alpha,
                            int n_beta, const double* mat, double scale, int accumulate,
                            void* stream);

/* ---- adjoint reductions (differentiation; no counterpart in the reference,
 * whose only gradient is finite differences, python/ffsim/optimize/_util.py:60-110) */

/* Phase backward of a diag-Coulomb gate: with w(a,b) = Im(conj(lam[a,b]) psi[a,b]),
 * writes into the DEVICE buffer out (dim_a + dim_b + dim_a*norb doubles):
 * row sums (dim_a), column sums (dim_b) and M[a][j] = sum over b with orbital j
 * occupied of w(a,b).  norb <= 32. */
int ffsim_diag_coulomb_grad_reduce(const void* psi, const void* lam, int dtype, int norb,
                                   int n_alpha, int n_beta, double* out, void* stream);

/* Transition one-body RDM rdm_sigma[p][q] = <lam| a+_{p sigma} a_{q sigma} |psi> for
 * sigma = alpha then beta, as 2 x norb x norb complex128 in the DEVICE buffer out
 * (overwritten).  Sums run in a fixed order: repeated calls give the same bits. */
int ffsim_transition_rdm(const void* lam, const void* psi, int dtype, int norb, int n_alpha,
                         int n_beta, double* out, void* stream);

/* A sequence of rotated diagonal-Coulomb evolutions, the gate pattern shared
 * by UCJOpSpinBalanced._apply_unitary_ (ucj_spin_balanced.py:657-696), the
 * double-factorized Trotter step (python/ffsim/trotter/double_factorized.py:
 * 89-127) and the split-operator Trotter step
 * (python/ffsim/trotter/diagonal_coulomb_split_op.py:87-127): for t in
 * 0..n_terms-1, rotate both spin sectors by
 * rotations[t] then evolve by the diag-Coulomb matrices mats[t] = (aa, ab, bb)
 * for times[t]; finally rotate by rotations[n_terms].
 * rotations: (n_terms + 1) x norb x norb complex128; mats: n_terms x 3 x norb x
 * norb real, prepared as for ffsim_apply_diag_coulomb_evolution (number
 * representation: aa/bb diagonals halved); times: n_terms.  In the number
 * representation, each evolution is applied inside the last pass
 * of the rotation before it when the beta sector runs transposed (results are
 * the same to rounding; no separate diagonal pass over the state). */
int ffsim_apply_diag_coulomb_sequence(void* vec, void* work, int dtype, int norb, int n_alpha,
                                      int n_beta, int n_terms, const double* rotations,
                                      const double* mats, const double* times,
                                      int z_representation, void* stream);

/* ---- ffsim's native calls one for one (device state, one rotation or one
 * phase per call, as ffsim's own Python schedule issues them) ------------- */

/* Replaces _lib.apply_givens_rotation_in_place (src/gates/orbital_rotation.rs:19-104,
 * python/ffsim/_lib.pyi:39-41).  vec: DEVICE n_rows x n_cols state; slice1 /
 * slice2: HOST arrays of n_pairs row indices (ffsim's usize tables).  For each
 * pair (r1, r2): r1' = c r1 + s r2, r2' = c r2 - conj(s) r1.  Rows must be
 * distinct across pairs (as ffsim's tables are).  An index >= n_rows is an
 * error (ffsim panics). */
int ffsim_apply_givens_rotation_in_place(void* vec, int dtype, int64_t n_rows, int64_t n_cols,
                                         double c, double s_re, double s_im,
                                         const uint64_t* slice1, const uint64_t* slice2,
                                         int64_t n_pairs, void* stream);

/* Replaces _lib.apply_phase_shift_in_place (src/gates/phase_shift.rs:16-30,
 * python/ffsim/_lib.pyi:36-38): rows indices[k] of the DEVICE n_rows x n_cols
 * state are multiplied by phase.  indices: HOST array of n_indices. */
int ffsim_apply_phase_shift_in_place(void* vec, int dtype, int64_t n_rows, int64_t n_cols,
                                     double phase_re, double phase_im, const uint64_t* indices,
                                     int64_t n_indices, void* stream);

/* One layer of disjoint adjacent Givens rotations on one spin sector, in one
 * call: the n_rot calls of _lib.apply_givens_rotation_in_place that
 * python/ffsim/gates/orbital_rotation.py:132-135 would issue for these
 * rotations (pair tables of gates/orbital_rotation.py:202-212).  The layer runs
 * as fused window passes over the state when the strings index rows, or index
 * columns and a whole string vector fits one on-chip tile; otherwise as one
 * launch per rotation over pair tables cached on the device.  Rotation r applies (c[r], s[2r] + i s[2r+1]) to
 * the strings with orbital i[r] occupied / j[r] empty (first) and their partners
 * (second) exactly as ffsim's call does; |i - j| must be 1 (AssertionError
 * otherwise, gates/orbital_rotation.py:195) and the orbital pairs disjoint.
 * strings_on_rows = 1: vec is C(norb, nocc) x n_other (the alpha sector);
 * 0: vec is n_other x C(norb, nocc) (the beta sector in place, no transpose). */
int ffsim_apply_givens_layer(void* vec, int dtype, int norb, int nocc, int64_t n_other,
                             int strings_on_rows, int n_rot, const double* c, const double* s,
                             const int32_t* i, const int32_t* j, void* stream);

/* ---- allocation-free variant -------------------------------------------- */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same result as ffsim_apply_diag_coulomb_evolution
 * (same arguments, same kernels), with the operator's parameters, the string
 * tables and the per-column factors in a static 8 MiB device array instead of
 * library-allocated memory; the call itself allocates nothing.  The tables of
 * a state must fit the array: 128 KiB + 8 (dim_a + dim_b) + 16 dim_b bytes
 * (rounded up to 256 bytes each) <= 8 MiB, e.g. norb=20 (10,10); larger states
 * fail with FFSIM_ERR_VALUE.  Every call on one GPU shares that GPU's array:
 * calls that overlap in time on one GPU race (calls on one stream are
 * ordered).  ffsim_amd_init(device) resolves the array's address on that
 * device; otherwise the first call on a device does. */
int ffsim_apply_diag_coulomb_evolution_noalloc(void* vec, int dtype, int norb, int n_alpha,
                                               int n_beta, const double* mat_aa,
                                               const double* mat_ab, const double* mat_bb,
                                               double time, int z_representation, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FFSIM_AMD_H */
