// Host-side algorithms of the ffsim_amd library: determinant-string tables,
// window fibers, the Clements Givens decomposition and the fused-pass planner.
// Everything here is plain FP64 / integer C++; the HIP side only consumes the
// tables and programs these routines produce.
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

namespace ffsim_amd {

using cdouble = std::complex<double>;

// ---------------------------------------------------------------- binomials
// C(n, k) for 0 <= n, k <= 64 (0 when k > n).
uint64_t binom(int n, int k);

// ------------------------------------------------------------ string tables
// All k-subsets of norb orbitals as bitmasks in ascending integer order.  The
// position of a string in this list is its address (the colex rank), the
// ordering ffsim uses for CI strings (python/ffsim/_cistring.py:21-24 via
// pyscf cistring.make_strings; tests/python/states/bitstring_test.py:24-97).
std::vector<uint64_t> make_strings(int norb, int nocc);
// Colex rank of a bitmask: sum over set bits t-th (from 0) at position c_t of
// C(c_t, t + 1).
uint64_t string_rank(uint64_t s);

// Fibers of a window of w consecutive orbitals [lo, lo + w).  A fiber is the
// set of strings that agree on every orbital outside the window; its local
// index is the colex rank of the window bits.  Any Givens rotation on an
// adjacent pair (p, p + 1) inside the window maps each fiber onto itself, and
// no orbital lies between p and p + 1, so no fermionic sign appears.
struct Fiber {
  int row_off;  // offset of this fiber in the window's row table
  int n_rows;   // C(w, kp)
  int kp;       // electrons inside the window
  int pad;
};

struct WindowLayout {
  int lo = 0, w = 0;
  std::vector<Fiber> fibers;      // active fibers (0 < kp < w) first, then trivial ones
  int n_active = 0;
  std::vector<int32_t> row_table; // local -> global address, fibers concatenated
  int max_rows = 0;               // largest fiber over all fibers
  int max_rows_active = 0;
  bool contiguous = false;        // every fiber is a contiguous address range
};
WindowLayout make_window_layout(int norb, int nocc, int lo, int w);

// Fibers of a window given as an orbital mask (not necessarily contiguous):
// strings grouped by their bits outside the mask, each fiber's strings in
// ascending order of their bits inside it (window-local strings deposited into
// the mask's orbitals in ascending order).  Every fiber is listed, kp = 0 and
// kp = w included.  Fiber::pad holds the parity of the fiber's fixed (outside)
// electrons on the orbitals of between_mask.
WindowLayout make_mask_layout(int norb, int nocc, uint64_t win_mask, uint64_t between_mask);

// One-body links of window-local strings (w orbitals, kp electrons): for each
// local string T in ascending order, the kp (w - kp + 1) pairs (p, q) with p in
// T and q = p or q not in T, and S = T - p + q, so that a+_p a_q |S> = sign |T>
// with sign the parity of S's electrons strictly between p and q (local
// orbitals).  Entries are packed S | p << 16 | q << 21 | (sign < 0) << 26;
// index[2 kp] is the offset of the kp block and index[2 kp + 1] the links per
// string; link l of string i is entry index[2 kp] + l * C(w, kp) + i (link
// major, so that consecutive strings read consecutive words).  w <= 31.
struct OneBodyLinks {
  int w = 0;
  std::vector<uint32_t> links;
  std::vector<int32_t> index;
};
OneBodyLinks make_one_body_links(int w);

// Window passes of the one-body contraction of one spin sector: with w the
// widest window whose fibers fit a tile, norb <= w is one pass over all
// orbitals; norb <= 1.5 w splits the orbitals into A = [0, norb - w),
// B = [norb - w, w), C = [w, norb) and runs the windows A+B (pairs inside it),
// B+C (pairs inside it and not inside A+B) and A+C (pairs between A and C, the
// electrons on B giving each fiber's sign).  Every ordered pair (p, q) is
// applied by exactly one pass.  Returns no passes for wider norb.
struct OneBodyPass {
  uint64_t mask = 0, between = 0;
  std::vector<char> take;  // norb x norb: pair (p, q) applied by this pass
};
std::vector<OneBodyPass> one_body_passes(int norb, int w);

// Local pair tables for a window of width w: for each kp in [0, w] and each
// adjacent local pair (p, p + 1), the list of (x, y) local ranks with x having
// p occupied / p + 1 empty and y = x with the electron moved to p + 1.
// Entries are packed as x | (y << 16).
struct PairTables {
  int w = 0;
  std::vector<uint32_t> pairs;
  std::vector<int32_t> index;  // [(kp * (w - 1) + p) * 2 + {0: offset, 1: count}]
};
PairTables make_pair_tables(int w);

// Register groups of a sub-window [lo2, lo2 + w2) inside a window of width w:
// for every kp and every sub-window start lo2 in [0, w - w2], the window-local
// strings with kp electrons are split into groups that agree outside the
// sub-window.  A group with k2 electrons inside the sub-window is of class k2: C(w2, k2) members in ascending order of their sub-window bits, so
// member m is the m-th k2-subset of w2 bits.  With merge_complements, a group
// with k2 > w2 / 2 is of class w2 - k2, complemented (descriptor bit 7): member
// m has the complement of the m-th (w2 - k2)-subset.  Groups with k2 == 0 or k2 == w2 are not
// stored (no rotation inside the sub-window changes them).  Groups are sorted
// by decreasing size, then class, plain before complemented.
struct GroupTables {
  int w = 0, w2 = 0, n_lo2 = 0;
  std::vector<uint32_t> desc;     // member_offset << 8 | complemented << 7 | class
  std::vector<uint16_t> members;  // window-local ranks
  // [(kp * n_lo2 + lo2) * 4 + {0: desc offset, 1: group count, 2: member offset,
  //                           3: member count}]; a (kp, lo2) set is contiguous
  std::vector<int32_t> index;
};
// merge_complements: store groups with k2 > w2 / 2 as complemented groups of
// class w2 - k2 (otherwise every group is of class k2, never complemented).
GroupTables make_group_tables(int w, int w2, bool merge_complements);

// Static wave schedule of the register groups: for every (kp, lo2) the
// groups are cut into chunks of at most `per_chunk` consecutive groups with
// the same class (one chunk fills a wave: per_chunk groups x W batch columns),
// and the chunks are assigned to the block's waves longest-first, balancing
// the rotation work per SIMD (wave i runs on SIMD i % 4).
struct WaveSchedule {
  int n_waves = 0;
  std::vector<uint32_t> chunks;  // first_group << 8 | n_groups
  std::vector<int32_t> index;    // [((kp * n_lo2 + lo2) * n_waves + wave) * 2 + {offset, count}]
};
WaveSchedule make_wave_schedule(const GroupTables& T, int n_waves, int per_chunk);

// ------------------------------------------------------ Givens decomposition
// Exact Clements decomposition U = D G_L^* ... G_1^* with each G acting on
// adjacent orbitals (i, j) as [[c, s], [-conj(s), c]] (c real), matching the
// output contract of ffsim.linalg.givens_decomposition
// (python/ffsim/linalg/givens.py:59-156, src/linalg/givens.rs:69-149).
struct GivensRotation {
  double c;
  cdouble s;
  int i, j;
};
struct GivensDecomposition {
  std::vector<GivensRotation> rotations;  // in application order G_1, G_2, ...
  std::vector<cdouble> diag;              // D
};
// mat is row-major n x n.
GivensDecomposition givens_decomposition(const cdouble* mat, int n, double tol);
// The same factorization form U = D G_N^* ... G_1^* in triangle (Reck) order:
// rows are nulled from the bottom, each by a left-to-right sweep of column
// rotations.  The fused passes run this order: at 12-orbital windows it needs
// 3 passes up to norb = 18 (Clements needs 4 from norb = 17) and at most as
// many register sub-passes.
GivensDecomposition triangle_decomposition(const cdouble* mat, int n, double tol);

// ----------------------------------------------------- fused-pass planning
// A rotation in kernel form: on the local pair x = (p occupied, p+1 empty),
// y = (p empty, p+1 occupied) it applies x' = c x + sigma y,
// y' = c y - conj(sigma) x.
struct KStep {
  int p;
  int pad;
  double c, sr, si;
};
static_assert(sizeof(KStep) == 32, "KStep layout is shared with the device");

// Converts a decomposition rotation (c, s, i, j) into the kernel step that
// reproduces ffsim's application (python/ffsim/gates/orbital_rotation.py:132-135
// applies (c, conj(s)) to the rows with i occupied / j occupied).
KStep kernel_step_from_rotation(const GivensRotation& g);

struct PlannedPass {
  int lo;                    // window start (orbital)
  std::vector<KStep> steps;  // window-local p
  bool last = false;         // final pass of the network
};
// A window pass is split into register sub-passes over sub-windows of width
// w2 <= kMaxSubWindow.  Inside a sub-pass the rotations are packed into
// brickwork layer pairs: an even layer (pairs p = 0, 2, 4 of the sub-window)
// followed by an odd layer (p = 1, 3); slot p of a layer pair holds the
// rotation on (p, p + 1) or is empty (bit p of `mask` clear).
// Orbitals whose phases a fiber pass can apply (strings are 64-bit masks).
constexpr int kMaxDiagOrbitals = 64;
constexpr int kMaxSubWindow = 6;
constexpr int kSlots = kMaxSubWindow - 1;

struct LayerPair {
  double c[kSlots], sr[kSlots], si[kSlots];
  int32_t mask;
  int32_t pad;
};
static_assert(sizeof(LayerPair) == 128, "LayerPair layout is shared with the device");

struct SubPass {
  int lo2;
  int lp0, n_lp;  // range in the pass's layer-pair list
  int slot;       // index of lo2 among the distinct sub-window starts of the pass
};
// Splits the steps of one window pass (window-local p) into sub-passes and
// packs each into layer pairs (ASAP on every orbital, layer parity = p parity).
void plan_sub_passes(const std::vector<KStep>& window_steps, int w, int w2,
                     std::vector<SubPass>& subs, std::vector<LayerPair>& layers);

// Greedy window schedule: alternate between the windows [0, w) and
// [norb - w, norb), each pass applying every rotation whose pair lies inside
// the window and whose predecessors (earlier rotations sharing an orbital)
// are already applied.  If end_low is set, the last pass is forced onto the
// window that starts at orbital 0 (its fibers are contiguous address ranges).
std::vector<PlannedPass> plan_passes(const std::vector<KStep>& global_steps, int norb, int w,
                                     bool end_low, int first_window);

}  // namespace ffsim_amd
