// CPU emulation of the fused Givens passes (test infrastructure).
//
// Runs the host planner (window passes, register sub-passes, layer pairs),
// the window fiber tables and the register group tables exactly as the GPU
// fiber kernel consumes them, but with plain loops, so that the host side of
// the orbital rotation can be checked against the oracle without a GPU.
// Built by tests/test_host_planner.py with g++.
#include "../../ffsim_amd/csrc/host_core.hpp"
#include <cstring>
#include <complex>
using namespace ffsim_amd;
typedef std::complex<double> C;

static std::vector<std::pair<int,int>> sub_pairs(int w2, int k2, int p) {
  std::vector<std::pair<int,int>> out; auto strs = make_strings(w2, k2);
  for (size_t r = 0; r < strs.size(); ++r) { uint64_t s = strs[r];
    if (((s >> p) & 1) && !((s >> (p + 1)) & 1)) out.push_back({(int)r, (int)string_rank(s ^ (3ull << p))}); }
  return out;
}
static void run_pass(const C* src, C* dst, long ld_src, long ld_dst, int n_batch, bool tstore,
                     const WindowLayout& L, bool all, const GroupTables& GT, const std::vector<KStep>& steps,
                     const std::vector<C>* diag, const std::vector<uint64_t>& strings) {
  std::vector<SubPass> subs; std::vector<LayerPair> lps;
  plan_sub_passes(steps, L.w, kMaxSubWindow, subs, lps);
  int nf = all ? (int)L.fibers.size() : L.n_active;
  for (int f = 0; f < nf; ++f) {
    const Fiber& fd = L.fibers[f];
    const int32_t* rows = L.row_table.data() + fd.row_off;
    for (int b = 0; b < n_batch; ++b) {
      std::vector<C> v(fd.n_rows);
      for (int i = 0; i < fd.n_rows; ++i) v[i] = src[(long)rows[i] * ld_src + b];
      for (auto& sp : subs) {
        int gi = (fd.kp * GT.n_lo2 + sp.lo2) * 4; int goff = GT.index[gi], ng = GT.index[gi + 1];
        for (int g = 0; g < ng; ++g) {
          uint32_t d = GT.desc[goff + g]; int k2 = d & 0x7f; bool flip = (d >> 7) & 1;
          const uint16_t* mem = GT.members.data() + (d >> 8);
          int N = binom(GT.w2, k2); std::vector<C> r(N);
          for (int m = 0; m < N; ++m) r[m] = v[mem[m]];
          for (int l = 0; l < sp.n_lp; ++l) { const LayerPair& LP = lps[sp.lp0 + l];
            for (int p : {0, 2, 4, 1, 3}) { if (p + 1 >= GT.w2 || !(LP.mask & (1 << p))) continue;
              // complemented group: x and y exchanged, i.e. sigma -> -conj(sigma)
              C sg(flip ? -LP.sr[p] : LP.sr[p], LP.si[p]);
              for (auto& xy : sub_pairs(GT.w2, k2, p)) { C a = r[xy.first], bb = r[xy.second];
                r[xy.first] = LP.c[p] * a + sg * bb; r[xy.second] = LP.c[p] * bb - std::conj(sg) * a; } } }
          for (int m = 0; m < N; ++m) v[mem[m]] = r[m];
        }
      }
      for (int i = 0; i < fd.n_rows; ++i) {
        C val = v[i];
        if (diag) { uint64_t bits = strings[rows[i]]; for (int o = 0; o < 64; ++o) if ((bits >> o) & 1) val *= (*diag)[o]; }
        if (tstore) dst[(long)b * ld_dst + rows[i]] = val; else dst[(long)rows[i] * ld_dst + b] = val;
      }
    }
  }
}

static GivensDecomposition decompose_in(const C* mat, int norb, int clements) {
  return clements ? givens_decomposition(mat, norb, 1e-12) : triangle_decomposition(mat, norb, 1e-12);
}

static void sector(C* state, long ld, int n_batch, int norb, int nocc, const C* mat, int w, C* tdst, long ld_t, int merge, int clements) {
  std::vector<KStep> steps; std::vector<C> diag(norb, 1.0); bool has_diag = false;
  if (mat) { auto g = decompose_in(mat, norb, clements); for (auto& r : g.rotations) steps.push_back(kernel_step_from_rotation(r)); diag = g.diag; has_diag = true; }
  auto passes = plan_passes(steps, norb, w, tdst != nullptr, -1);
  auto strings = make_strings(norb, nocc);
  auto GT = make_group_tables(w, kMaxSubWindow, merge != 0);
  for (auto& P : passes) {
    auto L = make_window_layout(norb, nocc, P.lo, w);
    bool ts = P.last && tdst; bool dg = P.last && has_diag;
    run_pass(state, ts ? tdst : state, ld, ts ? ld_t : ld, n_batch, ts, L, ts || dg, GT, P.steps, dg ? &diag : nullptr, strings);
  }
}

// clements = 0: the triangle (Reck) order; 1: ffsim's (Clements) order.  The library
// picks one per norb and window width, by the window passes and layer pairs
// that each needs.
extern "C" void emulate_order(double* vec, double* work, int norb, int na, int nb, const double* ma, const double* mb, int w, int merge, int clements) {
  long da = binom(norb, na), db = binom(norb, nb);
  C* v = (C*)vec; C* wk = (C*)work;
  sector(v, db, (int)db, norb, na, (const C*)ma, w, wk, da, merge, clements);  // alpha rows, tstore into work
  sector(wk, da, (int)da, norb, nb, (const C*)mb, w, v, db, merge, clements);  // beta rows on transposed, tstore back
}

extern "C" void emulate(double* vec, double* work, int norb, int na, int nb, const double* ma, const double* mb, int w, int merge) {
  emulate_order(vec, work, norb, na, nb, ma, mb, w, merge, 0);
}

// The decomposition the passes run: rotations (c, s, i, j) in application
// order and the diagonal; returns the rotation count (at most n(n-1)/2).
extern "C" int decompose(const double* mat, int n, int clements, double* c, double* s, int* i, int* j, double* diag) {
  auto g = decompose_in((const C*)mat, n, clements);
  for (size_t k = 0; k < g.rotations.size(); ++k) {
    c[k] = g.rotations[k].c; s[2 * k] = g.rotations[k].s.real(); s[2 * k + 1] = g.rotations[k].s.imag();
    i[k] = g.rotations[k].i; j[k] = g.rotations[k].j;
  }
  for (int k = 0; k < n; ++k) { diag[2 * k] = g.diag[k].real(); diag[2 * k + 1] = g.diag[k].imag(); }
  return (int)g.rotations.size();
}

// Window passes and register sub-passes the planner makes for one sector
// rotation by mat at window width w (with the transposing last pass).
extern "C" void plan_counts(const double* mat, int n, int w, int clements, int* n_passes, int* n_sub) {
  auto g = decompose_in((const C*)mat, n, clements);
  std::vector<KStep> steps; for (auto& r : g.rotations) steps.push_back(kernel_step_from_rotation(r));
  auto passes = plan_passes(steps, n, w, true, -1);
  *n_passes = (int)passes.size(); *n_sub = 0;
  for (auto& P : passes) { std::vector<SubPass> subs; std::vector<LayerPair> lps; plan_sub_passes(P.steps, w, kMaxSubWindow, subs, lps); *n_sub += (int)subs.size(); }
}

// The one-body window passes exactly as the library plans them (mask windows,
// link tables, per-pass coefficients, split-window fiber signs), with plain
// loops: out = sum_{pq} M_pq (E^alpha_pq + E^beta_pq) vec.  w: window width.
// Returns the number of passes per sector (0: the library uses its gather path).
extern "C" int emulate_one_body(const double* vec, double* out, int norb, int na, int nb,
                                const double* mat, int w) {
  auto passes = one_body_passes(norb, w);
  if (passes.empty()) return 0;
  const long da = binom(norb, na), db = binom(norb, nb);
  const C* v = (const C*)vec; C* o = (C*)out; const C* M = (const C*)mat;
  for (long k = 0; k < da * db; ++k) o[k] = 0.0;
  for (int sector = 0; sector < 2; ++sector) {
    const int nocc = sector == 0 ? na : nb;
    const long n_batch = sector == 0 ? db : da;
    for (auto& P : passes) {
      auto L = make_mask_layout(norb, nocc, P.mask, P.between);
      auto B = make_one_body_links(L.w);
      std::vector<int> loc;
      for (int b = 0; b < norb; ++b) if ((P.mask >> b) & 1) loc.push_back(b);
      for (auto& fd : L.fibers) {
        const int32_t* rows = L.row_table.data() + fd.row_off;
        const int off = B.index[2 * fd.kp], nl = B.index[2 * fd.kp + 1];
        for (int i = 0; i < fd.n_rows; ++i)
          for (long b = 0; b < n_batch; ++b) {
            C acc = 0.0;
            for (int l = 0; l < nl; ++l) {
              uint32_t u = B.links[off + (long)l * fd.n_rows + i];
              int S = u & 0xffff, p = (u >> 16) & 31, q = (u >> 21) & 31;
              int gp = loc[p], gq = loc[q];
              if (!P.take[(size_t)gp * norb + gq]) continue;
              C x = sector == 0 ? v[(long)rows[S] * db + b] : v[b * db + rows[S]];
              acc += ((u >> 26) ? -1.0 : 1.0) * M[(size_t)gp * norb + gq] * x;
            }
            if (fd.pad) acc = -acc;
            if (sector == 0) o[(long)rows[i] * db + b] += acc; else o[b * db + rows[i]] += acc;
          }
      }
    }
  }
  return (int)passes.size();
}
