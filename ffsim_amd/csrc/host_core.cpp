// Host-side algorithms of the ffsim_amd library (see host_core.hpp).
#include "host_core.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>

namespace ffsim_amd {

namespace {
struct BinomTable {
  uint64_t v[65][65];
  BinomTable() {
    for (int n = 0; n <= 64; ++n) {
      for (int k = 0; k <= 64; ++k) v[n][k] = 0;
      v[n][0] = 1;
      for (int k = 1; k <= n; ++k) v[n][k] = v[n - 1][k - 1] + (k <= n - 1 ? v[n - 1][k] : 0);
    }
  }
};
const BinomTable& binoms() {
  static const BinomTable t;
  return t;
}

// next bitmask with the same popcount (Gosper)
inline uint64_t next_same_popcount(uint64_t x) {
  uint64_t c = x & (~x + 1);
  uint64_t r = x + c;
  return (((r ^ x) >> 2) / c) | r;
}
}  // namespace

uint64_t binom(int n, int k) {
  if (n < 0 || k < 0 || n > 64 || k > 64 || k > n) return 0;
  return binoms().v[n][k];
}

std::vector<uint64_t> make_strings(int norb, int nocc) {
  std::vector<uint64_t> out;
  if (norb < 0 || nocc < 0 || nocc > norb || norb > 63) return out;
  uint64_t count = binom(norb, nocc);
  out.reserve(count);
  if (nocc == 0) {
    out.push_back(0);
    return out;
  }
  uint64_t s = (nocc == 64) ? ~0ull : ((1ull << nocc) - 1);
  for (uint64_t a = 0; a < count; ++a) {
    out.push_back(s);
    if (a + 1 < count) s = next_same_popcount(s);
  }
  return out;
}

uint64_t string_rank(uint64_t s) {
  uint64_t r = 0;
  int t = 0;
  while (s) {
    int c = __builtin_ctzll(s);
    r += binom(c, t + 1);
    ++t;
    s &= s - 1;
  }
  return r;
}

WindowLayout make_window_layout(int norb, int nocc, int lo, int w) {
  if (lo < 0 || w < 1 || lo + w > norb) throw std::invalid_argument("bad window");
  WindowLayout L;
  L.lo = lo;
  L.w = w;
  L.contiguous = (lo == 0);
  const uint64_t win_mask = (w == 64 ? ~0ull : ((1ull << w) - 1)) << lo;
  const uint64_t all_mask = (norb == 64 ? ~0ull : ((1ull << norb) - 1));
  const uint64_t out_mask = all_mask & ~win_mask;
  const int n_out = norb - w;
  struct Tmp {
    uint64_t fixed;
    int kp;
  };
  std::vector<Tmp> act, triv;
  // enumerate subsets of the outside orbitals in ascending order of the fixed pattern
  std::vector<int> out_pos;
  for (int b = 0; b < norb; ++b)
    if ((out_mask >> b) & 1) out_pos.push_back(b);
  for (uint64_t m = 0; m < (1ull << n_out); ++m) {
    uint64_t fixed = 0;
    for (int t = 0; t < n_out; ++t)
      if ((m >> t) & 1) fixed |= 1ull << out_pos[t];
    int e = __builtin_popcountll(fixed);
    int kp = nocc - e;
    if (kp < 0 || kp > w) continue;
    if (kp > 0 && kp < w)
      act.push_back({fixed, kp});
    else
      triv.push_back({fixed, kp});
  }
  // largest fibers first so that the heaviest tiles start early
  std::stable_sort(act.begin(), act.end(),
                   [&](const Tmp& a, const Tmp& b) { return binom(w, a.kp) > binom(w, b.kp); });
  auto emit = [&](const Tmp& t) {
    Fiber f;
    f.row_off = (int)L.row_table.size();
    f.kp = t.kp;
    f.n_rows = (int)binom(w, t.kp);
    f.pad = 0;
    std::vector<uint64_t> locals = make_strings(w, t.kp);
    for (uint64_t loc : locals) L.row_table.push_back((int32_t)string_rank(t.fixed | (loc << lo)));
    L.max_rows = std::max(L.max_rows, f.n_rows);
    L.fibers.push_back(f);
  };
  for (auto& t : act) {
    emit(t);
    L.max_rows_active = std::max(L.max_rows_active, (int)binom(w, t.kp));
  }
  L.n_active = (int)act.size();
  for (auto& t : triv) emit(t);
  return L;
}

WindowLayout make_mask_layout(int norb, int nocc, uint64_t win_mask, uint64_t between_mask) {
  const uint64_t all_mask = (norb == 64 ? ~0ull : ((1ull << norb) - 1));
  if ((win_mask & ~all_mask) || !win_mask) throw std::invalid_argument("bad window mask");
  WindowLayout L;
  L.lo = 0;
  L.w = __builtin_popcountll(win_mask);
  L.contiguous = false;
  const int w = L.w;
  std::vector<int> in_pos, out_pos;
  for (int b = 0; b < norb; ++b) ((win_mask >> b) & 1 ? in_pos : out_pos).push_back(b);
  const int n_out = (int)out_pos.size();
  auto deposit = [&](uint64_t loc) {
    uint64_t s = 0;
    for (int t = 0; t < w; ++t)
      if ((loc >> t) & 1) s |= 1ull << in_pos[t];
    return s;
  };
  for (uint64_t m = 0; m < (1ull << n_out); ++m) {
    uint64_t fixed = 0;
    for (int t = 0; t < n_out; ++t)
      if ((m >> t) & 1) fixed |= 1ull << out_pos[t];
    const int kp = nocc - __builtin_popcountll(fixed);
    if (kp < 0 || kp > w) continue;
    Fiber f;
    f.row_off = (int)L.row_table.size();
    f.kp = kp;
    f.n_rows = (int)binom(w, kp);
    f.pad = __builtin_popcountll(fixed & between_mask) & 1;
    for (uint64_t loc : make_strings(w, kp))
      L.row_table.push_back((int32_t)string_rank(fixed | deposit(loc)));
    L.max_rows = std::max(L.max_rows, f.n_rows);
    L.fibers.push_back(f);
  }
  L.n_active = (int)L.fibers.size();
  L.max_rows_active = L.max_rows;
  return L;
}

OneBodyLinks make_one_body_links(int w) {
  if (w < 1 || w > 31) throw std::invalid_argument("one-body window width must be in [1, 31]");
  OneBodyLinks B;
  B.w = w;
  B.index.assign((size_t)(w + 1) * 2, 0);
  for (int kp = 0; kp <= w; ++kp) {
    const size_t base = B.links.size();
    const int nl = kp * (w - kp + 1);
    const std::vector<uint64_t> locals = make_strings(w, kp);
    const size_t n_rows = locals.size();
    B.index[2 * kp] = (int32_t)base;
    B.index[2 * kp + 1] = nl;
    B.links.resize(base + n_rows * nl);
    for (size_t i = 0; i < n_rows; ++i) {
      const uint64_t t = locals[i];
      int l = 0;
      for (int p = 0; p < w; ++p) {
        if (!((t >> p) & 1)) continue;
        for (int q = 0; q < w; ++q) {
          if (q != p && ((t >> q) & 1)) continue;
          const uint64_t src = (t & ~(1ull << p)) | (1ull << q);
          int neg = 0;
          if (p != q) {
            const int lo = std::min(p, q), hi = std::max(p, q);
            const uint64_t between = ((1ull << hi) - 1) & ~((1ull << (lo + 1)) - 1);
            neg = __builtin_popcountll(src & between) & 1;
          }
          B.links[base + (size_t)l * n_rows + i] = (uint32_t)string_rank(src) |
                                                   ((uint32_t)p << 16) | ((uint32_t)q << 21) |
                                                   ((uint32_t)neg << 26);
          ++l;
        }
      }
    }
  }
  return B;
}

std::vector<OneBodyPass> one_body_passes(int norb, int w) {
  std::vector<OneBodyPass> out;
  if (norb < 1 || norb > 62 || w < 1) return out;
  const uint64_t all = (1ull << norb) - 1;
  auto pass = [&](uint64_t mask, uint64_t between, auto take) {
    OneBodyPass P;
    P.mask = mask;
    P.between = between;
    P.take.assign((size_t)norb * norb, 0);
    for (int p = 0; p < norb; ++p)
      for (int q = 0; q < norb; ++q) P.take[(size_t)p * norb + q] = take(p, q) ? 1 : 0;
    out.push_back(std::move(P));
  };
  if (norb <= w) {
    pass(all, 0, [](int, int) { return true; });
  } else if (2 * (norb - w) <= w) {
    const int a = norb - w;  // A = [0, a), B = [a, w), C = [w, norb)
    const uint64_t mA = (1ull << a) - 1, mB = ((1ull << w) - 1) & ~mA, mC = all & ~((1ull << w) - 1);
    pass(mA | mB, 0, [&](int p, int q) { return p < w && q < w; });
    pass(mB | mC, 0, [&](int p, int q) { return p >= a && q >= a && !(p < w && q < w); });
    pass(mA | mC, mB, [&](int p, int q) { return (p < a && q >= w) || (p >= w && q < a); });
  }
  return out;
}


PairTables make_pair_tables(int w) {
  PairTables T;
  T.w = w;
  int wm1 = std::max(w - 1, 1);
  T.index.assign((size_t)(w + 1) * wm1 * 2, 0);
  for (int kp = 0; kp <= w; ++kp) {
    std::vector<uint64_t> locals = make_strings(w, kp);
    for (int p = 0; p + 1 < w; ++p) {
      int off = (int)T.pairs.size();
      for (size_t r = 0; r < locals.size(); ++r) {
        uint64_t s = locals[r];
        if (((s >> p) & 1) && !((s >> (p + 1)) & 1)) {
          uint64_t y = s ^ (3ull << p);
          uint32_t ry = (uint32_t)string_rank(y);
          T.pairs.push_back((uint32_t)r | (ry << 16));
        }
      }
      T.index[((size_t)kp * wm1 + p) * 2 + 0] = off;
      T.index[((size_t)kp * wm1 + p) * 2 + 1] = (int)T.pairs.size() - off;
    }
  }
  return T;
}

GroupTables make_group_tables(int w, int w2, bool merge_complements) {
  GroupTables T;
  T.w = w;
  T.w2 = std::min(w2, w);
  T.n_lo2 = std::max(w - T.w2 + 1, 1);
  T.index.assign((size_t)(w + 1) * T.n_lo2 * 4, 0);
  if (T.w2 < 2) return T;
  const uint64_t sub_bits = (1ull << T.w2) - 1;
  for (int kp = 0; kp <= w; ++kp) {
    std::vector<uint64_t> locals = make_strings(w, kp);
    for (int lo2 = 0; lo2 + T.w2 <= w; ++lo2) {
      const uint64_t sub_mask = sub_bits << lo2;
      // distinct outside patterns in ascending order of first appearance
      std::vector<uint64_t> outs;
      for (uint64_t s : locals) {
        uint64_t o = s & ~sub_mask;
        if (std::find(outs.begin(), outs.end(), o) == outs.end()) outs.push_back(o);
      }
      // A group with k2 > w2 / 2 electrons in the sub-window is stored as the
      // complement of a class-(w2 - k2) group: member m is the string whose
      // sub-window bits are the complement of the m-th (w2 - k2)-subset.  A
      // rotation then pairs the same members with the roles of x and y swapped,
      // which the kernel absorbs by negating Re(sigma) for that lane, so groups
      // of k2 and w2 - k2 share one code path and can share a wave.
      struct G {
        uint64_t out;
        int cls;
        bool flip;
      };
      std::vector<G> gs;
      for (uint64_t o : outs) {
        int k2 = kp - __builtin_popcountll(o);
        if (k2 > 0 && k2 < T.w2) {
          const bool flip = merge_complements && 2 * k2 > T.w2;
          gs.push_back({o, flip ? T.w2 - k2 : k2, flip});
        }
      }
      std::stable_sort(gs.begin(), gs.end(), [&](const G& a, const G& b) {
        const uint64_t sa = binom(T.w2, a.cls), sb = binom(T.w2, b.cls);
        if (sa != sb) return sa > sb;
        if (a.cls != b.cls) return a.cls < b.cls;
        return a.flip < b.flip;
      });
      auto member_ranks = [&](const G& g) {
        std::vector<uint16_t> r;
        for (uint64_t sl : make_strings(T.w2, g.cls)) {
          const uint64_t bits = g.flip ? (~sl & sub_bits) : sl;
          r.push_back((uint16_t)string_rank(g.out | (bits << lo2)));
        }
        return r;
      };
      if (!merge_complements) {
        // c128 tiles hold one string per 128 bytes, and an LDS access serves the
        // strings of two neighbouring lane groups together: pair each group with
        // one whose member ranks differ in parity for as many members as
        // possible, so the two strings sit on different bank halves.
        std::vector<G> paired;
        paired.reserve(gs.size());
        for (size_t i = 0; i < gs.size();) {
          size_t j = i;
          while (j < gs.size() && gs[j].cls == gs[i].cls && gs[j].flip == gs[i].flip) ++j;
          std::vector<G> left(gs.begin() + i, gs.begin() + j);
          std::vector<std::vector<uint16_t>> ranks;
          for (const G& g : left) ranks.push_back(member_ranks(g));
          std::vector<bool> used(left.size(), false);
          for (size_t a = 0; a < left.size(); ++a) {
            if (used[a]) continue;
            used[a] = true;
            paired.push_back(left[a]);
            size_t best = left.size();
            int best_same = 1 << 30;
            for (size_t b = a + 1; b < left.size(); ++b) {
              if (used[b]) continue;
              int same = 0;
              for (size_t m = 0; m < ranks[a].size(); ++m)
                same += (ranks[a][m] & 1) == (ranks[b][m] & 1);
              if (same < best_same) {
                best_same = same;
                best = b;
              }
            }
            if (best < left.size()) {
              used[best] = true;
              paired.push_back(left[best]);
            }
          }
          i = j;
        }
        gs.swap(paired);
      }
      const int off = (int)T.desc.size();
      const int moff0 = (int)T.members.size();
      for (const G& g : gs) {
        const uint32_t moff = (uint32_t)T.members.size();
        for (uint16_t r : member_ranks(g)) T.members.push_back(r);
        T.desc.push_back((moff << 8) | (g.flip ? 0x80u : 0u) | (uint32_t)g.cls);
      }
      T.index[((size_t)kp * T.n_lo2 + lo2) * 4 + 0] = off;
      T.index[((size_t)kp * T.n_lo2 + lo2) * 4 + 1] = (int)gs.size();
      T.index[((size_t)kp * T.n_lo2 + lo2) * 4 + 2] = moff0;
      T.index[((size_t)kp * T.n_lo2 + lo2) * 4 + 3] = (int)T.members.size() - moff0;
    }
  }
  return T;
}

WaveSchedule make_wave_schedule(const GroupTables& T, int n_waves, int per_chunk) {
  WaveSchedule ws;
  ws.n_waves = n_waves;
  ws.index.assign((size_t)(T.w + 1) * T.n_lo2 * n_waves * 2, 0);
  const int n_simd = std::min(4, n_waves);
  for (int kp = 0; kp <= T.w; ++kp) {
    for (int lo2 = 0; lo2 < T.n_lo2; ++lo2) {
      const int* ix = T.index.data() + ((size_t)kp * T.n_lo2 + lo2) * 4;
      const int d0 = ix[0], nd = ix[1];
      struct Ch {
        int g0, ng;
        double cost;
      };
      std::vector<Ch> chs;
      for (int g = 0; g < nd;) {
        const int k2 = T.desc[d0 + g] & 0x7f;  // group class (flag bit 7: complemented)
        int e = g;
        while (e < nd && e - g < per_chunk && (int)(T.desc[d0 + e] & 0x7f) == k2) ++e;
        // rotation work per lane ~ pairs per pair position; plus the group's loads/stores
        const double pairs = (double)binom(T.w2 - 2, k2 - 1);
        const double size = (double)binom(T.w2, k2);
        chs.push_back({g, e - g, 6.0 * pairs + size});
        g = e;
      }
      std::stable_sort(chs.begin(), chs.end(),
                       [](const Ch& a, const Ch& b) { return a.cost > b.cost; });
      std::vector<double> simd_load(n_simd, 0.0), wave_load(n_waves, 0.0);
      std::vector<std::vector<uint32_t>> per_wave(n_waves);
      for (const Ch& c : chs) {
        int s = 0;
        for (int k = 1; k < n_simd; ++k)
          if (simd_load[k] < simd_load[s]) s = k;
        int wv = s;
        for (int k = s; k < n_waves; k += n_simd)
          if (wave_load[k] < wave_load[wv]) wv = k;
        simd_load[s] += c.cost;
        wave_load[wv] += c.cost;
        per_wave[wv].push_back(((uint32_t)c.g0 << 8) | (uint32_t)c.ng);
      }
      for (int wv = 0; wv < n_waves; ++wv) {
        int32_t* out = ws.index.data() + (((size_t)kp * T.n_lo2 + lo2) * n_waves + wv) * 2;
        out[0] = (int32_t)ws.chunks.size();
        out[1] = (int32_t)per_wave[wv].size();
        ws.chunks.insert(ws.chunks.end(), per_wave[wv].begin(), per_wave[wv].end());
      }
    }
  }
  return ws;
}

// Greedy sub-window schedule: at every step take the sub-window start whose
// closure (rotations inside it with all predecessors applied) is largest.
std::vector<PlannedPass> plan_greedy_max(const std::vector<KStep>& steps, int n, int w2) {
  std::vector<PlannedPass> out;
  if (w2 >= n) {
    PlannedPass P;
    P.lo = 0;
    P.steps = steps;
    out.push_back(P);
    return out;
  }
  std::vector<char> done(steps.size(), 0);
  size_t n_done = 0;
  while (n_done < steps.size()) {
    int best_lo = -1;
    std::vector<size_t> best;
    for (int lo = 0; lo + w2 <= n; ++lo) {
      std::vector<size_t> cur;
      std::vector<char> blocked(n + 1, 0);
      for (size_t r = 0; r < steps.size(); ++r) {
        if (done[r]) continue;
        const int p = steps[r].p;
        if (p >= lo && p + 1 < lo + w2 && !blocked[p] && !blocked[p + 1])
          cur.push_back(r);
        else
          blocked[p] = blocked[p + 1] = 1;
      }
      if (cur.size() > best.size()) {
        best = cur;
        best_lo = lo;
      }
    }
    if (best.empty()) throw std::runtime_error("sub-pass planner made no progress");
    PlannedPass P;
    P.lo = best_lo;
    for (size_t r : best) {
      KStep k = steps[r];
      k.p -= best_lo;
      P.steps.push_back(k);
      done[r] = 1;
      ++n_done;
    }
    out.push_back(P);
  }
  return out;
}

namespace {
// Rotations of sub-window [lo, lo + w2) that can run now: in order, those not
// yet applied whose orbitals no earlier pending rotation outside the window has
// touched.
std::vector<size_t> sub_window_closure(const std::vector<KStep>& steps,
                                       const std::vector<char>& done, int lo, int w2, int n) {
  std::vector<size_t> cur;
  std::vector<char> blocked(n + 1, 0);
  for (size_t r = 0; r < steps.size(); ++r) {
    if (done[r]) continue;
    const int p = steps[r].p;
    if (p >= lo && p + 1 < lo + w2 && !blocked[p] && !blocked[p + 1])
      cur.push_back(r);
    else
      blocked[p] = blocked[p + 1] = 1;
  }
  return cur;
}

// Beam search over sequences of sub-window starts for the fewest sub-passes
// (each step takes the whole closure of its sub-window).  Returns the starts.
std::vector<int> plan_sub_window_beam(const std::vector<KStep>& steps, int n, int w2,
                                      size_t width) {
  struct Node {
    std::vector<char> done;
    size_t n_done;
    std::vector<int> starts;
  };
  std::vector<Node> beam{{std::vector<char>(steps.size(), 0), 0, {}}};
  while (true) {
    for (const Node& b : beam)
      if (b.n_done == steps.size()) return b.starts;
    std::vector<Node> next;
    for (const Node& b : beam) {
      for (int lo = 0; lo + w2 <= n; ++lo) {
        std::vector<size_t> c = sub_window_closure(steps, b.done, lo, w2, n);
        if (c.empty()) continue;
        Node x = b;
        for (size_t r : c) x.done[r] = 1;
        x.n_done += c.size();
        x.starts.push_back(lo);
        next.push_back(std::move(x));
      }
    }
    if (next.empty()) throw std::runtime_error("sub-pass planner made no progress");
    std::stable_sort(next.begin(), next.end(),
                     [](const Node& a, const Node& b) { return a.n_done > b.n_done; });
    beam.clear();
    for (Node& x : next) {
      bool dup = false;
      for (const Node& u : beam)
        if (u.done == x.done) {
          dup = true;
          break;
        }
      if (!dup) beam.push_back(std::move(x));
      if (beam.size() >= width) break;
    }
  }
}

std::vector<PlannedPass> passes_from_starts(const std::vector<KStep>& steps, int n, int w2,
                                            const std::vector<int>& starts) {
  std::vector<PlannedPass> out;
  std::vector<char> done(steps.size(), 0);
  for (int lo : starts) {
    PlannedPass P;
    P.lo = lo;
    for (size_t r : sub_window_closure(steps, done, lo, w2, n)) {
      KStep k = steps[r];
      k.p -= lo;
      P.steps.push_back(k);
      done[r] = 1;
    }
    out.push_back(P);
  }
  return out;
}

// The schedule depends only on the sequence of pair positions, which is the
// same for every unitary of a size (Clements order), so searches are cached.
std::mutex g_beam_mu;
std::map<std::vector<int>, std::vector<int>> g_beam_cache;

std::vector<int> cached_beam_starts(const std::vector<KStep>& steps, int n, int w2) {
  std::vector<int> key;
  key.reserve(steps.size() + 2);
  key.push_back(n);
  key.push_back(w2);
  for (const KStep& k : steps) key.push_back(k.p);
  {
    std::lock_guard<std::mutex> lk(g_beam_mu);
    auto it = g_beam_cache.find(key);
    if (it != g_beam_cache.end()) return it->second;
  }
  std::vector<int> starts = plan_sub_window_beam(steps, n, w2, 32);
  std::lock_guard<std::mutex> lk(g_beam_mu);
  if (g_beam_cache.size() > 4096) g_beam_cache.clear();
  g_beam_cache[key] = starts;
  return starts;
}
}  // namespace

namespace {
// The sub-pass plan of a step list depends only on the steps' pair positions;
// plans are kept per position sequence with each used slot holding the index
// of its step, and filled with the coefficients of each call.
std::mutex g_sub_mu;
std::map<std::vector<int>, std::pair<std::vector<SubPass>, std::vector<LayerPair>>> g_sub_cache;

void plan_sub_passes_uncached(const std::vector<KStep>& window_steps, int w, int w2,
                              std::vector<SubPass>& subs, std::vector<LayerPair>& layers);
}  // namespace

void plan_sub_passes(const std::vector<KStep>& window_steps, int w, int w2,
                     std::vector<SubPass>& subs, std::vector<LayerPair>& layers) {
  subs.clear();
  layers.clear();
  if (window_steps.empty()) return;
  std::vector<int> key;
  key.reserve(window_steps.size() + 2);
  key.push_back(w);
  key.push_back(w2);
  for (const KStep& k : window_steps) key.push_back(k.p);
  bool found = false;
  {
    std::lock_guard<std::mutex> lk(g_sub_mu);
    auto it = g_sub_cache.find(key);
    if (it != g_sub_cache.end()) {
      subs = it->second.first;
      layers = it->second.second;
      found = true;
    }
  }
  if (!found) {
    std::vector<KStep> proxy(window_steps);
    for (size_t r = 0; r < proxy.size(); ++r) {
      proxy[r].c = (double)r;
      proxy[r].sr = 0.0;
      proxy[r].si = 0.0;
    }
    plan_sub_passes_uncached(proxy, w, w2, subs, layers);
    std::lock_guard<std::mutex> lk(g_sub_mu);
    if (g_sub_cache.size() > 4096) g_sub_cache.clear();
    g_sub_cache[key] = {subs, layers};
  }
  for (LayerPair& L : layers)
    for (int q = 0; q < kSlots; ++q)
      if ((L.mask >> q) & 1) {
        const KStep& k = window_steps[(size_t)L.c[q]];
        L.c[q] = k.c;
        L.sr[q] = k.sr;
        L.si[q] = k.si;
      }
}

namespace {
void plan_sub_passes_uncached(const std::vector<KStep>& window_steps, int w, int w2,
                              std::vector<SubPass>& subs, std::vector<LayerPair>& layers) {
  subs.clear();
  layers.clear();
  if (window_steps.empty()) return;
  w2 = std::min(w2, w);
  // zigzag over evenly spaced sub-windows or greedy best start: keep the
  // schedule with fewer sub-passes
  std::vector<PlannedPass> sp = plan_passes(window_steps, w, w2, false, -1);
  {
    std::vector<PlannedPass> alt = plan_greedy_max(window_steps, w, w2);
    if (alt.size() < sp.size()) sp.swap(alt);
  }
  if (w2 < w && sp.size() > 2) {
    std::vector<PlannedPass> alt =
        passes_from_starts(window_steps, w, w2, cached_beam_starts(window_steps, w, w2));
    if (alt.size() < sp.size()) sp.swap(alt);
  }
  for (const PlannedPass& P : sp) {
    if (P.steps.empty()) continue;
    // ASAP layering with layer parity fixed by the pair position
    std::vector<int> last(w2 + 1, -1);
    std::vector<int> lay(P.steps.size());
    int n_layers = 0;
    for (size_t r = 0; r < P.steps.size(); ++r) {
      const int p = P.steps[r].p;
      int L = std::max(last[p], last[p + 1]) + 1;
      if ((L & 1) != (p & 1)) ++L;
      lay[r] = L;
      last[p] = last[p + 1] = L;
      n_layers = std::max(n_layers, L + 1);
    }
    const int n_lp = (n_layers + 1) / 2;
    SubPass s{};
    s.lo2 = P.lo;
    s.lp0 = (int)layers.size();
    s.n_lp = n_lp;
    s.slot = -1;
    for (const SubPass& o : subs)
      if (o.lo2 == P.lo) s.slot = o.slot;
    if (s.slot < 0) {
      int n_slots = 0;
      for (const SubPass& o : subs) n_slots = std::max(n_slots, o.slot + 1);
      s.slot = n_slots;
    }
    LayerPair id{};
    for (int k = 0; k < kSlots; ++k) {
      id.c[k] = 1.0;
      id.sr[k] = 0.0;
      id.si[k] = 0.0;
    }
    id.mask = 0;
    layers.insert(layers.end(), n_lp, id);
    for (size_t r = 0; r < P.steps.size(); ++r) {
      const KStep& k = P.steps[r];
      LayerPair& L = layers[s.lp0 + lay[r] / 2];
      if (L.mask & (1 << k.p)) throw std::logic_error("layer slot used twice");
      L.c[k.p] = k.c;
      L.sr[k.p] = k.sr;
      L.si[k.p] = k.si;
      L.mask |= 1 << k.p;
    }
    subs.push_back(s);
  }
}
}  // namespace

// ------------------------------------------------------------------ Givens
namespace {
// Rotation that zeroes b against a: returns (c, s) with c real such that
// c*b - conj(s)*a == 0 for the column/row updates below (BLAS zrotg
// convention with thresholds, as the decomposition contract requires).
void zrotg_tol(cdouble a, cdouble b, double tol, double& c, cdouble& s) {
  double aa = std::abs(a), ab = std::abs(b);
  if (ab <= tol) {
    c = 1.0;
    s = 0.0;
    return;
  }
  if (aa <= tol) {
    c = 0.0;
    s = 1.0;
    return;
  }
  double r = std::hypot(aa, ab);
  c = std::min(1.0, std::max(-1.0, aa / r));
  s = (a / aa) * std::conj(b) / r;
}
}  // namespace

GivensDecomposition givens_decomposition(const cdouble* mat, int n, double tol) {
  GivensDecomposition out;
  std::vector<cdouble> M(mat, mat + (size_t)n * n);
  auto at = [&](int r, int c) -> cdouble& { return M[(size_t)r * n + c]; };
  // column update on (x, y): x <- c x + s y, y <- c y - conj(s) x  (M <- M G^T)
  auto rot_cols = [&](int x, int y, double c, cdouble s) {
    for (int r = 0; r < n; ++r) {
      cdouble xo = at(r, x), yo = at(r, y);
      at(r, x) = c * xo + s * yo;
      at(r, y) = c * yo - std::conj(s) * xo;
    }
  };
  // row update on (x, y) with the same 2x2 (M <- G M)
  auto rot_rows = [&](int x, int y, double c, cdouble s) {
    for (int k = 0; k < n; ++k) {
      cdouble xo = at(x, k), yo = at(y, k);
      at(x, k) = c * xo + s * yo;
      at(y, k) = c * yo - std::conj(s) * xo;
    }
  };
  std::vector<GivensRotation> right, left;
  // Clements et al., Optica 3, 1460 (2016): alternate nulling anti-diagonals
  // from the right (column rotations) and from the left (row rotations).
  for (int d = 0; d + 1 < n; ++d) {
    if (d % 2 == 0) {
      for (int t = 0; t <= d; ++t) {
        int col = d - t, row = n - 1 - t;
        if (std::abs(at(row, col)) > tol) {
          double c;
          cdouble s;
          zrotg_tol(at(row, col + 1), at(row, col), tol, c, s);
          right.push_back({c, s, col + 1, col});
          rot_cols(col + 1, col, c, s);
        }
      }
    } else {
      for (int t = 0; t <= d; ++t) {
        int row = n - 1 - d + t, col = t;
        if (std::abs(at(row, col)) > tol) {
          double c;
          cdouble s;
          zrotg_tol(at(row - 1, col), at(row, col), tol, c, s);
          left.push_back({c, s, row - 1, row});
          rot_rows(row - 1, row, c, s);
        }
      }
    }
  }
  // Now L_K ... L_1 U G_1^T ... G_J^T = diag(d), i.e.
  // U = L_1^H ... L_K^H diag(d) G_J^* ... G_1^*.  Move each L^H through the
  // diagonal: L^H diag(d) = diag(d) G'^* with G' = (c, -conj(s) d_i conj(d_j)).
  std::vector<cdouble> d(n);
  for (int k = 0; k < n; ++k) d[k] = at(k, k);
  out.rotations = right;
  for (auto it = left.rbegin(); it != left.rend(); ++it) {
    const GivensRotation& g = *it;
    cdouble sp = -std::conj(g.s) * d[g.i] * std::conj(d[g.j]);
    out.rotations.push_back({g.c, sp, g.i, g.j});
  }
  out.diag = d;
  return out;
}

GivensDecomposition triangle_decomposition(const cdouble* mat, int n, double tol) {
  GivensDecomposition out;
  std::vector<cdouble> M(mat, mat + (size_t)n * n);
  auto at = [&](int r, int c) -> cdouble& { return M[(size_t)r * n + c]; };
  // Reck et al., PRL 73, 58 (1994): null each row from the bottom, left to
  // right, by column rotations that push its weight into the diagonal
  // column.  Every rotation is a right factor, so U = D G_N^* ... G_1^*
  // with no rotation moved through the diagonal.
  for (int row = n - 1; row >= 1; --row) {
    for (int col = 0; col < row; ++col) {
      if (std::abs(at(row, col)) <= tol) continue;
      double c;
      cdouble s;
      zrotg_tol(at(row, col + 1), at(row, col), tol, c, s);
      out.rotations.push_back({c, s, col + 1, col});
      for (int r = 0; r < n; ++r) {
        const cdouble xo = at(r, col + 1), yo = at(r, col);
        at(r, col + 1) = c * xo + s * yo;
        at(r, col) = c * yo - std::conj(s) * xo;
      }
    }
  }
  out.diag.resize(n);
  for (int k = 0; k < n; ++k) out.diag[k] = at(k, k);
  return out;
}

KStep kernel_step_from_rotation(const GivensRotation& g) {
  KStep k{};
  cdouble sigma;
  if (g.j == g.i + 1) {
    k.p = g.i;
    sigma = std::conj(g.s);
  } else if (g.i == g.j + 1) {
    k.p = g.j;
    sigma = -g.s;
  } else {
    throw std::invalid_argument("Target orbitals must be adjacent.");
  }
  k.c = g.c;
  k.sr = sigma.real();
  k.si = sigma.imag();
  return k;
}

namespace {
std::vector<PlannedPass> plan_from(const std::vector<KStep>& steps, int norb, int w,
                                   const std::vector<int>& lows, int start) {
  std::vector<PlannedPass> passes;
  std::vector<char> done(steps.size(), 0);
  size_t n_done = 0;
  int m = (int)lows.size();
  // zigzag over the windows: start, start+1, ..., m-1, m-2, ..., 0, 1, ...
  std::vector<int> order;
  if (m == 1) {
    order = {0};
  } else {
    for (int k = 0; k < m; ++k) order.push_back(k);
    for (int k = m - 2; k >= 1; --k) order.push_back(k);
  }
  int pos = 0;
  for (int k = 0; k < (int)order.size(); ++k)
    if (order[k] == start) pos = k;
  int idle = 0;
  while (n_done < steps.size()) {
    int lo = lows[order[pos % order.size()]];
    ++pos;
    PlannedPass P;
    P.lo = lo;
    std::vector<char> blocked(norb + 1, 0);
    for (size_t r = 0; r < steps.size(); ++r) {
      if (done[r]) continue;
      int p = steps[r].p;
      bool inside = p >= lo && p + 1 < lo + w;
      if (inside && !blocked[p] && !blocked[p + 1]) {
        KStep k = steps[r];
        k.p = p - lo;
        P.steps.push_back(k);
        done[r] = 1;
        ++n_done;
      } else {
        blocked[p] = blocked[p + 1] = 1;
      }
    }
    if (!P.steps.empty()) {
      passes.push_back(std::move(P));
      idle = 0;
    } else if (++idle > 2 * m + 2) {
      throw std::runtime_error("pass planner made no progress");
    }
  }
  return passes;
}
}  // namespace

std::vector<PlannedPass> plan_passes(const std::vector<KStep>& steps, int norb, int w,
                                     bool end_low, int first_window) {
  if (w >= norb) {
    std::vector<PlannedPass> one(1);
    one[0].lo = 0;
    one[0].steps = steps;
    one[0].last = true;
    return one;
  }
  // window starts: evenly spaced from 0 to norb - w with overlapping windows
  int span = norb - w;
  int n_win = 2;
  while ((span + n_win - 2) / (n_win - 1) > w - 1) ++n_win;
  std::vector<int> lows;
  for (int k = 0; k < n_win; ++k) lows.push_back((int)((long)span * k / (n_win - 1)));
  std::vector<PlannedPass> best;
  long best_cost = -1;
  for (int start : {0, n_win - 1}) {
    if (first_window >= 0 && start != (first_window == 0 ? 0 : n_win - 1)) continue;
    std::vector<PlannedPass> p = plan_from(steps, norb, w, lows, start);
    long cost = (long)p.size();
    if (end_low && (p.empty() || p.back().lo != 0)) cost += 1;
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = std::move(p);
    }
  }
  if (end_low && (best.empty() || best.back().lo != 0)) {
    PlannedPass P;
    P.lo = 0;
    best.push_back(P);
  }
  if (best.empty()) {
    PlannedPass P;
    P.lo = 0;
    best.push_back(P);
  }
  best.back().last = true;
  return best;
}

}  // namespace ffsim_amd
