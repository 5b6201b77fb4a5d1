// Sanitizer driver for the library's host code (test infrastructure).
//
// Built with -fsanitize=address,undefined together with
// ffsim_amd/csrc/host_core.cpp and tests/native/pass_emulator.cpp by
// tests/test_sanitizers.py and run on the CPU.  It runs the host planner,
// window fibers, group tables, wave schedules and the emulated fused passes
// (the GPU kernel's programs executed with plain loops) over many shapes and
// window widths, and checks the emulated rotation against the compound-matrix
// closed form psi'[T] = sum_S det(U[T,S]) psi[S] (norb <= 7) or, for larger
// shapes, that a rotation followed by its inverse returns the state.  Prints OK
// and exits 0 when every check holds.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../ffsim_amd/csrc/host_core.hpp"

using namespace ffsim_amd;
typedef std::complex<double> C;

extern "C" void emulate(double* vec, double* work, int norb, int na, int nb, const double* ma,
                        const double* mb, int w, int merge);

static int g_fail = 0;
static void check(bool ok, const std::string& what, double err) {
  if (!ok) {
    std::printf("FAIL %s (err %.3e)\n", what.c_str(), err);
    g_fail = 1;
  }
}

static std::vector<C> random_unitary(int n, std::mt19937_64& g) {
  std::normal_distribution<double> nd;
  std::vector<C> u((size_t)n * n);
  for (auto& x : u) x = C(nd(g), nd(g));
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < i; ++j) {
      C d = 0;
      for (int k = 0; k < n; ++k) d += std::conj(u[j * n + k]) * u[i * n + k];
      for (int k = 0; k < n; ++k) u[i * n + k] -= d * u[j * n + k];
    }
    double nr = 0;
    for (int k = 0; k < n; ++k) nr += std::norm(u[i * n + k]);
    for (int k = 0; k < n; ++k) u[i * n + k] /= std::sqrt(nr);
  }
  return u;
}

static C det(std::vector<C> m, int k) {
  C d = 1;
  for (int c = 0; c < k; ++c) {
    int p = c;
    for (int r = c + 1; r < k; ++r)
      if (std::abs(m[r * k + c]) > std::abs(m[p * k + c])) p = r;
    if (std::abs(m[p * k + c]) == 0) return 0;
    if (p != c) {
      for (int t = 0; t < k; ++t) std::swap(m[p * k + t], m[c * k + t]);
      d = -d;
    }
    d *= m[c * k + c];
    for (int r = c + 1; r < k; ++r) {
      C f = m[r * k + c] / m[c * k + c];
      for (int t = c; t < k; ++t) m[r * k + t] -= f * m[c * k + t];
    }
  }
  return d;
}

static std::vector<C> compound(const std::vector<C>& u, int n, int k, const std::vector<C>& v,
                               long da, long db, bool alpha) {
  std::vector<uint64_t> s = make_strings(n, k);
  long dim = (long)s.size();
  std::vector<C> L((size_t)dim * dim);
  for (long t = 0; t < dim; ++t)
    for (long q = 0; q < dim; ++q) {
      std::vector<int> rt, cq;
      for (int i = 0; i < n; ++i) {
        if ((s[t] >> i) & 1) rt.push_back(i);
        if ((s[q] >> i) & 1) cq.push_back(i);
      }
      std::vector<C> m((size_t)k * k);
      for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) m[a * k + b] = u[rt[a] * n + cq[b]];
      L[t * dim + q] = k ? det(m, k) : C(1);
    }
  std::vector<C> out(v.size(), 0.0);
  for (long a = 0; a < da; ++a)
    for (long b = 0; b < db; ++b) {
      C acc = 0;
      if (alpha)
        for (long q = 0; q < da; ++q) acc += L[a * dim + q] * v[q * db + b];
      else
        for (long q = 0; q < db; ++q) acc += L[b * dim + q] * v[a * db + q];
      out[a * db + b] = acc;
    }
  return out;
}

static double maxdiff(const std::vector<C>& a, const std::vector<C>& b) {
  double m = 0;
  for (size_t t = 0; t < a.size(); ++t) m = std::max(m, std::abs(a[t] - b[t]));
  return m;
}

int main() {
  std::mt19937_64 g(11);
  std::normal_distribution<double> nd;
  struct Case {
    int n, na, nb, w;
  };
  const Case cases[] = {{2, 1, 1, 2},  {3, 2, 1, 3},  {4, 2, 2, 4},  {4, 0, 4, 2},  {5, 2, 3, 5},
                        {6, 3, 2, 4},  {6, 6, 1, 3},  {7, 3, 4, 4},  {7, 0, 3, 4},  {8, 4, 4, 5},
                        {9, 2, 7, 5},  {10, 5, 4, 6}, {10, 4, 5, 8}, {9, 4, 4, 9}};
  for (int merge = 0; merge < 2; ++merge)
    for (const Case& cs : cases) {
      const int n = cs.n;
      const long da = (long)make_strings(n, cs.na).size(), db = (long)make_strings(n, cs.nb).size();
      std::vector<C> v((size_t)(da * db));
      double nr = 0;
      for (auto& x : v) {
        x = C(nd(g), nd(g));
        nr += std::norm(x);
      }
      for (auto& x : v) x /= std::sqrt(nr);
      std::vector<C> ua = random_unitary(n, g), ub = random_unitary(n, g);
      std::vector<C> x = v, work(v.size());
      emulate(reinterpret_cast<double*>(x.data()), reinterpret_cast<double*>(work.data()), n,
              cs.na, cs.nb, reinterpret_cast<const double*>(ua.data()),
              reinterpret_cast<const double*>(ub.data()), cs.w, merge);
      const std::string tag = "n=" + std::to_string(n) + " w=" + std::to_string(cs.w) +
                              " merge=" + std::to_string(merge);
      if (n <= 7) {
        std::vector<C> ref =
            compound(ub, n, cs.nb, compound(ua, n, cs.na, v, da, db, true), da, db, false);
        const double e = maxdiff(x, ref);
        check(e < 1e-12, "emulated rotation vs compound matrix " + tag, e);
      }
      // the inverse rotation returns the state
      std::vector<C> ia((size_t)n * n), ib((size_t)n * n);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          ia[i * n + j] = std::conj(ua[j * n + i]);
          ib[i * n + j] = std::conj(ub[j * n + i]);
        }
      emulate(reinterpret_cast<double*>(x.data()), reinterpret_cast<double*>(work.data()), n,
              cs.na, cs.nb, reinterpret_cast<const double*>(ia.data()),
              reinterpret_cast<const double*>(ib.data()), cs.w, merge);
      const double e = maxdiff(x, v);
      check(e < 1e-12, "emulated rotation round trip " + tag, e);
    }
  // tables and schedules of every window width the library plans with
  for (int w = 2; w <= 12; ++w) {
    GroupTables T = make_group_tables(w, kMaxSubWindow, false);
    GroupTables M = make_group_tables(w, kMaxSubWindow, true);
    for (int waves : {4, 8}) {
      WaveSchedule a = make_wave_schedule(T, waves, 8);
      WaveSchedule b = make_wave_schedule(M, waves, 16);
      check(!a.index.empty() && !b.index.empty(), "wave schedule w=" + std::to_string(w), 0);
    }
    for (int norb = w; norb <= 16; ++norb)
      for (int k = 0; k <= norb; k += 3) {
        WindowLayout L = make_window_layout(norb, k, 0, w);
        long rows = 0;
        for (const Fiber& f : L.fibers) rows += f.n_rows;
        check(rows == (long)binom(norb, k), "window rows cover every string", 0);
      }
  }
  if (!g_fail) std::printf("OK\n");
  return g_fail;
}
