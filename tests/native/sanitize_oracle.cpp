// Sanitizer driver for the oracle's C kernels (test infrastructure).
//
// Built with -fsanitize=address,undefined together with oracle/ffsim_oracle.c
// by tests/test_sanitizers.py and run on the CPU.  It drives the kernels the
// way the Python oracle does (one call per Givens rotation and per orbital
// phase, transpose copies for the beta sector, the diagonal evolutions and
// contractions) with one and with several threads, and checks the results
// against closed forms computed here: the compound-matrix action of an
// orbital rotation, psi'[T] = sum_S det(U[T,S]) psi[S], and eigenphases
// from occupation numbers.  Prints OK and exits 0 when every check holds.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

typedef std::complex<double> C;

extern "C" {
int oracle_num_threads(void);
void oracle_apply_givens_rotation_in_place(double* vec, long dim_b, double c, double s_re,
                                           double s_im, const uint64_t* slice1,
                                           const uint64_t* slice2, long n_pairs);
void oracle_apply_phase_shift_in_place(double* vec, long dim_b, double ph_re, double ph_im,
                                       const uint64_t* indices, long n);
void oracle_apply_diag_coulomb_evolution_num_rep(double* vec, long dim_a, long dim_b, int norb,
                                                 const double* mat_exp_aa,
                                                 const double* mat_exp_ab,
                                                 const double* mat_exp_bb,
                                                 const uint64_t* occ_a, long na,
                                                 const uint64_t* occ_b, long nb);
void oracle_contract_diag_coulomb_num_rep(const double* vec, double* out, long dim_a,
                                          long dim_b, int norb, const double* mat_aa,
                                          const double* mat_ab, const double* mat_bb,
                                          const uint64_t* occ_a, long na,
                                          const uint64_t* occ_b, long nb);
void oracle_apply_num_op_sum_evolution_in_place(double* vec, long nrows, long ncols, long rs,
                                                long cs, const double* phases,
                                                const uint64_t* occupations, long nocc);
void oracle_transpose(const double* src, double* dst, long rows, long cols);
int oracle_givens_decomposition(const double* mat_in, int n, double tol, double* c_out,
                                double* s_out, int* i_out, int* j_out, double* diag_out);
}

static int g_fail = 0;
static void check(bool ok, const std::string& what, double err) {
  if (!ok) {
    std::printf("FAIL %s (err %.3e)\n", what.c_str(), err);
    g_fail = 1;
  }
}

static std::vector<uint64_t> strings(int n, int k) {
  std::vector<uint64_t> s;
  for (uint64_t x = 0; x < (1ull << n); ++x)
    if (__builtin_popcountll(x) == k) s.push_back(x);
  return s;
}
static std::vector<uint64_t> occ_lists(const std::vector<uint64_t>& s, int n, int k) {
  std::vector<uint64_t> o;
  for (uint64_t x : s)
    for (int i = 0; i < n; ++i)
      if ((x >> i) & 1) o.push_back((uint64_t)i);
  (void)k;
  return o;
}
static long rank_of(const std::vector<uint64_t>& s, uint64_t x) {
  for (size_t a = 0; a < s.size(); ++a)
    if (s[a] == x) return (long)a;
  return -1;
}

static std::vector<C> random_unitary(int n, std::mt19937_64& g) {
  std::normal_distribution<double> nd;
  std::vector<C> u((size_t)n * n);
  for (auto& x : u) x = C(nd(g), nd(g));
  for (int i = 0; i < n; ++i) {  // Gram-Schmidt on rows
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

// psi'[T] = sum_S det(U[T, S]) psi[S] along one sector of a dim_a x dim_b state
static std::vector<C> compound(const std::vector<C>& u, int n, int k, const std::vector<C>& v,
                               long da, long db, bool alpha) {
  std::vector<uint64_t> s = strings(n, k);
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

// the oracle's schedule on rows of a (dim x cols) matrix: one call per rotation
// with (c, conj(s)) on (i occupied / j empty, partner), then the phases
static void oracle_rotate_rows(std::vector<C>& v, long cols, const std::vector<C>& u, int n, int k) {
  int cap = n * (n - 1) / 2 + 1;
  std::vector<double> c(cap), s(2 * cap), d(2 * n);
  std::vector<int> ii(cap), jj(cap);
  int nr = oracle_givens_decomposition(reinterpret_cast<const double*>(u.data()), n, 1e-12,
                                       c.data(), s.data(), ii.data(), jj.data(), d.data());
  std::vector<uint64_t> st = strings(n, k);
  for (int r = 0; r < nr; ++r) {
    std::vector<uint64_t> s1, s2;
    for (uint64_t x : st)
      if (((x >> ii[r]) & 1) && !((x >> jj[r]) & 1)) {
        s1.push_back((uint64_t)rank_of(st, x));
        s2.push_back((uint64_t)rank_of(st, x ^ (1ull << ii[r]) ^ (1ull << jj[r])));
      }
    oracle_apply_givens_rotation_in_place(reinterpret_cast<double*>(v.data()), cols, c[r],
                                          s[2 * r], -s[2 * r + 1], s1.data(), s2.data(),
                                          (long)s1.size());
  }
  for (int i = 0; i < n; ++i) {
    std::vector<uint64_t> idx;
    for (size_t a = 0; a < st.size(); ++a)
      if ((st[a] >> i) & 1) idx.push_back(a);
    oracle_apply_phase_shift_in_place(reinterpret_cast<double*>(v.data()), cols, d[2 * i],
                                      d[2 * i + 1], idx.data(), (long)idx.size());
  }
}

static double maxdiff(const std::vector<C>& a, const std::vector<C>& b) {
  double m = 0;
  for (size_t t = 0; t < a.size(); ++t) m = std::max(m, std::abs(a[t] - b[t]));
  return m;
}

int main() {
  std::mt19937_64 g(7);
  std::normal_distribution<double> nd;
  struct Case {
    int n, na, nb;
  };
  const Case cases[] = {{2, 1, 1}, {3, 2, 1}, {4, 2, 2}, {5, 2, 3}, {6, 3, 3}, {6, 0, 4}, {6, 4, 1}};
  for (const char* threads : {"1", "3"}) {
    setenv("FFSIM_ORACLE_THREADS", threads, 1);
    for (const Case& cs : cases) {
      const int n = cs.n;
      const long da = (long)strings(n, cs.na).size(), db = (long)strings(n, cs.nb).size();
      std::vector<C> v((size_t)(da * db));
      double nr = 0;
      for (auto& x : v) {
        x = C(nd(g), nd(g));
        nr += std::norm(x);
      }
      for (auto& x : v) x /= std::sqrt(nr);
      std::vector<C> ua = random_unitary(n, g), ub = random_unitary(n, g);
      // orbital rotation: alpha on rows, beta on the transposed state
      std::vector<C> w = v;
      oracle_rotate_rows(w, db, ua, n, cs.na);
      std::vector<C> t((size_t)(da * db));
      oracle_transpose(reinterpret_cast<const double*>(w.data()), reinterpret_cast<double*>(t.data()), da, db);
      oracle_rotate_rows(t, da, ub, n, cs.nb);
      oracle_transpose(reinterpret_cast<const double*>(t.data()), reinterpret_cast<double*>(w.data()), db, da);
      std::vector<C> ref = compound(ub, n, cs.nb, compound(ua, n, cs.na, v, da, db, true), da, db, false);
      double e = maxdiff(w, ref);
      check(e < 1e-12, "orbital rotation vs compound matrix n=" + std::to_string(n), e);
      // diag-Coulomb evolution (number rep) and contraction vs eigenphases
      std::vector<double> J((size_t)n * n);
      for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) J[i * n + j] = J[j * n + i] = nd(g);
      const double time = 0.7;
      std::vector<C> ex((size_t)n * n);
      std::vector<double> Jh = J;
      for (int i = 0; i < n; ++i) Jh[i * n + i] *= 0.5;
      for (size_t q = 0; q < ex.size(); ++q) ex[q] = std::exp(C(0, -time * Jh[q]));
      std::vector<C> exab((size_t)n * n);
      for (size_t q = 0; q < exab.size(); ++q) exab[q] = std::exp(C(0, -time * J[q]));
      std::vector<uint64_t> sa = strings(n, cs.na), sb = strings(n, cs.nb);
      std::vector<uint64_t> oa = occ_lists(sa, n, cs.na), ob = occ_lists(sb, n, cs.nb);
      std::vector<C> dv = v;
      oracle_apply_diag_coulomb_evolution_num_rep(reinterpret_cast<double*>(dv.data()), da, db, n,
                                                  reinterpret_cast<double*>(ex.data()),
                                                  reinterpret_cast<double*>(exab.data()),
                                                  reinterpret_cast<double*>(ex.data()), oa.data(),
                                                  cs.na, ob.data(), cs.nb);
      std::vector<C> co((size_t)(da * db), 0.0);
      oracle_contract_diag_coulomb_num_rep(reinterpret_cast<const double*>(v.data()),
                                           reinterpret_cast<double*>(co.data()), da, db, n,
                                           Jh.data(), J.data(), Jh.data(), oa.data(), cs.na,
                                           ob.data(), cs.nb);
      double ed = 0, ec = 0;
      for (long a = 0; a < da; ++a)
        for (long b = 0; b < db; ++b) {
          double E = 0;
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
              const double ni = (double)((sa[a] >> i) & 1), nj = (double)((sa[a] >> j) & 1);
              const double mi = (double)((sb[b] >> i) & 1), mj = (double)((sb[b] >> j) & 1);
              E += 0.5 * J[i * n + j] * (ni * nj + mi * mj) + J[i * n + j] * ni * mj;
            }
          ed = std::max(ed, std::abs(dv[a * db + b] - v[a * db + b] * std::exp(C(0, -time * E))));
          ec = std::max(ec, std::abs(co[a * db + b] - E * v[a * db + b]));
        }
      check(ed < 1e-12, "diag-Coulomb evolution vs eigenphases n=" + std::to_string(n), ed);
      check(ec < 1e-11, "diag-Coulomb contraction vs energies n=" + std::to_string(n), ec);
      // num-op evolution on the alpha rows and on the beta columns (strided view)
      std::vector<double> lam(n);
      for (auto& x : lam) x = nd(g);
      std::vector<C> ph(n);
      for (int i = 0; i < n; ++i) ph[i] = std::exp(C(0, -time * lam[i]));
      std::vector<C> nv = v;
      oracle_apply_num_op_sum_evolution_in_place(reinterpret_cast<double*>(nv.data()), da, db, db, 1,
                                                 reinterpret_cast<double*>(ph.data()), oa.data(), cs.na);
      oracle_apply_num_op_sum_evolution_in_place(reinterpret_cast<double*>(nv.data()), db, da, 1, db,
                                                 reinterpret_cast<double*>(ph.data()), ob.data(), cs.nb);
      double en = 0;
      for (long a = 0; a < da; ++a)
        for (long b = 0; b < db; ++b) {
          double E = 0;
          for (int i = 0; i < n; ++i) E += lam[i] * (double)(((sa[a] >> i) & 1) + ((sb[b] >> i) & 1));
          en = std::max(en, std::abs(nv[a * db + b] - v[a * db + b] * std::exp(C(0, -time * E))));
        }
      check(en < 1e-12, "num-op evolution vs eigenphases n=" + std::to_string(n), en);
    }
    // the threaded Givens path (>= 128 row pairs): norm kept, same as one thread
    const int n = 12, k = 6;
    std::vector<uint64_t> st = strings(n, k);
    const long dim = (long)st.size(), cols = 5;
    std::vector<C> v((size_t)(dim * cols));
    for (auto& x : v) x = C(nd(g), nd(g));
    std::vector<uint64_t> s1, s2;
    for (uint64_t x : st)
      if (((x >> 4) & 1) && !((x >> 5) & 1)) {
        s1.push_back((uint64_t)rank_of(st, x));
        s2.push_back((uint64_t)rank_of(st, x ^ (3ull << 4)));
      }
    std::vector<C> w = v, w1 = v;
    oracle_apply_givens_rotation_in_place(reinterpret_cast<double*>(w.data()), cols, 0.6, 0.48, 0.64,
                                          s1.data(), s2.data(), (long)s1.size());
    const char* prev = getenv("FFSIM_ORACLE_THREADS");
    std::string keep = prev ? prev : "1";
    setenv("FFSIM_ORACLE_THREADS", "1", 1);
    oracle_apply_givens_rotation_in_place(reinterpret_cast<double*>(w1.data()), cols, 0.6, 0.48, 0.64,
                                          s1.data(), s2.data(), (long)s1.size());
    setenv("FFSIM_ORACLE_THREADS", keep.c_str(), 1);
    double n0 = 0, n1 = 0;
    for (size_t t = 0; t < v.size(); ++t) {
      n0 += std::norm(v[t]);
      n1 += std::norm(w[t]);
    }
    check(maxdiff(w, w1) == 0.0, std::string("threaded Givens equals one thread, threads=") + threads,
          maxdiff(w, w1));
    check(std::abs(n0 - n1) < 1e-9 * n0, "threaded Givens keeps the norm", std::abs(n0 - n1));
  }
  if (!g_fail) std::printf("OK\n");
  return g_fail;
}
