/*
 * ffsim_oracle.c -- CPU restatement of ffsim's Rust kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the
 * host-core baseline for the ffsim_amd GPU library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker / the timed CPU reference, never as the product path.
 *
 * Each function restates one kernel of the reference (qiskit-community/ffsim,
 * Rust, src/...), keeping its schedule and threading so that timing it on the
 * host is a fair "ffsim CPU path" baseline:
 *   - Givens rotation on row pairs, split into per-call threads over pairs
 *     when there are >= 128 pairs (src/gates/orbital_rotation.rs:19-104);
 *   - row phase shift (src/gates/phase_shift.rs:16-30);
 *   - num-op-sum evolution, parallel over rows (src/gates/num_op_sum.rs:18-36);
 *   - diag-Coulomb evolution in number and Z representation, three passes
 *     (src/gates/diag_coulomb.rs:19-190);
 *   - diag-Coulomb contraction, number and Z representation
 *     (src/contract/diag_coulomb.rs:19-174);
 *   - num-op-sum contraction for one spin (src/contract/num_op_sum.rs:18-39);
 *   - Givens (Clements) decomposition (src/linalg/givens.rs:19-149).
 * Complex numbers are interleaved doubles.  Matrices are row-major with an
 * explicit row stride (elements) and column stride so that the beta sector
 * can be addressed through a transposed view as the Python layer does
 * (python/ffsim/gates/num_op_sum.py:205-207).
 *
 * Threads: the count comes from FFSIM_ORACLE_THREADS, else RAYON_NUM_THREADS
 * (the variable ffsim reads, src/gates/orbital_rotation.rs:37-45), else the
 * number of online cores.
 */
#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef double _Complex cplx;

int oracle_num_threads(void) {
  const char* e = getenv("FFSIM_ORACLE_THREADS");
  if (!e || !*e) e = getenv("RAYON_NUM_THREADS");
  if (e && *e) {
    int n = atoi(e);
    if (n > 0) return n;
  }
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (int)n : 1;
}

/* ----------------------------------------------------------- parallel for */
typedef void (*range_fn)(void* ctx, long begin, long end);
struct par_job {
  range_fn fn;
  void* ctx;
  long begin, end;
};
static void* par_trampoline(void* p) {
  struct par_job* j = (struct par_job*)p;
  j->fn(j->ctx, j->begin, j->end);
  return NULL;
}
/* Splits [0, n) into nthreads contiguous chunks and runs them on threads
 * spawned for this call (the Givens kernel's std::thread::scope pattern; the
 * rayon row loops are served the same way). */
static void par_for(long n, int nthreads, range_fn fn, void* ctx) {
  if (n <= 0) return;
  if (nthreads > n) nthreads = (int)n;
  if (nthreads <= 1) {
    fn(ctx, 0, n);
    return;
  }
  long chunk = (n + nthreads - 1) / nthreads;
  pthread_t tid[256];
  struct par_job jobs[256];
  if (nthreads > 256) nthreads = 256;
  int started = 0;
  for (int k = 0; k < nthreads; ++k) {
    long b = k * chunk;
    if (b >= n) break;
    long e = b + chunk < n ? b + chunk : n;
    jobs[k].fn = fn;
    jobs[k].ctx = ctx;
    jobs[k].begin = b;
    jobs[k].end = e;
    if (pthread_create(&tid[k], NULL, par_trampoline, &jobs[k]) != 0) {
      fn(ctx, b, e);
      jobs[k].fn = NULL;
    }
    ++started;
  }
  for (int k = 0; k < started; ++k)
    if (jobs[k].fn) pthread_join(tid[k], NULL);
}

/* ------------------------------------------------------------ Givens rows */
struct givens_ctx {
  cplx* vec;
  long dim_b;
  double c;
  cplx s;
  const uint64_t* s1;
  const uint64_t* s2;
};
static void givens_range(void* p, long b, long e) {
  struct givens_ctx* g = (struct givens_ctx*)p;
  for (long k = b; k < e; ++k) {
    cplx* ri = g->vec + (long)g->s1[k] * g->dim_b;
    cplx* rj = g->vec + (long)g->s2[k] * g->dim_b;
    for (long t = 0; t < g->dim_b; ++t) {
      cplx io = ri[t], jo = rj[t];
      ri[t] = g->c * io + g->s * jo;
      rj[t] = g->c * jo - conj(g->s) * io;
    }
  }
}
/* vec: contiguous dim_a x dim_b; rows s1[k], s2[k] are rotated together. */
void oracle_apply_givens_rotation_in_place(double* vec, long dim_b, double c, double s_re,
                                           double s_im, const uint64_t* slice1,
                                           const uint64_t* slice2, long n_pairs) {
  if (n_pairs <= 0) return;
  struct givens_ctx g = {(cplx*)vec, dim_b, c, s_re + I * s_im, slice1, slice2};
  int nt = oracle_num_threads();
  if (nt > n_pairs) nt = (int)n_pairs;
  if (nt == 1 || n_pairs < 128)
    givens_range(&g, 0, n_pairs);
  else
    par_for(n_pairs, nt, givens_range, &g);
}

/* vec rows listed in indices are multiplied by phase (sequential). */
void oracle_apply_phase_shift_in_place(double* vec, long dim_b, double ph_re, double ph_im,
                                       const uint64_t* indices, long n) {
  cplx* v = (cplx*)vec;
  cplx ph = ph_re + I * ph_im;
  for (long k = 0; k < n; ++k) {
    cplx* row = v + (long)indices[k] * dim_b;
    for (long t = 0; t < dim_b; ++t) row[t] *= ph;
  }
}

/* --------------------------------------------------------- num-op evolve */
struct numop_ctx {
  cplx* vec;
  long rs, cs, ncols;
  const cplx* phases;
  const uint64_t* occ;
  long nocc;
};
static void numop_range(void* p, long b, long e) {
  struct numop_ctx* q = (struct numop_ctx*)p;
  for (long r = b; r < e; ++r) {
    cplx ph = 1.0;
    for (long k = 0; k < q->nocc; ++k) ph *= q->phases[q->occ[r * q->nocc + k]];
    cplx* row = q->vec + r * q->rs;
    for (long t = 0; t < q->ncols; ++t) row[t * q->cs] *= ph;
  }
}
/* Rows of the (possibly transposed) view: element (r, t) at vec[r*rs + t*cs]. */
void oracle_apply_num_op_sum_evolution_in_place(double* vec, long nrows, long ncols, long rs,
                                                long cs, const double* phases,
                                                const uint64_t* occupations, long nocc) {
  struct numop_ctx q = {(cplx*)vec, rs, cs, ncols, (const cplx*)phases, occupations, nocc};
  par_for(nrows, oracle_num_threads(), numop_range, &q);
}

/* ---------------------------------------------- diag-Coulomb, number rep */
struct dc_ctx {
  cplx* vec;
  long dim_a, dim_b;
  int norb;
  const cplx *eaa, *eab, *ebb;
  const uint64_t *occ_a, *occ_b;
  long na, nb;
  const int64_t *str_a, *str_b;
  const cplx *eaa_c, *eab_c, *ebb_c;
  cplx *alpha, *beta, *map;
};
static void dc_beta_range(void* p, long b, long e) {
  struct dc_ctx* d = (struct dc_ctx*)p;
  for (long r = b; r < e; ++r) {
    const uint64_t* o = d->occ_b + r * d->nb;
    cplx ph = 1.0;
    for (long j = 0; j < d->nb; ++j)
      for (long k = j; k < d->nb; ++k) ph *= d->ebb[o[j] * d->norb + o[k]];
    d->beta[r] = ph;
  }
}
static void dc_alpha_range(void* p, long b, long e) {
  struct dc_ctx* d = (struct dc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    const uint64_t* o = d->occ_a + r * d->na;
    cplx ph = 1.0;
    cplx* row = d->map + r * n;
    for (int t = 0; t < n; ++t) row[t] = 1.0;
    for (long j = 0; j < d->na; ++j) {
      for (int t = 0; t < n; ++t) row[t] *= d->eab[o[j] * n + t];
      for (long k = j; k < d->na; ++k) ph *= d->eaa[o[j] * n + o[k]];
    }
    d->alpha[r] = ph;
  }
}
static void dc_main_range(void* p, long b, long e) {
  struct dc_ctx* d = (struct dc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    cplx* row = d->vec + r * d->dim_b;
    const cplx* m = d->map + r * n;
    for (long c = 0; c < d->dim_b; ++c) {
      cplx ph = d->alpha[r] * d->beta[c];
      const uint64_t* o = d->occ_b + c * d->nb;
      for (long k = 0; k < d->nb; ++k) ph *= m[o[k]];
      row[c] *= ph;
    }
  }
}
void oracle_apply_diag_coulomb_evolution_num_rep(double* vec, long dim_a, long dim_b, int norb,
                                                 const double* mat_exp_aa,
                                                 const double* mat_exp_ab,
                                                 const double* mat_exp_bb,
                                                 const uint64_t* occ_a, long na,
                                                 const uint64_t* occ_b, long nb) {
  struct dc_ctx d;
  memset(&d, 0, sizeof d);
  d.vec = (cplx*)vec;
  d.dim_a = dim_a;
  d.dim_b = dim_b;
  d.norb = norb;
  d.eaa = (const cplx*)mat_exp_aa;
  d.eab = (const cplx*)mat_exp_ab;
  d.ebb = (const cplx*)mat_exp_bb;
  d.occ_a = occ_a;
  d.occ_b = occ_b;
  d.na = na;
  d.nb = nb;
  d.alpha = (cplx*)malloc(sizeof(cplx) * (dim_a > 0 ? dim_a : 1));
  d.beta = (cplx*)malloc(sizeof(cplx) * (dim_b > 0 ? dim_b : 1));
  d.map = (cplx*)malloc(sizeof(cplx) * (dim_a > 0 ? dim_a : 1) * (norb > 0 ? norb : 1));
  int nt = oracle_num_threads();
  par_for(dim_b, nt, dc_beta_range, &d);
  par_for(dim_a, nt, dc_alpha_range, &d);
  par_for(dim_a, nt, dc_main_range, &d);
  free(d.alpha);
  free(d.beta);
  free(d.map);
}

/* --------------------------------------------------- diag-Coulomb, Z rep */
static void dcz_beta_range(void* p, long b, long e) {
  struct dc_ctx* d = (struct dc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    int64_t s = d->str_b[r];
    cplx ph = 1.0;
    for (int j = 0; j < n; ++j) {
      int sj = (s >> j) & 1;
      for (int k = j + 1; k < n; ++k) {
        int sk = (s >> k) & 1;
        ph *= (sj ^ sk) ? d->ebb_c[j * n + k] : d->ebb[j * n + k];
      }
    }
    d->beta[r] = ph;
  }
}
static void dcz_alpha_range(void* p, long b, long e) {
  struct dc_ctx* d = (struct dc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    int64_t s = d->str_a[r];
    cplx ph = 1.0;
    cplx* row = d->map + r * n;
    for (int t = 0; t < n; ++t) row[t] = 1.0;
    for (int j = 0; j < n; ++j) {
      int sj = (s >> j) & 1;
      const cplx* src = sj ? d->eab_c + j * n : d->eab + j * n;
      for (int t = 0; t < n; ++t) row[t] *= src[t];
      for (int k = j + 1; k < n; ++k) {
        int sk = (s >> k) & 1;
        ph *= (sj ^ sk) ? d->eaa_c[j * n + k] : d->eaa[j * n + k];
      }
    }
    d->alpha[r] = ph;
  }
}
static void dcz_main_range(void* p, long b, long e) {
  struct dc_ctx* d = (struct dc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    cplx* row = d->vec + r * d->dim_b;
    const cplx* m = d->map + r * n;
    for (long c = 0; c < d->dim_b; ++c) {
      cplx ph = d->alpha[r] * d->beta[c];
      int64_t s = d->str_b[c];
      for (int j = 0; j < n; ++j) ph *= ((s >> j) & 1) ? conj(m[j]) : m[j];
      row[c] *= ph;
    }
  }
}
void oracle_apply_diag_coulomb_evolution_z_rep(double* vec, long dim_a, long dim_b, int norb,
                                               const double* eaa, const double* eab,
                                               const double* ebb, const double* eaa_c,
                                               const double* eab_c, const double* ebb_c,
                                               const int64_t* str_a, const int64_t* str_b) {
  struct dc_ctx d;
  memset(&d, 0, sizeof d);
  d.vec = (cplx*)vec;
  d.dim_a = dim_a;
  d.dim_b = dim_b;
  d.norb = norb;
  d.eaa = (const cplx*)eaa;
  d.eab = (const cplx*)eab;
  d.ebb = (const cplx*)ebb;
  d.eaa_c = (const cplx*)eaa_c;
  d.eab_c = (const cplx*)eab_c;
  d.ebb_c = (const cplx*)ebb_c;
  d.str_a = str_a;
  d.str_b = str_b;
  d.alpha = (cplx*)malloc(sizeof(cplx) * (dim_a > 0 ? dim_a : 1));
  d.beta = (cplx*)malloc(sizeof(cplx) * (dim_b > 0 ? dim_b : 1));
  d.map = (cplx*)malloc(sizeof(cplx) * (dim_a > 0 ? dim_a : 1) * (norb > 0 ? norb : 1));
  int nt = oracle_num_threads();
  par_for(dim_b, nt, dcz_beta_range, &d);
  par_for(dim_a, nt, dcz_alpha_range, &d);
  par_for(dim_a, nt, dcz_main_range, &d);
  free(d.alpha);
  free(d.beta);
  free(d.map);
}

/* ---------------------------------------------- diag-Coulomb contraction */
struct dcc_ctx {
  const cplx* vec;
  cplx* out;
  long dim_a, dim_b;
  int norb;
  const double *aa, *ab, *bb;
  const uint64_t *occ_a, *occ_b;
  long na, nb;
  const int64_t *str_a, *str_b;
  double *alpha, *beta, *map;
};
static void dcc_beta_range(void* p, long b, long e) {
  struct dcc_ctx* d = (struct dcc_ctx*)p;
  for (long r = b; r < e; ++r) {
    const uint64_t* o = d->occ_b + r * d->nb;
    double acc = 0.0;
    for (long j = 0; j < d->nb; ++j)
      for (long k = j; k < d->nb; ++k) acc += d->bb[o[j] * d->norb + o[k]];
    d->beta[r] = acc;
  }
}
static void dcc_alpha_range(void* p, long b, long e) {
  struct dcc_ctx* d = (struct dcc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    const uint64_t* o = d->occ_a + r * d->na;
    double acc = 0.0;
    double* row = d->map + r * n;
    for (int t = 0; t < n; ++t) row[t] = 0.0;
    for (long j = 0; j < d->na; ++j) {
      for (int t = 0; t < n; ++t) row[t] += d->ab[o[j] * n + t];
      for (long k = j; k < d->na; ++k) acc += d->aa[o[j] * n + o[k]];
    }
    d->alpha[r] = acc;
  }
}
static void dcc_main_range(void* p, long b, long e) {
  struct dcc_ctx* d = (struct dcc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    const cplx* src = d->vec + r * d->dim_b;
    cplx* dst = d->out + r * d->dim_b;
    const double* m = d->map + r * n;
    for (long c = 0; c < d->dim_b; ++c) {
      double coeff = d->alpha[r] + d->beta[c];
      const uint64_t* o = d->occ_b + c * d->nb;
      for (long k = 0; k < d->nb; ++k) coeff += m[o[k]];
      dst[c] += coeff * src[c];
    }
  }
}
void oracle_contract_diag_coulomb_num_rep(const double* vec, double* out, long dim_a,
                                          long dim_b, int norb, const double* mat_aa,
                                          const double* mat_ab, const double* mat_bb,
                                          const uint64_t* occ_a, long na,
                                          const uint64_t* occ_b, long nb) {
  struct dcc_ctx d;
  memset(&d, 0, sizeof d);
  d.vec = (const cplx*)vec;
  d.out = (cplx*)out;
  d.dim_a = dim_a;
  d.dim_b = dim_b;
  d.norb = norb;
  d.aa = mat_aa;
  d.ab = mat_ab;
  d.bb = mat_bb;
  d.occ_a = occ_a;
  d.occ_b = occ_b;
  d.na = na;
  d.nb = nb;
  d.alpha = (double*)malloc(sizeof(double) * (dim_a > 0 ? dim_a : 1));
  d.beta = (double*)malloc(sizeof(double) * (dim_b > 0 ? dim_b : 1));
  d.map = (double*)malloc(sizeof(double) * (dim_a > 0 ? dim_a : 1) * (norb > 0 ? norb : 1));
  int nt = oracle_num_threads();
  par_for(dim_b, nt, dcc_beta_range, &d);
  par_for(dim_a, nt, dcc_alpha_range, &d);
  par_for(dim_a, nt, dcc_main_range, &d);
  free(d.alpha);
  free(d.beta);
  free(d.map);
}

static void dccz_beta_range(void* p, long b, long e) {
  struct dcc_ctx* d = (struct dcc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    int64_t s = d->str_b[r];
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
      double sj = ((s >> j) & 1) ? -1.0 : 1.0;
      for (int k = j + 1; k < n; ++k) {
        double sk = ((s >> k) & 1) ? -1.0 : 1.0;
        acc += sj * sk * d->bb[j * n + k];
      }
    }
    d->beta[r] = acc;
  }
}
static void dccz_alpha_range(void* p, long b, long e) {
  struct dcc_ctx* d = (struct dcc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    int64_t s = d->str_a[r];
    double acc = 0.0;
    double* row = d->map + r * n;
    for (int t = 0; t < n; ++t) row[t] = 0.0;
    for (int j = 0; j < n; ++j) {
      double sj = ((s >> j) & 1) ? -1.0 : 1.0;
      for (int t = 0; t < n; ++t) row[t] += sj * d->ab[j * n + t];
      for (int k = j + 1; k < n; ++k) {
        double sk = ((s >> k) & 1) ? -1.0 : 1.0;
        acc += sj * sk * d->aa[j * n + k];
      }
    }
    d->alpha[r] = acc;
  }
}
static void dccz_main_range(void* p, long b, long e) {
  struct dcc_ctx* d = (struct dcc_ctx*)p;
  int n = d->norb;
  for (long r = b; r < e; ++r) {
    const cplx* src = d->vec + r * d->dim_b;
    cplx* dst = d->out + r * d->dim_b;
    const double* m = d->map + r * n;
    for (long c = 0; c < d->dim_b; ++c) {
      double coeff = d->alpha[r] + d->beta[c];
      int64_t s = d->str_b[c];
      for (int j = 0; j < n; ++j) coeff += (((s >> j) & 1) ? -1.0 : 1.0) * m[j];
      dst[c] += 0.25 * coeff * src[c];
    }
  }
}
void oracle_contract_diag_coulomb_z_rep(const double* vec, double* out, long dim_a, long dim_b,
                                        int norb, const double* mat_aa, const double* mat_ab,
                                        const double* mat_bb, const int64_t* str_a,
                                        const int64_t* str_b) {
  struct dcc_ctx d;
  memset(&d, 0, sizeof d);
  d.vec = (const cplx*)vec;
  d.out = (cplx*)out;
  d.dim_a = dim_a;
  d.dim_b = dim_b;
  d.norb = norb;
  d.aa = mat_aa;
  d.ab = mat_ab;
  d.bb = mat_bb;
  d.str_a = str_a;
  d.str_b = str_b;
  d.alpha = (double*)malloc(sizeof(double) * (dim_a > 0 ? dim_a : 1));
  d.beta = (double*)malloc(sizeof(double) * (dim_b > 0 ? dim_b : 1));
  d.map = (double*)malloc(sizeof(double) * (dim_a > 0 ? dim_a : 1) * (norb > 0 ? norb : 1));
  int nt = oracle_num_threads();
  par_for(dim_b, nt, dccz_beta_range, &d);
  par_for(dim_a, nt, dccz_alpha_range, &d);
  par_for(dim_a, nt, dccz_main_range, &d);
  free(d.alpha);
  free(d.beta);
  free(d.map);
}

/* ------------------------------------------- num-op-sum contraction (one spin) */
struct nc_ctx {
  const cplx* vec;
  cplx* out;
  long rs, cs, ncols;
  const double* coeffs;
  const uint64_t* occ;
  long nocc;
};
static void nc_range(void* p, long b, long e) {
  struct nc_ctx* q = (struct nc_ctx*)p;
  for (long r = b; r < e; ++r) {
    double coeff = 0.0;
    for (long k = 0; k < q->nocc; ++k) coeff += q->coeffs[q->occ[r * q->nocc + k]];
    for (long t = 0; t < q->ncols; ++t) q->out[r * q->rs + t * q->cs] += coeff * q->vec[r * q->rs + t * q->cs];
  }
}
void oracle_contract_num_op_sum_spin_into_buffer(const double* vec, double* out, long nrows,
                                                 long ncols, long rs, long cs,
                                                 const double* coeffs,
                                                 const uint64_t* occupations, long nocc) {
  struct nc_ctx q = {(const cplx*)vec, (cplx*)out, rs, cs, ncols, coeffs, occupations, nocc};
  par_for(nrows, oracle_num_threads(), nc_range, &q);
}

/* ------------------------------------------------------- transpose copy */
struct tr_ctx {
  const cplx* src;
  cplx* dst;
  long rows, cols;
};
static void tr_range(void* p, long b, long e) {
  struct tr_ctx* t = (struct tr_ctx*)p;
  for (long r = b; r < e; ++r)
    for (long c = 0; c < t->cols; ++c) t->dst[c * t->rows + r] = t->src[r * t->cols + c];
}
/* np.ascontiguousarray(vec.T): dst (cols x rows) = transpose of src (rows x cols).
 * numpy performs this copy single-threaded; so does this restatement. */
void oracle_transpose(const double* src, double* dst, long rows, long cols) {
  struct tr_ctx t = {(const cplx*)src, (cplx*)dst, rows, cols};
  tr_range(&t, 0, rows);
}

/* ------------------------------------------------- Givens decomposition */
static void zrotg_safe(cplx a, cplx b, double tol, double* c, cplx* s) {
  double abs_a = cabs(a), abs_b = cabs(b);
  if (abs_b <= tol) {
    *c = 1.0;
    *s = 0.0;
    return;
  }
  if (abs_a <= tol) {
    *c = 0.0;
    *s = 1.0;
    return;
  }
  double r = hypot(abs_a, abs_b);
  double cc = abs_a / r;
  if (cc > 1.0) cc = 1.0;
  if (cc < -1.0) cc = -1.0;
  *c = cc;
  *s = (a / abs_a) * conj(b) / r;
}
static void rot_cols(cplx* m, int n, int x, int y, double c, cplx s) {
  for (int r = 0; r < n; ++r) {
    cplx xo = m[r * n + x], yo = m[r * n + y];
    m[r * n + x] = c * xo + s * yo;
    m[r * n + y] = c * yo - conj(s) * xo;
  }
}
static void rot_rows(cplx* m, int n, int x, int y, double c, cplx s) {
  for (int k = 0; k < n; ++k) {
    cplx xo = m[x * n + k], yo = m[y * n + k];
    m[x * n + k] = c * xo + s * yo;
    m[y * n + k] = c * yo - conj(s) * xo;
  }
}
/* Output: up to n(n-1)/2 rotations (c, s, i, j) in application order and the
 * diagonal; returns the number of rotations. */
int oracle_givens_decomposition(const double* mat_in, int n, double tol, double* c_out,
                                double* s_out, int* i_out, int* j_out, double* diag_out) {
  cplx* m = (cplx*)malloc(sizeof(cplx) * (n > 0 ? n * n : 1));
  memcpy(m, mat_in, sizeof(cplx) * n * n);
  int cap = n * (n - 1) / 2 + 1;
  double* lc = (double*)malloc(sizeof(double) * cap);
  cplx* ls = (cplx*)malloc(sizeof(cplx) * cap);
  int* li = (int*)malloc(sizeof(int) * cap);
  int* lj = (int*)malloc(sizeof(int) * cap);
  int n_left = 0, n_right = 0;
  for (int i = 0; i + 1 < n; ++i) {
    if (i % 2 == 0) {
      for (int j = 0; j <= i; ++j) {
        int target = i - j, row = n - j - 1;
        if (cabs(m[row * n + target]) > tol) {
          double c;
          cplx s;
          zrotg_safe(m[row * n + target + 1], m[row * n + target], tol, &c, &s);
          c_out[n_right] = c;
          s_out[2 * n_right] = creal(s);
          s_out[2 * n_right + 1] = cimag(s);
          i_out[n_right] = target + 1;
          j_out[n_right] = target;
          ++n_right;
          rot_cols(m, n, target + 1, target, c, s);
        }
      }
    } else {
      for (int j = 0; j <= i; ++j) {
        int target = n - i + j - 1, col = j;
        if (cabs(m[target * n + col]) > tol) {
          double c;
          cplx s;
          zrotg_safe(m[(target - 1) * n + col], m[target * n + col], tol, &c, &s);
          lc[n_left] = c;
          ls[n_left] = s;
          li[n_left] = target - 1;
          lj[n_left] = target;
          ++n_left;
          rot_rows(m, n, target - 1, target, c, s);
        }
      }
    }
  }
  int nr = n_right;
  for (int k = n_left - 1; k >= 0; --k) {
    int i = li[k], j = lj[k];
    double c;
    cplx s;
    zrotg_safe(lc[k] * m[j * n + j], conj(ls[k]) * m[i * n + i], tol, &c, &s);
    c_out[nr] = c;
    s_out[2 * nr] = creal(-conj(s));
    s_out[2 * nr + 1] = cimag(-conj(s));
    i_out[nr] = i;
    j_out[nr] = j;
    ++nr;
    cplx di = m[i * n + i], dj = m[j * n + j];
    cplx g00 = c * di, g01 = -s * dj, g10 = conj(s) * di, g11 = c * dj;
    double cn;
    cplx sn;
    zrotg_safe(g11, g10, tol, &cn, &sn);
    m[i * n + i] = g00 * cn + g01 * (-conj(sn));
    m[j * n + j] = g10 * sn + g11 * cn;
  }
  for (int k = 0; k < n; ++k) {
    diag_out[2 * k] = creal(m[k * n + k]);
    diag_out[2 * k + 1] = cimag(m[k * n + k]);
  }
  free(m);
  free(lc);
  free(ls);
  free(li);
  free(lj);
  return nr;
}
