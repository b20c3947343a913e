/*
 * CPU oracle for the 2D pair-HMM forward/backward, expected transition
 * counts, expected emission counts and posteriors, in a compile-time
 * precision REAL (double, or long double for reference values on long pairs).
 *
 * TEST INFRASTRUCTURE ONLY (golden fixtures and their checks); the shipped
 * library never links or calls this file.
 *
 * The recursion is that of python/tkfmixdom/jax/dp/hmm.py (forward
 * :1186-1292, backward :1295-1437, counts :1930-1971, E column at the real
 * corner :1969-1971), in log space.  Each log-sum-exp over predecessor
 * states is evaluated in the max-shifted product form
 *     logsumexp_s (F[s] + T[s,k]) = m + log sum_s exp(F[s] - m) exp(T[s,k]),
 *     m = max_s F[s],
 * which is the same quantity (exp(T) is taken once per pair), so a cell
 * costs S exponentials and S^2 multiply-adds instead of S^2 exponentials;
 * the backward and the counts are shifted the same way.  Emission counts
 * follow emissions.emission_counts (train/likelihood.py:962-1028; M over
 * i, j >= 1, I over j >= 1, D over i >= 1).  The cells of one antidiagonal
 * are independent and are split over OpenMP threads; each thread keeps its
 * own count accumulators, summed at the end.
 *
 * Build: REAL defaults to double; -DREAL_LONG_DOUBLE gives long double.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef REAL_LONG_DOUBLE
typedef long double real;
#define EXP expl
#define LOG logl
#define FNAME oracle_fb2d_scaled_ld
#else
typedef double real;
#define EXP exp
#define LOG log
#define FNAME oracle_fb2d_scaled
#endif

#define NEG_INF (-1e30)
#define A_PAD 21
#define TAB (A_PAD * A_PAD + 2 * A_PAD)
enum { ST_S = 0, ST_M = 1, ST_I = 2, ST_D = 3, ST_E = 4 };

/* log emission of state type ty, group g at a cell with residues xi, yj;
 * the table layout is [g][21x21 match | 21 insert | 21 delete]. */
static inline int emit_entry(int ty, int xi, int yj) {
  if (ty == ST_M) return xi * A_PAD + yj;
  if (ty == ST_I) return A_PAD * A_PAD + yj;
  return A_PAD * A_PAD + A_PAD + xi;
}

static inline int type_ok(int ty, int i, int j) {
  return (ty == ST_M) ? (i >= 1 && j >= 1) : (ty == ST_I) ? (j >= 1) : (ty == ST_D) ? (i >= 1) : 0;
}

/*
 * One pair.  log_trans (S,S), log_tab (G,TAB) log values; x, y codes 0..20.
 * Outputs (double, rounded from REAL): *logP, n (S,S) (may be NULL),
 * ecnt (G,TAB) (may be NULL), and posteriors at n_samp cells
 * (samp_i[q], samp_j[q]) into post_s (n_samp,S) (may be NULL).
 * Returns 0, or -2 when out of memory.
 */
int FNAME(int S, const int32_t* types, const int32_t* grp, int G, const double* log_trans,
          const double* log_tab, const uint8_t* x, int Lx, const uint8_t* y, int Ly, int nthreads,
          double* out_logP, double* out_n, double* out_ecnt, int n_samp, const int32_t* samp_i,
          const int32_t* samp_j, double* out_post) {
  const int W = Ly + 1;
  int e_idx = -1;
  for (int k = 0; k < S; ++k)
    if (types[k] == ST_E) { e_idx = k; break; }
  if (e_idx < 0) return -1;
  if (nthreads < 1) nthreads = 1;
  const size_t cells = (size_t)(Lx + 1) * W;
  real* F = (real*)malloc(cells * S * sizeof(real));
  real* Bd = (real*)malloc((size_t)3 * (Lx + 1) * S * sizeof(real)); /* 3 antidiagonals of B */
  real* eT = (real*)malloc((size_t)S * S * sizeof(real));
  real* acc_n = (real*)calloc((size_t)nthreads * S * S, sizeof(real));
  real* acc_e = (real*)calloc((size_t)nthreads * G * TAB, sizeof(real));
  real* scratch = (real*)malloc((size_t)nthreads * 4 * S * sizeof(real));
  if (!F || !Bd || !eT || !acc_n || !acc_e || !scratch) {
    free(F); free(Bd); free(eT); free(acc_n); free(acc_e); free(scratch);
    return -2;
  }
  for (int q = 0; q < S * S; ++q) eT[q] = EXP((real)log_trans[q]);
#define FAT(i, j, k) F[(((size_t)(i)) * W + (j)) * S + (k)]
#define BAT(d, i, k) Bd[((size_t)((d) % 3) * (Lx + 1) + (i)) * S + (k)]
  /* ---- forward ---- */
  for (int d = 0; d <= Lx + Ly; ++d) {
    const int i0 = d - Ly > 0 ? d - Ly : 0, i1 = d < Lx ? d : Lx;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int i = i0; i <= i1; ++i) {
#ifdef _OPENMP
      real* u = scratch + (size_t)omp_get_thread_num() * 4 * S;
#else
      real* u = scratch;
#endif
      const int j = d - i;
      if (d == 0) {
        for (int k = 0; k < S; ++k) FAT(0, 0, k) = (k == 0) ? 0.0L : (real)NEG_INF;
        continue;
      }
      /* predecessors by type: M (i-1,j-1), I (i,j-1), D (i-1,j) */
      const int pi[4] = {0, i - 1, i, i - 1}, pj[4] = {0, j - 1, j - 1, j};
      real m[4];
      for (int t = ST_M; t <= ST_D; ++t) {
        if (!type_ok(t, i, j)) continue;
        real mx = -INFINITY;
        for (int s = 0; s < S; ++s) {
          const real v = FAT(pi[t], pj[t], s);
          if (v > mx) mx = v;
        }
        m[t] = mx;
        real* ut = u + (size_t)t * S;
        for (int s = 0; s < S; ++s) ut[s] = EXP(FAT(pi[t], pj[t], s) - mx);
      }
      const int xi = i >= 1 ? x[i - 1] : 0, yj = j >= 1 ? y[j - 1] : 0;
      for (int k = 0; k < S; ++k) {
        const int ty = types[k];
        real v = (real)NEG_INF;
        if (type_ok(ty, i, j)) {
          const real* ut = u + (size_t)ty * S;
          real acc = 0.0L;
          for (int s = 0; s < S; ++s) acc += ut[s] * eT[s * S + k];
          const real e = (real)log_tab[(size_t)grp[k] * TAB + emit_entry(ty, xi, yj)];
          v = (acc > 0.0L) ? m[ty] + LOG(acc) + e : (real)NEG_INF;
        }
        FAT(i, j, k) = v;
      }
    }
  }
  real mx = -INFINITY;
  for (int k = 0; k < S; ++k) {
    const real a = FAT(Lx, Ly, k) + (real)log_trans[k * S + e_idx];
    if (a > mx) mx = a;
  }
  real sum = 0.0L;
  for (int k = 0; k < S; ++k) sum += EXP(FAT(Lx, Ly, k) + (real)log_trans[k * S + e_idx] - mx);
  const real logP = mx + LOG(sum);
  *out_logP = (double)logP;
  /* ---- backward with counts and emission counts ---- */
  for (int d = Lx + Ly; d >= 0; --d) {
    const int i0 = d - Ly > 0 ? d - Ly : 0, i1 = d < Lx ? d : Lx;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int i = i0; i <= i1; ++i) {
#ifdef _OPENMP
      const int th = omp_get_thread_num();
#else
      const int th = 0;
#endif
      real* g = scratch + (size_t)th * 4 * S;   /* e(succ,k') + B(succ,k') */
      real* v = g + S;                          /* exp(g - mg) */
      real* uf = v + S;                         /* exp(F - mf) */
      real* an = acc_n + (size_t)th * S * S;
      real* ae = acc_e + (size_t)th * G * TAB;
      const int j = d - i;
      const int xi = i >= 1 ? x[i - 1] : 0, yj = j >= 1 ? y[j - 1] : 0;
      if (i == Lx && j == Ly) {
        for (int k = 0; k < S; ++k) BAT(d, i, k) = (real)log_trans[k * S + e_idx];
      } else {
        real mg = -INFINITY;
        for (int k2 = 0; k2 < S; ++k2) {
          const int ty = types[k2];
          int si = -1, sj = -1;
          if (ty == ST_M && i < Lx && j < Ly) { si = i + 1; sj = j + 1; }
          else if (ty == ST_I && j < Ly) { si = i; sj = j + 1; }
          else if (ty == ST_D && i < Lx) { si = i + 1; sj = j; }
          if (si >= 0) {
            const int xs = si >= 1 ? x[si - 1] : 0, ys = sj >= 1 ? y[sj - 1] : 0;
            g[k2] = (real)log_tab[(size_t)grp[k2] * TAB + emit_entry(ty, xs, ys)] + BAT(si + sj, si, k2);
          } else {
            g[k2] = (real)NEG_INF;
          }
          if (g[k2] > mg) mg = g[k2];
        }
        for (int k2 = 0; k2 < S; ++k2) v[k2] = EXP(g[k2] - mg);
        real mf = -INFINITY;
        for (int k = 0; k < S; ++k)
          if (FAT(i, j, k) > mf) mf = FAT(i, j, k);
        for (int k = 0; k < S; ++k) uf[k] = EXP(FAT(i, j, k) - mf);
        const real sc = EXP(mf + mg - logP);
        for (int k = 0; k < S; ++k) {
          real acc = 0.0L;
          for (int k2 = 0; k2 < S; ++k2) acc += eT[k * S + k2] * v[k2];
          BAT(d, i, k) = (acc > 0.0L) ? mg + LOG(acc) : (real)NEG_INF;
          if (out_n) {
            const real f = sc * uf[k];
            if (f != 0.0L)
              for (int k2 = 0; k2 < S; ++k2) an[k * S + k2] += f * eT[k * S + k2] * v[k2];
          }
        }
      }
      if (out_ecnt) {
        for (int k = 0; k < S; ++k) {
          const int ty = types[k];
          if (!type_ok(ty, i, j)) continue;
          const real pst = EXP(FAT(i, j, k) + BAT(d, i, k) - logP);
          ae[(size_t)grp[k] * TAB + emit_entry(ty, xi, yj)] += pst;
        }
      }
      /* sampled posteriors of this cell */
      for (int q = 0; q < n_samp && out_post; ++q)
        if (samp_i[q] == i && samp_j[q] == j)
          for (int k = 0; k < S; ++k) out_post[(size_t)q * S + k] = (double)EXP(FAT(i, j, k) + BAT(d, i, k) - logP);
    }
  }
  if (out_n) {
    for (int q = 0; q < S * S; ++q) {
      real t = 0.0L;
      for (int th = 0; th < nthreads; ++th) t += acc_n[(size_t)th * S * S + q];
      out_n[q] = (double)t;
    }
    for (int k = 0; k < S; ++k)
      out_n[k * S + e_idx] += (double)EXP(FAT(Lx, Ly, k) + (real)log_trans[k * S + e_idx] - logP);
  }
  if (out_ecnt)
    for (int q = 0; q < G * TAB; ++q) {
      real t = 0.0L;
      for (int th = 0; th < nthreads; ++th) t += acc_e[(size_t)th * G * TAB + q];
      out_ecnt[q] = (double)t;
    }
#undef FAT
#undef BAT
  free(F); free(Bd); free(eT); free(acc_n); free(acc_e); free(scratch);
  return 0;
}
