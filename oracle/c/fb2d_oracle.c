/*
 * FP64 CPU oracle for the 2D pair-HMM forward/backward/expected counts.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); the
 * shipped library never links or calls this file.
 *
 * A log-space restatement of python/tkfmixdom/jax/dp/hmm.py:
 *   forward   hmm.py:1186-1292  (F[0,0,0] = 0; predecessor by the type of
 *             the destination state: M (i-1,j-1), I (i,j-1), D (i-1,j))
 *   backward  hmm.py:1295-1437  (B[Lx,Ly,:] = T[:,E]; successor by the type
 *             of the destination state)
 *   counts    hmm.py:1930-1971  (n[k,k'] = sum_cells exp(F + T + e + B - logP),
 *             plus the E column at the real corner)
 * Emissions are looked up on the fly from per-group log tables:
 *   M: tabM[g][x_{i-1}][y_{j-1}], I: tabI[g][y_{j-1}], D: tabD[g][x_{i-1}],
 * the same values the reference materialises as its (Lx+1,Ly+1,S) table
 * (hmm.py:628-851).  NEG_INF = -1e30 sentinels as in hmm.py:108.
 * Pairs are independent: OpenMP parallel-for over pairs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NEG_INF (-1e30)
#define A_PAD 21
enum { ST_S = 0, ST_M = 1, ST_I = 2, ST_D = 3, ST_E = 4 };

static inline double lse2(double a, double b) {
  /* jnp.logaddexp */
  double m = a > b ? a : b;
  return m + log(exp(a - m) + exp(b - m));
}

static inline double emit_log(int type, int g, int xi, int yj,
                              const double* tabM, const double* tabI,
                              const double* tabD) {
  switch (type) {
    case ST_M: return tabM[(g * A_PAD + xi) * A_PAD + yj];
    case ST_I: return tabI[g * A_PAD + yj];
    case ST_D: return tabD[g * A_PAD + xi];
    default: return 0.0;
  }
}

/* One pair.  F, Bk are caller scratch of (Lx+1)(Ly+1)S doubles each. */
static void fb_pair(int S, const int32_t* types, const int32_t* grp,
                    const double* T, const double* tabM, const double* tabI,
                    const double* tabD, const uint8_t* x, int Lx,
                    const uint8_t* y, int Ly, int forward_only, double* F,
                    double* Bk, double* tmp, double* out_logP, double* out_n,
                    double* out_post) {
  const int W = Ly + 1;
  int e_idx = 0;
  for (int k = 0; k < S; ++k)
    if (types[k] == ST_E) { e_idx = k; break; }
#define AT(A, i, j, k) (A)[(((size_t)(i)) * W + (j)) * S + (k)]
  /* ---- forward ---- */
  for (int d = 0; d <= Lx + Ly; ++d) {
    int i0 = d - Ly > 0 ? d - Ly : 0, i1 = d < Lx ? d : Lx;
    for (int i = i0; i <= i1; ++i) {
      int j = d - i;
      for (int k = 0; k < S; ++k) {
        double v = NEG_INF;
        if (d == 0) {
          v = (k == 0) ? 0.0 : NEG_INF;
        } else {
          int ty = types[k], pi = -1, pj = -1;
          if (ty == ST_M && i >= 1 && j >= 1) { pi = i - 1; pj = j - 1; }
          else if (ty == ST_I && j >= 1) { pi = i; pj = j - 1; }
          else if (ty == ST_D && i >= 1) { pi = i - 1; pj = j; }
          if (pi >= 0) {
            double mx = -INFINITY;
            for (int s = 0; s < S; ++s) {
              double a = AT(F, pi, pj, s) + T[s * S + k];
              tmp[s] = a;
              if (a > mx) mx = a;
            }
            double sum = 0.0;
            for (int s = 0; s < S; ++s) sum += exp(tmp[s] - mx);
            v = mx + log(sum) +
                emit_log(ty, grp[k], i >= 1 ? x[i - 1] : 0, j >= 1 ? y[j - 1] : 0,
                         tabM, tabI, tabD);
          }
        }
        AT(F, i, j, k) = v;
      }
    }
  }
  double mx = -INFINITY;
  for (int k = 0; k < S; ++k) {
    double a = AT(F, Lx, Ly, k) + T[k * S + e_idx];
    tmp[k] = a;
    if (a > mx) mx = a;
  }
  double sum = 0.0;
  for (int k = 0; k < S; ++k) sum += exp(tmp[k] - mx);
  const double logP = mx + log(sum);
  *out_logP = logP;
  if (forward_only) return;
  /* ---- backward ---- */
  for (int d = Lx + Ly; d >= 0; --d) {
    int i0 = d - Ly > 0 ? d - Ly : 0, i1 = d < Lx ? d : Lx;
    for (int i = i0; i <= i1; ++i) {
      int j = d - i;
      if (i == Lx && j == Ly) {
        for (int k = 0; k < S; ++k) AT(Bk, i, j, k) = T[k * S + e_idx];
        continue;
      }
      /* g[k'] = e(succ,k') + B[succ,k'] for the type-shifted successor */
      for (int k2 = 0; k2 < S; ++k2) {
        int ty = types[k2], si = -1, sj = -1;
        if (ty == ST_M && i < Lx && j < Ly) { si = i + 1; sj = j + 1; }
        else if (ty == ST_I && j < Ly) { si = i; sj = j + 1; }
        else if (ty == ST_D && i < Lx) { si = i + 1; sj = j; }
        if (si >= 0) {
          int xc = si >= 1 ? x[si - 1] : 0, yc = sj >= 1 ? y[sj - 1] : 0;
          tmp[k2] = emit_log(ty, grp[k2], xc, yc, tabM, tabI, tabD) +
                    AT(Bk, si, sj, k2);
        } else {
          tmp[k2] = NEG_INF;
        }
      }
      for (int k = 0; k < S; ++k) {
        double m2 = -INFINITY;
        for (int k2 = 0; k2 < S; ++k2) {
          double a = T[k * S + k2] + tmp[k2];
          if (a > m2) m2 = a;
        }
        double s2 = 0.0;
        for (int k2 = 0; k2 < S; ++k2) s2 += exp(T[k * S + k2] + tmp[k2] - m2);
        AT(Bk, i, j, k) = m2 + log(s2);
        if (out_n) {
          double fk = AT(F, i, j, k);
          for (int k2 = 0; k2 < S; ++k2)
            out_n[k * S + k2] += exp(fk + T[k * S + k2] + tmp[k2] - logP);
        }
      }
    }
  }
  if (out_n)
    for (int k = 0; k < S; ++k)
      out_n[k * S + e_idx] += exp(AT(F, Lx, Ly, k) + T[k * S + e_idx] - logP);
  if (out_post)
    for (size_t c = 0; c < (size_t)(Lx + 1) * W * S; ++c)
      out_post[c] = exp(F[c] + Bk[c] - logP);
#undef AT
}

/*
 * Batched entry.  log_trans: (nT,S,S) with nT in {1,B}; tables: (G,21,21),
 * (G,21), (G,21) log values; x_cat/y_cat uint8 codes with per-pair offsets
 * and lengths.  out_n (B,S,S) and out_post (ragged sum (Lx+1)(Ly+1)S, with
 * post_off giving each pair's start) may be NULL.  Returns 0 on success.
 */
int oracle_fb2d_batch(int B, int S, const int32_t* types, const int32_t* grp,
                      const double* log_trans, int n_trans, const double* tabM,
                      const double* tabI, const double* tabD,
                      const uint8_t* x_cat, const int64_t* x_off,
                      const int32_t* x_len, const uint8_t* y_cat,
                      const int64_t* y_off, const int32_t* y_len,
                      int forward_only, double* out_logP, double* out_n,
                      double* out_post, const int64_t* post_off, int nthreads) {
  if (B < 0 || S <= 0) return -1;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int b = 0; b < B; ++b) {
    int Lx = x_len[b], Ly = y_len[b];
    size_t cells = (size_t)(Lx + 1) * (Ly + 1) * S;
    double* F = (double*)malloc(cells * sizeof(double));
    double* Bk = forward_only ? NULL : (double*)malloc(cells * sizeof(double));
    double* tmp = (double*)malloc((size_t)S * sizeof(double));
    if (!F || (!forward_only && !Bk) || !tmp) {
#pragma omp atomic write
      err = -2;
    } else {
      double* n = out_n ? out_n + (size_t)b * S * S : NULL;
      if (n) memset(n, 0, (size_t)S * S * sizeof(double));
      fb_pair(S, types, grp, log_trans + (n_trans > 1 ? (size_t)b * S * S : 0),
              tabM, tabI, tabD, x_cat + x_off[b], Lx, y_cat + y_off[b], Ly,
              forward_only, F, Bk, tmp, out_logP + b, n,
              out_post ? out_post + post_off[b] : NULL);
    }
    free(F);
    free(Bk);
    free(tmp);
  }
  return err;
}
