// tkfmi: null-state count restoration, sufficient statistics and BDI
// statistics on gfx950, FP64.
//
// Reference algorithm: python/tkfmixdom/jax/models/exact_suffstats_batch.py
// (exploded machine :61-253, elimination :272-282, ghost-usage restoration
// :285-339, extraction :347-458, chain :468-541) and core/bdi.py (smooth WFST
// likelihood :101-160, general / equal-rate BDI :682-778).
//
// Design:
//  * One workgroup (256 threads) per pair.  The exploded transition matrix of
//    the pair is assembled in its workspace slice from an "entry program" the
//    host derives once per call from the shared parameters: each nonzero
//    entry is coef * (1 or one entry of a per-pair TKF91 matrix), and the
//    TKF91 matrices of the main level and of each domain are computed in LDS
//    from t.
//  * The six eliminations run back to back: C = (I - T_ZZ)^-1 by Gauss-Jordan
//    (I - T_ZZ is a nonsingular M-matrix: T is substochastic on the null
//    group, so no pivoting is needed), W = C T_ZK, T' = T_KK + T_KZ W.  Every
//    intermediate T, C and W stays in the workspace for the restoration.
//  * Restoration walks the steps backwards with the ghost-usage products;
//    each output element is one thread's dot product.
//  * Statistics are gathered from the restored counts through a CSR program
//    (one output per thread, fixed order); BDI statistics per level are
//    computed by one thread each, with the derivatives of the smooth WFST
//    log-likelihood obtained exactly in hyper-dual arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "expm_lds.h"
#include "fb2d_internal.h"
#include "tkfmi.h"

namespace tkfmi_ss {

constexpr int S_ = 0, M_ = 1, I_ = 2, D_ = 3, E_ = 4;
constexpr int kMaxK = 16;
constexpr int kThreads = 256;

// ------------------------------------------------------------------ TKF91
// alpha, beta, gamma, kappa with the equal-rate (L'Hopital) branch for
// |1 - kappa| < 1e-4 (core/bdi.py:40-97); rows S, M, I use beta, row D gamma
// (core/params.py:29-54).
__host__ __device__ inline void tkf91_trans(double lam, double mu, double t, double* tau) {
  const double s = mu * t;
  const double alpha = exp(-s);
  const double kappa = lam / mu;
  double beta, gamma;
  if (fabs(1.0 - kappa) < 1e-4) {
    beta = s / (1.0 + s);
    const double phi = (s < 1e-10) ? 1.0 - s / 2.0 : (1.0 - alpha) / fmax(s, 1e-30);
    gamma = 1.0 - 1.0 / ((1.0 + s) * phi);
  } else {
    const double eta = exp(-lam * t);
    const double numer = lam * (eta - alpha);
    const double denom = mu * eta - lam * alpha;
    beta = (fabs(denom) < 1e-30) ? 0.0 : numer / denom;
    const double Phi = 1.0 - alpha;
    gamma = (fabs(Phi) < 1e-30) ? 0.0 : 1.0 - mu * beta / (lam * fmax(Phi, 1e-30));
  }
  for (int q = 0; q < 25; ++q) tau[q] = 0.0;
  for (int src = S_; src <= D_; ++src) {
    const double b = (src == D_) ? gamma : beta;
    tau[src * 5 + M_] = (1 - b) * kappa * alpha;
    tau[src * 5 + I_] = b;
    tau[src * 5 + D_] = (1 - b) * kappa * (1 - alpha);
    tau[src * 5 + E_] = (1 - b) * (1 - kappa);
  }
}

// ------------------------------------------------------ hyper-dual numbers
// x = a + b e1 + c e2 + d e1 e2 with e1^2 = e2^2 = 0: evaluating f at
// (lam + e1, mu + e2) yields f, df/dlam, df/dmu and d2f/dlam dmu exactly.
struct HD {
  double a, b, c, d;
};
__device__ inline HD hd(double a, double b = 0.0, double c = 0.0, double d = 0.0) { return HD{a, b, c, d}; }
__device__ inline HD operator+(HD x, HD y) { return HD{x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
__device__ inline HD operator-(HD x, HD y) { return HD{x.a - y.a, x.b - y.b, x.c - y.c, x.d - y.d}; }
__device__ inline HD operator*(HD x, HD y) {
  return HD{x.a * y.a, x.a * y.b + x.b * y.a, x.a * y.c + x.c * y.a,
            x.a * y.d + x.b * y.c + x.c * y.b + x.d * y.a};
}
__device__ inline HD operator*(double s, HD x) { return HD{s * x.a, s * x.b, s * x.c, s * x.d}; }
__device__ inline HD operator+(double s, HD x) { return HD{s + x.a, x.b, x.c, x.d}; }
__device__ inline HD operator-(double s, HD x) { return HD{s - x.a, -x.b, -x.c, -x.d}; }
// f(x) from f(a), f'(a), f''(a)
__device__ inline HD chain(HD x, double f0, double f1, double f2) {
  return HD{f0, f1 * x.b, f1 * x.c, f1 * x.d + f2 * x.b * x.c};
}
__device__ inline HD recip(HD x) {
  const double r = 1.0 / x.a;
  return chain(x, r, -r * r, 2.0 * r * r * r);
}
__device__ inline HD operator/(HD x, HD y) { return x * recip(y); }
__device__ inline HD hexp(HD x) {
  const double e = exp(x.a);
  return chain(x, e, e, e);
}
__device__ inline HD hexpm1(HD x) {
  const double e = exp(x.a);
  return chain(x, expm1(x.a), e, e);
}
__device__ inline HD hlog(HD x) { return chain(x, log(x.a), 1.0 / x.a, -1.0 / (x.a * x.a)); }
// max(x, c) with the derivative of the selected branch
__device__ inline HD hmaxc(HD x, double c) { return (x.a > c) ? x : hd(c); }

// Smooth WFST log-likelihood sum_g n_g log xi_g (core/bdi.py:101-160).
__device__ inline HD wfst_logp(HD lam, HD mu, double t, const double* g) {
  const HD alpha = hexp(-t * mu);
  const HD x = t * (mu - lam);
  HD r;
  if (fabs(x.a) < 1e-5) {
    r = 1.0 + 0.5 * x + (1.0 / 6.0) * (x * x) + (1.0 / 24.0) * (x * x * x);
  } else {
    r = hexpm1(x) / x;
  }
  const HD tr = t * r;
  const HD beta = (lam * tr) / (mu * tr + hd(1.0));
  const HD Phi = 1.0 - alpha;
  const HD gamma = 1.0 - (mu * beta) / (lam * hmaxc(Phi, 1e-30));
  const HD v[6] = {alpha, 1.0 - alpha, beta, 1.0 - beta, gamma, 1.0 - gamma};
  HD acc = hd(0.0);
  for (int q = 0; q < 6; ++q) acc = acc + g[q] * hlog(hmaxc(v[q], 1e-30));
  return acc;
}

// BDI statistics of one 5x5 count matrix (core/bdi.py:682-778).
__device__ inline void bdi_stats(const double* n, double lam, double mu, double t, double T,
                                 double* EB, double* ED, double* ES) {
  double g[6];
  g[0] = n[S_ * 5 + M_] + n[M_ * 5 + M_] + n[I_ * 5 + M_] + n[D_ * 5 + M_];
  g[1] = n[S_ * 5 + D_] + n[M_ * 5 + D_] + n[I_ * 5 + D_] + n[D_ * 5 + D_];
  g[2] = n[S_ * 5 + I_] + n[M_ * 5 + I_] + n[I_ * 5 + I_];
  g[3] = n[S_ * 5 + M_] + n[S_ * 5 + D_] + n[S_ * 5 + E_] + n[M_ * 5 + M_] + n[M_ * 5 + D_] +
         n[M_ * 5 + E_] + n[I_ * 5 + M_] + n[I_ * 5 + D_] + n[I_ * 5 + E_];
  g[4] = n[D_ * 5 + I_];
  g[5] = n[D_ * 5 + M_] + n[D_ * 5 + D_] + n[D_ * 5 + E_];
  double cons = 0.0;
  for (int a = 0; a < 5; ++a) cons += n[a * 5 + I_] - n[a * 5 + D_];
  const HD lp = wfst_logp(hd(lam, 1.0), hd(mu, 0.0, 1.0), t, g);
  const double lam_score = lam * lp.b, mu_score = mu * lp.c;
  double es;
  if (fabs(1.0 - lam / mu) < 1e-10) {
    // E_S = f'(mu), f(l) = cons + mu dlogP/dmu(l, mu) - l dlogP/dlam(l, mu) - l T
    const HD at = wfst_logp(hd(mu, 1.0), hd(mu, 0.0, 1.0), t, g);    // d/dlam, d2/dlam dmu
    const HD ll = wfst_logp(hd(mu, 1.0, 1.0), hd(mu), t, g);         // d2/dlam2
    es = mu * at.d - at.b - mu * ll.d - T;
  } else {
    es = (cons + mu_score - lam_score - lam * T) / (lam - mu);
  }
  *ES = es;
  *EB = lam_score + lam * es + lam * T;
  *ED = mu_score + mu * es;
}

// ------------------------------------------------------- restoration kernel
struct Entry {
  int32_t row, col, tau_sel, pad;  // value = coef * (tau_sel < 0 ? 1 : tau[tau_sel])
  double coef;
};

struct Step {
  int32_t n, nk, nz, pad;
  int32_t keep_off, elim_off, pos_off, pad2;  // into the index array
  int64_t T_off, C_off, W_off;                // doubles, into the pair's slice
};

struct RParams {
  int32_t K, F, N0, Nc;
  int32_t n_entries, n_stats, want_bdi, pad;
  double lam[kMaxK + 1], mu[kMaxK + 1];  // level 0 = main, 1 + k = domain k
  Step steps[6];
  int64_t nbuf_off, gj_off, sc_off, ct_off, cts_off, sh_off;  // scratch (doubles)
  int64_t slice;                                               // doubles per pair
  int32_t off_top, off_domM, off_kap, off_kap1;
  const Entry* entries;
  const int32_t* idx;      // keep / elim / position lists of every step
  const int32_t* g_ptr;    // CSR of the statistics gathers [n_stats + 1]
  const int32_t* g_src;    // flat source index src * N0 + dst
  const double* t;
  const double* n_chi;
  double* stats;
  double* bdi;
  double* ws;
};

// Assembles the exploded machine of one pair at time t in its workspace
// slice and runs the six eliminations.  Afterwards T_s sits at steps[s].T_off
// (s = 0..5), the C and W factors at steps[s].C_off / W_off, and the reduced
// matrix T_6 -- the collapsed chi in its state order -- right after T_5.
__device__ void explode_and_eliminate(const RParams& P, double* ws, double t, double* s_tau,
                                      double* s_fac) {
  const int tid = threadIdx.x;
  const int K = P.K;
  if (tid <= K) tkf91_trans(P.lam[tid], P.mu[tid], t, s_tau + tid * 25);

  // ---- exploded matrix
  double* T0 = ws + P.steps[0].T_off;
  const int N0 = P.N0;
  for (int q = tid; q < N0 * N0; q += kThreads) T0[q] = 0.0;
  __syncthreads();
  for (int q = tid; q < P.n_entries; q += kThreads) {
    const Entry e = P.entries[q];
    T0[e.row * N0 + e.col] = e.coef * (e.tau_sel < 0 ? 1.0 : s_tau[e.tau_sel]);
  }
  __syncthreads();

  // ---- eliminations
  double* A = ws + P.gj_off;
  for (int s = 0; s < 6; ++s) {
    const Step st = P.steps[s];
    const int n = st.n, nk = st.nk, nz = st.nz;
    const int32_t* kp = P.idx + st.keep_off;
    const int32_t* ez = P.idx + st.elim_off;
    const double* T = ws + st.T_off;
    double* C = ws + st.C_off;
    double* W = ws + st.W_off;
    double* Tn = ws + P.steps[s].T_off + (int64_t)n * n;  // T_{s+1} follows T_s
    double* Inv = A + (size_t)nz * nz;
    for (int q = tid; q < nz * nz; q += kThreads) {
      const int i = q / nz, j = q % nz;
      A[q] = (i == j ? 1.0 : 0.0) - T[ez[i] * n + ez[j]];
      Inv[q] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int p = 0; p < nz; ++p) {
      const double inv_piv = 1.0 / A[p * nz + p];
      __syncthreads();
      for (int j = tid; j < nz; j += kThreads) {
        A[p * nz + j] *= inv_piv;
        Inv[p * nz + j] *= inv_piv;
      }
      __syncthreads();
      for (int i = tid; i < nz; i += kThreads) s_fac[i] = A[i * nz + p];
      __syncthreads();
      for (int q = tid; q < nz * nz; q += kThreads) {
        const int i = q / nz, j = q % nz;
        if (i == p) continue;
        const double f = s_fac[i];
        A[q] -= f * A[p * nz + j];
        Inv[q] -= f * Inv[p * nz + j];
      }
      __syncthreads();
    }
    for (int q = tid; q < nz * nz; q += kThreads) C[q] = Inv[q];
    __syncthreads();
    for (int q = tid; q < nz * nk; q += kThreads) {  // W = C T_ZK
      const int a = q / nk, j = q % nk;
      double acc = 0.0;
      for (int c = 0; c < nz; ++c) acc = fma(C[a * nz + c], T[ez[c] * n + kp[j]], acc);
      W[q] = acc;
    }
    __syncthreads();
    for (int q = tid; q < nk * nk; q += kThreads) {  // T' = T_KK + T_KZ W
      const int i = q / nk, j = q % nk;
      double acc = T[kp[i] * n + kp[j]];
      for (int a = 0; a < nz; ++a) acc = fma(T[kp[i] * n + ez[a]], W[a * nk + j], acc);
      Tn[q] = acc;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void restore_kernel(RParams P) {
  __shared__ double s_tau[(kMaxK + 1) * 25];
  __shared__ double s_fac[1024];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int K = P.K;
  double* ws = P.ws + (size_t)b * P.slice;
  const double t = P.t[b];
  explode_and_eliminate(P, ws, t, s_tau, s_fac);
  const int N0 = P.N0;

  // ---- restoration (n_6 = n_chi, then back to n_0)
  double* nb0 = ws + P.nbuf_off;
  double* nb1 = nb0 + (size_t)N0 * N0;
  const double* nin = P.n_chi + (size_t)b * P.Nc * P.Nc;
  for (int q = tid; q < P.Nc * P.Nc; q += kThreads) nb0[q] = nin[q];
  __syncthreads();
  double* cur = nb0;
  double* nxt = nb1;
  double* Sc = ws + P.sc_off;
  double* Ct = ws + P.ct_off;
  double* CtS = ws + P.cts_off;
  double* SH = ws + P.sh_off;
  for (int s = 5; s >= 0; --s) {
    const Step st = P.steps[s];
    const int n = st.n, nk = st.nk, nz = st.nz;
    const int32_t* kp = P.idx + st.keep_off;
    const int32_t* ez = P.idx + st.elim_off;
    const int32_t* pos = P.idx + st.pos_off;  // >= 0: keep position, < 0: -(elim position) - 1
    const double* T = ws + st.T_off;
    const double* Tr = T + (int64_t)n * n;
    const double* C = ws + st.C_off;
    const double* W = ws + st.W_off;
    for (int q = tid; q < nk * nk; q += kThreads) {
      const double tr = Tr[q];
      Sc[q] = (fabs(tr) > 1e-30) ? cur[q] / tr : 0.0;
    }
    for (int q = tid; q < nk * nz; q += kThreads) {  // Ct = T_KZ C
      const int i = q / nz, a = q % nz;
      double acc = 0.0;
      for (int c = 0; c < nz; ++c) acc = fma(T[kp[i] * n + ez[c]], C[c * nz + a], acc);
      Ct[q] = acc;
    }
    __syncthreads();
    for (int q = tid; q < nk * nz; q += kThreads) {  // SH = Sc W^T
      const int i = q / nz, a = q % nz;
      double acc = 0.0;
      for (int j = 0; j < nk; ++j) acc = fma(Sc[i * nk + j], W[a * nk + j], acc);
      SH[q] = acc;
    }
    for (int q = tid; q < nz * nk; q += kThreads) {  // CtS = Ct^T Sc
      const int a = q / nk, j = q % nk;
      double acc = 0.0;
      for (int i = 0; i < nk; ++i) acc = fma(Ct[i * nz + a], Sc[i * nk + j], acc);
      CtS[q] = acc;
    }
    __syncthreads();
    for (int q = tid; q < n * n; q += kThreads) {
      const int u = q / n, v = q % n;
      const double tuv = T[q];
      double val = 0.0;
      if (tuv != 0.0) {
        const int pu = pos[u], pv = pos[v];
        if (pu >= 0 && pv >= 0) val = Sc[pu * nk + pv] * tuv;
        else if (pu >= 0) val = tuv * SH[pu * nz + (-pv - 1)];
        else if (pv >= 0) val = tuv * CtS[(-pu - 1) * nk + pv];
        else {
          const int a = -pu - 1, c = -pv - 1;
          double acc = 0.0;
          for (int j = 0; j < nk; ++j) acc = fma(CtS[a * nk + j], W[c * nk + j], acc);
          val = tuv * acc;
        }
      }
      nxt[q] = val;
    }
    __syncthreads();
    double* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }

  // ---- statistics
  double* out = P.stats + (size_t)b * P.n_stats;
  for (int o = tid; o < P.n_stats; o += kThreads) {
    double acc = 0.0;
    for (int q = P.g_ptr[o]; q < P.g_ptr[o + 1]; ++q) acc += cur[P.g_src[q]];
    out[o] = acc;
  }
  if (!P.want_bdi) return;
  __syncthreads();
  if (tid <= K) {
    const double* n5 = (tid == 0) ? out + P.off_top : out + P.off_domM + (tid - 1) * 25;
    double Tobs;
    if (tid == 0) {
      Tobs = 0.0;
      for (int a = 0; a < 5; ++a) Tobs += n5[a * 5 + E_];
    } else {
      Tobs = out[P.off_kap1 + tid - 1];
      for (int c = 0; c < 5; ++c) Tobs += n5[S_ * 5 + c];
    }
    Tobs *= t;
    double eb, ed, es;
    bdi_stats(n5, P.lam[tid], P.mu[tid], t, Tobs, &eb, &ed, &es);
    double* o4 = P.bdi + ((size_t)b * (K + 1) + tid) * 4;
    o4[0] = eb;
    o4[1] = ed;
    o4[2] = es;
    o4[3] = Tobs;
  }
}

__global__ void bdi_kernel(int B, double lam, double mu, const double* __restrict__ n5,
                           const double* __restrict__ t, const double* __restrict__ T,
                           double* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double eb, ed, es;
  bdi_stats(n5 + (size_t)b * 25, lam, mu, t[b], T[b], &eb, &ed, &es);
  out[b * 3 + 0] = eb;
  out[b * 3 + 1] = ed;
  out[b * 3 + 2] = es;
}

// TKF92 self-loop split, kappa coefficients and clamped BDI statistics
// (train/tkf92_svi_bw.py:75-123, 160-167, 312-334).
__global__ void tkf92_estep_kernel(int B, double lam, double mu, double ext,
                                   const double* __restrict__ t, const double* __restrict__ n_chi,
                                   double* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double tau[25];
  tkf91_trans(lam, mu, t[b], tau);
  double n[25];
  for (int q = 0; q < 25; ++q) n[q] = n_chi[(size_t)b * 25 + q];
  double ext_count = 0.0;
  for (int s = M_; s <= D_; ++s) {
    const double diag92 = ext + (1.0 - ext) * tau[s * 5 + s];
    double frac = (diag92 > 1e-30) ? ext / fmax(diag92, 1e-30) : 0.0;
    frac = fmin(fmax(frac, 0.0), 1.0);
    const double loop = n[s * 5 + s];
    ext_count += frac * loop;
    n[s * 5 + s] = (1.0 - frac) * loop;
  }
  double notext = 0.0;
  for (int s = M_; s <= D_; ++s)
    for (int c = 0; c < 5; ++c) notext += n[s * 5 + c];
  const double L = n[S_ * 5 + M_] + n[S_ * 5 + D_] + n[M_ * 5 + M_] + n[M_ * 5 + D_] +
                   n[I_ * 5 + M_] + n[I_ * 5 + D_] + n[D_ * 5 + M_] + n[D_ * 5 + D_];
  const double Mc = n[S_ * 5 + E_] + n[M_ * 5 + E_] + n[I_ * 5 + E_] + n[D_ * 5 + E_];
  double eb, ed, es;
  bdi_stats(n, lam, mu, t[b], t[b], &eb, &ed, &es);
  double* o = out + (size_t)b * TKFMI_TKF92_STATS;
  for (int q = 0; q < 25; ++q) o[q] = n[q];
  o[25] = ext_count;
  o[26] = notext;
  o[27] = L;
  o[28] = Mc;
  o[29] = fmax(eb, 0.0);
  o[30] = fmax(ed, 0.0);
  o[31] = fmax(es, 0.0);
  o[32] = t[b];
  o[33] = 0.0;
  o[34] = 0.0;
}


// Holmes-Rubin weighted statistics for one (pair, class) per workgroup
// (core/ctmc.py:173-326).  With S = D^1/2 Q D^-1/2 = V diag(l) V^T shared by
// every pair of a class, the weighted sums over the A^4 integral tensor fold
// into A^3 products:
//   R~ = V^T (mc / M * sqrt(pi_b / pi_a)) V,  Y = V (J o R~) V^T,
//   W_i = Y_ii,  U_ij = Q_ij sqrt(pi_i / pi_j) Y_ij (i != j),
// with J_kl = (e^{l_k t} - e^{l_l t}) / (l_k - l_l), t e^{l_k t} when
// |l_k - l_l| < 1e-10, and M = D^-1/2 V e^{l t} V^T D^1/2.
constexpr int kHrMaxA = 32;

__global__ __launch_bounds__(256) void hr_kernel(int C, int A, const double* __restrict__ V_all,
                                                 const double* __restrict__ lam_all,
                                                 const double* __restrict__ pi_all,
                                                 const double* __restrict__ Q_all,
                                                 const double* __restrict__ t_all,
                                                 const double* __restrict__ mc_all,
                                                 double* __restrict__ W_out, double* __restrict__ U_out) {
  __shared__ double sV[kHrMaxA * kHrMaxA], sX[kHrMaxA * kHrMaxA], sT[kHrMaxA * kHrMaxA];
  __shared__ double sl[kHrMaxA], se[kHrMaxA], ssp[kHrMaxA];
  const int bc = blockIdx.x;
  const int c = bc % C;
  const int tid = threadIdx.x;
  const double t = t_all[bc / C];
  const double* V = V_all + (size_t)c * A * A;
  const double* Q = Q_all + (size_t)c * A * A;
  const double* mc = mc_all + (size_t)bc * A * A;
  const int AA = A * A;
  for (int q = tid; q < AA; q += 256) sV[q] = V[q];
  for (int q = tid; q < A; q += 256) {
    sl[q] = lam_all[c * A + q];
    se[q] = exp(sl[q] * t);
    ssp[q] = sqrt(pi_all[c * A + q]);
  }
  __syncthreads();
  // R'_ab = mc_ab / M_ab * sqrt(pi_b / pi_a)
  for (int q = tid; q < AA; q += 256) {
    const int a = q / A, b = q % A;
    double m = 0.0;
    for (int k = 0; k < A; ++k) m = fma(sV[a * A + k] * se[k], sV[b * A + k], m);
    m *= ssp[b] / ssp[a];
    const double ms = (fabs(m) < 1e-30) ? 1.0 : m;
    sX[q] = mc[q] / ms * (ssp[b] / ssp[a]);
  }
  __syncthreads();
  for (int q = tid; q < AA; q += 256) {  // sT = R' V
    const int a = q / A, l = q % A;
    double acc = 0.0;
    for (int b = 0; b < A; ++b) acc = fma(sX[a * A + b], sV[b * A + l], acc);
    sT[q] = acc;
  }
  __syncthreads();
  for (int q = tid; q < AA; q += 256) {  // sX = J o (V^T sT)
    const int k = q / A, l = q % A;
    double acc = 0.0;
    for (int a = 0; a < A; ++a) acc = fma(sV[a * A + k], sT[a * A + l], acc);
    const double d = sl[k] - sl[l];
    const double J = (fabs(d) < 1e-10) ? t * se[k] : (se[k] - se[l]) / d;
    sX[q] = J * acc;
  }
  __syncthreads();
  for (int q = tid; q < AA; q += 256) {  // sT = V sX
    const int i = q / A, l = q % A;
    double acc = 0.0;
    for (int k = 0; k < A; ++k) acc = fma(sV[i * A + k], sX[k * A + l], acc);
    sT[q] = acc;
  }
  __syncthreads();
  double* U = U_out + (size_t)bc * AA;
  for (int q = tid; q < AA; q += 256) {  // Y = sT V^T
    const int i = q / A, j = q % A;
    double y = 0.0;
    for (int l = 0; l < A; ++l) y = fma(sT[i * A + l], sV[j * A + l], y);
    if (i == j) {
      W_out[(size_t)bc * A + i] = y;
      U[q] = 0.0;
    } else {
      U[q] = Q[q] * (ssp[i] / ssp[j]) * y;
    }
  }
}


// ----------------------------------------------------------- parameter build
// Per pair at its own t: the log transition matrix (TKF91 / conditioned
// TKF91 / TKF92 closed forms, or the collapsed MixDom chi as the reduced
// matrix of the exploded machine) and the log emission tables from
// P_c(t) = expm(Q_c t) (shared, per-domain or per-class).
constexpr int kMaxSub = 16;
constexpr int kMaxA = 20;
constexpr double kNegInf = -1e30;

struct BParams {
  RParams R;  // MixDom elimination program (model == TKFMI_MODEL_MIXDOM)
  int32_t model, emission, log_floor, C;
  int32_t A, S, G, F;
  double lam, mu, ext;
  const double* Q;   // [C, A, A] rate matrices (any Q; reversibility is not assumed)
  const double* pi;
  const double* cd;  // [K, F, C] class distribution (per-class emissions)
  double* log_trans;
  double* log_tab;
};

// max(x, f) that keeps a NaN x (jnp.maximum propagates NaN; fmax does not).
__device__ inline double max_nan(double x, double f) { return (x >= f || x != x) ? x : f; }
__device__ inline double floor_log(double x, int mode) {
  return mode ? log(max_nan(x, 1e-30)) : ((x > 0.0 || x != x) ? log(x) : kNegInf);
}
__device__ inline double lse2(double a, double b) {  // logaddexp
  const double m = fmax(a, b);
  return m + log1p(exp(-fabs(a - b)));
}

using namespace tkfmi_expm;

__global__ __launch_bounds__(kThreads) void build_kernel(BParams B) {
  __shared__ double s_tau[(kMaxK + 1) * 25];
  __shared__ double s_fac[1024];
  __shared__ double s_P[kMaxSub * kMaxA * kMaxA];
  __shared__ double s_w[kExpmBufs * kMaxA * kMaxA];
  __shared__ int s_sh[4];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const double t = B.R.t[b];
  const int S = B.S, A = B.A, C = B.C;
  double* lt = B.log_trans + (size_t)b * S * S;
  if (B.model == TKFMI_MODEL_MIXDOM) {
    double* ws = B.R.ws + (size_t)b * B.R.slice;
    explode_and_eliminate(B.R, ws, t, s_tau, s_fac);
    const Step st = B.R.steps[5];
    const double* chi = ws + st.T_off + (int64_t)st.n * st.n;
    for (int q = tid; q < S * S; q += kThreads) lt[q] = floor_log(chi[q], B.log_floor);
  } else if (tid == 0) {
    double tau[25];
    tkf91_trans(B.lam, B.mu, t, tau);
    if (B.model == TKFMI_MODEL_TKF91_COND) {  // kappa and 1 - kappa divided out (core/params.py:57-89)
      const double kappa = B.lam / B.mu;
      for (int r = S_; r <= D_; ++r) {
        tau[r * 5 + M_] /= kappa;
        tau[r * 5 + D_] /= kappa;
        tau[r * 5 + E_] /= (1.0 - kappa);
      }
    } else if (B.model == TKFMI_MODEL_TKF92) {  // core/params.py:92-108
      for (int r = M_; r <= D_; ++r) {
        for (int c = 0; c < 5; ++c) tau[r * 5 + c] *= (1.0 - B.ext);
        tau[r * 5 + r] += B.ext;
      }
    }
    for (int q = 0; q < 25; ++q) lt[q] = floor_log(tau[q], B.log_floor);
  }
  __syncthreads();
  // ---- substitution matrices P_c(t) = expm(Q_c t)
  for (int c = 0; c < C; ++c)
    expm_lds(B.Q + (size_t)c * A * A, t, A, s_P + c * A * A, s_w, s_sh);
  // ---- tables [G][21x21 match | 21 insert | 21 delete]
  double* tab = B.log_tab + (size_t)b * B.G * TKFMI_TAB_SZ;
  const int AP = TKFMI_A_PAD;
  for (int q = tid; q < B.G * TKFMI_TAB_SZ; q += kThreads) {
    const int g = q / TKFMI_TAB_SZ, e = q % TKFMI_TAB_SZ;
    int a, bb;  // characters of the entry (-1: not used)
    if (e < AP * AP) { a = e / AP; bb = e % AP; }
    else if (e < AP * AP + AP) { a = -1; bb = e - AP * AP; }
    else { a = e - AP * AP - AP; bb = -1; }
    double v;
    if (B.emission == TKFMI_EMIT_PER_CLASS) {
      v = kNegInf;
      for (int c = 0; c < C; ++c) {
        const double w = B.cd[g * C + c];
        const double lw = (w > 0.0) ? log(fmax(w, 1e-300)) : kNegInf;
        const int ch = (a >= 0) ? a : bb;  // pi factor: x for match / delete, y for insert
        double term = (ch < A) ? log(max_nan(B.pi[c * A + ch], 1e-30)) : 0.0;
        if (a >= 0 && bb >= 0 && a < A && bb < A) term += log(max_nan(s_P[(c * A + a) * A + bb], 1e-30));
        v = lse2(v, term + lw);
      }
    } else {
      const int c = (B.emission == TKFMI_EMIT_PER_DOMAIN) ? g : 0;
      const bool shared = B.emission == TKFMI_EMIT_SHARED;
      auto fl = [&](double x) { return shared ? log(x + 1e-30) : log(max_nan(x, 1e-30)); };
      const int ch = (a >= 0) ? a : bb;
      v = (ch < A) ? fl(B.pi[c * A + ch]) : 0.0;
      if (a >= 0 && bb >= 0 && a < A && bb < A) v += fl(s_P[(c * A + a) * A + bb]);
    }
    tab[q] = v;
  }
}

// ------------------------------------------------------------- host side
// State numbering of the exploded machine (the order of the reference's
// name list, exact_suffstats_batch.py:107-159).
struct Layout {
  int K, F, N0, Nc;
  int start() const { return 0; }
  int end() const { return 1; }
  int dom(int p) const { return 2 + p; }                               // Mat, Ins, Del
  int domtype(int k, int p) const { return 5 + 3 * k + p; }
  int frag(int k, int q) const { return 5 + 3 * K + 5 * k + q; }       // Mat, Ins, Del, I, D
  int fragtype(int k, int f, int q) const { return 5 + 8 * K + (k * F + f) * 5 + q; }
  int emit(int k, int q, int f) const { return 5 + 8 * K + 5 * K * F + k * 5 * F + q * F + f; }
  int fragend(int k, int f, int q) const { return 5 + 8 * K + 10 * K * F + (k * F + f) * 5 + q; }
  int domend(int p) const { return 5 + 8 * K + 15 * K * F + p; }
};

Layout make_layout(int K, int F) {
  Layout L;
  L.K = K;
  L.F = F;
  L.N0 = 8 + 8 * K + 15 * K * F;
  L.Nc = 2 + 5 * K * F;
  return L;
}

struct Program {
  Layout L;
  std::vector<Entry> entries;
  std::vector<int32_t> idx;
  std::vector<int32_t> g_ptr, g_src;
  Step steps[6];
  int n_stats = 0, off_top = 0, off_domM = 0, off_kap = 0, off_kap1 = 0;
  int64_t slice = 0, nbuf_off = 0, gj_off = 0, sc_off = 0, ct_off = 0, cts_off = 0, sh_off = 0;
};

int stats_size(int K, int F) { return 25 + 28 * K + K * F * (F + 2); }

bool build_entries(const tkfmi_mixdom_model* m, const Layout& L, std::vector<Entry>& out,
                   std::string& err) {
  std::map<std::pair<int, int>, Entry> cells;
  auto put = [&](int r, int c, double coef, int sel) {
    Entry e;
    e.row = r;
    e.col = c;
    e.tau_sel = sel;
    e.pad = 0;
    e.coef = coef;
    cells[{r, c}] = e;
  };
  const int K = L.K, F = L.F;
  auto tsel = [](int level, int r, int c) { return level * 25 + r * 5 + c; };
  const int rowsS[4] = {S_, S_, S_, D_};
  const int srcs[4] = {L.start(), L.domend(0), L.domend(1), L.domend(2)};
  for (int q = 0; q < 4; ++q) {
    put(srcs[q], L.dom(0), 1.0, tsel(0, rowsS[q], M_));
    put(srcs[q], L.dom(1), 1.0, tsel(0, rowsS[q], I_));
    put(srcs[q], L.dom(2), 1.0, tsel(0, rowsS[q], D_));
    put(srcs[q], L.end(), 1.0, tsel(0, rowsS[q], E_));
  }
  for (int k = 0; k < K; ++k) {
    const double kap = m->dom_ins[k] / m->dom_del[k];
    for (int p = 0; p < 3; ++p) put(L.dom(p), L.domtype(k, p), m->dom_w[k], -1);
    put(L.domtype(k, 0), L.frag(k, 0), 1.0, tsel(1 + k, S_, M_));
    put(L.domtype(k, 0), L.frag(k, 1), 1.0, tsel(1 + k, S_, I_));
    put(L.domtype(k, 0), L.frag(k, 2), 1.0, tsel(1 + k, S_, D_));
    put(L.domtype(k, 0), L.domend(0), 1.0, tsel(1 + k, S_, E_));
    put(L.domtype(k, 1), L.frag(k, 3), kap, -1);
    put(L.domtype(k, 1), L.domend(1), 1.0 - kap, -1);
    put(L.domtype(k, 2), L.frag(k, 4), kap, -1);
    put(L.domtype(k, 2), L.domend(2), 1.0 - kap, -1);
    for (int f = 0; f < F; ++f) {
      for (int q = 0; q < 5; ++q) {
        put(L.frag(k, q), L.fragtype(k, f, q), m->frag_w[k * F + f], -1);
        put(L.fragtype(k, f, q), L.emit(k, q, f), 1.0, -1);
        put(L.emit(k, q, f), L.fragend(k, f, q), 1.0, -1);
      }
      double ext_sum = 0.0;
      for (int g = 0; g < F; ++g) ext_sum += m->ext[(k * F + f) * F + g];
      const double notext = 1.0 - ext_sum;
      for (int g = 0; g < F; ++g) {
        const double e = m->ext[(k * F + f) * F + g];
        for (int q = 0; q < 5; ++q) put(L.fragend(k, f, q), L.fragtype(k, g, q), e, -1);
      }
      const int rows[3] = {M_, I_, D_};
      for (int q = 0; q < 3; ++q) {
        put(L.fragend(k, f, q), L.frag(k, 0), notext, tsel(1 + k, rows[q], M_));
        put(L.fragend(k, f, q), L.frag(k, 1), notext, tsel(1 + k, rows[q], I_));
        put(L.fragend(k, f, q), L.frag(k, 2), notext, tsel(1 + k, rows[q], D_));
        put(L.fragend(k, f, q), L.domend(0), notext, tsel(1 + k, rows[q], E_));
      }
      put(L.fragend(k, f, 3), L.frag(k, 3), notext * kap, -1);
      put(L.fragend(k, f, 3), L.domend(1), notext * (1.0 - kap), -1);
      put(L.fragend(k, f, 4), L.frag(k, 4), notext * kap, -1);
      put(L.fragend(k, f, 4), L.domend(2), notext * (1.0 - kap), -1);
    }
  }
  out.clear();
  for (auto& kv : cells) out.push_back(kv.second);
  (void)err;
  return true;
}

// Elimination groups in chain order (exact_suffstats_batch.py:468-482).
std::vector<std::vector<int>> elim_groups(const Layout& L) {
  std::vector<std::vector<int>> g(6);
  for (int k = 0; k < L.K; ++k) {
    for (int f = 0; f < L.F; ++f)
      for (int q = 0; q < 5; ++q) {
        g[0].push_back(L.fragtype(k, f, q));
        g[1].push_back(L.fragend(k, f, q));
      }
    for (int q = 0; q < 5; ++q) g[2].push_back(L.frag(k, q));
    for (int p = 0; p < 3; ++p) g[3].push_back(L.domtype(k, p));
  }
  for (int p = 0; p < 3; ++p) {
    g[4].push_back(L.dom(p));
    g[5].push_back(L.domend(p));
  }
  return g;
}

void build_gathers(const Layout& L, Program& pr) {
  const int K = L.K, F = L.F, N = L.N0;
  pr.n_stats = stats_size(K, F);
  pr.off_top = 0;
  pr.off_domM = 25;
  pr.off_kap = 25 + 25 * K;
  pr.off_kap1 = pr.off_kap + K;
  const int off_ext = pr.off_kap1 + K;
  const int off_term = off_ext + K * F * F;
  const int off_dw = off_term + K * F;
  const int off_fw = off_dw + K;
  std::vector<std::vector<int32_t>> src(pr.n_stats);
  auto add = [&](int o, int u, int v) { src[o].push_back(u * N + v); };
  const int dcol[3] = {M_, I_, D_};
  for (int p = 0; p < 3; ++p) add(pr.off_top + S_ * 5 + dcol[p], L.start(), L.dom(p));
  add(pr.off_top + S_ * 5 + E_, L.start(), L.end());
  for (int pu = 0; pu < 3; ++pu) {
    for (int p = 0; p < 3; ++p) add(pr.off_top + dcol[pu] * 5 + dcol[p], L.domend(pu), L.dom(p));
    add(pr.off_top + dcol[pu] * 5 + E_, L.domend(pu), L.end());
  }
  for (int k = 0; k < K; ++k) {
    const int o = pr.off_domM + 25 * k;
    const int dests[4] = {L.frag(k, 0), L.frag(k, 1), L.frag(k, 2), L.domend(0)};
    const int cols[4] = {M_, I_, D_, E_};
    for (int c = 0; c < 4; ++c) add(o + S_ * 5 + cols[c], L.domtype(k, 0), dests[c]);
    for (int f = 0; f < F; ++f)
      for (int q = 0; q < 3; ++q)
        for (int c = 0; c < 4; ++c) add(o + dcol[q] * 5 + cols[c], L.fragend(k, f, q), dests[c]);
    add(pr.off_kap + k, L.domtype(k, 1), L.frag(k, 3));
    add(pr.off_kap + k, L.domtype(k, 2), L.frag(k, 4));
    add(pr.off_kap1 + k, L.domtype(k, 1), L.domend(1));
    add(pr.off_kap1 + k, L.domtype(k, 2), L.domend(2));
    for (int f = 0; f < F; ++f) {
      add(pr.off_kap + k, L.fragend(k, f, 3), L.frag(k, 3));
      add(pr.off_kap + k, L.fragend(k, f, 4), L.frag(k, 4));
      add(pr.off_kap1 + k, L.fragend(k, f, 3), L.domend(1));
      add(pr.off_kap1 + k, L.fragend(k, f, 4), L.domend(2));
      for (int q = 0; q < 5; ++q) {
        const int fe = L.fragend(k, f, q);
        for (int g = 0; g < F; ++g) add(off_ext + (k * F + f) * F + g, fe, L.fragtype(k, g, q));
        for (int v = 0; v < N; ++v) {
          bool is_ft = false;
          for (int g = 0; g < F; ++g) is_ft |= (v == L.fragtype(k, g, q));
          if (!is_ft) add(off_term + k * F + f, fe, v);
        }
        add(off_fw + k * F + f, L.frag(k, q), L.fragtype(k, f, q));
      }
    }
    for (int p = 0; p < 3; ++p)
      for (int u = 0; u < N; ++u) add(off_dw + k, u, L.domtype(k, p));
  }
  pr.g_ptr.assign(pr.n_stats + 1, 0);
  pr.g_src.clear();
  for (int o = 0; o < pr.n_stats; ++o) {
    pr.g_ptr[o] = (int32_t)pr.g_src.size();
    pr.g_src.insert(pr.g_src.end(), src[o].begin(), src[o].end());
  }
  pr.g_ptr[pr.n_stats] = (int32_t)pr.g_src.size();
}

int build_program(const tkfmi_mixdom_model* m, Program& pr, std::string& err) {
  if (!m) { err = "null model"; return TKFMI_EINVAL; }
  if (m->K < 1 || m->K > kMaxK || m->F < 1 || m->F > 64) { err = "K must be in 1..16 and F in 1..64"; return TKFMI_EINVAL; }
  if (!m->dom_ins || !m->dom_del || !m->dom_w || !m->frag_w || !m->ext) { err = "model array missing"; return TKFMI_EINVAL; }
  pr.L = make_layout(m->K, m->F);
  const Layout& L = pr.L;
  build_entries(m, L, pr.entries, err);
  // steps: current state list, keep / elim positions, offsets
  std::vector<int> cur(L.N0);
  for (int i = 0; i < L.N0; ++i) cur[i] = i;
  const auto groups = elim_groups(L);
  int64_t off = 0;
  std::vector<int64_t> T_offs;
  int nkmax = 0, nzmax = 0;
  pr.idx.clear();
  for (int s = 0; s < 6; ++s) {
    std::vector<char> in_g(L.N0, 0);
    for (int v : groups[s]) in_g[v] = 1;
    std::vector<int> keep, elim, pos(cur.size());
    for (int i = 0; i < (int)cur.size(); ++i) {
      if (in_g[cur[i]]) { pos[i] = -(int)elim.size() - 1; elim.push_back(i); }
      else { pos[i] = (int)keep.size(); keep.push_back(i); }
    }
    Step& st = pr.steps[s];
    st.n = (int)cur.size();
    st.nk = (int)keep.size();
    st.nz = (int)elim.size();
    st.pad = st.pad2 = 0;
    st.keep_off = (int32_t)pr.idx.size();
    pr.idx.insert(pr.idx.end(), keep.begin(), keep.end());
    st.elim_off = (int32_t)pr.idx.size();
    pr.idx.insert(pr.idx.end(), elim.begin(), elim.end());
    st.pos_off = (int32_t)pr.idx.size();
    pr.idx.insert(pr.idx.end(), pos.begin(), pos.end());
    nkmax = std::max(nkmax, st.nk);
    nzmax = std::max(nzmax, st.nz);
    std::vector<int> nc;
    for (int i : keep) nc.push_back(cur[i]);
    cur = nc;
  }
  if ((int)cur.size() != L.Nc) { err = "internal: elimination left the wrong state count"; return TKFMI_EINVAL; }
  if (nzmax > 1024) { err = "null group too large"; return TKFMI_EINVAL; }
  // slice layout: T_0 T_1 ... T_6 contiguous (T_{s+1} right after T_s), then C, W, scratch
  for (int s = 0; s < 6; ++s) {
    pr.steps[s].T_off = off;
    off += (int64_t)pr.steps[s].n * pr.steps[s].n;
  }
  off += (int64_t)L.Nc * L.Nc;  // T_6
  for (int s = 0; s < 6; ++s) {
    pr.steps[s].C_off = off;
    off += (int64_t)pr.steps[s].nz * pr.steps[s].nz;
    pr.steps[s].W_off = off;
    off += (int64_t)pr.steps[s].nz * pr.steps[s].nk;
  }
  pr.nbuf_off = off;
  off += 2 * (int64_t)L.N0 * L.N0;
  pr.gj_off = off;
  off += 2 * (int64_t)nzmax * nzmax;
  pr.sc_off = off;
  off += (int64_t)nkmax * nkmax;
  pr.ct_off = off;
  off += (int64_t)nkmax * nzmax;
  pr.cts_off = off;
  off += (int64_t)nkmax * nzmax;
  pr.sh_off = off;
  off += (int64_t)nkmax * nzmax;
  pr.slice = (off + 31) / 32 * 32;
  build_gathers(L, pr);
  return TKFMI_OK;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

size_t program_bytes(const Program& pr) {
  return align256(pr.entries.size() * sizeof(Entry)) + align256(pr.idx.size() * 4) +
         align256(pr.g_ptr.size() * 4) + align256(pr.g_src.size() * 4);
}

}  // namespace tkfmi_ss

using namespace tkfmi_ss;

extern "C" {

int32_t tkfmi_suffstats_size(int32_t K, int32_t F) {
  if (K < 1 || F < 1) return 0;
  return stats_size(K, F);
}

size_t tkfmi_restore_workspace_bytes(const tkfmi_mixdom_model* m, int32_t B) {
  Program pr;
  std::string err;
  if (build_program(m, pr, err) != TKFMI_OK || B < 0) return 0;
  return program_bytes(pr) + (size_t)std::max(B, 1) * (size_t)pr.slice * 8;
}

int tkfmi_restore_counts_batched(const tkfmi_mixdom_model* m, int32_t B, const double* t,
                                 const double* n_chi, double* stats, double* bdi, void* workspace,
                                 size_t workspace_bytes, void* stream_v) {
  tkfmi_clear_error();
  Program pr;
  std::string err;
  int rc = build_program(m, pr, err);
  if (rc != TKFMI_OK) return tkfmi_fail(rc, err);
  if (B < 0) return tkfmi_fail(TKFMI_EINVAL, "B < 0");
  if (B == 0) return TKFMI_OK;
  if (!t || !n_chi || !stats) return tkfmi_fail(TKFMI_EINVAL, "device pointer missing");
  const size_t need = program_bytes(pr) + (size_t)B * (size_t)pr.slice * 8;
  if (!workspace || workspace_bytes < need)
    return tkfmi_fail(TKFMI_EWORKSPACE, "workspace too small: need " + std::to_string(need));
  for (int k = 0; k < m->K; ++k)
    if (!(m->dom_del[k] > 0.0)) return tkfmi_fail(TKFMI_EINVAL, "dom_del must be positive");
  if (!(m->main_del > 0.0)) return tkfmi_fail(TKFMI_EINVAL, "main_del must be positive");
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_v);
  // program upload: one host blob, one copy
  std::vector<char> blob(program_bytes(pr), 0);
  size_t o = 0;
  const size_t o_ent = o;
  memcpy(blob.data() + o, pr.entries.data(), pr.entries.size() * sizeof(Entry));
  o += align256(pr.entries.size() * sizeof(Entry));
  const size_t o_idx = o;
  memcpy(blob.data() + o, pr.idx.data(), pr.idx.size() * 4);
  o += align256(pr.idx.size() * 4);
  const size_t o_gp = o;
  memcpy(blob.data() + o, pr.g_ptr.data(), pr.g_ptr.size() * 4);
  o += align256(pr.g_ptr.size() * 4);
  const size_t o_gs = o;
  memcpy(blob.data() + o, pr.g_src.data(), pr.g_src.size() * 4);
  o += align256(pr.g_src.size() * 4);
  char* wsb = static_cast<char*>(workspace);
  rc = tkfmi_stage_upload(wsb, blob.data(), blob.size(), stream, err);
  if (rc != TKFMI_OK) return tkfmi_fail(rc, err);
  RParams P;
  memset(&P, 0, sizeof(P));
  P.K = m->K;
  P.F = m->F;
  P.N0 = pr.L.N0;
  P.Nc = pr.L.Nc;
  P.n_entries = (int32_t)pr.entries.size();
  P.n_stats = pr.n_stats;
  P.want_bdi = bdi != nullptr;
  P.lam[0] = m->main_ins;
  P.mu[0] = m->main_del;
  for (int k = 0; k < m->K; ++k) {
    P.lam[1 + k] = m->dom_ins[k];
    P.mu[1 + k] = m->dom_del[k];
  }
  for (int s = 0; s < 6; ++s) P.steps[s] = pr.steps[s];
  P.nbuf_off = pr.nbuf_off;
  P.gj_off = pr.gj_off;
  P.sc_off = pr.sc_off;
  P.ct_off = pr.ct_off;
  P.cts_off = pr.cts_off;
  P.sh_off = pr.sh_off;
  P.slice = pr.slice;
  P.off_top = pr.off_top;
  P.off_domM = pr.off_domM;
  P.off_kap = pr.off_kap;
  P.off_kap1 = pr.off_kap1;
  P.entries = reinterpret_cast<const Entry*>(wsb + o_ent);
  P.idx = reinterpret_cast<const int32_t*>(wsb + o_idx);
  P.g_ptr = reinterpret_cast<const int32_t*>(wsb + o_gp);
  P.g_src = reinterpret_cast<const int32_t*>(wsb + o_gs);
  P.t = t;
  P.n_chi = n_chi;
  P.stats = stats;
  P.bdi = bdi;
  P.ws = reinterpret_cast<double*>(wsb + program_bytes(pr));
  const int slot = tkfmi_ktime_start("restore_kernel", B, 0, stream, err);
  if (slot == -2) return tkfmi_fail(TKFMI_EHIP, err);
  hipLaunchKernelGGL(restore_kernel, dim3(B), dim3(kThreads), 0, stream, P);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("restore launch: ") + hipGetErrorString(e));
  if (tkfmi_ktime_stop(slot, stream, err) != TKFMI_OK) return tkfmi_fail(TKFMI_EHIP, err);
  return TKFMI_OK;
}

int tkfmi_bdi_stats_batched(int32_t B, double lam, double mu, const double* n5, const double* t,
                            const double* T, double* out, void* stream_v) {
  tkfmi_clear_error();
  if (B < 0) return tkfmi_fail(TKFMI_EINVAL, "B < 0");
  if (B == 0) return TKFMI_OK;
  if (!n5 || !t || !T || !out) return tkfmi_fail(TKFMI_EINVAL, "device pointer missing");
  if (!(mu > 0.0) || !(lam > 0.0)) return tkfmi_fail(TKFMI_EINVAL, "rates must be positive");
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_v);
  hipLaunchKernelGGL(bdi_kernel, dim3((B + 63) / 64), dim3(64), 0, stream, B, lam, mu, n5, t, T, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("bdi launch: ") + hipGetErrorString(e));
  return TKFMI_OK;
}

int tkfmi_tkf92_estep_stats(int32_t B, double lam, double mu, double ext, const double* t,
                            const double* n_chi, double* out, void* stream_v) {
  tkfmi_clear_error();
  if (B < 0) return tkfmi_fail(TKFMI_EINVAL, "B < 0");
  if (B == 0) return TKFMI_OK;
  if (!t || !n_chi || !out) return tkfmi_fail(TKFMI_EINVAL, "device pointer missing");
  if (!(mu > 0.0) || !(lam > 0.0)) return tkfmi_fail(TKFMI_EINVAL, "rates must be positive");
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_v);
  hipLaunchKernelGGL(tkf92_estep_kernel, dim3((B + 63) / 64), dim3(64), 0, stream, B, lam, mu, ext,
                     t, n_chi, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("tkf92 estep launch: ") + hipGetErrorString(e));
  return TKFMI_OK;
}

int tkfmi_hr_stats_batched(int32_t B, int32_t C, int32_t A, const double* eigvec,
                           const double* eigval, const double* pi, const double* Q, const double* t,
                           const double* mc, double* W, double* U, void* stream_v) {
  tkfmi_clear_error();
  if (B < 0 || C < 1 || A < 1 || A > kHrMaxA) return tkfmi_fail(TKFMI_EINVAL, "need B >= 0, C >= 1, 1 <= A <= 32");
  if (B == 0) return TKFMI_OK;
  if (!eigvec || !eigval || !pi || !Q || !t || !mc || !W || !U)
    return tkfmi_fail(TKFMI_EINVAL, "device pointer missing");
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_v);
  hipLaunchKernelGGL(hr_kernel, dim3((unsigned)B * (unsigned)C), dim3(256), 0, stream, C, A, eigvec,
                     eigval, pi, Q, t, mc, W, U);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("hr launch: ") + hipGetErrorString(e));
  return TKFMI_OK;
}

size_t tkfmi_build_workspace_bytes(const tkfmi_builder* bl, int32_t B) {
  if (!bl || B < 0) return 0;
  if (bl->model != TKFMI_MODEL_MIXDOM) return 256;
  Program pr;
  std::string err;
  if (build_program(bl->mix, pr, err) != TKFMI_OK) return 0;
  return program_bytes(pr) + (size_t)std::max(B, 1) * (size_t)pr.slice * 8;
}

int tkfmi_build_params_batched(const tkfmi_builder* bl, int32_t B, const double* t,
                               double* log_trans, double* log_tab, void* workspace,
                               size_t workspace_bytes, void* stream_v) {
  tkfmi_clear_error();
  if (!bl) return tkfmi_fail(TKFMI_EINVAL, "null builder");
  if (B < 0) return tkfmi_fail(TKFMI_EINVAL, "B < 0");
  if (bl->model < TKFMI_MODEL_TKF91 || bl->model > TKFMI_MODEL_MIXDOM)
    return tkfmi_fail(TKFMI_EINVAL, "unknown model");
  if (bl->emission < TKFMI_EMIT_SHARED || bl->emission > TKFMI_EMIT_PER_CLASS)
    return tkfmi_fail(TKFMI_EINVAL, "unknown emission mode");
  if (bl->A < 1 || bl->A > kMaxA || bl->C < 1 || bl->C > kMaxSub)
    return tkfmi_fail(TKFMI_EINVAL, "need 1 <= A <= 20 and 1 <= C <= 16");
  if (!(bl->mu > 0.0) && bl->model != TKFMI_MODEL_MIXDOM) return tkfmi_fail(TKFMI_EINVAL, "mu must be positive");
  if (B == 0) return TKFMI_OK;
  if (!t || !log_trans || !log_tab || !bl->Q || !bl->pi)
    return tkfmi_fail(TKFMI_EINVAL, "device pointer missing");
  BParams P;
  memset(&P, 0, sizeof(P));
  std::string err;
  int K = 1, F = 1;
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_v);
  char* wsb = static_cast<char*>(workspace);
  if (bl->model == TKFMI_MODEL_MIXDOM) {
    Program pr;
    int rc = build_program(bl->mix, pr, err);
    if (rc != TKFMI_OK) return tkfmi_fail(rc, err);
    const size_t need = program_bytes(pr) + (size_t)B * (size_t)pr.slice * 8;
    if (!workspace || workspace_bytes < need)
      return tkfmi_fail(TKFMI_EWORKSPACE, "workspace too small: need " + std::to_string(need));
    std::vector<char> blob(program_bytes(pr), 0);
    size_t o = 0;
    const size_t o_ent = o;
    memcpy(blob.data() + o, pr.entries.data(), pr.entries.size() * sizeof(Entry));
    o += align256(pr.entries.size() * sizeof(Entry));
    const size_t o_idx = o;
    memcpy(blob.data() + o, pr.idx.data(), pr.idx.size() * 4);
    rc = tkfmi_stage_upload(wsb, blob.data(), blob.size(), stream, err);
    if (rc != TKFMI_OK) return tkfmi_fail(rc, err);
    K = bl->mix->K;
    F = bl->mix->F;
    P.R.K = K;
    P.R.F = F;
    P.R.N0 = pr.L.N0;
    P.R.Nc = pr.L.Nc;
    P.R.n_entries = (int32_t)pr.entries.size();
    P.R.lam[0] = bl->mix->main_ins;
    P.R.mu[0] = bl->mix->main_del;
    for (int k = 0; k < K; ++k) {
      P.R.lam[1 + k] = bl->mix->dom_ins[k];
      P.R.mu[1 + k] = bl->mix->dom_del[k];
    }
    for (int s = 0; s < 6; ++s) P.R.steps[s] = pr.steps[s];
    P.R.gj_off = pr.gj_off;
    P.R.slice = pr.slice;
    P.R.entries = reinterpret_cast<const Entry*>(wsb + o_ent);
    P.R.idx = reinterpret_cast<const int32_t*>(wsb + o_idx);
    P.R.ws = reinterpret_cast<double*>(wsb + program_bytes(pr));
    P.S = pr.L.Nc;
  } else {
    P.S = 5;
  }
  if (bl->emission == TKFMI_EMIT_PER_CLASS) {
    if (!bl->class_dist) return tkfmi_fail(TKFMI_EINVAL, "class_dist missing");
    P.G = K * F;
  } else if (bl->emission == TKFMI_EMIT_PER_DOMAIN) {
    if (bl->C != K) return tkfmi_fail(TKFMI_EINVAL, "per-domain emissions need C == K");
    P.G = K;
  } else {
    if (bl->C != 1) return tkfmi_fail(TKFMI_EINVAL, "shared emissions need C == 1");
    P.G = 1;
  }
  P.R.t = t;
  P.model = bl->model;
  P.emission = bl->emission;
  P.log_floor = bl->log_floor;
  P.C = bl->C;
  P.A = bl->A;
  P.F = F;
  P.lam = bl->lam;
  P.mu = bl->mu;
  P.ext = bl->ext;
  P.Q = bl->Q;
  P.pi = bl->pi;
  P.cd = bl->class_dist;
  P.log_trans = log_trans;
  P.log_tab = log_tab;
  hipLaunchKernelGGL(build_kernel, dim3(B), dim3(kThreads), 0, stream, P);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("build launch: ") + hipGetErrorString(e));
  return TKFMI_OK;
}

}  // extern "C"
