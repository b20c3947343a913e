// tkfmi: 1D alignment-constrained forward/backward + expected counts on
// gfx950 (the MSA-guided DP).
// Reference algorithm: python/tkfmixdom/jax/dp/hmm.py forward_1d_scan
// (:404-437), backward_1d_scan (:440-468), forward_backward_1d_padded
// (:471-607), and the constrained emission builders (:854-1100).
//
// Design (one workgroup per pair, one thread per state):
//  * emit: a grid of workgroups turns the dense log emissions, or the column
//    types + characters + per-group tables, into probability-space emissions
//    E[row][state] in HBM.
//  * sweep: the forward recursion a_t = (a_{t-1} T) .* e_t runs with the
//    previous step's vector broadcast from LDS and T in LDS (S <= 128; read
//    from HBM with coalesced rows above that), rescaled every step by a
//    common power of two (the binary exponent of the step's maximum).  The
//    backward recursion b_t = T_bwd (e_{t+1} .* b_{t+1}) mirrors it.  The
//    forward lattice and its exponents go to HBM; the backward writes the
//    posteriors, the count operands a_t * scale_t (per step) and
//    u_{t+1} = e_{t+1} .* b_{t+1}.
//  * counts: the interior transition counts are T .* (sum_t a_t u_{t+1}^T),
//    a per-pair (S x L) (L x S) product on the matrix cores
//    (v_mfma_f64_16x16x4f64), and the emission counts a per-pair reduction of
//    the posteriors into LDS tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "fb2d_internal.h"
#include "tkfmi.h"

namespace tkfmi_1d {

constexpr int ST_M = 1, ST_I = 2, ST_D = 3;
constexpr int A_PAD = TKFMI_A_PAD;
constexpr int TAB_SZ = TKFMI_TAB_SZ;
constexpr double LN2 = 0.69314718055994530942;
constexpr int kMaxS = 256;
constexpr int kLdsT = 128;   // largest S whose transition matrix is staged in LDS
constexpr int kPF = 8;       // steps of emissions / lattice prefetched ahead

typedef double d4 __attribute__((ext_vector_type(4)));

struct F1Desc {
  int64_t row;      // first row of this pair in the workspace lattices / posteriors
  int64_t src_row;  // first row in the caller's row-indexed inputs
  int32_t L, pair, tr, tb;
};

struct F1Params {
  int32_t S, LD, s_idx, e_idx;
  int32_t flags, G, mode, pad;
  const F1Desc* desc;
  const int32_t* st;       // [S] state types (constrained mode)
  const int32_t* sg;       // [S] state groups (constrained mode)
  const double* tmat;      // [n_trans][2][S][S]: T, then TT[k][s] = T_bwd[s][k]
  const double* log_emit;  // mode 0: caller's [rows][S]
  const double* log_tab;   // mode 1: [n_tab][G][TAB_SZ]
  const uint8_t* col;      // mode 1: [rows]
  const uint8_t* anc;
  const uint8_t* dsc;
  double* em;              // [rows][S] emissions, then u_{t} for t >= 1
  double* al;              // [rows][S] forward lattice (normalised)
  double* as;              // [rows][S] count operand a_t * scale_t
  double* ea;              // [rows] forward exponents
  double* pinfo;           // [B][2] mantissa of P (0 if impossible), exponent
  double* post;            // [rows][S] posteriors (caller's or workspace), or null
  double* out_logP;
  double* out_n;
  double* out_ecnt;
};

__device__ inline int frexp_exp(double v) { return __builtin_amdgcn_frexp_exp(v); }

__device__ inline double wave_max(double v) {
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double wave_sum(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block-wide max / sum.  With more than one wave the per-wave results pass
// through one of two LDS slot rows (alternating), so a fast wave writing the
// next reduction never overwrites a row a slow wave is still reading.
template <int NW>
__device__ inline double block_max(double v, double* red, int& par) {
  v = wave_max(v);
  if constexpr (NW == 1) {
    return v;
  } else {
    double* r = red + par * NW;
    par ^= 1;
    if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = r[0];
#pragma unroll
    for (int q = 1; q < NW; ++q) m = fmax(m, r[q]);
    return m;
  }
}
template <int NW>
__device__ inline double block_sum(double v, double* red, int& par) {
  v = wave_sum(v);
  if constexpr (NW == 1) {
    return v;
  } else {
    double* r = red + par * NW;
    par ^= 1;
    if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = r[0];
#pragma unroll
    for (int q = 1; q < NW; ++q) m += r[q];
    return m;
  }
}

// Probability-space transition blocks: T and the transposed backward matrix
// (T with T[final,final] = 1, dp/hmm.py:532).
__global__ void fb1d_prep_kernel(const double* __restrict__ log_trans, int S, int e_idx,
                                 double* __restrict__ tmat) {
  const double* lt = log_trans + (size_t)blockIdx.x * S * S;
  double* T = tmat + (size_t)blockIdx.x * 2 * S * S;
  double* TT = T + (size_t)S * S;
  for (int q = threadIdx.x; q < S * S; q += blockDim.x) {
    const int s = q / S, k = q - s * S;
    const double v = exp(lt[q]);
    T[q] = v;
    TT[(size_t)k * S + s] = (s == e_idx && k == e_idx) ? 1.0 : v;
  }
}

// Probability-space emissions of one pair (one workgroup per pair).
__global__ void fb1d_emit_kernel(F1Params P) {
  const F1Desc d = P.desc[blockIdx.x];
  const int S = P.S;
  const int64_t n = (int64_t)d.L * S;
  for (int64_t q = threadIdx.x; q < n; q += blockDim.x) {
    const int64_t t = q / S;
    const int k = (int)(q - t * S);
    double v;
    if (P.mode == 0) {
      v = exp(P.log_emit[(d.src_row + t) * S + k]);
    } else {
      const int ct = P.col[d.src_row + t];
      if (P.st[k] != ct || ct < ST_M || ct > ST_D) {
        v = 0.0;
      } else {
        const int a = min((int)P.anc[d.src_row + t], A_PAD - 1);
        const int b = min((int)P.dsc[d.src_row + t], A_PAD - 1);
        const int e = ct == ST_M ? a * A_PAD + b : (ct == ST_I ? A_PAD * A_PAD + b : A_PAD * A_PAD + A_PAD + a);
        v = exp(P.log_tab[((size_t)d.tb * P.G + P.sg[k]) * TAB_SZ + e]);
      }
    }
    P.em[(d.row + t) * S + k] = v;
  }
}

// Forward and backward sweeps of one pair.
template <int NW, bool TL>
__global__ __launch_bounds__(NW * 64) void fb1d_sweep_kernel(F1Params P) {
  extern __shared__ double sm[];
  const int tid = threadIdx.x;
  const F1Desc d = P.desc[blockIdx.x];
  const int S = P.S, LD = P.LD, L = d.L;
  const int SP = (S + 1) & ~1;
  const bool act = tid < S;
  const int k = act ? tid : 0;
  const double* T = P.tmat + (size_t)d.tr * 2 * S * S;
  const double* TT = T + (size_t)S * S;
  double* sT = sm;                                  // TL: [S][LD]
  double* sv = sm + (TL ? (size_t)S * LD : 0);      // [2][SP] broadcast vectors
  double* red = sv + 2 * SP;                        // [2][NW] reduction slots
  if constexpr (TL) {
    for (int q = tid; q < S * S; q += NW * 64) {
      const int s = q / S;
      sT[s * LD + (q - s * S)] = T[q];
    }
    __syncthreads();
  }
  // T[s][c] (forward: threads read a column, coalesced) and T_bwd[r][c]
  // (backward: threads read a row; from TT when T lives in HBM).
  auto tf = [&](int s, int c) -> double { return TL ? sT[s * LD + c] : T[(size_t)s * S + c]; };
  auto tb = [&](int r, int c) -> double {
    if constexpr (TL) return (r == P.e_idx && c == P.e_idx) ? 1.0 : sT[r * LD + c];
    else return TT[(size_t)c * S + r];
  };
  const bool want_lat = !(P.flags & TKFMI_FORWARD_ONLY);
  const bool want_counts = want_lat && !(P.flags & TKFMI_NO_COUNTS) && P.out_n != nullptr;
  double* out_n = want_counts ? P.out_n + (size_t)d.pair * S * S : nullptr;
  const double* em = P.em + d.row * S;
  double* al = P.al + d.row * S;
  double* ea = P.ea + d.row;
  int par = 0;

  // ======================================================== forward sweep
  int EA = 0;
  double a = 0.0;
  double ecur[kPF];
#pragma unroll
  for (int u = 0; u < kPF; ++u) ecur[u] = (act && u < L) ? em[(size_t)u * S + k] : 0.0;
  for (int t0 = 0; t0 < L; t0 += kPF) {
    double enx[kPF];
#pragma unroll
    for (int u = 0; u < kPF; ++u) {
      const int tn = t0 + kPF + u;
      enx[u] = (act && tn < L) ? em[(size_t)tn * S + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPF; ++u) {
      const int t = t0 + u;
      if (t >= L) break;
      double h;
      if (t == 0) {
        h = tf(P.s_idx, k);
      } else {
        const double* vb = sv + ((t - 1) & 1) * SP;
        double h0 = 0.0, h1 = 0.0, h2 = 0.0, h3 = 0.0;
        int s = 0;
        for (; s + 3 < S; s += 4) {
          h0 = fma(vb[s], tf(s, k), h0);
          h1 = fma(vb[s + 1], tf(s + 1, k), h1);
          h2 = fma(vb[s + 2], tf(s + 2, k), h2);
          h3 = fma(vb[s + 3], tf(s + 3, k), h3);
        }
        for (; s < S; ++s) h0 = fma(vb[s], tf(s, k), h0);
        h = (h0 + h1) + (h2 + h3);
      }
      double v = act ? h * ecur[u] : 0.0;
      const double m = block_max<NW>(v, red, par);
      const int sh = (m > 0.0) ? frexp_exp(m) : 0;
      v = ldexp(v, -sh);
      EA += sh;
      if (act) {
        sv[(t & 1) * SP + k] = v;
        if (want_lat) al[(size_t)t * S + k] = v;
      }
      if (tid == 0 && want_lat) ea[t] = (double)EA;
      a = v;
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kPF; ++u) ecur[u] = enx[u];
  }
  // log P and the terminal counts (dp/hmm.py:579-590)
  double tot;
  if (L == 0) {
    tot = tf(P.s_idx, P.e_idx);
    EA = 0;
  } else {
    tot = block_sum<NW>(act ? a * tf(k, P.e_idx) : 0.0, red, par);
  }
  const int etot = (tot > 0.0) ? frexp_exp(tot) : 0;
  const double mP = (tot > 0.0) ? ldexp(tot, -etot) : 0.0;
  const int EP = EA + etot;
  if (tid == 0) {
    P.out_logP[d.pair] = (tot > 0.0) ? log(tot) + (double)EA * LN2 : -1e30;
    if (want_lat) {
      P.pinfo[2 * (size_t)d.pair] = mP;
      P.pinfo[2 * (size_t)d.pair + 1] = (double)EP;
    }
  }
  if (want_counts && tot > 0.0) {
    if (L == 0) {
      if (tid == 0) out_n[P.s_idx * S + P.e_idx] += 1.0;
    } else if (act) {
      out_n[k * S + P.e_idx] += a * tf(k, P.e_idx) / tot;
    }
  }
  if (!want_lat || L == 0 || !(tot > 0.0)) return;

  // ======================================================= backward sweep
  const double inv_mP = 1.0 / mP;
  double* as = P.as + d.row * S;
  double* emw = P.em + d.row * S;
  double* post = P.post ? P.post + d.row * S : nullptr;
  // b_{L-1}[k] = T_bwd[k][final]
  double b = act ? tb(k, P.e_idx) : 0.0;
  int EB;
  {
    const double m = block_max<NW>(b, red, par);
    const int sh = (m > 0.0) ? frexp_exp(m) : 0;
    b = ldexp(b, -sh);
    EB = sh;
  }
  // per step t: emission e_t, forward value a_t (and a_{t-1}), exponents
  double pe[kPF], pa[kPF], px[kPF];
#pragma unroll
  for (int u = 0; u < kPF; ++u) {
    const int t = L - 1 - u;
    pe[u] = (act && t >= 0) ? em[(size_t)t * S + k] : 0.0;
    pa[u] = (act && t >= 0) ? al[(size_t)t * S + k] : 0.0;
    px[u] = (t >= 0) ? ea[t] : 0.0;
  }
  for (int t0 = L - 1; t0 >= 0; t0 -= kPF) {
    double ne[kPF], na[kPF], nx[kPF];  // steps t0-kPF .. t0-2kPF+1
#pragma unroll
    for (int u = 0; u < kPF; ++u) {
      const int t = t0 - kPF - u;
      ne[u] = (act && t >= 0) ? em[(size_t)t * S + k] : 0.0;
      na[u] = (act && t >= 0) ? al[(size_t)t * S + k] : 0.0;
      nx[u] = (t >= 0) ? ea[t] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPF; ++u) {
      const int t = t0 - u;
      if (t < 0) break;
      const int EAt = (int)px[u];
      if (post && act) post[(size_t)t * S + k] = pa[u] * b * ldexp(inv_mP, EAt + EB - EP);
      const double uu = pe[u] * b;  // u_t = e_t .* b_t
      if (t >= 1) {
        // a_{t-1}: the next entry of this chunk, or the first of the next one
        const double a_prev = (u + 1 < kPF) ? pa[u + 1] : na[0];
        const int EAp = (int)((u + 1 < kPF) ? px[u + 1] : nx[0]);
        if (act) {
          sv[(t & 1) * SP + k] = uu;
          emw[(size_t)t * S + k] = uu;
          as[(size_t)(t - 1) * S + k] = a_prev * ldexp(inv_mP, EAp + EB - EP);
        }
        __syncthreads();
        const double* ub = sv + (t & 1) * SP;
        double h0 = 0.0, h1 = 0.0, h2 = 0.0, h3 = 0.0;
        int j = 0;
        for (; j + 3 < S; j += 4) {
          h0 = fma(tb(k, j), ub[j], h0);
          h1 = fma(tb(k, j + 1), ub[j + 1], h1);
          h2 = fma(tb(k, j + 2), ub[j + 2], h2);
          h3 = fma(tb(k, j + 3), ub[j + 3], h3);
        }
        for (; j < S; ++j) h0 = fma(tb(k, j), ub[j], h0);
        const double h = act ? (h0 + h1) + (h2 + h3) : 0.0;
        const double m = block_max<NW>(h, red, par);
        const int sh = (m > 0.0) ? frexp_exp(m) : 0;
        b = ldexp(h, -sh);
        EB += sh;
      } else if (want_counts && act) {
        // initial counts (dp/hmm.py:592-596)
        out_n[P.s_idx * S + k] += tf(P.s_idx, k) * uu * ldexp(inv_mP, EB - EP);
      }
    }
#pragma unroll
    for (int u = 0; u < kPF; ++u) {
      pe[u] = ne[u];
      pa[u] = na[u];
      px[u] = nx[u];
    }
  }
}

// Interior transition counts n += T .* sum_{t=0}^{L-2} as_t u_{t+1}^T
// (16x16 output tiles on the matrix cores, K = steps) and the emission
// counts (constrained mode) from the posteriors.  One workgroup per pair.
__global__ __launch_bounds__(256) void fb1d_counts_kernel(F1Params P) {
  extern __shared__ double s_ec[];
  const F1Desc d = P.desc[blockIdx.x];
  const int S = P.S, L = d.L;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (!(P.pinfo[2 * (size_t)d.pair] > 0.0)) return;  // impossible pair or L == 0
  const bool want_counts = !(P.flags & TKFMI_NO_COUNTS) && P.out_n != nullptr;
  if (want_counts && L >= 2) {
    const int NTL = (S + 15) >> 4;
    const double* as = P.as + d.row * S;
    const double* uu = P.em + (d.row + 1) * S;  // u_{t+1} at row t
    const double* T = P.tmat + (size_t)d.tr * 2 * S * S;
    double* out_n = P.out_n + (size_t)d.pair * S * S;
    const int K = L - 1;
    for (int tile = w; tile < NTL * NTL; tile += 4) {
      const int stl = tile / NTL, ktl = tile - stl * NTL;
      const int sA = 16 * stl + (lane & 15), kB = 16 * ktl + (lane & 15);
      const bool okA = sA < S, okB = kB < S;
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      for (int c = 0; c < K; c += 4) {
        const int tt = c + (lane >> 4);
        const bool ok = tt < K;
        const double av = (ok && okA) ? as[(size_t)tt * S + sA] : 0.0;
        const double bv = (ok && okB) ? uu[(size_t)tt * S + kB] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int s = 16 * stl + 4 * v + (lane >> 4), c2 = 16 * ktl + (lane & 15);
        if (s < S && c2 < S) out_n[s * S + c2] += T[(size_t)s * S + c2] * acc[v];
      }
    }
  }
  const bool want_ecnt = (P.flags & TKFMI_WANT_EMIT_COUNTS) && P.mode == 1 && P.out_ecnt && P.post;
  if (!want_ecnt) return;
  const int G = P.G;
  for (int q = threadIdx.x; q < G * TAB_SZ; q += 256) s_ec[q] = 0.0;
  __syncthreads();
  const double* post = P.post + d.row * S;
  const int64_t n = (int64_t)L * S;
  for (int64_t q = threadIdx.x; q < n; q += 256) {
    const int64_t t = q / S;
    const int kk = (int)(q - t * S);
    const int ct = P.col[d.src_row + t];
    if (P.st[kk] != ct || ct < ST_M || ct > ST_D) continue;
    const int a = min((int)P.anc[d.src_row + t], A_PAD - 1);
    const int b = min((int)P.dsc[d.src_row + t], A_PAD - 1);
    const int e = ct == ST_M ? a * A_PAD + b : (ct == ST_I ? A_PAD * A_PAD + b : A_PAD * A_PAD + A_PAD + a);
    const double pv = post[q];
    if (pv != 0.0) atomicAdd(&s_ec[P.sg[kk] * TAB_SZ + e], pv);
  }
  __syncthreads();
  double* oe = P.out_ecnt + (size_t)d.pair * G * TAB_SZ;
  for (int q = threadIdx.x; q < G * TAB_SZ; q += 256) oe[q] += s_ec[q];
}

// ---------------------------------------------------------------- host side
size_t align256(size_t v) { return (v + 255) / 256 * 256; }

struct Plan1D {
  bool ok = false;
  std::string why;
  int mode = 0, NW = 1, LD = 1;
  bool TL = true;
  int64_t rows = 0;
  std::vector<F1Desc> desc;
  size_t desc_b = 0, idx_b = 0, tm_b = 0, pinfo_b = 0, em_b = 0, al_b = 0, as_b = 0, ea_b = 0,
         post_b = 0, sweep_smem = 0, cnt_smem = 0;
  size_t total() const { return desc_b + idx_b + tm_b + pinfo_b + em_b + al_b + as_b + ea_b + post_b; }
};

Plan1D plan(const tkfmi_problem1d* p, uint32_t flags, bool have_post) {
  Plan1D pl;
  if (!p) { pl.why = "null problem"; return pl; }
  const int B = p->B, S = p->S;
  if (B < 0) { pl.why = "B < 0"; return pl; }
  if (S < 1 || S > kMaxS) { pl.why = "S must be in 1..256"; return pl; }
  if (p->s_idx < 0 || p->s_idx >= S || p->e_idx < 0 || p->e_idx >= S) { pl.why = "init/final state out of range"; return pl; }
  if (p->n_trans != 1 && p->n_trans != B) { pl.why = "n_trans must be 1 or B"; return pl; }
  if (B > 0 && (!p->h_off || !p->h_len)) { pl.why = "h_off / h_len missing"; return pl; }
  pl.mode = p->log_emit ? 0 : 1;
  if (pl.mode == 1) {
    if (!p->state_types || !p->state_group || p->n_groups < 1 || !p->log_tab) {
      pl.why = "constrained mode needs state_types, state_group, n_groups >= 1 and log_tab"; return pl;
    }
    if (p->n_tab != 1 && p->n_tab != B) { pl.why = "n_tab must be 1 or B"; return pl; }
    if (B > 0 && (!p->col_type || !p->anc || !p->desc)) { pl.why = "col_type / anc / desc missing"; return pl; }
    for (int s = 0; s < S; ++s)
      if (p->state_group[s] < 0 || p->state_group[s] >= p->n_groups) { pl.why = "state_group out of range"; return pl; }
    if ((size_t)p->n_groups * TAB_SZ * 8 > 160 * 1024) { pl.why = "too many emission groups"; return pl; }
  }
  const bool fwd = flags & TKFMI_FORWARD_ONLY;
  const bool ecnt = !fwd && (flags & TKFMI_WANT_EMIT_COUNTS) && pl.mode == 1;
  pl.desc.resize(B);
  int64_t row = 0;
  for (int b = 0; b < B; ++b) {
    const int L = p->h_len[b];
    if (L < 0 || L > (1 << 28)) { pl.why = "alignment length out of range"; return pl; }
    if (p->h_off[b] < 0) { pl.why = "negative row offset"; return pl; }
    F1Desc& d = pl.desc[b];
    d.row = row; d.src_row = p->h_off[b]; d.L = L; d.pair = b;
    d.tr = p->n_trans > 1 ? b : 0;
    d.tb = (pl.mode == 1 && p->n_tab > 1) ? b : 0;
    row += L;
  }
  pl.rows = row;
  // longest pairs first: their serial chains set the makespan
  std::stable_sort(pl.desc.begin(), pl.desc.end(), [](const F1Desc& x, const F1Desc& y) { return x.L > y.L; });
  pl.NW = (S + 63) / 64;
  pl.TL = S <= kLdsT;
  pl.LD = S | 1;
  const size_t SP = (size_t)((S + 1) & ~1);
  pl.sweep_smem = ((pl.TL ? (size_t)S * pl.LD : 0) + 2 * SP + 2 * (size_t)pl.NW) * 8;
  pl.cnt_smem = ecnt ? (size_t)p->n_groups * TAB_SZ * 8 : 0;
  const size_t cells = (size_t)row * S;
  pl.desc_b = align256(sizeof(F1Desc) * (size_t)std::max(B, 1));
  pl.idx_b = align256(2 * sizeof(int32_t) * (size_t)S);
  pl.tm_b = align256((size_t)p->n_trans * 2 * S * S * 8);
  pl.pinfo_b = align256(2 * 8 * (size_t)std::max(B, 1));
  pl.em_b = align256(cells * 8);
  if (!fwd) {
    pl.al_b = align256(cells * 8);
    pl.as_b = align256(cells * 8);
    pl.ea_b = align256((size_t)row * 8);
    if (ecnt && !have_post) pl.post_b = align256(cells * 8);
  }
  pl.ok = true;
  return pl;
}

template <int NW, bool TL>
int launch_sweep(const F1Params& prm, int B, size_t smem, hipStream_t s, std::string& err) {
  const void* fn = (const void*)fb1d_sweep_kernel<NW, TL>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) { err = std::string("hipFuncSetAttribute: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  hipLaunchKernelGGL((fb1d_sweep_kernel<NW, TL>), dim3(B), dim3(NW * 64), smem, s, prm);
  e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("fb1d sweep launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  return TKFMI_OK;
}

int run(const tkfmi_problem1d* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
        hipStream_t stream, uint32_t flags, std::string& err) {
  const bool fwd = flags & TKFMI_FORWARD_ONLY;
  const bool have_post = !fwd && (flags & TKFMI_WANT_POSTERIORS) && o->posteriors;
  const Plan1D pl = plan(p, flags, have_post);
  if (!pl.ok) { err = pl.why; return TKFMI_EINVAL; }
  if (!o || !o->log_prob) { err = "log_prob output missing"; return TKFMI_EINVAL; }
  if (pl.mode == 0 && p->B > 0 && pl.rows > 0 && !p->log_emit) { err = "log_emit missing"; return TKFMI_EINVAL; }
  if (!ws || ws_bytes < pl.total()) { err = "workspace too small: need " + std::to_string(pl.total()); return TKFMI_EWORKSPACE; }
  const int B = p->B, S = p->S;
  if (B == 0) return TKFMI_OK;
  char* w = static_cast<char*>(ws);
  F1Params prm;
  memset(&prm, 0, sizeof(prm));
  prm.S = S; prm.LD = pl.LD; prm.s_idx = p->s_idx; prm.e_idx = p->e_idx;
  prm.flags = (int32_t)flags; prm.mode = pl.mode; prm.G = pl.mode == 1 ? p->n_groups : 0;
  size_t off = 0;
  auto take = [&](size_t bytes) { char* q = w + off; off += bytes; return q; };
  prm.desc = reinterpret_cast<const F1Desc*>(take(pl.desc_b));
  int32_t* idx = reinterpret_cast<int32_t*>(take(pl.idx_b));
  prm.st = idx; prm.sg = idx + S;
  prm.tmat = reinterpret_cast<const double*>(take(pl.tm_b));
  prm.pinfo = reinterpret_cast<double*>(take(pl.pinfo_b));
  prm.em = reinterpret_cast<double*>(take(pl.em_b));
  prm.al = pl.al_b ? reinterpret_cast<double*>(take(pl.al_b)) : nullptr;
  prm.as = pl.as_b ? reinterpret_cast<double*>(take(pl.as_b)) : nullptr;
  prm.ea = pl.ea_b ? reinterpret_cast<double*>(take(pl.ea_b)) : nullptr;
  double* post_ws = pl.post_b ? reinterpret_cast<double*>(take(pl.post_b)) : nullptr;
  prm.post = have_post ? o->posteriors : post_ws;
  prm.log_emit = p->log_emit;
  prm.log_tab = p->log_tab;
  prm.col = p->col_type; prm.anc = p->anc; prm.dsc = p->desc;
  prm.out_logP = o->log_prob;
  prm.out_n = fwd ? nullptr : o->n_trans;
  prm.out_ecnt = fwd ? nullptr : o->emit_counts;
  // host-side descriptors and state metadata
  std::vector<int32_t> hidx(2 * (size_t)S, 0);
  if (pl.mode == 1)
    for (int s = 0; s < S; ++s) { hidx[s] = p->state_types[s]; hidx[S + s] = p->state_group[s]; }
  int rc = tkfmi_stage_upload((void*)prm.desc, pl.desc.data(), sizeof(F1Desc) * (size_t)B, stream, err);
  if (rc) return rc;
  rc = tkfmi_stage_upload(idx, hidx.data(), hidx.size() * sizeof(int32_t), stream, err);
  if (rc) return rc;
  int64_t cells = 0;
  for (int b = 0; b < B; ++b) cells += (int64_t)p->h_len[b];
  hipLaunchKernelGGL(fb1d_prep_kernel, dim3(p->n_trans), dim3(256), 0, stream, p->log_trans, S, p->e_idx,
                     const_cast<double*>(prm.tmat));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("fb1d prep launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  if (pl.rows > 0) {
    hipLaunchKernelGGL(fb1d_emit_kernel, dim3(B), dim3(256), 0, stream, prm);
    e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("fb1d emit launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  }
  const int slot = tkfmi_ktime_start("fb1d_sweep", B, cells, stream, err);
  if (slot == -2) return TKFMI_EHIP;
  switch (pl.NW) {
    case 1: rc = launch_sweep<1, true>(prm, B, pl.sweep_smem, stream, err); break;
    case 2: rc = launch_sweep<2, true>(prm, B, pl.sweep_smem, stream, err); break;
    case 3: rc = launch_sweep<3, false>(prm, B, pl.sweep_smem, stream, err); break;
    default: rc = launch_sweep<4, false>(prm, B, pl.sweep_smem, stream, err); break;
  }
  if (rc) return rc;
  if (slot >= 0 && tkfmi_ktime_stop(slot, stream, err)) return TKFMI_EHIP;
  const bool want_counts = !fwd && !(flags & TKFMI_NO_COUNTS) && o->n_trans;
  const bool want_ecnt = !fwd && (flags & TKFMI_WANT_EMIT_COUNTS) && pl.mode == 1 && o->emit_counts;
  if (want_counts || want_ecnt) {
    if (pl.cnt_smem > 64 * 1024) {
      e = hipFuncSetAttribute((const void*)fb1d_counts_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      if (e != hipSuccess) { err = std::string("hipFuncSetAttribute: ") + hipGetErrorString(e); return TKFMI_EHIP; }
    }
    const int cs = tkfmi_ktime_start("fb1d_counts", B, cells, stream, err);
    if (cs == -2) return TKFMI_EHIP;
    hipLaunchKernelGGL(fb1d_counts_kernel, dim3(B), dim3(256), pl.cnt_smem, stream, prm);
    e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("fb1d counts launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
    if (cs >= 0 && tkfmi_ktime_stop(cs, stream, err)) return TKFMI_EHIP;
  }
  return TKFMI_OK;
}

}  // namespace tkfmi_1d

extern "C" size_t tkfmi_fb1d_workspace_bytes(const tkfmi_problem1d* prob, uint32_t flags) {
  tkfmi_clear_error();
  // the posterior buffer is counted as workspace unless the caller provides
  // one (TKFMI_WANT_POSTERIORS), which the size query cannot see: assume not
  const bool fwd = flags & TKFMI_FORWARD_ONLY;
  const tkfmi_1d::Plan1D pl = tkfmi_1d::plan(prob, flags, !fwd && (flags & TKFMI_WANT_POSTERIORS));
  if (!pl.ok) {
    tkfmi_fail(TKFMI_EINVAL, pl.why);
    return 0;
  }
  return std::max<size_t>(pl.total(), 256);
}

extern "C" int tkfmi_fb1d_batched(const tkfmi_problem1d* prob, const tkfmi_outputs* out,
                                  void* workspace, size_t workspace_bytes, void* stream,
                                  uint32_t flags) {
  tkfmi_clear_error();
  std::string err;
  tkfmi_ktime_begin_call((flags & TKFMI_TIME_KERNELS) != 0);
  const int rc = tkfmi_1d::run(prob, out, workspace, workspace_bytes,
                               static_cast<hipStream_t>(stream), flags, err);
  if (rc != TKFMI_OK) return tkfmi_fail(rc, err);
  return TKFMI_OK;
}
