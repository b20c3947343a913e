// tkfmi: forward/backward + expected counts for wide state spaces (16 to 160
// emitting states, e.g. MixDom K=8 F=4: 2 + 160 states) on gfx950.
// Reference algorithm: python/tkfmixdom/jax/dp/hmm.py (forward :1186-1292 and
// the row-scan variant :1440-1533 used for S > 50, backward :1295-1437 /
// :1536-1692, counts :1930-1971).
//
// Design:
//  * One workgroup of 8 waves per pair sweeps the lattice by antidiagonals.
//    A cell's NB state values live in HBM (the forward lattice, NB doubles +
//    one binary exponent per cell); the values each cell pushes to its
//    successors sit in a 3-slot ring indexed by antidiagonal (d mod 3) and
//    row i.
//  * Per antidiagonal the transition contraction h = F P (forward) and
//    B = g P^T (backward) is a (cells x NBP) . (NBP x NBP) GEMM on the matrix
//    cores: work items are (16 cells, 16 states) output tiles, each NBP/4
//    v_mfma_f64_16x16x4f64 steps.
//  * The expected counts sum_c F_c g_c^T are a second GEMM per antidiagonal
//    (K = the antidiagonal's cells), accumulated over the whole backward
//    sweep in 16x16 register tiles owned round-robin by the waves.
//  * Emissions, per-cell exponent normalisation, posteriors and emission
//    counts are elementwise phases, one wave per cell, with the states spread
//    over the lanes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "fb2d_internal.h"
#include "tkfmi.h"

namespace tkfmi_wide {

constexpr int ST_S = 0, ST_M = 1, ST_I = 2, ST_D = 3, ST_E = 4;
constexpr int A_PAD = TKFMI_A_PAD;
constexpr int TAB_SZ = TKFMI_TAB_SZ;
constexpr int EXP_NONE = -(1 << 29);
constexpr int kWaves = 8;
constexpr int kMaxNB = 160;
constexpr double LN2 = 0.69314718055994530942;

typedef double d4 __attribute__((ext_vector_type(4)));

struct WDesc {
  int32_t pair, Lx, Ly;
  int32_t ck;        // checkpoint interval K in antidiagonals (0: full lattice)
  int64_t lat_off;   // doubles: forward lattice [(Lx+1)(Ly+1)][NBP], then exponents;
                     // with checkpoints: checkpoints [nck][2][Lx+1][NBP+1], then the
                     // segment lattice [K][Lx+1][NBP] + [K][Lx+1], then a ring
                     // [3][Lx+1][NBP] + [3][Lx+1] for the recomputation
  int64_t ring_off;  // doubles: ring [3][Lx+1][NBP], exponents [3][Lx+1]
  int64_t cur_off;   // doubles: 3 antidiagonal buffers [Lx+1][NBP]
  int64_t pm_off;    // doubles: transition block
  int64_t tab_off;   // doubles: probability-space emission tables [G][TAB_SZ]
  int64_t x_off, y_off, post_off;
};

struct WParams {
  int32_t S, NB, NBP, flags;
  int32_t e_idx, G, pad0, pad1;
  const int32_t* bidx;   // [NBP] state index of emitting slot k (padding: 0)
  const int32_t* btype;  // [NBP] type (0 for padding)
  const int32_t* bgrp;   // [NBP] emission group
  const uint8_t* x_cat;
  const uint8_t* y_cat;
  const double* log_emit;  // dense mode: pair b's [i][j][state] log table at desc.post_off
  const WDesc* desc;
  double* ws;
  double* out_logP;
  double* out_n;
  double* out_post;
  double* out_ecnt;
};

// Transition block per pair (probability space): P [NBP][NBP], PT [NBP][NBP],
// p_start [NBP], p_end [NBP], p_se.
__host__ __device__ inline int64_t pm_stride(int NBP) { return 2 * (int64_t)NBP * NBP + 2 * NBP + 8; }

__global__ void wide_prep_kernel(const double* __restrict__ log_trans, int S, int NB, int NBP,
                                 int e_idx, const int32_t* __restrict__ bidx, double* __restrict__ pm) {
  const int b = blockIdx.x;
  const double* lt = log_trans + (size_t)b * S * S;
  double* o = pm + (size_t)b * pm_stride(NBP);
  for (int q = threadIdx.x; q < NBP * NBP; q += blockDim.x) {
    const int s = q / NBP, k = q % NBP;
    const double v = (s < NB && k < NB) ? exp(lt[bidx[s] * S + bidx[k]]) : 0.0;
    o[q] = v;
    o[(size_t)NBP * NBP + (size_t)k * NBP + s] = v;
  }
  for (int k = threadIdx.x; k < NBP; k += blockDim.x) {
    o[2 * (size_t)NBP * NBP + k] = (k < NB) ? exp(lt[bidx[k]]) : 0.0;
    o[2 * (size_t)NBP * NBP + NBP + k] = (k < NB) ? exp(lt[bidx[k] * S + e_idx]) : 0.0;
  }
  if (threadIdx.x == 0) o[2 * (size_t)NBP * NBP + 2 * NBP] = exp(lt[e_idx]);
}

__global__ void wide_tab_kernel(const double* __restrict__ log_tab, int n, double* __restrict__ out) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < (size_t)n; q += (size_t)gridDim.x * blockDim.x)
    out[q] = exp(log_tab[q]);
}

__device__ inline double wave_max(double v) {
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double wave_sum(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline int frexp_exp(double v) { return __builtin_amdgcn_frexp_exp(v); }

// Emission probability of state slot k (type ty, group g) at cell (i, j).
__device__ inline double emit(const double* tab, int ty, int g, int xi, int yj, int i, int j) {
  const double* t = tab + (size_t)g * TAB_SZ;
  if (ty == ST_M) return (i >= 1 && j >= 1) ? t[xi * A_PAD + yj] : 0.0;
  if (ty == ST_I) return (j >= 1) ? t[A_PAD * A_PAD + yj] : 0.0;
  if (ty == ST_D) return (i >= 1) ? t[A_PAD * A_PAD + A_PAD + xi] : 0.0;
  return 0.0;
}

// Rows of the transition block held in LDS (the rest is read from global
// memory): as many 4-row chunks as fit in 159 KB, each LDS row padded by 16
// doubles so that the four rows one MFMA operand load touches fall in two
// bank halves (two passes, no further conflicts).
__host__ __device__ constexpr int lds_b_stride(int NBP) { return NBP + 16; }
__host__ __device__ constexpr int lds_b_rows(int NBP) {
  const int r = (159 * 1024) / (lds_b_stride(NBP) * 8) / 4 * 4;
  return r < NBP ? r : NBP;
}
// Output tiles per GEMM task: one 16-row block across CT column tiles, so
// each A fragment (a cell row, read from global memory) feeds CT MFMAs; the
// B fragments come mostly from LDS.  The backward, whose count tiles are live,
// uses the largest divisor of NT up to 5; the forward all NT tiles.
__host__ __device__ constexpr int gemm_ct(int NT) {
  return NT % 5 == 0 ? 5 : (NT % 4 == 0 ? 4 : (NT % 3 == 0 ? 3 : (NT % 2 == 0 ? 2 : 1)));
}

// C[cells x NBP] = A[cells x NBP] . Bm[NBP x NBP] for the rows i_lo..i_hi of
// an antidiagonal (A and C row-indexed by i, row length NBP); the result
// for row i_lo is replaced by `fix` when it is not null.  Rows 0..KL-1 of Bm
// are read from the LDS copy sB (row stride lds_b_stride).
template <int NT, int CT>
__device__ inline void gemm_rows(const double* __restrict__ A, const double* __restrict__ Bm,
                                 const double* __restrict__ sB, double* __restrict__ Cm, int i_lo, int i_hi,
                                 const double* fix) {
  constexpr int NBP = 16 * NT;
  static_assert(NT % CT == 0, "column tiles per task");
  constexpr int NCG = NT / CT;
  constexpr int KL4 = lds_b_rows(NBP) / 4;  // k-chunks read from LDS
  constexpr int SB = lds_b_stride(NBP);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = i_hi - i_lo + 1;
  const int nrb = (n + 15) >> 4;
  for (int task = w; task < nrb * NCG; task += kWaves) {
    const int rb = task / NCG, cg = task - rb * NCG;
    const int ra = i_lo + 16 * rb + (lane & 15);
    const bool va = ra <= i_hi;
    const double* ap = A + (size_t)(va ? ra : i_lo) * NBP + (lane >> 4);
    const int col0 = 16 * (cg * CT) + (lane & 15);
    const double* sbp = sB + (lane >> 4) * SB + col0;
    const double* bp = Bm + (size_t)(lane >> 4) * NBP + col0;
    d4 acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
    // A fragments (cell rows, 1280 B apart: one memory round trip each) are
    // loaded one batch of 4 k-chunks ahead of the MFMAs that use them
    static_assert(KL4 % 4 == 0, "LDS k-chunks in batches of 4");
    double an[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) an[q] = ap[4 * q];
#pragma unroll 1
    for (int kb = 0; kb < KL4; kb += 4) {
      double a[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) a[q] = va ? an[q] : 0.0;
      if (kb + 4 < NBP / 4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) an[q] = (kb + 4 + q < NBP / 4) ? ap[4 * (kb + 4 + q)] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int kk = kb + q;
#pragma unroll
        for (int c = 0; c < CT; ++c)
          acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], sbp[4 * kk * SB + 16 * c], acc[c], 0, 0, 0);
      }
    }
#pragma unroll 2
    for (int kk = KL4; kk < NBP / 4; ++kk) {
      const double a = va ? ap[4 * kk] : 0.0;
      double b[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) b[c] = bp[(size_t)4 * kk * NBP + 16 * c];
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[c], acc[c], 0, 0, 0);
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = i_lo + 16 * rb + 4 * v + (lane >> 4);
        if (row <= i_hi) {
          const int col = 16 * (cg * CT + c) + (lane & 15);
          Cm[(size_t)row * NBP + col] = (fix && row == i_lo) ? fix[col] : acc[c][v];
        }
      }
    }
  }
}

// Copies rows 0..lds_b_rows-1 of the NBP x NBP block Bm into LDS (whole
// workgroup; the caller synchronises).
template <int NT>
__device__ inline void stage_b(const double* __restrict__ Bm, double* __restrict__ sB) {
  constexpr int NBP = 16 * NT, KL = lds_b_rows(NBP), SB = lds_b_stride(NBP);
  for (int q = threadIdx.x; q < KL * NBP; q += kWaves * 64) {
    const int r = q / NBP, c = q - r * NBP;
    sB[r * SB + c] = Bm[q];
  }
}

// The NT x NT count tiles (16 x 16 state blocks: row kt holds forward-value
// fragment kt, column k2t g fragment k2t) are dealt out to the waves, tile
// w + kWaves q to wave w.  A wave's tiles use most fragments more than once,
// so a 4-cell step loads each fragment the wave needs once (at NT = 10: 15
// loads for 13 tiles instead of 26), all in flight together.
__host__ __device__ constexpr int count_ntile(int NT) { return (NT * NT + kWaves - 1) / kWaves; }
__host__ __device__ constexpr bool count_uses(int NT, int W0, int frag, bool row) {
  for (int tile = W0; tile < NT * NT; tile += kWaves)
    if ((row ? tile / NT : tile % NT) == frag) return true;
  return false;
}

// counts += F^T g over the cells i_lo..i_hi of one antidiagonal, for the
// tiles of wave W0.  cur0 rows hold the scaled forward values of the cells,
// cur1 rows their g; 4 cells per MFMA (lane >> 4 picks the cell, lane & 15
// the state within the fragment).
template <int NT, int W0, int NTILE>
__device__ __forceinline__ void count_antidiag(const double* __restrict__ cur0, const double* __restrict__ cur1,
                                               int i_lo, int i_hi, int lane, d4 (&cacc)[NTILE]) {
  constexpr int NBP = 16 * NT;
  const int nq = (i_hi - i_lo + 4) >> 2;
  for (int c4 = 0; c4 < nq; ++c4) {
    const int row = i_lo + 4 * c4 + (lane >> 4);
    const bool v = row <= i_hi;
    const double* r0 = cur0 + (size_t)(v ? row : i_lo) * NBP + (lane & 15);
    const double* r1 = cur1 + (size_t)(v ? row : i_lo) * NBP + (lane & 15);
    double fa[NT], fb[NT];
#pragma unroll
    for (int f = 0; f < NT; ++f) {
      if (count_uses(NT, W0, f, true)) { const double x = r0[16 * f]; fa[f] = v ? x : 0.0; }
      if (count_uses(NT, W0, f, false)) { const double y = r1[16 * f]; fb[f] = v ? y : 0.0; }
    }
#pragma unroll
    for (int q = 0; q < NTILE; ++q) {
      const int tile = W0 + kWaves * q;
      if (tile < NT * NT)
        cacc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[tile / NT], fb[tile % NT], cacc[q], 0, 0, 0);
    }
  }
}

// count_antidiag for the runtime wave index w (one instantiation per wave,
// so that fragment and tile indices are compile-time register indices).
template <int NT, int NTILE, int W0 = 0>
__device__ __forceinline__ void count_antidiag_w(int w, const double* __restrict__ cur0,
                                                 const double* __restrict__ cur1, int i_lo, int i_hi, int lane,
                                                 d4 (&cacc)[NTILE]) {
  if (w == W0) count_antidiag<NT, W0, NTILE>(cur0, cur1, i_lo, i_hi, lane, cacc);
  else if constexpr (W0 + 1 < kWaves) count_antidiag_w<NT, NTILE, W0 + 1>(w, cur0, cur1, i_lo, i_hi, lane, cacc);
}

// DENSE: emissions from the caller's dense log table (the reference's
// log_emit_table) instead of the per-group tables.
template <int NT, bool DENSE>
__global__ __launch_bounds__(kWaves * 64) void wide_kernel(WParams Pr) {
  constexpr int NBP = 16 * NT;
  constexpr int NTILE = count_ntile(NT);                   // count tiles per wave
  constexpr int KR = (NBP + 63) / 64;                      // state rounds per lane
  __shared__ double s_misc[4];
  extern __shared__ double s_B[];  // rows of P (forward) or PT (backward)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const WDesc pd = Pr.desc[blockIdx.x];
  const int Lx = pd.Lx, Ly = pd.Ly, W1 = Ly + 1, NB = Pr.NB, S = Pr.S;
  const double* pm = Pr.ws + pd.pm_off;
  const double* P = pm;
  const double* PT = pm + (size_t)NBP * NBP;
  const double* p_start = pm + 2 * (size_t)NBP * NBP;
  const double* p_end = p_start + NBP;
  const double p_se = p_start[2 * NBP];
  const double* tab = Pr.ws + pd.tab_off;
  const uint8_t* xs = DENSE ? nullptr : Pr.x_cat + pd.x_off;
  const uint8_t* ys = DENSE ? nullptr : Pr.y_cat + pd.y_off;
  const double* dem = DENSE ? Pr.log_emit + pd.post_off : nullptr;
  double* lat = Pr.ws + pd.lat_off;
  double* latE = lat + (size_t)(Lx + 1) * W1 * NBP;
  double* ring = Pr.ws + pd.ring_off;
  double* ringE = ring + 3 * (size_t)(Lx + 1) * NBP;
  double* cur0 = Pr.ws + pd.cur_off;
  double* cur1 = cur0 + (size_t)(Lx + 1) * NBP;
  double* cur2 = cur1 + (size_t)(Lx + 1) * NBP;
  const bool want_lat = !(Pr.flags & TKFMI_FORWARD_ONLY);
  const bool want_counts = want_lat && !(Pr.flags & TKFMI_NO_COUNTS) && Pr.out_n != nullptr;
  const bool want_post = want_lat && (Pr.flags & TKFMI_WANT_POSTERIORS) && Pr.out_post != nullptr;
  const bool want_ecnt = want_lat && (Pr.flags & TKFMI_WANT_EMIT_COUNTS) && Pr.out_ecnt != nullptr;
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  // slot metadata of this lane's states
  int ty[KR], gr[KR], sx[KR];
#pragma unroll
  for (int r = 0; r < KR; ++r) {
    const int k = lane + 64 * r;
    ty[r] = (k < NB) ? Pr.btype[k] : 0;
    gr[r] = (k < NB) ? Pr.bgrp[k] : 0;
    sx[r] = (k < NB) ? Pr.bidx[k] : 0;
  }
  auto xchar = [&](int i) { return (!DENSE && i >= 1) ? (int)min((int)xs[i - 1], A_PAD - 1) : 0; };
  auto ychar = [&](int j) { return (!DENSE && j >= 1) ? (int)min((int)ys[j - 1], A_PAD - 1) : 0; };
  // emission probability of slot round r (type t) at cell (i, j)
  auto em = [&](int r, int t, int xi, int yj, int i, int j) -> double {
    if constexpr (DENSE) {
      const bool ok = (t == ST_M) ? (i >= 1 && j >= 1) : (t == ST_I) ? (j >= 1) : (t == ST_D) ? (i >= 1) : false;
      return ok ? exp(dem[((size_t)i * W1 + j) * S + sx[r]]) : 0.0;
    } else {
      return emit(tab, t, gr[r], xi, yj, i, j);
    }
  };

  // ======================================================== forward sweep
  // One antidiagonal d of the forward: F from the ring slots of d-1 and d-2
  // (rg, rgE), written to the lattice (F of cell i at fb + i*fs, its
  // exponent at eb + i*es; fb null: not kept) and to cur0, then h = F P into
  // the ring slot of d.  `main`: the first sweep, which also produces log P
  // and the E column; a recomputation inside the backward skips them.
  // With checkpoints (pd.ck = K > 0) the first sweep keeps only the restart
  // state of every K-th antidiagonal, and the backward recomputes each
  // segment of K antidiagonals from its checkpoint when it enters it: the
  // same arithmetic in the same order, so the values are those of the full
  // lattice.
  const int ck = want_lat ? pd.ck : 0;
  const int nck = ck ? (Lx + Ly) / ck + 1 : 0;
  double* ckp = lat;                                               // checkpoints
  double* seg = lat + (size_t)nck * 2 * (Lx + 1) * (NBP + 1);      // segment lattice
  double* segE = seg + (size_t)ck * (Lx + 1) * NBP;
  double* rring = segE + (size_t)ck * (Lx + 1);                    // recomputation ring
  double* rringE = rring + 3 * (size_t)(Lx + 1) * NBP;
  auto fstep = [&](int d, double* rg, double* rgE, double* fb, size_t fs, double* eb, size_t es, bool main) {
    const int i_lo = max(0, d - Ly), i_hi = min(d, Lx);
    const int sd = d % 3, s1 = (d + 2) % 3, s2 = (d + 1) % 3;  // slots of d, d-1, d-2
    // The residue codes and predecessor exponents of the cells this wave
    // visits (i = i_lo + w + kWaves k) are loaded 64 cells at a time, lane l
    // holding the l-th, and handed out with readlane: a cell's chain of
    // dependent loads then starts at its emission-table reads.
    int bx = 0, by = 0, beM = EXP_NONE, beI = EXP_NONE, beD = EXP_NONE;
    // raw products (predecessor value x emission) of cell i, batch index kc;
    // the next cell's are loaded before this cell's stores, so its loads
    // overlap this cell's reductions
    auto load_raw = [&](int i, int kc, double (&raw)[KR]) {
      const int j = d - i;
      const int xi = __builtin_amdgcn_readlane(bx, kc & 63), yj = __builtin_amdgcn_readlane(by, kc & 63);
      const bool okM = i >= 1 && j >= 1, okI = j >= 1, okD = i >= 1;
#pragma unroll
      for (int r = 0; r < KR; ++r) {
        const int k = lane + 64 * r;
        double v = 0.0;
        if (ty[r] == ST_M && okM) v = rg[((size_t)s2 * (Lx + 1) + i - 1) * NBP + k] * em(r, ST_M, xi, yj, i, j);
        else if (ty[r] == ST_I && okI) v = rg[((size_t)s1 * (Lx + 1) + i) * NBP + k] * em(r, ST_I, xi, yj, i, j);
        else if (ty[r] == ST_D && okD) v = rg[((size_t)s1 * (Lx + 1) + i - 1) * NBP + k] * em(r, ST_D, xi, yj, i, j);
        raw[r] = v;
      }
    };
    double pre[KR];
    bool have_pre = false;
    int kc = 0;
    for (int i = i_lo + w; i <= i_hi; i += kWaves, ++kc) {
      if ((kc & 63) == 0) {
        const int ci = i + kWaves * lane, cj = d - ci;
        const bool in = ci <= i_hi;
        bx = in ? xchar(ci) : 0;
        by = in ? ychar(cj) : 0;
        // predecessor exponents: M from (i-1, j-1) [d-2], I from (i, j-1) [d-1], D from (i-1, j) [d-1]
        beM = (in && ci >= 1 && cj >= 1) ? (int)rgE[s2 * (Lx + 1) + (ci - 1)] : EXP_NONE;
        beI = (in && cj >= 1) ? (int)rgE[s1 * (Lx + 1) + ci] : EXP_NONE;
        beD = (in && ci >= 1) ? (int)rgE[s1 * (Lx + 1) + (ci - 1)] : EXP_NONE;
      }
      const int j = d - i;
      const int eM = __builtin_amdgcn_readlane(beM, kc & 63);
      const int eI = __builtin_amdgcn_readlane(beI, kc & 63);
      const int eD = __builtin_amdgcn_readlane(beD, kc & 63);
      double raw[KR];
      if (have_pre) {
#pragma unroll
        for (int r = 0; r < KR; ++r) raw[r] = pre[r];
      } else {
        load_raw(i, kc, raw);
      }
      have_pre = i + kWaves <= i_hi && ((kc + 1) & 63) != 0;
      if (have_pre) load_raw(i + kWaves, kc + 1, pre);
      double mM = 0.0, mI = 0.0, mD = 0.0;
#pragma unroll
      for (int r = 0; r < KR; ++r) {
        if (ty[r] == ST_M) mM = fmax(mM, raw[r]);
        else if (ty[r] == ST_I) mI = fmax(mI, raw[r]);
        else if (ty[r] == ST_D) mD = fmax(mD, raw[r]);
      }
      mM = wave_max(mM);
      mI = wave_max(mI);
      mD = wave_max(mD);
      const int cM = (mM > 0.0) ? eM + frexp_exp(mM) : EXP_NONE;
      const int cI = (mI > 0.0) ? eI + frexp_exp(mI) : EXP_NONE;
      const int cD = (mD > 0.0) ? eD + frexp_exp(mD) : EXP_NONE;
      int E = max(cM, max(cI, cD));
      if (E == EXP_NONE) E = 0;
      const bool origin = (i == 0 && j == 0);
      double F[KR];
#pragma unroll
      for (int r = 0; r < KR; ++r) {
        const int sh = (ty[r] == ST_M) ? eM - E : ((ty[r] == ST_I) ? eI - E : eD - E);
        F[r] = (origin || raw[r] == 0.0) ? 0.0 : ldexp(raw[r], sh);
      }
      if (origin) E = 0;
#pragma unroll
      for (int r = 0; r < KR; ++r) {
        const int k = lane + 64 * r;
        if (k < NBP) {
          cur0[(size_t)i * NBP + k] = F[r];
          if (fb) fb[(size_t)i * fs + k] = F[r];
        }
      }
      if (lane == 0) {
        rgE[sd * (Lx + 1) + i] = (double)E;
        if (eb) eb[(size_t)i * es] = (double)E;
      }
      if (main && i == Lx && j == Ly) {  // log P and the E column
        double part = 0.0;
#pragma unroll
        for (int r = 0; r < KR; ++r) {
          const int k = lane + 64 * r;
          if (k < NB) part = fma(F[r], p_end[k], part);
        }
        double tot = wave_sum(part);
        double lp;
        if (origin) { tot = p_se; lp = log(p_se); }
        else lp = log(tot) + (double)E * LN2;
        if (lane == 0) {
          Pr.out_logP[pd.pair] = (tot > 0.0) ? lp : -1e30;
          s_misc[0] = (tot > 0.0) ? ldexp(tot, -frexp_exp(tot)) : 0.0;
          s_misc[1] = (tot > 0.0) ? (double)(E + frexp_exp(tot)) : 0.0;
        }
        if (want_counts && tot > 0.0) {
          if (origin) { if (lane == 0) out_n[Pr.e_idx] = 1.0; }
          else {
#pragma unroll
            for (int r = 0; r < KR; ++r) {
              const int k = lane + 64 * r;
              if (k < NB) out_n[Pr.bidx[k] * S + Pr.e_idx] = F[r] * p_end[k] / tot;
            }
          }
        }
      }
    }
    __syncthreads();
    // h = F P (the origin pushes p_start)
    if (main)  // no count tiles live: one task covers all column tiles
      gemm_rows<NT, NT>(cur0, P, s_B, rg + (size_t)sd * (Lx + 1) * NBP, i_lo, i_hi, d == 0 ? p_start : nullptr);
    else
      gemm_rows<NT, gemm_ct(NT)>(cur0, P, s_B, rg + (size_t)sd * (Lx + 1) * NBP, i_lo, i_hi, d == 0 ? p_start : nullptr);
    __syncthreads();
  };
  // The state a forward restarts from at antidiagonal d: the ring slots of d-1
  // and d-2 (h values and exponents).
  auto ck_copy = [&](int d, double* rg, double* rgE, double* c, bool save) {
    for (int h = 0; h < 2; ++h) {
      const int sl = (d + 2 - h) % 3;  // h = 0: slot of d-1, h = 1: slot of d-2
      double* cs = c + (size_t)h * (Lx + 1) * (NBP + 1);
      for (int q = threadIdx.x; q < (Lx + 1) * NBP; q += kWaves * 64) {
        double* r = rg + (size_t)sl * (Lx + 1) * NBP + q;
        if (save) cs[q] = *r;
        else *r = cs[q];
      }
      for (int q = threadIdx.x; q <= Lx; q += kWaves * 64) {
        double* r = rgE + sl * (Lx + 1) + q;
        if (save) cs[(size_t)(Lx + 1) * NBP + q] = *r;
        else *r = cs[(size_t)(Lx + 1) * NBP + q];
      }
    }
    __syncthreads();
  };
  stage_b<NT>(P, s_B);
  __syncthreads();
  for (int d = 0; d <= Lx + Ly; ++d) {
    if (ck && d > 0 && d % ck == 0) ck_copy(d, ring, ringE, ckp + (size_t)(d / ck) * 2 * (Lx + 1) * (NBP + 1), true);
    fstep(d, ring, ringE, (want_lat && !ck) ? lat + (size_t)d * NBP : nullptr, (size_t)Ly * NBP,
          (want_lat && !ck) ? latE + d : nullptr, (size_t)Ly, true);
  }
  if (!want_lat) return;

  // ======================================================= backward sweep
  const double mP = s_misc[0];
  const int EP = (int)s_misc[1];
  if (!(mP > 0.0)) return;  // impossible pair: counts / posteriors stay zero
  const double inv_mP = 1.0 / mP;
  d4 cacc[NTILE];
#pragma unroll
  for (int q = 0; q < NTILE; ++q) cacc[q] = d4{0.0, 0.0, 0.0, 0.0};
  double* post = want_post ? Pr.out_post + pd.post_off : nullptr;
  double* ecnt = want_ecnt ? Pr.out_ecnt + (size_t)pd.pair * Pr.G * TAB_SZ : nullptr;
  __syncthreads();
  stage_b<NT>(PT, s_B);
  __syncthreads();
  for (int d = Lx + Ly; d >= 0; --d) {
    const int i_lo = max(0, d - Ly), i_hi = min(d, Lx);
    const int sd = d % 3, s1 = (d + 1) % 3, s2 = (d + 2) % 3;  // slots of d, d+1, d+2
    // ---- the forward values of antidiagonal d: the lattice, or (checkpoints)
    // the segment lattice of d's segment, recomputed when the sweep enters it
    const int d0 = ck ? d / ck * ck : 0;
    if (ck && d == min(Lx + Ly, d0 + ck - 1)) {
      if (d0 > 0) ck_copy(d0, rring, rringE, ckp + (size_t)(d0 / ck) * 2 * (Lx + 1) * (NBP + 1), false);
      stage_b<NT>(P, s_B);
      __syncthreads();
      for (int dd = d0; dd <= d; ++dd)
        fstep(dd, rring, rringE, seg + (size_t)(dd - d0) * (Lx + 1) * NBP, (size_t)NBP,
              segE + (size_t)(dd - d0) * (Lx + 1), 1, false);
      stage_b<NT>(PT, s_B);
      __syncthreads();
    }
    const double* fb = ck ? seg + (size_t)(d - d0) * (Lx + 1) * NBP : lat + (size_t)d * NBP;
    const size_t fs = ck ? (size_t)NBP : (size_t)Ly * NBP;
    const double* eb = ck ? segE + (size_t)(d - d0) * (Lx + 1) : latE + d;
    const size_t es = ck ? 1 : (size_t)Ly;
    // ---- g (successor-pulled q), scaled forward values
    // successor and forward exponents of the wave's cells, 64 at a time
    int beM = EXP_NONE, beI = EXP_NONE, beD = EXP_NONE, bEF = 0;
    int kc = 0;
    for (int i = i_lo + w; i <= i_hi; i += kWaves, ++kc) {
      if ((kc & 63) == 0) {
        const int ci = i + kWaves * lane, cj = d - ci;
        const bool in = ci <= i_hi;
        beM = (in && ci < Lx && cj < Ly) ? (int)ringE[s2 * (Lx + 1) + (ci + 1)] : EXP_NONE;
        beI = (in && cj < Ly) ? (int)ringE[s1 * (Lx + 1) + ci] : EXP_NONE;
        beD = (in && ci < Lx) ? (int)ringE[s1 * (Lx + 1) + (ci + 1)] : EXP_NONE;
        bEF = in ? (int)eb[(size_t)ci * es] : 0;
      }
      const int j = d - i;
      const bool okM = i < Lx && j < Ly, okI = j < Ly, okD = i < Lx;
      const int eM = __builtin_amdgcn_readlane(beM, kc & 63);
      const int eI = __builtin_amdgcn_readlane(beI, kc & 63);
      const int eD = __builtin_amdgcn_readlane(beD, kc & 63);
      int Eg = max(eM, max(eI, eD));
      if (Eg == EXP_NONE) Eg = 0;
      const int EF = __builtin_amdgcn_readlane(bEF, kc & 63);
      const double fsc = ldexp(inv_mP, EF + Eg - EP);
#pragma unroll
      for (int r = 0; r < KR; ++r) {
        const int k = lane + 64 * r;
        if (k >= NBP) continue;
        double g = 0.0;
        if (ty[r] == ST_M && okM) g = ldexp(ring[((size_t)s2 * (Lx + 1) + i + 1) * NBP + k], eM - Eg);
        else if (ty[r] == ST_I && okI) g = ldexp(ring[((size_t)s1 * (Lx + 1) + i) * NBP + k], eI - Eg);
        else if (ty[r] == ST_D && okD) g = ldexp(ring[((size_t)s1 * (Lx + 1) + i + 1) * NBP + k], eD - Eg);
        cur1[(size_t)i * NBP + k] = g;
        cur0[(size_t)i * NBP + k] = fb[(size_t)i * fs + k] * fsc;
        if (want_counts && i == 0 && j == 0 && k < NB)
          out_n[Pr.bidx[k]] = p_start[k] * g * ldexp(inv_mP, Eg - EP);
      }
      if (lane == 0) ringE[sd * (Lx + 1) + i] = (double)Eg;  // exponent of B (before normalising)
      if (want_post && i == 0 && j == 0) {
        double part = 0.0;
#pragma unroll
        for (int r = 0; r < KR; ++r) {
          const int k = lane + 64 * r;
          if (k < NB) part = fma(p_start[k], cur1[k], part);
        }
        part = wave_sum(part);
        if (lane == 0) post[0] = (i == Lx && j == Ly) ? 1.0 : part * ldexp(inv_mP, Eg - EP);
      }
    }
    __syncthreads();
    // ---- B = g P^T, counts += Fs^T g
    gemm_rows<NT, gemm_ct(NT)>(cur1, PT, s_B, cur2, i_lo, i_hi, nullptr);
    if (want_counts)
      count_antidiag_w<NT, NTILE>(w, cur0, cur1, i_lo, i_hi, lane, cacc);
    __syncthreads();
    // ---- normalise B, posteriors, emission counts, q = e . B for the predecessors
    int qx = 0, qy = 0;
    int kq = 0;
    for (int i = i_lo + w; i <= i_hi; i += kWaves, ++kq) {
      if ((kq & 63) == 0) {  // residue codes of the wave's cells, 64 at a time
        const int ci = i + kWaves * lane;
        const bool in = ci <= i_hi;
        qx = in ? xchar(ci) : 0;
        qy = in ? ychar(d - ci) : 0;
      }
      const int j = d - i;
      const bool corner = (i == Lx && j == Ly);
      const int xi = __builtin_amdgcn_readlane(qx, kq & 63), yj = __builtin_amdgcn_readlane(qy, kq & 63);
      double Bv[KR];
      double mx = 0.0;
#pragma unroll
      for (int r = 0; r < KR; ++r) {
        const int k = lane + 64 * r;
        Bv[r] = (k < NBP) ? (corner ? (k < NB ? p_end[k] : 0.0) : cur2[(size_t)i * NBP + k]) : 0.0;
        mx = fmax(mx, Bv[r]);
      }
      mx = wave_max(mx);
      const int sh = (mx > 0.0) ? frexp_exp(mx) : 0;
      const int EB = (corner ? 0 : (int)ringE[sd * (Lx + 1) + i]) + sh;
      const size_t cell = (size_t)i * W1 + j;
      const int EF = (int)eb[(size_t)i * es];
      const double psc = ldexp(inv_mP, EF + EB - EP);
#pragma unroll
      for (int r = 0; r < KR; ++r) {
        const int k = lane + 64 * r;
        if (k >= NBP) continue;
        const double b = ldexp(Bv[r], -sh);
        const double e = (k < NB) ? em(r, ty[r], xi, yj, i, j) : 0.0;
        ring[((size_t)sd * (Lx + 1) + i) * NBP + k] = e * b;
        if (k < NB && (want_post || want_ecnt)) {
          const double pv = fb[(size_t)i * fs + k] * b * psc;
          if (want_post) post[cell * S + Pr.bidx[k]] = pv;
          if (want_ecnt && pv != 0.0) {
            double* tb = ecnt + (size_t)gr[r] * TAB_SZ;
            if (ty[r] == ST_M) atomicAdd(tb + xi * A_PAD + yj, pv);
            else if (ty[r] == ST_I) atomicAdd(tb + A_PAD * A_PAD + yj, pv);
            else atomicAdd(tb + A_PAD * A_PAD + A_PAD + xi, pv);
          }
        }
      }
      if (lane == 0) ringE[sd * (Lx + 1) + i] = (double)EB;
    }
    __syncthreads();
  }
  if (!want_counts) return;
#pragma unroll
  for (int q = 0; q < NTILE; ++q) {
    const int tile = w + kWaves * q;
    if (tile >= NT * NT) break;
    const int kt = tile / NT, k2t = tile - kt * NT;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = 16 * kt + 4 * v + (lane >> 4), k2 = 16 * k2t + (lane & 15);
      if (k < NB && k2 < NB) {
        const double pr = P[(size_t)k * NBP + k2];
        out_n[Pr.bidx[k] * S + Pr.bidx[k2]] = (pr > 0.0) ? pr * cacc[q][v] : 0.0;
      }
    }
  }
}

// ---------------------------------------------------------------- host side
struct WPlan {
  int S = 0, NB = 0, NT = 0, NBP = 0, e_idx = -1, G = 0;
  std::vector<int32_t> bidx, btype, bgrp;
  std::vector<WDesc> desc;
  size_t idx_b = 0, desc_b = 0, pm_b = 0, tab_b = 0, data_b = 0;
  size_t total() const { return idx_b + desc_b + pm_b + tab_b + data_b; }
};

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

bool applies(const tkfmi_problem* p, WPlan& pl, std::string& why) {
  if (!p || p->S < 2 || !p->state_types || !p->state_group) { why = "bad problem"; return false; }
  pl.S = p->S;
  for (int s = 0; s < p->S; ++s) {
    const int ty = p->state_types[s];
    if (ty == ST_E && pl.e_idx < 0) pl.e_idx = s;
    if (ty == ST_M || ty == ST_I || ty == ST_D) {
      pl.bidx.push_back(s);
      pl.btype.push_back(ty);
      pl.bgrp.push_back(p->state_group[s]);
    }
  }
  pl.NB = (int)pl.bidx.size();
  if (pl.NB <= 15 || pl.NB > kMaxNB) { why = "state count outside the wide kernel's range"; return false; }
  if (p->state_types[0] != ST_S || pl.e_idx < 0) { why = "need state 0 of type S and a state of type E"; return false; }
  const int nt = (pl.NB + 15) / 16;
  pl.NT = (nt <= 2) ? 2 : (nt <= 4 ? 4 : (nt <= 6 ? 6 : (nt <= 8 ? 8 : 10)));
  pl.NBP = 16 * pl.NT;
  while ((int)pl.bidx.size() < pl.NBP) { pl.bidx.push_back(0); pl.btype.push_back(0); pl.bgrp.push_back(0); }
  pl.G = p->n_groups;
  for (int k = 0; k < pl.NB; ++k)
    if (pl.bgrp[k] < 0 || pl.bgrp[k] >= pl.G) { why = "state_group out of range"; return false; }
  return true;
}

bool plan(const tkfmi_problem* p, uint32_t flags, WPlan& pl, std::string& why) {
  if (!applies(p, pl, why)) return false;
  if (p->n_trans != 1 && p->n_trans != p->B) { why = "n_trans must be 1 or B"; return false; }
  const bool dense = p->log_emit != nullptr;
  if (!dense && p->n_tab != 1 && p->n_tab != p->B) { why = "n_tab must be 1 or B"; return false; }
  const int B = p->B;
  if (B > 0 && (!p->h_x_len || !p->h_y_len || !p->h_x_off || !p->h_y_off)) { why = "host arrays missing"; return false; }
  pl.idx_b = align256(3 * (size_t)pl.NBP * 4);
  pl.desc_b = align256(sizeof(WDesc) * (size_t)std::max(B, 1));
  pl.pm_b = align256((size_t)pm_stride(pl.NBP) * 8 * p->n_trans);
  pl.tab_b = dense ? 0 : align256((size_t)p->n_tab * pl.G * TAB_SZ * 8);
  // pairs in decreasing cost order
  std::vector<int> order(B);
  for (int b = 0; b < B; ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (int64_t)(p->h_x_len[a] + 1) * (p->h_y_len[a] + 1) > (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1);
  });
  std::vector<int64_t> post_off(B);
  int64_t ap = 0;
  for (int b = 0; b < B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    if (lx < 0 || ly < 0 || lx + ly > 64000) { why = "sequence lengths out of range"; return false; }
    post_off[b] = ap;
    ap += (int64_t)(lx + 1) * (ly + 1) * p->S;
  }
  const int64_t base = (int64_t)((pl.idx_b + pl.desc_b + pl.pm_b + pl.tab_b) / 8);
  int64_t off = base;
  pl.desc.resize(B);
  for (int q = 0; q < B; ++q) {
    const int b = order[q];
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    WDesc d;
    memset(&d, 0, sizeof(d));
    d.pair = b; d.Lx = lx; d.Ly = ly;
    d.pm_off = (int64_t)(pl.idx_b + pl.desc_b) / 8 + (p->n_trans > 1 ? (int64_t)b : 0) * pm_stride(pl.NBP);
    d.tab_off = (int64_t)(pl.idx_b + pl.desc_b + pl.pm_b) / 8 +
                (!dense && p->n_tab > 1 ? (int64_t)b : 0) * pl.G * TAB_SZ;
    d.x_off = p->h_x_off[b]; d.y_off = p->h_y_off[b];
    d.post_off = post_off[b];
    const int64_t cells = (int64_t)(lx + 1) * (ly + 1);
    d.lat_off = off;
    if (!(flags & TKFMI_FORWARD_ONLY)) {
      const int64_t full_bytes = (cells * pl.NBP + cells) * 8;
      const size_t budget = tkfmi_checkpoint_budget_bytes();
      if ((flags & TKFMI_CHECKPOINT) || (budget > 0 && (size_t)full_bytes > budget)) {
        // K ~ sqrt(2 (Lx + Ly + 1)) minimises checkpoints (2 rows each) + segment,
        // unless the caller fixed K (at most the pair's antidiagonals, at least 4)
        const int kset = tkfmi_checkpoint_segment_k();
        const int K = kset > 0 ? std::max(4, std::min(kset, lx + ly + 1))
                               : std::max(4, (int)std::ceil(std::sqrt(2.0 * (lx + ly + 1))));
        const int64_t nck = (lx + ly) / K + 1, row = (int64_t)(lx + 1) * (pl.NBP + 1);
        d.ck = K;
        off += nck * 2 * row + (int64_t)K * row + 3 * row;
      } else {
        off += cells * pl.NBP + cells;
      }
    }
    d.ring_off = off;
    off += 3 * (int64_t)(lx + 1) * pl.NBP + 3 * (int64_t)(lx + 1);
    d.cur_off = off;
    off += 3 * (int64_t)(lx + 1) * pl.NBP;
    off = (off + 31) / 32 * 32;
    pl.desc[q] = d;
  }
  pl.data_b = (size_t)(off - base) * 8;
  return true;
}

template <int NT, bool DENSE>
int launch_nt(const WParams& prm, int B, hipStream_t s, std::string& err) {
  const size_t smem = (size_t)lds_b_rows(16 * NT) * lds_b_stride(16 * NT) * 8;
  hipError_t e = hipFuncSetAttribute((const void*)wide_kernel<NT, DENSE>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) { err = std::string("hipFuncSetAttribute: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  hipLaunchKernelGGL((wide_kernel<NT, DENSE>), dim3(B), dim3(kWaves * 64), smem, s, prm);
  e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("wide launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  return TKFMI_OK;
}

template <bool DENSE>
int launch_d(int NT, const WParams& prm, int B, hipStream_t s, std::string& err) {
  switch (NT) {
    case 2: return launch_nt<2, DENSE>(prm, B, s, err);
    case 4: return launch_nt<4, DENSE>(prm, B, s, err);
    case 6: return launch_nt<6, DENSE>(prm, B, s, err);
    case 8: return launch_nt<8, DENSE>(prm, B, s, err);
    default: return launch_nt<10, DENSE>(prm, B, s, err);
  }
}

int launch(int NT, const WParams& prm, int B, hipStream_t s, std::string& err) {
  return prm.log_emit ? launch_d<true>(NT, prm, B, s, err) : launch_d<false>(NT, prm, B, s, err);
}

}  // namespace tkfmi_wide

using namespace tkfmi_wide;

int tkfmi_wide_prep(const tkfmi_problem* p, int NB, int NBP, int e_idx, const int32_t* d_bidx, double* pm,
                    int G, double* tab, hipStream_t stream, std::string& err) {
  hipLaunchKernelGGL(wide_prep_kernel, dim3(p->n_trans), dim3(256), 0, stream, p->log_trans, p->S, NB, NBP,
                     e_idx, d_bidx, pm);
  if (tab && p->log_tab) {
    const int ntab = p->n_tab * G * TAB_SZ;
    hipLaunchKernelGGL(wide_tab_kernel, dim3(std::min(1024, (ntab + 255) / 256)), dim3(256), 0, stream,
                       p->log_tab, ntab, tab);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("wide prep launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  return TKFMI_OK;
}

bool tkfmi_wide_applies(const tkfmi_problem* p) {
  WPlan pl;
  std::string why;
  return applies(p, pl, why);
}

size_t tkfmi_wide_workspace(const tkfmi_problem* p, uint32_t flags) {
  WPlan pl;
  std::string why;
  if (!plan(p, flags, pl, why)) return 0;
  return std::max<size_t>(pl.total(), 1);
}

int tkfmi_wide_run(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                   hipStream_t stream, uint32_t flags, std::string& err) {
  WPlan pl;
  if (!plan(p, flags, pl, err)) return TKFMI_EINVAL;
  if (!ws || ws_bytes < pl.total()) { err = "workspace too small: need " + std::to_string(pl.total()); return TKFMI_EWORKSPACE; }
  const int B = p->B;
  if (B == 0) return TKFMI_OK;
  char* wsb = static_cast<char*>(ws);
  // descriptors and state slot tables in one upload
  std::vector<char> blob(pl.idx_b + pl.desc_b, 0);
  memcpy(blob.data(), pl.bidx.data(), pl.NBP * 4);
  memcpy(blob.data() + pl.NBP * 4, pl.btype.data(), pl.NBP * 4);
  memcpy(blob.data() + 2 * pl.NBP * 4, pl.bgrp.data(), pl.NBP * 4);
  memcpy(blob.data() + pl.idx_b, pl.desc.data(), sizeof(WDesc) * B);
  int rc = (flags & TKFMI_REUSE_DESCRIPTORS) ? TKFMI_OK : tkfmi_stage_upload(wsb, blob.data(), blob.size(), stream, err);
  if (rc != TKFMI_OK) return rc;
  const int32_t* d_idx = reinterpret_cast<const int32_t*>(wsb);
  double* wsd = reinterpret_cast<double*>(wsb);
  hipLaunchKernelGGL(wide_prep_kernel, dim3(p->n_trans), dim3(256), 0, stream, p->log_trans, p->S, pl.NB,
                     pl.NBP, pl.e_idx, d_idx, wsd + (pl.idx_b + pl.desc_b) / 8);
  if (!p->log_emit) {
    const int ntab = p->n_tab * pl.G * TAB_SZ;
    hipLaunchKernelGGL(wide_tab_kernel, dim3(std::min(1024, (ntab + 255) / 256)), dim3(256), 0, stream,
                       p->log_tab, ntab, wsd + (pl.idx_b + pl.desc_b + pl.pm_b) / 8);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("wide prep launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  WParams prm;
  memset(&prm, 0, sizeof(prm));
  prm.S = p->S; prm.NB = pl.NB; prm.NBP = pl.NBP; prm.flags = (int32_t)flags;
  prm.e_idx = pl.e_idx; prm.G = pl.G;
  prm.bidx = d_idx; prm.btype = d_idx + pl.NBP; prm.bgrp = d_idx + 2 * pl.NBP;
  prm.x_cat = p->x_cat; prm.y_cat = p->y_cat; prm.log_emit = p->log_emit;
  prm.desc = reinterpret_cast<const WDesc*>(wsb + pl.idx_b);
  prm.ws = wsd;
  prm.out_logP = o->log_prob;
  prm.out_n = (flags & TKFMI_NO_COUNTS) ? nullptr : o->n_trans;
  prm.out_post = o->posteriors;
  prm.out_ecnt = (flags & TKFMI_WANT_EMIT_COUNTS) ? o->emit_counts : nullptr;
  int64_t cells = 0;
  for (int b = 0; b < B; ++b) cells += (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1);
  const std::string name = "wide_kernel<" + std::to_string(pl.NT) + ">";
  const int slot = tkfmi_ktime_start(name.c_str(), B, cells, stream, err);
  if (slot == -2) return TKFMI_EHIP;
  rc = launch(pl.NT, prm, B, stream, err);
  if (rc != TKFMI_OK) return rc;
  if (tkfmi_ktime_stop(slot, stream, err) != TKFMI_OK) return TKFMI_EHIP;
  return TKFMI_OK;
}
