// tkfmi: banded forward/backward + expected counts for wide state spaces
// (16 to 160 emitting states, e.g. MixDom K=8 F=4: 2 + 160 states; the
// reference's trained K=3..8, F=2 models: 32..82 states) on gfx950.
// Reference algorithm: python/tkfmixdom/jax/dp/hmm.py (forward :1186-1292 and
// the row-scan variant :1440-1533 used for S > 50, backward :1295-1437 /
// :1536-1692, counts :1930-1971).
//
// Design:
//  * A pair's lattice rows are cut into bands of R = 16 rows; each band is
//    one workgroup of 8 waves, so a long pair spreads over many compute units
//    instead of running on one.  Band g needs the last row of band g - 1
//    (forward) or the first row of band g + 1 (backward): it receives those
//    rows through a boundary buffer in global memory, written with
//    device-scope stores and published every CQ columns with a progress
//    counter (bounded waits; workgroups take tickets in dependency order, so a
//    wait is only ever on a workgroup that is already running).
//  * Inside a band, step t handles cell (row r, column t - r) of every row:
//    the 16 cells of a step are independent.  Per step:
//      A (elementwise, one wave per cell, states over the lanes): the cell's
//        state vector from the values its predecessors pushed to it, times
//        the emissions, normalised to one binary exponent per cell; written to
//        the HBM lattice (forward) and to LDS;
//      B (matrix cores): the transition contraction of all 16 cells as one
//        (16 x NBP) . (NBP x NBP) product with v_mfma_f64_16x16x4f64, P read
//        from L2; each output value is pushed straight to the one successor
//        that uses it (M-type states to (r+1, j+1), I-type to (r, j+1), D-type
//        to (r+1, j)), into a two-slot LDS ring -- or, for the band's last
//        row, to the boundary buffer;
//      backward: B = g P^T the same way, the expected counts sum_c F_c g_c^T
//        as a second product (K = the step's cells) into 16 x 16 register
//        tiles held for the whole sweep, then a third elementwise phase
//        (posteriors, emission counts, q = e . B pushed to the predecessors).
//    Only the lattice and the boundary rows go to HBM; the per-step buffers
//    stay in LDS.
//  * Probability space with one binary exponent per cell (as the other
//    kernels): no exp/log in the inner loops.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fb2d_internal.h"
#include "tkfmi.h"

namespace tkfmi_band {

constexpr int ST_M = 1, ST_I = 2, ST_D = 3;
constexpr int A_PAD = TKFMI_A_PAD;
constexpr int TAB_SZ = TKFMI_TAB_SZ;
constexpr int EXP_NONE = -(1 << 29);
constexpr int kWaves = 8;
constexpr int kThr = kWaves * 64;
constexpr int R = 16;                  // rows per band
constexpr int CQ = 16;                 // boundary columns per published chunk
constexpr int kSpinMax = 1 << 22;      // bounded waits (each spin sleeps ~0.2 us + a load)
constexpr double LN2 = 0.69314718055994530942;

typedef double d4 __attribute__((ext_vector_type(4)));

struct BDesc {
  int32_t pair, Lx, Ly, i0;    // pair geometry; first row of the band
  int32_t nr, g, G, T;         // rows in the band, band index, bands of the pair, steps
  int32_t fin, fout;           // progress counters in / out (ints into the flag region; -1: none)
  int32_t ain, aout;           // backward: count-accumulator flags in (from band g+1) / out (to g-1)
  int64_t bin_off, bout_off;   // doubles into ws: boundary rows in / out ((Ly+1) x (NBP+1); -1: none)
  int64_t lat_off;             // doubles into ws: the pair's lattice [(Lx+1)(Ly+1)][NBP], then exponents
  int64_t pm_off, tab_off;     // doubles into ws: transition block, probability-space tables
  int64_t x_off, y_off, post_off;
  int64_t acc_off;             // doubles into ws: the pair's count accumulator [NBP][NBP]
};

struct BParams {
  int32_t S, NB, NBP, flags;
  int32_t e_idx, G, spin_max, pad;
  const int32_t* bidx;   // [NBP] state index of emitting slot k (padding: 0)
  const int32_t* btype;  // [NBP] type (0 for padding)
  const int32_t* bgrp;   // [NBP] emission group
  const uint8_t* x_cat;
  const uint8_t* y_cat;
  const double* log_emit;  // dense mode: pair b's [i][j][state] log table at desc.post_off
  const BDesc* desc;
  double* ws;
  int32_t* flagv;   // boundary progress counters
  int32_t* ticket;  // this launch's ticket counter
  int32_t* abortf;  // raised by a timed-out wait
  double* pinfo;    // [B][2]: mantissa and binary exponent of P(x, y)
  double* out_logP;
  double* out_n;
  double* out_post;
  double* out_ecnt;
};

__host__ __device__ inline int64_t pm_stride(int NBP) { return 2 * (int64_t)NBP * NBP + 2 * NBP + 8; }

__device__ inline double wave_max(double v) {
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double wave_sum(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline int frexp_exp(double v) { return __builtin_amdgcn_frexp_exp(v); }

__device__ inline double emit(const double* tab, int ty, int g, int xi, int yj, int i, int j) {
  const double* t = tab + (size_t)g * TAB_SZ;
  if (ty == ST_M) return (i >= 1 && j >= 1) ? t[xi * A_PAD + yj] : 0.0;
  if (ty == ST_I) return (j >= 1) ? t[A_PAD * A_PAD + yj] : 0.0;
  if (ty == ST_D) return (i >= 1) ? t[A_PAD * A_PAD + A_PAD + xi] : 0.0;
  return 0.0;
}
__device__ inline int emit_entry(int ty, int xi, int yj) {
  return ty == ST_M ? xi * A_PAD + yj : (ty == ST_I ? A_PAD * A_PAD + yj : A_PAD * A_PAD + A_PAD + xi);
}

// ---- cross-workgroup boundary (as in the chained 17-state kernels): entries
// are device-scope relaxed atomics, performed at the device's coherence point,
// so publishing only waits for this wave's entry stores to complete
__device__ __forceinline__ void bnd_store(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double bnd_load(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish(int* flag, int count) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __hip_atomic_store(flag, count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits (whole wave) until *flag >= need; false if the launch was aborted.
__device__ bool wait_ge(int* flag, int need, int* abortf, int spin_max) {
  int v = __builtin_amdgcn_readfirstlane(__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
  int spins = 0;
  while (v < need) {
    if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
      return false;
    if (++spins > spin_max) {
      __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(8);
    v = __builtin_amdgcn_readfirstlane(__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
  }
  return true;
}

// The launch's next ticket, the same for the whole workgroup; a ticket past
// the grid (a counter that was not zeroed) raises the abort flag and ends the
// waves without touching a descriptor.
__device__ __forceinline__ int take_ticket(int32_t* counter, int32_t* abortf) {
  __shared__ int s_ticket;
  if (threadIdx.x == 0) s_ticket = atomicAdd(counter, 1);
  __syncthreads();
  const int t = __builtin_amdgcn_readfirstlane(s_ticket);
  if (t < 0 || t >= (int)gridDim.x) {
    if (threadIdx.x == 0) __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_endpgm();
  }
  return t;
}

// out[16 x NBP tile nt of rows 0..15] = A (LDS, rows stride NBPP) . Bm (global
// NBP x NBP, row-major) for column tile nt: acc[v] = C[4 v + lane/16][16 nt +
// lane%16].  B fragments are loaded one batch of 4 k-chunks ahead.
template <int NT>
__device__ __forceinline__ d4 tile_gemm(const double* __restrict__ sA, const double* __restrict__ Bm, int nt,
                                        int lane) {
  constexpr int NBP = 16 * NT, NBPP = NBP + 1, KC = NBP / 4;
  const double* ap = sA + (lane & 15) * NBPP + (lane >> 4);
  const double* bp = Bm + (size_t)(lane >> 4) * NBP + 16 * nt + (lane & 15);
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  double bn[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) bn[q] = bp[(size_t)4 * q * NBP];
#pragma unroll 1
  for (int kb = 0; kb < KC; kb += 4) {
    double b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) b[q] = bn[q];
    if (kb + 4 < KC) {
#pragma unroll
      for (int q = 0; q < 4; ++q) bn[q] = bp[(size_t)4 * (kb + 4 + q) * NBP];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ap[4 * (kb + q)], b[q], acc, 0, 0, 0);
  }
  return acc;
}

// ====================================================================== forward
template <int NT, bool DENSE>
__global__ __launch_bounds__(kThr) void band_fwd_kernel(BParams Pr) {
  constexpr int NBP = 16 * NT, NBPP = NBP + 1, KR = (NBP + 63) / 64;
  static_assert(NBP % 16 == 0 && (NBP / 4) % 4 == 0, "k-chunks in batches of 4");
  __shared__ double s_pre[2][R][NBPP];  // values pushed to the cells of steps t (slot t&1), t+1
  __shared__ double s_F[R][NBPP];       // this step's normalised forward values (GEMM operand)
  __shared__ double s_last[NBPP];       // the last row's pushes for the band below, and its exponent
  __shared__ int s_eM[2][R], s_eI[2][R], s_eD[2][R], s_E[R];
  __shared__ int s_ok;                  // cleared when a boundary wait gave up
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const BDesc pd = Pr.desc[take_ticket(Pr.ticket, Pr.abortf)];
  const int Lx = pd.Lx, Ly = pd.Ly, W1 = Ly + 1, NB = Pr.NB, S = Pr.S, nr = pd.nr, i0 = pd.i0;
  const double* pm = Pr.ws + pd.pm_off;
  const double* P = pm;
  const double* p_start = pm + 2 * (size_t)NBP * NBP;
  const double* p_end = p_start + NBP;
  const double p_se = p_start[2 * NBP];
  const double* tab = DENSE ? nullptr : Pr.ws + pd.tab_off;
  const double* dem = DENSE ? Pr.log_emit + pd.post_off : nullptr;
  const uint8_t* xs = DENSE ? nullptr : Pr.x_cat + pd.x_off;
  const uint8_t* ys = DENSE ? nullptr : Pr.y_cat + pd.y_off;
  const bool want_lat = !(Pr.flags & TKFMI_FORWARD_ONLY);
  const bool want_counts = want_lat && !(Pr.flags & TKFMI_NO_COUNTS) && Pr.out_n != nullptr;
  double* lat = Pr.ws + pd.lat_off;
  double* latE = lat + (size_t)(Lx + 1) * W1 * NBP;
  const double* bin = pd.bin_off >= 0 ? Pr.ws + pd.bin_off : nullptr;
  double* bout = pd.bout_off >= 0 ? Pr.ws + pd.bout_off : nullptr;
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  int ty[KR], gr[KR], sx[KR];
#pragma unroll
  for (int q = 0; q < KR; ++q) {
    const int k = lane + 64 * q;
    ty[q] = (k < NB) ? Pr.btype[k] : 0;
    gr[q] = (k < NB) ? Pr.bgrp[k] : 0;
    sx[q] = (k < NB) ? Pr.bidx[k] : 0;
  }
  for (int q = threadIdx.x; q < 2 * R * NBPP; q += kThr) (&s_pre[0][0][0])[q] = 0.0;
  for (int q = threadIdx.x; q < R * NBPP; q += kThr) (&s_F[0][0])[q] = 0.0;
  for (int q = threadIdx.x; q < NBPP; q += kThr) s_last[q] = 0.0;
  if (threadIdx.x == 0) s_ok = 1;
  if (threadIdx.x < 2 * R) {
    (&s_eM[0][0])[threadIdx.x] = EXP_NONE;
    (&s_eI[0][0])[threadIdx.x] = EXP_NONE;
    (&s_eD[0][0])[threadIdx.x] = EXP_NONE;
  }
  __syncthreads();
  bool ok_run = true;
  for (int t = 0; t < pd.T; ++t) {
    const int sl = t & 1, sl1 = sl ^ 1;
    ok_run = ok_run && s_ok;
    // ---- the last row's pushes of the previous step -> boundary (wave 7);
    // a band whose input wait gave up stops publishing, so that the bands
    // below it give up too instead of reading values computed from garbage
    if (bout && w == kWaves - 1 && ok_run) {
      const int jp = t - 1 - (nr - 1);
      if (jp >= 0 && jp <= Ly) {
        double* e = bout + (size_t)jp * (NBP + 1);
        for (int k = lane; k <= NBP; k += 64) bnd_store(e + k, s_last[k]);
        if (((jp + 1) % CQ) == 0 || jp == Ly) publish(Pr.flagv + pd.fout, jp + 1);
      }
    }
    // ---- A: the cells of this step
#pragma unroll 1
    for (int r = w; r < R; r += kWaves) {
      const int j = t - r, i = i0 + r;
      if (r >= nr || j < 0 || j > Ly) {
        for (int k = lane; k < NBP; k += 64) s_F[r][k] = 0.0;
        continue;
      }
      const bool top = (r == 0) && bin != nullptr;
      if (top && (j % CQ) == 0 && ok_run) {
        ok_run = wait_ge(Pr.flagv + pd.fin, min(j + CQ, Ly + 1), Pr.abortf, Pr.spin_max);
        if (!ok_run && lane == 0) s_ok = 0;
      }
      const bool okM = i >= 1 && j >= 1, okI = j >= 1, okD = i >= 1;
      // predecessor exponents: M from (i-1, j-1), I from (i, j-1), D from (i-1, j)
      int eM = EXP_NONE, eI = EXP_NONE, eD = EXP_NONE;
      if (okM) eM = top ? (int)bnd_load(bin + (size_t)(j - 1) * (NBP + 1) + NBP) : s_eM[sl][r];
      if (okI) eI = s_eI[sl][r];
      if (okD) eD = top ? (int)bnd_load(bin + (size_t)j * (NBP + 1) + NBP) : s_eD[sl][r];
      const int xi = (!DENSE && i >= 1) ? (int)min((int)xs[i - 1], A_PAD - 1) : 0;
      const int yj = (!DENSE && j >= 1) ? (int)min((int)ys[j - 1], A_PAD - 1) : 0;
      double raw[KR];
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = lane + 64 * q;
        double v = 0.0;
        if (ty[q] == ST_M && okM) v = top ? bnd_load(bin + (size_t)(j - 1) * (NBP + 1) + k) : s_pre[sl][r][k];
        else if (ty[q] == ST_I && okI) v = s_pre[sl][r][k];
        else if (ty[q] == ST_D && okD) v = top ? bnd_load(bin + (size_t)j * (NBP + 1) + k) : s_pre[sl][r][k];
        if (v != 0.0) {
          double e;
          if constexpr (DENSE) e = exp(dem[((size_t)i * W1 + j) * S + sx[q]]);
          else e = emit(tab, ty[q], gr[q], xi, yj, i, j);
          v *= e;
        }
        raw[q] = v;
      }
      double mM = 0.0, mI = 0.0, mD = 0.0;
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        if (ty[q] == ST_M) mM = fmax(mM, raw[q]);
        else if (ty[q] == ST_I) mI = fmax(mI, raw[q]);
        else if (ty[q] == ST_D) mD = fmax(mD, raw[q]);
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
      for (int q = 0; q < KR; ++q) {
        const int sh = (ty[q] == ST_M) ? eM - E : ((ty[q] == ST_I) ? eI - E : eD - E);
        F[q] = (origin || raw[q] == 0.0) ? 0.0 : ldexp(raw[q], sh);
      }
      if (origin) E = 0;
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = lane + 64 * q;
        if (k < NBP) {
          s_F[r][k] = F[q];
          if (want_lat) lat[((size_t)i * W1 + j) * NBP + k] = F[q];
        }
      }
      if (lane == 0) {
        s_E[r] = E;
        if (want_lat) latE[(size_t)i * W1 + j] = (double)E;
      }
      if (i == Lx && j == Ly) {  // log P and the E column
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < KR; ++q) {
          const int k = lane + 64 * q;
          if (k < NB) part = fma(F[q], p_end[k], part);
        }
        double tot = wave_sum(part);
        double lp;
        if (origin) { tot = p_se; lp = log(p_se); }
        else lp = log(tot) + (double)E * LN2;
        if (!ok_run) { tot = 0.0; lp = __builtin_nan(""); }
        if (lane == 0) {
          Pr.out_logP[pd.pair] = (tot > 0.0 || !ok_run) ? lp : -1e30;
          Pr.pinfo[2 * pd.pair] = (tot > 0.0) ? ldexp(tot, -frexp_exp(tot)) : 0.0;
          Pr.pinfo[2 * pd.pair + 1] = (tot > 0.0) ? (double)(E + frexp_exp(tot)) : 0.0;
        }
        if (want_counts && tot > 0.0) {
          if (origin) { if (lane == 0) out_n[Pr.e_idx] = 1.0; }
          else {
#pragma unroll
            for (int q = 0; q < KR; ++q) {
              const int k = lane + 64 * q;
              if (k < NB) out_n[Pr.bidx[k] * S + Pr.e_idx] = F[q] * p_end[k] / tot;
            }
          }
        }
      }
    }
    __syncthreads();
    // ---- B: h = F P for the step's cells, pushed to the successors
    const bool origin_step = (i0 == 0 && t == 0);
#pragma unroll 1
    for (int nt = w; nt < NT; nt += kWaves) {
      const d4 acc = tile_gemm<NT>(&s_F[0][0], P, nt, lane);
      const int k = 16 * nt + (lane & 15);
      const int tk = (k < NB) ? Pr.btype[k] : 0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 4 * v + (lane >> 4);
        const int j = t - r;
        if (r >= nr || j < 0 || j > Ly || tk == 0) continue;
        const double h = (origin_step && r == 0) ? p_start[k] : acc[v];
        if (tk == ST_I) s_pre[sl1][r][k] = h;
        else if (r + 1 < nr) s_pre[tk == ST_M ? sl : sl1][r + 1][k] = h;
        else s_last[k] = h;
      }
    }
    if (w == 0 && lane < nr) {  // the pushed values' exponents
      const int r = lane, j = t - r;
      if (j >= 0 && j <= Ly) {
        const int E = s_E[r];
        s_eI[sl1][r] = E;
        if (r + 1 < nr) {
          s_eD[sl1][r + 1] = E;
          s_eM[sl][r + 1] = E;
        } else {
          s_last[NBP] = (double)E;
        }
      }
    }
    __syncthreads();
  }
  // ---- the last pushes -> boundary
  ok_run = ok_run && s_ok;
  if (bout && w == kWaves - 1 && ok_run) {
    const int jp = pd.T - 1 - (nr - 1);
    if (jp >= 0 && jp <= Ly) {
      double* e = bout + (size_t)jp * (NBP + 1);
      for (int k = lane; k <= NBP; k += 64) bnd_store(e + k, s_last[k]);
      publish(Pr.flagv + pd.fout, jp + 1);
    }
  }
}

// ===================================================================== backward
// Count tiles (NT x NT blocks of 16 x 16) dealt out to the waves: tile
// w + kWaves q to wave w.
__host__ __device__ constexpr int count_ntile(int NT) { return (NT * NT + kWaves - 1) / kWaves; }

template <int NT, bool DENSE>
__global__ __launch_bounds__(kThr) void band_bwd_kernel(BParams Pr) {
  constexpr int NBP = 16 * NT, NBPP = NBP + 1, KR = (NBP + 63) / 64, NTILE = count_ntile(NT);
  __shared__ double s_pre[2][R][NBPP];  // q values pushed to the cells of steps t, t+1
  __shared__ double s_G[R][NBPP];       // g of this step's cells (successor-pulled q)
  __shared__ double s_FS[R][NBPP];      // their scaled forward values
  __shared__ double s_B[R][NBPP];       // B = g P^T
  __shared__ double s_last[NBPP];       // the first row's pushes for the band above, and its exponent
  __shared__ int s_eM[2][R], s_eI[2][R], s_eD[2][R], s_Eg[R];
  __shared__ int s_ok;                  // cleared when a boundary wait gave up
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const BDesc pd = Pr.desc[take_ticket(Pr.ticket, Pr.abortf)];
  const int Lx = pd.Lx, Ly = pd.Ly, W1 = Ly + 1, NB = Pr.NB, S = Pr.S, nr = pd.nr, i0 = pd.i0;
  const double* pm = Pr.ws + pd.pm_off;
  const double* P = pm;
  const double* PT = pm + (size_t)NBP * NBP;
  const double* p_start = pm + 2 * (size_t)NBP * NBP;
  const double* p_end = p_start + NBP;
  const double* tab = DENSE ? nullptr : Pr.ws + pd.tab_off;
  const double* dem = DENSE ? Pr.log_emit + pd.post_off : nullptr;
  const uint8_t* xs = DENSE ? nullptr : Pr.x_cat + pd.x_off;
  const uint8_t* ys = DENSE ? nullptr : Pr.y_cat + pd.y_off;
  const bool want_counts = !(Pr.flags & TKFMI_NO_COUNTS) && Pr.out_n != nullptr;
  const bool want_post = (Pr.flags & TKFMI_WANT_POSTERIORS) && Pr.out_post != nullptr;
  const bool want_ecnt = !DENSE && (Pr.flags & TKFMI_WANT_EMIT_COUNTS) && Pr.out_ecnt != nullptr;
  const double* lat = Pr.ws + pd.lat_off;
  const double* latE = lat + (size_t)(Lx + 1) * W1 * NBP;
  const double* bin = pd.bin_off >= 0 ? Pr.ws + pd.bin_off : nullptr;
  double* bout = pd.bout_off >= 0 ? Pr.ws + pd.bout_off : nullptr;
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  double* post = want_post ? Pr.out_post + pd.post_off : nullptr;
  double* ecnt = want_ecnt ? Pr.out_ecnt + (size_t)pd.pair * Pr.G * TAB_SZ : nullptr;
  const double mP = Pr.pinfo[2 * pd.pair];
  const int EP = (int)Pr.pinfo[2 * pd.pair + 1];
  if (!(mP > 0.0)) return;  // impossible or aborted pair: every band of it returns here
  const double inv_mP = 1.0 / mP;
  int ty[KR], gr[KR], sx[KR];
#pragma unroll
  for (int q = 0; q < KR; ++q) {
    const int k = lane + 64 * q;
    ty[q] = (k < NB) ? Pr.btype[k] : 0;
    gr[q] = (k < NB) ? Pr.bgrp[k] : 0;
    sx[q] = (k < NB) ? Pr.bidx[k] : 0;
  }
  for (int q = threadIdx.x; q < 2 * R * NBPP; q += kThr) (&s_pre[0][0][0])[q] = 0.0;
  for (int q = threadIdx.x; q < R * NBPP; q += kThr) {
    (&s_G[0][0])[q] = 0.0;
    (&s_FS[0][0])[q] = 0.0;
  }
  for (int q = threadIdx.x; q < NBPP; q += kThr) s_last[q] = 0.0;
  if (threadIdx.x == 0) s_ok = 1;
  if (threadIdx.x < 2 * R) {
    (&s_eM[0][0])[threadIdx.x] = EXP_NONE;
    (&s_eI[0][0])[threadIdx.x] = EXP_NONE;
    (&s_eD[0][0])[threadIdx.x] = EXP_NONE;
  }
  d4 cacc[NTILE];
#pragma unroll
  for (int q = 0; q < NTILE; ++q) cacc[q] = d4{0.0, 0.0, 0.0, 0.0};
  __syncthreads();
  constexpr int CPW = R / kWaves;  // cells per wave and step
  bool ok_run = true;
  // Reversed coordinates: rr = nr - 1 - r (rr = 0 is the band's last row),
  // jj = Ly - j; step t handles (rr, jj = t - rr).  Successors of (i, j):
  // M (i+1, j+1) = (rr-1, jj-1), I (i, j+1) = (rr, jj-1), D (i+1, j) = (rr-1, jj).
  for (int t = 0; t < pd.T; ++t) {
    const int sl = t & 1, sl1 = sl ^ 1;
    ok_run = ok_run && s_ok;
    // ---- the first row's pushes of the previous step -> boundary (wave 7)
    if (bout && w == kWaves - 1 && ok_run) {
      const int jjp = t - 1 - (nr - 1);
      if (jjp >= 0 && jjp <= Ly) {
        double* e = bout + (size_t)(Ly - jjp) * (NBP + 1);
        for (int k = lane; k <= NBP; k += 64) bnd_store(e + k, s_last[k]);
        if (((jjp + 1) % CQ) == 0 || jjp == Ly) publish(Pr.flagv + pd.fout, jjp + 1);
      }
    }
    // ---- A: g and the scaled forward values of this step's cells
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const int rr = w + kWaves * c;
      const int jj = t - rr, r = nr - 1 - rr, i = i0 + r, j = Ly - jj;
      if (rr >= nr || jj < 0 || jj > Ly) {
        for (int k = lane; k < NBP; k += 64) { s_G[rr][k] = 0.0; s_FS[rr][k] = 0.0; }
        continue;
      }
      const bool bot = (rr == 0) && bin != nullptr;
      if (bot && (jj % CQ) == 0 && ok_run) {
        ok_run = wait_ge(Pr.flagv + pd.fin, min(jj + CQ, Ly + 1), Pr.abortf, Pr.spin_max);
        if (!ok_run && lane == 0) s_ok = 0;
      }
      const bool okM = i < Lx && j < Ly, okI = j < Ly, okD = i < Lx;
      int eM = EXP_NONE, eI = EXP_NONE, eD = EXP_NONE;
      if (okM) eM = bot ? (int)bnd_load(bin + (size_t)(j + 1) * (NBP + 1) + NBP) : s_eM[sl][rr];
      if (okI) eI = s_eI[sl][rr];
      if (okD) eD = bot ? (int)bnd_load(bin + (size_t)j * (NBP + 1) + NBP) : s_eD[sl][rr];
      int Eg = max(eM, max(eI, eD));
      if (Eg == EXP_NONE) Eg = 0;
      const size_t cell = (size_t)i * W1 + j;
      const int EF = (int)latE[cell];
      const double fsc = ldexp(inv_mP, EF + Eg - EP);
      double g[KR];
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = lane + 64 * q;
        double v = 0.0;
        if (ty[q] == ST_M && okM)
          v = ldexp(bot ? bnd_load(bin + (size_t)(j + 1) * (NBP + 1) + k) : s_pre[sl][rr][k], eM - Eg);
        else if (ty[q] == ST_I && okI) v = ldexp(s_pre[sl][rr][k], eI - Eg);
        else if (ty[q] == ST_D && okD)
          v = ldexp(bot ? bnd_load(bin + (size_t)j * (NBP + 1) + k) : s_pre[sl][rr][k], eD - Eg);
        g[q] = v;
        if (k < NBP) {
          s_G[rr][k] = v;
          s_FS[rr][k] = lat[cell * NBP + k] * fsc;
        }
      }
      if (lane == 0) s_Eg[rr] = Eg;
      if (i == 0 && j == 0) {  // start state: S -> body counts, posterior of S at (0, 0)
        const double scale = ldexp(inv_mP, Eg - EP);
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < KR; ++q) {
          const int k = lane + 64 * q;
          if (k < NB) {
            part = fma(p_start[k], g[q], part);
            if (want_counts) out_n[Pr.bidx[k]] = p_start[k] * g[q] * scale;
          }
        }
        part = wave_sum(part);
        if (want_post && lane == 0) post[0] = (i == Lx && j == Ly) ? 1.0 : part * scale;
      }
    }
    __syncthreads();
    // ---- B: B = g P^T, counts += Fs^T g
#pragma unroll 1
    for (int nt = w; nt < NT; nt += kWaves) {
      const d4 bacc = tile_gemm<NT>(&s_G[0][0], PT, nt, lane);
#pragma unroll
      for (int v = 0; v < 4; ++v) s_B[4 * v + (lane >> 4)][16 * nt + (lane & 15)] = bacc[v];
    }
    if (want_counts) {
#pragma unroll
      for (int kc = 0; kc < R / 4; ++kc) {
        const double* fr = &s_FS[4 * kc + (lane >> 4)][lane & 15];
        const double* gr_ = &s_G[4 * kc + (lane >> 4)][lane & 15];
#pragma unroll
        for (int q = 0; q < NTILE; ++q) {
          const int tile = w + kWaves * q;
          if (tile < NT * NT)
            cacc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(fr[16 * (tile / NT)], gr_[16 * (tile % NT)], cacc[q], 0, 0, 0);
        }
      }
    }
    __syncthreads();
    // ---- C: normalise B, posteriors, emission counts, q = e . B to the predecessors
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const int rr = w + kWaves * c;
      const int jj = t - rr, r = nr - 1 - rr, i = i0 + r, j = Ly - jj;
      if (rr >= nr || jj < 0 || jj > Ly) continue;
      const bool corner = (i == Lx && j == Ly);
      double Bv[KR];
      double mx = 0.0;
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = lane + 64 * q;
        Bv[q] = (k < NBP) ? (corner ? (k < NB ? p_end[k] : 0.0) : s_B[rr][k]) : 0.0;
        mx = fmax(mx, Bv[q]);
      }
      mx = wave_max(mx);
      const int sh = (mx > 0.0) ? frexp_exp(mx) : 0;
      const int EB = (corner ? 0 : s_Eg[rr]) + sh;
      const int xi = (!DENSE && i >= 1) ? (int)min((int)xs[i - 1], A_PAD - 1) : 0;
      const int yj = (!DENSE && j >= 1) ? (int)min((int)ys[j - 1], A_PAD - 1) : 0;
      const size_t cellS = ((size_t)i * W1 + j) * S;
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = lane + 64 * q;
        if (k >= NBP) continue;
        const double b = ldexp(Bv[q], -sh);
        double e = 0.0;
        if (k < NB) {
          if constexpr (DENSE) {
            const bool ok = (ty[q] == ST_M) ? (i >= 1 && j >= 1) : (ty[q] == ST_I) ? (j >= 1) : (i >= 1);
            e = ok ? exp(dem[cellS + sx[q]]) : 0.0;
          } else {
            e = emit(tab, ty[q], gr[q], xi, yj, i, j);
          }
        }
        const double qv = e * b;
        if (k < NB && (want_post || want_ecnt)) {
          // F B / P(x, y) = Fs (scaled by 2^(EF + Eg - EP) / mP) times B
          // before its normalisation (relative to 2^Eg; the corner: Eg = 0)
          const double pv = s_FS[rr][k] * Bv[q];
          if (want_post) post[cellS + sx[q]] = pv;
          if (want_ecnt && pv != 0.0 && e != 0.0)
            atomicAdd(ecnt + (size_t)gr[q] * TAB_SZ + emit_entry(ty[q], xi, yj), pv);
        }
        // push q: M-type to (i-1, j-1), I-type to (i, j-1), D-type to (i-1, j)
        if (ty[q] == ST_I) s_pre[sl1][rr][k] = qv;
        else if (ty[q] == ST_M || ty[q] == ST_D) {
          if (rr + 1 < nr) s_pre[ty[q] == ST_M ? sl : sl1][rr + 1][k] = qv;
          else s_last[k] = qv;
        }
      }
      if (lane == 0) {
        s_eI[sl1][rr] = EB;
        if (rr + 1 < nr) {
          s_eD[sl1][rr + 1] = EB;
          s_eM[sl][rr + 1] = EB;
        } else {
          s_last[NBP] = (double)EB;
        }
      }
    }
    __syncthreads();
  }
  // ---- the last pushes -> boundary
  ok_run = ok_run && s_ok;
  if (bout && w == kWaves - 1 && ok_run) {
    const int jjp = pd.T - 1 - (nr - 1);
    if (jjp >= 0 && jjp <= Ly) {
      double* e = bout + (size_t)(Ly - jjp) * (NBP + 1);
      for (int k = lane; k <= NBP; k += 64) bnd_store(e + k, s_last[k]);
      publish(Pr.flagv + pd.fout, jjp + 1);
    }
  }
  if (!want_counts) {
    if (!ok_run && threadIdx.x == 0) Pr.out_logP[pd.pair] = __builtin_nan("");
    return;
  }
  // ---- counts n[k, k2] = P[k, k2] sum_c Fs_c[k] g_c[k2], summed over the
  // pair's bands in a fixed order (band G-1 first, band 0 last, which is the
  // order in which they finish), so that results do not depend on timing:
  // band g adds the accumulator of bands > g to its own tiles and passes it
  // on; band 0 writes n
  double* accb = Pr.ws + pd.acc_off;
  if (pd.ain >= 0 && ok_run) {
    ok_run = wait_ge(Pr.flagv + pd.ain, 1, Pr.abortf, Pr.spin_max);
    if (!ok_run && lane == 0) s_ok = 0;
  }
  __syncthreads();
  ok_run = ok_run && s_ok;
  if (!ok_run) {  // counts and posteriors of this pair are not valid
    if (threadIdx.x == 0) Pr.out_logP[pd.pair] = __builtin_nan("");
    return;
  }
#pragma unroll
  for (int q = 0; q < NTILE; ++q) {
    const int tile = w + kWaves * q;
    if (tile >= NT * NT) break;
    const int kt = tile / NT, k2t = tile - kt * NT;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = 16 * kt + 4 * v + (lane >> 4), k2 = 16 * k2t + (lane & 15);
      double a = cacc[q][v];
      if (pd.ain >= 0) a += bnd_load(accb + (size_t)k * NBP + k2);
      if (pd.aout >= 0) {
        bnd_store(accb + (size_t)k * NBP + k2, a);
      } else if (k < NB && k2 < NB) {
        const double pr = P[(size_t)k * NBP + k2];
        out_n[Pr.bidx[k] * S + Pr.bidx[k2]] = (pr > 0.0) ? pr * a : 0.0;
      }
    }
  }
  if (pd.aout >= 0) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(Pr.flagv + pd.aout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// status[0] = 1 if the call's abort flag was raised.
__global__ void raise_status_kernel(const int32_t* __restrict__ abortf, int32_t* __restrict__ status) {
  if (threadIdx.x == 0 && __hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    status[0] = 1;
}

__global__ void zero_ints_kernel(int32_t* p, int n) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) p[q] = 0;
}

// ---------------------------------------------------------------- host side
struct BPlan {
  bool ok = false;
  std::string why;
  int S = 0, NB = 0, NT = 0, NBP = 0, e_idx = -1, G = 0;
  std::vector<int32_t> bidx, btype, bgrp;
  std::vector<BDesc> fwd, bwd;
  int n_flags = 0;  // ints: 2 tickets, abort, pad, then fwd / bwd / accumulator counters per boundary
  size_t idx_b = 0, desc_b = 0, pm_b = 0, tab_b = 0, pinfo_b = 0, flag_b = 0, data_b = 0;
  size_t total() const { return idx_b + desc_b + pm_b + tab_b + pinfo_b + flag_b + data_b; }
};

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

BPlan plan(const tkfmi_problem* p, uint32_t flags) {
  BPlan pl;
  if (!p || p->S < 2 || !p->state_types || !p->state_group) { pl.why = "bad problem"; return pl; }
  pl.S = p->S;
  for (int s = 0; s < p->S; ++s) {
    const int ty = p->state_types[s];
    if (ty == 4 && pl.e_idx < 0) pl.e_idx = s;
    if (ty == ST_M || ty == ST_I || ty == ST_D) {
      pl.bidx.push_back(s);
      pl.btype.push_back(ty);
      pl.bgrp.push_back(p->state_group[s]);
    }
  }
  pl.NB = (int)pl.bidx.size();
  if (pl.NB <= 15 || pl.NB > 160) { pl.why = "state count outside 16..160"; return pl; }
  if (p->state_types[0] != 0 || pl.e_idx < 0) { pl.why = "need state 0 of type S and a state of type E"; return pl; }
  const int nt = (pl.NB + 15) / 16;
  pl.NT = (nt <= 2) ? 2 : (nt <= 4 ? 4 : (nt <= 6 ? 6 : (nt <= 8 ? 8 : 10)));
  pl.NBP = 16 * pl.NT;
  while ((int)pl.bidx.size() < pl.NBP) { pl.bidx.push_back(0); pl.btype.push_back(0); pl.bgrp.push_back(0); }
  pl.G = p->n_groups;
  for (int k = 0; k < pl.NB; ++k)
    if (pl.bgrp[k] < 0 || pl.bgrp[k] >= pl.G) { pl.why = "state_group out of range"; return pl; }
  const bool dense = p->log_emit != nullptr;
  if (p->n_trans != 1 && p->n_trans != p->B) { pl.why = "n_trans must be 1 or B"; return pl; }
  if (!dense && p->n_tab != 1 && p->n_tab != p->B) { pl.why = "n_tab must be 1 or B"; return pl; }
  const int B = p->B;
  if (B > 0 && (!p->h_x_len || !p->h_y_len || !p->h_x_off || !p->h_y_off)) { pl.why = "host arrays missing"; return pl; }
  const bool fwd_only = (flags & TKFMI_FORWARD_ONLY) != 0;
  const int NBP = pl.NBP;
  pl.idx_b = align256(3 * (size_t)NBP * 4);
  pl.pm_b = align256((size_t)pm_stride(NBP) * 8 * p->n_trans);
  pl.tab_b = dense ? 0 : align256((size_t)p->n_tab * pl.G * TAB_SZ * 8);
  pl.pinfo_b = align256((size_t)std::max(B, 1) * 16);
  // bands, boundaries, lattice
  int64_t off = 0;  // doubles into the data region
  int nbnd = 0;
  std::vector<int64_t> post_off(B);
  int64_t ap = 0;
  for (int b = 0; b < B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    if (lx < 0 || ly < 0 || lx + ly > 64000) { pl.why = "sequence lengths out of range"; return pl; }
    post_off[b] = ap;
    ap += (int64_t)(lx + 1) * (ly + 1) * p->S;
  }
  const bool counts = !fwd_only && !(flags & TKFMI_NO_COUNTS);
  struct PairBands { int b, G; int64_t lat, bnd0, acc; int flag0; };
  std::vector<PairBands> pb(B);
  for (int b = 0; b < B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    const int G = (lx + R) / R;  // ceil((Lx + 1) / R)
    pb[b].b = b;
    pb[b].G = G;
    pb[b].lat = off;
    if (!fwd_only) off += (int64_t)(lx + 1) * (ly + 1) * (NBP + 1);
    pb[b].bnd0 = off;
    off += (int64_t)(G - 1) * (ly + 1) * (NBP + 1);
    pb[b].acc = off;
    if (counts && G > 1) off += (int64_t)NBP * NBP;
    pb[b].flag0 = 4 + 3 * nbnd;
    nbnd += G - 1;
  }
  pl.n_flags = 4 + 3 * nbnd;
  pl.flag_b = align256((size_t)pl.n_flags * 4);
  for (int b = 0; b < B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b], G = pb[b].G;
    for (int g = 0; g < G; ++g) {
      BDesc d;
      memset(&d, 0, sizeof(d));
      d.pair = b; d.Lx = lx; d.Ly = ly; d.i0 = g * R; d.nr = std::min(R, lx + 1 - g * R);
      d.g = g; d.G = G; d.T = ly + d.nr;
      d.lat_off = pb[b].lat;
      d.pm_off = (p->n_trans > 1 ? (int64_t)b : 0) * pm_stride(NBP);
      d.tab_off = (!dense && p->n_tab > 1 ? (int64_t)b : 0) * pl.G * TAB_SZ;
      d.x_off = p->h_x_off[b]; d.y_off = p->h_y_off[b];
      d.post_off = post_off[b];
      d.acc_off = pb[b].acc;
      const int64_t bsz = (int64_t)(ly + 1) * (NBP + 1);
      // forward: boundary g-1 in, g out; backward: boundary g in, g-1 out
      BDesc f = d, k = d;
      const int fl = pb[b].flag0;
      f.bin_off = g > 0 ? pb[b].bnd0 + (int64_t)(g - 1) * bsz : -1;
      f.fin = g > 0 ? fl + 3 * (g - 1) : -1;
      f.bout_off = g < G - 1 ? pb[b].bnd0 + (int64_t)g * bsz : -1;
      f.fout = g < G - 1 ? fl + 3 * g : -1;
      f.ain = f.aout = -1;
      k.bin_off = g < G - 1 ? pb[b].bnd0 + (int64_t)g * bsz : -1;
      k.fin = g < G - 1 ? fl + 3 * g + 1 : -1;
      k.bout_off = g > 0 ? pb[b].bnd0 + (int64_t)(g - 1) * bsz : -1;
      k.fout = g > 0 ? fl + 3 * (g - 1) + 1 : -1;
      k.ain = (counts && g < G - 1) ? fl + 3 * g + 2 : -1;
      k.aout = (counts && g > 0) ? fl + 3 * (g - 1) + 2 : -1;
      pl.fwd.push_back(f);
      pl.bwd.push_back(k);
    }
  }
  // Ticket order: by the step at which a band can start working (a band
  // starts lag steps after the band it reads from), then by the length of
  // the chain still ahead of it.  Producers get lower tickets than their
  // consumers, and a band is not resident long before its input arrives.
  const int64_t lag = R + CQ;
  std::stable_sort(pl.fwd.begin(), pl.fwd.end(), [&](const BDesc& a, const BDesc& c) {
    if (a.g != c.g) return a.g < c.g;
    return (int64_t)(a.G - 1 - a.g) * lag + a.T > (int64_t)(c.G - 1 - c.g) * lag + c.T;
  });
  std::stable_sort(pl.bwd.begin(), pl.bwd.end(), [&](const BDesc& a, const BDesc& c) {
    if (a.G - a.g != c.G - c.g) return a.G - a.g < c.G - c.g;
    return (int64_t)a.g * lag + a.T > (int64_t)c.g * lag + c.T;
  });
  pl.desc_b = align256(sizeof(BDesc) * std::max<size_t>(pl.fwd.size() + pl.bwd.size(), 1));
  pl.data_b = (size_t)off * 8;
  // offsets of the data region, relative to ws (in doubles)
  const int64_t data0 = (int64_t)(pl.idx_b + pl.desc_b + pl.pm_b + pl.tab_b + pl.pinfo_b + pl.flag_b) / 8;
  const int64_t pm0 = (int64_t)(pl.idx_b + pl.desc_b) / 8;
  const int64_t tab0 = (int64_t)(pl.idx_b + pl.desc_b + pl.pm_b) / 8;
  for (std::vector<BDesc>* v : {&pl.fwd, &pl.bwd})
    for (BDesc& d : *v) {
      d.lat_off += data0;
      if (d.bin_off >= 0) d.bin_off += data0;
      if (d.bout_off >= 0) d.bout_off += data0;
      d.acc_off += data0;
      d.pm_off += pm0;
      d.tab_off += tab0;
    }
  pl.ok = true;
  return pl;
}

template <int NT, bool DENSE>
int launch_nt(bool bwd, const BParams& prm, int n, hipStream_t s, std::string& err) {
  if (n == 0) return TKFMI_OK;
  if (bwd) hipLaunchKernelGGL((band_bwd_kernel<NT, DENSE>), dim3(n), dim3(kThr), 0, s, prm);
  else hipLaunchKernelGGL((band_fwd_kernel<NT, DENSE>), dim3(n), dim3(kThr), 0, s, prm);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("band launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  return TKFMI_OK;
}

template <bool DENSE>
int launch_d(int NT, bool bwd, const BParams& prm, int n, hipStream_t s, std::string& err) {
  switch (NT) {
    case 2: return launch_nt<2, DENSE>(bwd, prm, n, s, err);
    case 4: return launch_nt<4, DENSE>(bwd, prm, n, s, err);
    case 6: return launch_nt<6, DENSE>(bwd, prm, n, s, err);
    case 8: return launch_nt<8, DENSE>(bwd, prm, n, s, err);
    default: return launch_nt<10, DENSE>(bwd, prm, n, s, err);
  }
}

bool needs_checkpoint(const tkfmi_problem* p, uint32_t flags, int NBP) {
  if (flags & TKFMI_FORWARD_ONLY) return false;
  if (flags & TKFMI_CHECKPOINT) return true;
  const size_t budget = tkfmi_checkpoint_budget_bytes();
  if (budget == 0) return false;
  for (int b = 0; b < p->B; ++b) {
    const size_t cells = (size_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1);
    if (cells * (NBP + 1) * 8 > budget) return true;
  }
  return false;
}

}  // namespace tkfmi_band

using namespace tkfmi_band;

bool tkfmi_band_applies(const tkfmi_problem* p, uint32_t flags) {
  if (flags & TKFMI_CLASS_KERNEL) return false;  // the one-workgroup-per-pair kernel
  const BPlan pl = plan(p, flags);
  if (!pl.ok) return false;
  return !needs_checkpoint(p, flags, pl.NBP);
}

size_t tkfmi_band_workspace(const tkfmi_problem* p, uint32_t flags) {
  const BPlan pl = plan(p, flags);
  return pl.ok ? std::max<size_t>(pl.total(), 1) : 0;
}

int tkfmi_band_run(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                   hipStream_t stream, uint32_t flags, std::string& err) {
  const BPlan pl = plan(p, flags);
  if (!pl.ok) { err = pl.why; return TKFMI_EINVAL; }
  if (!ws || ws_bytes < pl.total()) { err = "workspace too small: need " + std::to_string(pl.total()); return TKFMI_EWORKSPACE; }
  const int B = p->B;
  if (B == 0) return TKFMI_OK;
  char* wsb = static_cast<char*>(ws);
  const size_t nf = pl.fwd.size(), nbw = pl.bwd.size();
  std::vector<char> blob(pl.idx_b + pl.desc_b, 0);
  memcpy(blob.data(), pl.bidx.data(), pl.NBP * 4);
  memcpy(blob.data() + pl.NBP * 4, pl.btype.data(), pl.NBP * 4);
  memcpy(blob.data() + 2 * pl.NBP * 4, pl.bgrp.data(), pl.NBP * 4);
  memcpy(blob.data() + pl.idx_b, pl.fwd.data(), sizeof(BDesc) * nf);
  memcpy(blob.data() + pl.idx_b + sizeof(BDesc) * nf, pl.bwd.data(), sizeof(BDesc) * nbw);
  int rc = (flags & TKFMI_REUSE_DESCRIPTORS) ? TKFMI_OK : tkfmi_stage_upload(wsb, blob.data(), blob.size(), stream, err);
  if (rc != TKFMI_OK) return rc;
  const int32_t* d_idx = reinterpret_cast<const int32_t*>(wsb);
  double* wsd = reinterpret_cast<double*>(wsb);
  double* pm = wsd + (pl.idx_b + pl.desc_b) / 8;
  double* tab = p->log_emit ? nullptr : wsd + (pl.idx_b + pl.desc_b + pl.pm_b) / 8;
  rc = tkfmi_wide_prep(p, pl.NB, pl.NBP, pl.e_idx, d_idx, pm, pl.G, tab, stream, err);
  if (rc != TKFMI_OK) return rc;
  double* pinfo = wsd + (pl.idx_b + pl.desc_b + pl.pm_b + pl.tab_b) / 8;
  int32_t* flagv = reinterpret_cast<int32_t*>(wsb + pl.idx_b + pl.desc_b + pl.pm_b + pl.tab_b + pl.pinfo_b);
  hipLaunchKernelGGL(zero_ints_kernel, dim3(std::min(1024, (pl.n_flags + 255) / 256)), dim3(256), 0, stream,
                     flagv, pl.n_flags);
  if (hipGetLastError() != hipSuccess) { err = "zero launch failed"; return TKFMI_EHIP; }
  BParams prm;
  memset(&prm, 0, sizeof(prm));
  prm.S = p->S; prm.NB = pl.NB; prm.NBP = pl.NBP; prm.flags = (int32_t)flags;
  prm.e_idx = pl.e_idx; prm.G = pl.G;
  {
    const char* e = getenv("TKFMI_CHAIN_SPIN_LIMIT");
    prm.spin_max = e ? std::max(0, atoi(e)) : kSpinMax;
  }
  prm.bidx = d_idx; prm.btype = d_idx + pl.NBP; prm.bgrp = d_idx + 2 * pl.NBP;
  prm.x_cat = p->x_cat; prm.y_cat = p->y_cat; prm.log_emit = p->log_emit;
  prm.ws = wsd;
  prm.flagv = flagv; prm.abortf = flagv + 2;
  prm.pinfo = pinfo;
  prm.out_logP = o->log_prob;
  prm.out_n = (flags & TKFMI_NO_COUNTS) ? nullptr : o->n_trans;
  prm.out_post = o->posteriors;
  prm.out_ecnt = (flags & TKFMI_WANT_EMIT_COUNTS) ? o->emit_counts : nullptr;
  int64_t cells = 0;
  for (int b = 0; b < B; ++b) cells += (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1);
  const bool dense = p->log_emit != nullptr;
  for (int dir = 0; dir < ((flags & TKFMI_FORWARD_ONLY) ? 1 : 2); ++dir) {
    BParams pc = prm;
    pc.ticket = flagv + dir;
    pc.desc = reinterpret_cast<const BDesc*>(wsb + pl.idx_b) + (dir ? nf : 0);
    const std::string name = std::string(dir ? "band_bwd_kernel<" : "band_fwd_kernel<") + std::to_string(pl.NT) + ">";
    const int slot = tkfmi_ktime_start(name.c_str(), B, cells, stream, err);
    if (slot == -2) return TKFMI_EHIP;
    rc = dense ? launch_d<true>(pl.NT, dir == 1, pc, (int)(dir ? nbw : nf), stream, err)
               : launch_d<false>(pl.NT, dir == 1, pc, (int)(dir ? nbw : nf), stream, err);
    if (rc != TKFMI_OK) return rc;
    if (tkfmi_ktime_stop(slot, stream, err) != TKFMI_OK) return TKFMI_EHIP;
  }
  if (o->status) {
    hipLaunchKernelGGL(raise_status_kernel, dim3(1), dim3(64), 0, stream, prm.abortf, o->status);
    if (hipGetLastError() != hipSuccess) { err = "status launch failed"; return TKFMI_EHIP; }
  }
  return TKFMI_OK;
}
