// tkfmi: cell-per-lane 2D pair-HMM forward/backward + expected counts for
// gfx950 (MI355X), for the TKF91/TKF92 (S,M,I,D,E) and MixDom (2 + 5KF state)
// layouts.  Reference algorithm: python/tkfmixdom/jax/dp/hmm.py (forward
// hmm.py:1186-1292, backward hmm.py:1295-1437, counts hmm.py:1930-1971).
//
// Design:
//  * A wave holds 64 consecutive lattice rows (a band); lane r is row 64b+r
//    and, at step t, column t - r (forward) -- a skewed antidiagonal.  All
//    NB emitting-state values of the lane's cell live in its registers, so
//    the per-cell bookkeeping (normalisation, emissions, addressing) is
//    amortised over 64 cells per instruction.
//  * The transition contraction h = F . P (forward) / B = P . g (backward)
//    is NB^2 FP64 FMAs per lane with P read through the scalar cache (the
//    matrix is uniform across the workgroup, so it is an SGPR operand).
//  * Row r receives the type-shifted predecessor values from lane r-1 with a
//    DPP wave_shr:1 (wave_shl:1 in the backward); lane 0 receives them from
//    the band above through a small LDS ring written by lane 63 of the wave
//    that holds that band.
//  * A pair is processed by one workgroup of W waves (W in {1,2,4,8}, chosen
//    per pair): wave w holds bands w, w+W, ...; each wave wraps to its next
//    band without draining.  Waves are staggered by 64 steps and advance in
//    lock-step (LDS-only barrier per step).
//  * Expected counts sum_c F[c,k] g_c[k'] are accumulated on the matrix
//    cores: each step the wave stages F and g of its 64 cells in LDS and
//    runs 16 v_mfma_f64_16x16x4f64 (K = 4 cells each) into a 16x16
//    accumulator.
//  * Probability space with one binary exponent per cell; FP64 throughout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fb2d_cell.h"
#include "fb2d_internal.h"
#include "tkfmi.h"

namespace tkfmi_v2 {

// -------------------------------------------------------------- params
struct V2Desc {
  int32_t pair, Lx, Ly, nb;  // nb = bands of 64 rows
  int32_t P;                 // steps per round (>= Ly + 1, >= 64 W)
  int32_t ring;              // entries of the long ring (wave W-1 -> wave 0)
  int32_t T_f, T_b;          // forward / backward step counts
  int32_t ring_glob, pad;    // 1: the long ring lives in global memory
  int64_t ring_off;          // doubles, into the global ring region
  int64_t x_off, y_off;
  int64_t lat_off;           // doubles, into the lattice region
  int64_t post_off;          // doubles, into the posterior output
  int64_t pm_off;            // doubles, into the transition block region
};

struct V2Params {
  int32_t S, flags, n_tab, G_tab;
  int32_t gmap[16];  // layout group -> caller table group
  int32_t smem_ring_off, smem_stage_off, smem_seq_off, smem_ecnt_off;
  const double* log_tab;
  const uint8_t* x_cat;
  const uint8_t* y_cat;
  const V2Desc* desc;
  const double* pmat;
  double* lattice;
  double* gring;
  double* out_logP;
  double* out_n;
  double* out_post;
  double* out_ecnt;
};


// Position of a lane along its wave's band stream, tracked incrementally:
// p = position, rho = round, c = column index within the round (0..P-1).
struct Pos {
  int p, rho, c;
  __device__ __forceinline__ void init(int p0, int P) {
    p = p0;
    if (p0 >= 0) { rho = p0 / P; c = p0 - rho * P; }
    else { rho = -1; c = p0; }  // c counts up to 0 before the stream starts
  }
  __device__ __forceinline__ void step(int P) {
    ++p;
    ++c;
    if (c == P) { c = 0; ++rho; }
    if (rho < 0 && c == 0) rho = 0;
  }
};

// ------------------------------------------------------------------ kernel
// The transition blocks and descriptors are separate __restrict__ const
// arguments so the compiler can read them through the scalar cache (uniform
// addresses, never written by the kernel): P becomes an SGPR operand of the
// FP64 FMAs instead of a per-step vector load.
template <class L, int W>
__global__ __launch_bounds__(W * 64) void fbv2_kernel(V2Params Pr, const double* __restrict__ pmat,
                                                      const V2Desc* __restrict__ descs) {
  constexpr int NB = L::NB, NG = L::NG;
  constexpr int nM = count_type<L, ST_M>();
  constexpr int nD = count_type<L, ST_D>();
  constexpr int NX = nM + nD + 1;  // ring entry: M parts, D parts, exponent
  static_assert(NB <= SLOTS - 1, "exponent slot");
  extern __shared__ double smem[];

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const V2Desc pd = descs[blockIdx.x];
  const int Lx = pd.Lx, Ly = pd.Ly, NBND = pd.nb, PP = pd.P, RING = pd.ring;
  const int S = Pr.S;
  const int W1 = Ly + 1;
  const int BSTRIDE = Ly + 64;  // forward-local steps per band in the lattice

  double* s_tab = smem;                               // NG * TAB_SZ
  double* s_misc = s_tab + NG * TAB_SZ;               // 8
  double* s_ring = smem + Pr.smem_ring_off;           // (W-1)*2*NX short + RING*NX long
  double* s_stage = smem + Pr.smem_stage_off;         // W * 1024 doubles
  double* s_ecnt = smem + Pr.smem_ecnt_off;           // NG * TAB_SZ emission counts
  uint8_t* s_x = reinterpret_cast<uint8_t*>(smem) + Pr.smem_seq_off;
  uint8_t* s_y = s_x + ((Lx + 8) & ~7);

  // ---- stage tables, sequences
  const double* ltab = Pr.log_tab + (Pr.n_tab > 1 ? (size_t)pd.pair * Pr.G_tab * TAB_SZ : 0);
  for (int q = threadIdx.x; q < NG * TAB_SZ; q += W * 64) {
    const int g = q / TAB_SZ, o = q - g * TAB_SZ;
    s_tab[q] = exp(ltab[Pr.gmap[g] * TAB_SZ + o]);
  }
  for (int q = threadIdx.x; q < Lx; q += W * 64) s_x[q] = min((int)Pr.x_cat[pd.x_off + q], A_PAD - 1);
  for (int q = threadIdx.x; q < Ly; q += W * 64) s_y[q] = min((int)Pr.y_cat[pd.y_off + q], A_PAD - 1);
  if (threadIdx.x == 0) { s_x[Lx] = 0; s_y[Ly] = 0; }
  const bool want_ecnt = !(Pr.flags & TKFMI_FORWARD_ONLY) && (Pr.flags & TKFMI_WANT_EMIT_COUNTS) &&
                         Pr.out_ecnt != nullptr;
  if (want_ecnt)
    for (int q = threadIdx.x; q < NG * TAB_SZ; q += W * 64) s_ecnt[q] = 0.0;

  const double* __restrict__ pm = pmat + pd.pm_off;
  const double* __restrict__ Pf = pm;                    // P[s][k]
  const double* __restrict__ Pb = pm + NB * NB;          // PT[k'][k]
  const double* __restrict__ p_start = pm + 2 * NB * NB;
  const double* __restrict__ p_end = p_start + NB;
  const double p_se = pm[2 * NB * NB + 2 * NB];

  const bool want_lat = !(Pr.flags & TKFMI_FORWARD_ONLY);
  const bool want_counts = want_lat && !(Pr.flags & TKFMI_NO_COUNTS) && Pr.out_n != nullptr;
  const bool want_post = want_lat && (Pr.flags & TKFMI_WANT_POSTERIORS) && Pr.out_post != nullptr;
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  double* lat = Pr.lattice + pd.lat_off;
  constexpr int EIDX = L::EIDX;

  // Ring entries: short rings (wave pw -> pw+1, lag one step, 2 slots) in
  // LDS; the long ring (wave W-1 -> wave 0 of the next round) in LDS, or in
  // global memory when it does not fit (ring_glob).  A global ring entry is
  // written at least QF+1 steps before it is read, and a full barrier (which
  // waits for global stores) runs every QF steps.
  constexpr int QF = 32;
  const bool rglob = pd.ring_glob != 0;
  double* g_ring = rglob ? Pr.gring + pd.ring_off : nullptr;
  __syncthreads();

  // ======================================================== forward sweep
  {
    Pos ps;
    ps.init(-64 * w - lane, PP);
    int slot_p = wrap_mod(ps.p, RING);        // producer slot, long ring
    int slot_c = wrap_mod(ps.p - PP, RING);   // consumer key (p - P), long ring
    double hp[NB];       // own h of the previous step
    double rM_last[nM];  // M-type parts received last step
#pragma unroll
    for (int b = 0; b < NB; ++b) hp[b] = 0.0;
#pragma unroll
    for (int m = 0; m < nM; ++m) rM_last[m] = 0.0;
    int eH = EXP_NONE, eM_last = EXP_NONE;
    const int T = pd.T_f;
    for (int t = 0; t < T; ++t) {
      const int band = w + W * ps.rho;
      const int j = ps.c;
      const int i = 64 * band + lane;
      const bool valid = ps.p >= 0 && band < NBND && j <= Ly && i <= Lx;
      // ---- receive the predecessor row's h (lane-1, or the ring for lane 0)
      double rD[nD > 0 ? nD : 1], rM[nM > 0 ? nM : 1];
      int eR;
      {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (L::type(b) == ST_D) rD[rank_in_type<L, ST_D>(b)] = dpp_f64<DPP_WAVE_SHR1>(hp[b]);
          if (L::type(b) == ST_M) rM[rank_in_type<L, ST_M>(b)] = dpp_f64<DPP_WAVE_SHR1>(hp[b]);
        }
        eR = dpp_i32<DPP_WAVE_SHR1>(eH);
      }
      if (lane == 0 && valid && i >= 1) {
        if (w == 0 && rglob) {
          const double* e = g_ring + (size_t)slot_c * NX;
#pragma unroll
          for (int m = 0; m < nM; ++m) rM[m] = e[m];
#pragma unroll
          for (int d = 0; d < nD; ++d) rD[d] = e[nM + d];
          eR = (int)e[NX - 1];
        } else {
          const double* e = (w >= 1) ? s_ring + ((size_t)(w - 1) * 2 + (ps.p & 1)) * NX
                                     : s_ring + ((size_t)(W - 1) * 2 + slot_c) * NX;
#pragma unroll
          for (int m = 0; m < nM; ++m) rM[m] = e[m];
#pragma unroll
          for (int d = 0; d < nD; ++d) rD[d] = e[nM + d];
          eR = (int)e[NX - 1];
        }
      }
      // ---- emissions at (i, j)
      const int xi = valid ? s_x[i >= 1 ? i - 1 : 0] : 0;
      const int yj = valid ? s_y[j >= 1 ? j - 1 : 0] : 0;
      double emM[NG], emI[NG], emD[NG];
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        emM[g] = s_tab[g * TAB_SZ + xi * A_PAD + yj];
        emI[g] = s_tab[g * TAB_SZ + A_PAD * A_PAD + yj];
        emD[g] = s_tab[g * TAB_SZ + A_PAD * A_PAD + A_PAD + xi];
      }
      const bool okM = valid && i >= 1 && j >= 1;
      const bool okI = valid && j >= 1;
      const bool okD = valid && i >= 1;
      // ---- new cell: raw products and per-type maxima
      double f[NB];
      double mxM = 0.0, mxI = 0.0, mxD = 0.0;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        constexpr int dummy = 0;
        (void)dummy;
        const int ty = L::type(b), g = L::group(b);
        double v;
        if (ty == ST_M) { v = okM ? rM_last[rank_in_type<L, ST_M>(b)] * emM[g] : 0.0; mxM = fmax(mxM, v); }
        else if (ty == ST_I) { v = okI ? hp[b] * emI[g] : 0.0; mxI = fmax(mxI, v); }
        else { v = okD ? rD[rank_in_type<L, ST_D>(b)] * emD[g] : 0.0; mxD = fmax(mxD, v); }
        f[b] = v;
      }
      const int cM = (mxM > 0.0) ? eM_last + frexp_exp(mxM) : EXP_NONE;
      const int cI = (mxI > 0.0) ? eH + frexp_exp(mxI) : EXP_NONE;
      const int cD = (mxD > 0.0) ? eR + frexp_exp(mxD) : EXP_NONE;
      int E = max(cM, max(cI, cD));
      if (E == EXP_NONE) E = 0;
      const int shM = eM_last - E, shI = eH - E, shD = eR - E;
      double F[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ty = L::type(b);
        F[b] = ldexp(f[b], ty == ST_M ? shM : (ty == ST_I ? shI : shD));
      }
      const bool origin = valid && i == 0 && j == 0;
      if (origin) {
#pragma unroll
        for (int b = 0; b < NB; ++b) F[b] = 0.0;
        E = 0;
      }
      // ---- store the forward lattice (band-local step u = j + lane)
      if (want_lat && valid) {
        double* dst = lat + (((size_t)band * BSTRIDE + (j + lane)) * SLOTS) * 64 + lane;
#pragma unroll
        for (int b = 0; b < NB; ++b) dst[b * 64] = F[b];
        dst[(SLOTS - 1) * 64] = (double)E;
      }
      // ---- push: h[k] = sum_s F[s] P[s][k]
      double h[NB];
#pragma unroll
      for (int k = 0; k < NB; ++k) h[k] = 0.0;
#pragma unroll
      for (int s = 0; s < NB; ++s) {
        const_dptr row = launder(Pf + s * NB);
#pragma unroll
        for (int k = 0; k < NB; ++k) h[k] = fma(F[s], row[k], h[k]);
      }
      if (origin) {
#pragma unroll
        for (int k = 0; k < NB; ++k) h[k] = p_start[k];
      }
      // ---- corner: log P and the E column
      if (valid && i == Lx && j == Ly) {
        double tot = 0.0;
#pragma unroll
        for (int b = 0; b < NB; ++b) tot = fma(F[b], p_end[b], tot);
        double lp;
        if (origin) { tot = p_se; lp = log(p_se); }
        else lp = log(tot) + (double)E * LN2;
        const int ep = (tot > 0.0) ? E + frexp_exp(tot) : 0;
        const double mp = (tot > 0.0) ? ldexp(tot, -frexp_exp(tot)) : 0.0;
        Pr.out_logP[pd.pair] = (tot > 0.0) ? lp : -1e30;
        s_misc[0] = mp;
        s_misc[1] = (double)ep;
        if (want_counts && tot > 0.0) {
          if (origin) out_n[EIDX] = 1.0;
          else {
#pragma unroll
            for (int b = 0; b < NB; ++b) out_n[(b + L::SOFF) * S + EIDX] = F[b] * p_end[b] / tot;
          }
        }
      }
      // ---- publish this row's h for the band below (lane 63)
      if (lane == 63 && valid) {
        if (w == W - 1 && rglob) {
          double* e = g_ring + (size_t)slot_p * NX;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (L::type(b) == ST_M) e[rank_in_type<L, ST_M>(b)] = h[b];
            if (L::type(b) == ST_D) e[nM + rank_in_type<L, ST_D>(b)] = h[b];
          }
          e[NX - 1] = (double)E;
        } else {
          double* e = (w == W - 1) ? s_ring + ((size_t)(W - 1) * 2 + slot_p) * NX
                                   : s_ring + ((size_t)w * 2 + (ps.p & 1)) * NX;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (L::type(b) == ST_M) e[rank_in_type<L, ST_M>(b)] = h[b];
            if (L::type(b) == ST_D) e[nM + rank_in_type<L, ST_D>(b)] = h[b];
          }
          e[NX - 1] = (double)E;
        }
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) hp[b] = h[b];
#pragma unroll
      for (int m = 0; m < nM; ++m) rM_last[m] = rM[m];
      eH = E;
      eM_last = eR;
      ps.step(PP);
      if (++slot_p == RING) slot_p = 0;
      if (++slot_c == RING) slot_c = 0;
      if (rglob && (t & (QF - 1)) == QF - 1) __syncthreads();
      else lds_barrier();
    }
  }
  if (!want_lat) return;
  __syncthreads();  // forward lattice (global) and log P (LDS) visible to all waves

  // ======================================================= backward sweep
  const double mP = s_misc[0];
  const int EP = (int)s_misc[1];
  if (!(mP > 0.0)) return;  // impossible pair
  const double inv_mP = 1.0 / mP;
  d4 cacc = {0.0, 0.0, 0.0, 0.0};
  double* stage_f = s_stage + (size_t)w * 1024;  // [32 cells][16] F-scaled
  double* stage_g = stage_f + 512;               // [32 cells][16] g
  double* post = want_post ? Pr.out_post + pd.post_off : nullptr;
  {
    Pos ps;
    ps.init(-64 * w - (63 - lane), PP);
    int slot_p = wrap_mod(ps.p, RING);
    int slot_c = wrap_mod(ps.p - PP, RING);
    double qp[NB];
    double rM_last[nM];
#pragma unroll
    for (int b = 0; b < NB; ++b) qp[b] = 0.0;
#pragma unroll
    for (int m = 0; m < nM; ++m) rM_last[m] = 0.0;
    int eQ = EXP_NONE, eM_last = EXP_NONE;
    // prefetch of the forward lattice, one step ahead
    auto f_addr = [&](const Pos& q) -> const double* {
      const int band = NBND - 1 - (w + W * q.rho);
      const int j = Ly - q.c;
      const int i = 64 * band + lane;
      const bool v = q.p >= 0 && band >= 0 && q.c <= Ly && i <= Lx;
      return v ? lat + (((size_t)band * BSTRIDE + (j + lane)) * SLOTS) * 64 + lane : nullptr;
    };
    double fnext[SLOTS];
    {
      const double* a = f_addr(ps);
#pragma unroll
      for (int b = 0; b < NB; ++b) fnext[b] = a ? a[b * 64] : 0.0;
      fnext[SLOTS - 1] = a ? a[(SLOTS - 1) * 64] : 0.0;
    }
    const int T = pd.T_b;
    for (int t = 0; t < T; ++t) {
      const int band = NBND - 1 - (w + W * ps.rho);
      const int j = Ly - ps.c;
      const int i = 64 * band + lane;
      const bool valid = ps.p >= 0 && band >= 0 && ps.c <= Ly && i <= Lx;
      double fcur[SLOTS];
#pragma unroll
      for (int b = 0; b < NB; ++b) fcur[b] = fnext[b];
      fcur[SLOTS - 1] = fnext[SLOTS - 1];
      {
        Pos nq = ps;
        nq.step(PP);
        const double* a = f_addr(nq);
#pragma unroll
        for (int b = 0; b < NB; ++b) fnext[b] = a ? a[b * 64] : 0.0;
        fnext[SLOTS - 1] = a ? a[(SLOTS - 1) * 64] : 0.0;
      }
      // ---- receive the successor row's q (lane+1, or the ring for lane 63)
      double rD[nD > 0 ? nD : 1], rM[nM > 0 ? nM : 1];
      int eR;
      {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (L::type(b) == ST_D) rD[rank_in_type<L, ST_D>(b)] = dpp_f64<DPP_WAVE_SHL1>(qp[b]);
          if (L::type(b) == ST_M) rM[rank_in_type<L, ST_M>(b)] = dpp_f64<DPP_WAVE_SHL1>(qp[b]);
        }
        eR = dpp_i32<DPP_WAVE_SHL1>(eQ);
      }
      if (lane == 63 && valid && i < Lx) {
        if (w == 0 && rglob) {
          const double* e = g_ring + (size_t)slot_c * NX;
#pragma unroll
          for (int m = 0; m < nM; ++m) rM[m] = e[m];
#pragma unroll
          for (int d = 0; d < nD; ++d) rD[d] = e[nM + d];
          eR = (int)e[NX - 1];
        } else {
          const double* e = (w >= 1) ? s_ring + ((size_t)(w - 1) * 2 + (ps.p & 1)) * NX
                                     : s_ring + ((size_t)(W - 1) * 2 + slot_c) * NX;
#pragma unroll
          for (int m = 0; m < nM; ++m) rM[m] = e[m];
#pragma unroll
          for (int d = 0; d < nD; ++d) rD[d] = e[nM + d];
          eR = (int)e[NX - 1];
        }
      }
      const bool okM = valid && i < Lx && j < Ly;
      const bool okI = valid && j < Ly;
      const bool okD = valid && i < Lx;
      const int cM = okM ? eM_last : EXP_NONE;
      const int cI = okI ? eQ : EXP_NONE;
      const int cD = okD ? eR : EXP_NONE;
      int Eg = max(cM, max(cI, cD));
      if (Eg == EXP_NONE) Eg = 0;
      const int shM = eM_last - Eg, shI = eQ - Eg, shD = eR - Eg;
      double g[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ty = L::type(b);
        double v;
        if (ty == ST_M) v = okM ? ldexp(rM_last[rank_in_type<L, ST_M>(b)], shM) : 0.0;
        else if (ty == ST_I) v = okI ? ldexp(qp[b], shI) : 0.0;
        else v = okD ? ldexp(rD[rank_in_type<L, ST_D>(b)], shD) : 0.0;
        g[b] = v;
      }
      // ---- B[k] = sum_k' P[k][k'] g[k']
      double Bv[NB];
#pragma unroll
      for (int k = 0; k < NB; ++k) Bv[k] = 0.0;
#pragma unroll
      for (int k2 = 0; k2 < NB; ++k2) {
        const_dptr row = launder(Pb + k2 * NB);
#pragma unroll
        for (int k = 0; k < NB; ++k) Bv[k] = fma(row[k], g[k2], Bv[k]);
      }
      const bool corner = valid && i == Lx && j == Ly;
      int EBraw = Eg;
      if (corner) {
#pragma unroll
        for (int k = 0; k < NB; ++k) Bv[k] = p_end[k];
        EBraw = 0;
      }
      double mx = 0.0;
#pragma unroll
      for (int k = 0; k < NB; ++k) mx = fmax(mx, Bv[k]);
      const int sh = (mx > 0.0) ? frexp_exp(mx) : 0;
      const int EB = EBraw + sh;
#pragma unroll
      for (int k = 0; k < NB; ++k) Bv[k] = ldexp(Bv[k], -sh);
      // ---- counts: stage F-scaled and g, then MFMA (K = 4 cells per op)
      const int EF = (int)fcur[SLOTS - 1];
      const double fscale = valid ? ldexp(inv_mP, EF + Eg - EP) : 0.0;
      if (want_counts) {
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          if ((lane >> 5) == half) {
            const int c = lane & 31;
#pragma unroll
            for (int s = 0; s < 16; ++s) {
              const int o = c * 16 + (s ^ (c & 15));
              stage_f[o] = (s < NB) ? fcur[s < NB ? s : 0] * fscale : 0.0;
              stage_g[o] = (s < NB) ? g[s < NB ? s : 0] : 0.0;
            }
          }
          wave_lds_sync();
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int c = 4 * q + (lane >> 4);
            const int o = c * 16 + ((lane & 15) ^ (c & 15));
            cacc = __builtin_amdgcn_mfma_f64_16x16x4f64(stage_f[o], stage_g[o], cacc, 0, 0, 0);
          }
          wave_lds_sync();
        }
        if (valid && i == 0 && j == 0) {  // start state row: n[S, k]
          const double sc = ldexp(inv_mP, Eg - EP);
#pragma unroll
          for (int b = 0; b < NB; ++b) out_n[b + L::SOFF] = p_start[b] * g[b] * sc;
        }
      }
      const int xi = valid ? s_x[i >= 1 ? i - 1 : 0] : 0;
      const int yj = valid ? s_y[j >= 1 ? j - 1 : 0] : 0;
      // ---- expected emission counts: posterior mass per (group, type) onto
      // the table entry the cell's emission used
      if (want_ecnt && valid) {
        const double pscale = ldexp(inv_mP, EF + EB - EP);
        double sM[NG], sI[NG], sD[NG];
#pragma unroll
        for (int gq = 0; gq < NG; ++gq) { sM[gq] = 0.0; sI[gq] = 0.0; sD[gq] = 0.0; }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int ty = L::type(b), gq = L::group(b);
          const double pv = fcur[b] * Bv[b];
          if (ty == ST_M) sM[gq] += pv;
          else if (ty == ST_I) sI[gq] += pv;
          else sD[gq] += pv;
        }
#pragma unroll
        for (int gq = 0; gq < NG; ++gq) {
          double* tb = s_ecnt + gq * TAB_SZ;
          if (i >= 1 && j >= 1) atomicAdd(tb + xi * A_PAD + yj, sM[gq] * pscale);
          if (j >= 1) atomicAdd(tb + A_PAD * A_PAD + yj, sI[gq] * pscale);
          if (i >= 1) atomicAdd(tb + A_PAD * A_PAD + A_PAD + xi, sD[gq] * pscale);
        }
      }
      if (want_post && valid) {
        const double ps_ = ldexp(inv_mP, EF + EB - EP);
        double* dst = post + ((size_t)i * W1 + j) * S + L::SOFF;
#pragma unroll
        for (int b = 0; b < NB; ++b) dst[b] = fcur[b] * Bv[b] * ps_;
        if (i == 0 && j == 0) {
          double b00 = 0.0;
#pragma unroll
          for (int b = 0; b < NB; ++b) b00 = fma(p_start[b], g[b], b00);
          post[0] = corner ? 1.0 : b00 * ldexp(inv_mP, Eg - EP);
        }
      }
      // ---- q = e(i,j,k) B(i,j,k) for the predecessor row
      const bool eM_ok = valid && i >= 1 && j >= 1, eI_ok = valid && j >= 1, eD_ok = valid && i >= 1;
      double q[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ty = L::type(b), gq = L::group(b);
        double em;
        bool ok;
        if (ty == ST_M) { em = s_tab[gq * TAB_SZ + xi * A_PAD + yj]; ok = eM_ok; }
        else if (ty == ST_I) { em = s_tab[gq * TAB_SZ + A_PAD * A_PAD + yj]; ok = eI_ok; }
        else { em = s_tab[gq * TAB_SZ + A_PAD * A_PAD + A_PAD + xi]; ok = eD_ok; }
        q[b] = ok ? em * Bv[b] : 0.0;
      }
      // ---- publish this row's q for the band above (lane 0)
      if (lane == 0 && valid) {
        if (w == W - 1 && rglob) {
          double* e = g_ring + (size_t)slot_p * NX;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (L::type(b) == ST_M) e[rank_in_type<L, ST_M>(b)] = q[b];
            if (L::type(b) == ST_D) e[nM + rank_in_type<L, ST_D>(b)] = q[b];
          }
          e[NX - 1] = (double)EB;
        } else {
          double* e = (w == W - 1) ? s_ring + ((size_t)(W - 1) * 2 + slot_p) * NX
                                   : s_ring + ((size_t)w * 2 + (ps.p & 1)) * NX;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (L::type(b) == ST_M) e[rank_in_type<L, ST_M>(b)] = q[b];
            if (L::type(b) == ST_D) e[nM + rank_in_type<L, ST_D>(b)] = q[b];
          }
          e[NX - 1] = (double)EB;
        }
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) qp[b] = q[b];
#pragma unroll
      for (int m = 0; m < nM; ++m) rM_last[m] = rM[m];
      eQ = EB;
      eM_last = eR;
      ps.step(PP);
      if (++slot_p == RING) slot_p = 0;
      if (++slot_c == RING) slot_c = 0;
      if (rglob && (t & (QF - 1)) == QF - 1) __syncthreads();
      else lds_barrier();
    }
  }
  __syncthreads();
  if (want_ecnt) {  // layout group g -> caller group gmap[g]; each entry owned by one thread
    double* oe = Pr.out_ecnt + (size_t)pd.pair * Pr.G_tab * TAB_SZ;
    for (int q = threadIdx.x; q < TAB_SZ; q += W * 64) {
#pragma unroll
      for (int gq = 0; gq < NG; ++gq) oe[Pr.gmap[gq] * TAB_SZ + q] += s_ecnt[gq * TAB_SZ + q];
    }
  }
  if (!want_counts) return;
  // ---- reduce the MFMA accumulators over the waves (fixed order)
  double* s_red = s_stage;  // [W][16][16]
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = 4 * v + (lane >> 4), col = lane & 15;  // accumulator layout
    s_red[(w * 16 + row) * 16 + col] = cacc[v];
  }
  __syncthreads();
  for (int qq = threadIdx.x; qq < NB * NB; qq += W * 64) {
    const int k = qq / NB, k2 = qq % NB;
    double acc = 0.0;
    for (int ww = 0; ww < W; ++ww) acc += s_red[(ww * 16 + k) * 16 + k2];
    const double pr = Pf[k * NB + k2];
    out_n[(k + L::SOFF) * S + (k2 + L::SOFF)] = (pr > 0.0) ? pr * acc : 0.0;
  }
}

// ---------------------------------------------------------------- host side

int choose_W(int Lx, int Ly) {
  const int nb = (Lx + 64) / 64;
  int W = std::min(std::min(8, nb), std::max(1, (Ly + 1) / 64));
  if (W >= 8) return 8;
  if (W >= 4) return 4;
  if (W >= 2) return 2;
  return 1;
}

struct PairGeom {
  int W, nb, P, ring, T_f, T_b;
  size_t lat;  // doubles
};

PairGeom geom(int Lx, int Ly) {
  PairGeom g;
  g.W = choose_W(Lx, Ly);
  g.nb = (Lx + 64) / 64;
  g.P = std::max(Ly + 1, 64 * g.W);
  g.ring = (g.nb > g.W) ? g.P - 64 * g.W + 2 : 2;  // the long ring is used only on wrap
  const int bl = g.nb - 1, wl = bl % g.W, rl = bl / g.W, r_last = Lx - 64 * bl;
  g.T_f = rl * g.P + Ly + 64 * wl + r_last + 1;
  const int w0 = (g.nb - 1) % g.W, r0 = (g.nb - 1) / g.W;
  g.T_b = r0 * g.P + Ly + 64 * w0 + 64;
  g.lat = (size_t)g.nb * (Ly + 64) * SLOTS * 64;
  return g;
}



std::mutex g_streams_mu;
Streams g_stream_sets[64][kStreamSets];

int get_streams(hipStream_t caller, Streams*& out, std::string& err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) { err = "hipGetDevice failed"; return TKFMI_EHIP; }
  std::lock_guard<std::mutex> lk(g_streams_mu);
  Streams* row = g_stream_sets[dev];
  for (int i = 0; i < kStreamSets; ++i)
    if (row[i].init && row[i].key == caller) { out = &row[i]; return TKFMI_OK; }
  for (int i = 0; i < kStreamSets; ++i) {
    Streams& st = row[i];
    if (st.init) continue;
    for (int c = 0; c < 4; ++c) {
      if (hipStreamCreateWithFlags(&st.s[c], hipStreamNonBlocking) != hipSuccess) { err = "hipStreamCreate failed"; return TKFMI_EHIP; }
      if (hipEventCreateWithFlags(&st.done[c], hipEventDisableTiming) != hipSuccess) { err = "hipEventCreate failed"; return TKFMI_EHIP; }
    }
    if (hipEventCreateWithFlags(&st.start, hipEventDisableTiming) != hipSuccess) { err = "hipEventCreate failed"; return TKFMI_EHIP; }
    st.key = caller;
    st.init = true;
    out = &st;
    return TKFMI_OK;
  }
  out = &row[(reinterpret_cast<uintptr_t>(caller) >> 8) % kStreamSets];  // table full: share a set
  return TKFMI_OK;
}

template <class L>
int launch_class(int W, const V2Params& prm, int nblk, size_t smem, hipStream_t s, std::string& err) {
  const void* fn;
  switch (W) {
    case 1: fn = (const void*)fbv2_kernel<L, 1>; break;
    case 2: fn = (const void*)fbv2_kernel<L, 2>; break;
    case 4: fn = (const void*)fbv2_kernel<L, 4>; break;
    default: fn = (const void*)fbv2_kernel<L, 8>; break;
  }
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) { err = std::string("hipFuncSetAttribute: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  switch (W) {
    case 1: hipLaunchKernelGGL((fbv2_kernel<L, 1>), dim3(nblk), dim3(64), smem, s, prm, (const double*)prm.pmat, prm.desc); break;
    case 2: hipLaunchKernelGGL((fbv2_kernel<L, 2>), dim3(nblk), dim3(128), smem, s, prm, (const double*)prm.pmat, prm.desc); break;
    case 4: hipLaunchKernelGGL((fbv2_kernel<L, 4>), dim3(nblk), dim3(256), smem, s, prm, (const double*)prm.pmat, prm.desc); break;
    default: hipLaunchKernelGGL((fbv2_kernel<L, 8>), dim3(nblk), dim3(512), smem, s, prm, (const double*)prm.pmat, prm.desc); break;
  }
  e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("fbv2 launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  return TKFMI_OK;
}


// Full plan of one call: per pair geometry, classes by W, LDS layout per
// class, global-ring placement and workspace offsets.
struct CallPlan {
  bool ok = false;
  std::string why;
  std::vector<int> order;          // processing order (class, then cost)
  std::vector<V2Desc> desc;        // in processing order
  int64_t cls_cells[4] = {0, 0, 0, 0};
  int cls_begin[4] = {0, 0, 0, 0}, cls_count[4] = {0, 0, 0, 0};
  size_t smem[4] = {0, 0, 0, 0};
  int ring_off[4] = {0, 0, 0, 0}, stage_off[4] = {0, 0, 0, 0}, seq_off[4] = {0, 0, 0, 0};
  int ecnt_off[4] = {0, 0, 0, 0};
  size_t desc_b = 0, pm_b = 0, lat_b = 0, gring_b = 0;
  size_t total() const { return desc_b + pm_b + lat_b + gring_b; }
};

constexpr int kWs[4] = {8, 4, 2, 1};
int cls_of(int W) { return W == 8 ? 0 : (W == 4 ? 1 : (W == 2 ? 2 : 3)); }
constexpr size_t kLds = 160 * 1024;

size_t lds_layout(const Plan& pl, int W, int ring_cap, int max_lx, int max_ly, bool ecnt,
                  int& ring_off, int& stage_off, int& seq_off, int& ecnt_off) {
  const int NX = nx_of(pl);
  size_t d = (size_t)pl.NG * TAB_SZ + 8;
  ecnt_off = (int)d;
  if (ecnt) d += (size_t)pl.NG * TAB_SZ;
  ring_off = (int)d;
  d += (size_t)(W - 1) * 2 * NX + (size_t)ring_cap * NX;
  d = (d + 1) & ~(size_t)1;
  stage_off = (int)d;
  d += (size_t)W * 1024;
  seq_off = (int)(d * 8);
  return d * 8 + (size_t)((max_lx + 8) & ~7) + (size_t)((max_ly + 8) & ~7) + 16;
}

CallPlan plan_call(const tkfmi_problem* p, const Plan& pl, uint32_t flags) {
  CallPlan cp;
  const int B = p->B;
  const int NX = nx_of(pl);
  std::vector<PairGeom> gm(B);
  for (int b = 0; b < B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    if (lx < 0 || ly < 0 || lx + ly > 32000) { cp.why = "sequence lengths out of range"; return cp; }
    gm[b] = geom(lx, ly);
  }
  cp.order.resize(B);
  for (int b = 0; b < B; ++b) cp.order[b] = b;
  std::stable_sort(cp.order.begin(), cp.order.end(), [&](int a, int b) {
    if (gm[a].W != gm[b].W) return gm[a].W > gm[b].W;
    const int64_t ca = (int64_t)(p->h_x_len[a] + 1) * (p->h_y_len[a] + 1);
    const int64_t cb = (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1);
    return ca > cb;
  });
  // per class maxima
  int max_lx[4] = {0, 0, 0, 0}, max_ly[4] = {0, 0, 0, 0}, max_ring[4] = {2, 2, 2, 2};
  for (int q = 0; q < B; ++q) {
    const int b = cp.order[q];
    const int c = cls_of(gm[b].W);
    if (cp.cls_count[c] == 0) cp.cls_begin[c] = q;
    ++cp.cls_count[c];
    max_lx[c] = std::max(max_lx[c], p->h_x_len[b]);
    max_ly[c] = std::max(max_ly[c], p->h_y_len[b]);
    max_ring[c] = std::max(max_ring[c], gm[b].ring);
  }
  int ring_cap[4] = {2, 2, 2, 2};
  const bool ecnt = (flags & TKFMI_WANT_EMIT_COUNTS) && !(flags & TKFMI_FORWARD_ONLY);
  for (int c = 0; c < 4; ++c) {
    if (!cp.cls_count[c]) continue;
    int ro, so, qo, eo;
    const size_t base = lds_layout(pl, kWs[c], 0, max_lx[c], max_ly[c], ecnt, ro, so, qo, eo);
    if (base + (size_t)34 * NX * 8 > kLds) { cp.why = "LDS budget exceeded (sequences too long)"; return cp; }
    const size_t room = (kLds - base) / ((size_t)NX * 8);
    ring_cap[c] = (int)std::min<size_t>((size_t)max_ring[c], room);
    cp.smem[c] = lds_layout(pl, kWs[c], ring_cap[c], max_lx[c], max_ly[c], ecnt, cp.ring_off[c],
                            cp.stage_off[c], cp.seq_off[c], cp.ecnt_off[c]);
  }
  // The 8-wave class holds the longest pairs, whose serial step chains set
  // the makespan of the call.  When it has few pairs, its workgroups ask for
  // the whole LDS so that no other workgroup shares their CUs and slows their
  // steps (the other classes have ample CUs left).
  static const bool no_reserve = getenv("TKFMI_NO_CU_RESERVE") != nullptr;
  if (!no_reserve && cp.cls_count[0] > 0 && cp.cls_count[0] <= 128) cp.smem[0] = kLds;
  // descriptors and workspace offsets
  cp.desc.resize(B);
  std::vector<int64_t> lat_off(B), post_off(B);
  int64_t al = 0, ap = 0;
  for (int b = 0; b < B; ++b) {
    lat_off[b] = al;
    if (!(flags & TKFMI_FORWARD_ONLY)) al += (int64_t)gm[b].lat;
    post_off[b] = ap;
    ap += (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1) * p->S;
  }
  int64_t agr = 0;
  for (int q = 0; q < B; ++q) {
    const int b = cp.order[q];
    const PairGeom& g = gm[b];
    V2Desc d;
    memset(&d, 0, sizeof(d));
    d.pair = b; d.Lx = p->h_x_len[b]; d.Ly = p->h_y_len[b]; d.nb = g.nb;
    d.P = g.P; d.ring = g.ring; d.T_f = g.T_f; d.T_b = g.T_b;
    d.ring_glob = g.ring > ring_cap[cls_of(g.W)] ? 1 : 0;
    d.ring_off = d.ring_glob ? agr : 0;
    if (d.ring_glob) agr += (int64_t)g.ring * NX;
    d.x_off = p->h_x_off[b]; d.y_off = p->h_y_off[b];
    d.lat_off = lat_off[b]; d.post_off = post_off[b];
    d.pm_off = (p->n_trans > 1 ? (int64_t)b : 0) * (int64_t)stride_for(pl.NB);
    cp.desc[q] = d;
    cp.cls_cells[cls_of(g.W)] += (int64_t)(d.Lx + 1) * (d.Ly + 1);
  }
  cp.desc_b = align256(sizeof(V2Desc) * (size_t)std::max(B, 1));
  cp.pm_b = align256(stride_for(pl.NB) * 8 * (size_t)p->n_trans);
  cp.lat_b = align256((size_t)al * 8);
  cp.gring_b = align256((size_t)agr * 8);
  cp.ok = true;
  return cp;
}

}  // namespace tkfmi_v2

using namespace tkfmi_v2;

// Returns true (and fills *bytes) when the cell-per-lane path handles `p`.
bool tkfmi_v2_workspace(const tkfmi_problem* p, uint32_t flags, size_t* bytes) {
  Plan pl;
  if (!p || p->B < 0 || !p->state_types || !p->state_group || !detect(p, pl)) return false;
  if (p->B > 0 && (!p->h_x_len || !p->h_y_len || !p->h_x_off || !p->h_y_off)) return false;
  if (p->n_trans != 1 && p->n_trans != p->B) return false;
  const CallPlan cp = plan_call(p, pl, flags);
  if (!cp.ok) return false;
  *bytes = cp.total();
  return true;
}

int tkfmi_v2_run(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                 hipStream_t stream, uint32_t flags, std::string& err) {
  Plan pl;
  if (!detect(p, pl)) { err = "layout not handled"; return TKFMI_EINVAL; }
  const CallPlan cp = plan_call(p, pl, flags);
  if (!cp.ok) { err = cp.why; return TKFMI_EINVAL; }
  if (!ws || ws_bytes < cp.total()) {
    err = "workspace too small: need " + std::to_string(cp.total());
    return TKFMI_EWORKSPACE;
  }
  const int B = p->B;
  if (B == 0) return TKFMI_OK;
  char* wsb = static_cast<char*>(ws);
  V2Desc* ddesc = reinterpret_cast<V2Desc*>(wsb);
  double* pmat = reinterpret_cast<double*>(wsb + cp.desc_b);
  double* lattice = reinterpret_cast<double*>(wsb + cp.desc_b + cp.pm_b);
  double* gring = reinterpret_cast<double*>(wsb + cp.desc_b + cp.pm_b + cp.lat_b);
  // library class streams, or (allocation-free calls) the caller's stream only
  const bool one_stream = tkfmi_noalloc_active();
  Streams* sts = nullptr;
  int rc = TKFMI_OK;
  std::unique_lock<std::mutex> lk;
  if (!one_stream) {
    rc = get_streams(stream, sts, err);
    if (rc != TKFMI_OK) return rc;
    lk = std::unique_lock<std::mutex>(sts->mu);  // until the last join below
  }
  rc = (flags & TKFMI_REUSE_DESCRIPTORS) ? TKFMI_OK : tkfmi_stage_upload(ddesc, cp.desc.data(), sizeof(V2Desc) * (size_t)B, stream, err);
  if (rc != TKFMI_OK) return rc;
  switch (pl.lay) {
    case Layout::Tkf: rc = launch_prep<LTkf>(p, pmat, stream, err); break;
    case Layout::Mix11: rc = launch_prep<LMix<1, 1>>(p, pmat, stream, err); break;
    case Layout::Mix21: rc = launch_prep<LMix<2, 1>>(p, pmat, stream, err); break;
    case Layout::Mix31: rc = launch_prep<LMix<3, 1>>(p, pmat, stream, err); break;
    case Layout::Mix12: rc = launch_prep<LMix<1, 2>>(p, pmat, stream, err); break;
    case Layout::Mix13: rc = launch_prep<LMix<1, 3>>(p, pmat, stream, err); break;
    default: rc = TKFMI_EINVAL;
  }
  if (rc != TKFMI_OK) return rc;
  if (!one_stream && hipEventRecord(sts->start, stream) != hipSuccess) { err = "hipEventRecord failed"; return TKFMI_EHIP; }
  V2Params prm;
  memset(&prm, 0, sizeof(prm));
  prm.S = p->S; prm.flags = (int32_t)flags; prm.n_tab = p->n_tab; prm.G_tab = p->n_groups;
  for (int g = 0; g < 16; ++g) prm.gmap[g] = pl.gmap[g];
  prm.log_tab = p->log_tab; prm.x_cat = p->x_cat; prm.y_cat = p->y_cat;
  prm.pmat = pmat; prm.lattice = lattice; prm.gring = gring;
  prm.out_logP = o->log_prob;
  prm.out_n = (flags & TKFMI_NO_COUNTS) ? nullptr : o->n_trans;
  prm.out_post = o->posteriors;
  prm.out_ecnt = (flags & TKFMI_WANT_EMIT_COUNTS) ? o->emit_counts : nullptr;
  for (int c = 0; c < 4; ++c) {
    if (!cp.cls_count[c]) continue;
    hipStream_t s = one_stream ? stream : sts->s[c];
    if (!one_stream && hipStreamWaitEvent(s, sts->start, 0) != hipSuccess) { err = "hipStreamWaitEvent failed"; return TKFMI_EHIP; }
    V2Params pc = prm;
    pc.desc = ddesc + cp.cls_begin[c];
    pc.smem_ring_off = cp.ring_off[c]; pc.smem_stage_off = cp.stage_off[c]; pc.smem_seq_off = cp.seq_off[c];
    pc.smem_ecnt_off = cp.ecnt_off[c];
    const std::string kname = std::string("fbv2_kernel<") + layout_name(pl) + "," + std::to_string(kWs[c]) + ">";
    const int slot = tkfmi_ktime_start(kname.c_str(), cp.cls_count[c], cp.cls_cells[c], s, err);
    if (slot == -2) return TKFMI_EHIP;
    switch (pl.lay) {
      case Layout::Tkf: rc = launch_class<LTkf>(kWs[c], pc, cp.cls_count[c], cp.smem[c], s, err); break;
      case Layout::Mix11: rc = launch_class<LMix<1, 1>>(kWs[c], pc, cp.cls_count[c], cp.smem[c], s, err); break;
      case Layout::Mix21: rc = launch_class<LMix<2, 1>>(kWs[c], pc, cp.cls_count[c], cp.smem[c], s, err); break;
      case Layout::Mix31: rc = launch_class<LMix<3, 1>>(kWs[c], pc, cp.cls_count[c], cp.smem[c], s, err); break;
      case Layout::Mix12: rc = launch_class<LMix<1, 2>>(kWs[c], pc, cp.cls_count[c], cp.smem[c], s, err); break;
      case Layout::Mix13: rc = launch_class<LMix<1, 3>>(kWs[c], pc, cp.cls_count[c], cp.smem[c], s, err); break;
      default: rc = TKFMI_EINVAL;
    }
    if (rc != TKFMI_OK) return rc;
    if (tkfmi_ktime_stop(slot, s, err) != TKFMI_OK) return TKFMI_EHIP;
    if (one_stream) continue;
    if (hipEventRecord(sts->done[c], s) != hipSuccess) { err = "hipEventRecord failed"; return TKFMI_EHIP; }
    if (hipStreamWaitEvent(stream, sts->done[c], 0) != hipSuccess) { err = "hipStreamWaitEvent failed"; return TKFMI_EHIP; }
  }
  return TKFMI_OK;
}
